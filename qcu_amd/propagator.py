"""Quark propagators: 12 spin-colour solves as one batched solve.

Mirrors PyQuda's propagator driver (python/PyQuda/pyquda/core.py:44-75,
``invert`` and ``invert12``) and its sources (python/PyQuda/pyquda/utils/
source.py:7-48, ``point``, ``wall``, ``momentum``), on the handle API of
libqcu.so: the 12 right-hand sides are solved together in one batch, so every
kernel launch covers all of them and they share each link load.

Layouts (external, even sites first):
  fermion     complex128 [2][Lt][Lz][Ly][Lx/2][4][3]
  propagator  complex128 [2][Lt][Lz][Ly][Lx/2][4 sink spin][4 source spin][3 sink colour][3 source colour]

Normalisation: ``normalization="mass"`` returns 2 kappa x, the solution of the
mass-normalised Wilson-clover operator, as PyQuda's ``invert`` does
(python/PyQuda/pyquda/dslash/general.py:354-362: ``x.data *= 2 * kappa``);
``"kappa"`` returns x of (A - kappa D) x = b.
"""
from __future__ import annotations

import torch

from .lattice import FP32, Batch, Lattice, fermion_shape

Ns, Nc = 4, 3


def propagator_shape(L):
    Lx, Ly, Lz, Lt = L
    return (2, Lt, Lz, Ly, Lx // 2, Ns, Ns, Nc, Nc)


def point(L, t_srce, spin, color, device="cuda"):
    """1 at site t_srce = [x, y, z, t], spin, colour (source.py:7-22)"""
    Lx, Ly, Lz, Lt = L
    x, y, z, t = t_srce
    b = torch.zeros(fermion_shape(L), dtype=torch.complex128, device=device)
    b[(x + y + z + t) % 2, t, z, y, x // 2, spin, color] = 1
    return b


def wall(L, t_srce, spin, color, device="cuda"):
    """1 on every site of time slice t_srce (source.py:25-35)"""
    b = torch.zeros(fermion_shape(L), dtype=torch.complex128, device=device)
    b[:, t_srce, :, :, :, spin, color] = 1
    return b


def momentum(L, t_srce, phase, spin, color, device="cuda"):
    """phase[2][Lt][Lz][Ly][Lx/2] on time slice t_srce (source.py:38-48)"""
    b = torch.zeros(fermion_shape(L), dtype=torch.complex128, device=device)
    ph = torch.as_tensor(phase, dtype=torch.complex128, device=device)
    b[:, t_srce, :, :, :, spin, color] = ph[:, t_srce]
    return b


def source(L, source_type, t_srce, spin, color, phase=None, device="cuda"):
    if source_type == "point":
        return point(L, t_srce, spin, color, device)
    if source_type == "wall":
        return wall(L, t_srce, spin, color, device)
    if source_type == "momentum":
        return momentum(L, t_srce, phase, spin, color, device)
    raise ValueError(f"source type {source_type!r} is not supported (point, wall, momentum)")


def _solve12(lattice: Lattice, bs, tol, prec, reliable_delta, max_iter, normalization):
    if normalization not in ("mass", "kappa"):
        raise ValueError("normalization must be 'mass' or 'kappa'")
    batch = Batch([lattice], [Ns * Nc])
    try:
        xs, infos = batch.solve(bs, tol=tol, prec=prec, reliable_delta=reliable_delta, max_iter=max_iter)
    finally:
        batch.close()
    bad = [k for k, r in enumerate(infos) if not r.converged]
    if bad:
        raise RuntimeError(f"propagator: solves {bad} did not reach tol {tol}")
    L = lattice.L
    prop = torch.empty(propagator_shape(L), dtype=torch.complex128, device=bs[0].device)
    scale = 2 * lattice.kappa if normalization == "mass" else 1.0
    for spin in range(Ns):
        for color in range(Nc):
            prop[..., :, spin, :, color] = xs[spin * Nc + color] * scale
    return prop, infos


def invert(lattice: Lattice, source_type="point", t_srce=(0, 0, 0, 0), phase=None, tol=1e-10, prec=FP32,
           reliable_delta=0.01, max_iter=20000, normalization="mass"):
    """Propagator from the 12 spin-colour sources of one type (core.py:44-57).
    Returns (propagator, per-solve SolveResult list)."""
    bs = [source(lattice.L, source_type, t_srce, s, c, phase) for s in range(Ns) for c in range(Nc)]
    return _solve12(lattice, bs, tol, prec, reliable_delta, max_iter, normalization)


def invert12(b12: torch.Tensor, lattice: Lattice, tol=1e-10, prec=FP32, reliable_delta=0.01, max_iter=20000,
             normalization="mass"):
    """Propagator from a 12-column source in the propagator layout (core.py:60-75)."""
    if tuple(b12.shape) != propagator_shape(lattice.L):
        raise ValueError(f"b12 must have shape {propagator_shape(lattice.L)}")
    bs = [b12[..., :, s, :, c].contiguous() for s in range(Ns) for c in range(Nc)]
    return _solve12(lattice, bs, tol, prec, reliable_delta, max_iter, normalization)
