"""Single-fermion solves with QUDA's normalisation, on libqcu.so.

Mirrors ``invert`` and ``invertPC`` of PyQuda's dslash module
(python/PyQuda/pyquda/dslash/general.py:354-390).  Both return 2 kappa x,
where x solves (A - kappa D) x = b for the full-lattice fermion b
(complex128 [2][Lt][Lz][Ly][Lx/2][4][3], even sites first):
* ``invert`` hands the full system to the solver;
* ``invertPC`` states the even-odd reduction explicitly, as the reference does
  with cloverQuda / dslashQuda / invertQuda on the odd-odd system.
This library always solves the full system through the odd Schur complement
(b_o' = b_o + kappa A_o^-1 D_oe A_e^-1 b_e, then x_e from x_o), so the two
calls run the same solve and return the same field.  The FP64 true residual
|b - M x| / |b| <= tol is checked for every solve.
"""
from __future__ import annotations

import torch

from .lattice import FP64, Lattice


def _solve(lattice: Lattice, b: torch.Tensor, tol, prec, reliable_delta, max_iter, method):
    x, info = lattice.solve(b, tol=tol, prec=prec, reliable_delta=reliable_delta, max_iter=max_iter,
                            method=method)
    if not info.converged:
        raise RuntimeError(f"solve did not reach tol {tol}: {info}")
    return x * (2 * lattice.kappa), info


def invert(b: torch.Tensor, lattice: Lattice, tol=1e-10, prec=FP64, reliable_delta=0.01, max_iter=20000,
           method="bicgstab"):
    """2 kappa x for (A - kappa D) x = b (general.py:354-362). Returns (x, SolveResult)."""
    return _solve(lattice, b, tol, prec, reliable_delta, max_iter, method)


def invertPC(b: torch.Tensor, lattice: Lattice, tol=1e-10, prec=FP64, reliable_delta=0.01, max_iter=20000,
             method="bicgstab"):
    """The even-odd preconditioned form of ``invert`` (general.py:365-390): the same solution."""
    return _solve(lattice, b, tol, prec, reliable_delta, max_iter, method)
