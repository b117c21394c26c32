"""Propagators from one 12-right-hand-side batch: every column is the solution
of its spin-colour source (oracle FP64 true residual), equal to the single
solve, and carries PyQuda's 2 kappa mass normalisation; the single-fermion
invert / invertPC of qcu_amd.general carry it too."""
import numpy as np
import pytest
import torch

from gpu_util import dev, host, log, rel

pytestmark = pytest.mark.gpu


def test_point_propagator_columns(orc):
    from qcu_amd import propagator as prop
    from qcu_amd.lattice import FP32, Lattice
    L, kappa, csw = (4, 4, 4, 8), 0.124, 1.0
    U = orc.gauge_random(L, 0.3, 170)
    lat = Lattice(L, kappa, csw, -1)
    lat.load_gauge(dev(U))
    t_srce = [1, 2, 3, 5]
    P, infos = prop.invert(lat, "point", t_srce, prec=FP32)
    assert tuple(P.shape) == prop.propagator_shape(L)
    Ph = host(P)
    worst = 0.0
    for s in range(4):
        for c in range(3):
            b = prop.point(L, t_srce, s, c)
            x = Ph[..., :, s, :, c] / (2 * kappa)
            tr = orc.true_residual(np.ascontiguousarray(x), host(b), U, L, kappa, csw, -1)
            x1, _ = lat.solve(b, tol=1e-10, prec=FP32)
            worst = max(worst, rel(x, host(x1)))
            assert tr <= 1e-10, (s, c, tr)
    log(f"point propagator: iterations {[r.iters for r in infos]}, max |x - x_single|/|x| {worst:.1e}")
    assert worst < 1e-10


def test_invert12_matches_invert(orc):
    from qcu_amd import propagator as prop
    from qcu_amd.lattice import FP64, Lattice
    L, kappa = (4, 4, 4, 4), 0.12
    U = orc.gauge_random(L, 0.3, 171)
    lat = Lattice(L, kappa, 0.0, -1)
    lat.load_gauge(dev(U))
    P, _ = prop.invert(lat, "wall", 2, prec=FP64, normalization="kappa")
    b12 = torch.zeros(prop.propagator_shape(L), dtype=torch.complex128, device="cuda")
    for s in range(4):
        for c in range(3):
            b12[..., :, s, :, c] = prop.wall(L, 2, s, c)
    P12, _ = prop.invert12(b12, lat, prec=FP64, normalization="kappa")
    assert torch.equal(P, P12)


def test_general_invert_and_invert_pc(orc):
    """general.invert / invertPC: 2 kappa x with x the oracle-checked solution;
    the explicit even-odd form gives the same field"""
    from qcu_amd import general
    from qcu_amd.lattice import Lattice
    L, kappa, csw = (4, 4, 4, 8), 0.124, 1.0
    U = orc.gauge_random(L, 0.3, 172)
    lat = Lattice(L, kappa, csw, -1)
    lat.load_gauge(dev(U))
    b = orc.gauss((2, orc.vol(L) // 2, 4, 3), 173)
    y, info = general.invert(dev(b), lat)
    y_pc, _ = general.invertPC(dev(b), lat)
    y_cg, _ = general.invert(dev(b), lat, method="cgnr")
    tr = orc.true_residual(host(y) / (2 * kappa), b, U, L, kappa, csw, -1)
    tr_cg = orc.true_residual(host(y_cg) / (2 * kappa), b, U, L, kappa, csw, -1)
    log(f"invert: {info.iters} iters, oracle true residual of y / 2 kappa {tr:.2e} (cgnr {tr_cg:.2e})")
    assert tr <= 1e-10 and tr_cg <= 1e-10
    assert torch.equal(y, y_pc)
    assert rel(host(y_cg), host(y)) < 1e-8
