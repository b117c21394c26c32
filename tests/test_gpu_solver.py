"""GPU parity of the even-odd BiCGSTAB solvers (FP64 and mixed precision).

Every solve is checked by the FP64 true residual |b - M x| / |b| recomputed by
the host oracle; solutions are compared with the oracle's own BiCGSTAB
solution.  The stopping target is 1e-10 (the true residual required of every
solve); the solution tolerance follows from the condition number.
"""
import numpy as np
import pytest
import torch

from gpu_util import dev, host, log, rel

pytestmark = pytest.mark.gpu


def _setup(orc, L, eps, kappa, csw, tb, seed):
    from qcu_amd.lattice import Lattice
    U = orc.gauge_random(L, eps, seed)
    lat = Lattice(L, kappa, csw, tb)
    lat.load_gauge(dev(U))
    return U, lat


@pytest.mark.parametrize("csw", [0.0, 1.0])
def test_solve_fp64_matches_oracle(orc, csw):
    L, kappa = (8, 8, 8, 8), 0.124
    U, lat = _setup(orc, L, 0.3, kappa, csw, -1, 400)
    b = orc.gauss((2, orc.vol(L) // 2, 4, 3), 401)
    x, res = lat.solve(dev(b), tol=1e-10, prec=0)
    xh = host(x)
    tr = orc.true_residual(xh, b, U, L, kappa, csw, -1)
    xo, info = orc.bicgstab(b, U, L, kappa, csw, -1, tol=1e-12)
    log(f"fp64 csw={csw}: {res.iters} iters (oracle {info['iters']}), residual gpu {res.true_res:.2e} "
        f"oracle-recomputed {tr:.2e}, |x - x_oracle|/|x| {rel(xh, xo):.2e}")
    assert res.converged and tr < 1e-10 and res.true_res < 1e-10
    assert abs(res.true_res - tr) < 1e-12
    assert rel(xh, xo) < 1e-8
    assert abs(res.iters - info["iters"]) <= max(3, 0.1 * info["iters"])


@pytest.mark.parametrize("prec", [1, 2])
@pytest.mark.parametrize("csw", [0.0, 1.0])
def test_solve_mixed_precision(orc, prec, csw):
    L, kappa = (8, 8, 8, 16), 0.124
    U, lat = _setup(orc, L, 0.25, kappa, csw, -1, 410)
    b = orc.gauss((2, orc.vol(L) // 2, 4, 3), 411)
    x, res = lat.solve(dev(b), tol=1e-10, prec=prec, reliable_delta=0.1)
    tr = orc.true_residual(host(x), b, U, L, kappa, csw, -1)
    x64, r64 = lat.solve(dev(b), tol=1e-10, prec=0)
    log(f"mixed prec={prec} csw={csw}: {res.iters} iters, {res.reliable_updates} reliable updates "
        f"(fp64: {r64.iters} iters), oracle true residual {tr:.2e}")
    assert res.converged and tr < 1e-10
    assert rel(host(x), host(x64)) < 1e-8


def test_point_source_clover(orc):
    """the clover configuration's source: a point source at the origin, spin 0 colour 0"""
    from qcu_amd.lattice import point_source
    L, kappa = (8, 8, 8, 16), 0.124
    U, lat = _setup(orc, L, 0.3, kappa, 1.0, -1, 420)
    b = point_source(L)
    x, res = lat.solve(b, tol=1e-10, prec=0)
    tr = orc.true_residual(host(x), host(b), U, L, kappa, 1.0, -1)
    log(f"point source: {res.iters} iters, oracle true residual {tr:.2e}")
    assert res.converged and tr < 1e-10


def test_solve_without_graph_is_identical(orc):
    L, kappa = (4, 4, 4, 8), 0.124
    U, lat = _setup(orc, L, 0.3, kappa, 1.0, -1, 430)
    b = dev(orc.gauss((2, orc.vol(L) // 2, 4, 3), 431))
    x1, r1 = lat.solve(b, tol=1e-10, prec=0, use_graph=True)
    x2, r2 = lat.solve(b, tol=1e-10, prec=0, use_graph=False)
    x3, r3 = lat.solve(b, tol=1e-10, prec=0, use_graph=True)
    assert r1.iters == r2.iters == r3.iters
    # deterministic reductions: bitwise identical replays
    assert torch.equal(x1, x2) and torch.equal(x1, x3)


@pytest.mark.parametrize("source", ["gaussian", "ones"])
def test_legacy_bicgstab_entry_point(orc, source):
    """ncclBistabCgQcu: kappa 0.125, periodic, stop on Re<r,r> < 1e-9 (include/define.h:67-69);
    "ones" is the reference's own test source, every complex entry 1 + 1i
    (test/test.nccl.bistabcg.qcu-np1.py:30-31)"""
    from qcu_amd import pyqcu
    L = (8, 8, 8, 8)
    U = orc.gauge_random(L, 0.3, 440)
    if source == "ones":
        b = np.full((2, orc.vol(L) // 2, 4, 3), 1 + 1j)
    else:
        b = orc.gauss((2, orc.vol(L) // 2, 4, 3), 441)
    param = pyqcu.QcuParam()
    param.lattice_size = list(L)
    grid = pyqcu.QcuParam()
    grid.lattice_size = [1, 1, 1, 1]
    x = torch.zeros(b.shape, dtype=torch.complex128, device="cuda")
    bd = dev(b)
    pyqcu.ncclBistabCgQcu(x, bd, dev(U), param, grid)
    xh = host(x)
    r = orc.mat_full(xh, U, L, 0.125, 0.0, 1) - b
    rr = np.vdot(r, r).real
    log(f"ncclBistabCgQcu ({source} source) |b - Mx|^2 = {rr:.3e}")
    # the Schur residual is the odd part of b - Mx (Ao = 1); the even part vanishes
    assert rr < 1e-9
    assert np.array_equal(host(bd), b)


def test_nan_source_reports_not_converged(orc):
    L = (4, 4, 4, 4)
    U, lat = _setup(orc, L, 0.3, 0.124, 0.0, -1, 450)
    b = orc.gauss((2, orc.vol(L) // 2, 4, 3), 451)
    b[0, 3, 1, 2] = np.nan
    x, res = lat.solve(dev(b), tol=1e-10, prec=0, max_iter=50)
    log(f"nan source: converged={res.converged} iters={res.iters}")
    assert not res.converged


@pytest.mark.parametrize("csw", [0.0, 1.0])
def test_cg_normal_equations(orc, csw):
    """CG on the normal equations of the Schur system (FP64): the oracle's true
    residual meets 1e-10 and the solution is BiCGSTAB's; CG needs Mhat^+, so a
    wrong adjoint would not converge"""
    from qcu_amd.lattice import FP32, FP64, Lattice
    from qcu_amd._lib import QcuError
    L, kappa, tb = (8, 8, 4, 8), 0.122, -1
    U = orc.gauge_random(L, 0.3, 180)
    lat = Lattice(L, kappa, csw, tb)
    lat.load_gauge(dev(U))
    b = dev(orc.gauss((2, orc.vol(L) // 2, 4, 3), 181))
    x, r = lat.solve(b, tol=1e-10, method="cgnr")
    tr = orc.true_residual(host(x), host(b), U, L, kappa, csw, tb)
    xb, rb = lat.solve(b, tol=1e-10, prec=FP64)
    log(f"csw={csw}: CGNR {r.iters} iters (BiCGSTAB {rb.iters}), oracle true residual {tr:.1e}, "
        f"|x - x_bicgstab|/|x| {rel(host(x), host(xb)):.1e}")
    assert r.converged and tr <= 1e-10
    assert rel(host(x), host(xb)) < 1e-8
    with pytest.raises(QcuError):
        lat.solve(b, tol=1e-10, prec=FP32, method="cgnr")


def test_cg_normal_equations_batch(orc):
    from qcu_amd.lattice import Batch, Lattice
    shapes = [(4, 4, 4, 8), (6, 4, 4, 4)]
    lats, Us = [], []
    for k, L in enumerate(shapes):
        U = orc.gauge_random(L, 0.3, 190 + k)
        lat = Lattice(L, 0.124, 1.0, -1)
        lat.load_gauge(dev(U))
        lats.append(lat)
        Us.append(U)
    batch = Batch(lats, [2, 1])
    bs = [dev(orc.gauss((2, orc.vol(shapes[0 if k < 2 else 1]) // 2, 4, 3), 195 + k)) for k in range(3)]
    xs, res = batch.solve(bs, tol=1e-10, method="cgnr")
    for k in range(3):
        L, U = shapes[0 if k < 2 else 1], Us[0 if k < 2 else 1]
        tr = orc.true_residual(host(xs[k]), host(bs[k]), U, L, 0.124, 1.0, -1)
        assert res[k].converged and tr <= 1e-10, (k, tr)
