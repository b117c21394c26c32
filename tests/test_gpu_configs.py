"""The benchmark configurations at their stated sizes, judged by the oracle.

Like the reference's own solver test, which checks every solve through
MatQuda(x) - b (test/test.nccl.bistabcg.qcu-np1.py:66-74), every solve here is
checked by the FP64 true residual |b - M x| / |b| recomputed by the host oracle
(OpenMP FP64), never by the library's own report:

* 16^3x32 clover FP64 BiCGSTAB from a point source (eps 0.3, csw 1, kappa 0.124);
* 24^3x48 clover (eps 0.25, kappa 0.1245), Gaussian source, FP64 and FP32 /
  FP16-with-scale inner solves: each solution against the oracle's FP64
  BiCGSTAB solution, and each precision's hop against the oracle's hop;
* the 12 spin-colour point sources of that lattice in one batch, and 16 Wilson
  lattices {4 x 8^4, 4 x 12^4, 4 x 16^3x32, 4 x 20^3x40} in one batch;
* 32^3x64 clover (eps 0.2, kappa 0.1248): the learned preconditioner trained
  on the GPU and a learned + mixed-precision solve.

Bounds (stated per test): true residual <= 1e-10 for every system; solution
within 1e-8 (relative) of the oracle's FP64 solution, i.e. 100 x the
tolerance, which covers the condition number of these systems (the measured
differences are 1e-13 to 1e-10).
"""
import numpy as np
import pytest

from gpu_util import dev, host, log, rel

pytestmark = pytest.mark.gpu

TOL = 1e-10
SOL_BOUND = 1e-8


def _lattice(orc, L, eps, kappa, csw, seed, tb=-1):
    from qcu_amd.lattice import Lattice
    U = orc.gauge_random(L, eps, seed)
    lat = Lattice(L, kappa, csw, tb)
    lat.load_gauge(dev(U))
    return U, lat


def test_c1_point_source_clover_fp64(orc):
    """16^3x32 clover, FP64 BiCGSTAB, point source at the origin (spin 0, colour 0)"""
    from qcu_amd.lattice import FP64, point_source
    L, kappa, csw = (16, 16, 16, 32), 0.124, 1.0
    U, lat = _lattice(orc, L, 0.3, kappa, csw, 1601)
    b = point_source(L)
    x, res = lat.solve(b, tol=TOL, prec=FP64)
    bh, xh = host(b), host(x)
    tr = orc.true_residual(xh, bh, U, L, kappa, csw, -1)
    xo, info = orc.bicgstab(bh, U, L, kappa, csw, -1, tol=1e-12)
    e = rel(xh, xo)
    log(f"c1 16^3x32: {res.iters} iters (oracle {info['iters']}), oracle true residual {tr:.2e}, "
        f"|x - x_oracle|/|x| {e:.1e}")
    assert res.converged and tr <= TOL
    assert e < SOL_BOUND


@pytest.fixture(scope="module")
def c2(orc):
    L, kappa, csw = (24, 24, 24, 48), 0.1245, 1.0
    U, lat = _lattice(orc, L, 0.25, kappa, csw, 2401)
    b = orc.gauss((2, orc.vol(L) // 2, 4, 3), 2402)
    xo, info = orc.bicgstab(b, U, L, kappa, csw, -1, tol=1e-12)
    log(f"c2 24^3x48 oracle FP64 solution: {info['iters']} iters, true residual {info['true_res']:.1e}")
    return L, kappa, csw, U, lat, b, xo


@pytest.mark.parametrize("prec,name", [(0, "fp64"), (1, "fp32"), (2, "fp16")])
def test_c2_precisions_against_oracle(orc, c2, prec, name):
    """each inner precision reaches the FP64 true residual; its solution error
    against the oracle's FP64 solution is bounded by 100 x the tolerance"""
    L, kappa, csw, U, lat, b, xo = c2
    x, res = lat.solve(dev(b), tol=TOL, prec=prec, reliable_delta=0.01)
    xh = host(x)
    tr = orc.true_residual(xh, b, U, L, kappa, csw, -1)
    e = rel(xh, xo)
    log(f"c2 {name}: {res.iters} iters, {res.reliable_updates} reliable updates, oracle true residual "
        f"{tr:.2e}, |x - x_oracle_fp64|/|x| {e:.1e}")
    assert res.converged and tr <= TOL
    assert e < SOL_BOUND


@pytest.mark.parametrize("prec,bound", [(0, 1e-12), (1, 2e-6), (2, 3e-3)])
def test_c2_hop_precision_error(orc, c2, prec, bound):
    """the clover hop of each storage precision against the oracle's FP64 hop at 24^3x48"""
    L, kappa, csw, U, lat, b, xo = c2
    psi = b[1]
    out = host(lat.dslash(dev(psi), parity=0, clover=True, prec=prec))
    ref = orc.clover_dslash(psi, U, L, 0, kappa * csw, -1)
    e = rel(out, ref)
    log(f"c2 hop precision {prec}: relative error vs oracle {e:.2e} (bound {bound:g})")
    assert e < bound


def test_c3_twelve_point_sources_one_batch(orc, c2):
    """the 12 spin-colour point sources of the 24^3x48 lattice in one batch, FP32 inner"""
    from qcu_amd.lattice import FP32, Batch, point_source
    L, kappa, csw, U, lat, _, _ = c2
    bs = [point_source(L, s, c) for s in range(4) for c in range(3)]
    batch = Batch([lat], [12])
    xs, res = batch.solve(bs, tol=TOL, prec=FP32, reliable_delta=0.01)
    worst = 0.0
    for k, (x, b, r) in enumerate(zip(xs, bs, res)):
        tr = orc.true_residual(host(x), host(b), U, L, kappa, csw, -1)
        worst = max(worst, tr)
        assert r.converged and tr <= TOL, (k, tr)
    log(f"c3 12 RHS on 24^3x48: iterations {[r.iters for r in res]}, worst oracle true residual {worst:.2e}")
    batch.close()


def test_c3_sixteen_uneven_wilson_lattices_one_batch(orc):
    """4 each of 8^4, 12^4, 16^3x32 and 20^3x40 Wilson lattices (eps 0.3, kappa
    0.124, Gaussian sources) in one batch, FP32 inner with FP64 reliable updates"""
    from qcu_amd.lattice import FP32, Batch
    shapes = [(8, 8, 8, 8)] * 4 + [(12, 12, 12, 12)] * 4 + [(16, 16, 16, 32)] * 4 + [(20, 20, 20, 40)] * 4
    kappa = 0.124
    lats, Us, bs = [], [], []
    for k, L in enumerate(shapes):
        U, lat = _lattice(orc, L, 0.3, kappa, 0.0, 3100 + k)
        lats.append(lat)
        Us.append(U)
        bs.append(orc.gauss((2, orc.vol(L) // 2, 4, 3), 3200 + k))
    batch = Batch(lats, [1] * len(lats))
    xs, res = batch.solve([dev(b) for b in bs], tol=TOL, prec=FP32, reliable_delta=0.01)
    worst = 0.0
    for L, U, x, b, r in zip(shapes, Us, xs, bs, res):
        tr = orc.true_residual(host(x), b, U, L, kappa, 0.0, -1)
        worst = max(worst, tr)
        assert r.converged and tr <= TOL, (L, tr)
    log(f"c3 16 uneven lattices: iterations {[r.iters for r in res]}, worst oracle true residual {worst:.2e}")
    batch.close()


def test_c4_learned_mixed_solve_full_size(orc):
    """32^3x64 clover: the network is trained on the GPU (a few sources, one
    epoch) and applied as the right preconditioner of the FP16 mixed solve; the
    oracle's true residual judges the result, with the plain solve beside it"""
    from qcu_amd.lattice import FP16, LearnedPreconditioner, gaussian
    L, kappa, csw = (32, 32, 32, 64), 0.1248, 1.0
    U, lat = _lattice(orc, L, 0.2, kappa, csw, 4001)
    net = LearnedPreconditioner(lat, (1, 8, 8, 8, 1), seed=3)
    srcs = [gaussian((lat.vh, 4, 3), 4100 + k) for k in range(8)]
    losses, secs = net.train(srcs, epochs=1, batch=4, lr=2e-3)
    del srcs
    b = orc.gauss((2, orc.vol(L) // 2, 4, 3), 4002)
    bd = dev(b)
    x, res = lat.solve(bd, tol=TOL, prec=FP16, reliable_delta=0.01, precond=net)
    tr = orc.true_residual(host(x), b, U, L, kappa, csw, -1)
    _, res0 = lat.solve(bd, tol=TOL, prec=FP16, reliable_delta=0.01)
    log(f"c4 learned+mixed 32^3x64: loss {losses[0]:.4f} after {secs:.1f} s of training; {res.iters} iters "
        f"(mixed alone {res0.iters}), fallback {res.fallback}, oracle true residual {tr:.2e}")
    assert res.converged and tr <= TOL
    net.close()
