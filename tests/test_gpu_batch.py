"""GPU parity of batched solves: many right-hand sides sharing one lattice's
link loads, and several lattices of different shapes in one launch per kernel.

Every system is checked by its FP64 true residual recomputed by the host
oracle, and compared with the single-system solve of the same system.
"""
import numpy as np
import pytest
import torch

from gpu_util import dev, host, log, rel

pytestmark = pytest.mark.gpu


def test_twelve_point_sources_multi_rhs(orc):
    """12 spin-colour point sources on one clover lattice, FP32 inner solver"""
    from qcu_amd.lattice import Batch, Lattice, point_source
    L, kappa, csw = (8, 8, 8, 16), 0.1245, 1.0
    U = orc.gauge_random(L, 0.25, 500)
    lat = Lattice(L, kappa, csw, -1)
    lat.load_gauge(dev(U))
    bs = [point_source(L, s, c) for s in range(4) for c in range(3)]
    batch = Batch([lat], [12])
    assert batch.nsys == 12
    xs, res = batch.solve(bs, tol=1e-10, prec=1, reliable_delta=0.1)
    for k, (x, b, r) in enumerate(zip(xs, bs, res)):
        tr = orc.true_residual(host(x), host(b), U, L, kappa, csw, -1)
        x1, r1 = lat.solve(b, tol=1e-10, prec=1)
        log(f"rhs {k}: {r.iters} iters (single {r1.iters}), oracle true residual {tr:.2e}, "
            f"|x - x_single|/|x| {rel(host(x), host(x1)):.1e}")
        assert r.converged and tr < 1e-10
        assert rel(host(x), host(x1)) < 1e-8


def test_uneven_lattices_one_launch(orc):
    """several Wilson lattices of different shapes solved in the same launches"""
    from qcu_amd.lattice import Batch, Lattice
    shapes = [(4, 4, 4, 4), (6, 6, 6, 6), (4, 4, 4, 8), (8, 8, 8, 8), (4, 6, 4, 10)]
    kappa = 0.124
    lats, Us, bs = [], [], []
    for k, L in enumerate(shapes):
        U = orc.gauge_random(L, 0.3, 600 + k)
        lat = Lattice(L, kappa, 0.0, -1)
        lat.load_gauge(dev(U))
        lats.append(lat)
        Us.append(U)
        bs.append(dev(orc.gauss((2, orc.vol(L) // 2, 4, 3), 700 + k)))
    batch = Batch(lats, [1] * len(lats))
    for prec in (0, 2):
        xs, res = batch.solve(bs, tol=1e-10, prec=prec)
        for L, U, x, b, r, lat in zip(shapes, Us, xs, bs, res, lats):
            tr = orc.true_residual(host(x), host(b), U, L, kappa, 0.0, -1)
            x1, r1 = lat.solve(b, tol=1e-10, prec=prec)
            log(f"prec {prec} L={L}: {r.iters} iters (single {r1.iters}), true residual {tr:.2e}")
            assert r.converged and tr < 1e-10
            assert rel(host(x), host(x1)) < 1e-8


def test_mixed_rhs_counts_and_graph_replay(orc):
    """two lattices with 3 and 2 right-hand sides; graph and eager runs agree bit for bit"""
    from qcu_amd.lattice import Batch, Lattice
    shapes = [(4, 4, 4, 8), (6, 4, 4, 4)]
    lats, Us = [], []
    for k, L in enumerate(shapes):
        U = orc.gauge_random(L, 0.3, 800 + k)
        lat = Lattice(L, 0.124, 1.0, -1)
        lat.load_gauge(dev(U))
        lats.append(lat)
        Us.append(U)
    batch = Batch(lats, [3, 2])
    bs = [dev(orc.gauss((2, orc.vol(shapes[0 if k < 3 else 1]) // 2, 4, 3), 900 + k)) for k in range(5)]
    x1, r1 = batch.solve(bs, tol=1e-10, prec=0, use_graph=True)
    x2, r2 = batch.solve(bs, tol=1e-10, prec=0, use_graph=False)
    x3, r3 = batch.solve(bs, tol=1e-10, prec=0, use_graph=True)
    for k in range(5):
        L = shapes[0 if k < 3 else 1]
        U = Us[0 if k < 3 else 1]
        tr = orc.true_residual(host(x1[k]), host(bs[k]), U, L, 0.124, 1.0, -1)
        assert r1[k].converged and tr < 1e-10
        assert torch.equal(x1[k], x2[k]) and torch.equal(x1[k], x3[k])


def test_twelve_sources_fp16_inner(orc):
    """12 right-hand sides interleaved on one lattice with FP16 vectors (per-site
    scales per right-hand side): every system reaches the oracle's true residual"""
    from qcu_amd.lattice import FP16, Batch, Lattice, point_source
    L, kappa, csw = (8, 8, 8, 16), 0.124, 1.0
    U = orc.gauge_random(L, 0.3, 180)
    lat = Lattice(L, kappa, csw, -1)
    lat.load_gauge(dev(U))
    bs = [point_source(L, s, c) for s in range(4) for c in range(3)]
    batch = Batch([lat], [12])
    xs, res = batch.solve(bs, tol=1e-10, prec=FP16, reliable_delta=0.01)
    worst = 0.0
    for b, x, r in zip(bs, xs, res):
        tr = orc.true_residual(host(x), host(b), U, L, kappa, csw, -1)
        worst = max(worst, tr)
        assert r.converged and tr <= 1e-10
    log(f"12 RHS FP16: iterations {[r.iters for r in res]}, worst oracle true residual {worst:.2e}")
    batch.close()
