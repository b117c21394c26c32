"""Every launch shape the autotuner can choose, checked against the oracle.

The reference launches its kernels with one fixed shape (32-thread blocks,
grid Vh/32: include/define.h:5, include/lattice_set.h:75,110).  Here the tuner
picks among 18 shapes per lattice and precision: block size 64/128/256, the
XCD-aware block order on or off, and three instruction-scheduling groupings
of the hop kernel.  The block size also sets the layout of the fused
reduction partials (<r~, v> and <t, s>, |t|^2) and of the finishing kernel.
Each shape is pinned here with qcuTuneSet / qcuBatchTuneSet and checked:

* the hop (plain and clover) of each precision against the oracle's FP64 hop;
* FP64, FP32-inner and FP16-inner BiCGSTAB solves (all three fused-reduction
  variants, the xpay epilogue, reliable updates and the true residual) judged
  by the oracle's FP64 true residual;
* a 12-RHS batch (per-system partial slots) judged the same way.
"""
import itertools

import pytest

from gpu_util import dev, host, log, rel

pytestmark = pytest.mark.gpu

SHAPES = list(itertools.product((64, 128, 256), (0, 1), (0, 1, 2)))
HOP_BOUND = {0: 1e-12, 1: 2e-6, 2: 3e-3}


@pytest.fixture(scope="module")
def lattices(orc):
    from qcu_amd.lattice import Lattice
    L = (8, 8, 8, 16)
    out = {}
    for csw in (0.0, 1.0):
        U = orc.gauge_random(L, 0.3, 5100)
        lat = Lattice(L, 0.124, csw, -1)
        lat.load_gauge(dev(U))
        out[csw] = (L, U, lat)
    return out


@pytest.mark.parametrize("block,remap,group", SHAPES)
def test_every_tuner_shape_matches_oracle(orc, lattices, block, remap, group):
    for csw, (L, U, lat) in lattices.items():
        for prec in (0, 1, 2):
            lat.set_launch_shape(prec, block, remap, group)
            assert lat.launch_shape(prec) == {"block": block, "remap": remap, "group": group}
        psi = orc.gauss((orc.vol(L) // 2, 4, 3), 5101)
        for prec in (0, 1, 2):
            for clover in ((False, True) if csw else (False,)):
                out = host(lat.dslash(dev(psi), parity=1, clover=clover, prec=prec))
                ref = (orc.clover_dslash(psi, U, L, 1, 0.124 * csw, -1) if clover
                       else orc.hop(psi, U, L, 1, -1))
                e = rel(out, ref)
                assert e < HOP_BOUND[prec], (csw, prec, clover, e)
        b = orc.gauss((2, orc.vol(L) // 2, 4, 3), 5102)
        bd = dev(b)
        worst = 0.0
        for prec in (0, 1, 2):
            x, res = lat.solve(bd, tol=1e-10, prec=prec, reliable_delta=0.1)
            tr = orc.true_residual(host(x), b, U, L, 0.124, csw, -1)
            worst = max(worst, tr)
            assert res.converged and tr <= 1e-10, (csw, prec, tr)
        log(f"shape block {block} remap {remap} group {group} csw {csw}: hops ok, "
            f"worst oracle true residual {worst:.1e}")


@pytest.mark.parametrize("block,remap,group", SHAPES)
def test_every_tuner_shape_batch(orc, lattices, block, remap, group):
    """12 right-hand sides interleaved on the clover lattice, FP32 inner solver"""
    from qcu_amd.lattice import FP32, Batch, point_source
    L, U, lat = lattices[1.0]
    bs = [point_source(L, s, c) for s in range(4) for c in range(3)]
    batch = Batch([lat], [12])
    for prec in (0, 1):
        batch.set_launch_shape(prec, block, remap, group)
    xs, res = batch.solve(bs, tol=1e-10, prec=FP32, reliable_delta=0.1)
    worst = 0.0
    for x, b, r in zip(xs, bs, res):
        tr = orc.true_residual(host(x), host(b), U, L, 0.124, 1.0, -1)
        worst = max(worst, tr)
        assert r.converged and tr <= 1e-10
    log(f"batch shape block {block} remap {remap} group {group}: worst oracle true residual {worst:.1e}")
    batch.close()


def test_tuner_result_is_a_valid_shape(orc, lattices):
    """qcuTune times the 18 shapes and keeps one of them; the solve after it
    meets the oracle's true residual"""
    L, U, lat = lattices[1.0]
    for prec in (0, 2):
        r = lat.tune(prec)
        assert (r["block"], r["remap"], r["group"]) in SHAPES and r["ms"] > 0
    b = orc.gauss((2, orc.vol(L) // 2, 4, 3), 5103)
    x, res = lat.solve(dev(b), tol=1e-10, prec=2, reliable_delta=0.01)
    tr = orc.true_residual(host(x), b, U, L, 0.124, 1.0, -1)
    log(f"tuned shapes {lat.launch_shape(0)} / {lat.launch_shape(2)}: oracle true residual {tr:.1e}")
    assert res.converged and tr <= 1e-10
