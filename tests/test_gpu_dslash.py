"""GPU parity of the hop, clover and layout paths against the FP64 host oracle.

Inputs are generated on the host by the oracle (seeded), uploaded bit for bit,
and the HIP results are compared at the stated tolerances: 1e-12 relative for
FP64 (the target of the FP64 dslash), and the storage precision for FP32 /
FP16-with-scale.
"""
import numpy as np
import pytest
import torch

from gpu_util import dev, host, log, rel

pytestmark = pytest.mark.gpu

L0 = (8, 8, 8, 8)  # the single-hop configuration: 8^4, eps 0.3, kappa 0.125


@pytest.fixture(scope="module")
def cfg0(orc):
    U = orc.gauge_random(L0, 0.3, 100)
    psi = orc.gauss((2, orc.vol(L0) // 2, 4, 3), 101)
    return U, psi


@pytest.mark.parametrize("tb", [1, -1])
@pytest.mark.parametrize("parity", [0, 1])
def test_hop_fp64_matches_oracle(orc, cfg0, tb, parity):
    from qcu_amd.lattice import Lattice, FP64
    U, psi = cfg0
    log(f"hop fp64 L={L0} tb={tb} parity={parity}")
    lat = Lattice(L0, 0.125, 0.0, tb)
    lat.load_gauge(dev(U))
    inp = np.ascontiguousarray(psi[1 - parity])
    out = host(lat.dslash(dev(inp), parity, prec=FP64))
    ref = orc.hop(inp, U, L0, parity, tb)
    err = rel(out, ref)
    log(f"  rel err {err:.3e}")
    assert err < 1e-12


@pytest.mark.parametrize("prec,tol", [(1, 2e-6), (2, 3e-3)])
def test_hop_low_precision_error(orc, cfg0, prec, tol):
    from qcu_amd.lattice import Lattice
    U, psi = cfg0
    lat = Lattice(L0, 0.125, 0.0, -1)
    lat.load_gauge(dev(U))
    inp = np.ascontiguousarray(psi[0])
    out = host(lat.dslash(dev(inp), 1, prec=prec))
    err = rel(out, orc.hop(inp, U, L0, 1, -1))
    log(f"hop prec={prec} rel err vs FP64 oracle {err:.3e}")
    assert err < tol


@pytest.mark.parametrize("L", [(4, 4, 4, 4), (2, 4, 6, 8), (6, 2, 4, 4), (16, 8, 8, 16)])
def test_hop_shapes(orc, L):
    """uneven and small extents, including a lattice smaller than one block"""
    from qcu_amd.lattice import Lattice, FP64
    U = orc.gauge_random(L, 0.3, 7)
    psi = orc.gauss((orc.vol(L) // 2, 4, 3), 8)
    lat = Lattice(L, 0.124, 0.0, -1)
    lat.load_gauge(dev(U))
    for parity in (0, 1):
        err = rel(host(lat.dslash(dev(psi), parity, prec=FP64)), orc.hop(psi, U, L, parity, -1))
        log(f"hop L={L} parity={parity} rel err {err:.3e}")
        assert err < 1e-12


def test_legacy_dslash_entry_points(orc, cfg0):
    """dslashQcu / ncclDslashQcu / mpiDslashQcu: periodic, inputs left bit-identical"""
    from qcu_amd import pyqcu
    U, psi = cfg0
    param = pyqcu.QcuParam()
    param.lattice_size = list(L0)
    grid = pyqcu.QcuParam()
    grid.lattice_size = [1, 1, 1, 1]
    Ud = dev(U)
    for parity in (0, 1):
        inp = dev(psi[1 - parity])
        ref = orc.hop(np.ascontiguousarray(psi[1 - parity]), U, L0, parity, 1)
        for name in ("dslashQcu", "ncclDslashQcu", "mpiDslashQcu"):
            out = torch.zeros_like(inp)
            fn = getattr(pyqcu, name)
            if name == "dslashQcu":
                fn(out, inp, Ud, param, parity)
            else:
                fn(out, inp, Ud, param, parity, grid)
            err = rel(host(out), ref)
            log(f"{name} parity={parity} rel err {err:.3e}")
            assert err < 1e-12
        assert np.array_equal(host(inp), psi[1 - parity])
    assert np.array_equal(host(Ud), U)


def test_legacy_errors_are_reported(orc, cfg0):
    from qcu_amd import pyqcu
    from qcu_amd._lib import QcuError
    U, psi = cfg0
    param = pyqcu.QcuParam()
    param.lattice_size = list(L0)
    grid = pyqcu.QcuParam()
    grid.lattice_size = [2, 1, 1, 1]
    inp = dev(psi[0])
    out = torch.zeros_like(inp)
    with pytest.raises(QcuError, match="grid"):
        pyqcu.ncclDslashQcu(out, inp, dev(U), param, 1, grid)
    with pytest.raises(QcuError, match="device"):
        pyqcu.dslashQcu(out, psi[0].ctypes.data, dev(U), param, 1)
    bad = pyqcu.QcuParam()
    bad.lattice_size = [7, 8, 8, 8]
    with pytest.raises(QcuError):
        pyqcu.dslashQcu(out, inp, dev(U), bad, 1)


def test_legacy_clover_dslash(orc, cfg0):
    """dslashCloverQcu: A_p^-1 D with kappa*csw = 1, periodic (src/nccl_clover_dslash.cu:3-79)"""
    from qcu_amd import pyqcu
    U, psi = cfg0
    param = pyqcu.QcuParam()
    param.lattice_size = list(L0)
    for parity in (0, 1):
        inp = np.ascontiguousarray(psi[1 - parity])
        out = torch.zeros(inp.shape, dtype=torch.complex128, device="cuda")
        pyqcu.dslashCloverQcu(out, dev(inp), dev(U), param, parity)
        ref = orc.clover_dslash(inp, U, L0, parity, 1.0, 1)
        err = rel(host(out), ref)
        log(f"dslashCloverQcu parity={parity} rel err {err:.3e}")
        # kappa*csw = 1 at eps = 0.3 makes some clover blocks ill conditioned;
        # the bound scales with the worst condition number of this parity
        A = orc.clover_build(U, L0, parity, 1.0)
        cond = np.linalg.cond(A).max()
        log(f"  max cond(A) {cond:.1e}")
        assert err < 1e-15 * cond * 100


def test_legacy_gauge_reuse_follows_contents(orc):
    """the reference-named entry points keep converted links and the clover term
    while the caller's gauge is unchanged; new contents at the same pointer,
    or another pointer, are picked up on the next call"""
    from qcu_amd import pyqcu
    L = (4, 4, 4, 8)
    U1, U2 = orc.gauge_random(L, 0.2, 60), orc.gauge_random(L, 0.2, 61)
    psi = orc.gauss((orc.vol(L) // 2, 4, 3), 62)
    param = pyqcu.QcuParam()
    param.lattice_size = list(L)
    g, inp = dev(U1), dev(psi)

    def run(fn, gauge):
        out = torch.zeros_like(inp)
        fn(out, inp, gauge, param, 0)
        return host(out)

    # kappa*csw = 1: the clover bound scales with the worst block condition number
    cond = max(np.linalg.cond(orc.clover_build(U, L, 0, 1.0)).max() for U in (U1, U2))
    for fn, ref_of, tol in ((pyqcu.dslashCloverQcu, lambda U: orc.clover_dslash(psi, U, L, 0, 1.0, 1),
                             max(1e-12, 1e-15 * cond * 100)),
                            (pyqcu.dslashQcu, lambda U: orc.hop(psi, U, L, 0, 1), 1e-12)):
        g.copy_(dev(U1))
        a, b = run(fn, g), run(fn, g)
        assert np.array_equal(a, b) and rel(a, ref_of(U1)) < tol
        g.copy_(dev(U2))  # same pointer, new contents
        c = run(fn, g)
        assert rel(c, ref_of(U2)) < tol
        d = run(fn, dev(U1))  # another pointer
        assert rel(d, ref_of(U1)) < tol
        log(f"{fn.__name__}: reuse {rel(a, ref_of(U1)):.1e}, new contents {rel(c, ref_of(U2)):.1e}")


@pytest.mark.parametrize("prec,tol", [(0, 1e-12), (1, 2e-6), (2, 5e-3)])
def test_clover_hop(orc, prec, tol):
    from qcu_amd.lattice import Lattice
    L = (8, 8, 8, 16)
    kappa, csw = 0.124, 1.0
    U = orc.gauge_random(L, 0.3, 200)
    psi = orc.gauss((orc.vol(L) // 2, 4, 3), 201)
    lat = Lattice(L, kappa, csw, -1)
    lat.load_gauge(dev(U))
    for parity in (0, 1):
        out = host(lat.dslash(dev(psi), parity, clover=True, prec=prec))
        ref = orc.clover_dslash(psi, U, L, parity, kappa * csw, -1)
        err = rel(out, ref)
        log(f"clover hop prec={prec} parity={parity} rel err {err:.3e}")
        assert err < tol


@pytest.mark.parametrize("csw", [0.0, 1.0])
def test_mat_full(orc, csw):
    from qcu_amd.lattice import Lattice
    L = (4, 4, 4, 8)
    U = orc.gauge_random(L, 0.25, 300)
    x = orc.gauss((2, orc.vol(L) // 2, 4, 3), 301)
    lat = Lattice(L, 0.1245, csw, -1)
    lat.load_gauge(dev(U))
    err = rel(host(lat.mat(dev(x))), orc.mat_full(x, U, L, 0.1245, csw, -1))
    log(f"mat full csw={csw} rel err {err:.3e}")
    assert err < 1e-13


def test_generators_match_oracle(orc):
    from qcu_amd.lattice import gaussian, random_gauge
    L = (4, 4, 4, 8)
    U = host(random_gauge(L, 0.3, 555)).reshape(4, 2, -1, 3, 3)
    Uo = orc.gauge_random(L, 0.3, 555)
    err = np.abs(U - Uo).max()
    log(f"gauge generator max abs diff {err:.3e}")
    assert err < 1e-13
    g = host(gaussian((1000,), 9))
    assert np.abs(g - orc.gauss(1000, 9)).max() < 1e-13
