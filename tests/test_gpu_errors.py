"""Errors of the handle API on the GPU.

Each failing call returns its include/qcu.h code with a message, never aborts
the process, and leaves the handles usable: the last solve of each test
still reaches the oracle's FP64 true residual.
"""
import numpy as np
import pytest
import torch

from gpu_util import dev, host, log

pytestmark = pytest.mark.gpu

ERR_ARG, ERR_STATE = 1, 7


def _code(fn):
    from qcu_amd._lib import QcuError
    with pytest.raises(QcuError) as e:
        fn()
    log(f"  -> {e.value}")
    return e.value.code


def test_handle_api_errors_leave_handles_usable(orc):
    from qcu_amd.lattice import FP16, FP64, Lattice, LearnedPreconditioner
    L, kappa, csw = (4, 4, 4, 8), 0.124, 1.0
    lat = Lattice(L, kappa, csw, -1)
    bh = orc.gauss((2, orc.vol(L) // 2, 4, 3), 300)
    b = dev(bh)
    # before the gauge field is loaded
    assert _code(lambda: lat.solve(b)) == ERR_STATE
    assert _code(lambda: LearnedPreconditioner(lat, (1, 4, 1))) == ERR_STATE
    U = orc.gauge_random(L, 0.3, 301)
    lat.load_gauge(dev(U))
    # host memory where device pointers are required: the Python front end
    # refuses CPU tensors, and the C entry point refuses host pointers
    with pytest.raises(ValueError):
        lat.solve(torch.from_numpy(bh))
    import ctypes as C
    from qcu_amd._lib import SolveInfo, SolveParam
    xh = np.zeros_like(bh)
    prm = SolveParam(tol=1e-10, max_iter=100, precision=FP64, reliable_delta=0.1, chunk=8, use_graph=1)
    rc = lat._lib.qcuSolve(lat.h, xh.ctypes.data_as(C.c_void_p), bh.ctypes.data_as(C.c_void_p), C.byref(prm),
                           C.byref(SolveInfo()))
    log(f"  -> host pointers: {rc} {lat._lib.qcuGetLastErrorString().decode()}")
    assert rc == ERR_ARG and not xh.any()
    # network shapes: one input and one output channel, at most 8 per layer
    for chan in ((2, 4, 1), (1, 9, 1), (1, 4, 2)):
        assert _code(lambda: LearnedPreconditioner(lat, chan)) == ERR_ARG
    # the normal-equation CG runs in FP64 without a preconditioner
    net = LearnedPreconditioner(lat, (1, 4, 1), seed=302)
    assert _code(lambda: lat.solve(b, prec=FP16, method="cgnr")) == ERR_ARG
    assert _code(lambda: lat.solve(b, precond=net, method="cgnr")) == ERR_ARG
    # a network belongs to the lattice it was built on
    other = Lattice(L, kappa, csw, -1)
    other.load_gauge(dev(U))
    assert _code(lambda: other.solve(b, precond=net)) == ERR_ARG
    # every handle still works
    for la, kw in ((lat, {}), (lat, {"precond": net, "prec": FP16}), (other, {"prec": FP64})):
        x, res = la.solve(b, tol=1e-10, **kw)
        tr = orc.true_residual(host(x), bh, U, L, kappa, csw, -1)
        log(f"after the errors: {res.iters} iters, oracle true residual {tr:.2e}")
        assert res.converged and tr <= 1e-10
    net.close()
    other.close()
    lat.close()


def test_network_entry_points_refuse_host_pointers(orc):
    """qcuNetForward / Backward / Train / Loss check every pointer before any
    launch: numpy (host) buffers return ERR_ARG and leave the handle usable"""
    import ctypes as C
    from qcu_amd.lattice import Lattice, LearnedPreconditioner
    L = (4, 4, 4, 8)
    lat = Lattice(L, 0.124, 1.0, -1)
    U = orc.gauge_random(L, 0.3, 310)
    lat.load_gauge(dev(U))
    net = LearnedPreconditioner(lat, (1, 4, 1), seed=311)
    lib = lat._lib
    vp = C.c_void_p
    full = orc.gauss((2, orc.vol(L) // 2, 4, 3), 312)
    half = np.ascontiguousarray(full[1])
    hp = full.ctypes.data_as(vp)
    fd = dev(full)
    dp = vp(fd.data_ptr())
    out = torch.empty_like(fd)
    op = vp(out.data_ptr())
    srcs_host = (vp * 2)(half.ctypes.data, half.ctypes.data)
    loss = C.c_double()
    for name, call in (("forward out", lambda: lib.qcuNetForward(net.handle, hp, dp)),
                       ("forward in", lambda: lib.qcuNetForward(net.handle, op, hp)),
                       ("backward grad_out", lambda: lib.qcuNetBackward(net.handle, None, hp, dp)),
                       ("backward in", lambda: lib.qcuNetBackward(net.handle, None, dp, hp)),
                       ("backward grad_in", lambda: lib.qcuNetBackward(net.handle, hp, dp, dp)),
                       ("train", lambda: lib.qcuNetTrain(net.handle, srcs_host, 2, 1, 1, 1e-3, None, None)),
                       ("loss", lambda: lib.qcuNetLoss(net.handle, srcs_host, 2, C.byref(loss)))):
        rc = call()
        log(f"  {name}: {rc} {lib.qcuGetLastErrorString().decode()}")
        assert rc == ERR_ARG, name
    # a null source in the list is refused as well
    srcs_null = (vp * 2)(fd.data_ptr(), None)
    assert lib.qcuNetLoss(net.handle, srcs_null, 2, C.byref(loss)) == ERR_ARG
    y = net.forward(fd)
    assert torch.isfinite(y).all()
    net.close()
    lat.close()


def test_failed_gauge_reload_leaves_lattice_unloaded(orc):
    """a reload that fails (a NaN link makes a clover block singular) leaves no
    half-built state behind: solves and the network report ERR_STATE until a
    good field is loaded, after which the solve meets the oracle's residual"""
    from qcu_amd.lattice import FP16, Lattice, LearnedPreconditioner
    L, kappa, csw = (4, 4, 4, 8), 0.124, 1.0
    lat = Lattice(L, kappa, csw, -1)
    U = orc.gauge_random(L, 0.3, 320)
    lat.load_gauge(dev(U))
    net = LearnedPreconditioner(lat, (1, 4, 1), seed=321)
    b = orc.gauss((2, orc.vol(L) // 2, 4, 3), 322)
    bd = dev(b)
    x, res = lat.solve(bd, tol=1e-10)
    assert res.converged
    Ubad = U.copy()
    Ubad[2, 1, 5, 1, 1] = np.nan
    assert _code(lambda: lat.load_gauge(dev(Ubad))) == ERR_STATE
    assert _code(lambda: lat.solve(bd, tol=1e-10)) == ERR_STATE
    assert _code(lambda: lat.solve(bd, tol=1e-10, prec=FP16, precond=net)) == ERR_STATE
    assert _code(lambda: net.forward(dev(np.zeros((1,) + b.shape, complex)))) == ERR_STATE
    assert _code(lambda: lat.dslash(bd[0].contiguous(), parity=0)) == ERR_STATE
    lat.load_gauge(dev(U))
    for kw in ({}, {"prec": FP16, "precond": net, "reliable_delta": 0.01}):
        x, res = lat.solve(bd, tol=1e-10, **kw)
        tr = orc.true_residual(host(x), b, U, L, kappa, csw, -1)
        log(f"after a good reload {kw.get('prec', 0)}: oracle true residual {tr:.1e}")
        assert res.converged and tr <= 1e-10
    net.close()
    lat.close()
