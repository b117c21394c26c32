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
