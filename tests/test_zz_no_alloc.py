"""qcuDslashNoAlloc: the hop of qcuDslash with its staging in the library's
own device array instead of allocated buffers.

Called through the C ABI (ctypes) and compared bit for bit with qcuDslash,
which runs the same kernels on the same data, for every precision, both
parities, with and without the clover term; FP64 also against the oracle's
hop (1e-12).  Two lattices in a row, the second smaller, share the array:
the second call must not see the first one's contents.  A lattice whose
staging exceeds the 32 MiB array is refused with QCU_ERR_DIMS before any
launch.
"""
import ctypes as C

import numpy as np
import pytest
import torch

from gpu_util import dev, host, log, rel

pytestmark = pytest.mark.gpu

ERR_DIMS = 2


def _no_alloc(lat, inp, parity, clover, prec):
    out = torch.empty_like(inp)
    rc = lat._lib.qcuDslashNoAlloc(lat.h, C.c_void_p(out.data_ptr()), C.c_void_p(inp.data_ptr()),
                                   int(parity), int(clover), int(prec))
    return rc, out


def test_no_alloc_dslash_matches_dslash(orc):
    from qcu_amd.lattice import Lattice
    kappa, csw = 0.124, 1.0
    for L in ((8, 8, 8, 16), (4, 4, 4, 8)):  # the second call is smaller
        U = orc.gauge_random(L, 0.3, 700 + L[0])
        lat = Lattice(L, kappa, csw, -1)
        lat.load_gauge(dev(U))
        psi = orc.gauss((orc.vol(L) // 2, 4, 3), 710 + L[0])
        pd = dev(psi)
        for prec in (0, 1, 2):
            for clover in (0, 1):
                for parity in (0, 1):
                    ref = lat.dslash(pd, parity=parity, clover=bool(clover), prec=prec)
                    rc, got = _no_alloc(lat, pd, parity, clover, prec)
                    torch.cuda.synchronize()
                    assert rc == 0, lat._lib.qcuGetLastErrorString()
                    assert torch.equal(got, ref), (L, prec, clover, parity)
        o = orc.clover_dslash(psi, U, L, 1, kappa * csw, -1)
        rc, got = _no_alloc(lat, pd, 1, 1, 0)
        e = rel(host(got), o)
        log(f"L={L}: no-alloc hop equals qcuDslash in every precision; FP64 clover vs oracle {e:.1e}")
        assert rc == 0 and e < 1e-12
        lat.close()


def test_no_alloc_dslash_refuses_large_lattices(orc):
    """16^3x48 in FP64 stages 2 x 98304 x 192 bytes (37.7 MB) > 32 MiB"""
    from qcu_amd.lattice import Lattice
    L = (16, 16, 16, 48)
    lat = Lattice(L, 0.124, 0.0, -1)
    x = torch.zeros((orc.vol(L) // 2, 4, 3), dtype=torch.complex128, device="cuda")
    rc, _ = _no_alloc(lat, x, 0, 0, 0)
    msg = lat._lib.qcuGetLastErrorString().decode()
    log(f"16^3x48 FP64: code {rc}: {msg}")
    assert rc == ERR_DIMS and "scratch" in msg
    lat.close()
