"""ctypes front end of the FP64 host oracle (TEST INFRASTRUCTURE).

Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py may
import this module.  The product (qcu_amd) never does.  All arrays are numpy
complex128 in the external PyQuda layouts described in qcu_oracle.h.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None


def build(force: bool = False) -> str:
    """Compile liboracle.so with the Makefile beside this file (make rebuilds
    it when the sources are newer; force rebuilds unconditionally)."""
    subprocess.run(["make", "-s"] + (["-B"] if force else []) + ["-C", _HERE, "liboracle.so"],
                   check=True, stdout=subprocess.DEVNULL)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


_P = np.ctypeslib.ndpointer(dtype=np.complex128, flags="C_CONTIGUOUS")
_I4 = C.c_int * 4


class SolveParam(C.Structure):
    _fields_ = [("kappa", C.c_double), ("csw", C.c_double), ("t_boundary", C.c_int),
                ("tol", C.c_double), ("max_iter", C.c_int), ("mode", C.c_int),
                ("iters", C.c_int), ("schur_res", C.c_double), ("true_res", C.c_double),
                ("setup_s", C.c_double), ("iter_s", C.c_double)]


def _declare(L):
    L.or_gauge_random.argtypes = [_P, _I4, C.c_double, C.c_uint64]
    L.or_gauss_fill.argtypes = [_P, C.c_long, C.c_uint64]
    L.or_uniform.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64]
    L.or_uniform.restype = C.c_double
    L.or_su3_exp.argtypes = [np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS"),
                             C.c_double, _P]
    L.or_hop.argtypes = [_P, _P, _P, _I4, C.c_int, C.c_int]
    L.or_hop_full.argtypes = [_P, _P, _P, _I4, C.c_int]
    L.or_clover_build.argtypes = [_P, _P, _I4, C.c_int, C.c_double]
    L.or_clover_invert.argtypes = [_P, _P, C.c_long]
    L.or_clover_invert.restype = C.c_int
    L.or_clover_apply.argtypes = [_P, _P, _P, C.c_long]
    L.or_clover_dslash.argtypes = [_P, _P, _P, _I4, C.c_int, C.c_double, C.c_int]
    L.or_mat_full.argtypes = [_P, _P, _P, C.c_void_p, _I4, C.c_double, C.c_int]
    L.or_bicgstab.argtypes = [_P, _P, _P, _I4, C.POINTER(SolveParam)]
    L.or_true_residual.argtypes = [_P, _P, _P, _I4, C.c_double, C.c_double, C.c_int]
    L.or_true_residual.restype = C.c_double
    L.or_net_weight_count.argtypes = [C.c_int, C.POINTER(C.c_int)]
    L.or_net_weight_count.restype = C.c_long
    L.or_net_forward.argtypes = [_P, _P, _P, C.c_int, C.POINTER(C.c_int), _P, _I4, C.c_int, _P]
    L.or_net_backward.argtypes = [C.c_void_p, _P, _P, _P, C.c_int, C.POINTER(C.c_int), _P, _I4,
                                  C.c_int, _P]
    L.or_dot_out.argtypes = [_P, _P, _P, C.c_long]
    L.or_norm2.argtypes = [_P, C.c_long]
    L.or_norm2.restype = C.c_double
    L.or_num_threads.restype = C.c_int
    L.or_set_num_threads.argtypes = [C.c_int]


def _l4(L):
    return _I4(*[int(v) for v in L])


def vol(L):
    return int(L[0] * L[1] * L[2] * L[3])


def gauge_random(L, eps, seed):
    U = np.empty((4, 2, vol(L) // 2, 3, 3), np.complex128)
    lib().or_gauge_random(U, _l4(L), float(eps), int(seed))
    return U


def gauss(shape, seed):
    v = np.empty(shape, np.complex128)
    lib().or_gauss_fill(v, v.size, int(seed))
    return v


def uniform(seed, stream, k):
    return lib().or_uniform(int(seed), int(stream), int(k))


def su3_exp(omega, eps):
    out = np.empty((3, 3), np.complex128)
    lib().or_su3_exp(np.ascontiguousarray(omega, np.float64), float(eps), out)
    return out


def hop(inp, U, L, parity, t_boundary=1):
    out = np.empty_like(inp)
    lib().or_hop(out, np.ascontiguousarray(inp), U, _l4(L), int(parity), int(t_boundary))
    return out


def hop_full(inp, U, L, t_boundary=1):
    out = np.empty_like(inp)
    lib().or_hop_full(out, np.ascontiguousarray(inp), U, _l4(L), int(t_boundary))
    return out


def clover_build(U, L, parity, coeff):
    A = np.empty((vol(L) // 2, 12, 12), np.complex128)
    lib().or_clover_build(A, U, _l4(L), int(parity), float(coeff))
    return A


def clover_invert(A):
    Ainv = np.empty_like(A)
    bad = lib().or_clover_invert(Ainv, np.ascontiguousarray(A), A.shape[0])
    if bad:
        raise ArithmeticError(f"{bad} zero pivots")
    return Ainv


def clover_apply(A, v):
    out = np.empty_like(v)
    lib().or_clover_apply(out, np.ascontiguousarray(A), np.ascontiguousarray(v), A.shape[0])
    return out


def clover_dslash(inp, U, L, parity, coeff=1.0, t_boundary=1):
    out = np.empty_like(inp)
    lib().or_clover_dslash(out, np.ascontiguousarray(inp), U, _l4(L), int(parity), float(coeff),
                           int(t_boundary))
    return out


def mat_full(x, U, L, kappa, csw=0.0, t_boundary=1):
    """M x = A x - kappa D x on the full lattice (A = 1 for csw = 0)."""
    out = np.empty_like(x)
    Af = None
    if csw != 0.0:
        Af = np.ascontiguousarray(np.stack([clover_build(U, L, p, kappa * csw) for p in (0, 1)]))
    lib().or_mat_full(out, np.ascontiguousarray(x), U,
                      None if Af is None else Af.ctypes.data_as(C.c_void_p), _l4(L),
                      float(kappa), int(t_boundary))
    return out


def bicgstab(b, U, L, kappa, csw=0.0, t_boundary=1, tol=1e-10, max_iter=10000, mode=1):
    x = np.zeros_like(b)
    prm = SolveParam(kappa=kappa, csw=csw, t_boundary=t_boundary, tol=tol, max_iter=max_iter,
                     mode=mode)
    lib().or_bicgstab(x, np.ascontiguousarray(b), U, _l4(L), C.byref(prm))
    return x, {"iters": prm.iters, "schur_res": prm.schur_res, "true_res": prm.true_res,
               "setup_s": prm.setup_s, "iter_s": prm.iter_s}


def true_residual(x, b, U, L, kappa, csw=0.0, t_boundary=1):
    return lib().or_true_residual(np.ascontiguousarray(x), np.ascontiguousarray(b), U, _l4(L),
                                  float(kappa), float(csw), int(t_boundary))


def dot(a, b):
    """<a, b> = sum conj(a) b"""
    a = np.ascontiguousarray(a, np.complex128).reshape(-1)
    b = np.ascontiguousarray(b, np.complex128).reshape(-1)
    out = np.zeros(1, np.complex128)
    lib().or_dot_out(out, a, b, a.size)
    return complex(out[0])


def norm2(a):
    a = np.ascontiguousarray(a, np.complex128).reshape(-1)
    return float(lib().or_norm2(a, a.size))


def net_weight_count(chan):
    ch = (C.c_int * len(chan))(*chan)
    return lib().or_net_weight_count(len(chan) - 1, ch)


def net_forward(inp, W, chan, U, L, t_boundary=1):
    """inp: [chan[0]][2][Vh][12] -> out [chan[-1]][2][Vh][12]; returns (out, acts)."""
    fs = vol(L) * 12
    nl = len(chan) - 1
    ch = (C.c_int * len(chan))(*chan)
    out = np.empty(chan[-1] * fs, np.complex128)
    acts = np.empty(sum(chan[:nl]) * fs, np.complex128)
    lib().or_net_forward(out, np.ascontiguousarray(inp).reshape(-1), np.ascontiguousarray(W), nl,
                         ch, U, _l4(L), int(t_boundary), acts)
    return out, acts


def net_backward(grad_out, W, chan, U, L, acts, t_boundary=1):
    fs = vol(L) * 12
    nl = len(chan) - 1
    ch = (C.c_int * len(chan))(*chan)
    gW = np.zeros_like(W)
    gin = np.empty(chan[0] * fs, np.complex128)
    lib().or_net_backward(gin.ctypes.data_as(C.c_void_p), gW,
                          np.ascontiguousarray(grad_out).reshape(-1), np.ascontiguousarray(W), nl,
                          ch, U, _l4(L), int(t_boundary), acts)
    return gin, gW


def set_threads(n):
    lib().or_set_num_threads(int(n))


def threads():
    return lib().or_num_threads()
