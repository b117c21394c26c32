"""ctypes loader of libqcu.so (the HIP implementation) and its C ABI.

The library is built in-tree (qcu_amd/lib/libqcu.so) by qcu_amd.build.build().
There is no CPU fallback: if the library is missing, importing the operator
modules raises ImportError with the build command to run.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libqcu.so")

# the reference entry points (include/qcu.h:21-30 of qcu) and the new ones
REFERENCE_SYMBOLS = ("dslashQcu", "dslashCloverQcu", "mpiDslashQcu", "mpiBistabCgQcu",
                     "ncclDslashQcu", "ncclBistabCgQcu")


class QcuError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"qcu error {code}: {msg}")
        self.code = code


class QcuParamStruct(C.Structure):
    _fields_ = [("lattice_size", C.c_int * 4)]


class SolveParam(C.Structure):
    _fields_ = [("tol", C.c_double), ("max_iter", C.c_int), ("precision", C.c_int),
                ("reliable_delta", C.c_double), ("chunk", C.c_int), ("use_graph", C.c_int),
                ("precond", C.c_void_p), ("legacy_stop", C.c_int), ("method", C.c_int)]


class SolveInfo(C.Structure):
    _fields_ = [("iters", C.c_int), ("reliable_updates", C.c_int), ("converged", C.c_int),
                ("fallback", C.c_int), ("true_res", C.c_double), ("seconds", C.c_double),
                ("restarts", C.c_int), ("pad", C.c_int)]


_lib = None

_VP = C.c_void_p
_PP = C.POINTER(QcuParamStruct)
_I4 = C.c_int * 4

SIGNATURES = {
    # reference entry points
    "dslashQcu": (None, [_VP, _VP, _VP, _PP, C.c_int]),
    "dslashCloverQcu": (None, [_VP, _VP, _VP, _PP, C.c_int]),
    "mpiDslashQcu": (None, [_VP, _VP, _VP, _PP, C.c_int, _PP]),
    "mpiBistabCgQcu": (None, [_VP, _PP, _PP]),
    "ncclDslashQcu": (None, [_VP, _VP, _VP, _PP, C.c_int, _PP]),
    "ncclBistabCgQcu": (None, [_VP, _VP, _VP, _PP, _PP]),
    # errors
    "qcuGetLastError": (C.c_int, []),
    "qcuGetLastErrorString": (C.c_char_p, []),
    "qcuClearError": (None, []),
    # handle API
    "qcuLatticeCreate": (C.c_int, [C.POINTER(_VP), _I4, C.c_double, C.c_double, C.c_int]),
    "qcuLatticeDestroy": (C.c_int, [_VP]),
    "qcuLatticeLoadGauge": (C.c_int, [_VP, _VP]),
    "qcuGaugeRandom": (C.c_int, [_VP, _I4, C.c_double, C.c_uint64]),
    "qcuGaussian": (C.c_int, [_VP, C.c_long, C.c_uint64]),
    "qcuDslash": (C.c_int, [_VP, _VP, _VP, C.c_int, C.c_int, C.c_int]),
    "qcuDslashNoAlloc": (C.c_int, [_VP, _VP, _VP, C.c_int, C.c_int, C.c_int]),
    "qcuMatFull": (C.c_int, [_VP, _VP, _VP]),
    "qcuTrueResidual": (C.c_int, [_VP, _VP, _VP, C.POINTER(C.c_double)]),
    "qcuSolve": (C.c_int, [_VP, _VP, _VP, C.POINTER(SolveParam), C.POINTER(SolveInfo)]),
    "qcuBenchHop": (C.c_int, [_VP, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                              C.POINTER(C.c_double)]),
    "qcuTune": (C.c_int, [_VP, C.c_int, C.POINTER(C.c_double)]),
    "qcuBatchCreate": (C.c_int, [C.POINTER(_VP), C.POINTER(_VP), C.POINTER(C.c_int), C.c_int]),
    "qcuBatchDestroy": (C.c_int, [_VP]),
    "qcuBatchSystems": (C.c_int, [_VP, C.POINTER(C.c_int)]),
    "qcuBatchSolve": (C.c_int, [_VP, C.POINTER(_VP), C.POINTER(_VP), C.POINTER(SolveParam),
                                C.POINTER(SolveInfo)]),
    "qcuBatchBenchHop": (C.c_int, [_VP, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_double)]),
    "qcuBatchTune": (C.c_int, [_VP, C.c_int, C.POINTER(C.c_double)]),
    "qcuTuneGet": (C.c_int, [_VP, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
                             C.POINTER(C.c_int)]),
    "qcuTuneSet": (C.c_int, [_VP, C.c_int, C.c_int, C.c_int, C.c_int]),
    "qcuBatchTuneSet": (C.c_int, [_VP, C.c_int, C.c_int, C.c_int, C.c_int]),
    "qcuTuneSave": (C.c_int, [C.c_char_p]),
    "qcuTuneLoad": (C.c_int, [C.c_char_p]),
    # learned preconditioner
    "qcuNetCreate": (C.c_int, [C.POINTER(_VP), _VP, C.c_int, C.POINTER(C.c_int), C.c_uint64, C.c_double]),
    "qcuNetDestroy": (C.c_int, [_VP]),
    "qcuNetWeightCount": (C.c_long, [_VP]),
    "qcuNetGetWeights": (C.c_int, [_VP, C.POINTER(C.c_double), C.c_long]),
    "qcuNetSetWeights": (C.c_int, [_VP, C.POINTER(C.c_double), C.c_long]),
    "qcuNetGetGradient": (C.c_int, [_VP, C.POINTER(C.c_double), C.c_long]),
    "qcuNetForward": (C.c_int, [_VP, _VP, _VP]),
    "qcuNetBackward": (C.c_int, [_VP, _VP, _VP, _VP]),
    "qcuNetTrain": (C.c_int, [_VP, C.POINTER(_VP), C.c_int, C.c_int, C.c_int, C.c_double,
                              C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "qcuNetLoss": (C.c_int, [_VP, C.POINTER(_VP), C.c_int, C.POINTER(C.c_double)]),
    "qcuNetApplyOdd": (C.c_int, [_VP, _VP, _VP]),
    "qcuNetLossGrad": (C.c_int, [_VP, _VP, C.POINTER(C.c_double)]),
    "qcuNetBenchLayer": (C.c_int, [_VP, C.c_int, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "qcuNetBenchApply": (C.c_int, [_VP, C.c_int, C.POINTER(C.c_double)]),
    "qcuNetBenchPrecondLayer": (C.c_int, [_VP, C.c_int, C.c_int, C.POINTER(C.c_double)]),
    "qcuNetSave": (C.c_int, [_VP, C.c_char_p]),
    "qcuNetLoad": (C.c_int, [_VP, C.c_char_p]),
}


def header_symbols():
    """Function names declared in include/qcu.h."""
    import re
    path = os.path.join(os.path.dirname(HERE), "include", "qcu.h")
    txt = open(path).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(?:void|int|long|const char \*)\s*\*?\s*(\w+)\s*\(", txt)))


def load():
    """Load libqcu.so (raises ImportError if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: run `python -c 'import __graft_entry__ as g; "
                          f"g.build()'` (or `make -C qcu_amd/csrc`) first")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def check(rc):
    if rc:
        lib = load()
        msg = lib.qcuGetLastErrorString()
        raise QcuError(rc, msg.decode() if msg else "")
    return rc


def check_last():
    """Raise if the last void reference entry point recorded an error."""
    lib = load()
    rc = lib.qcuGetLastError()
    if rc:
        msg = lib.qcuGetLastErrorString()
        raise QcuError(rc, msg.decode() if msg else "")
