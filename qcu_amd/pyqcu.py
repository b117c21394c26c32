"""PyQcu: the Python names of qcu's Cython binding, over libqcu.so.

Same names, argument order and meaning as python/PyQuda/pyquda/src/pyqcu.pyx:4-33
(QcuParam with a lattice_size property; six functions taking Pointer objects).
Arguments may also be torch tensors on the GPU or raw device addresses.  Where
the reference library prints and exits on a failure, these functions raise
QcuError carrying the library's message.
"""
from __future__ import annotations

import ctypes as C

from . import _lib
from ._lib import QcuError, QcuParamStruct  # noqa: F401
from .pointer import Pointer, Pointers, Pointerss, address, ndarrayDataPointer  # noqa: F401


class QcuParam:
    def __init__(self):
        self.param = QcuParamStruct()

    @property
    def lattice_size(self):
        return list(self.param.lattice_size)

    @lattice_size.setter
    def lattice_size(self, value):
        v = [int(x) for x in value]
        if len(v) != 4:
            raise ValueError("lattice_size needs 4 entries (Lx, Ly, Lz, Lt)")
        self.param.lattice_size = (C.c_int * 4)(*v)


def _p(x):
    return C.c_void_p(address(x))


def _call(name, *args):
    lib = _lib.load()
    getattr(lib, name)(*args)
    _lib.check_last()


def dslashQcu(fermion_out, fermion_in, gauge, param: QcuParam, parity: int):
    _call("dslashQcu", _p(fermion_out), _p(fermion_in), _p(gauge), C.byref(param.param), int(parity))


def dslashCloverQcu(fermion_out, fermion_in, gauge, param: QcuParam, parity: int):
    _call("dslashCloverQcu", _p(fermion_out), _p(fermion_in), _p(gauge), C.byref(param.param),
          int(parity))


def mpiDslashQcu(fermion_out, fermion_in, gauge, param: QcuParam, parity: int, grid: QcuParam):
    _call("mpiDslashQcu", _p(fermion_out), _p(fermion_in), _p(gauge), C.byref(param.param),
          int(parity), C.byref(grid.param))


def mpiBistabCgQcu(gauge, param: QcuParam, grid: QcuParam):
    _call("mpiBistabCgQcu", _p(gauge), C.byref(param.param), C.byref(grid.param))


def ncclDslashQcu(fermion_out, fermion_in, gauge, param: QcuParam, parity: int, grid: QcuParam):
    _call("ncclDslashQcu", _p(fermion_out), _p(fermion_in), _p(gauge), C.byref(param.param),
          int(parity), C.byref(grid.param))


def ncclBistabCgQcu(fermion_out, fermion_in, gauge, param: QcuParam, grid: QcuParam):
    _call("ncclBistabCgQcu", _p(fermion_out), _p(fermion_in), _p(gauge), C.byref(param.param),
          C.byref(grid.param))
