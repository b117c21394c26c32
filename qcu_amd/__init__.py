"""qcu_amd: MI355X-native even-odd Wilson / clover dslash and BiCGSTAB.

Drop-in for qcu's hot path: the C ABI of include/qcu.h (libqcu.so, hand-written
HIP for gfx950) and the PyQcu Python names (qcu_amd.pyqcu).
"""
from ._lib import LIB_PATH, QcuError, load  # noqa: F401

__all__ = ["LIB_PATH", "QcuError", "load"]
