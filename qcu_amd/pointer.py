"""Pointer objects with the names of PyQuda's pointer module.

Mirrors python/PyQuda/pyquda/src/pointer.pyx:8-58 (Pointer, Pointers,
Pointerss) and ndarrayDataPointer (:60-162): a Pointer carries a device
address in `.ptr`.  Torch tensors and objects exposing
__cuda_array_interface__ are accepted; host arrays are refused by the library
(the C entry points check that every address is device memory).
"""
from __future__ import annotations


class Pointer:
    def __init__(self, dtype: str, ptr: int = 0):
        self.dtype = dtype
        self.ptr = int(ptr)

    def __int__(self):
        return self.ptr

    def __repr__(self):
        return f"Pointer({self.dtype!r}, 0x{self.ptr:x})"


class Pointers(Pointer):
    def __init__(self, dtype: str, ptrs):
        super().__init__(dtype, 0)
        self.ptrs = [int(p) for p in ptrs]


class Pointerss(Pointer):
    def __init__(self, dtype: str, ptrss):
        super().__init__(dtype, 0)
        self.ptrss = [[int(p) for p in row] for row in ptrss]


_DTYPES = {"complex128": "<c16", "float64": "<f8", "int32": "<i4"}


def ndarrayDataPointer(ndarray, as_void: bool = False) -> Pointer:
    """Pointer to the data of a device array (torch tensor or CUDA-array-interface object)."""
    try:
        import torch
        if isinstance(ndarray, torch.Tensor):
            if not ndarray.is_contiguous():
                raise ValueError("ndarrayDataPointer needs a contiguous tensor")
            name = str(ndarray.dtype).replace("torch.", "")
            return Pointer("void" if as_void else _DTYPES.get(name, name), ndarray.data_ptr())
    except ImportError:
        pass
    cai = getattr(ndarray, "__cuda_array_interface__", None)
    if cai is not None:
        return Pointer("void" if as_void else cai["typestr"], cai["data"][0])
    raise ValueError(f"{type(ndarray).__name__} does not expose device memory")


def address(obj) -> int:
    """Device address of a Pointer, a torch tensor or a plain integer."""
    if isinstance(obj, Pointer):
        return obj.ptr
    if isinstance(obj, int):
        return obj
    return ndarrayDataPointer(obj).ptr
