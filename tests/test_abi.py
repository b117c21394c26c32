"""CPU checks of the C ABI library: it loads and exports every declared symbol.

No compute call is made (there is no GPU in the build container); the
reference-named entry points must match include/qcu.h:21-30 of the reference.
Argument errors are checked before any device call, so their codes and
messages are tested here too.
"""
import ctypes as C
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def qlib():
    from qcu_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        from qcu_amd.build import build
        build()
    return _lib


def test_library_loads_and_exports_header(qlib):
    lib = qlib.load()
    syms = qlib.header_symbols()
    for s in qlib.REFERENCE_SYMBOLS:
        assert s in syms
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing
    # every exported function is declared in the header with a ctypes signature
    assert set(syms) == set(qlib.SIGNATURES)


def test_exported_symbols_are_c_linkage(qlib):
    out = subprocess.run(["nm", "-D", "--defined-only", qlib.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    names = {line.split()[-1] for line in out.splitlines() if " T " in line}
    for s in qlib.header_symbols():
        assert s in names, s


def test_header_matches_reference_signatures():
    """the six reference prototypes, token for token (include/qcu.h:21-30)"""
    import re
    txt = open(os.path.join(ROOT, "include", "qcu.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    flat = " ".join(txt.split())
    expect = [
        "void dslashQcu(void *fermion_out, void *fermion_in, void *gauge, QcuParam *param, int parity);",
        "void dslashCloverQcu(void *fermion_out, void *fermion_in, void *gauge, QcuParam *param, int parity);",
        "void mpiDslashQcu(void *fermion_out, void *fermion_in, void *gauge, QcuParam *param, int parity, QcuParam *grid);",
        "void mpiBistabCgQcu(void *gauge, QcuParam *param, QcuParam *grid);",
        "void ncclDslashQcu(void *fermion_out, void *fermion_in, void *gauge, QcuParam *param, int parity, QcuParam *grid);",
        "void ncclBistabCgQcu(void *fermion_out, void *fermion_in, void *gauge, QcuParam *param, QcuParam *grid);",
        "typedef struct QcuParam_s { int lattice_size[4]; } QcuParam;",
    ]
    for e in expect:
        assert e in flat, e


def test_pyqcu_names_and_param():
    from qcu_amd import pyqcu
    for name in ("QcuParam", "dslashQcu", "dslashCloverQcu", "mpiDslashQcu", "mpiBistabCgQcu",
                 "ncclDslashQcu", "ncclBistabCgQcu"):
        assert hasattr(pyqcu, name)
    p = pyqcu.QcuParam()
    p.lattice_size = [8, 4, 6, 10]
    assert p.lattice_size == [8, 4, 6, 10]
    assert C.sizeof(p.param) == 16
    with pytest.raises(ValueError):
        p.lattice_size = [1, 2, 3]


def test_pointer_shim():
    import torch
    from qcu_amd.pointer import Pointer, address, ndarrayDataPointer
    t = torch.zeros(4, dtype=torch.complex128)
    p = ndarrayDataPointer(t)
    assert p.ptr == t.data_ptr() and p.dtype == "<c16"
    assert address(Pointer("void", 1234)) == 1234
    with pytest.raises(ValueError):
        ndarrayDataPointer([1, 2, 3])


def test_argument_errors_are_reported_without_a_device(qlib):
    """entry points validate their arguments before any device call: codes of
    include/qcu.h, a message, and no exception or exit across the ABI"""
    from qcu_amd.lattice import Lattice
    ERR_ARG, ERR_DIMS = 1, 2
    for L, tb, code, words in (((3, 4, 4, 4), -1, ERR_DIMS, "Lx must be even"),
                               ((4, 0, 4, 4), -1, ERR_DIMS, "positive"),
                               ((4, 4, 4, 4), 0, ERR_ARG, "t_boundary"),
                               ((2048, 2048, 1024, 1024), -1, ERR_DIMS, "too large")):
        with pytest.raises(qlib.QcuError) as e:
            Lattice(L, 0.12, 0.0, tb)
        assert e.value.code == code and words in str(e.value)
    lib = qlib.load()
    null = C.c_void_p()
    assert lib.qcuSolve(null, null, null, None, None) == ERR_ARG
    assert lib.qcuLatticeLoadGauge(null, null) == ERR_ARG
    assert lib.qcuNetForward(null, null, null) == ERR_ARG
    assert lib.qcuBatchSolve(null, None, None, None, None) == ERR_ARG
    assert lib.qcuGetLastError() == ERR_ARG and lib.qcuGetLastErrorString()
    chan = (C.c_int * 3)(1, 8, 1)
    assert lib.qcuNetCreate(C.byref(C.c_void_p()), null, 2, chan, 1, 0.05) == ERR_ARG


def test_lattice_size_limit_of_32_bit_offsets(qlib):
    """kernels address fields with 32-bit per-lane byte offsets of 16 bytes per
    site, so a lattice needs Vh < 2^28: 64x64x256x512 (Vh = 2^28) is refused
    with ERR_DIMS before any device call; one just below passes the check"""
    ERR_DIMS = 2
    lib = qlib.load()
    h = C.c_void_p()
    assert lib.qcuLatticeCreate(C.byref(h), (C.c_int * 4)(64, 64, 256, 512), 0.12, 0.0, -1) == ERR_DIMS
    assert "2^28" in lib.qcuGetLastErrorString().decode()
    rc = lib.qcuLatticeCreate(C.byref(h), (C.c_int * 4)(64, 64, 256, 510), 0.12, 0.0, -1)
    assert rc != ERR_DIMS  # without a GPU: ERR_STATE (no device); with one: created
    if rc == 0:
        lib.qcuLatticeDestroy(h)


def test_new_entry_points_validate_without_a_device(qlib):
    lib = qlib.load()
    null = C.c_void_p()
    ERR_ARG = 1
    assert lib.qcuTuneSet(null, 0, 64, 1, 0) == ERR_ARG
    assert lib.qcuBatchTuneSet(null, 0, 64, 1, 0) == ERR_ARG
    assert lib.qcuNetSave(null, b"/nonexistent") == ERR_ARG
    assert lib.qcuNetLoad(null, b"/nonexistent") == ERR_ARG
