"""ILDG/LIME gauge reader (qcu_amd.gauge_io) against the reference's own fixture.

tests/golden/weak_field.lime is the 4^3x8 weak-field configuration shipped with
the reference's tests (python/PyQuda/tests/weak_field.lime, written by Chroma).
Its SciDAC checksum record pins the byte order and site order of the reader;
the layout conversion is checked against the independent checkerboard map of
tests/indep.py; write/read round trips pin the writer.
"""
import os

import numpy as np
import pytest

import indep
from qcu_amd import gauge_io

HERE = os.path.dirname(os.path.abspath(__file__))
LIME = os.path.join(HERE, "golden", "weak_field.lime")


def test_weak_field_reads_and_checksum_matches():
    U, L = gauge_io.read_ildg(LIME)  # raises if the SciDAC checksum differs
    assert L == [4, 4, 4, 8]
    assert U.shape == (4, 2, 8, 4, 4, 2, 3, 3)
    u = U.reshape(4, -1, 3, 3)
    assert np.abs(np.einsum("msba,msbc->msac", u.conj(), u) - np.eye(3)).max() < 1e-12
    assert np.abs(np.linalg.det(u) - 1).max() < 1e-12
    # the checksum values themselves, recomputed from the raw record
    data = open(LIME, "rb").read()
    recs = {n: (s, k) for n, s, k in gauge_io.lime_records(data)}
    s, n = recs["ildg-binary-data"]
    raw = np.frombuffer(data, np.uint8, count=n, offset=s).reshape(4 * 4 * 4 * 8, -1)
    assert gauge_io.scidac_checksum(raw, L) == (0xA2C41090, 0x11193C39)


def test_weak_field_plaquette_is_weak():
    U, L = gauge_io.read_ildg(LIME)
    Ul = indep.gauge_to_lexico(U.reshape(4, 2, -1, 3, 3), L)
    plaq = []
    for mu in range(4):
        for nu in range(mu + 1, 4):
            a = Ul[mu]
            b = np.roll(Ul[nu], -1, axis=indep.AX[mu])
            c = np.roll(Ul[mu], -1, axis=indep.AX[nu])
            d = Ul[nu]
            p = np.einsum("...ab,...bc,...dc,...ed->...ae", a, b, c.conj(), d.conj())
            plaq.append(np.trace(p, axis1=-2, axis2=-1).real.mean() / 3)
    assert 0.9 < np.mean(plaq) < 1.0


def test_layout_matches_independent_checkerboard(orc):
    L = (4, 2, 2, 6)
    U = orc.gauge_random(L, 0.5, 3)
    lex = gauge_io.external_to_lexico(U, L)  # [t][z][y][x][mu][3][3]
    ref = np.moveaxis(indep.gauge_to_lexico(U, L), 0, 4)
    assert np.array_equal(lex, ref)
    assert np.array_equal(gauge_io.lexico_to_external(lex, L).reshape(U.shape), U)


@pytest.mark.parametrize("precision,tol", [(64, 0.0), (32, 1e-7)])
def test_write_read_round_trip(orc, tmp_path, precision, tol):
    L = [4, 2, 4, 2]
    U = orc.gauge_random(L, 0.4, 4).reshape(4, 2, 2, 4, 2, 2, 3, 3)
    path = str(tmp_path / "u.lime")
    gauge_io.write_ildg(path, U, L, precision)
    V, L2 = gauge_io.read_ildg(path)
    assert L2 == L
    assert np.abs(V - U).max() <= tol


def test_corrupted_data_is_rejected(tmp_path):
    data = bytearray(open(LIME, "rb").read())
    recs = {n: (s, k) for n, s, k in gauge_io.lime_records(bytes(data))}
    s, _ = recs["ildg-binary-data"]
    data[s + 1000] ^= 0x01
    path = tmp_path / "bad.lime"
    path.write_bytes(bytes(data))
    with pytest.raises(ValueError, match="checksum"):
        gauge_io.read_ildg(str(path))
    with pytest.raises(ValueError, match="magic"):
        bad = bytearray(data)
        bad[0] ^= 0xFF
        gauge_io.lime_records(bytes(bad))
