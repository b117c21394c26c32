"""ILDG / LIME gauge files into the external gauge layout of the C ABI.

    from qcu_amd.gauge_io import read_ildg
    U, L = read_ildg("weak_field.lime")   # complex128 [4][2][Lt][Lz][Ly][Lx/2][3][3], [Lx, Ly, Lz, Lt]

Same entry points and meaning as PyQuda's gauge reader
(python/PyQuda/pyquda/utils/gauge_utils.py:13-77): ``read_ildg`` parses the
LIME records of an ILDG file (``readIldg``), ``read_ildg_bin`` reads raw
ILDG-ordered binary data (``readIldgBin``).  Both return host numpy arrays;
``torch.from_numpy(U).cuda()`` feeds them to ``Lattice.load_gauge``.

File layout (ILDG): big-endian complex, site-major with x fastest,
``[t][z][y][x][mu][row][col]``.  The external layout splits each direction's
field into even and odd sites, x stored at x/2 of its parity
(python/PyQuda/pyquda/field.py:22-63).

The SciDAC checksum record, when present, is verified: for every site of
lexicographic rank r, the CRC-32 of its bytes c gives
``suma ^= rotl(c, r % 29)`` and ``sumb ^= rotl(c, r % 31)``.
No record payload is executed or unpickled; XML records are parsed with
ElementTree from the bytes of the file.
"""
from __future__ import annotations

import re
import struct
import zlib
from xml.etree import ElementTree as ET

import numpy as np

LIME_MAGIC = 0x456789AB


def lime_records(data: bytes):
    """(name, payload offset, payload length) of every LIME record."""
    out = []
    off = 0
    while off < len(data):
        if off + 144 > len(data):
            raise ValueError("truncated LIME header")
        magic, version = struct.unpack(">IH", data[off:off + 6])
        if magic != LIME_MAGIC:
            raise ValueError(f"bad LIME magic at byte {off}")
        length = struct.unpack(">Q", data[off + 8:off + 16])[0]
        name = data[off + 16:off + 144].split(b"\x00", 1)[0].decode("ascii")
        start = off + 144
        if start + length > len(data):
            raise ValueError(f"LIME record {name!r} runs past the end of the file")
        out.append((name, start, length))
        off = start + (length + 7) // 8 * 8
    return out


def _xml(payload: bytes):
    root = ET.fromstring(payload.rstrip(b"\x00").decode("utf-8"))
    ns = re.match(r"\{.*\}", root.tag)
    return root, (ns.group(0) if ns else "")


def scidac_checksum(site_bytes: np.ndarray, L) -> tuple[int, int]:
    """SciDAC (suma, sumb) of per-site records: site_bytes is uint8 [V][bytes
    per site] in lexicographic order (x fastest)."""
    suma = sumb = 0
    for r in range(site_bytes.shape[0]):
        c = zlib.crc32(site_bytes[r].tobytes()) & 0xFFFFFFFF
        a, b = r % 29, r % 31
        suma ^= ((c << a) | (c >> (32 - a))) & 0xFFFFFFFF if a else c
        sumb ^= ((c << b) | (c >> (32 - b))) & 0xFFFFFFFF if b else c
    return suma, sumb


def lexico_to_external(lex: np.ndarray, L) -> np.ndarray:
    """[t][z][y][x][mu][3][3] -> [mu][2][Lt][Lz][Ly][Lx/2][3][3] (even sites first)."""
    Lx, Ly, Lz, Lt = L
    if Lx % 2:
        raise ValueError("Lx must be even")
    t, z, y, x = np.meshgrid(np.arange(Lt), np.arange(Lz), np.arange(Ly), np.arange(Lx), indexing="ij")
    par = (x + y + z + t) & 1
    out = np.empty((4, 2, Lt, Lz, Ly, Lx // 2, 3, 3), np.complex128)
    for p in (0, 1):
        m = par == p
        out[:, p, t[m], z[m], y[m], x[m] // 2] = np.moveaxis(lex[m], 1, 0)
    return out


def external_to_lexico(U: np.ndarray, L) -> np.ndarray:
    Lx, Ly, Lz, Lt = L
    U = U.reshape(4, 2, Lt, Lz, Ly, Lx // 2, 3, 3)
    t, z, y, x = np.meshgrid(np.arange(Lt), np.arange(Lz), np.arange(Ly), np.arange(Lx), indexing="ij")
    par = (x + y + z + t) & 1
    lex = U[:, par, t, z, y, x // 2]  # [mu][t][z][y][x][3][3]
    return np.ascontiguousarray(np.moveaxis(lex, 0, 4))


def read_ildg(path: str, verify_checksum: bool = True):
    """(gauge in the external layout, [Lx, Ly, Lz, Lt]) from an ILDG file."""
    with open(path, "rb") as f:
        data = f.read()
    recs = {name: (start, length) for name, start, length in lime_records(data)}
    for need in ("ildg-format", "ildg-binary-data"):
        if need not in recs:
            raise ValueError(f"{path}: no {need} record")
    s, n = recs["ildg-format"]
    root, ns = _xml(data[s:s + n])
    prec = int(root.find(f"{ns}precision").text)
    L = [int(root.find(f"{ns}{k}").text) for k in ("lx", "ly", "lz", "lt")]
    if root.find(f"{ns}field").text.strip() != "su3gauge" or prec not in (32, 64):
        raise ValueError(f"{path}: unsupported ILDG field or precision")
    s, n = recs["ildg-binary-data"]
    V = L[0] * L[1] * L[2] * L[3]
    word = prec // 8
    if n != V * 4 * 9 * 2 * word:
        raise ValueError(f"{path}: binary data has {n} bytes, expected {V * 72 * word}")
    raw = np.frombuffer(data, np.uint8, count=n, offset=s)
    if verify_checksum and "scidac-checksum" in recs:
        cs, cn = recs["scidac-checksum"]
        croot, cns = _xml(data[cs:cs + cn])
        want = (int(croot.find(f"{cns}suma").text, 16), int(croot.find(f"{cns}sumb").text, 16))
        got = scidac_checksum(raw.reshape(V, -1), L)
        if got != want:
            raise ValueError(f"{path}: SciDAC checksum {got[0]:08x} {got[1]:08x} != {want[0]:08x} {want[1]:08x}")
    lex = raw.view(f">c{2 * word}").astype(np.complex128).reshape(L[3], L[2], L[1], L[0], 4, 3, 3)
    return lexico_to_external(lex, L), L


def read_ildg_bin(path: str, dtype: str, L):
    """Raw ILDG-ordered binary data of the given numpy dtype (e.g. '>c16')."""
    Lx, Ly, Lz, Lt = L
    lex = np.fromfile(path, dtype).astype(np.complex128).reshape(Lt, Lz, Ly, Lx, 4, 3, 3)
    return lexico_to_external(lex, L)


def write_ildg(path: str, U: np.ndarray, L, precision: int = 64):
    """Writes an ILDG file (format, binary data and SciDAC checksum records)."""
    if precision not in (32, 64):
        raise ValueError("precision must be 32 or 64")
    lex = external_to_lexico(U, L)
    V = L[0] * L[1] * L[2] * L[3]
    binary = lex.astype(f">c{precision // 4}").tobytes()
    suma, sumb = scidac_checksum(np.frombuffer(binary, np.uint8).reshape(V, -1), L)
    fmt = ('<?xml version="1.0" encoding="UTF-8"?><ildgFormat xmlns="http://www.lqcd.org/ildg">'
           f"<version>1.0</version><field>su3gauge</field><precision>{precision}</precision>"
           f"<lx>{L[0]}</lx><ly>{L[1]}</ly><lz>{L[2]}</lz><lt>{L[3]}</lt></ildgFormat>").encode()
    chk = ('<?xml version="1.0" encoding="UTF-8"?><scidacChecksum><version>1.0</version>'
           f"<suma>{suma:x}</suma><sumb>{sumb:x}</sumb></scidacChecksum>").encode()
    recs = [("ildg-format", fmt), ("ildg-binary-data", binary), ("scidac-checksum", chk)]
    with open(path, "wb") as f:
        for k, (name, payload) in enumerate(recs):
            flags = (0x8000 if k == 0 else 0) | (0x4000 if k == len(recs) - 1 else 0)
            f.write(struct.pack(">IHH", LIME_MAGIC, 1, flags))
            f.write(struct.pack(">Q", len(payload)))
            f.write(name.encode().ljust(128, b"\x00"))
            f.write(payload)
            f.write(b"\x00" * ((len(payload) + 7) // 8 * 8 - len(payload)))
