"""Python side of the CPU oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this
package, and only as the checker.  The product library never imports it.

It restates, from the reference's behaviour, the setup half of the CoSMo hot path
(orbital table, Oxbash reader, file formats) in Python, and drives the C restatement in
cosmo_oracle.c (liboracle.so) for the heavy parts.  Each function is pinned by a test
against fixtures the reference's own programs wrote (tests/golden) or by a mathematical
check; tests/test_oracle.py names the pin of each.
"""
from __future__ import annotations

import ctypes
import hashlib
import os
import re
import struct
import subprocess
import xml.etree.ElementTree as ET
from fractions import Fraction

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE, "liboracle"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.c_void_p
        L.orc_cg.restype = ctypes.c_double
        L.orc_cg.argtypes = [ctypes.c_int] * 6
        L.orc_jm_vspx.restype = ctypes.c_long
        L.orc_jm_vspx.argtypes = [ctypes.c_int] + [P] * 7 + [ctypes.c_long, P, P, ctypes.c_long, P, P]
        L.orc_basis.restype = ctypes.c_long
        L.orc_basis.argtypes = [ctypes.c_int, ctypes.c_int, P, P, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                ctypes.c_int, ctypes.c_long, P]
        L.orc_h_rows.restype = ctypes.c_long
        L.orc_h_rows.argtypes = [ctypes.c_int, ctypes.c_long, P, ctypes.c_int, ctypes.c_long, P, P, ctypes.c_long,
                                 P, P, ctypes.c_long, ctypes.c_long, ctypes.c_int, P, P, P, P]
        L.orc_csb_times.restype = None
        L.orc_csb_times.argtypes = [ctypes.c_long, ctypes.c_int, P, P, P, ctypes.c_int, P, P]
        L.orc_hh_times.restype = None
        L.orc_hh_times.argtypes = [ctypes.c_long, P, P, P, ctypes.c_int, P, P]
        L.orc_csb_build.restype = P
        L.orc_csb_build.argtypes = [ctypes.c_long, ctypes.c_int, P, P, P]
        L.orc_csb_free.restype = None
        L.orc_csb_free.argtypes = [P]
        L.orc_csb_apply.restype = None
        L.orc_csb_apply.argtypes = [P, ctypes.c_long, ctypes.c_int, P, P]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


# ----------------------------------------------------------------------------------------
# XML and orbital table

def _half(s: str) -> int:
    """'5/2' -> 5, '2' -> 4 (doubled)."""
    return int(Fraction(s.strip()) * 2)


def read_system_xml(path: str):
    """Orbitals (sorted by index) and the <system> attributes of a CoSMo XML file."""
    root = ET.parse(path).getroot()
    vs = root.find("valencespace")
    orbs = []
    for o in vs.findall("orbital"):
        orbs.append(dict(name=o.get("name"), n=int(o.get("n", 0)), l=int(o.get("l", 0)),
                         j2=_half(o.get("j")), index=int(o.get("index")), type=o.get("type", "pn")))
    orbs.sort(key=lambda d: d["index"])
    sysn = root.find("system")
    attrs = dict(sysn.attrib)
    ham = root.find("hamiltonian")
    files = []
    for f in ham.findall("file"):
        scale = {1: 1.0, 2: 1.0}
        for s in f.findall("scale"):
            scale[int(s.get("rank"))] *= float(s.text)
        files.append((f.find("name").text.strip(), scale))
    return orbs, attrs, files


def sps_table(orbs):
    """m-scheme orbital table: neutrons first, then protons; m = -j..j step 2.
    Restates SM::Make (include/xsm/xSingleParticleStates.cxx:26-104)."""
    neu, pro = [], []
    for o in orbs:
        base = dict(level=o["index"] - 1, j2=o["j2"], n=o["n"], l=o["l"], par=o["l"] & 1, t2=1)
        for m in range(-o["j2"], o["j2"] + 1, 2):
            if o["type"] in ("pn", "n"):
                neu.append(dict(base, m2=m, tz2=1))
            if o["type"] in ("pn", "p"):
                pro.append(dict(base, m2=m, tz2=-1))
    return neu + pro, len(neu), len(pro)


# ----------------------------------------------------------------------------------------
# Oxbash reader

def read_oxbash(path: str, av: int, spescale: float = 1.0, vscale: float = 1.0):
    """Restates ReadOxbashINT (include/sm/sphericalint.cxx:145-246) and
    CheckSphericalINT (:43-84).  Numbers go through a float32 parse, as std::stof does.
    Returns (spe list, sorted list of ((l0,l1,l2,l3,J,T), V))."""
    f32 = lambda s: float(np.float32(s))
    spe = []
    tb = {}
    lc = 0
    vsc = vscale
    with open(path) as fh:
        for line in fh:
            line = line.lstrip()
            if not line or line[0] in "!#$%":
                continue
            toks = [t for t in re.split(r"[\s,]+", line) if t != ""]
            if not toks:
                continue
            if lc == 0:
                nl = int(toks[0])
                words = len(toks) - 1 - (3 if nl < 0 else 0)
                spe = [f32(toks[i + 1]) * spescale for i in range(words)]
                if nl < 0:
                    acore, ascale, spow = f32(toks[words + 1]), f32(toks[words + 2]), f32(toks[words + 3])
                    vsc *= (ascale / (av + acore)) ** spow
            else:
                if len(toks) < 7:
                    lc += 1
                    continue
                k = (int(toks[2]) - 1, int(toks[3]) - 1, int(toks[0]) - 1, int(toks[1]) - 1, int(toks[4]),
                     int(toks[5]))
                tb[k] = tb.get(k, 0.0) + vsc * f32(toks[6])
            lc += 1
    items = sorted(tb.items())
    # zero later duplicates (same J, T and the same unordered pairs, either orientation)
    vals = [v for _, v in items]
    keys = [k for k, _ in items]
    for i, k1 in enumerate(keys):
        a = (min(k1[0], k1[1]), max(k1[0], k1[1]), min(k1[2], k1[3]), max(k1[2], k1[3]))
        for j in range(i + 1, len(keys)):
            k2 = keys[j]
            if k1[4] != k2[4] or k1[5] != k2[5]:
                continue
            b = (min(k2[0], k2[1]), max(k2[0], k2[1]), min(k2[2], k2[3]), max(k2[2], k2[3]))
            if a == b or (a[0] == b[2] and a[1] == b[3] and a[2] == b[0] and a[3] == b[1]):
                vals[j] = 0.0
    return spe, list(zip(keys, vals))


def cg(j1, m1, j2, m2, J, M):
    return lib().orc_cg(j1, m1, j2, m2, J, M)


class System:
    """Everything the hot path needs, derived from a system XML by the oracle."""

    def __init__(self, xml_path: str):
        d = os.path.dirname(os.path.abspath(xml_path))
        self.orbs, self.attrs, files = read_system_xml(xml_path)
        self.sps, self.nn, self.np_ = sps_table(self.orbs)
        self.nsps = len(self.sps)
        self.nv = int(self.attrs["Nv"])
        self.zv = int(self.attrs["Zv"])
        self.av = int(self.attrs.get("Av", self.nv + self.zv))
        self.twoM = _half(self.attrs.get("Jz", "0"))
        self.parity = 1 if self.attrs.get("P", "+").strip().startswith("-") else 0
        self.name = self.attrs["name"]
        nlev = len(self.orbs)
        ent = [0.0] * nlev
        jv = {}
        for fname, scale in files:
            spe, tb = read_oxbash(os.path.join(d, fname), self.av, scale[1], scale[2])
            for i, e in enumerate(spe):
                ent[i] += e
            for k, v in tb:
                jv[k] = jv.get(k, 0.0) + v
        self.spe = ent
        self.jv = sorted(jv.items())
        self.esp_keys, self.esp_vals = self._esp()
        self.vspx_keys, self.vspx_vals = self._vspx()

    def _arr(self, key):
        return np.array([s[key] for s in self.sps], dtype=np.int32)

    def _esp(self):
        """ESP(i,i) = e[level(i)] (include/sm/MakeSPO.cxx:195-202)."""
        k = np.array([[i, i] for i in range(self.nsps)], dtype=np.uint16)
        v = np.array([self.spe[s["level"]] if s["level"] < len(self.spe) else 0.0 for s in self.sps])
        return k, v

    def _vspx(self):
        L = lib()
        lv, j2, m2, tz2, t2, nq, par = (self._arr(x) for x in ("level", "j2", "m2", "tz2", "t2", "n", "par"))
        jk = np.array([k for k, _ in self.jv], dtype=np.int32).reshape(-1, 6)
        jvv = np.array([v for _, v in self.jv], dtype=np.float64)
        cap = 1 << 16
        while True:
            ok = np.zeros((cap, 4), dtype=np.uint16)
            ov = np.zeros(cap)
            r = L.orc_jm_vspx(self.nsps, _p(lv), _p(j2), _p(m2), _p(tz2), _p(t2), _p(nq), _p(par), len(jk),
                              _p(jk), _p(jvv), cap, _p(ok), _p(ov))
            if r >= 0:
                return ok[:r].copy(), ov[:r].copy()
            cap = -r

    def basis(self):
        L = lib()
        m2 = self._arr("m2")
        par = self._arr("par")
        A = self.nv + self.zv
        cap = 1 << 16
        while True:
            out = np.zeros((cap, max(A, 1)), dtype=np.uint16)
            r = L.orc_basis(self.nn, self.np_, _p(m2), _p(par), self.nv, self.zv, self.twoM, self.parity, cap,
                            _p(out))
            if r >= 0:
                return out[:r, :A].copy()
            cap = -r

    def h_upper(self, basis=None, drop_small=True, r0=0, r1=None):
        """Upper-triangle H rows [r0,r1) in CSR (rowptr, cols, vals)."""
        L = lib()
        if basis is None:
            basis = self.basis()
        basis = np.ascontiguousarray(basis, dtype=np.uint16)
        n, N = basis.shape
        r1 = n if r1 is None else r1
        cnt = np.zeros(r1 - r0, dtype=np.int64)
        args = (N, n, _p(basis), self.nsps, len(self.esp_keys), _p(self.esp_keys), _p(self.esp_vals),
                len(self.vspx_keys), _p(self.vspx_keys), _p(self.vspx_vals), r0, r1, int(drop_small))
        L.orc_h_rows(*args, _p(cnt), None, None, None)
        rowptr = np.zeros(r1 - r0 + 1, dtype=np.int64)
        np.cumsum(cnt, out=rowptr[1:])
        cols = np.zeros(rowptr[-1], dtype=np.int64)
        vals = np.zeros(rowptr[-1])
        L.orc_h_rows(*args, _p(cnt), _p(rowptr[:-1].copy()), _p(cols), _p(vals))
        return rowptr, cols, vals


def csb_times(n, rowptr, cols, vals, X, br=None):
    """Reference CSB product of an upper-triangle CSR matrix with X (n x m, columns)."""
    L = lib()
    br = int(np.sqrt(1.0 * n)) + 1 if br is None else br
    X = np.asfortranarray(X, dtype=np.float64)
    m = X.shape[1]
    Y = np.zeros((n, m), order="F")
    L.orc_csb_times(n, br, _p(rowptr), _p(cols), _p(vals), m, _p(X), _p(Y))
    return Y


def hh_times(n, rowptr, cols, vals, X):
    L = lib()
    X = np.asfortranarray(X, dtype=np.float64)
    m = X.shape[1]
    Y = np.zeros((n, m), order="F")
    L.orc_hh_times(n, _p(rowptr), _p(cols), _p(vals), m, _p(X), _p(Y))
    return Y


def csb_port_times(n, rowptr, cols, vals, X, br=None):
    """The CPU-baseline form of the CSB product (orc_csb_build / orc_csb_apply: the same
    block traversal over an n x m row-major copy of X); returns Y (n x m)."""
    L = lib()
    br = int(np.sqrt(1.0 * n)) + 1 if br is None else br
    rp = np.ascontiguousarray(rowptr, dtype=np.int64)
    c = np.ascontiguousarray(cols, dtype=np.int64)
    v = np.ascontiguousarray(vals, dtype=np.float64)
    Xr = np.ascontiguousarray(X, dtype=np.float64)
    m = Xr.shape[1]
    Yr = np.zeros((n, m))
    h = L.orc_csb_build(n, br, _p(rp), _p(c), _p(v))
    try:
        L.orc_csb_apply(h, n, m, _p(Xr), _p(Yr))
    finally:
        L.orc_csb_free(ctypes.c_void_p(h))
    return Yr


def dense_from_upper(n, rowptr, cols, vals):
    H = np.zeros((n, n))
    for r in range(n):
        c = cols[rowptr[r]:rowptr[r + 1]]
        H[r, c] = vals[rowptr[r]:rowptr[r + 1]]
    return H + np.triu(H, 1).T


# ----------------------------------------------------------------------------------------
# file formats (Appendix-A layouts of the reference)

def read_mbs(path):
    """[i32 N][u64 n][n*N u16] (include/sm/ManyBodyStates.cxx:199-246)."""
    with open(path, "rb") as f:
        N, = struct.unpack("<i", f.read(4))
        n, = struct.unpack("<Q", f.read(8))
        z = np.frombuffer(f.read(2 * N * n), dtype=np.uint16).reshape(n, N)
    return z


def read_hh(path):
    """[u64 n] then per row [u32 nz][nz u64 col][nz f64] (include/tav/sparse.cxx:401-445)."""
    data = open(path, "rb").read()
    n, = struct.unpack_from("<Q", data, 0)
    off = 8
    rowptr = np.zeros(n + 1, dtype=np.int64)
    cols, vals = [], []
    for r in range(n):
        nz, = struct.unpack_from("<I", data, off)
        off += 4
        cols.append(np.frombuffer(data, dtype=np.uint64, count=nz, offset=off).astype(np.int64))
        off += 8 * nz
        vals.append(np.frombuffer(data, dtype=np.float64, count=nz, offset=off))
        off += 8 * nz
        rowptr[r + 1] = rowptr[r] + nz
    return rowptr, np.concatenate(cols) if cols else np.zeros(0, np.int64), \
        np.concatenate(vals) if vals else np.zeros(0)


def read_csb(path):
    """CSB file of SparseMatrixCSB::Write (include/tav/csbmatrix.cxx:553-576, per block
    CSBBlock::Write :246-254): [bool sym][u32 n][u16 BR][u16 BC][u16 nBR][u16 nBC], then
    for every block (row-block major, all nBR x nBC blocks) [bool isDiagonal][u32 nnz]
    [nnz u16 col][nnz u16 row][nnz f64].  Returns (header dict, list of blocks
    (i, j, isDiagonal, rows, cols, vals)) and the SHA-256 of the file without the values."""
    data = open(path, "rb").read()
    sym, n, br, bc, nbr, nbc = struct.unpack_from("<?IHHHH", data, 0)
    off = 13
    blocks = []
    h = hashlib.sha256(data[:13])
    for i in range(nbr):
        for j in range(nbc):
            isd, nnz = struct.unpack_from("<?I", data, off)
            h.update(data[off:off + 5 + 4 * nnz])
            off += 5
            c = np.frombuffer(data, dtype=np.uint16, count=nnz, offset=off)
            off += 2 * nnz
            r = np.frombuffer(data, dtype=np.uint16, count=nnz, offset=off)
            off += 2 * nnz
            v = np.frombuffer(data, dtype=np.float64, count=nnz, offset=off)
            off += 8 * nnz
            blocks.append((i, j, bool(isd), r, c, v))
    if off != len(data):
        raise ValueError("CSB file %s: %d trailing bytes" % (path, len(data) - off))
    return dict(symmetric=bool(sym), n=n, br=br, bc=bc, nbr=nbr, nbc=nbc), blocks, h.hexdigest()


def csb_bytes(n, rowptr, cols, vals, br=None):
    """The CSB file of an upper-row matrix as SparseMatrixCSB::Write lays it out after
    PPMatrix2CSBBlk (include/sm/PPMatrix.cxx:571-656): BR = BC = uint16(sqrt(n)) + 1
    (include/xsm/DEXHVcsb.cxx:63); block (I, J) holds rows of block row I and columns of
    block column J in ascending (row, col) order; every block is written, isDiagonal on
    I == J (include/tav/csbmatrix.cxx:246-254,553-576)."""
    rowptr = np.asarray(rowptr, np.int64)
    cols = np.asarray(cols, np.int64)
    vals = np.asarray(vals, np.float64)
    if br is None:
        br = int(np.sqrt(float(n)) + 1)
    nb = (n + br - 1) // br
    out = [struct.pack("<?IHHHH", True, n, br, br, nb, nb)]
    rows = np.repeat(np.arange(n, dtype=np.int64), np.diff(rowptr))
    for i in range(nb):
        lo, hi = rowptr[i * br], rowptr[min(n, (i + 1) * br)]
        r, c, v = rows[lo:hi], cols[lo:hi], vals[lo:hi]
        jb = c // br
        for j in range(nb):
            sel = jb == j
            out.append(struct.pack("<?I", i == j, int(sel.sum())))
            out.append((c[sel] - j * br).astype("<u2").tobytes())
            out.append((r[sel] - i * br).astype("<u2").tobytes())
            out.append(v[sel].astype("<f8").tobytes())
    return b"".join(out)


def csb_to_upper(hdr, blocks):
    """Upper rows (rowptr, cols, vals) of a symmetric CSB: entries ordered by row, then
    column (blocks of one block row are stored in ascending block column)."""
    n, br, bc = hdr["n"], hdr["br"], hdr["bc"]
    R = [np.asarray(r, np.int64) + i * br for (i, j, _, r, c, v) in blocks]
    C = [np.asarray(c, np.int64) + j * bc for (i, j, _, r, c, v) in blocks]
    V = [np.asarray(v) for (_, _, _, _, _, v) in blocks]
    R = np.concatenate(R) if R else np.zeros(0, np.int64)
    C = np.concatenate(C) if C else np.zeros(0, np.int64)
    V = np.concatenate(V) if V else np.zeros(0)
    o = np.lexsort((C, R))
    R, C, V = R[o], C[o], V[o]
    rowptr = np.zeros(n + 1, np.int64)
    np.add.at(rowptr, R + 1, 1)
    return np.cumsum(rowptr), C, V


def read_ee(path):
    """[u32 count] then per pair [f64 E][f64 dE][f64 resid][u64 n][n f64]
    (include/kdav/cdavidson.cpp:327-368)."""
    data = open(path, "rb").read()
    cnt, = struct.unpack_from("<I", data, 0)
    off = 4
    E, dE, R, V = [], [], [], []
    for _ in range(cnt):
        e, de, r, n = struct.unpack_from("<dddQ", data, off)
        off += 32
        V.append(np.frombuffer(data, dtype=np.float64, count=n, offset=off))
        off += 8 * n
        E.append(e); dE.append(de); R.append(r)
    return np.array(E), np.array(dE), np.array(R), V
