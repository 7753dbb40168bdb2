"""cosmo_amd -- MI355X-native configuration-interaction hot path (CoSMo drop-in).

Python binding of the C ABI in include/cosmo_hip.h (libcosmo_hip.so, built in-tree by
``python -c "import cosmo_amd.build as b; b.build()"``).  The library is the product: host
C++ for the model-space setup and hand-written HIP kernels for the GPU Hamiltonian
builder, the stored and on-the-fly H.X and the block Davidson.  Importing this module
never falls back to a CPU path: a missing library raises ImportError.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcosmo_hip.so")

OK = 0
EINVAL, ENOMEM, EHIP, EIO, EUNSUP, ECONV = -1, -2, -3, -4, -5, -6

OP_STORED, OP_VIRTUAL, OP_KEEP_SMALL, OP_NO_FP32, OP_FACTORED = 0x0, 0x1, 0x2, 0x4, 0x8
PREC_FP64, PREC_FP32, PREC_FP64_ORDERED = 0, 1, 2


class CosmoError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("cosmo_hip error %d: %s" % (code, msg))
        self.code = code


class NotConverged(CosmoError):
    """The eigensolver stopped with pairs above the residual target (COSMO_ECONV).
    ``results`` holds what the call would have returned."""

    def __init__(self, msg, results):
        super().__init__(ECONV, msg)
        self.results = results


class DavParams(ctypes.Structure):
    _fields_ = [("nume", ctypes.c_int), ("mblock", ctypes.c_int), ("lim", ctypes.c_int), ("maxiter", ctypes.c_int),
                ("crite", ctypes.c_double), ("critc", ctypes.c_double), ("critr", ctypes.c_double),
                ("ortho", ctypes.c_double), ("precision", ctypes.c_int), ("mixed_switch", ctypes.c_double),
                ("ninit", ctypes.c_int), ("verbose", ctypes.c_int), ("strict", ctypes.c_int)]


class DavStats(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int), ("matvecs", ctypes.c_int), ("matvecs_fp64", ctypes.c_int),
                ("restarts", ctypes.c_int), ("ierr", ctypes.c_int), ("ms_total", ctypes.c_double),
                ("ms_apply", ctypes.c_double), ("ms_ortho", ctypes.c_double), ("ms_rr", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ServerParams(ctypes.Structure):
    _fields_ = [("hbm_budget", ctypes.c_int64), ("max_active", ctypes.c_int), ("precision", ctypes.c_int)]


class Result(ctypes.Structure):
    _fields_ = [("id", ctypes.c_int64), ("n", ctypes.c_int64), ("nume", ctypes.c_int), ("status", ctypes.c_int),
                ("t_submit", ctypes.c_double), ("t_admit", ctypes.c_double), ("t_done", ctypes.c_double),
                ("evals", ctypes.c_double * 64), ("resid", ctypes.c_double * 64)]


# every symbol of include/cosmo_hip.h with (restype, argtypes)
_P = ctypes.c_void_p
_I = ctypes.c_int
_L = ctypes.c_int64
_D = ctypes.c_double
_S = ctypes.c_char_p
SIGNATURES = {
    "cosmo_last_error": (ctypes.c_char_p, []),
    "cosmo_set_device": (_I, [_I]),
    "cosmo_system_from_xml": (_I, [_S, _S, _P]),
    "cosmo_system_create": (_I, [_I] + [_P] * 6 + [_I, _I, _I, _I, _L, _P, _P, _L, _P, _P, _P]),
    "cosmo_system_destroy": (_I, [_P]),
    "cosmo_system_dim": (_L, [_P]),
    "cosmo_system_info": (_I, [_P, _P, _P, _P, _P, _P, _P]),
    "cosmo_system_basis": (_I, [_P, _P]),
    "cosmo_system_tensors": (_I, [_P, _P, _P, _P, _P]),
    "cosmo_system_orbitals": (_I, [_P, _P, _P, _P, _P]),
    "cosmo_system_write_mbs": (_I, [_P, _S]),
    "cosmo_system_labels": (_I, [_P, _P, _I, _P, _P]),
    "cosmo_twice_spin": (_I, [_D, _I]),
    "cosmo_system_check_mbs": (_I, [_P, _S]),
    "cosmo_op_create": (_I, [_P, _I, _P]),
    "cosmo_op_create_upper": (_I, [_L, _P, _P, _P, _I, _P]),
    "cosmo_op_read_hh": (_I, [_S, _I, _P]),
    "cosmo_op_read_csb": (_I, [_S, _I, _P]),
    "cosmo_op_create_csb": (_I, [_L, _I, _I, _I, _I, _P, _P, _P, _P, _I, _P]),
    "cosmo_op_destroy": (_I, [_P]),
    "cosmo_op_size": (_L, [_P]),
    "cosmo_op_nnz": (_I, [_P, _P, _P]),
    "cosmo_op_bytes": (_L, [_P]),
    "cosmo_op_footprint": (_I, [_P, _I, _P]),
    "cosmo_op_diagonal": (_I, [_P, _P]),
    "cosmo_op_apply": (_I, [_P, _P, _P, _I, _I]),
    "cosmo_op_apply_device": (_I, [_P, _P, _P, _I, _I, _I, _P]),
    "cosmo_op_apply_device_noalloc": (_I, [_P, _P, _P, _I, _I, _I, _P]),
    "cosmo_op_upper_rows": (_I, [_P, _L, _L, _P, _P, _P, _P]),
    "cosmo_op_write_hh": (_I, [_P, _S]),
    "cosmo_op_write_csb": (_I, [_P, _S, _I]),
    "cosmo_op_timings": (_I, [_P, _P, _P]),
    "cosmo_op_set_layout": (_I, [_P, _I]),
    "cosmo_op_dense_eigen": (_I, [_P, _P, _P]),
    "cosmo_dav_defaults": (_I, [_P, _I]),
    "cosmo_davidson": (_I, [_P, _P, _P, _P, _P, _P, _P, _P]),
    "cosmo_write_ee": (_I, [_S, _L, _I, _P, _P, _P, _P]),
    "cosmo_read_ee": (_I, [_S, _L, _I, _P, _P, _P]),
    "cosmo_batch_create": (_I, [_I, _P, _I, _P]),
    "cosmo_batch_solve": (_I, [_P, _P, _P, _P, _P, _P]),
    "cosmo_batch_set_single_launch": (_I, [_P, ctypes.c_int64]),
    "cosmo_batch_destroy": (_I, [_P]),
    "cosmo_server_create": (_I, [_P, _P]),
    "cosmo_server_submit": (_I, [_P, _P, _I, _D, _P]),
    "cosmo_server_poll": (_I, [_P, _P, _I]),
    "cosmo_server_drain": (_I, [_P, _P, _I, _P]),
    "cosmo_server_destroy": (_I, [_P]),
}

_lib = None


def lib():
    """The loaded product library (ImportError when it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("libcosmo_hip.so is not built (run cosmo_amd.build.build())")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def check(rc):
    if rc < 0:
        raise CosmoError(rc, lib().cosmo_last_error().decode())
    return rc


class System:
    """Model space + interaction + basis specification (host side)."""

    def __init__(self, handle):
        self.h = ctypes.c_void_p(handle) if not isinstance(handle, ctypes.c_void_p) else handle
        N, ns, nn, np_, ne, nvx = (ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int64(),
                                   ctypes.c_int64())
        check(lib().cosmo_system_info(self.h, ctypes.byref(N), ctypes.byref(ns), ctypes.byref(nn), ctypes.byref(np_),
                                      ctypes.byref(ne), ctypes.byref(nvx)))
        self.nparticles, self.nsps, self.nn, self.np, self.nesp, self.nvspx = (N.value, ns.value, nn.value, np_.value,
                                                                                ne.value, nvx.value)
        self.dim = lib().cosmo_system_dim(self.h)

    @classmethod
    def from_xml(cls, path, sysname=None):
        h = ctypes.c_void_p()
        check(lib().cosmo_system_from_xml(path.encode(), sysname.encode() if sysname else None, ctypes.byref(h)))
        return cls(h)

    @classmethod
    def from_tables(cls, level, j2, m2, tz2, nv, zv, twoM, parity, esp_keys, esp_vals, vspx_keys, vspx_vals,
                    nq=None, lq=None):
        a = [np.ascontiguousarray(x, dtype=np.int32) for x in (level, j2, m2, tz2)]
        nq = np.ascontiguousarray(nq if nq is not None else np.zeros(len(a[0])), dtype=np.int32)
        lq = np.ascontiguousarray(lq if lq is not None else np.zeros(len(a[0])), dtype=np.int32)
        ek = np.ascontiguousarray(esp_keys, dtype=np.uint16)
        ev = np.ascontiguousarray(esp_vals, dtype=np.float64)
        vk = np.ascontiguousarray(vspx_keys, dtype=np.uint16)
        vv = np.ascontiguousarray(vspx_vals, dtype=np.float64)
        h = ctypes.c_void_p()
        check(lib().cosmo_system_create(len(a[0]), *[_p(x) for x in a], _p(nq), _p(lq), nv, zv, twoM, parity, len(ev),
                                        _p(ek), _p(ev), len(vv), _p(vk), _p(vv), ctypes.byref(h)))
        return cls(h)

    def basis(self):
        out = np.zeros((self.dim, self.nparticles), dtype=np.uint16)
        check(lib().cosmo_system_basis(self.h, _p(out)))
        return out

    def tensors(self):
        ek = np.zeros((self.nesp, 2), np.uint16)
        ev = np.zeros(self.nesp)
        vk = np.zeros((self.nvspx, 4), np.uint16)
        vv = np.zeros(self.nvspx)
        check(lib().cosmo_system_tensors(self.h, _p(ek), _p(ev), _p(vk), _p(vv)))
        return ek, ev, vk, vv

    def labels(self, evecs):
        """(<J-J+>, <T-T+>) of the columns of evecs (n x count)"""
        Z = np.ascontiguousarray(np.asarray(evecs, dtype=np.float64).T)
        c = Z.shape[0]
        jj = np.zeros(c)
        tt = np.zeros(c)
        check(lib().cosmo_system_labels(self.h, _p(Z), c, _p(jj), _p(tt)))
        return jj, tt

    def write_mbs(self, path):
        check(lib().cosmo_system_write_mbs(self.h, path.encode()))

    def check_mbs(self, path):
        check(lib().cosmo_system_check_mbs(self.h, path.encode()))

    def close(self):
        if self.h:
            lib().cosmo_system_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Operator:
    """Device-resident Hamiltonian (MatB::MatrixBase semantics)."""

    def __init__(self, handle):
        self.h = handle
        self.n = lib().cosmo_op_size(self.h)

    @classmethod
    def build(cls, system: System, flags=OP_STORED):
        h = ctypes.c_void_p()
        check(lib().cosmo_op_create(system.h, flags, ctypes.byref(h)))
        return cls(h)

    @classmethod
    def from_upper(cls, n, rowptr, cols, vals, flags=OP_STORED):
        rp = np.ascontiguousarray(rowptr, dtype=np.int64)
        c = np.ascontiguousarray(cols, dtype=np.int64)
        v = np.ascontiguousarray(vals, dtype=np.float64)
        h = ctypes.c_void_p()
        check(lib().cosmo_op_create_upper(n, _p(rp), _p(c), _p(v), flags, ctypes.byref(h)))
        return cls(h)

    @classmethod
    def read_hh(cls, path, flags=OP_STORED):
        h = ctypes.c_void_p()
        check(lib().cosmo_op_read_hh(path.encode(), flags, ctypes.byref(h)))
        return cls(h)

    @classmethod
    def from_csb(cls, n, br, bc, nbr, nbc, blocks, flags=OP_STORED):
        """Stored operator from symmetric CSB blocks: blocks maps (i, j) -> (rows, cols,
        vals), in-block indices (uint16) and values."""
        nb = nbr * nbc
        cnt = np.zeros(nb, np.uint32)
        keep = []
        R, C, V = (ctypes.c_void_p * nb)(), (ctypes.c_void_p * nb)(), (ctypes.c_void_p * nb)()
        for (i, j), (r, c, v) in blocks.items():
            r = np.ascontiguousarray(r, np.uint16)
            c = np.ascontiguousarray(c, np.uint16)
            v = np.ascontiguousarray(v, np.float64)
            keep += [r, c, v]
            b = i * nbc + j
            cnt[b] = len(v)
            R[b], C[b], V[b] = r.ctypes.data, c.ctypes.data, v.ctypes.data
        h = ctypes.c_void_p()
        check(lib().cosmo_op_create_csb(n, br, bc, nbr, nbc, _p(cnt), ctypes.cast(R, ctypes.c_void_p),
                                        ctypes.cast(C, ctypes.c_void_p), ctypes.cast(V, ctypes.c_void_p), flags,
                                        ctypes.byref(h)))
        return cls(h)

    @classmethod
    def read_csb(cls, path, flags=OP_STORED):
        """Stored operator from a symmetric CSB file (SparseMatrixCSB::Write layout)."""
        h = ctypes.c_void_p()
        check(lib().cosmo_op_read_csb(path.encode(), flags, ctypes.byref(h)))
        return cls(h)

    def nnz(self):
        u, f = ctypes.c_int64(), ctypes.c_int64()
        check(lib().cosmo_op_nnz(self.h, ctypes.byref(u), ctypes.byref(f)))
        return u.value, f.value

    def bytes(self):
        return lib().cosmo_op_bytes(self.h)

    def diagonal(self):
        d = np.zeros(self.n)
        check(lib().cosmo_op_diagonal(self.h, _p(d)))
        return d

    def apply(self, X, precision=PREC_FP64):
        """Y = H X for X of shape (n, m) (columns are the vectors); host round trip."""
        X = np.asarray(X, dtype=np.float64)
        if X.ndim == 1:
            X = X[:, None]
        m = X.shape[1]
        cols = [np.ascontiguousarray(X[:, j]) for j in range(m)]
        outs = [np.zeros(self.n) for _ in range(m)]
        xp = (ctypes.c_void_p * m)(*[c.ctypes.data for c in cols])
        yp = (ctypes.c_void_p * m)(*[o.ctypes.data for o in outs])
        check(lib().cosmo_op_apply(self.h, yp, xp, m, precision))
        return np.stack(outs, axis=1)

    def apply_device(self, X, Y, precision=PREC_FP64, stream=None):
        """Device form: X, Y are torch CUDA tensors of shape (n, m) row-major (float64 or
        float32 per precision); stream is a torch.cuda.Stream or None."""
        m, ld = X.shape[1], X.stride(0)
        s = None if stream is None else ctypes.c_void_p(stream.cuda_stream)
        check(lib().cosmo_op_apply_device(self.h, ctypes.c_void_p(X.data_ptr()), ctypes.c_void_p(Y.data_ptr()), m, ld,
                                          precision, s))

    def apply_device_noalloc(self, X, Y, precision=PREC_FP64, stream=None):
        """apply_device allocating nothing (product-form operators; the scratch is a fixed
        64 MiB device array per GPU, shared by every call on it)."""
        m, ld = X.shape[1], X.stride(0)
        s = None if stream is None else ctypes.c_void_p(stream.cuda_stream)
        check(lib().cosmo_op_apply_device_noalloc(self.h, ctypes.c_void_p(X.data_ptr()), ctypes.c_void_p(Y.data_ptr()),
                                                  m, ld, precision, s))

    def upper_rows(self, r0=0, r1=None):
        r1 = self.n if r1 is None else r1
        cnt = np.zeros(r1 - r0, np.int64)
        check(lib().cosmo_op_upper_rows(self.h, r0, r1, _p(cnt), None, None, None))
        rp = np.zeros(r1 - r0 + 1, np.int64)
        np.cumsum(cnt, out=rp[1:])
        cols = np.zeros(rp[-1], np.int64)
        vals = np.zeros(rp[-1])
        check(lib().cosmo_op_upper_rows(self.h, r0, r1, _p(cnt), _p(rp[:-1].copy()), _p(cols), _p(vals)))
        return rp, cols, vals

    def write_hh(self, path):
        check(lib().cosmo_op_write_hh(self.h, path.encode()))

    def write_csb(self, path, block=0):
        """Upper triangle as a CSB file (block 0: uint16(sqrt(n)) + 1 rows per block)."""
        check(lib().cosmo_op_write_csb(self.h, path.encode(), block))

    def dense_eigen(self, vectors=True):
        w = np.zeros(self.n)
        V = np.zeros((self.n, self.n)) if vectors else None
        check(lib().cosmo_op_dense_eigen(self.h, _p(w), _p(V)))
        return w, (V.T if vectors else None)

    def set_layout(self, layout):
        """Stored product kernel: 0 automatic, 1 rows, 2 tiled."""
        check(lib().cosmo_op_set_layout(self.h, int(layout)))

    def timings(self):
        b, a = ctypes.c_double(), ctypes.c_double()
        check(lib().cosmo_op_timings(self.h, ctypes.byref(b), ctypes.byref(a)))
        return b.value, a.value

    def close(self):
        if self.h:
            lib().cosmo_op_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def dav_params(nume, **kw):
    p = DavParams()
    check(lib().cosmo_dav_defaults(ctypes.byref(p), nume))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def davidson(op: Operator, nume, init=None, want_vectors=True, check_converged=True, **kw):
    """Lowest nume eigenpairs: returns (evals, dE, resid, evecs (n x nume) or None, stats dict).

    A solve that stops with a pair above critr (maxiter, orthogonalisation failure, ...)
    raises NotConverged carrying the results, unless check_converged is False; the stats
    dict's ``converged`` says which case it was."""
    p = dav_params(nume, **kw)
    n = op.n
    ev = np.zeros(nume)
    de = np.zeros(nume)
    rs = np.zeros(nume)
    vec = np.zeros((nume, n)) if want_vectors else None
    ini = None
    if init is not None:
        ini = np.ascontiguousarray(np.asarray(init, dtype=np.float64).T)  # nume x n
        p.ninit = ini.shape[0]
    st = DavStats()
    rc = lib().cosmo_davidson(op.h, ctypes.byref(p), _p(ini), _p(ev), _p(de), _p(rs), _p(vec), ctypes.byref(st))
    if rc != ECONV:
        check(rc)
    out = (ev, de, rs, (vec.T if vec is not None else None), dict(st.as_dict(), converged=(rc == OK)))
    if rc == ECONV and check_converged:
        raise NotConverged(lib().cosmo_last_error().decode(), out)
    return out


def write_ee(path, evals, dE, resid, evecs):
    evecs = np.ascontiguousarray(np.asarray(evecs, dtype=np.float64).T)
    n = evecs.shape[1]
    check(lib().cosmo_write_ee(path.encode(), n, len(evals), _p(np.asarray(evals, float)), _p(np.asarray(dE, float)),
                               _p(np.asarray(resid, float)), _p(evecs)))


def footprint(system: System, flags=OP_STORED):
    """Exact device bytes a stored operator of `system` would take."""
    b = ctypes.c_int64()
    check(lib().cosmo_op_footprint(system.h, flags, ctypes.byref(b)))
    return b.value


class Batch:
    """Independent solves run together on the device (one stream per worker)."""

    def __init__(self, systems, flags=OP_STORED):
        self.systems = list(systems)
        arr = (ctypes.c_void_p * len(self.systems))(*[s.h.value for s in self.systems])
        self.h = ctypes.c_void_p()
        check(lib().cosmo_batch_create(len(self.systems), arr, flags, ctypes.byref(self.h)))

    def set_single_launch(self, nmax):
        """Dimension bound of the one-launch path for small solves (0: off)."""
        check(lib().cosmo_batch_set_single_launch(self.h, ctypes.c_int64(int(nmax))))

    def solve(self, nume, check_converged=True, **kw):
        """nume: int or per-system list.  Returns (list of eigenvalue arrays, list of
        residual arrays, stats dict); raises NotConverged (results attached) when a solve
        stopped above its residual target, unless check_converged is False."""
        nl = [nume] * len(self.systems) if np.isscalar(nume) else list(nume)
        p = dav_params(max(nl), **kw)
        cnt = np.array(nl, dtype=np.int32)
        tot = int(cnt.sum())
        ev = np.zeros(tot)
        rs = np.zeros(tot)
        st = DavStats()
        rc = lib().cosmo_batch_solve(self.h, ctypes.byref(p), _p(cnt), _p(ev), _p(rs), ctypes.byref(st))
        if rc != ECONV:
            check(rc)
        off = np.concatenate([[0], np.cumsum(cnt)])
        out = ([ev[off[i]:off[i + 1]] for i in range(len(nl))], [rs[off[i]:off[i + 1]] for i in range(len(nl))],
               dict(st.as_dict(), converged=(rc == OK)))
        if rc == ECONV and check_converged:
            raise NotConverged(lib().cosmo_last_error().decode(), out)
        return out

    def close(self):
        if self.h:
            lib().cosmo_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Request loop: admission by HBM footprint, up to max_active solves overlapping."""

    def __init__(self, hbm_budget, max_active=8, precision=PREC_FP32):
        p = ServerParams(int(hbm_budget), int(max_active), int(precision))
        self.h = ctypes.c_void_p()
        check(lib().cosmo_server_create(ctypes.byref(p), ctypes.byref(self.h)))
        self._keep = {}

    def submit(self, system: System, nume, tol=1e-6):
        i = ctypes.c_int64()
        check(lib().cosmo_server_submit(self.h, system.h, int(nume), float(tol), ctypes.byref(i)))
        self._keep[i.value] = system  # the system must outlive its request
        return i.value

    @staticmethod
    def _conv(r):
        k = r.nume
        return dict(id=r.id, n=r.n, nume=k, status=r.status, t_submit=r.t_submit, t_admit=r.t_admit, t_done=r.t_done,
                    evals=np.array(r.evals[:k]), resid=np.array(r.resid[:k]))

    def poll(self, cap=256):
        buf = (Result * cap)()
        k = check(lib().cosmo_server_poll(self.h, buf, cap))
        out = [self._conv(buf[i]) for i in range(k)]
        for r in out:
            self._keep.pop(r["id"], None)
        return out

    def drain(self, cap=4096):
        buf = (Result * cap)()
        w = ctypes.c_int()
        check(lib().cosmo_server_drain(self.h, buf, cap, ctypes.byref(w)))
        out = [self._conv(buf[i]) for i in range(w.value)]
        for r in out:
            self._keep.pop(r["id"], None)
        return out

    def close(self):
        if self.h:
            lib().cosmo_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
