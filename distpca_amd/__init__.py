"""distpca_amd — Python binding of libdpca (include/dpca.h) through ctypes.

The compute path is the HIP library ``distpca_amd/libdpca.so`` built for
gfx950; this module only marshals arguments.  It raises at import of the
library when the shared object is missing: there is no CPU fallback.

Mirrors the reference's process interface (DistPCA.exe flags, defaults and
outputs, src/distpca.cpp:47-644) as ``Params`` / ``PCA``.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libdpca.so")
EXE_PATH = os.path.join(_HERE, "DistPCA.exe")

FP64, MIXED = 0, 1
HOST_PINNED, HOST_PAGEABLE, HOST_FILE = 0, 1, 2


class _Params(ctypes.Structure):
    _fields_ = [
        ("nsv", ctypes.c_int), ("nrhs", ctypes.c_int), ("power", ctypes.c_int),
        ("crit", ctypes.c_int), ("tol", ctypes.c_double), ("bsize", ctypes.c_int),
        ("miter", ctypes.c_int), ("verbose", ctypes.c_int), ("fwrite", ctypes.c_int),
        ("fullsvd", ctypes.c_int), ("memory", ctypes.c_double),
        ("prefix", ctypes.c_char * 1024), ("precision", ctypes.c_int),
        ("device", ctypes.c_int), ("device_cap", ctypes.c_longlong),
        ("strict_passes", ctypes.c_int), ("host_mode", ctypes.c_int),
        ("slices", ctypes.c_int), ("zt_image", ctypes.c_int),
    ]


class Report(ctypes.Structure):
    _fields_ = [
        ("M", ctypes.c_int), ("N", ctypes.c_longlong), ("nsv", ctypes.c_int),
        ("nrhs", ctypes.c_int), ("power", ctypes.c_int), ("bsize", ctypes.c_int),
        ("crit", ctypes.c_int), ("total_its", ctypes.c_int), ("passes", ctypes.c_int),
        ("incore", ctypes.c_int), ("streamed", ctypes.c_int), ("qr_shifted", ctypes.c_int),
        ("t_generate_rhs", ctypes.c_double), ("t_load", ctypes.c_double),
        ("t_mm", ctypes.c_double), ("t_mm_a", ctypes.c_double), ("t_mm_at", ctypes.c_double),
        ("t_gs", ctypes.c_double), ("t_rr", ctypes.c_double), ("t_other", ctypes.c_double),
        ("wall", ctypes.c_double), ("last_crit", ctypes.c_double),
        ("t_h2d", ctypes.c_double), ("h2d_bytes", ctypes.c_double),
        ("host_mode", ctypes.c_int), ("qr_repaired", ctypes.c_int),
        ("resident_bytes", ctypes.c_double), ("graph_its", ctypes.c_int),
        ("zt_image", ctypes.c_int),
    ]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class GenSpec(ctypes.Structure):
    _fields_ = [
        ("M", ctypes.c_int), ("N", ctypes.c_longlong), ("npop", ctypes.c_int),
        ("fst", ctypes.c_double), ("admix_frac", ctypes.c_double),
        ("missing", ctypes.c_double), ("seed", ctypes.c_ulonglong),
        ("pmin", ctypes.c_double), ("pmax", ctypes.c_double),
    ]


class ServerParams(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int), ("hbm_budget", ctypes.c_longlong),
                ("batch_capacity", ctypes.c_int), ("workers", ctypes.c_int)]


class TicketInfo(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("status", ctypes.c_int), ("n_out", ctypes.c_longlong),
                ("t_submit", ctypes.c_double), ("t_start", ctypes.c_double),
                ("t_done", ctypes.c_double), ("batch", ctypes.c_longlong),
                ("batch_samples", ctypes.c_int), ("footprint", ctypes.c_longlong)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class ServerCounters(ctypes.Structure):
    _fields_ = [(f, ctypes.c_longlong) for f in ("batches", "batch_samples", "projections_done",
                                                  "jobs_done", "jobs_rejected", "budget_free")]


TICKET_PROJECT, TICKET_JOB = 1, 2

class FullSvdStats(ctypes.Structure):
    _fields_ = [("seconds", ctypes.c_double), ("frob_norm_angle", ctypes.c_double),
                ("cos_error", ctypes.c_double), ("info", ctypes.c_int)]


EXPORTS = [
    "dpca_params_init", "dpca_open", "dpca_open_packed", "dpca_open_synthetic",
    "dpca_write_synthetic", "dpca_resolve", "dpca_get_dims", "dpca_omega", "dpca_snp_stats", "dpca_decode",
    "dpca_copy_packed", "dpca_pass", "dpca_run", "dpca_crit_history", "dpca_write_outputs",
    "dpca_orthonormalize", "dpca_syev", "dpca_bench_pass", "dpca_close", "dpca_last_error",
    "dpca_footprint", "dpca_model_from_run", "dpca_model_info", "dpca_model_project",
    "dpca_model_free", "dpca_server_params_init", "dpca_server_create",
    "dpca_server_submit_project", "dpca_server_submit_job", "dpca_server_wait",
    "dpca_server_stats", "dpca_server_destroy", "dpca_host_alloc", "dpca_host_free",
    "dpca_full_svd", "dpca_write_fullsvd", "dpca_orthonormalize_noalloc",
]

_lib = None


def lib():
    """Load libdpca.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"libdpca.so not built ({LIB_PATH}); run __graft_entry__.build()")
    L = ctypes.CDLL(LIB_PATH)
    vp, i32, i64, f64 = ctypes.c_void_p, ctypes.c_int, ctypes.c_longlong, ctypes.c_double
    P = ctypes.POINTER
    L.dpca_last_error.restype = ctypes.c_char_p
    L.dpca_params_init.argtypes = [P(_Params)]
    L.dpca_open.argtypes = [ctypes.c_char_p, P(_Params), P(vp)]
    L.dpca_open_packed.argtypes = [vp, i32, i64, P(_Params), P(vp)]
    L.dpca_open_synthetic.argtypes = [P(GenSpec), P(_Params), P(vp)]
    L.dpca_write_synthetic.argtypes = [ctypes.c_char_p, P(GenSpec)]
    L.dpca_resolve.argtypes = [i32, i64, P(_Params), P(_Params)]
    L.dpca_get_dims.argtypes = [vp, P(i32), P(i64), P(i32), P(i32), P(i32)]
    L.dpca_omega.argtypes = [i32, i32, vp]
    L.dpca_snp_stats.argtypes = [vp, vp, vp, vp]
    L.dpca_decode.argtypes = [vp, i64, i64, vp]
    L.dpca_copy_packed.argtypes = [vp, i64, i64, vp]
    L.dpca_pass.argtypes = [vp, vp, vp]
    L.dpca_run.argtypes = [vp, vp, vp, P(Report)]
    L.dpca_crit_history.argtypes = [vp, vp, i32]
    L.dpca_write_outputs.argtypes = [vp, ctypes.c_char_p, vp, vp]
    L.dpca_orthonormalize.argtypes = [vp, i64, i32, vp, P(i32), i32]
    L.dpca_syev.argtypes = [vp, i32, vp, vp, i32]
    L.dpca_orthonormalize_noalloc.argtypes = [vp, i64, i32, vp, P(i32), i32, vp]
    L.dpca_bench_pass.argtypes = [vp, vp, i32, vp]
    L.dpca_close.argtypes = [vp]
    L.dpca_close.restype = None
    L.dpca_footprint.argtypes = [i32, i64, P(_Params), P(i64)]
    L.dpca_model_from_run.argtypes = [vp, vp, vp, i32, P(vp)]
    L.dpca_model_info.argtypes = [vp, P(i64), P(i32), P(i64)]
    L.dpca_model_project.argtypes = [vp, vp, i32, vp]
    L.dpca_model_free.argtypes = [vp]
    L.dpca_model_free.restype = None
    L.dpca_server_params_init.argtypes = [P(ServerParams)]
    L.dpca_server_create.argtypes = [P(ServerParams), vp, P(vp)]
    L.dpca_server_submit_project.argtypes = [vp, vp, i32, P(i64)]
    L.dpca_server_submit_job.argtypes = [vp, P(GenSpec), P(_Params), P(i64)]
    L.dpca_server_wait.argtypes = [vp, i64, vp, i64, P(TicketInfo)]
    L.dpca_server_stats.argtypes = [vp, P(ServerCounters)]
    L.dpca_server_destroy.argtypes = [vp]
    L.dpca_server_destroy.restype = None
    L.dpca_host_alloc.argtypes = [i64]
    L.dpca_host_alloc.restype = vp
    L.dpca_host_free.argtypes = [vp]
    L.dpca_host_free.restype = None
    L.dpca_full_svd.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp, P(FullSvdStats)]
    L.dpca_write_fullsvd.argtypes = [ctypes.c_char_p, i32, i32, vp, vp, vp, vp]
    _lib = L
    return L


def _check(rc):
    if rc != 0:
        raise RuntimeError(f"libdpca error {rc}: {lib().dpca_last_error().decode()}")


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


@dataclass
class Params:
    """DistPCA.exe flags (defaults of src/utilities.cpp:195-232)."""
    nsv: int = 10
    nrhs: int = 20
    power: int = 1
    crit: int = 1
    tol: float = 1e-3
    bsize: int = 0
    miter: int = 100
    verbose: int = 1
    fwrite: int = 0
    memory: float = 0.0
    prefix: str = "noprefix"
    precision: int = FP64
    device: int = 0
    device_cap: int = 0
    strict_passes: int = 1
    host_mode: int = HOST_PINNED
    slices: int = 0
    zt_image: int = 1

    def c(self) -> _Params:
        p = _Params()
        lib().dpca_params_init(ctypes.byref(p))
        for f in ("nsv", "nrhs", "power", "crit", "tol", "bsize", "miter", "verbose",
                  "fwrite", "memory", "precision", "device", "device_cap", "strict_passes", "host_mode",
                  "slices", "zt_image"):
            setattr(p, f, getattr(self, f))
        p.prefix = self.prefix.encode()
        return p


def resolve(M: int, N: int, params: Params) -> dict:
    """bsize / nsv / nrhs / power as dpca_open resolves them (dpca_resolve,
    src/distpca.cpp:239-352); host-only, no device is touched."""
    p, out = params.c(), _Params()
    _check(lib().dpca_resolve(M, N, ctypes.byref(p), ctypes.byref(out)))
    return {f: getattr(out, f) for f in ("nsv", "nrhs", "power", "bsize")}


def omega(M: int, k: int) -> np.ndarray:
    """Reference-exact Omega (host side of libdpca)."""
    out = np.empty((M, k), np.float64)
    _check(lib().dpca_omega(M, k, _ptr(out)))
    return out


def write_synthetic(prefix: str, spec: GenSpec):
    _check(lib().dpca_write_synthetic(prefix.encode(), ctypes.byref(spec)))


def gen_spec(M, N, npop=3, fst=0.05, admix_frac=0.0, missing=0.01, seed=0,
             pmin=0.05, pmax=0.5) -> GenSpec:
    return GenSpec(M, N, npop, fst, admix_frac, missing, seed, pmin, pmax)


class PCA:
    """One libdpca context (dpca_ctx)."""

    def __init__(self, handle):
        self._h = handle
        M, N, nsv, nrhs, b = (ctypes.c_int(), ctypes.c_longlong(), ctypes.c_int(),
                              ctypes.c_int(), ctypes.c_int())
        _check(lib().dpca_get_dims(self._h, ctypes.byref(M), ctypes.byref(N),
                                   ctypes.byref(nsv), ctypes.byref(nrhs), ctypes.byref(b)))
        self.M, self.N, self.nsv, self.nrhs, self.bsize = M.value, N.value, nsv.value, nrhs.value, b.value

    @classmethod
    def from_bfile(cls, bfile: str, params: Params):
        h = ctypes.c_void_p()
        p = params.c()
        _check(lib().dpca_open(bfile.encode(), ctypes.byref(p), ctypes.byref(h)))
        return cls(h)

    @classmethod
    def from_packed(cls, packed: np.ndarray, M: int, N: int, params: Params):
        packed = np.ascontiguousarray(packed, dtype=np.uint8)
        assert packed.size >= N * ((M + 3) // 4)
        h = ctypes.c_void_p()
        p = params.c()
        _check(lib().dpca_open_packed(_ptr(packed), M, N, ctypes.byref(p), ctypes.byref(h)))
        return cls(h)

    @classmethod
    def synthetic(cls, spec: GenSpec, params: Params):
        h = ctypes.c_void_p()
        p = params.c()
        _check(lib().dpca_open_synthetic(ctypes.byref(spec), ctypes.byref(p), ctypes.byref(h)))
        return cls(h)

    def close(self):
        if self._h:
            lib().dpca_close(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def snp_stats(self):
        s = np.empty(self.N, np.int64)
        c = np.empty(self.N, np.int64)
        z = np.empty((self.N, 4), np.float64)
        _check(lib().dpca_snp_stats(self._h, _ptr(s), _ptr(c), _ptr(z)))
        return s, c, z

    def decode(self, snp0: int, rows: int) -> np.ndarray:
        out = np.empty((rows, self.M), np.uint8)
        _check(lib().dpca_decode(self._h, snp0, rows, _ptr(out)))
        return out

    def packed(self, snp0: int = 0, rows: int | None = None) -> np.ndarray:
        rows = self.N - snp0 if rows is None else rows
        out = np.empty((rows, (self.M + 3) // 4), np.uint8)
        _check(lib().dpca_copy_packed(self._h, snp0, rows, _ptr(out)))
        return out

    def pass_(self, Q: np.ndarray) -> np.ndarray:
        Q = np.ascontiguousarray(Q, dtype=np.float64)
        assert Q.shape == (self.M, self.nrhs)
        out = np.empty_like(Q)
        _check(lib().dpca_pass(self._h, _ptr(Q), _ptr(out)))
        return out

    def run(self, rhs: np.ndarray | None = None):
        """Returns (sigma[nsv], rhs[M, nrhs], report dict)."""
        if rhs is None:
            rhs = omega(self.M, self.nrhs)
        rhs = np.array(rhs, dtype=np.float64, order="C", copy=True)
        assert rhs.shape == (self.M, self.nrhs)
        sigma = np.empty(self.nsv, np.float64)
        rep = Report()
        _check(lib().dpca_run(self._h, _ptr(rhs), _ptr(sigma), ctypes.byref(rep)))
        return sigma, rhs, rep.as_dict()

    def bench_pass(self, Q: np.ndarray, npass: int):
        """Event-timed passes: (ms per pass, ms Z Q kernel, ms Z^T T kernel, ms other,
        host wall ms per pass, host-to-device copy ms per pass, small-kernel ms,
        host mode of a streamed matrix or -1, host-to-device bytes per pass)."""
        Q = np.ascontiguousarray(Q, dtype=np.float64)
        ms = np.zeros(9, np.float64)
        _check(lib().dpca_bench_pass(self._h, _ptr(Q), npass, _ptr(ms)))
        return ms

    def crit_history(self):
        buf = np.empty(1024, np.float64)
        n = lib().dpca_crit_history(self._h, _ptr(buf), 1024)
        return buf[:n].copy()

    def full_svd(self, rhs: np.ndarray, sigma: np.ndarray):
        """-fullSVD diagnostic against a finished run: dict with true_sigma,
        true_U (M x nsv), sigma_relerr, vec_relerr, cos_values and the stats."""
        M, nsv = self.M, self.nsv
        rhs = np.ascontiguousarray(rhs, dtype=np.float64)
        sigma = np.ascontiguousarray(sigma, dtype=np.float64)
        out = {"true_sigma": np.empty(nsv), "true_U": np.empty((M, nsv)),
               "sigma_relerr": np.empty(nsv), "vec_relerr": np.empty((M, nsv)),
               "cos_values": np.empty(nsv)}
        st = FullSvdStats()
        _check(lib().dpca_full_svd(self._h, _ptr(rhs), _ptr(sigma), _ptr(out["true_sigma"]),
                                   _ptr(out["true_U"]), _ptr(out["sigma_relerr"]),
                                   _ptr(out["vec_relerr"]), _ptr(out["cos_values"]),
                                   ctypes.byref(st)))
        out.update({f: getattr(st, f) for f, _ in st._fields_})
        return out

    def write_outputs(self, prefix: str, rhs: np.ndarray, sigma: np.ndarray):
        rhs = np.ascontiguousarray(rhs, dtype=np.float64)
        sigma = np.ascontiguousarray(sigma, dtype=np.float64)
        _check(lib().dpca_write_outputs(self._h, prefix.encode(), _ptr(rhs), _ptr(sigma)))


def orthonormalize(A: np.ndarray, device: int = 0):
    A = np.ascontiguousarray(A, dtype=np.float64)
    m, k = A.shape
    Q = np.empty_like(A)
    sh = ctypes.c_int()
    _check(lib().dpca_orthonormalize(_ptr(A), m, k, _ptr(Q), ctypes.byref(sh), device))
    return Q, bool(sh.value)


def syev(B: np.ndarray, device: int = 0):
    B = np.ascontiguousarray(B, dtype=np.float64)
    k = B.shape[0]
    w = np.empty(k, np.float64)
    X = np.empty((k, k), np.float64)
    _check(lib().dpca_syev(_ptr(B), k, _ptr(w), _ptr(X), device))
    return w, X


def footprint(M: int, N: int, params: Params) -> int:
    b = ctypes.c_longlong()
    p = params.c()
    _check(lib().dpca_footprint(M, N, ctypes.byref(p), ctypes.byref(b)))
    return b.value


class Model:
    """Resident SNP loadings + training statistics (dpca_model)."""

    def __init__(self, pca: PCA, rhs: np.ndarray, sigma: np.ndarray, nsv: int):
        rhs = np.ascontiguousarray(rhs, dtype=np.float64)
        sigma = np.ascontiguousarray(sigma, dtype=np.float64)
        h = ctypes.c_void_p()
        _check(lib().dpca_model_from_run(pca._h, _ptr(rhs), _ptr(sigma), nsv, ctypes.byref(h)))
        self._h = h
        N, k, b = ctypes.c_longlong(), ctypes.c_int(), ctypes.c_longlong()
        lib().dpca_model_info(h, ctypes.byref(N), ctypes.byref(k), ctypes.byref(b))
        self.N, self.nsv, self.bytes = N.value, k.value, b.value

    def project(self, packed: np.ndarray, n: int) -> np.ndarray:
        packed = np.ascontiguousarray(packed, dtype=np.uint8)
        assert packed.size >= self.N * ((n + 3) // 4)
        out = np.empty((n, self.nsv), np.float64)
        _check(lib().dpca_model_project(self._h, _ptr(packed), n, _ptr(out)))
        return out

    def close(self):
        if self._h:
            lib().dpca_model_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """dpca_server: continuous-batching projections + HBM-admitted PCA jobs."""

    def __init__(self, model: "Model | None", hbm_budget: int = 64 << 30,
                 batch_capacity: int = 8192, workers: int = 2, device: int = 0):
        p = ServerParams()
        lib().dpca_server_params_init(ctypes.byref(p))
        p.device, p.hbm_budget, p.batch_capacity, p.workers = device, hbm_budget, batch_capacity, workers
        h = ctypes.c_void_p()
        _check(lib().dpca_server_create(ctypes.byref(p), model._h if model else None,
                                        ctypes.byref(h)))
        self._h = h
        self.model = model
        self._keep = {}

    def submit_project(self, packed: np.ndarray, n: int) -> int:
        packed = np.ascontiguousarray(packed, dtype=np.uint8)
        t = ctypes.c_longlong()
        _check(lib().dpca_server_submit_project(self._h, _ptr(packed), n, ctypes.byref(t)))
        self._keep[t.value] = (packed, n)
        return t.value

    def submit_job(self, spec: GenSpec, params: Params) -> int:
        t = ctypes.c_longlong()
        p = params.c()
        _check(lib().dpca_server_submit_job(self._h, ctypes.byref(spec), ctypes.byref(p),
                                            ctypes.byref(t)))
        self._keep[t.value] = (None, params.nsv)
        return t.value

    def wait(self, ticket: int):
        payload, n = self._keep.pop(ticket)
        cap = n * self.model.nsv if payload is not None else n
        out = np.empty(max(cap, 1), np.float64)
        info = TicketInfo()
        rc = lib().dpca_server_wait(self._h, ticket, _ptr(out), cap, ctypes.byref(info))
        _check(rc)
        res = out[:info.n_out]
        if payload is not None:
            res = res.reshape(n, self.model.nsv)
        return res, info.as_dict()

    def counters(self):
        c = ServerCounters()
        lib().dpca_server_stats(self._h, ctypes.byref(c))
        return {f: getattr(c, f) for f, _ in c._fields_}

    def close(self):
        if self._h:
            lib().dpca_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class PinnedArray:
    """A page-locked host uint8 buffer (dpca_host_alloc) viewed as numpy."""

    def __init__(self, shape):
        n = int(np.prod(shape))
        self._p = lib().dpca_host_alloc(n)
        if not self._p:
            raise RuntimeError(lib().dpca_last_error().decode())
        buf = (ctypes.c_uint8 * n).from_address(self._p)
        self.array = np.frombuffer(buf, dtype=np.uint8).reshape(shape)

    def close(self):
        if self._p:
            self.array = None
            lib().dpca_host_free(self._p)
            self._p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
