"""ORACLE — test infrastructure, not product code.

CPU restatement of DistPCA's randomized block subspace iteration
(``BlockSubspaceIter_MPI_OOC_double_buffering``, src/methods.cpp:702-1135, and
the in-core ``SubspaceIteration_MPI``, src/methods.cpp:24-325) for one rank.
The bit-exact pieces (Omega, 2-bit decode, standardisation, cached per-SNP
table) are in dpca_oracle.c; this module drives them and does the linear
algebra with the same LAPACK routines the reference calls: dgeqrf + dorgqr for
the orthonormalisation and dsyev('V','U') for the Rayleigh-Ritz problem
(scipy.linalg.lapack), dgemm through numpy.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg may
import this module, and only as the checker or the timed CPU baseline.

How each function is pinned (tests/test_oracle.py):
  omega / rand_val / norm2   bit-identical to the reference's own
                             src/gennorm.c compiled into oracle/_ref, incl. the
                             int32 seed wrap at M >= 46341; Park-Miller known
                             answer (seed 1, 10000 steps -> 1043618065).
  decode / decode_precomp    exhaustive over all 256 byte values against the
                             PLINK .bed code table (00 hom, 01 missing, 10 het,
                             11 hom-alt) written out independently in the test.
  standardize / snp_stats    closed form recomputed in numpy from genotype
                             counts; monomorphic and all-missing SNPs -> 0.
  decode_block               first-touch and cached passes agree bit for bit.
  OOCPass                    brute-force dense Z^T (Z Q) on small inputs.
  resolve                    against the library's own resolution
                             (dpca_resolve) over the auto-bsize branches
                             (fixed -memory: positive and non-positive
                             available memory, block >= N) and the clamps.
  write_outputs              byte-identical to the head of the reference-held
                             docs/results/accuracy/1000_genomes_
                             leftSingularVectors.txt rewritten from its own
                             values and IIDs.
  orthonormalise             Q^T Q = I, Q R = A, matches numpy's QR up to signs.
  rayleigh_ritz              eigenvalues match numpy.linalg.eigh.
  rsi (whole loop)           HGDP example, nsv 20, nrhs 40, crit 2, tol 1e-3,
                             bsize 100 stops after 11 outer iterations, as in
                             the reference's published log
                             (example/hgdp.qc.pruned.log:40); converged
                             eigenvalues match a dense eigendecomposition of
                             Z^T Z.
  full_svd_diag              the comparisons of -fullSVD: exact cases (the
                             true vectors themselves, sign flips, a known
                             plane rotation by theta: ||.||_F =
                             2 sqrt(1 - cos theta), ||.||_2 = 2 sin(theta/2))
                             and numpy's SVD reconstructing the matrix.
"""
from __future__ import annotations

import ctypes
import math
import os
import time
from dataclasses import dataclass, field

import numpy as np
import scipy.linalg.lapack as lapack

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def lib():
    """Load oracle/liboracle.so (built by oracle/Makefile)."""
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle library missing: {path} (run make -C oracle)")
        L = ctypes.CDLL(path)
        vp, i32, i64, u64, f64 = (ctypes.c_void_p, ctypes.c_int, ctypes.c_int64,
                                  ctypes.c_uint64, ctypes.c_double)
        L.orc_omega.argtypes = [i32, i32, vp]
        L.orc_lcg_after.argtypes = [i32, i32]
        L.orc_lcg_after.restype = ctypes.c_long
        L.orc_decode.argtypes = [vp, vp, ctypes.c_uint]
        L.orc_decode_precomp.argtypes = [vp, vp, ctypes.c_uint]
        L.orc_standardize.argtypes = [vp, vp, i32, vp, vp, f64]
        L.orc_decode_block.argtypes = [vp, i32, i32, u64, i64, f64, vp, vp, vp, vp, vp]
        L.orc_snp_stats.argtypes = [vp, i64, i32, vp, vp, vp, vp, vp]
        _LIB = L
    return _LIB


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


# --------------------------------------------------------------- generator --

def omega(M: int, k: int) -> np.ndarray:
    """Omega, M x k row-major, src/distpca.cpp:384-397 (gennorm.c:21-72)."""
    out = np.empty((M, k), dtype=np.float64)
    lib().orc_omega(M, k, _p(out))
    return out


def lcg_after(seed: int, n: int) -> int:
    return lib().orc_lcg_after(seed, n)


# ------------------------------------------------------------------ decode --

def decode(packed_row: np.ndarray) -> np.ndarray:
    """{0,3,1,2} decode of n bytes -> 4n values (src/utilities.cpp:351-414)."""
    b = np.ascontiguousarray(packed_row, dtype=np.uint8)
    out = np.empty(4 * b.size, dtype=np.uint8)
    lib().orc_decode(_p(out), _p(b), b.size)
    return out


def decode_precomp(packed_row: np.ndarray) -> np.ndarray:
    """{3,1,2,0} decode (src/utilities.cpp:417-480)."""
    b = np.ascontiguousarray(packed_row, dtype=np.uint8)
    out = np.empty(4 * b.size, dtype=np.uint8)
    lib().orc_decode_precomp(_p(out), _p(b), b.size)
    return out


def standardize(g: np.ndarray, inv_sqrtN: float):
    """src/utilities.cpp:272-306 -> (z, avg, sd)."""
    g = np.ascontiguousarray(g, dtype=np.uint8)
    z = np.empty(g.size, dtype=np.float64)
    avg = ctypes.c_double()
    sd = ctypes.c_double()
    lib().orc_standardize(_p(z), _p(g), g.size, ctypes.byref(avg), ctypes.byref(sd),
                          float(inv_sqrtN))
    return z, avg.value, sd.value


def snp_stats(packed: np.ndarray, N: int, M: int):
    """Per-SNP (sum, ctr, avg, sd, zval[N,4]) in the reference's arithmetic."""
    packed = np.ascontiguousarray(packed, dtype=np.uint8)
    np_ = (M + 3) // 4
    assert packed.size >= N * np_
    s = np.empty(N, np.int64)
    c = np.empty(N, np.int64)
    a = np.empty(N, np.float64)
    d = np.empty(N, np.float64)
    z = np.empty((N, 4), np.float64)
    lib().orc_snp_stats(_p(packed), N, M, _p(s), _p(c), _p(a), _p(d), _p(z))
    return s, c, a, d, z


class BlockDecoder:
    """Decode + standardise state of one run: norm_precomp[4N] and seen[N]
    (src/methods.cpp:776-779), filled on first touch of each SNP."""

    def __init__(self, M: int, N: int):
        self.M, self.N = M, N
        self.np = (M + 3) // 4
        self.inv_sqrtN = 1.0 / math.sqrt(float(N))
        self.norm_precomp = np.zeros(4 * N, np.float64)
        self.seen = np.zeros(N, np.uint8)
        self._dec = np.empty(4 * self.np, np.uint8)
        self._tmp = np.empty(M, np.float64)

    def block(self, packed_rows: np.ndarray, snp0: int, rows: int,
              out: np.ndarray | None = None) -> np.ndarray:
        if out is None:
            out = np.empty((rows, self.M), np.float64)
        pr = np.ascontiguousarray(packed_rows, dtype=np.uint8)
        assert pr.size >= rows * self.np and out.shape[0] >= rows
        lib().orc_decode_block(_p(pr), rows, self.M, self.np, snp0, self.inv_sqrtN,
                               _p(self.norm_precomp), _p(self.seen), _p(out),
                               _p(self._dec), _p(self._tmp))
        return out[:rows]


def dense_z(packed: np.ndarray, N: int, M: int) -> np.ndarray:
    """The whole standardised N x M matrix (first-touch decode of every SNP,
    as Read_Bed_Local does, src/utilities.cpp:482-518)."""
    dec = BlockDecoder(M, N)
    np_ = (M + 3) // 4
    return dec.block(packed.reshape(-1)[: N * np_], 0, N).copy()


# ------------------------------------------------------------ PLINK files --

def count_lines(path: str):
    """Line handling of src/distpca.cpp:122-147: every newline-terminated line
    counts; an unterminated last line is dropped."""
    with open(path, "rb") as f:
        data = f.read()
    lines = data.split(b"\n")
    return [ln.decode() for ln in lines[:-1]]


def read_plink(prefix: str):
    fam = count_lines(prefix + ".fam")
    bim = count_lines(prefix + ".bim")
    iids = [ln.split()[1] for ln in fam]
    snps = [ln.split()[1] for ln in bim]
    M, N = len(fam), len(bim)
    np_ = (M + 3) // 4
    with open(prefix + ".bed", "rb") as f:
        f.seek(3)
        packed = np.frombuffer(f.read(N * np_), dtype=np.uint8)
    return M, N, iids, snps, packed


# ------------------------------------------------------------- parameters --

@dataclass
class Params:
    """Defaults of initialize_structure (src/utilities.cpp:195-232)."""
    nsv: int = 10
    nrhs: int = 20
    power: int = 1
    crit: int = 1
    tol: float = 1e-3
    miter: int = 100
    bsize: int = 0
    memory: float = 0.0


def resolve(p: Params, M: int, N: int, ram_gb: float = 64.0) -> Params:
    """bsize and NSV/NRHS clamps of src/distpca.cpp:218-352 for one rank."""
    q = Params(**p.__dict__)
    mem = (ram_gb / 10.0) if q.memory == 0.0 else q.memory
    if q.bsize <= 0:
        membuff = (3.0 * N * 8.0 + M * q.nsv * 8.0 + 2.0 * N * q.nrhs
                   + 3.0 * M * 8.0 + 2048.0 * 100000.0)
        avail = mem * 1e9 - membuff
        blk = avail / (8.0 * M) if avail > 0.0 else (membuff + M * 8.0) / 1e7
        q.bsize = int(blk)
        if q.bsize <= 0:
            q.bsize = q.nsv
        if q.bsize >= N:
            q.bsize = N
    else:
        if q.bsize > N:
            q.bsize = N
        if q.bsize < 1:
            q.bsize = q.nsv
    if q.nsv > min(M, N):
        q.nsv = min(M, N)
    if q.nrhs <= 0:
        q.nrhs = min(2 * q.nsv, min(M, N))
    if q.nrhs > min(M, N):
        q.nrhs = min(M, N)
    if q.nrhs < q.nsv:
        q.nrhs = min(2 * q.nsv, min(M, N))
    return q


# ------------------------------------------------------------ linear algebra --

def orthonormalise(A: np.ndarray) -> np.ndarray:
    """LAPACKE_dgeqrf + LAPACKE_dorgqr on the M x k matrix (src/methods.cpp:947-948)."""
    qr, tau, _, info = lapack.dgeqrf(np.asfortranarray(A))
    assert info == 0
    q, _, info = lapack.dorgqr(qr, tau)
    assert info == 0
    return np.ascontiguousarray(q)


def rayleigh_ritz(Q: np.ndarray, AQ: np.ndarray):
    """B2 = Q^T (A A^T Q); keep the upper triangle; dsyev('V','U'); reverse to
    descending order (src/methods.cpp:970-1000).  Returns (w ascending, X
    with columns in descending eigenvalue order)."""
    B2 = Q.T @ AQ
    B2 = np.triu(B2)
    w, v, info = lapack.dsyev(B2, compute_v=1, lower=0)
    assert info == 0
    return w, v[:, ::-1].copy()


# ------------------------------------------------------------------ driver --

@dataclass
class Result:
    sigma: np.ndarray
    rhs: np.ndarray                    # M x k final Ritz vectors
    total_its: int
    delta_iter: list = field(default_factory=list)
    crit_values: list = field(default_factory=list)
    passes: int = 0
    timers: dict = field(default_factory=dict)


class OOCPass:
    """One pass A A^T Q over the SNP blocks (src/methods.cpp:786-936): decode
    block, T = Z_b Q, ARHS += Z_b^T T."""

    def __init__(self, packed: np.ndarray, M: int, N: int, bsize: int, progress=None):
        # progress(j, nblk), when given, is called after each block (long
        # CPU runs report their progress through it)
        self.progress = progress
        self.packed = packed.reshape(-1)
        self.M, self.N, self.b = M, N, bsize
        self.np = (M + 3) // 4
        self.dec = BlockDecoder(M, N)
        self.buf = np.empty((bsize, M), np.float64)
        self.t_load = 0.0
        self.t_mm_a = 0.0
        self.t_mm_at = 0.0
        self.count = 0

    def __call__(self, Q: np.ndarray, max_blocks: int | None = None) -> np.ndarray:
        M, N, b, np_ = self.M, self.N, self.b, self.np
        ARHS = np.zeros((M, Q.shape[1]), np.float64)
        nblk = (N + b - 1) // b
        if max_blocks is not None:
            nblk = min(nblk, max_blocks)
        for j in range(nblk):
            s0 = j * b
            rows = min(b, N - s0)
            t0 = time.perf_counter()
            Zb = self.dec.block(self.packed[s0 * np_:(s0 + rows) * np_], s0, rows, self.buf)
            t1 = time.perf_counter()
            T = Zb @ Q
            t2 = time.perf_counter()
            ARHS += Zb.T @ T
            t3 = time.perf_counter()
            self.t_load += t1 - t0
            self.t_mm_a += t2 - t1
            self.t_mm_at += t3 - t2
            if self.progress:
                self.progress(j, nblk)
        self.count += 1
        return ARHS


def rsi(packed: np.ndarray, M: int, N: int, p: Params, rhs0: np.ndarray | None = None,
        verbose: bool = False) -> Result:
    """BlockSubspaceIter_MPI_OOC_double_buffering for one rank
    (src/methods.cpp:702-1135); p must already be resolved."""
    k, nsv = p.nrhs, p.nsv
    RHS2 = omega(M, k) if rhs0 is None else np.array(rhs0, dtype=np.float64, copy=True)
    op = OOCPass(packed, M, N, p.bsize)
    res = Result(sigma=np.zeros(nsv), rhs=RHS2, total_its=0)
    t_gs = t_rr = 0.0
    RHS2_old = RHS2.copy()
    s_old = np.zeros(nsv)
    ii = 0
    while ii < p.miter:
        ARHS = op(RHS2)
        for _ in range(p.power - 1):
            t0 = time.perf_counter()
            Q = orthonormalise(ARHS)
            t_gs += time.perf_counter() - t0
            RHS2 = Q
            ARHS = op(RHS2)
        t0 = time.perf_counter()
        Q = orthonormalise(ARHS)
        t_gs += time.perf_counter() - t0
        RHS2 = Q
        ARHS = op(RHS2)
        t0 = time.perf_counter()
        w, X = rayleigh_ritz(Q, ARHS)
        t_rr += time.perf_counter() - t0
        RHS2 = Q @ X
        sig = np.sqrt(w[::-1][:nsv])
        res.delta_iter.append(float(np.sum(sig)))
        stop = False
        if ii > 0:
            if p.crit == 0:
                e = abs(res.delta_iter[ii - 1] - res.delta_iter[ii]) / res.delta_iter[ii]
                res.crit_values.append(e)
                stop = e <= p.tol
            elif p.crit == 1:
                conv = int(np.sum(np.abs((sig - s_old) / sig) <= p.tol))
                res.crit_values.append(conv)
                stop = conv == nsv
            elif p.crit == 2:
                dots = np.einsum("ij,ij->j", RHS2_old[:, :nsv], RHS2[:, :nsv])
                mev = 1.0 - float(np.sum(dots * dots)) / nsv
                res.crit_values.append(mev)
                stop = mev <= p.tol
            if verbose:
                print(f"iteration {ii}: {res.crit_values[-1]}")
        res.sigma = sig
        if stop:
            ii += 1
            break
        s_old = sig.copy()
        RHS2_old = RHS2.copy()
        ii += 1
    res.total_its = ii
    res.rhs = RHS2
    res.passes = op.count
    res.timers = dict(load=op.t_load, mm_a=op.t_mm_a, mm_at=op.t_mm_at, gs=t_gs, rr=t_rr)
    return res


def rsi_incore(Z: np.ndarray, p: Params, rhs0: np.ndarray | None = None) -> Result:
    """SubspaceIteration_MPI for one rank (src/methods.cpp:24-325): power
    passes, then QR, a half pass RHS_local = Z Q and B = RHS_local^T RHS_local."""
    N, M = Z.shape
    k, nsv = p.nrhs, p.nsv
    RHS2 = omega(M, k) if rhs0 is None else np.array(rhs0, dtype=np.float64, copy=True)
    res = Result(sigma=np.zeros(nsv), rhs=RHS2, total_its=0)
    RHS2_old = RHS2.copy()
    s_old = np.zeros(nsv)
    ii = 0
    while ii < p.miter:
        ARHS = Z.T @ (Z @ RHS2)
        for _ in range(p.power - 1):
            RHS2 = orthonormalise(ARHS)
            ARHS = Z.T @ (Z @ RHS2)
        Q = orthonormalise(ARHS)
        RL = Z @ Q
        B = np.triu(RL.T @ RL)
        w, v, info = lapack.dsyev(B, compute_v=1, lower=0)
        X = v[:, ::-1]
        RHS2 = Q @ X
        sig = np.sqrt(w[::-1][:nsv])
        res.delta_iter.append(float(np.sum(sig)))
        stop = False
        if ii > 0:
            if p.crit == 0:
                e = abs(res.delta_iter[ii - 1] - res.delta_iter[ii]) / res.delta_iter[ii]
                res.crit_values.append(e)
                stop = e <= p.tol
            elif p.crit == 1:
                conv = int(np.sum(np.abs((sig - s_old) / sig) <= p.tol))
                res.crit_values.append(conv)
                stop = conv == nsv
            elif p.crit == 2:
                dots = np.einsum("ij,ij->j", RHS2_old[:, :nsv], RHS2[:, :nsv])
                mev = 1.0 - float(np.sum(dots * dots)) / nsv
                res.crit_values.append(mev)
                stop = mev <= p.tol
        res.sigma = sig
        if stop:
            ii += 1
            break
        s_old = sig.copy()
        RHS2_old = RHS2.copy()
        ii += 1
    res.total_its = ii
    res.rhs = RHS2
    return res


# ----------------------------------------------------------------- metrics --

def full_svd_diag(Zt: np.ndarray, Uhat: np.ndarray, sigma: np.ndarray,
                  U: np.ndarray | None = None, s: np.ndarray | None = None) -> dict:
    """-fullSVD diagnostic (src/distpca.cpp:449-570, computeCosineError in
    src/utilities.cpp:234-270): economy SVD of the M x N standardised matrix Zt
    (dgesvd there, numpy here unless U, s are given), then the comparisons
    with the run's nsv Ritz vectors Uhat (M x nsv) and sigma (nsv)."""
    nsv = Uhat.shape[1]
    if U is None:
        U, s, _ = np.linalg.svd(Zt, full_matrices=False)
    U, s = U[:, :nsv], s[:nsv]
    den = np.where(s != 0, s, 1.0)
    sig_err = np.abs(s - sigma) / den
    vec_err = np.abs(np.abs(U) - np.abs(Uhat)) / (np.abs(U) + 1e-12)
    C = Uhat.T @ U
    d = np.diag(C).copy()
    frob = math.sqrt(float(np.sum((np.abs(d) - 1.0) ** 2) + np.sum(C ** 2) - np.sum(d ** 2)))
    C2 = C.copy()
    np.fill_diagonal(C2, np.abs(d) - 1.0)
    cos_err = float(np.linalg.svd(C2, compute_uv=False)[0])
    return {"true_sigma": s, "true_U": U, "sigma_relerr": sig_err, "vec_relerr": vec_err,
            "cos_values": d, "frob_norm_angle": frob, "cos_error": cos_err}


def principal_angles(U: np.ndarray, V: np.ndarray) -> np.ndarray:
    """Principal angles (radians) between span(U) and span(V), ascending.
    Uses the sine form for small angles (accurate below 1e-8)."""
    qu, _ = np.linalg.qr(U)
    qv, _ = np.linalg.qr(V)
    c = np.linalg.svd(qu.T @ qv, compute_uv=False)
    r = qv - qu @ (qu.T @ qv)
    s = np.linalg.svd(r, compute_uv=False)
    s = np.sort(s)
    ang_s = np.arcsin(np.clip(s, 0, 1))
    ang_c = np.arccos(np.clip(np.sort(c)[::-1], -1, 1))
    return np.where(ang_s < 0.5, ang_s, ang_c)


def write_outputs(prefix: str, sigma, rhs, nsv: int, iids):
    """<prefix>_singularValues.txt and <prefix>_leftSingularVectors.txt in the
    reference's formats (src/methods.cpp:1077-1118)."""
    with open(prefix + "_singularValues.txt", "w") as f:
        for s in sigma[:nsv]:
            f.write("%2.13f\n" % s)
    with open(prefix + "_leftSingularVectors.txt", "w") as f:
        f.write("FID" + "".join("\tPC%d" % j for j in range(nsv)) + "\n")
        for i, iid in enumerate(iids):
            f.write("%8s" % iid + "".join("\t%2.13f" % rhs[i, j] for j in range(nsv)) + "\n")
