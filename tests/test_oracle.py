"""Pins the oracle (CPU, no GPU): each check is chosen so that a plausible
slip in the restatement (wrong table, dropped term, transposed operand, wrong
seed wrap) fails it."""
import ctypes
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np
import pytest

from conftest import REF_EXAMPLE, ROOT
from oracle import oracle as O

REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libref.so")


def _ref_omega(M, k):
    # fresh copy per call: the reference keeps its generator state in a
    # function-static, which must start at 0 as in a fresh process
    d = tempfile.mkdtemp()
    p = os.path.join(d, "libref.so")
    shutil.copy(REF_LIB, p)
    lib = ctypes.CDLL(p)
    out = np.empty((M, k))
    lib.ref_omega(M, k, out.ctypes.data_as(ctypes.c_void_p))
    shutil.rmtree(d)
    return out


def test_park_miller_known_answer():
    # Park & Miller (1988): seed 1, 10,000 steps -> 1,043,618,065
    assert O.lcg_after(1, 10000) == 1043618065


@pytest.mark.skipif(not os.path.exists(REF_LIB), reason="oracle/_ref not built")
@pytest.mark.parametrize("M,k", [(942, 40), (1024, 20), (50000, 20), (100000, 40),
                                 (400000, 40)])
def test_omega_matches_compiled_reference(M, k):
    a = O.omega(M, k)
    b = _ref_omega(M, k)
    assert np.array_equal(a.view(np.int64), b.view(np.int64))


def test_omega_first_element_is_plus_zero():
    # seed 0 at (0,0) lands on the LCG fixed point: u = 1, r = -0.0, +0 after "+ mean"
    a = O.omega(4, 3)
    assert a[0, 0] == 0.0 and not np.signbit(a[0, 0])
    assert a.std() > 0.1  # not a constant


def test_omega_wrap_chain():
    # for M >= 46341 the seed jj*M+kk wraps; rows past 2^31/M continue the chain
    M, k = 50000, 4
    a = O.omega(M, k)
    jj = 2 ** 31 // M + 1
    seeds = (np.uint32(jj) * np.uint32(M) + np.arange(k, dtype=np.uint32)).view(np.int32)
    assert (seeds <= 0).any()
    assert np.isfinite(a).all() and len(np.unique(a[jj:jj + 50])) > 150


PLINK = {0b00: 0, 0b01: 3, 0b10: 1, 0b11: 2}  # hom A1, missing, het, hom A2 (count of A2)


def test_decode_exhaustive():
    b = np.arange(256, dtype=np.uint8)
    out = O.decode(b).reshape(256, 4)
    pre = O.decode_precomp(b).reshape(256, 4)
    for v in range(256):
        for j in range(4):
            code = (v >> (2 * j)) & 3
            assert out[v, j] == PLINK[code]
            # precomp index: 3 <- g=0, 2 <- g=1, 0 <- g=2, 1 <- missing
            assert pre[v, j] == {0: 3, 3: 1, 1: 2, 2: 0}[PLINK[code]]


def test_standardize_closed_form():
    rng = np.random.default_rng(1)
    g = rng.choice([0, 1, 2, 3], size=1001, p=[0.5, 0.3, 0.15, 0.05]).astype(np.uint8)
    N = 12345
    z, avg, sd = O.standardize(g, 1.0 / np.sqrt(N))
    ok = g != 3
    a = g[ok].sum() / ok.sum()
    P = a / 2.0
    s = np.sqrt(2.0 * P * (1 - P))
    assert avg == a and sd == s
    ref = np.where(ok, (g - a) / s * (1.0 / np.sqrt(N)), 0.0)
    assert np.array_equal(z, ref)
    # binomial sd, not the sample sd
    assert abs(sd - g[ok].std()) > 1e-6


def test_standardize_degenerate():
    z, avg, sd = O.standardize(np.full(10, 2, np.uint8), 0.1)
    assert sd == 0 and (z == 0).all()
    z, avg, sd = O.standardize(np.full(10, 3, np.uint8), 0.1)
    assert np.isnan(avg) and (z == 0).all()


def _random_packed(M, N, seed=0, miss=0.02):
    rng = np.random.default_rng(seed)
    codes = rng.choice([0, 1, 2, 3], size=(N, M), p=[0.4, miss, 0.35, 0.25 - miss])
    codes[3] = 0          # monomorphic SNP
    codes[5] = 1          # all missing
    np_ = (M + 3) // 4
    pad = np.zeros((N, 4 * np_), np.uint8)
    pad[:, :M] = codes
    packed = (pad[:, 0::4] | (pad[:, 1::4] << 2) | (pad[:, 2::4] << 4) | (pad[:, 3::4] << 6))
    return packed.astype(np.uint8), codes


def test_snp_stats_and_blocks():
    M, N = 37, 40
    packed, codes = _random_packed(M, N)
    s, c, a, d, z = O.snp_stats(packed, N, M)
    g = np.vectorize(PLINK.get)(codes)
    ok = g != 3
    assert np.array_equal(s, np.where(ok, g, 0).sum(1))
    assert np.array_equal(c, ok.sum(1))
    assert (z[3] == 0).all() and (z[5] == 0).all()
    # first-touch and cached decodes give the same bits, and match zval
    dec = O.BlockDecoder(M, N)
    first = dec.block(packed, 0, N).copy()
    again = dec.block(packed, 0, N).copy()
    assert np.array_equal(first, again)
    zz = z[np.arange(N)[:, None], codes]
    assert np.array_equal(first, zz)


def test_pass_brute_force():
    M, N = 53, 300
    packed, _ = _random_packed(M, N, seed=2)
    Z = O.dense_z(packed, N, M)
    Q = np.random.default_rng(3).standard_normal((M, 7))
    dense = np.einsum("si,sj,jk->ik", Z, Z, Q)  # sum_s Z[s,i] Z[s,j] Q[j,k]
    op = O.OOCPass(packed, M, N, 64)
    assert np.allclose(op(Q), dense, rtol=1e-12, atol=1e-12)


def test_orthonormalise_and_rr():
    rng = np.random.default_rng(4)
    A = rng.standard_normal((200, 9))
    Q = O.orthonormalise(A)
    assert np.allclose(Q.T @ Q, np.eye(9), atol=1e-13)
    R = Q.T @ A
    assert np.allclose(np.tril(R, -1), 0, atol=1e-12) and np.allclose(Q @ R, A)
    q2, _ = np.linalg.qr(A)
    assert np.allclose(np.abs(q2.T @ Q), np.eye(9), atol=1e-12)
    S = rng.standard_normal((200, 200))
    S = S @ S.T
    AQ = S @ Q
    w, X = O.rayleigh_ritz(Q, AQ)
    ref = np.linalg.eigvalsh(Q.T @ S @ Q)
    assert np.allclose(w, ref, rtol=1e-12)
    assert np.all(np.diff(w) >= 0)
    B = Q.T @ AQ
    Bs = np.triu(B) + np.triu(B, 1).T
    assert np.allclose(Bs @ X, X * w[::-1], atol=1e-9)


def test_rsi_converges_to_dense_eigs():
    M, N = 60, 2000
    packed, _ = _random_packed(M, N, seed=5)
    Z = O.dense_z(packed, N, M)
    ev = np.linalg.eigvalsh(Z.T @ Z)[::-1]
    p = O.resolve(O.Params(nsv=4, nrhs=12, power=3, crit=1, tol=1e-14, miter=200, bsize=500), M, N)
    r = O.rsi(packed, M, N, p)
    assert np.allclose(r.sigma ** 2, ev[:4], rtol=1e-10)
    # in-core variant: same subspace, same values
    ri = O.rsi_incore(Z, p)
    assert np.allclose(ri.sigma ** 2, ev[:4], rtol=1e-10)


def test_principal_angles_metric():
    rng = np.random.default_rng(6)
    U = np.linalg.qr(rng.standard_normal((100, 5)))[0]
    assert O.principal_angles(U, U @ np.diag([1, -1, 1, 1, -1])).max() < 1e-14
    th = 1e-9
    V = U.copy()
    V[:, 0] = np.cos(th) * U[:, 0] + np.sin(th) * np.linalg.qr(
        np.column_stack([U, rng.standard_normal(100)]))[0][:, 5]
    assert abs(O.principal_angles(U, V).max() - th) < 1e-15


@pytest.mark.skipif(not os.path.exists(REF_EXAMPLE + ".bed"), reason="reference tree absent")
def test_hgdp_stops_after_11_iterations_like_published_log(hgdp_golden):
    # example/hgdp.qc.pruned.log:38-40: power 1, "fetched 22" = 11 x (1+1)
    M, N, iids, snps, packed = O.read_plink(REF_EXAMPLE)
    assert (M, N) == (942, 133594)
    p = O.resolve(O.Params(nsv=20, nrhs=40, crit=2, tol=1e-3, bsize=100), M, N)
    assert p.bsize == 100
    r = O.rsi(packed, M, N, p)
    assert r.total_its == 11 and r.passes == 22
    assert np.allclose(r.sigma, hgdp_golden["sigma"], rtol=1e-12)


def test_golden_fixture_consistency(hgdp_golden, c1_golden):
    assert int(hgdp_golden["total_its"]) == 11
    assert hgdp_golden["crit"][-1] <= 1e-3 < hgdp_golden["crit"][-2]
    assert int(c1_golden["total_its"]) == 20


def test_full_svd_diag_closed_forms():
    # -fullSVD comparisons (src/distpca.cpp:510-570, src/utilities.cpp:234-270)
    rng = np.random.default_rng(3)
    Zt = rng.standard_normal((60, 400))
    U, s, Vt = np.linalg.svd(Zt, full_matrices=False)
    assert np.abs(U * s @ Vt - Zt).max() < 1e-12 * np.abs(Zt).max()
    k = 4
    d = O.full_svd_diag(Zt, U[:, :k], s[:k])
    assert d["frob_norm_angle"] < 1e-12 and d["cos_error"] < 1e-12
    assert np.allclose(d["cos_values"], 1) and d["sigma_relerr"].max() < 1e-14
    # column signs do not matter except in the cosine values themselves
    flip = U[:, :k] * np.array([1, -1, 1, -1])
    d = O.full_svd_diag(Zt, flip, s[:k])
    assert d["frob_norm_angle"] < 1e-12 and d["cos_error"] < 1e-12
    assert np.allclose(d["cos_values"], [1, -1, 1, -1]) and d["vec_relerr"].max() < 1e-12
    # a rotation by theta inside the plane of PCs 1 and 2
    th = 0.3
    Uh = U[:, :k].copy()
    Uh[:, 0] = np.cos(th) * U[:, 0] - np.sin(th) * U[:, 1]
    Uh[:, 1] = np.sin(th) * U[:, 0] + np.cos(th) * U[:, 1]
    d = O.full_svd_diag(Zt, Uh, s[:k] * 1.01)
    assert abs(d["frob_norm_angle"] - 2 * np.sqrt(1 - np.cos(th))) < 1e-12
    assert abs(d["cos_error"] - 2 * np.sin(th / 2)) < 1e-12
    assert np.allclose(d["sigma_relerr"], 0.01)


def _fixture(*parts):
    return os.path.join(ROOT, "tests", "golden", *parts)


def test_write_outputs_matches_reference_file_layout(tmp_path):
    # the head of docs/results/accuracy/1000_genomes_leftSingularVectors.txt
    # (reference-held output, 10 PCs): rewriting its own IIDs and values with
    # the oracle's writer must give the same bytes ("%8s" IID, "\t%2.13f")
    ref = open(_fixture("formats", "leftSingularVectors_head.txt")).read()
    lines = ref.splitlines()
    nsv = len(lines[0].split("\t")) - 1
    iids = [ln.split("\t")[0].strip() for ln in lines[1:]]
    rhs = np.array([[float(v) for v in ln.split("\t")[1:]] for ln in lines[1:]])
    sigma = np.linspace(3.0, 1.0, nsv)
    pre = str(tmp_path / "res")
    O.write_outputs(pre, sigma, rhs, nsv, iids)
    assert open(pre + "_leftSingularVectors.txt").read() == ref
    sv = open(pre + "_singularValues.txt").read().splitlines()
    assert sv == ["%.13f" % s for s in sigma] and sv[0] == "3.0000000000000"


def _ram_gb():
    for ln in open("/proc/meminfo"):
        if ln.startswith("MemTotal:"):
            return int(ln.split()[1]) / 1e6
    raise RuntimeError("no MemTotal")


@pytest.mark.parametrize("M,N,kw", [
    (1024, 50000, dict(nsv=10, nrhs=20, bsize=4096)),             # given, < N
    (1024, 50000, dict(nsv=10, nrhs=20, bsize=60000)),            # > N -> N (in-core)
    (1024, 50000, dict(nsv=10, nrhs=20, bsize=-3)),               # auto
    (100000, 1000000, dict(nsv=20, nrhs=40, memory=2.0)),         # auto from -memory
    (100000, 1000000, dict(nsv=20, nrhs=40, memory=0.05)),        # available <= 0
    (20000, 500000, dict(nsv=10, nrhs=20, memory=500.0)),         # auto block >= N
    (942, 133594, dict(nsv=20, nrhs=40)),                         # RAM / 10
    (30, 5000, dict(nsv=40, nrhs=0, bsize=100)),                  # nsv > min, nrhs <= 0
    (30, 5000, dict(nsv=8, nrhs=50, bsize=100)),                  # nrhs > min
    (500, 5000, dict(nsv=12, nrhs=5, bsize=100)),                 # nrhs < nsv
    (500, 20, dict(nsv=12, nrhs=30, bsize=0, memory=1.0)),        # N < M
])
def test_resolve_matches_library(M, N, kw):
    # src/distpca.cpp:239-352: the oracle's restatement and the library's
    # dpca_resolve (host-only) agree on every branch
    import distpca_amd as D
    fields = dict(kw)
    ram = _ram_gb()
    ref = O.resolve(O.Params(**fields), M, N, ram_gb=ram)
    got = D.resolve(M, N, D.Params(**fields))
    assert (got["bsize"], got["nsv"], got["nrhs"]) == (ref.bsize, ref.nsv, ref.nrhs), (got, ref)


def test_resolve_auto_bsize_closed_form():
    # the auto branch by hand at 100k x 1M, -memory 2 (GB): (2e9 - membuff) / (8 M)
    M, N, nsv, nrhs = 100000, 1000000, 20, 40
    membuff = 3 * N * 8 + M * nsv * 8 + 2 * N * nrhs + 3 * M * 8 + 2048 * 100000
    assert O.resolve(O.Params(nsv=nsv, nrhs=nrhs, memory=2.0), M, N).bsize == \
        int((2e9 - membuff) / (8 * M)) == 2091
    # non-positive available memory: (membuff + 8 M) / 1e7
    assert O.resolve(O.Params(nsv=nsv, nrhs=nrhs, memory=0.05), M, N).bsize == \
        int((membuff + 8 * M) / 1e7)
