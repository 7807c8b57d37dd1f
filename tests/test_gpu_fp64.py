"""GPU parity of the plain FP64 form against the oracle, through the C ABI.

Tolerances (FP64 form vs the reference's CPU arithmetic, restated by the
oracle):  decode and per-SNP statistics bit-exact; one pass within 1e-12 of
the pass norm.  For whole runs: eigenvalues (sigma^2) within 1e-10 relative
and principal angles of every top-j subspace within 1e-8 rad -- or, for the
components the iteration has not converged (the noise bulk), within 20x of
how far the oracle itself moves when only its summation order changes
(the alt_* runs of the golden fixtures, bsize 1000 instead of 100/4096).
"""
import os
import subprocess

import numpy as np
import pytest

import distpca_amd as D
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _banner(msg):
    print(f"[gpu] {msg}", flush=True)


def _packed_random(M, N, seed, miss=0.02):
    rng = np.random.default_rng(seed)
    codes = rng.choice([0, 1, 2, 3], size=(N, M), p=[0.4, miss, 0.35, 0.25 - miss]).astype(np.uint8)
    codes[min(3, N - 1)] = 0   # monomorphic
    codes[min(5, N - 1)] = 1   # all missing
    np_ = (M + 3) // 4
    pad = np.zeros((N, 4 * np_), np.uint8)
    pad[:, :M] = codes
    packed = pad[:, 0::4] | (pad[:, 1::4] << 2) | (pad[:, 2::4] << 4) | (pad[:, 3::4] << 6)
    return np.ascontiguousarray(packed.astype(np.uint8))


def test_stats_and_decode_bitexact_hgdp(hgdp_prefix, hgdp_golden):
    _banner("stats/decode HGDP 942 x 133594")
    with D.PCA.from_bfile(hgdp_prefix, D.Params(nsv=20, nrhs=40, bsize=100)) as pca:
        s, c, z = pca.snp_stats()
        assert np.array_equal(s, hgdp_golden["sum"])
        assert np.array_equal(c, hgdp_golden["ctr"])
        assert np.array_equal(z.view(np.int64), hgdp_golden["zval"].view(np.int64))
        M, N, _, _, packed = O.read_plink(hgdp_prefix)
        rows = pca.decode(1000, 64)
        np_ = (M + 3) // 4
        for r in range(64):
            ref = O.decode(packed[(1000 + r) * np_:(1001 + r) * np_])[:M]
            assert np.array_equal(rows[r], ref)
        back = pca.packed(0, 300)
        assert np.array_equal(back.reshape(-1), packed[:300 * np_])


@pytest.mark.parametrize("M,N", [(1001, 777), (5, 300), (256, 256), (1024, 50000)])
def test_stats_edge_shapes(M, N):
    _banner(f"stats edge M={M} N={N}")
    packed = _packed_random(M, N, seed=M + N)
    s0, c0, a0, d0, z0 = O.snp_stats(packed, N, M)
    with D.PCA.from_packed(packed, M, N, D.Params(nsv=2, nrhs=4, bsize=64)) as pca:
        s, c, z = pca.snp_stats()
    assert np.array_equal(s, s0) and np.array_equal(c, c0)
    assert np.array_equal(z.view(np.int64), z0.view(np.int64))


def test_stats_every_count_pair():
    # zval depends only on (sum, ctr): cover every pair reachable at M = 64
    M = 64
    rows = []
    for ctr in range(0, M + 1):
        for sm in range(0, 2 * ctr + 1):
            n2 = max(0, sm - ctr)
            n1 = sm - 2 * n2
            n0 = ctr - n1 - n2
            codes = [0] * n0 + [2] * n1 + [3] * n2 + [1] * (M - ctr)
            rows.append(codes)
    codes = np.array(rows, np.uint8)
    N = codes.shape[0]
    packed = (codes[:, 0::4] | (codes[:, 1::4] << 2) | (codes[:, 2::4] << 4) |
              (codes[:, 3::4] << 6)).astype(np.uint8)
    s0, c0, a0, d0, z0 = O.snp_stats(packed, N, M)
    with D.PCA.from_packed(packed, M, N, D.Params(nsv=2, nrhs=4, bsize=64)) as pca:
        s, c, z = pca.snp_stats()
    assert np.array_equal(z.view(np.int64), z0.view(np.int64))


@pytest.mark.parametrize("M,N,k", [(1001, 3000, 20), (942, 5000, 40), (300, 1000, 7)])
def test_pass_matches_oracle(M, N, k):
    _banner(f"pass M={M} N={N} k={k}")
    packed = _packed_random(M, N, seed=k)
    Q = np.random.default_rng(k).standard_normal((M, k))
    ref = O.OOCPass(packed, M, N, 512)(Q)
    with D.PCA.from_packed(packed, M, N, D.Params(nsv=min(k, 5), nrhs=k, bsize=512)) as pca:
        got = pca.pass_(Q)
    err = np.abs(got - ref).max() / np.abs(ref).max()
    assert err < 1e-12, err


@pytest.mark.parametrize("k", [1, 4, 5, 12, 13, 16, 17, 28, 29, 33, 44, 45, 52, 60, 64])
def test_pass_column_groups(k):
    # every split of k columns into whole 16-column tiles and 4-column groups
    # (k mod 16 in 1..12 -> ceil(k mod 16 / 4) groups, 13..15 -> a tile)
    M, N = 300, 1500
    _banner(f"pass column groups k={k}")
    packed = _packed_random(M, N, seed=100 + k)
    Q = np.random.default_rng(100 + k).standard_normal((M, k))
    ref = O.OOCPass(packed, M, N, 512)(Q)
    with D.PCA.from_packed(packed, M, N, D.Params(nsv=1, nrhs=k, bsize=512)) as pca:
        got = pca.pass_(Q)
    assert got.shape == (M, k)
    err = np.abs(got - ref).max() / np.abs(ref).max()
    assert err < 1e-12, err


@pytest.mark.parametrize("k", [40, 64])
def test_pass_wide_tiles_streamed(k, capfd, monkeypatch):
    # two or more whole column tiles take the 4-SNP-tile FP64 Z Q where it
    # keeps two workgroups per CU; streamed 1024-row blocks also split the
    # samples into segments (short blocks): both against the oracle.  The
    # launch plan line confirms that the 4-tile form is the one that ran at
    # nrhs 40; at 64 (four whole tiles) its accumulators leave one workgroup
    # per CU, so the 2-tile form runs there.
    M, N = 2000, 6000
    _banner(f"wide-tile pass k={k}, resident and streamed")
    packed = _packed_random(M, N, seed=300 + k)
    Q = np.random.default_rng(300 + k).standard_normal((M, k))
    ref = O.OOCPass(packed, M, N, 512)(Q)
    base = dict(nsv=2, nrhs=k, bsize=512)
    monkeypatch.setenv("DPCA_SHOW_PLAN", "1")
    capfd.readouterr()
    with D.PCA.from_packed(packed, M, N, D.Params(**base)) as res, \
            D.PCA.from_packed(packed, M, N, D.Params(device_cap=1 << 20, **base)) as stm:
        a, b = res.pass_(Q), stm.pass_(Q)
    plans = [l for l in capfd.readouterr().err.splitlines() if l.startswith("dpca plan:")]
    assert len(plans) == 2, plans
    for l in plans:
        assert "fp64 splits" in l and f"Z Q tiles {4 if k == 40 else 2};" in l, l
    for got in (a, b):
        err = np.abs(got - ref).max() / np.abs(ref).max()
        assert err < 1e-12, err


def test_orthonormalize_cholqr2():
    rng = np.random.default_rng(7)
    A = rng.standard_normal((5000, 40)) @ np.diag(np.logspace(0, 4, 40))
    Q, shifted = D.orthonormalize(A)
    assert not shifted
    assert np.abs(Q.T @ Q - np.eye(40)).max() < 1e-13
    R = Q.T @ A
    assert np.abs(np.tril(R, -1)).max() < 1e-9 * np.abs(R).max()
    assert np.all(np.diag(R) > 0)
    # nearly collinear columns (kappa 1e12): the shifted fallback keeps Q
    # orthonormal and spanning the same space
    U = np.linalg.qr(rng.standard_normal((4000, 30)))[0]
    V = np.linalg.qr(rng.standard_normal((30, 30)))[0]
    B = U @ np.diag(np.logspace(0, -12, 30)) @ V.T
    Q2, shifted2 = D.orthonormalize(B)
    assert shifted2
    assert np.abs(Q2.T @ Q2 - np.eye(30)).max() < 1e-12
    # the span is only determined to ~kappa * eps in its weakest directions;
    # Householder QR (numpy) is the yardstick for that loss
    qn, _ = np.linalg.qr(B)
    assert O.principal_angles(U, Q2).max() <= 10 * O.principal_angles(U, qn).max() + 1e-12
    # exactly rank-deficient (rank 25 of 32, dependent columns interleaved):
    # dgeqrf + dorgqr never fail; the library completes the span with
    # orthonormal directions as well
    C = rng.standard_normal((3000, 25))
    mix = rng.standard_normal((25, 7))
    D_ = np.concatenate([C[:, :10], C @ mix, C[:, 10:]], axis=1)
    Q3, _ = D.orthonormalize(D_)
    assert np.abs(Q3.T @ Q3 - np.eye(32)).max() < 1e-12
    assert np.abs(Q3 @ (Q3.T @ D_) - D_).max() < 1e-10 * np.abs(D_).max()


@pytest.mark.parametrize("k", [1, 7, 20, 40, 64])
def test_syev_jacobi(k):
    rng = np.random.default_rng(k)
    S = rng.standard_normal((k, k))
    S = S @ S.T + np.diag(np.arange(k, dtype=float))
    B = np.triu(S) + np.tril(rng.standard_normal((k, k)), -1)  # lower part must be ignored
    w, X = D.syev(B)
    ref = np.linalg.eigvalsh(S)[::-1]
    assert np.allclose(w, ref, rtol=1e-12, atol=1e-12 * ref[0])
    assert np.abs(X.T @ X - np.eye(k)).max() < 1e-12
    assert np.abs(S @ X - X * w).max() < 1e-11 * ref[0]


def check_run(sigma, rhs, g, nsv, fac=20.0):
    """Per-component parity against the golden oracle run (see module doc)."""
    ev, ev_alt = g["sigma"] ** 2, g["alt_sigma"] ** 2
    spread = np.abs(ev_alt - ev) / ev
    rel = np.abs(sigma ** 2 - ev) / ev
    tol_e = np.maximum(1e-10, fac * spread)
    print("eig rel err", rel, "\n  oracle spread", spread)
    assert np.all(rel <= tol_e), (rel, tol_e)
    for j in range(1, nsv + 1):
        a = O.principal_angles(rhs[:, :j], g["rhs"][:, :j]).max()
        sp = O.principal_angles(g["alt_rhs"][:, :j], g["rhs"][:, :j]).max()
        assert a <= max(1e-8, fac * sp), (j, a, sp)
    return rel


def test_run_hgdp_matches_reference_settings(hgdp_prefix, hgdp_golden):
    # example/example.sh:5-12 settings; the published log implies 11 iterations
    _banner("run HGDP nsv 20 nrhs 40 crit 2 tol 1e-3 bsize 100")
    p = D.Params(nsv=20, nrhs=40, crit=2, tol=1e-3, bsize=100)
    with D.PCA.from_bfile(hgdp_prefix, p) as pca:
        sigma, rhs, rep = pca.run()
        crit = pca.crit_history()
    assert rep["total_its"] == 11 and rep["passes"] == 22
    # the MEV criterion follows the noise components; the oracle's own
    # trajectory moves by ~1.4e-4 relative when only bsize changes
    assert np.allclose(crit, hgdp_golden["crit"], rtol=2e-3)
    rel = check_run(sigma, rhs, hgdp_golden, 20)
    assert rel[:10].max() < 1e-10


def test_run_c1_synthetic(c1_golden):
    _banner("run c1 1024 x 50000 on device-generated data")
    spec = D.gen_spec(1024, 50000, npop=3, fst=0.05, missing=0.01, seed=0)
    p = D.Params(nsv=10, nrhs=20, power=4, crit=1, tol=1e-6, miter=20, bsize=4096)
    with D.PCA.synthetic(spec, p) as pca:
        import hashlib
        sha = np.frombuffer(hashlib.sha256(pca.packed().tobytes()).digest(), np.uint8)
        assert np.array_equal(sha, c1_golden["bed_sha"]), "device generator != host generator"
        sigma, rhs, rep = pca.run()
    assert rep["total_its"] == int(c1_golden["total_its"])
    # PCs 1-2 carry the 3-population structure and converge; PCs 3-10 sit in
    # the noise bulk and are not converged after 20 iterations
    rel = check_run(sigma, rhs, c1_golden, 10)
    assert rel[:2].max() < 1e-10
    assert O.principal_angles(rhs[:, :2], c1_golden["rhs"][:, :2]).max() < 1e-8


@pytest.mark.parametrize("incore", [False, True])
def test_crit0_trace_criterion(tmp_path, incore):
    # crit 0: |delta trace| / trace <= tol (src/methods.cpp:1025-1031); the
    # oracle's trajectory crosses tol 3e-4 between 3.30e-4 and 2.66e-4, so
    # both runs stop after 12 iterations
    _banner(f"crit 0 run, incore {incore}")
    pre = str(tmp_path / "crit0")
    D.write_synthetic(pre, D.gen_spec(600, 8000, npop=3, fst=0.05, missing=0.01, seed=21))
    M, N, _, _, packed = O.read_plink(pre)
    kw = dict(nsv=5, nrhs=10, power=2, crit=0, tol=3e-4, miter=40, bsize=N if incore else 1000)
    p = O.resolve(O.Params(**kw), M, N)
    ref = O.rsi_incore(O.dense_z(packed, N, M), p) if incore else O.rsi(packed, M, N, p)
    with D.PCA.from_bfile(pre, D.Params(**kw)) as pca:
        sigma, rhs, rep = pca.run()
        crit = pca.crit_history()
    assert rep["total_its"] == ref.total_its == 12 and rep["incore"] == int(incore)
    # |delta trace| / trace differences two traces that agree to ~1e-3, so
    # the summation-order differences of the noise components (~1e-12) show
    # up ~1e3 larger; the stop itself has a 10% margin
    dev = np.abs(np.asarray(crit) - ref.crit_values) / np.asarray(ref.crit_values)
    print("crit 0 trajectory max rel dev", dev.max(), flush=True)
    assert dev.max() < 1e-6
    assert np.allclose(sigma, ref.sigma, rtol=1e-10)
    assert O.principal_angles(rhs[:, :2], ref.rhs[:, :2]).max() < 1e-8


@pytest.mark.parametrize("precision", [D.FP64, D.MIXED])
def test_pass_saving_schedule(c1_golden, precision):
    # A A^T (Q X) = (A A^T Q) X: from the second iteration on the first pass
    # is the previous last pass rotated by the Ritz vectors.  Same mathematics
    # as the reference's schedule, so the converged components must agree
    # with the oracle's strict run.
    _banner(f"run c1 with one pass saved per iteration, precision {precision}")
    spec = D.gen_spec(1024, 50000, npop=3, fst=0.05, missing=0.01, seed=0)
    p = D.Params(nsv=10, nrhs=20, power=4, crit=1, tol=1e-6, miter=20, bsize=4096,
                 precision=precision, strict_passes=0)
    with D.PCA.synthetic(spec, p) as pca:
        sigma, rhs, rep = pca.run()
    its = rep["total_its"]
    assert rep["passes"] == 5 + 4 * (its - 1)
    ev = c1_golden["sigma"][:2] ** 2
    tol_e, tol_a = (1e-10, 1e-8) if precision == D.FP64 else (1e-6, 1e-4)
    assert np.abs(sigma[:2] ** 2 - ev).max() / ev.min() < tol_e
    assert O.principal_angles(rhs[:, :2], c1_golden["rhs"][:, :2]).max() < tol_a


def test_incore_variant_matches_oracle():
    M, N = 400, 2000
    packed = _packed_random(M, N, seed=11)
    p = D.Params(nsv=5, nrhs=10, power=2, crit=1, tol=1e-12, miter=60, bsize=N)
    Z = O.dense_z(packed, N, M)
    ref = O.rsi_incore(Z, O.resolve(O.Params(nsv=5, nrhs=10, power=2, crit=1, tol=1e-12,
                                             miter=60, bsize=N), M, N))
    with D.PCA.from_packed(packed, M, N, p) as pca:
        sigma, rhs, rep = pca.run()
    assert rep["incore"] == 1 and rep["total_its"] == ref.total_its
    assert np.allclose(sigma, ref.sigma, rtol=1e-12)


def test_cli_dropin(tmp_path):
    pre = str(tmp_path / "cli")
    spec = D.gen_spec(600, 4000, npop=3, fst=0.05, missing=0.01, seed=5)
    D.write_synthetic(pre, spec)
    out = subprocess.run([D.EXE_PATH, "-bfile", pre, "-nsv", "4", "-nrhs", "8", "-power", "2",
                          "-crit", "2", "-tol", "1e-6", "-bsize", "500", "-fwrite", "1",
                          "-prefix", str(tmp_path / "res"), "-verbose", "2"],
                         capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    lines = out.stdout.splitlines()
    i = lines.index("[The following times are listed in seconds]")
    tline = next(l for l in lines[i:] if l.startswith("Total wall-clock time elapsed:"))
    float(tline.split()[4])  # field 5, as example/example.sh parses it
    sv = np.loadtxt(str(tmp_path / "res_singularValues.txt"))
    M, N, iids, _, packed = O.read_plink(pre)
    r = O.rsi(packed, M, N, O.resolve(O.Params(nsv=4, nrhs=8, power=2, crit=2, tol=1e-6,
                                               bsize=500), M, N))
    ra = O.rsi(packed, M, N, O.resolve(O.Params(nsv=4, nrhs=8, power=2, crit=2, tol=1e-6,
                                                bsize=4000), M, N))
    txt = open(str(tmp_path / "res_leftSingularVectors.txt")).read().splitlines()
    assert txt[0] == "FID\tPC0\tPC1\tPC2\tPC3"
    assert txt[1].split("\t")[0].strip() == iids[0] and len(txt) == M + 1
    U = np.array([[float(x) for x in l.split("\t")[1:]] for l in txt[1:]])
    # the files carry 13 decimals (src/methods.cpp:1093,1114)
    rel = np.abs(sv ** 2 - r.sigma ** 2) / r.sigma ** 2
    spread = np.abs(ra.sigma ** 2 - r.sigma ** 2) / r.sigma ** 2
    assert np.all(rel <= np.maximum(1e-10, 20 * spread)), (rel, spread)
    for j in range(1, 5):
        a = O.principal_angles(U[:, :j], r.rhs[:, :j]).max()
        sp = O.principal_angles(ra.rhs[:, :j], r.rhs[:, :j]).max()
        assert a <= max(1e-8, 20 * sp), (j, a, sp)
    # byte layout: the oracle's writer (pinned on the reference-held 1000
    # Genomes output) rewrites the parsed values to the same bytes
    O.write_outputs(str(tmp_path / "orc"), sv, U, 4, iids)
    for f in ("_singularValues.txt", "_leftSingularVectors.txt"):
        assert open(str(tmp_path / ("res" + f))).read() == open(str(tmp_path / ("orc" + f))).read()


def _fetch_line(stdout):
    key = "Total number of times which the data matrix was fetched from the memory: "
    return int(next(l for l in stdout.splitlines() if l.startswith(key))[len(key):])


@pytest.mark.parametrize("strict", [1, 0])
def test_cli_fetch_count(tmp_path, strict):
    # the console's fetch count: total_its * (power + 1) under the
    # reference's schedule (src/utilities.cpp:160-161); the passes actually
    # made under -strict 0 (power + 1 for the first iteration, power after)
    _banner(f"CLI fetch count, strict {strict}")
    pre = str(tmp_path / "fc")
    D.write_synthetic(pre, D.gen_spec(600, 4000, npop=3, fst=0.05, missing=0.01, seed=5))
    kw = dict(nsv=4, nrhs=8, power=3, crit=1, tol=1e-6, bsize=500, strict_passes=strict)
    out = subprocess.run([D.EXE_PATH, "-bfile", pre, "-nsv", "4", "-nrhs", "8", "-power", "3",
                          "-crit", "1", "-tol", "1e-6", "-bsize", "500", "-strict", str(strict)],
                         capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    with D.PCA.from_bfile(pre, D.Params(**kw)) as pca:
        assert (pca.M, pca.N, pca.nsv, pca.nrhs, pca.bsize) == \
            (600, 4000) + tuple(D.resolve(600, 4000, D.Params(**kw))[f]
                                for f in ("nsv", "nrhs", "bsize"))
        _, _, rep = pca.run()
    its = rep["total_its"]
    assert its > 1
    want = its * 4 if strict else 4 + 3 * (its - 1)
    assert rep["passes"] == want
    assert _fetch_line(out.stdout) == want


@pytest.mark.parametrize("incore", [True, False])
def test_rank_deficient_basis(incore):
    # nrhs = M: centred columns leave A A^T rank M - 1, so every basis the
    # iteration orthonormalises spans at most M - 1 dimensions; the reference's
    # Householder QR completes it with some orthonormal direction.  The
    # library completes it too: the shifted CholeskyQR3 fallback turns the
    # rounding noise of the dependent direction into one, or, where that
    # fallback meets a non-positive pivot (which of the two happens depends
    # on rounding at the 1e-16 level), the run is repeated with the dependent
    # column replaced and the completion reported.  Either way the leading
    # components are those of the oracle.
    _banner(f"rank-deficient basis (nrhs = M), incore {incore}")
    M, N = 30, 3000
    packed = _packed_random(M, N, seed=21)
    p = O.Params(nsv=10, nrhs=30, power=2, crit=1, tol=1e-12, miter=6,
                 bsize=N if incore else 1000)
    ref = O.rsi(packed, M, N, O.resolve(p, M, N))
    with D.PCA.from_packed(packed, M, N, D.Params(**vars(p))) as pca:
        sigma, rhs, rep = pca.run()
        assert np.abs(rhs.T @ rhs - np.eye(30)).max() < 1e-10
        if rep["qr_repaired"] > 0:
            s2, _, rep2 = pca.run()          # robust from the start on this context
            assert rep2["qr_repaired"] > 0 and np.array_equal(s2, sigma)
    assert rep["total_its"] == ref.total_its
    rel = np.abs(sigma[:9] ** 2 - ref.sigma[:9] ** 2) / ref.sigma[:9] ** 2
    assert rel.max() < 1e-10, rel
    assert O.principal_angles(rhs[:, :9], ref.rhs[:, :9]).max() < 1e-8


@pytest.mark.parametrize("flags,msgs,nsv,nrhs", [
    # src/distpca.cpp:272-352 on a 40 x 2000 cohort (min(M, N) = 40)
    (["-nsv", "50", "-nrhs", "0", "-bsize", "5000"],
     ["Warning: Specified block size (5000) exceeds local SNPs (2000). Adjusting to local_N.",
      "The value of NSV given was larger than min(M, N). Adjusting to NSV = min(M,N)...",
      "NRHS was either zero or negative. Adjusting to NRHS = min(2xNSV, min(M,N))..."], 40, 40),
    (["-nsv", "6", "-nrhs", "3"],
     ["NRHS can not be smaller than NSV. Adjusting to NRHS = min(2xNSV,min(M,N))..."], 6, 12),
    (["-nsv", "4", "-nrhs", "60"],
     ["The value of NRHS given was larger than min(M,N). Adjusting to NRHS = min(M,N)..."],
     4, 40),
])
def test_cli_clamp_messages(tmp_path, flags, msgs, nsv, nrhs):
    _banner(f"CLI clamps {flags}")
    pre = str(tmp_path / "clamp")
    D.write_synthetic(pre, D.gen_spec(40, 2000, npop=2, fst=0.05, missing=0.01, seed=2))
    out = subprocess.run([D.EXE_PATH, "-bfile", pre, "-miter", "3"] + flags,
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    for m in msgs:
        assert m in out.stdout, (m, out.stdout)
    for m in ("Estimated memory per rank:", "Block size (in MB):", "Memory per rank:"):
        assert m in out.stdout
    assert f"Total number of singular pairs sought: {nsv}\n" in out.stdout
    assert f"Total number of right-hand sides used: {nrhs}\n" in out.stdout


@pytest.mark.parametrize("precision", [D.FP64, D.MIXED])
def test_streamed_matches_resident(precision):
    # device_cap of 1 MB forces 1024-row blocks streamed from pinned host
    # memory through two HBM buffers (the config-4 path at small scale)
    _banner(f"streamed vs resident, precision {precision}")
    M, N = 1000, 20000
    spec = D.gen_spec(M, N, npop=3, fst=0.05, missing=0.01, seed=9)
    base = dict(nsv=4, nrhs=8, power=2, crit=1, tol=1e-10, miter=8, bsize=2000,
                precision=precision)
    with D.PCA.synthetic(spec, D.Params(**base)) as res, \
            D.PCA.synthetic(spec, D.Params(device_cap=1 << 20, **base)) as stm:
        assert np.array_equal(res.packed(), stm.packed())
        assert np.array_equal(res.decode(777, 5), stm.decode(777, 5))
        s1, c1, z1 = res.snp_stats()
        s2, c2, z2 = stm.snp_stats()
        assert np.array_equal(z1.view(np.int64), z2.view(np.int64))
        Q = np.linalg.qr(np.random.default_rng(1).standard_normal((M, 8)))[0]
        a, b = res.pass_(Q), stm.pass_(Q)
        tol = 1e-12 if precision == D.FP64 else 2e-6
        assert np.abs(a - b).max() / np.abs(a).max() < tol
        if precision == D.FP64:
            ref = O.OOCPass(res.packed().reshape(-1), M, N, 2000)(Q)
            assert np.abs(b - ref).max() / np.abs(ref).max() < 1e-12
        sa, Ua, ra = res.run()
        sb, Ub, rb = stm.run()
    assert rb["streamed"] == 1 and ra["streamed"] == 0
    # the rest of the 1 MB cap keeps the leading 1024-row blocks in HBM after
    # the first pass (the single pass above): the run streams only the others
    ldp, npad = -(-((M + 3) // 4) // 256) * 256, -(-N // 1024) * 1024
    assert rb["resident_bytes"] == 2 * 1024 * ldp
    assert rb["h2d_bytes"] == rb["passes"] * (npad * ldp - rb["resident_bytes"])
    assert ra["total_its"] == rb["total_its"]
    assert np.allclose(sa, sb, rtol=1e-10 if precision == D.FP64 else 1e-6)
    assert O.principal_angles(Ua[:, :2], Ub[:, :2]).max() < (1e-8 if precision == D.FP64 else 1e-4)


@pytest.mark.parametrize("mode", [D.HOST_PAGEABLE, D.HOST_FILE])
def test_streamed_host_modes(tmp_path, mode):
    # the matrix kept pageable (staged through pinned block buffers) or left
    # in the .bed file and re-read every pass (raw rows repacked on the
    # device); both must give the resident context's bytes, statistics and
    # FP64 results
    _banner(f"streamed host mode {mode}")
    M, N = 1001, 12345                     # ragged last byte, partial last block
    pre = str(tmp_path / "hm")
    spec = D.gen_spec(M, N, npop=3, fst=0.05, missing=0.02, seed=13)
    D.write_synthetic(pre, spec)
    base = dict(nsv=3, nrhs=6, power=2, crit=1, tol=1e-10, miter=6, bsize=3000)
    if mode == D.HOST_PAGEABLE:
        # device-generated blocks parked in pageable memory through staging
        with D.PCA.synthetic(spec, D.Params(device_cap=1 << 20, host_mode=mode, **base)) as g, \
                D.PCA.from_bfile(pre, D.Params(**base)) as res:
            assert np.array_equal(g.packed(), res.packed())
    with D.PCA.from_bfile(pre, D.Params(**base)) as res, \
            D.PCA.from_bfile(pre, D.Params(device_cap=1 << 20, host_mode=mode, **base)) as stm:
        s1, c1, z1 = res.snp_stats()
        s2, c2, z2 = stm.snp_stats()
        assert np.array_equal(s1, s2) and np.array_equal(c1, c2)
        assert np.array_equal(z1.view(np.int64), z2.view(np.int64))
        assert np.array_equal(res.packed(), stm.packed())
        assert np.array_equal(res.decode(N - 7, 7), stm.decode(N - 7, 7))
        Q = np.linalg.qr(np.random.default_rng(2).standard_normal((M, 6)))[0]
        a, b = res.pass_(Q), stm.pass_(Q)
        assert np.abs(a - b).max() / np.abs(a).max() < 1e-12
        sa, Ua, ra = res.run()
        sb, Ub, rb = stm.run()
    assert rb["streamed"] == 1 and rb["host_mode"] == mode
    assert ra["total_its"] == rb["total_its"]
    assert np.allclose(sa, sb, rtol=1e-10)
    assert O.principal_angles(Ua[:, :2], Ub[:, :2]).max() < 1e-8
    # against the oracle itself (the reference's MPI-IO block loop,
    # src/methods.cpp:795-811, restated): per-SNP statistics bit-exact, the
    # streamed pass within 1e-12, and the streamed run's iteration count,
    # eigenvalues and the 2 population PCs within the FP64 tolerances (the
    # third PC lies in the noise bulk and is compared with the resident run
    # above)
    M2, N2, _, _, packed = O.read_plink(pre)
    os_, oc_, _, _, oz_ = O.snp_stats(packed, N2, M2)
    assert np.array_equal(s2, os_) and np.array_equal(c2, oc_)
    assert np.array_equal(z2.view(np.int64), oz_.view(np.int64))
    ref_pass = O.OOCPass(packed, M, N, 3000)(Q)
    assert np.abs(b - ref_pass).max() / np.abs(ref_pass).max() < 1e-12
    r = O.rsi(packed, M, N, O.resolve(O.Params(nsv=3, nrhs=6, power=2, crit=1, tol=1e-10,
                                               miter=6, bsize=3000), M, N))
    assert rb["total_its"] == r.total_its and rb["passes"] == r.passes
    assert np.allclose(sb[:2] ** 2, r.sigma[:2] ** 2, rtol=1e-10, atol=0)
    assert O.principal_angles(Ub[:, :2], r.rhs[:, :2]).max() < 1e-8
