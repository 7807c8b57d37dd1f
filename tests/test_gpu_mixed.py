"""GPU checks of the mixed-precision form (exact int8 codes x three int8
slices of the FP64 operand on int8 MFMA, int32 accumulation, FP64
reductions and small dense algebra) against the FP64 oracle.

Stated tolerances: one pass within 2e-5 of the pass norm; for the principal
components whose Ritz values have converged in the FP64 run (the population
structure), eigenvalue relative error <= 1e-5 and principal angles <= 1e-4
rad.  Components inside the noise bulk are reported, not asserted: neither
form converges them, and any perturbation rotates them freely.
"""
import numpy as np
import pytest

import distpca_amd as D
from oracle import oracle as O
from test_gpu_fp64 import _packed_random

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("M,N,k", [(1001, 3000, 20), (942, 5000, 40), (300, 1000, 7),
                                   (2000, 20000, 40)])
def test_mixed_pass(M, N, k):
    print(f"[gpu] mixed pass M={M} N={N} k={k}", flush=True)
    packed = _packed_random(M, N, seed=k + 1)
    Q = np.linalg.qr(np.random.default_rng(k).standard_normal((M, k)))[0]
    ref = O.OOCPass(packed, M, N, 512)(Q)
    p = D.Params(nsv=min(k, 5), nrhs=k, bsize=512, precision=D.MIXED)
    with D.PCA.from_packed(packed, M, N, p) as pca:
        got = pca.pass_(Q)
        got2 = pca.pass_(Q)
    err = np.abs(got - ref).max() / np.abs(ref).max()
    print("mixed pass rel err", err)
    assert err < 2e-5, err
    assert np.array_equal(got, got2)  # fixed reduction order: bitwise repeatable


# one pass against the oracle: 3 slices within 2e-5 of the pass norm, 2
# slices (quantisation error 2^7 larger) within 5e-4
PASS_TOL = {3: 2e-5, 2: 5e-4}


@pytest.mark.parametrize("slices,k", [(3, k) for k in (1, 5, 10, 16, 22, 26, 37, 40, 44, 48, 53,
                                                        58, 64)] +
                         [(2, k) for k in (1, 8, 9, 16, 20, 24, 31, 40, 47, 56, 64)])
def test_mixed_pass_column_tile_shapes(slices, k):
    # every packing of the slices into column tiles.  3 slices: whole tiles
    # only (16, 48, 64; 44 padded to 48), one shared remainder tile (1, 5,
    # 37, 53) and two (10, 26, 58), two with the second holding only slice 2,
    # which the mask product of Z Q leaves out (22, 40).  2 slices: whole
    # tiles (16, 64; 9, 31, 47 padded), one shared remainder tile (1, 8, 20,
    # 24, 40, 56)
    M, N = 300, 2000
    print(f"[gpu] mixed pass tile shape slices={slices} k={k}", flush=True)
    packed = _packed_random(M, N, seed=100 + k)
    Q = np.linalg.qr(np.random.default_rng(k).standard_normal((M, k)))[0]
    ref = O.OOCPass(packed, M, N, 512)(Q)
    p = D.Params(nsv=min(k, 2), nrhs=k, bsize=512, precision=D.MIXED, slices=slices)
    with D.PCA.from_packed(packed, M, N, p) as pca:
        got = pca.pass_(Q)
        got2 = pca.pass_(Q)
    err = np.abs(got - ref).max() / np.abs(ref).max()
    print("mixed pass rel err", slices, k, err)
    assert err < PASS_TOL[slices], (slices, k, err)
    assert np.array_equal(got, got2)


@pytest.mark.parametrize("slices", [2, 3])
@pytest.mark.parametrize("zt,za", [(2, 1), (4, 2), (8, 4)])
def test_mixed_pass_tiles_per_wave(slices, zt, za, monkeypatch):
    # every instantiated tiles-per-wave form of both int8 products (the
    # launch plan's choice is occupancy-driven; the knobs force the others)
    # gives the same pass within the tolerance; forms whose accumulators do
    # not fit are skipped by the library (the plan keeps its own choice)
    M, N, k = 1500, 4096, 20
    print(f"[gpu] mixed pass slices={slices} Z^T T tiles {zt} Z Q tiles {za}", flush=True)
    packed = _packed_random(M, N, seed=7)
    Q = np.linalg.qr(np.random.default_rng(7).standard_normal((M, k)))[0]
    ref = O.OOCPass(packed, M, N, 512)(Q)
    monkeypatch.setenv("DPCA_ZT_TILES", str(zt))
    monkeypatch.setenv("DPCA_ZA_TILES", str(za))
    p = D.Params(nsv=4, nrhs=k, bsize=512, precision=D.MIXED, slices=slices)
    with D.PCA.from_packed(packed, M, N, p) as pca:
        got = pca.pass_(Q)
    err = np.abs(got - ref).max() / np.abs(ref).max()
    print("mixed pass rel err", err)
    assert err < PASS_TOL[slices], err


@pytest.mark.parametrize("slices,k,zt", [(3, 40, 0), (3, 20, 2), (3, 20, 4), (2, 40, 0),
                                         (2, 20, 2), (3, 33, 0), (3, 64, 0)])
def test_zt_image_bitwise(slices, k, zt, monkeypatch):
    # Z^T T from the matrix's tile-order copy (by DMA) against Z^T T staging
    # the codes itself: the same integer sums, so passes and whole runs are
    # bitwise equal; sample tiles of 64 and 128 bytes (zt = 2, 4 sample tiles
    # per wave; 0 = the launch plan's choice)
    M, N = 1500, 6000
    print(f"[gpu] zt image slices={slices} k={k} zt={zt}", flush=True)
    if zt:
        monkeypatch.setenv("DPCA_ZT_TILES", str(zt))
    packed = _packed_random(M, N, seed=11)
    Q = np.linalg.qr(np.random.default_rng(11).standard_normal((M, k)))[0]
    out = {}
    for img in (0, 1):
        p = D.Params(nsv=4, nrhs=k, power=2, miter=3, bsize=512, precision=D.MIXED,
                     slices=slices, zt_image=img)
        with D.PCA.from_packed(packed, M, N, p) as pca:
            a = pca.pass_(Q)
            b = pca.pass_(a / np.linalg.norm(a, axis=0))
            sigma, rhs, rep = pca.run()
        assert rep["zt_image"] == img
        out[img] = (a, b, sigma, rhs)
    for x, y in zip(out[0], out[1]):
        assert np.array_equal(x.view(np.int64), y.view(np.int64))


def test_zt_image_not_used(monkeypatch):
    # no copy for a streamed matrix, a device_cap that holds one copy only,
    # the FP64 form or 4-wave workgroups; one where the cap holds both
    M, N = 1000, 8192
    print("[gpu] zt image: when it is not used", flush=True)
    packed = _packed_random(M, N, seed=12)
    nbytes = 8192 * 256  # Npad x ldp
    cases = [(dict(precision=D.MIXED, device_cap=nbytes // 2), None, 0),
             (dict(precision=D.MIXED, device_cap=nbytes + (nbytes >> 1)), None, 0),
             (dict(precision=D.MIXED, device_cap=2 * nbytes), None, 1),
             (dict(precision=D.FP64), None, 0),
             (dict(precision=D.MIXED), "4", 0)]
    for kw, waves, used in cases:
        if waves:
            monkeypatch.setenv("DPCA_I8_WAVES", waves)
        p = D.Params(nsv=2, nrhs=8, miter=1, bsize=1024, **kw)
        with D.PCA.from_packed(packed, M, N, p) as pca:
            _, _, rep = pca.run()
        assert rep["zt_image"] == used, (kw, waves, rep["zt_image"])


@pytest.mark.parametrize("N", [1536, 2560, 3584])
def test_three_buffer_short_ranges(N, capfd, monkeypatch):
    # the three-buffer forms of both int8 products (chunk i+2's copies in
    # flight, the last chunk copied again near its range's end) where SNP
    # splits and sample segments hold only a few chunks: against the oracle,
    # and Z^T T from the tile-order copy bitwise against the staging kernel.
    # On an MI355X the plan takes 6 splits of 384 / 512 / 768 SNPs (3 / 4 / 6
    # chunks of 128; last splits of 1 and 2 chunks) and sample segments of
    # 256 / 384 / 512 packed bytes (4 / 6 / 8 chunks of 64; a last one of 2)
    M, k = 100000, 40
    print(f"[gpu] three-buffer forms, short ranges, M={M} N={N}", flush=True)
    packed = _packed_random(M, N, seed=N)
    Q = np.linalg.qr(np.random.default_rng(N).standard_normal((M, k)))[0]
    ref = O.OOCPass(packed, M, N, 1024)(Q)
    monkeypatch.setenv("DPCA_SHOW_PLAN", "1")
    capfd.readouterr()
    out = {}
    for img in (0, 1):
        p = D.Params(nsv=4, nrhs=k, bsize=1024, precision=D.MIXED, zt_image=img)
        with D.PCA.from_packed(packed, M, N, p) as pca:
            out[img] = pca.pass_(Q)
    plans = [l for l in capfd.readouterr().err.splitlines() if l.startswith("dpca plan:")]
    print(plans[-1])
    err = np.abs(out[1] - ref).max() / np.abs(ref).max()
    print("pass rel err", err)
    assert err < 2e-5, err
    assert np.array_equal(out[0].view(np.int64), out[1].view(np.int64))


def test_cli_ztimage_flag(tmp_path):
    # DistPCA.exe -precision mixed with and without the tile-order copy
    # (-ztimage 0) writes the same files, byte for byte
    import subprocess
    pre = str(tmp_path / "zti")
    D.write_synthetic(pre, D.gen_spec(800, 6000, npop=3, fst=0.05, missing=0.01, seed=9))
    print("[gpu] CLI -ztimage", flush=True)
    outs = {}
    for tag, extra in (("img", []), ("noimg", ["-ztimage", "0"])):
        res = subprocess.run([D.EXE_PATH, "-bfile", pre, "-nsv", "4", "-nrhs", "8", "-power", "2",
                              "-bsize", "1000", "-fwrite", "1", "-precision", "mixed",
                              "-prefix", str(tmp_path / tag)] + extra,
                             capture_output=True, text=True, timeout=300)
        assert res.returncode == 0, res.stdout + res.stderr
        outs[tag] = [open(str(tmp_path / (tag + f))).read()
                     for f in ("_singularValues.txt", "_leftSingularVectors.txt")]
    assert outs["img"] == outs["noimg"]


def test_mixed_c1(c1_golden):
    print("[gpu] mixed c1 run", flush=True)
    spec = D.gen_spec(1024, 50000, npop=3, fst=0.05, missing=0.01, seed=0)
    p = D.Params(nsv=10, nrhs=20, power=4, crit=1, tol=1e-6, miter=20, bsize=4096,
                 precision=D.MIXED)
    with D.PCA.synthetic(spec, p) as pca:
        sigma, rhs, rep = pca.run()
    ev0 = c1_golden["sigma"] ** 2
    rel = np.abs(sigma ** 2 - ev0) / ev0
    print("mixed c1 eig rel err per PC", rel)
    print("mixed c1 angles top2", O.principal_angles(rhs[:, :2], c1_golden["rhs"][:, :2]))
    assert rel[:2].max() <= 1e-5
    assert O.principal_angles(rhs[:, :2], c1_golden["rhs"][:, :2]).max() <= 1e-4


def test_mixed_hgdp(hgdp_prefix, hgdp_golden):
    print("[gpu] mixed HGDP run", flush=True)
    p = D.Params(nsv=20, nrhs=40, crit=2, tol=1e-3, bsize=100, precision=D.MIXED)
    with D.PCA.from_bfile(hgdp_prefix, p) as pca:
        sigma, rhs, rep = pca.run()
    ev0 = hgdp_golden["sigma"] ** 2
    rel = np.abs(sigma ** 2 - ev0) / ev0
    print("mixed HGDP its", rep["total_its"], "eig rel err per PC", rel)
    ang = [O.principal_angles(rhs[:, :j], hgdp_golden["rhs"][:, :j]).max() for j in (1, 2, 4, 8, 20)]
    print("mixed HGDP angles top 1/2/4/8/20", ang)
    assert rel[:4].max() <= 1e-5
    assert ang[2] <= 1e-4
