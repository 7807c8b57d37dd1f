"""GPU checks of the -fullSVD diagnostic (src/distpca.cpp:449-635).

The library's singular values and left singular vectors of the whole
standardised matrix (Gram eigendecomposition on the device) are compared
with numpy's SVD of the oracle's dense matrix: singular values within 1e-11
relative, each vector within 1e-9 rad of numpy's (up to sign).  The
comparison quantities (relative errors, cosine values, ||Uhat^T U - I||_F and
_2) are then checked against the oracle's restatement computed from numpy's
vectors with the device's column signs: within 1e-9 absolute.
"""
import subprocess

import numpy as np
import pytest

import distpca_amd as D
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _check_against_numpy(got, Zt, rhs, sigma, nsv):
    U, s, _ = np.linalg.svd(Zt, full_matrices=False)
    U, s = U[:, :nsv], s[:nsv]
    assert np.abs(got["true_sigma"] - s).max() < 1e-11 * s[0]
    for j in range(nsv):
        c = abs(float(got["true_U"][:, j] @ U[:, j]))
        assert np.arccos(min(1.0, c)) < 1e-9 or np.linalg.norm(
            got["true_U"][:, j] - np.sign(got["true_U"][:, j] @ U[:, j]) * U[:, j]) < 1e-9, j
    sgn = np.sign(np.sum(got["true_U"] * U, axis=0))
    ref = O.full_svd_diag(Zt, np.ascontiguousarray(rhs[:, :nsv]), sigma, U * sgn, s)
    for key in ("sigma_relerr", "cos_values"):
        assert np.abs(got[key] - ref[key]).max() < 1e-9, key
    # vec_relerr divides by |U_ij| + 1e-12: compare the numerators
    den = np.abs(ref["true_U"]) + 1e-12
    assert np.abs((got["vec_relerr"] - ref["vec_relerr"]) * den).max() < 1e-11
    assert abs(got["frob_norm_angle"] - ref["frob_norm_angle"]) < 1e-9
    assert abs(got["cos_error"] - ref["cos_error"]) < 1e-9
    assert got["info"] == 0


@pytest.mark.parametrize("M,N", [(400, 6000), (1500, 900)])
def test_full_svd_library(M, N):
    # samples < SNPs (Gram over samples) and SNPs < samples (Gram over SNPs)
    print(f"[gpu] full SVD M={M} N={N}", flush=True)
    spec = D.gen_spec(M, N, npop=3, fst=0.05, missing=0.02, seed=17)
    with D.PCA.synthetic(spec, D.Params(nsv=5, nrhs=10, power=2, crit=1, tol=1e-10,
                                        miter=40)) as pca:
        sigma, rhs, _ = pca.run()
        got = pca.full_svd(rhs, sigma)
        packed = pca.packed()
    Zt = O.dense_z(packed.reshape(-1), N, M).T
    _check_against_numpy(got, Zt, rhs, sigma, 5)


def test_full_svd_cli(tmp_path):
    print("[gpu] full SVD through DistPCA.exe", flush=True)
    pre = str(tmp_path / "fs")
    M, N = 300, 2000
    D.write_synthetic(pre, D.gen_spec(M, N, npop=3, fst=0.05, missing=0.01, seed=8))
    out = subprocess.run([D.EXE_PATH, "-bfile", pre, "-nsv", "4", "-nrhs", "8", "-power", "2",
                          "-tol", "1e-10", "-miter", "40",
                          "-bsize", str(N), "-fullSVD", "1", "-fwrite", "1", "-verbose", "2",
                          "-prefix", str(tmp_path / "r")], capture_output=True, text=True,
                         timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    lines = out.stdout.splitlines()
    # in-core run: the statistics block carries the full-SVD lines, printed
    # before the SVD is computed (zeros), as the reference does
    assert "||Uhat^TU - I||_F: 0.0000000000000" in lines
    assert any(l.startswith("FULL SVD computed in ") and l.endswith("info = 0") for l in lines)
    assert sum(l.startswith("RelError Sigma(") for l in lines) == 4
    assert sum(l.startswith("Cosine(") for l in lines) == 4
    vec = np.loadtxt(str(tmp_path / "r_realLeftSingularVectors.txt"))
    val = np.loadtxt(str(tmp_path / "r_realSingularValues.txt"))
    acc = np.loadtxt(str(tmp_path / "r_singvecs_accuracy.txt"))
    cos = np.loadtxt(str(tmp_path / "r_cosineValues.txt"))
    assert vec.shape == (M, 4) and val.shape == (4,) and acc.shape == (M * 4,) and cos.shape == (4,)
    _, _, _, _, packed = O.read_plink(pre)
    s = np.linalg.svd(O.dense_z(packed, N, M).T, compute_uv=False)[:4]
    assert np.abs(val - s).max() < 1e-11 * s[0]
    assert np.all(np.abs(np.abs(cos[:2]) - 1) < 1e-9)   # the two population PCs
