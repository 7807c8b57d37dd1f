"""Writes the golden fixtures under tests/golden/ (run once, in a container
that has the reference tree; the outputs are committed).

  hgdp/hgdp.qc.pruned.bed.gz, hgdp/hgdp.qc.pruned.fam
      the reference's own example genotypes (example/ of the reference),
      942 samples x 133,594 SNPs; the .bim is regenerated by the tests (only
      its line count is read).
  hgdp_golden.npz, c1_golden.npz also hold alt_* results: the same oracle run
      with bsize 1000, i.e. only the summation order changed; the GPU parity
      tests measure their tolerance for unconverged components against that
      spread.
  hgdp_golden.npz
      oracle run of the reference's example settings (example/example.sh:5-12:
      nsv 20, nrhs 40, crit 2, tol 1e-3, bsize 100, power 1): sigma, Ritz
      vectors, criterion trajectory, iteration count (11, as the reference's
      published log example/hgdp.qc.pruned.log:40 implies), per-SNP stats.
  c1_golden.npz
      oracle run of the 1,024 x 50,000 synthetic workload (3 populations,
      Fst 0.05, 1% missing, seed 0; nsv 10, nrhs 20, power 4, crit 1,
      tol 1e-6, miter 20, bsize 4096) on the bytes of the library's host
      generator.
"""
import gzip
import os
import shutil
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
from oracle import oracle as O  # noqa: E402
import distpca_amd as D  # noqa: E402

REF_EXAMPLE = "/root/reference/example/hgdp.qc.pruned"


def hgdp():
    os.makedirs(os.path.join(HERE, "hgdp"), exist_ok=True)
    dst = os.path.join(HERE, "hgdp", "hgdp.qc.pruned")
    with open(REF_EXAMPLE + ".bed", "rb") as f, gzip.open(dst + ".bed.gz", "wb", 9) as g:
        shutil.copyfileobj(f, g)
    shutil.copy(REF_EXAMPLE + ".fam", dst + ".fam")
    M, N, iids, snps, packed = O.read_plink(REF_EXAMPLE)
    p = O.resolve(O.Params(nsv=20, nrhs=40, crit=2, tol=1e-3, bsize=100), M, N)
    r = O.rsi(packed, M, N, p)
    s, c, a, d, z = O.snp_stats(packed, N, M)
    # the same run with another block size: how far the reference's own result
    # moves when only its summation order changes
    pa = O.resolve(O.Params(nsv=20, nrhs=40, crit=2, tol=1e-3, bsize=1000), M, N)
    ra = O.rsi(packed, M, N, pa)
    np.savez_compressed(os.path.join(HERE, "hgdp_golden.npz"), sigma=r.sigma, rhs=r.rhs,
                        crit=np.array(r.crit_values), total_its=r.total_its,
                        passes=r.passes, sum=s, ctr=c, zval=z, alt_sigma=ra.sigma,
                        alt_rhs=ra.rhs, alt_crit=np.array(ra.crit_values))
    print("hgdp", r.total_its, r.sigma[:4])


def c1():
    spec = D.gen_spec(1024, 50000, npop=3, fst=0.05, missing=0.01, seed=0)
    d = tempfile.mkdtemp()
    pre = os.path.join(d, "c1")
    D.write_synthetic(pre, spec)
    M, N, iids, snps, packed = O.read_plink(pre)
    p = O.resolve(O.Params(nsv=10, nrhs=20, power=4, crit=1, tol=1e-6, miter=20,
                           bsize=4096), M, N)
    r = O.rsi(packed, M, N, p)
    s, c, a, dd, z = O.snp_stats(packed, N, M)
    pa = O.resolve(O.Params(nsv=10, nrhs=20, power=4, crit=1, tol=1e-6, miter=20,
                            bsize=1000), M, N)
    ra = O.rsi(packed, M, N, pa)
    np.savez_compressed(os.path.join(HERE, "c1_golden.npz"), sigma=r.sigma, rhs=r.rhs,
                        alt_sigma=ra.sigma, alt_rhs=ra.rhs,
                        crit=np.array(r.crit_values), total_its=r.total_its,
                        passes=r.passes, sum=s, ctr=c, zval=z,
                        bed_sha=np.frombuffer(__import__("hashlib").sha256(
                            packed.tobytes()).digest(), np.uint8))
    shutil.rmtree(d)
    print("c1", r.total_its, r.sigma[:4])


if __name__ == "__main__":
    hgdp()
    c1()
