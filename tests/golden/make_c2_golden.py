"""Writes tests/golden/c2_golden.npz: the oracle's run of the 20,000-sample x
500,000-SNP drop-in workload (5 populations, Fst 0.01, 1% missing, seed 1;
nsv 10, nrhs 20, power 4, bsize 16384, reference defaults otherwise: crit 1,
tol 1e-3, miter 100) on the bytes of the library's host generator, plus the
same run at bsize 50,000 (summation-order spread).  CPU only; ~1 hour."""
import hashlib
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle import oracle as O  # noqa: E402
import distpca_amd as D  # noqa: E402

if __name__ == "__main__":
    d = tempfile.mkdtemp(dir=os.environ.get("DPCA_TMP"))
    pre = os.path.join(d, "c2")
    D.write_synthetic(pre, D.gen_spec(20000, 500000, npop=5, fst=0.01, missing=0.01, seed=1))
    M, N, iids, snps, packed = O.read_plink(pre)
    sha = np.frombuffer(hashlib.sha256(packed.tobytes()).digest(), np.uint8)
    out = {"bed_sha": sha}
    for tag, b in (("", 16384), ("alt_", 50000)):
        p = O.resolve(O.Params(nsv=10, nrhs=20, power=4, bsize=b), M, N)
        r = O.rsi(packed, M, N, p, verbose=True)
        out[tag + "sigma"] = r.sigma
        out[tag + "rhs"] = r.rhs
        out[tag + "crit"] = np.array(r.crit_values)
        out[tag + "total_its"] = r.total_its
        print(tag, "its", r.total_its, r.sigma[:4], flush=True)
    np.savez_compressed(os.path.join(HERE, "c2_golden.npz"), **out)
    os.remove(pre + ".bed")
