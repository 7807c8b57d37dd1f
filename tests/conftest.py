import gzip
import os
import shutil
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")
REF_EXAMPLE = "/root/reference/example/hgdp.qc.pruned"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libdpca on the GPU)")


@pytest.fixture(scope="session")
def hgdp_prefix(tmp_path_factory):
    """The reference's HGDP example as a PLINK triple in a temp dir (the .bed
    from the committed gzip, a .bim regenerated with the right line count)."""
    d = tmp_path_factory.mktemp("hgdp")
    pre = str(d / "hgdp.qc.pruned")
    with gzip.open(os.path.join(GOLDEN, "hgdp", "hgdp.qc.pruned.bed.gz"), "rb") as f, \
            open(pre + ".bed", "wb") as g:
        shutil.copyfileobj(f, g)
    shutil.copy(os.path.join(GOLDEN, "hgdp", "hgdp.qc.pruned.fam"), pre + ".fam")
    N = (os.path.getsize(pre + ".bed") - 3) // ((942 + 3) // 4)
    with open(pre + ".bim", "w") as f:
        for s in range(N):
            f.write(f"1\trs{s}\t0\t{s + 1}\tA\tG\n")
    return pre


@pytest.fixture(scope="session")
def hgdp_golden():
    return dict(np.load(os.path.join(GOLDEN, "hgdp_golden.npz")))


@pytest.fixture(scope="session")
def c1_golden():
    return dict(np.load(os.path.join(GOLDEN, "c1_golden.npz")))
