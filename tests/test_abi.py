"""CPU-side checks of the C ABI: the library loads, exports what
include/dpca.h declares, and its host-only pieces (Omega, the file
generator, defaults, the CLI's -about) behave.  No GPU call is made."""
import hashlib
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import ROOT
import distpca_amd as D
from oracle import oracle as O


def test_exports_match_header():
    hdr = open(os.path.join(ROOT, "include", "dpca.h")).read()
    declared = set(re.findall(r"\b(dpca_[a-z_]+)\s*\(", hdr))
    assert declared, "no declarations parsed"
    L = D.lib()
    for name in declared:
        assert hasattr(L, name), name
    assert declared == set(D.EXPORTS)


def test_params_defaults():
    p = D.Params().c()
    assert (p.nsv, p.nrhs, p.power, p.crit, p.tol, p.bsize, p.miter) == (10, 20, 1, 1, 1e-3, 0, 100)
    assert (p.verbose, p.fwrite, p.fullsvd) == (1, 0, 0)
    assert p.prefix == b"noprefix"


@pytest.mark.parametrize("M,k", [(5, 3), (942, 40), (1024, 20), (50000, 20), (100000, 40),
                                 (400000, 40)])
def test_library_omega_is_reference_exact(M, k):
    a = D.omega(M, k)
    b = O.omega(M, k)
    assert np.array_equal(a.view(np.int64), b.view(np.int64))


def test_host_generator_files(tmp_path, c1_golden):
    pre = str(tmp_path / "c1")
    D.write_synthetic(pre, D.gen_spec(1024, 50000, npop=3, fst=0.05, missing=0.01, seed=0))
    assert open(pre + ".bed", "rb").read(3) == b"\x6c\x1b\x01"
    M, N, iids, snps, packed = O.read_plink(pre)
    assert (M, N) == (1024, 50000)
    sha = np.frombuffer(hashlib.sha256(packed.tobytes()).digest(), np.uint8)
    assert np.array_equal(sha, c1_golden["bed_sha"])
    s, c, a, d, z = O.snp_stats(packed, N, M)
    miss = 1 - c.sum() / (N * M)
    assert abs(miss - 0.01) < 1e-3
    # population structure: per-population allele frequencies differ (Fst ~ 0.05)
    dec = np.stack([O.decode(packed[i * 256:(i + 1) * 256]).reshape(-1)[:M] for i in range(1)])
    assert dec.shape[1] == M


def test_cli_about_runs_without_gpu():
    exe = D.EXE_PATH
    out = subprocess.run([exe, "-about"], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0 and "DistPCA" in out.stdout


def _cli(args):
    return subprocess.run([D.EXE_PATH] + args, capture_output=True, text=True, timeout=60)


@pytest.fixture(scope="module")
def tiny_bfile(tmp_path_factory):
    pre = str(tmp_path_factory.mktemp("cli") / "tiny")
    D.write_synthetic(pre, D.gen_spec(40, 300, npop=2, fst=0.05, missing=0.0, seed=3))
    return pre


def test_cli_missing_files(tmp_path, tiny_bfile):
    # src/distpca.cpp:105-125: the .bed is opened first, then the .fam, .bim
    out = _cli(["-bfile", str(tmp_path / "none")])
    assert out.returncode == 1
    assert f"[Data::read_bed] Error reading file {tmp_path / 'none'}.bed" in out.stdout
    pre = str(tmp_path / "nofam")
    os.symlink(tiny_bfile + ".bed", pre + ".bed")
    out = _cli(["-bfile", pre])
    assert out.returncode == 1
    assert f"Error reading file '{pre}.fam': No such file or directory" in out.stdout
    os.symlink(tiny_bfile + ".fam", pre + ".fam")
    out = _cli(["-bfile", pre])
    assert out.returncode == 1
    assert f"Error reading file '{pre}.bim': No such file or directory" in out.stdout


@pytest.mark.parametrize("args,msg", [
    (["-nsv", "0"], "NSV was either zero or negative. Aborting..."),
    (["-nsv", "-3"], "NSV was either zero or negative. Aborting..."),
    (["-crit", "3"], "crit must be 0, 1 or 2"),
    (["-precision", "half"], "Unknown -precision 'half' (fp64 | mixed)"),
    (["-hostmem", "disk"], "Unknown -hostmem 'disk' (pinned | pageable | file)"),
])
def test_cli_parameter_errors_before_any_device_call(tiny_bfile, args, msg):
    # parameters are resolved before the device is touched, so these fail
    # the same way with or without a GPU
    out = _cli(["-bfile", tiny_bfile] + args)
    assert out.returncode == 1
    assert msg in out.stdout, out.stdout


def test_footprint_without_device():
    # the HBM footprint of a context is host arithmetic (no device call): the
    # resident 100k x 1M matrix plus its work buffers (the mixed form by
    # default also the matrix's tile-order copy for Z^T T); above device_cap
    # the two block buffers and the resident blocks stay within the cap
    M, N = 100000, 1000000
    packed = -(-N // 1024) * 1024 * (-(-((M + 3) // 4) // 256) * 256)
    mixed = D.footprint(M, N, D.Params(nsv=20, nrhs=40, power=6, precision=D.MIXED, zt_image=0))
    mixed_img = D.footprint(M, N, D.Params(nsv=20, nrhs=40, power=6, precision=D.MIXED))
    fp64 = D.footprint(M, N, D.Params(nsv=20, nrhs=40, power=6, precision=D.FP64))
    assert packed < mixed < packed + (3 << 30)
    assert mixed_img == mixed + packed
    assert packed < fp64 < packed + (3 << 30)
    # a device_cap that holds one copy but not two: no image
    one = D.footprint(M, N, D.Params(nsv=20, nrhs=40, power=6, precision=D.MIXED,
                                     device_cap=packed + (1 << 30)))
    assert one == mixed
    cap = 2 * 10**9
    streamed = D.footprint(M, N, D.Params(nsv=20, nrhs=40, power=6, precision=D.MIXED,
                                          device_cap=cap))
    work = mixed - packed
    assert streamed <= cap + work + (1 << 30)
    assert streamed < mixed
    # more SNPs, more bytes; the same matrix in either form holds the same packed bytes
    assert D.footprint(M, 2 * N, D.Params(nsv=20, nrhs=40)) > D.footprint(M, N, D.Params(nsv=20, nrhs=40))
