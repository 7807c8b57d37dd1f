"""Benchmark of the hot path: one data pass A A^T Q = Z^T (Z Q) of randomized
subspace-iteration PCA over an HBM-resident packed PLINK matrix.

Default workload: 100,000 samples x 1,000,000 SNPs (25 GB packed), 8
Balding-Nichols populations (Fst 0.005) plus 20% Dirichlet-admixed samples,
2% missing, seed 2, nrhs 40, generated directly in HBM.  A step is one pass
in the mixed-precision form; the FP64 form's pass is timed beside it.

  python bench.py [--steps K] [--warmup W] [--workload c3|c2|c1|...] [--no-full]
                  [--slices 2|3] [--cpu-plan [--cpu-passes P]]

Prints one JSON line.  value = packed-genotype GB/s through the mixed pass.
Unless --no-full, the line also carries the wall time of a full PCA run in
each form on the same matrix (and with the pass-saving schedule, and with
two int8 slices instead of three) and the mixed forms' eigenvalue and
principal-angle errors against the FP64 run.  `roofline` rates the slower
int8 product kernel against the HBM roof (its useful work sits below the
int8 ridge), with the MFMA fractions beside; `cpu_baseline` is the oracle
timed on a bounded sample, extrapolated (and labelled so) to a full pass and
run.  --cpu-plan runs only the CPU plan: the oracle's whole run for c1/c2,
one outer iteration for c3.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# run flags beyond nsv/nrhs/power/bsize: the reference's defaults (crit 1,
# tol 1e-3, miter 100, src/utilities.cpp:215-217) unless a workload names its
# own
REF_RUN = dict(crit=1, tol=1e-3, miter=100)
WORKLOADS = {
    # name: (M, N, npop, fst, admix, missing, seed, nsv, nrhs, power, bsize, run flags)
    "c1": (1024, 50000, 3, 0.05, 0.0, 0.01, 0, 10, 20, 4, 4096, dict(crit=1, tol=1e-6, miter=20)),
    "c2": (20000, 500000, 5, 0.01, 0.0, 0.01, 1, 10, 20, 4, 16384, REF_RUN),
    "c3": (100000, 1000000, 8, 0.005, 0.2, 0.02, 2, 20, 40, 6, 0, REF_RUN),
    # one GPU's SNP shard of the 100k x 1M cohort split over 8 GPUs
    "c3-shard8": (100000, 125000, 8, 0.005, 0.2, 0.02, 2, 20, 40, 6, 0, REF_RUN),
    # out-of-core: host memory (pinned, or pageable staged through pinned
    # blocks when the driver cannot pin it all), 16 GB device cap, ~512 MB
    # blocks; the second is the same pipeline at 1/8 of the SNPs (2 GB cap)
    "c4": (400000, 2000000, 10, 0.005, 0.0, 0.01, 3, 20, 40, 4, 65536, REF_RUN),
    "c4s": (400000, 250000, 10, 0.005, 0.0, 0.01, 3, 20, 40, 4, 65536, REF_RUN),
    # the 20k x 500k cohort written to a .bed file and re-read from it every
    # pass (the reference's on-disk block loop), 0.5 GB device cap
    "c2-file": (20000, 500000, 5, 0.01, 0.0, 0.01, 1, 10, 20, 4, 16384, REF_RUN),
}
STREAMED = {"c4": 16e9, "c4s": 16e9, "c2-file": 0.5e9}
PEAK_INT8 = 5.0e15   # dense int8 MFMA, ops/s
PEAK_FP64 = 78.6e12
PEAK_HBM = 8.0e12


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def layout_bytes(M, N, host_mode):
    """Bytes one pass copies when nothing stays resident: the padded device
    layout (rows of ceil(M/4) rounded up to 256 bytes, 1024-row multiples)
    from host memory, or the .bed rows themselves in file mode."""
    np_ = (M + 3) // 4
    if host_mode == 2:
        return N * np_
    return -(-N // 1024) * 1024 * (-(-np_ // 256) * 256)


class stdout_to_stderr:
    """Sends what the library prints on fd 1 (its per-iteration lines at
    verbose 2) to stderr while a full run is in progress, so the JSON line
    stays the only thing on stdout and a long run shows its progress."""

    def __enter__(self):
        sys.stdout.flush()
        self.saved = os.dup(1)
        os.dup2(2, 1)

    def __exit__(self, *exc):
        os.dup2(self.saved, 1)
        os.close(self.saved)


def max_principal_angle(U, V):
    """Largest principal angle (radians) between span(U) and span(V), from
    the sines (accurate for small angles)."""
    qu, _ = np.linalg.qr(U)
    qv, _ = np.linalg.qr(V)
    s = np.linalg.svd(qv - qu @ (qu.T @ qv), compute_uv=False)
    return float(np.arcsin(min(1.0, float(s.max()))))


def host_cpu_info():
    """Threads the CPU baseline uses and what the host offers: BLAS threads
    (the decode loop is serial, as one reference rank's), the CPUs this
    process may run on, and the host's physical cores (/proc/cpuinfo)."""
    try:
        from threadpoolctl import threadpool_info
        blas = max([t.get("num_threads", 1) for t in threadpool_info()] + [1])
    except Exception:
        blas = int(os.environ.get("OMP_NUM_THREADS", "1"))
    share = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    phys = set()
    try:
        cur = {}
        for ln in open("/proc/cpuinfo"):
            if ":" in ln:
                k, v = [x.strip() for x in ln.split(":", 1)]
                cur[k] = v
            elif cur:
                phys.add((cur.get("physical id"), cur.get("core id")))
                cur = {}
        if cur:
            phys.add((cur.get("physical id"), cur.get("core id")))
    except OSError:
        pass
    model = ""
    try:
        model = next(ln.split(":", 1)[1].strip() for ln in open("/proc/cpuinfo")
                     if ln.startswith("model name"))
    except (OSError, StopIteration):
        pass
    return {"blas_threads": int(blas), "cpu_share": int(share),
            "host_physical_cores": len(phys) or None, "host_cpu": model}


def cpu_bsize(M, N, nsv, nrhs, bsize):
    """Block size of the CPU run: the workload's own, else the reference's
    automatic rule (src/distpca.cpp:239-266) at -memory 16 (GB per rank)."""
    from oracle import oracle as O
    return O.resolve(O.Params(nsv=nsv, nrhs=nrhs, bsize=bsize, memory=16.0), M, N).bsize


def cpu_baseline(pca, M, N, k, bsize, gpu_passes=None, target_s=15.0, host_bytes=0):
    """The oracle's pass (the reference's serial cached decode,
    src/methods.cpp:838-843, plus threaded host dgemm) timed on a bounded SNP
    sample of the same matrix; seconds per full pass and per full run are
    extrapolated from it and labelled so.  The sample's decoded FP64 cache
    (rows x M x 8 bytes) stays within 64 GB less what the workload already
    holds in host memory (host_bytes: a matrix streamed from the host), and
    at least 3.2 GB (1,000 SNPs at 400,000 samples)."""
    from oracle import oracle as O
    info = host_cpu_info()
    np_ = (M + 3) // 4
    Q = np.linalg.qr(np.random.default_rng(0).standard_normal((M, k)))[0]
    max_rows = max(256, int(max(3.2e9, 64e9 - host_bytes) / (8.0 * M)))
    rows = min(N, 1024, max_rows)
    while True:
        packed = pca.packed(0, rows)
        op = O.OOCPass(packed.reshape(-1), M, rows, min(rows, bsize))
        op(Q)                                    # first touch: stats + table
        op.t_load = op.t_mm_a = op.t_mm_at = 0.0
        t0 = time.perf_counter()
        op(Q)                                    # cached pass, as later passes
        dt = time.perf_counter() - t0
        if dt * 2 > target_s or rows >= N or rows >= min(65536, max_rows):
            break
        rows = min(N, 65536, max_rows, rows * max(2, int(target_s / max(dt, 1e-3) / 2)))
    gbps = rows * np_ / dt / 1e9
    pass_s = dt * N / rows
    out = {"value": round(gbps, 4), "unit": "GB/s", "cores": info["blas_threads"], "kind": "port",
           "sample": f"one cached pass over the first {rows} of {N} SNPs ({M} samples, nrhs {k}, "
                     f"bsize {min(rows, bsize)}) of the same matrix: {dt:.2f} s (decode "
                     f"{op.t_load:.2f} s serial, Z Q {op.t_mm_a:.2f} s, Z^T T {op.t_mm_at:.2f} s "
                     f"on {info['blas_threads']} BLAS threads)",
           "extrapolated_pass_s": round(pass_s, 2)}
    out.update(info)
    if gpu_passes:
        out["extrapolated_run_s"] = round(pass_s * gpu_passes, 1)
        out["extrapolation"] = (f"linear in SNPs from the sample: {N / rows:.1f} x the sample per "
                                f"pass, x {gpu_passes} passes of the GPU run")
    return out


def cpu_plan(a, D):
    """The CPU plan beside the GPU: the oracle's whole run (wall clock and the
    reference's stage timers) for the 1,024 x 50,000 and 20,000 x 500,000
    workloads; for the 100k x 1M one outer iteration (power + 1 passes,
    the first with its first-touch decode) and its extrapolation to the GPU
    run's pass count (--cpu-passes)."""
    from oracle import oracle as O
    M, N, npop, fst, admix, miss, seed, nsv, nrhs, power, bsize, runf = WORKLOADS[a.workload]
    spec = D.gen_spec(M, N, npop=npop, fst=fst, admix_frac=admix, missing=miss, seed=seed)
    log(f"cpu plan {a.workload}: generating {M} x {N} and copying it to host memory")
    with D.PCA.synthetic(spec, D.Params(nsv=2, nrhs=4)) as g:
        packed = g.packed().reshape(-1)
    b = cpu_bsize(M, N, nsv, nrhs, bsize)
    p = O.resolve(O.Params(nsv=nsv, nrhs=nrhs, power=power, bsize=b, **runf), M, N)
    info = host_cpu_info()
    res = {"workload": a.workload, "samples": M, "snps": N, "nsv": p.nsv, "nrhs": p.nrhs,
           "power": p.power, "bsize": p.bsize, **{k: getattr(p, k) for k in ("crit", "tol", "miter")},
           **info, "kind": "port"}
    t0 = time.perf_counter()
    if a.workload in ("c1", "c2"):
        log(f"cpu plan: full oracle run (bsize {p.bsize})")
        with stdout_to_stderr():  # one progress line per iteration
            r = O.rsi(packed, M, N, p, verbose=True)
        res.update({"what": "whole run", "wall_s": round(time.perf_counter() - t0, 2),
                    "total_its": r.total_its, "passes": r.passes,
                    "stage_timers_s": {k: round(v, 2) for k, v in r.timers.items()}})
    else:
        log(f"cpu plan: one outer iteration, {power + 1} passes (bsize {p.bsize})")
        nblk_log = max(1, ((N + p.bsize - 1) // p.bsize) // 8)
        op = O.OOCPass(packed, M, N, p.bsize,
                       progress=lambda j, n: (j + 1) % nblk_log == 0 and log(
                           f"cpu plan: block {j + 1}/{n}"))
        Q = O.omega(M, p.nrhs)
        for _ in range(power + 1):
            t1 = time.perf_counter()
            Q = O.orthonormalise(op(Q))
            log(f"cpu plan: pass {op.count} {time.perf_counter() - t1:.1f} s")
        wall = time.perf_counter() - t0
        per = wall / (power + 1)
        res.update({"what": f"one outer iteration ({power + 1} passes incl. the first-touch one)",
                    "wall_s": round(wall, 1), "s_per_pass": round(per, 2),
                    "stage_timers_s": {"load": round(op.t_load, 2), "mm_a": round(op.t_mm_a, 2),
                                       "mm_at": round(op.t_mm_at, 2)}})
        if a.cpu_passes:
            res["extrapolated_run_s"] = round(per * a.cpu_passes, 1)
            res["extrapolation"] = (f"labelled extrapolation: {per:.1f} s per pass x "
                                    f"{a.cpu_passes} passes of the GPU run")
    print(json.dumps({"cpu_plan": res}), flush=True)


def serve_bench(a, D):
    """Serving workload: seeded Poisson arrivals; 70% projections of 1-4,096
    new samples (log-uniform) onto a resident 1,000,000-SNP nsv-20 loading
    set (5-population, Fst 0.01 generator, 20,000 training samples); 30%
    cohort-PCA jobs on 2,000-50,000 samples x 100,000-1,000,000 SNPs
    (log-uniform), nsv 10, nrhs 20.  Same trace for the mixed and FP64 forms."""
    import threading
    rng = np.random.default_rng(a.seed)
    Ns = 1_000_000
    n_req = int(a.rate * a.duration)
    gaps = rng.exponential(1.0 / a.rate, n_req)
    arrivals = np.cumsum(gaps)
    kinds = rng.random(n_req) < 0.7
    sizes = np.unique(np.round(np.exp(rng.uniform(0, np.log(4096), 16))).astype(int))
    req_size = rng.choice(sizes, n_req)
    job_M = np.round(np.exp(rng.uniform(np.log(2000), np.log(50000), n_req))).astype(int)
    job_N = np.round(np.exp(rng.uniform(np.log(100000), np.log(1000000), n_req))).astype(int)
    log(f"serve: {n_req} requests over {arrivals[-1]:.1f} s, "
        f"{int(kinds.sum())} projections, {int((~kinds).sum())} jobs")
    new = D.PCA.synthetic(D.gen_spec(4096, Ns, npop=5, fst=0.01, missing=0.01, seed=11),
                          D.Params(nsv=2, nrhs=4))
    pool = new.packed()
    new.close()
    def pinned(x):
        buf = D.PinnedArray(x.shape)
        buf.array[...] = x
        return buf.array, buf

    payloads = {int(n): pinned(pool[:, :(int(n) + 3) // 4]) for n in sizes}
    del pool
    results = {}
    for form_name, form in (("mixed", D.MIXED), ("fp64", D.FP64)):
        log(f"serve[{form_name}]: training the loading set (20000 x {Ns})")
        tr = D.PCA.synthetic(D.gen_spec(20000, Ns, npop=5, fst=0.01, missing=0.01, seed=1),
                             D.Params(nsv=20, nrhs=40, power=2, crit=1, tol=1e-3,
                                      precision=form))
        sig, U, rep = tr.run()
        model = D.Model(tr, U, sig, 20)
        tr.close()
        srv = D.Server(model, hbm_budget=int(200e9), batch_capacity=16384, workers=4)
        tickets = []
        lat = {"project": [], "job": []}
        done_evt = threading.Event()

        def waiter():
            for kind, t in iter_tickets():
                _, info = srv.wait(t)
                lat[kind].append(info["t_done"] - info["t_submit"])
            done_evt.set()

        q = []
        cond = threading.Condition()

        def iter_tickets():
            i = 0
            while i < n_req:
                with cond:
                    while len(q) <= i:
                        cond.wait()
                    item = q[i]
                i += 1
                yield item

        th = threading.Thread(target=waiter)
        th.start()
        t0 = time.perf_counter()
        for i in range(n_req):
            delay = arrivals[i] - (time.perf_counter() - t0)
            if delay > 0:
                time.sleep(delay)
            if kinds[i]:
                pay = payloads[int(req_size[i])][0]
                t = srv.submit_project(pay, int(req_size[i]))
                item = ("project", t)
            else:
                spec = D.gen_spec(int(job_M[i]), int(job_N[i]), npop=5, fst=0.01,
                                  missing=0.01, seed=1000 + i)
                t = srv.submit_job(spec, D.Params(nsv=10, nrhs=20, precision=form))
                item = ("job", t)
            with cond:
                q.append(item)
                cond.notify_all()
            if (i + 1) % 50 == 0:
                log(f"serve[{form_name}]: {i + 1}/{n_req} submitted")
        while th.is_alive():
            th.join(timeout=30)
            if th.is_alive():
                log(f"serve[{form_name}]: draining, {len(lat['project']) + len(lat['job'])}"
                    f"/{n_req} done")
        wall = time.perf_counter() - t0
        cnt = srv.counters()
        srv.close()
        model.close()
        res = {"wall_s": round(wall, 2), "requests_per_s": round(n_req / wall, 3),
               "jobs_per_s": round(len(lat["job"]) / wall, 3),
               "batches": cnt["batches"],
               "mean_batch_samples": round(cnt["batch_samples"] / max(1, cnt["batches"]), 1)}
        for kind in ("project", "job"):
            v = np.array(lat[kind]) * 1e3
            if v.size:
                res[f"{kind}_p50_ms"] = round(float(np.percentile(v, 50)), 2)
                res[f"{kind}_p99_ms"] = round(float(np.percentile(v, 99)), 2)
                res[f"{kind}_n"] = int(v.size)
        log(f"serve[{form_name}]: {res}")
        results[form_name] = res
    out = {"metric": "serving p50/p99 latency and jobs/s, mixed vs FP64",
           "config": {"workload": "c5 serving", "rate_per_s": a.rate, "duration_s": a.duration,
                      "seed": a.seed, "projection_sizes": [int(x) for x in sizes],
                      "loading_set": "20000 x 1000000, nsv 20", "workers": 4,
                      "batch_capacity": 16384},
           "mixed": results["mixed"], "fp64": results["fp64"]}
    print(json.dumps(out), flush=True)


def load_traffic(name):
    path = os.path.join(ROOT, "profiles", f"traffic_{name}.json")
    if os.path.exists(path):
        with open(path) as f:
            return json.load(f)
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="c3", choices=sorted(WORKLOADS) + ["c5"])
    ap.add_argument("--rate", type=float, default=5.0, help="c5: arrivals per second")
    ap.add_argument("--duration", type=float, default=20.0, help="c5: seconds of arrivals")
    ap.add_argument("--seed", type=int, default=5)
    ap.add_argument("--fp64-steps", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-full", action="store_true",
                    help="skip the full PCA runs of both forms (and their accuracy comparison)")
    ap.add_argument("--full", action="store_true", help=argparse.SUPPRESS)  # the default
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--slices", type=int, default=0,
                    help="int8 slices of the mixed form's FP64 operand (2 or 3; 0 = default)")
    ap.add_argument("--cpu-plan", action="store_true",
                    help="run only the CPU plan (oracle whole run for c1/c2, one outer "
                         "iteration for c3) and print its own JSON line")
    ap.add_argument("--cpu-passes", type=int, default=0,
                    help="--cpu-plan: GPU-run pass count to extrapolate one iteration to")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run only a plain 1 GiB device copy (HBM counter calibration)")
    a = ap.parse_args()
    if a.calibrate_copy:
        import torch
        n = 1 << 30
        src = torch.ones(n // 4, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"calibration": "device copy", "bytes_read": n, "bytes_written": n,
                          "launches": 3}))
        return
    if a.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {a.gpus} is not supported "
              f"(the SNP-sharded multi-GPU design is described, not built)", file=sys.stderr)
        sys.exit(2)

    import distpca_amd as D
    if a.workload == "c5":
        serve_bench(a, D)
        return
    if a.cpu_plan:
        cpu_plan(a, D)
        return
    M, N, npop, fst, admix, miss, seed, nsv, nrhs, power, bsize, runf = WORKLOADS[a.workload]
    spec = D.gen_spec(M, N, npop=npop, fst=fst, admix_frac=admix, missing=miss, seed=seed)
    np_ = (M + 3) // 4
    packed_bytes = N * np_

    t0 = time.perf_counter()
    cap = int(STREAMED.get(a.workload, 0))
    if cap and a.workload == "c4s":
        cap = int(2e9)   # same block pipeline on a matrix 8x smaller than c4
    bed = None
    if a.workload.endswith("-file"):
        import tempfile
        tmp = tempfile.TemporaryDirectory()
        bed = os.path.join(tmp.name, a.workload)
        log(f"writing {a.workload}: {M} x {N} ({packed_bytes / 1e9:.2f} GB) to {bed}.bed")
        D.write_synthetic(bed, spec)
        log(f"written in {time.perf_counter() - t0:.1f} s")
    else:
        log(f"generating {a.workload}: {M} x {N} ({packed_bytes / 1e9:.2f} GB packed) in HBM")

    def open_ctx(params):
        if bed:
            params.host_mode = D.HOST_FILE
            return D.PCA.from_bfile(bed, params)
        return D.PCA.synthetic(spec, params)

    pm = D.Params(nsv=nsv, nrhs=nrhs, power=power, bsize=bsize, precision=D.MIXED,
                  device_cap=cap, verbose=2, slices=a.slices, **runf)
    mixed = open_ctx(pm)
    log(f"generated + stats in {time.perf_counter() - t0:.1f} s")
    Q = np.linalg.qr(D.omega(M, nrhs))[0]

    log(f"mixed: warmup {a.warmup}, timed {a.steps} passes")
    if a.warmup:
        mixed.bench_pass(Q, a.warmup)
    # bench_pass uploads Q, drains the stream, then times K passes from the
    # first launch to the drained stream on the host clock (ms[4]) and with
    # events on the library's stream (ms[0..3])
    ms = mixed.bench_pass(Q, a.steps)
    ms_step = ms[4]
    wall = ms_step * a.steps * 1e-3
    value = packed_bytes * a.steps / wall / 1e9
    log(f"mixed pass {ms_step:.2f} ms (host clock), events {ms[:4]}")

    stream = None
    if cap:
        compute = ms[1] + ms[2] + ms[6]
        # the leading blocks that fit the cap stay in HBM after the first
        # pass; ms[8] is what a pass copies
        stream = {"device_cap_bytes": cap,
                  "host_memory": {0: "pinned", 1: "pageable, staged through pinned blocks",
                                  2: ".bed file"}.get(int(ms[7]), "?"),
                  "h2d_bytes_per_pass": int(ms[8]),
                  "resident_fraction": round(1.0 - ms[8] / layout_bytes(M, N, int(ms[7])), 4),
                  "h2d_ms_per_pass": round(ms[5], 2),
                  "h2d_GBps_copy_engine": round(ms[8] / (ms[5] * 1e-3) / 1e9, 2),
                  "h2d_GBps_end_to_end": round(ms[8] / (ms_step * 1e-3) / 1e9, 2),
                  "matrix_GBps": round(value, 2),
                  "compute_ms_per_pass": round(compute, 2),
                  "overlap_fraction": round(max(0.0, (ms[5] + compute - ms_step)) /
                                            max(1e-9, min(ms[5], compute)), 3)}
        log(f"streamed: {stream}")
    # Roofline of the dominant product kernel.  Per launch it reads the packed
    # matrix once plus its FP64-equivalent operand (Q or T) and writes its
    # result: N ceil(M/4) + (M + N) k 8 bytes against 2 M N k useful int8 ops,
    # about 315 op/B at 100k x 1M, k 40 -- below the int8 ridge (5.0 POP/s /
    # 8 TB/s = 625 op/B), so HBM is the binding roof for the useful work.
    flops_kernel = 2.0 * M * N * nrhs            # useful ops of one product (Z Q or Z^T T)
    bytes_kernel = packed_bytes + (M + N) * nrhs * 8.0
    bytes_pass = packed_bytes + 2.0 * M * nrhs * 8  # algorithmic bytes of a whole pass
    dom = 1 if ms[1] >= ms[2] else 2
    dom_name = "k_za_i8" if dom == 1 else "k_zt_i8"
    sec = ms[dom] * 1e-3
    traffic = load_traffic(f"{a.workload}_{dom_name}")
    pass_traffic = load_traffic(f"{a.workload}_pass")
    # issued int8 MFMA work: two code slots (h, m) x the 16-wide column tiles
    # that hold the slices of nrhs columns (Z Q's mask slot leaves out the
    # tiles holding only the last slice)
    ns = a.slices or 3
    F, R = divmod(nrhs, 16)
    RT = (ns * R + 15) // 16
    if RT >= ns:
        F, R, RT = F + 1, 0, 0
    vt = ns * F + RT
    mask_tiles = vt - F - (1 if RT > 0 and (ns - 1) * R <= 16 * (RT - 1) else 0)
    issued = {"k_zt_i8": 2.0 * M * N * 16 * vt * 2, "k_za_i8": 2.0 * M * N * 16 * (vt + mask_tiles)}
    roofline = {"bound": "hbm", "kernel": dom_name,
                "achieved": round(bytes_kernel / sec / 1e9, 1), "peak": PEAK_HBM / 1e9,
                "unit": "GB/s", "frac": round(bytes_kernel / sec / PEAK_HBM, 4),
                "traffic": traffic.get("bytes_per_launch") if traffic else None,
                "algorithmic_bytes": bytes_kernel,
                "traffic_ratio": round(traffic["bytes_per_launch"] / bytes_kernel, 3)
                if traffic else None,
                "useful_mfma_frac": round(flops_kernel / sec / PEAK_INT8, 4),
                "issued_mfma_frac": round(issued[dom_name] / sec / PEAK_INT8, 4),
                "kernel_ms": round(ms[dom], 3),
                "za_ms": round(ms[1], 3), "zt_ms": round(ms[2], 3), "aux_ms": round(ms[3], 3),
                "pass_algorithmic_bytes": bytes_pass,
                "pass_traffic_bytes": pass_traffic.get("bytes_per_pass") if pass_traffic else None,
                "pass_traffic_ratio": round(pass_traffic["bytes_per_pass"] / bytes_pass, 3)
                if pass_traffic else None,
                "slices": ns}

    # One context at a time after the mixed pass timing: HIP spreads streams
    # over a few hardware queues, and a second streamed context's streams can
    # share a queue with the first's copy stream, serialising its copies with
    # its kernels.
    extra = {}
    full = not a.no_fp64 and not a.no_full
    if full:
        log("full PCA run, mixed")
        with stdout_to_stderr():
            s_mx, U_mx, rep = mixed.run()
        extra["mixed_pca_wall_s"] = round(rep["wall"], 3)
        extra["mixed_pca_its"] = rep["total_its"]
        extra["mixed_pca_passes"] = rep["passes"]
        extra["mixed_sigma"] = [float(x) for x in s_mx]
    cpu = None
    if not a.no_cpu:
        log("cpu baseline (oracle pass on a bounded sample)")
        # a matrix streamed from host memory (not from a file) holds its bytes there
        host_bytes = packed_bytes if (cap and not bed) else 0
        cpu = cpu_baseline(mixed, M, N, nrhs, cpu_bsize(M, N, nsv, nrhs, bsize),
                           extra.get("mixed_pca_passes"), a.cpu_seconds, host_bytes)
    mixed.close()

    if not a.no_fp64:
        log(f"fp64: {a.fp64_steps} passes")
        pf = D.Params(nsv=nsv, nrhs=nrhs, power=power, bsize=bsize, precision=D.FP64,
                      device_cap=cap, verbose=2, **runf)
        f64 = open_ctx(pf)
        f64.bench_pass(Q, 1)
        ms64 = f64.bench_pass(Q, a.fp64_steps)
        wall64 = ms64[4]
        extra["fp64_pass_ms"] = round(wall64, 3)
        extra["fp64_za_ms"] = round(ms64[1], 3)
        extra["fp64_zt_ms"] = round(ms64[2], 3)
        extra["fp64_frac_of_peak"] = round(2 * flops_kernel / (wall64 * 1e-3) / PEAK_FP64, 4)
        extra["mixed_speedup_vs_fp64"] = round(wall64 / ms_step, 3)
        if full:
            log("full PCA run, fp64")
            with stdout_to_stderr():
                s, U, rep = f64.run()
            extra["fp64_pca_wall_s"] = round(rep["wall"], 3)
            extra["fp64_pca_its"] = rep["total_its"]
            extra["fp64_pca_passes"] = rep["passes"]
            extra["fp64_sigma"] = [float(x) for x in s]
        f64.close()
        if full:
            # pass-saving schedule: (A A^T Q) X reused as the next first pass
            log("full PCA run, mixed with one pass saved per iteration")
            pm.strict_passes = 0
            ctx = open_ctx(pm)
            with stdout_to_stderr():
                s_sv, U_sv, rep = ctx.run()
            ctx.close()
            extra["mixed_saving_pca_wall_s"] = round(rep["wall"], 3)
            extra["mixed_saving_pca_its"] = rep["total_its"]
            extra["mixed_saving_pca_passes"] = rep["passes"]
            # the population PCs (npop - 1 of them) are the converged ones; the
            # rest lie in the noise bulk, where neither form converges them
            npc = min(nsv, max(1, npop - 1))
            angle_js = sorted({1, npc, min(4, nsv), min(7, nsv), nsv})
            extra["population_pcs"] = npc
            for tag, (s_x, U_x) in (("mixed", (s_mx, U_mx)), ("mixed_saving", (s_sv, U_sv))):
                rel = np.abs(s_x ** 2 - s ** 2) / s ** 2
                extra[f"{tag}_vs_fp64_eig_relerr_max"] = float(rel.max())
                extra[f"{tag}_vs_fp64_eig_relerr_population"] = float(rel[:npc].max())
                extra[f"{tag}_vs_fp64_angles_top"] = {
                    j: max_principal_angle(U_x[:, :j], U[:, :j]) for j in angle_js}
            extra["mixed_vs_fp64_eig_relerr"] = [
                float(x) for x in np.abs(s_mx ** 2 - s ** 2) / s ** 2]
            if ns == 3:
                # the two-slice variant of the mixed form on the same matrix
                log("mixed pass and full PCA run with two int8 slices")
                pm.strict_passes = 1
                pm.slices = 2
                ctx = open_ctx(pm)
                ctx.bench_pass(Q, 1)
                ms2 = ctx.bench_pass(Q, a.steps)
                with stdout_to_stderr():
                    s_2, U_2, rep = ctx.run()
                ctx.close()
                rel = np.abs(s_2 ** 2 - s ** 2) / s ** 2
                extra["mixed2_pass_ms"] = round(ms2[4], 3)
                extra["mixed2_za_ms"] = round(ms2[1], 3)
                extra["mixed2_zt_ms"] = round(ms2[2], 3)
                extra["mixed2_pca_wall_s"] = round(rep["wall"], 3)
                extra["mixed2_pca_its"] = rep["total_its"]
                extra["mixed2_vs_fp64_eig_relerr_max"] = float(rel.max())
                extra["mixed2_vs_fp64_eig_relerr_population"] = float(rel[:npc].max())
                extra["mixed2_vs_fp64_angles_top"] = {
                    j: max_principal_angle(U_2[:, :j], U[:, :j]) for j in angle_js}

    out = {
        "metric": "PCA wall-s & HBM GB/s, mixed vs FP64 at equal PC error; p50/p99 serve latency",
        "value": round(value, 3), "unit": "GB/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": round(ms_step, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": f"int8x{ns}-int32acc",
        "data": "synthetic (Balding-Nichols genotypes generated in HBM)",
        "config": {"workload": f"{a.workload}: {M} samples x {N} SNPs, nrhs {nrhs}, "
                               f"mixed-precision pass A A^T Q",
                   "samples": M, "snps": N, "nrhs": nrhs, "packed_bytes": packed_bytes,
                   "parallelism": "single GPU"},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    if stream:
        out["streaming"] = stream
    out.update(extra)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
