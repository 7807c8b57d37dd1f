"""GPU checks of loadings, projection and the serving loop.

Projection arithmetic is pinned against the oracle: scores = z^T W with
W = Z U diag(1/sigma^2) and z standardised by the training table (dense
numpy from the oracle's decode).  Tolerances: FP64 form 1e-12 of the score
scale, mixed form 1e-5.  Serving: a request's scores are bitwise the same
alone and inside a mixed batch, and a cohort job's sigma is bitwise the same
alone and run concurrently with other jobs and projections.
"""
import numpy as np
import pytest

import distpca_amd as D
from oracle import oracle as O

pytestmark = pytest.mark.gpu

M, N = 800, 20000


def _train(precision):
    spec = D.gen_spec(M, N, npop=3, fst=0.05, missing=0.01, seed=21)
    p = D.Params(nsv=4, nrhs=10, power=3, crit=1, tol=1e-12, miter=80, bsize=5000,
                 precision=precision)
    pca = D.PCA.synthetic(spec, p)
    sigma, rhs, rep = pca.run()
    return pca, sigma, rhs


def _oracle_scores(zval, packed_new, n, W):
    codes = np.stack([(packed_new[:, i >> 2] >> (2 * (i & 3))) & 3 for i in range(n)], axis=1)
    z = zval[np.arange(N)[:, None], codes]      # N x n, training standardisation
    return z.T @ W


@pytest.mark.parametrize("precision", [D.FP64, D.MIXED])
def test_projection_matches_oracle_and_training_rows(precision):
    print(f"[gpu] projection precision {precision}", flush=True)
    pca, sigma, rhs = _train(precision)
    packed = pca.packed()
    s, c, zval = pca.snp_stats()
    model = D.Model(pca, rhs, sigma, 4)
    Z = O.dense_z(packed.reshape(-1), N, M)
    W = (Z @ rhs[:, :4]) / sigma[:4] ** 2
    # training samples project onto their own rows of U (converged PCs 1-2)
    got = model.project(packed, M)
    ref = Z.T @ W
    tol = 1e-12 if precision == D.FP64 else 1e-5
    assert np.abs(got - ref).max() <= tol * np.abs(ref).max()
    assert np.abs(got[:, :2] - rhs[:, :2]).max() < 1e-6 * np.abs(rhs[:, :2]).max() + (
        0 if precision == D.FP64 else 1e-5)
    # new samples from the same populations, odd count (ragged last byte)
    new = D.PCA.synthetic(D.gen_spec(37, N, npop=3, fst=0.05, missing=0.02, seed=99),
                          D.Params(nsv=2, nrhs=4))
    pn = new.packed()
    gn = model.project(pn, 37)
    rn = _oracle_scores(zval, pn, 37, W)
    assert np.abs(gn - rn).max() <= tol * np.abs(rn).max()
    model.close()
    pca.close()
    new.close()


@pytest.mark.parametrize("precision", [D.FP64, D.MIXED])
def test_server_same_alone_and_in_mix(precision):
    print(f"[gpu] server precision {precision}", flush=True)
    pca, sigma, rhs = _train(precision)
    model = D.Model(pca, rhs, sigma, 4)
    new = D.PCA.synthetic(D.gen_spec(600, N, npop=3, fst=0.05, missing=0.02, seed=7),
                          D.Params(nsv=2, nrhs=4))
    pool = new.packed()
    srv = D.Server(model, hbm_budget=8 << 30, batch_capacity=1024, workers=2)
    reqs = {n: np.ascontiguousarray(pool[:, :(n + 3) // 4]) for n in (1, 5, 130, 400)}
    alone = {}
    for n, pay in reqs.items():
        t = srv.submit_project(pay, n)
        alone[n], info = srv.wait(t)
        assert info["batch_samples"] >= n
    jspec = D.gen_spec(500, 8000, npop=3, fst=0.05, missing=0.01, seed=5)
    jpar = D.Params(nsv=3, nrhs=6, power=2, crit=1, tol=1e-8, miter=30, precision=precision)
    tj = srv.submit_job(jspec, jpar)
    sig_alone, _ = srv.wait(tj)
    # now everything at once: two jobs and all requests in flight together
    tjs = [srv.submit_job(jspec, jpar), srv.submit_job(
        D.gen_spec(700, 12000, npop=2, fst=0.05, missing=0.01, seed=6), jpar)]
    ts = {n: srv.submit_project(pay, n) for n, pay in reqs.items()}
    batches = set()
    for n, t in ts.items():
        sc, info = srv.wait(t)
        batches.add(info["batch"])
        assert np.array_equal(sc, alone[n]), n
        assert np.array_equal(sc, model.project(reqs[n], n)), n
    sig_mix, jinfo = srv.wait(tjs[0])
    srv.wait(tjs[1])
    assert np.array_equal(sig_mix, sig_alone)
    assert jinfo["footprint"] > 0
    assert len(batches) < len(reqs)  # several requests shared a batch
    cnt = srv.counters()
    assert cnt["jobs_done"] == 3 and cnt["projections_done"] == 8
    srv.close()
    model.close()
    pca.close()
    new.close()


@pytest.mark.parametrize("precision", [D.FP64, D.MIXED])
def test_projection_batch_invariant(precision):
    # a loading set wide enough (200,000 SNPs) that batch classes of 256 to
    # 4,096 samples would pick different SNP splits if the split followed the
    # batch width; scores must be bitwise the same whatever the batch size
    print(f"[gpu] projection batch invariance precision {precision}", flush=True)
    spec = D.gen_spec(600, 200000, npop=3, fst=0.05, missing=0.01, seed=31)
    p = D.Params(nsv=3, nrhs=6, power=2, crit=1, tol=1e-6, miter=10, precision=precision)
    with D.PCA.synthetic(spec, p) as pca:
        sigma, rhs, _ = pca.run()
        model = D.Model(pca, rhs, sigma, 3)
    new = D.PCA.synthetic(D.gen_spec(4000, 200000, npop=3, fst=0.05, missing=0.02, seed=32),
                          D.Params(nsv=2, nrhs=4))
    pool = new.packed()
    new.close()
    full = model.project(np.ascontiguousarray(pool[:, :1000]), 4000)
    for n in (1, 7, 300, 1100):
        part = model.project(np.ascontiguousarray(pool[:, :(n + 3) // 4]), n)
        assert np.array_equal(part, full[:n]), n
    model.close()


def test_server_rejects_job_beyond_budget():
    srv = D.Server(None, hbm_budget=64 << 20, workers=1)
    with pytest.raises(RuntimeError, match="exceeds the HBM budget"):
        srv.submit_job(D.gen_spec(20000, 200000, npop=2), D.Params(nsv=2, nrhs=4))
    assert srv.counters()["jobs_rejected"] == 1
    srv.close()


def test_serving_at_full_loading_set_size():
    # the serving workload's own shapes: a resident 1,000,000-SNP, nsv-20
    # loading set (5 populations, Fst 0.01, 20,000 training samples, mixed
    # form), projection requests from 1 to 4,096 new samples and a cohort job
    # of 2,000 x 100,000 running beside them.  Each request's scores are
    # bitwise the same alone and in the mix; training samples project onto
    # their own rows of U on the converged population PCs.
    Nl, Ml = 1000000, 20000
    print(f"[gpu] serving at {Ml} x {Nl}, nsv 20, mixed", flush=True)
    spec = D.gen_spec(Ml, Nl, npop=5, fst=0.01, missing=0.01, seed=1)
    p = D.Params(nsv=20, nrhs=40, power=2, crit=1, tol=1e-6, miter=30, precision=D.MIXED)
    with D.PCA.synthetic(spec, p) as pca:
        sigma, rhs, rep = pca.run()
        model = D.Model(pca, rhs, sigma, 20)
        train = np.ascontiguousarray(pca.packed()[:, :16])       # samples 0..63
    own = model.project(train, 64)
    # 5 populations: PCs 1-4 carry the structure and are converged
    scale = np.abs(rhs[:64, :4]).max()
    assert np.abs(own[:, :4] - rhs[:64, :4]).max() <= 1e-4 * scale
    new = D.PCA.synthetic(D.gen_spec(4096, Nl, npop=5, fst=0.01, missing=0.02, seed=12),
                          D.Params(nsv=2, nrhs=4))
    pool = new.packed()
    new.close()
    srv = D.Server(model, hbm_budget=int(100e9), batch_capacity=16384, workers=2)
    sizes = (1, 3, 37, 300, 1500, 4096)
    reqs = {n: np.ascontiguousarray(pool[:, :(n + 3) // 4]) for n in sizes}
    alone = {}
    for n, pay in reqs.items():
        alone[n], _ = srv.wait(srv.submit_project(pay, n))
    jspec = D.gen_spec(2000, 100000, npop=5, fst=0.01, missing=0.01, seed=1005)
    jpar = D.Params(nsv=10, nrhs=20, precision=D.MIXED)
    sig_alone, _ = srv.wait(srv.submit_job(jspec, jpar))
    tj = srv.submit_job(jspec, jpar)
    ts = {n: srv.submit_project(pay, n) for n, pay in reqs.items()}
    for n, t in ts.items():
        sc, info = srv.wait(t)
        assert np.array_equal(sc, alone[n]), n
    sig_mix, jinfo = srv.wait(tj)
    assert np.array_equal(sig_mix, sig_alone)
    print(f"[gpu] serving: job sigma[:3] {sig_mix[:3]}, footprint {jinfo['footprint']}", flush=True)
    srv.close()
    model.close()
