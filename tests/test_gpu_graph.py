"""Graph-replayed outer iterations against eager ones, through the C ABI.

On a resident context every outer iteration after the first is replayed
from a captured HIP graph (one graph per state of the buffers that trade
places within an iteration); with DPCA_NO_GRAPH=1 the same work is enqueued
launch by launch.  The graph holds the same launches with the same
arguments in the same order, so sigma, the Ritz vectors, the iteration and
pass counts and the criterion history must agree bit for bit: over several
replays of each graph, with the host's copies and synchronisation between
them, and again on a second run of the same context (graphs reused).
"""
import os

import numpy as np
import pytest

import distpca_amd as D

pytestmark = pytest.mark.gpu

# (schedule, crit): the reference's block schedule with the pass-saving
# rotation, the strict block schedule, and the in-core variant
CASES = [("saving", 1), ("strict", 2), ("incore", 0)]


def _runs(spec, params, graph):
    if graph:
        os.environ.pop("DPCA_NO_GRAPH", None)
    else:
        os.environ["DPCA_NO_GRAPH"] = "1"
    try:
        out = []
        with D.PCA.synthetic(spec, params) as pca:
            for _ in range(2):
                sigma, rhs, rep = pca.run()
                out.append((sigma, rhs, rep, pca.crit_history()))
        return out
    finally:
        os.environ.pop("DPCA_NO_GRAPH", None)


@pytest.mark.parametrize("precision", [D.FP64, D.MIXED])
@pytest.mark.parametrize("schedule,crit", CASES)
def test_graph_iterations_bitwise_equal(precision, schedule, crit):
    M, N = 1000, 20000
    print(f"[gpu] graph vs eager: precision {precision} schedule {schedule} crit {crit}",
          flush=True)
    spec = D.gen_spec(M, N, npop=3, fst=0.05, missing=0.01, seed=21)
    params = D.Params(nsv=4, nrhs=12, power=3, crit=crit, tol=1e-14, miter=7,
                      bsize=N if schedule == "incore" else 4096, precision=precision,
                      strict_passes=0 if schedule == "saving" else 1)
    eager = _runs(spec, params, graph=False)
    graph = _runs(spec, params, graph=True)
    for (s0, u0, r0, h0), (s1, u1, r1, h1) in zip(eager, graph):
        assert r0["graph_its"] == 0
        # iterations 1 .. total_its - 1 replay graphs (two are captured: the
        # buffers' state alternates between iterations)
        assert r1["graph_its"] == r1["total_its"] - 1 >= 4
        assert r0["total_its"] == r1["total_its"] and r0["passes"] == r1["passes"]
        assert r0["incore"] == r1["incore"] == (schedule == "incore")
        assert np.array_equal(s0.view(np.int64), s1.view(np.int64))
        assert np.array_equal(u0.view(np.int64), u1.view(np.int64))
        assert np.array_equal(np.asarray(h0).view(np.int64), np.asarray(h1).view(np.int64))
    # the second run of a context starts from the same Omega: same answer
    assert np.array_equal(graph[0][0], graph[1][0])
    assert np.array_equal(graph[0][1], graph[1][1])


def test_graph_stage_timers_add_up():
    # replayed iterations are timed as a whole and split over the stage
    # timers in the eager first iteration's proportions
    spec = D.gen_spec(1000, 20000, npop=3, fst=0.05, missing=0.01, seed=22)
    params = D.Params(nsv=4, nrhs=12, power=3, crit=1, tol=1e-14, miter=6, bsize=4096,
                      precision=D.MIXED, strict_passes=0)
    with D.PCA.synthetic(spec, params) as pca:
        _, _, rep = pca.run()
    assert rep["graph_its"] == rep["total_its"] - 1
    parts = rep["t_mm"] + rep["t_mm_a"] + rep["t_mm_at"] + rep["t_gs"] + rep["t_rr"]
    assert 0 < parts <= rep["wall"]
    assert rep["t_mm_a"] > 0 and rep["t_mm_at"] > 0 and rep["t_gs"] > 0
