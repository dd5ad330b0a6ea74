"""The headline index (N = 10M FP64 points, D = 8, ratios 100,50,20, sigma divisor 2.2, seed
20260330) through the batched engine: a sample of the measured workload -- 64 boxes with
selectivity log-uniform in [1e-5, 1e-2] and 64 KNN points with K 10 and 100 -- on every result
path (captured graph, plain launches, plain launches with statistics), against the oracle's
exact refine of the reference traversal.  The oracle queries run in a thread pool (its C calls
release the GIL)."""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import oracle_cases as OC
import oracle_py as O
import tblaeq
from tblaeq import workload as W

pytestmark = pytest.mark.gpu

C2 = W.C2


@pytest.fixture(scope="module")
def c2():
    print("building the 10M-point index (handler and oracle)", flush=True)
    h = tblaeq.Handler.random(C2["N"], C2["D"], C2["vmin"], C2["vmax"], C2["is_int"], C2["ratios"], C2["seed"],
                              C2["sigma_div"])
    o = O.Index.generate(C2["N"], C2["D"], C2["vmin"], C2["vmax"], C2["is_int"], C2["ratios"], C2["seed"],
                         C2["sigma_div"])
    h.prepare("L0")
    h.batch_init(256, 1 << 26, 1 << 26)
    yield h, o
    h.close()


def _pool_map(fn, items):
    with ThreadPoolExecutor(16) as ex:
        return list(ex.map(fn, items))


def _paths(h, qtype, q, Ks=None):
    h.batch_set(qtype, q, Ks)
    h.batch_run()
    graph = h.batch_results()
    h.batch_use_graph(False)
    try:
        h.batch_run()
        plain = h.batch_results()
    finally:
        h.batch_use_graph(True)
    h.batch_run(stats=True)
    return graph, plain, h.batch_results(), h.batch_stats()


def test_c2_range_sample(c2):
    h, o = c2
    data = h.data()
    boxes, _ = W.range_boxes(data, 64, 20260331, 1e-5, 1e-2)
    _, pts = OC.points_outside_radius(o, limit=16)
    boxes = np.concatenate([boxes, OC.degenerate_boxes(pts)])
    graph, plain, res, (surv, fine) = _paths(h, "range", boxes)
    D = h.D

    def expect(row):
        r = o.range(row[:D], row[D:])
        return r.refine_range(row[:D], row[D:]), r.levels()

    exp = _pool_map(expect, boxes)
    for i in range(len(boxes)):
        e, (_, s, out) = exp[i]
        for got, name in ((graph[i], "graph"), (plain[i], "plain"), (res[i], "statistics")):
            assert np.array_equal(got, e), "box %d hits differ (%s)" % (i, name)
        assert np.array_equal(surv[i], s) and fine[i] == out[-1], "box %d survivors differ" % i
    assert sum(len(x) for x in res[:64]) > 10000


def test_c2_knn_sample(c2):
    h, o = c2
    pts = W.knn_points(h.data(), 64, 20260332, h.mesh_sizes, C2["vmin"], C2["vmax"])
    Ks = np.where(np.arange(64) % 2 == 0, 10, 100).astype(np.uint64)
    graph, plain, res, (surv, fine) = _paths(h, "knn", pts, Ks)

    def expect(i):
        r = o.knn(pts[i], int(Ks[i]))
        return r.refine_knn(pts[i], int(Ks[i])), r.levels()

    exp = _pool_map(expect, range(len(pts)))
    for i in range(len(pts)):
        (eid, ed2), (_, s, out) = exp[i]
        for (gid, gd2), name in ((graph[i], "graph"), (plain[i], "plain"), (res[i], "statistics")):
            assert np.array_equal(gid, eid), "query %d ids differ (%s)" % (i, name)
            assert np.array_equal(gd2.view(np.uint64), ed2.view(np.uint64)), "query %d d2 differ (%s)" % (i, name)
        assert np.array_equal(surv[i], s) and fine[i] == out[-1], "query %d survivors differ" % i
