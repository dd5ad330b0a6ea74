"""Batched range, KNN and learned KNN at D = 4, 8 and 16: the dimensions with exact-D kernel
instantiations (the range prune and filter templates and the KNN D == DM branch), D = 8 being
the dimension of the headline measurement.  Every path's answers equal the oracle's."""
import numpy as np
import pytest

import oracle_cases as OC
import tblaeq
from batch_checks import _OPEN
from batch_checks import check_knn_batch, check_range_batch, make
from tblaeq import learn
from tblaeq import workload as W

pytestmark = pytest.mark.gpu

DIMS = (4, 8, 16)


def _cfg(D):
    return dict(N=40000, D=D, vmin=1.0, vmax=100.0, is_int=False, ratios=(25, 10, 8), seed=5, sigma_div=2.2)


@pytest.fixture(scope="module", params=DIMS, ids=["D%d" % d for d in DIMS])
def idx(request):
    h, o = make(_cfg(request.param), max_q=1024)
    yield h, o
    _OPEN.remove(h)
    h.close()


@pytest.mark.parametrize("rowq", [0, 16, 64])
def test_range_dims(idx, rowq):
    """rowq: clusters kept by at most this many queries go to the row-major filter (0: none, 64:
    every cluster of one slice)."""
    h, o = idx
    h.batch_set_param(tblaeq.PARAM_RANGE_ROW_QUERIES, rowq)
    try:
        _range_dims(h, o)
    finally:
        h.batch_set_param(tblaeq.PARAM_RANGE_ROW_QUERIES, 16)


def _range_dims(h, o):
    data = h.data()
    boxes, _ = W.range_boxes(data, 200, 40 + h.D, 1e-4, 5e-2)
    _, pts = OC.points_outside_radius(o, limit=64)
    assert len(pts) >= 32
    D = h.D
    edges = np.array([
        np.r_[np.full(D, -1e9), np.full(D, 1e9)],   # everything
        np.r_[np.full(D, 60.0), np.full(D, 40.0)],  # inverted
        np.r_[data[3], data[3]],                    # degenerate on a data point
        np.r_[np.full(D, -np.inf), np.full(D, np.inf)],                    # everything, infinite
        np.r_[np.full(D - 1, -1e9), np.nan, np.full(D, 1e9)],              # NaN lower bound
    ])
    res = check_range_batch(h, o, np.concatenate([boxes, OC.degenerate_boxes(pts), edges]))
    assert sum(len(x) for x in res[:200]) > 0
    assert len(res[-5]) == h.N and len(res[-2]) == h.N and len(res[-1]) == 0


@pytest.mark.parametrize("rowq,leaf", [(0, 0), (1, 0), (4, 1), (0, 1)])
def test_knn_dims(idx, rowq, leaf):
    """rowq: clusters refined by at most this many queries go to the row-major refine (0: none);
    leaf: the other clusters' work items go through per (query, kd leaf) box tests too."""
    h, o = idx
    h.batch_set_param(tblaeq.PARAM_KNN_ROW_QUERIES, rowq)
    h.batch_set_param(tblaeq.PARAM_KNN_LEAF_REFINE, leaf)
    try:
        pts = W.knn_points(h.data(), 120, 50 + h.D, h.mesh_sizes, 1.0, 100.0, noise=0.2)
        Ks = np.random.default_rng(h.D).choice([0, 1, 10, 64, 100, 200], len(pts))
        check_knn_batch(h, o, pts, Ks)
    finally:
        h.batch_set_param(tblaeq.PARAM_KNN_ROW_QUERIES, 4)
        h.batch_set_param(tblaeq.PARAM_KNN_LEAF_REFINE, 0)


def test_knn_query_major_dims(idx):
    h, o = idx
    h.batch_set_param(tblaeq.PARAM_KNN_MODE, 1)
    try:
        pts = W.knn_points(h.data(), 60, 60 + h.D, h.mesh_sizes, 1.0, 100.0, noise=0.1)
        Ks = np.random.default_rng(h.D + 1).choice([1, 10, 100], len(pts))
        check_knn_batch(h, o, pts, Ks)
    finally:
        h.batch_set_param(tblaeq.PARAM_KNN_MODE, 0)


def test_learned_knn_dims(idx):
    """Learned radii (a briefly trained network) with certification and fallback: answers equal
    the seed method's and the oracle's at every D."""
    h, o = idx
    mlp = tblaeq.Mlp(h, seed=13)
    try:
        mlp.set_norm(1.0, 99.0)
        X, T = learn.traces(h, 2048, 70 + h.D, 1.0, 100.0, chunk=1024)
        learn.fit(mlp, X, T, epochs=10, batch=256, lr=3e-3)
        pts = W.knn_points(h.data(), 96, 80 + h.D, h.mesh_sizes, 1.0, 100.0)
        Ks = np.random.default_rng(9).choice([10, 100], len(pts)).astype(np.uint64)
        h.batch_set("knn", pts, Ks)
        h.batch_run()
        seed_res = h.batch_results()
        h.batch_set("knn", pts, Ks)
        fb = h.batch_knn_learned(mlp, factor=1.15)
        res = h.batch_results()
        assert fb < len(pts)
        for i, q in enumerate(pts):
            eid, ed2 = o.knn(q, int(Ks[i])).refine_knn(q, int(Ks[i]))
            for gid, gd2 in (seed_res[i], res[i]):
                assert np.array_equal(gid, eid), "query %d ids differ" % i
                assert np.array_equal(gd2.view(np.uint64), ed2.view(np.uint64)), "query %d distances differ" % i
    finally:
        mlp.close()
