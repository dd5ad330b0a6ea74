"""GPU parity of the batched engine: per-query results equal the oracle's exact
refine of the reference traversal (range: sorted id sets; KNN: top-K ids and
squared distances bit-exact), and the per-level survivor counts and candidate
superset sizes equal the oracle's."""
import numpy as np
import pytest

import oracle_cases as OC
import tblaeq
from batch_checks import _OPEN
from batch_checks import check_knn_batch as _check_knn_batch
from batch_checks import check_range_batch as _check_range_batch
from batch_checks import make as _make
from tblaeq import workload as W

pytestmark = pytest.mark.gpu

CFG = dict(N=60000, D=6, vmin=1.0, vmax=100.0, is_int=False, ratios=(30, 10, 8), seed=21, sigma_div=2.5)


@pytest.fixture(scope="module")
def idx():
    h, o = _make(CFG)
    yield h, o
    _OPEN.remove(h)
    h.close()


def test_batch_range_matches_oracle(idx):
    h, o = idx
    data = h.data()
    boxes, _ = W.range_boxes(data, 300, 5, 1e-4, 5e-2)
    res = _check_range_batch(h, o, boxes)
    assert sum(len(x) for x in res) > 0


def test_batch_range_edges(idx):
    h, o = idx
    D = h.D
    boxes = np.array([
        np.r_[np.full(D, -1e6), np.full(D, -1e6 + 1)],   # nothing
        np.r_[np.full(D, -1e9), np.full(D, 1e9)],         # everything
        np.r_[np.full(D, 60.0), np.full(D, 40.0)],        # inverted
        np.r_[h.data()[7], h.data()[7]],                  # degenerate on a point
        np.r_[np.full(D, -np.inf), np.full(D, np.inf)],   # everything, infinite bounds
        np.r_[np.nan, np.full(D - 1, -1e9), np.full(D, 1e9)],          # NaN lower bound
        np.r_[np.full(D, -1e9), 1e9, np.nan, np.full(D - 2, 1e9)],     # NaN upper bound
    ])
    res = _check_range_batch(h, o, boxes)
    assert len(res[1]) == h.N and len(res[0]) == 0
    assert len(res[4]) == h.N and len(res[5]) == 0 and len(res[6]) == 0


def test_batch_rejects_empty_batch(idx):
    h, _ = idx
    with pytest.raises(tblaeq.TblaeqError):
        h.batch_set("range", np.zeros((0, 2 * h.D)))
    with pytest.raises(tblaeq.TblaeqError):
        h.batch_set("knn", np.zeros((0, h.D)), np.zeros(0, np.uint64))


def test_batch_range_points_outside_stored_radius(idx):
    """Degenerate boxes on reconstructed finest points that lie outside their last-level
    cluster's stored radius (radii come from the original coordinates): the reference's ball
    test drops the cluster, so none of these boxes may return the point, on any path."""
    h, o = idx
    ids, pts = OC.points_outside_radius(o, limit=256)
    assert len(ids) >= 100
    boxes = OC.degenerate_boxes(pts)
    res = _check_range_batch(h, o, boxes)
    assert sum(len(x) for x in res) == 0


def test_batch_range_integer_grid():
    """Integer index: degenerate boxes on points outside their cluster's radius and boxes with
    integer corners (points on the box faces, equal coordinates everywhere)."""
    h, o = _make(dict(CFG, N=20000, is_int=True, seed=8), max_q=512)
    _, pts = OC.points_outside_radius(o, limit=128)
    assert len(pts) >= 50
    boxes = np.concatenate([OC.degenerate_boxes(pts), OC.integer_boxes(h.data(), 200, 3)])
    res = _check_range_batch(h, o, boxes)
    assert sum(len(x) for x in res) > 0


@pytest.mark.parametrize("blocks,stage,radix,rowq", [(1, 256, 1, 0), (2, 1024, 32, 64), (64, 512, 2, 1),
                                                     (7, 256, 5, 8), (16, 1024, 4, 16)])
def test_batch_range_launch_params_vs_oracle(blocks, stage, radix, rowq):
    """Range launch parameters (prune and filter grid, LDS hit stage, radix grid, the query count
    up to which a cluster is filtered row-major) change only how the work is spread: every
    setting's answers equal the oracle's."""
    h, o = _make(dict(CFG, N=30000, seed=19), max_q=256)
    h.batch_set_param(tblaeq.PARAM_RANGE_BLOCKS_PER_CU, blocks)
    h.batch_set_param(tblaeq.PARAM_HIT_STAGE, stage)
    h.batch_set_param(tblaeq.PARAM_RADIX_BLOCKS_PER_CU, radix)
    h.batch_set_param(tblaeq.PARAM_RANGE_ROW_QUERIES, rowq)
    assert h.batch_get_param(tblaeq.PARAM_RANGE_ROW_QUERIES) == rowq
    data = h.data()
    boxes, _ = W.range_boxes(data, 160, 23, 1e-4, 0.2)
    _, pts = OC.points_outside_radius(o, limit=40)
    _check_range_batch(h, o, np.concatenate([boxes, OC.degenerate_boxes(pts)]))


def test_batch_range_single_and_uneven(idx):
    h, o = idx
    data = h.data()
    boxes, _ = W.range_boxes(data, 1, 9, 1e-2, 1e-2)
    _check_range_batch(h, o, boxes)
    # very uneven selectivities in one batch
    b1, _ = W.range_boxes(data, 64, 10, 1e-5, 1e-5)
    b2, _ = W.range_boxes(data, 64, 11, 0.2, 0.5)
    _check_range_batch(h, o, np.concatenate([b1, b2]))


def test_batch_range_overflow_splits():
    """Capacities far below the batch's needs: the driver splits and still matches."""
    h, o = _make(dict(CFG, N=20000, seed=3), max_q=128, pair_cap=3000, hit_cap=40000)
    data = h.data()
    boxes, _ = W.range_boxes(data, 96, 12, 1e-3, 2e-2)
    _check_range_batch(h, o, boxes)


@pytest.mark.parametrize("ratios", [(40,), (30, 10), (20, 8, 4, 3)])
def test_batch_range_hierarchy_heights(ratios):
    """Range batches over one level (the final filter reads the coarsest clusters' children), two and
    four levels."""
    h, o = _make(dict(CFG, N=24000, ratios=ratios, seed=14), max_q=96)
    data = h.data()
    boxes, _ = W.range_boxes(data, 80, 15, 1e-4, 5e-2)
    _check_range_batch(h, o, boxes)


@pytest.mark.parametrize("K", [1, 10, 100])
def test_batch_knn_matches_oracle(idx, K):
    h, o = idx
    pts = W.knn_points(h.data(), 64, 20 + K, h.mesh_sizes, CFG["vmin"], CFG["vmax"])
    _check_knn_batch(h, o, pts, np.full(64, K))


def test_batch_knn_mixed_k(idx):
    h, o = idx
    pts = W.knn_points(h.data(), 97, 31, h.mesh_sizes, CFG["vmin"], CFG["vmax"], noise=0.5)
    Ks = np.random.default_rng(4).choice([0, 1, 5, 10, 37, 100, 200, 256, 1000], 97)
    assert (Ks == 1000).any() and (Ks == 256).any()
    _check_knn_batch(h, o, pts, Ks)


def test_batch_knn_large_k_only(idx):
    """Every query above the batched kernels' K limit: the whole batch runs per query."""
    h, o = idx
    pts = W.knn_points(h.data(), 12, 33, h.mesh_sizes, CFG["vmin"], CFG["vmax"], noise=0.5)
    _check_knn_batch(h, o, pts, np.full(12, 2000))


def test_batch_knn_far_and_ties():
    """Integer grid (many equal lower bounds and distances) and far-away queries."""
    h, o = _make(dict(CFG, N=20000, is_int=True, seed=8), max_q=64)
    data = h.data()
    pts = np.concatenate([data[:8], data[100:108] + 0.5, np.full((4, h.D), 1e4), np.full((4, h.D), -50.0)])
    _check_knn_batch(h, o, pts, np.full(len(pts), 25))


def test_batch_matches_plain_c1():
    cfg = W.C1
    h, o = _make(cfg, max_q=256, pair_cap=1 << 24, hit_cap=1 << 24)
    data = h.data()
    boxes, _ = W.range_boxes(data, 100, 12345, 1e-3, 1e-3)
    h.batch_set("range", boxes)
    h.batch_run()
    res = h.batch_results()
    plain = h.query("range", boxes, refine=True)
    for i in range(len(boxes)):
        assert np.array_equal(res[i], plain.refined(i)[0])
    pts = W.knn_points(data, 100, 12346, h.mesh_sizes, cfg["vmin"], cfg["vmax"])
    h.batch_set("knn", pts, np.full(100, 10))
    h.batch_run()
    kres = h.batch_results()
    kplain = h.query("knn", pts, K=10, refine=True)
    for i in range(len(pts)):
        pid, pd2 = kplain.refined(i)
        assert np.array_equal(kres[i][0], pid)
        assert np.array_equal(kres[i][1].view(np.uint64), pd2.view(np.uint64))


@pytest.mark.parametrize("seed,wg,rowq,leaf", [(1, 1, 0, 1), (7, 16, 1, 0), (256, 2, 4, 1), (64, 4, 2, 0)])
def test_batch_knn_launch_params(seed, wg, rowq, leaf):
    """Seed cluster count, grid shape, the row-major refine's query limit (rowq: clusters
    refined by at most this many queries go row-major) and the leaf-tested refine of the other
    clusters change only the work done, never the answer."""
    h, o = _make(dict(CFG, N=30000, is_int=True, seed=9), max_q=128)
    h.batch_set_param(tblaeq.PARAM_KNN_SEED, seed)
    h.batch_set_param(tblaeq.PARAM_KNN_WG_PER_CU, wg)
    h.batch_set_param(tblaeq.PARAM_KNN_ROW_QUERIES, rowq)
    h.batch_set_param(tblaeq.PARAM_KNN_LEAF_REFINE, leaf)
    assert h.batch_get_param(tblaeq.PARAM_KNN_ROW_QUERIES) == rowq
    assert h.batch_get_param(tblaeq.PARAM_KNN_LEAF_REFINE) == leaf
    data = h.data()
    pts = np.concatenate([data[:40] + 0.5, W.knn_points(data, 40, 77, h.mesh_sizes, CFG["vmin"], CFG["vmax"])])
    Ks = np.random.default_rng(seed).choice([1, 10, 64, 150], len(pts))
    _check_knn_batch(h, o, pts, Ks)


def test_batch_knn_row_refine_takes_sparse_clusters():
    """Few spread-out queries: most last-level clusters are refined by one to four of them and
    go to the row-major refine, whose distance evaluations (only the rows of kd leaves that the
    pair's own threshold cannot rule out) differ from the query-major count; answers equal the
    oracle's either way."""
    h, o = _make(dict(CFG, N=60000, seed=31), max_q=64)
    pts = W.knn_points(h.data(), 24, 33, h.mesh_sizes, CFG["vmin"], CFG["vmax"], noise=0.5)
    Ks = np.random.default_rng(34).choice([1, 10, 100], len(pts)).astype(np.uint64)
    work = {}
    for rowq in (0, 4):
        h.batch_set_param(tblaeq.PARAM_KNN_ROW_QUERIES, rowq)
        _check_knn_batch(h, o, pts, Ks)
        h.batch_set("knn", pts, Ks)
        h.batch_use_graph(False)
        try:
            h.batch_run()
        finally:
            h.batch_use_graph(True)
        work[rowq] = h.batch_work()[0]
    assert work[0] > 0 and work[4] > 0
    assert work[4] != work[0]


def test_batch_param_rejects_bad_values(idx):
    h, _ = idx
    with pytest.raises(tblaeq.TblaeqError):
        h.batch_set_param(tblaeq.PARAM_KNN_SEED, 0)
    with pytest.raises(tblaeq.TblaeqError):
        h.batch_set_param(tblaeq.PARAM_KNN_WG_PER_CU, 17)
    with pytest.raises(tblaeq.TblaeqError):
        h.batch_set_param(tblaeq.PARAM_KNN_ROW_QUERIES, 5)
    with pytest.raises(tblaeq.TblaeqError):
        h.batch_set_param(99, 1)


def test_batch_knn_query_major_mode():
    """The query-major refine (every workgroup expands its own clusters) gives the same answers."""
    h, o = _make(dict(CFG, N=30000, seed=12), max_q=128)
    h.batch_set_param(tblaeq.PARAM_KNN_MODE, 1)
    data = h.data()
    pts = W.knn_points(data, 70, 5, h.mesh_sizes, CFG["vmin"], CFG["vmax"], noise=0.3)
    Ks = np.random.default_rng(6).choice([0, 1, 10, 100], len(pts))
    _check_knn_batch(h, o, pts, Ks)


def test_batch_knn_overflow_splits():
    """Pair and record capacities far below the batch's needs: the driver splits and still matches."""
    h, o = _make(dict(CFG, N=20000, seed=4), max_q=128, pair_cap=1500, hit_cap=1500)
    data = h.data()
    pts = W.knn_points(data, 48, 8, h.mesh_sizes, CFG["vmin"], CFG["vmax"], noise=0.2)
    _check_knn_batch(h, o, pts, np.full(len(pts), 20))


@pytest.mark.parametrize("ratios", [(40,), (30, 10), (20, 8, 4, 3)])
def test_batch_knn_hierarchy_heights(ratios):
    """One level (no parent level to split on), two levels (the split parent is the coarsest level)
    and four levels."""
    h, o = _make(dict(CFG, N=24000, ratios=ratios, seed=13), max_q=96)
    data = h.data()
    pts = W.knn_points(data, 60, 17, h.mesh_sizes, CFG["vmin"], CFG["vmax"], noise=0.2)
    Ks = np.random.default_rng(3).choice([0, 1, 7, 50, 120], len(pts))
    _check_knn_batch(h, o, pts, Ks)


def test_autotune_keeps_results_and_params_roundtrip(tmp_path):
    """The tuner only changes launch parameters: answers after tuning (graph and plain launches) still
    equal the oracle's, and a saved tuning table restores the chosen values."""
    h, o = _make(dict(CFG, N=30000, seed=17), max_q=128)
    data = h.data()
    boxes, _ = W.range_boxes(data, 64, 3, 1e-4, 2e-2)
    pts = W.knn_points(data, 64, 4, h.mesh_sizes, CFG["vmin"], CFG["vmax"])
    Ks = np.random.default_rng(2).choice([1, 10, 60], len(pts)).astype(np.uint64)
    ms = h.batch_autotune(boxes, pts, Ks, reps=1)
    assert ms["range_ms"][1] <= ms["range_ms"][0] and ms["knn_ms"][1] <= ms["knn_ms"][0]
    params = {p: h.batch_get_param(p) for p in range(1, 11)}
    assert params[tblaeq.PARAM_HIT_STAGE] in (256, 512, 1024)
    _check_range_batch(h, o, boxes)
    h.batch_set("range", boxes)
    h.batch_run()  # captured graph with the tuned parameters
    graph_res = h.batch_results()
    for i, row in enumerate(boxes):
        assert np.array_equal(graph_res[i], o.range(row[:h.D], row[h.D:]).refine_range(row[:h.D], row[h.D:]))
    _check_knn_batch(h, o, pts, Ks)
    path = tmp_path / "tuning.txt"
    h.batch_save_params(path)
    h.batch_set_param(tblaeq.PARAM_HIT_STAGE, 256 if params[tblaeq.PARAM_HIT_STAGE] != 256 else 1024)
    h.batch_set_param(tblaeq.PARAM_KNN_SEED, 3)
    h.batch_load_params(path)
    assert {p: h.batch_get_param(p) for p in range(1, 11)} == params
    bad = tmp_path / "bad.txt"
    bad.write_text("hit_stage 300\n")
    with pytest.raises(tblaeq.TblaeqError):
        h.batch_load_params(bad)


@pytest.mark.parametrize("mode", [0, 1])
def test_batch_knn_many_workgroups(mode):
    """4096 KNN queries in one batch -- over a thousand query-major workgroups, several resident
    per CU, in both refine modes.  Each workgroup's LDS queue counters must start from zero (a
    counter not reset per workgroup once made a workgroup read past its queue); every answer of
    the captured graph and of plain launches equals the oracle's."""
    from concurrent.futures import ThreadPoolExecutor

    h, o = _make(dict(CFG, N=60000, seed=23), max_q=4096, pair_cap=1 << 24, hit_cap=1 << 24)
    h.batch_set_param(tblaeq.PARAM_KNN_MODE, mode)
    n = 4096
    pts = W.knn_points(h.data(), n, 40 + mode, h.mesh_sizes, CFG["vmin"], CFG["vmax"], noise=0.2)
    Ks = np.random.default_rng(41 + mode).choice([1, 10, 100, 200], n).astype(np.uint64)
    h.batch_set("knn", pts, Ks)
    h.batch_run()
    graph = h.batch_results()
    h.batch_use_graph(False)
    try:
        h.batch_run()
        plain = h.batch_results()
    finally:
        h.batch_use_graph(True)

    def expect(i):
        return o.knn(pts[i], int(Ks[i])).refine_knn(pts[i], int(Ks[i]))

    with ThreadPoolExecutor(16) as ex:
        exp = list(ex.map(expect, range(n)))
    for i in range(n):
        eid, ed2 = exp[i]
        for gid, gd2 in (graph[i], plain[i]):
            assert np.array_equal(gid, eid), "query %d ids differ" % i
            assert np.array_equal(gd2.view(np.uint64), ed2.view(np.uint64)), "query %d distances differ" % i


@pytest.mark.parametrize("caps", [(1 << 22, 1 << 22, 1 << 22, 1 << 22, (1, 10, 50, 200)),
                                  (3000, 40000, 1500, 1500, (1, 10, 20))])
def test_batch_slots_run_concurrently(caps):
    """A range batch in slot 0 and a KNN batch in slot 1 replayed concurrently on two streams:
    every answer equals the oracle's, also when small capacities force both slots back onto the
    splitting path and when the KNN batch holds K above the batched limit (per-query route)."""
    from concurrent.futures import ThreadPoolExecutor

    rp, rh, kp, kh, kchoice = caps
    h, o = _make(dict(CFG, N=40000, seed=25), max_q=512, pair_cap=rp, hit_cap=rh)
    try:
        h.batch_slot(1)
        h.batch_init(512, kp, kh)
        data = h.data()
        boxes, _ = W.range_boxes(data, 200, 26, 1e-4, 2e-2)
        pts = W.knn_points(data, 150, 27, h.mesh_sizes, CFG["vmin"], CFG["vmax"], noise=0.3)
        for Ks in (np.random.default_rng(28).choice(kchoice, 150),
                   np.random.default_rng(29).choice([5, 300], 150)):
            h.batch_slot(0)
            h.batch_set("range", boxes)
            h.batch_slot(1)
            h.batch_set("knn", pts, Ks)
            for _ in range(2):  # capture, then replay
                hits, nn = h.batch_run_slots((0, 1))
            h.batch_slot(0)
            rres = h.batch_results()
            h.batch_slot(1)
            kres = h.batch_results()
            D = h.D

            def exp_range(i):
                lo, hi = boxes[i][:D], boxes[i][D:]
                return o.range(lo, hi).refine_range(lo, hi)

            def exp_knn(i):
                return o.knn(pts[i], int(Ks[i])).refine_knn(pts[i], int(Ks[i]))

            with ThreadPoolExecutor(16) as ex:
                er = list(ex.map(exp_range, range(len(boxes))))
                ek = list(ex.map(exp_knn, range(len(pts))))
            assert hits == sum(len(x) for x in er) and nn == sum(len(x[0]) for x in ek)
            for i in range(len(boxes)):
                assert np.array_equal(rres[i], er[i]), "box %d" % i
            for i in range(len(pts)):
                assert np.array_equal(kres[i][0], ek[i][0]), "point %d ids" % i
                assert np.array_equal(kres[i][1].view(np.uint64), ek[i][1].view(np.uint64)), "point %d d2" % i
    finally:
        h.batch_slot(0)
        _OPEN.remove(h)
        h.close()
