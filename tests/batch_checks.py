"""Shared checks of the batched engine against the CPU oracle (test infrastructure, not
collected by pytest): every result path of a batch -- captured graph, plain launches, plain
launches with statistics -- is compared with the oracle's exact refine of the reference
traversal, and the per-level survivor counts and candidate-superset sizes with the oracle's."""
import numpy as np

import oracle_py as O
import tblaeq

_OPEN = []


def make(cfg, max_q=512, pair_cap=1 << 22, hit_cap=1 << 22):
    h = tblaeq.Handler.random(cfg["N"], cfg["D"], cfg["vmin"], cfg["vmax"], cfg["is_int"], cfg["ratios"], cfg["seed"],
                              cfg["sigma_div"])
    _OPEN.append(h)
    o = O.Index.generate(cfg["N"], cfg["D"], cfg["vmin"], cfg["vmax"], cfg["is_int"], cfg["ratios"], cfg["seed"],
                         cfg["sigma_div"])
    h.prepare("L0")
    h.batch_init(max_q, pair_cap, hit_cap)
    return h, o


def close_new(n):
    """Destroys the handlers made since len(_OPEN) was n."""
    while len(_OPEN) > n:
        _OPEN.pop().close()


def check_range_batch(h, o, boxes):
    """Results of the three range paths against the oracle: the captured graph without
    statistics (the default, timed path), plain launches without statistics, and plain
    launches with statistics (per-level survivors and candidate-superset sizes too)."""
    h.batch_set("range", boxes)
    h.batch_run()
    graph = h.batch_results()
    h.batch_use_graph(False)
    try:
        h.batch_run()
        plain = h.batch_results()
    finally:
        h.batch_use_graph(True)
    h.batch_run(stats=True)
    res = h.batch_results()
    surv, fine = h.batch_stats()
    D = h.D
    for i, row in enumerate(boxes):
        lo, hi = row[:D], row[D:]
        r = o.range(lo, hi)
        exp = r.refine_range(lo, hi)
        assert np.array_equal(graph[i], exp), "query %d hits differ (graph, no statistics)" % i
        assert np.array_equal(plain[i], exp), "query %d hits differ (no statistics)" % i
        assert np.array_equal(res[i], exp), "query %d hits differ" % i
        _, s, out = r.levels()
        assert np.array_equal(surv[i], s), "query %d survivors per level differ" % i
        assert fine[i] == out[-1]
    return res


def check_knn_batch(h, o, pts, Ks):
    """Results of the captured graph and of plain launches without statistics (the timed
    path) and with them, against the oracle."""
    h.batch_set("knn", pts, Ks)
    h.batch_run()
    graph = h.batch_results()
    h.batch_use_graph(False)
    try:
        h.batch_run()
        plain = h.batch_results()
    finally:
        h.batch_use_graph(True)
    h.batch_run(stats=True)
    res = h.batch_results()
    surv, fine = h.batch_stats()
    for i, q in enumerate(pts):
        K = int(Ks[i])
        r = o.knn(q, K)
        eid, ed2 = r.refine_knn(q, K)
        for gid, gd2 in (graph[i], plain[i], res[i]):
            assert np.array_equal(gid, eid), "query %d ids differ" % i
            assert np.array_equal(gd2.view(np.uint64), ed2.view(np.uint64)), "query %d distances differ" % i
        _, s, out = r.levels()
        assert np.array_equal(surv[i], s), "query %d survivors per level differ" % i
        assert fine[i] == out[-1]
