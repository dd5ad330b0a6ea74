"""GPU tests of the continuous-batching serving loop: every request's result is
the same alone and in a concurrent mix (and equals the oracle's), admission by
the HBM budget defers requests without losing any, and batches that exceed the
engine capacities are split."""
import random
import threading

import numpy as np
import pytest

import oracle_py as O
import tblaeq
from tblaeq import workload as W

pytestmark = pytest.mark.gpu

CFG = dict(N=40000, D=6, vmin=1.0, vmax=100.0, is_int=False, ratios=(30, 10, 8), seed=31, sigma_div=2.5)


@pytest.fixture(scope="module")
def idx():
    h = tblaeq.Handler.random(CFG["N"], CFG["D"], CFG["vmin"], CFG["vmax"], CFG["is_int"], CFG["ratios"],
                              CFG["seed"], CFG["sigma_div"])
    o = O.Index.generate(CFG["N"], CFG["D"], CFG["vmin"], CFG["vmax"], CFG["is_int"], CFG["ratios"], CFG["seed"],
                         CFG["sigma_div"])
    h.prepare("L0")
    yield h, o
    h.close()


def _requests(h, n, seed):
    data = h.data()
    boxes, _ = W.range_boxes(data, n, seed, 1e-4, 2e-2)
    pts = W.knn_points(data, n, seed + 1, h.mesh_sizes, CFG["vmin"], CFG["vmax"], noise=0.3)
    Ks = np.random.default_rng(seed).choice([0, 1, 5, 10, 40, 120, 256, 1000], n)
    reqs = [("range", b, 0) for b in boxes] + [("knn", p, int(k)) for p, k in zip(pts, Ks)]
    random.Random(seed).shuffle(reqs)
    return reqs


def _alone(srv, reqs):
    out = []
    for t, q, k in reqs:
        tk = srv.submit(t, q, k)
        srv.wait(tk, 60.0)
        out.append(srv.result(tk))
    return out


def _mixed(srv, reqs, threads=4):
    tickets = [None] * len(reqs)

    def feed(part):
        rng = random.Random(part)
        for i in range(part, len(reqs), threads):
            t, q, k = reqs[i]
            tickets[i] = srv.submit(t, q, k)
            if rng.random() < 0.3:
                threading.Event().wait(rng.expovariate(20000.0))  # Poisson-like gaps

    ths = [threading.Thread(target=feed, args=(p,)) for p in range(threads)]
    for th in ths:
        th.start()
    for th in ths:
        th.join()
    out = []
    for tk in tickets:
        srv.wait(tk, 60.0)
        out.append(srv.result(tk))
    return out


def _same(a, b):
    return all(np.array_equal(x[0], y[0]) and np.array_equal(x[1].view(np.uint64), y[1].view(np.uint64))
               for x, y in zip(a, b))


def test_serve_same_alone_and_in_mix(idx):
    h, o = idx
    reqs = _requests(h, 80, 5)
    srv = tblaeq.Server(h, max_batch=64)
    try:
        alone = _alone(srv, reqs)
        srv.reset_stats()
        mixed = _mixed(srv, reqs)
        st = srv.stats()
    finally:
        srv.close()
    assert _same(alone, mixed)
    assert st["completed"] == len(reqs) and st["mean_batch"] > 1.0
    assert st["p99_us"] >= st["p50_us"] > 0
    D = h.D
    for (t, q, k), (ids, d2) in list(zip(reqs, mixed))[:40]:
        if t == "range":
            assert np.array_equal(ids, o.range(q[:D], q[D:]).refine_range(q[:D], q[D:]))
        else:
            eid, ed2 = o.knn(q, k).refine_knn(q, k)
            assert np.array_equal(ids, eid) and np.array_equal(d2.view(np.uint64), ed2.view(np.uint64))


def test_serve_admission_defers_by_hbm_budget(idx):
    """Each range request is estimated at 1/3 of the budget: at most three are in flight, the rest
    wait in the queue and are admitted as batches complete."""
    h, o = idx
    reqs = [r for r in _requests(h, 60, 9) if r[0] == "range"]
    probe = tblaeq.Server(h, max_batch=64)
    budget = probe.stats()["budget_bytes"]
    probe.close()
    srv = tblaeq.Server(h, max_batch=64, range_request_bytes=int(budget // 3))
    try:
        mixed = _mixed(srv, reqs, threads=2)
        st = srv.stats()
    finally:
        srv.close()
    assert st["completed"] == len(reqs)
    assert st["deferred"] > 0 and st["peak_inflight_bytes"] <= st["budget_bytes"] + budget // 3
    D = h.D
    for (t, q, k), (ids, _) in zip(reqs, mixed):
        assert np.array_equal(ids, o.range(q[:D], q[D:]).refine_range(q[:D], q[D:]))


def test_serve_splits_batches_over_capacity(idx):
    """Capacities far below a full batch's needs: overflowing batches are split and rerun, the
    workers' later batches start at the size that fit, and every answer equals the roomy
    server's."""
    h, o = idx
    reqs = _requests(h, 40, 13)
    srv = tblaeq.Server(h, max_batch=64, pair_cap=3000, hit_cap=4000)
    try:
        mixed = _mixed(srv, reqs, threads=2)
        again = _mixed(srv, reqs, threads=2)
        reruns = srv.kind_stats("range")["reruns"] + srv.kind_stats("knn")["reruns"]
    finally:
        srv.close()
    big = tblaeq.Server(h, max_batch=64)
    try:
        alone = _alone(big, reqs)
    finally:
        big.close()
    assert reruns > 0
    assert _same(alone, mixed)
    assert _same(alone, again)
