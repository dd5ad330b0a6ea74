"""GPU parity of the plain per-query path (QueryHandler -> QueryEngine -> L0..L3)
against the CPU oracle: candidate grids bit-exact (ids, FP64 values and order),
per-level sizes equal, exact refine equal."""
import numpy as np
import pytest

import oracle_py as O
import tblaeq

pytestmark = pytest.mark.gpu

CFG_SMALL = dict(N=20000, D=4, vmin=1.0, vmax=100.0, is_int=False, ratios=(20, 10, 5), seed=7, sigma_div=3.0)
CFG_C1 = dict(N=500000, D=5, vmin=1.0, vmax=100.0, is_int=False, ratios=(100, 50, 20), seed=12345, sigma_div=3.0)


def _pair(cfg, policy="L0"):
    h = tblaeq.Handler.random(cfg["N"], cfg["D"], cfg["vmin"], cfg["vmax"], cfg["is_int"], cfg["ratios"], cfg["seed"],
                              cfg["sigma_div"])
    o = O.Index.generate(cfg["N"], cfg["D"], cfg["vmin"], cfg["vmax"], cfg["is_int"], cfg["ratios"], cfg["seed"],
                         cfg["sigma_div"])
    h.prepare(policy)
    return h, o


@pytest.fixture(scope="module")
def small():
    return _pair(CFG_SMALL)


@pytest.fixture(scope="module")
def c1():
    return _pair(CFG_C1)


def _boxes(data, n, seed, scale):
    rng = np.random.default_rng(seed)
    std = data.std(0)
    out = []
    for _ in range(n):
        c = data[rng.integers(data.shape[0])]
        h = scale * rng.uniform(0.2, 1.0)
        out.append((c - h * std, c + h * std))
    return out


def _same_bits(a, b):
    return a.shape == b.shape and np.array_equal(np.ascontiguousarray(a).view(np.uint64),
                                                 np.ascontiguousarray(b).view(np.uint64))


def _check_range(h, o, lo, hi):
    r = h.range(lo, hi)
    ro = o.range(lo, hi)
    ids, vals = r.fine_mesh()
    oids, ovals = ro.grid()
    assert np.array_equal(ids, oids), "candidate ids/order differ"
    assert _same_bits(vals, ovals), "candidate coordinates differ"
    _, surv, out = ro.levels()
    assert np.array_equal(r.info(0)["levels"], out)
    assert np.array_equal(r.levels(0)[0], surv)
    assert np.array_equal(r.refined()[0], ro.refine_range(lo, hi))
    return ids.size


def _check_knn(h, o, q, K):
    r = h.knn(q, K)
    ro = o.knn(q, K)
    ids, vals = r.fine_mesh()
    oids, ovals = ro.grid()
    assert np.array_equal(ids, oids), "candidate ids/order differ"
    assert _same_bits(vals, ovals), "candidate coordinates differ"
    _, surv, out = ro.levels()
    assert np.array_equal(r.info(0)["levels"], out)
    assert np.array_equal(r.levels(0)[0], surv)
    rid, rd2 = r.refined()
    eid, ed2 = ro.refine_knn(q, K)
    assert np.array_equal(rid, eid)
    assert _same_bits(rd2, ed2)
    return ids.size


def test_range_small(small):
    h, o = small
    data = h.data()
    sizes = [_check_range(h, o, lo, hi) for lo, hi in _boxes(data, 12, 1, 0.6)]
    assert max(sizes) > 0


def test_range_c1(c1):
    h, o = c1
    data = h.data()
    sizes = [_check_range(h, o, lo, hi) for lo, hi in _boxes(data, 8, 2, 0.4)]
    assert max(sizes) > 1000


def test_range_edges(small):
    h, o = small
    D = h.D
    # nothing survives the first level
    assert _check_range(h, o, np.full(D, -1e6), np.full(D, -1e6 + 1)) == 0
    # inverted box (lo > hi): only clusters whose sphere reaches the empty box
    _check_range(h, o, np.full(D, 60.0), np.full(D, 40.0))
    # the whole domain: every point is a candidate
    assert _check_range(h, o, np.full(D, -1e9), np.full(D, 1e9)) == h.N
    # degenerate box on a data point
    p = h.data()[123]
    _check_range(h, o, p, p)
    # infinite bounds: every point; a NaN bound: candidates as the reference's
    # conditional distance keeps them, no refined hit
    assert _check_range(h, o, np.full(D, -np.inf), np.full(D, np.inf)) == h.N
    _check_range(h, o, np.r_[np.nan, np.full(D - 1, -1e9)], np.full(D, 1e9))
    _check_range(h, o, np.full(D, -1e9), np.r_[np.full(D - 1, 1e9), np.nan])


@pytest.mark.parametrize("K", [0, 1, 10, 100, 1000])
def test_knn_small(small, K):
    h, o = small
    data = h.data()
    rng = np.random.default_rng(K + 3)
    for _ in range(4):
        q = data[rng.integers(data.shape[0])] + rng.normal(0, 0.5, h.D)
        _check_knn(h, o, q, K)


def test_knn_k_exceeds_n(small):
    h, o = small
    q = h.data()[5]
    assert _check_knn(h, o, q, h.N + 7) == h.N


@pytest.mark.parametrize("K", [10, 100])
def test_knn_c1(c1, K):
    h, o = c1
    data = h.data()
    rng = np.random.default_rng(100 + K)
    for _ in range(3):
        q = data[rng.integers(data.shape[0])] + rng.normal(0, 0.2, h.D)
        _check_knn(h, o, q, K)


@pytest.mark.parametrize("policy", ["L1", "L2", "L3", ["L0", "L1", "L3"], ["L3", "L2", "L0"]])
def test_policies_match(policy):
    h, o = _pair(CFG_SMALL, policy)
    tblaeq.lib().tblaeq_handler_set_l3_tile_bytes(64 * 1024)
    try:
        data = h.data()
        for lo, hi in _boxes(data, 4, 11, 0.6):
            _check_range(h, o, lo, hi)
        rng = np.random.default_rng(12)
        for K in (1, 10, 50):
            _check_knn(h, o, data[rng.integers(h.N)] + 0.3, K)
    finally:
        tblaeq.lib().tblaeq_handler_set_l3_tile_bytes(0)


def test_integer_grid_ties():
    """isInt generator: many exactly equal coordinates and KNN lower bounds."""
    cfg = dict(CFG_SMALL, is_int=True, seed=99)
    h, o = _pair(cfg)
    data = h.data()
    rng = np.random.default_rng(5)
    for _ in range(4):
        q = data[rng.integers(h.N)].copy()
        _check_knn(h, o, q, 25)
    for lo, hi in _boxes(data, 4, 6, 0.5):
        _check_range(h, o, np.round(lo), np.round(hi))


def test_saved_index_roundtrip(tmp_path):
    """An index written in the reference layout by the oracle loads and queries identically."""
    o = O.Index.generate(5000, 3, 1.0, 50.0, False, (10, 5), 3, 2.0)
    path = str(tmp_path / "idx")
    o.save(path)
    h = tblaeq.Handler.load(path)
    h.prepare("L0")
    data = O.Index.load(path).recon()
    for lo, hi in _boxes(data, 4, 8, 0.7):
        _check_range(h, o, lo, hi)
    _check_knn(h, o, data[17] + 0.1, 10)
