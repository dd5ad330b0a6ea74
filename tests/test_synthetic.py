"""The parallel synthetic hierarchy (tblaeq_handler_create_synthetic): the same index for any
thread count; the generator's shape (child i under parent i % parents, rows of a column
ascending, coordinates in [vmin, vmax], integers when asked); P-tensor values and radii that
reproduce the finest data from the parent coordinates; and, on the GPU, queries over it equal the
oracle's on the index files it saves."""
import filecmp
import os

import numpy as np
import pytest

import oracle_py as O
import tblaeq

ARGS = dict(vmin=1.0, vmax=100.0, ratios=(10, 6, 4), seed=5, sigma_div=2.0)


def _save(h, path):
    h.save(str(path))
    return str(path)


@pytest.mark.parametrize("is_int", [False, True])
def test_thread_count_does_not_change_the_index(tmp_path, is_int):
    a = tblaeq.Handler.synthetic(150_001, 3, is_int=is_int, threads=1, **ARGS)
    b = tblaeq.Handler.synthetic(150_001, 3, is_int=is_int, threads=6, **ARGS)
    assert np.array_equal(a.data().view(np.uint64), b.data().view(np.uint64))
    pa, pb = _save(a, tmp_path / "a"), _save(b, tmp_path / "b")
    for f in sorted(os.listdir(pa)):
        assert filecmp.cmp(os.path.join(pa, f), os.path.join(pb, f), shallow=False), f
    c = tblaeq.Handler.synthetic(150_001, 3, is_int=is_int, threads=3, **dict(ARGS, seed=6))
    assert not np.array_equal(a.data(), c.data())


@pytest.mark.parametrize("is_int", [False, True])
def test_generator_shape(tmp_path, is_int):
    N, D = 40_000, 4
    h = tblaeq.Handler.synthetic(N, D, is_int=is_int, threads=4, **ARGS)
    x = h.data()
    assert h.mesh_sizes == [40000, 4000, 666, 166]
    assert x.min() >= ARGS["vmin"] and x.max() <= ARGS["vmax"]
    if is_int:
        assert np.array_equal(x, np.round(x))
    rows = np.array([0, 17, 39_999])
    assert np.array_equal(h.data_rows(rows), x[rows])
    o = O.Index.load(_save(h, tmp_path / "syn"))
    assert o.mesh_sizes == h.mesh_sizes
    L = o.intervals
    for l in range(L):
        # interval l maps level l (mesh_sizes[L - l] clusters) to level l + 1
        n_par, n_ch = o.mesh_sizes[L - l], o.mesh_sizes[L - l - 1]
        col, mp = o.array(l, 0).astype(np.int64), o.array(l, 3).astype(np.int64)
        counts = np.bincount(np.arange(n_ch) % n_par, minlength=n_par)
        assert np.array_equal(np.diff(col), counts)
        # a stable argsort of the labels i % parents
        assert np.array_equal(mp, np.argsort(np.arange(n_ch) % n_par, kind="stable"))
    # finest values and radii against the parent level's reconstructed coordinates
    par = o.recon_level(L - 1)
    col, mp = o.array(L - 1, 0).astype(np.int64), o.array(L - 1, 3).astype(np.int64)
    vals, rad = o.array(L - 1, 4).reshape(-1, D), o.array(L - 1, 5)
    owner = np.repeat(np.arange(par.shape[0]), np.diff(col))
    np.testing.assert_allclose(vals * par[owner], x[mp], rtol=1e-12)
    dist = np.sqrt(((x[mp] - par[owner]) ** 2).sum(1))
    np.testing.assert_allclose(rad, np.maximum.reduceat(dist, col[:-1]), rtol=1e-9)


@pytest.mark.gpu
def test_queries_over_a_synthetic_index_match_the_oracle(tmp_path):
    h = tblaeq.Handler.synthetic(60_000, 5, threads=8, **ARGS)
    o = O.Index.load(_save(h, tmp_path / "syn"))
    h.prepare("L0")
    x = h.data()
    rng = np.random.default_rng(8)
    for i in range(6):
        c = x[rng.integers(x.shape[0])]
        lo, hi = c - 6.0, c + 6.0
        r = h.range(lo, hi)
        ro = o.range(lo, hi)
        assert np.array_equal(r.fine_mesh()[0], ro.grid()[0]), i
        assert np.array_equal(r.refined()[0], ro.refine_range(lo, hi)), i
        q = c + 0.5
        k = h.knn(q, 12)
        ko = o.knn(q, 12)
        rid, rd2 = k.refined()
        eid, ed2 = ko.refine_knn(q, 12)
        assert np.array_equal(rid, eid) and np.array_equal(rd2.view(np.uint64), ed2.view(np.uint64)), i
        r.close()
        k.close()
