"""Host-side logic of the shipped library on the CPU (no kernels run):
generator and index assembly byte-identical to the oracle's, the on-disk
format round trip, and the memory-policy scheduler."""
import filecmp
import os

import numpy as np
import pytest

import oracle_py as O
import tblaeq


@pytest.mark.parametrize("args", [
    (20000, 4, 1.0, 100.0, False, (20, 10, 5), 7, 3.0),
    (5000, 5, 1.0, 50.0, True, (10, 5), 99, 2.2),
    (3000, 2, 2.0, 9.0, False, (50, 100), 1, 1.0),     # plan truncated where a level collapses
])
def test_generated_index_files_identical_to_oracle(tmp_path, args):
    h = tblaeq.Handler.random(*args)
    o = O.Index.generate(*args)
    a, b = str(tmp_path / "prod"), str(tmp_path / "oracle")
    h.save(a)
    o.save(b)
    assert h.mesh_sizes == o.mesh_sizes
    for f in sorted(os.listdir(b)):
        assert filecmp.cmp(os.path.join(a, f), os.path.join(b, f), shallow=False), f


def test_load_save_roundtrip(tmp_path):
    o = O.Index.generate(4000, 3, 1.0, 100.0, False, (10, 5), 5, 3.0)
    a, b = str(tmp_path / "a"), str(tmp_path / "b")
    o.save(a)
    tblaeq.Handler.load(a).save(b)
    for f in sorted(os.listdir(a)):
        assert filecmp.cmp(os.path.join(a, f), os.path.join(b, f), shallow=False), f


def test_load_missing_index_is_an_error(tmp_path):
    with pytest.raises(tblaeq.TblaeqError):
        tblaeq.Handler.load(str(tmp_path / "nope"))


def test_kmeans_build_matches_oracle(tmp_path):
    rng = np.random.default_rng(0)
    data = rng.uniform(1, 100, (3000, 3))
    h = tblaeq.Handler.from_data(data, (10, 5))
    o = O.Index.build(data, (10, 5), 7)
    a, b = str(tmp_path / "a"), str(tmp_path / "b")
    h.save(a)
    o.save(b)
    for f in sorted(os.listdir(b)):
        assert filecmp.cmp(os.path.join(a, f), os.path.join(b, f), shallow=False), f


def test_scheduler_reference_formulas_match_oracle():
    h = tblaeq.Handler.random(20000, 4, 1.0, 100.0, False, (20, 10, 5), 7, 3.0)
    o = O.Index.generate(20000, 4, 1.0, 100.0, False, (20, 10, 5), 7, 3.0)
    ms = o.mesh_sizes  # finest first
    L = o.intervals
    nnz = [ms[L - 1 - l] for l in range(L)]
    cols = [ms[L - l] for l in range(L)]
    coarsest = ms[-1] * 8 + ms[-1] * 4 * 8
    maps = sum(ms[:-1]) * 8
    for free in np.unique(np.geomspace(1e3, 5e7, 60).astype(np.int64)):
        exp = ["L%d" % p for p in O.policy_recommend(nnz, cols, 4, coarsest, maps, int(free))]
        assert h.recommend(int(free), count_work_buffers=False) == exp, free


def test_scheduler_counts_work_buffers():
    h = tblaeq.Handler.random(20000, 4, 1.0, 100.0, False, (20, 10, 5), 7, 3.0)
    big = 1 << 40
    assert h.recommend(big) == ["L0", "L0", "L0"]
    assert h.recommend(1000) == ["L3", "L3", "L3"]
    # monotone: more memory never gives a more streaming policy at any level
    prev = None
    for free in np.geomspace(1e5, 1e9, 40).astype(np.int64):
        cur = [int(p[1]) for p in h.recommend(int(free))]
        if prev is not None:
            assert all(c <= p or p == 2 for c, p in zip(cur, prev))
        prev = cur


def _clustered(n, D, seed):
    rng = np.random.default_rng(seed)
    cen = rng.uniform(1.0, 100.0, (max(1, n // 50), D))
    return np.ascontiguousarray(cen[rng.integers(0, cen.shape[0], n)] + rng.normal(0.0, 1.5, (n, D)))


def test_host_build_options_match_oracle_kmeans():
    """The host build with other iteration and seed settings (the index-build configuration's 20
    iterations from seed 7) equals the oracle's k-means with the same settings, and its reported
    inertia is that clustering's."""
    x = _clustered(6000, 5, 1)
    h = tblaeq.Handler.build(x, (40,), "opt", force_cpu=True, max_iter=20, seed=7)
    lab, cen, it = O.kmeans(x, 150, 20, seed=7)
    st = h.build_stats()
    assert st.shape == (1, 6) and int(st[0, 0]) == it
    d = x - cen[lab.astype(np.int64)]
    np.testing.assert_allclose(h.build_inertia()[0], float(np.sum(d * d)), rtol=1e-12)
    # another seed gives another clustering
    lab2, _, _ = O.kmeans(x, 150, 20, seed=42)
    assert not np.array_equal(lab, lab2)
    h.close()


def test_host_build_fp32_input_equals_fp64_of_it(tmp_path):
    x32 = _clustered(3000, 4, 2).astype(np.float32)
    a = tblaeq.Handler.build(x32, (30,), "f32", force_cpu=True)
    b = tblaeq.Handler.build(x32.astype(np.float64), (30,), "f32", force_cpu=True, max_iter=7)
    a.save(str(tmp_path / "a"))
    b.save(str(tmp_path / "b"))
    for f in sorted(os.listdir(tmp_path / "a")):
        assert filecmp.cmp(str(tmp_path / "a" / f), str(tmp_path / "b" / f), shallow=False), f
    a.close()
    b.close()


def test_host_build_rejects_kmeans_parallel_seeding():
    with pytest.raises(tblaeq.TblaeqError, match="GPU only"):
        tblaeq.Handler.build(_clustered(500, 3, 3), (10,), force_cpu=True, seeding=1)
