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
