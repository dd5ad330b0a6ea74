"""The allocation-free range compaction (tblaeq_range_prune_compact_noalloc, scratch in a fixed
device array) against tblaeq_range_prune_compact (the same kernel with a context's allocations),
through the C ABI: survivor ids, values, prolongation offsets and counts bit-identical, on two
calls in a row (the second with a smaller grid), and both equal to the oracle's range prune."""
import ctypes as C

import numpy as np
import pytest

import oracle_py as O
import tblaeq

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


class Level(C.Structure):
    """tblaeq_level (include/tblaeq.h)."""
    _fields_ = [("D", C.c_uint64), ("cols", C.c_uint64), ("nnz", C.c_uint64), ("col_res", C.c_void_p),
                ("nnz_per_col", C.c_void_p), ("pvals", C.c_void_p), ("radius", C.c_void_p),
                ("child_ids", C.c_void_p), ("empty_cols", C.c_uint64)]


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _case(n, D, seed):
    """A level-l grid of n points (ids a permutation of the cluster ids), radii and column
    pointers, and one box."""
    rng = np.random.default_rng(seed)
    cols = n + 37
    ids = rng.permutation(cols)[:n].astype(np.uint64)
    vals = rng.uniform(0.0, 100.0, (n, D))
    radius = rng.uniform(0.0, 25.0, cols)
    radius[rng.integers(0, cols, 50)] = 0.0
    counts = rng.integers(0, 40, cols).astype(np.uint64)
    col_res = np.concatenate([[0], np.cumsum(counts)]).astype(np.uint64)
    lo = rng.uniform(20.0, 45.0, D)
    hi = lo + rng.uniform(10.0, 40.0, D)
    return ids, vals, radius, col_res, lo, hi


def _expected(ids, vals, radius, col_res, lo, hi):
    keep = np.array([O.range_dist2(v, lo, hi) <= radius[i] * radius[i] for i, v in zip(ids, vals)], bool)
    sid = ids[keep]
    w = (col_res[sid.astype(np.int64) + 1] - col_res[sid.astype(np.int64)]).astype(np.uint64)
    off = np.concatenate([[0], np.cumsum(w)[:-1]]).astype(np.uint64) if len(w) else w
    return sid, vals[keep], off, int(w.sum())


def _run(ctx, case):
    """Runs both entry points on the same device inputs; returns their host results."""
    ids, vals, radius, col_res, lo, hi = case
    n, D = vals.shape
    d_ids, d_vals, d_rad, d_cr = _dev(ids.view(np.int64)), _dev(vals), _dev(radius), _dev(col_res.view(np.int64))
    d_lo, d_hi = _dev(lo), _dev(hi)
    d_count = _dev(np.array([n], np.int64))
    lv = Level(D, len(radius), int(col_res[-1]), d_cr.data_ptr(), None, None, d_rad.data_ptr(), None, 0)
    L = tblaeq.lib()
    out = []
    for kind in ("ctx", "noalloc"):
        s_ids = torch.full((n,), -1, dtype=torch.int64, device="cuda")
        s_vals = torch.full((n, D), float("nan"), dtype=torch.float64, device="cuda")
        s_off = torch.full((n,), -1, dtype=torch.int64, device="cuda")
        s_cnt = torch.full((1,), -1, dtype=torch.int64, device="cuda")
        o_cnt = torch.full((1,), -1, dtype=torch.int64, device="cuda")
        args = (C.byref(lv), d_ids.data_ptr(), d_vals.data_ptr(), d_count.data_ptr(), n, d_lo.data_ptr(),
                d_hi.data_ptr(), s_ids.data_ptr(), s_vals.data_ptr(), s_off.data_ptr(), s_cnt.data_ptr(),
                o_cnt.data_ptr())
        torch.cuda.synchronize()
        if kind == "ctx":
            tblaeq.check(L.tblaeq_range_prune_compact(ctx, *args), "tblaeq_range_prune_compact")
            tblaeq.check(L.tblaeq_ctx_sync(ctx), "ctx_sync")
        else:
            stream = torch.cuda.current_stream().cuda_stream
            tblaeq.check(L.tblaeq_range_prune_compact_noalloc(stream, *args), "tblaeq_range_prune_compact_noalloc")
            torch.cuda.synchronize()
        m = int(s_cnt.cpu()[0])
        out.append((s_ids.cpu().numpy()[:m].view(np.uint64), s_vals.cpu().numpy()[:m], s_off.cpu().numpy()[:m]
                    .view(np.uint64), int(o_cnt.cpu()[0]), s_ids.cpu().numpy()[m:]))
    return out


def test_noalloc_matches_context_entry_point():
    ctx = C.c_void_p()
    L = tblaeq.lib()
    tblaeq.check(L.tblaeq_ctx_create(0, None, C.byref(ctx)), "ctx_create")
    try:
        for n, D, seed in ((300_000, 6, 1), (20_000, 6, 2)):  # the second call with a smaller grid
            print("noalloc vs context: n=%d D=%d seed=%d" % (n, D, seed), flush=True)
            case = _case(n, D, seed)
            a, b = _run(ctx, case)
            assert np.array_equal(a[0], b[0]) and np.array_equal(a[2], b[2]) and a[3] == b[3]
            assert np.array_equal(a[1].view(np.uint64), b[1].view(np.uint64))
            assert np.all(b[4] == -1)  # nothing written past the survivors
            sid, sval, off, total = _expected(*case)
            assert np.array_equal(b[0], sid) and np.array_equal(b[2], off) and b[3] == total
            assert np.array_equal(b[1].view(np.uint64), sval.view(np.uint64))
            assert 0 < len(sid) < n
    finally:
        L.tblaeq_ctx_destroy(ctx)


def test_noalloc_rejects_grids_beyond_its_scratch():
    L = tblaeq.lib()
    lv = Level(3, 1, 1, 1, None, None, 1, None, 0)  # never dereferenced: the size check comes first
    rc = L.tblaeq_range_prune_compact_noalloc(None, C.byref(lv), 1, 1, 1, (1 << 24) + 1, 1, 1, 1, 1, 1, 1, 1)
    assert rc == 4  # TBLAEQ_ERR_CAPACITY
    assert b"exceeds the fixed scratch" in L.tblaeq_last_error()
