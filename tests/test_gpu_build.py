"""GPU index build (k-means on the matrix cores + exact FP64 re-score, CSC
P-tensor on the device) against the CPU build: the same labels and centroids,
the same P-tensor up to the order of the rows inside a column, and an index
that answers queries exactly like the oracle over it."""
import os

import numpy as np
import pytest

import oracle_py as O
import tblaeq
from tblaeq import workload as W

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gen():
    h = tblaeq.Handler.random(40000, 6, 1.0, 100.0, False, (40, 10), 5, 2.5)
    yield h
    h.close()


def _inertia(x, lab, cen):
    d = x - cen[lab.astype(np.int64)]
    return float(np.sum(d * d))


@pytest.mark.parametrize("D,k", [(5, 150), (8, 400), (32, 60)])
def test_kmeans_equals_cpu_kmeans(gen, D, k):
    """Seeding from the same mt19937_64 stream, FP32-MFMA ranking with the exact FP64 re-score of
    near ties, point-order centroid sums: labels and centroids equal the CPU build's."""
    rng = np.random.default_rng(D)
    base = gen.data()[:20000, : min(D, 6)]
    x = np.concatenate([base, rng.uniform(1, 100, (base.shape[0], max(0, D - base.shape[1])))], axis=1)
    x = np.ascontiguousarray(x[:, :D])
    glab, gcen, st = gen.kmeans(x, k, 7, 42)
    clab, ccen, it = O.kmeans(x, k, 7)
    assert st[0] == it
    assert np.array_equal(glab, clab)
    assert np.array_equal(gcen.view(np.uint64), ccen.view(np.uint64))
    gi, ci = _inertia(x, glab, gcen), _inertia(x, clab, ccen)
    assert abs(gi - ci) <= 1e-6 * ci


def test_kmeans_ties_take_the_lowest_index(gen):
    """Integer points on a coarse grid: many exact distance ties, all settled by the FP64 re-score
    with the CPU build's rule (strict <, lowest center index)."""
    rng = np.random.default_rng(3)
    x = rng.integers(0, 6, (6000, 4)).astype(np.float64)
    glab, gcen, st = gen.kmeans(x, 40, 7, 42)
    clab, ccen, _ = O.kmeans(x, 40, 7)
    assert st[1] > 0  # near ties were re-scored
    assert np.array_equal(glab, clab)
    assert np.array_equal(gcen.view(np.uint64), ccen.view(np.uint64))


def _read_index(path):
    """The on-disk index format (metadata, P tensors, radii, maps, coarsest mesh)."""
    def u64(f, n=1):
        return np.fromfile(f, np.uint64, n)

    with open(os.path.join(path, "metadata.bin"), "rb") as f:
        D, N, height, intervals = u64(f, 4)
        f.read(1)
        ratios = u64(f, int(intervals))
        msz = int(u64(f)[0])
        mesh = u64(f, msz)
    P = []
    for i in range(int(intervals)):
        with open(os.path.join(path, "P_Tensor_%d.bin" % i), "rb") as f:
            pD, rows, cols, nnz = (int(v) for v in u64(f, 4))
            f.read(1)
            rowIds = u64(f, nnz)
            nnzPerCol = u64(f, cols)
            colRes = u64(f, cols + 1)
            vals = np.fromfile(f, np.float64, nnz * pD).reshape(nnz, pD)
        P.append(dict(rowIds=rowIds, nnzPerCol=nnzPerCol, colRes=colRes, vals=vals))
    R = []
    for i in range(int(intervals)):
        with open(os.path.join(path, "MaxRadius_%d.bin" % i), "rb") as f:
            n = int(u64(f)[0])
            R.append(np.fromfile(f, np.float64, n))
    M = {}
    for i in range(1, int(height)):
        p = os.path.join(path, "Map_%d.bin" % i)
        if os.path.exists(p):
            with open(p, "rb") as f:
                n = int(u64(f)[0])
                M[i] = u64(f, n)
    with open(os.path.join(path, "Coreast_Mesh.bin"), "rb") as f:
        rows, dim, nnz = (int(v) for v in u64(f, 3))
        f.read(1)
        cids = u64(f, nnz)
        cvals = np.fromfile(f, np.float64, nnz * dim)
    return dict(D=int(D), mesh=mesh, ratios=ratios, P=P, R=R, M=M, cids=cids, cvals=cvals)


def test_gpu_build_index_equals_cpu_build(gen, tmp_path):
    data = np.ascontiguousarray(gen.data()[:30000])
    hg = tblaeq.Handler.build(data, (30, 10), "g", force_cpu=False)
    hc = tblaeq.Handler.build(data, (30, 10), "c", force_cpu=True)
    st = hg.build_stats()
    assert st.shape == (2, 6) and np.all(st[:, 0] >= 1)
    hg.save(str(tmp_path / "g"))
    hc.save(str(tmp_path / "c"))
    g, c = _read_index(str(tmp_path / "g")), _read_index(str(tmp_path / "c"))
    assert np.array_equal(g["mesh"], c["mesh"])
    assert np.array_equal(g["cvals"].view(np.uint64), c["cvals"].view(np.uint64))
    assert np.array_equal(g["cids"], c["cids"])
    for i in range(len(g["P"])):
        pg, pc = g["P"][i], c["P"][i]
        assert np.array_equal(pg["colRes"], pc["colRes"]) and np.array_equal(pg["nnzPerCol"], pc["nnzPerCol"])
        assert np.array_equal(pg["rowIds"], pc["rowIds"])
        assert np.array_equal(g["R"][i].view(np.uint64), c["R"][i].view(np.uint64)), "radius level %d" % i
        # the same children (original ids) with the same ratios in every column; the row order
        # inside a column may differ (stable sort here, std::sort in the CPU build)
        mg, mc = g["M"][i + 1], c["M"][i + 1]  # P tensor i (0 = coarsest) maps mesh i + 1
        cr = pg["colRes"]
        for col in range(len(cr) - 1):
            a, b = int(cr[col]), int(cr[col + 1])
            og, oc = np.argsort(mg[a:b], kind="stable"), np.argsort(mc[a:b], kind="stable")
            assert np.array_equal(mg[a:b][og], mc[a:b][oc])
            assert np.array_equal(pg["vals"][a:b][og].view(np.uint64), pc["vals"][a:b][oc].view(np.uint64))
        # rows of a column in point order
        for col in range(len(cr) - 1):
            seg = mg[int(cr[col]):int(cr[col + 1])]
            assert np.all(np.diff(seg.astype(np.int64)) > 0)
    hg.close()
    hc.close()


def test_gpu_built_index_answers_like_the_oracle(gen, tmp_path):
    data = np.ascontiguousarray(gen.data()[:30000])
    h = tblaeq.Handler.build(data, (30, 10), "g", force_cpu=False)
    h.save(str(tmp_path / "g"))
    o = O.Index.load(str(tmp_path / "g"))
    h.prepare("L0")
    boxes, _ = W.range_boxes(data, 40, 8, 1e-3, 3e-2)
    D = data.shape[1]
    for row in boxes[:20]:
        lo, hi = row[:D], row[D:]
        r = h.range(lo, hi)
        ids, _ = r.refined()
        assert np.array_equal(np.sort(ids), o.range(lo, hi).refine_range(lo, hi))
    h.batch_init(64, 1 << 20, 1 << 20)
    h.batch_set("range", boxes)
    h.batch_run()
    for i, row in enumerate(h.batch_results()):
        lo, hi = boxes[i][:D], boxes[i][D:]
        assert np.array_equal(row, o.range(lo, hi).refine_range(lo, hi))
    pts = data[np.random.default_rng(4).choice(len(data), 24, replace=False)] + 0.01
    Ks = np.full(len(pts), 10, np.uint64)
    h.batch_set("knn", pts, Ks)
    h.batch_run()
    for i, (gid, gd2) in enumerate(h.batch_results()):
        eid, ed2 = o.knn(pts[i], 10).refine_knn(pts[i], 10)
        assert np.array_equal(gid, eid) and np.array_equal(gd2.view(np.uint64), ed2.view(np.uint64))
    h.close()


def _clustered(n, D, seed, sigma=1.5):
    rng = np.random.default_rng(seed)
    cen = rng.uniform(1.0, 100.0, (max(1, n // 60), D))
    return np.ascontiguousarray(cen[rng.integers(0, cen.shape[0], n)] + rng.normal(0.0, sigma, (n, D)))


def test_gpu_build_options_and_fp32_input_match_host_build(tmp_path):
    """The index-build configuration's settings (20 iterations, seed 7) on FP32 input at D = 32: the
    GPU build (FP32 values converted to FP64 on the device) gives the host build's mesh sizes,
    coarsest mesh and radii, and the same inertia to within the summation order."""
    x32 = _clustered(20000, 32, 5).astype(np.float32)
    g = tblaeq.Handler.build(x32, (50, 10), "c4", force_cpu=False, max_iter=20, seed=7)
    c = tblaeq.Handler.build(x32, (50, 10), "c4", force_cpu=True, max_iter=20, seed=7)
    gs, cs = g.build_stats(), c.build_stats()
    assert np.array_equal(gs[:, 0], cs[:, 0])  # Lloyd iterations per level
    np.testing.assert_allclose(g.build_inertia(), c.build_inertia(), rtol=1e-12)
    g.save(str(tmp_path / "g"))
    c.save(str(tmp_path / "c"))
    ig, ic = _read_index(str(tmp_path / "g")), _read_index(str(tmp_path / "c"))
    assert np.array_equal(ig["mesh"], ic["mesh"])
    assert np.array_equal(ig["cvals"].view(np.uint64), ic["cvals"].view(np.uint64))
    for l in range(len(ig["R"])):
        assert np.array_equal(ig["R"][l].view(np.uint64), ic["R"][l].view(np.uint64)), l
    g.close()
    c.close()


def test_gpu_kmeans_parallel_seeding(tmp_path):
    """k-means|| seeding: deterministic (two builds give the same index), more candidates than
    centres, an inertia close to k-means++'s on the same data, and an index that answers queries
    exactly like the oracle over its saved files."""
    x = _clustered(60000, 8, 6)
    opts = dict(max_iter=20, seed=7, seeding=1, oversampling=0.5, rounds=4)
    a = tblaeq.Handler.build(x, (60, 10), "kpar", **opts)
    b = tblaeq.Handler.build(x, (60, 10), "kpar", **opts)
    pp = tblaeq.Handler.build(x, (60, 10), "kpp", max_iter=20, seed=7)
    sa, sb = a.build_stats(), b.build_stats()
    assert np.array_equal(sa, sb)
    assert sa[0, 5] > 1000  # candidates of the first level (k = 1000)
    assert pp.build_stats()[0, 5] == 0
    ia, ip = a.build_inertia()[0], pp.build_inertia()[0]
    assert ia <= 1.1 * ip, (ia, ip)
    a.save(str(tmp_path / "a"))
    o = O.Index.load(str(tmp_path / "a"))
    a.prepare("L0")
    rng = np.random.default_rng(9)
    for i in range(6):
        c = x[rng.integers(x.shape[0])]
        lo, hi = c - 3.0, c + 3.0
        r = a.range(lo, hi)
        assert np.array_equal(r.refined()[0], o.range(lo, hi).refine_range(lo, hi)), i
        k = a.knn(c + 0.3, 15)
        eid, ed2 = o.knn(c + 0.3, 15).refine_knn(c + 0.3, 15)
        rid, rd2 = k.refined()
        assert np.array_equal(rid, eid) and np.array_equal(rd2.view(np.uint64), ed2.view(np.uint64)), i
        r.close()
        k.close()
    for hh in (a, b, pp):
        hh.close()


def test_kmeans_inertia_stat(gen):
    x = _clustered(30000, 6, 7)
    h = tblaeq.Handler.build(x, (100,), "inertia", max_iter=10, seed=3)
    lab, cen, _ = O.kmeans(x, 300, 10, seed=3)
    d = x - cen[lab.astype(np.int64)]
    np.testing.assert_allclose(h.build_inertia()[0], float(np.sum(d * d)), rtol=1e-12)
    h.close()
