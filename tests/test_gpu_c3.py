"""The learned-radius index (N = 10M FP64 points, D = 16, ratios 100,50,20, sigma divisor 2.2,
seed 20260330): a network trained briefly on batched-engine traces proposes radii for a sample of
KNN points (K 10 and 100, noise 0.01 spacing); the batched learned path, the batched seed method
and the per-query learned traversal all return the oracle's exact refine of the reference
traversal, whatever the proposals' quality.  The oracle queries run in a thread pool."""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import oracle_py as O
import tblaeq
from tblaeq import learn
from tblaeq import workload as W

pytestmark = pytest.mark.gpu

C3 = W.C3


@pytest.fixture(scope="module")
def c3():
    print("building the 10M-point D=16 index (handler and oracle)", flush=True)
    h = tblaeq.Handler.random(C3["N"], C3["D"], C3["vmin"], C3["vmax"], C3["is_int"], C3["ratios"], C3["seed"],
                              C3["sigma_div"])
    o = O.Index.generate(C3["N"], C3["D"], C3["vmin"], C3["vmax"], C3["is_int"], C3["ratios"], C3["seed"],
                         C3["sigma_div"])
    h.prepare("L0")
    h.batch_init(4096, 1 << 26, 1 << 26)
    print("training on 8192 traces", flush=True)
    X, T = learn.traces(h, 8192, 5, C3["vmin"], C3["vmax"])
    mlp = tblaeq.Mlp(h, seed=3)
    mlp.set_norm(C3["vmin"], C3["vmax"] - C3["vmin"])
    learn.fit(mlp, X, T, epochs=4, batch=1024, lr=2e-3)
    yield h, o, mlp
    mlp.close()
    h.close()


def test_c3_learned_knn_sample(c3):
    h, o, mlp = c3
    pts = W.knn_points(h.data(), 64, 20260333, h.mesh_sizes, C3["vmin"], C3["vmax"], noise=0.01)
    Ks = np.where(np.arange(64) % 2 == 0, 10, 100).astype(np.uint64)
    with ThreadPoolExecutor(16) as ex:
        exp = list(ex.map(lambda i: o.knn(pts[i], int(Ks[i])).refine_knn(pts[i], int(Ks[i])), range(64)))
    h.batch_set("knn", pts, Ks)
    h.batch_run()
    seed = h.batch_results()
    # a tight factor forces fallbacks, a loose one none: answers stay exact either way
    for factor in (0.5, 1.0, 3.0):
        fb = h.batch_knn_learned(mlp, factor=factor)
        got = h.batch_results()
        for i in range(64):
            eid, ed2 = exp[i]
            for (gid, gd2), name in ((seed[i], "seed"), (got[i], "learned x%.1f" % factor)):
                assert np.array_equal(gid, eid), "query %d ids differ (%s)" % (i, name)
                assert np.array_equal(gd2.view(np.uint64), ed2.view(np.uint64)), "query %d d2 differ (%s)" % (i, name)
        print("factor %.1f: %d fallbacks" % (factor, fb), flush=True)
    radii = learn.propose_radii(h, mlp, pts[:8], Ks[:8], C3["vmin"], C3["vmax"])
    for K in (10, 100):
        sel = np.nonzero(Ks[:8] == K)[0]
        r, _ = h.query_learned(pts[sel], K, radii[sel])
        for j, i in enumerate(sel):
            rid, rd2 = r.refined(j)
            assert np.array_equal(rid, exp[i][0]) and np.array_equal(rd2.view(np.uint64), exp[i][1].view(np.uint64))
        r.close()
