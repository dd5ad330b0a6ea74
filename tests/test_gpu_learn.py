"""GPU tests of the learned KNN radius model: the BF16-MFMA forward and backward
kernels against a NumPy reference of the same network (same BF16 rounding
points, float64 accumulation), training, and the certified learned KNN paths:
answers equal the oracle's with and without fallbacks, and with a trained
network the per-query traversal prolongs far fewer points than without it."""
import numpy as np
import pytest

import oracle_py as O
import tblaeq
from tblaeq import learn
from tblaeq import workload as W

pytestmark = pytest.mark.gpu

IN, H, OP, OUT = 48, 64, 16, 4


def bf(x):
    """Round float32 to bfloat16 (nearest even) and back."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def unpack(w):
    o = 0
    out = {}
    for name, shape in (("W1", (IN, H)), ("b1", (H,)), ("W2", (H, H)), ("b2", (H,)), ("W3", (H, OP)), ("b3", (OP,))):
        n = int(np.prod(shape))
        out[name] = w[o:o + n].reshape(shape).astype(np.float64)
        o += n
    return out


def ref_forward(w, X):
    p = unpack(w)
    Xb = bf(X).astype(np.float64)
    W1, W2, W3 = (bf(p[k].astype(np.float32)).astype(np.float64) for k in ("W1", "W2", "W3"))
    H1 = np.maximum(Xb @ W1 + p["b1"], 0.0)
    H1b = bf(H1.astype(np.float32)).astype(np.float64)
    H2 = np.maximum(H1b @ W2 + p["b2"], 0.0)
    H2b = bf(H2.astype(np.float32)).astype(np.float64)
    Y = H2b @ W3 + p["b3"]
    return Y, (Xb, W1, W2, W3, H1b, H2b)


def ref_gradients(w, X, T, lw=(1, 1, 1, 1)):
    n = X.shape[0]
    Y, (Xb, W1, W2, W3, H1b, H2b) = ref_forward(w, X)
    lw = np.asarray(lw, np.float64)
    d = Y[:, :OUT] - T
    loss = float(np.sum(lw * d * d) / n)
    dY = np.zeros((n, OP))
    dY[:, :OUT] = 2.0 * lw * d / n
    dY = bf(dY.astype(np.float32)).astype(np.float64)
    g = {"W3": H2b.T @ dY, "b3": dY.sum(0)}
    dH2 = np.where(H2b > 0, bf((dY @ W3.T).astype(np.float32)).astype(np.float64), 0.0)
    g["W2"], g["b2"] = H1b.T @ dH2, dH2.sum(0)
    dH1 = np.where(H1b > 0, bf((dH2 @ W2.T).astype(np.float32)).astype(np.float64), 0.0)
    g["W1"], g["b1"] = Xb.T @ dH1, dH1.sum(0)
    return np.concatenate([g[k].ravel() for k in ("W1", "b1", "W2", "b2", "W3", "b3")]), loss


CFG = dict(N=60000, D=6, vmin=1.0, vmax=100.0, is_int=False, ratios=(30, 10, 8), seed=41, sigma_div=2.5)


@pytest.fixture(scope="module")
def idx():
    h = tblaeq.Handler.random(CFG["N"], CFG["D"], CFG["vmin"], CFG["vmax"], CFG["is_int"], CFG["ratios"],
                              CFG["seed"], CFG["sigma_div"])
    o = O.Index.generate(CFG["N"], CFG["D"], CFG["vmin"], CFG["vmax"], CFG["is_int"], CFG["ratios"], CFG["seed"],
                         CFG["sigma_div"])
    h.prepare("L0")
    h.batch_init(4096, 1 << 24, 1 << 24)
    yield h, o
    h.close()


def _random_net(mlp, seed):
    rng = np.random.default_rng(seed)
    w = mlp.get_params()
    w = (w + rng.normal(0, 0.05, w.size)).astype(np.float32)
    p = unpack(w)
    # keep the padding columns of W3 / b3 at zero (as the layout requires)
    W3 = p["W3"]
    W3[:, OUT:] = 0
    b3 = p["b3"]
    b3[OUT:] = 0
    w = np.concatenate([p[k].ravel() for k in ("W1", "b1", "W2", "b2", "W3", "b3")]).astype(np.float32)
    mlp.set_params(w)
    return w


def test_mlp_forward_matches_reference(idx):
    h, _ = idx
    mlp = tblaeq.Mlp(h, seed=3)
    w = _random_net(mlp, 1)
    X = np.random.default_rng(2).normal(0, 1, (203, IN)).astype(np.float32)  # not a multiple of 16
    Y = mlp.forward(X)
    R, _ = ref_forward(w, X)
    err = np.abs(Y - R[:, :OUT])
    # BF16 activations: the GPU and the reference round the same values; float32 vs float64
    # accumulation can flip a rounding, so allow a few BF16 ulps of the output scale
    assert err.max() <= 2e-2 * (np.abs(R[:, :OUT]).max() + 1.0)
    assert err.mean() <= 2e-3 * (np.abs(R[:, :OUT]).mean() + 1.0)
    mlp.close()


def test_mlp_gradients_match_cpu_reference(idx):
    h, _ = idx
    mlp = tblaeq.Mlp(h, seed=4)
    w = _random_net(mlp, 5)
    rng = np.random.default_rng(6)
    X = rng.normal(0, 1, (517, IN)).astype(np.float32)
    T = rng.normal(0, 1, (517, OUT)).astype(np.float32)
    g, loss = mlp.gradients(X, T)
    gr, lr_ = ref_gradients(w, X, T)
    assert abs(loss - lr_) <= 1e-2 * abs(lr_) + 1e-6
    o = 0
    for name, n in (("W1", IN * H), ("b1", H), ("W2", H * H), ("b2", H), ("W3", H * OP), ("b3", OP)):
        a, b = g[o:o + n], gr[o:o + n]
        o += n
        scale = np.abs(b).max() + 1e-12
        assert np.abs(a - b).max() <= 3e-2 * scale, name
        assert np.abs(a - b).mean() <= 3e-3 * scale, name
    # padding outputs never receive gradient
    W3g = g[IN * H + H + H * H + H:IN * H + H + H * H + H + H * OP].reshape(H, OP)
    assert np.all(W3g[:, OUT:] == 0)
    mlp.close()


def test_mlp_training_reduces_loss(idx):
    h, _ = idx
    mlp = tblaeq.Mlp(h, seed=8)
    rng = np.random.default_rng(9)
    X = rng.normal(0, 1, (4096, IN)).astype(np.float32)
    A = rng.normal(0, 0.3, (IN, OUT))
    T = np.tanh(X.astype(np.float64) @ A).astype(np.float32)
    losses = mlp.train(X, T, epochs=15, batch=256, lr=3e-3)
    assert losses[-1] < 0.3 * losses[0]
    mlp.close()


def _trained(h, n=6000, epochs=30):
    mlp = tblaeq.Mlp(h, seed=11)
    scale = CFG["vmax"] - CFG["vmin"]
    mlp.set_norm(CFG["vmin"], scale)
    X, T = learn.traces(h, n, 21, CFG["vmin"], CFG["vmax"])
    losses = learn.fit(mlp, X, T, epochs=epochs, batch=256, lr=3e-3)
    return mlp, losses


def _check(h, o, pts, Ks, res):
    for i, q in enumerate(pts):
        eid, ed2 = o.knn(q, int(Ks[i])).refine_knn(q, int(Ks[i]))
        assert np.array_equal(res[i][0], eid), "query %d ids differ" % i
        assert np.array_equal(res[i][1].view(np.uint64), ed2.view(np.uint64)), "query %d distances differ" % i


def test_batched_learned_knn_exact_with_and_without_fallback(idx):
    """The batched engine with a proposed radius per query (no seed expansion) certifies or falls
    back, and its answers always equal the seed method's and the oracle's."""
    h, o = idx
    mlp, losses = _trained(h)
    assert losses[-1] < losses[0]
    data = h.data()
    pts = W.knn_points(data, 300, 77, h.mesh_sizes, CFG["vmin"], CFG["vmax"])
    Ks = np.random.default_rng(1).choice([10, 100], len(pts)).astype(np.uint64)
    h.batch_set("knn", pts, Ks)
    h.batch_run()
    seed_res = h.batch_results()
    h.batch_set("knn", pts, Ks)
    fb = h.batch_knn_learned(mlp, factor=1.15)
    learned_res = h.batch_results()
    assert all(np.array_equal(a[0], b[0]) and np.array_equal(a[1].view(np.uint64), b[1].view(np.uint64))
               for a, b in zip(seed_res, learned_res))
    _check(h, o, pts[:60], Ks[:60], learned_res[:60])
    assert fb < len(pts) // 4
    # a radius far too small: every query falls back and the answers still hold
    h.batch_set("knn", pts[:80], Ks[:80])
    assert h.batch_knn_learned(mlp, factor=1e-6) == 80
    _check(h, o, pts[:40], Ks[:40], h.batch_results()[:40])
    # with statistics the learned path reports the same survivors as the oracle
    h.batch_set("knn", pts[:40], Ks[:40])
    h.batch_knn_learned(mlp, factor=1.15, stats=True)
    surv, fine = h.batch_stats()
    for i, q in enumerate(pts[:40]):
        _, s, out = o.knn(q, int(Ks[i])).levels()
        assert np.array_equal(surv[i], s) and fine[i] == out[-1]
    mlp.close()


def test_plain_learned_knn_gains_over_plain_traversal(idx):
    """Per-query KNN: with the trained network's radius the last level prolongs only the survivors
    within it -- far fewer points than the plain traversal -- and the certified answers are the
    plain traversal's; a hopeless radius makes every query rerun and still answers exactly."""
    h, o = idx
    mlp, _ = _trained(h)
    data = h.data()
    pts = W.knn_points(data, 120, 91, h.mesh_sizes, CFG["vmin"], CFG["vmax"])
    K = 10
    radii = learn.propose_radii(h, mlp, pts, np.full(len(pts), K), CFG["vmin"], CFG["vmax"], factor=1.15)
    plain = h.query("knn", pts, K=K, refine=True)
    learned, fb = h.query_learned(pts, K, radii)
    assert fb < len(pts) // 4
    fine_plain = fine_learned = 0
    for i in range(len(pts)):
        pid, pd2 = plain.refined(i)
        lid, ld2 = learned.refined(i)
        assert np.array_equal(pid, lid) and np.array_equal(pd2.view(np.uint64), ld2.view(np.uint64))
        fine_plain += plain.info(i)["fine"]
        fine_learned += learned.info(i)["fine"]
    assert fine_learned * 2 < fine_plain, (fine_learned, fine_plain)
    _check(h, o, pts[:30], np.full(30, K), [learned.refined(i) for i in range(30)])
    bad, fb_all = h.query_learned(pts[:20], K, radii[:20] * 1e-6)
    assert fb_all == 3 * 20  # three growing radii miss, then the full traversal
    for i in range(20):
        assert np.array_equal(bad.refined(i)[0], plain.refined(i)[0])
    mlp.close()
