"""Training traces and training of the learned KNN radius model.

A trace is one KNN query run through the batched engine with statistics: its
features (the network inputs), the K-th nearest candidate distance and the
survivors of the first levels (the targets).  The batched engine returns the
same candidate sets and answers as the per-query traversal (both are checked
against the oracle), so its traces are the plain traversal's traces.
"""
import numpy as np

from . import workload as W


def traces(h, n, seed, vmin, vmax, Ks=(10, 100), noise=0.01, chunk=4096):
    """n traces from query points near the data (K drawn from Ks); returns X [n x 48], T [n x 4]."""
    data = h.data()
    rng = np.random.default_rng(seed)
    scale = vmax - vmin
    Xs, Ts = [], []
    for c0 in range(0, n, chunk):
        m = min(chunk, n - c0)
        pts = W.knn_points(data, m, int(rng.integers(1 << 31)), h.mesh_sizes, vmin, vmax, noise=noise)
        K = rng.choice(np.asarray(Ks, np.uint64), m)
        h.batch_set("knn", pts, K)
        h.batch_run(stats=True)
        res = h.batch_results()
        surv, _ = h.batch_stats()
        dk = np.array([np.sqrt(r[1][-1]) if len(r[1]) else 0.0 for r in res])
        T = np.zeros((m, 4), np.float32)
        T[:, 0] = np.log(np.maximum(dk, 1e-9 * scale) / scale)
        for l in range(min(3, surv.shape[1])):
            T[:, 1 + l] = np.log1p(surv[:, l].astype(np.float64)) / 10.0
        Xs.append(h.knn_features(pts, K, vmin, scale))
        Ts.append(T)
    return np.concatenate(Xs), np.concatenate(Ts)


def fit(mlp, X, T, epochs=20, batch=1024, lr=2e-3, seed=0):
    """Shuffles the traces once and trains; returns the per-epoch losses."""
    order = np.random.default_rng(seed).permutation(X.shape[0])
    return mlp.train(X[order], T[order], epochs=epochs, batch=batch, lr=lr)


def propose_radii(h, mlp, points, Ks, vmin, vmax, factor=1.0):
    """Learned search radius per query: scale * exp(y0) * factor (batch_init first)."""
    scale = vmax - vmin
    X = h.knn_features(points, Ks, vmin, scale)
    Y = mlp.forward(X)
    return scale * np.exp(Y[:, 0].astype(np.float64)) * factor
