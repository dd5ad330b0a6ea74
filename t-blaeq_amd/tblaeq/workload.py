"""Synthetic query workloads over a generated hierarchy.

Range boxes are centred on data points with half-widths h*std per dimension;
h is the exact k-th order statistic of max_d |x_d - c_d| / std_d over a random
subsample of the data, so that a box holds about `selectivity` of the points.
KNN queries are data points plus N(0, (noise * spacing)^2) with spacing the
finest-level centre spacing of the generator, (vmax - vmin) / nParents^(1/D).
"""
import numpy as np

# The index of the headline measurement: N = 10M FP64 points, D = 8, height 4,
# ratios 100,50,20, sigma divisor 2.2, seed 20260330 in [1, 100]^8.
C2 = dict(N=10_000_000, D=8, vmin=1.0, vmax=100.0, is_int=False, ratios=(100, 50, 20), seed=20260330,
          sigma_div=2.2)
# The learned-radius index: N = 10M FP64 points, D = 16, height 4, ratios
# 100,50,20 (generator constants as for the headline index).
C3 = dict(N=10_000_000, D=16, vmin=1.0, vmax=100.0, is_int=False, ratios=(100, 50, 20), seed=20260330,
          sigma_div=2.2)
C1 = dict(N=500_000, D=5, vmin=1.0, vmax=100.0, is_int=False, ratios=(100, 50, 20), seed=12345, sigma_div=3.0)


def _kth_threshold(sub, centers, std, ks, progress=None):
    """Per centre, the k-th smallest normalised Chebyshev distance to the subsample."""
    out = np.zeros(len(centers))
    # float32 is ample for choosing a half-width; the boxes themselves are FP64
    s32 = np.ascontiguousarray((sub / std).T, dtype=np.float32)   # D x n
    c32 = (centers / std).astype(np.float32)
    t = np.empty(s32.shape[1], np.float32)
    tmp = np.empty(s32.shape[1], np.float32)
    for i in range(len(centers)):
        np.subtract(s32[0], c32[i, 0], out=t)
        np.abs(t, out=t)
        for d in range(1, s32.shape[0]):
            np.subtract(s32[d], c32[i, d], out=tmp)
            np.abs(tmp, out=tmp)
            np.maximum(t, tmp, out=t)
        k = int(ks[i])
        out[i] = float(np.partition(t, k - 1)[k - 1])
        if progress and i % 1024 == 1023:
            progress("calibrated %d/%d boxes" % (i + 1, len(centers)))
    return out


def range_boxes(data, n, seed, sel_lo=1e-5, sel_hi=1e-2, subsample=200_000, progress=None):
    """n boxes (rows lo|hi) with target selectivity log-uniform in [sel_lo, sel_hi]."""
    rng = np.random.default_rng(seed)
    N, D = data.shape
    sub = data[rng.choice(N, min(N, subsample), replace=False)] if N > subsample else data
    std = data.std(axis=0)
    centers = data[rng.integers(0, N, n)]
    sel = np.exp(rng.uniform(np.log(sel_lo), np.log(sel_hi), n))
    ks = np.maximum(1, np.round(sel * sub.shape[0])).astype(np.int64)
    h = _kth_threshold(sub, centers, std, ks, progress)
    lo = centers - h[:, None] * std
    hi = centers + h[:, None] * std
    return np.ascontiguousarray(np.concatenate([lo, hi], axis=1)), sel


def knn_points(data, n, seed, mesh_sizes, vmin, vmax, noise=0.01):
    """n query points near data points (noise in units of the finest cluster spacing)."""
    rng = np.random.default_rng(seed)
    N, D = data.shape
    n_parents = mesh_sizes[1] if len(mesh_sizes) > 1 else 1
    spacing = (vmax - vmin) / n_parents ** (1.0 / D)
    q = data[rng.integers(0, N, n)] + rng.normal(0.0, noise * spacing, (n, D))
    return np.ascontiguousarray(q)
