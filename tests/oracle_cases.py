"""Query cases built with the CPU oracle (test infrastructure, not collected by pytest).

points_outside_radius: finest points whose reconstructed coordinates lie outside the
stored radius of their last-level cluster.  Radii are computed from the original data
(Setup.cuh:148-171) while every query sees the reconstructed coordinates (the P-ratio
products of SpMSpV.cuh:47-67), so a degenerate box placed on such a point is the case
where the reference's last-level ball test (RangePruning.cu:5-25) drops a cluster that
holds a point inside the box.
"""
import numpy as np

import oracle_py as O


def parent_ids(o, l):
    """Level-l cluster id of every level-(l+1) point (by level-(l+1) id): the CSC column of
    interval l that holds it, mapped through map[l+1] (Setup.cuh:115-118)."""
    col_res = o.array(l, 0)
    row_ids = o.array(l, 2)
    fmap = o.array(l, 3)
    parent = np.empty(len(fmap), np.uint64)
    counts = np.diff(col_res).astype(np.int64)
    parent[fmap[row_ids]] = np.repeat(np.arange(len(counts), dtype=np.uint64), counts)
    return parent


def last_level_parent(o):
    """Last-level cluster id of every finest point (by finest id)."""
    return parent_ids(o, o.intervals - 1)


def points_outside_radius(o, limit=None):
    """Finest ids whose reconstructed point x has rangeDist2(c, [x, x]) > r_c^2 for its
    last-level cluster c (the oracle's own fma chain decides), and their coordinates."""
    L = o.intervals
    fine = o.recon_level(L)
    cen = o.recon_level(L - 1)
    rad = o.array(L - 1, 5)
    parent = last_level_parent(o).astype(np.int64)
    c = cen[parent]
    r2 = rad[parent] * rad[parent]
    d2 = ((fine - c) ** 2).sum(axis=1)
    cand = np.nonzero(d2 >= r2 * (1.0 - 1e-9))[0]
    out = []
    for i in cand:
        x = fine[i]
        if O.range_dist2(c[i], x, x) > r2[i]:
            out.append(i)
            if limit is not None and len(out) >= limit:
                break
    ids = np.asarray(out, np.uint64)
    return ids, fine[ids.astype(np.int64)] if len(out) else np.zeros((0, o.D))


def degenerate_boxes(points):
    """Boxes lo = hi = point (rows lo|hi)."""
    return np.ascontiguousarray(np.concatenate([points, points], axis=1))


def integer_boxes(data, n, seed, max_half=3):
    """Boxes with integer corners around integer data points (exact ties on box faces)."""
    rng = np.random.default_rng(seed)
    c = data[rng.integers(0, data.shape[0], n)]
    h = rng.integers(0, max_half + 1, (n, data.shape[1])).astype(np.float64)
    return np.ascontiguousarray(np.concatenate([c - h, c + h], axis=1))
