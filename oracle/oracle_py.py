"""ctypes binding of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() as the
checker, and by bench.py's CPU-baseline leg.  The shipped library never loads it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

u64p = C.POINTER(C.c_uint64)
f64p = C.POINTER(C.c_double)
u8p = C.POINTER(C.c_uint8)


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        vp = C.c_void_p
        sig = {
            "oracle_gen_index": (vp, [C.c_uint64, C.c_uint64, C.c_double, C.c_double, C.c_int, u64p, C.c_uint64,
                                      C.c_uint64, C.c_double]),
            "oracle_build_index": (vp, [f64p, C.c_uint64, C.c_uint64, u64p, C.c_uint64, C.c_uint64]),
            "oracle_load_index": (vp, [C.c_char_p]),
            "oracle_save_index": (C.c_int, [vp, C.c_char_p]),
            "oracle_free_index": (None, [vp]),
            "oracle_index_meta": (C.c_uint64, [vp, u64p, u64p, C.c_uint64]),
            "oracle_index_data": (C.c_uint64, [vp, f64p, C.c_uint64]),
            "oracle_index_scale_radius": (None, [vp, C.c_double]),
            "oracle_index_layer": (C.c_uint64, [vp, C.c_uint64, f64p]),
            "oracle_index_labels": (C.c_uint64, [vp, C.c_uint64, u64p]),
            "oracle_index_array": (C.c_uint64, [vp, C.c_uint64, C.c_int, vp]),
            "oracle_range": (vp, [vp, f64p, f64p]),
            "oracle_knn": (vp, [vp, f64p, C.c_uint64]),
            "oracle_free_result": (None, [vp]),
            "oracle_result_size": (C.c_uint64, [vp]),
            "oracle_result_copy": (None, [vp, u64p, f64p]),
            "oracle_result_levels": (None, [vp, u64p, u64p, u64p]),
            "oracle_result_survivors": (C.c_uint64, [vp, C.c_uint64, u64p, C.c_uint64]),
            "oracle_refine_range": (C.c_uint64, [vp, f64p, f64p, u64p, C.c_uint64]),
            "oracle_refine_knn": (C.c_uint64, [vp, f64p, C.c_uint64, u64p, f64p]),
            "oracle_recon_finest": (None, [vp, f64p]),
            "oracle_recon_level": (C.c_uint64, [vp, C.c_uint64, f64p]),
            "oracle_brute_range": (C.c_uint64, [f64p, C.c_uint64, C.c_uint64, f64p, f64p, u64p, C.c_uint64]),
            "oracle_brute_knn": (C.c_uint64, [f64p, C.c_uint64, C.c_uint64, f64p, C.c_uint64, u64p, f64p]),
            "oracle_sqrt_rd": (C.c_double, [C.c_double]),
            "oracle_sqrt_rd_fenv": (C.c_double, [C.c_double]),
            "oracle_range_dist2": (C.c_double, [f64p, f64p, f64p, C.c_uint64]),
            "oracle_knn_lb": (C.c_double, [f64p, f64p, C.c_double, C.c_uint64]),
            "oracle_step_sequential": (C.c_uint64, [f64p, u64p, f64p, C.c_uint64, C.c_uint64, u8p]),
            "oracle_policy_recommend": (None, [u64p, u64p, C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64,
                                               C.c_uint64, u8p]),
            "oracle_cpu_batch": (C.c_double, [vp, C.c_int, f64p, C.c_uint64, C.c_uint64, C.c_int, f64p, u64p,
                                              C.POINTER(C.c_int)]),
            "oracle_kmeans": (C.c_uint64, [f64p, C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64, u64p, f64p]),
            "oracle_kmeans_seeded": (C.c_uint64, [f64p, C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64,
                                                  u64p, f64p]),
            "oracle_kmeans_assign": (C.c_double, [f64p, C.c_uint64, C.c_uint64, f64p, C.c_uint64, u64p]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def _p(a, t):
    return a.ctypes.data_as(t)


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class Index:
    """Oracle index (host arrays).  Same semantics as the reference IndexData."""

    def __init__(self, handle):
        if not handle:
            raise RuntimeError("oracle index construction failed")
        self.h = handle
        meta = np.zeros(4, np.uint64)
        ms = np.zeros(16, np.uint64)
        n = lib().oracle_index_meta(self.h, _p(meta, u64p), _p(ms, u64p), 16)
        self.D, self.N, self.height, self.intervals = (int(x) for x in meta)
        self.mesh_sizes = [int(x) for x in ms[:n]]

    @classmethod
    def generate(cls, N, D, vmin=1.0, vmax=100.0, is_int=False, ratios=(100, 50, 20), seed=12345, sigma_div=3.0):
        r = np.asarray(ratios, np.uint64)
        return cls(lib().oracle_gen_index(N, D, vmin, vmax, int(is_int), _p(r, u64p), len(r), seed, sigma_div))

    @classmethod
    def build(cls, data, ratios=(100, 50, 20), max_iter=7):
        data = f64(data)
        r = np.asarray(ratios, np.uint64)
        return cls(lib().oracle_build_index(_p(data, f64p), data.shape[0], data.shape[1], _p(r, u64p), len(r),
                                            max_iter))

    @classmethod
    def load(cls, path):
        return cls(lib().oracle_load_index(path.encode()))

    def save(self, path):
        if lib().oracle_save_index(self.h, path.encode()) != 0:
            raise RuntimeError("save failed")

    def data(self):
        out = np.zeros((self.N, self.D))
        if lib().oracle_index_data(self.h, _p(out, f64p), out.size) == 0:
            return None
        return out

    def layer(self, i):
        n = lib().oracle_index_layer(self.h, i, None)
        out = np.zeros(n)
        lib().oracle_index_layer(self.h, i, _p(out, f64p))
        return out

    def labels(self, i):
        n = lib().oracle_index_labels(self.h, i, None)
        out = np.zeros(n, np.uint64)
        lib().oracle_index_labels(self.h, i, _p(out, u64p))
        return out

    def array(self, l, which):
        n = lib().oracle_index_array(self.h, l, which, None)
        out = np.zeros(n, np.float64 if which >= 4 else np.uint64)
        lib().oracle_index_array(self.h, l, which, out.ctypes.data_as(C.c_void_p))
        return out

    def scale_radius(self, f):
        lib().oracle_index_scale_radius(self.h, f)

    def recon(self):
        out = np.zeros((self.N, self.D))
        lib().oracle_recon_finest(self.h, _p(out, f64p))
        return out

    def recon_level(self, level):
        """Reconstructed coordinates of level `level` (0 = coarsest, intervals = finest), by id."""
        n = lib().oracle_recon_level(self.h, level, None)
        out = np.zeros((n, self.D))
        lib().oracle_recon_level(self.h, level, _p(out, f64p))
        return out

    def range(self, lo, hi):
        return Result(lib().oracle_range(self.h, _p(f64(lo), f64p), _p(f64(hi), f64p)), self)

    def knn(self, q, K):
        return Result(lib().oracle_knn(self.h, _p(f64(q), f64p), K), self)

    def cpu_batch(self, qtype, qs, K=0, nthreads=0):
        qs = f64(qs)
        nq = qs.shape[0]
        us = np.zeros(nq)
        cnt = np.zeros(nq, np.uint64)
        th = C.c_int(0)
        wall = lib().oracle_cpu_batch(self.h, qtype, _p(qs, f64p), nq, K, nthreads, _p(us, f64p), _p(cnt, u64p),
                                      C.byref(th))
        return wall, us, cnt, th.value

    def __del__(self):
        if getattr(self, "h", None) and _LIB is not None:
            _LIB.oracle_free_index(self.h)
            self.h = None


class Result:
    def __init__(self, h, index):
        self.h = h
        self.D = index.D
        self.L = index.intervals

    def grid(self):
        n = lib().oracle_result_size(self.h)
        ids = np.zeros(n, np.uint64)
        vals = np.zeros((n, self.D))
        lib().oracle_result_copy(self.h, _p(ids, u64p), _p(vals, f64p))
        return ids, vals

    def levels(self):
        c = np.zeros(self.L, np.uint64)
        s = np.zeros(self.L, np.uint64)
        o = np.zeros(self.L, np.uint64)
        lib().oracle_result_levels(self.h, _p(c, u64p), _p(s, u64p), _p(o, u64p))
        return c, s, o

    def survivors(self, l):
        n = lib().oracle_result_survivors(self.h, l, None, 0)
        out = np.zeros(n, np.uint64)
        lib().oracle_result_survivors(self.h, l, _p(out, u64p), n)
        return out

    def refine_range(self, lo, hi):
        n = lib().oracle_refine_range(self.h, _p(f64(lo), f64p), _p(f64(hi), f64p), None, 0)
        out = np.zeros(n, np.uint64)
        lib().oracle_refine_range(self.h, _p(f64(lo), f64p), _p(f64(hi), f64p), _p(out, u64p), n)
        return out

    def refine_knn(self, q, K):
        ids = np.zeros(K, np.uint64)
        d2 = np.zeros(K)
        n = lib().oracle_refine_knn(self.h, _p(f64(q), f64p), K, _p(ids, u64p), _p(d2, f64p))
        return ids[:n], d2[:n]

    def __del__(self):
        if getattr(self, "h", None) and _LIB is not None:
            _LIB.oracle_free_result(self.h)
            self.h = None


def brute_range(pts, lo, hi):
    pts = f64(pts)
    n = lib().oracle_brute_range(_p(pts, f64p), pts.shape[0], pts.shape[1], _p(f64(lo), f64p), _p(f64(hi), f64p),
                                 None, 0)
    out = np.zeros(n, np.uint64)
    lib().oracle_brute_range(_p(pts, f64p), pts.shape[0], pts.shape[1], _p(f64(lo), f64p), _p(f64(hi), f64p),
                             _p(out, u64p), n)
    return out


def brute_knn(pts, q, K):
    pts = f64(pts)
    ids = np.zeros(K, np.uint64)
    d2 = np.zeros(K)
    n = lib().oracle_brute_knn(_p(pts, f64p), pts.shape[0], pts.shape[1], _p(f64(q), f64p), K, _p(ids, u64p),
                               _p(d2, f64p))
    return ids[:n], d2[:n]


def sqrt_rd(x):
    return lib().oracle_sqrt_rd(float(x))


def sqrt_rd_fenv(x):
    return lib().oracle_sqrt_rd_fenv(float(x))


def range_dist2(c, lo, hi):
    c = f64(c)
    return lib().oracle_range_dist2(_p(c, f64p), _p(f64(lo), f64p), _p(f64(hi), f64p), c.size)


def knn_lb(c, q, r):
    c = f64(c)
    return lib().oracle_knn_lb(_p(c, f64p), _p(f64(q), f64p), float(r), c.size)


def step_sequential(lb_sorted, w_sorted, r_sorted, K):
    lb = f64(lb_sorted)
    w = np.ascontiguousarray(w_sorted, np.uint64)
    r = f64(r_sorted)
    sel = np.zeros(lb.size, np.uint8)
    lib().oracle_step_sequential(_p(lb, f64p), _p(w, u64p), _p(r, f64p), lb.size, K, _p(sel, u8p))
    return sel.astype(bool)


def policy_recommend(nnz, cols, D, coarsest_bytes, maps_bytes, free_mem):
    nnz = np.asarray(nnz, np.uint64)
    cols = np.asarray(cols, np.uint64)
    out = np.zeros(nnz.size, np.uint8)
    lib().oracle_policy_recommend(_p(nnz, u64p), _p(cols, u64p), nnz.size, D, coarsest_bytes, maps_bytes, free_mem,
                                  _p(out, u8p))
    return out


def kmeans(data, k, max_iter, seed=42):
    data = f64(data)
    N, D = data.shape
    lab = np.zeros(N, np.uint64)
    cent = np.zeros((k, D))
    it = lib().oracle_kmeans_seeded(_p(data, f64p), N, D, k, max_iter, seed, _p(lab, u64p), _p(cent, f64p))
    return lab, cent, it


def kmeans_assign(data, cent):
    data = f64(data)
    cent = f64(cent)
    lab = np.zeros(data.shape[0], np.uint64)
    inertia = lib().oracle_kmeans_assign(_p(data, f64p), data.shape[0], data.shape[1], _p(cent, f64p),
                                         cent.shape[0], _p(lab, u64p))
    return lab, inertia
