"""Writes tests/golden/ref_vectors.npz from the reference's own C++ headers.

oracle/_ref/libref.so is compiled (oracle/Makefile target `ref`) from the
reference's RandomKmeans.cuh, CPUKmeans.cuh and func.hpp where they lie in the
reference tree; this script calls it on small inputs and stores inputs and
outputs as data, so the oracle can be checked where the reference is absent.
"""
import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
LIB = os.path.join(ROOT, "oracle", "_ref", "libref.so")

u64p = C.POINTER(C.c_uint64)
f64p = C.POINTER(C.c_double)


def load():
    L = C.CDLL(LIB)
    L.ref_gen_create.restype = C.c_void_p
    L.ref_gen_create.argtypes = [C.c_uint64, C.c_uint64, u64p, C.c_uint64, C.c_double, C.c_double, C.c_int,
                                 C.c_uint64, C.c_double]
    L.ref_gen_layer.restype = C.c_uint64
    L.ref_gen_layer.argtypes = [C.c_void_p, C.c_uint64, f64p]
    L.ref_gen_labels.restype = C.c_uint64
    L.ref_gen_labels.argtypes = [C.c_void_p, C.c_uint64, u64p]
    L.ref_gen_free.argtypes = [C.c_void_p]
    L.ref_argsort.argtypes = [u64p, C.c_uint64, u64p]
    L.ref_euclidean.restype = C.c_double
    L.ref_euclidean.argtypes = [f64p, f64p, C.c_uint64]
    L.ref_kmeans.argtypes = [f64p, C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64, u64p, f64p]
    return L


def p(a, t):
    return a.ctypes.data_as(t)


def generate(L, N, D, ratios, vmin, vmax, is_int, seed, sd):
    r = np.asarray(ratios, np.uint64)
    h = L.ref_gen_create(N, D, p(r, u64p), len(r), vmin, vmax, int(is_int), seed, sd)
    layers, labels = [], []
    for i in range(len(r) + 1):
        n = L.ref_gen_layer(h, i, None)
        a = np.zeros(n)
        L.ref_gen_layer(h, i, p(a, f64p))
        layers.append(a)
    for i in range(len(r)):
        n = L.ref_gen_labels(h, i, None)
        a = np.zeros(n, np.uint64)
        L.ref_gen_labels(h, i, p(a, u64p))
        labels.append(a)
    L.ref_gen_free(h)
    return layers, labels


def main():
    if not os.path.exists(LIB):
        print("build oracle/_ref/libref.so first (make -C oracle ref)", file=sys.stderr)
        return 1
    L = load()
    out = {}
    cases = {"g1": (3000, 3, (10, 5), 1.0, 100.0, False, 12345, 3.0),
             "g2": (2000, 5, (20, 10), 1.0, 50.0, True, 7, 2.2)}
    for name, args in cases.items():
        layers, labels = generate(L, *args)
        out[name + "_args"] = np.array([args[0], args[1], args[3], args[4], int(args[5]), args[6], args[7]], float)
        out[name + "_ratios"] = np.asarray(args[2], np.uint64)
        for i, a in enumerate(layers):
            out["%s_layer%d" % (name, i)] = a
        for i, a in enumerate(labels):
            out["%s_labels%d" % (name, i)] = a
            s = np.zeros(a.size, np.uint64)
            L.ref_argsort(p(np.ascontiguousarray(a), u64p), a.size, p(s, u64p))
            out["%s_argsort%d" % (name, i)] = s
    rng = np.random.default_rng(3)
    a = rng.uniform(1, 100, (64, 7))
    b = rng.uniform(1, 100, (64, 7))
    out["euclid_a"], out["euclid_b"] = a, b
    out["euclid"] = np.array([L.ref_euclidean(p(np.ascontiguousarray(a[i]), f64p), p(np.ascontiguousarray(b[i]), f64p),
                                               7) for i in range(64)])
    data = rng.uniform(1, 100, (1500, 4))
    lab = np.zeros(1500, np.uint64)
    cen = np.zeros((30, 4))
    L.ref_kmeans(p(data, f64p), 1500, 4, 30, 7, p(lab, u64p), p(cen, f64p))
    out["km_data"], out["km_labels"], out["km_centroids"] = data, lab, cen
    np.savez_compressed(os.path.join(HERE, "ref_vectors.npz"), **out)
    print("wrote", os.path.join(HERE, "ref_vectors.npz"))
    return 0


if __name__ == "__main__":
    sys.exit(main())
