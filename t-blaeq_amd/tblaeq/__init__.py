"""ctypes binding of libtblaeq.so (the gfx950 HIP query path and its C++ host mirror).

The library is built in-tree (``make -C t-blaeq_amd``).  There is no fallback:
if the shared library is missing, importing any entry point raises.
"""
import atexit
import ctypes as C
import json
import os
import weakref

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "libtblaeq.so")
_LIB = None

u64p = C.POINTER(C.c_uint64)
f64p = C.POINTER(C.c_double)
u8p = C.POINTER(C.c_uint8)
f32p = C.POINTER(C.c_float)
vp = C.c_void_p

# Every function declared in include/tblaeq.h and include/tblaeq_handler.h.
SIGNATURES = {
    # tblaeq.h
    "tblaeq_ctx_create": (C.c_int, [C.c_int, vp, C.POINTER(vp)]),
    "tblaeq_ctx_destroy": (C.c_int, [vp]),
    "tblaeq_ctx_sync": (C.c_int, [vp]),
    "tblaeq_ctx_stream": (vp, [vp]),
    "tblaeq_status_string": (C.c_char_p, [C.c_int]),
    "tblaeq_last_error": (C.c_char_p, []),
    "tblaeq_ctx_set_timing": (C.c_int, [vp, C.c_int]),
    "tblaeq_ctx_set_tag": (C.c_int, [vp, C.c_int]),
    "tblaeq_ctx_timing": (C.c_int, [vp, C.c_int, C.c_int, u64p, f64p]),
    "tblaeq_ctx_timing_reset": (C.c_int, [vp]),
    "tblaeq_range_prune_compact": (C.c_int, [vp] * 13),
    "tblaeq_range_prune_compact_noalloc": (C.c_int, [vp, vp, vp, vp, vp, C.c_uint64, vp, vp, vp, vp, vp, vp, vp]),
    "tblaeq_knn_lower_bound_sort_step": (C.c_int, [vp, vp, vp, vp, vp, C.c_uint64, vp, C.c_uint64, vp, vp, vp, vp,
                                                   vp]),
    "tblaeq_prolong_refactor": (C.c_int, [vp, vp, vp, vp, vp, vp, C.c_uint64, vp, vp]),
    "tblaeq_prolong_refactor_packed": (C.c_int, [vp, vp, vp, vp, vp, vp, C.c_uint64, vp, vp, vp, vp]),
    "tblaeq_knn_lower_bounds": (C.c_int, [vp, C.c_uint64, vp, vp, C.c_uint64, vp, vp, vp]),
    "tblaeq_range_prune_compact_tile": (C.c_int, [vp, vp, vp, vp, C.c_uint64, vp, vp, vp, vp, vp, vp, vp]),
    "tblaeq_refine_range": (C.c_int, [vp, C.c_uint64, vp, vp, C.c_uint64, vp, vp, vp, C.c_uint64, u64p]),
    "tblaeq_refine_knn_topk": (C.c_int, [vp, C.c_uint64, vp, vp, C.c_uint64, vp, C.c_uint64, vp, vp, u64p]),
    "tblaeq_sort_pairs": (C.c_int, [vp, vp, vp, C.c_uint64]),
    # tblaeq_handler.h
    "tblaeq_handler_create_random": (C.c_int, [C.c_uint64, C.c_uint64, C.c_double, C.c_double, C.c_int, C.c_uint64,
                                               u64p, C.c_uint64, C.c_double, C.POINTER(vp)]),
    "tblaeq_handler_create_synthetic": (C.c_int, [C.c_uint64, C.c_uint64, C.c_double, C.c_double, C.c_int,
                                                  C.c_uint64, u64p, C.c_uint64, C.c_double, C.c_uint64,
                                                  C.POINTER(vp)]),
    "tblaeq_handler_data_rows": (C.c_int, [vp, u64p, C.c_uint64, f64p]),
    "tblaeq_handler_create_from_data": (C.c_int, [f64p, C.c_uint64, C.c_uint64, C.c_uint64, u64p, C.c_char_p,
                                                  C.POINTER(vp)]),
    "tblaeq_handler_load": (C.c_int, [C.c_char_p, C.POINTER(vp)]),
    # tblaeq_build.h
    "tblaeq_kmeans": (C.c_int, [vp, vp, C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64, vp, vp, u64p]),
    "tblaeq_kmeans_host": (C.c_int, [vp, f64p, C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64, f64p,
                                     u64p, u64p]),
    "tblaeq_build_level": (C.c_int, [vp, vp, C.c_uint64, C.c_uint64, vp, C.c_uint64, vp, vp, vp, vp, vp, vp]),
    "tblaeq_kmeans_options_init": (None, [vp]),
    "tblaeq_kmeans_ex": (C.c_int, [vp, vp, C.c_uint64, C.c_uint64, C.c_uint64, vp, vp, vp, u64p]),
    "tblaeq_convert_f32_f64": (C.c_int, [vp, vp, C.c_uint64, vp]),
    "tblaeq_handler_build_ex": (C.c_int, [vp, C.c_int, C.c_uint64, C.c_uint64, C.c_uint64, u64p, C.c_char_p,
                                          C.c_int, vp, C.POINTER(vp)]),
    "tblaeq_set_progress": (None, [vp, vp]),
    "tblaeq_handler_build": (C.c_int, [f64p, C.c_uint64, C.c_uint64, C.c_uint64, u64p, C.c_char_p, C.c_int,
                                       C.POINTER(vp)]),
    "tblaeq_handler_build_stats": (C.c_int, [vp, u64p, C.c_uint64, u64p]),
    "tblaeq_handler_save": (C.c_int, [vp, C.c_char_p]),
    "tblaeq_handler_destroy": (C.c_int, [vp]),
    "tblaeq_handler_meta": (C.c_int, [vp, u64p, u64p, C.c_uint64]),
    "tblaeq_handler_data": (C.c_int, [vp, f64p, C.c_uint64]),
    "tblaeq_handler_prepare": (C.c_int, [vp, u8p, C.c_uint64, C.c_uint64]),
    "tblaeq_handler_policy": (C.c_int, [vp, u8p, C.c_uint64]),
    "tblaeq_handler_recommend": (C.c_int, [vp, C.c_uint64, C.c_int, u8p, C.c_uint64]),
    "tblaeq_handler_ctx": (C.c_int, [vp, C.POINTER(vp)]),
    "tblaeq_handler_set_l3_tile_bytes": (C.c_int, [C.c_uint64]),
    "tblaeq_handler_query": (C.c_int, [vp, C.c_int, f64p, C.c_uint64, C.c_uint64, C.c_int, C.c_int, C.POINTER(vp)]),
    "tblaeq_handler_query_file": (C.c_int, [vp, C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_uint64, C.POINTER(vp)]),
    "tblaeq_dataset_load": (C.c_int, [C.c_char_p, u64p, u64p, f64p, C.c_uint64]),
    "tblaeq_query_file_load": (C.c_int, [C.c_char_p, C.c_int, u64p, u64p, f64p, C.c_uint64, C.c_char_p,
                                         C.c_uint64]),
    "tblaeq_result_destroy": (C.c_int, [vp]),
    "tblaeq_result_info": (C.c_int, [vp, u64p, u64p, f64p]),
    "tblaeq_result_query": (C.c_int, [vp, C.c_uint64, u64p, f64p, u64p, u64p]),
    "tblaeq_result_levels": (C.c_int, [vp, C.c_uint64, u64p, u64p]),
    "tblaeq_result_fine_mesh": (C.c_int, [vp, C.c_uint64, u64p, f64p]),
    "tblaeq_result_refined": (C.c_int, [vp, C.c_uint64, u64p, f64p]),
    "tblaeq_result_save_csv": (C.c_int, [vp, C.c_char_p]),
    "tblaeq_batch_create": (C.c_int, [vp, vp, C.c_uint64, vp, vp, C.c_uint64, vp, C.POINTER(vp)]),
    "tblaeq_batch_destroy": (C.c_int, [vp]),
    "tblaeq_batch_reserve": (C.c_int, [vp, C.c_uint64, C.c_uint64, C.c_uint64]),
    "tblaeq_batch_range": (C.c_int, [vp, vp, C.c_uint64, vp, vp, C.c_uint64, vp, vp, vp]),
    "tblaeq_batch_knn": (C.c_int, [vp, vp, vp, C.c_uint64, C.c_uint64, vp, vp, vp, vp, vp, vp]),
    "tblaeq_handler_batch_init": (C.c_int, [vp, C.c_uint64, C.c_uint64, C.c_uint64]),
    "tblaeq_handler_batch_set_queries": (C.c_int, [vp, C.c_int, f64p, u64p, C.c_uint64]),
    "tblaeq_handler_batch_run": (C.c_int, [vp, C.c_int, u64p]),
    "tblaeq_handler_batch_slot": (C.c_int, [vp, C.c_int]),
    "tblaeq_handler_batch_run_slots": (C.c_int, [vp, C.c_uint64, u64p]),
    "tblaeq_handler_batch_results": (C.c_int, [vp, u64p, u64p, f64p, u64p]),
    "tblaeq_handler_batch_stats": (C.c_int, [vp, u64p, u64p]),
    "tblaeq_handler_batch_free": (C.c_int, [vp]),
    "tblaeq_handler_batch_use_graph": (C.c_int, [vp, C.c_int]),
    "tblaeq_batch_range_graph": (C.c_int, [vp, vp, C.c_uint64, vp, vp, C.c_uint64, vp]),
    "tblaeq_batch_knn_graph": (C.c_int, [vp, vp, vp, C.c_uint64, C.c_uint64, vp, vp, vp, vp]),
    "tblaeq_batch_set_param": (C.c_int, [vp, C.c_int, C.c_uint64]),
    "tblaeq_batch_get_param": (C.c_int, [vp, C.c_int, u64p]),
    "tblaeq_handler_batch_set_param": (C.c_int, [vp, C.c_int, C.c_uint64]),
    "tblaeq_handler_batch_work": (C.c_int, [vp, u64p]),
    "tblaeq_handler_batch_leaf_tests": (C.c_int, [vp, u64p]),
    "tblaeq_batch_leaf_tests": (C.c_int, [vp, u64p]),
    "tblaeq_batch_knn_max_k": (C.c_int, [vp, u64p]),
    "tblaeq_batch_autotune": (C.c_int, [vp, vp, C.c_uint64, vp, vp, C.c_uint64, C.c_uint64, C.c_int, f64p]),
    "tblaeq_batch_save_params": (C.c_int, [vp, C.c_char_p]),
    "tblaeq_batch_load_params": (C.c_int, [vp, C.c_char_p]),
    "tblaeq_handler_batch_autotune": (C.c_int, [vp, f64p, C.c_uint64, f64p, u64p, C.c_uint64, C.c_int, f64p]),
    "tblaeq_handler_batch_save_params": (C.c_int, [vp, C.c_char_p]),
    "tblaeq_handler_batch_get_param": (C.c_int, [vp, C.c_int, u64p]),
    # tblaeq_learn.h
    "tblaeq_mlp_param_count": (C.c_uint64, []),
    "tblaeq_mlp_create": (C.c_int, [vp, C.c_uint64, C.POINTER(vp)]),
    "tblaeq_mlp_destroy": (C.c_int, [vp]),
    "tblaeq_mlp_set_params": (C.c_int, [vp, f32p]),
    "tblaeq_mlp_get_params": (C.c_int, [vp, f32p]),
    "tblaeq_mlp_set_norm": (C.c_int, [vp, C.c_double, C.c_double]),
    "tblaeq_mlp_get_norm": (C.c_int, [vp, f64p, f64p]),
    "tblaeq_mlp_set_loss_weights": (C.c_int, [vp, f32p]),
    "tblaeq_mlp_forward": (C.c_int, [vp, f32p, C.c_uint64, f32p]),
    "tblaeq_mlp_forward_device": (C.c_int, [vp, vp, C.c_uint64, vp]),
    "tblaeq_mlp_gradients": (C.c_int, [vp, f32p, f32p, C.c_uint64, f32p, f32p]),
    "tblaeq_mlp_train": (C.c_int, [vp, f32p, f32p, C.c_uint64, C.c_uint64, C.c_uint64, C.c_float, f32p]),
    "tblaeq_knn_features": (C.c_int, [vp, vp, vp, C.c_uint64, C.c_double, C.c_double, vp]),
    "tblaeq_mlp_propose_radius": (C.c_int, [vp, vp, C.c_uint64, C.c_double, C.c_double, vp]),
    "tblaeq_batch_knn_radius": (C.c_int, [vp, vp, vp, C.c_uint64, C.c_uint64, vp, vp, vp, vp, vp, vp, vp, vp]),
    "tblaeq_handler_batch_knn_learned": (C.c_int, [vp, vp, C.c_double, C.c_int, u64p]),
    "tblaeq_handler_knn_features": (C.c_int, [vp, f64p, u64p, C.c_uint64, C.c_double, C.c_double, f32p]),
    "tblaeq_handler_query_learned": (C.c_int, [vp, f64p, C.c_uint64, C.c_uint64, f64p, u64p, C.POINTER(vp)]),
    "tblaeq_knn_radius_compact": (C.c_int, [vp] * 13),
    # tblaeq_serve.h
    "tblaeq_server_config_init": (None, [vp]),
    "tblaeq_server_create": (C.c_int, [vp, vp, C.POINTER(vp)]),
    "tblaeq_server_destroy": (C.c_int, [vp]),
    "tblaeq_server_submit": (C.c_int, [vp, C.c_int, f64p, C.c_uint64, u64p]),
    "tblaeq_server_wait": (C.c_int, [vp, C.c_uint64, C.c_double]),
    "tblaeq_server_result": (C.c_int, [vp, C.c_uint64, u64p, C.c_uint64, f64p, u64p, C.c_int]),
    "tblaeq_server_stats": (C.c_int, [vp, f64p]),
    "tblaeq_server_reset_stats": (C.c_int, [vp]),
    "tblaeq_server_kind_stats": (C.c_int, [vp, C.c_int, f64p]),
    "tblaeq_server_query_json": (C.c_int, [vp, C.c_char_p, C.c_double, C.POINTER(C.c_void_p), C.POINTER(C.c_int)]),
    "tblaeq_free_string": (None, [C.c_void_p]),
    "tblaeq_handler_batch_load_params": (C.c_int, [vp, C.c_char_p]),
}

PARAM_KNN_SEED = 1
PARAM_KNN_WG_PER_CU = 2
PARAM_KNN_MODE = 3
PARAM_RANGE_BLOCKS_PER_CU = 4
PARAM_STEP_BLOCKS_PER_CU = 5
PARAM_RADIX_BLOCKS_PER_CU = 6
PARAM_HIT_STAGE = 7
PARAM_RANGE_ROW_QUERIES = 8
PARAM_KNN_ROW_QUERIES = 9
PARAM_KNN_LEAF_REFINE = 10


class TblaeqError(RuntimeError):
    pass


def lib():
    """Load libtblaeq.so (raises if it has not been built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise TblaeqError(f"{LIB_PATH} not built: run `make -C t-blaeq_amd` (or __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def check(rc, what):
    if rc != 0:
        L = lib()
        raise TblaeqError(f"{what}: {L.tblaeq_status_string(rc).decode()} ({L.tblaeq_last_error().decode()})")


def _p(a, t):
    return a.ctypes.data_as(t)


# Native objects still open when the interpreter exits are closed by an atexit
# hook, while the HIP runtime and every Python object they use still exist:
# servers first (their worker threads use the handler's index), then learned
# models and results, then handlers.  Without it, objects kept alive until
# shutdown (e.g. by a failed test's traceback) would be finalised in arbitrary
# order, possibly with server threads still running during runtime teardown.
_LIVE = {"server": weakref.WeakSet(), "mlp": weakref.WeakSet(), "result": weakref.WeakSet(),
         "handler": weakref.WeakSet()}


def _track(kind, obj):
    _LIVE[kind].add(obj)


@atexit.register
def close_all():
    """Closes every native object still open (servers, models, results, then handlers)."""
    for kind in ("server", "mlp", "result", "handler"):
        for obj in list(_LIVE[kind]):
            try:
                obj.close()
            except Exception:
                pass


class _Closing:
    """Context-manager support: `with Handler.random(...) as h:` closes on exit."""

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
        return False


LEVEL_POLICIES = {"L0": 0, "L1": 1, "L2": 2, "L3": 3}


class Handler(_Closing):
    """Mirror of the reference QueryHandler (src/core/QueryHandler.cuh:50-247)."""

    def __init__(self, h):
        self.h = h
        _track("handler", self)
        meta = np.zeros(4, np.uint64)
        ms = np.zeros(16, np.uint64)
        check(lib().tblaeq_handler_meta(self.h, _p(meta, u64p), _p(ms, u64p), 16), "meta")
        self.D, self.N, self.height, self.intervals = (int(x) for x in meta)
        self.mesh_sizes = [int(x) for x in ms[: self.height]]

    @classmethod
    def random(cls, N, D, vmin=1.0, vmax=100.0, is_int=False, ratios=(100, 50, 20), seed=12345, sigma_div=3.0):
        r = np.asarray(ratios, np.uint64)
        h = vp()
        check(lib().tblaeq_handler_create_random(N, D, vmin, vmax, int(is_int), len(r) + 1, _p(r, u64p), seed,
                                                 sigma_div, C.byref(h)), "create_random")
        return cls(h)

    @classmethod
    def synthetic(cls, N, D, vmin=1.0, vmax=100.0, is_int=False, ratios=(100, 50, 20), seed=12345, sigma_div=3.0,
                  threads=0):
        """The generator's hierarchy shape on host threads, from per-chunk streams (not the reference's
        sequence; independent of the thread count); for very large indexes."""
        r = np.asarray(ratios, np.uint64)
        h = vp()
        check(lib().tblaeq_handler_create_synthetic(N, D, vmin, vmax, int(is_int), len(r) + 1, _p(r, u64p), seed,
                                                    sigma_div, int(threads), C.byref(h)), "create_synthetic")
        return cls(h)

    @classmethod
    def from_data(cls, data, ratios=(100, 50, 20), name="in-memory"):
        data = np.ascontiguousarray(data, np.float64)
        r = np.asarray(ratios, np.uint64)
        h = vp()
        check(lib().tblaeq_handler_create_from_data(_p(data, f64p), data.shape[0], data.shape[1], len(r) + 1,
                                                    _p(r, u64p), name.encode(), C.byref(h)), "create_from_data")
        return cls(h)

    @classmethod
    def build(cls, data, ratios=(100, 50, 20), name="in-memory", force_cpu=False, **options):
        """QueryHandler(forceCPU, PointCloud, name, height, ratios): k-means build, on the GPU unless
        force_cpu (IndexBuilder.cu:44-151).  FP32 data, or k-means options (max_iter, seed, seeding 0 =
        k-means++ / 1 = k-means||, oversampling, rounds), go through tblaeq_handler_build_ex."""
        r = np.asarray(ratios, np.uint64)
        h = vp()
        if np.asarray(data).dtype == np.float32 or options:
            f32 = np.asarray(data).dtype == np.float32
            data = np.ascontiguousarray(data, np.float32 if f32 else np.float64)
            opt = KmeansOptions.make(**options)
            check(lib().tblaeq_handler_build_ex(data.ctypes.data_as(C.c_void_p), int(f32), data.shape[0],
                                                data.shape[1], len(r) + 1, _p(r, u64p), name.encode(),
                                                int(bool(force_cpu)), C.byref(opt), C.byref(h)), "handler_build_ex")
            return cls(h)
        data = np.ascontiguousarray(data, np.float64)
        check(lib().tblaeq_handler_build(_p(data, f64p), data.shape[0], data.shape[1], len(r) + 1, _p(r, u64p),
                                         name.encode(), int(bool(force_cpu)), C.byref(h)), "handler_build")
        return cls(h)

    def build_stats(self):
        """Per level of a build (finest first): Lloyd iterations, points re-scored in FP64, FP32 matrix-core
        score evaluations, milliseconds in the assignment kernels, inertia bits, k-means|| candidates (the
        host build fills iterations and inertia)."""
        n = C.c_uint64(0)
        check(lib().tblaeq_handler_build_stats(self.h, None, 0, C.byref(n)), "build_stats")
        out = np.zeros(n.value, np.uint64)
        check(lib().tblaeq_handler_build_stats(self.h, _p(out, u64p), n.value, C.byref(n)), "build_stats")
        return out.reshape(-1, 6)

    def build_inertia(self):
        """Per level of a build (finest first): the k-means inertia."""
        return self.build_stats()[:, 4].copy().view(np.float64)

    def kmeans(self, x, k, max_iter=7, seed=42):
        """tblaeq_kmeans on host arrays through this handler's context: (labels, centers, stats)."""
        x = np.ascontiguousarray(x, np.float64)
        n, D = x.shape
        cen = np.zeros((k, D))
        lab = np.zeros(n, np.uint64)
        st = np.zeros(4, np.uint64)
        check(lib().tblaeq_kmeans_host(self.ctx(), _p(x, f64p), n, D, k, max_iter, seed, _p(cen, f64p),
                                       _p(lab, u64p), _p(st, u64p)), "kmeans")
        return lab, cen, st

    @classmethod
    def load(cls, path):
        h = vp()
        check(lib().tblaeq_handler_load(path.encode(), C.byref(h)), "load")
        return cls(h)

    def save(self, path):
        check(lib().tblaeq_handler_save(self.h, path.encode()), "save")

    def data(self):
        out = np.zeros((self.N, self.D))
        check(lib().tblaeq_handler_data(self.h, _p(out, f64p), out.size), "data")
        return out

    def data_rows(self, rows):
        """Rows of the generated finest data (a sample of a large index without copying all of it)."""
        r = np.ascontiguousarray(rows, np.uint64)
        out = np.zeros((r.size, self.D))
        check(lib().tblaeq_handler_data_rows(self.h, _p(r, u64p), r.size, _p(out, f64p)), "data_rows")
        return out

    def prepare(self, policy=None, free_mem=0):
        if policy is None:
            check(lib().tblaeq_handler_prepare(self.h, None, 0, free_mem), "prepare")
        else:
            if isinstance(policy, str):
                policy = [policy] * self.intervals
            lv = np.asarray([LEVEL_POLICIES[p] if isinstance(p, str) else p for p in policy], np.uint8)
            check(lib().tblaeq_handler_prepare(self.h, _p(lv, u8p), lv.size, 0), "prepare")

    def policy(self):
        out = np.zeros(self.intervals, np.uint8)
        check(lib().tblaeq_handler_policy(self.h, _p(out, u8p), out.size), "policy")
        return ["L%d" % x for x in out]

    def recommend(self, free_mem, count_work_buffers=True):
        out = np.zeros(self.intervals, np.uint8)
        check(lib().tblaeq_handler_recommend(self.h, free_mem, int(count_work_buffers), _p(out, u8p), out.size),
              "recommend")
        return ["L%d" % x for x in out]

    def query(self, qtype, queries, K=0, refine=False, save_fine_mesh=False):
        """qtype 'range' (rows lo|hi) or 'knn' (rows = points)."""
        q = np.ascontiguousarray(queries, np.float64)
        if q.ndim == 1:
            q = q[None, :]
        t = 0 if qtype == "range" else 1
        r = vp()
        check(lib().tblaeq_handler_query(self.h, t, _p(q, f64p), q.shape[0], K, int(refine), int(save_fine_mesh),
                                         C.byref(r)), "query")
        return Result(r, self.D, self.intervals)

    def query_file(self, path, qtype, max_queries=2 ** 31 - 1, K=0, save_fine_mesh=False):
        """QueryHandler::performQuery(queryPath, ...): a query file through the per-query traversal."""
        r = vp()
        check(lib().tblaeq_handler_query_file(self.h, str(path).encode(), 0 if qtype == "range" else 1,
                                              int(save_fine_mesh), int(max_queries), int(K), C.byref(r)),
              "query_file")
        return Result(r, self.D, self.intervals)

    def range(self, lo, hi, refine=True, save_fine_mesh=True):
        return self.query("range", np.concatenate([np.asarray(lo, float), np.asarray(hi, float)]), 0, refine,
                          save_fine_mesh)

    def knn(self, q, K, refine=True, save_fine_mesh=True):
        return self.query("knn", np.asarray(q, float), K, refine, save_fine_mesh)

    # --- batched engine -------------------------------------------------------
    def batch_init(self, max_queries, pair_cap, hit_cap):
        check(lib().tblaeq_handler_batch_init(self.h, max_queries, pair_cap, hit_cap), "batch_init")

    # the uploaded batch of each slot: (type, count, largest K)
    @property
    def _bq(self):
        return self.__dict__.setdefault("_slots", {}).get(self.__dict__.get("_slot", 0))

    def batch_slot(self, slot):
        """Selects the batched engine (slot 0 or 1, each on its own stream) the batch calls act on."""
        check(lib().tblaeq_handler_batch_slot(self.h, int(slot)), "batch_slot")
        self._slot = int(slot)

    def batch_free(self):
        """Releases the selected slot's engine, its stream and scratch."""
        check(lib().tblaeq_handler_batch_free(self.h), "batch_free")
        self.__dict__.setdefault("_slots", {}).pop(self.__dict__.get("_slot", 0), None)

    def batch_run_slots(self, slots=(0, 1)):
        """Runs the uploaded batches of several slots concurrently; returns their totals."""
        mask = sum(1 << int(i) for i in slots)
        t = np.zeros(2, np.uint64)
        check(lib().tblaeq_handler_batch_run_slots(self.h, mask, _p(t, u64p)), "batch_run_slots")
        return [int(t[i]) for i in slots]

    def batch_set(self, qtype, queries, Ks=None):
        q = np.ascontiguousarray(queries, np.float64)
        if qtype == "range":
            check(lib().tblaeq_handler_batch_set_queries(self.h, 0, _p(q, f64p), None, q.shape[0]), "batch_set")
            kmax = 0
        else:
            k = np.ascontiguousarray(np.broadcast_to(np.asarray(Ks, np.uint64), (q.shape[0],)))
            kmax = int(k.max()) if k.size else 0
            check(lib().tblaeq_handler_batch_set_queries(self.h, 1, _p(q, f64p), _p(k, u64p), q.shape[0]),
                  "batch_set")
        self.__dict__.setdefault("_slots", {})[self.__dict__.get("_slot", 0)] = (qtype, q.shape[0], kmax)

    def batch_use_graph(self, on=True):
        check(lib().tblaeq_handler_batch_use_graph(self.h, int(on)), "batch_use_graph")

    def batch_autotune(self, boxes=None, points=None, Ks=None, reps=3):
        """Tunes the batched engine's launch parameters on sample queries; returns
        {'range_ms': (before, after), 'knn_ms': (before, after)}."""
        D = self.D
        b = np.ascontiguousarray(boxes if boxes is not None else np.zeros((0, 2 * D)), np.float64)
        p = np.ascontiguousarray(points if points is not None else np.zeros((0, D)), np.float64)
        k = np.ascontiguousarray(np.broadcast_to(np.asarray(Ks if Ks is not None else 0, np.uint64), (p.shape[0],)))
        out = np.zeros(4)
        check(lib().tblaeq_handler_batch_autotune(self.h, _p(b, f64p), b.shape[0], _p(p, f64p), _p(k, u64p),
                                                  p.shape[0], int(reps), _p(out, f64p)), "batch_autotune")
        return {"range_ms": (out[0], out[1]), "knn_ms": (out[2], out[3])}

    def batch_save_params(self, path):
        check(lib().tblaeq_handler_batch_save_params(self.h, str(path).encode()), "batch_save_params")

    def batch_load_params(self, path):
        check(lib().tblaeq_handler_batch_load_params(self.h, str(path).encode()), "batch_load_params")

    def batch_get_param(self, param):
        v = C.c_uint64(0)
        check(lib().tblaeq_handler_batch_get_param(self.h, int(param), C.byref(v)), "batch_get_param")
        return v.value

    def knn_features(self, points, Ks, vmin, scale):
        """Learned-radius features [n x 48] of host KNN queries (batch_init first)."""
        p = np.ascontiguousarray(points, np.float64)
        k = np.ascontiguousarray(np.broadcast_to(np.asarray(Ks, np.uint64), (p.shape[0],)))
        X = np.zeros((p.shape[0], 48), np.float32)
        check(lib().tblaeq_handler_knn_features(self.h, _p(p, f64p), _p(k, u64p), p.shape[0], float(vmin),
                                                float(scale), _p(X, f32p)), "knn_features")
        return X

    def query_learned(self, points, K, radii):
        """Per-query KNN with a learned radius per query (certified, reruns on a miss);
        returns (Result with refined top-K, number of reruns)."""
        q = np.ascontiguousarray(points, np.float64)
        if q.ndim == 1:
            q = q[None, :]
        r = np.ascontiguousarray(np.broadcast_to(np.asarray(radii, np.float64), (q.shape[0],)))
        fb = C.c_uint64(0)
        out = vp()
        check(lib().tblaeq_handler_query_learned(self.h, _p(q, f64p), q.shape[0], int(K), _p(r, f64p),
                                                 C.byref(fb), C.byref(out)), "query_learned")
        return Result(out, self.D, self.intervals), fb.value

    def batch_knn_learned(self, mlp, factor=1.0, stats=False):
        """Runs the uploaded KNN batch with radii proposed by `mlp`; returns the number of
        queries that failed certification and were rerun by the seed method."""
        f = C.c_uint64(0)
        check(lib().tblaeq_handler_batch_knn_learned(self.h, mlp.m, float(factor), int(stats), C.byref(f)),
              "batch_knn_learned")
        return f.value

    def batch_work(self):
        """(tests, candidate rows) of the last batch_run's final stage."""
        w = np.zeros(2, np.uint64)
        check(lib().tblaeq_handler_batch_work(self.h, _p(w, u64p)), "batch_work")
        return int(w[0]), int(w[1])

    def batch_leaf_tests(self):
        """Leaf x box tests of the exact filter in the last range (sub-)batch."""
        v = C.c_uint64(0)
        check(lib().tblaeq_handler_batch_leaf_tests(self.h, C.byref(v)), "batch_leaf_tests")
        return v.value

    def batch_set_param(self, param, value):
        """Launch parameter of the batched engine (PARAM_*); results do not depend on it."""
        check(lib().tblaeq_handler_batch_set_param(self.h, int(param), int(value)), "batch_set_param")

    def batch_run(self, stats=False):
        t = C.c_uint64(0)
        check(lib().tblaeq_handler_batch_run(self.h, int(stats), C.byref(t)), "batch_run")
        return t.value

    def batch_results(self):
        qtype, nq, K = self._bq
        if qtype == "range":
            off = np.zeros(nq + 1, np.uint64)
            check(lib().tblaeq_handler_batch_results(self.h, _p(off, u64p), None, None, None), "batch_results")
            ids = np.zeros(int(off[-1]), np.uint64)
            check(lib().tblaeq_handler_batch_results(self.h, _p(off, u64p), _p(ids, u64p), None, None),
                  "batch_results")
            return [ids[int(off[i]):int(off[i + 1])] for i in range(nq)]
        n = np.zeros(nq, np.uint64)
        ids = np.zeros(nq * K, np.uint64)
        d2 = np.zeros(nq * K)
        check(lib().tblaeq_handler_batch_results(self.h, None, _p(ids, u64p), _p(d2, f64p), _p(n, u64p)),
              "batch_results")
        return [(ids[i * K:i * K + int(n[i])], d2[i * K:i * K + int(n[i])]) for i in range(nq)]

    def batch_stats(self):
        nq = self._bq[1]
        s = np.zeros(nq * self.intervals, np.uint64)
        f = np.zeros(nq, np.uint64)
        check(lib().tblaeq_handler_batch_stats(self.h, _p(s, u64p), _p(f, u64p)), "batch_stats")
        return s.reshape(nq, self.intervals), f

    def ctx(self):
        c = vp()
        check(lib().tblaeq_handler_ctx(self.h, C.byref(c)), "handler_ctx")
        return c

    def close(self):
        """Destroys the index (servers still serving it are closed first)."""
        if self.h:
            for srv in list(_LIVE["server"]):
                if srv.handler is self:
                    srv.close()
            lib().tblaeq_handler_destroy(self.h)
            self.h = None

    def __del__(self):
        if getattr(self, "h", None) and _LIB is not None:
            _LIB.tblaeq_handler_destroy(self.h)
            self.h = None


def load_dataset(path):
    """loadFromFile (File.cu:8-50): the dataset text file as an N x D array."""
    n, d = C.c_uint64(0), C.c_uint64(0)
    check(lib().tblaeq_dataset_load(str(path).encode(), C.byref(n), C.byref(d), None, 0), "dataset_load")
    out = np.zeros((n.value, d.value))
    check(lib().tblaeq_dataset_load(str(path).encode(), C.byref(n), C.byref(d), _p(out, f64p), out.size),
          "dataset_load")
    return out


def load_query_file(path, qtype):
    """loadQueryRangeFromFile / loadQueryPointFromFile (File.cu:139-216): rows (lo|hi or points) and, for
    range files, the CSV label of the file name."""
    t = 0 if qtype == "range" else 1
    n, d = C.c_uint64(0), C.c_uint64(0)
    info = C.create_string_buffer(256)
    check(lib().tblaeq_query_file_load(str(path).encode(), t, C.byref(n), C.byref(d), None, 0, info, 256),
          "query_file_load")
    out = np.zeros((n.value, d.value * (2 if t == 0 else 1)))
    check(lib().tblaeq_query_file_load(str(path).encode(), t, C.byref(n), C.byref(d), _p(out, f64p), out.size, info,
                                       256), "query_file_load")
    return out, info.value.decode()


class Mlp(_Closing):
    """The learned-radius network (tblaeq_learn.h) on the handler's device."""

    def __init__(self, handler, seed=7):
        self.m = vp()
        self.handler = handler  # the handler outlives the network
        check(lib().tblaeq_mlp_create(handler.ctx(), int(seed), C.byref(self.m)), "mlp_create")
        self.n_params = int(lib().tblaeq_mlp_param_count())
        _track("mlp", self)

    def get_params(self):
        w = np.zeros(self.n_params, np.float32)
        check(lib().tblaeq_mlp_get_params(self.m, _p(w, f32p)), "mlp_get_params")
        return w

    def set_params(self, w):
        w = np.ascontiguousarray(w, np.float32)
        assert w.size == self.n_params
        check(lib().tblaeq_mlp_set_params(self.m, _p(w, f32p)), "mlp_set_params")

    def set_norm(self, vmin, scale):
        check(lib().tblaeq_mlp_set_norm(self.m, float(vmin), float(scale)), "mlp_set_norm")

    def set_loss_weights(self, w4):
        w = np.ascontiguousarray(w4, np.float32)
        check(lib().tblaeq_mlp_set_loss_weights(self.m, _p(w, f32p)), "mlp_set_loss_weights")

    def forward(self, X):
        X = np.ascontiguousarray(X, np.float32)
        Y = np.zeros((X.shape[0], 4), np.float32)
        check(lib().tblaeq_mlp_forward(self.m, _p(X, f32p), X.shape[0], _p(Y, f32p)), "mlp_forward")
        return Y

    def gradients(self, X, T):
        X = np.ascontiguousarray(X, np.float32)
        T = np.ascontiguousarray(T, np.float32)
        g = np.zeros(self.n_params, np.float32)
        loss = np.zeros(1, np.float32)
        check(lib().tblaeq_mlp_gradients(self.m, _p(X, f32p), _p(T, f32p), X.shape[0], _p(g, f32p),
                                         _p(loss, f32p)), "mlp_gradients")
        return g, float(loss[0])

    def train(self, X, T, epochs=10, batch=1024, lr=1e-3):
        X = np.ascontiguousarray(X, np.float32)
        T = np.ascontiguousarray(T, np.float32)
        losses = np.zeros(epochs, np.float32)
        check(lib().tblaeq_mlp_train(self.m, _p(X, f32p), _p(T, f32p), X.shape[0], int(epochs), int(batch),
                                     float(lr), _p(losses, f32p)), "mlp_train")
        return losses

    def close(self):
        if self.m:
            lib().tblaeq_mlp_destroy(self.m)
            self.m = None

    def __del__(self):
        if getattr(self, "m", None) and _LIB is not None:
            _LIB.tblaeq_mlp_destroy(self.m)
            self.m = None


class KmeansOptions(C.Structure):
    """tblaeq_kmeans_options (tblaeq_build.h)."""
    _fields_ = [("max_iter", C.c_uint64), ("seed", C.c_uint64), ("seeding", C.c_int), ("oversampling", C.c_double),
                ("rounds", C.c_uint64)]

    @classmethod
    def make(cls, **kw):
        o = cls()
        lib().tblaeq_kmeans_options_init(C.byref(o))
        for k, v in kw.items():
            setattr(o, k, v)
        return o


_PROGRESS_FN = C.CFUNCTYPE(None, C.c_char_p, C.c_void_p)
_progress_keep = None


def set_progress(fn):
    """fn(message) receives the library's progress messages (long builds); None turns them off."""
    global _progress_keep
    if fn is None:
        lib().tblaeq_set_progress(None, None)
        _progress_keep = None
        return
    cb = _PROGRESS_FN(lambda msg, user: fn(msg.decode()))
    _progress_keep = cb
    lib().tblaeq_set_progress(C.cast(cb, C.c_void_p), None)


class ServerConfig(C.Structure):
    """tblaeq_server_config (tblaeq_serve.h)."""
    _fields_ = [("max_batch", C.c_uint64), ("max_k", C.c_uint64), ("pair_cap", C.c_uint64), ("hit_cap", C.c_uint64),
                ("hbm_fraction", C.c_double), ("range_request_bytes", C.c_uint64),
                ("knn_request_bytes", C.c_uint64), ("max_wait_us", C.c_double)]


class Server(_Closing):
    """Continuous-batching serving loop over a handler's index (tblaeq_serve.h).
    Close it before the handler."""

    def __init__(self, handler, **cfg):
        c = ServerConfig()
        lib().tblaeq_server_config_init(C.byref(c))
        for k, v in cfg.items():
            setattr(c, k, v)
        self.handler = handler
        self.D = handler.D
        self.s = vp()
        check(lib().tblaeq_server_create(handler.h, C.byref(c), C.byref(self.s)), "server_create")
        _track("server", self)

    def submit(self, qtype, q, K=0):
        q = np.ascontiguousarray(q, np.float64)
        t = C.c_uint64(0)
        check(lib().tblaeq_server_submit(self.s, 0 if qtype == "range" else 1, _p(q, f64p), int(K), C.byref(t)),
              "server_submit")
        return t.value

    def wait(self, ticket, timeout=-1.0):
        check(lib().tblaeq_server_wait(self.s, ticket, float(timeout)), "server_wait")

    def result(self, ticket, release=True):
        n = C.c_uint64(0)
        check(lib().tblaeq_server_result(self.s, ticket, None, 0, None, C.byref(n), 0), "server_result")
        ids = np.zeros(n.value, np.uint64)
        d2 = np.zeros(n.value)
        check(lib().tblaeq_server_result(self.s, ticket, _p(ids, u64p), n.value, _p(d2, f64p), C.byref(n),
                                         int(release)), "server_result")
        return ids, d2

    def release(self, ticket):
        """Forgets a finished request without copying its result."""
        n = C.c_uint64(0)
        check(lib().tblaeq_server_result(self.s, ticket, None, 0, None, C.byref(n), 1), "server_result")
        return n.value

    def stats(self):
        o = np.zeros(8)
        check(lib().tblaeq_server_stats(self.s, _p(o, f64p)), "server_stats")
        keys = ("completed", "p50_us", "p99_us", "batches", "mean_batch", "deferred", "peak_inflight_bytes",
                "budget_bytes")
        return dict(zip(keys, o.tolist()))

    def reset_stats(self):
        check(lib().tblaeq_server_reset_stats(self.s), "server_reset_stats")

    def kind_stats(self, qtype):
        """Per request kind: completed, batches, reruns after overflow, per-query traversals, busy us, hits per
        range request."""
        o = np.zeros(6)
        check(lib().tblaeq_server_kind_stats(self.s, 0 if qtype == "range" else 1, _p(o, f64p)), "server_kind_stats")
        return dict(zip(("completed", "batches", "reruns", "traversals", "busy_us", "hits_per_request"), o.tolist()))

    def query_json(self, request, timeout=-1.0):
        """The reference's POST /query body (a JSON text or a dict) through the server: (HTTP status, response
        dict)."""
        if not isinstance(request, (str, bytes)):
            request = json.dumps(request)
        if isinstance(request, str):
            request = request.encode()
        out = C.c_void_p()
        status = C.c_int(0)
        check(lib().tblaeq_server_query_json(self.s, request, float(timeout), C.byref(out), C.byref(status)),
              "server_query_json")
        try:
            text = C.string_at(out.value).decode()
        finally:
            lib().tblaeq_free_string(out)
        return status.value, json.loads(text)

    def close(self):
        if self.s:
            lib().tblaeq_server_destroy(self.s)
            self.s = None

    def __del__(self):
        if getattr(self, "s", None) and _LIB is not None:
            _LIB.tblaeq_server_destroy(self.s)
            self.s = None


class Result(_Closing):
    """QueryResult of the plain traversal (host copies of the per-query results)."""

    def __init__(self, h, D, L):
        self.h = h
        self.D = D
        self.L = L
        nq = C.c_uint64(0)
        check(lib().tblaeq_result_info(self.h, C.byref(nq), None, None), "result_info")
        self.nq = nq.value
        _track("result", self)

    def close(self):
        if self.h:
            lib().tblaeq_result_destroy(self.h)
            self.h = None

    def info(self, i):
        lv = np.zeros(self.L, np.uint64)
        us = C.c_double(0)
        fs = C.c_uint64(0)
        rc = C.c_uint64(0)
        check(lib().tblaeq_result_query(self.h, i, _p(lv, u64p), C.byref(us), C.byref(fs), C.byref(rc)), "result")
        return {"levels": lv, "us": us.value, "fine": fs.value, "refined": rc.value}

    def levels(self, i=0):
        s = np.zeros(self.L, np.uint64)
        o = np.zeros(self.L, np.uint64)
        check(lib().tblaeq_result_levels(self.h, i, _p(s, u64p), _p(o, u64p)), "result_levels")
        return s, o

    def fine_mesh(self, i=0):
        n = self.info(i)["fine"]
        ids = np.zeros(n, np.uint64)
        vals = np.zeros((n, self.D))
        check(lib().tblaeq_result_fine_mesh(self.h, i, _p(ids, u64p), _p(vals, f64p)), "fine_mesh")
        return ids, vals

    def refined(self, i=0):
        n = self.info(i)["refined"]
        ids = np.zeros(n, np.uint64)
        d2 = np.zeros(n)
        check(lib().tblaeq_result_refined(self.h, i, _p(ids, u64p), _p(d2, f64p)), "refined")
        return ids, d2

    def save_csv(self, path):
        check(lib().tblaeq_result_save_csv(self.h, path.encode()), "save_csv")

    def __del__(self):
        if getattr(self, "h", None) and _LIB is not None:
            _LIB.tblaeq_result_destroy(self.h)
            self.h = None


KERNEL_KINDS = {"compact_range": 0, "compact_knn": 1, "compact_box": 2, "prolong": 3, "tilesort": 4, "merge": 5,
                "step": 6, "knn_lb": 7}


class Timing:
    """HIP-event timing of the kernels a context launches (on its own stream)."""

    def __init__(self, ctx):
        self.ctx = ctx

    def enable(self, on=True):
        check(lib().tblaeq_ctx_set_timing(self.ctx, int(on)), "set_timing")

    def reset(self):
        check(lib().tblaeq_ctx_timing_reset(self.ctx), "timing_reset")

    def tag(self, t):
        check(lib().tblaeq_ctx_set_tag(self.ctx, t), "set_tag")

    def get(self, kind, tag=-1):
        k = KERNEL_KINDS.get(kind, kind)
        n = C.c_uint64(0)
        ms = C.c_double(0)
        check(lib().tblaeq_ctx_timing(self.ctx, k, tag, C.byref(n), C.byref(ms)), "timing")
        return n.value, ms.value
