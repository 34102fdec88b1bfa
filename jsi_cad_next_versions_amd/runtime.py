"""Python host side of libjsi_mi355x.so: a thin ctypes layer over the C ABI
declared in include/jsi_mi355x.h.

It mirrors the reference's query API (JSI_CAD_NEXT_VERSION5/include/jsi_cad.hpp):
``Context.minimum_distance(s1, r1, s2, r2)`` is minimum_distance,
``Context.intersection(...)`` is intersection, ``Context.eval_grid(...)`` is
evalWithCuda.  There is no CPU fallback: if the library is missing or no GPU
is present, loading raises.
"""
from __future__ import annotations

import ctypes as C
import json
import os

import numpy as np

from ._abi import (CURVE_DTYPE, JSI_ERR_BUSY, JSI_OK, MAX_ROUNDS, QUERY_DISTANCE, QUERY_INTERSECT,  # noqa: F401
                   CurvePoint, Pair, Params, Query, Region, Result, curve_points_to_numpy, default_params,
                   pairs_to_numpy)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libjsi_mi355x.so")
# autotuned launch choices per (degree, nuv) (autotune.py writes it; jsi_tune_set)
TUNING_PATH = os.path.join(HERE, "tuning.json")
TUNING_KNOBS = ("seed_pairs", "filter_walk", "r1_scale", "r1_seeds")

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")

# Every extern "C" symbol of include/jsi_mi355x.h (test_library_exports checks them).
EXPORTS = [
    "jsi_params_default", "jsi_status_string", "jsi_last_error", "jsi_ctx_create", "jsi_ctx_destroy",
    "jsi_ctx_sync", "jsi_ctx_profile", "jsi_ctx_kernel_stats", "jsi_surface_upload", "jsi_surface_register_device", "jsi_surface_release",
    "jsi_eval_grid_f32", "jsi_eval_grid_f64", "jsi_min_distance", "jsi_min_distance_noalloc", "jsi_intersect",
    "jsi_intersect_curves", "jsi_run_batch",
    "jsi_server_create", "jsi_server_destroy", "jsi_server_surface", "jsi_server_submit", "jsi_server_poll",
    "jsi_server_counters", "jsi_tune_set", "jsi_tune_clear",
]
# include/jsi_ascend_hotspots.h
OP_EXPORTS = ["jsi_ascend_distance_topk_mask", "jsi_ascend_intersection_mask"]

_lib = None


class JsiError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"jsi status {status}: {message}")
        self.status = status


def load_library(path: str = LIB_PATH):
    """Loads libjsi_mi355x.so (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it with python -m jsi_cad_next_versions_amd.build "
                           "(or __graft_entry__.build())")
    L = C.CDLL(path)
    L.jsi_params_default.argtypes = [C.POINTER(Params)]
    L.jsi_params_default.restype = None
    L.jsi_status_string.argtypes = [C.c_int]
    L.jsi_status_string.restype = C.c_char_p
    L.jsi_last_error.restype = C.c_char_p
    L.jsi_ctx_create.argtypes = [C.c_int, C.c_size_t, C.POINTER(C.c_void_p)]
    L.jsi_ctx_destroy.argtypes = [C.c_void_p]
    L.jsi_ctx_sync.argtypes = [C.c_void_p]
    L.jsi_ctx_profile.argtypes = [C.c_void_p, C.c_int]
    L.jsi_ctx_kernel_stats.argtypes = [C.c_void_p, C.c_char_p, C.POINTER(C.c_int64), C.POINTER(C.c_double),
                                       C.POINTER(C.c_double)]
    L.jsi_surface_upload.argtypes = [C.c_void_p, _i32p, _f32p, C.c_int, C.POINTER(C.c_int)]
    L.jsi_surface_register_device.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.POINTER(C.c_int)]
    L.jsi_surface_release.argtypes = [C.c_void_p, C.c_int]
    L.jsi_eval_grid_f32.argtypes = [C.c_void_p, C.c_int, C.c_float, C.c_float, C.c_float, C.c_float, C.c_int, _f32p]
    L.jsi_eval_grid_f64.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int,
                                    _f64p]
    q = [C.c_void_p, C.c_int, C.POINTER(Region), C.c_int, C.POINTER(Region), C.POINTER(Params), C.POINTER(Result),
         C.c_void_p, C.c_int]
    L.jsi_min_distance.argtypes = q
    L.jsi_min_distance_noalloc.argtypes = q
    L.jsi_intersect.argtypes = q
    L.jsi_intersect_curves.argtypes = q[:6] + [C.c_double, C.c_double, C.POINTER(Result), C.c_void_p, C.c_int,
                                               C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.jsi_run_batch.argtypes = [C.c_void_p, C.c_int, C.POINTER(Query), C.POINTER(Result), C.POINTER(C.c_int),
                                C.c_int]
    L.jsi_server_create.argtypes = [C.c_void_p, C.c_int, C.c_size_t, C.c_int, C.c_size_t, C.POINTER(C.c_void_p)]
    L.jsi_server_destroy.argtypes = [C.c_void_p]
    L.jsi_server_surface.argtypes = [C.c_void_p, C.c_uint64, _i32p, _f32p, C.c_int, C.POINTER(C.c_int)]
    L.jsi_server_submit.argtypes = [C.c_void_p, C.POINTER(Query), C.POINTER(C.c_int64)]
    L.jsi_server_poll.argtypes = [C.c_void_p, C.c_int64, C.POINTER(Result), C.c_int]
    L.jsi_server_counters.argtypes = [C.c_void_p] + [C.POINTER(C.c_int64)] * 5
    L.jsi_tune_set.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_char_p, C.c_int]
    L.jsi_tune_clear.argtypes = [C.c_void_p]
    L.jsi_ascend_distance_topk_mask.argtypes = [_f32p, C.c_size_t, _f32p, C.c_size_t, C.c_int, _i32p, _i32p,
                                                C.POINTER(C.c_float)]
    L.jsi_ascend_intersection_mask.argtypes = [_f32p, _f32p, C.c_int, _f32p, _f32p, C.c_int, _i32p, _i32p]
    _lib = L
    return L


def _check(L, rc):
    if rc != JSI_OK:
        raise JsiError(rc, L.jsi_last_error().decode(errors="replace"))


class Context:
    """One device context with resident surfaces (jsi_ctx)."""

    def __init__(self, device: int = 0, hbm_cap_bytes: int = 0, tuning: bool = True):
        """tuning: load the autotuned launch table (TUNING_PATH, written by
        autotune.py) unless JSI_TUNING=0; it changes launch shapes only."""
        self.L = load_library()
        h = C.c_void_p()
        _check(self.L, self.L.jsi_ctx_create(device, hbm_cap_bytes, C.byref(h)))
        self.h = h
        self._handles = {}
        if tuning and os.environ.get("JSI_TUNING", "1") != "0" and os.path.exists(TUNING_PATH):
            self.load_tuning(TUNING_PATH)

    def tune_set(self, degree, nuv, knob, value):
        """One autotuned choice for queries of this (larger) degree and round-1 nuv."""
        _check(self.L, self.L.jsi_tune_set(self.h, int(degree), int(nuv), knob.encode(), int(value)))

    def tune_clear(self):
        """Back to the built-in launch choices."""
        _check(self.L, self.L.jsi_tune_clear(self.h))

    def load_tuning(self, path):
        """Loads a table written by autotune.py: entries of (degree, nuv, knobs)."""
        with open(path) as f:
            table = json.load(f)
        for e in table["entries"]:
            for knob in TUNING_KNOBS:
                if knob in e["choice"]:
                    self.tune_set(e["degree"], e["nuv"], knob, e["choice"][knob])

    def close(self):
        if self.h:
            self.L.jsi_ctx_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sync(self):
        _check(self.L, self.L.jsi_ctx_sync(self.h))

    def profile(self, enable=True):
        _check(self.L, self.L.jsi_ctx_profile(self.h, int(enable)))

    def kernel_stats(self, name):
        """(launches, summed ms, summed algorithmic units) of a profiled kernel."""
        n, ms, u = C.c_int64(), C.c_double(), C.c_double()
        _check(self.L, self.L.jsi_ctx_kernel_stats(self.h, name.encode(), C.byref(n), C.byref(ms), C.byref(u)))
        return n.value, ms.value, u.value

    def upload(self, s) -> int:
        """Uploads a Surface once; later calls with the same object reuse it."""
        key = id(s)
        if key in self._handles and self._handles[key][0] is s:
            return self._handles[key][1]
        h = C.c_int()
        _check(self.L, self.L.jsi_surface_upload(self.h, s.ivec, s.data, s.ndata, C.byref(h)))
        self._handles[key] = (s, h.value)
        return h.value

    def release(self, s):
        key = id(s)
        if key in self._handles:
            _check(self.L, self.L.jsi_surface_release(self.h, self._handles.pop(key)[1]))

    def eval_grid(self, s, ustep, vstep, ustart, vstart, nuv, dtype=np.float32):
        """evalWithCuda: 18 values per grid point, point (idx, idy) at offset idx + idy*nuv."""
        h = self.upload(s)
        out = np.empty(nuv * nuv * 18, dtype)
        if dtype == np.float32:
            _check(self.L, self.L.jsi_eval_grid_f32(self.h, h, ustep, vstep, ustart, vstart, nuv, out))
        else:
            _check(self.L, self.L.jsi_eval_grid_f64(self.h, h, ustep, vstep, ustart, vstart, nuv, out))
        return out

    def _query(self, fn, s1, r1, s2, r2, params, cap):
        h1, h2 = self.upload(s1), self.upload(s2)
        res = Result()
        buf = (Pair * max(cap, 1))()
        _check(self.L, fn(self.h, h1, C.byref(r1), h2, C.byref(r2), C.byref(params), C.byref(res),
                          C.cast(buf, C.c_void_p), cap))
        n = int(min(cap, res.npairs))
        return res, pairs_to_numpy(buf, n)

    def minimum_distance(self, s1, r1, s2, r2, params=None, cap=None):
        """Returns (Result, last-round top-k pairs)."""
        params = params or default_params()
        cap = params.k if cap is None else cap
        return self._query(self.L.jsi_min_distance, s1, r1, s2, r2, params, cap)

    def minimum_distance_noalloc(self, s1, r1, s2, r2, params=None, cap=None):
        """minimum_distance through the allocation-free entry point."""
        params = params or default_params()
        cap = params.k if cap is None else cap
        return self._query(self.L.jsi_min_distance_noalloc, s1, r1, s2, r2, params, cap)

    def intersection(self, s1, r1, s2, r2, params=None, cap=0):
        """Returns (Result, last-round overlap pairs (up to cap))."""
        params = params or default_params()
        return self._query(self.L.jsi_intersect, s1, r1, s2, r2, params, cap)

    def intersection_curves(self, s1, r1, s2, r2, params=None, spacing=0.0, tol=0.0, cap=1 << 20):
        """Intersection rounds, then the FP64 curve samples: (Result, ordered
        samples as a numpy record array, number of curves)."""
        params = params or default_params()
        h1, h2 = self.upload(s1), self.upload(s2)
        res = Result()
        dt = np.dtype(CURVE_DTYPE)
        assert dt.itemsize == C.sizeof(CurvePoint)
        buf = np.empty(max(cap, 1), dtype=dt)  # not cleared: the library writes the first min(cap, n) records
        n, nc = C.c_int(), C.c_int()
        _check(self.L, self.L.jsi_intersect_curves(self.h, h1, C.byref(r1), h2, C.byref(r2), C.byref(params),
                                                   spacing, tol, C.byref(res), C.c_void_p(buf.ctypes.data), cap,
                                                   C.byref(n), C.byref(nc)))
        return res, buf[: min(cap, n.value)].copy(), nc.value

    def query(self, kind, s1, r1, s2, r2, params=None):
        """A jsi_query for run_batch / Server.submit (uploads the surfaces)."""
        return Query(kind, self.upload(s1), self.upload(s2), r1, r2, params or default_params())

    def run_batch(self, queries, concurrency=8):
        """Runs the queries with up to `concurrency` in flight; returns (results, statuses)."""
        n = len(queries)
        qa = (Query * max(n, 1))(*queries)
        ra = (Result * max(n, 1))()
        sa = (C.c_int * max(n, 1))()
        self.L.jsi_run_batch(self.h, n, qa, ra, sa, concurrency)
        return [ra[i] for i in range(n)], [sa[i] for i in range(n)]


class Server:
    """Serving loop (jsi_server): `slots` worker streams, admission by device
    memory, an LRU cache of resident surfaces keyed by the caller."""

    def __init__(self, ctx: Context, slots=8, hbm_cap_bytes=0, max_pending=0, cache_cap_bytes=0):
        self.ctx, self.L = ctx, ctx.L
        h = C.c_void_p()
        _check(self.L, self.L.jsi_server_create(ctx.h, slots, hbm_cap_bytes, max_pending, cache_cap_bytes,
                                                C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            self.L.jsi_server_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def surface(self, key, s):
        h = C.c_int()
        _check(self.L, self.L.jsi_server_surface(self.h, key, s.ivec, s.data, s.ndata, C.byref(h)))
        return h.value

    def submit(self, query):
        """Returns a ticket, or None when admission refuses (JSI_ERR_BUSY)."""
        t = C.c_int64()
        rc = self.L.jsi_server_submit(self.h, C.byref(query), C.byref(t))
        if rc == JSI_ERR_BUSY:
            return None
        _check(self.L, rc)
        return t.value

    def poll(self, ticket, wait=True):
        """Result when done, None while pending."""
        res = Result()
        rc = self.L.jsi_server_poll(self.h, ticket, C.byref(res), int(wait))
        if rc == 0:
            return None
        if rc < 0:
            raise JsiError(rc, self.L.jsi_last_error().decode(errors="replace"))
        return res

    def counters(self):
        v = [C.c_int64() for _ in range(5)]
        _check(self.L, self.L.jsi_server_counters(self.h, *[C.byref(x) for x in v]))
        return dict(zip(["admitted", "refused", "completed", "cache_hits", "cache_evictions"], [x.value for x in v]))


# ---------------------------------------------------------------------------
# Hotspot operator ABI (include/jsi_ascend_hotspots.h)
# ---------------------------------------------------------------------------
def distance_topk_mask(c1, c2, topk):
    """k-th smallest centre distance and the endpoint masks
    (jsi_ascend_distance_topk_mask).  c1, c2: float32 [n, 3]."""
    L = load_library()
    c1 = np.ascontiguousarray(c1, np.float32).reshape(-1)
    c2 = np.ascontiguousarray(c2, np.float32).reshape(-1)
    m1 = np.zeros(c1.size // 3, np.int32)
    m2 = np.zeros(c2.size // 3, np.int32)
    kth = C.c_float()
    rc = L.jsi_ascend_distance_topk_mask(c1, c1.size // 3, c2, c2.size // 3, int(topk), m1, m2, C.byref(kth))
    if rc != 0:
        raise JsiError(rc, "jsi_ascend_distance_topk_mask failed")
    return kth.value, m1, m2


def intersection_mask(c1, r1, c2, r2):
    L = load_library()
    a = [np.ascontiguousarray(x, np.float32).reshape(-1) for x in (c1, r1, c2, r2)]
    m1 = np.zeros(a[0].size // 3, np.int32)
    m2 = np.zeros(a[2].size // 3, np.int32)
    rc = L.jsi_ascend_intersection_mask(a[0], a[1], a[0].size // 3, a[2], a[3], a[2].size // 3, m1, m2)
    if rc != 0:
        raise JsiError(rc, "jsi_ascend_intersection_mask failed")
    return m1, m2
