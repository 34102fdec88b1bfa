"""ctypes mirrors of the C ABI structs in include/jsi_mi355x.h.

The field order and types must match the header exactly; test_abi_layout
checks the sizes against the values the library reports.
"""
import ctypes as C

JSI_OK = 0
JSI_ERR_INVALID = -1
JSI_ERR_NOMEM = -2
JSI_ERR_DEVICE = -3
JSI_ERR_CAPACITY = -4
JSI_ERR_UNSUPPORTED = -5
JSI_ERR_BUSY = -6

PREC_EXACT32 = 0
PREC_FP64 = 1
PREC_FP32SCREEN = 2
PRECISIONS = {"exact32": PREC_EXACT32, "fp64": PREC_FP64, "fp32screen": PREC_FP32SCREEN}

MAX_ROUNDS = 16
MAX_NUV = 1025
MAX_K = 8192


class Region(C.Structure):
    """jsi_region: the value part of EvalAndConstructTask (jsi_cad.hpp:11-23)."""
    _fields_ = [
        ("ustart", C.c_float),
        ("vstart", C.c_float),
        ("ustop", C.c_float),
        ("vstop", C.c_float),
        ("nuv", C.c_int),
        ("side", C.c_int),
    ]


class Params(C.Structure):
    _fields_ = [
        ("k", C.c_int),
        ("dist_refine_nuv", C.c_int),
        ("dist_max_rounds", C.c_int),
        ("isect_max_rounds", C.c_int),
        ("isect_refine_nuv", C.c_int),
        ("isect_halve_above", C.c_int),
        ("isect_min_nuv", C.c_int),
        ("use_compute_k", C.c_int),
        ("precision", C.c_int),
        ("cull", C.c_int),
        ("trace", C.c_int),
        ("cand_cap", C.c_int),
        ("polish", C.c_int),
        ("screen_out", C.c_int),
        ("param_tol", C.c_double),
    ]


class RoundInfo(C.Structure):
    _fields_ = [
        ("npairs", C.c_int64),
        ("slots", C.c_int64),
        ("units1", C.c_int64),
        ("units2", C.c_int64),
        ("kth_d2", C.c_double),
        ("hash", C.c_uint64),
        ("nuv", C.c_int),
        ("pad_", C.c_int),
    ]


class Closest(C.Structure):
    """jsi_closest: FP64 closest points after the polish (jsi_params.polish > 0)."""
    _fields_ = [
        ("distance", C.c_double),
        ("u1", C.c_double),
        ("v1", C.c_double),
        ("u2", C.c_double),
        ("v2", C.c_double),
        ("p1", C.c_double * 3),
        ("p2", C.c_double * 3),
        ("candidate", C.c_int32),
        ("iterations", C.c_int32),
        ("polished", C.c_int32),
        ("flags", C.c_int32),
    ]

    def as_dict(self):
        return dict(distance=self.distance, u1=self.u1, v1=self.v1, u2=self.u2, v2=self.v2, p1=tuple(self.p1),
                    p2=tuple(self.p2), candidate=self.candidate, iterations=self.iterations,
                    polished=self.polished, flags=self.flags)


POLISH_STATIONARY = 1
POLISH_STALLED = 2
POLISH_FIXED = 4
POLISH_CONVERGED = 8


class Result(C.Structure):
    _fields_ = [
        ("distance", C.c_double),
        ("npairs", C.c_int64),
        ("rounds", C.c_int),
        ("isect_refine_nuv", C.c_int),
        ("ms_total", C.c_double),
        ("ms_eval", C.c_double),
        ("ms_screen", C.c_double),
        ("round", RoundInfo * MAX_ROUNDS),
        ("closest", Closest),
        ("isect_k_raw", C.c_double * 2),
        ("isect_k_scale", C.c_double * 2),
    ]

    def rounds_list(self):
        out = []
        for r in range(self.rounds):
            ri = self.round[r]
            out.append(dict(npairs=ri.npairs, slots=ri.slots, units1=ri.units1, units2=ri.units2,
                            kth_d2=ri.kth_d2, hash=ri.hash, nuv=ri.nuv))
        return out


class Pair(C.Structure):
    _fields_ = [("d2", C.c_double), ("slot", C.c_int32), ("i", C.c_int32), ("j", C.c_int32), ("pad_", C.c_int32)]


def default_params(**overrides):
    """VERSION5 defaults (distance_analysis.cu:33-36, intersection.cu:36,368-370)."""
    p = Params()
    p.k = 8
    p.dist_refine_nuv = 16
    p.dist_max_rounds = 4
    p.isect_max_rounds = 6
    p.isect_refine_nuv = 16
    p.isect_halve_above = 512
    p.isect_min_nuv = 4
    p.use_compute_k = 1
    p.precision = PREC_EXACT32
    p.cull = 1
    p.trace = 0
    p.cand_cap = 0
    p.polish = 0
    p.param_tol = 0.0
    for key, val in overrides.items():
        if key == "precision" and isinstance(val, str):
            val = PRECISIONS[val]
        setattr(p, key, float(val) if key == "param_tol" else int(val))
    return p


def pairs_to_numpy(buf, n):
    import numpy as np
    dt = np.dtype([("d2", "<f8"), ("slot", "<i4"), ("i", "<i4"), ("j", "<i4"), ("pad_", "<i4")])
    return np.frombuffer(bytes(memoryview(buf).cast("B")[: n * C.sizeof(Pair)]), dtype=dt).copy()


class CurvePoint(C.Structure):
    """jsi_curve_point: one ordered sample of an intersection curve."""
    _fields_ = [("x", C.c_double), ("y", C.c_double), ("z", C.c_double), ("u1", C.c_double), ("v1", C.c_double),
                ("u2", C.c_double), ("v2", C.c_double), ("residual", C.c_double), ("curve", C.c_int32),
                ("flags", C.c_int32)]


CURVE_CLOSED = 1
CURVE_DTYPE = [("x", "<f8"), ("y", "<f8"), ("z", "<f8"), ("u1", "<f8"), ("v1", "<f8"), ("u2", "<f8"),
               ("v2", "<f8"), ("residual", "<f8"), ("curve", "<i4"), ("flags", "<i4")]


def curve_points_to_numpy(buf, n):
    import numpy as np
    return np.frombuffer(bytes(memoryview(buf).cast("B")[: n * C.sizeof(CurvePoint)]),
                         dtype=np.dtype(CURVE_DTYPE)).copy()


QUERY_DISTANCE = 0
QUERY_INTERSECT = 1


class Query(C.Structure):
    """jsi_query: one query of a batch or a server."""
    _fields_ = [("kind", C.c_int), ("h1", C.c_int), ("h2", C.c_int), ("r1", Region), ("r2", Region),
                ("params", Params)]
