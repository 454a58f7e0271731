"""ctypes binding of libadtomo_hip.so (the C ABI in include/adtomo_hip.h).

The product path has no CPU fallback: if the HIP library is missing this
module raises at load time.
"""
from __future__ import annotations

import ctypes
import os

_PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(_PKG_ROOT, "libadtomo_hip.so")
HEADER_PATH = os.path.join(os.path.dirname(_PKG_ROOT), "include", "adtomo_hip.h")

ADT_OK = 0
ADT_HOST_POINTERS = 0x1
ADT_FORWARD = 0x2
ADT_ADJOINT = 0x4
ADT_F32_COMPUTE_F64 = 0x8
MAX_LOGGED_ITERS = 100

_i32 = ctypes.c_int32
_i64 = ctypes.c_int64
_d = ctypes.c_double
_f = ctypes.c_float
_p = ctypes.c_void_p


class AdtInfo(ctypes.Structure):
    _fields_ = [
        ("iterations", _i32),
        ("converged", _i32),
        ("err", _d),
        ("ambiguous", _i32),
        ("adj_sweeps", _i32),
        ("adj_converged", _i32),
        ("fault", _i32),
        ("n_singular", _i64),
        ("err_log", _d * MAX_LOGGED_ITERS),
    ]

    def as_dict(self):
        k = max(0, min(self.iterations, MAX_LOGGED_ITERS))
        return {
            "iterations": self.iterations,
            "converged": bool(self.converged),
            "err": self.err,
            "ambiguous": bool(self.ambiguous),
            "adj_sweeps": self.adj_sweeps,
            "adj_converged": bool(self.adj_converged),
            "fault": self.fault,
            "n_singular": self.n_singular,
            "err_log": list(self.err_log[:k]),
        }


class AdtEvent3d(ctypes.Structure):
    _fields_ = [
        ("m", _i64), ("n", _i64), ("l", _i64),
        ("h", _d), ("tol", _d),
        ("u0", _p), ("f", _p), ("grad_u", _p),
        ("u", _p), ("grad_u0", _p), ("grad_f", _p),
    ]


class AdtSource3d(ctypes.Structure):
    _fields_ = [("x", _d), ("y", _d), ("z", _d), ("scale", _d), ("u0", _p)]


class AdtMisfit3d(ctypes.Structure):
    _fields_ = [
        ("m", _i64), ("n", _i64), ("l", _i64),
        ("u", _p), ("tobs", _p), ("weight", _p),
        ("tcal", _p), ("grad_u", _p), ("loss", _p),
        ("huber_delta", _d),
    ]


class AdtTune(ctypes.Structure):
    _fields_ = [("diag_max_l", _i32), ("adj_wg", _i32), ("order", _i32), ("fuse_max_n", _i32),
                ("fwd_ms", ctypes.c_float), ("adj_ms", ctypes.c_float), ("adj_algo", _i32),
                ("fwd_segs", _i32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class AdtTuneClass(ctypes.Structure):
    _fields_ = [("l_blocks", _i32), ("bands", _i32), ("log2n", _i32), ("tune", AdtTune)]

    def as_dict(self):
        d = {"l_blocks": self.l_blocks, "bands": self.bands, "log2n": self.log2n}
        d.update(self.tune.as_dict())
        return d


class AdtError(RuntimeError):
    pass


_lib = None

# name -> (restype, argtypes); REAL is filled per dtype suffix.
_SIGS = {
    "adt_eikonal2d_fwd": [_p, _i64, _i64, _i64, _i64, "R", _p, _p, _i32, _p],
    "adt_eikonal2d_bwd": [_p, _p, _p, _i64, _i64, _i64, _i64, "R", _p, _p, _i32, _p],
    "adt_eikonal3d_fwd": [_p, _p, "R", _i64, _i64, _i64, "R", _i32, _p, _p, _i32, _p],
    "adt_eikonal3d_bwd": [_p, _p, _p, _p, "R", _i64, _i64, _i64, _p, _p, _p, _i32, _p],
    "adt_eikonal2d_fwd_strided": [_p, _i64, _i64, _i64, _i64, "R", _p, _p, _p, _i32, _p],
    "adt_eikonal2d_bwd_strided": [_p, _p, _p, _i64, _i64, _i64, _i64, "R", _p, _p, _p, _i32, _p],
    "adt_eikonal3d_fwd_strided": [_p, _p, "R", _i64, _i64, _i64, "R", _i32, _p, _p, _p, _i32, _p],
    "adt_eikonal3d_bwd_strided": [_p, _p, _p, _p, "R", _i64, _i64, _i64, _p, _p, _p, _p, _i32,
                                  _p],
    "adt_eikonal3d_batched": [_p, _i32, _i32, _p, _p],
    "adt_eikonal3d_misfit_batched": [_p, _p, _i32, _p, _p, _p, _i64, _i32, _p, _p],
    "adt_selftest_arith": [_i32, _p, _p, _p, _i64, _p],
    "adt_source_u0_3d_batched": [_p, _i32, _p, _i64, _i64, _i64, "R", _p],
    "adt_misfit3d_batched": [_p, _i32, _p, _p, _p, _i64, _p],
    "adt_model3d": [_p, _p, _i64, _p, _p, _p],
    "adt_model3d_grad": [_p, _p, _p, _p, _i64, _p, _p],
    "adt_smooth3d_l1": [_p, _i64, _i64, _i64, _i32, _i32, "R", _p, _p, _p],
    "adt_sum_arrays": [_p, _i32, _i64, _p, _p],
    "adt_slowness3d": [_p, _i64, "R", _p, _p],
    "adt_slowness3d_grad": [_p, _i64, "R", _p, _p, _p, _p],
}


def lib():
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise AdtError(f"libadtomo_hip.so not found at {LIB_PATH}; build it with "
                       f"`make -C adtomo.jl_amd` (or __graft_entry__.build())")
    L = ctypes.CDLL(LIB_PATH)
    for base, args in _SIGS.items():
        for sfx, real in (("_f64", _d), ("_f32", _f)):
            fn = getattr(L, base + sfx)
            fn.argtypes = [real if a == "R" else a for a in args]
            fn.restype = _i32
    L.adt_eikonal3d_fwd_noalloc_f64.argtypes = [_p, _p, _d, _i64, _i64, _i64, _d, _i32, _p, _p,
                                                _p]
    L.adt_eikonal3d_fwd_noalloc_f64.restype = _i32
    L.adt_eikonal3d_fwd_plain_f64.argtypes = [_p, _p, _d, _i64, _i64, _i64, _d, _p, _p, _p]
    L.adt_eikonal3d_fwd_plain_f64.restype = _i32
    L.adt_eikonal3d_bwd_plain_f64.argtypes = [_p, _p, _p, _p, _d, _i64, _i64, _i64, _p, _p, _p, _p]
    L.adt_eikonal3d_bwd_plain_f64.restype = _i32
    L.adt_graph_create.argtypes = [_p, _i32, _i32, _i32, ctypes.POINTER(_p), _p]
    L.adt_graph_create.restype = _i32
    L.adt_graph_create_misfit.argtypes = [_p, _p, _i32, _p, _p, _p, _i64, _i32, _i32,
                                          ctypes.POINTER(_p), _p]
    L.adt_graph_create_misfit.restype = _i32
    L.adt_graph_launch.argtypes = [_p, _p]
    L.adt_graph_launch.restype = _i32
    L.adt_graph_infos.argtypes = [_p, _p, _p]
    L.adt_graph_infos.restype = _i32
    L.adt_graph_destroy.argtypes = [_p]
    L.adt_graph_destroy.restype = _i32
    L.adt_get_tune.argtypes = [_i32, _p]
    L.adt_get_tune.restype = _i32
    L.adt_set_tune.argtypes = [_i32, _p]
    L.adt_set_tune.restype = _i32
    L.adt_autotune.argtypes = [_p, _i32, _i32, _i32, _i32, _p, _p]
    L.adt_autotune.restype = _i32
    L.adt_tune_class_set.argtypes = [_i32, _i64, _i64, _i64, _p]
    L.adt_tune_class_set.restype = _i32
    L.adt_tune_class_count.argtypes = [_i32]
    L.adt_tune_class_count.restype = _i32
    L.adt_tune_class_get.argtypes = [_i32, _i32, _p]
    L.adt_tune_class_get.restype = _i32
    L.adt_tune_class_clear.argtypes = [_i32]
    L.adt_tune_class_clear.restype = _i32
    L.adt_tune_save.argtypes = [ctypes.c_char_p]
    L.adt_tune_save.restype = _i32
    L.adt_tune_load.argtypes = [ctypes.c_char_p]
    L.adt_tune_load.restype = _i32
    L.adt_version.restype = ctypes.c_char_p
    L.adt_last_error.restype = ctypes.c_char_p
    _lib = L
    return L


def check(rc: int, what: str):
    if rc != ADT_OK:
        msg = lib().adt_last_error().decode(errors="replace")
        raise AdtError(f"{what} failed (rc={rc}): {msg}")


def declared_symbols(header: str = HEADER_PATH):
    """Function names declared in include/adtomo_hip.h."""
    import re
    txt = open(header).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(adt_[a-z0-9_]+)\s*\(", txt)))
