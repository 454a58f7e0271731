"""adtomo_hip: MI355X (gfx950) HIP implementation of ADTomo.jl's Eikonal ops.

Public API (mirrors src/eikonal_op.jl and the registered gradient ops):
    eikonal, eikonal_grad, eikonal3d, eikonal3d_grad,
    eikonal3d_batched, EikonalGraph, eikonal3d_plain, eikonal3d_grad_plain;
    eikonal_strided, eikonal_grad_strided, eikonal3d_strided,
    eikonal3d_grad_strided take arrays of any common layout.
The objective around the solves (scripts/inversion.jl) is in
``adtomo_hip.objective``; the inversion driver in ``adtomo_hip.inversion``.
"""
from . import objective
from .ops import (EikonalGraph, autotune, eikonal, eikonal3d, eikonal3d_batched, eikonal3d_grad,
                  eikonal3d_grad_plain, eikonal3d_grad_strided, eikonal3d_plain,
                  eikonal3d_strided, eikonal_grad, eikonal_grad_strided, eikonal_strided,
                  clear_tune_classes, get_tune, last_info, load_tune, save_tune, selftest_arith,
                  set_tune, set_tune_class, tune_classes)
from ._lib import LIB_PATH, AdtError, lib

__all__ = ["eikonal", "eikonal_grad", "eikonal3d", "eikonal3d_grad", "eikonal3d_batched",
           "EikonalGraph", "eikonal3d_plain", "eikonal3d_grad_plain", "selftest_arith",
           "last_info", "lib", "LIB_PATH", "AdtError", "objective", "autotune", "get_tune",
           "set_tune", "set_tune_class", "tune_classes", "clear_tune_classes", "save_tune",
           "load_tune", "eikonal_strided", "eikonal_grad_strided", "eikonal3d_strided",
           "eikonal3d_grad_strided"]
