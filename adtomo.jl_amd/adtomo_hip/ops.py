"""Python mirror of ADTomo.jl's Eikonal op API over the HIP C ABI.

Same names, argument meaning and layouts as src/eikonal_op.jl:

  eikonal(f, srcx, srcy, h)                     src/eikonal_op.jl:3-21
  eikonal3d(u0, f, h, m, n, l, tol, verbose)    src/eikonal_op.jl:24-32

and the gradient ops the reference registers (EikonalGrad,
Eikonal.cpp:48-63; EikonalThreeDGrad, EikonalThreeD.cpp:47-65).

Inputs may be torch tensors on a HIP device (device pointers are handed to
the library, outputs are torch tensors, and autograd flows through the
adjoint kernels like the reference's ADCME gradient) or numpy arrays (the
library stages them through device memory; outputs are numpy arrays).
There is no CPU fallback: every call runs the HIP kernels.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import _lib as L

last_info: dict = {}


def _torch():
    import torch
    return torch


def _is_torch(x) -> bool:
    try:
        import torch
    except ImportError:  # pragma: no cover
        return False
    return isinstance(x, torch.Tensor)


def _stream():
    torch = _torch()
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _checked(info: dict) -> dict:
    """Raises if a kernel reported an internal fault for this event (a
    bounded wait between its waves timed out; its results are invalid)."""
    if info.get("fault"):
        raise L.AdtError("kernel fault reported for an event (wave synchronisation timed out)")
    return info


def _sfx(dtype) -> str:
    name = str(dtype)
    return "_f32" if name.endswith("float32") else "_f64"


def _np_dtype(x):
    return np.float32 if np.asarray(x).dtype == np.float32 else np.float64


def _set_info(info: L.AdtInfo):
    global last_info
    last_info = _checked(info.as_dict())
    return last_info


# ------------------------------------------------------------------- 2D ---

def _eik2d_fwd_raw(f, srcx: int, srcy: int, h: float):
    n1, m1 = f.shape
    m, n = m1 - 1, n1 - 1
    info = L.AdtInfo()
    if _is_torch(f):
        f = f.contiguous()
        u = _torch().empty_like(f)
        rc = getattr(L.lib(), "adt_eikonal2d_fwd" + _sfx(f.dtype))(
            f.data_ptr(), m, n, srcx, srcy, h, u.data_ptr(), ctypes.byref(info), 0, _stream())
    else:
        f = np.ascontiguousarray(f, dtype=_np_dtype(f))
        u = np.empty_like(f)
        rc = getattr(L.lib(), "adt_eikonal2d_fwd" + _sfx(f.dtype))(
            f.ctypes.data, m, n, srcx, srcy, h, u.ctypes.data, ctypes.byref(info),
            L.ADT_HOST_POINTERS, None)
    L.check(rc, "eikonal")
    _set_info(info)
    return u


def eikonal_grad(grad_u, u, f, srcx: int, srcy: int, h: float):
    """EikonalGrad (Eikonal.cpp:141-221): returns grad_f, shape of f."""
    n1, m1 = f.shape
    m, n = m1 - 1, n1 - 1
    info = L.AdtInfo()
    if _is_torch(f):
        g, u, f = grad_u.contiguous(), u.contiguous(), f.contiguous()
        gf = _torch().empty_like(f)
        rc = getattr(L.lib(), "adt_eikonal2d_bwd" + _sfx(f.dtype))(
            g.data_ptr(), u.data_ptr(), f.data_ptr(), m, n, srcx, srcy, h, gf.data_ptr(),
            ctypes.byref(info), 0, _stream())
    else:
        dt = _np_dtype(f)
        g, u, f = (np.ascontiguousarray(a, dtype=dt) for a in (grad_u, u, f))
        gf = np.empty_like(f)
        rc = getattr(L.lib(), "adt_eikonal2d_bwd" + _sfx(dt))(
            g.ctypes.data, u.ctypes.data, f.ctypes.data, m, n, srcx, srcy, h, gf.ctypes.data,
            ctypes.byref(info), L.ADT_HOST_POINTERS, None)
    L.check(rc, "eikonal_grad")
    _set_info(info)
    return gf


def _make_eik2d_fn():
    torch = _torch()

    class Eikonal2dFn(torch.autograd.Function):
        @staticmethod
        def forward(ctx, f, srcx, srcy, h):
            u = _eik2d_fwd_raw(f.detach(), srcx, srcy, h)
            ctx.save_for_backward(u, f)
            ctx.args = (srcx, srcy, h)
            return u

        @staticmethod
        def backward(ctx, grad_u):
            u, f = ctx.saved_tensors
            srcx, srcy, h = ctx.args
            gf = eikonal_grad(grad_u.contiguous(), u, f.detach(), srcx, srcy, h)
            return gf, None, None, None

    return Eikonal2dFn


_Eik2dFn = None


def eikonal(f, srcx: int, srcy: int, h: float):
    """2D first-arrival travel times (src/eikonal_op.jl:3-21).

    ``f`` is the slowness on an (n+1) x (m+1) grid (rows = depth), ``srcx``
    and ``srcy`` the 1-based source column and row, ``h`` the spacing.
    Returns u with the shape of ``f``."""
    global _Eik2dFn
    if _is_torch(f):
        if _Eik2dFn is None:
            _Eik2dFn = _make_eik2d_fn()
        return _Eik2dFn.apply(f, int(srcx), int(srcy), float(h))
    return _eik2d_fwd_raw(np.asarray(f), int(srcx), int(srcy), float(h))


# ------------------------------------------------------------------- 3D ---

def _eik3d_fwd_raw(u0, f, h, m, n, l, tol, verbose):
    info = L.AdtInfo()
    if _is_torch(f):
        u0 = u0.reshape(m, n, l).contiguous()
        f = f.reshape(m, n, l).contiguous()
        u = _torch().empty_like(f)
        rc = getattr(L.lib(), "adt_eikonal3d_fwd" + _sfx(f.dtype))(
            u0.data_ptr(), f.data_ptr(), h, m, n, l, tol, int(verbose), u.data_ptr(),
            ctypes.byref(info), 0, _stream())
    else:
        dt = _np_dtype(f)
        u0 = np.ascontiguousarray(np.reshape(u0, (m, n, l)), dtype=dt)
        f = np.ascontiguousarray(np.reshape(f, (m, n, l)), dtype=dt)
        u = np.empty_like(f)
        rc = getattr(L.lib(), "adt_eikonal3d_fwd" + _sfx(dt))(
            u0.ctypes.data, f.ctypes.data, h, m, n, l, tol, int(verbose), u.ctypes.data,
            ctypes.byref(info), L.ADT_HOST_POINTERS, None)
    L.check(rc, "eikonal3d")
    _set_info(info)
    return u


def eikonal3d_grad(grad_u, u, u0, f, h, m, n, l, tol=None, verbose=False):
    """EikonalThreeDGrad (EikonalThreeD.cpp:154-250): returns (grad_u0, grad_f)."""
    info = L.AdtInfo()
    if _is_torch(f):
        g, u, u0, f = (a.reshape(m, n, l).contiguous() for a in (grad_u, u, u0, f))
        gu0 = _torch().empty_like(f)
        gf = _torch().empty_like(f)
        rc = getattr(L.lib(), "adt_eikonal3d_bwd" + _sfx(f.dtype))(
            g.data_ptr(), u.data_ptr(), u0.data_ptr(), f.data_ptr(), h, m, n, l,
            gu0.data_ptr(), gf.data_ptr(), ctypes.byref(info), 0, _stream())
    else:
        dt = _np_dtype(f)
        g, u, u0, f = (np.ascontiguousarray(np.reshape(a, (m, n, l)), dtype=dt)
                       for a in (grad_u, u, u0, f))
        gu0 = np.empty_like(f)
        gf = np.empty_like(f)
        rc = getattr(L.lib(), "adt_eikonal3d_bwd" + _sfx(dt))(
            g.ctypes.data, u.ctypes.data, u0.ctypes.data, f.ctypes.data, h, m, n, l,
            gu0.ctypes.data, gf.ctypes.data, ctypes.byref(info), L.ADT_HOST_POINTERS, None)
    L.check(rc, "eikonal3d_grad")
    _set_info(info)
    return gu0, gf


def _make_eik3d_fn():
    torch = _torch()

    class Eikonal3dFn(torch.autograd.Function):
        @staticmethod
        def forward(ctx, u0, f, h, m, n, l, tol, verbose):
            u = _eik3d_fwd_raw(u0.detach(), f.detach(), h, m, n, l, tol, verbose)
            ctx.save_for_backward(u, u0, f)
            ctx.args = (h, m, n, l, tol, verbose)
            return u

        @staticmethod
        def backward(ctx, grad_u):
            u, u0, f = ctx.saved_tensors
            h, m, n, l, tol, verbose = ctx.args
            gu0, gf = eikonal3d_grad(grad_u.contiguous(), u, u0.detach(), f.detach(), h, m, n, l)
            # The reference returns gradients only for u0 and f
            # (EikonalThreeD.cpp:154-250); h, dims, tol, verbose get none.
            return (gu0.reshape(u0.shape), gf.reshape(f.shape), None, None, None, None, None,
                    None)

    return Eikonal3dFn


_Eik3dFn = None


def eikonal3d(u0, f, h, m, n, l, tol, verbose=False):
    """3D travel times by fast sweeping (src/eikonal_op.jl:24-32).

    ``u0`` (initial times, e.g. 1000 with the source times set) and ``f``
    (slowness) hold m*n*l values, row-major (m, n, l).  Returns (m, n, l)."""
    global _Eik3dFn
    m, n, l = int(m), int(n), int(l)
    if _is_torch(f):
        if _Eik3dFn is None:
            _Eik3dFn = _make_eik3d_fn()
        return _Eik3dFn.apply(u0, f, float(h), m, n, l, float(tol), bool(verbose))
    return _eik3d_fwd_raw(np.asarray(u0), np.asarray(f), float(h), m, n, l, float(tol), verbose)


# ------------------------------------------------------ strided layouts ---
#
# The _strided entry points take arrays of any common layout: a Fortran-
# ordered numpy array, a permuted or sliced torch view.  All arrays of a
# call share one shape and one set of strides; outputs get the same strides.

def _layout(arrays, ndim):
    """(strides in elements, device flag) shared by ``arrays``; raises
    unless they agree and every stride of an axis longer than one is
    positive."""
    ref = arrays[0]
    if ref.ndim != ndim:
        raise ValueError(f"strided calls take {ndim}-D arrays")
    dev = _is_torch(ref)
    st = None
    for a in arrays:
        if _is_torch(a) != dev or tuple(a.shape) != tuple(ref.shape) or a.dtype != ref.dtype:
            raise ValueError("arrays of a strided call must share kind, shape and dtype")
        if dev:
            if not a.is_cuda:
                raise ValueError("torch tensors must be on the HIP device")
            s = tuple(int(x) for x in a.stride())
        else:
            if any(x % a.itemsize for x in a.strides):
                raise ValueError("numpy strides must be whole elements")
            s = tuple(int(x) // a.itemsize for x in a.strides)
        if st is None:
            st = s
        elif s != st:
            raise ValueError("arrays of a strided call must share their strides")
    if any(x <= 0 for x, d in zip(st, ref.shape) if d > 1):
        raise ValueError("strides must be positive")
    return st, dev


def _empty_like_strided(a, st):
    """Uninitialised array with a's shape and element strides ``st``."""
    if _is_torch(a):
        return _torch().empty_strided(tuple(a.shape), st, dtype=a.dtype, device=a.device)
    span = 1 + sum((d - 1) * x for d, x in zip(a.shape, st) if d > 1)
    base = np.empty(span, dtype=a.dtype)
    return np.lib.stride_tricks.as_strided(base, a.shape, tuple(x * a.itemsize for x in st))


def _ptr(a):
    return a.data_ptr() if _is_torch(a) else a.ctypes.data


def _call_strided(name, dev, dtype, args, info):
    fn = getattr(L.lib(), name + _sfx(dtype))
    rc = fn(*args, ctypes.byref(info), 0 if dev else L.ADT_HOST_POINTERS,
            _stream() if dev else None)
    L.check(rc, name)
    _set_info(info)


def eikonal3d_strided(u0, f, h, tol, verbose=False):
    """EikonalThreeD (EikonalThreeD.cpp:14-45) on (m, n, l) arrays of any
    common layout (adt_eikonal3d_fwd_strided_*); returns u with the same
    strides."""
    st, dev = _layout([u0, f], 3)
    m, n, l = (int(x) for x in f.shape)
    u = _empty_like_strided(f, st)
    info = L.AdtInfo()
    _call_strided("adt_eikonal3d_fwd_strided", dev, f.dtype,
                  [_ptr(u0), _ptr(f), float(h), m, n, l, float(tol), int(verbose),
                   (ctypes.c_int64 * 3)(*st), _ptr(u)], info)
    return u


def eikonal3d_grad_strided(grad_u, u, u0, f, h):
    """EikonalThreeDGrad (EikonalThreeD.cpp:154-250) on (m, n, l) arrays of
    any common layout; returns (grad_u0, grad_f) with the same strides."""
    st, dev = _layout([grad_u, u, u0, f], 3)
    m, n, l = (int(x) for x in f.shape)
    gu0, gf = _empty_like_strided(f, st), _empty_like_strided(f, st)
    info = L.AdtInfo()
    _call_strided("adt_eikonal3d_bwd_strided", dev, f.dtype,
                  [_ptr(grad_u), _ptr(u), _ptr(u0), _ptr(f), float(h), m, n, l,
                   (ctypes.c_int64 * 3)(*st), _ptr(gu0), _ptr(gf)], info)
    return gu0, gf


def eikonal_strided(f, srcx, srcy, h):
    """Eikonal (Eikonal.cpp:20-46) on an (n+1, m+1) array (rows = depth j,
    columns = width i, as src/eikonal_op.jl:3-21 takes it) of any layout;
    returns u with the same strides."""
    st, dev = _layout([f], 2)
    n, m = int(f.shape[0]) - 1, int(f.shape[1]) - 1
    u = _empty_like_strided(f, st)
    info = L.AdtInfo()
    _call_strided("adt_eikonal2d_fwd_strided", dev, f.dtype,
                  [_ptr(f), m, n, int(srcx), int(srcy), float(h),
                   (ctypes.c_int64 * 2)(st[1], st[0]), _ptr(u)], info)
    return u


def eikonal_grad_strided(grad_u, u, f, srcx, srcy, h):
    """EikonalGrad (Eikonal.cpp:48-63) on (n+1, m+1) arrays of any common
    layout; returns grad_f with the same strides."""
    st, dev = _layout([grad_u, u, f], 2)
    n, m = int(f.shape[0]) - 1, int(f.shape[1]) - 1
    gf = _empty_like_strided(f, st)
    info = L.AdtInfo()
    _call_strided("adt_eikonal2d_bwd_strided", dev, f.dtype,
                  [_ptr(grad_u), _ptr(u), _ptr(f), m, n, int(srcx), int(srcy), float(h),
                   (ctypes.c_int64 * 2)(st[1], st[0]), _ptr(gf)], info)
    return gf


# -------------------------------------------------------------- batched ---

def _event_table(u0s, fs, hs, tols, grad_us, us, grad_u0s, grad_fs):
    nev = len(fs)
    tab = (L.AdtEvent3d * max(nev, 1))()
    dtype = None
    for e in range(nev):
        f = fs[e]
        if not _is_torch(f) or not f.is_cuda or not f.is_contiguous():
            raise ValueError("batched calls take contiguous torch tensors on the HIP device")
        if dtype is None:
            dtype = f.dtype
        elif f.dtype != dtype:
            raise ValueError("all events must share one dtype")
        m, n, l = f.shape
        ev = tab[e]
        ev.m, ev.n, ev.l = m, n, l
        ev.h = float(hs[e])
        ev.tol = float(tols[e])
        ev.f = f.data_ptr()
        ev.u0 = u0s[e].data_ptr() if u0s is not None else None
        ev.u = us[e].data_ptr()
        ev.grad_u = grad_us[e].data_ptr() if grad_us is not None else None
        ev.grad_u0 = grad_u0s[e].data_ptr() if grad_u0s is not None else None
        ev.grad_f = grad_fs[e].data_ptr() if grad_fs is not None else None
    return tab, dtype


def _flags(forward, grad_us, compute_f64):
    return ((L.ADT_FORWARD if forward else 0) | (L.ADT_ADJOINT if grad_us is not None else 0)
            | (L.ADT_F32_COMPUTE_F64 if compute_f64 else 0))


def eikonal3d_batched(u0s, fs, hs, tols, grad_us=None, forward=True, us=None,
                      return_info=True, compute_f64=False):
    """Many independent 3D events (boxes of any size) in one persistent launch.

    Lists of contiguous (m, n, l) device tensors.  With ``grad_us`` the
    adjoint runs too.  ``compute_f64`` (float32 tensors): FP32 storage with
    every forward local solve in FP64.  Returns dict(u, grad_u0, grad_f, info)."""
    torch = _torch()
    nev = len(fs)
    if us is None:
        us = [torch.empty_like(f) for f in fs]
    flags = _flags(forward, grad_us, compute_f64)
    gu0s = [torch.empty_like(f) for f in fs] if grad_us is not None else None
    gfs = [torch.empty_like(f) for f in fs] if grad_us is not None else None
    tab, dtype = _event_table(u0s, fs, hs, tols, grad_us, us, gu0s, gfs)
    infos = (L.AdtInfo * max(nev, 1))() if return_info else None
    rc = getattr(L.lib(), "adt_eikonal3d_batched" + _sfx(dtype))(
        ctypes.byref(tab), nev, flags, ctypes.byref(infos) if infos is not None else None,
        _stream())
    L.check(rc, "eikonal3d_batched")
    return {"u": us, "grad_u0": gu0s, "grad_f": gfs,
            "info": [_checked(infos[e].as_dict()) for e in range(nev)] if infos is not None else None}


def _dtype_code(dtype):
    return 1 if str(dtype).endswith("float32") else 0


def get_tune(dtype="float64"):
    """Current launch setting of the 3D kernels for one precision."""
    t = L.AdtTune()
    L.check(L.lib().adt_get_tune(_dtype_code(dtype), ctypes.byref(t)), "get_tune")
    return t.as_dict()


_TUNE_FIELDS = ("diag_max_l", "adj_wg", "order", "fuse_max_n", "adj_algo", "fwd_segs")


def _apply(t, fields):
    for k, v in fields.items():
        if k not in _TUNE_FIELDS:
            raise TypeError(f"unknown launch setting field {k!r}")
        if v is not None:
            setattr(t, k, int(v))


def set_tune(dtype="float64", **fields):
    """Change the global launch setting (unspecified fields keep their
    value): ``diag_max_l``, ``adj_wg``, ``order``, ``fuse_max_n``,
    ``adj_algo`` (0 = adjoint over 8x8x8 tiles, the default; 1 = node by
    node) and ``fwd_segs`` (c segments of the diagonal forward, 1..8; 0 =
    8)."""
    t = L.AdtTune()
    L.check(L.lib().adt_get_tune(_dtype_code(dtype), ctypes.byref(t)), "get_tune")
    _apply(t, fields)
    L.check(L.lib().adt_set_tune(_dtype_code(dtype), ctypes.byref(t)), "set_tune")
    return t.as_dict()


def set_tune_class(shape, dtype="float64", clear=False, **fields):
    """Give the box-shape class of ``shape`` = (m, n, l) its own per-event
    fields (``diag_max_l``, ``fwd_segs``, ``adj_algo``; the others stay
    global), starting from the global setting; ``clear=True`` removes the
    class's setting."""
    m, n, l = (int(x) for x in shape)
    if clear:
        L.check(L.lib().adt_tune_class_set(_dtype_code(dtype), m, n, l, None), "tune_class_set")
        return None
    t = L.AdtTune()
    L.check(L.lib().adt_get_tune(_dtype_code(dtype), ctypes.byref(t)), "get_tune")
    _apply(t, fields)
    L.check(L.lib().adt_tune_class_set(_dtype_code(dtype), m, n, l, ctypes.byref(t)),
            "tune_class_set")
    return t.as_dict()


def tune_classes(dtype="float64"):
    """The class settings of one precision, each a dict with the class key
    (``l_blocks``, ``bands``, ``log2n``) and the setting's fields."""
    lib = L.lib()
    code = _dtype_code(dtype)
    n = lib.adt_tune_class_count(code)
    L.check(min(n, 0), "tune_class_count")
    out = []
    for i in range(n):
        c = L.AdtTuneClass()
        L.check(lib.adt_tune_class_get(code, i, ctypes.byref(c)), "tune_class_get")
        out.append(c.as_dict())
    return out


def clear_tune_classes(dtype="float64"):
    L.check(L.lib().adt_tune_class_clear(_dtype_code(dtype)), "tune_class_clear")


def save_tune(path):
    """Write every launch setting (global and per class, both precisions)
    to a JSON file."""
    L.check(L.lib().adt_tune_save(os.fsencode(path)), "tune_save")


def load_tune(path):
    """Read settings written by save_tune (replacing the global settings it
    holds and every class setting)."""
    L.check(L.lib().adt_tune_load(os.fsencode(path)), "tune_load")


def autotune(u0s, fs, hs, tols, grad_us=None, reps=3, compute_f64=False, cache=None):
    """Time the launch settings on these events and keep the fastest for
    their precision: per box-shape class among them, the forward kernel and
    its segment count and the adjoint traversal; over all of them, the event
    order, the node-level adjoint's workgroup size and the fusion threshold.
    ``cache``: a JSON file of settings; read instead of timing when it
    exists, written after timing when it does not.  Returns the global
    setting with its measured times."""
    dtype = "float32" if fs[0].dtype == _torch().float32 else "float64"
    if cache is not None and os.path.exists(cache):
        load_tune(cache)
        return get_tune(dtype)
    torch = _torch()
    us = [torch.empty_like(f) for f in fs]
    gu0s = [torch.empty_like(f) for f in fs] if grad_us is not None else None
    gfs = [torch.empty_like(f) for f in fs] if grad_us is not None else None
    tab, dtype = _event_table(u0s, fs, hs, tols, grad_us, us, gu0s, gfs)
    flags = _flags(True, grad_us, compute_f64)
    best = L.AdtTune()
    L.check(L.lib().adt_autotune(ctypes.byref(tab), len(fs), flags, _dtype_code(dtype), int(reps),
                                 ctypes.byref(best), _stream()), "autotune")
    if cache is not None:
        save_tune(cache)
    return best.as_dict()


class EikonalGraph:
    """hipGraph of the batched forward and/or adjoint over fixed buffers.

    Build once, then ``launch()`` repeatedly; the event buffers (inputs and
    outputs, all kept alive by this object) are read/written in place."""

    def __init__(self, u0s, fs, hs, tols, grad_us=None, forward=True, us=None, compute_f64=False):
        torch = _torch()
        self.u0s, self.fs, self.grad_us = u0s, fs, grad_us
        # travel times: written by the forward, read by the adjoint (pass the
        # forward's buffers to build an adjoint-only graph)
        self.us = list(us) if us is not None else [torch.empty_like(f) for f in fs]
        self.grad_u0s = [torch.empty_like(f) for f in fs] if grad_us is not None else None
        self.grad_fs = [torch.empty_like(f) for f in fs] if grad_us is not None else None
        self.nev = len(fs)
        flags = _flags(forward, grad_us, compute_f64)
        tab, dtype = _event_table(u0s, fs, hs, tols, grad_us, self.us, self.grad_u0s, self.grad_fs)
        self._g = ctypes.c_void_p()
        rc = L.lib().adt_graph_create(ctypes.byref(tab), self.nev, flags,
                                      1 if str(dtype).endswith("float32") else 0,
                                      ctypes.byref(self._g), _stream())
        L.check(rc, "adt_graph_create")

    def launch(self):
        L.check(L.lib().adt_graph_launch(self._g, _stream()), "adt_graph_launch")

    def infos(self):
        arr = (L.AdtInfo * max(self.nev, 1))()
        L.check(L.lib().adt_graph_infos(self._g, ctypes.byref(arr), _stream()), "adt_graph_infos")
        return [_checked(arr[e].as_dict()) for e in range(self.nev)]

    def close(self):
        if self._g:
            L.lib().adt_graph_destroy(self._g)
            self._g = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------- plain baseline ---

def eikonal3d_plain(u0, f, h, tol):
    """Per-event launch loop (one launch per hyperplane, host-synchronised
    stop test per iteration); FP64 torch device tensors."""
    torch = _torch()
    u = torch.empty_like(f)
    m, n, l = f.shape
    info = L.AdtInfo()
    rc = L.lib().adt_eikonal3d_fwd_plain_f64(u0.data_ptr(), f.data_ptr(), float(h), m, n, l,
                                             float(tol), u.data_ptr(), ctypes.byref(info),
                                             _stream())
    L.check(rc, "eikonal3d_plain")
    _set_info(info)
    return u


def eikonal3d_grad_plain(grad_u, u, u0, f, h):
    torch = _torch()
    m, n, l = f.shape
    gu0 = torch.empty_like(f)
    gf = torch.empty_like(f)
    info = L.AdtInfo()
    rc = L.lib().adt_eikonal3d_bwd_plain_f64(grad_u.data_ptr(), u.data_ptr(), u0.data_ptr(),
                                             f.data_ptr(), float(h), m, n, l, gu0.data_ptr(),
                                             gf.data_ptr(), ctypes.byref(info), _stream())
    L.check(rc, "eikonal3d_grad_plain")
    _set_info(info)
    return gu0, gf


def selftest_arith(op: int, a, b=None):
    """Runs one device arithmetic op of the path on torch device tensors
    (op 6: plain copy, the reference kernel for memory counters; ops 7, 8:
    the square root and the 3D local solve as the forward's step loops
    take them)."""
    torch = _torch()
    n = a.numel() // (3 if op in (4, 8) else 2 if op == 5 else 1)
    out = torch.empty(n, dtype=a.dtype, device=a.device)
    fn = getattr(L.lib(), "adt_selftest_arith" + _sfx(a.dtype))
    rc = fn(op, a.data_ptr(), b.data_ptr() if b is not None else None, out.data_ptr(), n,
            _stream())
    L.check(rc, "selftest_arith")
    return out
