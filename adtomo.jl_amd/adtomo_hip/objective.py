"""The tomography objective around the Eikonal solves, on the device.

Mirrors scripts/inversion.jl (the caller of src/eikonal_op.jl):

  source_u0     :48-60   8-corner initial travel times per station
  misfit        :64-105  trilinear gather of the travel times at the events,
                         quality-weighted L2 misfit (tobs == -1 skipped) and
                         its adjoint source dL/du for the Eikonal adjoint
  model         :42-43   fvar = 2 sigmoid(x) - 1 + vel0, f = 1 / fvar
  model_grad             chain rule back to x
  slowness               f_s = vp/vs / fvar (scripts/inversion_joint.jl:79)
  smooth_l1     :107-121 lambda * sum |fvar - box(fvar)| (periodic box mean)
  sum_arrays             sum of per-station gradients (fixed order)

Arguments are torch tensors on a HIP device (float64 or float32, contiguous,
(m, n, l) row-major); coordinates are the reference's 1-based reals.  Every
call runs HIP kernels; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib as L
from .ops import _checked, _event_table, _flags, _sfx, _stream, _torch


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _check_dev(*ts):
    for t in ts:
        if t is not None and (not t.is_cuda or not t.is_contiguous()):
            raise ValueError("expected contiguous device tensors")


def source_u0(stations, vel0, h: float, scale: float = 1.0, out=None):
    """Initial travel times for each station (x, y, z) (1-based reals).

    Returns a list of (m, n, l) tensors (or fills ``out``)."""
    torch = _torch()
    _check_dev(vel0)
    m, n, l = vel0.shape
    stations = [tuple(map(float, s)) for s in stations]
    if out is None:
        out = [torch.empty_like(vel0) for _ in stations]
    tab = (L.AdtSource3d * max(len(stations), 1))()
    for e, (x, y, z) in enumerate(stations):
        tab[e] = L.AdtSource3d(x, y, z, float(scale), out[e].data_ptr())
    fn = getattr(L.lib(), "adt_source_u0_3d_batched" + _sfx(vel0.dtype))
    L.check(fn(ctypes.byref(tab), len(stations), _ptr(vel0), m, n, l, h, _stream()), "source_u0")
    return out


def misfit(us, rx, ry, rz, tobs, weights, grad=True, tcal=True, grad_out=None, huber_delta=0.0):
    """Receiver gather + misfit for each station.

    ``us``: list of travel-time tensors; ``rx, ry, rz``: event coordinates
    (host sequences, 1-based); ``tobs``/``weights``: per-station tensors of
    length len(rx).  ``grad_out``: optional list of tensors to receive
    dL/du.  ``huber_delta`` > 0 replaces the squared residual by the Huber
    form (r^2 inside +-delta, 2 delta |r| - delta^2 outside).  Returns
    dict(loss=tensor[nstation], tcal=[...], grad_u=[...])."""
    torch = _torch()
    ns = len(us)
    dt = us[0].dtype
    ax, ay, az = _check_picks(dt, ns, rx, ry, rz, tobs, weights)
    nrec = len(ax)
    _check_dev(*us, *tobs, *weights)
    loss = torch.zeros(ns, dtype=dt, device=us[0].device)
    tcals = [torch.empty(nrec, dtype=dt, device=us[0].device) for _ in range(ns)] if tcal else None
    if grad:
        grads = grad_out if grad_out is not None else [torch.empty_like(u) for u in us]
        _check_dev(*grads)
    else:
        grads = None
    tab = (L.AdtMisfit3d * max(ns, 1))()
    esz = loss.element_size()
    for e, u in enumerate(us):
        m, n, l = u.shape
        tab[e] = L.AdtMisfit3d(m, n, l, u.data_ptr(), tobs[e].data_ptr(), weights[e].data_ptr(),
                               tcals[e].data_ptr() if tcal else None,
                               grads[e].data_ptr() if grad else None, loss.data_ptr() + e * esz,
                               float(huber_delta))
    fn = getattr(L.lib(), "adt_misfit3d_batched" + _sfx(dt))
    rc = fn(ctypes.byref(tab), ns, ax.ctypes.data_as(ctypes.c_void_p),
            ay.ctypes.data_as(ctypes.c_void_p), az.ctypes.data_as(ctypes.c_void_p), nrec, _stream())
    L.check(rc, "misfit")
    return {"loss": loss, "tcal": tcals, "grad_u": grads}


def _step_items(us, tobs, weights, grad_us, losses, huber_delta):
    items = (L.AdtMisfit3d * max(len(us), 1))()
    esz = losses.element_size()
    for e, u in enumerate(us):
        m, n, l = u.shape
        items[e] = L.AdtMisfit3d(m, n, l, u.data_ptr(), tobs[e].data_ptr(), weights[e].data_ptr(),
                                 None, grad_us[e].data_ptr(), losses.data_ptr() + e * esz,
                                 float(huber_delta))
    return items


def _coords(rx, ry, rz):
    """Host coordinate arrays of the receivers (equal lengths)."""
    ax, ay, az = (np.ascontiguousarray(a, dtype=np.float64).reshape(-1) for a in (rx, ry, rz))
    if not (len(ax) == len(ay) == len(az)):
        raise ValueError(f"rx, ry and rz must have equal lengths (got {len(ax)}, {len(ay)}, "
                         f"{len(az)})")
    return ax, ay, az


def _check_picks(dtype, nsolve, rx, ry, rz, tobs, weights):
    """Validates the receivers and picks of ``nsolve`` solves before any
    device work (the library cannot check device array lengths): equal
    coordinate lengths, one tobs and one weights tensor per solve, each of
    the travel times' dtype with one value per receiver.  Returns the host
    coordinate arrays."""
    ax, ay, az = _coords(rx, ry, rz)
    nrec = len(ax)
    if len(tobs) != nsolve or len(weights) != nsolve:
        raise ValueError(f"expected one tobs and one weights tensor per solve ({nsolve}), got "
                         f"{len(tobs)} and {len(weights)}")
    for name, ts in (("tobs", tobs), ("weights", weights)):
        for e, t in enumerate(ts):
            if t.dtype != dtype:
                raise ValueError(f"{name}[{e}] has dtype {t.dtype}, the travel times {dtype}")
            if t.numel() != nrec:
                raise ValueError(f"{name}[{e}] has {t.numel()} values for {nrec} receivers")
    return ax, ay, az


def inversion_step(u0s, fs, hs, tols, rx, ry, rz, tobs, weights, huber_delta=0.0,
                   compute_f64=False):
    """One inversion step in one call (adt_eikonal3d_misfit_batched_*): for
    each solve its forward, the receiver misfit of its travel times (its
    gradient becomes the adjoint source) and the adjoint.  Arguments as for
    ``InversionStepGraph``.  Returns dict(u, grad_u, grad_u0, grad_f, loss,
    info)."""
    torch = _torch()
    n = len(fs)
    ax, ay, az = _check_picks(fs[0].dtype, n, rx, ry, rz, tobs, weights)
    us = [torch.empty_like(f) for f in fs]
    gus = [torch.empty_like(f) for f in fs]
    gu0s = [torch.empty_like(f) for f in fs]
    gfs = [torch.empty_like(f) for f in fs]
    losses = torch.zeros(max(n, 1), dtype=fs[0].dtype, device=fs[0].device)
    _check_dev(*tobs, *weights)
    tab, dtype = _event_table(u0s, fs, hs, tols, gus, us, gu0s, gfs)
    items = _step_items(us, tobs, weights, gus, losses, huber_delta)
    infos = (L.AdtInfo * max(n, 1))()
    rc = getattr(L.lib(), "adt_eikonal3d_misfit_batched" + _sfx(dtype))(
        ctypes.byref(tab), ctypes.byref(items), n, ax.ctypes.data_as(ctypes.c_void_p),
        ay.ctypes.data_as(ctypes.c_void_p), az.ctypes.data_as(ctypes.c_void_p), len(ax),
        _flags(True, gus, compute_f64), ctypes.byref(infos), _stream())
    L.check(rc, "adt_eikonal3d_misfit_batched")
    return {"u": us, "grad_u": gus, "grad_u0": gu0s, "grad_f": gfs, "loss": losses[:n],
            "info": [_checked(infos[e].as_dict()) for e in range(n)]}


class InversionStepGraph:
    """hipGraph of one inversion step over fixed buffers
    (adt_graph_create_misfit): for each solve its forward, the receiver
    misfit of its travel times (scripts/inversion.jl:64-105; its gradient
    becomes the adjoint source) and the adjoint, in one launch.

    ``u0s``/``fs``/``hs``/``tols`` as for ``EikonalGraph``; ``rx, ry, rz``
    host event coordinates (1-based); ``tobs``/``weights`` per-solve
    tensors.  After ``launch()``: ``losses`` (one per solve), ``us``,
    ``grad_us`` (dL/du), ``grad_fs`` and ``grad_u0s``.  Every buffer is kept
    alive by this object."""

    def __init__(self, u0s, fs, hs, tols, rx, ry, rz, tobs, weights, huber_delta=0.0,
                 compute_f64=False):
        torch = _torch()
        self.u0s, self.fs, self.tobs, self.weights = list(u0s), list(fs), list(tobs), list(weights)
        self.nev = len(fs)
        ax, ay, az = _check_picks(fs[0].dtype, self.nev, rx, ry, rz, self.tobs, self.weights)
        self.us = [torch.empty_like(f) for f in fs]
        self.grad_us = [torch.empty_like(f) for f in fs]
        self.grad_u0s = [torch.empty_like(f) for f in fs]
        self.grad_fs = [torch.empty_like(f) for f in fs]
        dt = fs[0].dtype
        self.losses = torch.zeros(max(self.nev, 1), dtype=dt, device=fs[0].device)
        _check_dev(*self.tobs, *self.weights)
        tab, dtype = _event_table(self.u0s, self.fs, hs, tols, self.grad_us, self.us,
                                  self.grad_u0s, self.grad_fs)
        items = _step_items(self.us, self.tobs, self.weights, self.grad_us, self.losses,
                            huber_delta)
        flags = _flags(True, self.grad_us, compute_f64)
        self._g = ctypes.c_void_p()
        rc = L.lib().adt_graph_create_misfit(
            ctypes.byref(tab), ctypes.byref(items), self.nev, ax.ctypes.data_as(ctypes.c_void_p),
            ay.ctypes.data_as(ctypes.c_void_p), az.ctypes.data_as(ctypes.c_void_p), len(ax), flags,
            1 if str(dtype).endswith("float32") else 0, ctypes.byref(self._g), _stream())
        L.check(rc, "adt_graph_create_misfit")

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


def model(x, vel0, slowness=True, out=None):
    """(fvar, f): fvar = 2 sigmoid(x) - 1 + vel0, f = 1/fvar.  ``out``: an
    optional (fvar, f) pair of tensors to fill."""
    torch = _torch()
    _check_dev(x, vel0)
    if out is not None:
        fvar, f = out
        _check_dev(fvar, f)
    else:
        fvar = torch.empty_like(x)
        f = torch.empty_like(x) if slowness else None
    fn = getattr(L.lib(), "adt_model3d" + _sfx(x.dtype))
    L.check(fn(_ptr(x), _ptr(vel0), x.numel(), _ptr(fvar), _ptr(f), _stream()), "model3d")
    return fvar, f


def model_grad(x, fvar, grad_f=None, grad_fvar=None):
    """d/dx through fvar(x) and f = 1/fvar."""
    torch = _torch()
    _check_dev(x, fvar, grad_f, grad_fvar)
    gx = torch.empty_like(x)
    fn = getattr(L.lib(), "adt_model3d_grad" + _sfx(x.dtype))
    L.check(fn(_ptr(x), _ptr(fvar), _ptr(grad_f), _ptr(grad_fvar), x.numel(), _ptr(gx), _stream()),
            "model3d_grad")
    return gx


def slowness(fvar, scale: float, out=None):
    """f_s = scale / fvar (the S slowness pvs ./ fvar)."""
    torch = _torch()
    _check_dev(fvar, out)
    out = torch.empty_like(fvar) if out is None else out
    fn = getattr(L.lib(), "adt_slowness3d" + _sfx(fvar.dtype))
    L.check(fn(_ptr(fvar), fvar.numel(), float(scale), _ptr(out), _stream()), "slowness3d")
    return out


def slowness_grad(fvar, scale: float, grad_fs, grad_fvar=None, grad_scale=True):
    """Adjoint of ``slowness``: adds -scale grad_fs / fvar^2 to ``grad_fvar``
    in place (when given) and returns d/dscale as a device scalar (or None)."""
    torch = _torch()
    _check_dev(fvar, grad_fs, grad_fvar)
    gs = torch.zeros(1, dtype=fvar.dtype, device=fvar.device) if grad_scale else None
    fn = getattr(L.lib(), "adt_slowness3d_grad" + _sfx(fvar.dtype))
    L.check(fn(_ptr(fvar), fvar.numel(), float(scale), _ptr(grad_fs), _ptr(grad_fvar), _ptr(gs),
               _stream()), "slowness3d_grad")
    return gs


def smooth_l1(fvar, sh: int, sv: int, lam: float, grad=True):
    """(value, grad_fvar) of lambda * sum |fvar - box(fvar)|."""
    torch = _torch()
    _check_dev(fvar)
    m, n, l = fvar.shape
    value = torch.zeros(1, dtype=fvar.dtype, device=fvar.device)
    g = torch.empty_like(fvar) if grad else None
    fn = getattr(L.lib(), "adt_smooth3d_l1" + _sfx(fvar.dtype))
    L.check(fn(_ptr(fvar), m, n, l, sh, sv, lam, _ptr(g), _ptr(value), _stream()), "smooth3d_l1")
    return value, g


def sum_arrays(arrays, out=None):
    """Sum of equally shaped device tensors, in list order."""
    torch = _torch()
    _check_dev(*arrays)
    if out is None:
        out = torch.empty_like(arrays[0])
    ptrs = (ctypes.c_void_p * len(arrays))(*[a.data_ptr() for a in arrays])
    fn = getattr(L.lib(), "adt_sum_arrays" + _sfx(arrays[0].dtype))
    L.check(fn(ptrs, len(arrays), arrays[0].numel(), _ptr(out), _stream()), "sum_arrays")
    return out
