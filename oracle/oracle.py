"""ctypes front end of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline
leg import this module, and only as the checker / reported CPU baseline.  The
product package (``adtomo.jl_amd/adtomo_hip``) never imports it.

Layouts follow the reference's Julia API (src/eikonal_op.jl):
  * 2D: ``f`` has shape (n+1, m+1) (rows = depth j, cols = width i), flattened
    row-major so flat = j*(m+1) + i; ``srcx``/``srcy`` are 1-based column/row
    (deps/CustomOps/Eikonal/Eikonal.cpp:125 subtracts 1).
  * 3D: ``u0``/``f`` have shape (m, n, l), flat = i*n*l + j*l + k
    (deps/CustomOps/Eikonal3D/Eikonal3D.cpp:6-8).

Besides the C restatement, this module holds the brute-force assembly of the
reference's sparse adjoint system (``assemble_2d``/``assemble_3d``), written
from Eikonal.h:105-189 and Eikonal3D.cpp:118-187, solved with SciPy's SuperLU
(the closest available stand-in for Eigen::SparseLU).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libadt_oracle.so")
_lib = None

_d = ctypes.c_double
_i64 = ctypes.c_int64
_int = ctypes.c_int
_p = ctypes.c_void_p


def build() -> str:
    """Compile the oracle with its Makefile (gcc, -ffp-contract=off)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        for sfx, real in (("_f64", _d), ("_f32", ctypes.c_float)):
            fn = getattr(L, "adt_oracle_eikonal2d_fwd" + sfx)
            fn.argtypes = [_p, _i64, _i64, real, _i64, _i64, _int, _p, _p,
                           ctypes.POINTER(_int), ctypes.POINTER(_d)]
            fn.restype = _int
            fn = getattr(L, "adt_oracle_eikonal2d_bwd" + sfx)
            fn.argtypes = [_p, _p, _p, _i64, _i64, real, _i64, _i64, _p, _p,
                           ctypes.POINTER(_i64)]
            fn.restype = _int
            fn = getattr(L, "adt_oracle_eikonal3d_fwd" + sfx)
            fn.argtypes = [_p, _p, real, _i64, _i64, _i64, real, _int, _int, _p, _p,
                           ctypes.POINTER(_int), ctypes.POINTER(_d)]
            fn.restype = _int
            fn = getattr(L, "adt_oracle_eikonal3d_bwd" + sfx)
            fn.argtypes = [_p, _p, _p, _p, real, _i64, _i64, _i64, _p, _p, _p]
            fn.restype = _int
        L.adt_oracle_eikonal3d_fwd_m32.argtypes = [_p, _p, _d, _i64, _i64, _i64, ctypes.c_float,
                                                   _int, _int, _p, _p, ctypes.POINTER(_int),
                                                   ctypes.POINTER(_d)]
        L.adt_oracle_eikonal3d_fwd_m32.restype = _int
        L.adt_oracle_eikonal3d_batch_f64.argtypes = [
            _int, _p, _p, _p, _p, _p, _d, _p, _p, _p, _p, _int]
        L.adt_oracle_eikonal3d_batch_f64.restype = _int
        _lib = L
    return _lib


def _sfx(dtype):
    return "_f32" if np.dtype(dtype) == np.float32 else "_f64"


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(_p)


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


# ------------------------------------------------------------------- 2D ---

def eikonal2d(f, srcx: int, srcy: int, h: float, dtype=np.float64, max_iter: int = 100):
    """Forward of the 2D op (Eikonal.h:54-93).  Returns (u, iters, err, converged)."""
    f = _c(f, dtype)
    n1, m1 = f.shape
    m, n = m1 - 1, n1 - 1
    u = np.empty_like(f)
    uold = np.empty_like(f)
    it = _int(0)
    err = _d(0.0)
    rc = getattr(lib(), "adt_oracle_eikonal2d_fwd" + _sfx(dtype))(
        _ptr(f), m, n, h, srcx - 1, srcy - 1, max_iter, _ptr(u), _ptr(uold),
        ctypes.byref(it), ctypes.byref(err))
    return u, it.value, err.value, rc == 0


def eikonal2d_grad(grad_u, u, f, srcx: int, srcy: int, h: float, dtype=np.float64):
    """Adjoint of the 2D op (Eikonal.h:95-199) as an exact back-substitution.
    Returns (grad_f, lambda, n_singular)."""
    f = _c(f, dtype)
    u = _c(u, dtype)
    g = _c(grad_u, dtype)
    n1, m1 = f.shape
    gf = np.empty_like(f)
    lam = np.empty_like(f)
    ns = _i64(0)
    getattr(lib(), "adt_oracle_eikonal2d_bwd" + _sfx(dtype))(
        _ptr(g), _ptr(u), _ptr(f), m1 - 1, n1 - 1, h, srcx - 1, srcy - 1,
        _ptr(gf), _ptr(lam), ctypes.byref(ns))
    return gf, lam, ns.value


# ------------------------------------------------------------------- 3D ---

def eikonal3d(u0, f, h: float, tol: float, dtype=np.float64, max_iter: int = 20,
              verbose: bool = False):
    """Forward of the 3D op (Eikonal3D.cpp:71-94).  Returns (u, iters, err)."""
    u0 = _c(u0, dtype)
    f = _c(f, dtype)
    m, n, l = u0.shape
    u = np.empty_like(u0)
    uold = np.empty_like(u0)
    it = _int(0)
    err = _d(0.0)
    getattr(lib(), "adt_oracle_eikonal3d_fwd" + _sfx(dtype))(
        _ptr(u0), _ptr(f), h, m, n, l, tol, max_iter, int(verbose), _ptr(u), _ptr(uold),
        ctypes.byref(it), ctypes.byref(err))
    return u, it.value, err.value


def eikonal3d_m32(u0, f, h: float, tol: float, max_iter: int = 20):
    """FP32-storage forward with FP64 local solves (the library's
    ADT_F32_COMPUTE_F64 variant).  Returns (u float32, iters, err)."""
    u0 = _c(u0, np.float32)
    f = _c(f, np.float32)
    m, n, l = u0.shape
    u = np.empty_like(u0)
    uold = np.empty_like(u0)
    it = _int(0)
    err = _d(0.0)
    lib().adt_oracle_eikonal3d_fwd_m32(_ptr(u0), _ptr(f), h, m, n, l, tol, max_iter, 0, _ptr(u),
                                       _ptr(uold), ctypes.byref(it), ctypes.byref(err))
    return u, it.value, err.value


def eikonal3d_grad(grad_u, u, u0, f, h: float, dtype=np.float64):
    """Adjoint of the 3D op (Eikonal3D.cpp:96-198) as an exact
    back-substitution.  Returns (grad_u0, grad_f, lambda)."""
    g = _c(grad_u, dtype)
    u = _c(u, dtype)
    u0 = _c(u0, dtype)
    f = _c(f, dtype)
    m, n, l = u.shape
    gu0 = np.empty_like(u)
    gf = np.empty_like(u)
    lam = np.empty_like(u)
    getattr(lib(), "adt_oracle_eikonal3d_bwd" + _sfx(dtype))(
        _ptr(g), _ptr(u), _ptr(u0), _ptr(f), h, m, n, l, _ptr(gu0), _ptr(gf), _ptr(lam))
    return gu0, gf, lam


def eikonal3d_batch(u0s, fs, gs, hs, tol: float, nthreads: int):
    """Serial-per-event forward (+ adjoint when ``gs`` is given) of many
    events spread one per thread over ``nthreads`` host threads."""
    nev = len(u0s)
    u0s = [_c(a, np.float64) for a in u0s]
    fs = [_c(a, np.float64) for a in fs]
    us = [np.empty_like(a) for a in u0s]
    dims = np.array([a.shape for a in u0s], dtype=np.int64).reshape(-1)
    hv = np.asarray(hs, dtype=np.float64)
    iters = np.zeros(nev, dtype=np.int32)

    def parr(lst):
        return (ctypes.c_void_p * nev)(*[a.ctypes.data for a in lst])

    if gs is not None:
        gs = [_c(a, np.float64) for a in gs]
        gu0 = [np.empty_like(a) for a in u0s]
        gf = [np.empty_like(a) for a in u0s]
        gp, gu0p, gfp = parr(gs), parr(gu0), parr(gf)
    else:
        gu0 = gf = None
        gp = gu0p = gfp = None
    lib().adt_oracle_eikonal3d_batch_f64(
        nev, parr(u0s), parr(fs), gp, _ptr(dims), _ptr(hv), tol, parr(us), gu0p, gfp,
        _ptr(iters), nthreads)
    return us, gu0, gf, iters


# ------------------------------------------- brute-force sparse systems ---

def _smin(x, y):
    """std::min(x, y) == (y < x) ? y : x, elementwise."""
    return np.where(y < x, y, x)


def assemble_2d(u, f, srcx: int, srcy: int, h: float):
    """The reference's 2D Jacobian A (before its transpose), assembled exactly
    as Eikonal.h:105-189 does: identity row at the source, and for every
    other node a diagonal 2(u - a) plus off-diagonal 2(a - u) at the argmin
    neighbour of each active dimension.  Returns (A as scipy CSC, dFdf)."""
    import scipy.sparse as sp
    u = np.asarray(u, dtype=np.float64)
    n1, m1 = u.shape
    m, n = m1 - 1, n1 - 1
    ix, jx = srcx - 1, srcy - 1
    W = m + 1
    rows, cols, vals = [], [], []
    jj, ii = np.meshgrid(np.arange(n + 1), np.arange(m + 1), indexing="ij")
    idx = jj * W + ii
    uf = u.reshape(-1)
    notsrc = ~((ii == ix) & (jj == jx))
    # x dimension
    left = np.where(ii > 0, idx - 1, idx + 1)
    right = np.where(ii < m, idx + 1, idx - 1)
    ux_r, ux_l = uf[right], uf[left]
    choose_left = ux_r > ux_l
    nbx = np.where(ii == 0, idx + 1, np.where(ii == m, idx - 1, np.where(choose_left, idx - 1, idx + 1)))
    a = uf[nbx]
    act = notsrc & (uf[idx] > a)
    rows += [idx[act], idx[act]]
    cols += [idx[act], nbx[act]]
    vals += [2 * (uf[idx][act] - a[act]), 2 * (a[act] - uf[idx][act])]
    # y dimension
    upn = np.where(jj > 0, idx - W, idx + W)
    dn = np.where(jj < n, idx + W, idx - W)
    choose_up = uf[dn] > uf[upn]
    nby = np.where(jj == 0, idx + W, np.where(jj == n, idx - W, np.where(choose_up, idx - W, idx + W)))
    b = uf[nby]
    act = notsrc & (uf[idx] > b)
    rows += [idx[act], idx[act]]
    cols += [idx[act], nby[act]]
    vals += [2 * (uf[idx][act] - b[act]), 2 * (b[act] - uf[idx][act])]
    src = jx * W + ix
    rows.append(np.array([src]))
    cols.append(np.array([src]))
    vals.append(np.array([1.0]))
    N = W * (n + 1)
    A = sp.coo_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                      shape=(N, N)).tocsc()
    dFdf = (-2 * np.asarray(f, dtype=np.float64).reshape(-1) * h * h)
    dFdf[src] = 0.0
    return A, dFdf


def assemble_3d(u, u0, f, h: float):
    """The reference's 3D Jacobian A (before its transpose), assembled as
    Eikonal3D.cpp:118-187 does, including the Z fix-up (rows/columns of Z
    zeroed, unit diagonal, g zeroed on Z).  Returns (A as CSC, rhs, Zmask)."""
    import scipy.sparse as sp
    u = np.asarray(u, dtype=np.float64)
    u0 = np.asarray(u0, dtype=np.float64)
    m, n, l = u.shape
    N = m * n * l
    uf = u.reshape(-1)
    I, J, K = np.meshgrid(np.arange(m), np.arange(n), np.arange(l), indexing="ij")
    r = (I * n * l + J * l + K).reshape(-1)
    Zu = (uf == u0.reshape(-1))
    rows, cols, vals = [], [], []
    anyact = np.zeros(N, dtype=bool)
    for coord, ext, stride in ((I.reshape(-1), m, n * l), (J.reshape(-1), n, l), (K.reshape(-1), l, 1)):
        plus = np.where(coord < ext - 1, r + stride, r - stride)
        minus = np.where(coord > 0, r - stride, r + stride)
        umin = np.where(coord == 0, uf[plus], np.where(coord == ext - 1, uf[minus], _smin(uf[plus], uf[minus])))
        nb = np.where(coord == 0, r + stride,
                      np.where(coord == ext - 1, r - stride,
                               np.where(uf[plus] > uf[minus], r - stride, r + stride)))
        act = (~Zu) & (uf > umin)
        anyact |= act
        rows += [r[act], r[act]]
        cols += [r[act], nb[act]]
        vals += [2.0 * (uf[act] - umin[act]), -2.0 * (uf[act] - umin[act])]
    Z = Zu | ~anyact
    rows = np.concatenate(rows)
    cols = np.concatenate(cols)
    vals = np.concatenate(vals)
    keep = ~(Z[rows] | Z[cols])
    zi = np.nonzero(Z)[0]
    rows = np.concatenate([rows[keep], zi])
    cols = np.concatenate([cols[keep], zi])
    vals = np.concatenate([vals[keep], np.ones(len(zi))])
    A = sp.coo_matrix((vals, (rows, cols)), shape=(N, N)).tocsc()
    rhs = -2 * np.asarray(f, dtype=np.float64).reshape(-1) * h * h
    return A, rhs, Z


def lu_grad_2d(grad_u, u, f, srcx, srcy, h):
    """grad_f of the 2D op by the reference's route: A^T lambda = g with a
    sparse LU (SciPy SuperLU), grad_f = -lambda * dFdf."""
    import scipy.sparse.linalg as spla
    A, dFdf = assemble_2d(u, f, srcx, srcy, h)
    lam = spla.splu(A.T.tocsc()).solve(np.asarray(grad_u, dtype=np.float64).reshape(-1))
    return (-lam * dFdf).reshape(np.shape(u)), lam.reshape(np.shape(u))


def lu_grad_3d(grad_u, u, u0, f, h):
    """(grad_u0, grad_f) of the 3D op by the reference's route with SuperLU."""
    import scipy.sparse.linalg as spla
    A, rhs, Z = assemble_3d(u, u0, f, h)
    g = np.array(grad_u, dtype=np.float64).reshape(-1)
    gu0 = np.where(np.asarray(u).reshape(-1) == np.asarray(u0).reshape(-1), g, 0.0)
    g[Z] = 0.0
    lam = spla.splu(A.T.tocsc()).solve(g)
    shp = np.shape(u)
    return gu0.reshape(shp), (-lam * rhs).reshape(shp), lam.reshape(shp)
