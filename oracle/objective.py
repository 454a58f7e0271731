"""NumPy restatement of the tomography objective around the Eikonal path
(TEST INFRASTRUCTURE ONLY: imported by ``tests/`` as the checker; the
product package never imports it).

Follows scripts/inversion.jl (and scripts/inversion_joint.jl for the S-wave
scale):
  * ``source_u0``   :48-60   8-corner initial travel times of one station
  * ``misfit``      :64-105  trilinear gather at the events, quality-weighted
                             L2 misfit skipping the -1 sentinel, and its
                             gradient with respect to the travel times
  * ``model``       :42-43   fvar = 2 sigmoid(x) - 1 + vel0, f = 1 / fvar
  * ``smooth_l1``   :107-121 lambda * sum |fvar - box(fvar)|, box = mean over
                             an sh x sh x sv window with periodic padding
  * ``model_grad``           chain rule of the two maps above
  * ``slowness``    inversion_joint.jl:79 S slowness vp/vs / fvar, and its
                             adjoint ``slowness_grad``
  * ``huber``                Huber residual (delta > 0) for the robust misfit

Arrays are (m, n, l) row-major (flat = i*n*l + j*l + k, the layout of
src/eikonal_op.jl:24-32); coordinates are the reference's 1-based reals.
Expressions keep the reference's operation order where the device result
is compared bit for bit (source_u0, the interpolated times).
"""
from __future__ import annotations

import math

import numpy as np


def source_u0(shape, x, y, z, vel0, h, scale=1.0):
    """scripts/inversion.jl:48-60 (u0 = 1000 * ones(m,n,l), then the 8
    corners ixu/ixd x iyu/iyd x izu/izd in this order)."""
    u0 = np.full(shape, 1000.0)
    xs = (math.ceil(x), math.floor(x))
    ys = (math.ceil(y), math.floor(y))
    zs = (math.ceil(z), math.floor(z))
    for ix in xs:
        for iy in ys:
            for iz in zs:
                dx, dy, dz = x - ix, y - iy, z - iz
                dist = math.sqrt(dx * dx + dy * dy + dz * dz)
                u0[ix - 1, iy - 1, iz - 1] = dist * h / vel0[ix - 1, iy - 1, iz - 1] * scale
    return u0


def _corners(jx, jy, jz):
    x1, x2 = math.floor(jx), math.ceil(jx)
    y1, y2 = math.floor(jy), math.ceil(jy)
    z1, z2 = math.floor(jz), math.ceil(jz)
    return (x1, x2), (y1, y2), (z1, z2)


def interpolate(u, jx, jy, jz):
    """scripts/inversion.jl:67-92: travel time at a 1-based point."""
    (x1, x2), (y1, y2), (z1, z2) = _corners(jx, jy, jz)

    def at(i, j, k):
        return u[i - 1, j - 1, k - 1]

    if x1 == x2:
        tx11, tx12 = at(x1, y1, z1), at(x1, y1, z2)
        tx21, tx22 = at(x1, y2, z1), at(x1, y2, z2)
    else:
        tx11 = (x2 - jx) * at(x1, y1, z1) + (jx - x1) * at(x2, y1, z1)
        tx12 = (x2 - jx) * at(x1, y1, z2) + (jx - x1) * at(x2, y1, z2)
        tx21 = (x2 - jx) * at(x1, y2, z1) + (jx - x1) * at(x2, y2, z1)
        tx22 = (x2 - jx) * at(x1, y2, z2) + (jx - x1) * at(x2, y2, z2)
    if y1 == y2:
        txy1, txy2 = tx11, tx12
    else:
        txy1 = (y2 - jy) * tx11 + (jy - y1) * tx21
        txy2 = (y2 - jy) * tx12 + (jy - y1) * tx22
    if z1 == z2:
        return txy1
    return (z2 - jz) * txy1 + (jz - z1) * txy2


def interpolation_weights(shape, jx, jy, jz):
    """d t / d u of ``interpolate`` as a dict {(i, j, k) 0-based: weight}."""
    (x1, x2), (y1, y2), (z1, z2) = _corners(jx, jy, jz)
    wx = [(x1, 1.0)] if x1 == x2 else [(x1, x2 - jx), (x2, jx - x1)]
    wy = [(y1, 1.0)] if y1 == y2 else [(y1, y2 - jy), (y2, jy - y1)]
    wz = [(z1, 1.0)] if z1 == z2 else [(z1, z2 - jz), (z2, jz - z1)]
    out = {}
    for k, c in wz:
        for j, b in wy:
            for i, a in wx:
                key = (i - 1, j - 1, k - 1)
                out[key] = out.get(key, 0.0) + (c * b) * a
    return out


def interpolation_adjoint(g, jx, jy, jz):
    """g * d t / d u of ``interpolate`` as {(i, j, k) 0-based: value}, in the
    order the chain rule of the interpolation takes it (backpropagation of
    scripts/inversion.jl:67-92): the z weight first, then y, then x,
    ((g * wz) * wy) * wx; a dimension that is not interpolated contributes
    the exact factor 1."""
    (x1, x2), (y1, y2), (z1, z2) = _corners(jx, jy, jz)
    wx = [(x1, 1.0)] if x1 == x2 else [(x1, x2 - jx), (x2, jx - x1)]
    wy = [(y1, 1.0)] if y1 == y2 else [(y1, y2 - jy), (y2, jy - y1)]
    wz = [(z1, 1.0)] if z1 == z2 else [(z1, z2 - jz), (z2, jz - z1)]
    return {(i - 1, j - 1, k - 1): ((g * c) * b) * a
            for k, c in wz for j, b in wy for i, a in wx}


def huber(r, delta):
    """(rho, d rho / d r): r^2 inside [-delta, delta], 2 delta |r| - delta^2
    outside; delta <= 0 gives the reference's r^2."""
    if delta > 0 and abs(r) > delta:
        return 2.0 * delta * abs(r) - delta * delta, (2.0 * delta if r > 0 else -(2.0 * delta))
    return r * r, 2.0 * r


def misfit(u, rx, ry, rz, tobs, w, delta=0.0):
    """scripts/inversion.jl:64-105 for one station: (tcal, loss, grad_u);
    delta > 0 selects the Huber residual."""
    tcal = np.array([interpolate(u, a, b, c) for a, b, c in zip(rx, ry, rz)])
    loss = 0.0
    grad = np.zeros_like(u)
    for j in range(len(tcal)):
        if tobs[j] == -1:
            continue
        r = tobs[j] - tcal[j]
        rho, drho = huber(r, delta)
        loss += w[j] * rho
        g = -(w[j] * drho)
        # receivers in order, each node's terms added into the zeroed grad
        for key, val in interpolation_adjoint(g, rx[j], ry[j], rz[j]).items():
            grad[key] += val
    return tcal, loss, grad


def sigmoid(x):
    return 1.0 / (1.0 + np.exp(-x))


def model(x, vel0):
    """scripts/inversion.jl:42-43,61: fvar = 2 sigmoid(x) - 1 + vel0, f = 1/fvar."""
    fvar = 2.0 * sigmoid(x) - 1.0 + vel0
    return fvar, 1.0 / fvar


def slowness(fvar, scale):
    """scripts/inversion_joint.jl:79: S slowness pvs ./ fvar."""
    return scale / fvar


def slowness_grad(fvar, scale, grad_fs):
    """(d/dfvar, d/dscale) of sum(grad_fs * scale / fvar)."""
    return -scale * grad_fs / (fvar * fvar), float((grad_fs / fvar).sum())


def box_mean(a, sh, sv):
    """Mean over an sh x sh x sv window with periodic wrap (the padded VALID
    conv3d with a uniform filter of scripts/inversion.jl:107-118)."""
    rh, rv = sh // 2, sv // 2
    s = np.zeros_like(a)
    wgt = 1.0 / (sh * sh * sv)
    for di in range(-rh, rh + 1):
        for dj in range(-rh, rh + 1):
            for dk in range(-rv, rv + 1):
                s += wgt * np.roll(a, (-di, -dj, -dk), axis=(0, 1, 2))
    return s


def smooth_l1(fvar, sh, sv, lam):
    """scripts/inversion.jl:107-121: (value, d value / d fvar)."""
    d = fvar - box_mean(fvar, sh, sv)
    value = lam * np.abs(d).sum()
    sgn = np.sign(d)
    grad = lam * (sgn - box_mean(sgn, sh, sv))
    return value, grad


def model_grad(x, fvar, grad_f, grad_fvar):
    """Chain rule: d/dx of L(f(fvar(x)), fvar(x))."""
    s = sigmoid(x)
    g = grad_fvar - grad_f / (fvar * fvar)
    return g * (2.0 * s * (1.0 - s))
