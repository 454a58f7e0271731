"""Independent numpy formulation of the reference forward, in hyperplane order.

The reference sweeps lexicographically (deps/CustomOps/Eikonal3D/Eikonal3D.cpp:35-57,
deps/CustomOps/Eikonal/Eikonal.h:22-52).  A node's update reads its upwind
neighbours after they were updated in the same sweep and its downwind
neighbours before; processing the hyperplanes i'+j'+k' = s (primed = index
counted along the sweep direction) in increasing s keeps exactly those
reads, so the result must equal the lexicographic sweep bit for bit.  This
module implements that order with vectorised numpy (IEEE binary64, no FMA)
and is used by the tests to pin the C oracle; it is also the schedule the HIP
kernels follow.
"""
import numpy as np


def smin(x, y):
    return np.where(y < x, y, x)


def solution3d(a1, a2, a3, f, h):
    # Eikonal3D.cpp:11-28, branch values evaluated then selected.
    a1, a2 = np.where(a1 > a2, a2, a1), np.where(a1 > a2, a1, a2)
    a1, a3 = np.where(a1 > a3, a3, a1), np.where(a1 > a3, a1, a3)
    a2, a3 = np.where(a2 > a3, a3, a2), np.where(a2 > a3, a2, a3)
    with np.errstate(invalid="ignore"):
        x1 = a1 + f * h
        B = -(a1 + a2)
        C = (a1 * a1 + a2 * a2 - f * f * h * h) / 2.0
        x2 = (-B + np.sqrt(B * B - 4 * C)) / 2.0
        B3 = -2.0 * (a1 + a2 + a3) / 3.0
        C3 = (a1 * a1 + a2 * a2 + a3 * a3 - f * f * h * h) / 3.0
        x3 = (-B3 + np.sqrt(B3 * B3 - 4 * C3)) / 2.0
    return np.where(x1 <= a2, x1, np.where(x2 <= a3, x2, x3))


DIRS3 = [(1, 1, 1), (-1, 1, 1), (-1, -1, 1), (1, -1, 1),
         (1, -1, -1), (1, 1, -1), (-1, 1, -1), (-1, -1, -1)]


def _planes(shape):
    m, n, l = shape
    a, b, c = np.meshgrid(np.arange(m), np.arange(n), np.arange(l), indexing="ij")
    s = (a + b + c).reshape(-1)
    order = np.argsort(s, kind="stable")
    bounds = np.searchsorted(s[order], np.arange(m + n + l - 1))
    bounds = np.append(bounds, len(s))
    return a.reshape(-1)[order], b.reshape(-1)[order], c.reshape(-1)[order], bounds


def sweep3d(u, f, h, d, mixed=False):
    """One sweep in hyperplane order.  mixed: u and f are binary32 arrays,
    each local solve runs in binary64 and is rounded once to binary32."""
    m, n, l = u.shape
    A, Bb, Cc, bounds = _planes(u.shape)
    for s in range(m + n + l - 2):
        lo, hi = bounds[s], bounds[s + 1]
        a, b, c = A[lo:hi], Bb[lo:hi], Cc[lo:hi]
        i = a if d[0] > 0 else m - 1 - a
        j = b if d[1] > 0 else n - 1 - b
        k = c if d[2] > 0 else l - 1 - c

        def nbmin(coord, ext, get):
            plus = get(np.minimum(coord + 1, ext - 1))
            minus = get(np.maximum(coord - 1, 0))
            return np.where(coord == 0, plus, np.where(coord == ext - 1, minus, smin(plus, minus)))

        ux = nbmin(i, m, lambda t: u[t, j, k])
        uy = nbmin(j, n, lambda t: u[i, t, k])
        uz = nbmin(k, l, lambda t: u[i, j, t])
        if mixed:
            un = solution3d(ux.astype(np.float64), uy.astype(np.float64), uz.astype(np.float64),
                            f[i, j, k].astype(np.float64), h).astype(np.float32)
        else:
            un = solution3d(ux, uy, uz, f[i, j, k], h)
        cur = u[i, j, k]
        u[i, j, k] = np.where(cur < un, cur, un)


def forward3d(u0, f, h, tol, max_iter=20, mixed=False):
    """mixed: binary32 storage and stop rule, binary64 local solves."""
    dt = np.float32 if mixed else np.float64
    u = np.array(u0, dtype=dt, copy=True)
    f = np.asarray(f, dtype=dt)
    tol = dt(tol)
    it = 0
    err = dt(0.0)
    for it in range(1, max_iter + 1):
        uold = u.copy()
        for d in DIRS3:
            sweep3d(u, f, h, d, mixed)
        diff = np.abs(u - uold)
        err = dt(0.0)
        for v in diff.reshape(-1):
            err = v if err < v else err
        if err < tol:
            break
    return u, it, err


def solution2d(a, b, f, h):
    # Eikonal.h:14-20
    with np.errstate(invalid="ignore"):
        d = np.abs(a - b)
        x1 = smin(a, b) + f * h
        x2 = (a + b + np.sqrt(2 * f * f * h * h - (a - b) * (a - b))) / 2
    return np.where(d >= f * h, x1, x2)


def forward2d(f, srcx, srcy, h, max_iter=100):
    f = np.asarray(f, dtype=np.float64)
    n1, m1 = f.shape
    m, n = m1 - 1, n1 - 1
    ix, jx = srcx - 1, srcy - 1
    u = np.full(f.shape, 100000.0)
    u[jx, ix] = 0.0
    uold = u.copy()
    ii, jj = np.meshgrid(np.arange(m + 1), np.arange(n + 1), indexing="ij")
    s = (ii + jj).reshape(-1)
    order = np.argsort(s, kind="stable")
    I0, J0 = ii.reshape(-1)[order], jj.reshape(-1)[order]
    bounds = np.append(np.searchsorted(s[order], np.arange(m + n + 1)), len(s))
    it = 0
    conv = False
    for it in range(1, max_iter + 1):
        for di, dj in ((1, 1), (-1, 1), (-1, -1), (1, -1)):
            for sidx in range(m + n + 1):
                lo, hi = bounds[sidx], bounds[sidx + 1]
                i = I0[lo:hi] if di > 0 else m - I0[lo:hi]
                j = J0[lo:hi] if dj > 0 else n - J0[lo:hi]
                keep = ~((i == ix) & (j == jx))
                i, j = i[keep], j[keep]
                a = np.where(i == 0, u[j, np.minimum(i + 1, m)],
                             np.where(i == m, u[j, np.maximum(i - 1, 0)],
                                      smin(u[j, np.minimum(i + 1, m)], u[j, np.maximum(i - 1, 0)])))
                b = np.where(j == 0, u[np.minimum(j + 1, n), i],
                             np.where(j == n, u[np.maximum(j - 1, 0), i],
                                      smin(u[np.maximum(j - 1, 0), i], u[np.minimum(j + 1, n), i])))
                un = solution2d(a, b, f[j, i], h)
                cur = u[j, i]
                u[j, i] = np.where(un < cur, un, cur)
        sd = 0.0
        so = 0.0
        for p in range(u.size):
            dd = u.flat[p] - uold.flat[p]
            sd = sd + dd * dd
            so = so + uold.flat[p] * uold.flat[p]
        err = np.sqrt(sd) / np.sqrt(so)
        if err < 1e-8:
            conv = True
            break
        uold = u.copy()
    return u, it, conv
