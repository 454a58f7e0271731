"""The reference inversion's optimiser on the host: L-BFGS with a static
initial step and a backtracking line search.

src/mpi_optimize.jl:35-39 builds
``Optim.LBFGS(alphaguess=InitialStatic(), linesearch=BackTracking())`` and
runs ``Optim.optimize(f, g!, x0, method, Options(iterations=...))``
(:46-48, scripts/inversion.jl:129-131).  Optim.jl and LineSearches.jl are
not vendored in the reference (Project.toml lists Optim without a pinned
version), so this module restates their published algorithms with their
default parameters:

* L-BFGS (Nocedal & Wright, Numerical Optimization, 2nd ed., Algorithm 7.4,
  the two-loop recursion) with a history of m = 10 pairs, the initial inverse
  Hessian scaled by s'y / y'y of the newest pair (their eq. 7.20) from the
  second iteration on, a pair dropped (and the history restarted) when
  1 / s'y is infinite, and the direction reset to -g (history restarted)
  when it is not a descent direction.
* InitialStatic: every line search starts at alpha = 1.
* BackTracking (order 3): c1 = 1e-4, rho_hi = 0.5, rho_lo = 0.1, at most
  1000 backtracks.  A non-finite value halves alpha (at most 52 times);
  while phi(alpha) > phi(0) + c1 alpha phi'(0) the step is replaced by the
  minimiser of the quadratic interpolant of phi(0), phi'(0), phi(alpha) on
  the first backtrack and of the cubic interpolant through the last two
  trial values afterwards, clamped to [rho_lo alpha, rho_hi alpha].
* Stopping (Optim.Options defaults): max |g| <= 1e-8 (also tested at x0),
  x or f unchanged (zero tolerances; f must stay unchanged on two
  successive iterations), the iteration limit, or a failed line search.

Call pattern, as Optim drives a pair of separate f and g! (NLSolversBase's
OnceDifferentiable): the value and the gradient at x0, the value alone at
each trial point of a line search, and the gradient with the value at each
accepted point.  ``fg(x)`` returns both (one gradient call), ``f(x)`` the
value.
"""
from __future__ import annotations

import math

import numpy as np

_EPS = np.finfo(np.float64).eps


class LineSearchFailed(Exception):
    def __init__(self, msg, alpha):
        super().__init__(msg)
        self.alpha = alpha


def backtracking(phi, alpha0, phi0, dphi0, c1=1e-4, rho_hi=0.5, rho_lo=0.1, iterations=1000,
                 order=3):
    """LineSearches.BackTracking on phi(alpha) = f(x + alpha s): returns
    (alpha, phi(alpha)).  Raises LineSearchFailed (carrying the last alpha)
    after ``iterations`` backtracks."""
    a1 = a2 = float(alpha0)
    p_old = p_new = phi0
    p_new = phi(a1)
    # halve until the value is finite
    it_finite = 0
    it_finite_max = int(-math.log2(_EPS))
    while not math.isfinite(p_new) and it_finite < it_finite_max:
        it_finite += 1
        a1 = a2
        a2 = a1 / 2
        p_new = phi(a2)
    it = 0
    while p_new > phi0 + c1 * a2 * dphi0:
        it += 1
        if it > iterations:
            raise LineSearchFailed(f"line search reached {iterations} backtracks", a2)
        if order == 2 or it == 1:
            # minimiser of the quadratic through phi(0), phi'(0), phi(a2)
            a_tmp = -(dphi0 * a2 ** 2) / (2 * (p_new - phi0 - dphi0 * a2))
        else:
            # minimiser of the cubic through phi(0), phi'(0), phi(a1), phi(a2)
            div = 1.0 / (a1 ** 2 * a2 ** 2 * (a2 - a1))
            r2 = p_new - phi0 - dphi0 * a2
            r1 = p_old - phi0 - dphi0 * a1
            a = (a1 ** 2 * r2 - a2 ** 2 * r1) * div
            b = (-a1 ** 3 * r2 + a2 ** 3 * r1) * div
            if abs(a) <= _EPS:
                a_tmp = dphi0 / (2 * b)
            else:
                d = max(b * b - 3 * a * dphi0, 0.0)
                a_tmp = (-b + math.sqrt(d)) / (3 * a)
        a1 = a2
        a_tmp = _nanmin(a_tmp, a2 * rho_hi)  # not too small a reduction
        a2 = _nanmax(a_tmp, a2 * rho_lo)     # not too large a reduction
        p_old, p_new = p_new, phi(a2)
    return a2, p_new


def _nanmin(x, y):
    return y if math.isnan(x) else (x if math.isnan(y) else min(x, y))


def _nanmax(x, y):
    return y if math.isnan(x) else (x if math.isnan(y) else max(x, y))


def two_loop(g, dx_hist, dg_hist, rho, pseudo, m, scale_h0=True):
    """Search direction -H g of the two-loop recursion over the stored pairs
    (slots (index-1) % m for index = pseudo-m .. pseudo-1)."""
    upper = pseudo - 1
    lower = upper - m + 1
    q = np.array(g, dtype=np.float64, copy=True)
    alpha = np.zeros(m)
    for index in range(upper, lower - 1, -1):
        if index < 1:
            continue
        i = (index - 1) % m
        alpha[i] = rho[i] * np.dot(dx_hist[i], q)
        q -= alpha[i] * dg_hist[i]
    if scale_h0 and pseudo > 1:
        i = (upper - 1) % m
        s = (np.dot(dx_hist[i], dg_hist[i]) / np.dot(dg_hist[i], dg_hist[i])) * q
    else:
        s = q
    for index in range(lower, upper + 1):
        if index < 1:
            continue
        i = (index - 1) % m
        beta = rho[i] * np.dot(dg_hist[i], s)
        s = s + dx_hist[i] * (alpha[i] - beta)
    return -s


class Result:
    def __init__(self, x, fx, g, iterations, f_calls, g_calls, converged, stop, history):
        self.x, self.fun, self.jac = x, fx, g
        self.iterations, self.f_calls, self.g_calls = iterations, f_calls, g_calls
        self.converged, self.stop, self.history = converged, stop, history

    def __repr__(self):
        return (f"Result(fun={self.fun!r}, iterations={self.iterations}, f_calls={self.f_calls}, "
                f"g_calls={self.g_calls}, stop={self.stop!r})")


def lbfgs(f, fg, x0, iterations=1000, m=10, g_abstol=1e-8, successive_f_tol=1,
          linesearch=None, callback=None):
    """Minimises f from x0 as Optim.LBFGS(InitialStatic, BackTracking) does.

    ``f(x)`` -> value; ``fg(x)`` -> (value, gradient).  ``callback(k, x, fx,
    g)`` runs after each iteration k (1-based).  Returns a Result with
    ``x``, ``fun``, ``jac``, ``iterations``, ``f_calls``, ``g_calls``,
    ``stop`` and ``history`` (the value after each iteration, x0's first)."""
    ls = linesearch or backtracking
    x = np.array(x0, dtype=np.float64, copy=True).reshape(-1)
    n_f = n_g = 0
    fx, g = fg(x)
    g = np.asarray(g, dtype=np.float64).reshape(-1)
    n_f += 1
    n_g += 1
    history = [float(fx)]
    if np.max(np.abs(g), initial=0.0) <= g_abstol:
        return Result(x, fx, g, 0, n_f, n_g, True, "gradient", history)
    dx_hist = [np.zeros_like(x) for _ in range(m)]
    dg_hist = [np.zeros_like(x) for _ in range(m)]
    rho = np.zeros(m)
    pseudo = 1
    counter_f = 0
    stop = "iterations"
    converged = False
    k = 0
    while k < iterations:
        k += 1
        s = two_loop(g, dx_hist, dg_hist, rho, pseudo, m)
        g_prev = g.copy()
        dphi0 = float(np.dot(g, s))
        if dphi0 >= 0:  # not a descent direction: restart from -g
            pseudo = 1
            s = -g
            dphi0 = float(np.dot(g, s))
        phi0 = fx
        x_prev, f_prev = x.copy(), fx

        def phi(a):
            nonlocal n_f
            n_f += 1
            return float(f(x + a * s))

        ok = True
        try:
            alpha, _ = ls(phi, 1.0, phi0, dphi0)
        except LineSearchFailed as e:
            alpha, ok = e.alpha, False
        dx = alpha * s
        x = x + dx
        if not ok:
            stop = "line search"
            break
        fx, g = fg(x)
        g = np.asarray(g, dtype=np.float64).reshape(-1)
        n_f += 1
        n_g += 1
        history.append(float(fx))
        x_conv = float(np.max(np.abs(x - x_prev), initial=0.0)) <= 0.0
        f_conv = abs(fx - f_prev) <= 0.0
        g_conv = float(np.max(np.abs(g), initial=0.0)) <= g_abstol
        counter_f = counter_f + 1 if f_conv else 0
        converged = x_conv or g_conv or counter_f > successive_f_tol
        if not converged:
            dg = g - g_prev
            sy = float(np.dot(dx, dg))
            r = math.inf if sy == 0.0 else 1.0 / sy
            if math.isinf(r):
                pseudo = 1
            else:
                i = (pseudo - 1) % m
                dx_hist[i] = dx.copy()
                dg_hist[i] = dg
                rho[i] = r
                pseudo += 1
        if callback is not None:
            callback(k, x, fx, g)
        if converged:
            stop = "gradient" if g_conv else ("x" if x_conv else "f")
            break
        if not np.all(np.isfinite(g)):
            stop = "non-finite gradient"
            break
    return Result(x, fx, g, k, n_f, n_g, converged, stop, history)
