"""The host optimiser (adtomo_hip.optim): L-BFGS with InitialStatic and
order-3 BackTracking, the method src/mpi_optimize.jl:35-39 builds.

Pinned by mathematics: a first iterate derived by hand, the two-loop
recursion against the dense BFGS inverse-Hessian update, the cubic backtrack
against the minimiser of a cubic found by calculus, and the minimiser of a
convex quadratic.  (Optim.jl and LineSearches.jl are not vendored in the
reference and Julia is absent here, so no trajectory of theirs is compared.)
"""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path[:0] = [ROOT, os.path.join(ROOT, "adtomo.jl_amd")]

from adtomo_hip import optim  # noqa: E402


def _quad2(x):
    # f = (x - 1)^2 + 10 (y - 2)^2
    return (x[0] - 1.0) ** 2 + 10.0 * (x[1] - 2.0) ** 2


def _quad2_g(x):
    return np.array([2.0 * (x[0] - 1.0), 20.0 * (x[1] - 2.0)])


def test_first_iterate_by_hand():
    # From (0, 0): g = (-2, -40), s = (2, 40), phi(0) = 41, phi'(0) = -1604.
    # alpha = 1: phi = f(2, 40) = 14441 fails the sufficient decrease, so the
    # quadratic interpolant's minimiser 1604 / 32008 ~ 0.0501 is clamped up to
    # 0.1 * 1; phi(0.1) = f(0.2, 4) = 40.64 <= 41 - 1e-4 * 0.1 * 1604 passes.
    calls = {"f": 0, "fg": 0}

    def f(x):
        calls["f"] += 1
        return _quad2(x)

    def fg(x):
        calls["fg"] += 1
        return _quad2(x), _quad2_g(x)

    r = optim.lbfgs(f, fg, [0.0, 0.0], iterations=1)
    assert r.x == pytest.approx([0.2, 4.0], abs=1e-15)
    assert r.fun == pytest.approx(40.64, abs=1e-12)
    assert r.iterations == 1 and calls == {"f": 2, "fg": 2}
    assert r.f_calls == 4 and r.g_calls == 2
    assert r.history[0] == 41.0


def test_two_loop_equals_dense_bfgs_update():
    # -H g with H built by the dense BFGS recursion
    #   H <- (I - rho s y^T) H (I - rho y s^T) + rho s s^T
    # from H0 = (s'y / y'y) I of the newest pair, oldest pair first
    rng = np.random.default_rng(5)
    n, m = 7, 4
    A = rng.normal(size=(n, n))
    A = A @ A.T + n * np.eye(n)  # SPD so that every s'y > 0
    pairs = []
    for _ in range(6):  # more pairs than the memory: only the last m count
        s = rng.normal(size=n)
        pairs.append((s, A @ s))
    dx = [np.zeros(n) for _ in range(m)]
    dg = [np.zeros(n) for _ in range(m)]
    rho = np.zeros(m)
    pseudo = 1
    for s, y in pairs:
        i = (pseudo - 1) % m
        dx[i], dg[i], rho[i] = s, y, 1.0 / s.dot(y)
        pseudo += 1
    g = rng.normal(size=n)
    d = optim.two_loop(g, dx, dg, rho, pseudo, m)
    kept = pairs[-m:]
    s_new, y_new = kept[-1]
    H = (s_new.dot(y_new) / y_new.dot(y_new)) * np.eye(n)
    eye = np.eye(n)
    for s, y in kept:
        r = 1.0 / s.dot(y)
        H = (eye - r * np.outer(s, y)) @ H @ (eye - r * np.outer(y, s)) + r * np.outer(s, s)
    assert np.allclose(d, -H @ g, rtol=1e-12, atol=1e-12)
    # first iteration (no pairs): the steepest descent direction
    assert np.array_equal(optim.two_loop(g, dx, dg, rho, 1, m), -g)


def test_cubic_backtrack_lands_on_the_cubic_minimiser():
    # phi(a) = -a + 1000 a^3: alpha = 1 fails, the quadratic step is clamped
    # to 0.1, which fails again (phi(0.1) = 0.9); the cubic through phi(0),
    # phi'(0), phi(1), phi(0.1) is phi itself, whose minimiser is
    # 1/sqrt(3000) (phi' = -1 + 3000 a^2), inside the clamp [0.01, 0.05]
    phi = lambda a: -a + 1000.0 * a ** 3  # noqa: E731
    a, pa = optim.backtracking(phi, 1.0, 0.0, -1.0)
    assert a == pytest.approx(1.0 / np.sqrt(3000.0), rel=1e-12)
    assert pa == pytest.approx(phi(a), rel=1e-15)


def test_backtrack_clamps_and_nonfinite_halving():
    # a value that is infinite beyond 0.3 halves alpha (1 -> 0.5 -> 0.25)
    # before any interpolation
    seen = []

    def phi(a):
        seen.append(a)
        return np.inf if a > 0.3 else -a

    a, _ = optim.backtracking(phi, 1.0, 0.0, -1.0)
    assert seen[:3] == [1.0, 0.5, 0.25] and a == 0.25
    # accepted at once when alpha = 1 already decreases enough
    a, pa = optim.backtracking(lambda a: -a, 1.0, 0.0, -1.0)
    assert (a, pa) == (1.0, -1.0)


def test_minimiser_of_a_convex_quadratic():
    rng = np.random.default_rng(11)
    n = 12
    A = rng.normal(size=(n, n))
    A = A @ A.T + 0.5 * np.eye(n)
    b = rng.normal(size=n)
    f = lambda x: 0.5 * x @ A @ x - b @ x  # noqa: E731
    fg = lambda x: (f(x), A @ x - b)  # noqa: E731
    r = optim.lbfgs(f, fg, np.zeros(n), iterations=500)
    # it stops once the backtracking cannot lower the value any further in
    # binary64 (x unchanged) or the gradient is below 1e-8
    assert r.converged and r.stop in ("gradient", "x")
    assert np.max(np.abs(r.jac)) <= 1e-6
    assert np.allclose(r.x, np.linalg.solve(A, b), rtol=0, atol=1e-6)
    assert all(h1 <= h0 for h0, h1 in zip(r.history, r.history[1:]))


def test_gradient_already_small_at_x0():
    r = optim.lbfgs(lambda x: 0.0, lambda x: (0.0, np.zeros(3)), np.ones(3))
    assert r.iterations == 0 and r.converged and r.g_calls == 1


def test_rosenbrock_reaches_the_minimum():
    def f(x):
        return (1 - x[0]) ** 2 + 100 * (x[1] - x[0] ** 2) ** 2

    def fg(x):
        g = np.array([-2 * (1 - x[0]) - 400 * x[0] * (x[1] - x[0] ** 2),
                      200 * (x[1] - x[0] ** 2)])
        return f(x), g

    r = optim.lbfgs(f, fg, [-1.2, 1.0], iterations=1000)
    assert r.converged
    assert np.allclose(r.x, [1.0, 1.0], atol=1e-6)
