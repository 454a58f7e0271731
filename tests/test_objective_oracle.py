"""Pins of the objective oracle (oracle/objective.py) on the CPU.

The reference objective (scripts/inversion.jl:42-121) runs inside Julia +
TensorFlow, which are absent here, and the reference holds no vectors for
it, so each function is pinned by what the mathematics fixes:
  * source_u0: grid-node stations (all 8 corners coincide), half-cell
    stations (closed-form distances), the 1000 far value elsewhere;
  * interpolate: exact on trilinear fields (a wrong corner, weight or axis
    breaks it), equals node values at integer points, and agrees with the
    independent weight table of interpolation_weights;
  * misfit: central differences of the loss (quadratic in u, so differences
    are exact up to rounding), the -1 sentinel, and the closed form for a
    single event;
  * model / sigmoid: scipy.special.expit, fvar = vel0 at x = 0;
  * box_mean: scipy.ndimage.uniform_filter with periodic wrap;
  * smooth_l1 and model_grad: central differences of the composite.
"""
import numpy as np
import pytest
from scipy import ndimage, special

from oracle import objective as O


def _rng(seed):
    return np.random.default_rng(seed)


# ---------------------------------------------------------------- source_u0

def test_source_u0_grid_node_station():
    v = _rng(0).uniform(3.0, 8.0, (6, 5, 4))
    u0 = O.source_u0(v.shape, 3.0, 2.0, 4.0, v, 0.5)
    assert u0[2, 1, 3] == 0.0
    mask = np.ones(v.shape, bool)
    mask[2, 1, 3] = False
    assert np.all(u0[mask] == 1000.0)


def test_source_u0_half_cell_station_closed_form():
    v = _rng(1).uniform(3.0, 8.0, (6, 5, 4))
    h, scale = 0.25, 1.7
    u0 = O.source_u0(v.shape, 2.5, 3.5, 1.5, v, h, scale)
    d = np.sqrt(3 * 0.25)
    idx = [(i, j, k) for i in (1, 2) for j in (2, 3) for k in (0, 1)]
    for i, j, k in idx:
        assert u0[i, j, k] == pytest.approx(d * h / v[i, j, k] * scale, rel=1e-15)
    assert np.count_nonzero(u0 != 1000.0) == 8


def test_source_u0_mixed_axes():
    # integer in x only: 4 distinct corners
    v = np.full((5, 5, 5), 4.0)
    u0 = O.source_u0(v.shape, 3.0, 2.25, 4.75, v, 1.0)
    assert np.count_nonzero(u0 != 1000.0) == 4
    assert u0[2, 1, 4] == pytest.approx(np.hypot(0.25, 0.25) / 4.0, rel=1e-15)
    assert u0[2, 2, 3] == pytest.approx(np.hypot(0.75, 0.75) / 4.0, rel=1e-15)


# ---------------------------------------------------------- interpolation

def _trilinear_field(shape, c):
    i, j, k = np.meshgrid(*(np.arange(1, s + 1, dtype=float) for s in shape), indexing="ij")
    return (c[0] + c[1] * i + c[2] * j + c[3] * k + c[4] * i * j + c[5] * j * k + c[6] * i * k
            + c[7] * i * j * k), lambda x, y, z: (c[0] + c[1] * x + c[2] * y + c[3] * z
                                                   + c[4] * x * y + c[5] * y * z + c[6] * x * z
                                                   + c[7] * x * y * z)


def test_interpolate_exact_on_trilinear_fields():
    r = _rng(2)
    shape = (7, 6, 5)
    for _ in range(5):
        c = r.normal(size=8)
        u, fn = _trilinear_field(shape, c)
        for _ in range(40):
            x, y, z = (r.uniform(1, s) for s in shape)
            # integer coordinates on some axes too
            if r.random() < 0.3:
                x = float(r.integers(1, shape[0] + 1))
            if r.random() < 0.3:
                z = float(r.integers(1, shape[2] + 1))
            assert O.interpolate(u, x, y, z) == pytest.approx(fn(x, y, z), rel=1e-12, abs=1e-12)


def test_interpolate_node_values_and_weights():
    r = _rng(3)
    u = r.normal(size=(5, 6, 7))
    for i, j, k in [(1, 1, 1), (5, 6, 7), (3, 2, 4)]:
        assert O.interpolate(u, float(i), float(j), float(k)) == u[i - 1, j - 1, k - 1]
    for _ in range(50):
        x, y, z = r.uniform(1, 5), r.uniform(1, 6), r.uniform(1, 7)
        w = O.interpolation_weights(u.shape, x, y, z)
        assert sum(w.values()) == pytest.approx(1.0, rel=1e-14)
        assert all(v >= 0 for v in w.values())
        lin = sum(wt * u[key] for key, wt in w.items())
        assert O.interpolate(u, x, y, z) == pytest.approx(lin, rel=1e-13, abs=1e-13)
        # the chain-rule form carries the same weights, scaled by the seed
        g = r.normal()
        adj = O.interpolation_adjoint(g, x, y, z)
        assert adj.keys() == w.keys()
        for key, wt in w.items():
            assert adj[key] == pytest.approx(g * wt, rel=1e-15, abs=1e-300)


# ------------------------------------------------------------------ misfit

def _events(r, shape, count):
    return (r.uniform(1, shape[0], count), r.uniform(1, shape[1], count),
            r.uniform(1, shape[2], count))


def test_misfit_gradient_central_differences():
    r = _rng(4)
    shape = (5, 4, 6)
    u = r.uniform(0, 5, shape)
    rx, ry, rz = _events(r, shape, 9)
    rx[0] = 3.0  # integer coordinate on x
    tobs = r.uniform(0, 5, 9)
    tobs[2] = -1.0
    w = r.uniform(0.5, 2.0, 9)
    _, loss, grad = O.misfit(u, rx, ry, rz, tobs, w)
    eps = 1e-5
    for idx in np.ndindex(shape):
        up = u.copy()
        up[idx] += eps
        dn = u.copy()
        dn[idx] -= eps
        fd = (O.misfit(up, rx, ry, rz, tobs, w)[1] - O.misfit(dn, rx, ry, rz, tobs, w)[1]) / (2 * eps)
        assert fd == pytest.approx(grad[idx], rel=1e-6, abs=1e-8)


def test_misfit_sentinel_and_single_event():
    r = _rng(5)
    u = r.uniform(0, 5, (4, 4, 4))
    x, y, z = 2.25, 3.5, 1.75
    t = O.interpolate(u, x, y, z)
    tcal, loss, grad = O.misfit(u, [x], [y], [z], [t + 0.5], [3.0])
    assert tcal[0] == t
    assert loss == pytest.approx(3.0 * 0.25, rel=1e-14)
    assert grad.sum() == pytest.approx(-3.0 * 2 * 0.5, rel=1e-13)
    _, loss, grad = O.misfit(u, [x], [y], [z], [-1.0], [3.0])
    assert loss == 0.0 and not grad.any()


# ------------------------------------------------------- model / smoothing

def test_model_sigmoid():
    r = _rng(6)
    x = r.normal(scale=3, size=(4, 5, 6))
    vel0 = r.uniform(3, 8, x.shape)
    np.testing.assert_allclose(O.sigmoid(x), special.expit(x), rtol=1e-15)
    fvar, f = O.model(x, vel0)
    np.testing.assert_allclose(fvar, 2 * special.expit(x) - 1 + vel0, rtol=1e-15)
    np.testing.assert_allclose(f * fvar, 1.0, rtol=1e-15)
    fvar0, _ = O.model(np.zeros_like(x), vel0)
    np.testing.assert_array_equal(fvar0, vel0)


@pytest.mark.parametrize("sh,sv", [(1, 1), (3, 3), (5, 3), (3, 1), (7, 5)])
def test_box_mean_is_periodic_uniform_filter(sh, sv):
    a = _rng(7).normal(size=(9, 8, 7))
    ref = ndimage.uniform_filter(a, size=(sh, sh, sv), mode="wrap")
    np.testing.assert_allclose(O.box_mean(a, sh, sv), ref, rtol=1e-12, atol=1e-13)


def test_smooth_l1_and_model_grad_central_differences():
    r = _rng(8)
    shape = (5, 6, 4)
    vel0 = r.uniform(3, 8, shape)
    x = r.normal(scale=0.5, size=shape)
    a = r.normal(size=shape)  # dL/df of a linear term sum(a * f)
    lam, sh, sv = 0.7, 3, 3

    def total(xv):
        fvar, f = O.model(xv, vel0)
        return (a * f).sum() + O.smooth_l1(fvar, sh, sv, lam)[0]

    fvar, _ = O.model(x, vel0)
    _, gfvar = O.smooth_l1(fvar, sh, sv, lam)
    gx = O.model_grad(x, fvar, a, gfvar)
    eps = 1e-6
    for idx in list(np.ndindex(shape))[::3]:
        xp = x.copy()
        xp[idx] += eps
        xm = x.copy()
        xm[idx] -= eps
        fd = (total(xp) - total(xm)) / (2 * eps)
        assert fd == pytest.approx(gx[idx], rel=1e-5, abs=1e-7)


def test_smooth_l1_constant_field_is_zero():
    v = np.full((4, 5, 6), 6.0)
    val, g = O.smooth_l1(v, 3, 3, 2.0)
    assert val == pytest.approx(0.0, abs=1e-12)


# ------------------------------------------------- Huber and S slowness

def test_huber_matches_square_inside_and_is_c1():
    for r in (-0.05, 0.0, 0.07):
        assert O.huber(r, 0.1) == (r * r, 2 * r)
        assert O.huber(r, 0.0) == (r * r, 2 * r)
    d = 0.1
    for r in (0.3, -0.25):
        rho, drho = O.huber(r, d)
        assert rho == pytest.approx(2 * d * abs(r) - d * d, rel=1e-15)
        assert drho == pytest.approx(2 * d * np.sign(r), rel=1e-15)
    # continuity of value and slope at |r| = delta
    eps = 1e-9
    for s in (1, -1):
        a, b = O.huber(s * (d - eps), d), O.huber(s * (d + eps), d)
        assert a[0] == pytest.approx(b[0], abs=1e-9) and a[1] == pytest.approx(b[1], abs=1e-8)


def test_huber_misfit_gradient_central_differences():
    r = _rng(9)
    shape = (4, 5, 4)
    u = r.uniform(0, 2, shape)
    rx, ry, rz = _events(r, shape, 12)
    tobs = np.array([O.interpolate(u, a, b, c) for a, b, c in zip(rx, ry, rz)])
    tobs += r.choice([-1, 1], 12) * r.uniform(0.0, 0.3, 12)  # both sides of delta
    w = r.uniform(0.5, 2.0, 12)
    _, loss, grad = O.misfit(u, rx, ry, rz, tobs, w, delta=0.1)
    eps = 1e-7
    for idx in list(np.ndindex(shape))[::2]:
        up = u.copy()
        up[idx] += eps
        dn = u.copy()
        dn[idx] -= eps
        fd = (O.misfit(up, rx, ry, rz, tobs, w, delta=0.1)[1]
              - O.misfit(dn, rx, ry, rz, tobs, w, delta=0.1)[1]) / (2 * eps)
        assert fd == pytest.approx(grad[idx], rel=1e-5, abs=1e-7)
    # a huge delta gives the squared misfit
    assert O.misfit(u, rx, ry, rz, tobs, w, delta=1e9)[1] == O.misfit(u, rx, ry, rz, tobs, w)[1]


def test_slowness_grad_central_differences():
    r = _rng(10)
    fvar = r.uniform(3, 8, (3, 4, 5))
    a = r.normal(size=fvar.shape)
    s = 1.73
    gfvar, gs = O.slowness_grad(fvar, s, a)
    eps = 1e-6
    fd_s = ((a * O.slowness(fvar, s + eps)).sum() - (a * O.slowness(fvar, s - eps)).sum()) / (2 * eps)
    assert fd_s == pytest.approx(gs, rel=1e-8)
    for idx in list(np.ndindex(fvar.shape))[::7]:
        fp = fvar.copy()
        fp[idx] += eps
        fm = fvar.copy()
        fm[idx] -= eps
        fd = ((a * O.slowness(fp, s)).sum() - (a * O.slowness(fm, s)).sum()) / (2 * eps)
        assert fd == pytest.approx(gfvar[idx], rel=1e-6, abs=1e-10)
