"""Pins of the CPU oracle (oracle/eikonal_oracle.c) on the CPU.

The reference op cannot be built or run here (Eigen, TensorFlow, ADCME and
Julia are absent) and its own tests hold no golden vectors, so the oracle is
pinned by checks that a plausible transcription mistake would fail:
  * an independent numpy hyperplane-order formulation (tests/_hyperplane.py)
    must agree bit for bit (wrong index/edge rule/sweep order/argument order);
  * closed forms (1D axis times, the 2D diagonal formula, symmetry);
  * invariants (u <= u0, u >= 0, source time, monotone iterations);
  * the adjoint against the reference's own sparse system assembled from
    Eikonal.h:105-189 / Eikonal3D.cpp:118-187 and solved with SuperLU and
    dense LU (wrong sign, dropped term, transposed system);
  * central finite differences and a Taylor-remainder test of the gradient,
    mirroring deps/CustomOps/Eikonal{,3D}/gradtest.jl.
"""
import numpy as np
import pytest

from oracle import oracle
from adtomo_hip import synth

import _hyperplane as hp


# ------------------------------------------------------------------ forward 3D

@pytest.mark.parametrize("shape,seed", [((7, 6, 5), 0), ((9, 5, 11), 1), ((2, 3, 4), 2), ((12, 10, 8), 3)])
def test_forward3d_matches_hyperplane_order(shape, seed):
    r = synth.rng(seed)
    f = r.uniform(0.5, 2.0, size=shape)
    src = tuple(int(r.integers(0, s)) for s in shape)
    u0 = synth.point_source_u0(shape, src)
    for tol in (1e-3, 0.0):
        u, it, err = oracle.eikonal3d(u0, f, 0.7, tol)
        uh, ith, errh = hp.forward3d(u0, f, 0.7, tol)
        assert it == ith
        assert err == errh
        np.testing.assert_array_equal(u.view(np.uint64), uh.view(np.uint64))


@pytest.mark.parametrize("shape,seed", [((7, 6, 5), 20), ((9, 5, 11), 21), ((2, 3, 4), 22)])
def test_forward3d_mixed_matches_hyperplane_order(shape, seed):
    # FP32 storage, FP64 local solves: the independent numpy formulation
    # (binary32 arrays, binary64 solve rounded once) agrees bit for bit
    r = synth.rng(seed)
    f = r.uniform(0.5, 2.0, size=shape).astype(np.float32)
    src = tuple(int(r.integers(0, s)) for s in shape)
    u0 = synth.point_source_u0(shape, src).astype(np.float32)
    for tol in (1e-3, 0.0):
        u, it, err = oracle.eikonal3d_m32(u0, f, 0.7, tol)
        uh, ith, errh = hp.forward3d(u0, f, 0.7, tol, mixed=True)
        assert u.dtype == np.float32 and it == ith and err == float(errh)
        np.testing.assert_array_equal(u.view(np.uint32), uh.view(np.uint32))


def test_forward3d_mixed_is_fp64_rounded_within_storage_error():
    # on a grid where the binary32 twin loses accuracy to cancellation in the
    # 3D quadratic (travel times large against f*h), the mixed variant stays
    # within a few binary32 ulps of FP64 while the binary32 twin does not
    shape = (40, 36, 20)
    f = synth.rng(23).uniform(0.5, 0.6, size=shape)
    u0 = synth.point_source_u0(shape, (0, 0, 0))
    h = 0.05
    u64, _, _ = oracle.eikonal3d(u0, f, h, 0.0, max_iter=6)
    um, _, _ = oracle.eikonal3d_m32(u0, f, h, 0.0, max_iter=6)
    u32, _, _ = oracle.eikonal3d(u0, f, h, 0.0, dtype=np.float32, max_iter=6)
    em = np.abs(um.astype(np.float64) - u64).max() / u64.max()
    e32 = np.abs(u32.astype(np.float64) - u64).max() / u64.max()
    assert em < 2e-6
    assert e32 > 10 * em


def test_forward3d_random_u0_matches_hyperplane():
    # gradtest.jl style: u0 = rand * 1000 (deps/CustomOps/Eikonal3D/gradtest.jl:43)
    r = synth.rng(11)
    shape = (6, 7, 5)
    u0 = r.uniform(0, 1000, size=shape)
    f = np.ones(shape)
    u, it, err = oracle.eikonal3d(u0, f, 0.01, 1e-6)
    uh, ith, errh = hp.forward3d(u0, f, 0.01, 1e-6)
    assert it == ith
    np.testing.assert_array_equal(u, uh)


def test_forward3d_axis_closed_form():
    # Constant slowness, source on a node: along the grid axes through the
    # source the scheme is exactly 1D, t = repeated addition of f*h.
    shape = (15, 13, 11)
    src = (5, 6, 4)
    fh = 0.3 * 0.7
    u, it, err = oracle.eikonal3d(synth.point_source_u0(shape, src), np.full(shape, 0.3), 0.7, 0.0)
    for ax in range(3):
        for sign in (1, -1):
            t = 0.0
            p = list(src)
            while True:
                p[ax] += sign
                if p[ax] < 0 or p[ax] >= shape[ax]:
                    break
                t = t + fh
                assert u[tuple(p)] == t
    assert u[src] == 0.0


def test_forward3d_diagonal_2d_closed_form():
    # Neighbour of the source across a face diagonal: two equal upwind
    # values a = f*h and an inactive third -> root of 2(x-a)^2 = (fh)^2,
    # x = a + fh/sqrt(2) (to rounding of the quadratic formula).
    shape = (9, 9, 9)
    src = (4, 4, 4)
    u, *_ = oracle.eikonal3d(synth.point_source_u0(shape, src), np.ones(shape), 1.0, 0.0)
    assert abs(u[5, 5, 4] - (1.0 + 1.0 / np.sqrt(2.0))) < 1e-14
    # Body diagonal: three equal upwind values b = 1 + 1/sqrt(2) -> b + 1/sqrt(3).
    b = 1.0 + 1.0 / np.sqrt(2.0)
    assert abs(u[5, 5, 5] - (b + 1.0 / np.sqrt(3.0))) < 1e-14


def test_forward3d_symmetry_and_invariants():
    shape = (11, 11, 9)
    src = (5, 5, 4)
    r = synth.rng(5)
    f = r.uniform(0.8, 1.2, size=shape)
    f = 0.5 * (f + f[::-1, :, :])  # mirror-symmetric in i about the source plane
    u0 = synth.point_source_u0(shape, src)
    u, it, err = oracle.eikonal3d(u0, f, 0.5, 0.0)
    assert err == 0.0  # converged to the exact discrete fixed point
    np.testing.assert_array_equal(u, u[::-1, :, :])
    assert (u >= 0).all() and (u <= u0).all() and u[src] == 0.0


def test_forward3d_iterations_monotone():
    shape = (10, 9, 8)
    src = (3, 4, 2)
    r = synth.rng(9)
    f = r.uniform(0.5, 2, size=shape)
    u0 = synth.point_source_u0(shape, src)
    prev = u0
    for k in range(1, 5):
        u, it, err = oracle.eikonal3d(u0, f, 1.0, -1.0, max_iter=k)
        assert it == k
        assert (u <= prev).all()
        prev = u


def test_forward3d_f32_close_to_f64():
    shape = (16, 14, 12)
    f = 1.0 / synth.layered_grf_model(shape, 0.5, seed=3)
    u0 = synth.point_source_u0(shape, (8, 7, 3))
    u64, *_ = oracle.eikonal3d(u0, f, 0.5, 1e-6)
    u32, *_ = oracle.eikonal3d(u0, f, 0.5, 1e-6, dtype=np.float32)
    assert u32.dtype == np.float32
    rel = np.abs(u32 - u64).max() / np.abs(u64).max()
    assert rel < 1e-5


# ------------------------------------------------------------------ forward 2D

@pytest.mark.parametrize("m,n,seed", [(12, 9, 0), (5, 17, 1), (1, 1, 2), (20, 20, 3)])
def test_forward2d_matches_hyperplane_order(m, n, seed):
    r = synth.rng(seed)
    f = r.uniform(0.5, 2.0, size=(n + 1, m + 1))
    srcx = int(r.integers(1, m + 2))
    srcy = int(r.integers(1, n + 2))
    u, it, err, conv = oracle.eikonal2d(f, srcx, srcy, 0.3)
    uh, ith, convh = hp.forward2d(f, srcx, srcy, 0.3)
    assert (it, conv) == (ith, convh)
    np.testing.assert_array_equal(u, uh)


def test_forward2d_axis_and_diagonal_closed_form():
    m, n = 20, 14
    f = np.full((n + 1, m + 1), 0.25)
    h = 0.4
    srcx, srcy = 8, 6  # 1-based column / row
    u, it, err, conv = oracle.eikonal2d(f, srcx, srcy, h)
    assert conv
    ix, jx = srcx - 1, srcy - 1
    assert u[jx, ix] == 0.0
    t = 0.0
    for i in range(ix + 1, m + 1):
        t = t + 0.25 * h
        assert u[jx, i] == t
    # diagonal neighbour: a = b = fh -> (2a + sqrt(2 f^2 h^2)) / 2
    fh = 0.25 * h
    assert abs(u[jx + 1, ix + 1] - (fh + fh / np.sqrt(2.0))) < 1e-15


def test_forward2d_nonconvergence_reported():
    f = np.ones((6, 7))
    u, it, err, conv = oracle.eikonal2d(f, 2, 3, 1.0, max_iter=1)
    assert not conv and it == 1


# -------------------------------------------------------------- adjoint 3D

def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("shape,seed", [((8, 7, 6), 0), ((13, 11, 9), 1), ((21, 21, 21), 2)])
def test_adjoint3d_vs_sparse_lu(shape, seed):
    r = synth.rng(seed)
    f = r.uniform(0.5, 1.5, size=shape)
    src = tuple(int(r.integers(1, s - 1)) for s in shape)
    u0 = synth.point_source_u0(shape, src)
    u, it, err = oracle.eikonal3d(u0, f, 0.5, 1e-6)
    g = r.normal(size=shape)
    gu0, gf, lam = oracle.eikonal3d_grad(g, u, u0, f, 0.5)
    gu0_lu, gf_lu, lam_lu = oracle.lu_grad_3d(g, u, u0, f, 0.5)
    np.testing.assert_array_equal(gu0, gu0_lu)
    assert _rel(lam, lam_lu) < 1e-10
    assert _rel(gf, gf_lu) < 1e-10


def test_adjoint3d_vs_dense_solve_with_random_u0():
    # gradtest.jl:43-49 style random u0 gives a large Z set.
    r = synth.rng(4)
    shape = (5, 6, 4)
    u0 = r.uniform(0, 1000, size=shape)
    f = r.uniform(0.5, 1.5, size=shape)
    u, *_ = oracle.eikonal3d(u0, f, 3.0, 1e-9)
    g = r.normal(size=shape)
    gu0, gf, lam = oracle.eikonal3d_grad(g, u, u0, f, 3.0)
    A, rhs, Z = oracle.assemble_3d(u, u0, f, 3.0)
    gz = g.reshape(-1).copy()
    gz[Z] = 0.0
    lam_d = np.linalg.solve(A.toarray().T, gz)
    assert Z.any() and (~Z).any()
    assert _rel(lam.reshape(-1), lam_d) < 1e-10
    np.testing.assert_array_equal(gu0.reshape(-1), np.where(u.reshape(-1) == u0.reshape(-1), g.reshape(-1), 0))


def _loss3d(f, u0, h, g):
    u, it, err = oracle.eikonal3d(u0, f, h, 0.0, max_iter=20)
    return float(np.sum(g * u)), err


def test_adjoint3d_central_fd():
    # Checkerboard-layered model (config 2 family, reduced size), tight tol.
    shape = (20, 18, 12)
    v = synth.checkerboard_model(shape, period=8)
    f = 1.0 / v
    src = (9, 8, 5)
    u0 = synth.point_source_u0(shape, src)
    h = 1.0
    g = synth.sparse_adjoint_source(shape, 40, seed=21, scale=1.0)
    u, it, err = oracle.eikonal3d(u0, f, h, 0.0)
    assert err == 0.0
    gu0, gf, lam = oracle.eikonal3d_grad(g, u, u0, f, h)
    r = synth.rng(22)
    cand = np.argwhere(np.abs(gf) > 1e-3 * np.abs(gf).max())
    picks = cand[r.choice(len(cand), size=20, replace=False)]
    checked = 0
    for p in picks:
        p = tuple(p)
        dlt = 1e-6 * f[p]
        fp = f.copy(); fp[p] += dlt
        fm = f.copy(); fm[p] -= dlt
        lp, ep = _loss3d(fp, u0, h, g)
        lm, em = _loss3d(fm, u0, h, g)
        fd = (lp - lm) / (2 * dlt)
        assert abs(fd - gf[p]) <= 1e-5 * abs(gf[p]) + 1e-9, (p, fd, gf[p])
        checked += 1
    assert checked == 20


def test_adjoint3d_taylor_remainder():
    # deps/CustomOps/Eikonal3D/gradtest.jl:51-63 structure with a slowness
    # perturbation: |L(f+gv)-L(f)| = O(g), |L(f+gv)-L(f)-g<v,grad>| = O(g^2).
    shape = (12, 11, 10)
    r = synth.rng(31)
    f = r.uniform(0.8, 1.2, size=shape)
    u0 = synth.point_source_u0(shape, (6, 5, 4))
    h = 0.5
    g = r.normal(size=shape)
    u, *_ = oracle.eikonal3d(u0, f, h, 0.0)
    _, gf, _ = oracle.eikonal3d_grad(g, u, u0, f, h)
    v = r.uniform(size=shape)
    L0, _ = _loss3d(f, u0, h, g)
    ws = []
    for gam in (1e-3, 1e-4, 1e-5):
        Lg, _ = _loss3d(f + gam * v, u0, h, g)
        ws.append(abs(Lg - L0 - gam * np.sum(v * gf)))
    assert ws[1] < ws[0] / 30 and ws[2] < ws[1] / 30


# -------------------------------------------------------------- adjoint 2D

@pytest.mark.parametrize("m,n,seed", [(10, 7, 0), (30, 20, 1), (60, 30, 2)])
def test_adjoint2d_vs_sparse_lu(m, n, seed):
    r = synth.rng(seed)
    f = r.uniform(0.5, 1.5, size=(n + 1, m + 1))
    srcx, srcy = int(r.integers(2, m)), int(r.integers(2, n))
    h = 0.1
    u, it, err, conv = oracle.eikonal2d(f, srcx, srcy, h)
    g = r.normal(size=f.shape)
    gf, lam, ns = oracle.eikonal2d_grad(g, u, f, srcx, srcy, h)
    assert ns == 0
    gf_lu, lam_lu = oracle.lu_grad_2d(g, u, f, srcx, srcy, h)
    assert _rel(lam, lam_lu) < 1e-10
    assert _rel(gf, gf_lu) < 1e-10
    assert gf[srcy - 1, srcx - 1] == 0.0


def test_adjoint2d_gradtest_layer_fd():
    # deps/CustomOps/Eikonal/gradtest.jl:41-50: m=60, n=30, h=0.1, slow layer
    # f[12:18,:] = 10, source (30, 3); loss sum(u^2).
    m, n, h = 60, 30, 0.1
    f = np.ones((n + 1, m + 1))
    f[11:18, :] = 10.0
    srcx, srcy = 30, 3
    u, it, err, conv = oracle.eikonal2d(f, srcx, srcy, h)
    assert conv
    gf, lam, ns = oracle.eikonal2d_grad(2 * u, u, f, srcx, srcy, h)
    r = synth.rng(3)
    cand = np.argwhere(np.abs(gf) > 1e-3 * np.abs(gf).max())
    for p in cand[r.choice(len(cand), size=12, replace=False)]:
        p = tuple(p)
        d = 1e-6 * f[p]
        fp = f.copy(); fp[p] += d
        fm = f.copy(); fm[p] -= d
        up, *_ = oracle.eikonal2d(fp, srcx, srcy, h)
        um, *_ = oracle.eikonal2d(fm, srcx, srcy, h)
        fd = (np.sum(up ** 2) - np.sum(um ** 2)) / (2 * d)
        assert abs(fd - gf[p]) <= 1e-4 * abs(gf[p]) + 1e-8, (p, fd, gf[p])
