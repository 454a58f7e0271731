"""GPU parity of the device objective (adtomo_hip.objective) and of one
inversion step (adtomo_hip.inversion) against the CPU oracles.

Bar: initial travel times, interpolated travel times and dL/du bit-exact
in FP64 (same operation order, no contraction; each node's terms are summed
in receiver order on both sides); the misfit loss within 1e-13 (the device
sums it over threads in a tree); model / smoothing / chain rule
within 1e-13 relative (exp and the box sums differ in the last bits);
the whole objective gradient within 1e-9 relative (the Eikonal adjoint's
own bar is 1e-10).
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from oracle import objective as OO  # noqa: E402
from oracle import oracle  # noqa: E402


def _torch():
    import torch
    return torch


def dev(a, dtype=np.float64):
    return _torch().from_numpy(np.ascontiguousarray(np.asarray(a, dtype=dtype))).to("cuda")


def host(t):
    return t.detach().cpu().numpy()


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def _stations(r, shape, count):
    out = [(3.0, 4.0, 1.0), (2.5, 3.5, 1.5)]  # grid node, half cell
    for _ in range(count - 2):
        out.append(tuple(float(r.uniform(1, s)) for s in shape))
    return out


def test_source_u0_bit_exact():
    from adtomo_hip import objective as O
    r = np.random.default_rng(10)
    shape = (17, 13, 9)
    vel0 = r.uniform(3, 8, shape)
    st = _stations(r, shape, 6)
    for scale in (1.0, 1.73):
        us = O.source_u0(st, dev(vel0), 0.5, scale)
        for s, u in zip(st, us):
            ref = OO.source_u0(shape, *s, vel0, 0.5, scale)
            np.testing.assert_array_equal(host(u), ref)


def test_source_u0_f32():
    from adtomo_hip import objective as O
    r = np.random.default_rng(11)
    shape = (9, 8, 7)
    vel0 = r.uniform(3, 8, shape)
    st = _stations(r, shape, 4)
    us = O.source_u0(st, dev(vel0, np.float32), 0.5)
    for s, u in zip(st, us):
        ref = OO.source_u0(shape, *s, vel0, 0.5)
        np.testing.assert_allclose(host(u), ref, rtol=1e-6)


def _misfit_case(seed, shape, ns, ne):
    r = np.random.default_rng(seed)
    us = [r.uniform(0, 10, shape) for _ in range(ns)]
    rx, ry, rz = (r.uniform(1, s, ne) for s in shape)
    rx[:3] = np.floor(rx[:3])          # integer coordinate on x
    ry[3:5] = np.ceil(ry[3:5])         # and on y
    rx[5], ry[5], rz[5] = shape        # the far corner
    rx[6], ry[6], rz[6] = 1.0, 1.0, 1.0
    tobs = r.uniform(0, 10, (ns, ne))
    tobs[:, 7] = -1.0                  # missing picks
    tobs[0, ::5] = -1.0
    w = r.uniform(0.2, 2.0, (ns, ne))
    return us, (rx, ry, rz), tobs, w


@pytest.mark.parametrize("shape,ns,ne", [((11, 9, 7), 3, 40), ((32, 24, 16), 5, 700),
                                         ((4, 5, 3), 1, 9)])
def test_misfit_parity(shape, ns, ne):
    from adtomo_hip import objective as O
    us, (rx, ry, rz), tobs, w = _misfit_case(ns * 100 + ne, shape, ns, ne)
    out = O.misfit([dev(u) for u in us], rx, ry, rz, [dev(t) for t in tobs], [dev(x) for x in w])
    loss = host(out["loss"])
    for s in range(ns):
        tcal, L, g = OO.misfit(us[s], rx, ry, rz, tobs[s], w[s])
        np.testing.assert_array_equal(host(out["tcal"][s]), tcal)
        assert abs(loss[s] - L) <= 1e-13 * abs(L)
        np.testing.assert_array_equal(host(out["grad_u"][s]), g)


def test_misfit_f32():
    from adtomo_hip import objective as O
    us, (rx, ry, rz), tobs, w = _misfit_case(5, (10, 8, 6), 2, 30)
    out = O.misfit([dev(u, np.float32) for u in us], rx, ry, rz,
                   [dev(t, np.float32) for t in tobs], [dev(x, np.float32) for x in w])
    for s in range(2):
        tcal, L, g = OO.misfit(us[s], rx, ry, rz, tobs[s], w[s])
        np.testing.assert_allclose(host(out["tcal"][s]), tcal, rtol=1e-5, atol=1e-5)
        assert abs(host(out["loss"])[s] - L) <= 1e-4 * abs(L)
        assert rel(host(out["grad_u"][s]), g) <= 1e-4


def test_misfit_rejects_out_of_grid_events():
    import adtomo_hip as A
    from adtomo_hip import objective as O
    u = dev(np.zeros((5, 5, 5)))
    t = dev(np.zeros(1))
    for bad in ((0.5, 2.0, 2.0), (2.0, 5.5, 2.0), (2.0, 2.0, float("nan"))):
        with pytest.raises(A.AdtError):
            O.misfit([u], [bad[0]], [bad[1]], [bad[2]], [t], [t])


def test_model_and_grad_parity():
    from adtomo_hip import objective as O
    r = np.random.default_rng(12)
    shape = (21, 17, 13)
    x = r.normal(scale=2.0, size=shape)
    vel0 = r.uniform(3, 8, shape)
    gf = r.normal(size=shape)
    gv = r.normal(size=shape)
    fvar, f = O.model(dev(x), dev(vel0))
    rfvar, rf = OO.model(x, vel0)
    assert rel(host(fvar), rfvar) <= 1e-15
    assert rel(host(f), rf) <= 1e-15
    gx = O.model_grad(dev(x), fvar, dev(gf), dev(gv))
    assert rel(host(gx), OO.model_grad(x, rfvar, gf, gv)) <= 1e-13
    gx = O.model_grad(dev(x), fvar, grad_f=dev(gf))
    assert rel(host(gx), OO.model_grad(x, rfvar, gf, np.zeros(shape))) <= 1e-13


@pytest.mark.parametrize("sh,sv", [(3, 3), (5, 3), (1, 1), (7, 5)])
def test_smooth_l1_parity(sh, sv):
    from adtomo_hip import objective as O
    r = np.random.default_rng(13 + sh)
    shape = (19, 15, 11)
    fvar = r.uniform(3, 8, shape)
    val, g = O.smooth_l1(dev(fvar), sh, sv, 0.3)
    rval, rg = OO.smooth_l1(fvar, sh, sv, 0.3)
    if sh == 1 and sv == 1:
        assert host(val)[0] == 0.0 and not host(g).any()
        return
    assert abs(host(val)[0] - rval) <= 1e-13 * abs(rval)
    assert rel(host(g), rg) <= 1e-13
    val2, g2 = O.smooth_l1(dev(fvar), sh, sv, 0.3, grad=False)
    assert g2 is None and host(val2)[0] == host(val)[0]


def test_smooth_rejects_even_or_oversized_windows():
    import adtomo_hip as A
    from adtomo_hip import objective as O
    f = dev(np.ones((4, 4, 3)))
    for sh, sv in ((2, 3), (3, 2), (5, 3), (3, 5)):
        with pytest.raises(A.AdtError):
            O.smooth_l1(f, sh, sv, 1.0)


def test_sum_arrays_bit_exact():
    from adtomo_hip import objective as O
    r = np.random.default_rng(14)
    arrs = [r.normal(size=(9, 10, 11)) for _ in range(5)]
    out = host(O.sum_arrays([dev(a) for a in arrs]))
    ref = arrs[0].copy()
    for a in arrs[1:]:
        ref = ref + a
    np.testing.assert_array_equal(out, ref)


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-12), (np.float32, 1e-4)])
def test_misfit_huber_parity(dtype, tol):
    from adtomo_hip import objective as O
    us, (rx, ry, rz), tobs, w = _misfit_case(31, (12, 10, 8), 3, 60)
    r = np.random.default_rng(32)
    for s in range(3):  # residuals on both sides of delta
        t = np.array([OO.interpolate(us[s], *p) for p in zip(rx, ry, rz)])
        keep = tobs[s] == -1
        tobs[s] = t + r.choice([-1, 1], len(t)) * r.uniform(0, 0.3, len(t))
        tobs[s][keep] = -1.0
    out = O.misfit([dev(u, dtype) for u in us], rx, ry, rz, [dev(t, dtype) for t in tobs],
                   [dev(x, dtype) for x in w], huber_delta=0.1)
    for s in range(3):
        _, L, g = OO.misfit(us[s], rx, ry, rz, tobs[s], w[s], delta=0.1)
        assert abs(host(out["loss"])[s] - L) <= tol * abs(L)
        if dtype == np.float64:
            np.testing.assert_array_equal(host(out["grad_u"][s]), g)
        else:
            assert rel(host(out["grad_u"][s]), g) <= tol


def test_slowness_and_grad_parity():
    from adtomo_hip import objective as O
    r = np.random.default_rng(33)
    fvar = r.uniform(3, 8, (13, 11, 9))
    gfs = r.normal(size=fvar.shape)
    base = r.normal(size=fvar.shape)
    fs = O.slowness(dev(fvar), 1.73)
    np.testing.assert_array_equal(host(fs), OO.slowness(fvar, 1.73))
    gfvar = dev(base)
    gs = O.slowness_grad(dev(fvar), 1.73, dev(gfs), grad_fvar=gfvar)
    rg, rs = OO.slowness_grad(fvar, 1.73, gfs)
    assert rel(host(gfvar), base + rg) <= 1e-14
    assert abs(host(gs)[0] - rs) <= 1e-12 * abs(rs)
    assert O.slowness_grad(dev(fvar), 1.73, dev(gfs), grad_scale=False) is None


# --------------------------------------------------- one inversion step ---

def _inversion_case(seed=20, shape=(20, 18, 12), ns=3, ne=25):
    r = np.random.default_rng(seed)
    m, n, l = shape
    vel0 = np.repeat(np.linspace(4.0, 7.0, l)[None, None, :], m * n, 0).reshape(shape)
    true = vel0 * (1 + 0.03 * np.sin(np.arange(m) / 3.0)[:, None, None])
    stations = [(float(r.uniform(2, m - 1)), float(r.uniform(2, n - 1)), 1.0) for _ in range(ns)]
    ev = tuple(r.uniform(2, s - 1, ne) for s in shape)
    h = 1.0
    tobs = np.empty((ns, ne))
    for s, st in enumerate(stations):
        u0 = OO.source_u0(shape, *st, true, h)
        u, _, _ = oracle.eikonal3d(u0, 1.0 / true, h, 1e-3)
        tobs[s] = [OO.interpolate(u, *p) for p in zip(*ev)]
    tobs[1, 4] = -1.0
    w = r.uniform(0.5, 1.5, (ns, ne))
    return vel0, h, stations, ev, tobs, w


def _oracle_objective(x, vel0, h, stations, ev, tobs, w, sh, sv, lam):
    shape = vel0.shape
    fvar, f = OO.model(x, vel0)
    loss = 0.0
    grad_f = np.zeros(shape)
    for s, st in enumerate(stations):
        u0 = OO.source_u0(shape, *st, vel0, h)
        u, _, _ = oracle.eikonal3d(u0, f, h, 1e-3)
        _, L, gu = OO.misfit(u, *ev, tobs[s], w[s])
        loss += L
        _, gf, _ = oracle.eikonal3d_grad(gu, u, u0, f, h)
        grad_f += gf
    reg, greg = OO.smooth_l1(fvar, sh, sv, lam)
    return loss + reg, OO.model_grad(x, fvar, grad_f, greg)


def test_inversion_objective_matches_oracle():
    from adtomo_hip.inversion import TravelTimeInversion
    vel0, h, stations, ev, tobs, w = _inversion_case()
    inv = TravelTimeInversion(vel0, h, stations, ev, tobs, w, smooth=(3, 3), lam=0.05)
    try:
        r = np.random.default_rng(21)
        for x in (np.zeros(vel0.shape), r.normal(scale=0.2, size=vel0.shape)):
            L, g = inv.objective(x.reshape(-1))
            RL, Rg = _oracle_objective(x, vel0, h, stations, ev, tobs, w, 3, 3, 0.05)
            assert abs(L - RL) <= 1e-12 * abs(RL)
            assert rel(g, Rg.reshape(-1)) <= 1e-9
    finally:
        inv.close()


@pytest.mark.parametrize("chunk", [None, 2])
def test_inversion_objective_fp32_storage_fp64_solves(chunk):
    # the precision of large grids: FP32 arrays with every forward local
    # solve in FP64 (one launch per evaluation, or chunks of 2 stations),
    # against the FP64 oracle composite; the bound is binary32 storage
    # (about 6e-8 relative per value) carried through the sums
    from adtomo_hip.inversion import TravelTimeInversion
    vel0, h, stations, ev, tobs, w = _inversion_case(seed=60)
    inv = TravelTimeInversion(vel0, h, stations, ev, tobs, w, smooth=(3, 3), lam=0.05,
                              dtype="float32", compute_f64=True, chunk=chunk)
    try:
        x = np.random.default_rng(61).normal(scale=0.2, size=vel0.shape)
        L, g = inv.objective(x.reshape(-1))
        RL, Rg = _oracle_objective(x, vel0, h, stations, ev, tobs, w, 3, 3, 0.05)
        assert abs(L - RL) <= 1e-5 * abs(RL), (L, RL)
        assert rel(g, Rg.reshape(-1)) <= 1e-3, rel(g, Rg.reshape(-1))
    finally:
        inv.close()


def test_inversion_run_decreases_loss_and_checkpoints(tmp_path):
    from adtomo_hip.inversion import TravelTimeInversion
    vel0, h, stations, ev, tobs, w = _inversion_case(seed=22)
    inv = TravelTimeInversion(vel0, h, stations, ev, tobs, w, smooth=(3, 3), lam=0.0)
    try:
        x, hist = inv.run(iterations=4, steps=2, checkpoint_dir=str(tmp_path))
    finally:
        inv.close()
    assert hist[-1] < hist[0]
    files = sorted(os.listdir(tmp_path))
    assert files and all(f.startswith("iter_") and f.endswith(".npz") for f in files)
    with np.load(tmp_path / files[0]) as z:
        assert z["x"].shape == vel0.shape


def _joint_case(seed=40, shape=(18, 16, 12), ns=5, ne=20):
    r = np.random.default_rng(seed)
    vel0, h, stations, ev, tobs_p, w_p = _inversion_case(seed, shape, ns, ne)
    true_s = vel0 / 1.8
    tobs_s = np.empty_like(tobs_p)
    for s, st in enumerate(stations):
        u0 = OO.source_u0(shape, *st, vel0, h, 1.7)
        u, _, _ = oracle.eikonal3d(u0, 1.0 / true_s, h, 1e-3)
        tobs_s[s] = [OO.interpolate(u, *p) for p in zip(*ev)]
    tobs_s += r.normal(scale=0.05, size=tobs_s.shape)
    tobs_s[2, 3] = -1.0
    w_s = r.uniform(0.5, 1.5, tobs_s.shape)
    return vel0, h, stations, ev, {"P": tobs_p, "S": tobs_s}, {"P": w_p, "S": w_s}


def _oracle_joint(params, vel0, h, stations, ev, tobs, w, sh, sv, lam, delta):
    shape = vel0.shape
    N = vel0.size
    x, vpvs = params[:N].reshape(shape), params[N]
    fvar, fp = OO.model(x, vel0)
    fs = OO.slowness(fvar, vpvs)
    loss = 0.0
    gf = {"P": np.zeros(shape), "S": np.zeros(shape)}
    for ph, f, scale in (("P", fp, 1.0), ("S", fs, 1.7)):
        for s, st in enumerate(stations):
            u0 = OO.source_u0(shape, *st, vel0, h, scale)
            u, _, _ = oracle.eikonal3d(u0, f, h, 1e-3)
            _, L, gu = OO.misfit(u, *ev, tobs[ph][s], w[ph][s], delta=delta)
            loss += L
            gf[ph] += oracle.eikonal3d_grad(gu, u, u0, f, h)[1]
    reg, greg = OO.smooth_l1(fvar, sh, sv, lam)
    g_fvar_s, g_vpvs = OO.slowness_grad(fvar, vpvs, gf["S"])
    gx = OO.model_grad(x, fvar, gf["P"], greg + g_fvar_s)
    return loss + reg, np.concatenate([gx.reshape(-1), [g_vpvs]])


@pytest.mark.parametrize("chunk", [None, 2])
def test_joint_inversion_objective_matches_oracle(chunk):
    from adtomo_hip.inversion import TravelTimeInversion
    vel0, h, stations, ev, tobs, w = _joint_case()
    inv = TravelTimeInversion(vel0, h, stations, ev, tobs, w, smooth=(3, 3), lam=0.02,
                              phases=("P", "S"), vpvs=1.7, huber_delta=0.1, chunk=chunk)
    try:
        r = np.random.default_rng(41)
        p0 = inv.initial_params()
        p1 = p0.copy()
        p1[:-1] = r.normal(scale=0.2, size=vel0.size)
        p1[-1] = 1.76
        for params in (p0, p1):
            L, g = inv.objective(params)
            RL, Rg = _oracle_joint(params, vel0, h, stations, ev, tobs, w, 3, 3, 0.02, 0.1)
            assert abs(L - RL) <= 1e-12 * abs(RL)
            assert rel(g[:-1], Rg[:-1]) <= 1e-9
            assert abs(g[-1] - Rg[-1]) <= 1e-9 * abs(Rg[-1])
    finally:
        inv.close()


# --------------------------------------- inversion step in one launch ---

@pytest.mark.parametrize("fuse,dtype", [(1, "float64"), (0, "float64"), (1, "float32")])
def test_inversion_step_matches_separate_launches(fuse, dtype):
    # forward, misfit and adjoint of each station in one launch (graph and
    # single call; fused kernel or three kernels) against the forward graph,
    # the misfit call and the adjoint graph on the same inputs: travel times,
    # losses, dL/du and dL/df bit for bit, and the same again on a second
    # launch of the step graph; float32 runs FP32 storage with FP64 solves
    import adtomo_hip as A
    from adtomo_hip import objective as O
    torch = _torch()
    vel0, h, stations, ev, tobs, w = _inversion_case(seed=50, ns=4, ne=30)
    n = len(stations)
    npd = np.float64 if dtype == "float64" else np.float32
    mix = dtype == "float32"
    base = A.get_tune(dtype)
    graphs = []
    try:
        # the fused kernel (forward, misfit, adjoint in one workgroup) runs
        # with the node-level adjoint
        A.set_tune(dtype, fuse_max_n=(1 << 30) if fuse else 0, adj_algo=1 if fuse else 0)
        f = dev(1.0 / vel0, npd)
        u0s = O.source_u0(stations, dev(vel0, npd), h)
        tb = [dev(tobs[s], npd) for s in range(n)]
        wb = [dev(w[s], npd) for s in range(n)]
        args = (u0s, [f] * n, [h] * n, [1e-3] * n)
        step = O.InversionStepGraph(*args, *ev, tb, wb, huber_delta=0.05, compute_f64=mix)
        graphs.append(step)
        step.launch()
        torch.cuda.synchronize()
        first = [host(x).copy() for x in step.grad_us + step.grad_fs]
        step.launch()
        torch.cuda.synchronize()
        for a, b in zip(first, [host(x) for x in step.grad_us + step.grad_fs]):
            assert np.array_equal(a, b)
        one = O.inversion_step(*args, *ev, tb, wb, huber_delta=0.05, compute_f64=mix)
        fwd = A.EikonalGraph(*args, compute_f64=mix)
        graphs.append(fwd)
        fwd.launch()
        gus = [torch.empty_like(f) for _ in range(n)]
        mis = O.misfit(fwd.us, *ev, tb, wb, tcal=False, grad_out=gus, huber_delta=0.05)
        adj = A.EikonalGraph(*args, grad_us=gus, forward=False, us=fwd.us)
        graphs.append(adj)
        adj.launch()
        torch.cuda.synchronize()
        for res_u, res_gu, res_gf, res_loss in (
                (step.us, step.grad_us, step.grad_fs, step.losses[:n]),
                (one["u"], one["grad_u"], one["grad_f"], one["loss"])):
            for e in range(n):
                assert np.array_equal(host(res_u[e]), host(fwd.us[e])), e
                assert np.array_equal(host(res_gu[e]), host(gus[e])), e
                assert np.array_equal(host(res_gf[e]), host(adj.grad_fs[e])), e
            assert np.array_equal(host(res_loss), host(mis["loss"]))
        assert all(i["converged"] and i["adj_converged"] for i in step.infos())
        assert all(i["converged"] and i["adj_converged"] for i in one["info"])
    finally:
        A.set_tune(dtype, fuse_max_n=base["fuse_max_n"], adj_algo=base["adj_algo"])
        for g in graphs:
            g.close()
