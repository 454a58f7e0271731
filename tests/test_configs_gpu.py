"""The benchmark configurations 3, 4 and 5 at their stated workloads, on the
GPU, against the CPU oracle on sampled events.

  3  256 events on boxes drawn from 48x48x24 .. 128x128x64 (h = 0.5 km,
     layered 3.5-7.5 km/s + 3% GRF), forward + adjoint in one hipGraph, in
     FP64, FP32 and FP32 storage with FP64 solves
  4  joint Vp/Vs on 200x200x80 (h = 0.5 km): 100 surface stations x {P, S},
     500 events, Huber delta 0.1 s, picks with N(0, 0.05 s) noise; the
     objective and its gradient against the oracle composite on a station
     subset, and one L-BFGS step of the full problem
  5  400x400x160 (25.6M nodes, h = 0.25 km) in FP32 storage with FP64 local
     solves, 2000 source fields (the events, by reciprocity) and 150
     receivers: more fields than fit in memory, so the gradient recomputes
     them in chunks; FP64 oracle on 2 events (scripts/inversion.jl:46-62,
     the per-station eikonal3d calls)

Bars: FP64 travel times bit-exact with the oracle and the same iteration
counts; FP64 gradients within 1e-10 relative per solve; composite
objectives within 1e-12 (loss) and 1e-9 (gradient) as in
tests/test_objective_gpu.py; FP32 errors against FP64 below the bounds
written in each test.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from oracle import objective as OO  # noqa: E402
from oracle import oracle  # noqa: E402
from adtomo_hip import synth  # noqa: E402

GRAD_RTOL = 1e-10


def _torch():
    import torch
    return torch


def dev(a, dtype=np.float64):
    return _torch().from_numpy(np.ascontiguousarray(np.asarray(a, dtype=dtype))).to("cuda")


def host(t):
    return t.detach().cpu().numpy()


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def bits_equal(a, b):
    return np.array_equal(np.asarray(a).view(np.uint64 if a.dtype == np.float64 else np.uint32),
                          np.asarray(b).view(np.uint64 if b.dtype == np.float64 else np.uint32))


def _threads():
    return max(1, min(16, os.cpu_count() or 1))


# ------------------------------------------------------------ config 3 ---

def test_config3_uneven_boxes_256_events():
    import adtomo_hip as A
    torch = _torch()
    nev, h, tol = 256, 0.5, 1e-3
    fs, u0s, gs, srcs = synth.uneven_box_set(nev, h)
    sizes = np.array([f.size for f in fs])
    r = np.random.default_rng(33)
    picks = sorted({int(sizes.argmin()), int(sizes.argmax())} |
                   {int(x) for x in r.choice(nev, 2, replace=False)})
    print(f"config 3: {nev} boxes, {sizes.sum() / 1e6:.1f}M nodes, oracle on events {picks}",
          flush=True)
    res = {}
    for name, dt, mix in (("f64", torch.float64, False), ("f32", torch.float32, False),
                          ("mixed", torch.float32, True)):
        npd = np.float64 if dt == torch.float64 else np.float32
        G = A.EikonalGraph([dev(a, npd) for a in u0s], [dev(a, npd) for a in fs], [h] * nev,
                           [tol] * nev, grad_us=[dev(a, npd) for a in gs], compute_f64=mix)
        G.launch()
        torch.cuda.synchronize()
        infos = G.infos()
        assert all(i["adj_converged"] and i["fault"] == 0 for i in infos), name
        if name != "f32":  # FP32 arithmetic may reach the iteration cap
            assert all(i["converged"] for i in infos), name
        for e in range(nev):  # 0 <= u <= u0, 0 at the source
            u = G.us[e]
            assert bool((u >= 0).all()) and bool((u <= G.u0s[e]).all()), (name, e)
            assert float(u[srcs[e]]) == 0.0, (name, e)
        res[name] = G
        print(f"  {name}: K mean {np.mean([i['iterations'] for i in infos]):.2f}", flush=True)
    G64 = res["f64"]
    info64 = G64.infos()
    for e in picks:
        u_ref, it, _ = oracle.eikonal3d(u0s[e], fs[e], h, tol)
        gu0_ref, gf_ref, _ = oracle.eikonal3d_grad(gs[e], u_ref, u0s[e], fs[e], h)
        assert info64[e]["iterations"] == it, e
        assert bits_equal(host(G64.us[e]), u_ref), e
        assert bits_equal(host(G64.grad_u0s[e]), gu0_ref), e
        assert rel(host(G64.grad_fs[e]), gf_ref) <= GRAD_RTOL, e
        # the FP32 variants against the oracle's binary32 twins (bit for bit)
        u32, _, _ = oracle.eikonal3d(u0s[e], fs[e], h, tol, dtype=np.float32)
        assert bits_equal(host(res["f32"].us[e]), u32), e
        um32, _, _ = oracle.eikonal3d_m32(u0s[e], fs[e], h, tol)
        assert bits_equal(host(res["mixed"].us[e]), um32), e

    def errs(name):
        G = res[name]
        eu = eg = umax = gmax = ssu = ssg = 0.0
        for e in range(nev):
            du = G.us[e].double() - G64.us[e]
            dg = G.grad_fs[e].double() - G64.grad_fs[e]
            eu, eg = max(eu, float(du.abs().max())), max(eg, float(dg.abs().max()))
            umax = max(umax, float(G64.us[e].abs().max()))
            gmax = max(gmax, float(G64.grad_fs[e].abs().max()))
            ssu += float((du * du).sum())
            ssg += float((dg * dg).sum())
        n = float(sizes.sum())
        return eu / umax, (ssu / n) ** 0.5, eg / gmax, (ssg / n) ** 0.5

    e32, em = errs("f32"), errs("mixed")
    print(f"  FP32 vs FP64: u max rel {e32[0]:.2e} rms {e32[1]:.2e} s, grad max rel {e32[2]:.2e}",
          flush=True)
    print(f"  FP32 storage + FP64 solves: u max rel {em[0]:.2e} rms {em[1]:.2e} s, grad max rel "
          f"{em[2]:.2e}", flush=True)
    # binary32 loses digits to cancellation in the 3D quadratic (travel
    # times large against f*h); FP64 solves on FP32 storage keep ~1 ulp
    assert e32[0] <= 5e-3 and e32[1] <= 5e-3 and e32[2] <= 5e-2
    assert em[0] <= 2e-6 and em[1] <= 2e-6 and em[2] <= 1e-4
    for G in res.values():
        G.close()


# ------------------------------------------------------------ config 4 ---

def _picks(vel, h, stations, events, scale=1.0):
    """Travel times (stations x events) on model ``vel`` from GPU solves."""
    import adtomo_hip as A
    from adtomo_hip import objective as O
    torch = _torch()
    v = dev(vel)
    f = O.slowness(v, 1.0)
    ne = len(events[0])
    zero = torch.zeros(ne, dtype=torch.float64, device="cuda")
    u0s = O.source_u0(stations, v, h, scale)
    r = A.eikonal3d_batched(u0s, [f] * len(stations), [h] * len(stations),
                            [1e-3] * len(stations), return_info=False)
    mis = O.misfit(r["u"], *events, [zero] * len(stations), [zero] * len(stations), grad=False)
    return np.stack([host(t) for t in mis["tcal"]])


def _oracle_composite(params, vel0, h, stations, ev, tobs, w, sh, sv, lam, delta, phases, vpvs0,
                      us_given=None):
    """Loss and gradient of the inversion objective from the oracles (the
    stations' forwards on host threads, exact adjoints).  ``us_given``: the
    oracle forward fields of the P phase when already computed."""
    shape = vel0.shape
    N = vel0.size
    joint = "S" in phases
    x = params[:N].reshape(shape)
    fvar, fp = OO.model(x, vel0)
    vpvs = params[N] if joint else None
    fph = {"P": fp}
    if joint:
        fph["S"] = OO.slowness(fvar, vpvs)
    loss = 0.0
    gf = {p: np.zeros(shape) for p in phases}
    for ph in phases:
        scale = 1.0 if ph == "P" else vpvs0
        u0s = [OO.source_u0(shape, *st, vel0, h, scale) for st in stations]
        if ph == "P" and us_given is not None:
            us = us_given
        else:
            us, _, _, _ = oracle.eikonal3d_batch(u0s, [fph[ph]] * len(u0s), None,
                                                 [h] * len(u0s), 1e-3, _threads())
        for s in range(len(stations)):
            _, L, gu = OO.misfit(us[s], *ev, tobs[ph][s], w[ph][s], delta=delta)
            loss += L
            gf[ph] += oracle.eikonal3d_grad(gu, us[s], u0s[s], fph[ph], h)[1]
    reg, greg = OO.smooth_l1(fvar, sh, sv, lam)
    if joint:
        g_fvar_s, g_vpvs = OO.slowness_grad(fvar, vpvs, gf["S"])
        gx = OO.model_grad(x, fvar, gf["P"], greg + g_fvar_s)
        return loss + reg, np.concatenate([gx.reshape(-1), [g_vpvs]])
    gx = OO.model_grad(x, fvar, gf["P"], greg)
    return loss + reg, gx.reshape(-1)


def test_config4_joint_vp_vs_full_grid():
    from adtomo_hip.inversion import TravelTimeInversion
    shape, h = (200, 200, 80), 0.5
    vp = synth.layered_grf_model(shape, h, 5.0, 7.0, 0.04, 5.0, seed=41)
    vs = vp / 1.73 * (1.0 + synth.gaussian_random_field(shape, 5.0 / h, 0.02, seed=42))
    vel0 = synth.layered_model(shape, 5.0, 7.0)
    r = synth.rng(43)
    stations = [(float(r.uniform(1, shape[0])), float(r.uniform(1, shape[1])), 1.0)
                for _ in range(100)]
    events = tuple(r.uniform(1, s, 500) for s in shape)
    tp = _picks(vp, h, stations, events) + r.normal(0, 0.05, (100, 500))
    ts = _picks(vs, h, stations, events) + r.normal(0, 0.05, (100, 500))
    tp[3, 7] = ts[5, 11] = -1.0  # missing picks
    w = r.uniform(0.5, 1.5, (100, 500))
    tobs, wts = {"P": tp, "S": ts}, {"P": w, "S": w}
    kw = dict(smooth=(3, 3), lam=0.1, phases=("P", "S"), vpvs=1.7, huber_delta=0.1)
    # the objective and gradient of a station subset against the oracles
    sub = [0, 57]
    sub_t = {p: tobs[p][sub] for p in tobs}
    sub_w = {p: wts[p][sub] for p in wts}
    inv = TravelTimeInversion(vel0, h, [stations[s] for s in sub], events, sub_t, sub_w, **kw)
    try:
        p1 = inv.initial_params()
        p1[:-1] = synth.rng(44).normal(scale=0.2, size=vel0.size)
        p1[-1] = 1.74
        L, g = inv.objective(p1)
    finally:
        inv.close()
    RL, Rg = _oracle_composite(p1, vel0, h, [stations[s] for s in sub], events, sub_t, sub_w,
                               3, 3, 0.1, 0.1, ("P", "S"), 1.7)
    print(f"config 4 subset: loss {L:.6f} (oracle {RL:.6f}), grad rel {rel(g[:-1], Rg[:-1]):.2e}",
          flush=True)
    assert abs(L - RL) <= 1e-12 * abs(RL)
    assert rel(g[:-1], Rg[:-1]) <= 1e-9
    assert abs(g[-1] - Rg[-1]) <= 1e-9 * abs(Rg[-1])
    # one L-BFGS step of the whole problem lowers the loss
    inv = TravelTimeInversion(vel0, h, stations, events, tobs, wts, **kw)
    try:
        x1, hist = inv.run(iterations=1)
        res = inv.result
    finally:
        inv.close()
    print(f"config 4 full: one L-BFGS step {hist[0]:.4f} -> {hist[-1]:.4f} in {res.f_calls} "
          f"values / {res.g_calls} gradients, vp/vs {x1[-1]:.4f}", flush=True)
    assert res.iterations == 1 and res.g_calls == 2 and hist[-1] < hist[0]


# ------------------------------------------------------------ config 5 ---

def test_config5_large_grid_chunked_fp32_storage():
    from adtomo_hip.inversion import TravelTimeInversion
    torch = _torch()
    shape, h = (400, 400, 160), 0.25
    vel0 = synth.layered_model(shape, 3.5, 7.5)
    r = synth.rng(52)
    nsrc, nrec = 2000, 150
    sources = [tuple(float(r.uniform(1, s)) for s in shape) for _ in range(nsrc)]
    recv = (r.uniform(1, shape[0], nrec), r.uniform(1, shape[1], nrec), np.ones(nrec))
    # straight-ray picks on the mean velocity, with noise: the objective is
    # what is exercised here, not the recovery
    src = np.array(sources)
    dist = np.sqrt(sum((src[:, a:a + 1] - recv[a][None, :]) ** 2 for a in range(3))) * h
    tobs = dist / 5.5 + r.normal(0, 0.05, (nsrc, nrec))
    w = np.ones((nsrc, nrec))
    inv = TravelTimeInversion(vel0, h, sources, recv, tobs, w, smooth=(3, 3), lam=0.1,
                              dtype="float32", compute_f64=True)
    try:
        chunks = -(-nsrc // inv.chunk)
        print(f"config 5: {nsrc} source fields of {np.prod(shape) / 1e6:.1f}M nodes in chunks of "
              f"{inv.chunk} ({chunks} chunks)", flush=True)
        assert inv.chunk < nsrc  # the fields cannot all be held at once
        p0 = inv.initial_params()
        L, g = inv.objective(p0)
        assert np.isfinite(L) and np.all(np.isfinite(g)) and np.abs(g).max() > 0
        # the recomputed fields of the last chunk, two of its sources
        last0 = (chunks - 1) * inv.chunk
        pick = [last0, nsrc - 1]
        us_chunk = [host(inv.travel_times()[j - last0]).astype(np.float64) for j in pick]
    finally:
        inv.close()
        del inv
        torch.cuda.empty_cache()
    # the FP64 oracle on those two sources (model at x = 0)
    f64 = OO.model(np.zeros(shape), vel0)[1]
    u0s = [OO.source_u0(shape, *sources[j], vel0, h) for j in pick]
    print("  FP64 oracle on two sources ...", flush=True)
    refs, _, _, _ = oracle.eikonal3d_batch(u0s, [f64] * 2, None, [h] * 2, 1e-3, 2)
    for j, uc, ref in zip(pick, us_chunk, refs):
        d = uc - ref
        emax = float(np.abs(d).max() / np.abs(ref).max())
        erms = float(np.sqrt(np.mean(d * d)))
        print(f"  source {j}: travel time max rel {emax:.2e}, rms {erms:.2e} s", flush=True)
        assert emax <= 1e-5 and erms <= 1e-5
    # the same two sources' objective and gradient (FP32 storage, FP64
    # solves) against the FP64 oracle composite (no regulariser: at x = 0 on
    # a layered model its sign is rounding noise)
    two = TravelTimeInversion(vel0, h, [sources[j] for j in pick], recv, tobs[pick], w[pick],
                              smooth=(3, 3), lam=0.0, dtype="float32", compute_f64=True)
    try:
        L2, g2 = two.objective(two.initial_params())
    finally:
        two.close()
    RL, Rg = _oracle_composite(np.zeros(vel0.size), vel0, h, [sources[j] for j in pick], recv,
                               {"P": tobs[pick]}, {"P": w[pick]}, 3, 3, 0.0, 0.0, ("P",), 1.7,
                               us_given=refs)
    eg = rel(g2, Rg)
    print(f"  two sources: loss rel {abs(L2 - RL) / abs(RL):.2e}, gradient max rel {eg:.2e}",
          flush=True)
    assert abs(L2 - RL) <= 1e-4 * abs(RL) and eg <= 1e-4
