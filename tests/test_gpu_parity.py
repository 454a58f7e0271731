"""GPU parity of the HIP path against the CPU oracle and the golden fixtures.

Bar: FP64 travel times bit-exact (and the same iteration count K); FP64
gradients within 1e-10 relative (max-norm) of the oracle; FP32 bit-exact
against the oracle's binary32 twin, with its error against FP64 reported.
All calls go through the C ABI (adtomo_hip -> libadtomo_hip.so).
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from oracle import oracle  # noqa: E402
from adtomo_hip import synth  # noqa: E402

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
GRAD_RTOL = 1e-10


def _torch():
    import torch
    return torch


def dev(a, dtype=None):
    torch = _torch()
    a = np.asarray(a)
    t = torch.from_numpy(np.ascontiguousarray(a if dtype is None else a.astype(dtype)))
    return t.to("cuda")


def host(t):
    return t.detach().cpu().numpy()


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def bits_equal(a, b):
    """Bit-for-bit equality; any NaN equals any NaN (payloads differ between
    x86 and CDNA and carry no meaning for the op)."""
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    assert a.shape == b.shape and a.dtype == b.dtype
    w = np.uint64 if a.dtype == np.float64 else np.uint32
    na, nb = np.isnan(a), np.isnan(b)
    return np.array_equal(na, nb) and np.array_equal(a[~na].view(w), b[~nb].view(w))


# ---------------------------------------------------------- arithmetic ---

def test_device_arith_is_ieee():
    import adtomo_hip as A
    r = synth.rng(1)
    x = np.concatenate([
        r.uniform(0, 1e4, 200000), r.uniform(0, 1e-6, 20000), np.exp(r.uniform(-700, 700, 20000)),
        np.array([0.0, 1.0, 2.0, 3.0, 4.0, 1e-310, 5e-324, 1e308, np.inf]),
        np.nextafter(np.arange(1, 2000, dtype=np.float64) ** 2, 0),
        np.nextafter(np.arange(1, 2000, dtype=np.float64) ** 2, np.inf)])
    x = np.concatenate([x, -x[:1000]])
    # the range edges of the fast square root and division by 3, signed
    # zeros, NaN, and random bit patterns over every exponent
    edges = np.array([2.0 ** -767, 2.0 ** -1000, 2.0 ** 1000, 2.0 ** 1023, 1.7976931348623157e308,
                      -0.0, np.nan, -np.inf, 3.0 * 2.0 ** -1000, 3.0 * 2.0 ** 999])
    edges = np.concatenate([edges, np.nextafter(edges, np.inf), np.nextafter(edges, -np.inf)])
    bits = r.integers(0, 2 ** 63, 400000, dtype=np.int64).view(np.float64)
    x = np.concatenate([x, edges, -edges, bits, -bits])
    # the division by 3 of the local solve is IEEE's except that -0 gives +0
    div3 = lambda v: np.where(v == 0, np.zeros_like(v), v / 3.0)  # noqa: E731
    for op, ref in ((0, np.sqrt), (7, np.sqrt), (1, div3), (2, lambda v: v / 2.0)):
        with np.errstate(invalid="ignore"):
            want = ref(x)
        got = host(A.selftest_arith(op, dev(x)))
        assert bits_equal(got, want), op
    x = x[np.isfinite(x)]
    y = r.uniform(0.1, 100, x.size)
    got = host(A.selftest_arith(3, dev(x), dev(y)))
    assert bits_equal(got, x / y)
    with np.errstate(invalid="ignore", over="ignore"):
        x32 = np.concatenate([x.astype(np.float32),
                              r.integers(0, 2 ** 31, 400000, dtype=np.int32).view(np.float32)])
        x32 = np.concatenate([x32, -x32])
        assert bits_equal(host(A.selftest_arith(0, dev(x32))), np.sqrt(x32))
        assert bits_equal(host(A.selftest_arith(1, dev(x32))),
                          np.where(x32 == 0, np.float32(0), x32 / np.float32(3)))
    # every binary32 subnormal and the lowest binades, both signs
    sub = np.arange(1, 1 << 25, dtype=np.uint32).view(np.float32)
    sub = np.concatenate([sub, -sub])
    assert bits_equal(host(A.selftest_arith(1, dev(sub))), sub / np.float32(3))


def test_selftest_copy():
    import adtomo_hip as A
    x = synth.rng(9).standard_normal(1 << 20)
    assert bits_equal(host(A.selftest_arith(6, dev(x))), x)


def test_device_local_solvers_match_numpy():
    import adtomo_hip as A
    import _hyperplane as hp
    r = synth.rng(2)
    n = 100000
    a = r.uniform(0, 10, (n, 3))
    a[: n // 4, 1] = a[: n // 4, 0]  # ties
    a[n // 4: n // 2, 2] = a[n // 4: n // 2, 0] + r.uniform(0, 1e-3, n // 4)
    fh = np.stack([r.uniform(0.1, 2, n), r.uniform(0.1, 3, n)], 1)
    want = hp.solution3d(a[:, 0], a[:, 1], a[:, 2], fh[:, 0], fh[:, 1])
    for op in (4, 8):  # 8: the fast square roots of the step loops
        got = host(A.selftest_arith(op, dev(a.reshape(-1)), dev(fh.reshape(-1))))
        assert bits_equal(got, want), op
    got2 = host(A.selftest_arith(5, dev(a[:, :2].reshape(-1)), dev(fh.reshape(-1))))
    want2 = hp.solution2d(a[:, 0], a[:, 1], fh[:, 0], fh[:, 1])
    assert bits_equal(got2, want2)
    # every combination of special neighbour values (signed zeros,
    # subnormals, values whose squares underflow or overflow, infinities,
    # NaN) with special f and h: the 3D candidate's divisions by 3 must not
    # change the result
    sv = np.array([0.0, -0.0, 5e-324, 1e-310, 1e-160, 3e-155, 1.0, 7.0, -1.0, 1e154, 1e200,
                   1.7976931348623157e308, np.inf, -np.inf, np.nan])
    fv = np.array([0.0, 5e-324, 1e-160, 0.5, 1.0, 1e155, 1e200, np.inf])
    g = np.stack(np.meshgrid(sv, sv, sv, fv, fv, indexing="ij"), -1).reshape(-1, 5)
    with np.errstate(all="ignore"):
        want3 = hp.solution3d(g[:, 0], g[:, 1], g[:, 2], g[:, 3], g[:, 4])
    for op in (4, 8):
        got3 = host(A.selftest_arith(op, dev(g[:, :3].reshape(-1)), dev(g[:, 3:].reshape(-1))))
        assert bits_equal(got3, want3), op
    # discriminants on both sides of the fast square root's range edge
    # 2^-767: a1 = a2 = 0 gives D2 = 2 (f h)^2 (selected with a3 = 1, while
    # D3 < 0), a1 = a2 = a3 = 0 gives D3 = 4 (f h)^2 / 3 (selected), up to
    # rounding, so f h around 2^-384 puts them at the edge
    fhv = np.ldexp(1.0, np.arange(-400, -370))
    fhv = np.concatenate([fhv, np.nextafter(fhv, 0), np.nextafter(fhv, 1)])
    z = np.zeros_like(fhv)
    e = np.concatenate([np.stack([z, z, a3 + z, fhv, z + 1.0], 1) for a3 in (0.0, 1.0, 1e300)])
    e = np.concatenate([e, e[:, [2, 0, 1, 3, 4]]])
    with np.errstate(all="ignore"):
        want4 = hp.solution3d(e[:, 0], e[:, 1], e[:, 2], e[:, 3], e[:, 4])
    got4 = host(A.selftest_arith(8, dev(e[:, :3].reshape(-1)), dev(e[:, 3:].reshape(-1))))
    assert bits_equal(got4, want4)


# ------------------------------------------------------------ golden ---

@pytest.mark.parametrize("name", ["gradtest3d", "small3d"])
def test_golden_3d(name):
    import adtomo_hip as A
    d = dict(np.load(os.path.join(GOLD, name + ".npz")))
    m, n, l = d["u0"].shape
    u = A.eikonal3d(dev(d["u0"]), dev(d["f"]), float(d["h"]), m, n, l, float(d["tol"]), False)
    assert A.ops.last_info["iterations"] == int(d["iters"])
    assert bits_equal(host(u), d["u"])
    gu0, gf = A.eikonal3d_grad(dev(d["grad_u"]), u, dev(d["u0"]), dev(d["f"]), float(d["h"]), m, n, l)
    assert bits_equal(host(gu0), d["grad_u0"])
    assert rel(host(gf), d["grad_f"]) <= GRAD_RTOL


@pytest.mark.parametrize("name", ["gradtest2d", "config1_2d"])
def test_golden_2d(name):
    import adtomo_hip as A
    d = dict(np.load(os.path.join(GOLD, name + ".npz")))
    u = A.eikonal(dev(d["f"]), int(d["srcx"]), int(d["srcy"]), float(d["h"]))
    assert A.ops.last_info["iterations"] == int(d["iters"])
    assert not A.ops.last_info["ambiguous"]
    assert bits_equal(host(u), d["u"])
    gf = A.eikonal_grad(dev(d["grad_u"]), u, dev(d["f"]), int(d["srcx"]), int(d["srcy"]), float(d["h"]))
    assert rel(host(gf), d["grad_f"]) <= GRAD_RTOL


# ------------------------------------------------------- 3D vs oracle ---

@pytest.mark.parametrize("shape,tol,seed", [
    ((2, 2, 2), 1e-3, 0), ((2, 5, 3), 1e-6, 1), ((7, 6, 5), 1e-3, 2), ((17, 9, 33), 1e-6, 3),
    ((31, 40, 12), 1e-3, 4)])
def test_forward3d_bit_exact_small(shape, tol, seed):
    import adtomo_hip as A
    r = synth.rng(seed)
    f = r.uniform(0.3, 3.0, size=shape)
    u0 = synth.point_source_u0(shape, tuple(int(r.integers(0, s)) for s in shape))
    u_ref, it, err = oracle.eikonal3d(u0, f, 0.5, tol)
    u = A.eikonal3d(dev(u0), dev(f), 0.5, *shape, tol, False)
    assert A.ops.last_info["iterations"] == it
    assert A.ops.last_info["err"] == err
    assert bits_equal(host(u), u_ref)


@pytest.mark.parametrize("shape", [
    (64, 64, 16), (48, 65, 23), (70, 40, 9), (20, 71, 33), (100, 72, 5), (130, 66, 17),
    (3, 129, 40), (129, 3, 2), (65, 65, 57), (66, 140, 64), (128, 128, 3), (20, 24, 70),
    (9, 66, 130)])
def test_forward3d_column_mappings(shape):
    # exercises the column mappings of the forward kernels: one band (shorter
    # non-k axis <= 64) or several, either non-k axis longer, k segments of
    # 1..8 nodes with idle waves (l <= 64, diagonal streaming) and l > 64
    # (k-streaming)
    import adtomo_hip as A
    r = synth.rng(sum(shape))
    f = 1.0 / synth.layered_grf_model(shape, 0.5, seed=sum(shape))
    u0 = synth.point_source_u0(shape, tuple(int(r.integers(0, s)) for s in shape))
    u_ref, it, err = oracle.eikonal3d(u0, f, 0.5, 1e-4)
    u = A.eikonal3d(dev(u0), dev(f), 0.5, *shape, 1e-4, False)
    print(shape, "K", it, flush=True)
    assert A.ops.last_info["iterations"] == it
    assert bits_equal(host(u), u_ref)


def _tiny_slowness_events():
    # discriminants below the fast square root's range (2^-767): a block of
    # slowness 1e-170 (travel times flat across it, D = 2 (f h)^2 or a
    # difference of nearly equal values: about half of all updates, counted
    # on the CPU), and an all-zero field (every update); k segments of 2 and
    # 8 nodes, chained and single sweeps
    out = []
    for shape, seed in (((20, 24, 16), 5), ((40, 48, 64), 6)):
        r = synth.rng(seed)
        f = r.uniform(0.3, 3.0, size=shape)
        f[:, : shape[1] // 2, :] = 1e-170
        out.append((synth.point_source_u0(shape, (3, 5, 7)), f))
    out.append((np.zeros((24, 20, 64)), np.zeros((24, 20, 64))))
    return out


def test_forward_redo_on_tiny_discriminants():
    # the step loops' exact redo (the wave recomputes a step with the IEEE
    # square root when a lane's discriminant lies outside the fast range)
    import adtomo_hip as A
    ev = _tiny_slowness_events()
    refs = [oracle.eikonal3d(u0, f, 0.5, 1e-4) for u0, f in ev]
    for (u0, f), (u_ref, it, err) in zip(ev, refs):
        u = A.eikonal3d(dev(u0), dev(f), 0.5, *u0.shape, 1e-4, False)
        assert A.ops.last_info["iterations"] == it, u0.shape
        assert bits_equal(host(u), u_ref), u0.shape
    res = A.eikonal3d_batched([dev(u0) for u0, _ in ev], [dev(f) for _, f in ev],
                              [0.5] * len(ev), [1e-4] * len(ev))
    for e, (u_ref, it, _) in enumerate(refs):
        assert res["info"][e]["iterations"] == it, e
        assert bits_equal(host(res["u"][e]), u_ref), e
    # FP32 storage, FP64 solves
    u32 = [dev(u0, np.float32) for u0, _ in ev]
    f32 = [dev(f, np.float32) for _, f in ev]
    res = A.eikonal3d_batched(u32, f32, [0.5] * len(ev), [1e-4] * len(ev), compute_f64=True)
    for e, (u0, f) in enumerate(ev):
        um, it, _ = oracle.eikonal3d_m32(u0, f, 0.5, 1e-4)
        assert res["info"][e]["iterations"] == it, e
        assert bits_equal(host(res["u"][e]), um), e


def _odd_value_events():
    # NaN and -0 in the fields: the step loops' single-instruction minima
    # order such values differently from the reference's comparisons, so a
    # step that lands, receives or produces one is redone exactly.  -0 at the
    # source and at two other nodes (they stay -0); a NaN block of slowness
    # (NaN travel times spread from it), a NaN and a +inf start value
    out = []
    r = synth.rng(31)
    shape = (20, 24, 16)
    f = r.uniform(0.3, 3.0, size=shape)
    u0 = synth.point_source_u0(shape, (3, 5, 7))
    u0[3, 5, 7] = u0[10, 10, 3] = u0[15, 2, 12] = -0.0
    out.append((u0, f))
    r = synth.rng(32)
    shape = (40, 48, 64)
    f = r.uniform(0.3, 3.0, size=shape)
    u0 = synth.point_source_u0(shape, (30, 9, 40))
    f[:6, :7, :9] = np.nan
    u0[20, 30, 5] = np.nan
    u0[35, 40, 60] = np.inf
    u0[30, 9, 40] = -0.0
    out.append((u0, f))
    # k in two chunks (l = 80: chunk-edge values come from memory) and
    # segments of 5 nodes (l = 40, the general step loop)
    for shape, seed, src, kb in (((20, 24, 80), 33, (4, 6, 70), 60),
                                 ((26, 18, 40), 34, (20, 3, 2), 30)):
        r = synth.rng(seed)
        f = r.uniform(0.3, 3.0, size=shape)
        u0 = synth.point_source_u0(shape, src)
        u0[src] = -0.0
        f[10:13, 9:12, kb:kb + 3] = np.nan
        u0[15, 4, kb + 5] = np.nan
        out.append((u0, f))
    return out


def _odd_ref(kind, u0, f):
    if kind == "f64":
        return oracle.eikonal3d(u0, f, 0.5, 1e-4)
    if kind == "m32":
        return oracle.eikonal3d_m32(u0, f, 0.5, 1e-4)
    return oracle.eikonal3d(u0, f, 0.5, 1e-4, dtype=np.float32)


def test_forward_nan_and_negative_zero_inputs():
    import adtomo_hip as A
    ev = _odd_value_events()
    r0, r1 = oracle.eikonal3d(*ev[0], 0.5, 1e-4)[0], oracle.eikonal3d(*ev[1], 0.5, 1e-4)[0]
    assert np.isnan(r1).sum() > 100 and np.signbit(r0[r0 == 0]).sum() == 3

    def check(kind, e, u, it):
        u_ref, it_ref, _ = _odd_ref(kind, *ev[e])
        assert it == it_ref, (kind, e)
        assert bits_equal(host(u), u_ref), (kind, e)
        assert np.array_equal(np.signbit(host(u)), np.signbit(u_ref)), (kind, e)

    for e, (u0, f) in enumerate(ev):
        u = A.eikonal3d(dev(u0), dev(f), 0.5, *u0.shape, 1e-4, False)
        check("f64", e, u, A.ops.last_info["iterations"])
    res = A.eikonal3d_batched([dev(u0) for u0, _ in ev], [dev(f) for _, f in ev],
                              [0.5] * len(ev), [1e-4] * len(ev))
    for e in range(len(ev)):
        check("f64", e, res["u"][e], res["info"][e]["iterations"])
    u32 = [dev(u0, np.float32) for u0, _ in ev]
    f32 = [dev(f, np.float32) for _, f in ev]
    # FP32 storage, FP64 solves
    res = A.eikonal3d_batched(u32, f32, [0.5] * len(ev), [1e-4] * len(ev), compute_f64=True)
    for e in range(len(ev)):
        check("m32", e, res["u"][e], res["info"][e]["iterations"])
    # FP32
    res = A.eikonal3d_batched(u32, f32, [0.5] * len(ev), [1e-4] * len(ev))
    for e in range(len(ev)):
        check("f32", e, res["u"][e], res["info"][e]["iterations"])


def test_forward_stop_rule_in_memory_order(tmp_path):
    # events 2 and 3 have NaN differences late in memory order: the
    # reference folds err = std::max(fabs(u - u_old), err) in memory order
    # (Eikonal3D.cpp:79-81), so a NaN difference restarts the maximum and
    # err is the maximum after the last one (NaN if that is the last node);
    # a maximum with NaN differences left out is larger and takes more
    # iterations (11 against 9 on event 2).  Iterations and the last err
    # as the oracle's.
    import adtomo_hip as A
    ev = _odd_value_events()
    for e in (2, 3):
        u0, f = ev[e]
        _, it, err = oracle.eikonal3d(u0, f, 0.5, 1e-4)
        A.eikonal3d(dev(u0), dev(f), 0.5, *u0.shape, 1e-4, False)
        assert A.ops.last_info["iterations"] == it, e
        assert bits_equal(np.array([A.ops.last_info["err"]]), np.array([err])), e
        # the plain per-hyperplane loop (lastnan_kernel + maxdiff_kernel)
        A.eikonal3d_plain(dev(u0), dev(f), 0.5, 1e-4)
        assert A.ops.last_info["iterations"] == it, e
        assert bits_equal(np.array([A.ops.last_info["err"]]), np.array([err])), e
    # the k-streaming forward (launch setting diag_max_l = 0)
    saved = str(tmp_path / "tune.json")
    A.save_tune(saved)
    try:
        A.set_tune("float64", diag_max_l=0)
        for e in (2, 3):
            u0, f = ev[e]
            u_ref, it, err = oracle.eikonal3d(u0, f, 0.5, 1e-4)
            u = A.eikonal3d(dev(u0), dev(f), 0.5, *u0.shape, 1e-4, False)
            assert A.ops.last_info["iterations"] == it, e
            assert bits_equal(np.array([A.ops.last_info["err"]]), np.array([err])), e
            assert bits_equal(host(u), u_ref), e
    finally:
        A.load_tune(saved)


@pytest.mark.parametrize("tol", [1e-3, 1e-6])
def test_config2_forward_adjoint(tol):
    # 64x64x32, h=1, 4->8 km/s layered x (1 +/- 5%) checkerboard (period 16),
    # seeded source, 50 top-surface receivers.
    import adtomo_hip as A
    shape = (64, 64, 32)
    f = 1.0 / synth.checkerboard_model(shape)
    src = synth.interior_source(shape, seed=22)
    u0 = synth.point_source_u0(shape, src)
    u_ref, it, err = oracle.eikonal3d(u0, f, 1.0, tol)
    u = A.eikonal3d(dev(u0), dev(f), 1.0, *shape, tol, False)
    print("config2 tol", tol, "K", it, "gpu K", A.ops.last_info["iterations"], flush=True)
    assert A.ops.last_info["iterations"] == it
    assert bits_equal(host(u), u_ref)
    g = synth.sparse_adjoint_source(shape, 50, seed=23, scale=1.0)
    gu0_ref, gf_ref, _ = oracle.eikonal3d_grad(g, u_ref, u0, f, 1.0)
    gu0, gf = A.eikonal3d_grad(dev(g), u, dev(u0), dev(f), 1.0, *shape)
    assert bits_equal(host(gu0), gu0_ref)
    e = rel(host(gf), gf_ref)
    print("config2 grad rel err", e, "adjoint sweeps", A.ops.last_info["adj_sweeps"], flush=True)
    assert e <= GRAD_RTOL


@pytest.mark.parametrize("shape", [(128, 128, 64), (129, 128, 64)])
def test_adjoint_at_queue_entry_packing_limit(shape):
    # 128x128x64 is exactly 2^20 nodes, the largest event whose queue entries
    # carry the node's code and feeder mask beside its index (the last index
    # fills all 20 index bits); 129x128x64 is past it and stores plain indices
    import adtomo_hip as A
    f = 1.0 / synth.layered_grf_model(shape, 0.5, seed=shape[0])
    u0 = synth.point_source_u0(shape, synth.interior_source(shape, seed=shape[0] + 1))
    u_ref, it, err = oracle.eikonal3d(u0, f, 0.5, 1e-3)
    u = A.eikonal3d(dev(u0), dev(f), 0.5, *shape, 1e-3, False)
    assert bits_equal(host(u), u_ref)
    g = synth.sparse_adjoint_source(shape, 60, seed=shape[0] + 2, scale=1.0)
    gu0_ref, gf_ref, _ = oracle.eikonal3d_grad(g, u_ref, u0, f, 0.5)
    gu0, gf = A.eikonal3d_grad(dev(g), u, dev(u0), dev(f), 0.5, *shape)
    assert A.ops.last_info["adj_converged"]
    assert bits_equal(host(gu0), gu0_ref)
    e = rel(host(gf), gf_ref)
    print(shape, "grad rel err", e, "levels", A.ops.last_info["adj_sweeps"], flush=True)
    assert e <= GRAD_RTOL


def test_headline_single_event_bit_exact():
    import adtomo_hip as A
    f, u0s, gs, srcs = synth.headline_event_set(1)
    u_ref, it, err = oracle.eikonal3d(u0s[0], f, 0.5, 1e-3)
    u = A.eikonal3d(dev(u0s[0]), dev(f), 0.5, 128, 128, 64, 1e-3, False)
    print("headline K", it, flush=True)
    assert A.ops.last_info["iterations"] == it
    assert bits_equal(host(u), u_ref)


def test_headline_batch_full_size():
    # the full headline batch, 256 events of 128x128x64, forward + adjoint in
    # one call with the default launch setting (fused kernel): every event
    # converged (forward and adjoint) with the op's invariants (0 <= u <= u0,
    # u = 0 at the source, grad_u0 = grad_u exactly on the set u == u0 and 0
    # elsewhere), and two sampled events against the oracle
    import adtomo_hip as A
    torch = _torch()
    nev = 256
    f, u0s, gs, srcs = synth.headline_event_set(nev)
    fd = dev(f)
    u0d = [dev(a) for a in u0s]
    gd = [dev(a) for a in gs]
    res = A.eikonal3d_batched(u0d, [fd] * nev, [0.5] * nev, [1e-3] * nev, grad_us=gd)
    infos = res["info"]
    assert all(i["converged"] and i["adj_converged"] and not i["fault"] for i in infos)
    for e in range(nev):
        u = res["u"][e]
        assert bool((u <= u0d[e]).all()) and bool((u >= 0).all())
        assert float(u[srcs[e]]) == 0.0
        z = torch.where(u == u0d[e], gd[e], torch.zeros_like(gd[e]))
        assert torch.equal(res["grad_u0"][e], z)
    for e in (0, 173):
        u_ref, it, _ = oracle.eikonal3d(u0s[e], f, 0.5, 1e-3)
        assert infos[e]["iterations"] == it and bits_equal(host(res["u"][e]), u_ref)
        _, gf_ref, _ = oracle.eikonal3d_grad(gs[e], u_ref, u0s[e], f, 0.5)
        assert rel(host(res["grad_f"][e]), gf_ref) <= GRAD_RTOL
    print("headline batch K:", sorted(set(i["iterations"] for i in infos)), flush=True)


def test_plain_path_matches_oracle():
    import adtomo_hip as A
    shape = (24, 20, 16)
    r = synth.rng(7)
    f = r.uniform(0.5, 2, size=shape)
    u0 = synth.point_source_u0(shape, (5, 11, 8))
    u_ref, it, err = oracle.eikonal3d(u0, f, 0.5, 1e-6)
    u = A.eikonal3d_plain(dev(u0), dev(f), 0.5, 1e-6)
    assert A.ops.last_info["iterations"] == it
    assert bits_equal(host(u), u_ref)
    g = r.normal(size=shape)
    gu0_ref, gf_ref, _ = oracle.eikonal3d_grad(g, u_ref, u0, f, 0.5)
    gu0, gf = A.eikonal3d_grad_plain(dev(g), u, dev(u0), dev(f), 0.5)
    assert bits_equal(host(gu0), gu0_ref)
    assert rel(host(gf), gf_ref) <= GRAD_RTOL


def test_autograd_rrule_twin():
    import adtomo_hip as A
    torch = _torch()
    shape = (16, 12, 10)
    r = synth.rng(8)
    fnp = r.uniform(0.5, 1.5, size=shape)
    u0np = synth.point_source_u0(shape, (8, 6, 5))
    g = r.normal(size=shape)
    f = dev(fnp).requires_grad_(True)
    u0 = dev(u0np).requires_grad_(True)
    u = A.eikonal3d(u0, f, 0.5, *shape, 1e-6, False)
    (u * dev(g)).sum().backward()
    u_ref, *_ = oracle.eikonal3d(u0np, fnp, 0.5, 1e-6)
    gu0_ref, gf_ref, _ = oracle.eikonal3d_grad(g, u_ref, u0np, fnp, 0.5)
    assert rel(host(f.grad), gf_ref) <= GRAD_RTOL
    assert bits_equal(host(u0.grad), gu0_ref)
    # 2D twin
    d = dict(np.load(os.path.join(GOLD, "gradtest2d.npz")))
    f2 = dev(d["f"]).requires_grad_(True)
    u2 = A.eikonal(f2, int(d["srcx"]), int(d["srcy"]), float(d["h"]))
    (u2 ** 2).sum().backward()
    assert rel(host(f2.grad), d["grad_f"]) <= GRAD_RTOL
    torch.cuda.synchronize()


def test_host_pointer_mode_numpy():
    import adtomo_hip as A
    d = dict(np.load(os.path.join(GOLD, "small3d.npz")))
    m, n, l = d["u0"].shape
    u = A.eikonal3d(d["u0"], d["f"], float(d["h"]), m, n, l, float(d["tol"]), False)
    assert isinstance(u, np.ndarray) and bits_equal(u, d["u"])
    gu0, gf = A.eikonal3d_grad(d["grad_u"], u, d["u0"], d["f"], float(d["h"]), m, n, l)
    assert rel(gf, d["grad_f"]) <= GRAD_RTOL


# ------------------------------------------------------ strided layouts ---

def _perm(a, dtype=None):
    """Device tensor with the Julia column-major layout: strides (1, m, m*n)."""
    return dev(np.ascontiguousarray(a.transpose(2, 1, 0)), dtype).permute(2, 1, 0)


def test_strided_3d_layouts_match_rowmajor():
    import adtomo_hip as A
    shape = (20, 24, 18)
    r = synth.rng(31)
    f = r.uniform(0.3, 3.0, size=shape)
    u0 = synth.point_source_u0(shape, (5, 7, 9))
    g = r.normal(size=shape)
    u_ref, it, err = oracle.eikonal3d(u0, f, 0.5, 1e-6)
    ud = A.eikonal3d(dev(u0), dev(f), 0.5, *shape, 1e-6, False)
    assert bits_equal(host(ud), u_ref)
    gu0d, gfd = (host(a) for a in A.eikonal3d_grad(dev(g), ud, dev(u0), dev(f), 0.5, *shape))
    # host arrays in the Julia column-major layout
    F = np.asfortranarray
    u = A.eikonal3d_strided(F(u0), F(f), 0.5, 1e-6)
    assert u.flags.f_contiguous and bits_equal(u, u_ref)
    assert A.ops.last_info["iterations"] == it
    gu0, gf = A.eikonal3d_grad_strided(F(g), F(u_ref), F(u0), F(f), 0.5)
    assert gf.flags.f_contiguous and bits_equal(gu0, gu0d) and bits_equal(gf, gfd)
    # device tensors with the same layout (permuted views)
    ut = A.eikonal3d_strided(_perm(u0), _perm(f), 0.5, 1e-6)
    assert ut.stride() == (1, shape[0], shape[0] * shape[1])
    assert bits_equal(host(ut), u_ref)
    gu0t, gft = A.eikonal3d_grad_strided(_perm(g), ut, _perm(u0), _perm(f), 0.5)
    assert bits_equal(host(gu0t), gu0d) and bits_equal(host(gft), gfd)


@pytest.mark.parametrize("host_ptrs", [False, True])
def test_strided_padded_layout_leaves_gaps(host_ptrs):
    # every array of the call is a (9, 11, 13) corner of a (12, 13, 18)
    # buffer: results land in the corner, the rest of the output buffer keeps
    # its values
    import ctypes
    import adtomo_hip as A
    from adtomo_hip import _lib as L
    torch = _torch()
    shape, big = (9, 11, 13), (12, 13, 18)
    r = synth.rng(32)
    f = r.uniform(0.5, 2.0, size=shape)
    u0 = synth.point_source_u0(shape, (4, 5, 6))
    g = r.normal(size=shape)
    u_ref, it, err = oracle.eikonal3d(u0, f, 0.5, 1e-6)
    gu0_ref, gf_ref, _ = oracle.eikonal3d_grad(g, u_ref, u0, f, 0.5)
    corner = (slice(0, shape[0]), slice(0, shape[1]), slice(0, shape[2]))

    def embed(a):
        b = np.full(big, -7.0)
        b[corner] = a
        return b

    bufs = [embed(u0), embed(f), np.full(big, -7.0), embed(g), embed(u_ref),
            np.full(big, -7.0), np.full(big, -7.0)]
    st = (ctypes.c_int64 * 3)(big[1] * big[2], big[2], 1)
    lib, info = L.lib(), L.AdtInfo()
    if host_ptrs:
        ptr = [b.ctypes.data for b in bufs]
        flags, stream = L.ADT_HOST_POINTERS, None
    else:
        tb = [dev(b) for b in bufs]
        ptr = [t.data_ptr() for t in tb]
        flags, stream = 0, A.ops._stream()
    rc = lib.adt_eikonal3d_fwd_strided_f64(ptr[0], ptr[1], 0.5, *shape, 1e-6, 0, st, ptr[2],
                                           ctypes.byref(info), flags, stream)
    L.check(rc, "eikonal3d strided")
    assert info.iterations == it
    rc = lib.adt_eikonal3d_bwd_strided_f64(ptr[3], ptr[4], ptr[0], ptr[1], 0.5, *shape, st,
                                           ptr[5], ptr[6], ctypes.byref(info), flags, stream)
    L.check(rc, "eikonal3d grad strided")
    if not host_ptrs:
        torch.cuda.synchronize()
        bufs = [host(t) for t in tb]
    gap = np.ones(big, bool)
    gap[corner] = False
    for k in (2, 5, 6):
        assert (bufs[k][gap] == -7.0).all()
    assert bits_equal(np.ascontiguousarray(bufs[2][corner]), u_ref)
    assert bits_equal(np.ascontiguousarray(bufs[5][corner]), gu0_ref)
    assert rel(bufs[6][corner], gf_ref) <= GRAD_RTOL


def test_strided_2d_column_major_matches_rowmajor():
    import adtomo_hip as A
    r = synth.rng(33)
    m, n = 30, 20
    f = r.uniform(0.5, 2.0, size=(n + 1, m + 1))
    srcx, srcy = 7, 5
    u_ref, it, err, conv = oracle.eikonal2d(f, srcx, srcy, 0.3)
    F = np.asfortranarray
    u = A.eikonal_strided(F(f), srcx, srcy, 0.3)
    assert u.flags.f_contiguous and bits_equal(u, u_ref)
    g = r.normal(size=f.shape)
    gf_row = A.eikonal_grad(g, u_ref, f, srcx, srcy, 0.3)
    gf = A.eikonal_grad_strided(F(g), F(u_ref), F(f), srcx, srcy, 0.3)
    assert bits_equal(gf, gf_row)
    # device: the transposed view of an (m+1, n+1) tensor
    ut = A.eikonal_strided(dev(np.ascontiguousarray(f.T)).t(), srcx, srcy, 0.3)
    assert ut.stride() == (1, n + 1) and bits_equal(host(ut), u_ref)


def test_strided_f32_matches_rowmajor_f32():
    import adtomo_hip as A
    shape = (16, 12, 10)
    r = synth.rng(34)
    f = r.uniform(0.5, 2.0, size=shape)
    u0 = synth.point_source_u0(shape, (3, 4, 5))
    g = r.normal(size=shape)
    u32 = A.eikonal3d(dev(u0, np.float32), dev(f, np.float32), 0.5, *shape, 1e-4, False)
    us = A.eikonal3d_strided(_perm(u0, np.float32), _perm(f, np.float32), 0.5, 1e-4)
    assert bits_equal(host(us), host(u32))
    ga, gb = A.eikonal3d_grad(dev(g, np.float32), u32, dev(u0, np.float32), dev(f, np.float32),
                              0.5, *shape)
    gc, gd = A.eikonal3d_grad_strided(_perm(g, np.float32), us, _perm(u0, np.float32),
                                      _perm(f, np.float32), 0.5)
    assert bits_equal(host(gc), host(ga)) and bits_equal(host(gd), host(gb))


# --------------------------------------------------------------- 2D ---

@pytest.mark.parametrize("m,n,seed", [(1, 1, 0), (12, 9, 1), (60, 30, 2), (100, 100, 3)])
def test_2d_vs_oracle(m, n, seed):
    import adtomo_hip as A
    r = synth.rng(seed)
    f = r.uniform(0.5, 2.0, size=(n + 1, m + 1))
    srcx, srcy = int(r.integers(1, m + 2)), int(r.integers(1, n + 2))
    u_ref, it, err, conv = oracle.eikonal2d(f, srcx, srcy, 0.3)
    u = A.eikonal(dev(f), srcx, srcy, 0.3)
    assert A.ops.last_info["iterations"] == it and A.ops.last_info["converged"] == conv
    assert bits_equal(host(u), u_ref)
    g = r.normal(size=f.shape)
    gf_ref, _, ns = oracle.eikonal2d_grad(g, u_ref, f, srcx, srcy, 0.3)
    gf = A.eikonal_grad(dev(g), u, dev(f), srcx, srcy, 0.3)
    assert A.ops.last_info["n_singular"] == ns
    assert rel(host(gf), gf_ref) <= GRAD_RTOL


# ---------------------------------------------------------- batched ---

def _uneven(nev, seed):
    fs, u0s, gs, srcs = synth.uneven_box_set(nev, seed=seed, lo=(8, 8, 4), hi=(40, 36, 24))
    return fs, u0s, gs


def test_batched_uneven_fp64_bit_exact():
    import adtomo_hip as A
    fs, u0s, gs = _uneven(12, 5)
    res = A.eikonal3d_batched([dev(a) for a in u0s], [dev(a) for a in fs], [0.5] * 12,
                              [1e-3] * 12, grad_us=[dev(a) for a in gs])
    for e in range(12):
        u_ref, it, err = oracle.eikonal3d(u0s[e], fs[e], 0.5, 1e-3)
        assert res["info"][e]["iterations"] == it, e
        assert bits_equal(host(res["u"][e]), u_ref), e
        gu0_ref, gf_ref, _ = oracle.eikonal3d_grad(gs[e], u_ref, u0s[e], fs[e], 0.5)
        assert bits_equal(host(res["grad_u0"][e]), gu0_ref), e
        assert rel(host(res["grad_f"][e]), gf_ref) <= GRAD_RTOL, e


def test_batched_mixed_k_extents_bit_exact():
    # one batch with events for both forward kernels (l <= 64 and l > 64)
    import adtomo_hip as A
    shapes = [(10, 12, 70), (20, 9, 30), (8, 8, 100), (12, 14, 64), (6, 7, 65)]
    fs, u0s = [], []
    for e, shp in enumerate(shapes):
        r = synth.rng(100 + e)
        fs.append(r.uniform(0.3, 3.0, size=shp))
        u0s.append(synth.point_source_u0(shp, tuple(int(r.integers(0, s)) for s in shp)))
    res = A.eikonal3d_batched([dev(a) for a in u0s], [dev(a) for a in fs], [0.5] * len(shapes),
                              [1e-4] * len(shapes))
    for e in range(len(shapes)):
        u_ref, it, err = oracle.eikonal3d(u0s[e], fs[e], 0.5, 1e-4)
        assert res["info"][e]["iterations"] == it, e
        assert bits_equal(host(res["u"][e]), u_ref), e


def test_batched_fp32_vs_oracle_f32_and_fp64_error():
    import adtomo_hip as A
    fs, u0s, gs = _uneven(6, 6)
    res = A.eikonal3d_batched([dev(a, np.float32) for a in u0s], [dev(a, np.float32) for a in fs],
                              [0.5] * 6, [1e-3] * 6, grad_us=[dev(a, np.float32) for a in gs])
    worst_u, worst_g = 0.0, 0.0
    for e in range(6):
        u32, it32, _ = oracle.eikonal3d(u0s[e], fs[e], 0.5, 1e-3, dtype=np.float32)
        assert res["info"][e]["iterations"] == it32
        assert bits_equal(host(res["u"][e]), u32), e
        u64, it, _ = oracle.eikonal3d(u0s[e], fs[e], 0.5, 1e-3)
        _, gf64, _ = oracle.eikonal3d_grad(gs[e], u64, u0s[e], fs[e], 0.5)
        worst_u = max(worst_u, rel(host(res["u"][e]).astype(np.float64), u64))
        worst_g = max(worst_g, rel(host(res["grad_f"][e]).astype(np.float64), gf64))
    print(f"FP32 vs FP64: max rel travel-time err {worst_u:.3e}, grad err {worst_g:.3e}", flush=True)
    # binary32 accumulates ~1e-7 per update along O(100)-node paths
    assert worst_u < 1e-4 and worst_g < 1e-2


def test_graph_replays_match_eager():
    import adtomo_hip as A
    torch = _torch()
    fs, u0s, gs = _uneven(5, 7)
    dfs = [dev(a) for a in fs]
    du0 = [dev(a) for a in u0s]
    dgs = [dev(a) for a in gs]
    eager = A.eikonal3d_batched(du0, dfs, [0.5] * 5, [1e-3] * 5, grad_us=dgs)
    G = A.EikonalGraph(du0, dfs, [0.5] * 5, [1e-3] * 5, grad_us=dgs)
    for rep in range(4):
        for t in G.us + G.grad_fs:
            t.fill_(-1.0)
        G.launch()
        torch.cuda.synchronize()
        x = torch.ones(1000, device="cuda").sum()  # eager work between replays
        torch.cuda.synchronize()
        assert float(x) == 1000.0
        for e in range(5):
            assert torch.equal(G.us[e], eager["u"][e]), (rep, e)
            assert torch.equal(G.grad_fs[e], eager["grad_f"][e]), (rep, e)
        infos = G.infos()
        assert [i["iterations"] for i in infos] == [i["iterations"] for i in eager["info"]]
    G.close()


# --------------------------------------------------------------- tuning ---

def _mixed_batch(seed=300):
    shapes = [(10, 12, 70), (20, 9, 30), (8, 8, 100), (12, 14, 64), (6, 7, 65), (9, 11, 7)]
    fs, u0s, gs = [], [], []
    for e, shp in enumerate(shapes):
        r = synth.rng(seed + e)
        fs.append(r.uniform(0.3, 3.0, size=shp))
        u0s.append(synth.point_source_u0(shp, tuple(int(r.integers(0, s)) for s in shp)))
        gs.append(synth.sparse_adjoint_source(shp, 6, seed + e))
    return fs, u0s, gs


@pytest.mark.parametrize("setting", [
    {"diag_max_l": 1 << 30, "adj_wg": 1024, "order": 1, "fuse_max_n": 1 << 30, "adj_algo": 1},
    {"diag_max_l": 1 << 30, "adj_wg": 512, "order": 0, "fuse_max_n": 0, "adj_algo": 1},
    {"diag_max_l": 64, "adj_wg": 1024, "order": 1, "fuse_max_n": 1 << 30, "adj_algo": 1},
    {"diag_max_l": 0, "adj_wg": 256, "order": 0, "fuse_max_n": 1 << 30, "adj_algo": 1},
    {"diag_max_l": 32, "adj_wg": 512, "order": 1, "fuse_max_n": 0, "adj_algo": 1},
    {"diag_max_l": 7, "adj_wg": 256, "order": 1, "fuse_max_n": 1 << 30, "adj_algo": 1},
    {"diag_max_l": 1 << 30, "adj_wg": 1024, "order": 1, "fuse_max_n": 1 << 30, "adj_algo": 0},
    {"diag_max_l": 0, "adj_wg": 256, "order": 0, "fuse_max_n": 0, "adj_algo": 0},
    {"diag_max_l": 32, "adj_wg": 512, "order": 1, "fuse_max_n": 0, "adj_algo": 0},
    # segment counts of the diagonal forward (k chunks of 8, 24, 40, 56 nodes)
    {"diag_max_l": 1 << 30, "adj_wg": 256, "order": 1, "fuse_max_n": 0, "adj_algo": 0,
     "fwd_segs": 1},
    {"diag_max_l": 1 << 30, "adj_wg": 512, "order": 0, "fuse_max_n": 1 << 30, "adj_algo": 1,
     "fwd_segs": 3},
    {"diag_max_l": 1 << 30, "adj_wg": 256, "order": 1, "fuse_max_n": 0, "adj_algo": 0,
     "fwd_segs": 5},
    {"diag_max_l": 64, "adj_wg": 1024, "order": 0, "fuse_max_n": 0, "adj_algo": 1,
     "fwd_segs": 7},
])
def test_every_tune_setting_matches_oracle(setting, tmp_path):
    import adtomo_hip as A
    saved = str(tmp_path / "tune.json")
    A.save_tune(saved)
    fs, u0s, gs = _mixed_batch()
    try:
        A.set_tune("float64", **setting)
        res = A.eikonal3d_batched([dev(a) for a in u0s], [dev(a) for a in fs], [0.5] * len(fs),
                                  [1e-4] * len(fs), grad_us=[dev(a) for a in gs])
        G = A.EikonalGraph([dev(a) for a in u0s], [dev(a) for a in fs], [0.5] * len(fs),
                           [1e-4] * len(fs), grad_us=[dev(a) for a in gs])
        G.launch()
        ginfo = G.infos()
        for e in range(len(fs)):
            u_ref, it, _ = oracle.eikonal3d(u0s[e], fs[e], 0.5, 1e-4)
            _, gf_ref, _ = oracle.eikonal3d_grad(gs[e], u_ref, u0s[e], fs[e], 0.5)
            for r, info in ((res, res["info"][e]), (G, ginfo[e])):
                us = r["u"] if r is res else r.us
                gfs = r["grad_f"] if r is res else r.grad_fs
                assert info["iterations"] == it, (setting, e)
                assert bits_equal(host(us[e]), u_ref), (setting, e)
                assert rel(host(gfs[e]), gf_ref) <= GRAD_RTOL, (setting, e)
        G.close()
    finally:
        A.load_tune(saved)


def test_class_settings_mix_in_one_graph(tmp_path):
    # events of several box-shape classes, each class with its own forward
    # kernel, segment count and adjoint traversal, in one call and one graph:
    # travel times bit for bit and gradients to the bar of the oracle
    import adtomo_hip as A
    saved = str(tmp_path / "tune.json")
    A.save_tune(saved)
    fs, u0s, gs = _mixed_batch(310)
    per_class = [{"diag_max_l": 0, "adj_algo": 1}, {"fwd_segs": 3, "adj_algo": 0},
                 {"fwd_segs": 8, "adj_algo": 1}, {"fwd_segs": 2, "adj_algo": 0},
                 {"diag_max_l": 1 << 30, "fwd_segs": 6, "adj_algo": 0}, {"fwd_segs": 1}]
    try:
        A.clear_tune_classes("float64")
        A.set_tune("float64", diag_max_l=1 << 30, fwd_segs=0, adj_algo=0, fuse_max_n=0)
        for f, cfg in zip(fs, per_class):
            A.set_tune_class(f.shape, "float64", **cfg)
        assert len(A.tune_classes("float64")) == len({(-(-f.shape[2] // 8), -(-min(f.shape[:2]) // 64),
                                                         int(np.log2(f.size))) for f in fs})
        res = A.eikonal3d_batched([dev(a) for a in u0s], [dev(a) for a in fs], [0.5] * len(fs),
                                  [1e-4] * len(fs), grad_us=[dev(a) for a in gs])
        G = A.EikonalGraph([dev(a) for a in u0s], [dev(a) for a in fs], [0.5] * len(fs),
                           [1e-4] * len(fs), grad_us=[dev(a) for a in gs])
        G.launch()
        G.launch()
        ginfo = G.infos()
        for e in range(len(fs)):
            u_ref, it, _ = oracle.eikonal3d(u0s[e], fs[e], 0.5, 1e-4)
            _, gf_ref, _ = oracle.eikonal3d_grad(gs[e], u_ref, u0s[e], fs[e], 0.5)
            for us, gfs, info in ((res["u"], res["grad_f"], res["info"][e]),
                                  (G.us, G.grad_fs, ginfo[e])):
                assert info["iterations"] == it, e
                assert bits_equal(host(us[e]), u_ref), e
                assert rel(host(gfs[e]), gf_ref) <= GRAD_RTOL, e
        G.close()
    finally:
        A.load_tune(saved)


@pytest.mark.parametrize("dtype,compute_f64", [("float64", False), ("float32", False),
                                                ("float32", True)])
def test_fused_forward_adjoint_matches_separate_kernels(dtype, compute_f64):
    # one kernel per event (forward, then adjoint) against the two kernels:
    # travel times, iteration counts and both gradients bit for bit
    import adtomo_hip as A
    torch = _torch()
    base = A.get_tune(dtype)
    fs, u0s, gs = _mixed_batch(600)
    npd = np.float64 if dtype == "float64" else np.float32
    args = ([dev(a, npd) for a in u0s], [dev(a, npd) for a in fs], [0.5] * len(fs),
            [1e-4] * len(fs))
    out = {}
    try:
        for fuse in (1, 0):
            # the fused kernel runs the node-level traversal
            A.set_tune(dtype, diag_max_l=1 << 30, fuse_max_n=(1 << 30) if fuse else 0, adj_algo=1)
            G = A.EikonalGraph(*args, grad_us=[dev(a, npd) for a in gs], compute_f64=compute_f64)
            G.launch()
            G.launch()
            torch.cuda.synchronize()
            out[fuse] = ([host(u) for u in G.us], [host(g) for g in G.grad_fs],
                         [host(g) for g in G.grad_u0s], [i["iterations"] for i in G.infos()])
            G.close()
    finally:
        A.set_tune(dtype, diag_max_l=base["diag_max_l"], fuse_max_n=base["fuse_max_n"],
                   adj_algo=base["adj_algo"])
    for k in range(3):
        for a, b in zip(out[1][k], out[0][k]):
            assert bits_equal(a, b), (dtype, compute_f64, k)
    assert out[1][3] == out[0][3]


@pytest.mark.parametrize("dtype", ["float64", "float32"])
def test_tile_adjoint_matches_node_adjoint_bitwise(dtype):
    # both traversals resolve every node with the same expression in the same
    # operand order, so they agree bit for bit (uneven boxes, partial tiles)
    import adtomo_hip as A
    base = A.get_tune(dtype)
    npd = np.float64 if dtype == "float64" else np.float32
    fs, u0s, gs = _mixed_batch(700)
    u = [dev(oracle.eikonal3d(u0s[e], fs[e], 0.5, 1e-4, dtype=npd)[0], npd) for e in range(len(fs))]
    out = {}
    try:
        for algo in (0, 1):
            A.set_tune(dtype, adj_algo=algo)
            r = A.eikonal3d_batched([dev(a, npd) for a in u0s], [dev(a, npd) for a in fs],
                                    [0.5] * len(fs), [1e-4] * len(fs),
                                    grad_us=[dev(a, npd) for a in gs], us=u, forward=False)
            assert all(i["adj_converged"] for i in r["info"]), (dtype, algo)
            out[algo] = ([host(x) for x in r["grad_f"]], [host(x) for x in r["grad_u0"]])
    finally:
        A.set_tune(dtype, adj_algo=base["adj_algo"])
    for k in range(2):
        for e, (a, b) in enumerate(zip(out[0][k], out[1][k])):
            assert bits_equal(a, b), (dtype, k, e)


def test_tile_adjoint_random_u0_large_z_set():
    # random u0 makes many nodes keep u == u0 (the Z set, lambda = 0) and cuts
    # the upwind graph into many pieces; against the oracle's exact solve
    import adtomo_hip as A
    shp = (21, 34, 19)
    r = synth.rng(811)
    f = r.uniform(0.5, 2.0, size=shp)
    u0 = r.uniform(0.0, 6.0, size=shp)
    g = r.normal(size=shp)
    u, _, _ = oracle.eikonal3d(u0, f, 0.5, 1e-6)
    gu0_ref, gf_ref, _ = oracle.eikonal3d_grad(g, u, u0, f, 0.5)
    nz = int(np.sum(u == u0))
    print(f"Z nodes {nz} of {u.size}", flush=True)
    gu0, gf = A.eikonal3d_grad(dev(g), dev(u), dev(u0), dev(f), 0.5, *shp)
    info = A.ops.last_info
    assert info["adj_converged"] == 1 and info["fault"] == 0
    assert rel(host(gf), gf_ref) <= GRAD_RTOL
    assert bits_equal(host(gu0), gu0_ref)


def test_tile_adjoint_large_tile_count_variant():
    # more than 8192 tiles of 8^3 (here 34 x 32 x 8 = 8704): the variant with
    # fewer waves and larger scheduling bitmaps
    import adtomo_hip as A
    shp = (272, 256, 64)
    h, tol = 0.5, 1e-3
    f = 1.0 / synth.layered_grf_model(shp, h, seed=21)
    u0 = synth.point_source_u0(shp, (200, 40, 30))
    g = synth.sparse_adjoint_source(shp, 80, seed=22)
    print(f"event {shp}: oracle", flush=True)
    u, it, _ = oracle.eikonal3d(u0, f, h, tol)
    _, gf_ref, _ = oracle.eikonal3d_grad(g, u, u0, f, h)
    print(f"event {shp}: device, K={it}", flush=True)
    _, gf = A.eikonal3d_grad(dev(g), dev(u), dev(u0), dev(f), h, *shp)
    info = A.ops.last_info
    print(f"tile visits {info['adj_sweeps']}", flush=True)
    assert info["adj_converged"] == 1 and info["fault"] == 0
    assert rel(host(gf), gf_ref) <= GRAD_RTOL


def test_autotune_stores_a_valid_setting(tmp_path):
    # per-class settings for each class of the batch, a global setting, the
    # JSON cache written after timing and read instead of timing
    import adtomo_hip as A
    saved = str(tmp_path / "tune.json")
    cache = str(tmp_path / "cache.json")
    A.save_tune(saved)
    fs, u0s, gs = _mixed_batch(400)
    args = ([dev(a) for a in u0s], [dev(a) for a in fs], [0.5] * len(fs), [1e-3] * len(fs))
    try:
        A.clear_tune_classes("float64")
        best = A.autotune(*args, grad_us=[dev(a) for a in gs], reps=2, cache=cache)
        print("autotune:", best, flush=True)
        assert best["adj_wg"] in (256, 512, 1024)
        assert best["order"] in (0, 1) and best["fwd_ms"] > 0 and best["adj_ms"] > 0
        assert best["fuse_max_n"] >= 0
        cur = A.get_tune("float64")
        assert all(cur[k] == best[k] for k in ("adj_wg", "order", "fuse_max_n"))
        classes = A.tune_classes("float64")
        keys = {(-(-f.shape[2] // 8), -(-min(f.shape[:2]) // 64), int(np.log2(f.size))) for f in fs}
        assert {(c["l_blocks"], c["bands"], c["log2n"]) for c in classes} == keys
        for c in classes:
            print("class:", c, flush=True)
            assert c["diag_max_l"] in (0, 1 << 30) and c["adj_algo"] in (0, 1)
            assert 1 <= c["fwd_segs"] <= 8 and c["fwd_ms"] > 0 and c["adj_ms"] > 0
        res = A.eikonal3d_batched(*args)
        for e in range(len(fs)):
            assert bits_equal(host(res["u"][e]), oracle.eikonal3d(u0s[e], fs[e], 0.5, 1e-3)[0])
        # the cache holds what was chosen and is read back instead of timing
        assert os.path.exists(cache)
        A.clear_tune_classes("float64")
        again = A.autotune(*args, grad_us=[dev(a) for a in gs], reps=2, cache=cache)
        assert all(again[k] == best[k] for k in ("adj_wg", "order", "fuse_max_n"))
        assert A.tune_classes("float64") == classes
    finally:
        A.load_tune(saved)


# ------------------------------------------- FP32 storage, FP64 solves ---

def test_fp32_storage_fp64_solve_bit_exact():
    # both forward kernels (l <= 64 and l > 64), eager and graph, against the
    # oracle's mixed twin; the adjoint then runs in FP32 on these times
    import adtomo_hip as A
    fs, u0s, gs = _mixed_batch(500)
    f32 = [dev(a, np.float32) for a in fs]
    u32 = [dev(a, np.float32) for a in u0s]
    g32 = [dev(a, np.float32) for a in gs]
    res = A.eikonal3d_batched(u32, f32, [0.5] * len(fs), [1e-4] * len(fs), grad_us=g32,
                              compute_f64=True)
    G = A.EikonalGraph(u32, f32, [0.5] * len(fs), [1e-4] * len(fs), compute_f64=True)
    G.launch()
    ginfo = G.infos()
    worst = 0.0
    for e in range(len(fs)):
        um, it, _ = oracle.eikonal3d_m32(u0s[e], fs[e], 0.5, 1e-4)
        assert res["info"][e]["iterations"] == it and ginfo[e]["iterations"] == it, e
        assert bits_equal(host(res["u"][e]), um), e
        assert bits_equal(host(G.us[e]), um), e
        u64, _, _ = oracle.eikonal3d(u0s[e], fs[e], 0.5, 1e-4)
        worst = max(worst, rel(um.astype(np.float64), u64))
        _, gf32, _ = oracle.eikonal3d_grad(gs[e], um, u0s[e], fs[e], 0.5, dtype=np.float32)
        assert rel(host(res["grad_f"][e]), gf32) <= 1e-5, e
    G.close()
    print(f"FP32 storage / FP64 solve vs FP64: max rel travel-time err {worst:.3e}", flush=True)
    assert worst < 1e-6


# ------------------------------------------------------------ edge cases ---

def test_iteration_cap_reports_not_converged():
    # tol 0 never stops early: the reference runs its 20 iterations and
    # prints nothing; the library reports converged = 0 with the same field
    import adtomo_hip as A
    shape = (9, 7, 6)
    r = synth.rng(61)
    f = r.uniform(0.3, 3.0, size=shape)
    u0 = synth.point_source_u0(shape, (4, 3, 2))
    u_ref, it, err = oracle.eikonal3d(u0, f, 0.5, 0.0)
    u = A.eikonal3d(dev(u0), dev(f), 0.5, *shape, 0.0, False)
    assert it == 20 and A.ops.last_info["iterations"] == 20
    assert A.ops.last_info["converged"] == 0
    assert bits_equal(host(u), u_ref)
    res = A.eikonal3d_batched([dev(u0)], [dev(f)], [0.5], [0.0])
    assert res["info"][0]["iterations"] == 20 and res["info"][0]["converged"] == 0
    assert bits_equal(host(res["u"][0]), u_ref)


def test_no_source_field_is_all_z():
    # u0 = 1000 everywhere: no node improves, every node is in Z (u == u0),
    # so lambda = 0, grad_f = 0 and grad_u0 = grad_u
    import adtomo_hip as A
    shape = (8, 9, 10)
    f = synth.rng(62).uniform(0.3, 3.0, size=shape)
    u0 = np.full(shape, 1000.0)
    g = synth.sparse_adjoint_source(shape, 12, seed=63, scale=1.0)
    res = A.eikonal3d_batched([dev(u0)], [dev(f)], [0.5], [1e-3], grad_us=[dev(g)])
    assert bits_equal(host(res["u"][0]), u0)
    assert not host(res["grad_f"][0]).any()
    assert bits_equal(host(res["grad_u0"][0]), g)
    gu0_ref, gf_ref, _ = oracle.eikonal3d_grad(g, u0, u0, f, 0.5)
    assert bits_equal(host(res["grad_u0"][0]), gu0_ref) and not gf_ref.any()


def test_empty_batch_is_a_no_op():
    import adtomo_hip as A
    res = A.eikonal3d_batched([], [], [], [])
    assert res["u"] == [] and res["info"] == []


def test_config2_gradient_central_differences_batched():
    # config 2 (64x64x32, h = 1, checkerboard-layered model, 50 surface
    # receivers): the device gradient against central differences of the
    # device forward (step 1e-6 relative) at 20 seeded nodes; the 40
    # perturbed forwards run as one batched launch, to the fixed point
    import adtomo_hip as A
    torch = _torch()
    shape = (64, 64, 32)
    f = 1.0 / synth.checkerboard_model(shape)
    u0 = synth.point_source_u0(shape, synth.interior_source(shape, seed=22))
    g = synth.sparse_adjoint_source(shape, 50, seed=23, scale=1.0)
    res = A.eikonal3d_batched([dev(u0)], [dev(f)], [1.0], [0.0], grad_us=[dev(g)])
    assert res["info"][0]["err"] == 0.0
    gf = host(res["grad_f"][0])
    r = synth.rng(24)
    cand = np.argwhere(np.abs(gf) > 1e-3 * np.abs(gf).max())
    picks = [tuple(p) for p in cand[r.choice(len(cand), size=20, replace=False)]]
    fs, dl = [], []
    for p in picks:
        d = 1e-6 * f[p]
        for sgn in (1.0, -1.0):
            fp = f.copy()
            fp[p] += sgn * d
            fs.append(dev(fp))
        dl.append(d)
    u0d = dev(u0)
    out = A.eikonal3d_batched([u0d] * len(fs), fs, [1.0] * len(fs), [0.0] * len(fs))
    gd = dev(g)
    loss = [float((u * gd).sum()) for u in out["u"]]
    # rounding of the loss differences: the loss sums terms of size up to
    # S = sum |g u|, so a difference quotient carries ~ eps S / step of noise
    S = float(np.abs(g * host(res["u"][0])).sum())
    worst = 0.0
    for k, p in enumerate(picks):
        fd = (loss[2 * k] - loss[2 * k + 1]) / (2 * dl[k])
        noise = 8 * np.finfo(np.float64).eps * S / dl[k]
        assert abs(fd - gf[p]) <= 1e-5 * abs(gf[p]) + noise, (p, fd, gf[p], noise)
        worst = max(worst, abs(fd - gf[p]) / abs(gf[p]))
    print(f"config2 central differences: worst rel diff {worst:.2e} over 20 nodes", flush=True)
