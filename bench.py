"""Benchmark: batched FP64 Eikonal forward + adjoint event-solves/s.

Workload (one "step"): 256 source events on a 128x128x64 grid (h = 0.5 km,
layered 3.5->7.5 km/s + 3% Gaussian random field, seeded interior point
sources, u0 = 1000 with 0 at the source, tol 1e-3, sparse adjoint sources at
30-120 surface receivers), forward + adjoint of every event, as one
hipGraph launch.  Inputs are resident in HBM before timing starts.

Also reported: the plain per-event launch loop on the same events, the
achieved bandwidth of the kernels against the HBM roofline (algorithmic
bytes: forward 8*K*N*(s_f + 2*s_u) per event, adjoint
N*(s_u+s_u0+s_f+s_g+s_gradf+s_gradu0)): "roofline" is the step's kernel when
forward and adjoint run fused in one kernel (the default setting), else the
longer of the two; the separate forward and adjoint kernels are timed on
their own as well; and the CPU oracle timed on a bounded sample of the same
events on the host cores, with, on one 64x64x32 event, the oracle's
triangular adjoint solve against the reference's sparse-LU route (SciPy
SuperLU standing in for Eigen's SparseLU).

Other workloads (``--workload``), each printing its own JSON line:
  uneven  256 events on boxes drawn from 48x48x24 .. 128x128x64 (FP64 and
          FP32, errors against FP64), one hipGraph launch per step
  joint   joint P/S gradient evaluation on 200x200x80 (h = 0.5 km): 100
          surface stations x {P, S} = 200 solves, 500 events, Huber residual
          (delta 0.1 s), picks with N(0, 0.05 s) noise; plus one L-BFGS step
  large   400x400x160 FP32 storage with FP64 local solves (h = 0.25 km):
          150 stations (as many per launch as fit in memory, or --chunk),
          2000 events; against FP64 (and pure FP32) on a 16-station subset

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--nev 256]
                       [--workload headline|uneven|joint|large]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [ROOT, os.path.join(ROOT, "adtomo.jl_amd")]

HBM_PEAK_GBS = 8000.0
FWD_KERNEL = "fwd3_diag_kernel<double, false>"
ADJ_KERNEL = "adj3_kahn_kernel<double>"
ADJ_TILE_KERNEL = "adj3_tile_kernel<double, 8, 8192>"
FUSED_KERNEL = "fwd_adj3_diag_kernel<double, false>"
# HBM bytes (read + write) per launch of each kernel on this workload (256
# events, default setting), measured with rocprofv3 FETCH_SIZE / WRITE_SIZE
# scaled by a copy kernel of known size and checked against the TCC EA
# request counters; kernels not profiled at the current code report null
TRAFFIC_SOURCE = "profiles/pmc_traffic_fwd3_diag_adj3_tile_256x128x128x64.md"
TRAFFIC_BYTES_256 = {FWD_KERNEL: 122.87e9 + 53.99e9, ADJ_TILE_KERNEL: 37.91e9 + 7.98e9}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def _class_key(shape):
    m, n, l = shape
    return (-(-l // 8), -(-min(m, n) // 64), (m * n * l).bit_length() - 1)


def effective_setting(A, dtype, shape):
    """The setting events of this box run with: the global setting with its
    per-event fields (diag_max_l, fwd_segs, adj_algo) taken from the box's
    class setting when there is one."""
    tune = A.get_tune(dtype)
    for c in A.tune_classes(dtype):
        if (c["l_blocks"], c["bands"], c["log2n"]) == _class_key(shape):
            tune.update({k: c[k] for k in ("diag_max_l", "fwd_segs", "adj_algo")})
            tune["class"] = {k: c[k] for k in ("l_blocks", "bands", "log2n", "fwd_ms", "adj_ms")}
    return tune


def class_report(A, dtype):
    return [{k: c[k] for k in ("l_blocks", "bands", "log2n", "diag_max_l", "fwd_segs", "adj_algo",
                               "fwd_ms", "adj_ms")} for c in A.tune_classes(dtype)]


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def _cpu_adjoint_solvers(oracle, synth, shape=(64, 64, 32)):
    """The two CPU adjoint routes, timed single-threaded on one event of the
    64x64x32 grid (the largest where a sparse LU finishes in seconds): the
    oracle's exact triangular solve, and the reference's route (assemble the
    upwind system, Eikonal3D.cpp:118-187, and solve it with a sparse LU,
    :186-193), here SciPy's SuperLU standing in for Eigen's SparseLU."""
    import numpy as np
    f = 1.0 / synth.layered_grf_model(shape, 0.5, seed=1)
    u0 = synth.point_source_u0(shape, synth.interior_source(shape, seed=2))
    u, _, _ = oracle.eikonal3d(u0, f, 0.5, 1e-3)
    g = synth.sparse_adjoint_source(shape, 60, seed=3)
    t = time.perf_counter()
    _, gf_tri, _ = oracle.eikonal3d_grad(g, u, u0, f, 0.5)
    t_tri = time.perf_counter() - t
    t = time.perf_counter()
    _, gf_lu, _ = oracle.lu_grad_3d(g, u, u0, f, 0.5)
    t_lu = time.perf_counter() - t
    return {"grid": list(shape), "triangular_s": t_tri, "sparse_lu_s": t_lu,
            "sparse_lu": "scipy SuperLU (stand-in for Eigen SparseLU)",
            "max_rel_diff": float(np.abs(gf_lu - gf_tri).max() / np.abs(gf_tri).max())}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nev", type=int, default=256)
    ap.add_argument("--plain-events", type=int, default=1)
    ap.add_argument("--cpu-events", type=int, default=0, help="0 = 20 per host thread")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-superlu", action="store_true",
                    help="skip timing the sparse-LU adjoint route on the host")
    ap.add_argument("--no-fp32", action="store_true", help="skip the FP32 variant")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default launch setting instead of timing the candidates")
    ap.add_argument("--workload", default="headline",
                    choices=["headline", "uneven", "joint", "large"])
    ap.add_argument("--chunk", type=int, default=0,
                    help="stations per batched launch (0: as many as fit in device memory)")
    ap.add_argument("--calib-copy", action="store_true",
                    help="run a 1 GiB device copy first (reference kernel for HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "events would shard one-per-rank across GPUs, see the design notes", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import adtomo_hip as A
    from adtomo_hip import synth

    assert torch.cuda.is_available(), "bench.py needs a HIP device"
    if args.workload != "headline":
        return {"uneven": run_uneven, "joint": run_joint, "large": run_large}[args.workload](args)
    if args.calib_copy:
        # 2^27 doubles = 1 GiB read + 1 GiB written by the library's plain
        # copy kernel (arith_kernel<double>, op 6)
        a = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        b = A.selftest_arith(6, a)
        torch.cuda.synchronize()
        del a, b
    shape = (128, 128, 64)
    m, n, l = shape
    N = m * n * l
    h, tol = 0.5, 1e-3
    nev = args.nev
    t0 = time.time()
    f, u0s, gs, srcs = synth.headline_event_set(nev, shape, h)
    log(f"workload built in {time.time() - t0:.1f}s")
    fd = torch.from_numpy(f).cuda()
    fs = [fd] * nev  # one shared model, as in scripts/inversion.jl:61
    u0d = [torch.from_numpy(a).cuda() for a in u0s]
    gd = [torch.from_numpy(a).cuda() for a in gs]
    del u0s, gs

    # launch setting: timed on this workload by the library's autotuner
    if not args.no_autotune:
        t0 = time.time()
        A.autotune(u0d, fs, [h] * nev, [tol] * nev, grad_us=gd, reps=1)
        log(f"autotune ({time.time() - t0:.1f}s): {A.get_tune('float64')}, "
            f"classes {class_report(A, 'float64')}")
    tune = effective_setting(A, "float64", shape)

    G = A.EikonalGraph(u0d, fs, [h] * nev, [tol] * nev, grad_us=gd)
    for _ in range(args.warmup):
        G.launch()
    torch.cuda.synchronize()
    infos = G.infos()
    Ks = np.array([i["iterations"] for i in infos])
    sweeps = np.array([i["adj_sweeps"] for i in infos])
    assert all(i["converged"] for i in infos) and all(i["adj_converged"] for i in infos)
    log(f"K: mean {Ks.mean():.2f} min {Ks.min()} max {Ks.max()}; adjoint levels mean {sweeps.mean():.1f}")

    s = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t_wall = time.perf_counter()
    e0.record(s)
    for _ in range(args.steps):
        G.launch()
    e1.record(s)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_wall
    ms = e0.elapsed_time(e1) / args.steps
    value = nev / (ms / 1e3)
    log(f"batched fwd+adj: {ms:.2f} ms/step, {value:.1f} event-solves/s (wall {wall:.2f}s)")

    # the two kernels timed alone: forward-only and adjoint-only graphs on
    # the same buffers, HIP events on the launch stream
    Gf = A.EikonalGraph(u0d, fs, [h] * nev, [tol] * nev, us=G.us)
    Ga = A.EikonalGraph(u0d, fs, [h] * nev, [tol] * nev, grad_us=gd, forward=False, us=G.us)

    def time_graph(graph, reps):
        graph.launch()
        torch.cuda.synchronize()
        b0, b1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        b0.record(s)
        for _ in range(reps):
            graph.launch()
        b1.record(s)
        torch.cuda.synchronize()
        return b0.elapsed_time(b1) / reps

    fwd_ms = time_graph(Gf, max(2, args.steps))
    adj_ms = time_graph(Ga, max(2, args.steps))
    Gf.close()
    Ga.close()
    s_el = 8
    fwd_bytes = float(np.sum(8 * Ks * N * (s_el + 2 * s_el)))
    adj_bytes = float(nev * N * 6 * s_el)
    fwd_gbs = fwd_bytes / (fwd_ms / 1e3) / 1e9
    adj_gbs = adj_bytes / (adj_ms / 1e3) / 1e9
    total_gbs = (fwd_bytes + adj_bytes) / (ms / 1e3) / 1e9
    node_updates = float(np.sum(8 * Ks * N))
    log(f"forward kernel {fwd_ms:.2f} ms -> {fwd_gbs:.1f} GB/s algorithmic; "
        f"adjoint kernel {adj_ms:.2f} ms -> {adj_gbs:.1f} GB/s")

    # FP32 variant of the same step, with its error against FP64 (the FP64
    # device results are bit-exact with the oracle's travel times)
    fp32 = mixed = None
    if not args.no_fp32:
        ff = fd.float()
        u0f = [a.float() for a in u0d]
        gdf = [a.float() for a in gd]

        def variant(mix):
            if not args.no_autotune:
                A.autotune(u0f, [ff] * nev, [h] * nev, [tol] * nev, grad_us=gdf, reps=1,
                           compute_f64=mix)
            G32 = A.EikonalGraph(u0f, [ff] * nev, [h] * nev, [tol] * nev, grad_us=gdf,
                                 compute_f64=mix)
            ms32 = time_graph(G32, max(2, args.steps))
            K32 = np.array([i["iterations"] for i in G32.infos()])
            eu_max = eu_ss = u_max = eg_max = eg_ss = g_max = 0.0
            for e in range(nev):
                du = G32.us[e].double() - G.us[e]
                dg = G32.grad_fs[e].double() - G.grad_fs[e]
                eu_max = max(eu_max, float(du.abs().max()))
                eg_max = max(eg_max, float(dg.abs().max()))
                eu_ss += float((du * du).sum())
                eg_ss += float((dg * dg).sum())
                u_max = max(u_max, float(G.us[e].abs().max()))
                g_max = max(g_max, float(G.grad_fs[e].abs().max()))
            G32.close()
            out = {"event_solves_per_s": nev / (ms32 / 1e3), "ms_per_step": ms32,
                   "K_mean": float(K32.mean()),
                   "u_err_max_rel": eu_max / u_max, "u_err_rms_s": (eu_ss / (nev * N)) ** 0.5,
                   "grad_f_err_max_rel": eg_max / g_max,
                   "grad_f_err_rms": (eg_ss / (nev * N)) ** 0.5}
            log(f"{'fp32 storage, fp64 solves' if mix else 'fp32'}: {ms32:.2f} ms/step -> "
                f"{out['event_solves_per_s']:.1f} event-solves/s; travel time max rel err "
                f"{out['u_err_max_rel']:.2e}, rms {out['u_err_rms_s']:.2e} s; "
                f"grad_f max rel err {out['grad_f_err_max_rel']:.2e}")
            return out

        fp32 = variant(False)
        mixed = variant(True)
        del u0f, gdf, ff

    plain = None
    if not args.no_plain and args.plain_events > 0:
        pe = args.plain_events
        torch.cuda.synchronize()
        tp = time.perf_counter()
        for e in range(pe):
            u = A.eikonal3d_plain(u0d[e], fd, h, tol)
            A.eikonal3d_grad_plain(gd[e], u, u0d[e], fd, h)
        torch.cuda.synchronize()
        tp = time.perf_counter() - tp
        plain = pe / tp
        log(f"plain per-event loop: {tp / pe * 1e3:.1f} ms/event -> {plain:.2f} event-solves/s")

    cpu = None
    if not args.no_cpu:
        from oracle import oracle
        threads = min(16, os.cpu_count() or 1)
        ce = args.cpu_events or 20 * threads  # about 10 s on the host
        sub_u0 = [synth.point_source_u0(shape, srcs[e % nev]) for e in range(ce)]
        sub_g = [synth.sparse_adjoint_source(shape, 60, seed=2024 * 1000 + e) for e in range(ce)]
        tc = time.perf_counter()
        oracle.eikonal3d_batch(sub_u0, [f] * ce, sub_g, [h] * ce, tol, threads)
        tc = time.perf_counter() - tc
        cpu = {"value": ce / tc, "unit": "event-solves/s", "cores": threads, "kind": "port",
               "sample": f"{ce} events of the same workload (128x128x64, FP64 fwd+adjoint), "
                         f"one event per thread, {tc:.1f}s",
               "host_cpus": os.cpu_count(), "cpu_model": _cpu_model()}
        log(f"cpu oracle: {ce} events in {tc:.1f}s on {threads} threads -> {cpu['value']:.3f}/s")
        if not args.no_superlu:
            cpu["adjoint_solvers"] = _cpu_adjoint_solvers(oracle, synth)
            log(f"cpu adjoint on {cpu['adjoint_solvers']['grid']}: "
                f"{cpu['adjoint_solvers']['triangular_s']:.3f}s triangular solve, "
                f"{cpu['adjoint_solvers']['sparse_lu_s']:.2f}s sparse LU")

    def roof(kernel, gbs, kms):
        return {"bound": "hbm", "kernel": kernel, "achieved": round(gbs, 2), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 5), "launch_ms": round(kms, 3),
                "traffic": TRAFFIC_BYTES_256.get(kernel) if nev == 256 else None,
                "traffic_source": TRAFFIC_SOURCE if nev == 256 and kernel in TRAFFIC_BYTES_256 else None}

    adj_kernel = ADJ_TILE_KERNEL if tune.get("adj_algo", 1) == 0 else ADJ_KERNEL
    roofs = [roof(FWD_KERNEL, fwd_gbs, fwd_ms), roof(adj_kernel, adj_gbs, adj_ms)]
    # with the fused setting (node-level adjoint only) the step is one kernel
    # doing each event's forward and then its adjoint: the dominant kernel
    fused = (tune.get("adj_algo", 1) == 1 and N <= tune["fuse_max_n"]
             and tune["diag_max_l"] >= shape[2])
    if fused:
        roofs.insert(0, roof(FUSED_KERNEL, total_gbs, ms))
    dominant = roofs[0] if fused else max(roofs, key=lambda r: r["launch_ms"])
    out = {
        "metric": "event-solves/s fwd+adjoint at 128x128x64 (batched), HBM GB/s",
        "value": round(value, 3),
        "unit": "event-solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "256 point-source events, 128x128x64, h=0.5 km, layered "
                               "3.5-7.5 km/s + 3% GRF, tol 1e-3, fwd+adjoint, one hipGraph",
                   "events": nev, "grid": list(shape), "K_mean": float(Ks.mean()),
                   "serial_depth_per_sweep": int(sum(shape) - 2),
                   "adj_sweeps_mean": float(sweeps.mean()),
                   "launch_setting": {k: tune[k] for k in ("diag_max_l", "adj_wg", "order", "fuse_max_n",
                                                           "adj_algo", "fwd_segs") + (
                                                               ("class",) if "class" in tune else ())}},
        "roofline": dominant,
        "roofline_kernels": roofs,
        "cpu_baseline": cpu,
        "fp32": fp32,
        "fp32_storage_fp64_solve": mixed,
        "plain_event_solves_per_s": plain,
        "batched_vs_plain": (value / plain) if plain else None,
        "node_updates_per_s": node_updates / (fwd_ms / 1e3),
        "fwd_ms": fwd_ms, "adj_ms": adj_ms,
        "adjoint_gbs_algorithmic": adj_gbs,
        "step_gbs_algorithmic": total_gbs,
    }
    print(json.dumps(out), flush=True)


def _line(metric, value, unit, ms, args, dtype, config, **extra):
    out = {"metric": metric, "value": round(value, 4), "unit": unit, "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 3),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": dtype,
           "data": "synthetic", "config": config}
    out.update(extra)
    print(json.dumps(out), flush=True)


def _time(fn, steps):
    import torch
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t) / steps * 1e3


def _rel_err(a, b):
    """(max |a-b| / max |b|, rms(a-b)) over lists of device tensors."""
    mx = ref = ss = cnt = 0.0
    for x, y in zip(a, b):
        d = x.double() - y.double()
        mx = max(mx, float(d.abs().max()))
        ref = max(ref, float(y.double().abs().max()))
        ss += float((d * d).sum())
        cnt += d.numel()
    return mx / max(ref, 1e-300), (ss / max(cnt, 1)) ** 0.5


def run_uneven(args):
    """256 events on uneven boxes, FP64 and FP32 (one graph launch per step)."""
    import numpy as np
    import torch
    import adtomo_hip as A
    from adtomo_hip import synth
    nev = args.nev
    fs, u0s, gs, _ = synth.uneven_box_set(nev)
    h, tol = 0.5, 1e-3
    nodes = float(sum(f.size for f in fs))
    res = {}
    for name, dt, mix in (("f64", torch.float64, False), ("f32", torch.float32, False),
                          ("mixed", torch.float32, True)):
        fd = [torch.from_numpy(a).to("cuda", dt) for a in fs]
        ud = [torch.from_numpy(a).to("cuda", dt) for a in u0s]
        gd = [torch.from_numpy(a).to("cuda", dt) for a in gs]
        tune = A.autotune(ud, fd, [h] * nev, [tol] * nev, grad_us=gd, reps=1, compute_f64=mix)
        tune["classes"] = class_report(A, "float32" if name != "f64" else "float64")
        G = A.EikonalGraph(ud, fd, [h] * nev, [tol] * nev, grad_us=gd, compute_f64=mix)
        for _ in range(args.warmup):
            G.launch()
        ms = _time(G.launch, args.steps)
        K = np.array([i["iterations"] for i in G.infos()])
        res[name] = (G, ms, K, tune)
        log(f"uneven {name}: {ms:.2f} ms/step -> {nev / ms * 1e3:.1f} event-solves/s, K mean {K.mean():.2f}, {tune}")
    G64, ms64, K64, tune64 = res["f64"]

    def variant(name):
        G, ms_, K, _ = res[name]
        eu = _rel_err(G.us, G64.us)
        eg = _rel_err(G.grad_fs, G64.grad_fs)
        return {"event_solves_per_s": nev / ms_ * 1e3, "ms_per_step": ms_, "K_mean": float(K.mean()),
                "u_err_max_rel": eu[0], "u_err_rms_s": eu[1], "grad_f_err_max_rel": eg[0],
                "grad_f_err_rms": eg[1]}
    _line("event-solves/s fwd+adjoint, 256 events on uneven boxes 48x48x24..128x128x64 (batched)",
          nev / ms64 * 1e3, "event-solves/s", ms64, args, "f64",
          {"workload": "256 events, boxes uniform in 48x48x24..128x128x64, h=0.5 km, layered "
                       "3.5-7.5 km/s + 3% GRF, tol 1e-3, fwd+adjoint, one hipGraph",
           "events": nev, "nodes_total": nodes, "K_mean": float(K64.mean()),
           "launch_setting": {k: tune64[k] for k in ("diag_max_l", "adj_wg", "order", "fuse_max_n",
                                                     "classes")}},
          fp32=variant("f32"), fp32_storage_fp64_solve=variant("mixed"))


def _picks(vel, h, stations, events, chunk):
    """Travel times (FP64, device) from each station to every event on
    model ``vel``: (nstation, nevent) numpy array."""
    import numpy as np
    import torch
    import adtomo_hip as A
    from adtomo_hip import objective as O
    v = torch.from_numpy(np.ascontiguousarray(vel)).cuda()
    f = O.slowness(v, 1.0)
    ne = len(events[0])
    zero = torch.zeros(ne, dtype=torch.float64, device="cuda")
    out = np.empty((len(stations), ne))
    for s0 in range(0, len(stations), chunk):
        st = stations[s0:s0 + chunk]
        u0s = O.source_u0(st, v, h)
        r = A.eikonal3d_batched(u0s, [f] * len(st), [h] * len(st), [1e-3] * len(st),
                                return_info=False)
        mis = O.misfit(r["u"], *events, [zero] * len(st), [zero] * len(st), grad=False)
        for k in range(len(st)):
            out[s0 + k] = mis["tcal"][k].cpu().numpy()
        del r, u0s, mis
    return out


def run_joint(args):
    """Joint P/S gradient evaluations and one L-BFGS step on 200x200x80."""
    import numpy as np
    from adtomo_hip import synth
    from adtomo_hip.inversion import TravelTimeInversion
    shape, h = (200, 200, 80), 0.5
    t0 = time.time()
    vp = synth.layered_grf_model(shape, h, 5.0, 7.0, 0.04, 5.0, seed=41)
    vs = vp / 1.73 * (1.0 + synth.gaussian_random_field(shape, 5.0 / h, 0.02, seed=42))
    vel0 = synth.layered_model(shape, 5.0, 7.0)
    r = synth.rng(43)
    stations = [(float(r.uniform(1, shape[0])), float(r.uniform(1, shape[1])), 1.0)
                for _ in range(100)]
    events = tuple(r.uniform(1, s, 500) for s in shape)
    tp = _picks(vp, h, stations, events, 100) + r.normal(0, 0.05, (100, 500))
    ts = _picks(vs, h, stations, events, 100) + r.normal(0, 0.05, (100, 500))
    w = np.ones((100, 500))
    log(f"joint workload built in {time.time() - t0:.1f}s")
    inv = TravelTimeInversion(vel0, h, stations, events, {"P": tp, "S": ts}, {"P": w, "S": w},
                              smooth=(3, 3), lam=0.1, phases=("P", "S"), vpvs=1.7,
                              huber_delta=0.1, chunk=args.chunk or None)
    chunk = inv.chunk
    p0 = inv.initial_params()
    L0 = None
    for _ in range(args.warmup):
        L0, g0 = inv.objective(p0)
    ms = _time(lambda: inv.objective(p0), args.steps)
    log(f"joint: {ms:.1f} ms per gradient evaluation (200 solves fwd+adj + objective), loss {L0}")
    t = time.perf_counter()
    x1, hist = inv.run(iterations=1)
    t = time.perf_counter() - t
    res = inv.result
    log(f"one L-BFGS step: {res.f_calls} values, {res.g_calls} gradients in {t:.2f}s, loss "
        f"{hist[0]:.4g} -> {hist[-1]:.4g}, vp/vs {p0[-1]:.3f} -> {x1[-1]:.4f}")
    inv.close()
    _line("joint P/S gradient evaluations: solves/s (fwd+adjoint) at 200x200x80",
          200 / ms * 1e3, "solves/s", ms, args, "f64",
          {"workload": "joint Vp/Vs gradient, 200x200x80, h=0.5 km, 100 stations x {P,S}, "
                       "500 events, Huber delta 0.1 s, noise N(0, 0.05 s)",
           "solves": 200, "events": 500, "chunk": chunk},
          lbfgs_step={"values": res.f_calls, "gradients": res.g_calls, "seconds": t,
                      "loss_before": hist[0], "loss_after": hist[-1], "vpvs_after": float(x1[-1])})


def run_large(args):
    """400x400x160 FP32 gradient in station chunks; FP64 check on 16 stations."""
    import numpy as np
    import torch
    from adtomo_hip import synth
    from adtomo_hip.inversion import TravelTimeInversion
    shape, h = (400, 400, 160), 0.25
    t0 = time.time()
    vtrue = synth.layered_grf_model(shape, h, 3.5, 7.5, 0.03, 5.0, seed=51)
    vel0 = synth.layered_model(shape, 3.5, 7.5)
    r = synth.rng(52)
    ns, ne = 150, 2000
    stations = [(float(r.uniform(1, shape[0])), float(r.uniform(1, shape[1])), 1.0)
                for _ in range(ns)]
    events = tuple(r.uniform(1, s, ne) for s in shape)
    tobs = _picks(vtrue, h, stations, events, 50) + r.normal(0, 0.05, (ns, ne))
    del vtrue
    w = np.ones((ns, ne))
    log(f"large workload built in {time.time() - t0:.1f}s")
    # chunk: --chunk, else as many stations per launch as fit in device memory
    inv = TravelTimeInversion(vel0, h, stations, events, tobs, w, smooth=(3, 3), lam=0.1,
                              dtype="float32", chunk=args.chunk or None, compute_f64=True)
    chunk = inv.chunk
    p0 = inv.initial_params()
    for _ in range(args.warmup):
        inv.objective(p0)
    ms = _time(lambda: inv.objective(p0), args.steps)
    peak = torch.cuda.max_memory_allocated() / 2**30
    log(f"large: {ms / 1e3:.2f} s per FP32-storage gradient evaluation ({ns} solves, chunks of "
        f"{chunk}, FP64 local solves), peak torch device memory {peak:.1f} GiB")
    inv.close()
    del inv
    torch.cuda.empty_cache()
    # FP32 against FP64 on a 16-station subset
    sub = slice(0, 16)
    g = {}
    for name, dt, mix in (("mixed", "float32", True), ("float32", "float32", False),
                          ("float64", "float64", False)):
        # no regulariser here: at x = 0 on a layered model fvar - box(fvar)
        # is rounding noise, and its sign would dominate the comparison
        iv = TravelTimeInversion(vel0, h, stations[sub], events, tobs[sub], w[sub], smooth=(3, 3),
                                 lam=0.0, dtype=dt, compute_f64=mix)
        t = time.perf_counter()
        L, gr = iv.objective(p0)
        t = time.perf_counter() - t
        g[name] = (L, gr, [u.double() for u in iv.travel_times()], t)
        iv.close()
        del iv
        torch.cuda.empty_cache()
        log(f"16-station subset, {name}: {t:.2f} s")

    def errs(name):
        eu = _rel_err(g[name][2], g["float64"][2])
        gd = np.abs(g[name][1] - g["float64"][1])
        return {"u_err_max_rel": eu[0], "u_err_rms_s": eu[1],
                "loss_rel": abs(g[name][0] - g["float64"][0]) / abs(g["float64"][0]),
                "grad_err_max_rel": float(gd.max() / np.abs(g["float64"][1]).max()),
                "grad_err_rms": float((gd ** 2).mean() ** 0.5), "seconds": g[name][3]}
    _line("FP32-storage gradient evaluations: solves/s (fwd+adjoint) at 400x400x160",
          ns / ms * 1e3, "solves/s", ms, args, "f32",
          {"workload": "400x400x160 (25.6M nodes), h=0.25 km, layered 3.5-7.5 km/s + 3% GRF, "
                       f"{ns} stations in chunks of {chunk}, {ne} events, FP32 storage with "
                       "FP64 local solves",
           "solves": ns, "events": ne, "chunk": chunk},
          peak_torch_device_gib=peak,
          vs_fp64_16_stations={"fp32_storage_fp64_solve": errs("mixed"),
                               "fp32": errs("float32"), "fp64_seconds": g["float64"][3]})


if __name__ == "__main__":
    main()
