"""Travel-time tomography driver: the inversion loop of scripts/inversion.jl
(P only) and scripts/inversion_joint.jl (joint P/S) on the device, with a
host L-BFGS and periodic checkpoints (src/mpi_optimize.jl:3-72).

One objective evaluation, all on the GPU except the optimiser state:

    fvar, f   = model(x)                      inversion.jl:42-43,61
    f_S       = vpvs / fvar                   inversion_joint.jl:79 (joint only)
    u_s       = eikonal3d(u0_s, f_phase)      inversion.jl:47-62 (batched launches)
    loss_s, dL/du_s = misfit(u_s, events)     inversion.jl:64-105
    dL/df_s   = adjoint(u_s, dL/du_s)         (the EikonalThreeDGrad op)
    reg, dR/dfvar = smooth_l1(fvar)           inversion.jl:107-121
    loss = sum_s loss_s + reg,  dL/dx (and dL/dvpvs) by the chain rule

Stations are solved ``chunk`` at a time, all phases in the same launch:
each chunk's forward, misfit and adjoint run back to back over one set of
slot buffers, and the per-station dL/df are summed into one accumulator per
phase, so device memory holds ``chunk`` x phases travel-time fields (plus
their adjoint workspace), not one per station.  When every station fits in
one chunk, forward, misfit and adjoint of all stations are one hipGraph
launch (``objective.InversionStepGraph``: each station's adjoint starts as
soon as its own forward and misfit are done); with several chunks the three
run as separate launches per chunk.
By default ``chunk`` is every station when their slots fit in 80% of the
free device memory, else the largest count that fits.
Without chunking (``chunk`` = number of stations) the initial times u0_s,
which depend on the fixed starting model only (inversion.jl:48-60), are
built once; with chunking they are rebuilt on the device per chunk.

The optimiser is the reference's: L-BFGS (history 10) with a static unit
initial step and order-3 backtracking (``adtomo_hip.optim``, restating
Optim.LBFGS(alphaguess=InitialStatic(), linesearch=BackTracking()) of
src/mpi_optimize.jl:35-39), driven with the value alone at line-search trial
points and value plus gradient at accepted points, as Optim drives the
separate f and g! of src/mpi_optimize.jl:10-33.
Checkpoints: every ``steps`` gradient calls the point of that call is saved
as ``iter_<k>.npz`` (keys ``x`` and, for the joint inversion, ``vpvs``)
under ``checkpoint_dir`` (the reference writes ``iter_<k>.h5`` from its g!,
src/mpi_optimize.jl:22-28; h5py is not installed).
"""
from __future__ import annotations

import os
import warnings

import numpy as np

from . import _lib as L
from . import objective as O
from .ops import EikonalGraph, _torch


def _check_solves(infos, forward=True, adjoint=True):
    """Checks the per-solve diagnostics of a launch before its loss or
    gradient is used: a kernel fault (``infos`` raises) or an adjoint that
    did not resolve every node invalidates the results and raises; a forward
    that reached the reference's 20-iteration cap without meeting the tolerance
    is what the reference op also returns silently (Eikonal3D.cpp:74-88), so
    it only warns."""
    if forward and any(not i["converged"] for i in infos):
        bad = sum(1 for i in infos if not i["converged"])
        warnings.warn(f"{bad} of {len(infos)} forward solves reached the iteration cap "
                      "without meeting the tolerance", RuntimeWarning, stacklevel=3)
    if adjoint and any(not i["adj_converged"] for i in infos):
        bad = [e for e, i in enumerate(infos) if not i["adj_converged"]]
        raise L.AdtError(f"adjoint did not resolve every node for solves {bad[:8]}")


class TravelTimeInversion:
    """Travel-time inversion for a set of stations (the sources, by
    reciprocity) and events (the receivers).

    ``tobs`` / ``weights``: (nstation, nevent) arrays for P only, or dicts
    {"P": ..., "S": ...} with ``phases=("P", "S")``; -1 marks a missing pick.
    ``vpvs``: the starting vp/vs ratio of the joint inversion; it scales the
    S initial times (fixed, inversion_joint.jl:65-73) and is a trainable
    parameter of the S slowness (inversion_joint.jl:50-51,79).
    ``huber_delta`` > 0 uses the Huber residual instead of the square.
    ``compute_f64`` with ``dtype="float32"``: FP32 storage with FP64 local
    solves in the forward.
    """

    def __init__(self, vel0, h, stations, events, tobs, weights, smooth=(3, 3), lam=0.0,
                 tol=1e-3, dtype="float64", phases=("P",), vpvs=1.7, huber_delta=0.0,
                 chunk=None, compute_f64=False):
        torch = _torch()
        self.torch = torch
        self.dt = getattr(torch, dtype)
        dev = torch.device("cuda")
        self.shape = tuple(int(s) for s in vel0.shape)
        self.N = int(np.prod(self.shape))
        self.h = float(h)
        self.vel0 = torch.as_tensor(np.ascontiguousarray(vel0), dtype=self.dt, device=dev)
        self.stations = [tuple(map(float, s)) for s in stations]
        ns = len(self.stations)
        self.events = tuple(np.ascontiguousarray(a, dtype=np.float64) for a in events)
        ne = len(self.events[0])
        self.phases = tuple(phases)
        if any(p not in ("P", "S") for p in self.phases) or not self.phases:
            raise ValueError("phases must be a non-empty subset of ('P', 'S')")
        self.joint = "S" in self.phases
        if not isinstance(tobs, dict):
            tobs, weights = {self.phases[0]: tobs}, {self.phases[0]: weights}
        self.vpvs0 = float(vpvs)
        self.huber = float(huber_delta)
        self.smooth = smooth
        self.lam = float(lam)
        self.tobs, self.weights = {}, {}
        for p in self.phases:
            t = np.asarray(tobs[p], dtype=np.float64).reshape(ns, ne)
            w = np.asarray(weights[p], dtype=np.float64).reshape(ns, ne)
            self.tobs[p] = [torch.as_tensor(t[s], dtype=self.dt, device=dev) for s in range(ns)]
            self.weights[p] = [torch.as_tensor(w[s], dtype=self.dt, device=dev) for s in range(ns)]
        # model buffers
        self.x = torch.zeros(self.shape, dtype=self.dt, device=dev)
        self.fvar = torch.empty_like(self.x)
        self.f = {"P": torch.empty_like(self.x)}
        if self.joint:
            self.f["S"] = torch.empty_like(self.x)
        self.grad_f = {p: torch.empty_like(self.x) for p in self.phases}
        # solve slots: `chunk` stations per batched launch, for every phase
        # at once (slots [q*C, (q+1)*C) belong to phase q), so that one
        # launch carries all phases' solves
        if chunk is None:
            chunk = self.fitting_chunk(ns, len(self.phases), self.N, self.x.element_size())
        self.chunk = max(1, min(int(chunk), ns))
        C = self.chunk
        P = len(self.phases)
        self.u0s = [torch.empty_like(self.x) for _ in range(P * C)]
        fs = [self.f[p] for p in self.phases for _ in range(C)]
        hs, tols = [self.h] * (P * C), [tol] * (P * C)
        self.step = self.fwd = self.adj = None
        if C == ns:
            for q, p in enumerate(self.phases):
                O.source_u0(self.stations, self.vel0, self.h, self._u0_scale(p),
                            out=self.u0s[q * C:(q + 1) * C])
            self.step = O.InversionStepGraph(
                self.u0s, fs, hs, tols, *self.events,
                [self.tobs[p][s] for p in self.phases for s in range(ns)],
                [self.weights[p][s] for p in self.phases for s in range(ns)],
                huber_delta=self.huber, compute_f64=compute_f64)
        else:
            self.grad_us = [torch.empty_like(self.x) for _ in range(P * C)]
            self.fwd = EikonalGraph(self.u0s, fs, hs, tols, compute_f64=compute_f64)
            self.adj = EikonalGraph(self.u0s, fs, hs, tols, grad_us=self.grad_us, forward=False,
                                    us=self.fwd.us)
        self.nevals = 0
        self.last_loss = {}

    @staticmethod
    def slot_bytes(N, elem):
        """Device bytes of one solve slot: u0, u, dL/du, dL/du0 and dL/df
        (N elements each) plus the batched workspace: the forward snapshot
        and the tile adjoint's multipliers (N elements each), its upwind
        codes (N bytes) and per-tile counts (N/512 int32)."""
        return N * (7 * elem + 1) + 4 * (N // 512 + 64)

    def fitting_chunk(self, ns, nphases, N, elem, budget=0.8):
        """All stations if their slots fit in ``budget`` of the free device
        memory, else the largest number of stations per launch that does."""
        free, _ = self.torch.cuda.mem_get_info()
        per_station = nphases * self.slot_bytes(N, elem)
        return max(1, min(ns, int(budget * free // per_station)))

    def _u0_scale(self, phase):
        return 1.0 if phase == "P" else self.vpvs0

    @property
    def nparams(self):
        return self.N + (1 if self.joint else 0)

    def initial_params(self):
        """x = 0 (the starting model) and, for the joint inversion, vpvs."""
        x0 = np.zeros(self.nparams)
        if self.joint:
            x0[-1] = self.vpvs0
        return x0

    def _solve(self, grad=True):
        """Forward, misfit and (with ``grad``) adjoint of every station and
        phase; returns the summed misfit per phase and leaves sum_s dL/df_s
        in self.grad_f."""
        C, ns = self.chunk, len(self.stations)
        if self.step is not None:
            # one launch: forward, misfit and adjoint of every station (the
            # adjoint is computed also when only the loss is asked for)
            self.step.launch()
            _check_solves(self.step.infos())
            loss = {}
            for q, p in enumerate(self.phases):
                loss[p] = float(self.step.losses[q * C:(q + 1) * C].cpu().numpy().sum())
                if grad:
                    O.sum_arrays(self.step.grad_fs[q * C:(q + 1) * C], out=self.grad_f[p])
            return loss
        loss = {p: 0.0 for p in self.phases}
        for start in range(0, ns, C):
            k = min(C, ns - start)
            if C < ns:
                for q, p in enumerate(self.phases):
                    O.source_u0(self.stations[start:start + k], self.vel0, self.h,
                                self._u0_scale(p), out=self.u0s[q * C:q * C + k])
            self.fwd.launch()
            _check_solves(self.fwd.infos()[:len(self.u0s)], adjoint=False)
            for q, p in enumerate(self.phases):
                sl = slice(q * C, q * C + k)
                mis = O.misfit(self.fwd.us[sl], *self.events, self.tobs[p][start:start + k],
                               self.weights[p][start:start + k], grad=grad, tcal=False,
                               grad_out=self.grad_us[sl] if grad else None,
                               huber_delta=self.huber)
                loss[p] += float(mis["loss"].cpu().numpy().sum())
            if not grad:
                continue
            self.adj.launch()
            _check_solves(self.adj.infos(), forward=False)
            for q, p in enumerate(self.phases):
                parts = self.adj.grad_fs[q * C:q * C + k]
                O.sum_arrays(parts if start == 0 else [self.grad_f[p]] + parts, out=self.grad_f[p])
        return loss

    def objective(self, params, grad=True):
        """(loss, d loss / d params) at the flattened parameters."""
        torch = self.torch
        params = np.asarray(params, dtype=np.float64)
        self.x.copy_(torch.as_tensor(params[:self.N].reshape(self.shape), dtype=self.dt))
        vpvs = float(params[-1]) if self.joint else None
        O.model(self.x, self.vel0, out=(self.fvar, self.f["P"]))
        if self.joint:
            O.slowness(self.fvar, vpvs, out=self.f["S"])
        self.last_loss = self._solve(grad)
        loss = sum(self.last_loss.values())
        reg_val, grad_fvar = O.smooth_l1(self.fvar, self.smooth[0], self.smooth[1], self.lam)
        self.last_loss["reg"] = float(reg_val.cpu().numpy()[0])
        loss += self.last_loss["reg"]
        self.nevals += 1
        if not grad:
            return loss
        g = np.empty(self.nparams)
        if self.joint:
            gv = O.slowness_grad(self.fvar, vpvs, self.grad_f["S"], grad_fvar=grad_fvar)
            g[-1] = float(gv.cpu().numpy()[0])
        gf = self.grad_f["P"] if "P" in self.phases else None
        gx = O.model_grad(self.x, self.fvar, grad_f=gf, grad_fvar=grad_fvar)
        g[:self.N] = gx.cpu().numpy().reshape(-1)
        return loss, g

    def travel_times(self, phase="P"):
        """Travel times of the last chunk's forward, one (m, n, l) tensor per slot."""
        q = self.phases.index(phase)
        us = self.step.us if self.step is not None else self.fwd.us
        return us[q * self.chunk:(q + 1) * self.chunk]

    def save_checkpoint(self, path, params):
        params = np.asarray(params)
        extra = {"vpvs": params[-1]} if self.joint else {}
        np.savez(path, x=params[:self.N].reshape(self.shape), **extra)

    def run(self, iterations=10, steps=None, checkpoint_dir=None, x0=None, callback=None):
        """L-BFGS over the parameters as src/mpi_optimize.jl:35-48 runs it
        (``adtomo_hip.optim.lbfgs``: the value alone at line-search trial
        points, value and gradient at accepted points); every ``steps``
        gradient calls the point of the call is checkpointed before its
        gradient is computed (src/mpi_optimize.jl:22-28).  Returns (params,
        loss after each iteration); the full optimiser result is kept in
        ``self.result``."""
        from .optim import lbfgs
        calls = {"grad": 0}

        def f(xf):
            return self.objective(xf, grad=False)

        def fg(xf):
            calls["grad"] += 1
            if checkpoint_dir and steps and calls["grad"] % steps == 0:
                os.makedirs(checkpoint_dir, exist_ok=True)
                self.save_checkpoint(os.path.join(checkpoint_dir, f"iter_{calls['grad']}.npz"), xf)
            return self.objective(xf)

        x0 = self.initial_params() if x0 is None else np.asarray(x0, dtype=np.float64).reshape(-1)
        self.result = lbfgs(f, fg, x0, iterations=iterations, callback=callback)
        return self.result.x, self.result.history

    def close(self):
        for g in (self.step, self.fwd, self.adj):
            if g is not None:
                g.close()
