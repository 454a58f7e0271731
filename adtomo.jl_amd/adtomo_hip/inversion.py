"""Travel-time tomography driver: the inversion loop of scripts/inversion.jl
on the device, with a host L-BFGS and periodic checkpoints
(src/mpi_optimize.jl:3-72).

One objective evaluation, all on the GPU except the optimiser state:

    fvar, f   = model(x)                      scripts/inversion.jl:42-43,61
    u_s       = eikonal3d(u0_s, f)            :47-62 (one batched launch)
    loss_s, dL/du_s = misfit(u_s, events)     :64-105
    dL/df_s   = adjoint(u_s, dL/du_s)         (the EikonalThreeDGrad op)
    reg, dR/dfvar = smooth_l1(fvar)           :107-121
    loss = sum_s loss_s + reg,  dL/dx = model_grad(...)

The u0_s depend on the fixed starting model vel0 only (:48-60) and are
built once.  The optimiser is SciPy's L-BFGS-B on the flattened x (the
reference uses Optim.jl's LBFGS with a backtracking line search, which is
not available here); the objective and its gradient are the reference's.
Checkpoints: every ``steps`` gradient evaluations, x is saved as
``iter_<k>.npy`` under ``checkpoint_dir`` (the reference writes
``iter_<k>.h5``, src/mpi_optimize.jl:24-28).
"""
from __future__ import annotations

import os

import numpy as np

from . import objective as O
from .ops import EikonalGraph, _torch


class TravelTimeInversion:
    """P-wave travel-time inversion for one set of stations (sources)."""

    def __init__(self, vel0, h, stations, events, tobs, weights, smooth=(3, 3), lam=0.0,
                 tol=1e-3, dtype="float64"):
        torch = _torch()
        self.torch = torch
        self.dt = getattr(torch, dtype)
        dev = torch.device("cuda")
        self.shape = tuple(vel0.shape)
        self.h = float(h)
        self.vel0 = torch.as_tensor(np.ascontiguousarray(vel0), dtype=self.dt, device=dev)
        self.stations = [tuple(map(float, s)) for s in stations]
        ns = len(self.stations)
        ex, ey, ez = (np.ascontiguousarray(a, dtype=np.float64) for a in events)
        self.events = (ex, ey, ez)
        tobs = np.asarray(tobs, dtype=np.float64).reshape(ns, len(ex))
        weights = np.asarray(weights, dtype=np.float64).reshape(ns, len(ex))
        self.tobs = [torch.as_tensor(tobs[s], dtype=self.dt, device=dev) for s in range(ns)]
        self.weights = [torch.as_tensor(weights[s], dtype=self.dt, device=dev) for s in range(ns)]
        self.smooth = smooth
        self.lam = float(lam)
        # fixed buffers shared by the captured graphs
        self.x = torch.zeros(self.shape, dtype=self.dt, device=dev)
        self.fvar = torch.empty_like(self.x)
        self.f = torch.empty_like(self.x)
        self.u0s = O.source_u0(self.stations, self.vel0, self.h)
        self.grad_us = [torch.empty_like(self.x) for _ in range(ns)]
        fs = [self.f] * ns
        self.fwd = EikonalGraph(self.u0s, fs, [self.h] * ns, [tol] * ns)
        self.adj = EikonalGraph(self.u0s, fs, [self.h] * ns, [tol] * ns, grad_us=self.grad_us,
                                forward=False, us=self.fwd.us)
        self.nevals = 0

    def _load_x(self, xflat):
        self.x.copy_(self.torch.as_tensor(np.asarray(xflat, dtype=np.float64).reshape(self.shape),
                                          dtype=self.dt))

    def objective(self, xflat, grad=True):
        """(loss, dL/dx flattened) at model parameters x."""
        torch = self.torch
        self._load_x(xflat)
        fvar, f = O.model(self.x, self.vel0)
        self.fvar.copy_(fvar)
        self.f.copy_(f)
        self.fwd.launch()
        mis = O.misfit(self.fwd.us, *self.events, self.tobs, self.weights, grad=grad, tcal=False,
                       grad_out=self.grad_us if grad else None)
        loss = float(mis["loss"].sum().item())
        reg_val, reg_grad = O.smooth_l1(self.fvar, self.smooth[0], self.smooth[1], self.lam,
                                        grad=grad)
        loss += float(reg_val.item())
        self.nevals += 1
        if not grad:
            return loss
        self.adj.launch()
        grad_f = O.sum_arrays(self.adj.grad_fs)
        gx = O.model_grad(self.x, self.fvar, grad_f=grad_f, grad_fvar=reg_grad)
        torch.cuda.synchronize()
        return loss, gx.double().cpu().numpy().reshape(-1)

    def travel_times(self):
        """Travel times of the last forward, one (m, n, l) tensor per station."""
        return self.fwd.us

    def run(self, iterations=10, steps=None, checkpoint_dir=None, x0=None):
        """L-BFGS over x; returns (x, loss history)."""
        from scipy.optimize import minimize
        history = []
        calls = {"grad": 0}

        def fun(xf):
            L, g = self.objective(xf)
            history.append(L)
            calls["grad"] += 1
            if checkpoint_dir and steps and calls["grad"] % steps == 0:
                os.makedirs(checkpoint_dir, exist_ok=True)
                np.save(os.path.join(checkpoint_dir, f"iter_{calls['grad']}.npy"),
                        np.asarray(xf).reshape(self.shape))
            return L, g

        x0 = np.zeros(int(np.prod(self.shape))) if x0 is None else np.asarray(x0).reshape(-1)
        res = minimize(fun, x0, jac=True, method="L-BFGS-B", options={"maxiter": iterations})
        return res.x.reshape(self.shape), history

    def close(self):
        self.fwd.close()
        self.adj.close()
