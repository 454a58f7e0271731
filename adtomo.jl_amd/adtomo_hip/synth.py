"""Seeded synthetic tomography workloads (velocity models, sources, receivers).

All generators use numpy ``Generator(PCG64(seed))`` so every workload is
reproducible from its seed.  Axis conventions follow the reference:
3D fields have shape (m, n, l) with k (the last, contiguous axis) the depth
axis (layered models vary along k, receivers sit on k = 0); 2D fields have
shape (n+1, m+1) with rows = depth j (src/eikonal_op.jl:5).

The source initial field follows scripts/inversion.jl:52-60 (1000 everywhere,
straight-line times on the 8 nodes around the station) or tests/test3d.jl:21-22
(1000 everywhere, 0 at a grid-node source).
"""
from __future__ import annotations

import numpy as np


def rng(seed: int) -> np.random.Generator:
    return np.random.Generator(np.random.PCG64(seed))


# ----------------------------------------------------------- velocity ---

def gaussian_random_field(shape, corr_nodes: float, std: float, seed: int) -> np.ndarray:
    """White noise smoothed by a Gaussian of ``corr_nodes`` nodes (periodic,
    by FFT), rescaled to standard deviation ``std``."""
    r = rng(seed)
    w = r.standard_normal(shape)
    if corr_nodes > 0:
        fw = np.fft.fftn(w)
        k2 = np.zeros(shape)
        for ax, nax in enumerate(shape):
            k = np.fft.fftfreq(nax)
            sh = [1] * len(shape)
            sh[ax] = nax
            k2 = k2 + (2 * np.pi * k.reshape(sh)) ** 2
        w = np.real(np.fft.ifftn(fw * np.exp(-0.5 * corr_nodes ** 2 * k2)))
    s = w.std()
    return w * (std / s) if s > 0 else w


def layered_model(shape, v_top: float, v_bot: float) -> np.ndarray:
    """Velocity rising linearly with depth (the last axis)."""
    l = shape[-1]
    prof = np.linspace(v_top, v_bot, l)
    return np.broadcast_to(prof, shape).copy()


def layered_grf_model(shape, h: float, v_top=3.5, v_bot=7.5, std=0.03, corr_km=5.0,
                      seed: int = 0) -> np.ndarray:
    """Layered velocity times (1 + GRF) with correlation length ``corr_km``."""
    v = layered_model(shape, v_top, v_bot)
    return v * (1.0 + gaussian_random_field(shape, corr_km / h, std, seed))


def checkerboard_model(shape, v_top=4.0, v_bot=8.0, amp=0.05, period=16) -> np.ndarray:
    """Layered 4->8 km/s times (1 +/- amp) on a checkerboard of ``period`` nodes."""
    v = layered_model(shape, v_top, v_bot)
    idx = np.indices(shape)
    parity = sum((ix // (period // 2)) for ix in idx) % 2
    return v * np.where(parity == 0, 1.0 + amp, 1.0 - amp)


def gaussians_2d(shape, h: float, v0=6.0, count=3, amp=0.10, sigma_km=8.0, seed: int = 0):
    """2D background ``v0`` plus ``count`` Gaussian anomalies of +/-``amp``."""
    r = rng(seed)
    n1, m1 = shape
    jj, ii = np.meshgrid(np.arange(n1), np.arange(m1), indexing="ij")
    v = np.full(shape, v0, dtype=np.float64)
    for _ in range(count):
        cj, ci = r.uniform(0, n1 - 1), r.uniform(0, m1 - 1)
        sgn = 1.0 if r.uniform() < 0.5 else -1.0
        d2 = ((jj - cj) ** 2 + (ii - ci) ** 2) * h * h
        v = v * (1.0 + sgn * amp * np.exp(-0.5 * d2 / sigma_km ** 2))
    return v


# -------------------------------------------------------- sources etc ---

def point_source_u0(shape, src, far: float = 1000.0) -> np.ndarray:
    """u0 = ``far`` everywhere and 0 at the grid-node source."""
    u0 = np.full(shape, far, dtype=np.float64)
    u0[tuple(src)] = 0.0
    return u0


def station_u0(v: np.ndarray, h: float, x: float, y: float, z: float,
               far: float = 1000.0) -> np.ndarray:
    """u0 of scripts/inversion.jl:52-60 for a station at fractional 0-based
    grid position (x, y, z): straight-line times on the 8 surrounding nodes."""
    u0 = np.full(v.shape, far, dtype=np.float64)
    xs = {int(np.floor(x)), int(np.ceil(x))}
    ys = {int(np.floor(y)), int(np.ceil(y))}
    zs = {int(np.floor(z)), int(np.ceil(z))}
    for a in xs:
        for b in ys:
            for c in zs:
                u0[a, b, c] = np.sqrt((x - a) ** 2 + (y - b) ** 2 + (z - c) ** 2) * h / v[a, b, c]
    return u0


def interior_source(shape, seed: int, margin: int = 2):
    r = rng(seed)
    return tuple(int(r.integers(margin, s - margin)) for s in shape)


def surface_receivers(shape, count: int, seed: int):
    """``count`` distinct receivers on the k = 0 surface; returns index arrays."""
    r = rng(seed)
    m, n = shape[0], shape[1]
    flat = r.choice(m * n, size=min(count, m * n), replace=False)
    return flat // n, flat % n


def sparse_adjoint_source(shape, count: int, seed: int, scale: float = 0.05) -> np.ndarray:
    """grad_u with seeded residuals at ``count`` surface receivers, zero elsewhere."""
    r = rng(seed)
    g = np.zeros(shape, dtype=np.float64)
    ri, rj = surface_receivers(shape, count, seed + 1)
    g[ri, rj, 0] = r.normal(0.0, scale, size=len(ri))
    return g


# ------------------------------------------------------------ workloads ---

def headline_event_set(nev: int, shape=(128, 128, 64), h: float = 0.5, seed: int = 2024,
                       model_seed: int = 7):
    """``nev`` events on one shared layered+3% GRF model (3.5->7.5 km/s,
    5 km correlation), each with its own seeded interior point source and
    30-120 seeded surface receivers.  Returns (f, [u0], [g], [src])."""
    v = layered_grf_model(shape, h, 3.5, 7.5, 0.03, 5.0, seed=model_seed)
    f = 1.0 / v
    r = rng(seed)
    u0s, gs, srcs = [], [], []
    for e in range(nev):
        src = tuple(int(r.integers(2, s - 2)) for s in shape)
        nrec = int(r.integers(30, 121))
        u0s.append(point_source_u0(shape, src))
        gs.append(sparse_adjoint_source(shape, nrec, seed=seed * 1000 + e))
        srcs.append(src)
    return f, u0s, gs, srcs


def uneven_box_set(nev: int, h: float = 0.5, seed: int = 3, lo=(48, 48, 24), hi=(128, 128, 64)):
    """Events on boxes with dims drawn uniformly in [lo, hi] per axis, each a
    crop of one shared layered+3% GRF model.  Returns lists (f, u0, g, src)."""
    big = layered_grf_model(hi, h, 3.5, 7.5, 0.03, 5.0, seed=seed + 11)
    r = rng(seed)
    fs, u0s, gs, srcs = [], [], [], []
    for e in range(nev):
        dims = tuple(int(r.integers(lo[a], hi[a] + 1)) for a in range(3))
        off = tuple(int(r.integers(0, hi[a] - dims[a] + 1)) for a in range(3))
        v = big[off[0]:off[0] + dims[0], off[1]:off[1] + dims[1], off[2]:off[2] + dims[2]]
        src = tuple(int(r.integers(1, d - 1)) for d in dims)
        fs.append(np.ascontiguousarray(1.0 / v))
        u0s.append(point_source_u0(dims, src))
        gs.append(sparse_adjoint_source(dims, int(r.integers(30, 121)), seed=seed * 7919 + e))
        srcs.append(src)
    return fs, u0s, gs, srcs
