"""Writes the golden fixtures of the Eikonal path (run from the repo root).

The reference op cannot run in this image (it needs Eigen, TensorFlow, ADCME
and Julia), and its own tests hold no golden vectors, so these fixtures are
produced by the pinned CPU oracle (oracle/eikonal_oracle.c; see
tests/test_oracle.py for what pins it) on seeded inputs shaped like the
reference's own test cases:

  gradtest2d  deps/CustomOps/Eikonal/gradtest.jl:41-50 (m=60, n=30, h=0.1,
              slow layer f[12:18,:]=10, source (30,3)), loss sum(u^2)
  gradtest3d  deps/CustomOps/Eikonal3D/gradtest.jl:14-23 (21^3, f=1, h=0.01,
              tol 1e-6, source at flat index 10*n*l+10*l+11 1-based)
  config1     2D 101x101, h=1, 6 km/s + 3 Gaussian anomalies, 20 receivers
  small3d     13x11x9 random slowness, random u0 in [0,1000) (large Z set)

Each fixture stores the inputs, the forward output, its iteration count and
the adjoint outputs for a seeded grad_u.
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path[:0] = [ROOT, os.path.join(ROOT, "adtomo.jl_amd")]

from oracle import oracle  # noqa: E402
from adtomo_hip import synth  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def gradtest2d():
    m, n, h = 60, 30, 0.1
    f = np.ones((n + 1, m + 1))
    f[11:18, :] = 10.0
    srcx, srcy = 30, 3
    u, it, err, conv = oracle.eikonal2d(f, srcx, srcy, h)
    g = 2 * u
    gf, lam, ns = oracle.eikonal2d_grad(g, u, f, srcx, srcy, h)
    np.savez_compressed(os.path.join(OUT, "gradtest2d.npz"), f=f, srcx=srcx, srcy=srcy, h=h,
                        u=u, iters=it, converged=conv, grad_u=g, grad_f=gf)


def config1():
    shape = (101, 101)
    h = 1.0
    v = synth.gaussians_2d(shape, h, 6.0, 3, 0.10, 8.0, seed=101)
    f = 1.0 / v
    r = synth.rng(102)
    srcx, srcy = int(r.integers(1, 102)), int(r.integers(30, 102))
    rec = r.choice(101, size=20, replace=False)
    u, it, err, conv = oracle.eikonal2d(f, srcx, srcy, h)
    u_bg, *_ = oracle.eikonal2d(np.full(shape, 1 / 6.0), srcx, srcy, h)
    tobs = u_bg[0, rec] + r.normal(0, 0.01, size=20)
    g = np.zeros(shape)
    g[0, rec] = u[0, rec] - tobs  # d/du of 0.5*sum(r^2)
    gf, lam, ns = oracle.eikonal2d_grad(g, u, f, srcx, srcy, h)
    np.savez_compressed(os.path.join(OUT, "config1_2d.npz"), f=f, srcx=srcx, srcy=srcy, h=h,
                        u=u, iters=it, converged=conv, grad_u=g, grad_f=gf, receivers=rec)


def gradtest3d():
    m = n = l = 21
    u0 = np.full(m * n * l, 1000.0)
    u0[10 * n * l + 10 * l + 11 - 1] = 0.0
    u0 = u0.reshape(m, n, l)
    f = np.ones((m, n, l))
    h, tol = 0.01, 1e-6
    u, it, err = oracle.eikonal3d(u0, f, h, tol)
    g = 2 * u  # loss sum(u^2)
    gu0, gf, lam = oracle.eikonal3d_grad(g, u, u0, f, h)
    np.savez_compressed(os.path.join(OUT, "gradtest3d.npz"), u0=u0, f=f, h=h, tol=tol, u=u,
                        iters=it, err=err, grad_u=g, grad_u0=gu0, grad_f=gf)


def small3d():
    r = synth.rng(303)
    shape = (13, 11, 9)
    u0 = r.uniform(0, 1000, size=shape)
    f = r.uniform(0.5, 1.5, size=shape)
    h, tol = 2.0, 1e-3
    u, it, err = oracle.eikonal3d(u0, f, h, tol)
    g = r.normal(size=shape)
    gu0, gf, lam = oracle.eikonal3d_grad(g, u, u0, f, h)
    np.savez_compressed(os.path.join(OUT, "small3d.npz"), u0=u0, f=f, h=h, tol=tol, u=u,
                        iters=it, err=err, grad_u=g, grad_u0=gu0, grad_f=gf)


if __name__ == "__main__":
    gradtest2d()
    config1()
    gradtest3d()
    small3d()
    print("fixtures written to", OUT)
