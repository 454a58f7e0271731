// Receiver gather + misfit of one solve (scripts/inversion.jl:64-105) as a
// per-workgroup device function: used by misfit_kernel and by the fused
// forward + misfit + adjoint kernel.
#pragma once

#include "kernels.hpp"

namespace adt {

// Travel time at 1-based point (jx, jy, jz) by the reference's trilinear
// interpolation (scripts/inversion.jl:67-92): a dimension whose floor and
// ceil coincide is not interpolated.  The adjoint adds g * dt/du to grad_u,
// applying the weights in reverse order of the interpolation (z, y, x).
template <typename T>
struct Trilinear {
  int64_t q[2][2][2];  // [x][y][z] node indices
  T wx[2], wy[2], wz[2];
  bool ix, iy, iz;     // interpolate along x / y / z
};

template <typename T>
__device__ __forceinline__ Trilinear<T> trilinear_setup(double jx, double jy, double jz, int n,
                                                        int l) {
  Trilinear<T> t;
  const int x1 = (int)__builtin_floor(jx), x2 = (int)__builtin_ceil(jx);
  const int y1 = (int)__builtin_floor(jy), y2 = (int)__builtin_ceil(jy);
  const int z1 = (int)__builtin_floor(jz), z2 = (int)__builtin_ceil(jz);
  const int xs[2] = {x1, x2}, ys[2] = {y1, y2}, zs[2] = {z1, z2};
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b)
      for (int c = 0; c < 2; ++c)
        t.q[a][b][c] = (int64_t)(xs[a] - 1) * n * l + (int64_t)(ys[b] - 1) * l + (zs[c] - 1);
  t.ix = x1 != x2;
  t.iy = y1 != y2;
  t.iz = z1 != z2;
  t.wx[0] = (T)((double)x2 - jx);
  t.wx[1] = (T)(jx - (double)x1);
  t.wy[0] = (T)((double)y2 - jy);
  t.wy[1] = (T)(jy - (double)y1);
  t.wz[0] = (T)((double)z2 - jz);
  t.wz[1] = (T)(jz - (double)z1);
  return t;
}

template <typename T>
__device__ __forceinline__ T trilinear_eval(const Trilinear<T>& t, const T* __restrict__ u) {
  T tx[2][2];  // [y][z]
  for (int b = 0; b < 2; ++b)
    for (int c = 0; c < 2; ++c)
      tx[b][c] = t.ix ? t.wx[0] * u[t.q[0][b][c]] + t.wx[1] * u[t.q[1][b][c]] : u[t.q[0][b][c]];
  T txy[2];
  for (int c = 0; c < 2; ++c) txy[c] = t.iy ? t.wy[0] * tx[0][c] + t.wy[1] * tx[1][c] : tx[0][c];
  return t.iz ? t.wz[0] * txy[0] + t.wz[1] * txy[1] : txy[0];
}

// d t / d u at corner (x side a, y side b, z side c) times the seed g, as
// the chain rule of trilinear_eval gives it: the weights applied in reverse
// order of the interpolation (z, y, x); a dimension that is not
// interpolated has the single corner 0 and no weight.
template <typename T>
__device__ __forceinline__ T trilinear_adjoint_at(const Trilinear<T>& t, T g, int a, int b,
                                                  int c) {
  const T gz = t.iz ? g * t.wz[c] : g;
  const T gy = t.iy ? gz * t.wy[b] : gz;
  return t.ix ? gy * t.wx[a] : gy;
}

// Misfit of one solve (station): loss = sum over receivers with tobs != -1
// of w * rho(tobs - t), rho(r) = r^2 (scripts/inversion.jl:97-105) or, with
// huber_delta > 0, r^2 inside [-delta, delta] and 2 delta |r| - delta^2
// outside (continuous with its derivative), summed per thread in receiver
// order and then over threads in a fixed tree order; grad_u = d loss / d u,
// zero except at the interpolation nodes.
//
// grad_u is deterministic: each receiver's seed g_j = -w_j rho'(r_j) goes
// to nodes.gj, and then one thread per interpolation node sums the node's
// terms in increasing receiver order (nodes lists them so), starting from
// +0, exactly as a sequential loop over the receivers adding into a zeroed
// grad_u does.  A receiver without a pick has seed +0; its terms are +0
// (the weights are >= 0) and leave the sum unchanged (a sum that starts at
// +0 is never -0).
//
// Run by one workgroup of at least 256 threads: every thread zeroes grad_u,
// the first 256 take the receivers (receiver j on thread j mod 256), so the
// loss is summed in the same order whatever the workgroup size.
template <typename T>
__device__ __forceinline__ void misfit_event(const MisfitDesc& s, const MisfitNodes& nodes,
                                             const double* __restrict__ rx,
                                             const double* __restrict__ ry,
                                             const double* __restrict__ rz, int nrec) {
  constexpr int kRecThreads = 256;
  __shared__ T part[kRecThreads];
  const T* __restrict__ u = (const T*)s.u;
  const T* __restrict__ tobs = (const T*)s.tobs;
  const T* __restrict__ w = (const T*)s.weight;
  T* grad_u = (T*)s.grad_u;
  T* tcal = (T*)s.tcal;
  T* gj = (T*)nodes.gj;
  const T delta = (T)s.huber_delta;
  const int64_t N = (int64_t)s.m * s.n * s.l;
  if (grad_u)
    for (int64_t p = threadIdx.x; p < N; p += blockDim.x) grad_u[p] = T(0);
  __syncthreads();
  T acc = T(0);
  const int rt = threadIdx.x < kRecThreads ? (int)threadIdx.x : nrec;  // others take none
  for (int j = rt; j < nrec; j += kRecThreads) {
    const Trilinear<T> tr = trilinear_setup<T>(rx[j], ry[j], rz[j], s.n, s.l);
    const T t = trilinear_eval(tr, u);
    if (tcal) tcal[j] = t;
    const T ob = tobs[j];
    T seed = T(0);
    if (ob != T(-1)) {
      const T r = ob - t;
      T rho, drho;  // rho(r) and d rho / d r
      if (delta > T(0) && (r > delta || r < -delta)) {
        const T ar = r < T(0) ? -r : r;
        rho = T(2.0) * delta * ar - delta * delta;
        drho = r < T(0) ? -(T(2.0) * delta) : T(2.0) * delta;
      } else {
        rho = r * r;
        drho = T(2.0) * r;
      }
      acc = acc + w[j] * rho;
      seed = -(w[j] * drho);
    }
    if (grad_u) gj[j] = seed;
  }
  if (threadIdx.x < kRecThreads) part[threadIdx.x] = acc;
  __syncthreads();  // seeds (and the zeroed grad_u) before the node sums
  if (grad_u) {
    for (int64_t q = threadIdx.x; q < nodes.nuniq; q += blockDim.x) {
      T sum = T(0);
      for (int64_t e = nodes.off[q]; e < nodes.off[q + 1]; ++e) {
        const int64_t en = nodes.ent[e];
        const int64_t j = en >> 3;
        const Trilinear<T> tr = trilinear_setup<T>(rx[j], ry[j], rz[j], s.n, s.l);
        sum = sum + trilinear_adjoint_at(tr, gj[j], (int)(en >> 2) & 1, (int)(en >> 1) & 1,
                                         (int)en & 1);
      }
      grad_u[nodes.node[q]] = sum;
    }
  }
  for (int off = kRecThreads / 2; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) part[threadIdx.x] = part[threadIdx.x] + part[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) *(T*)s.loss = part[0];
}

}  // namespace adt
