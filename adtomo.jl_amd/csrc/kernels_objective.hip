// The tomography objective around the Eikonal path (scripts/inversion.jl):
// source initialisation (:48-60), receiver gather + misfit with its sparse
// adjoint source (:64-105), and the model parametrisation + smoothing
// regulariser (:42-43, :107-121).  All on the device, so an inversion step
// needs no host round trip between the solves.
#include "eikonal_device.hpp"
#include "kernels.hpp"
#include "misfit3.hpp"

namespace adt {

// ------------------------------------------------------- source init ---

// u0 = 1000 everywhere, then the 8 grid nodes around a station at 1-based
// coordinates (x, y, z) get dist * h / vel0 * scale, in the reference's
// assignment order (scripts/inversion.jl:48-60; scale = vp/vs for S,
// scripts/inversion_joint.jl:65-73).  One workgroup per station.
template <typename T>
__global__ void __launch_bounds__(256) source_u0_kernel(const SourceDesc* __restrict__ src,
                                                        const T* __restrict__ vel0, int m, int n,
                                                        int l, T h) {
  const SourceDesc& s = src[blockIdx.x];
  T* __restrict__ u0 = (T*)s.u0;
  const int64_t nl = (int64_t)n * l;
  const int64_t N = (int64_t)m * nl;
  for (int64_t p = threadIdx.x; p < N; p += blockDim.x) u0[p] = T(1000);
  __syncthreads();
  if (threadIdx.x == 0) {
    const T x = (T)s.x, y = (T)s.y, z = (T)s.z, sc = (T)s.scale;
    const int xs[2] = {(int)__builtin_ceil(s.x), (int)__builtin_floor(s.x)};
    const int ys[2] = {(int)__builtin_ceil(s.y), (int)__builtin_floor(s.y)};
    const int zs[2] = {(int)__builtin_ceil(s.z), (int)__builtin_floor(s.z)};
    // (u,u,u), (u,u,d), (u,d,u), (u,d,d), (d,u,u), (d,u,d), (d,d,u), (d,d,d)
    for (int c = 0; c < 8; ++c) {
      const int ix = xs[c >> 2], iy = ys[(c >> 1) & 1], iz = zs[c & 1];
      const T dx = x - (T)ix, dy = y - (T)iy, dz = z - (T)iz;
      const int64_t q = (int64_t)(ix - 1) * nl + (int64_t)(iy - 1) * l + (iz - 1);
      const T dist = Real<T>::sqrt_(dx * dx + dy * dy + dz * dz);
      u0[q] = dist * h / vel0[q] * sc;
    }
  }
}

// ------------------------------------------- receiver gather + misfit ---

// One workgroup of 256 threads per solve (station).
template <typename T>
__global__ void __launch_bounds__(256) misfit_kernel(const MisfitDesc* __restrict__ d,
                                                     const MisfitNodes* __restrict__ nodes,
                                                     const double* __restrict__ rx,
                                                     const double* __restrict__ ry,
                                                     const double* __restrict__ rz, int nrec) {
  misfit_event<T>(d[blockIdx.x], nodes[blockIdx.x], rx, ry, rz, nrec);
}

// ------------------------------------------------ model parametrisation ---

// fvar = 2 sigmoid(x) - 1 + vel0, f = 1 / fvar (scripts/inversion.jl:42-43,61).
template <typename T>
__global__ void model_kernel(const T* __restrict__ x, const T* __restrict__ vel0, int64_t N,
                             T* __restrict__ fvar, T* __restrict__ f) {
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < N;
       p += (int64_t)gridDim.x * blockDim.x) {
    const T sg = T(1) / (T(1) + (T)exp(-(double)x[p]));
    const T v = T(2) * sg - T(1) + vel0[p];
    fvar[p] = v;
    if (f) f[p] = T(1) / v;
  }
}

// out = sum over k of in[k] (fixed order), e.g. dL/df over the stations.
template <typename T>
__global__ void sum_arrays_kernel(const T* const* __restrict__ in, int nin, int64_t N,
                                  T* __restrict__ out) {
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < N;
       p += (int64_t)gridDim.x * blockDim.x) {
    T s = in[0][p];
    for (int k = 1; k < nin; ++k) s = s + in[k][p];
    out[p] = s;
  }
}

// Box mean of a over a (2*rh+1)^2 x (2*rv+1) window with periodic wrap
// (the padded VALID conv3d of scripts/inversion.jl:107-118).
template <typename T>
__device__ __forceinline__ T box_mean(const T* __restrict__ a, int i, int j, int k, int m, int n,
                                      int l, int rh, int rv, T wgt) {
  T s = T(0);
  for (int di = -rh; di <= rh; ++di) {
    const int ii = (i + di + m) % m;
    for (int dj = -rh; dj <= rh; ++dj) {
      const int jj = (j + dj + n) % n;
      const int64_t base = ((int64_t)ii * n + jj) * l;
      for (int dk = -rv; dk <= rv; ++dk) s = s + wgt * a[base + (k + dk + l) % l];
    }
  }
  return s;
}

// Pass 1: d = fvar - box(fvar); sgn = sign(d); partial sums of |d| per block.
template <typename T>
__global__ void __launch_bounds__(256) smooth_pass1_kernel(const T* __restrict__ fvar, int m,
                                                           int n, int l, int rh, int rv, T wgt,
                                                           T* __restrict__ sgn,
                                                           T* __restrict__ partial) {
  __shared__ T part[256];
  const int64_t N = (int64_t)m * n * l;
  T acc = T(0);
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < N;
       p += (int64_t)gridDim.x * blockDim.x) {
    const int k = (int)(p % l);
    const int j = (int)((p / l) % n);
    const int i = (int)(p / ((int64_t)n * l));
    const T dv = fvar[p] - box_mean(fvar, i, j, k, m, n, l, rh, rv, wgt);
    acc = acc + (dv < T(0) ? -dv : dv);
    sgn[p] = dv > T(0) ? T(1) : (dv < T(0) ? T(-1) : T(0));
  }
  part[threadIdx.x] = acc;
  __syncthreads();
  for (int off = blockDim.x / 2; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) part[threadIdx.x] = part[threadIdx.x] + part[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = part[0];
}

// Pass 2: grad_fvar = lambda (sgn - box(sgn)) (the box with periodic wrap
// and a symmetric window is its own adjoint); value = lambda * sum(partial).
template <typename T>
__global__ void __launch_bounds__(256) smooth_pass2_kernel(const T* __restrict__ sgn, int m,
                                                           int n, int l, int rh, int rv, T wgt,
                                                           T lambda, T* __restrict__ grad,
                                                           const T* __restrict__ partial,
                                                           int npart, T* __restrict__ value) {
  const int64_t N = grad ? (int64_t)m * n * l : 0;  // value only: no gradient pass
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < N;
       p += (int64_t)gridDim.x * blockDim.x) {
    const int k = (int)(p % l);
    const int j = (int)((p / l) % n);
    const int i = (int)(p / ((int64_t)n * l));
    grad[p] = lambda * (sgn[p] - box_mean(sgn, i, j, k, m, n, l, rh, rv, wgt));
  }
  if (blockIdx.x == 0 && threadIdx.x == 0 && value) {
    T s = T(0);
    for (int b = 0; b < npart; ++b) s = s + partial[b];
    *value = lambda * s;
  }
}

// grad_x = (grad_fvar - grad_f / fvar^2) * 2 sigmoid(x) (1 - sigmoid(x)):
// the chain rule through f = 1/fvar and fvar = 2 sigmoid(x) - 1 + vel0.
template <typename T>
__global__ void model_grad_kernel(const T* __restrict__ x, const T* __restrict__ fvar,
                                  const T* __restrict__ grad_f, const T* __restrict__ grad_fvar,
                                  int64_t N, T* __restrict__ grad_x) {
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < N;
       p += (int64_t)gridDim.x * blockDim.x) {
    const T sg = T(1) / (T(1) + (T)exp(-(double)x[p]));
    T gv = grad_fvar ? grad_fvar[p] : T(0);
    if (grad_f) gv = gv - grad_f[p] / (fvar[p] * fvar[p]);
    grad_x[p] = gv * (T(2) * sg * (T(1) - sg));
  }
}

// f_s = scale / fvar: the S-wave slowness pvs ./ fvar of
// scripts/inversion_joint.jl:79 (one division, as written there).
template <typename T>
__global__ void slowness_kernel(const T* __restrict__ fvar, int64_t N, T scale,
                                T* __restrict__ f_s) {
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < N;
       p += (int64_t)gridDim.x * blockDim.x)
    f_s[p] = scale / fvar[p];
}

// Adjoint of f_s = scale / fvar: grad_fvar += -scale grad_fs / fvar^2, and
// per-block partial sums of d/dscale = sum grad_fs / fvar (fixed tree order).
template <typename T>
__global__ void __launch_bounds__(256) slowness_grad_kernel(const T* __restrict__ fvar, int64_t N,
                                                            T scale, const T* __restrict__ grad_fs,
                                                            T* __restrict__ grad_fvar,
                                                            T* __restrict__ partial) {
  __shared__ T part[256];
  T acc = T(0);
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < N;
       p += (int64_t)gridDim.x * blockDim.x) {
    const T v = fvar[p], g = grad_fs[p];
    if (grad_fvar) grad_fvar[p] = grad_fvar[p] - scale * g / (v * v);
    acc = acc + g / v;
  }
  part[threadIdx.x] = acc;
  __syncthreads();
  for (int off = blockDim.x / 2; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) part[threadIdx.x] = part[threadIdx.x] + part[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = part[0];
}

// out = sum of partial[0..npart) in order (one thread).
template <typename T>
__global__ void sum_partials_kernel(const T* __restrict__ partial, int npart, T* __restrict__ out) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    T s = T(0);
    for (int b = 0; b < npart; ++b) s = s + partial[b];
    *out = s;
  }
}

#define ADT_OBJ_INSTANTIATE(T)                                                                    \
  template __global__ void source_u0_kernel<T>(const SourceDesc*, const T*, int, int, int, T);    \
  template __global__ void misfit_kernel<T>(const MisfitDesc*, const MisfitNodes*, const double*, const double*,      \
                                            const double*, int);                                  \
  template __global__ void model_kernel<T>(const T*, const T*, int64_t, T*, T*);                 \
  template __global__ void sum_arrays_kernel<T>(const T* const*, int, int64_t, T*);              \
  template __global__ void smooth_pass1_kernel<T>(const T*, int, int, int, int, int, T, T*, T*); \
  template __global__ void smooth_pass2_kernel<T>(const T*, int, int, int, int, int, T, T, T*,   \
                                                  const T*, int, T*);                             \
  template __global__ void slowness_kernel<T>(const T*, int64_t, T, T*);                         \
  template __global__ void slowness_grad_kernel<T>(const T*, int64_t, T, const T*, T*, T*);       \
  template __global__ void sum_partials_kernel<T>(const T*, int, T*);                            \
  template __global__ void model_grad_kernel<T>(const T*, const T*, const T*, const T*, int64_t, \
                                                T*);

ADT_OBJ_INSTANTIATE(double)
ADT_OBJ_INSTANTIATE(float)

}  // namespace adt
