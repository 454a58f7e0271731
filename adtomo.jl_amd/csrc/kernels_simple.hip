// Hyperplane-ordered fast sweeping: the per-hyperplane ("plain") kernels and
// the one-workgroup-per-event persistent kernels.
//
// A reference sweep (Eikonal3D.cpp:35-57, Eikonal.h:22-52) is a lexicographic
// Gauss-Seidel pass.  Each node reads its upwind neighbours (already updated
// in this sweep) and its downwind neighbours (not yet updated).  Taking the
// hyperplanes a+b+c = s (a,b,c counted along the sweep direction) in
// increasing s keeps exactly those reads, so every node sees the same
// operand values as in the serial loop and the result is bit-identical.
// Nodes of one hyperplane are independent and are updated in parallel.
#include "eikonal_device.hpp"
#include "kernels.hpp"
#include "adjoint_device.hpp"

namespace adt {

// Clears the per-event diagnostics before a launch writes its fields (a
// kernel, not a memset, so that it can sit in a captured graph).
__global__ void info_clear_kernel(adt_info* __restrict__ info, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    info[i] = adt_info{};
}


// Copies between a strided array and its row-major twin: element p of the
// row-major (d0, d1, d2) array is (i, j, k) = (p / (d1 d2), p / d2 mod d1,
// p mod d2) and sits at i*s0 + j*s1 + k*s2 of the strided one.  to_rowmajor
// gathers (dst[p] = src[offset]), else it scatters (dst[offset] = src[p]).
template <typename T>
__global__ void strided_copy_kernel(const T* __restrict__ src, T* __restrict__ dst, int64_t N,
                                    int64_t d1, int64_t d2, int64_t s0, int64_t s1, int64_t s2,
                                    int to_rowmajor) {
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < N;
       p += (int64_t)gridDim.x * blockDim.x) {
    const int64_t q = p / d2, k = p - q * d2;
    const int64_t i = q / d1, j = q - i * d1;
    const int64_t off = i * s0 + j * s1 + k * s2;
    if (to_rowmajor)
      dst[p] = src[off];
    else
      dst[off] = src[p];
  }
}
template __global__ void strided_copy_kernel<double>(const double*, double*, int64_t, int64_t,
                                                     int64_t, int64_t, int64_t, int64_t, int);
template __global__ void strided_copy_kernel<float>(const float*, float*, int64_t, int64_t,
                                                    int64_t, int64_t, int64_t, int64_t, int);


// ------------------------------------------------------------------ 3D ---

template <typename T>
__device__ __forceinline__ void fwd3_node(T* __restrict__ u, const T* __restrict__ f, T h,
                                          int m, int n, int l, int i, int j, int k) {
  const int64_t nl = (int64_t)n * l;
  const int64_t p = (int64_t)i * nl + (int64_t)j * l + k;
  T ux = i == 0 ? u[p + nl] : (i == m - 1 ? u[p - nl] : smin(u[p + nl], u[p - nl]));
  T uy = j == 0 ? u[p + l] : (j == n - 1 ? u[p - l] : smin(u[p + l], u[p - l]));
  T uz = k == 0 ? u[p + 1] : (k == l - 1 ? u[p - 1] : smin(u[p + 1], u[p - 1]));
  T un = solution3d(ux, uy, uz, f[p], h);
  u[p] = smin(un, u[p]);
}

// Range of the first canonical coordinate on hyperplane s.
__device__ __host__ __forceinline__ void plane_a_range(int s, int m, int n, int l, int* lo, int* hi) {
  int a_lo = s - (n - 1) - (l - 1);
  *lo = a_lo > 0 ? a_lo : 0;
  *hi = s < m - 1 ? s : m - 1;
}

template <typename T>
__global__ void fwd3_plane_kernel(T* __restrict__ u, const T* __restrict__ f, T h, int m, int n,
                                  int l, int dir, int s, int a_lo, int a_cnt) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)a_cnt * n) return;
  int a = a_lo + (int)(t / n), b = (int)(t % n), c = s - a - b;
  if (c < 0 || c >= l) return;
  int i = (dir & 1) ? m - 1 - a : a;
  int j = (dir & 2) ? n - 1 - b : b;
  int k = (dir & 4) ? l - 1 - c : c;
  fwd3_node(u, f, h, m, n, l, i, j, k);
}

// Element-wise helpers for the per-launch path.
template <typename T>
__global__ void copy_kernel(T* __restrict__ dst, const T* __restrict__ src, int64_t N) {
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < N;
       p += (int64_t)gridDim.x * blockDim.x)
    dst[p] = src[p];
}

// The stop rule's err as the reference folds it (Eikonal3D.cpp:79-81):
// err = std::max(fabs(a[p] - b[p]), err) over p in memory order, where a
// NaN difference makes err NaN and the next difference replaces it, so err
// is the maximum over p >= *from, *from = 1 + the index of the last NaN
// difference (lastnan_kernel; err is NaN when that is N).
template <typename T>
__global__ void lastnan_kernel(const T* __restrict__ a, const T* __restrict__ b, int64_t N,
                               unsigned long long* last) {
  unsigned long long m = 0;
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < N;
       p += (int64_t)gridDim.x * blockDim.x) {
    const T d = Real<T>::fabs_(a[p] - b[p]);
    if (d != d) m = (unsigned long long)p + 1;
  }
  if (m) atomicMax(last, m);
}

template <typename T>
__global__ void maxdiff_kernel(const T* __restrict__ a, const T* __restrict__ b, int64_t N,
                               const unsigned long long* __restrict__ from,
                               unsigned long long* out) {
  const int64_t p0 = (int64_t)*from;
  uint64_t key = 0;
  for (int64_t p = p0 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < N;
       p += (int64_t)gridDim.x * blockDim.x) {
    uint64_t q = maxkey(Real<T>::fabs_(a[p] - b[p]));
    key = q > key ? q : key;
  }
  for (int off = 32; off > 0; off >>= 1) {
    uint64_t o = __shfl_xor(key, off);
    key = o > key ? o : key;
  }
  if ((threadIdx.x & 63) == 0 && key) atomicMax(out, (unsigned long long)key);
}

// ------------------------------------------------------------- 3D adjoint ---
// Gauss-Seidel sweeps of the transposed upwind system, one launch per
// hyperplane (the plain path); codes as described in adjoint_device.hpp.

template <typename T>
__global__ void adj3_prep_kernel(const T* __restrict__ u, const T* __restrict__ u0,
                                 const T* __restrict__ g, int m, int n, int l,
                                 uint8_t* __restrict__ code, T* __restrict__ lam,
                                 T* __restrict__ grad_u0) {
  const int64_t N = (int64_t)m * n * l;
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < N;
       p += (int64_t)gridDim.x * blockDim.x)
    adj3_prep_node(u, u0, g, m, n, l, p, code, lam, grad_u0);
}

// lambda_r = (g_r + sum_c 2(u_c - u_r) lambda_c) / D_r, pulled from the
// neighbours c (order i-1, i+1, j-1, j+1, k-1, k+1) whose argmin neighbour
// along the connecting dimension is r; D_r sums 2(u_r - umin_d) over the
// active dimensions of r in x, y, z order.  Returns true if lambda_r changed.
template <typename T>
__device__ __forceinline__ bool adj3_node(const T* __restrict__ u, const T* __restrict__ g,
                                          const uint8_t* __restrict__ code, T* __restrict__ lam,
                                          int m, int n, int l, int i, int j, int k) {
  const int64_t nl = (int64_t)n * l;
  const int64_t p = (int64_t)i * nl + (int64_t)j * l + k;
  const uint8_t cr = code[p];
  if (cr & 0x40) return false;
  const T ur = u[p];
  T num = g[p];
  const int co[3] = {i, j, k};
  const int ex[3] = {m, n, l};
  const int64_t st[3] = {nl, l, 1};
  T D = T(0);
  bool have = false;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    // neighbour at -1 contributes if its argmin along d is at +1 (i.e. r)
    if (co[d] > 0) {
      const int64_t c = p - st[d];
      const uint8_t cc = code[c];
      if (!(cc & 0x40) && ((cc >> (2 * d)) & 3) == 2) num = num + (T(2.0) * (u[c] - ur)) * lam[c];
    }
    if (co[d] < ex[d] - 1) {
      const int64_t c = p + st[d];
      const uint8_t cc = code[c];
      if (!(cc & 0x40) && ((cc >> (2 * d)) & 3) == 1) num = num + (T(2.0) * (u[c] - ur)) * lam[c];
    }
    const int sd = (cr >> (2 * d)) & 3;
    if (sd) {
      const T t = T(2.0) * (ur - u[sd == 1 ? p - st[d] : p + st[d]]);
      D = have ? D + t : t;
      have = true;
    }
  }
  const T v = num / D;
  if (bits_of(v) != bits_of(lam[p])) {
    lam[p] = v;
    return true;
  }
  return false;
}

template <typename T>
__global__ void adj3_plane_kernel(const T* __restrict__ u, const T* __restrict__ g,
                                  const uint8_t* __restrict__ code, T* __restrict__ lam, int m,
                                  int n, int l, int dir, int s, int a_lo, int a_cnt,
                                  int* changed) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  bool ch = false;
  if (t < (int64_t)a_cnt * n) {
    int a = a_lo + (int)(t / n), b = (int)(t % n), c = s - a - b;
    if (c >= 0 && c < l) {
      int i = (dir & 1) ? m - 1 - a : a;
      int j = (dir & 2) ? n - 1 - b : b;
      int k = (dir & 4) ? l - 1 - c : c;
      ch = adj3_node(u, g, code, lam, m, n, l, i, j, k);
    }
  }
  if (__any(ch) && (threadIdx.x & 63) == 0) *changed = 1;
}

template <typename T>
__global__ void adj3_finish_kernel(const T* __restrict__ lam, const T* __restrict__ f, T h,
                                   int64_t N, T* __restrict__ grad_f) {
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < N;
       p += (int64_t)gridDim.x * blockDim.x) {
    const T rhs = T(-2.0) * f[p] * h * h;
    grad_f[p] = -lam[p] * rhs;
  }
}

// ------------------------------------------------------------------ 2D ---
// Flat index j*(m+1)+i, i in [0,m], j in [0,n] (Eikonal.h:22-52).

template <typename T>
__device__ __forceinline__ void fwd2_node(T* __restrict__ u, const T* __restrict__ f, T h, int m,
                                          int n, int i, int j) {
  const int64_t W = (int64_t)m + 1;
  T a, b;
  if (i == 0)
    a = u[j * W + 1];
  else if (i == m)
    a = u[j * W + m - 1];
  else
    a = smin(u[j * W + i + 1], u[j * W + i - 1]);
  if (j == 0)
    b = u[W + i];
  else if (j == n)
    b = u[(int64_t)(n - 1) * W + i];
  else
    b = smin(u[(int64_t)(j - 1) * W + i], u[(int64_t)(j + 1) * W + i]);
  const int64_t p = (int64_t)j * W + i;
  T un = solution2d(a, b, f[p], h);
  u[p] = smin(u[p], un);
}

template <typename T>
__device__ void fwd2_iteration_wg(T* __restrict__ u, const T* __restrict__ f, T h, int m, int n,
                                  int ix, int jx) {
  for (int sw = 0; sw < 4; ++sw) {
    const int dir = kDir2[sw];
    for (int s = 0; s <= m + n; ++s) {
      const int a_lo = s - n > 0 ? s - n : 0, a_hi = s < m ? s : m;
      for (int a = a_lo + threadIdx.x; a <= a_hi; a += blockDim.x) {
        const int b = s - a;
        const int i = (dir & 1) ? m - a : a;
        const int j = (dir & 2) ? n - b : b;
        if (i == ix && j == jx) continue;
        fwd2_node(u, f, h, m, n, i, j);
      }
      __syncthreads();
    }
  }
}

template <typename T>
__device__ T wg_sum(T v, T* red) {
  // Deterministic: fixed xor-butterfly in each wave, then waves in order.
  for (int off = 32; off > 0; off >>= 1) v = v + __shfl_xor(v, off);
  const int w = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[w] = v;
  __syncthreads();
  T r = T(0);
  for (int q = 0; q < nw; ++q) r = r + red[q];
  return r;
}

// Persistent 2D forward (Eikonal.h:54-93): init 1e5 / 0 at the source, up to
// 100 iterations of 4 sweeps, stop when ||u - u_old|| / ||u_old|| < 1e-8.
template <typename T>
__global__ void __launch_bounds__(1024) fwd2_event_kernel(const Event2Desc* __restrict__ evs,
                                                          int max_iter) {
  __shared__ T red[16];
  const Event2Desc& ev = evs[blockIdx.x];
  const int m = ev.m, n = ev.n, ix = ev.ix, jx = ev.jx;
  const int64_t W = (int64_t)m + 1, N = W * (n + 1);
  T* u = (T*)ev.u;
  T* uold = (T*)ev.uold;
  const T* f = (const T*)ev.f;
  const T h = (T)ev.h;
  for (int64_t p = threadIdx.x; p < N; p += blockDim.x) {
    const T v = (p == (int64_t)jx * W + ix) ? T(0.0) : T(100000.0);
    u[p] = v;
    uold[p] = v;
  }
  __syncthreads();
  int it;
  bool conv = false;
  T err = T(0);
  for (it = 0; it < max_iter; ++it) {
    fwd2_iteration_wg(u, f, h, m, n, ix, jx);
    T sd = T(0), so = T(0);
    for (int64_t p = threadIdx.x; p < N; p += blockDim.x) {
      const T d = u[p] - uold[p];
      sd = sd + d * d;
      so = so + uold[p] * uold[p];
    }
    sd = wg_sum(sd, red);
    so = wg_sum(so, red);
    err = Real<T>::sqrt_(sd) / Real<T>::sqrt_(so);
    if (threadIdx.x == 0 && it < kMaxLoggedIters) ev.info->err_log[it] = (double)err;
    if (err < T(1e-8)) {
      conv = true;
      break;
    }
    for (int64_t p = threadIdx.x; p < N; p += blockDim.x) uold[p] = u[p];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    ev.info->iterations = conv ? it + 1 : it;
    ev.info->converged = conv ? 1 : 0;
    ev.info->err = (double)err;
    // The stop rule compares a sum whose rounding depends on the summation
    // order; flag results within 1e-12 (relative) of the threshold.
    const double e = (double)err, thr = (double)T(1e-8);
    ev.info->ambiguous = (e > thr * (1 - 1e-12) && e < thr * (1 + 1e-12)) ? 1 : 0;
  }
}

// 2D upwind codes: bits 0-1 x side, bits 2-3 y side (1: -1, 2: +1), bit 6
// singular (non-source node without an active dimension), bit 7 source.
// ------------------------------------------------------------ self-test ---
// Device arithmetic used on the path, for checks against host IEEE results.
template <typename T>
__global__ void arith_kernel(int op, const T* __restrict__ a, const T* __restrict__ b,
                             T* __restrict__ out, int64_t N) {
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < N;
       p += (int64_t)gridDim.x * blockDim.x) {
    T x = a[p];
    T r;
    switch (op) {
      case 0: r = Real<T>::sqrt_(x); break;
      case 1: r = div3(x); break;  // the division the local solve uses
      case 2: r = x / T(2.0); break;
      case 3: r = x / b[p]; break;
      case 4: r = solution3d(a[3 * p], a[3 * p + 1], a[3 * p + 2], b[2 * p], b[2 * p + 1]); break;
      case 5: r = solution2d(a[2 * p], a[2 * p + 1], b[2 * p], b[2 * p + 1]); break;
      case 7: r = Real<T>::risky_(x) ? Real<T>::sqrt_(x) : Real<T>::refine_(x); break;
      case 8: {
        // inputs the fast sort may order differently are redone, as in the
        // step loops
        bool unsafe = (int)Real<T>::odd_(a[3 * p]) | (int)Real<T>::odd_(a[3 * p + 1]) |
                      (int)Real<T>::odd_(a[3 * p + 2]);
        r = solution3d<T, true>(a[3 * p], a[3 * p + 1], a[3 * p + 2], b[2 * p], b[2 * p + 1],
                                unsafe);
        if (unsafe)
          r = solution3d(a[3 * p], a[3 * p + 1], a[3 * p + 2], b[2 * p], b[2 * p + 1]);
        break;
      }
      default: r = x; break;
    }
    out[p] = r;
  }
}

// ------------------------------------------------------ explicit instances ---
#define ADT_INSTANTIATE(T)                                                                      \
  template __global__ void fwd3_plane_kernel<T>(T*, const T*, T, int, int, int, int, int, int, \
                                                int);                                           \
  template __global__ void copy_kernel<T>(T*, const T*, int64_t);                               \
  template __global__ void lastnan_kernel<T>(const T*, const T*, int64_t, unsigned long long*); \
  template __global__ void maxdiff_kernel<T>(const T*, const T*, int64_t,                       \
                                             const unsigned long long*, unsigned long long*);   \
  template __global__ void adj3_prep_kernel<T>(const T*, const T*, const T*, int, int, int,     \
                                               uint8_t*, T*, T*);                               \
  template __global__ void adj3_plane_kernel<T>(const T*, const T*, const uint8_t*, T*, int,    \
                                                int, int, int, int, int, int, int*);            \
  template __global__ void adj3_finish_kernel<T>(const T*, const T*, T, int64_t, T*);           \
  template __global__ void fwd2_event_kernel<T>(const Event2Desc*, int);                        \
  template __global__ void arith_kernel<T>(int, const T*, const T*, T*, int64_t);

ADT_INSTANTIATE(double)
ADT_INSTANTIATE(float)

}  // namespace adt
