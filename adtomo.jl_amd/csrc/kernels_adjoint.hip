// Adjoint kernels: the 3D topological traversal (device code in
// adjoint3.hpp), one workgroup per event, and the 2D one below.
//
// The reference (Eikonal3D.cpp:96-198, Eikonal.h:95-199) assembles the
// upwind Jacobian A and solves A^T lambda = g with a sparse LU; the upwind
// graph is acyclic, so Kahn's algorithm solves the system exactly once per
// node, pulling each node's feeders in a fixed order.
#include "adjoint3.hpp"
#include "adjoint3_tile.hpp"
#include "adjoint_device.hpp"
#include "kernels.hpp"

namespace adt {

// Decrements byte `idx` of a byte array through a 32-bit atomic on its word;
// returns the byte's previous value.  Counts never go below zero, so no
// borrow crosses into neighbouring bytes.
__device__ __forceinline__ unsigned dec_byte(uint8_t* base, int64_t idx) {
  unsigned* w = (unsigned*)(base + (idx & ~int64_t(3)));
  const int sh = int(idx & 3) * 8;
  const unsigned old = atomicSub(w, 1u << sh);
  return (old >> sh) & 0xffu;
}

// ----------------------------------------------------------------- 3D ---

// Each adjoint kernel takes the events whose launch setting names its
// traversal and leaves the others to the other kernel.
template <typename T>
__global__ void __launch_bounds__(1024) adj3_kahn_kernel(const EventDesc* __restrict__ evs) {
  const EventDesc& ev = evs[blockIdx.x];
  if (ev.adj_algo != 1) return;
  adj3_kahn_event<T, 2, 2>(ev);
}

template <typename T, int NW, int MAXT>
__global__ void __launch_bounds__(NW * 64) adj3_tile_kernel(const EventDesc* __restrict__ evs) {
  const EventDesc& ev = evs[blockIdx.x];
  if (ev.adj_algo != 0) return;
  adj3_tile_event<T, NW, MAXT>(ev);
}

// ----------------------------------------------------------------- 2D ---

template <typename T>
__global__ void __launch_bounds__(1024) adj2_kahn_kernel(const Event2Desc* __restrict__ evs) {
  __shared__ int qtail;
  __shared__ int nact;
  const Event2Desc& ev = evs[blockIdx.x];
  const int m = ev.m, n = ev.n;
  const int64_t W = (int64_t)m + 1, N = W * (n + 1);
  const T* __restrict__ u = (const T*)ev.u;
  const T* __restrict__ g = (const T*)ev.grad_u;
  uint8_t* __restrict__ code = ev.code;
  uint8_t* __restrict__ indeg = ev.indeg;
  int32_t* __restrict__ queue = ev.queue;
  T* __restrict__ lam = (T*)ev.lam;
  if (threadIdx.x == 0) {
    qtail = 0;
    nact = 0;
    ev.info->n_singular = 0;
  }
  __syncthreads();
  for (int64_t p = threadIdx.x; p < N; p += blockDim.x)
    adj2_prep_node(u, m, n, ev.ix, ev.jx, p, code, lam, (unsigned long long*)&ev.info->n_singular);
  __syncthreads();
  const int ex[2] = {m, n};
  const int64_t st[2] = {1, W};
  int my_act = 0;
  for (int64_t p = threadIdx.x; p < N; p += blockDim.x) {
    const uint8_t cr = code[p];
    if (cr & 0x40) continue;  // singular: lambda stays 0
    ++my_act;
    const int co[2] = {(int)(p % W), (int)(p / W)};
    int cnt = 0;
#pragma unroll
    for (int d = 0; d < 2; ++d) {
      if (co[d] > 0) {
        const uint8_t cc = code[p - st[d]];
        cnt += (!(cc & 0xC0) && ((cc >> (2 * d)) & 3) == 2) ? 1 : 0;
      }
      if (co[d] < ex[d]) {
        const uint8_t cc = code[p + st[d]];
        cnt += (!(cc & 0xC0) && ((cc >> (2 * d)) & 3) == 1) ? 1 : 0;
      }
    }
    indeg[p] = (uint8_t)cnt;
    if (cnt == 0) queue[atomicAdd(&qtail, 1)] = (int32_t)p;
  }
  atomicAdd(&nact, my_act);
  __syncthreads();
  int head = 0, tail = qtail, levels = 0;
  while (head < tail) {
    for (int idx = head + threadIdx.x; idx < tail; idx += blockDim.x) {
      const int64_t r = queue[idx];
      const uint8_t cr = code[r];
      const int co[2] = {(int)(r % W), (int)(r / W)};
      const T ur = u[r];
      T num = g[r];
      T D = T(0);
      bool have = false;
#pragma unroll
      for (int d = 0; d < 2; ++d) {
        if (co[d] > 0) {
          const int64_t c = r - st[d];
          const uint8_t cc = code[c];
          if (!(cc & 0xC0) && ((cc >> (2 * d)) & 3) == 2)
            num = num + (T(2.0) * (u[c] - ur)) * lam[c];
        }
        if (co[d] < ex[d]) {
          const int64_t c = r + st[d];
          const uint8_t cc = code[c];
          if (!(cc & 0xC0) && ((cc >> (2 * d)) & 3) == 1)
            num = num + (T(2.0) * (u[c] - ur)) * lam[c];
        }
        const int sd = (cr >> (2 * d)) & 3;
        if (sd) {
          const T t = T(2.0) * (ur - u[sd == 1 ? r - st[d] : r + st[d]]);
          D = have ? D + t : t;
          have = true;
        }
      }
      if (cr & 0x80) D = T(1.0);
      lam[r] = num / D;
#pragma unroll
      for (int d = 0; d < 2; ++d) {
        const int sd = (cr >> (2 * d)) & 3;
        if (!sd) continue;
        const int64_t t = sd == 1 ? r - st[d] : r + st[d];
        if (code[t] & 0x40) continue;
        if (dec_byte(indeg, t) == 1u) queue[atomicAdd(&qtail, 1)] = (int32_t)t;
      }
    }
    __syncthreads();
    head = tail;
    tail = qtail;
    ++levels;
    __syncthreads();
  }
  const T h = (T)ev.h;
  const T* __restrict__ f = (const T*)ev.f;
  T* __restrict__ gf = (T*)ev.grad_f;
  const int64_t src = (int64_t)ev.jx * W + ev.ix;
  for (int64_t p = threadIdx.x; p < N; p += blockDim.x) {
    T dFdf = T(-2.0) * f[p] * h * h;
    if (p == src) dFdf = T(0.0);
    gf[p] = -lam[p] * dFdf;
  }
  if (threadIdx.x == 0) {
    ev.info->adj_sweeps = levels;
    ev.info->adj_converged = (tail == nact) ? 1 : 0;
  }
}

template __global__ void adj3_kahn_kernel<double>(const EventDesc*);
template __global__ void adj3_kahn_kernel<float>(const EventDesc*);
template __global__ void adj3_tile_kernel<double, kTileWavesSmall, kTileSmallMax>(const EventDesc*);
template __global__ void adj3_tile_kernel<double, kTileWavesLarge64, kTileLargeMax>(const EventDesc*);
template __global__ void adj3_tile_kernel<float, kTileWavesSmall, kTileSmallMax>(const EventDesc*);
template __global__ void adj3_tile_kernel<float, kTileWavesLarge32, kTileLargeMax>(const EventDesc*);
template __global__ void adj2_kahn_kernel<double>(const Event2Desc*);
template __global__ void adj2_kahn_kernel<float>(const Event2Desc*);

}  // namespace adt
