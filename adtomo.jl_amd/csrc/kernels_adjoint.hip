// Adjoint of the fast-sweeping forward as an exact topological traversal.
//
// The reference (Eikonal3D.cpp:96-198, Eikonal.h:95-199) assembles the
// upwind Jacobian A and solves A^T lambda = g with a sparse LU.  Every
// off-diagonal entry of row c points to c's argmin neighbour r along an
// active dimension, whose u is strictly smaller, so the graph "lambda_r needs
// lambda_c" is acyclic.  Kahn's algorithm over that graph solves the system
// exactly once per node: a node is computed as soon as all nodes feeding it
// are final, with
//     lambda_r = (g_r + sum_c 2(u_c - u_r) lambda_c) / D_r
// pulled from its feeding neighbours in the fixed order (i-1, i+1, j-1, j+1,
// k-1, k+1), D_r = sum over the active dimensions of r (x, y, z order) of
// 2(u_r - umin_d).  The result does not depend on the traversal order.
//
// One workgroup per event.  A node enters the queue once; the levels of the
// traversal are separated by workgroup barriers.
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

// Per-node record of the 3D traversal: everything a node reads from a
// neighbour (u, lambda, upwind code) sits in one cache line.
//   meta bits 0-7: upwind code (see adjoint_device.hpp), bits 8-15: number
//   of feeding neighbours not yet final.
template <typename T> struct Rec;
template <> struct alignas(32) Rec<double> {
  double u;
  double lam;
  uint32_t meta;
  uint32_t pad;
  double g;
};
template <> struct alignas(16) Rec<float> {
  float u;
  float lam;
  uint32_t meta;
  float g;
};

// Upwind code of node p (Eikonal3D.cpp:105-171): 0x40 for Z, else the
// argmin side of every active dimension.
template <typename T>
__device__ __forceinline__ uint32_t code3(gptr<const T> u, gptr<const T> u0, int m, int n, int l,
                                         int64_t p) {
  const int64_t nl = (int64_t)n * l;
  const int co[3] = {(int)(p / nl), (int)((p / l) % n), (int)(p % l)};
  const int ex[3] = {m, n, l};
  const int64_t st[3] = {nl, l, 1};
  const T up = u[p];
  if (up == u0[p]) return 0x40u;
  uint32_t cd = 0;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    T umin;
    uint32_t side;
    if (co[d] == 0) {
      umin = u[p + st[d]];
      side = 2;
    } else if (co[d] == ex[d] - 1) {
      umin = u[p - st[d]];
      side = 1;
    } else {
      const T vp = u[p + st[d]], vm = u[p - st[d]];
      umin = smin(vp, vm);
      side = (vp > vm) ? 1u : 2u;
    }
    if (up > umin) cd |= side << (2 * d);
  }
  return cd ? cd : 0x40u;
}

// Values node r reads to resolve itself: its own record and its six
// neighbours' (u, lambda, code).  Loaded for several nodes before any is
// resolved, so the loads overlap.
template <typename T>
struct Gathered {
  int64_t r;
  bool valid;
  int co[3];
  T ur, g;
  uint32_t cr;
  T nu[6], nlam[6];
  uint32_t nmeta[6];
};

template <typename T>
__device__ __forceinline__ void gather3(gptr<const Rec<T>> rec, int m, int n, int l, int64_t r,
                                        bool valid, Gathered<T>& G) {
  const int64_t nl = (int64_t)n * l;
  G.r = valid ? r : 0;
  G.valid = valid;
  r = G.r;
  G.co[0] = (int)(r / nl);
  G.co[1] = (int)((r / l) % n);
  G.co[2] = (int)(r % l);
  const int ex[3] = {m, n, l};
  const int64_t st[3] = {nl, l, 1};
  G.ur = rec[r].u;
  G.g = rec[r].g;
  G.cr = rec[r].meta & 0xffu;
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    const int d = q >> 1;
    const bool plus = q & 1;
    const bool in = plus ? (G.co[d] < ex[d] - 1) : (G.co[d] > 0);
    const int64_t c = in ? (plus ? r + st[d] : r - st[d]) : r;
    G.nu[q] = rec[c].u;
    G.nlam[q] = rec[c].lam;
    const uint32_t mc = rec[c].meta & 0xffu;
    G.nmeta[q] = in ? mc : 0x40u;
  }
}

// Resolves a gathered node: lambda_r from its feeding neighbours (all
// final) is stored, and the node's upwind neighbours are released: old[d]
// receives the previous unresolved count of the upwind neighbour along d
// (0 if none was released), tgt[d] its index.
template <typename T>
__device__ __forceinline__ void resolve3(gptr<Rec<T>> rec, int n, int l, const Gathered<T>& G,
                                         uint32_t (&old)[3], int64_t (&tgt)[3]) {
  const int64_t nl = (int64_t)n * l;
  const int64_t st[3] = {nl, l, 1};
  const int64_t r = G.r;
  T num = G.g;
  T D = T(0);
  bool have = false;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    // neighbour at -1 feeds r if its argmin along d is at +1, and vice versa
    if (!(G.nmeta[2 * d] & 0x40u) && ((G.nmeta[2 * d] >> (2 * d)) & 3u) == 2u)
      num = num + (T(2.0) * (G.nu[2 * d] - G.ur)) * G.nlam[2 * d];
    if (!(G.nmeta[2 * d + 1] & 0x40u) && ((G.nmeta[2 * d + 1] >> (2 * d)) & 3u) == 1u)
      num = num + (T(2.0) * (G.nu[2 * d + 1] - G.ur)) * G.nlam[2 * d + 1];
    const uint32_t sd = (G.cr >> (2 * d)) & 3u;
    if (sd) {
      const T t = T(2.0) * (G.ur - G.nu[sd == 1 ? 2 * d : 2 * d + 1]);
      D = have ? D + t : t;
      have = true;
    }
  }
  if (G.valid) rec[r].lam = num / D;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const uint32_t sd = (G.cr >> (2 * d)) & 3u;
    const int q = sd == 1 ? 2 * d : 2 * d + 1;
    const bool rel = G.valid && sd && !(G.nmeta[q] & 0x40u);
    tgt[d] = sd == 1 ? r - st[d] : r + st[d];
    old[d] = rel ? __hip_atomic_fetch_sub(&rec[tgt[d]].meta, 1u << 8, __ATOMIC_RELAXED,
                                          __HIP_MEMORY_SCOPE_AGENT)
                 : 0u;
  }
}

__device__ __forceinline__ void push_ready(const uint32_t (&old)[3], const int64_t (&tgt)[3],
                                           gptr<int32_t> queue, int* qtail) {
#pragma unroll
  for (int d = 0; d < 3; ++d)
    if (((old[d] >> 8) & 0xffu) == 1u) queue[atomicAdd(qtail, 1)] = (int32_t)tgt[d];
}

template <typename T>
__global__ void __launch_bounds__(1024) adj3_kahn_kernel(const EventDesc* __restrict__ evs) {
  __shared__ int qtail;
  __shared__ int nonz;
  const EventDesc& ev = evs[blockIdx.x];
  const int m = ev.m, n = ev.n, l = ev.l;
  const int64_t nl = (int64_t)n * l;
  const int64_t N = (int64_t)m * nl;
  const gptr<const T> u = (gptr<const T>)ev.u;
  const gptr<const T> u0 = (gptr<const T>)ev.u0;
  const gptr<const T> g = (gptr<const T>)ev.grad_u;
  const gptr<T> gu0 = (gptr<T>)ev.grad_u0;
  const gptr<int32_t> queue = (gptr<int32_t>)ev.queue;
  const gptr<Rec<T>> rec = (gptr<Rec<T>>)ev.lam;
  if (threadIdx.x == 0) {
    qtail = 0;
    nonz = 0;
  }
  for (int64_t p = threadIdx.x; p < N; p += blockDim.x) {
    const T gp = g[p];
    if (gu0) gu0[p] = (u[p] == u0[p]) ? gp : T(0);  // Eikonal3D.cpp:105-110
    rec[p].u = u[p];
    rec[p].lam = T(0);
    rec[p].meta = code3<T>(u, u0, m, n, l, p);
    rec[p].g = gp;
  }
  __syncthreads();
  const int ex[3] = {m, n, l};
  const int64_t st[3] = {nl, l, 1};
  int my_nonz = 0;
  for (int64_t p = threadIdx.x; p < N; p += blockDim.x) {
    const uint32_t cr = rec[p].meta;
    if (cr & 0x40u) continue;
    ++my_nonz;
    const int co[3] = {(int)(p / nl), (int)((p / l) % n), (int)(p % l)};
    uint32_t cnt = 0;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      if (co[d] > 0) {
        const uint32_t cc = rec[p - st[d]].meta;
        cnt += (!(cc & 0x40u) && ((cc >> (2 * d)) & 3u) == 2u) ? 1u : 0u;
      }
      if (co[d] < ex[d] - 1) {
        const uint32_t cc = rec[p + st[d]].meta;
        cnt += (!(cc & 0x40u) && ((cc >> (2 * d)) & 3u) == 1u) ? 1u : 0u;
      }
    }
    rec[p].meta = cr | (cnt << 8);
    if (cnt == 0) queue[atomicAdd(&qtail, 1)] = (int32_t)p;
  }
  atomicAdd(&nonz, my_nonz);
  __syncthreads();
  int head = 0, tail = qtail, levels = 0;
  const int bs = blockDim.x;
  while (head < tail) {
    for (int base = head; base < tail; base += 2 * bs) {
      // two queue entries per thread, resolved with their loads overlapped
      const int i0 = base + threadIdx.x, i1 = base + bs + threadIdx.x;
      const bool v0 = i0 < tail, v1 = i1 < tail;
      const int64_t r0 = queue[v0 ? i0 : head], r1 = queue[v1 ? i1 : head];
      Gathered<T> G0, G1;
      gather3((gptr<const Rec<T>>)rec, m, n, l, r0, v0, G0);
      gather3((gptr<const Rec<T>>)rec, m, n, l, r1, v1, G1);
      uint32_t o0[3], o1[3];
      int64_t t0[3], t1[3];
      resolve3(rec, n, l, G0, o0, t0);
      resolve3(rec, n, l, G1, o1, t1);
      push_ready(o0, t0, queue, &qtail);
      push_ready(o1, t1, queue, &qtail);
    }
    __syncthreads();
    head = tail;
    tail = qtail;
    ++levels;
    __syncthreads();
  }
  const T h = (T)ev.h;
  const gptr<const T> f = (gptr<const T>)ev.f;
  const gptr<T> gf = (gptr<T>)ev.grad_f;
  for (int64_t p = threadIdx.x; p < N; p += blockDim.x) {
    const T rhs = T(-2.0) * f[p] * h * h;  // Eikonal3D.cpp:113-116, 194-196
    gf[p] = -rec[p].lam * rhs;
  }
  if (threadIdx.x == 0) {
    ev.info->adj_sweeps = levels;
    ev.info->adj_converged = (tail == nonz) ? 1 : 0;
  }
}

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
template __global__ void adj2_kahn_kernel<double>(const Event2Desc*);
template __global__ void adj2_kahn_kernel<float>(const Event2Desc*);

}  // namespace adt
