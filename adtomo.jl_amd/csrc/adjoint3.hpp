// 3D adjoint of one event as an exact topological traversal (the device
// code of adj3_kahn_kernel, and of the fused forward + adjoint kernel).
//
// The reference (Eikonal3D.cpp:96-198) assembles the upwind Jacobian A and
// solves A^T lambda = g with a sparse LU.  Every off-diagonal entry of row c
// points to c's argmin neighbour r along an active dimension, whose u is
// strictly smaller, so the graph "lambda_r needs lambda_c" is acyclic.
// Kahn's algorithm over that graph solves the system exactly once per node:
// a node is computed as soon as all nodes feeding it are final, with
//     lambda_r = (g_r + sum_c 2(u_c - u_r) lambda_c) / D_r
// pulled from its feeding neighbours in the fixed order (i-1, i+1, j-1, j+1,
// k-1, k+1), D_r = sum over the active dimensions of r (x, y, z order) of
// 2(u_r - umin_d).  The result does not depend on the traversal order.
//
// One workgroup per event.  A node enters the queue once; the levels of the
// traversal are separated by workgroup barriers.
#pragma once

#include "eikonal_device.hpp"
#include "kernels.hpp"

namespace adt {

// ----------------------------------------------------------------- 3D ---
//
// Per-node state of the 3D traversal (built by a coalesced prep pass):
//   rec[p] = {lambda, w_x, w_y, w_z}: lambda starts as g (0 on Z) and is
//            replaced by the solution; w_d = 2 (u_p - u_parent_d) is p's
//            coefficient toward its upwind neighbour along an active
//            dimension d (0 otherwise).  A node r resolves as
//              lambda_r = (g_r + sum_{feeders c} w_{c,d} lambda_c) / D_r,
//              D_r = w_x + w_y + w_z over r's active dimensions,
//            reading only its own record and its feeders' records.
//   meta[p] (16 bits): bits 0-5 upwind code (2 bits per dimension: 0
//            inactive, 1 argmin at -1, 2 argmin at +1), bit 6 Z, bits 7-12
//            feeder mask in neighbour order (i-1, i+1, j-1, j+1, k-1, k+1),
//            bits 13-15 number of feeders not yet final.  Z nodes start
//            with 7, so releases (at most 6) never make them ready.
// Releasing a neighbour is one 32-bit atomic on the word holding its meta.
//
// Queue entries: for events of at most 2^20 nodes an entry carries, beside
// the node index (bits 12-31), the node's upwind code (bits 0-5) and feeder
// mask (bits 6-11), both final once the prep passes end, so that resolving
// a node does not read its meta word (one scattered read less per node).
// Larger events store the plain index and read the meta word.
template <typename T>
struct alignas(kAdjRecBytes<T>) Rec {
  T lam;
  T w[3];
};
static_assert(sizeof(Rec<double>) == kAdjRecBytes<double>, "record size");
static_assert(sizeof(Rec<float>) == kAdjRecBytes<float>, "record size");

// Coordinates (i, j, k) of node p, advanced by a fixed stride without
// divisions (a grid-stride loop over the flat index).
struct Walk {
  int co[3];
  int step[3];
  int ex[3];
  Walk() = default;
  __device__ Walk(int64_t p, int64_t stride, int m, int n, int l) {
    const int64_t nl = (int64_t)n * l;
    co[0] = (int)(p / nl);
    co[1] = (int)((p / l) % n);
    co[2] = (int)(p % l);
    step[0] = (int)(stride / nl);
    step[1] = (int)((stride / l) % n);
    step[2] = (int)(stride % l);
    ex[0] = m;
    ex[1] = n;
    ex[2] = l;
  }
  __device__ __forceinline__ void advance() {
    co[2] += step[2];
    int c = co[2] >= ex[2];
    co[2] -= c ? ex[2] : 0;
    co[1] += step[1] + c;
    c = co[1] >= ex[1];
    co[1] -= c ? ex[1] : 0;
    co[0] += step[0] + c;
  }
};

constexpr uint32_t kCntShift = 13;
constexpr uint32_t kZ = 0x40u;
constexpr int64_t kPackedMaxNodes = int64_t(1) << 20;

__device__ __forceinline__ uint32_t meta_shift(int64_t p) { return (p & 1) ? 16u : 0u; }

// Upwind code of a node (Eikonal3D.cpp:105-171) from its value up, its
// initial value u0p and its neighbours nb[2d] (at -1) and nb[2d+1] (at +1)
// along each dimension d: kZ for Z, else the argmin side of every active
// dimension.  On a grid edge only the inner neighbour counts.
template <typename T>
__device__ __forceinline__ uint32_t code3(T up, T u0p, const T (&nb)[6], const int (&co)[3],
                                         const int (&ex)[3]) {
  if (up == u0p) return kZ;
  uint32_t cd = 0;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const T vm = nb[2 * d], vp = nb[2 * d + 1];
    T umin;
    uint32_t side;
    if (co[d] == 0) {
      umin = vp;
      side = 2;
    } else if (co[d] == ex[d] - 1) {
      umin = vm;
      side = 1;
    } else {
      umin = smin(vp, vm);
      side = (vp > vm) ? 1u : 2u;
    }
    if (up > umin) cd |= side << (2 * d);
  }
  return cd ? cd : kZ;
}

// P nodes p0 + k * blockDim.x of the prep passes' grid-stride loop, with
// their coordinates; a node past N is replaced by p0 (its loads stay in
// range, its results are not stored).
template <int P>
struct PrepGroup {
  int64_t q[P];
  bool valid[P];
  int co[P][3];
};

template <int P>
__device__ __forceinline__ PrepGroup<P> prep_group(int64_t p0, int64_t N, Walk (&wk)[P]) {
  PrepGroup<P> G;
#pragma unroll
  for (int k = 0; k < P; ++k) {
    const int64_t pk = p0 + (int64_t)k * blockDim.x;
    G.valid[k] = pk < N;
    G.q[k] = G.valid[k] ? pk : p0;
#pragma unroll
    for (int d = 0; d < 3; ++d) G.co[k][d] = G.valid[k] ? wk[k].co[d] : wk[0].co[d];
  }
#pragma unroll
  for (int k = 0; k < P; ++k) wk[k].advance();
  return G;
}

// Values node r reads to resolve itself, loaded for several nodes before
// any is resolved so that the loads overlap.
template <typename T>
struct Gathered {
  int64_t r;
  bool valid;
  uint32_t mr;
  Rec<T> own;
  T flam[6], fw[6];
};

template <typename T>
__device__ __forceinline__ void gather3(gptr<const Rec<T>> rec, gptr<const uint16_t> meta, int n,
                                        int l, uint32_t e, bool packed, bool valid, Gathered<T>& G) {
  const int64_t nl = (int64_t)n * l;
  const int64_t st[3] = {nl, l, 1};
  const int64_t r = !valid ? 0 : packed ? (int64_t)(e >> 12) : (int64_t)e;
  G.r = r;
  G.valid = valid;
  G.mr = !valid ? 0u : packed ? ((e & 63u) | (((e >> 6) & 63u) << 7)) : (uint32_t)meta[r];
  G.own.lam = rec[r].lam;
#pragma unroll
  for (int d = 0; d < 3; ++d) G.own.w[d] = rec[r].w[d];
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    const int d = q >> 1;
    const bool feeds = (G.mr >> (7 + q)) & 1u;
    const int64_t c = feeds ? ((q & 1) ? r + st[d] : r - st[d]) : r;
    G.flam[q] = feeds ? rec[c].lam : T(0);
    G.fw[q] = feeds ? rec[c].w[d] : T(0);
  }
}

// Resolves a gathered node (Eikonal3D.cpp:149-196 restated as the
// triangular solve): lambda_r is stored and the upwind neighbours are
// released; old[d] receives the released neighbour's previous meta word
// (0 if none), tgt[d] its index.
template <typename T>
__device__ __forceinline__ void resolve3(gptr<Rec<T>> rec, gptr<uint32_t> meta_w, int n, int l,
                                         const Gathered<T>& G, uint32_t (&old)[3],
                                         int64_t (&tgt)[3]) {
  const int64_t nl = (int64_t)n * l;
  const int64_t st[3] = {nl, l, 1};
  const int64_t r = G.r;
  T num = G.own.lam;
#pragma unroll
  for (int q = 0; q < 6; ++q)
    if ((G.mr >> (7 + q)) & 1u) num = num + G.fw[q] * G.flam[q];
  T D = T(0);
  bool have = false;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    if ((G.mr >> (2 * d)) & 3u) {
      D = have ? D + G.own.w[d] : G.own.w[d];
      have = true;
    }
  }
  if (G.valid) rec[r].lam = num / D;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const uint32_t sd = (G.mr >> (2 * d)) & 3u;
    tgt[d] = sd == 1 ? r - st[d] : r + st[d];
    const bool rel = G.valid && sd;
    old[d] = rel ? __hip_atomic_fetch_sub(&meta_w[tgt[d] >> 1], 1u << (kCntShift + meta_shift(tgt[d])),
                                          __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                 : 0u;
  }
}

// Queue entry of node p whose meta is mt (the packed form for small events).
__device__ __forceinline__ uint32_t qentry(int64_t p, uint32_t mt, bool packed) {
  return packed ? ((uint32_t)p << 12) | (mt & 63u) | (((mt >> 7) & 63u) << 6) : (uint32_t)p;
}

__device__ __forceinline__ void push_ready(const uint32_t (&old)[3], const int64_t (&tgt)[3],
                                           gptr<uint32_t> queue, int* qtail, bool packed) {
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const uint32_t mt = old[d] >> meta_shift(tgt[d]);
    if (((mt >> kCntShift) & 7u) == 1u) queue[atomicAdd(qtail, 1)] = qentry(tgt[d], mt, packed);
  }
}

// The whole adjoint of one event, run by one workgroup of any size.  kIlp
// queue entries per thread and round are resolved with their loads in
// flight together (more registers per entry, fewer rounds per level); the
// prep passes take kPrep nodes per thread and iteration, issuing all their
// loads before using any (more requests in flight per wave).
template <typename T, int kIlp, int kPrep>
__device__ __forceinline__ void adj3_kahn_event(const EventDesc& ev) {
  __shared__ int qtail;
  __shared__ int nonz;
  const int m = ev.m, n = ev.n, l = ev.l;
  const int64_t nl = (int64_t)n * l;
  const int64_t N = (int64_t)m * nl;
  const bool packed = N <= kPackedMaxNodes;
  const gptr<const T> u = (gptr<const T>)ev.u;
  const gptr<const T> u0 = (gptr<const T>)ev.u0;
  const gptr<const T> g = (gptr<const T>)ev.grad_u;
  const gptr<T> gu0 = (gptr<T>)ev.grad_u0;
  const gptr<uint32_t> queue = (gptr<uint32_t>)ev.queue;
  const gptr<Rec<T>> rec = (gptr<Rec<T>>)ev.lam;
  const gptr<uint16_t> meta = (gptr<uint16_t>)ev.code;
  const gptr<uint32_t> meta_w = (gptr<uint32_t>)ev.code;
  if (threadIdx.x == 0) {
    qtail = 0;
    nonz = 0;
  }
  // pass 1: upwind codes; grad_u0 = g where u == u0 (Eikonal3D.cpp:105-110).
  // Each thread takes kPrep nodes per iteration and issues all their
  // loads before using any, so that more requests are in flight per wave.
  const int ex[3] = {m, n, l};
  const int64_t st[3] = {nl, l, 1};
  const int64_t stride = kPrep * (int64_t)blockDim.x;
  {
    Walk wk[kPrep];
#pragma unroll
    for (int k = 0; k < kPrep; ++k) wk[k] = Walk(threadIdx.x + k * blockDim.x, stride, m, n, l);
    for (int64_t p0 = threadIdx.x; p0 < N; p0 += stride) {
      const PrepGroup<kPrep> G = prep_group<kPrep>(p0, N, wk);
      T uc[kPrep], u0c[kPrep], gc[kPrep], nb[kPrep][6];
#pragma unroll
      for (int k = 0; k < kPrep; ++k) {
        const int64_t q = G.q[k];
        uc[k] = u[q];
        u0c[k] = u0[q];
        gc[k] = g[q];
#pragma unroll
        for (int d = 0; d < 3; ++d) {  // clamped into the grid; edges use one side
          nb[k][2 * d] = u[G.co[k][d] > 0 ? q - st[d] : q + st[d]];
          nb[k][2 * d + 1] = u[G.co[k][d] < ex[d] - 1 ? q + st[d] : q - st[d]];
        }
      }
#pragma unroll
      for (int k = 0; k < kPrep; ++k) {
        const uint32_t code = code3<T>(uc[k], u0c[k], nb[k], G.co[k], ex);
        if (G.valid[k]) {
          if (gu0) gu0[G.q[k]] = (uc[k] == u0c[k]) ? gc[k] : T(0);
          meta[G.q[k]] = (uint16_t)code;
        }
      }
    }
  }
  __syncthreads();
  // pass 2: feeders, counters, coefficients; nodes without feeders start.
  // Other threads rewrite neighbouring meta entries meanwhile, but only
  // bits 7-15; the code bits read here never change.
  int my_nonz = 0;
  {
    Walk wk[kPrep];
#pragma unroll
    for (int k = 0; k < kPrep; ++k) wk[k] = Walk(threadIdx.x + k * blockDim.x, stride, m, n, l);
    for (int64_t p0 = threadIdx.x; p0 < N; p0 += stride) {
      const PrepGroup<kPrep> G = prep_group<kPrep>(p0, N, wk);
      uint32_t cr[kPrep], nm[kPrep][6];
      T uc[kPrep], gc[kPrep], nu[kPrep][6];
#pragma unroll
      for (int k = 0; k < kPrep; ++k) {
        const int64_t q = G.q[k];
        cr[k] = meta[q];
        uc[k] = u[q];
        gc[k] = g[q];
#pragma unroll
        for (int d = 0; d < 3; ++d) {  // a missing neighbour reads the node itself
          const int64_t lo = G.co[k][d] > 0 ? q - st[d] : q;
          const int64_t hi = G.co[k][d] < ex[d] - 1 ? q + st[d] : q;
          nm[k][2 * d] = meta[lo];
          nm[k][2 * d + 1] = meta[hi];
          nu[k][2 * d] = u[lo];
          nu[k][2 * d + 1] = u[hi];
        }
      }
#pragma unroll
      for (int k = 0; k < kPrep; ++k) {
        uint32_t fm = 0;  // neighbours whose upwind neighbour this node is
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          const uint32_t cm = nm[k][2 * d], cp = nm[k][2 * d + 1];
          if (G.co[k][d] > 0)
            fm |= (!(cm & kZ) && ((cm >> (2 * d)) & 3u) == 2u) ? (1u << (2 * d)) : 0u;
          if (G.co[k][d] < ex[d] - 1)
            fm |= (!(cp & kZ) && ((cp >> (2 * d)) & 3u) == 1u) ? (2u << (2 * d)) : 0u;
        }
        Rec<T> rp;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          const uint32_t sd = (cr[k] >> (2 * d)) & 3u;
          // 2 (u_r - u_argmin): Eikonal3D.cpp:149-166
          rp.w[d] = sd ? T(2.0) * (uc[k] - (sd == 1 ? nu[k][2 * d] : nu[k][2 * d + 1])) : T(0);
        }
        const bool z = cr[k] & kZ;
        rp.lam = z ? T(0) : gc[k];
        if (G.valid[k]) {
          const int64_t p = G.q[k];
          rec[p].lam = rp.lam;
#pragma unroll
          for (int d = 0; d < 3; ++d) rec[p].w[d] = rp.w[d];
          const uint32_t cnt = z ? 7u : (uint32_t)__builtin_popcount(fm);
          const uint32_t mt = (cr[k] & 0x7fu) | (fm << 7) | (cnt << kCntShift);
          meta[p] = (uint16_t)mt;
          if (!z) {
            ++my_nonz;
            if (cnt == 0) queue[atomicAdd(&qtail, 1)] = qentry(p, mt, packed);
          }
        }
      }
    }
  }
  atomicAdd(&nonz, my_nonz);
  __syncthreads();
  int head = 0, tail = qtail, levels = 0;
  const int bs = blockDim.x;
  while (head < tail) {
    for (int base = head; base < tail; base += kIlp * bs) {
      // kIlp queue entries per thread, resolved with their loads overlapped
      Gathered<T> G[kIlp];
#pragma unroll
      for (int k = 0; k < kIlp; ++k) {
        const int i = base + k * bs + threadIdx.x;
        const bool v = i < tail;
        gather3((gptr<const Rec<T>>)rec, (gptr<const uint16_t>)meta, n, l, queue[v ? i : head],
                packed, v, G[k]);
      }
      uint32_t o[kIlp][3];
      int64_t t[kIlp][3];
#pragma unroll
      for (int k = 0; k < kIlp; ++k) resolve3(rec, meta_w, n, l, G[k], o[k], t[k]);
#pragma unroll
      for (int k = 0; k < kIlp; ++k) push_ready(o[k], t[k], queue, &qtail, packed);
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

}  // namespace adt
