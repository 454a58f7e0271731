// 3D adjoint of one event as a tile-ordered topological solve (device code
// of adj3_tile_kernel).
//
// The system is the one of Eikonal3D.cpp:96-198: the reference assembles the
// upwind Jacobian and solves A^T lambda = g with a sparse LU.  Every edge of
// the upwind graph (node c "feeds" node r when r is c's argmin neighbour
// along one of c's active dimensions) goes from a larger to a smaller travel
// time, so A is triangular in decreasing u and each node resolves once all
// its feeders have:
//     lambda_r = (g_r + sum_c 2 (u_c - u_r) lambda_c) / D_r,
//     D_r = sum over r's active dimensions (x, y, z order) of 2 (u_r - u_up),
// the feeders pulled in the fixed order (i-1, i+1, j-1, j+1, k-1, k+1).
// The result does not depend on the order in which nodes are resolved.
//
// Instead of one global traversal (levels of scattered single-node
// transactions), the event is cut into 8x8x8 tiles.  A wave visits one tile
// at a time: it loads the tile with a one-node halo into its own LDS boxes
// (coalesced 64-byte rows), resolves every node of the tile that can be
// resolved given the multipliers already final outside the tile (a
// topological traversal in LDS), and writes the multipliers and dL/df of the
// nodes it resolved.  Tiles are scheduled by a per-tile count of the edges
// that enter them from unresolved nodes of other tiles: a tile whose count
// reaches zero resolves completely in one visit.  Tiles whose external
// feeders depend on each other in a cycle (the node graph is acyclic, the
// tile graph need not be) are marked dirty whenever one of their external
// feeders resolves; a wave that finds no ready tile visits a dirty one as
// it is.  That always leaves some work reachable: a tile that is neither
// ready nor dirty nor being visited has had a visit since its last
// delivery, which resolved all it could then, and the unresolved node of
// largest u (all its feeders final) lies in a tile that has not.
//
// Multipliers flow from large to small travel times, so tiles are taken in
// decreasing order of the largest travel time at their corners (ranked
// once, by a counting sort over 1024 bins): the bitmaps are indexed by
// that rank and a wave takes the ready (else dirty) tile of lowest rank.
// Most tiles then resolve in a single visit.
//
// One workgroup per event: the scheduling state (bitmaps, counters) lives
// in LDS and global memory is shared only among the waves of one compute
// unit, so workgroup-scope release/acquire order the multiplier stores of a
// visit before the loads of the visits they enable.
#pragma once

#include "adjoint3.hpp"
#include "buffer_ops.hpp"
#include "eikonal_device.hpp"
#include "kernels.hpp"

namespace adt {
namespace tile {

constexpr int kT = 8;                  // tile edge (nodes)
constexpr int kTN = kT * kT * kT;      // nodes per tile
constexpr int kB = kT + 2;             // box edge with a one-node halo
constexpr int kBox = kB * kB * kB;     // box entries
constexpr int kBI = kB * kB, kBJ = kB;  // box strides (k stride 1)
constexpr uint8_t kCZ = 0x40;          // code flag: Z (lambda = 0, no upwind); in
                                       // LDS also nodes outside the grid, which
                                       // like Z nodes feed and release nothing
constexpr uint8_t kCF = 0x80;          // code flag: lambda final (set by the visit
                                       // that resolves the node)
// Per-node record of a visit (16 bits): bits 0-3 the number of feeders not
// yet final, 4-9 the feeder mask (neighbour order i-1, i+1, j-1, j+1, k-1,
// k+1), 10-12 which upwind neighbours (x, y, z) are present and not Z (they
// are released when the node resolves), 14 not part of the visit (Z,
// absent, or final before it), 15 resolved in this visit.
constexpr uint32_t kPCount = 0xFu;
constexpr int kPFeedShift = 4, kPRelShift = 10;
constexpr uint32_t kPNA = 1u << 14, kPDone = 1u << 15;

// LDS of one wave: the tile and its halo (u, multipliers, codes), the
// pending-feeder count of each tile node and the traversal queue.
template <typename T>
struct WaveBuf {
  T u[kBox];
  T lam[kBox];
  uint8_t code[kBox];
  uint16_t pend[kTN];
  uint16_t queue[kTN];
};

// Scheduling state of the workgroup (LDS): five bitmaps over the tiles and
// a few counters.
enum { kReady = 0, kDirty = 1, kBusy = 2, kDone = 3, kVisited = 4, kMaps = 5 };
struct Sched {
  int remaining;   // non-Z nodes not final
  int active;      // waves holding a tile (diagnostics)
  int fault;       // a bounded wait gave up
  int visits;      // tile visits (diagnostics)
  unsigned kmin, kmax;  // order-preserving bits of the smallest / largest tile key
};
constexpr int kBins = 1024;  // counting-sort bins of the tile order (+1 for done tiles)

// Bounded idle polling of a wave: past this many polls without any node
// resolving, the kernel reports a fault instead of waiting on.
constexpr int kIdleCap = 1 << 22;
constexpr int kVisitCapPerTile = 64;

// Per-event constants.
template <typename T>
struct Ev {
  int m, n, l;
  int nti, ntj, ntk, ntiles, nwords;
  int64_t sI;  // n * l
  T h;
  gptr<const T> u, u0, g, f;
  gptr<T> gu0, gf, lam;
  gptr<uint8_t> code;
  // per tile (ntiles each): count of unresolved external feeding edges (-1:
  // no non-Z node), rank in the visiting order and its inverse, and the
  // ordering key (largest travel time at the tile's corners)
  gptr<int> ext, t2r, r2t;
  gptr<float> key;
};

// Node of slot s (0..13) of a lane in the tile box: slots 0-7 are the tile
// slabs ii = s, 8/9 the i faces (ii = -1 / 8), 10/11 the j faces, 12/13 the
// k faces.  Returns the box index; (gi, gj, gk) the grid coordinates.
__device__ __forceinline__ int slot_node(int s, int lane, int i0, int j0, int k0, int& gi,
                                         int& gj, int& gk) {
  const int a = lane >> 3, b = lane & 7;
  int ii, jj, kk;
  if (s < 8) {
    ii = s; jj = a; kk = b;
  } else if (s < 10) {
    ii = s == 8 ? -1 : kT; jj = a; kk = b;
  } else if (s < 12) {
    ii = a; jj = s == 10 ? -1 : kT; kk = b;
  } else {
    ii = a; jj = b; kk = s == 12 ? -1 : kT;
  }
  gi = i0 + ii;
  gj = j0 + jj;
  gk = k0 + kk;
  return (ii + 1) * kBI + (jj + 1) * kBJ + (kk + 1);
}

__device__ __forceinline__ int box_of_id(int id) {
  return ((id >> 6) + 1) * kBI + (((id >> 3) & 7) + 1) * kBJ + (id & 7) + 1;
}

// Neighbour q (i-1, i+1, j-1, j+1, k-1, k+1) of a box index.
__device__ __forceinline__ int box_nb(int b, int q) {
  const int st = (q >> 1) == 0 ? kBI : ((q >> 1) == 1 ? kBJ : 1);
  return (q & 1) ? b + st : b - st;
}

// Does the node with code cn, lying in direction q of node r, feed r?  It
// must be present, not Z, and its argmin side along q's axis must point back
// to r (a neighbour at -1 has its upwind at +1, and the reverse).
__device__ __forceinline__ bool feeds(uint8_t cn, int q) {
  const int d = q >> 1;
  return !(cn & kCZ) && (((cn >> (2 * d)) & 3u) == ((q & 1) ? 1u : 2u));
}

// Wave-level ordering of LDS accesses by different lanes (the wave's LDS
// instructions execute in order; this keeps the compiler from moving them).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ int wave_sum(int x) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off);
  return x;
}

__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }

// Subtracts `by` (0 or 1) from the pending count of record `id` through a
// 32-bit atomic on the word holding it; returns the record's previous value.
// Counts never go below zero, so no borrow crosses into the other half of
// the word.  Lanes with nothing to release subtract 0, so every lane issues
// its atomics unconditionally and the wave waits once.
__device__ __forceinline__ unsigned pend_dec(uint16_t* pend, int id, unsigned by) {
  unsigned* w = (unsigned*)(pend + (id & ~1));
  const int sh = (id & 1) * 16;
  return (atomicSub(w, by << sh) >> sh) & 0xffffu;
}

// ---------------------------------------------------------------- phases ---

// Phase 0 (whole workgroup, grid-stride over k rows of kT nodes): upwind
// codes (the reference's Z rule, Eikonal3D.cpp:105-171; no node final yet),
// grad_u0 = g where u == u0 (:105-110; g is read only there), and
// dL/df = -0 * (-2 f h h) on Z (the value the reference's final loop gives
// those nodes, :194-196).  Z multipliers are 0 and never read: a Z node
// feeds nothing.  A thread reads its row of u and u0 and the four
// neighbouring rows (i -+ 1, j -+ 1; on a grid edge the inner row twice,
// whose value the edge rule ignores) in 16-byte pieces, and the two k
// neighbours past the row's ends.  Returns this thread's count of non-Z
// nodes.

// kT values from flat index q on (k0 + b < l; the rest 0); rows that lie
// inside the grid and start 16-byte aligned are read in 16-byte pieces.
template <typename T>
__device__ __forceinline__ void row_load(gptr<const T> p, int64_t q, int k0, int l, T (&x)[kT]) {
  constexpr int kPer = 16 / (int)sizeof(T);
  if (k0 + kT <= l && !(q & (kPer - 1))) {
    using v4 = __attribute__((ext_vector_type(4))) unsigned int;
    gptr<const v4> v = (gptr<const v4>)(p + q);
#pragma unroll
    for (int c = 0; c < kT / kPer; ++c) {
      const v4 w = v[c];
      const unsigned ww[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
      for (int e = 0; e < kPer; ++e) {
        if constexpr (sizeof(T) == 8)
          x[c * kPer + e] = __longlong_as_double((long long)(((uint64_t)ww[2 * e + 1] << 32) | ww[2 * e]));
        else
          x[c * kPer + e] = __uint_as_float(ww[e]);
      }
    }
  } else {
#pragma unroll
    for (int e = 0; e < kT; ++e) x[e] = k0 + e < l ? p[q + e] : T(0);
  }
}

template <typename T>
__device__ int phase_codes(const Ev<T>& E) {
  const int m = E.m, n = E.n, l = E.l;
  const int nkb = (l + kT - 1) / kT;
  const int64_t nrows = (int64_t)m * n * nkb;
  const int ex[3] = {m, n, l};
  int nonz = 0;
  // row r = (i n + j) nkb + kb, advanced by the stride blockDim.x =
  // (di n + dj) nkb + dkb with one carry per digit (a 64-bit division per
  // row would cost more than the row's codes)
  const int stride = (int)blockDim.x;
  const int dkb = stride % nkb, dij = stride / nkb;
  const int dj = dij % n, di = dij / n;
  int kb = (int)threadIdx.x % nkb, j, i;
  {
    const int ij = (int)threadIdx.x / nkb;
    j = ij % n;
    i = ij / n;
  }
  for (int64_t r = threadIdx.x; r < nrows; r += stride) {
    const int k0 = kb * kT;
    const int64_t q0 = (int64_t)i * E.sI + (int64_t)j * l + k0;
    T own[kT], u0r[kT], im[kT], ip[kT], jm[kT], jp[kT];
    row_load(E.u, q0, k0, l, own);
    row_load(E.u0, q0, k0, l, u0r);
    row_load(E.u, i > 0 ? q0 - E.sI : q0 + E.sI, k0, l, im);
    row_load(E.u, i < m - 1 ? q0 + E.sI : q0 - E.sI, k0, l, ip);
    row_load(E.u, j > 0 ? q0 - l : q0 + l, k0, l, jm);
    row_load(E.u, j < n - 1 ? q0 + l : q0 - l, k0, l, jp);
    const T kprev = k0 > 0 ? E.u[q0 - 1] : T(0);
    const T knext = k0 + kT < l ? E.u[q0 + kT] : T(0);
    uint64_t codes = 0;
    T gz[kT];  // grad_u0 of the row, stored in 16-byte pieces where aligned
#pragma unroll
    for (int e = 0; e < kT; ++e) {
      const int k = k0 + e;
      const T nb[6] = {im[e], ip[e], jm[e], jp[e], e > 0 ? own[e - 1] : kprev,
                       e + 1 < kT ? own[e + 1] : knext};
      const int co[3] = {i, j, k};
      const uint32_t code = k < l ? code3<T>(own[e], u0r[e], nb, co, ex) : kCZ;
      codes |= (uint64_t)code << (8 * e);
      gz[e] = T(0);
      if (k < l) {
        const int64_t q = q0 + e;
        if (E.gu0 && own[e] == u0r[e]) gz[e] = E.g[q];
        if (code & kZ) {
          const T rhs = T(-2.0) * E.f[q] * E.h * E.h;
          E.gf[q] = -T(0) * rhs;
        } else {
          ++nonz;
        }
      }
    }
    if (k0 + kT <= l && !(q0 & 7)) {
      *(gptr<uint64_t>)(E.code + q0) = codes;
    } else {
#pragma unroll
      for (int e = 0; e < kT; ++e)
        if (k0 + e < l) E.code[q0 + e] = (uint8_t)(codes >> (8 * e));
    }
    if (E.gu0) {
      constexpr int kPer = 16 / (int)sizeof(T);
      if (k0 + kT <= l && !(q0 & (kPer - 1))) {
        using v4 = __attribute__((ext_vector_type(4))) unsigned int;
        gptr<v4> dst = (gptr<v4>)(E.gu0 + q0);
#pragma unroll
        for (int c = 0; c < kT / kPer; ++c) {
          unsigned ww[4];
#pragma unroll
          for (int e = 0; e < kPer; ++e) {
            if constexpr (sizeof(T) == 8) {
              const uint64_t b = (uint64_t)__double_as_longlong(gz[c * kPer + e]);
              ww[2 * e] = (unsigned)b;
              ww[2 * e + 1] = (unsigned)(b >> 32);
            } else {
              ww[e] = __float_as_uint(gz[c * kPer + e]);
            }
          }
          v4 w;
          w.x = ww[0];
          w.y = ww[1];
          w.z = ww[2];
          w.w = ww[3];
          dst[c] = w;
        }
      } else {
#pragma unroll
        for (int e = 0; e < kT; ++e)
          if (k0 + e < l) E.gu0[q0 + e] = gz[e];
      }
    }
    kb += dkb;
    const int c1 = kb >= nkb ? 1 : 0;
    kb -= c1 * nkb;
    j += dj + c1;  // < 2n: one carry
    const int c2 = j >= n ? 1 : 0;
    j -= c2 * n;
    i += di + c2;
  }
  return nonz;
}

// Phase 1 (waves over tiles, kPT tiles per iteration with their loads in
// flight together): the count of edges that enter each tile from non-Z
// nodes of other tiles (its external feeders, all unresolved at the
// start), -1 for tiles without non-Z nodes, and the tile's ordering key.
// Lane (s, a) holds the tile's k row (s, a, 0..7), eight codes in one
// 64-bit word, checks its faces against the halo rows (i faces: s = 0, 7;
// j faces: a = 0, 7) and the two k-face halo nodes, and counts its row's
// non-Z nodes; lanes 0, 7, 56, 63 read the tile's corner travel times.
template <typename T>
__device__ __forceinline__ uint8_t code_at(const Ev<T>& E, int i, int j, int k) {
  const bool in = i >= 0 && i < E.m && j >= 0 && j < E.n && k >= 0 && k < E.l;
  return in ? E.code[(int64_t)i * E.sI + (int64_t)j * E.l + k] : kCZ;
}

// The codes (i, j, k0 .. k0+7) as bytes of one word (outside the grid: Z).
template <typename T>
__device__ __forceinline__ uint64_t code_row(const Ev<T>& E, int i, int j, int k0) {
  if (i < 0 || i >= E.m || j < 0 || j >= E.n) return 0x4040404040404040ull;
  const int64_t q = (int64_t)i * E.sI + (int64_t)j * E.l + k0;
  if (k0 + kT <= E.l && !(q & 7)) return *(gptr<const uint64_t>)(E.code + q);
  uint64_t r = 0;
#pragma unroll
  for (int b = 0; b < kT; ++b)
    r |= (uint64_t)(k0 + b < E.l ? E.code[q + b] : kCZ) << (8 * b);
  return r;
}

// Edges from the halo row `hal` into the tile row `row` across a face of
// direction q (the halo node lies at q of the tile node, byte by byte).
__device__ __forceinline__ int face_edges(uint64_t row, uint64_t hal, int q) {
  int n = 0;
#pragma unroll
  for (int b = 0; b < kT; ++b) {
    const uint8_t tn = (uint8_t)(row >> (8 * b)), hn = (uint8_t)(hal >> (8 * b));
    n += (!(tn & kCZ) && feeds(hn, q)) ? 1 : 0;
  }
  return n;
}

// Order-preserving unsigned bits of a float (NaN aside).
__device__ __forceinline__ unsigned key_bits(float x) {
  const unsigned b = __float_as_uint(x);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float key_float(unsigned k) {
  return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
}

template <typename T>
__device__ void phase_counts(const Ev<T>& E, Sched& S, int nw, int w) {
  constexpr int kPT = 4;
  const int lane = lane_id();
  const int si = lane >> 3, a = lane & 7;
  for (int t0 = w * kPT; t0 < E.ntiles; t0 += nw * kPT) {
    uint64_t row[kPT], hi[kPT], hj[kPT];
    uint8_t hk[kPT][2];
    float cu[kPT][2];
#pragma unroll
    for (int x = 0; x < kPT; ++x) {
      const int t = t0 + x < E.ntiles ? t0 + x : t0;
      const int i0 = t / (E.ntj * E.ntk) * kT, j0 = (t / E.ntk) % E.ntj * kT, k0 = t % E.ntk * kT;
      const int i = i0 + si, j = j0 + a;
      row[x] = code_row(E, i, j, k0);
      // halo rows of the i faces (si = 0: i0 - 1; si = 7: i0 + 8) and of the
      // j faces (a = 0: j0 - 1; a = 7: j0 + 8)
      hi[x] = (si == 0 || si == kT - 1) ? code_row(E, si == 0 ? i0 - 1 : i0 + kT, j, k0)
                                        : 0x4040404040404040ull;
      hj[x] = (a == 0 || a == kT - 1) ? code_row(E, i, a == 0 ? j0 - 1 : j0 + kT, k0)
                                      : 0x4040404040404040ull;
      hk[x][0] = code_at(E, i, j, k0 - 1);
      hk[x][1] = code_at(E, i, j, k0 + kT);
      // corners (clamped into the grid) on lanes (0|7, 0|7)
      const bool corner = (si == 0 || si == kT - 1) && (a == 0 || a == kT - 1);
      const int ci = i < E.m ? i : E.m - 1, cj = j < E.n ? j : E.n - 1;
      const int ck = k0 + kT - 1 < E.l ? k0 + kT - 1 : E.l - 1;
      const int64_t cq = (int64_t)ci * E.sI + (int64_t)cj * E.l;
      cu[x][0] = corner ? (float)E.u[cq + k0] : -3.4e38f;
      cu[x][1] = corner ? (float)E.u[cq + ck] : -3.4e38f;
    }
#pragma unroll
    for (int x = 0; x < kPT; ++x) {
      const int t = t0 + x;
      // non-Z nodes of the row: bytes with bit 6 clear
      int nz = __builtin_popcountll(~row[x] & 0x4040404040404040ull);
      int ext = 0;
      if (si == 0) ext += face_edges(row[x], hi[x], 0);
      if (si == kT - 1) ext += face_edges(row[x], hi[x], 1);
      if (a == 0) ext += face_edges(row[x], hj[x], 2);
      if (a == kT - 1) ext += face_edges(row[x], hj[x], 3);
      const uint8_t first = (uint8_t)row[x], last = (uint8_t)(row[x] >> 56);
      ext += (!(first & kCZ) && feeds(hk[x][0], 4)) ? 1 : 0;
      ext += (!(last & kCZ) && feeds(hk[x][1], 5)) ? 1 : 0;
      nz = wave_sum(nz);
      ext = wave_sum(ext);
      float kmax = cu[x][0] > cu[x][1] ? cu[x][0] : cu[x][1];
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) {
        const float o = __shfl_xor(kmax, off);
        kmax = o > kmax ? o : kmax;
      }
      if (lane == 0 && t < E.ntiles) {
        E.ext[t] = nz == 0 ? -1 : ext;
        E.key[t] = kmax;
        if (nz) {
          atomicMin(&S.kmin, key_bits(kmax));
          atomicMax(&S.kmax, key_bits(kmax));
        }
      }
    }
  }
}

// Counting-sort bin of a tile: tiles with non-Z nodes by decreasing key
// over [kmin, kmax] in kBins bins, tiles without in bin kBins.
__device__ __forceinline__ int tile_bin(float key, int ext, float lo, float hi) {
  if (ext < 0) return kBins;
  const float span = hi - lo;
  const float x = span > 0.f ? (hi - key) / span * (float)(kBins - 1) : 0.f;
  return x >= 0.f ? (x < (float)(kBins - 1) ? (int)x : kBins - 1) : 0;  // NaN -> 0
}

// Rank of every tile in the visiting order (rank2tile / tile2rank) and the
// initial bitmaps at those ranks: done for tiles without non-Z nodes, dirty
// for the others, ready for those without external feeders.  hist / cur:
// kBins + 1 counters of LDS scratch each.
template <typename T>
__device__ void phase_rank(const Ev<T>& E, const Sched& S, uint32_t* maps, unsigned* hist,
                           unsigned* cur) {
  const float lo = key_float(S.kmin), hi = key_float(S.kmax);
  for (int i = threadIdx.x; i <= kBins; i += blockDim.x) hist[i] = 0;
  __syncthreads();
  for (int t = threadIdx.x; t < E.ntiles; t += blockDim.x)
    atomicAdd(&hist[tile_bin(E.key[t], E.ext[t], lo, hi)], 1u);
  __syncthreads();
  if (threadIdx.x < 64) {  // exclusive scan by one wave, 17 bins per lane
    constexpr int kPer = (kBins + 1 + 63) / 64;
    const int lane = lane_id();
    unsigned v[kPer], sum = 0;
#pragma unroll
    for (int q = 0; q < kPer; ++q) {
      const int i = lane * kPer + q;
      v[q] = i <= kBins ? hist[i] : 0u;
      sum += v[q];
    }
    unsigned incl = sum;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const unsigned o = __shfl_up(incl, off);
      incl += lane >= off ? o : 0u;
    }
    unsigned run = incl - sum;
#pragma unroll
    for (int q = 0; q < kPer; ++q) {
      const int i = lane * kPer + q;
      if (i <= kBins) cur[i] = run;
      run += v[q];
    }
  }
  __syncthreads();
  for (int t = threadIdx.x; t < E.ntiles; t += blockDim.x) {
    const int ext = E.ext[t];
    const int r = (int)atomicAdd(&cur[tile_bin(E.key[t], ext, lo, hi)], 1u);
    E.t2r[t] = r;
    E.r2t[r] = t;
    const uint32_t bit = 1u << (r & 31);
    uint32_t* wd = maps + (r >> 5);
    if (ext < 0) {
      atomicOr(wd + kDone * E.nwords, bit);
    } else {
      atomicOr(wd + kDirty * E.nwords, bit);
      if (ext == 0) atomicOr(wd + kReady * E.nwords, bit);
    }
  }
}

// One visit of tile t by one wave.  Resolves every node of the tile whose
// feeders are final (inside the tile: transitively in this visit), stores
// their multipliers and dL/df, and returns the number resolved; `left` = the
// tile's nodes still unresolved; face[q] = edges from the resolved nodes to
// non-Z nodes of the neighbouring tile in direction q.
template <typename T>
__device__ int visit(const Ev<T>& E, WaveBuf<T>& B, int t, bool revisit, int& left,
                     int (&face)[6]) {
  const int lane = lane_id();
  const int ti = t / (E.ntj * E.ntk), tj = (t / E.ntk) % E.ntj, tk = t % E.ntk;
  const int i0 = ti * kT, j0 = tj * kT, k0 = tk * kT;
#pragma unroll
  for (int q = 0; q < 6; ++q) face[q] = 0;
  // --- codes and multipliers of the box (halo always, tile on revisits)
  int bx[14];
  int64_t gx[14];
  bool in[14];
  {
    uint8_t c[14];
    T lm[14];
#pragma unroll
    for (int s = 0; s < 14; ++s) {
      int gi, gj, gk;
      bx[s] = slot_node(s, lane, i0, j0, k0, gi, gj, gk);
      in[s] = gi >= 0 && gi < E.m && gj >= 0 && gj < E.n && gk >= 0 && gk < E.l;
      gx[s] = in[s] ? (int64_t)gi * E.sI + (int64_t)gj * E.l + gk : 0;
      c[s] = in[s] ? E.code[gx[s]] : kCZ;
    }
    // multipliers of final nodes only (of the tile itself on revisits),
    // loaded after their flags
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
#pragma unroll
    for (int s = 0; s < 14; ++s) {
      const bool need = (s >= 8 || revisit) && (c[s] & kCF);
      lm[s] = need ? E.lam[gx[s]] : T(0);
    }
#pragma unroll
    for (int s = 0; s < 14; ++s) {
      B.code[bx[s]] = c[s];
      B.lam[bx[s]] = lm[s];
    }
  }
  wave_sync();
  // --- records of the tile's nodes (lane owns slab nodes s*64+lane)
  uint32_t todo = 0;  // bit s: node (s, lane) unresolved and not Z
  int nready = 0;
#pragma unroll
  for (int s = 0; s < 8; ++s) {
    const int bi = bx[s];
    const uint8_t c = B.code[bi];
    uint8_t cn[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) cn[q] = B.code[box_nb(bi, q)];
    uint32_t rec = kPNA;
    if (!(c & (kCZ | kCF))) {
      uint32_t cnt = 0, fm = 0, rel = 0;
#pragma unroll
      for (int q = 0; q < 6; ++q) {
        const bool fd = feeds(cn[q], q);
        fm |= fd ? (1u << q) : 0u;
        cnt += (fd && !(cn[q] & kCF)) ? 1u : 0u;
      }
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const uint32_t sd = (c >> (2 * d)) & 3u;
        const uint8_t cx = cn[2 * d + (sd == 2 ? 1 : 0)];
        rel |= (sd && !(cx & kCZ)) ? (1u << d) : 0u;
      }
      rec = cnt | (fm << kPFeedShift) | (rel << kPRelShift);
      todo |= 1u << s;
      nready += cnt == 0 ? 1 : 0;
    }
    B.pend[s * 64 + lane] = (uint16_t)rec;
  }
  left = wave_sum(__builtin_popcount(todo));
  if (wave_sum(nready) == 0) return 0;  // nothing can resolve yet
  // --- travel times of the box, adjoint sources and slowness of the tile
  T fr[8];
  {
    T uv[14], gv[8];
    const rsrc_t ru = make_rsrc((const void*)E.u, (uint64_t)E.m * E.sI * sizeof(T));
#pragma unroll
    for (int s = 0; s < 14; ++s)
      uv[s] = bload(ru, in[s] ? (uint32_t)(gx[s] * (int64_t)sizeof(T)) : kOOB, T(0));
#pragma unroll
    for (int s = 0; s < 8; ++s) {
      const bool need = (todo >> s) & 1u;
      gv[s] = need ? E.g[gx[s]] : T(0);
      fr[s] = need ? E.f[gx[s]] : T(0);
    }
#pragma unroll
    for (int s = 0; s < 14; ++s) B.u[bx[s]] = uv[s];
    // lambda of an unresolved tile node starts as its source g
#pragma unroll
    for (int s = 0; s < 8; ++s)
      if ((todo >> s) & 1u) B.lam[bx[s]] = gv[s];
  }
  wave_sync();
  // --- initial queue: unresolved nodes with no pending feeder
  int qtail = 0;
#pragma unroll
  for (int s = 0; s < 8; ++s) {
    const bool r = ((todo >> s) & 1u) && (B.pend[s * 64 + lane] & kPCount) == 0;
    const uint64_t bal = __ballot(r);
    const int pos = qtail + (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(bal >> 32),
                                                        __builtin_amdgcn_mbcnt_lo((uint32_t)bal, 0));
    if (r && pos < kTN) B.queue[pos] = (uint16_t)(s * 64 + lane);
    qtail += __builtin_popcountll(bal);
  }
  wave_sync();
  // --- traversal: each round resolves up to 64 queued nodes, all of whose
  // feeders are final, then releases their upwind neighbours.  A round reads
  // a node's record, its own and its six neighbours' values in one batch.
  int qhead = 0;
  int fcnt[6] = {0, 0, 0, 0, 0, 0};
  while (qhead < qtail) {
    const int qend = qhead + 64 < qtail ? qhead + 64 : qtail;  // this round's entries
    const int qi = qhead + lane;
    const bool act = qi < qend;
    const int id = act ? (int)B.queue[qi] : 0;
    const int br = box_of_id(id);
    const uint32_t rec = B.pend[id];
    const uint8_t cr = B.code[br];
    const T ur = B.u[br];
    const T g0 = B.lam[br];
    T un[6], ln[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      un[q] = B.u[box_nb(br, q)];
      ln[q] = B.lam[box_nb(br, q)];
    }
    T num = g0;
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      const T c = num + T(2.0) * (un[q] - ur) * ln[q];
      num = ((rec >> (kPFeedShift + q)) & 1u) ? c : num;
    }
    T D = T(0);
    bool have = false;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const uint32_t sd = (cr >> (2 * d)) & 3u;
      const T w = T(2.0) * (ur - (sd == 2 ? un[2 * d + 1] : un[2 * d]));
      D = sd ? (have ? D + w : w) : D;
      have = have || sd;
    }
    const T lr = num / D;
    wave_sync();  // every lane has read its feeders before any writes
    if (act) {
      B.lam[br] = lr;
      B.pend[id] = (uint16_t)(rec | kPDone);
    }
    wave_sync();
    // release the upwind neighbours (pend decrement; halo ones are counted
    // for the neighbouring tile's external count)
    bool push[3];
    int pid[3];
    unsigned old[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const uint32_t sd = (cr >> (2 * d)) & 3u;
      const int co = d == 0 ? (id >> 6) : (d == 1 ? ((id >> 3) & 7) : (id & 7));
      const int cn = sd == 1 ? co - 1 : co + 1;
      const bool rel = act && ((rec >> (kPRelShift + d)) & 1u);
      const bool inside = cn >= 0 && cn < kT;
      const bool dec = rel && inside;
      pid[d] = dec ? id + (cn - co) * (d == 0 ? 64 : (d == 1 ? 8 : 1)) : id;
      old[d] = pend_dec(B.pend, pid[d], dec ? 1u : 0u);
      push[d] = dec && (old[d] & kPCount) == 1u;
      fcnt[2 * d + (sd == 2 ? 1 : 0)] += (rel && !inside) ? 1 : 0;
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const uint64_t bal = __ballot(push[d]);
      const int pos = qtail + (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(bal >> 32),
                                                          __builtin_amdgcn_mbcnt_lo((uint32_t)bal, 0));
      if (push[d] && pos < kTN) B.queue[pos] = (uint16_t)pid[d];
      qtail += __builtin_popcountll(bal);
    }
    qhead = qend;
    qtail = qtail < kTN ? qtail : kTN;  // each node enters once: never clamps
    wave_sync();
  }
  // --- results of the nodes resolved in this visit (owners store coalesced):
  // multipliers and dL/df, then the final flags.  Another wave may read a
  // code byte at any time and reads the multiplier only after it saw the
  // flag; within the workgroup the release keeps the flag behind the value.
  int nres = 0;
  uint32_t fin = 0;
#pragma unroll
  for (int s = 0; s < 8; ++s) {
    if (((todo >> s) & 1u) && (B.pend[s * 64 + lane] & kPDone)) {
      const T lr = B.lam[bx[s]];
      E.lam[gx[s]] = (lr == lr) ? lr : T(__builtin_nan(""));
      const T rhs = T(-2.0) * fr[s] * E.h * E.h;  // Eikonal3D.cpp:113-116, 194-196
      E.gf[gx[s]] = -lr * rhs;
      fin |= 1u << s;
      ++nres;
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
#pragma unroll
  for (int s = 0; s < 8; ++s)
    if ((fin >> s) & 1u) E.code[gx[s]] = (uint8_t)(B.code[bx[s]] | kCF);
#pragma unroll
  for (int q = 0; q < 6; ++q) face[q] = wave_sum(fcnt[q]);
  nres = wave_sum(nres);
  left -= nres;
  return nres;
}

// Tile of neighbouring direction q of tile t, or -1 outside the grid.
template <typename T>
__device__ __forceinline__ int tile_nb(const Ev<T>& E, int t, int q) {
  int c[3] = {t / (E.ntj * E.ntk), (t / E.ntk) % E.ntj, t % E.ntk};
  const int ext[3] = {E.nti, E.ntj, E.ntk};
  const int d = q >> 1;
  c[d] += (q & 1) ? 1 : -1;
  if (c[d] < 0 || c[d] >= ext[d]) return -1;
  return (c[0] * E.ntj + c[1]) * E.ntk + c[2];
}

// Takes the tile of lowest rank marked in map `which` (kReady or kDirty)
// that no wave is visiting and that is not done: clears that bit (and the
// dirty bit: the visit sees every delivery made so far), marks the tile
// busy and counts the wave as active.  Returns its rank, or -1 if none was
// found (wave-uniform).
__device__ __forceinline__ int pop_tile(uint32_t* maps, int nwords, Sched& S, int which) {
  const int lane = lane_id();
  for (int base = 0; base < nwords; base += 64) {
    const int wi = base + lane;
    uint32_t cand = 0;
    if (wi < nwords)
      cand = maps[which * nwords + wi] & ~maps[kBusy * nwords + wi] & ~maps[kDone * nwords + wi];
    const uint64_t bal = __ballot(cand != 0);
    if (!bal) continue;
    const int src = __builtin_ctzll(bal);
    int t = -1;
    if (lane == src) {
      const uint32_t bit = cand & (0u - cand);
      const uint32_t old = atomicAnd(&maps[which * nwords + wi], ~bit);
      if (old & bit) {
        const uint32_t ob = atomicOr(&maps[kBusy * nwords + wi], bit);
        if (ob & bit) {
          atomicOr(&maps[which * nwords + wi], bit);  // being visited: leave the mark
        } else {
          atomicAdd(&S.active, 1);
          if (which != kDirty) atomicAnd(&maps[kDirty * nwords + wi], ~bit);
          t = wi * 32 + __builtin_ctz(bit);
        }
      }
    }
    return __shfl(t, src);
  }
  return -1;
}

}  // namespace tile

// The whole adjoint of one event by one workgroup of NW waves (at most
// MAXT tiles).  Returns nothing; results and diagnostics go to the event.
template <typename T, int NW, int MAXT>
__device__ void adj3_tile_event(const EventDesc& ev) {
  using namespace tile;
  constexpr int MW = MAXT / 32;
  __shared__ WaveBuf<T> wb[NW];
  __shared__ uint32_t maps[kMaps * MW];
  __shared__ Sched S;
  Ev<T> E;
  E.m = ev.m;
  E.n = ev.n;
  E.l = ev.l;
  E.nti = (E.m + kT - 1) / kT;
  E.ntj = (E.n + kT - 1) / kT;
  E.ntk = (E.l + kT - 1) / kT;
  E.ntiles = E.nti * E.ntj * E.ntk;
  E.nwords = (E.ntiles + 31) / 32;
  E.sI = (int64_t)E.n * E.l;
  E.h = (T)ev.h;
  E.u = (gptr<const T>)ev.u;
  E.u0 = (gptr<const T>)ev.u0;
  E.g = (gptr<const T>)ev.grad_u;
  E.f = (gptr<const T>)ev.f;
  E.gu0 = (gptr<T>)ev.grad_u0;
  E.gf = (gptr<T>)ev.grad_f;
  E.lam = (gptr<T>)ev.lam;
  E.code = (gptr<uint8_t>)ev.code;
  E.ext = (gptr<int>)ev.queue;
  E.t2r = E.ext + E.ntiles;
  E.r2t = E.t2r + E.ntiles;
  E.key = (gptr<float>)(E.r2t + E.ntiles);
  const int w = threadIdx.x >> 6;
  const int lane = lane_id();
  if (E.ntiles > MAXT) {  // the host routes such events elsewhere
    if (threadIdx.x == 0) ev.info->fault = 1;
    return;
  }
  for (int i = threadIdx.x; i < kMaps * E.nwords; i += blockDim.x) maps[i] = 0;
  if (threadIdx.x == 0) {
    S.remaining = 0;
    S.active = 0;
    S.fault = 0;
    S.visits = 0;
    S.kmin = 0xffffffffu;
    S.kmax = 0u;
  }
  __syncthreads();
  const int nz = phase_codes<T>(E);
  atomicAdd(&S.remaining, nz);
  __syncthreads();
  phase_counts<T>(E, S, NW, w);
  __syncthreads();
  // the waves' boxes are free until the traversal: counting-sort scratch
  phase_rank<T>(E, S, maps, (unsigned*)wb[0].u, (unsigned*)wb[0].lam);
  __syncthreads();
  // --- tile visits until every non-Z node is final
  int idle = 0;
  int last_rem = -1;
  for (;;) {
    const int rem = __hip_atomic_load(&S.remaining, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    if (rem == 0 || __hip_atomic_load(&S.fault, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP))
      break;
    // a visit either resolves nodes or clears the bit that scheduled it, so
    // the visits are bounded; this cap only turns a logic error into a fault
    if (__hip_atomic_load(&S.visits, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) >
        kVisitCapPerTile * E.ntiles + 1024) {
      if (lane == 0) atomicExch(&S.fault, 1);
      break;
    }
    int r = pop_tile(maps, E.nwords, S, kReady);
    if (r < 0) r = pop_tile(maps, E.nwords, S, kDirty);
    if (r >= 0) {
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
      const int t = E.r2t[r];
      const uint32_t bit = 1u << (r & 31);
      const bool revisit = maps[kVisited * E.nwords + (r >> 5)] & bit;
      int left = 0, face[6];
      const int nres = visit<T>(E, wb[w], t, revisit, left, face);
      // the visit's stores before the bits and counts that publish them
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
      if (lane < 6 && face[lane] > 0) {
        const int t2 = tile_nb(E, t, lane);
        if (t2 >= 0) {
          const int old = __hip_atomic_fetch_sub(&E.ext[t2], face[lane], __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_WORKGROUP);
          const int r2 = E.t2r[t2];
          const uint32_t b2 = 1u << (r2 & 31);
          atomicOr(&maps[kDirty * E.nwords + (r2 >> 5)], b2);
          if (old == face[lane]) atomicOr(&maps[kReady * E.nwords + (r2 >> 5)], b2);
        }
      }
      if (lane == 0) {
        atomicOr(&maps[kVisited * E.nwords + (r >> 5)], bit);
        if (left == 0) atomicOr(&maps[kDone * E.nwords + (r >> 5)], bit);
        atomicAnd(&maps[kBusy * E.nwords + (r >> 5)], ~bit);
        atomicAdd(&S.visits, 1);
        if (nres) atomicSub(&S.remaining, nres);
        atomicSub(&S.active, 1);
      }
      wave_sync();
      idle = 0;
      continue;
    }
    // nothing ready or dirty: other waves hold the work; poll
    if (rem != last_rem) {
      last_rem = rem;
      idle = 0;
    }
    if (++idle > kIdleCap) {
      if (lane == 0) atomicExch(&S.fault, 1);
      break;
    }
    __builtin_amdgcn_s_sleep(1);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    ev.info->adj_sweeps = S.visits;
    ev.info->adj_converged = (S.remaining == 0 && !S.fault) ? 1 : 0;
    if (S.fault) ev.info->fault = 1;
  }
}

}  // namespace adt
