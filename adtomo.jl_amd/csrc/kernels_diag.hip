// 3D forward: diagonal-streaming fast sweeping, one workgroup per event.
//
// Schedule.  A sweep (Eikonal3D.cpp:35-57) is equivalent to processing the
// hyperplanes a+b+c = s in increasing s, with (a, b, c) counted along the
// sweep direction (c along k, the contiguous axis): a node's upwind
// neighbours are final and its downwind neighbours still old when it is
// updated, exactly as in the reference's loop nest.
//
// The (a, b) plane is cut into "virtual anti-diagonals": b runs along the
// longer of the two non-k axes (N_b nodes), a along the other (M_a nodes),
// split into `bands` bands of H <= 64 rows.  Lane p of diagonal v owns column
//     a = beta*H + p,  b = (v - p) mod N_b,  beta = (v - p) div N_b,
// so (a-1, b) is lane p-1 and (a, b-1) lane p of diagonal v-1; (a+1, b) is
// lane p+1 and (a, b+1) lane p of diagonal v+1.  Across a band edge (lane 0
// or lane H-1), (a-1, b) lies on diagonal v - delta and (a+1, b) on
// v + delta, delta = N_b - H + 1 >= 33.
//
// The c axis is cut into S <= kDW segments of Lw <= kLW nodes, one per wave
// (S is the event's launch setting, kDW by default; waves beyond S idle and
// the k axis is swept in chunks of S * kLW nodes).  At step T wave w
// processes diagonal v = T - w over its segment, node by node in
// increasing c.  Each node's neighbours then come from:
//   (a-1, b), (a, b-1): the new values of diagonal v-1 that lanes p-1 and p
//       left in the wave's LDS plane n, or, across a band edge, memory
//       (stored delta - 1 >= 32 steps before);
//   (a+1, b), (a, b+1): old values of diagonal v+1 (lanes p+1 and p), loaded
//       one step ahead into plane x; across a band edge from memory (not yet
//       updated);
//   (a, b, c-1): the lane's previous node, or for the first node of the
//       segment wave w-1's last node of diagonal v (step T-1, through LDS);
//   (a, b, c+1): the lane's next node, or one past the segment (loaded with
//       the segment; wave w+1 updates it only at step T+1).
// One workgroup barrier per step; every lane updates Lw nodes per step.
//
// Loads for diagonals v+1 (slowness, band-edge values) and v+2 (old travel
// times) are issued at the start of step T and used at step T+1, so each
// step hides one step of memory latency.  Segments move between memory and
// the planes in 16-byte pieces, several lanes per contiguous run.  Memory
// instructions are never skipped: masked lanes use an out-of-range buffer
// offset.
//
// Chained sweeps.  A sweep's first H-1 diagonals leave lanes p > v idle and
// its last H-1 leave lanes p <= v - bands*N_b idle, exactly complementary:
// when the next sweep keeps the c direction (it flips the a or the b axis),
// sweep k+1's diagonal v runs in the same step as sweep k's diagonal
// v + bands*N_b, lanes p <= v on sweep k+1 and the others on sweep k.
// Every global access of wave w touches k-segment w only, so the only
// ordering needed is the wave's own program order: a column's old values
// for sweep k+1 must have been stored by sweep k at least two steps before
// they are loaded, and sweep k's last columns must be read before sweep k+1
// stores over their neighbours.  With diagonal indices both hold when the
// next sweep flips a and N_b >= H + 3 (its first columns are sweep k's
// columns of diagonals <= (bands-1) N_b + H), or flips b and bands >= 2
// (they are sweep k's first band).  In-plane neighbours across the lane
// boundary of the two sweeps are grid edges, whose one-sided minimum never
// reads them.  Chains stop at the first sweep of the first iteration (it
// reads u0), at a flip of c, and for events swept in several k chunks.  A
// chain may end with the last sweep of an iteration (it takes the stop
// rule): the steps that reach it store its lanes' values to the next
// snapshot and fold their error, lane by lane.
//
// Iteration (Eikonal3D.cpp:71-88): u = u0 is a copy pass at the start;
// err = max|u - u_old| is a coalesced pass after the eighth sweep against
// the snapshot (u0 for the first iteration), which also writes the next
// snapshot u_old = u.
#include <type_traits>

#include "adjoint3.hpp"
#include "buffer_ops.hpp"
#include "misfit3.hpp"
#include "eikonal_device.hpp"
#include "kernels.hpp"

namespace adt {

namespace {

template <typename T>
using lptr = __attribute__((address_space(3))) T*;

constexpr int kDW = 8;  // waves per workgroup = c segments
constexpr int kLW = 8;  // maximum segment length (nodes per lane per step)

struct DGeom {
  int Ma, Nb, L, H, bands, V, Lw;
  int sA, sB;  // memory strides of the a and b axes (offsets fit 32 bits)
  bool a_is_i;
};

__device__ inline DGeom make_dgeom(int m, int n, int l) {
  DGeom g;
  g.a_is_i = (m <= n);
  g.Ma = g.a_is_i ? m : n;
  g.Nb = g.a_is_i ? n : m;
  g.sA = g.a_is_i ? n * l : l;
  g.sB = g.a_is_i ? l : n * l;
  g.L = l;
  g.Lw = (l + kDW - 1) / kDW;
  g.bands = (g.Ma + 63) / 64;
  g.H = (g.Ma + g.bands - 1) / g.bands;
  g.V = g.H - 1 + g.bands * g.Nb;
  return g;
}

// Column of lane p on a diagonal: canonical (a, b) of its sweep, whether
// that sweep flips a / b, and the flat index of canonical (a, b, c = 0).
// The flags are bits of an integer (a lane value in a vector register, not
// a lane mask in scalar registers, which the step loop cannot spare).
struct DCol {
  int a, b;
  uint32_t fl;  // bit 0: valid, bit 1: fa, bit 2: fb, bit 3: the chain's last sweep
  int node0;
  __device__ __forceinline__ bool valid() const { return fl & 1; }
  __device__ __forceinline__ bool fa() const { return fl & 2; }
  __device__ __forceinline__ bool fb() const { return fl & 4; }
  __device__ __forceinline__ bool last() const { return fl & 8; }
};

constexpr int kChainMax = 4;  // sweeps of one chain (the c direction fixed)

template <typename T>
struct DSweep {
  DGeom g;
  int tC;
  bool flipC;
  // the chained sweeps (1..kChainMax): bits 2q, 2q+1 of cfl = sweep q
  // flips a, b; a sweep's canonical (a, b, c = 0) lies at
  // ra sA + rb sB + coff, ra = fa ? Ma-1-a : a, rb = fb ? Nb-1-b : b
  int nch;
  int V0;  // bands * N_b: diagonal offset between consecutive chained sweeps
  int cfl;
  int coff;
  T h;
  double hd;    // h for the FP64 local solve of the FP32-storage variant
  int cs;       // canonical c of this pass's first node (k chunks)
  int Lfull;    // k extent of the event (g.L is the chunk's)
  int nseg;     // c segments (active waves), 1..kDW
  bool first;   // the event's first sweep: its old values are u0
  rsrc_t u;     // travel times: new values written (and read back) here
  rsrc_t src;   // old values: u, or u0 in the very first sweep (u = u0,
                // Eikonal3D.cpp:72, folded into that sweep)
  rsrc_t f;
  rsrc_t u0;
  rsrc_t snap_src;  // last sweep of an iteration: snapshot at its start
  rsrc_t snap_dst;  // ... and the snapshot for the next iteration
};

// Position of lane p on chain diagonal t: with q = t - p, the lane's
// chained sweep is k = q div V0 (lanes p > t - kk V0 of diagonal t, kk =
// t div V0, still run sweep kk-1's last diagonals), its band beta and
// canonical b come from r = q mod V0 = beta N_b + b, and a = beta H + p.
// Lanes keep (k, beta, b) and advance it by one diagonal per step; before
// the chain's first column reaches the lane (q < 0) b holds q.
struct DCur {
  int k, beta, b;
};

template <typename T>
__device__ __forceinline__ DCur cur_at(const DSweep<T>& S, int t, int p) {
  const int q = t - p;
  DCur c;
  if (q < 0) {
    c.k = 0;
    c.beta = 0;
    c.b = q;
  } else {
    c.k = q / S.V0;
    const int r = q - c.k * S.V0;
    c.beta = r / S.g.Nb;
    c.b = r - c.beta * S.g.Nb;
  }
  return c;
}

template <typename T>
__device__ __forceinline__ void cur_next(const DSweep<T>& S, DCur& c) {
  const int b1 = c.b + 1;
  const bool wb = b1 == S.g.Nb;
  c.b = wb ? 0 : b1;
  const int beta1 = c.beta + (wb ? 1 : 0);
  const bool wk = beta1 == S.g.bands;
  c.beta = wk ? 0 : beta1;
  c.k += wk ? 1 : 0;
}

template <typename T>
__device__ __forceinline__ DCol cur_col(const DSweep<T>& S, const DCur& c, int p) {
  const DGeom& g = S.g;
  DCol col;
  const int kc = c.k < kChainMax ? c.k : kChainMax;  // bits beyond the chain are 0
  const uint32_t fl = (S.cfl >> (2 * kc)) & 3;
  col.a = c.beta * g.H + p;
  col.b = c.b;
  const bool valid = (p < g.H) && (c.b >= 0) && (c.k < S.nch) && (col.a < g.Ma);
  col.fl = (fl << 1) | (valid ? 1u : 0u) | (c.k == S.nch - 1 ? 8u : 0u);
  const int ra = (fl & 1) ? g.Ma - 1 - col.a : col.a;
  const int rb = (fl & 2) ? g.Nb - 1 - col.b : col.b;
  col.node0 = S.coff + ra * g.sA + rb * g.sB;
  return col;
}

// cnt values of column `col` from canonical c0 on (element j <- c0 + j),
// one load per element.
template <typename T, int NE>
__device__ __forceinline__ void load_seg(T (&buf)[NE], rsrc_t r, const DCol& col, int delta,
                                         bool ok, int c0, int cnt, int L, int tC) {
#pragma unroll
  for (int j = 0; j < NE; ++j) {
    const int c = c0 + j;
    buf[j] = bload(r, boff<T>(ok && j < cnt && c < L, col.node0 + delta + c * tC), T(0));
  }
}

// ---- per-wave LDS planes and coalesced segment transfers ------------------
//
// Each wave keeps three planes in LDS, each holding one value per (node of
// the segment, column of the wave): node j (canonical order along the
// sweep) of lane q's segment sits at row j, position q + 1.  Position 0
// stands for lane -1 and position H + 1 for lane H: the band-edge
// neighbours of lanes 0 and H-1 are written there, so every lane reads its
// a-neighbours at positions p and p + 2 without exceptions.  Row kLW is
// scratch for writes that have no row.  Rows are padded so the piece writes
// spread over the banks.
//
// Segments move between memory and a plane in 16-byte pieces: lane L of
// transfer i handles piece L % LPC of the segment of lane i*CPI + L / LPC,
// so the LPC lanes of one segment read or write one contiguous run.  A
// segment whose length is not a multiple of the piece is finished by its
// own lane, element by element.  FULL: every segment has kLW nodes.

using u32x4 = __attribute__((ext_vector_type(4))) unsigned int;

template <typename T>
struct Pieces {
  static constexpr int EP = 16 / (int)sizeof(T);      // elements per piece
  static constexpr int LPC = kLW / EP;                // lanes per segment
  static constexpr int CPI = 64 / LPC;                // segments per transfer
  static constexpr int RS = 64 + 32 / (int)sizeof(T);  // plane row stride (elements)
  static constexpr int PLANE = (kLW + 1) * RS;        // kLW rows + a scratch row
};

__device__ __forceinline__ u32x4 bload4(rsrc_t r, uint32_t off) {
  return __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0);
}
__device__ __forceinline__ void bstore4(rsrc_t r, uint32_t off, u32x4 v) {
  __builtin_amdgcn_raw_buffer_store_b128(v, r, off, 0, 0);
}

template <typename T>
__device__ __forceinline__ T piece_get(const u32x4& v, int e);
template <>
__device__ __forceinline__ double piece_get<double>(const u32x4& v, int e) {
  const uint32_t lo = e == 0 ? v.x : v.z, hi = e == 0 ? v.y : v.w;
  return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}
template <>
__device__ __forceinline__ float piece_get<float>(const u32x4& v, int e) {
  return __uint_as_float(e == 0 ? v.x : e == 1 ? v.y : e == 2 ? v.z : v.w);
}

template <typename T>
__device__ __forceinline__ u32x4 piece_make(const T (&x)[16 / sizeof(T)]);
template <>
__device__ __forceinline__ u32x4 piece_make<double>(const double (&x)[2]) {
  const uint64_t a = (uint64_t)__double_as_longlong(x[0]), b = (uint64_t)__double_as_longlong(x[1]);
  u32x4 v;
  v.x = (uint32_t)a;
  v.y = (uint32_t)(a >> 32);
  v.z = (uint32_t)b;
  v.w = (uint32_t)(b >> 32);
  return v;
}
template <>
__device__ __forceinline__ u32x4 piece_make<float>(const float (&x)[4]) {
  u32x4 v;
  v.x = __float_as_uint(x[0]);
  v.y = __float_as_uint(x[1]);
  v.z = __float_as_uint(x[2]);
  v.w = __float_as_uint(x[3]);
  return v;
}

// Loads in flight for one segment transfer.
template <typename T>
struct Flight {
  u32x4 pc[Pieces<T>::LPC];
  T tail[Pieces<T>::EP - 1];
};

// Canonical row of memory element m of a segment of cnt nodes.
template <bool FULL>
__device__ __forceinline__ int canon(int m, int cnt, bool fwd) {
  return fwd ? m : (FULL ? kLW : cnt) - 1 - m;
}

// Issues the loads of every lane's segment: `seg` = memory index of the
// lowest node of the lane's segment (-1: none), cnt = nodes per segment.
template <typename T, bool FULL>
__device__ __forceinline__ void seg_issue(Flight<T>& fl, rsrc_t r, int seg, int cnt, int p) {
  using P = Pieces<T>;
  if (FULL) cnt = kLW;
#pragma unroll
  for (int i = 0; i < P::LPC; ++i) {
    const int q = i * P::CPI + p / P::LPC;
    const int m0 = (p % P::LPC) * P::EP;
    const int sq = __shfl(seg, q);
    const bool ok = sq >= 0 && (FULL || m0 + P::EP <= cnt);
    fl.pc[i] = bload4(r, ok ? (uint32_t)(sq + m0) * (uint32_t)sizeof(T) : kOOB);
  }
  if (!FULL) {
    const int nfull = (cnt / P::EP) * P::EP;
#pragma unroll
    for (int t = 0; t < P::EP - 1; ++t) {
      const int m = nfull + t;
      fl.tail[t] = bload(r, boff<T>(seg >= 0 && m < cnt, (int64_t)seg + m), T(0));
    }
  }
}

// Writes arrived loads into a plane (canonical rows).
template <typename T, bool FULL>
__device__ __forceinline__ void seg_land(const Flight<T>& fl, lptr<T> pl, int cnt, bool fwd,
                                         int p) {
  using P = Pieces<T>;
#pragma unroll
  for (int i = 0; i < P::LPC; ++i) {
    const int q = i * P::CPI + p / P::LPC;
    const int m0 = (p % P::LPC) * P::EP;
#pragma unroll
    for (int e = 0; e < P::EP; ++e) {
      const int m = m0 + e;
      const int j = (FULL || m < cnt) ? canon<FULL>(m, cnt, fwd) : kLW;
      pl[j * P::RS + q + 1] = piece_get<T>(fl.pc[i], e);
    }
  }
  if (!FULL) {
    const int nfull = (cnt / P::EP) * P::EP;
#pragma unroll
    for (int t = 0; t < P::EP - 1; ++t) {
      const int m = nfull + t;
      const int j = m < cnt ? canon<FULL>(m, cnt, fwd) : kLW;
      pl[j * P::RS + p + 1] = fl.tail[t];
    }
  }
}

// Stores every lane's segment from a plane (canonical rows).
template <typename T, bool FULL>
__device__ __forceinline__ void seg_store(lptr<const T> pl, rsrc_t r, int seg, int cnt, bool fwd,
                                          int p) {
  using P = Pieces<T>;
  if (FULL) cnt = kLW;
#pragma unroll
  for (int i = 0; i < P::LPC; ++i) {
    const int q = i * P::CPI + p / P::LPC;
    const int m0 = (p % P::LPC) * P::EP;
    const int sq = __shfl(seg, q);
    const bool ok = sq >= 0 && (FULL || m0 + P::EP <= cnt);
    T x[P::EP];
#pragma unroll
    for (int e = 0; e < P::EP; ++e) {
      const int m = m0 + e;
      const int j = (FULL || m < cnt) ? canon<FULL>(m, cnt, fwd) : kLW;
      x[e] = pl[j * P::RS + q + 1];
    }
    bstore4(r, ok ? (uint32_t)(sq + m0) * (uint32_t)sizeof(T) : kOOB, piece_make<T>(x));
  }
  if (!FULL) {
    const int nfull = (cnt / P::EP) * P::EP;
#pragma unroll
    for (int t = 0; t < P::EP - 1; ++t) {
      const int m = nfull + t;
      const int j = m < cnt ? canon<FULL>(m, cnt, fwd) : kLW;
      bstore(r, boff<T>(seg >= 0 && m < cnt, (int64_t)seg + m), pl[j * P::RS + p + 1]);
    }
  }
}

// seg_store for the last sweep of an iteration: the segments of lanes on
// that sweep (segs: seg, or -1 for lanes still on the sweep before it in a
// chain) also go to the next snapshot, and max |new - snapshot| over them is
// folded into `key` (the stop rule's err, Eikonal3D.cpp:79-81), `snap`
// holding the snapshot's values at the same offsets (loaded with seg_issue
// one step earlier).  The stores are issued before the snapshot values are
// waited for.
template <typename T, bool FULL>
__device__ __forceinline__ void seg_store_err(lptr<const T> pl, rsrc_t r, rsrc_t sdst, int seg,
                                              int segs, int cnt, bool fwd, int p,
                                              const Flight<T>& snap, uint64_t& key) {
  using P = Pieces<T>;
  if (FULL) cnt = kLW;
  T x[P::LPC][P::EP];
  bool ok[P::LPC];
#pragma unroll
  for (int i = 0; i < P::LPC; ++i) {
    const int q = i * P::CPI + p / P::LPC;
    const int m0 = (p % P::LPC) * P::EP;
    const int sq = __shfl(seg, q);
    const int ss = __shfl(segs, q);
    const bool whole = FULL || m0 + P::EP <= cnt;
    ok[i] = ss >= 0 && whole;
#pragma unroll
    for (int e = 0; e < P::EP; ++e) {
      const int m = m0 + e;
      const int j = (FULL || m < cnt) ? canon<FULL>(m, cnt, fwd) : kLW;
      x[i][e] = pl[j * P::RS + q + 1];
    }
    const uint32_t off = (sq >= 0 && whole) ? (uint32_t)(sq + m0) * (uint32_t)sizeof(T) : kOOB;
    const uint32_t offs = ok[i] ? (uint32_t)(ss + m0) * (uint32_t)sizeof(T) : kOOB;
    const u32x4 v = piece_make<T>(x[i]);
    bstore4(r, off, v);
    bstore4(sdst, offs, v);
  }
  T xt[P::EP - 1];
  if (!FULL) {
    const int nfull = (cnt / P::EP) * P::EP;
#pragma unroll
    for (int t = 0; t < P::EP - 1; ++t) {
      const int m = nfull + t;
      const int j = m < cnt ? canon<FULL>(m, cnt, fwd) : kLW;
      xt[t] = pl[j * P::RS + p + 1];
      bstore(r, boff<T>(seg >= 0 && m < cnt, (int64_t)seg + m), xt[t]);
      bstore(sdst, boff<T>(segs >= 0 && m < cnt, (int64_t)segs + m), xt[t]);
    }
  }
#pragma unroll
  for (int i = 0; i < P::LPC; ++i) {
#pragma unroll
    for (int e = 0; e < P::EP; ++e) {
      const uint64_t k = maxkey(Real<T>::fabs_(x[i][e] - piece_get<T>(snap.pc[i], e)));
      key = (ok[i] && k > key) ? k : key;
    }
  }
  if (!FULL) {
    const int nfull = (cnt / P::EP) * P::EP;
#pragma unroll
    for (int t = 0; t < P::EP - 1; ++t) {
      const uint64_t k = maxkey(Real<T>::fabs_(xt[t] - snap.tail[t]));
      key = (segs >= 0 && nfull + t < cnt && k > key) ? k : key;
    }
  }
}

// Does a transfer in flight (once arrived) hold NaN or -0?  Masked loads
// read +0.
template <typename T>
__device__ __forceinline__ bool flight_odd(const Flight<T>& fl, bool full) {
  using P = Pieces<T>;
  bool odd = false;
#pragma unroll
  for (int i = 0; i < P::LPC; ++i)
#pragma unroll
    for (int e = 0; e < P::EP; ++e) odd = (int)odd | (int)Real<T>::odd_(piece_get<T>(fl.pc[i], e));
  if (!full) {
#pragma unroll
    for (int t = 0; t < P::EP - 1; ++t) odd = (int)odd | (int)Real<T>::odd_(fl.tail[t]);
  }
  return odd;
}

// Does a transfer in flight (once arrived) hold a value with which the
// solve may produce NaN (Real::wild_)?
template <typename T>
__device__ __forceinline__ bool flight_wild(const Flight<T>& fl, bool full) {
  using P = Pieces<T>;
  bool wild = false;
#pragma unroll
  for (int i = 0; i < P::LPC; ++i)
#pragma unroll
    for (int e = 0; e < P::EP; ++e) wild = (int)wild | (int)Real<T>::wild_(piece_get<T>(fl.pc[i], e));
  if (!full) {
#pragma unroll
    for (int t = 0; t < P::EP - 1; ++t) wild = (int)wild | (int)Real<T>::wild_(fl.tail[t]);
  }
  return wild;
}

// Publishes canonical node 0 of every lane's segment from a transfer in
// flight (once it has arrived) to dst[lane]: the old value the wave
// upstream along k needs one past its own segment.
template <typename T, bool FULL>
__device__ __forceinline__ void seg_publish(const Flight<T>& fl, lptr<T> dst, int cnt, bool fwd,
                                            int p) {
  using P = Pieces<T>;
  if (FULL) cnt = kLW;
  const int mp = fwd ? 0 : cnt - 1;  // memory element holding canonical node 0
  const int nfull = (cnt / P::EP) * P::EP;
  if (FULL || mp < nfull) {
    const int kp = mp / P::EP, ep = mp % P::EP;
    if (p % P::LPC == kp) {
#pragma unroll
      for (int i = 0; i < P::LPC; ++i) {
        const int q = i * P::CPI + p / P::LPC;
        T x = piece_get<T>(fl.pc[i], 0);
#pragma unroll
        for (int e = 1; e < P::EP; ++e)
          if (e == ep) x = piece_get<T>(fl.pc[i], e);
        dst[q] = x;
      }
    }
  } else {
    T x = fl.tail[0];
#pragma unroll
    for (int t = 1; t < P::EP - 1; ++t)
      if (t == mp - nfull) x = fl.tail[t];
    dst[p] = x;
  }
}

// Band-edge segments of a full-length wave: lane 0's (a-1, b) column and
// lane H-1's (a+1, b) column, both kLW nodes, moved as 2 x LPC pieces by
// lanes 0 .. 2*LPC-1 of one transfer.  `eseg` = memory index of the lowest
// node of the lane's band-edge segment (lanes 0 and H-1; -1: none).
// Lane 0's segment holds new values (read from r_new), lane H-1's old
// values (read from r_old).
template <typename T>
__device__ __forceinline__ u32x4 edge_issue(rsrc_t r_new, rsrc_t r_old, int eseg, int hm1, int p) {
  using P = Pieces<T>;
  const int src = p < P::LPC ? 0 : hm1;
  const int sq = __shfl(eseg, src);
  const int m0 = (p % P::LPC) * P::EP;
  const bool ok = p < 2 * P::LPC && sq >= 0;
  const uint32_t off = ok ? (uint32_t)(sq + m0) * (uint32_t)sizeof(T) : kOOB;
  const u32x4 a = bload4(r_new, p < P::LPC ? off : kOOB);
  const u32x4 b = bload4(r_old, p < P::LPC ? kOOB : off);
  return p < P::LPC ? a : b;
}

// Writes the band-edge pieces: lane 0's segment to plane n position 0,
// lane H-1's to plane x position H + 1 (canonical rows).
template <typename T>
__device__ __forceinline__ void edge_land(const u32x4& pc, lptr<T> pn, lptr<T> px, int H,
                                          bool fwd, int p) {
  using P = Pieces<T>;
  if (p < 2 * P::LPC) {
    const int m0 = (p % P::LPC) * P::EP;
    lptr<T> dst = p < P::LPC ? pn : px + H + 1;
#pragma unroll
    for (int e = 0; e < P::EP; ++e) dst[canon<true>(m0 + e, kLW, fwd) * P::RS] = piece_get<T>(pc, e);
  }
}

}  // namespace

// Point-to-point step synchronisation of the waves of a sweep.  Wave w's
// step t reads the slots its neighbours w-1 (xch) and w+1 (xold) wrote at
// their step t-1, and at its end overwrites the slots they read at their
// step t-1: so before step t a wave waits until both neighbours have
// completed t steps (prog[] counts completed steps), and after its own
// writes it publishes t+1.  Waves drift at most one step apart instead of
// meeting at a workgroup barrier every step.  Every wait is bounded: past
// kSpinCap polls it sets *abort, after which no wave waits any more, the
// kernel runs to its end and reports a fault.
constexpr int kSpinCap = 1 << 16;

__device__ __forceinline__ void wait_neighbours(int* prog, int* abort_flag, int w, int nseg,
                                                int t) {
  if (t == 0) return;
  for (int spin = 0;; ++spin) {
    const int a = w > 0 ? __hip_atomic_load(&prog[w - 1], __ATOMIC_RELAXED,
                                            __HIP_MEMORY_SCOPE_WORKGROUP)
                        : t;
    const int b = w + 1 < nseg ? __hip_atomic_load(&prog[w + 1], __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_WORKGROUP)
                               : t;
    if (a >= t && b >= t) break;
    if (__hip_atomic_load(abort_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) break;
    if (spin >= kSpinCap) {
      __hip_atomic_store(abort_flag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      break;
    }
    __builtin_amdgcn_s_sleep(1);
  }
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

__device__ __forceinline__ void publish_progress(int* prog, int w, int done) {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __hip_atomic_store(&prog[w], done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// One sweep of one event by the whole workgroup (the steps of wave w).
//
// Every global memory instruction of the step loop is issued on every step
// (inactive lanes and steps use out-of-range offsets), and a step's new
// values are stored at the start of the next step, before any load is
// waited for.  So the compiler's wait counts stay exact on every path into
// the loop body: landing a load never waits for a store issued after it.
// FC: the sweep runs c downwards (flips the k axis).
template <typename T, bool FULL, bool MIX, bool ERR, bool CH, bool FC>
__device__ __forceinline__ uint64_t diag_sweep_wave(const DSweep<T>& S, int w, int cnt,
                                                    lptr<T> xch, lptr<T> xold, lptr<T> planes,
                                                    int* prog, int* flags) {
  using P = Pieces<T>;
  const DGeom& g = S.g;
  const int p = threadIdx.x & 63;
  const int L = g.L, Lw = g.Lw;
  if (FULL) cnt = kLW;
  const int c0 = w * Lw;      // canonical start of this wave's segment
  constexpr bool fwd = !FC;  // the segment runs forward in memory
  const bool lane_first = (p == 0);
  const bool lane_last = (p == g.H - 1);
  const bool has_x = c0 + Lw < L && w + 1 < S.nseg;  // a node exists one past the segment
  // planes of this wave: new values of v-1 (then v), old values of v+1,
  // slowness of v
  lptr<T> pn = planes + w * 3 * P::PLANE;
  lptr<T> px = pn + P::PLANE;
  lptr<T> pf = px + P::PLANE;

  // Memory index of the lowest node of the segment of column `col`.
  auto seg_of = [&](const DCol& col) -> int {
    const int lo = fwd ? col.node0 + c0 : col.node0 - c0 - (cnt - 1);
    return (col.valid() && cnt > 0) ? lo : -1;
  };
  // Band-edge neighbour of lane 0 ((a-1, b), new) and lane H-1 ((a+1, b),
  // old) of column `col`, one load per node (canonical order).
  auto load_edge = [&](T(&buf)[kLW], const DCol& col) {
    const bool eu = lane_first && col.valid() && col.a > 0;
    const bool ed = lane_last && col.valid() && col.a + 1 < g.Ma;
    T bo[kLW];
    const int tA = col.fa() ? -g.sA : g.sA;
    load_seg(buf, S.u, col, -tA, eu, c0, cnt, L, S.tC);
    load_seg(bo, S.src, col, tA, ed, c0, cnt, L, S.tC);
#pragma unroll
    for (int j = 0; j < kLW; ++j) buf[j] = eu ? buf[j] : bo[j];
  };

  // Registers: OWN = old values of diagonal v in, new values out.  FD, FF,
  // EF / LE are the loads in flight for the next step.  The old value one
  // past the segment comes from the next wave, which publishes canonical
  // node 0 of its transfer in flight (old values of our diagonal v at
  // step T-1) to xold.
  T OWN[kLW], LE[kLW];  // LE: band-edge values (general variant)
  u32x4 EF;             // band-edge pieces (full-length variant)
  Flight<T> FD, FF;
  Flight<T> FS;         // ERR: snapshot values of the segment stored next step
  uint64_t key = 0;     // ERR: max |new - snapshot| key of this lane
  // CH (k in several chunks): wave 0's c-1 neighbour (new, end of the
  // previous chunk) and the last segment's c+1 neighbour (old, start of the
  // next chunk) come from memory, one value per lane and step, loaded one
  // step ahead.  Without chunks both are grid edges.
  const bool need_cm = CH && w == 0 && S.cs > 0;
  const bool need_cp = CH && !has_x && cnt > 0 && S.cs + L < S.Lfull;
  auto load_cm = [&](const DCol& col) -> T {
    if constexpr (!CH) return T(0);
    return bload(S.u, boff<T>(need_cm && col.valid(), col.node0 - S.tC), T(0));
  };
  auto load_cp = [&](const DCol& col) -> T {
    if constexpr (!CH) return T(0);
    return bload(S.src, boff<T>(need_cp && col.valid(), col.node0 + L * S.tC), T(0));
  };
  T XCM, XCP;  // values for the current diagonal, landed at step start
  // band-edge segment of lane 0 / lane H-1 for column `col` (full variant)
  auto eseg_of = [&](const DCol& col) -> int {
    const bool eu = lane_first && col.valid() && col.a > 0;
    const bool ed = lane_last && col.valid() && col.a + 1 < g.Ma;
    const int tA = col.fa() ? -g.sA : g.sA;
    const int d = eu ? -tA : tA;
    const int lo = fwd ? col.node0 + d + c0 : col.node0 + d - c0 - (kLW - 1);
    return (eu || ed) ? lo : -1;
  };
  auto issue_edge = [&](const DCol& col) {
    if constexpr (FULL)
      EF = edge_issue<T>(S.u, S.src, eseg_of(col), g.H - 1, p);
    else
      load_edge(LE, col);
  };

  const int v0 = -w;  // diagonal of step 0
  DCur cur = cur_at(S, v0, p);
  DCol cv = cur_col(S, cur, p);
  cur_next(S, cur);
  DCol cv1 = cur_col(S, cur, p);
  cur_next(S, cur);
  DCol cv2 = cur_col(S, cur, p);
  cur_next(S, cur);  // at diagonal v0 + 3
  // prologue: plane x = old(v0); in flight: old(v0+1), f(v0), edges(v0);
  // node 0 of old(v0+1) published for the wave upstream (its step 0)
  seg_issue<T, FULL>(FD, S.src, seg_of(cv), cnt, p);
  seg_land<T, FULL>(FD, px, cnt, fwd, p);
  // a step whose inputs hold NaN or -0 takes the exact node pass (the fast
  // one orders such values differently); so does the step after it, which
  // reads the values that step landed or produced
  int redo_prev = __builtin_amdgcn_ballot_w64(flight_odd(FD, FULL)) != 0;
  if (S.first && __builtin_amdgcn_ballot_w64(flight_wild(FD, FULL)) != 0 && p == 0)
    __hip_atomic_store(&flags[1], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  seg_issue<T, FULL>(FD, S.src, seg_of(cv1), cnt, p);
  seg_issue<T, FULL>(FF, S.f, seg_of(cv), cnt, p);
  issue_edge(cv);
  T XCM_n = load_cm(cv), XCP_n = load_cp(cv);
  if constexpr (ERR) seg_issue<T, FULL>(FS, S.snap_src, -1, cnt, p);
  seg_publish<T, FULL>(FD, xold + kDW * 64 + w * 64, cnt, fwd, p);
  if (p == 0) __hip_atomic_store(&prog[w], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  __syncthreads();

  const int steps = (S.nch - 1) * S.V0 + g.V + S.nseg - 1;
  const int vend = (S.nch - 1) * S.V0 + g.V;  // chain diagonals
  int seg_prev = -1;   // segment of the previous step's new values (-1: none)
  int segs_prev = -1;  // ... if its column is on the chain's last sweep
  // ES: a step of the stop-rule sweep (ERR), which the chain reaches at
  // diagonal (nch-1) V0 (this wave's step (nch-1) V0 + w); earlier steps
  // of its chain take the plain step
  auto step = [&](int t, auto es) {
    constexpr bool ES = decltype(es)::value;
    const int v = v0 + t;
    const bool active = v >= 0 && v < vend;
    // new values of v-1 (plane n) to memory
    if constexpr (ES) {
      seg_store_err<T, FULL>(pn, S.u, S.snap_dst, seg_prev, segs_prev, cnt, fwd, p, FS, key);
      // snapshot values of diagonal v (its last-sweep lanes), stored at the
      // next step
      seg_issue<T, FULL>(FS, S.snap_src, (active && cv.last()) ? seg_of(cv) : -1, cnt, p);
    } else {
      seg_store<T, FULL>(pn, S.u, seg_prev, cnt, fwd, p);
    }
    // rotate by one diagonal: OWN <- old(v) from plane x, which then
    // receives old(v+1); plane f receives f(v)
#pragma unroll
    for (int j = 0; j < kLW; ++j)
      if (FULL || j < cnt) OWN[j] = px[j * P::RS + p + 1];
    bool odd = flight_odd(FD, FULL);
    // the event's first sweep lands every node's u0 and f once: a value with
    // which the solve may produce NaN sets flags[1]
    if (S.first && __builtin_amdgcn_ballot_w64((int)flight_wild(FD, FULL) |
                                               (int)flight_wild(FF, FULL)) != 0 && p == 0)
      __hip_atomic_store(&flags[1], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    seg_land<T, FULL>(FD, px, cnt, fwd, p);
    seg_land<T, FULL>(FF, pf, cnt, fwd, p);
    XCM = XCM_n;
    XCP = XCP_n;
    // band-edge neighbours of v at lanes -1 (plane n) and H (plane x)
    if constexpr (FULL) {
#pragma unroll
      for (int e = 0; e < P::EP; ++e) odd = (int)odd | (int)Real<T>::odd_(piece_get<T>(EF, e));
      edge_land<T>(EF, pn, px, g.H, fwd, p);
    } else {
#pragma unroll
      for (int j = 0; j < kLW; ++j)
        if (j < cnt) odd = (int)odd | (int)Real<T>::odd_(LE[j]);
      if (lane_first) {
#pragma unroll
        for (int j = 0; j < kLW; ++j)
          if (j < cnt) pn[j * P::RS] = LE[j];
      }
      if (lane_last) {
#pragma unroll
        for (int j = 0; j < kLW; ++j)
          if (j < cnt) px[j * P::RS + g.H + 1] = LE[j];
      }
    }

    // Node updates of diagonal v (lanes of inactive steps hold no valid
    // column and skip them).  The loads for the next step (old values of
    // v+2, slowness and band edges of v+1) are issued with the first
    // nodes, every step, so they have nearly a whole step to arrive.
    T xo = T(0);
    wait_neighbours(prog, flags, w, S.nseg, t);
    {
      const T xc_in = w > 0 ? xch[((t - 1) & 1) * kDW * 64 + (w - 1) * 64 + p] : XCM;
      const T xd_in = has_x ? xold[((t - 1) & 1) * kDW * 64 + (w + 1) * 64 + p] : XCP;
      // canonical c-1 of the segment's first node / c+1 of its last node
      // lie outside the grid: the axis minimum then takes the one existing
      // neighbour (Eikonal3D.cpp:51-52), as smin(v, v) = v bit for bit
      const bool c_first = w == 0 && S.cs == 0;
      const bool c_last = !has_x && S.cs + L == S.Lfull;
      const int ra = cv.fa() ? g.Ma - 1 - cv.a : cv.a;
      const int rb = cv.fb() ? g.Nb - 1 - cv.b : cv.b;
      const bool af = ra == 0, al = ra == g.Ma - 1;
      const bool bf = rb == 0, bl = rb == g.Nb - 1;
      // The a and b neighbours are read straight from the planes: the sweep
      // direction and which of a, b is the x axis only choose addresses.
      // Canonical a-1 (new) / a+1 (old) sit at positions p / p+2, canonical
      // b-1 (new) / b+1 (old) at position p+1 of planes n / x.
      const lptr<T> a_up = pn + p, a_dn = px + p + 2;
      const lptr<T> b_up = pn + p + 1, b_dn = px + p + 1;
      const lptr<T> a_plus = cv.fa() ? a_up : a_dn, a_minus = cv.fa() ? a_dn : a_up;
      const lptr<T> b_plus = cv.fb() ? b_up : b_dn, b_minus = cv.fb() ? b_dn : b_up;
      const lptr<T> x_plus0 = g.a_is_i ? a_plus : b_plus, x_minus0 = g.a_is_i ? a_minus : b_minus;
      const lptr<T> y_plus0 = g.a_is_i ? b_plus : a_plus, y_minus0 = g.a_is_i ? b_minus : a_minus;
      const bool xf = g.a_is_i ? af : bf, xl = g.a_is_i ? al : bl;
      const bool yf = g.a_is_i ? bf : af, yl = g.a_is_i ? bl : al;
      // One-sided neighbours at the grid edges (Eikonal3D.cpp:47-52): both
      // operands of the minimum read the one existing neighbour, and
      // smin(v, v) is v bit for bit (NaN included), so the per-node edge
      // selects become one address choice per step.
      const lptr<T> x_plus = xl ? x_minus0 : x_plus0, x_minus = xf ? x_plus0 : x_minus0;
      const lptr<T> y_plus = yl ? y_minus0 : y_plus0, y_minus = yf ? y_plus0 : y_minus0;
      // The node updates take single-instruction minima and the fast square
      // roots (Real::refine_).  If a lane met a discriminant the square
      // roots may get wrong, or NaN or -0 among the values the minima order
      // (landed this step, produced by the previous step or by this one),
      // the whole wave redoes the step's updates from the old values (still
      // in memory: this step's new values are stored at the next step) with
      // the reference's comparisons and the IEEE square root.
      bool unsafe = (int)odd | redo_prev | (int)Real<T>::odd_(xc_in) | (int)Real<T>::odd_(xd_in);
      auto nodes = [&](auto fast, auto issue) {
        constexpr bool FAST = decltype(fast)::value, ISSUE = decltype(issue)::value;
        // the updates in two blocks of half a segment each (a lane's nodes
        // share their column, so one branch covers a block, and the plane
        // reads of a node can be issued during the updates before it); the
        // next step's loads are issued before the first block (old values)
        // and between the blocks (slowness, band edges)
        auto half = [&](auto lo) {
          constexpr int E0 = decltype(lo)::value;
          if (cv.valid()) {
#pragma unroll
            for (int e = E0; e < E0 + kLW / 2; ++e) {
              if (FULL || e < cnt) {
                const int rw = e * P::RS;
                // std::min(u(+1), u(-1)) per axis, one-sided at the edges
                // (Eikonal3D.cpp:47-52)
                const T xp = x_plus[rw], xm = x_minus[rw], yp = y_plus[rw], ym = y_minus[rw];
                const T fe = pf[rw + p + 1];
                T uc = (e == 0) ? xc_in : OWN[e == 0 ? 0 : e - 1];
                T dc = (FULL ? e + 1 < kLW : e + 1 < cnt) ? OWN[e + 1 < kLW ? e + 1 : e] : xd_in;
                if (e == 0 && c_first) uc = dc;
                if ((FULL ? e == kLW - 1 : e == cnt - 1) && c_last) dc = uc;
                const T ux = FAST ? Real<T>::min1_(xp, xm) : smin(xp, xm);
                const T uy = FAST ? Real<T>::min1_(yp, ym) : smin(yp, ym);
                // std::min(u(k+1), u(k-1)): canonical c+1 is k+1 unless the
                // sweep runs k downwards
                const T mc = FAST ? Real<T>::min1_(uc, dc) : FC ? smin(uc, dc) : smin(dc, uc);
                T cand;
                if constexpr (MIX)  // FP32 storage, FP64 local solve rounded once
                  cand = (T)solution3d<double, FAST>(ux, uy, mc, fe, S.hd, unsafe);
                else
                  cand = solution3d<T, FAST>(ux, uy, mc, fe, S.h, unsafe);
                // a NaN or -0 candidate (which the single-instruction minimum
                // would order differently) sends the step to the exact pass
                if constexpr (FAST) unsafe = (int)unsafe | (int)Real<T>::odd_(cand);
                OWN[e] = FAST ? Real<T>::min1_(cand, OWN[e]) : smin(cand, OWN[e]);
              }
            }
          }
        };
        if constexpr (ISSUE) seg_issue<T, FULL>(FD, S.src, seg_of(cv2), cnt, p);
        half(std::integral_constant<int, 0>{});
        if constexpr (ISSUE) {
          seg_issue<T, FULL>(FF, S.f, seg_of(cv1), cnt, p);
          issue_edge(cv1);
          XCM_n = load_cm(cv1);
          XCP_n = load_cp(cv1);
        }
        half(std::integral_constant<int, kLW / 2>{});
#pragma unroll
        for (int e = 0; e < kLW; ++e)
          if (FULL ? e == kLW - 1 : e == cnt - 1) xo = OWN[e];
      };
      nodes(std::true_type{}, std::true_type{});
      redo_prev = __builtin_amdgcn_ballot_w64(unsafe) != 0;
      if (__builtin_expect(redo_prev, 0)) {
        load_seg(OWN, S.src, cv, 0, cv.valid(), c0, cnt, L, S.tC);
        nodes(std::false_type{}, std::false_type{});
        // a NaN result (a NaN candidate always comes here) sets flags[1]
        bool nan_out = false;
#pragma unroll
        for (int e = 0; e < kLW; ++e)
          if (FULL || e < cnt) nan_out = (int)nan_out | (int)(cv.valid() && OWN[e] != OWN[e]);
        if (__builtin_amdgcn_ballot_w64(nan_out) != 0 && p == 0)
          __hip_atomic_store(&flags[1], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      }
    }
    // plane n: new values of v (all reads of v-1 are done)
#pragma unroll
    for (int j = 0; j < kLW; ++j)
      if (FULL || j < cnt) pn[j * P::RS + p + 1] = OWN[j];
    // node 0 of old(v+2), which the wave upstream needs at step T+1
    seg_publish<T, FULL>(FD, xold + (t & 1) * kDW * 64 + w * 64, cnt, fwd, p);
    seg_prev = active ? seg_of(cv) : -1;
    segs_prev = cv.last() ? seg_prev : -1;
    xch[(t & 1) * kDW * 64 + w * 64 + p] = xo;
    cv = cv1;
    cv1 = cv2;
    cv2 = cur_col(S, cur, p);  // diagonal v + 3
    cur_next(S, cur);
    if (p == 0) publish_progress(prog, w, t + 1);
  };
  int t = 0;
  if constexpr (ERR) {
    const int t1 = (S.nch - 1) * S.V0 + w;
    for (; t < steps && t < t1; ++t) step(t, std::false_type{});
    for (; t < steps; ++t) step(t, std::true_type{});
    seg_store_err<T, FULL>(pn, S.u, S.snap_dst, seg_prev, segs_prev, cnt, fwd, p, FS, key);
  } else {
    for (; t < steps; ++t) step(t, std::false_type{});
    seg_store<T, FULL>(pn, S.u, seg_prev, cnt, fwd, p);
  }
  __syncthreads();  // the next sweep reads these values
  return key;
}

// One sweep of one event by the whole workgroup.  Waves whose segment has
// kLW nodes take the fixed-length variant; the others (and empty segments)
// the general one.  Both run the same number of steps and barriers.  Waves
// beyond the event's segment count only meet the sweep's two workgroup
// barriers.
// ERR (last sweep of an iteration): returns this lane's max |u - snapshot|
// key and writes the next snapshot.
template <typename T, bool MIX, bool ERR>
__device__ __forceinline__ uint64_t diag_sweep(const DSweep<T>& S, lptr<T> xch, lptr<T> xold,
                                               lptr<T> planes, int* prog, int* flags) {
  // the wave index is uniform across the wave: keep it (and everything
  // derived from it) in scalar registers
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  // with all kDW segments active, waves wv and wv + kDW/2 (which the
  // hardware places on one SIMD) take neighbouring segments
  const int w = S.nseg == kDW ? (wv % (kDW / 2)) * 2 + wv / (kDW / 2) : wv;
  if (w >= S.nseg) {  // the prologue's and the end's barriers of the active waves
    __syncthreads();
    __syncthreads();
    return 0;
  }
  int cnt = S.g.L - w * S.g.Lw;  // nodes in this wave's segment
  cnt = cnt < 0 ? 0 : (cnt > S.g.Lw ? S.g.Lw : cnt);
  auto run = [&](auto full, auto ch, auto fc) {
    return diag_sweep_wave<T, decltype(full)::value, MIX, ERR, decltype(ch)::value,
                           decltype(fc)::value>(S, w, cnt, xch, xold, planes, prog, flags);
  };
  auto by_c = [&](auto full, auto ch) {
    return S.flipC ? run(full, ch, std::true_type{}) : run(full, ch, std::false_type{});
  };
  auto by_len = [&](auto ch) {
    return cnt == kLW ? by_c(std::true_type{}, ch) : by_c(std::false_type{}, ch);
  };
  return S.Lfull > S.nseg * kLW ? by_len(std::true_type{}) : by_len(std::false_type{});
}

// The whole forward of one event, run by one workgroup of kDW waves.
// The stop rule's err as the reference folds it (Eikonal3D.cpp:79-81):
// err = std::max(fabs(u[j] - u_old[j]), err) over j in memory order, where a
// NaN difference makes err NaN and the next difference replaces it, so err
// is the maximum over the nodes after the last NaN difference (NaN if that
// is the last node).  Two passes of the workgroup: the last NaN difference,
// then the maximum after it, writing the next snapshot dst = u (dst may be
// snap: each node is read before it is written, by the same thread).
template <typename T>
__device__ T stop_err_in_order(const T* u, const T* snap, T* dst, int64_t N,
                               unsigned long long* last_nan, uint64_t* red) {
  if (threadIdx.x == 0) *last_nan = 0;
  __syncthreads();
  unsigned long long last = 0;  // 1 + index of the last NaN difference seen
  for (int64_t j = threadIdx.x; j < N; j += blockDim.x) {
    const T d = Real<T>::fabs_(u[j] - snap[j]);
    if (d != d) last = (unsigned long long)j + 1;
  }
  if (last) atomicMax(last_nan, last);
  __syncthreads();
  const unsigned long long from = *last_nan;  // the first index that counts
  uint64_t key = 0;
  for (int64_t j = threadIdx.x; j < N; j += blockDim.x) {
    const T uj = u[j];
    if ((unsigned long long)j >= from) {
      const uint64_t k = maxkey(Real<T>::fabs_(uj - snap[j]));
      key = k > key ? k : key;
    }
    dst[j] = uj;
  }
  for (int off = 32; off > 0; off >>= 1) {
    const uint64_t o = __shfl_xor(key, off);
    key = o > key ? o : key;
  }
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = key;
  __syncthreads();
  uint64_t r = 0;
  for (int q = 0; q < (int)(blockDim.x >> 6); ++q) r = red[q] > r ? red[q] : r;
  __syncthreads();
  return from == (unsigned long long)N ? T(__builtin_nan("")) : key_to_real(r, T(0));
}

template <typename T, bool MIX>
__device__ __forceinline__ void fwd3_diag_event(const EventDesc& ev, int max_iter) {
  __shared__ T xch[2 * kDW * 64];
  __shared__ T xold[2 * kDW * 64];
  __shared__ T planes[kDW * 3 * Pieces<T>::PLANE];
  __shared__ uint64_t red[kDW];
  __shared__ int prog[kDW];
  // [0]: a bounded wait ran out (the kernel reports a fault); [1]: NaN may
  // occur in u (an input Real::wild_ flags, or a NaN result): the stop rule
  // is then evaluated in memory order, as the reference does
  __shared__ int flags[2];
  __shared__ unsigned long long last_nan;
  const int m = ev.m, n = ev.n, l = ev.l;
  if (!ev.fwd_diag) return;  // handled by the k-streaming kernel
  const int64_t N = (int64_t)m * n * l;
  const uint64_t bytes = (uint64_t)N * sizeof(T);
  if (threadIdx.x < 2) flags[threadIdx.x] = 0;
  __syncthreads();
  DSweep<T> S;
  S.g = make_dgeom(m, n, l);
  S.Lfull = l;
  S.nseg = (ev.fwd_segs >= 1 && ev.fwd_segs <= kDW) ? ev.fwd_segs : kDW;
  // the k axis is swept in chunks of nseg * kLW nodes (the last one
  // shorter), in sweep order: a chunk's nodes read the previous chunk's as
  // new values and the next chunk's as old ones, as the lexicographic sweep
  // does
  const int kChunk = S.nseg * kLW;
  const int nchunk = (l + kChunk - 1) / kChunk;
  // h in a scalar pair of its own: read straight from the descriptor it is
  // one part of a wider scalar load, and the step loop would reload the
  // whole wider tuple from spill slots for every node
  double hd = ev.h;
  asm("" : "+s"(hd));
  S.h = (T)hd;
  S.hd = hd;
  // (read by the waves only after the barriers of the first sweep)
  if (threadIdx.x == 0 && Real<T>::wild_(S.h)) flags[1] = 1;
  S.u = make_rsrc(ev.u, bytes);
  S.u0 = make_rsrc(ev.u0, bytes);
  S.f = make_rsrc(ev.f, bytes);
  S.snap_dst = make_rsrc(ev.uold, bytes);
  const T tol = (T)ev.tol;
  int it = 0;
  T err = T(0);
  S.V0 = S.g.bands * S.g.Nb;
  // may sweep sw+1 run chained after sweep sw (see "Chained sweeps")?
  auto chains = [&](int it_, int sw) -> bool {
    if (sw + 1 > 7 || (it_ == 0 && sw == 0) || nchunk != 1) return false;
    // the stop-rule sweep runs alone when it must be evaluated in order
    // (flags[1] is read after a barrier: the same value in every lane)
    if (sw + 1 == 7 && __builtin_amdgcn_readfirstlane(flags[1])) return false;
    const int flip = kDir3[sw] ^ kDir3[sw + 1];
    if (flip & 4) return false;
    const bool flips_a = (flip & 1) ? S.g.a_is_i : !S.g.a_is_i;
    return flips_a ? S.g.Nb >= S.g.H + 3 : S.g.bands >= 2;
  };
  for (it = 0; it < max_iter; ++it) {
    uint64_t key = 0;
    bool in_order = false;  // err from stop_err_in_order
    for (int sw = 0; sw < 8;) {
      int nch = 1;
      while (nch < kChainMax && chains(it, sw + nch - 1)) ++nch;
      const int fk = kDir3[sw] & 4;
      S.flipC = fk;
      S.tC = S.flipC ? -1 : 1;
      S.nch = nch;
      // u = u0 (Eikonal3D.cpp:72): the first sweep reads its old values
      // from u0; every node is written once per sweep
      S.src = (it == 0 && sw == 0) ? S.u0 : S.u;
      const bool last = sw + nch - 1 == 7;  // the chain ends with the stop-rule sweep
      if (last) S.snap_src = make_rsrc(it == 0 ? ev.u0 : ev.uold, bytes);
      if (last && __builtin_amdgcn_readfirstlane(flags[1])) in_order = true;
      S.first = it == 0 && sw == 0;
      for (int ch = 0; ch < nchunk; ++ch) {
        S.cs = ch * kChunk;
        const int Lc = l - S.cs < kChunk ? l - S.cs : kChunk;
        S.g.L = Lc;
        S.g.Lw = (Lc + S.nseg - 1) / S.nseg;
        S.cfl = 0;
        for (int q = 0; q < nch; ++q) {
          const int dir = kDir3[sw + q];
          const bool fi = dir & 1, fj = dir & 2;
          const bool fa = S.g.a_is_i ? fi : fj, fb = S.g.a_is_i ? fj : fi;
          S.cfl |= ((fa ? 1 : 0) | (fb ? 2 : 0)) << (2 * q);
        }
        S.coff = S.flipC ? l - 1 - S.cs : S.cs;
        if (!last || in_order) {
          diag_sweep<T, MIX, false>(S, (lptr<T>)xch, (lptr<T>)xold, (lptr<T>)planes, prog,
                                    flags);
        } else {
          // err = max |u - u_old| (Eikonal3D.cpp:79-81), taken as the last
          // sweep stores each node, which also writes u_old = u for the
          // next iteration (u_old of the first iteration is u0 itself)
          const uint64_t kc =
              diag_sweep<T, MIX, true>(S, (lptr<T>)xch, (lptr<T>)xold, (lptr<T>)planes, prog,
                                       flags);
          key = kc > key ? kc : key;
        }
      }
      sw += nch;
    }
    if (in_order) {
      err = stop_err_in_order((const T*)ev.u, (const T*)(it == 0 ? ev.u0 : ev.uold),
                              (T*)ev.uold, N, &last_nan, red);
    } else {
      for (int off = 32; off > 0; off >>= 1) {
        uint64_t o = __shfl_xor(key, off);
        key = o > key ? o : key;
      }
      if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = key;
      __syncthreads();
      uint64_t r = 0;
      for (int q = 0; q < kDW; ++q) r = red[q] > r ? red[q] : r;
      __syncthreads();
      err = key_to_real(r, T(0));
    }
    if (threadIdx.x == 0 && it < kMaxLoggedIters) ev.info->err_log[it] = (double)err;
    if (err < tol) {
      ++it;
      break;
    }
  }
  if (threadIdx.x == 0) {
    ev.info->iterations = it > max_iter ? max_iter : it;
    ev.info->converged = (err < tol) ? 1 : 0;
    ev.info->err = (double)err;
    if (flags[0]) ev.info->fault = 1;
  }
}

template <typename T, bool MIX>
__global__ void __launch_bounds__(kDW * 64) fwd3_diag_kernel(const EventDesc* __restrict__ evs,
                                                             int max_iter) {
  fwd3_diag_event<T, MIX>(evs[blockIdx.x], max_iter);
}

// Forward and then adjoint of one event per workgroup (with the event's
// receiver misfit in between when one is attached: its gradient is then
// the adjoint source), for batches whose every event takes the diagonal
// forward: an event whose forward converges
// in fewer iterations starts its adjoint while others still sweep, so the
// memory-bound traversal overlaps the latency-bound sweeps.  The adjoint
// runs with kDW waves, four queue entries and four prep-pass nodes per
// thread (the forward's register budget leaves room for them).
template <typename T, bool MIX>
__global__ void __launch_bounds__(kDW * 64) fwd_adj3_diag_kernel(const EventDesc* __restrict__ evs,
                                                                 int max_iter) {
  const EventDesc& ev = evs[blockIdx.x];
  fwd3_diag_event<T, MIX>(ev, max_iter);
  // the next phase reads the travel times the waves of this workgroup stored
  __threadfence();
  __syncthreads();
  __threadfence();
  if (ev.mis) {
    // receiver misfit of these travel times: its gradient is the adjoint
    // source (same for every thread of the workgroup)
    misfit_event<T>(*ev.mis, *ev.mnodes, ev.rx, ev.ry, ev.rz, (int)ev.nrec);
    __threadfence();
    __syncthreads();
    __threadfence();
  }
  adj3_kahn_event<T, 4, 4>(ev);
}

template __global__ void fwd3_diag_kernel<double, false>(const EventDesc*, int);
template __global__ void fwd3_diag_kernel<float, false>(const EventDesc*, int);
template __global__ void fwd3_diag_kernel<float, true>(const EventDesc*, int);
template __global__ void fwd_adj3_diag_kernel<double, false>(const EventDesc*, int);
template __global__ void fwd_adj3_diag_kernel<float, false>(const EventDesc*, int);
template __global__ void fwd_adj3_diag_kernel<float, true>(const EventDesc*, int);

}  // namespace adt
