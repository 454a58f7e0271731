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
// The c axis is cut into kDW segments of Lw <= kLW nodes, one per wave.  At
// step T wave w processes diagonal v = T - w over its segment, node by node
// in increasing c.  Each node's neighbours then come from:
//   (a-1, b), (a, b-1): the lane's and its left neighbour's new values of
//       diagonal v-1 (registers of the previous step; DPP lane shift), or,
//       across a band edge, memory (stored delta - 1 >= 32 steps before);
//   (a+1, b), (a, b+1): old values of diagonal v+1, loaded one step ahead
//       (DPP lane shift); across a band edge from memory (not yet updated);
//   (a, b, c-1): the lane's previous node, or for the first node of the
//       segment wave w-1's last node of diagonal v (step T-1, through LDS);
//   (a, b, c+1): the lane's next node, or one past the segment (loaded with
//       the segment; wave w+1 updates it only at step T+1).
// One workgroup barrier per step; every lane updates Lw nodes per step.
//
// Loads for diagonals v+1 (slowness, band-edge values) and v+2 (old travel
// times) are issued at the start of step T and used at step T+1, so each
// step hides one step of memory latency.  Memory instructions are never
// skipped: masked lanes use an out-of-range buffer offset.
//
// Iteration (Eikonal3D.cpp:71-88): u = u0 is a copy pass at the start; the
// snapshot u_old is written during the first sweep of each later
// iteration; err = max|u - u_old| is a coalesced pass after the eighth sweep
// against the snapshot (u0 for the first iteration).
#include "buffer_ops.hpp"
#include "eikonal_device.hpp"
#include "kernels.hpp"

namespace adt {

namespace {

template <typename T>
using lptr = __attribute__((address_space(3))) T*;

constexpr int kDW = 8;  // waves per workgroup = c segments
constexpr int kLW = 8;  // maximum segment length (nodes per lane per step)

// DPP wavefront shifts (gfx9 family): lane i reads lane i-1 / i+1.
constexpr int kDppWaveShr1 = 0x138;
constexpr int kDppWaveShl1 = 0x130;

template <int CTRL>
__device__ __forceinline__ double lane_shift(double x) {
  const uint64_t b = (uint64_t)__double_as_longlong(x);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)b, CTRL, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)(b >> 32), CTRL, 0xf, 0xf, false);
  return __longlong_as_double((long long)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo));
}
template <int CTRL>
__device__ __forceinline__ float lane_shift(float x) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), CTRL, 0xf, 0xf, false));
}

struct DGeom {
  int Ma, Nb, L, H, bands, V, Lw;
  int64_t sA, sB;  // memory strides of the a and b axes
  bool a_is_i;
};

__device__ inline DGeom make_dgeom(int m, int n, int l) {
  DGeom g;
  g.a_is_i = (m <= n);
  g.Ma = g.a_is_i ? m : n;
  g.Nb = g.a_is_i ? n : m;
  g.sA = g.a_is_i ? (int64_t)n * l : (int64_t)l;
  g.sB = g.a_is_i ? (int64_t)l : (int64_t)n * l;
  g.L = l;
  g.Lw = (l + kDW - 1) / kDW;
  g.bands = (g.Ma + 63) / 64;
  g.H = (g.Ma + g.bands - 1) / g.bands;
  g.V = g.H - 1 + g.bands * g.Nb;
  return g;
}

// Column of lane p on diagonal v.
struct DCol {
  int a, b;
  bool valid;
  int64_t node0;  // flat index of canonical (a, b, c = 0)
};

__device__ __forceinline__ DCol dcol(const DGeom& g, int v, int p, int64_t base0, int64_t tA,
                                     int64_t tB) {
  DCol c;
  const int q = v - p;
  const int qq = q > 0 ? q : 0;
  const int beta = qq / g.Nb;
  c.b = qq - beta * g.Nb;
  c.a = beta * g.H + p;
  c.valid = (p < g.H) && (q >= 0) && (beta < g.bands) && (c.a < g.Ma);
  c.node0 = base0 + (int64_t)c.a * tA + (int64_t)c.b * tB;
  return c;
}

template <typename T>
struct DSweep {
  DGeom g;
  int64_t tA, tB, tC, base0;
  bool flipA, flipB, flipC;
  T h;
  rsrc_t u;     // travel times (read old / write new in place)
  rsrc_t f;
  rsrc_t snap;  // snapshot written during the first sweep of iteration >= 1
};

// cnt values of column `col` from canonical c0 on (element j <- c0 + j).
template <typename T, int NE>
__device__ __forceinline__ void load_seg(T (&buf)[NE], rsrc_t r, const DCol& col, int64_t delta,
                                         bool ok, int c0, int cnt, int L, int64_t tC) {
#pragma unroll
  for (int j = 0; j < NE; ++j) {
    const int c = c0 + j;
    buf[j] = bload(r, boff<T>(ok && j < cnt && c < L, col.node0 + delta + (int64_t)c * tC), T(0));
  }
}

template <typename T>
__device__ __forceinline__ void store_seg(const T (&buf)[kLW + 1], rsrc_t r, const DCol& col,
                                          int c0, int cnt, int L, int64_t tC) {
#pragma unroll
  for (int j = 0; j < kLW; ++j) {
    const int c = c0 + j;
    bstore(r, boff<T>(col.valid && j < cnt && c < L, col.node0 + (int64_t)c * tC), buf[j]);
  }
}

}  // namespace

// One sweep of one event by the whole workgroup.  `snap`: also store every
// node's old value to the snapshot.
template <typename T>
__device__ __forceinline__ void diag_sweep(const DSweep<T>& S_in, bool snap, lptr<T> xch) {
  // Local copy: keeps the fields in registers.
  const DSweep<T> S = S_in;
  const DGeom& g = S.g;
  const int w = threadIdx.x >> 6, p = threadIdx.x & 63;
  const int L = g.L, Lw = g.Lw;
  const int c0 = w * Lw;  // canonical start of this wave's segment
  const bool lane_first = (p == 0);
  const bool lane_last = (p == g.H - 1);

  // Registers: PREV = new values of diagonal v-1, OWN = diagonal v (old in,
  // new out; element Lw is the old value one past the segment), NEXT = old
  // values of v+1, F / EDGE = slowness and band-edge values of v.  LD, LF,
  // LE are the loads in flight for the next step.
  T PREV[kLW], OWN[kLW + 1], NEXT[kLW + 1], F[kLW], EDGE[kLW];
  T LD[kLW + 1], LF[kLW], LE[kLW];

  // Band-edge neighbour of lane 0 ((a-1, b), new) and lane H-1 ((a+1, b),
  // old) of column `col`, loaded into `buf`.
  auto load_edge = [&](T(&buf)[kLW], const DCol& col) {
    const bool eu = lane_first && col.valid && col.a > 0;
    const bool ed = lane_last && col.valid && col.a + 1 < g.Ma;
    load_seg(buf, S.u, col, eu ? -S.tA : S.tA, eu || ed, c0, Lw, L, S.tC);
  };

  const int v0 = -w;  // diagonal of step 0
  DCol cv = dcol(g, v0, p, S.base0, S.tA, S.tB);
  DCol cv1 = dcol(g, v0 + 1, p, S.base0, S.tA, S.tB);
  DCol cv2 = dcol(g, v0 + 2, p, S.base0, S.tA, S.tB);
  // prologue: OWN(v0) arrives through NEXT, NEXT(v0+1) through LD
  load_seg(NEXT, S.u, cv, 0, cv.valid, c0, Lw + 1, L, S.tC);
  load_seg(LD, S.u, cv1, 0, cv1.valid, c0, Lw + 1, L, S.tC);
  load_seg(LF, S.f, cv, 0, cv.valid, c0, Lw, L, S.tC);
  load_edge(LE, cv);
#pragma unroll
  for (int e = 0; e < kLW; ++e) {
    OWN[e] = T(0);
    PREV[e] = T(0);
  }
  OWN[kLW] = T(0);

  const int steps = g.V + kDW - 1;
  for (int t = 0; t < steps; ++t) {
    const int v = v0 + t;
    // rotate the register pipeline by one diagonal
#pragma unroll
    for (int e = 0; e < kLW; ++e) {
      PREV[e] = OWN[e];
      OWN[e] = NEXT[e];
      NEXT[e] = LD[e];
      F[e] = LF[e];
      EDGE[e] = LE[e];
    }
    OWN[kLW] = NEXT[kLW];
    NEXT[kLW] = LD[kLW];
    // loads for the next step: old values of v+2, slowness and edges of v+1
    load_seg(LD, S.u, cv2, 0, cv2.valid, c0, Lw + 1, L, S.tC);
    load_seg(LF, S.f, cv1, 0, cv1.valid, c0, Lw, L, S.tC);
    load_edge(LE, cv1);

    if (v >= 0 && v < g.V) {
      if (snap) store_seg(OWN, S.snap, cv, c0, Lw, L, S.tC);
      const T xc_in = w > 0 ? xch[((t - 1) & 1) * kDW * 64 + (w - 1) * 64 + p] : T(0);
      const int ra = S.flipA ? g.Ma - 1 - cv.a : cv.a;
      const int rb = S.flipB ? g.Nb - 1 - cv.b : cv.b;
      const bool af = ra == 0, al = ra == g.Ma - 1;
      const bool bf = rb == 0, bl = rb == g.Nb - 1;
      T xo = T(0);
#pragma unroll
      for (int e = 0; e < kLW; ++e) {
        if (e < Lw) {
          const int c = c0 + e;
          // cross-lane moves run with the full wave active
          const T sl = lane_shift<kDppWaveShr1>(PREV[e]);
          const T sr = lane_shift<kDppWaveShl1>(NEXT[e]);
          const T ua = lane_first ? EDGE[e] : sl;
          const T da = lane_last ? EDGE[e] : sr;
          const T ub = PREV[e], db = NEXT[e];
          const T uc = (e == 0) ? xc_in : OWN[e == 0 ? 0 : e - 1];
          const T dc = OWN[e + 1];
          if (cv.valid && c < L) {
            const int rc = S.flipC ? L - 1 - c : c;
            const T ma = axis_min(ua, da, S.flipA, af, al);
            const T mb = axis_min(ub, db, S.flipB, bf, bl);
            const T mc = axis_min(uc, dc, S.flipC, rc == 0, rc == L - 1);
            const T ux = g.a_is_i ? ma : mb;
            const T uy = g.a_is_i ? mb : ma;
            const T cand = solution3d(ux, uy, mc, F[e], S.h);
            OWN[e] = smin(cand, OWN[e]);
          }
          if (e == Lw - 1) xo = OWN[e];
        }
      }
      store_seg(OWN, S.u, cv, c0, Lw, L, S.tC);
      xch[(t & 1) * kDW * 64 + w * 64 + p] = xo;
    }
    cv = cv1;
    cv1 = cv2;
    cv2 = dcol(g, v + 3, p, S.base0, S.tA, S.tB);
    __syncthreads();
  }
}

template <typename T>
__global__ void __launch_bounds__(kDW * 64) fwd3_diag_kernel(const EventDesc* __restrict__ evs,
                                                             int max_iter) {
  __shared__ T xch[2 * kDW * 64];
  __shared__ uint64_t red[kDW];
  const EventDesc& ev = evs[blockIdx.x];
  const int m = ev.m, n = ev.n, l = ev.l;
  if (l > kDW * kLW) return;  // handled by the k-streaming kernel
  const int64_t N = (int64_t)m * n * l;
  const uint64_t bytes = (uint64_t)N * sizeof(T);
  {  // u = u0 (Eikonal3D.cpp:72)
    const gptr<const T> s = (gptr<const T>)ev.u0;
    const gptr<T> d = (gptr<T>)ev.u;
    for (int64_t q = threadIdx.x; q < N; q += blockDim.x) d[q] = s[q];
  }
  __syncthreads();
  DSweep<T> S;
  S.g = make_dgeom(m, n, l);
  S.h = (T)ev.h;
  S.u = make_rsrc(ev.u, bytes);
  S.f = make_rsrc(ev.f, bytes);
  S.snap = make_rsrc(ev.uold, bytes);
  const T tol = (T)ev.tol;
  int it = 0;
  T err = T(0);
  for (it = 0; it < max_iter; ++it) {
    for (int sw = 0; sw < 8; ++sw) {
      const int dir = kDir3[sw];
      const bool fi = dir & 1, fj = dir & 2, fk = dir & 4;
      S.flipA = S.g.a_is_i ? fi : fj;
      S.flipB = S.g.a_is_i ? fj : fi;
      S.flipC = fk;
      S.tA = S.flipA ? -S.g.sA : S.g.sA;
      S.tB = S.flipB ? -S.g.sB : S.g.sB;
      S.tC = S.flipC ? -1 : 1;
      S.base0 = (S.flipA ? (int64_t)(S.g.Ma - 1) * S.g.sA : 0) +
                (S.flipB ? (int64_t)(S.g.Nb - 1) * S.g.sB : 0) + (S.flipC ? (int64_t)(l - 1) : 0);
      diag_sweep<T>(S, sw == 0 && it > 0, (lptr<T>)xch);
    }
    // err = max |u - u_old| over the event (coalesced pass after the 8 sweeps)
    uint64_t key = 0;
    {
      const gptr<const T> ua = (gptr<const T>)ev.u;
      const gptr<const T> sr = (gptr<const T>)((it == 0) ? ev.u0 : ev.uold);
      for (int64_t q = threadIdx.x; q < N; q += blockDim.x) {
        const uint64_t kq = maxkey(Real<T>::fabs_(ua[q] - sr[q]));
        key = kq > key ? kq : key;
      }
    }
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
  }
}

template __global__ void fwd3_diag_kernel<double>(const EventDesc*, int);
template __global__ void fwd3_diag_kernel<float>(const EventDesc*, int);

}  // namespace adt
