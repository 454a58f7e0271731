// 3D forward: k-streaming fast sweeping, one workgroup per event.
//
// Schedule.  A sweep (Eikonal3D.cpp:35-57) is equivalent to processing the
// hyperplanes a+b+c = s in increasing s, where (a, b, c) are the indices
// counted along the sweep direction (c along k, the contiguous axis).  Each
// lane owns a column (a, b) and streams it along c, one node per step,
// keeping the column in registers (16-node windows).  The columns of one
// wave lie on one "virtual anti-diagonal" v, so all lanes of a wave are at
// the same c at every step; wave v runs one step behind wave v-1, which is
// the hyperplane order.  Neighbour values cross waves through LDS once per
// step (one workgroup barrier per step).
//
// Column mapping.  Let b run along the longer of the two non-k axes (N_b
// nodes) and a along the other (M_a nodes).  Lanes p < H (H <= 64) of a
// wave cover rows a = H*beta + p of band beta; lane p of virtual diagonal v
// owns column b = (v - p) mod N_b of band beta = (v - p) div N_b.  Its
// upwind neighbours (a-1, b) and (a, b-1) then sit on diagonal v-1 (lane p-1
// and p) or, across a band edge, on diagonal v - (N_b - H + 1); its downwind
// neighbours on v+1 or v + (N_b - H + 1).  Waves 0..W-1 of the workgroup take
// diagonals v = g*W .. g*W+W-1 for g = 0, 1, ... (a new group starts when a
// wave finishes its column), so within a group neighbours are one step apart
// (LDS), and across groups or bands they are at least Lp - W + 1 steps apart
// and go through global memory, where they have been stored already.
//
// Memory instructions are never skipped: loads and stores of lanes or nodes
// outside the event use an out-of-range buffer offset, which the hardware
// turns into a zero load / dropped store.  Unconditional memory operations
// keep the compiler's wait counts exact, so a wait for a prefetched value
// does not also wait for younger loads.
//
// Iteration (Eikonal3D.cpp:71-88): u = u0 is fused into the first sweep
// (it reads u0 and writes u); the snapshot u_old is written during the first
// sweep of each later iteration; err = max|u - u_old| is a coalesced pass
// after the eighth sweep against the snapshot (u0 for the first iteration).
#include "buffer_ops.hpp"
#include "eikonal_device.hpp"
#include "kernels.hpp"

namespace adt {

namespace {

template <typename T>
using lptr = __attribute__((address_space(3))) T*;

constexpr int kW = 8;       // waves per workgroup
constexpr int kC = 16;      // window length (nodes along k)
constexpr int kPF = 4;      // prefetch distance (steps) of the global neighbour reads

struct Geom {
  int Ma, Nb, L, H, Lp, G, delta;
  int64_t sA, sB;  // memory strides of the a and b axes
  bool a_is_i;
};

__device__ inline Geom make_geom(int m, int n, int l) {
  Geom g;
  g.a_is_i = (m <= n);
  g.Ma = g.a_is_i ? m : n;
  g.Nb = g.a_is_i ? n : m;
  g.sA = g.a_is_i ? (int64_t)n * l : (int64_t)l;
  g.sB = g.a_is_i ? (int64_t)l : (int64_t)n * l;
  g.L = l;
  if (g.Nb <= 64) {
    g.H = g.Nb;  // band edge neighbours on the adjacent diagonal
  } else if (g.Nb >= 64 + kW - 1) {
    g.H = 64;
  } else {
    g.H = g.Nb - kW + 1;  // keep band-edge neighbours >= W diagonals apart
  }
  g.delta = g.Nb - g.H + 1;
  int Lp = ((l + kC - 1) / kC) * kC;
  const int minLp = ((kW + kC + kPF + 2 + kC - 1) / kC) * kC;  // stores visible across groups
  g.Lp = Lp > minLp ? Lp : minLp;
  const int bands = (g.Ma + g.H - 1) / g.H;
  const int V = (g.H - 1) + (g.Nb - 1) + (bands - 1) * g.Nb + 1;  // virtual diagonals
  g.G = (V + kW - 1) / kW;
  return g;
}

// Column of lane p on virtual diagonal v, with the canonical-origin node index.
struct Col {
  int a, b;
  bool valid;
  int64_t node0;  // flat index of canonical (a, b, c = 0)
};

__device__ __forceinline__ Col col_of(const Geom& g, int v, int p, int64_t base0, int64_t tA,
                                      int64_t tB) {
  Col c;
  const int q = v - p;
  c.valid = (p < g.H) && (q >= 0);
  const int qq = q > 0 ? q : 0;
  const int beta = qq / g.Nb;
  c.b = qq - beta * g.Nb;
  c.a = beta * g.H + p;
  c.valid = c.valid && (c.a < g.Ma);
  c.node0 = base0 + (int64_t)c.a * tA + (int64_t)c.b * tB;
  return c;
}

template <typename T>
struct Sweep {
  Geom g;
  int64_t tA, tB, tC, base0;
  bool flipA, flipB, flipC;
  T h;
  double hd;      // h for the FP64 local solve of the FP32-storage variant
  rsrc_t src;     // old values (u0 in the very first sweep)
  rsrc_t dst;     // u
  rsrc_t f;
  rsrc_t snap_w;  // snapshot written at chunk start (first sweep of it >= 1)
};

template <typename T, int NE>
__device__ __forceinline__ void load_n(T (&buf)[NE], rsrc_t r, const Col& col, int c0, int L,
                                       int64_t tC) {
#pragma unroll
  for (int e = 0; e < NE; ++e)
    buf[e] = bload(r, boff<T>(col.valid && c0 + e < L, col.node0 + (int64_t)(c0 + e) * tC), T(0));
}

template <typename T>
__device__ __forceinline__ void store_chunk(const T (&buf)[kC], rsrc_t r, const Col& col, int c0,
                                            int L, int64_t tC) {
#pragma unroll
  for (int e = 0; e < kC; ++e)
    bstore(r, boff<T>(col.valid && c0 + e < L, col.node0 + (int64_t)(c0 + e) * tC), buf[e]);
}

// Per-wave state of one sweep.
template <typename T>
struct Lane {
  int w, p;
  bool up_glob_a, up_glob_b, dn_glob_a, dn_glob_b;
  int sp, dp;   // LDS lanes of the upwind-a / downwind-a neighbours
  int wu, wd;   // LDS waves of the upwind / downwind diagonal (clamped)
  int step;     // global step index (parity of the LDS buffers)
  // prefetch cursor (kPF steps ahead of the stream)
  Col pcol;
  int pgroup, pc;
  bool nUA, nUB, nDA, nDB;  // which neighbours of pcol this lane reads globally
  T pUA[kPF], pUB[kPF], pDA[kPF], pDB[kPF];
};

template <typename T>
__device__ __forceinline__ void set_prefetch_col(const Sweep<T>& S, Lane<T>& ln, int group) {
  const Geom& g = S.g;
  ln.pgroup = group;
  ln.pc = 0;
  ln.pcol = col_of(g, group * kW + ln.w, ln.p, S.base0, S.tA, S.tB);
  const bool v = ln.pcol.valid && group < g.G;
  ln.nUA = v && ln.up_glob_a && ln.pcol.a > 0;
  ln.nUB = v && ln.up_glob_b && ln.pcol.b > 0;
  ln.nDA = v && ln.dn_glob_a && ln.pcol.a + 1 < g.Ma;
  ln.nDB = v && ln.dn_glob_b && ln.pcol.b + 1 < g.Nb;
}

// Loads the globally-read neighbours of the cursor node into `slot`, then
// advances the cursor by one stream position.
template <typename T>
__device__ __forceinline__ void glob_fetch(const Sweep<T>& S, Lane<T>& ln, int slot) {
  const Geom& g = S.g;
  const bool inc = ln.pc < g.L;
  const int64_t node = ln.pcol.node0 + (int64_t)ln.pc * S.tC;
  ln.pUA[slot] = bload(S.dst, boff<T>(inc && ln.nUA, node - S.tA), T(0));
  ln.pUB[slot] = bload(S.dst, boff<T>(inc && ln.nUB, node - S.tB), T(0));
  ln.pDA[slot] = bload(S.src, boff<T>(inc && ln.nDA, node + S.tA), T(0));
  ln.pDB[slot] = bload(S.src, boff<T>(inc && ln.nDB, node + S.tB), T(0));
  if (++ln.pc == g.Lp) set_prefetch_col(S, ln, ln.pgroup + 1);
}

// Processes one 16-node chunk of the lane's stream held in A (old values in,
// new values out); B holds the next chunk (old values).  The slowness comes
// in two 8-node halves: F0 (elements 0-7, loaded beforehand) and F1
// (elements 8-15, loaded at element 0); at element 8, F0 is refilled with
// the first half of the next chunk (column ncol, offset nc0).
template <typename T, bool MIX>
__device__ __forceinline__ void process_chunk(const Sweep<T>& S, Lane<T>& ln, const Col& col,
                                              bool afirst, bool alast, bool bfirst, bool blast,
                                              int c0, T (&A)[kC], const T (&B)[kC],
                                              T (&F0)[kC / 2], T (&F1)[kC / 2], const Col& ncol,
                                              int nc0, T& prev_new, lptr<T> lds_new,
                                              lptr<T> lds_old) {
  const Geom& g = S.g;
  const int w = ln.w, p = ln.p;
#pragma unroll
  for (int e = 0; e < kC; ++e) {
    const int c = c0 + e;
    const int slot = e % kPF;
    if (e == 0) load_n(F1, S.f, col, c0 + kC / 2, g.L, S.tC);
    if (e == kC / 2) load_n(F0, S.f, ncol, nc0, g.L, S.tC);
    const T gUA = ln.pUA[slot], gUB = ln.pUB[slot], gDA = ln.pDA[slot], gDB = ln.pDB[slot];
    glob_fetch(S, ln, slot);
    const int rp = (ln.step & 1) ^ 1;  // buffers written in the previous step
    const int wp = ln.step & 1;
    const T own = A[e];
    const T od_c = (e + 1 < kC) ? A[(e + 1) % kC] : B[0];
    const T od_c2 = (e + 2 < kC) ? A[(e + 2) % kC] : B[(e + 2) % kC];
    const T lua = lds_new[(rp * kW + ln.wu) * 64 + ln.sp];
    const T lub = lds_new[(rp * kW + ln.wu) * 64 + p];
    const T lda = lds_old[(rp * kW + ln.wd) * 64 + ln.dp];
    const T ldb = lds_old[(rp * kW + ln.wd) * 64 + p];
    const T ua = ln.up_glob_a ? gUA : lua;
    const T ub = ln.up_glob_b ? gUB : lub;
    const T da = ln.dn_glob_a ? gDA : lda;
    const T db = ln.dn_glob_b ? gDB : ldb;
    T un = own;
    if (col.valid && c < g.L) {
      const int rc = S.flipC ? g.L - 1 - c : c;
      const T ma = axis_min(ua, da, S.flipA, afirst, alast);
      const T mb = axis_min(ub, db, S.flipB, bfirst, blast);
      const T mc = axis_min(prev_new, od_c, S.flipC, rc == 0, rc == g.L - 1);
      const T ux = g.a_is_i ? ma : mb;
      const T uy = g.a_is_i ? mb : ma;
      const T fe = e < kC / 2 ? F0[e % (kC / 2)] : F1[e % (kC / 2)];
      T cand;
      if constexpr (MIX)  // FP32 storage, FP64 local solve rounded once
        cand = (T)solution3d<double>(ux, uy, mc, fe, S.hd);
      else
        cand = solution3d(ux, uy, mc, fe, S.h);
      un = smin(cand, own);
    }
    A[e] = un;
    prev_new = un;
    lds_new[(wp * kW + w) * 64 + p] = un;
    lds_old[(wp * kW + w) * 64 + p] = od_c2;
    ++ln.step;
    __syncthreads();
  }
}

}  // namespace

// One sweep of one event by the whole workgroup.
template <typename T, bool SNAP, bool MIX>
__device__ void stream_sweep(const Sweep<T>& S_in, lptr<T> lds_new, lptr<T> lds_old) {
  // Local copy: keeps the fields in registers instead of re-reading them
  // through a generic pointer after every store.
  const Sweep<T> S = S_in;
  const Geom& g = S.g;
  Lane<T> ln;
  ln.w = threadIdx.x >> 6;
  ln.p = threadIdx.x & 63;
  const int w = ln.w, p = ln.p;
  ln.up_glob_b = (w == 0);
  ln.dn_glob_b = (w == kW - 1);
  ln.up_glob_a = (w == 0) || (p == 0 && g.delta > 1);
  ln.dn_glob_a = (w == kW - 1) || (p == g.H - 1 && g.delta > 1);
  ln.sp = (p == 0) ? g.H - 1 : p - 1;
  ln.dp = (p == g.H - 1) ? 0 : p + 1;
  ln.wu = w > 0 ? w - 1 : 0;
  ln.wd = w < kW - 1 ? w + 1 : kW - 1;
  const int L = g.L, Lp = g.Lp;
  const int chunks = g.G * (Lp / kC);

  T X[kC], Y[kC], F0[kC / 2], F1[kC / 2];
  Col cur = col_of(g, w, p, S.base0, S.tA, S.tB);        // group 0
  Col nxt = col_of(g, kW + w, p, S.base0, S.tA, S.tB);   // group 1
  load_n(X, S.src, cur, 0, L, S.tC);
  load_n(F0, S.f, cur, 0, L, S.tC);
  load_n(Y, S.src, cur, kC, L, S.tC);  // Lp >= 2 * kC: chunk 1 is in the same column
  set_prefetch_col(S, ln, 0);
#pragma unroll
  for (int k = 0; k < kPF; ++k) glob_fetch(S, ln, k);

  // Wave w processes stream position s at global step T = w + 1 + s and, at
  // step T, publishes old(s + 2) for the upwind wave's next step.  Before its
  // first step it publishes old(0) (step w - 1) and old(1) (step w).
  ln.step = 0;
  for (; ln.step < w - 1; ++ln.step) __syncthreads();
  if (w >= 1) {
    lds_old[((ln.step & 1) * kW + w) * 64 + p] = X[0];
    ++ln.step;
    __syncthreads();
  }
  lds_old[((ln.step & 1) * kW + w) * 64 + p] = X[1];
  ++ln.step;
  __syncthreads();

  T prev_new = T(0);
  int gcur = 0;
  // boundary flags of the current column along the real a and b axes
  auto edges = [&](const Col& c, bool& af, bool& al, bool& bf, bool& bl) {
    const int ra = S.flipA ? g.Ma - 1 - c.a : c.a;
    const int rb = S.flipB ? g.Nb - 1 - c.b : c.b;
    af = ra == 0;
    al = ra == g.Ma - 1;
    bf = rb == 0;
    bl = rb == g.Nb - 1;
  };
  bool af, al, bf, bl;
  edges(cur, af, al, bf, bl);
  // One chunk: A holds it (old values), B the finished previous chunk (new
  // values, stored here) and then the following chunk (old values).
  auto do_chunk = [&](int jj, T(&A)[kC], T(&B)[kC]) {
    const int s0 = jj * kC;
    const int gs = s0 / Lp;
    const int c0 = s0 - gs * Lp;
    const Col prev = cur;
    if (gs != gcur) {  // this chunk starts a new column
      gcur = gs;
      cur = nxt;
      nxt = col_of(g, (gs + 1) * kW + w, p, S.base0, S.tA, S.tB);
      edges(cur, af, al, bf, bl);
    }
    const int sn = s0 + kC;
    const int gn = sn / Lp;
    const Col& cn = (gn == gs) ? cur : nxt;
    const int cn0 = sn - gn * Lp;
    if (jj > 0) {
      const int sp0 = s0 - kC;
      const int gp = sp0 / Lp;
      store_chunk(B, S.dst, gp == gs ? cur : prev, sp0 - gp * Lp, L, S.tC);
      load_n(B, S.src, cn, cn0, L, S.tC);
    }
    if (SNAP) store_chunk(A, S.snap_w, cur, c0, L, S.tC);
    process_chunk<T, MIX>(S, ln, cur, af, al, bf, bl, c0, A, B, F0, F1, cn, cn0, prev_new, lds_new,
                     lds_old);
  };
  for (int j = 0; j < chunks; j += 2) {
    do_chunk(j, X, Y);
    if (j + 1 < chunks) do_chunk(j + 1, Y, X);
  }
  {  // store the last chunk
    const int jl = chunks - 1;
    const int sl = jl * kC;
    const int gl = sl / Lp;
    if (jl & 1)
      store_chunk(Y, S.dst, cur, sl - gl * Lp, L, S.tC);
    else
      store_chunk(X, S.dst, cur, sl - gl * Lp, L, S.tC);
  }
  // trailing idle steps (every wave passes G*Lp + W barriers), then one more
  // so that the final stores are visible to the next sweep
  for (int t = w; t < kW - 1; ++t) __syncthreads();
  __syncthreads();
}

template <typename T, bool MIX>
__global__ void __launch_bounds__(kW * 64) fwd3_stream_kernel(const EventDesc* __restrict__ evs,
                                                              int max_iter) {
  __shared__ T lds_new[2 * kW * 64];
  __shared__ T lds_old[2 * kW * 64];
  __shared__ uint64_t red[kW];
  __shared__ unsigned long long last_nan;
  const EventDesc& ev = evs[blockIdx.x];
  const int m = ev.m, n = ev.n, l = ev.l;
  if (ev.fwd_diag) return;  // handled by the diagonal-streaming kernel
  const int64_t N = (int64_t)m * n * l;
  const uint64_t bytes = (uint64_t)N * sizeof(T);
  Sweep<T> S;
  S.g = make_geom(m, n, l);
  S.h = (T)ev.h;
  S.hd = ev.h;
  S.f = make_rsrc(ev.f, bytes);
  S.dst = make_rsrc(ev.u, bytes);
  S.snap_w = make_rsrc(ev.uold, bytes);
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
      S.src = make_rsrc((it == 0 && sw == 0) ? ev.u0 : ev.u, bytes);
      if (sw == 0 && it > 0)
        stream_sweep<T, true, MIX>(S, (lptr<T>)lds_new, (lptr<T>)lds_old);
      else
        stream_sweep<T, false, MIX>(S, (lptr<T>)lds_new, (lptr<T>)lds_old);
    }
    // err = max |u - u_old| over the event (coalesced pass after the 8
    // sweeps) as the reference folds it in memory order (Eikonal3D.cpp:79-81):
    // a NaN difference restarts the maximum, so with NaN differences err is
    // the maximum after the last one (NaN if that is the last node), taken
    // by a second pass
    uint64_t key = 0;
    unsigned long long from = 0;  // 1 + index of the last NaN difference
    {
      const gptr<const T> ua = (gptr<const T>)ev.u;
      const gptr<const T> sr = (gptr<const T>)((it == 0) ? ev.u0 : ev.uold);
      if (threadIdx.x == 0) last_nan = 0;
      __syncthreads();
      unsigned long long last = 0;
      for (int64_t q = threadIdx.x; q < N; q += blockDim.x) {
        const T d = Real<T>::fabs_(ua[q] - sr[q]);
        if (d != d) last = (unsigned long long)q + 1;
        const uint64_t kq = maxkey(d);
        key = kq > key ? kq : key;
      }
      if (last) atomicMax(&last_nan, last);
      __syncthreads();
      from = last_nan;
      if (from) {
        key = 0;
        for (int64_t q = (int64_t)from + threadIdx.x; q < N; q += blockDim.x) {
          const uint64_t kq = maxkey(Real<T>::fabs_(ua[q] - sr[q]));
          key = kq > key ? kq : key;
        }
      }
    }
    for (int off = 32; off > 0; off >>= 1) {
      uint64_t o = __shfl_xor(key, off);
      key = o > key ? o : key;
    }
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = key;
    __syncthreads();
    uint64_t r = 0;
    for (int q = 0; q < kW; ++q) r = red[q] > r ? red[q] : r;
    __syncthreads();
    err = from == (unsigned long long)N ? T(__builtin_nan("")) : key_to_real(r, T(0));
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

template __global__ void fwd3_stream_kernel<double, false>(const EventDesc*, int);
template __global__ void fwd3_stream_kernel<float, false>(const EventDesc*, int);
template __global__ void fwd3_stream_kernel<float, true>(const EventDesc*, int);

}  // namespace adt
