// Internal kernel descriptors and launch declarations.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/adtomo_hip.h"

namespace adt {

constexpr int kMaxLoggedIters = ADT_MAX_LOGGED_ITERS;

struct MisfitNodes;

// One 3D event as seen by the persistent kernels (device resident).
struct EventDesc {
  int m, n, l;
  // launch setting of this event (its box class's, or the global one):
  // forward on the diagonal-streaming kernel (else k-streaming), c segments
  // of that kernel (1..8), adjoint traversal (0 tiles, 1 node by node)
  int8_t fwd_diag, fwd_segs, adj_algo, pad;
  double h, tol;
  const void* u0;
  const void* f;
  const void* grad_u;
  void* u;
  void* uold;      // workspace: iteration snapshot [N]
  void* lam;       // workspace: adjoint multipliers [N]
  uint8_t* code;   // workspace: upwind codes [N]
  uint8_t* indeg;  // workspace: unresolved incoming count per node [N] (word-aligned)
  int32_t* queue;  // workspace: topological order of the adjoint [N]
  void* grad_u0;
  void* grad_f;
  adt_info* info;  // device
  // Optional (device, NULL if absent): the receiver misfit whose gradient
  // becomes grad_u between the forward and the adjoint (fused kernel), and
  // the receiver coordinates it reads.
  const adt_misfit3d* mis;
  const MisfitNodes* mnodes;
  const double* rx;
  const double* ry;
  const double* rz;
  int64_t nrec;
};

// One 2D event.
struct Event2Desc {
  int m, n, ix, jx;
  double h;
  const void* f;
  const void* grad_u;
  void* u;
  void* uold;
  void* lam;
  uint8_t* code;
  uint8_t* indeg;
  int32_t* queue;
  void* grad_f;
  adt_info* info;
};

// Objective kernels use the public descriptors directly (plain C layout).
using SourceDesc = adt_source3d;
using MisfitDesc = adt_misfit3d;

// A solve's receivers grouped by interpolation node (built on the host for
// each box of a receiver set): node[u] (flat index, increasing, u < nuniq)
// takes the entries ent[off[u] .. off[u+1]), each (receiver j) * 8 +
// corner (bit 2: the x1 / x2 side, bit 1: y, bit 0: z), in increasing j.
// gj: the solve's scratch of nrec adjoint seeds d loss / d t_j.
struct MisfitNodes {
  const int64_t* node;
  const int64_t* off;
  const int64_t* ent;
  int64_t nuniq;
  void* gj;
};

template <typename T>
__global__ void source_u0_kernel(const SourceDesc* src, const T* vel0, int m, int n, int l, T h);
template <typename T>
__global__ void misfit_kernel(const MisfitDesc* d, const MisfitNodes* nodes, const double* rx,
                              const double* ry, const double* rz, int nrec);
template <typename T>
__global__ void model_kernel(const T* x, const T* vel0, int64_t N, T* fvar, T* f);
template <typename T>
__global__ void sum_arrays_kernel(const T* const* in, int nin, int64_t N, T* out);
template <typename T>
__global__ void smooth_pass1_kernel(const T* fvar, int m, int n, int l, int rh, int rv, T wgt,
                                    T* sgn, T* partial);
template <typename T>
__global__ void smooth_pass2_kernel(const T* sgn, int m, int n, int l, int rh, int rv, T wgt,
                                    T lambda, T* grad, const T* partial, int npart, T* value);
template <typename T>
__global__ void slowness_kernel(const T* fvar, int64_t N, T scale, T* f_s);
template <typename T>
__global__ void slowness_grad_kernel(const T* fvar, int64_t N, T scale, const T* grad_fs,
                                     T* grad_fvar, T* partial);
template <typename T>
__global__ void sum_partials_kernel(const T* partial, int npart, T* out);
template <typename T>
__global__ void model_grad_kernel(const T* x, const T* fvar, const T* grad_f, const T* grad_fvar,
                                  int64_t N, T* grad_x);

template <typename T>
__global__ void fwd3_plane_kernel(T* u, const T* f, T h, int m, int n, int l, int dir, int s,
                                  int a_lo, int a_cnt);
// MIX (T = float only): FP32 storage with every local solve in FP64,
// rounded once to FP32.
template <typename T, bool MIX>
__global__ void fwd3_stream_kernel(const EventDesc* evs, int max_iter);
// Bytes of one node's record in the 3D adjoint workspace (lambda and three
// coefficients).
template <typename T>
constexpr size_t kAdjRecBytes = 4 * sizeof(T);
// Launch setting diag_max_l meaning "every k extent on fwd3_diag_kernel"
// (it sweeps the k axis in chunks of 64 nodes).
constexpr int kDiagAll = 1 << 30;
template <typename T, bool MIX>
__global__ void fwd3_diag_kernel(const EventDesc* evs, int max_iter);
// Forward then adjoint of each event in one workgroup (every event on the
// diagonal forward).
template <typename T, bool MIX>
__global__ void fwd_adj3_diag_kernel(const EventDesc* evs, int max_iter);
__global__ void info_clear_kernel(adt_info* info, int n);
template <typename T>
__global__ void strided_copy_kernel(const T* src, T* dst, int64_t N, int64_t d1, int64_t d2,
                                    int64_t s0, int64_t s1, int64_t s2, int to_rowmajor);
template <typename T>
__global__ void copy_kernel(T* dst, const T* src, int64_t N);
template <typename T>
__global__ void maxdiff_kernel(const T* a, const T* b, int64_t N, const unsigned long long* from,
                               unsigned long long* out);
template <typename T>
__global__ void lastnan_kernel(const T* a, const T* b, int64_t N, unsigned long long* last);
template <typename T>
__global__ void adj3_prep_kernel(const T* u, const T* u0, const T* g, int m, int n, int l,
                                 uint8_t* code, T* lam, T* grad_u0);
template <typename T>
__global__ void adj3_plane_kernel(const T* u, const T* g, const uint8_t* code, T* lam, int m,
                                  int n, int l, int dir, int s, int a_lo, int a_cnt, int* changed);
template <typename T>
__global__ void adj3_finish_kernel(const T* lam, const T* f, T h, int64_t N, T* grad_f);
template <typename T>
__global__ void fwd2_event_kernel(const Event2Desc* evs, int max_iter);
template <typename T>
__global__ void adj3_kahn_kernel(const EventDesc* evs);
// Tile-ordered 3D adjoint: NW waves per event, at most MAXT tiles of 8^3
// nodes.  Workspace per event: lam = multipliers [N] (T), code = upwind codes
// [N] (bytes), queue = per-tile external-edge counts [tiles] (int32).
template <typename T, int NW, int MAXT>
__global__ void adj3_tile_kernel(const EventDesc* evs);
constexpr int kTileEdge = 8;
constexpr int kTileSmallMax = 8192;   // tiles of the 8-wave FP64 variant
constexpr int kTileLargeMax = 65536;  // tiles of the large-event variants
constexpr int kTileWavesSmall = 8, kTileWavesLarge64 = 6, kTileWavesLarge32 = 8;
template <typename T>
__global__ void adj2_kahn_kernel(const Event2Desc* evs);
template <typename T>
__global__ void arith_kernel(int op, const T* a, const T* b, T* out, int64_t N);

}  // namespace adt
