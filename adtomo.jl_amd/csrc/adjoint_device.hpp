// Node-level pieces of the adjoint shared by the adjoint kernels.
//
// Per-node upwind code (one byte), computed from the converged travel times:
//   3D: bit 6 = Z (u == u0, or no active dimension; Eikonal3D.cpp:126-130,
//       168-171); bits 2d..2d+1 for dimension d (x, y, z): 0 inactive,
//       1 argmin neighbour at -1, 2 argmin neighbour at +1 (Eikonal3D.cpp:139-166).
//   2D: bits 0-1 x side, bits 2-3 y side, bit 6 singular (non-source node
//       without an active dimension), bit 7 source (Eikonal.h:107-185).
#pragma once

#include "eikonal_device.hpp"

namespace adt {

template <typename T>
__device__ __forceinline__ void adj3_prep_node(const T* __restrict__ u, const T* __restrict__ u0,
                                               const T* __restrict__ g, int m, int n, int l,
                                               int64_t p, uint8_t* __restrict__ code,
                                               T* __restrict__ lam, T* __restrict__ grad_u0) {
  const int64_t nl = (int64_t)n * l;
  const int i = (int)(p / nl), j = (int)((p / l) % n), k = (int)(p % l);
  const T up = u[p];
  lam[p] = T(0);
  if (grad_u0) grad_u0[p] = (up == u0[p]) ? g[p] : T(0);
  if (up == u0[p]) {
    code[p] = 0x40;
    return;
  }
  uint8_t cd = 0;
  const int co[3] = {i, j, k};
  const int ex[3] = {m, n, l};
  const int64_t st[3] = {nl, l, 1};
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    T umin;
    int side;
    if (co[d] == 0) {
      umin = u[p + st[d]];
      side = 2;
    } else if (co[d] == ex[d] - 1) {
      umin = u[p - st[d]];
      side = 1;
    } else {
      T vp = u[p + st[d]], vm = u[p - st[d]];
      umin = smin(vp, vm);
      side = (vp > vm) ? 1 : 2;
    }
    if (up > umin) cd |= (uint8_t)(side << (2 * d));
  }
  code[p] = cd ? cd : (uint8_t)0x40;
}

template <typename T>
__device__ __forceinline__ void adj2_prep_node(const T* __restrict__ u, int m, int n, int ix,
                                               int jx, int64_t p, uint8_t* __restrict__ code,
                                               T* __restrict__ lam,
                                               unsigned long long* n_singular) {
  const int64_t W = (int64_t)m + 1;
  const int i = (int)(p % W), j = (int)(p / W);
  lam[p] = T(0);
  if (i == ix && j == jx) {
    code[p] = 0x80;
    return;
  }
  uint8_t cd = 0;
  const T up = u[p];
  T a;
  int sa;
  if (i == 0) {
    a = u[j * W + 1];
    sa = 2;
  } else if (i == m) {
    a = u[j * W + m - 1];
    sa = 1;
  } else {
    const bool left = u[j * W + i + 1] > u[j * W + i - 1];
    a = left ? u[j * W + i - 1] : u[j * W + i + 1];
    sa = left ? 1 : 2;
  }
  if (up > a) cd |= (uint8_t)sa;
  T b;
  int sb;
  if (j == 0) {
    b = u[W + i];
    sb = 2;
  } else if (j == n) {
    b = u[(int64_t)(n - 1) * W + i];
    sb = 1;
  } else {
    const bool upn = u[(int64_t)(j + 1) * W + i] > u[(int64_t)(j - 1) * W + i];
    b = upn ? u[(int64_t)(j - 1) * W + i] : u[(int64_t)(j + 1) * W + i];
    sb = upn ? 1 : 2;
  }
  if (up > b) cd |= (uint8_t)(sb << 2);
  if (!cd) {
    cd = 0x40;
    atomicAdd(n_singular, 1ull);
  }
  code[p] = cd;
}

}  // namespace adt
