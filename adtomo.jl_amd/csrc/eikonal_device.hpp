// Device-side numerics of the Eikonal forward/adjoint (gfx950, CDNA4).
//
// Every expression keeps the operand order of the reference op so that, built
// with -ffp-contract=off (no FMA contraction) and IEEE-correct sqrt/divide,
// FP64 results are the reference CPU op's results bit for bit.
//   std::min(x, y) in the reference is `(y < x) ? y : x`  -> smin(x, y)
//   std::max(x, y) in the reference is `(x < y) ? y : x`  -> smax(x, y)
// The argument order is kept at each call site: it decides which operand
// survives a NaN.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace adt {

// Pointers into device (global) memory.  Pointers read from descriptor
// tables are generic; generic (flat) accesses share the wait counter with
// LDS accesses, so kernels convert them to this type once.
template <typename T>
using gptr = __attribute__((address_space(1))) T*;

template <typename T> struct Real;
template <> struct Real<double> {
  static __device__ __forceinline__ double sqrt_(double x) { return __builtin_sqrt(x); }
  // Fast IEEE square root for the local solve's discriminants.  The
  // compiler expands __builtin_sqrt into a reciprocal square root estimate
  // refined by Newton steps (g ~ sqrt x, h ~ 1 / (2 sqrt x), then two
  // residual corrections), wrapped in a rescaling for x < 2^-767 and a
  // fix-up for +-0 and +inf.  refine_ runs the same sequence of operations
  // without the wrapper: the same bits for every x with |x| >= 2^-767 below
  // +inf (the wrapper is then the identity; negative x gives NaN in both).
  // risky_ flags the other operands (|x| < 2^-767, +inf), for which the
  // caller takes __builtin_sqrt instead, and NaN (the fast local solve's
  // discriminant is NaN where its shortcuts could differ from the
  // reference's expressions).
  static __device__ __forceinline__ double refine_(double x) {
    const double r = __builtin_amdgcn_rsq(x);
    double g = x * r, h = r * 0.5;
    const double e = __builtin_fma(-h, g, 0.5);
    g = __builtin_fma(g, e, g);
    h = __builtin_fma(h, e, h);
    double d = __builtin_fma(-g, g, x);
    g = __builtin_fma(d, h, g);
    d = __builtin_fma(-g, g, x);
    return __builtin_fma(d, h, g);
  }
  static __device__ __forceinline__ bool risky_(double x) {
    // class bits sNaN, qNaN, +inf
    return (int)(__builtin_fabs(x) < 0x1p-767) | (int)__builtin_amdgcn_class(x, 0x203);
  }
  static constexpr bool kFastSqrt = true;
  static __device__ __forceinline__ double fabs_(double x) { return __builtin_fabs(x); }
  // Minimum / maximum in one instruction (__builtin_fmin adds a NaN-quieting
  // instruction per loaded operand).  They equal smin / smax bit for bit
  // unless an operand is NaN or the operands are zeros of opposite signs.
  static __device__ __forceinline__ double min1_(double x, double y) {
    double r;
    asm("v_min_f64 %0, %1, %2" : "=v"(r) : "v"(x), "v"(y));
    return r;
  }
  static __device__ __forceinline__ double max1_(double x, double y) {
    double r;
    asm("v_max_f64 %0, %1, %2" : "=v"(r) : "v"(x), "v"(y));
    return r;
  }
  // NaN or -0 (class bits sNaN, qNaN, -0): the values on which min1_ /
  // max1_ may differ from the compare-and-select forms
  static __device__ __forceinline__ bool odd_(double x) { return __builtin_amdgcn_class(x, 0x23); }
  // NaN, +-inf or |x| >= 2^240: an input (u0, f, h) with which the solve
  // may produce NaN.  Below it (f h < 2^480, and an update never raises a
  // value), travel times stay far below 2^511 on any grid that fits in
  // memory, so every square in the solve stays finite.
  static __device__ __forceinline__ bool wild_(double x) { return !(__builtin_fabs(x) < 0x1p240); }
};
template <> struct Real<float> {
  static __device__ __forceinline__ float sqrt_(float x) { return __builtin_sqrtf(x); }
  static __device__ __forceinline__ float refine_(float x) { return __builtin_sqrtf(x); }
  static __device__ __forceinline__ bool risky_(float) { return false; }
  static constexpr bool kFastSqrt = false;
  static __device__ __forceinline__ float fabs_(float x) { return __builtin_fabsf(x); }
  static __device__ __forceinline__ float min1_(float x, float y) {
    float r;
    asm("v_min_f32 %0, %1, %2" : "=v"(r) : "v"(x), "v"(y));
    return r;
  }
  static __device__ __forceinline__ float max1_(float x, float y) {
    float r;
    asm("v_max_f32 %0, %1, %2" : "=v"(r) : "v"(x), "v"(y));
    return r;
  }
  static __device__ __forceinline__ bool odd_(float x) { return __builtin_amdgcn_classf(x, 0x23); }
  static __device__ __forceinline__ bool wild_(float x) { return !(__builtin_fabsf(x) < 0x1p20f); }
};

template <typename T> __device__ __forceinline__ T smin(T x, T y) { return (y < x) ? y : x; }
template <typename T> __device__ __forceinline__ T smax(T x, T y) { return (x < y) ? y : x; }

// deps/CustomOps/Eikonal/Eikonal.h:14-20 (`solution`).
template <typename T>
__device__ __forceinline__ T solution2d(T a, T b, T f, T h) {
  T d = Real<T>::fabs_(a - b);
  if (d >= f * h) return smin(a, b) + f * h;
  return (a + b + Real<T>::sqrt_(T(2.0) * f * f * h * h - (a - b) * (a - b))) / T(2.0);
}

// x / 3 without a division: with y = RN(1/3), q0 = RN(x y), the residual
// r = x - 3 q0 (exact in an FMA: it is a small multiple of q0's quantum) and
// q = RN(q0 + r y).  q = RN(x / 3) for every finite x except -0: x is a
// multiple of a quantum Q at least that of the result, so x / 3 = k Q / 3
// lies at least (result quantum) / 6 from every rounding midpoint, while
// q0 + r y differs from x / 3 by r (y - 1/3), below 2^-50 of that quantum;
// the argument holds in gradual underflow too.  For x = +-inf the residual
// is inf - inf = NaN, which the minimum with 2^1023 (IEEE minNum returns the
// number) replaces by a finite value, so q = q0 = +-inf; the minimum leaves
// every finite residual as it is (|r| is at most about 2^-52 |x| < 2^972),
// and NaN stays NaN.  (2^1023 is an instruction literal; the largest finite
// number would take a scalar register pair through the step loops.)
// Checked against IEEE division on every binary32 input and on 2e8 binary64
// inputs (half of them subnormal or in the lowest 60 binades).  The one
// exception: -0 gives +0.
template <typename T> struct Div3;
template <> struct Div3<double> {
  static __device__ __forceinline__ double go(double x) {
    const double y = 0x1.5555555555555p-2;
    const double q0 = x * y;
    const double r = __builtin_fmin(__builtin_fma(-3.0, q0, x), 0x1p1023);
    return __builtin_fma(r, y, q0);
  }
  // Without the minimum: the same for finite x, NaN for x = +-inf.
  static __device__ __forceinline__ double raw(double x) {
    const double y = 0x1.5555555555555p-2;
    const double q0 = x * y;
    return __builtin_fma(__builtin_fma(-3.0, q0, x), y, q0);
  }
};
template <> struct Div3<float> {
  static __device__ __forceinline__ float go(float x) {
    const float y = 0x1.555556p-2f;
    const float q0 = x * y;
    const float r = __builtin_fminf(__builtin_fmaf(-3.0f, q0, x), 0x1.fffffep127f);
    return __builtin_fmaf(r, y, q0);
  }
};
template <typename T>
__device__ __forceinline__ T div3(T x) { return Div3<T>::go(x); }

// deps/CustomOps/Eikonal3D/Eikonal3D.cpp:11-28 (`calculate_unique_solution`).
// The three candidates (1D, 2D, 3D) are evaluated as written and the first
// that the reference would return is selected: nearly every update of a wave
// needs the 3D one, so evaluating all of them costs nothing and lets the 2D
// and 3D square roots run side by side.  FAST: the three values sorted by
// minima and maxima (the reference's compare-and-swap order gives the same
// bits when none of a1, a2, a3 is NaN or -0, which the caller checks), and
// in FP64 the square roots through Real::refine_, with `unsafe` set for a
// lane whose discriminants they may get wrong; the caller then redoes the
// update with FAST off.
template <typename T, bool FAST = false>
__device__ __forceinline__ T solution3d(T a1, T a2, T a3, T f, T h, bool& unsafe) {
  if constexpr (FAST) {
    // lo = min, hi = max, median = max(min(a1, a2), min(max(a1, a2), a3))
    const T mn = Real<T>::min1_(a1, a2), mx = Real<T>::max1_(a1, a2);
    const T md = Real<T>::max1_(mn, Real<T>::min1_(mx, a3));
    a1 = Real<T>::min1_(mn, a3);
    a3 = Real<T>::max1_(mx, a3);
    a2 = md;
  } else {
    T temp;
    if (a1 > a2) { temp = a1; a1 = a2; a2 = temp; }
    if (a1 > a3) { temp = a1; a1 = a3; a3 = temp; }
    if (a2 > a3) { temp = a2; a2 = a3; a3 = temp; }
  }
  if constexpr (FAST && Real<T>::kFastSqrt) {
    // The reference's expressions (below) with three shortcuts that keep
    // every unflagged result (D2, D3 >= 2^-767, finite, not NaN):
    //   D2 = B2 B2 - 4 ((q - fhh) / 2) as fma(-2, q - fhh, B2 B2): the same
    //     whenever halving q - fhh is exact; it is not only for
    //     |q - fhh| < 2^-1021, which moves the exact difference by less than
    //     2^-1019 and so changes no rounding of a D2 >= 2^-767;
    //   D3 = B3 B3 - 4 C3 as fma(-4, C3, B3 B3): 4 C3 is exact unless it
    //     overflows, and then D3 is +inf (flagged) or negative in both;
    //   div3 without its minimum: the same for finite arguments, and an
    //     infinite one makes D3 NaN (flagged).
    const T x1 = a1 + f * h;
    const T fhh = f * f * h * h;
    const T q = a1 * a1 + a2 * a2;
    const T B2 = -(a1 + a2);
    const T B3 = Div3<T>::raw(T(-2.0) * (a1 + a2 + a3));
    const T C3 = Div3<T>::raw(q + a3 * a3 - fhh);
    const T D2 = __builtin_fma(T(-2.0), q - fhh, B2 * B2);
    const T D3 = __builtin_fma(T(-4.0), C3, B3 * B3);
    const T s2 = Real<T>::refine_(D2), s3 = Real<T>::refine_(D3);
    unsafe = (int)unsafe | (int)Real<T>::risky_(D2) | (int)Real<T>::risky_(D3);
    const T x2 = (-B2 + s2) / T(2.0);
    const T x3 = (-B3 + s3) / T(2.0);
    return (x1 <= a2) ? x1 : ((x2 <= a3) ? x2 : x3);
  }
  const T x1 = a1 + f * h;
  const T B2 = -(a1 + a2);
  const T C2 = (a1 * a1 + a2 * a2 - f * f * h * h) / T(2.0);
  // div3's one exception (-0 -> +0) leaves x3 unchanged: B3 enters x3 as
  // B3 * B3 and as -B3 + s with s = sqrt(...) never -0, and C3's argument is
  // never -0 (a difference of two non-negative numbers)
  const T B3 = div3(T(-2.0) * (a1 + a2 + a3));
  const T C3 = div3(a1 * a1 + a2 * a2 + a3 * a3 - f * f * h * h);
  const T D2 = B2 * B2 - T(4.0) * C2, D3 = B3 * B3 - T(4.0) * C3;
  const T s2 = Real<T>::sqrt_(D2), s3 = Real<T>::sqrt_(D3);
  const T x2 = (-B2 + s2) / T(2.0);
  const T x3 = (-B3 + s3) / T(2.0);
  return (x1 <= a2) ? x1 : ((x2 <= a3) ? x2 : x3);
}
template <typename T>
__device__ __forceinline__ T solution3d(T a1, T a2, T a3, T f, T h) {
  bool unused = false;
  return solution3d<T, false>(a1, a2, a3, f, h, unused);
}

// Combines the canonical upwind (new) / downwind (old) values of one axis
// into the reference's neighbour minimum along the real axis
// (std::min(u(+1), u(-1)), one-sided at the edges).
template <typename T>
__device__ __forceinline__ T axis_min(T up, T dn, bool flip, bool first, bool last) {
  const T plus = flip ? up : dn;
  const T minus = flip ? dn : up;
  return first ? plus : (last ? minus : smin(plus, minus));
}

// Bit pattern of a value (for exact "changed?" tests that also hold for NaN).
__device__ __forceinline__ uint64_t bits_of(double x) { return __double_as_longlong(x); }
__device__ __forceinline__ uint64_t bits_of(float x) { return (uint32_t)__float_as_uint(x); }

// Order-preserving key of a travel-time difference for max-reductions with
// integer atomics.  Non-negative IEEE values order like their bit patterns;
// NaN maps to 0, so NaN differences are left out of the maximum: the
// reference's fold `err = std::max(fabs(u - u_old), err)` (Eikonal3D.cpp:79-81,
// in memory order) whenever no NaN difference occurs.  (With NaN differences
// the reference's err is the maximum after the last one; the diagonal forward
// then evaluates it in memory order instead.)
__device__ __forceinline__ uint64_t maxkey(double x) {
  return (x == x) ? (uint64_t)__double_as_longlong(x) : 0ull;
}
__device__ __forceinline__ uint64_t maxkey(float x) {
  return (x == x) ? (uint64_t)__float_as_uint(x) : 0ull;
}
__device__ __forceinline__ double key_to_real(uint64_t k, double) { return __longlong_as_double((long long)k); }
__device__ __forceinline__ float key_to_real(uint64_t k, float) { return __uint_as_float((uint32_t)k); }

// The 8 sweep orderings of Eikonal3D.cpp:59-68, in order.  Packed as bits
// (bit 0: dI<0, bit 1: dJ<0, bit 2: dK<0).
__device__ __constant__ static const int kDir3[8] = {0, 1, 3, 2, 6, 4, 5, 7};
// The 4 orderings of Eikonal.h:74-77: (I,J), (iI,J), (iI,iJ), (I,iJ).
__device__ __constant__ static const int kDir2[4] = {0, 1, 3, 2};

}  // namespace adt
