// Device-side building blocks of the two-loop hot path (gfx950, HIP).
//
// Kernel base functions follow the reference's compile-time kernels
// (kernel/base_m6.f90:36-97, base_m4.f90:133-188, base_gauss.f90:228-271) in
// branch-free form: with t_k = max(k - q, 0) the piecewise polynomials collapse
// to one expression whose value in every piece equals the reference's (the
// dropped pieces contribute an exact 0).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sph_types.h"

namespace sph {

enum { KT_M4 = 0, KT_M6 = 1, KT_GAUSS = 2 };

// Reference constants (const.f90:4-8; base_*.f90 knorm/krad/fwc/maxneibnum).
constexpr double kPi = 3.141592653589793;
constexpr double kLightspeed = 2.997924e10;
constexpr double kRg = 8.314e7;
constexpr double kSigmaB = 5.67051e-5;

__host__ __device__ constexpr double krad_of(int kt) { return kt == KT_M4 ? 2.0 : (kt == KT_M6 ? 3.0 : 10.0); }

// Packed image code of a neighbour entry: (sx+1) + 3(sy+1) + 9(sz+1), 13 = no shift.
constexpr uint32_t kNoImage = 13;
constexpr int kCodeBits = 5;

__device__ __forceinline__ uint32_t nb_pack(uint32_t j, uint32_t code) { return (j << kCodeBits) | code; }
__device__ __forceinline__ uint32_t nb_index(uint32_t e) { return e >> kCodeBits; }
__device__ __forceinline__ uint32_t nb_code(uint32_t e) { return e & 31u; }

// Two FP32 values in the 8 bytes of one double slot (the mixed path's gathered
// j-records carry (x, y, z) in FP64 and two FP32 scalars in the fourth slot).
__device__ __forceinline__ double pack_f2(float lo, float hi) {
  return __hiloint2double(__float_as_int(hi), __float_as_int(lo));
}
__device__ __forceinline__ float f2_lo(double d) { return __int_as_float(__double2loint(d)); }
__device__ __forceinline__ float f2_hi(double d) { return __int_as_float(__double2hiint(d)); }

// Reciprocal and square root of the pair math: FP32 uses the hardware
// instructions (v_rcp_f32, v_sqrt_f32, within 1 ulp), FP64 the IEEE operations.
__device__ __forceinline__ float rcp_pair(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ double rcp_pair(double x) { return 1.0 / x; }
__device__ __forceinline__ float sqrt_pair(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ double sqrt_pair(double x) { return sqrt(x); }
// a / b in the pair arithmetic: the FP32 path multiplies by the 1-ulp
// hardware reciprocal, the FP64 path divides (one rounding, as the reference)
__device__ __forceinline__ float div_pair(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
__device__ __forceinline__ double div_pair(double a, double b) { return a / b; }

// 1/sqrt(x) in FP64 from the hardware estimate (v_rsq_f64, about 23 correct
// bits) and two Newton-Raphson steps y <- y (3/2 - x y^2 / 2), each doubling
// the correct bits: within a few ulp for normal x > 0.  x = 0 gives NaN.
__device__ __forceinline__ double rsq_nr(double x) {
  const double hx = 0.5 * x;
  double y = __builtin_amdgcn_rsq(x);
  y = fma(y, fma(-hx, y * y, 0.5), y);
  return fma(y, fma(-hx, y * y, 0.5), y);
}

// Periodic image coordinate exactly as the reference ghost copies compute it
// (BC.f90:339,366,396,419,445,461): +1 -> emax + (x - emin), -1 -> emin + (x - emax).
__device__ __forceinline__ double image_coord(double x, int s, double emin, double emax) {
  return s > 0 ? emax + (x - emin) : (s < 0 ? emin + (x - emax) : x);
}

template <typename T>
__device__ __forceinline__ void image_shift(T &x, T &y, T &z, uint32_t code, const Box &b) {
  int sx = (int)(code % 3) - 1, sy = (int)((code / 3) % 3) - 1, sz = (int)(code / 9) - 1;
  x = image_coord(x, sx, b.emin[0], b.emax[0]);
  y = image_coord(y, sy, b.emin[1], b.emax[1]);
  z = image_coord(z, sz, b.emin[2], b.emax[2]);
}

// Kernel value f, slope df and curvature ddf at q >= 0 (q < 0 or NaN are
// flagged by the callers, as base_m6.f90:48-53 stops on them).
template <int KT, typename T>
struct KernelFn;

template <typename T>
struct KernelFn<KT_M6, T> {
  static __device__ __forceinline__ void f_df(T q, T &f, T &df) {
    T t3 = fmax(T(3) - q, T(0)), t2 = fmax(T(2) - q, T(0)), t1 = fmax(T(1) - q, T(0));
    T a2 = t3 * t3, b2 = t2 * t2, c2 = t1 * t1;
    T a4 = a2 * a2, b4 = b2 * b2, c4 = c2 * c2;
    f = a4 * t3 - T(6) * (b4 * t2) + T(15) * (c4 * t1);
    df = T(-5) * a4 + T(30) * b4 - T(75) * c4;
  }
  static __device__ __forceinline__ T f(T q) {
    T a, b;
    f_df(q, a, b);
    return a;
  }
  static __device__ __forceinline__ T df(T q) {
    T t3 = fmax(T(3) - q, T(0)), t2 = fmax(T(2) - q, T(0)), t1 = fmax(T(1) - q, T(0));
    T a2 = t3 * t3, b2 = t2 * t2, c2 = t1 * t1;
    return T(-5) * (a2 * a2) + T(30) * (b2 * b2) - T(75) * (c2 * c2);
  }
  static __device__ __forceinline__ void df_ddf(T q, T &df, T &ddf) {
    T t3 = fmax(T(3) - q, T(0)), t2 = fmax(T(2) - q, T(0)), t1 = fmax(T(1) - q, T(0));
    T a2 = t3 * t3, b2 = t2 * t2, c2 = t1 * t1;
    df = T(-5) * (a2 * a2) + T(30) * (b2 * b2) - T(75) * (c2 * c2);
    ddf = T(20) * (a2 * t3) - T(120) * (b2 * t2) + T(300) * (c2 * t1);
  }
};

template <typename T>
struct KernelFn<KT_M4, T> {
  static __device__ __forceinline__ void f_df(T q, T &f, T &df) {
    T t2 = fmax(T(2) - q, T(0)), t1 = fmax(T(1) - q, T(0));
    T b2 = t2 * t2, c2 = t1 * t1;
    f = T(0.25) * (b2 * t2) - c2 * t1;
    df = T(-0.75) * b2 + T(3) * c2;
  }
  static __device__ __forceinline__ T f(T q) {
    T a, b;
    f_df(q, a, b);
    return a;
  }
  static __device__ __forceinline__ T df(T q) {
    T t2 = fmax(T(2) - q, T(0)), t1 = fmax(T(1) - q, T(0));
    return T(-0.75) * (t2 * t2) + T(3) * (t1 * t1);
  }
  static __device__ __forceinline__ void df_ddf(T q, T &df, T &ddf) {
    T t2 = fmax(T(2) - q, T(0)), t1 = fmax(T(1) - q, T(0));
    df = T(-0.75) * (t2 * t2) + T(3) * (t1 * t1);
    ddf = T(1.5) * t2 - T(6) * t1;
  }
};

template <typename T>
struct KernelFn<KT_GAUSS, T> {
  static __device__ __forceinline__ void f_df(T q, T &f, T &df) {
    f = exp(-q * q);
    df = T(-2) * f * q;
  }
  static __device__ __forceinline__ T f(T q) { return exp(-q * q); }
  static __device__ __forceinline__ T df(T q) { return T(-2) * exp(-q * q) * q; }
  static __device__ __forceinline__ void df_ddf(T q, T &df, T &ddf) {
    T e = exp(-q * q);
    df = T(-2) * e * q;
    ddf = T(-2) * e + T(4) * e * q * q;
  }
};

template <int N, typename T>
__host__ __device__ __forceinline__ T ipow(T x) {
  if constexpr (N == 0) return T(1);
  else if constexpr (N == 1) return x;
  else if constexpr (N % 2 == 0) { T h = ipow<N / 2>(x); return h * h; }
  else return ipow<N - 1>(x) * x;
}

// Wave-aggregated atomic increment (one atomic per wavefront).
__device__ __forceinline__ void wave_count(int *ctr, bool pred) {
  unsigned long long m = __ballot(pred);
  if (m && (threadIdx.x % warpSize) == (unsigned)(__ffsll((long long)m) - 1))
    atomicAdd(ctr, __popcll(m));
}

}  // namespace sph
