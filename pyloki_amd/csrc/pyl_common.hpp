// Shared helpers of the pyloki HIP library: error reporting, launch checks and the
// FP64 phase/index math that every kernel of the search path uses.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/pyloki_hip.h"

namespace pyl {

constexpr double kCVal = 299792458.0;  // speed of light (utils/misc.py:37)
constexpr double kFloatEps = 1e-6;      // index guard (utils/misc.py:39)

int fail(int code, const char* fmt, ...);

#define PYL_CHECK_HIP(call)                                                          \
  do {                                                                               \
    hipError_t e_ = (call);                                                          \
    if (e_ != hipSuccess)                                                            \
      return ::pyl::fail(PYL_ERR_HIP, "%s failed: %s", #call, hipGetErrorString(e_)); \
  } while (0)

#define PYL_CHECK_LAUNCH(name)                                                        \
  do {                                                                                \
    hipError_t e_ = hipGetLastError();                                                \
    if (e_ != hipSuccess)                                                             \
      return ::pyl::fail(PYL_ERR_HIP, "launch of %s failed: %s", name,                \
                         hipGetErrorString(e_));                                      \
  } while (0)

#define PYL_REQUIRE(cond, ...)                                  \
  do {                                                          \
    if (!(cond)) return ::pyl::fail(PYL_ERR_INVALID, __VA_ARGS__); \
  } while (0)

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// psr_utils.py:12-46 get_phase_idx: fractional bin of (t - delay) * f in [0, nbins).
__device__ __host__ inline double phase_idx(double t, double f, int nbins, double delay) {
  double phase = (t - delay) * f;
  phase -= floor(phase);
  double iphase = phase * (double)nbins;
  if (iphase >= (double)nbins) iphase = 0.0;
  return iphase;
}

// psr_utils.py:49-56 / common.py:313-318: int(f32(x) + 0.5f), nbins -> 0.
__device__ __host__ inline int shift_int(double iphase, int nbins) {
  float xf = (float)iphase;
  float r = xf + 0.5f;
  int ibin = (int)r;
  if (ibin == nbins) ibin = 0;
  return ibin;
}

// psr_utils.py:466-514: truncation of the grid position with a 1e-6 guard.
__device__ __host__ inline double nearest_raw(double val, int64_t n_grid, double vmin,
                                              double vmax) {
  return (double)n_grid * (val - vmin) / (vmax - vmin);
}

__device__ __host__ inline int64_t idx_of_raw(double raw, int64_t n_grid) {
  int64_t idx = (int64_t)(raw + kFloatEps);
  if (idx < 0) idx = 0;
  else if (idx >= n_grid) idx = n_grid - 1;
  return idx;
}

__device__ __host__ inline int64_t nearest_idx(double val, int64_t n_grid, double vmin,
                                               double vmax) {
  if (n_grid <= 1) return 0;
  return idx_of_raw(nearest_raw(val, n_grid, vmin, vmax), n_grid);
}

// Whether a +-1 ulp change of the FP64 value an integer decision is taken on would change
// it: the grid index (int(raw + 1e-6)) and the bin shift (int(f32(x) + 0.5f)).
__device__ inline bool idx_flips(double raw, int64_t n_grid) {
  const int64_t i = idx_of_raw(raw, n_grid);
  return idx_of_raw(nextafter(raw, -INFINITY), n_grid) != i ||
         idx_of_raw(nextafter(raw, INFINITY), n_grid) != i;
}

__device__ inline bool shift_flips(double x, int nbins) {
  const int b = shift_int(x, nbins);
  return shift_int(nextafter(x, -INFINITY), nbins) != b ||
         shift_int(nextafter(x, INFINITY), nbins) != b;
}

__device__ __host__ inline double fact_d(int n) {
  double r = 1.0;
  for (int i = 2; i <= n; ++i) r *= (double)i;
  return r;
}

// e / sqrtf(v), both operations IEEE round-to-nearest. For operands well inside the normal
// range (|e| in [2^-40, 2^40), v in [2^-50, 2^50); norm_in_range) norm_fast runs the
// arithmetic of the compiler's correctly rounded square root (v_sqrt_f32, then the +-1 ulp
// correction by FMA residuals) and divide (reciprocal, one Newton step, two residual
// corrections) without their operand scaling and special-value steps, which are identities
// there, so the result is the same to the bit; other operands take the library expansion.
__device__ __forceinline__ bool norm_in_range(float e, float v) {
  const uint32_t eb = __float_as_uint(e) & 0x7fffffffu;
  const uint32_t vb = __float_as_uint(v);
  return (eb - 0x2b800000u < 0x28000000u) & (vb - 0x26800000u < 0x32000000u);
}

__device__ __forceinline__ float norm_fast(float e, float v) {
  float s = __builtin_amdgcn_sqrtf(v);
  const float sd = __uint_as_float(__float_as_uint(s) - 1u);
  const float su = __uint_as_float(__float_as_uint(s) + 1u);
  const float rd = __builtin_fmaf(-sd, s, v);
  const float ru = __builtin_fmaf(-su, s, v);
  s = (rd <= 0.0f) ? sd : s;
  s = (ru > 0.0f) ? su : s;
  float y = __builtin_amdgcn_rcpf(s);
  y = __builtin_fmaf(__builtin_fmaf(-s, y, 1.0f), y, y);
  float q = e * y;
  q = __builtin_fmaf(__builtin_fmaf(-s, q, e), y, q);
  return __builtin_fmaf(__builtin_fmaf(-s, q, e), y, q);
}

__device__ __forceinline__ float norm_ev(float e, float v) {
  if (norm_in_range(e, v)) return norm_fast(e, v);
  return e / sqrtf(v);
}

// Four bins at once: one range test, so the four fast sequences interleave.
__device__ __forceinline__ float4 norm_ev4(float4 e, float4 v) {
  if ((int)norm_in_range(e.x, v.x) & (int)norm_in_range(e.y, v.y) &
      (int)norm_in_range(e.z, v.z) & (int)norm_in_range(e.w, v.w))
    return make_float4(norm_fast(e.x, v.x), norm_fast(e.y, v.y), norm_fast(e.z, v.z),
                       norm_fast(e.w, v.w));
  return make_float4(e.x / sqrtf(v.x), e.y / sqrtf(v.y), e.z / sqrtf(v.z), e.w / sqrtf(v.w));
}

// max(m, a, b) as the reference's sequential `if d > m: m = d` over a then b leaves it,
// for every maximum other than a zero: both skip NaN operands (arithmetic NaNs are quiet,
// and IEEE maxNum returns the other operand) and agree on every non-zero value. A zero
// maximum may differ in sign (the sequential form keeps the first zero it meets), so
// callers recompute a zero maximum sequentially. One v_max3_f32.
__device__ __forceinline__ float max3_seq(float m, float a, float b) {
  return fmaxf(m, fmaxf(a, b));
}

// x**n for an integer n by binary powering. The reference raises floats to integer
// arrays inside numba fastmath code (`delta_t**powers`, transforms.py:43;
// `(tobs - t_ref)**(k+1)`, psr_utils.py:89); LLVM's fast-math simplification turns
// pow(x, sitofp(n)) into llvm.powi, whose runtime routine is this multiply chain.
__device__ __host__ inline double powi(double a, int b) {
  const bool recip = b < 0;
  if (recip) b = -b;
  double r = 1.0;
  while (true) {
    if (b & 1) r *= a;
    b /= 2;
    if (b == 0) break;
    a *= a;
  }
  return recip ? 1.0 / r : r;
}

// transforms.py:41-43: T[i][j] = dt**(i-j) / (i-j)! for j <= i, else 0 (n x n, row-major).
void taylor_matrix(int n, double dt, double* out);

// BF16 storage (upper 16 bits of an FP32), round-to-nearest-even on the way down; the
// FFA's reduced-precision fold storage. Values are always finite here.
__device__ __host__ inline uint16_t f32_to_bf16(float f) {
  uint32_t u;
  __builtin_memcpy(&u, &f, 4);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

__device__ __host__ inline float bf16_to_f32(uint16_t h) {
  uint32_t u = (uint32_t)h << 16;
  float f;
  __builtin_memcpy(&f, &u, 4);
  return f;
}

// Storage-type traits of a fold element: FP32 or BF16 (FP32 arithmetic either way).
template <typename T>
struct Store;
template <>
struct Store<float> {
  __device__ __host__ static float load(float v) { return v; }
  __device__ __host__ static float store(float v) { return v; }
};
template <>
struct Store<uint16_t> {
  __device__ __host__ static float load(uint16_t v) { return bf16_to_f32(v); }
  __device__ __host__ static uint16_t store(float v) { return f32_to_bf16(v); }
};

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

}  // namespace pyl
