// Counter-based normal deviates for the Monte-Carlo kernels (threshold design, signal
// simulation).
//
// Uniform bits: Philox4x32-10 (Salmon, Moraes, Dror & Shaw, "Parallel random numbers: as
// easy as 1, 2, 3", SC'11): a 128-bit counter and a 64-bit key give four 32-bit words, so
// any (stream, row, quad) can be drawn independently and drawn again later with the same
// result. Counter layout: {quad, row, stream_lo, stream_hi}; key: the 64-bit seed.
//
// Normals: Box-Muller on the pairs (u0, u1) and (u2, u3), u = (x + 0.5) 2^-32 in (0, 1):
//   z0 = r cos(2 pi u1), z1 = r sin(2 pi u1), r = sqrt(-2 ln u0); likewise z2, z3.
// ln and sincos(2 pi u) are evaluated here with IEEE basic operations only (frexp by bit
// manipulation, a fixed polynomial, correctly rounded +, -, *, /, sqrt, no contraction),
// so a CPU build of the same sequence reproduces every deviate bit for bit. Accuracy:
// ln within 2 ulp on (0, 1); sin and cos within 1.2e-16 absolute (Taylor series to the
// 17th / 18th power on [0, pi/4] after an exact quadrant reduction of 4u).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace pyl {

struct U4 {
  uint32_t x, y, z, w;
};

__device__ __host__ inline uint32_t mulhi32(uint32_t a, uint32_t b) {
  return (uint32_t)(((uint64_t)a * (uint64_t)b) >> 32);
}

__device__ __host__ inline U4 philox4x32_10(U4 c, uint32_t k0, uint32_t k1) {
  constexpr uint32_t kM0 = 0xD2511F53u, kM1 = 0xCD9E8D57u;
  constexpr uint32_t kW0 = 0x9E3779B9u, kW1 = 0xBB67AE85u;
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint32_t hi0 = mulhi32(kM0, c.x), lo0 = kM0 * c.x;
    const uint32_t hi1 = mulhi32(kM1, c.z), lo1 = kM1 * c.z;
    c = U4{hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
    k0 += kW0;
    k1 += kW1;
  }
  return c;
}

__device__ __host__ inline uint64_t bits_of(double x) { return __builtin_bit_cast(uint64_t, x); }
__device__ __host__ inline double double_of(uint64_t b) { return __builtin_bit_cast(double, b); }

// ln(x) for a normal x in (0, 1): x = m 2^e with m in [sqrt(1/2), sqrt(2)),
// ln m = 2 s (1 + s^2/3 + s^4/5 + ... + s^20/21), s = (m - 1) / (m + 1).
__device__ __host__ inline double ln_det(double x) {
  const uint64_t b = bits_of(x);
  int e = (int)((b >> 52) & 0x7ff) - 1022;
  double m = double_of((b & 0x000fffffffffffffull) | 0x3fe0000000000000ull);  // [0.5, 1)
  if (m < 0x1.6a09e667f3bcdp-1) {
    m = m * 2.0;
    e -= 1;
  }
  const double s = (m - 1.0) / (m + 1.0);
  const double s2 = s * s;
  double p = 0x1.8618618618618p-5;  // 1/21
  p = p * s2 + 0x1.af286bca1af28p-5;  // 1/19
  p = p * s2 + 0x1.e1e1e1e1e1e1ep-5;  // 1/17
  p = p * s2 + 0x1.1111111111111p-4;  // 1/15
  p = p * s2 + 0x1.3b13b13b13b14p-4;  // 1/13
  p = p * s2 + 0x1.745d1745d1746p-4;  // 1/11
  p = p * s2 + 0x1.c71c71c71c71cp-4;  // 1/9
  p = p * s2 + 0x1.2492492492492p-3;  // 1/7
  p = p * s2 + 0x1.999999999999ap-3;  // 1/5
  p = p * s2 + 0x1.5555555555555p-2;  // 1/3
  p = p * s2 + 1.0;
  return (double)e * 0x1.62e42fefa39efp-1 + 2.0 * s * p;
}

// sin and cos of 2 pi u for u in [0, 1): t = 4u = q + f exactly (q the quadrant), the
// first-quadrant angle a = g pi/2 with g = f or 1 - f (so a <= pi/4), then the quadrant.
__device__ __host__ inline void sincos_2pi(double u, double* sn, double* cs) {
  const double t = 4.0 * u;
  const int q = (int)t;
  const double f = t - (double)q;
  const bool swap = f > 0.5;
  const double g = swap ? 1.0 - f : f;
  const double a = g * 0x1.921fb54442d18p+0;
  const double a2 = a * a;
  double ps = 0x1.952c77030ad4ap-49;   //  1/17!
  ps = ps * a2 + -0x1.ae7f3e733b81fp-41;  // -1/15!
  ps = ps * a2 + 0x1.6124613a86d09p-33;   //  1/13!
  ps = ps * a2 + -0x1.ae64567f544e4p-26;  // -1/11!
  ps = ps * a2 + 0x1.71de3a556c734p-19;   //  1/9!
  ps = ps * a2 + -0x1.a01a01a01a01ap-13;  // -1/7!
  ps = ps * a2 + 0x1.1111111111111p-7;    //  1/5!
  ps = ps * a2 + -0x1.5555555555555p-3;   // -1/3!
  ps = ps * a2 + 1.0;
  double pc = -0x1.6827863b97d97p-53;  // -1/18!
  pc = pc * a2 + 0x1.ae7f3e733b81fp-45;   //  1/16!
  pc = pc * a2 + -0x1.93974a8c07c9dp-37;  // -1/14!
  pc = pc * a2 + 0x1.1eed8eff8d898p-29;   //  1/12!
  pc = pc * a2 + -0x1.27e4fb7789f5cp-22;  // -1/10!
  pc = pc * a2 + 0x1.a01a01a01a01ap-16;   //  1/8!
  pc = pc * a2 + -0x1.6c16c16c16c17p-10;  // -1/6!
  pc = pc * a2 + 0x1.5555555555555p-5;    //  1/4!
  pc = pc * a2 + -0.5;
  pc = pc * a2 + 1.0;
  double s0 = a * ps, c0 = pc;
  if (swap) {
    const double tmp = s0;
    s0 = c0;
    c0 = tmp;
  }
  switch (q & 3) {
    case 0: *sn = s0; *cs = c0; break;
    case 1: *sn = c0; *cs = -s0; break;
    case 2: *sn = -s0; *cs = -c0; break;
    default: *sn = -c0; *cs = s0; break;
  }
}

__device__ __host__ inline double unit_open(uint32_t x) {
  return ((double)x + 0.5) * 0x1p-32;
}

// Four N(0, 1) deviates (FP64) of counter {quad, row, stream}.
__device__ __host__ inline void normal4(uint64_t seed, uint64_t stream, uint32_t row,
                                        uint32_t quad, double z[4]) {
  const U4 r = philox4x32_10(U4{quad, row, (uint32_t)stream, (uint32_t)(stream >> 32)},
                             (uint32_t)seed, (uint32_t)(seed >> 32));
  double sn, cs;
  const double r0 = sqrt(-2.0 * ln_det(unit_open(r.x)));
  sincos_2pi(unit_open(r.y), &sn, &cs);
  z[0] = r0 * cs;
  z[1] = r0 * sn;
  const double r1 = sqrt(-2.0 * ln_det(unit_open(r.z)));
  sincos_2pi(unit_open(r.w), &sn, &cs);
  z[2] = r1 * cs;
  z[3] = r1 * sn;
}

}  // namespace pyl
