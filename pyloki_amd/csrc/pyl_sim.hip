// Seeded pulsar signal simulation on the GPU (simulation/pulse.py:310-406 generate,
// simulation/modulate.py:23-259 proper time).
//
// One pass over the samples (sim_series_kernel), one thread per quad of samples:
//  - proper time t' = t - delay(t - t_ref) of the modulation (FP64, the reference's
//    operation order; t = i dt as numpy.arange forms it);
//  - the pulse template: the shape's CDF between the phases of t' and t' + dt, linearly
//    interpolated in the FP32 CDF table (pulse.py:39-66), in FP64, stored as FP32;
//  - the noise: N(0, 1) deviates of the counter-based generator (pyl_rng.hpp, counter
//    {quad, 0, stream}), the same values pyl_rng_normal writes for one row;
//  - the calibration folds of the noise, the template and the unit variance at
//    f = 1/period with int(f32(phase_bin) + 0.5f) bins (fold.py:57-105 brutefold, one
//    segment). The sums are exact fixed-point integers (value * 2^32, rounded to
//    nearest-even, summed in int64), so they do not depend on the order of the atomics:
//    a launch is deterministic, and the host reads each sum as an FP32 value.
// The host then forms the S/N calibration on the folded profiles (a few thousand values),
// and sim_combine_kernel writes ts_e = f32(ns * noise) + f32(f32(cs) * template) and
// ts_v = f32(ns * ns) (the reference's FP32 arithmetic).
#include <algorithm>
#include <cmath>

#include "pyl_common.hpp"
#include "pyl_rng.hpp"

namespace {

constexpr int kSimBlock = 256;
constexpr int kSimLdsBins = 2048;  // the three fold arrays in LDS up to this many bins

// Python's float modulo (result carries the divisor's sign), as numba compiles `%`.
__device__ inline double py_mod(double a, double b) {
  double r = fmod(a, b);
  if (r != 0.0) {
    if ((b < 0.0) != (r < 0.0)) r += b;
  } else {
    r = copysign(0.0, b);
  }
  return r;
}

// modulate.py: t_proper for barycentric time t.
__device__ inline double proper_time(const pyl_modulation& m, double t) {
  constexpr double kTwoPi = 2.0 * M_PI;
  switch (m.kind) {
    case PYL_MOD_DERIVATIVE: {  // modulate.py:34-54
      const double x = t - m.t_ref;
      const double delay =
          m.p[0] + x * (m.p[1] + x * (m.p[2] * 0.5 +
                                      x * (m.p[3] * (1.0 / 6.0) + x * m.p[4] * (1.0 / 24.0))));
      return t - delay / pyl::kCVal;
    }
    case PYL_MOD_DERIVATIVE_SERIES: {  // modulate.py:92-97: sum_k (d_k / k!) x^k, k order
      const double x = t - m.t_ref;
      double delay = m.coeffs[0] * pow(x, 0.0);
      for (int k = 1; k < m.n_coeffs; ++k) delay = delay + m.coeffs[k] * pow(x, (double)k);
      return t - delay / pyl::kCVal;
    }
    case PYL_MOD_CIRCULAR: {  // modulate.py:147-150 (delay in light-seconds)
      const double omega = kTwoPi / m.p[0];
      const double delay = m.p[2] * sin(omega * (t - m.t_ref) + m.p[1]);
      return t - delay;
    }
    case PYL_MOD_CIRCULAR_T0: {  // modulate.py:222-226
      const double omega = kTwoPi / m.p[1];
      const double phi = kTwoPi * py_mod(m.t_ref - m.p[2], m.p[1]) / m.p[1];
      const double delay = m.p[0] * sin(omega * (t - m.t_ref) + phi);
      return t - delay / pyl::kCVal;
    }
    default: {  // PYL_MOD_KEPLERIAN, modulate.py:240-250 via kepler.py:64-140
      const double p_orb = m.p[0], ecc = fabs(m.p[1]), phi = m.p[2], a = m.p[3];
      const double aop = m.p[4], inc = m.p[5];
      const double tn = py_mod(t - m.t_ref, p_orb) / p_orb;
      const double ma = kTwoPi * tn + phi;
      double ea = ma;
      for (int it = 0; it < 8; ++it) ea += (ma + ecc * sin(ea) - ea) / (1.0 - ecc * cos(ea));
      const double nu = 2.0 * atan(sqrt((1.0 + ecc) / (1.0 - ecc)) * tan(ea / 2.0));
      const double r = a * (1.0 - ecc * ecc) / (1.0 + ecc * cos(nu));
      const double z = r * sin(nu + aop) * sin(inc);
      return t - z / pyl::kCVal;
    }
  }
}

// pulse.py:57-64: one end of the CDF difference. p lies in [0, 1] after the modulo; p
// rounded up to 1 reads the wrap element with weight 1. A non-finite p (a proper time
// that is not a number) reads cell 0, so no index leaves the table.
__device__ inline double cdf_at(const float* __restrict__ lut, int ngrid, double p) {
  double x = p * (double)ngrid;
  if (!(x >= 0.0)) x = 0.0;
  if (x > (double)ngrid) x = (double)ngrid;
  const int64_t i = (int64_t)x;
  if (i >= ngrid) return (double)lut[ngrid];
  const double f = x - (double)i;
  return (1.0 - f) * (double)lut[i] + f * (double)lut[i + 1];
}

__device__ inline long long to_fixed(float x) {
  return __double2ll_rn((double)x * 4294967296.0);
}

template <bool kLds>
__global__ void __launch_bounds__(kSimBlock)
sim_series_kernel(pyl_modulation mod, int64_t nsamps, double dt, double period,
                  const float* __restrict__ lut, int ngrid, uint64_t seed, uint64_t stream,
                  int nbins, float* __restrict__ noise, float* __restrict__ tmpl,
                  unsigned long long* __restrict__ folds) {
  // folds: [3][nbins] = noise sums, template sums, sample counts.
  __shared__ unsigned long long s_f[kLds ? 3 * kSimLdsBins : 1];
  if (kLds) {
    for (int b = threadIdx.x; b < 3 * nbins; b += kSimBlock) s_f[b] = 0ull;
    __syncthreads();
  }
  unsigned long long* acc = kLds ? s_f : folds;
  const double inv_period = 1.0 / period;
  const double freq = 1.0 / period;
  const int64_t nquad = (nsamps + 3) / 4;
  for (int64_t q = (int64_t)blockIdx.x * kSimBlock + threadIdx.x; q < nquad;
       q += (int64_t)gridDim.x * kSimBlock) {
    double z[4];
    pyl::normal4(seed, stream, 0u, (uint32_t)q, z);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int64_t i = 4 * q + j;
      if (i >= nsamps) break;
      const double t0 = proper_time(mod, (double)i * dt);
      const double t1 = t0 + dt;
      const double p0 = py_mod(t0, period) * inv_period;
      const double p1 = py_mod(t1, period) * inv_period;
      const float tv = p1 > p0 ? (float)(cdf_at(lut, ngrid, p1) - cdf_at(lut, ngrid, p0)) : 0.0f;
      const float nv = (float)z[j];
      tmpl[i] = tv;
      noise[i] = nv;
      int b = pyl::shift_int(pyl::phase_idx(t0, freq, nbins, 0.0), nbins);
      if (b < 0 || b >= nbins) b = 0;  // only for a non-finite proper time
      atomicAdd(&acc[b], (unsigned long long)to_fixed(nv));
      atomicAdd(&acc[nbins + b], (unsigned long long)to_fixed(tv));
      atomicAdd(&acc[2 * nbins + b], 1ull);
    }
  }
  if (kLds) {
    __syncthreads();
    for (int b = threadIdx.x; b < 3 * nbins; b += kSimBlock)
      if (s_f[b] != 0ull) atomicAdd(&folds[b], s_f[b]);
  }
}

__global__ void sim_combine_kernel(const float* __restrict__ noise, const float* __restrict__ tmpl,
                                   int64_t n, float ns, float cs, float var,
                                   float* __restrict__ ts_e, float* __restrict__ ts_v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const float a = ns * noise[i];
    const float b = cs * tmpl[i];
    ts_e[i] = a + b;
    ts_v[i] = var;
  }
}

__global__ void sim_zero_kernel(unsigned long long* __restrict__ p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0ull;
}

}  // namespace

extern "C" int pyl_sim_series(const pyl_modulation* mod, int64_t nsamps, double dt,
                              double period, const float* cdf_lut_dev, int32_t ngrid,
                              uint64_t seed, uint64_t stream_id, int32_t nbins,
                              float* noise_dev, float* tmpl_dev, int64_t* folds_dev,
                              void* stream) {
  PYL_REQUIRE(mod != nullptr, "null modulation");
  PYL_REQUIRE(mod->kind >= PYL_MOD_DERIVATIVE && mod->kind <= PYL_MOD_KEPLERIAN,
              "unknown modulation kind %d", mod->kind);
  PYL_REQUIRE(mod->kind != PYL_MOD_DERIVATIVE_SERIES ||
                  (mod->n_coeffs >= 1 && mod->n_coeffs <= PYL_MOD_MAX_COEFFS),
              "derivative series needs 1..%d coefficients, got %d", PYL_MOD_MAX_COEFFS,
              mod->n_coeffs);
  PYL_REQUIRE(nsamps > 0 && nsamps <= PYL_SIM_MAX_SAMPS,
              "nsamps must be in [1, %lld] (fixed-point fold range), got %lld",
              (long long)PYL_SIM_MAX_SAMPS, (long long)nsamps);
  PYL_REQUIRE(dt > 0 && period > 0, "dt and period must be positive");
  PYL_REQUIRE(ngrid >= 1, "the CDF table needs at least one cell");
  PYL_REQUIRE(nbins >= 1, "the calibration fold needs at least one bin");
  PYL_REQUIRE(cdf_lut_dev && noise_dev && tmpl_dev && folds_dev, "null device pointer");
  hipStream_t s = pyl::as_stream(stream);
  auto* acc = reinterpret_cast<unsigned long long*>(folds_dev);
  const int64_t nf = 3 * (int64_t)nbins;
  sim_zero_kernel<<<(unsigned)std::min<int64_t>(pyl::ceil_div(nf, 256), 1024), 256, 0, s>>>(acc,
                                                                                           nf);
  PYL_CHECK_LAUNCH("sim_zero_kernel");
  const int64_t nquad = (nsamps + 3) / 4;
  // A few quads per thread, so that the LDS partial folds are flushed by few blocks.
  const int64_t blocks = std::max<int64_t>(1, std::min<int64_t>(pyl::ceil_div(nquad, kSimBlock * 8),
                                                                2048));
  if (nbins <= kSimLdsBins) {
    sim_series_kernel<true><<<(unsigned)blocks, kSimBlock, 0, s>>>(
        *mod, nsamps, dt, period, cdf_lut_dev, ngrid, seed, stream_id, nbins, noise_dev,
        tmpl_dev, acc);
  } else {
    sim_series_kernel<false><<<(unsigned)blocks, kSimBlock, 0, s>>>(
        *mod, nsamps, dt, period, cdf_lut_dev, ngrid, seed, stream_id, nbins, noise_dev,
        tmpl_dev, acc);
  }
  PYL_CHECK_LAUNCH("sim_series_kernel");
  return PYL_OK;
}

extern "C" int pyl_sim_combine(const float* noise_dev, const float* tmpl_dev, int64_t nsamps,
                               float noise_scale, float signal_scale, float variance,
                               float* ts_e_dev, float* ts_v_dev, void* stream) {
  PYL_REQUIRE(nsamps >= 0, "negative sample count");
  if (nsamps == 0) return PYL_OK;
  PYL_REQUIRE(noise_dev && tmpl_dev && ts_e_dev && ts_v_dev, "null device pointer");
  const int64_t blocks = std::min<int64_t>(pyl::ceil_div(nsamps, 256), 8192);
  sim_combine_kernel<<<(unsigned)blocks, 256, 0, pyl::as_stream(stream)>>>(
      noise_dev, tmpl_dev, nsamps, noise_scale, signal_scale, variance, ts_e_dev, ts_v_dev);
  PYL_CHECK_LAUNCH("sim_combine_kernel");
  return PYL_OK;
}
