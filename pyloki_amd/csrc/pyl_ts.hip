// Time-series conditioning for io/timeseries.py:197-235 TimeSeries.from_tim: the running-
// median baseline removal ("deredden", median method, fast form) and the normalisation to
// zero location and unit scale, on the GPU.
//
// Order statistics by radix select. Every float maps to a 32-bit key in the same order
// (sign bit flipped for x >= 0, all bits flipped for x < 0); the k-th smallest key of a
// segment is found 8 bits at a time: each pass counts the segment's keys whose higher bits
// equal the prefix found so far, by their next 8 bits (per-block histograms in LDS, summed
// with integer atomics, so the counts do not depend on the order of the atomics), and one
// thread per selection walks its histogram to the bin holding the k-th key. Four passes
// give the key exactly. A quantile q of a segment of m values is numpy's "linear" one:
// position p = q (m - 1), the order statistics at floor(p) and ceil(p), interpolated in
// FP64. The key may be |x - c_s| (c_s per segment) for median absolute deviations.
//
// Dereddening: the medians of consecutive blocks of w samples (the last block may be
// shorter), placed at the block centres, are interpolated linearly (held constant before
// the first and after the last centre) and subtracted from the series.
//
// Normalisation: location = mean (FP64 partial sums per block, then one block sums the
// partials in index order: deterministic), scale = the standard deviation (FP64, the same
// two-level sum), the interquartile range / 1.3489795003921634 or the median absolute
// deviation x 1.482602218505602 (the normal distribution's scale factors); the series
// becomes (x - location) / scale in FP64, stored as FP32.
#include <algorithm>
#include <cmath>

#include "pyl_common.hpp"

namespace {

constexpr int kT = 256;
constexpr int kSelMax = 4;       // selections per segment in one call
constexpr int kRedBlocks = 1024;  // blocks of the two-level FP64 sums

__device__ __forceinline__ uint32_t key_of(float x) {
  const uint32_t u = __float_as_uint(x);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

__device__ __forceinline__ float value_of(uint32_t k) {
  const uint32_t u = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
  return __uint_as_float(u);
}

struct SelState {
  uint32_t prefix;  // key bits found so far (high bits)
  uint32_t k;       // rank still to find among the keys with that prefix
};

// Segment s: samples [s * seg_len, min(n, (s + 1) * seg_len)).
__device__ __forceinline__ int64_t seg_size(int64_t s, int64_t seg_len, int64_t n) {
  return std::min<int64_t>(seg_len, n - s * seg_len);
}

// Selection j of segment s: order statistic floor or ceil (j & 1) of quantile q[j / 2].
__global__ void sel_init_kernel(SelState* __restrict__ st, const double* __restrict__ q,
                                int nsel, int64_t nseg, int64_t seg_len, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (i >= nseg * nsel) return;
  const int64_t s = i / nsel;
  const int j = (int)(i - s * nsel);
  const int64_t m = seg_size(s, seg_len, n);
  const double p = q[j / 2] * (double)(m - 1);
  const int64_t lo = (int64_t)floor(p);
  const int64_t r = (j & 1) ? std::min<int64_t>(lo + 1, m - 1) : lo;
  st[i] = SelState{0u, (uint32_t)r};
}

// One pass: histogram of the next 8 key bits (shift = 24, 16, 8, 0) of the keys whose
// higher bits equal each selection's prefix. hist: u32 [nseg][nsel][256], zero on entry.
__global__ void __launch_bounds__(kT)
sel_hist_kernel(const float* __restrict__ x, int64_t n, int64_t seg_len,
                const float* __restrict__ center, const SelState* __restrict__ st, int nsel,
                int shift, uint32_t* __restrict__ hist) {
  __shared__ uint32_t h[kSelMax * 256];
  const int64_t s = blockIdx.y;
  for (int i = threadIdx.x; i < nsel * 256; i += kT) h[i] = 0u;
  __syncthreads();
  uint32_t pre[kSelMax];
#pragma unroll
  for (int j = 0; j < kSelMax; ++j) pre[j] = j < nsel ? st[s * nsel + j].prefix : 0u;
  const int64_t m = seg_size(s, seg_len, n);
  const float* xs = x + s * seg_len;
  const float c = center != nullptr ? center[s] : 0.0f;
  const bool use_c = center != nullptr;
  for (int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x; i < m; i += (int64_t)gridDim.x * kT) {
    const float v = use_c ? fabsf(xs[i] - c) : xs[i];
    const uint32_t key = key_of(v);
    const uint32_t hi = shift == 24 ? 0u : key >> (shift + 8);
    const uint32_t d = (key >> shift) & 0xffu;
#pragma unroll
    for (int j = 0; j < kSelMax; ++j)
      if (j < nsel && hi == pre[j]) atomicAdd(&h[j * 256 + d], 1u);
  }
  __syncthreads();
  uint32_t* hs = hist + s * nsel * 256;
  for (int i = threadIdx.x; i < nsel * 256; i += kT)
    if (h[i] != 0u) atomicAdd(&hs[i], h[i]);
}

// One thread per selection: the digit whose bin holds rank k, then the histogram cleared
// for the next pass.
__global__ void sel_pick_kernel(SelState* __restrict__ st, uint32_t* __restrict__ hist,
                                int64_t nselt) {
  const int64_t i = (int64_t)blockIdx.x * 64 + threadIdx.x;
  if (i >= nselt) return;
  uint32_t* hb = hist + i * 256;
  SelState s = st[i];
  uint32_t acc = 0;
  int d = 255;
  for (int b = 0; b < 256; ++b) {
    const uint32_t c = hb[b];
    if (s.k < acc + c) {
      d = b;
      break;
    }
    acc += c;
  }
  s.k -= acc;
  s.prefix = (s.prefix << 8) | (uint32_t)d;
  st[i] = s;
  for (int b = 0; b < 256; ++b) hb[b] = 0u;
}

__global__ void zero_u32_kernel(uint32_t* __restrict__ p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kT)
    p[i] = 0u;
}

// out[s][t] = quantile q[t] of segment s (FP64 interpolation of its two order statistics).
__global__ void sel_finish_kernel(const SelState* __restrict__ st, const double* __restrict__ q,
                                  int nq, int64_t nseg, int64_t seg_len, int64_t n,
                                  float* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (i >= nseg * nq) return;
  const int64_t s = i / nq;
  const int t = (int)(i - s * nq);
  const int64_t m = seg_size(s, seg_len, n);
  const double p = q[t] * (double)(m - 1);
  const double f = p - floor(p);
  const double a = (double)value_of(st[s * 2 * nq + 2 * t].prefix);
  const double b = (double)value_of(st[s * 2 * nq + 2 * t + 1].prefix);
  out[i] = (float)(a + (b - a) * f);
}

// x - the baseline interpolated between the block medians at the block centres.
__global__ void deredden_kernel(const float* __restrict__ x, int64_t n, int64_t w,
                                const float* __restrict__ med, int64_t nblk,
                                float* __restrict__ out) {
  for (int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kT) {
    // centre of block j: j w + (m_j - 1) / 2
    const int64_t j = i / w;
    const double cj = (double)(j * w) + 0.5 * (double)(seg_size(j, w, n) - 1);
    int64_t j0, j1;
    if ((double)i >= cj) {
      j0 = j;
      j1 = std::min<int64_t>(j + 1, nblk - 1);
    } else {
      j0 = std::max<int64_t>(j - 1, 0);
      j1 = j;
    }
    double base;
    if (j0 == j1) {
      base = (double)med[j0];
    } else {
      const double c0 = (double)(j0 * w) + 0.5 * (double)(seg_size(j0, w, n) - 1);
      const double c1 = (double)(j1 * w) + 0.5 * (double)(seg_size(j1, w, n) - 1);
      const double t = ((double)i - c0) / (c1 - c0);
      base = (double)med[j0] + ((double)med[j1] - (double)med[j0]) * t;
    }
    out[i] = (float)((double)x[i] - base);
  }
}

// Two-level FP64 sums: block b sums its grid-stride elements of (x - c)^p (p = 1 or 2).
__global__ void __launch_bounds__(kT)
sum_part_kernel(const float* __restrict__ x, int64_t n, double c, int pw,
                double* __restrict__ part) {
  __shared__ double sh[kT];
  double acc = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kT) {
    const double d = (double)x[i] - c;
    acc += pw == 2 ? d * d : d;
  }
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int off = kT / 2; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

__global__ void sum_final_kernel(const double* __restrict__ part, int nb,
                                 double* __restrict__ out) {
  __shared__ double sh[kT];
  double acc = 0.0;
  for (int i = threadIdx.x; i < nb; i += kT) acc += part[i];
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int off = kT / 2; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = sh[0];
}

__global__ void scale_kernel(const float* __restrict__ x, int64_t n, double loc, double scale,
                             float* __restrict__ out) {
  for (int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kT)
    out[i] = (float)(((double)x[i] - loc) / scale);
}

inline unsigned grid_for(int64_t n, int64_t cap) {
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>(pyl::ceil_div(n, kT), cap));
}

// Quantiles of up to 65535 segments (the launch grid's y extent) into out (nseg x nq).
int quantiles_chunk(const float* x, int64_t n, int64_t seg_len, const double* q_host, int nq,
                    const float* center, float* out, hipStream_t s) {
  const int64_t nseg = pyl::ceil_div(n, seg_len);
  const int nsel = 2 * nq;
  const int64_t nselt = nseg * nsel;
  char* work = nullptr;
  const size_t bytes = sizeof(SelState) * nselt + sizeof(uint32_t) * nselt * 256 +
                       sizeof(double) * 4;
  PYL_CHECK_HIP(hipMallocAsync(reinterpret_cast<void**>(&work), bytes, s));
  auto* st = reinterpret_cast<SelState*>(work);
  auto* hist = reinterpret_cast<uint32_t*>(work + sizeof(SelState) * nselt);
  auto* qd = reinterpret_cast<double*>(work + sizeof(SelState) * nselt +
                                       sizeof(uint32_t) * nselt * 256);
  int rc = PYL_OK;
  hipError_t e = hipMemcpyAsync(qd, q_host, sizeof(double) * nq, hipMemcpyHostToDevice, s);
  if (e != hipSuccess) rc = pyl::fail(PYL_ERR_HIP, "hipMemcpyAsync: %s", hipGetErrorString(e));
  if (rc == PYL_OK) {
    // q per selection pair: sel_init reads q[j / 2], so lay out q as given
    zero_u32_kernel<<<grid_for(nselt * 256, 4096), kT, 0, s>>>(hist, nselt * 256);
    sel_init_kernel<<<grid_for(nselt, 65535), kT, 0, s>>>(st, qd, nsel, nseg, seg_len, n);
    const int64_t bps = std::max<int64_t>(1, std::min<int64_t>(pyl::ceil_div(seg_len, kT * 16),
                                                               std::max<int64_t>(1, 4096 / nseg)));
    for (int shift = 24; shift >= 0; shift -= 8) {
      sel_hist_kernel<<<dim3((unsigned)bps, (unsigned)nseg), kT, 0, s>>>(x, n, seg_len, center,
                                                                         st, nsel, shift, hist);
      sel_pick_kernel<<<(unsigned)pyl::ceil_div(nselt, 64), 64, 0, s>>>(st, hist, nselt);
    }
    sel_finish_kernel<<<grid_for(nseg * nq, 65535), kT, 0, s>>>(st, qd, nq, nseg, seg_len, n,
                                                                  out);
    e = hipGetLastError();
    if (e != hipSuccess) rc = pyl::fail(PYL_ERR_HIP, "select launch: %s", hipGetErrorString(e));
  }
  e = hipFreeAsync(work, s);
  if (rc == PYL_OK && e != hipSuccess)
    rc = pyl::fail(PYL_ERR_HIP, "hipFreeAsync: %s", hipGetErrorString(e));
  return rc;
}

// Quantiles of every segment into out_dev (nseg x nq), on stream s.
int quantiles(const float* x, int64_t n, int64_t seg_len, const double* q_host, int nq,
              const float* center, float* out, hipStream_t s) {
  PYL_REQUIRE(x && out, "null device pointer");
  PYL_REQUIRE(n >= 1 && seg_len >= 1, "empty series or segment");
  PYL_REQUIRE(nq >= 1 && 2 * nq <= kSelMax, "1 or 2 quantiles per call, got %d", nq);
  for (int t = 0; t < nq; ++t)
    PYL_REQUIRE(q_host[t] >= 0.0 && q_host[t] <= 1.0, "quantile %g outside [0, 1]", q_host[t]);
  PYL_REQUIRE(seg_len <= (int64_t)UINT32_MAX, "segment too long");
  const int64_t nseg = pyl::ceil_div(n, seg_len);
  for (int64_t s0 = 0; s0 < nseg; s0 += 65535) {
    const int64_t cnt = std::min<int64_t>(65535, nseg - s0);
    const int64_t nc = std::min<int64_t>(n - s0 * seg_len, cnt * seg_len);
    const int rc = quantiles_chunk(x + s0 * seg_len, nc, seg_len, q_host, nq,
                                   center != nullptr ? center + s0 : nullptr, out + s0 * nq, s);
    if (rc != PYL_OK) return rc;
  }
  return PYL_OK;
}

int fp64_sum(const float* x, int64_t n, double c, int pw, double* out_host, hipStream_t s) {
  double* buf = nullptr;
  PYL_CHECK_HIP(hipMallocAsync(reinterpret_cast<void**>(&buf), sizeof(double) * (kRedBlocks + 1),
                               s));
  sum_part_kernel<<<kRedBlocks, kT, 0, s>>>(x, n, c, pw, buf);
  sum_final_kernel<<<1, kT, 0, s>>>(buf, kRedBlocks, buf + kRedBlocks);
  int rc = PYL_OK;
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) rc = pyl::fail(PYL_ERR_HIP, "sum launch: %s", hipGetErrorString(e));
  if (rc == PYL_OK) {
    e = hipMemcpyAsync(out_host, buf + kRedBlocks, sizeof(double), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) rc = pyl::fail(PYL_ERR_HIP, "sum read-back: %s", hipGetErrorString(e));
  }
  (void)hipFreeAsync(buf, s);
  return rc;
}

}  // namespace

extern "C" int pyl_ts_quantiles(const float* x_dev, int64_t n, int64_t seg_len,
                                const double* q, int32_t nq, const float* center_dev,
                                float* out_dev, void* stream) {
  PYL_REQUIRE(q != nullptr, "null quantiles");
  return quantiles(x_dev, n, seg_len, q, nq, center_dev, out_dev, pyl::as_stream(stream));
}

extern "C" int pyl_ts_deredden_median(const float* x_dev, int64_t n, int64_t window,
                                      float* medians_dev, float* out_dev, void* stream) {
  PYL_REQUIRE(window >= 1, "the window must hold at least one sample");
  PYL_REQUIRE(x_dev && medians_dev && out_dev, "null device pointer");
  PYL_REQUIRE(n >= 1, "empty series");
  hipStream_t s = pyl::as_stream(stream);
  const double half = 0.5;
  const int64_t nblk = pyl::ceil_div(n, window);
  int rc = quantiles(x_dev, n, window, &half, 1, nullptr, medians_dev, s);
  if (rc != PYL_OK) return rc;
  deredden_kernel<<<grid_for(n, 8192), kT, 0, s>>>(x_dev, n, window, medians_dev, nblk, out_dev);
  PYL_CHECK_LAUNCH("deredden_kernel");
  return PYL_OK;
}

extern "C" int pyl_ts_normalise(const float* x_dev, int64_t n, int32_t scale_method,
                                float* out_dev, double* loc_scale, void* stream) {
  PYL_REQUIRE(x_dev && out_dev && loc_scale, "null pointer argument");
  PYL_REQUIRE(n >= 2, "normalisation needs at least two samples");
  PYL_REQUIRE(scale_method >= PYL_TS_SCALE_STD && scale_method <= PYL_TS_SCALE_MAD,
              "unknown scale method %d", scale_method);
  hipStream_t s = pyl::as_stream(stream);
  double sum = 0.0;
  int rc = fp64_sum(x_dev, n, 0.0, 1, &sum, s);
  if (rc != PYL_OK) return rc;
  const double loc = sum / (double)n;
  double scale = 0.0;
  if (scale_method == PYL_TS_SCALE_STD) {
    double ss = 0.0;
    rc = fp64_sum(x_dev, n, loc, 2, &ss, s);
    if (rc != PYL_OK) return rc;
    scale = sqrt(ss / (double)n);
  } else {
    float* qv = nullptr;
    PYL_CHECK_HIP(hipMallocAsync(reinterpret_cast<void**>(&qv), sizeof(float) * 2, s));
    float hq[2] = {0.0f, 0.0f};
    if (scale_method == PYL_TS_SCALE_IQR) {
      const double q[2] = {0.25, 0.75};
      rc = quantiles(x_dev, n, n, q, 2, nullptr, qv, s);
      if (rc == PYL_OK) {
        hipError_t e = hipMemcpyAsync(hq, qv, sizeof(float) * 2, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) rc = pyl::fail(PYL_ERR_HIP, "read-back: %s", hipGetErrorString(e));
      }
      scale = ((double)hq[1] - (double)hq[0]) / 1.3489795003921634;
    } else {  // MAD: the median of |x - median(x)|
      const double half = 0.5;
      rc = quantiles(x_dev, n, n, &half, 1, nullptr, qv, s);
      if (rc == PYL_OK) rc = quantiles(x_dev, n, n, &half, 1, qv, qv + 1, s);
      if (rc == PYL_OK) {
        hipError_t e = hipMemcpyAsync(hq, qv, sizeof(float) * 2, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) rc = pyl::fail(PYL_ERR_HIP, "read-back: %s", hipGetErrorString(e));
      }
      scale = (double)hq[1] * 1.482602218505602;
    }
    (void)hipFreeAsync(qv, s);
    if (rc != PYL_OK) return rc;
  }
  PYL_REQUIRE(scale > 0.0 && std::isfinite(scale), "the series has zero scale (%g)", scale);
  loc_scale[0] = loc;
  loc_scale[1] = scale;
  scale_kernel<<<grid_for(n, 8192), kT, 0, s>>>(x_dev, n, loc, scale, out_dev);
  PYL_CHECK_LAUNCH("scale_kernel");
  return PYL_OK;
}
