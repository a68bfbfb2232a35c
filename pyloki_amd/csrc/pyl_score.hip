// Boxcar matched-filter S/N (detection/scoring.py:42-165, 249-282).
//
// One thread owns one profile. Its circular prefix sum must be the reference's
// sequential FP32 scan (out[i] = out[i-1] + x[i]) to be bit-exact, so it is computed by
// that thread alone; the scan lives in LDS in a [i][thread] layout, so that the 64 lanes
// of a wavefront touch 64 consecutive words per step (no bank conflicts). Profiles are
// first read row-by-row with coalesced loads (lane = bin) and transposed through LDS
// (row stride padded to 65 words).
//
// Per width w: dmax = max_i (S[i+w] - S[i]) in FP32, then
//   snr = ((h + b) * dmax - b * S[B-1]) * inv_stdnoise     in FP64 (no contraction),
// with h = sqrt((B-w)/(B*w)), b = w*h/(B-w) computed on the host in FP64.
#include <algorithm>
#include <cfloat>
#include <cmath>

#include "pyl_common.hpp"

namespace {

constexpr int kMaxW = 64;
constexpr int kProfilesPerBlock = 64;
constexpr int kStride = kProfilesPerBlock + 1;  // padded LDS row: conflict-free both ways

struct ScoreArgs {
  int32_t nw;
  int32_t wmax;
  int32_t widths[kMaxW];
  double hpb[kMaxW];  // h + b
  double bb[kMaxW];   // b
  double inv_std;
};

__global__ void __launch_bounds__(kProfilesPerBlock)
score_boxcar_kernel(const float* __restrict__ folds, int64_t n, int nbins, int layout,
                    ScoreArgs a, int mode, float* __restrict__ out) {
  extern __shared__ float S[];  // [(nbins + wmax)][kStride]
  const int tid = threadIdx.x;
  const int64_t p0 = (int64_t)blockIdx.x * kProfilesPerBlock;
  const int nprof = (int)std::min<int64_t>(kProfilesPerBlock, n - p0);
  const int64_t rowlen = layout == PYL_LAYOUT_EV ? 2 * (int64_t)nbins : (int64_t)nbins;
  // Coalesced load + normalisation + transpose: lane = bin, loop over the block's rows.
  for (int r = 0; r < nprof; ++r) {
    const float* row = folds + (p0 + r) * rowlen;
    for (int b = tid; b < nbins; b += kProfilesPerBlock) {
      float x;
      if (layout == PYL_LAYOUT_EV) x = row[b] / sqrtf(row[nbins + b]);
      else x = row[b];
      S[b * kStride + r] = x;
    }
  }
  __syncthreads();
  if (tid >= nprof) return;
  // Sequential inclusive scan, in place (scoring.py:53-55).
  float run = S[tid];
  for (int i = 1; i < nbins; ++i) {
    run = run + S[i * kStride + tid];
    S[i * kStride + tid] = run;
  }
  const float total = run;
  // One wrap (scoring.py:68-71): out[B + j] = out[j] + last_sum.
  for (int j = 0; j < a.wmax; ++j)
    S[(nbins + j) * kStride + tid] = S[j * kStride + tid] + total;
  double best = -(double)FLT_MAX;
  for (int k = 0; k < a.nw; ++k) {
    const int w = a.widths[k];
    float dmax = -FLT_MAX;
    for (int i = 0; i < nbins; ++i) {
      const float d = S[(i + w) * kStride + tid] - S[i * kStride + tid];
      dmax = d > dmax ? d : dmax;
    }
    const double t1 = a.hpb[k] * (double)dmax;
    const double t2 = a.bb[k] * (double)total;
    const double snr = t1 - t2;
    if (mode == PYL_SCORE_PER_WIDTH) {
      out[(p0 + tid) * a.nw + k] = (float)(snr * a.inv_std);
    } else {
      best = snr > best ? snr : best;
    }
  }
  if (mode == PYL_SCORE_MAX) out[p0 + tid] = (float)best;
}

}  // namespace

extern "C" int pyl_score_boxcar(const float* folds, int64_t n, int32_t nbins, int32_t layout,
                                const int64_t* widths, int32_t nw, double stdnoise, int32_t mode,
                                float* out, void* stream) {
  PYL_REQUIRE(folds && out && widths, "null pointer argument");
  PYL_REQUIRE(n >= 0, "negative profile count");
  PYL_REQUIRE(nbins > 1, "nbins must be > 1");
  PYL_REQUIRE(nw >= 1 && nw <= kMaxW, "number of widths must be in [1, %d], got %d", kMaxW, nw);
  PYL_REQUIRE(layout == PYL_LAYOUT_EV || layout == PYL_LAYOUT_NORM, "invalid layout %d", layout);
  PYL_REQUIRE(mode == PYL_SCORE_PER_WIDTH || mode == PYL_SCORE_MAX, "invalid mode %d", mode);
  PYL_REQUIRE(stdnoise > 0, "stdnoise must be positive");
  if (n == 0) return PYL_OK;
  ScoreArgs a{};
  a.nw = nw;
  a.wmax = 0;
  for (int k = 0; k < nw; ++k) {
    const int64_t w = widths[k];
    PYL_REQUIRE(w >= 1 && w < nbins, "width %lld outside [1, nbins-1]", (long long)w);
    a.widths[k] = (int32_t)w;
    a.wmax = std::max<int32_t>(a.wmax, (int32_t)w);
    // scoring.py:135-137 (size_w is an exact small integer in FP32 there).
    const int64_t size_w = nbins - w;
    const double h = std::sqrt((double)size_w / (double)(nbins * w));
    const double b = (double)w * h / (double)size_w;
    a.hpb[k] = h + b;
    a.bb[k] = b;
  }
  a.inv_std = 1.0 / stdnoise;
  const size_t lds = sizeof(float) * (size_t)(nbins + a.wmax) * kStride;
  PYL_REQUIRE(lds <= 64 * 1024, "nbins too large for the LDS scorer (%d)", nbins);
  hipStream_t s = pyl::as_stream(stream);
  score_boxcar_kernel<<<(unsigned)pyl::ceil_div(n, kProfilesPerBlock), kProfilesPerBlock, lds,
                        s>>>(folds, n, nbins, layout, a, mode, out);
  PYL_CHECK_LAUNCH("score_boxcar_kernel");
  return PYL_OK;
}
