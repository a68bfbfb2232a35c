// Boxcar matched-filter S/N of stored profiles (detection/scoring.py:42-165, 249-282).
//
// One thread owns one profile; its circular prefix sum must be the reference's sequential
// FP32 scan to be bit-exact, so that thread alone computes it (pyl_score.hpp). Profiles are
// first read row by row with coalesced loads (lane = bin), normalised (e / sqrt(v) in FP32,
// ffa.py:411 / scoring.py:265-266) and transposed into LDS.
#include <algorithm>
#include <cmath>

#include "pyl_score.hpp"

namespace pyl {

int build_score_const(int nbins, const int64_t* widths, int nw, double stdnoise,
                      ScoreConst* sc) {
  PYL_REQUIRE(widths != nullptr, "null widths");
  PYL_REQUIRE(nbins > 1, "nbins must be > 1");
  PYL_REQUIRE(nw >= 1 && nw <= kScoreMaxW, "number of widths must be in [1, %d], got %d",
              kScoreMaxW, nw);
  PYL_REQUIRE(stdnoise > 0, "stdnoise must be positive");
  *sc = ScoreConst{};
  sc->nw = nw;
  sc->wmax = 0;
  for (int k = 0; k < nw; ++k) {
    const int64_t w = widths[k];
    PYL_REQUIRE(w >= 1 && w < nbins, "width %lld outside [1, nbins-1]", (long long)w);
    sc->widths[k] = (int32_t)w;
    sc->wmax = std::max<int32_t>(sc->wmax, (int32_t)w);
    // scoring.py:135-137 (size_w is an exact small integer in FP32 there).
    const int64_t size_w = nbins - w;
    const double h = std::sqrt((double)size_w / (double)(nbins * w));
    const double b = (double)w * h / (double)size_w;
    sc->hpb[k] = h + b;
    sc->bb[k] = b;
  }
  sc->inv_std = 1.0 / stdnoise;
  const size_t lds = sizeof(float) * (size_t)(nbins + sc->wmax) * kScoreStride;
  PYL_REQUIRE(lds <= 64 * 1024, "nbins too large for the LDS scorer (%d)", nbins);
  return PYL_OK;
}

}  // namespace pyl

namespace {

__global__ void __launch_bounds__(pyl::kScoreBlock)
score_boxcar_kernel(const float* __restrict__ folds, int64_t n, int nbins, int layout,
                    pyl::ScoreConst k, int mode, float* __restrict__ out) {
  extern __shared__ float S[];  // [(nbins + wmax)][kScoreStride]
  const int tid = threadIdx.x;
  const int64_t p0 = (int64_t)blockIdx.x * pyl::kScoreBlock;
  const int nprof = (int)std::min<int64_t>(pyl::kScoreBlock, n - p0);
  const int64_t rowlen = layout == PYL_LAYOUT_NORM ? (int64_t)nbins : 2 * (int64_t)nbins;
  for (int r = 0; r < nprof; ++r) {
    for (int b = tid; b < nbins; b += pyl::kScoreBlock) {
      float x;
      if (layout == PYL_LAYOUT_EV) {
        const float* row = folds + (p0 + r) * rowlen;
        x = pyl::norm_ev(row[b], row[nbins + b]);
      } else if (layout == PYL_LAYOUT_EV_BF16) {
        const uint16_t* row = reinterpret_cast<const uint16_t*>(folds) + (p0 + r) * rowlen;
        x = pyl::norm_ev(pyl::bf16_to_f32(row[b]), pyl::bf16_to_f32(row[nbins + b]));
      } else {
        x = folds[(p0 + r) * rowlen + b];
      }
      S[b * pyl::kScoreStride + r] = x;
    }
  }
  __syncthreads();
  if (tid >= nprof) return;
  if (mode == PYL_SCORE_PER_WIDTH) {
    pyl::score_column<true>(k, nbins, S, tid, out + (p0 + tid) * k.nw);
  } else {
    out[p0 + tid] = (float)pyl::score_column<false>(k, nbins, S, tid, nullptr);
  }
}

}  // namespace

extern "C" int pyl_score_boxcar(const float* folds, int64_t n, int32_t nbins, int32_t layout,
                                const int64_t* widths, int32_t nw, double stdnoise, int32_t mode,
                                float* out, void* stream) {
  PYL_REQUIRE(widths, "null widths");
  PYL_REQUIRE(n >= 0, "negative profile count");
  PYL_REQUIRE(layout == PYL_LAYOUT_EV || layout == PYL_LAYOUT_NORM || layout == PYL_LAYOUT_EV_BF16,
              "invalid layout %d", layout);
  PYL_REQUIRE(mode == PYL_SCORE_PER_WIDTH || mode == PYL_SCORE_MAX, "invalid mode %d", mode);
  pyl::ScoreConst k;
  int rc = pyl::build_score_const(nbins, widths, nw, stdnoise, &k);
  if (rc != PYL_OK) return rc;
  if (n == 0) return PYL_OK;
  PYL_REQUIRE(folds && out, "null device pointer");
  const size_t lds = sizeof(float) * (size_t)(nbins + k.wmax) * pyl::kScoreStride;
  score_boxcar_kernel<<<(unsigned)pyl::ceil_div(n, pyl::kScoreBlock), pyl::kScoreBlock, lds,
                        pyl::as_stream(stream)>>>(folds, n, nbins, layout, k, mode, out);
  PYL_CHECK_LAUNCH("score_boxcar_kernel");
  return PYL_OK;
}
