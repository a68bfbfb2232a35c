// Boxcar S/N pieces shared by the scorer, the fused FFA level + score kernel and the EP
// kernels (detection/scoring.py:42-165, 249-282).
//
// A block owns up to 64 profiles; profile r's normalised values sit in column r of an LDS
// array S[(nbins + wmax)][kScoreStride] (stride padded to 65 words so that both the
// row-wise fills, lane = bin, and the column-wise scans, lane = profile, are free of bank
// conflicts). The scan is the reference's sequential FP32 prefix sum; per width
//   snr = ((h + b) * max_i(S[i+w] - S[i]) - b * S[B-1]) * inv_stdnoise   (FP64, no FMA),
// with h = sqrt((B-w)/(B*w)) and b = w*h/(B-w) computed on the host in FP64.
#pragma once

#include <cfloat>

#include "pyl_common.hpp"

namespace pyl {

constexpr int kScoreBlock = 64;
constexpr int kScoreStride = kScoreBlock + 1;
constexpr int kScoreMaxW = 64;

struct ScoreConst {
  int32_t nw;
  int32_t wmax;
  int32_t widths[kScoreMaxW];
  double hpb[kScoreMaxW];  // h + b
  double bb[kScoreMaxW];   // b
  double inv_std;
};

// Host: validate widths and fill the per-width constants (defined in pyl_score.hip).
int build_score_const(int nbins, const int64_t* widths, int nw, double stdnoise,
                      ScoreConst* sc);

// In-place sequential scan of column `tid` and its one-wrap extension, then the S/N of
// every width: written to out[w] when kPerWidth, and the FP64 maximum returned
// (snr_score_batch_func's max, before the final FP32 rounding).
template <bool kPerWidth>
__device__ inline double score_column(const ScoreConst& k, int nbins, float* S, int tid,
                                      float* out) {
  float run = S[tid];
  for (int i = 1; i < nbins; ++i) {
    run = run + S[i * kScoreStride + tid];
    S[i * kScoreStride + tid] = run;
  }
  const float total = run;
  for (int j = 0; j < k.wmax; ++j)
    S[(nbins + j) * kScoreStride + tid] = S[j * kScoreStride + tid] + total;
  double best = -(double)FLT_MAX;
  for (int w = 0; w < k.nw; ++w) {
    const int wd = k.widths[w];
    float dmax = -FLT_MAX;
    for (int i = 0; i < nbins; ++i) {
      const float d = S[(i + wd) * kScoreStride + tid] - S[i * kScoreStride + tid];
      dmax = d > dmax ? d : dmax;
    }
    const double t1 = k.hpb[w] * (double)dmax;
    const double t2 = k.bb[w] * (double)total;
    const double snr = t1 - t2;
    if (kPerWidth) out[w] = (float)(snr * k.inv_std);
    best = snr > best ? snr : best;
  }
  return best;
}

// ---- Register-resident scoring for small compile-time nbins (32, 64) ----------------
// The profile's prefix sums live in B registers of the owning thread; the one-wrap
// extension S[B + j] = S[j] + S[B-1] is formed on the fly (same FP32 addition as the
// stored extension), and the max-difference loop is fully unrolled for each width, so
// no LDS traffic is needed after the transpose.

// In-place sequential inclusive scan (scoring.py:53-55).
template <int B>
__device__ __forceinline__ void scan_regs(float (&S)[B]) {
#pragma unroll
  for (int i = 1; i < B; ++i) S[i] = S[i - 1] + S[i];
}

template <int B, int W>
__device__ __forceinline__ float maxdiff_regs(const float (&S)[B]) {
  float dmax = -FLT_MAX;
#pragma unroll
  for (int i = 0; i < B; ++i) {
    const float hi = (i + W < B) ? S[i + W] : (S[i + W - B] + S[B - 1]);
    const float d = hi - S[i];
    dmax = d > dmax ? d : dmax;
  }
  return dmax;
}

template <int B, int W = 1>
__device__ __forceinline__ float maxdiff_dispatch(const float (&S)[B], int w) {
  if constexpr (W >= B) {
    return -FLT_MAX;  // unreachable: widths are validated to lie in [1, B-1]
  } else {
    if (w == W) return maxdiff_regs<B, W>(S);
    return maxdiff_dispatch<B, W + 1>(S, w);
  }
}

// S/N of every width from register prefix sums; per width to out[w] (kPerWidth) and the
// FP64 maximum returned.
template <int B, bool kPerWidth>
__device__ __forceinline__ double score_regs(const ScoreConst& k, const float (&S)[B],
                                             float* out) {
  const float total = S[B - 1];
  double best = -(double)FLT_MAX;
  for (int w = 0; w < k.nw; ++w) {
    const float dmax = maxdiff_dispatch<B>(S, k.widths[w]);
    const double t1 = k.hpb[w] * (double)dmax;
    const double t2 = k.bb[w] * (double)total;
    const double snr = t1 - t2;
    if (kPerWidth) out[w] = (float)(snr * k.inv_std);
    best = snr > best ? snr : best;
  }
  return best;
}

// ---- Complex (Fourier-domain) rows --------------------------------------------------
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// Twiddles for irfft_bin: tw[j] = cos(2 pi j / B), tw[B + j] = sin(2 pi j / B), FP32 of FP64.
__device__ inline void fill_twiddles(float* tw, int nbins, int tid, int nthreads) {
  for (int j = tid; j < nbins; j += nthreads) {
    const double a = 2.0 * M_PI * (double)j / (double)nbins;
    tw[j] = (float)cos(a);
    tw[nbins + j] = (float)sin(a);
  }
}

// Real bin b of the inverse real FFT of one row of B/2 + 1 harmonics (power-of-two B):
// x[b] = (X_0 + (-1)^b X_{B/2} + 2 sum_{k=1}^{B/2-1} Re(X_k e^{2 pi i k b / B})) / B, FP32,
// the imaginary parts of X_0 and X_{B/2} ignored (as a real inverse transform does).
__device__ inline float irfft_bin(const float2* X, int nbins, int b, const float* tw) {
  const int nb_f = nbins / 2 + 1;
  float acc = X[0].x;
  const float ny = X[nb_f - 1].x;
  acc += (b & 1) ? -ny : ny;
  for (int kk = 1; kk < nb_f - 1; ++kk) {
    const int idx = (kk * b) & (nbins - 1);
    acc += 2.0f * (X[kk].x * tw[idx] - X[kk].y * tw[nbins + idx]);
  }
  return acc * (1.0f / (float)nbins);
}

}  // namespace pyl
