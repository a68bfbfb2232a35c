// Boxcar S/N threshold screen on the matrix cores: the boxcar scores as a product of
// profiles and templates (detection/scoring.py:383-471 builds the same templates for its
// matched filter), with exact rescoring of every profile the screen cannot reject.
//
// A box sum of profile y over the w bins starting at phase i is y . M_w[i], with the 0/1
// template M_w[i][j] = ((j - i) mod B < w): for one width, the B phases of 32 profiles are
// a (B x B) x (B x 32) product. The templates are exact in BF16; the normalised profile
// is rounded to BF16 (unit roundoff 2^-8), and v_mfma_f32_32x32x8_bf16 accumulates in
// FP32. Against the exact scorer's FP32 prefix differences the screen's box sum is off by
// at most (2^-8 + 3 B 2^-24) sum|y| <= 2^-7 sum|y|, so with
//   ub = max_w [ (h_w + b_w) (max_i box_screen + 2^-7 sum|y|) - b_w total ]
// (total = the exact scorer's FP32 sum, h and b as in the exact scorer) the exact score is
// at most ub. A profile whose ub is below the threshold (with a 2^-20 relative margin for
// the final FP32 rounding) is rejected and gets ub; every other profile is scored by the
// exact path (pyl_score.hpp score_column: the reference's sequential scan), so the
// profiles at or above the threshold and their scores are those of pyl_score_boxcar
// (mode PYL_SCORE_MAX) bit for bit.
//
// Block: 256 threads, 32 profiles. The profiles are normalised into LDS twice: FP32
// columns for the exact path and BF16 rows for the products; wavefront v takes the widths
// v, v + 4, ... and, per width, the B/32 phase tiles (A = template rows made in registers,
// B = the BF16 profile rows), keeping each profile's maximum box sum.
#include <algorithm>
#include <cfloat>
#include <cmath>

#include "pyl_score.hpp"

namespace {

typedef short sc_bf16x4 __attribute__((ext_vector_type(4)));
typedef float sc_f32x16 __attribute__((ext_vector_type(16)));

constexpr int kP = 32;  // profiles per block
constexpr int kThreads = 256;
constexpr int kScreenMaxW = 16;

__device__ __forceinline__ uint16_t bf16_rne(float x) {
  uint32_t u = __float_as_uint(x);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

template <int B>
__global__ void __launch_bounds__(kThreads)
screen_kernel(const float* __restrict__ folds, int64_t n, pyl::ScoreConst k, double thr,
              float* __restrict__ out, unsigned long long* __restrict__ nres) {
  extern __shared__ float S[];  // [(B + wmax)][kScoreStride] FP32 columns (exact path)
  __shared__ uint16_t yb[kP * B];  // BF16 rows
  __shared__ float smax[kScreenMaxW * kP];
  __shared__ float s_total[kP], s_abs[kP];
  const int tid = threadIdx.x;
  const int64_t p0 = (int64_t)blockIdx.x * kP;
  const int np = (int)std::min<int64_t>(kP, n - p0);
  for (int idx = tid; idx < kP * B; idx += kThreads) {
    const int r = idx / B, b = idx - r * B;
    float y = 0.0f;
    if (r < np) {
      const float* row = folds + (p0 + r) * 2 * B;
      y = pyl::norm_ev(row[b], row[B + b]);
    }
    S[b * pyl::kScoreStride + r] = y;
    yb[r * B + b] = bf16_rne(y);
  }
  __syncthreads();
  if (tid < kP) {
    float run = 0.0f, ab = 0.0f;
    for (int b = 0; b < B; ++b) {
      const float y = S[b * pyl::kScoreStride + tid];
      run = b == 0 ? y : run + y;  // the exact scan's total
      ab += fabsf(y);
    }
    s_total[tid] = run;
    s_abs[tid] = ab;
  }
  // Box sums of every width and phase on the matrix cores.
  const int lane = tid & 63, wave = tid >> 6;
  const int row = lane & 31, half = lane >> 5;
  for (int m = wave; m < k.nw; m += kThreads / 64) {
    const int w = k.widths[m];
    float best = -FLT_MAX;
#pragma unroll
    for (int it = 0; it < B / 32; ++it) {
      const int i = it * 32 + row;  // this lane's template row (phase)
      sc_f32x16 acc;
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[r] = 0.0f;
#pragma unroll
      for (int kb = 0; kb < B / 8; ++kb) {
        const int j0 = 8 * kb + 4 * half;
        sc_bf16x4 a;
#pragma unroll
        for (int c = 0; c < 4; ++c) a[c] = (short)((((j0 + c - i) & (B - 1)) < w) ? 0x3F80 : 0);
        const sc_bf16x4 bv = *reinterpret_cast<const sc_bf16x4*>(yb + row * B + j0);
        acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, bv, acc, 0, 0, 0);
      }
      // acc[r]: phase it*32 + 8 (r / 4) + 4 half + r % 4 of profile `row`
#pragma unroll
      for (int r = 0; r < 16; ++r) best = fmaxf(best, acc[r]);
    }
    best = fmaxf(best, __shfl_xor(best, 32));
    if (half == 0) smax[m * kP + row] = best;
  }
  __syncthreads();
  if (tid >= np) return;
  const double total = (double)s_total[tid];
  const double slack = 0x1p-7 * (double)s_abs[tid];
  double ub = -DBL_MAX;
  for (int m = 0; m < k.nw; ++m) {
    const double v = k.hpb[m] * ((double)smax[m * kP + tid] + slack) - k.bb[m] * total;
    ub = v > ub ? v : ub;
  }
  if (ub >= thr - 0x1p-20 * fabs(thr)) {
    out[p0 + tid] = (float)pyl::score_column<false>(k, B, S, tid, nullptr);
    if (nres != nullptr) atomicAdd(nres, 1ull);
  } else {
    out[p0 + tid] = (float)ub;
  }
}

}  // namespace

extern "C" int pyl_score_boxcar_screen(const float* folds, int64_t n, int32_t nbins,
                                       const int64_t* widths, int32_t nw, double threshold,
                                       float* out, unsigned long long* nrescored_dev,
                                       void* stream) {
  PYL_REQUIRE(nbins == 32 || nbins == 64, "the screen takes nbins 32 or 64, got %d", nbins);
  PYL_REQUIRE(nw >= 1 && nw <= kScreenMaxW, "the screen takes 1..%d widths, got %d",
              kScreenMaxW, nw);
  PYL_REQUIRE(n >= 0, "negative profile count");
  PYL_REQUIRE(std::isfinite(threshold), "the threshold must be finite");
  pyl::ScoreConst k;
  int rc = pyl::build_score_const(nbins, widths, nw, 1.0, &k);
  if (rc != PYL_OK) return rc;
  if (n == 0) return PYL_OK;
  PYL_REQUIRE(folds && out, "null device pointer");
  const size_t lds = sizeof(float) * (size_t)(nbins + k.wmax) * pyl::kScoreStride;
  const unsigned grid = (unsigned)pyl::ceil_div(n, kP);
  hipStream_t s = pyl::as_stream(stream);
  if (nbins == 32)
    screen_kernel<32><<<grid, kThreads, lds, s>>>(folds, n, k, threshold, out, nrescored_dev);
  else
    screen_kernel<64><<<grid, kThreads, lds, s>>>(folds, n, k, threshold, out, nrescored_dev);
  PYL_CHECK_LAUNCH("screen_kernel");
  return PYL_OK;
}
