// Monte-Carlo threshold design (detection/thresholding.py:45-157, 483-700, 1010-1110).
//
// One wavefront per task (one simulated fold set of simulate_folds): each lane owns one
// trial (row) at a time, draws its nbins normal deviates from the counter-based generator
// (pyl_rng.hpp), forms the fold in registers, normalises it and scores it with the
// register-resident boxcar scorer of the FFA path (pyl_score.hpp). Survivor counts for a
// list of thresholds come from wave ballots; survivors are written in trial order with
// a ballot prefix (the stable order of np.nonzero in prune_folds).
#include <algorithm>
#include <cfloat>
#include <cmath>

#include "pyl_rng.hpp"
#include "pyl_score.hpp"

namespace {

constexpr int kSimWave = 64;
constexpr int kSimMaxThr = 256;

template <int B, int mode>
__global__ void __launch_bounds__(kSimWave)
thresh_sim_kernel(const float* __restrict__ src, const PylSimTask* __restrict__ tasks,
                  const double* __restrict__ prof_scaled, double std_add, uint64_t seed,
                  pyl::ScoreConst k, const double* __restrict__ thr, float* __restrict__ dst,
                  float* __restrict__ scores, int32_t* __restrict__ counts) {
  __shared__ double prof[B];
  __shared__ double thr_s[kSimMaxThr];
  __shared__ int cnt[kSimMaxThr];
  const PylSimTask t = tasks[blockIdx.x];
  const int lane = threadIdx.x;
  for (int j = lane; j < B; j += kSimWave) prof[j] = prof_scaled[t.hyp * B + j];
  const int nthr = mode == PYL_SIM_COUNT ? min(max(t.thr_count, 0), kSimMaxThr) : 0;
  for (int j = lane; j < nthr; j += kSimWave) {
    thr_s[j] = thr[t.thr_first + j];
    cnt[j] = 0;
  }
  __syncthreads();
  const double sd = sqrt(t.var_norm);
  int64_t kept = 0;
  for (int base = 0; base < t.n_sim; base += kSimWave) {
    const int i = base + lane;
    const bool live = i < t.n_sim;
    float f[B];
    double score = -(double)FLT_MAX;
    if (live) {
      const float* srow = t.src_row >= 0 ? src + (t.src_row + i % t.n_prev) * (int64_t)B : nullptr;
      if (t.flags & PYL_SIM_NO_NOISE) {
#pragma unroll
        for (int q = 0; q < B / 4; ++q) {
          const float4 s4 = srow ? reinterpret_cast<const float4*>(srow)[q]
                                 : make_float4(0.f, 0.f, 0.f, 0.f);
          f[4 * q] = s4.x;
          f[4 * q + 1] = s4.y;
          f[4 * q + 2] = s4.z;
          f[4 * q + 3] = s4.w;
        }
      } else {
#pragma unroll
        for (int q = 0; q < B / 4; ++q) {
          double z[4];
          pyl::normal4(seed, t.stream, (uint32_t)i, (uint32_t)q, z);
          const float4 s4 = srow ? reinterpret_cast<const float4*>(srow)[q]
                                 : make_float4(0.f, 0.f, 0.f, 0.f);
          const float sv[4] = {s4.x, s4.y, s4.z, s4.w};
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const float noise = (float)(std_add * z[r]);
            const float a = sv[r] + noise;
            f[4 * q + r] = (float)((double)a + prof[4 * q + r]);
          }
        }
      }
      float S[B];
#pragma unroll
      for (int j = 0; j < B; ++j) S[j] = (float)((double)f[j] / sd);
      pyl::scan_regs<B>(S);
      score = pyl::score_regs<B, false>(k, S, nullptr);
    }
    const float sc = (float)score;
    if constexpr (mode == PYL_SIM_COUNT) {
      if (live && scores) scores[t.score_off + i] = sc;
      for (int j = 0; j < nthr; ++j) {
        const uint64_t m = __ballot(live && (double)sc > thr_s[j]);
        if (lane == 0) cnt[j] += __popcll(m);
      }
    } else {
      const bool keep = live && (double)sc > t.thr;
      const uint64_t m = __ballot(keep);
      const uint64_t below = lane == 0 ? 0ull : (m & (~0ull >> (64 - lane)));
      if (keep) {
        const int64_t row = kept + __popcll(below);
        float4* drow = reinterpret_cast<float4*>(dst + (t.dst_row + row) * (int64_t)B);
#pragma unroll
        for (int q = 0; q < B / 4; ++q)
          drow[q] = make_float4(f[4 * q], f[4 * q + 1], f[4 * q + 2], f[4 * q + 3]);
        if (scores) scores[t.score_off + row] = sc;
      }
      kept += __popcll(m);
    }
  }
  __syncthreads();
  if constexpr (mode == PYL_SIM_COUNT) {
    for (int j = lane; j < nthr; j += kSimWave) counts[t.count_off + j] = cnt[j];
  } else if (lane == 0) {
    counts[t.count_off] = (int32_t)kept;
  }
}

__global__ void rng_normal_kernel(uint64_t seed, uint64_t stream, int64_t n_rows, int row_len,
                                  double sd, float* __restrict__ out) {
  const int64_t nq = (int64_t)(row_len / 4);
  const int64_t total = n_rows * nq;
  for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < total;
       g += (int64_t)gridDim.x * blockDim.x) {
    const int64_t row = g / nq;
    const int64_t q = g - row * nq;
    double z[4];
    pyl::normal4(seed, stream, (uint32_t)row, (uint32_t)q, z);
    reinterpret_cast<float4*>(out + row * (int64_t)row_len)[q] =
        make_float4((float)(sd * z[0]), (float)(sd * z[1]), (float)(sd * z[2]),
                    (float)(sd * z[3]));
  }
}

__global__ void rows_gather_kernel(const float* __restrict__ src, const int64_t* __restrict__ idx,
                                   int64_t n, int row_len, float* __restrict__ dst) {
  const int64_t total = n * row_len;
  for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < total;
       g += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = g / row_len;
    const int64_t j = g - r * row_len;
    dst[g] = src[idx[r] * row_len + j];
  }
}

template <int B>
int launch_sim(const float* src, const PylSimTask* tasks, int ntasks, const double* prof,
               double std_add, uint64_t seed, const pyl::ScoreConst& k, const double* thr,
               int mode, float* dst, float* scores, int32_t* counts, hipStream_t s) {
  if (mode == PYL_SIM_COUNT)
    thresh_sim_kernel<B, PYL_SIM_COUNT><<<ntasks, kSimWave, 0, s>>>(src, tasks, prof, std_add,
                                                                    seed, k, thr, dst, scores,
                                                                    counts);
  else
    thresh_sim_kernel<B, PYL_SIM_COMPACT><<<ntasks, kSimWave, 0, s>>>(src, tasks, prof, std_add,
                                                                      seed, k, thr, dst, scores,
                                                                      counts);
  PYL_CHECK_LAUNCH("thresh_sim_kernel");
  return PYL_OK;
}

}  // namespace

extern "C" int pyl_thresh_simulate(const float* src_dev, const PylSimTask* tasks_dev,
                                   int32_t ntasks, int32_t nbins, const double* prof_scaled_dev,
                                   double var_add, uint64_t seed, const int64_t* widths,
                                   int32_t nw, const double* thr_dev, int32_t mode, float* dst_dev,
                                   float* scores_dev, int32_t* counts_dev, void* stream) {
  PYL_REQUIRE(ntasks >= 0, "negative task count");
  PYL_REQUIRE(nbins == 16 || nbins == 32 || nbins == 64,
              "the simulation kernel supports nbins 16, 32 and 64, got %d", nbins);
  PYL_REQUIRE(mode == PYL_SIM_COUNT || mode == PYL_SIM_COMPACT, "invalid mode %d", mode);
  PYL_REQUIRE(var_add >= 0, "var_add must be non-negative");
  pyl::ScoreConst k;
  int rc = pyl::build_score_const(nbins, widths, nw, 1.0, &k);
  if (rc != PYL_OK) return rc;
  if (ntasks == 0) return PYL_OK;
  PYL_REQUIRE(tasks_dev && prof_scaled_dev && counts_dev, "null device pointer");
  PYL_REQUIRE(mode != PYL_SIM_COMPACT || dst_dev, "PYL_SIM_COMPACT needs dst_dev");
  PYL_REQUIRE(mode != PYL_SIM_COUNT || thr_dev, "PYL_SIM_COUNT needs thr_dev");
  const double std_add = sqrt(var_add);
  const hipStream_t s = pyl::as_stream(stream);
  switch (nbins) {
    case 16:
      return launch_sim<16>(src_dev, tasks_dev, ntasks, prof_scaled_dev, std_add, seed, k,
                            thr_dev, mode, dst_dev, scores_dev, counts_dev, s);
    case 32:
      return launch_sim<32>(src_dev, tasks_dev, ntasks, prof_scaled_dev, std_add, seed, k,
                            thr_dev, mode, dst_dev, scores_dev, counts_dev, s);
    default:
      return launch_sim<64>(src_dev, tasks_dev, ntasks, prof_scaled_dev, std_add, seed, k,
                            thr_dev, mode, dst_dev, scores_dev, counts_dev, s);
  }
}

extern "C" int pyl_rng_normal(uint64_t seed, uint64_t stream, int64_t n_rows, int32_t row_len,
                              double std, float* out_dev, void* stream_) {
  PYL_REQUIRE(n_rows >= 0 && n_rows <= (int64_t)UINT32_MAX + 1, "n_rows outside [0, 2^32]");
  PYL_REQUIRE(row_len > 0 && row_len % 4 == 0, "row_len must be a positive multiple of 4");
  PYL_REQUIRE(std >= 0, "std must be non-negative");
  if (n_rows == 0) return PYL_OK;
  PYL_REQUIRE(out_dev, "null device pointer");
  const int64_t total = n_rows * (row_len / 4);
  const int64_t blocks = std::min<int64_t>(pyl::ceil_div(total, 256), 1 << 16);
  rng_normal_kernel<<<(unsigned)blocks, 256, 0, pyl::as_stream(stream_)>>>(seed, stream, n_rows,
                                                                         row_len, std, out_dev);
  PYL_CHECK_LAUNCH("rng_normal_kernel");
  return PYL_OK;
}

extern "C" int pyl_rows_gather(const float* src_dev, const int64_t* idx_dev, int64_t n,
                               int32_t row_len, float* dst_dev, void* stream) {
  PYL_REQUIRE(n >= 0 && row_len > 0, "invalid gather shape");
  if (n == 0) return PYL_OK;
  PYL_REQUIRE(src_dev && idx_dev && dst_dev, "null device pointer");
  const int64_t blocks = std::min<int64_t>(pyl::ceil_div(n * row_len, 256), 1 << 16);
  rows_gather_kernel<<<(unsigned)blocks, 256, 0, pyl::as_stream(stream)>>>(src_dev, idx_dev, n,
                                                                         row_len, dst_dev);
  PYL_CHECK_LAUNCH("rows_gather_kernel");
  return PYL_OK;
}
