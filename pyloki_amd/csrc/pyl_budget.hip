// Survivor budget of an EP level (no counterpart in the reference): keep the `budget`
// survivors of highest key (S/N, or the ranking network's logit), in their order.
//
// The cut is an exact radix select on the keys' order-preserving 32-bit images: four
// passes of an 8-bit digit histogram over the candidates that share the prefix found so
// far, each followed by a one-block pick of the digit where the running count from the
// top reaches the budget. The kept rows are then the keys above the cut value plus the
// first rows equal to it (in row order) that fill the budget; one packed chunked scan
// (count above in the low half, count equal in the high half) gives every kept row its
// position. Counts come from device memory, so the level loop enqueues the cut without
// the host.
//
// The network's features need the level's S/N mean, standard deviation and rank: the
// statistics are reduced in a fixed order (per-chunk partials, then a fixed tree over the
// chunks), the rank comes from a 1024-bin histogram between the level's smallest and
// largest S/N (integer counts), so the features do not depend on scheduling.
//
// A level's ranking is a chain of small dependent launches, so each launch does what the
// next step needs from the previous one's complete output: the histogram blocks reduce the
// statistics' partials themselves, the feature (key) blocks build the histogram's prefix
// in LDS, and the last block of each digit histogram (found with a device-scope counter)
// picks the digit.
#include <cfloat>

#include "pyl_budget.hpp"
#include "pyl_loop.hpp"

namespace {

using namespace pyl_ep;

constexpr int kT = 256;
constexpr int kBins = PYL_BUDGET_HIST_BINS;
constexpr int kSelPrefix = 0, kSelMask = 1, kSelRemain = 2, kSelAll = 3, kSelDone = 4,
              kSelHist = 16;
constexpr int kFeat = 8;
static_assert(kSelHist + 256 == PYL_BUDGET_SEL_WORDS, "select state layout");
static_assert(4 * kScanBlocks + 8 == PYL_BUDGET_STATS, "statistics layout");

__device__ __forceinline__ int64_t count_of(int64_t n, const int64_t* n_dev) {
  return n_dev ? *n_dev : n;
}

// Rows of the block's tree and the first of them. One tree: [0, count_of(n, n_dev)). A
// forest (off not null): tree blockIdx.y's rows [off[y], off[y + 1]), or none when n_dev
// reads 0 (no tree needs ranking). Per-tree state is offset by blockIdx.y (0 for one tree).
__device__ __forceinline__ int64_t tree_rows(int64_t n, const int64_t* n_dev,
                                             const int64_t* off, int64_t* base) {
  if (!off) {
    *base = 0;
    return count_of(n, n_dev);
  }
  const int y = blockIdx.y;
  *base = off[y];
  if (n_dev && *n_dev == 0) return 0;
  return off[y + 1] - off[y];
}

// Order-preserving image of a float: larger key, larger unsigned value.
__device__ __forceinline__ uint32_t key_bits(float x) {
  const uint32_t b = __float_as_uint(x);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

// ---- Features --------------------------------------------------------------------------

// Per chunk of the rows (kScanBlocks chunks): sum, sum of squares (FP64), min, max.
__global__ void __launch_bounds__(kT)
feat_stats_part_kernel(const float* __restrict__ scores, int64_t n_arg,
                       const int64_t* __restrict__ n_dev, double* __restrict__ stats,
                       uint32_t* __restrict__ shist, const int64_t* __restrict__ off = nullptr) {
  __shared__ double s_sum[kT], s_sq[kT];
  __shared__ float s_mn[kT], s_mx[kT];
  int64_t base;
  const int64_t n = tree_rows(n_arg, n_dev, off, &base);
  scores += base;
  stats += (int64_t)blockIdx.y * PYL_BUDGET_STATS;
  shist += (int64_t)blockIdx.y * 2 * kBins;
  if (blockIdx.x == 0)  // the histogram of the next launch starts from zero
    for (int j = threadIdx.x; j < 2 * kBins; j += kT) shist[j] = 0;
  int64_t b0, b1;
  chunk_bounds(n, &b0, &b1);
  double su = 0.0, sq = 0.0;
  float mn = FLT_MAX, mx = -FLT_MAX;
  for (int64_t i = b0 + threadIdx.x; i < b1; i += kT) {
    const float x = scores[i];
    su += (double)x;
    sq += (double)x * (double)x;
    mn = fminf(mn, x);
    mx = fmaxf(mx, x);
  }
  const int t = threadIdx.x;
  s_sum[t] = su;
  s_sq[t] = sq;
  s_mn[t] = mn;
  s_mx[t] = mx;
  __syncthreads();
  for (int o = kT / 2; o > 0; o >>= 1) {
    if (t < o) {
      s_sum[t] += s_sum[t + o];
      s_sq[t] += s_sq[t + o];
      s_mn[t] = fminf(s_mn[t], s_mn[t + o]);
      s_mx[t] = fmaxf(s_mx[t], s_mx[t + o]);
    }
    __syncthreads();
  }
  if (t == 0) {
    double* o = stats + 4 * (int64_t)blockIdx.x;
    o[0] = s_sum[0];
    o[1] = s_sq[0];
    o[2] = (double)s_mn[0];
    o[3] = (double)s_mx[0];
  }
}

// The chunks' partials in a fixed tree -> mean, unbiased standard deviation, min, max in
// f[0..3] (every thread of the block; the same values in every block and every run).
__device__ inline void stats_final(const double* __restrict__ stats, int64_t n, double* f) {
  static_assert(kScanBlocks == kT, "one chunk per thread");
  __shared__ double s_su[kT], s_sq[kT], s_mn[kT], s_mx[kT];
  const int t = threadIdx.x;
  s_su[t] = stats[4 * t];
  s_sq[t] = stats[4 * t + 1];
  s_mn[t] = stats[4 * t + 2];
  s_mx[t] = stats[4 * t + 3];
  __syncthreads();
  for (int o = kT / 2; o > 0; o >>= 1) {
    if (t < o) {
      s_su[t] += s_su[t + o];
      s_sq[t] += s_sq[t + o];
      s_mn[t] = fmin(s_mn[t], s_mn[t + o]);
      s_mx[t] = fmax(s_mx[t], s_mx[t + o]);
    }
    __syncthreads();
  }
  const double su = s_su[0], sq = s_sq[0];
  const double mean = n > 0 ? su / (double)n : 0.0;
  const double var = n > 1 ? (sq - su * mean) / (double)(n - 1) : 0.0;
  f[0] = mean;
  f[1] = sqrt(fmax(var, 0.0));
  f[2] = s_mn[0];
  f[3] = s_mx[0];
}

__device__ __forceinline__ int score_bin(float x, const double* f) {
  const double lo = f[2], hi = f[3];
  if (!(hi > lo)) return 0;
  int b = (int)(((double)x - lo) / (hi - lo) * (double)kBins);
  return b < 0 ? 0 : (b >= kBins ? kBins - 1 : b);
}

__global__ void __launch_bounds__(kT)
feat_hist_kernel(const float* __restrict__ scores, int64_t n_arg,
                 const int64_t* __restrict__ n_dev, double* __restrict__ stats,
                 uint32_t* __restrict__ shist, const int64_t* __restrict__ off = nullptr) {
  __shared__ uint32_t h[kBins];
  int64_t base;
  const int64_t n = tree_rows(n_arg, n_dev, off, &base);
  scores += base;
  stats += (int64_t)blockIdx.y * PYL_BUDGET_STATS;
  shist += (int64_t)blockIdx.y * 2 * kBins;
  if (blockIdx.x > 0 && (int64_t)blockIdx.x * kT >= n) return;  // the whole block
  double f[4];
  stats_final(stats, n, f);
  if (blockIdx.x == 0 && threadIdx.x < 4) stats[4 * kScanBlocks + threadIdx.x] = f[threadIdx.x];
  for (int j = threadIdx.x; j < kBins; j += kT) h[j] = 0;
  __syncthreads();
  for (int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kT)
    atomicAdd(&h[score_bin(scores[i], f)], 1u);
  __syncthreads();
  for (int j = threadIdx.x; j < kBins; j += kT)
    if (h[j]) atomicAdd(&shist[j], h[j]);
}

// The histogram and its exclusive prefix into LDS (every block; integer counts, so the
// prefix is exact whatever the order).
__device__ inline void hist_cdf(const uint32_t* __restrict__ shist, uint32_t* cnt,
                                uint32_t* cdf) {
  constexpr int kPer = kBins / kT;
  const int t = threadIdx.x;
  uint32_t acc = 0;
  for (int j = 0; j < kPer; ++j) {
    const uint32_t c = shist[t * kPer + j];
    cnt[t * kPer + j] = c;
    acc += c;
  }
  int64_t tot;
  uint32_t pre = (uint32_t)(block_scan_i64((int64_t)acc, &tot) - (int64_t)acc);
  for (int j = 0; j < kPer; ++j) {
    cdf[t * kPer + j] = pre;
    pre += cnt[t * kPer + j];
  }
  __syncthreads();
}

__global__ void __launch_bounds__(kT)
feat_write_kernel(const double* __restrict__ leaves, const float* __restrict__ scores,
                  int64_t n_arg, const int64_t* __restrict__ n_dev, int P, int level,
                  int nlevels, const double* __restrict__ stats,
                  const uint32_t* __restrict__ shist, float* __restrict__ feats,
                  const int64_t* __restrict__ off = nullptr) {
  __shared__ uint32_t cnt[kBins], cdf[kBins];
  int64_t base;
  const int64_t n = tree_rows(n_arg, n_dev, off, &base);
  leaves += base * (P + 2) * 2;
  scores += base;
  feats += base * kFeat;
  stats += (int64_t)blockIdx.y * PYL_BUDGET_STATS;
  shist += (int64_t)blockIdx.y * 2 * kBins;
  if ((int64_t)blockIdx.x * kT >= n) return;  // the whole block
  hist_cdf(shist, cnt, cdf);
  const double* f = stats + 4 * kScanBlocks;
  const int kw = P < kFeat - 4 ? P : kFeat - 4;
  for (int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kT) {
    const float x = scores[i];
    float* o = feats + i * kFeat;
    o[0] = x / 10.0f;
    o[1] = (float)level / (float)(nlevels > 1 ? nlevels : 1);
    if (n > 1) {
      o[2] = (float)(((double)x - f[0]) / (f[1] + 1e-6));
      const int b = score_bin(x, f);
      o[3] = (float)(((double)cdf[b] + 0.5 * ((double)cnt[b] - 1.0)) / (double)(n - 1));
    } else {
      o[2] = 0.0f;
      o[3] = 0.0f;
    }
    const double* lf = leaves + i * (P + 2) * 2;
    for (int j = 0; j < kFeat - 4; ++j)
      o[4 + j] = j < kw ? (float)(log10(1.0 + fabs(lf[2 * (P - kw + j) + 1])) / 10.0) : 0.0f;
  }
}

// Hybrid key: the larger of the row's rank fraction by the network's logit (from the
// histogram just built over the logits) and by S/N (feature 3), so that the cut keeps the
// rows that either ranking puts near the top.
__global__ void __launch_bounds__(kT)
hybrid_key_kernel(float* __restrict__ keys, const float* __restrict__ feats, int64_t n_arg,
                  const int64_t* __restrict__ n_dev, const double* __restrict__ stats,
                  const uint32_t* __restrict__ shist, const int64_t* __restrict__ off = nullptr) {
  __shared__ uint32_t cnt[kBins], cdf[kBins];
  int64_t base;
  const int64_t n = tree_rows(n_arg, n_dev, off, &base);
  keys += base;
  feats += base * kFeat;
  stats += (int64_t)blockIdx.y * PYL_BUDGET_STATS;
  shist += (int64_t)blockIdx.y * 2 * kBins;
  if ((int64_t)blockIdx.x * kT >= n) return;  // the whole block
  hist_cdf(shist, cnt, cdf);
  const double* f = stats + 4 * kScanBlocks;
  for (int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kT) {
    float r = 0.0f;
    if (n > 1) {
      const int b = score_bin(keys[i], f);
      r = (float)(((double)cdf[b] + 0.5 * ((double)cnt[b] - 1.0)) / (double)(n - 1));
    }
    keys[i] = fmaxf(r, feats[i * kFeat + 3]);
  }
}

inline unsigned stride_grid(int64_t cap) {
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>((cap + kT - 1) / kT, 512));
}

// Blocks per tree of a forest's strided launch: about 512 in all (the launch's count is
// fixed at enqueue time, and most trees hold far fewer rows than the capacity).
inline unsigned forest_grid(int64_t cap, int ntree) {
  return std::min<unsigned>(stride_grid(cap), (unsigned)std::max(1, 512 / ntree));
}

// off / ntree: a forest's trees (off = their survivor offsets), each with its own
// statistics; n_cap then bounds one tree's rows.
int features_enqueue(const double* leaves, const float* scores, int64_t n_cap,
                     const int64_t* n_dev, int P, int level, int nlevels, double* stats,
                     uint32_t* shist, float* feats, hipStream_t s,
                     const int64_t* off = nullptr, int ntree = 1) {
  const unsigned T = (unsigned)ntree;
  const unsigned gx = forest_grid(n_cap, ntree);
  feat_stats_part_kernel<<<dim3(kScanBlocks, T), kT, 0, s>>>(scores, n_cap, n_dev, stats,
                                                             shist, off);
  feat_hist_kernel<<<dim3(gx, T), kT, 0, s>>>(scores, n_cap, n_dev, stats, shist, off);
  feat_write_kernel<<<dim3(gx, T), kT, 0, s>>>(
      leaves, scores, n_cap, n_dev, P, level, nlevels, stats, shist, feats, off);
  PYL_CHECK_LAUNCH("feat_write_kernel");
  return PYL_OK;
}

// ---- The cut -------------------------------------------------------------------------

// The select state, and the count of rows to rank: none when all of them fit the budget
// (the features, the network and the hybrid key then skip the level).
__global__ void sel_init_kernel(int64_t* __restrict__ rec, int64_t budget,
                                uint32_t* __restrict__ sel) {
  const int64_t m = rec[kRecNsurv];
  if (threadIdx.x == 0) {
    sel[kSelPrefix] = 0;
    sel[kSelMask] = 0;
    sel[kSelRemain] = (uint32_t)std::min<int64_t>(budget, m);
    sel[kSelAll] = m <= budget ? 1u : 0u;
    sel[kSelDone] = 0;
    rec[PYL_EP_REC_NRANK] = m <= budget ? 0 : m;
  }
  sel[kSelHist + threadIdx.x] = 0;
}

// One digit pass: the histogram of the digit at `shift` over the candidates that share the
// prefix found so far; the last block to finish (a device-scope counter, after a release
// fence in every block) then picks the digit where the count of candidates from the top
// reaches the remaining budget, and clears the histogram and the counter for the next pass.
__global__ void __launch_bounds__(kT)
sel_hist_kernel(const float* __restrict__ keys, const int64_t* __restrict__ m_dev,
                uint32_t* __restrict__ sel, int shift, const int64_t* __restrict__ off = nullptr) {
  __shared__ uint32_t h[256];
  __shared__ uint32_t suf[257];
  __shared__ int s_last;
  const int t = threadIdx.x;
  sel += (int64_t)blockIdx.y * PYL_BUDGET_SEL_WORDS;  // a forest: the block's tree
  if (sel[kSelAll] != 0) return;  // every block: all rows are kept
  h[t] = 0;
  __syncthreads();
  int64_t base;
  const int64_t m = tree_rows(0, m_dev, off, &base);
  keys += base;
  const uint32_t prefix = sel[kSelPrefix], mask = sel[kSelMask];
  for (int64_t i = (int64_t)blockIdx.x * kT + t; i < m; i += (int64_t)gridDim.x * kT) {
    const uint32_t u = key_bits(keys[i]);
    if ((u & mask) == prefix) atomicAdd(&h[(u >> shift) & 255u], 1u);
  }
  __syncthreads();
  if (h[t]) atomicAdd(&sel[kSelHist + t], h[t]);
  __threadfence();
  __syncthreads();
  if (t == 0) s_last = atomicAdd(&sel[kSelDone], 1u) == gridDim.x - 1;
  __syncthreads();
  if (!s_last) return;
  __threadfence();
  if (t == 0) {
    uint32_t acc = 0;
    suf[256] = 0;
    for (int d = 255; d >= 0; --d) {
      acc += __hip_atomic_load(&sel[kSelHist + d], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      suf[d] = acc;
    }
  }
  __syncthreads();
  const uint32_t rem = sel[kSelRemain];
  __syncthreads();  // every thread has read the state before it changes
  if (suf[t + 1] < rem && rem <= suf[t]) {
    sel[kSelPrefix] = prefix | ((uint32_t)t << shift);
    sel[kSelMask] = mask | (0xffu << shift);
    sel[kSelRemain] = rem - suf[t + 1];
  }
  sel[kSelHist + t] = 0;
  if (t == 0) sel[kSelDone] = 0;
}

struct KeepSrc {  // packed: 1 = above the cut, 1 << 32 = at the cut
  const float* keys;
  const uint32_t* sel;
  __device__ int64_t operator()(int64_t i) const {
    if (sel[kSelAll]) return 1;
    const uint32_t u = key_bits(keys[i]);
    const uint32_t t = sel[kSelPrefix];
    return u > t ? 1 : (u == t ? ((int64_t)1 << 32) : 0);
  }
};

struct KeepSink {
  int64_t* kept;
  const uint32_t* sel;
  int64_t* rec;
  __device__ void operator()(int64_t i, int64_t excl, int64_t x) const {
    if (!x) return;
    const int64_t above = excl & 0xffffffffLL, at = excl >> 32;
    if (sel[kSelAll]) {
      kept[above] = i;
      return;
    }
    const int64_t need = sel[kSelRemain];
    if (x & 0xffffffffLL) kept[above + (at < need ? at : need)] = i;
    else if (at < need) kept[above + at] = i;
  }
  __device__ void finish(int64_t tot) const {
    const int64_t m = rec[kRecNsurv];
    const int64_t need = sel[kSelRemain];
    const int64_t k = sel[kSelAll] ? m : (tot & 0xffffffffLL) + std::min<int64_t>(tot >> 32, need);
    rec[PYL_EP_REC_NKEPT] = k;
    rec[kRecWords + kRecNpar] = k;
    rec[kRecWords + kRecHist] = rec[kRecHist] + k;
  }
};

// One wavefront per kept row: leaves, fold row, score, and the backtrack row into the
// history at the level's rows.
__global__ void __launch_bounds__(kT)
budget_gather_kernel(const int64_t* __restrict__ kept, const int64_t* __restrict__ rec,
                     int P, int row_floats, const double* __restrict__ leaves,
                     const float* __restrict__ folds, const float* __restrict__ scores,
                     const int32_t* __restrict__ bt, double* __restrict__ o_leaves,
                     float* __restrict__ o_folds, float* __restrict__ o_scores,
                     int32_t* __restrict__ o_hist) {
  const int64_t k = rec[kRecWords + kRecNpar];
  const int lane = threadIdx.x & 63;
  const int64_t r = (int64_t)blockIdx.x * (kT / 64) + (threadIdx.x >> 6);
  if (r >= k) return;
  const int64_t i = kept[r];
  const int nl = (P + 2) * 2;
  const int nb = P + 2;
  if (lane < nl) o_leaves[r * nl + lane] = leaves[i * nl + lane];
  for (int j = lane; j < row_floats; j += 64)
    o_folds[r * row_floats + j] = folds[i * row_floats + j];
  if (lane == 0) o_scores[r] = scores[i];
  if (lane < nb) o_hist[(rec[kRecHist] + r) * nb + lane] = bt[i * nb + lane];
}

// A level enqueued without the cut: one wavefront per survivor copies it as the gather
// does (row r = survivor r) when the survivors fit the budget; else the level stops (no
// kept rows, no parents for the next level) and runs per level.
__global__ void __launch_bounds__(kT)
budget_pass_kernel(int64_t* __restrict__ rec, int64_t budget, int P, int row_floats,
                   const double* __restrict__ leaves, const float* __restrict__ folds,
                   const float* __restrict__ scores, const int32_t* __restrict__ bt,
                   double* __restrict__ o_leaves, float* __restrict__ o_folds,
                   float* __restrict__ o_scores, int32_t* __restrict__ o_hist) {
  const int64_t m = rec[kRecNsurv];
  const int64_t h = rec[kRecHist];
  const bool over = m > budget;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    const int64_t k = over ? 0 : m;
    if (over) rec[kRecStatus] = PYL_EP_ST_BUDGET;
    rec[PYL_EP_REC_NKEPT] = k;
    rec[kRecWords + kRecNpar] = k;
    rec[kRecWords + kRecHist] = h + k;
  }
  const int64_t r = (int64_t)blockIdx.x * (kT / 64) + (threadIdx.x >> 6);
  if (over || r >= m) return;
  const int lane = threadIdx.x & 63;
  const int nl = (P + 2) * 2;
  const int nb = P + 2;
  if (lane < nl) o_leaves[r * nl + lane] = leaves[r * nl + lane];
  for (int j = lane; j < row_floats; j += 64)
    o_folds[r * row_floats + j] = folds[r * row_floats + j];
  if (lane == 0) o_scores[r] = scores[r];
  if (lane < nb) o_hist[(h + r) * nb + lane] = bt[r * nb + lane];
}

__global__ void budget_record_kernel(int64_t* __restrict__ rec, int64_t n) {
  const int w = threadIdx.x;
  if (w < 2 * kRecWords) rec[w] = 0;
  if (w == kRecNsurv) rec[w] = n;
}

// ---- Forests: the budget per tree ----------------------------------------------------
// Tree t's survivors are rows [so[t], so[t + 1]) of the level's materialised rows; each
// tree is ranked and cut on its own (its own select state, statistics and histogram), its
// kept rows' indices go to kept[so[t]..], and the kept rows of all trees are gathered one
// tree after the other (ko = their offsets, the next level's parents).

// One block per tree. The level's rows are ranked (NRANK = NSURV) when any tree's
// survivors exceed the budget; the record starts at zero, so the blocks that find one
// all write the same value.
__global__ void fsel_init_kernel(int64_t* __restrict__ rec, const int64_t* __restrict__ so,
                                 int64_t budget, uint32_t* __restrict__ sel) {
  const int y = blockIdx.x;
  const int64_t m = so[y + 1] - so[y];
  uint32_t* st = sel + (int64_t)y * PYL_BUDGET_SEL_WORDS;
  if (threadIdx.x == 0) {
    st[kSelPrefix] = 0;
    st[kSelMask] = 0;
    st[kSelRemain] = (uint32_t)std::min<int64_t>(budget, m);
    st[kSelAll] = m <= budget ? 1u : 0u;
    st[kSelDone] = 0;
    if (m > budget) rec[PYL_EP_REC_NRANK] = rec[kRecNsurv];
  }
  st[kSelHist + threadIdx.x] = 0;
}

struct FKeepSrc {
  const float* keys;
  const uint32_t* sel;
  __device__ int64_t operator()(int64_t i, int y) const {
    return KeepSrc{keys, sel + (int64_t)y * PYL_BUDGET_SEL_WORDS}(i);
  }
};

struct FKeepSink {
  int64_t* kept;
  const uint32_t* sel;
  const int64_t* so;
  int64_t* kcount;
  __device__ void operator()(int64_t i, int64_t excl, int64_t x, int y) const {
    if (!x) return;
    const uint32_t* st = sel + (int64_t)y * PYL_BUDGET_SEL_WORDS;
    int64_t* kt = kept + so[y];
    const int64_t above = excl & 0xffffffffLL, at = excl >> 32;
    if (st[kSelAll]) {
      kt[above] = i;
      return;
    }
    const int64_t need = st[kSelRemain];
    if (x & 0xffffffffLL) kt[above + (at < need ? at : need)] = i;
    else if (at < need) kt[above + at] = i;
  }
  __device__ void finish(int y, int64_t tot) const {
    const uint32_t* st = sel + (int64_t)y * PYL_BUDGET_SEL_WORDS;
    const int64_t need = st[kSelRemain];
    kcount[y] = st[kSelAll] ? so[y + 1] - so[y]
                            : (tot & 0xffffffffLL) + std::min<int64_t>(tot >> 32, need);
  }
};

// The kept rows' offsets per tree (the next level's parents) and the level's totals. With
// pass set (a level enqueued without the cut) every tree keeps its survivors, or, when
// any tree's exceed the budget, the level stops with PYL_EP_ST_BUDGET and keeps none.
__global__ void fkeep_offsets_kernel(int64_t* __restrict__ rec, int64_t* __restrict__ tab,
                                     int64_t* __restrict__ tab_next, int T,
                                     const int64_t* __restrict__ kcount, int pass,
                                     int64_t budget) {
  if (threadIdx.x != 0) return;
  const int64_t* so = tab + PYL_EP_TAB_SO * (T + 1);
  int64_t* ko = tab + PYL_EP_TAB_KO * (T + 1);
  bool over = false;
  if (pass)
    for (int t = 0; t < T; ++t) over = over || so[t + 1] - so[t] > budget;
  if (over && rec[kRecStatus] == 0) rec[kRecStatus] = PYL_EP_ST_BUDGET;
  int64_t acc = 0;
  for (int t = 0; t < T; ++t) {
    ko[t] = acc;
    tab_next[t] = acc;  // PO of the next level
    acc += over ? 0 : (pass ? so[t + 1] - so[t] : kcount[t]);
  }
  ko[T] = acc;
  tab_next[T] = acc;
  rec[PYL_EP_REC_NKEPT] = acc;
  rec[kRecWords + kRecNpar] = acc;
  rec[kRecWords + kRecHist] = rec[kRecHist] + acc;
}

// One wavefront per kept row r (tree t, ko[t] <= r < ko[t + 1]): row kept[so[t] + r -
// ko[t]] of the materialised rows (row r itself when kept is null), into o_* and history
// row HIST + r.
__global__ void __launch_bounds__(kT)
fgather_kernel(const int64_t* __restrict__ tab, int T, const int64_t* __restrict__ kept,
               const int64_t* __restrict__ rec, int P, int row_floats,
               const double* __restrict__ leaves, const float* __restrict__ folds,
               const float* __restrict__ scores, const int32_t* __restrict__ bt,
               double* __restrict__ o_leaves, float* __restrict__ o_folds,
               float* __restrict__ o_scores, int32_t* __restrict__ o_hist) {
  const int64_t k = rec[kRecWords + kRecNpar];
  const int lane = threadIdx.x & 63;
  const int64_t r = (int64_t)blockIdx.x * (kT / 64) + (threadIdx.x >> 6);
  if (r >= k) return;
  int64_t i = r;
  if (kept) {
    const int64_t* so = tab + PYL_EP_TAB_SO * (T + 1);
    const int64_t* ko = tab + PYL_EP_TAB_KO * (T + 1);
    int lo = 0, hi = T;
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (ko[mid] <= r) lo = mid;
      else hi = mid;
    }
    i = kept[so[lo] + r - ko[lo]];
  }
  const int nl = (P + 2) * 2;
  const int nb = P + 2;
  if (lane < nl) o_leaves[r * nl + lane] = leaves[i * nl + lane];
  for (int j = lane; j < row_floats; j += 64)
    o_folds[r * row_floats + j] = folds[i * row_floats + j];
  if (lane == 0) o_scores[r] = scores[i];
  if (lane < nb) o_hist[(rec[kRecHist] + r) * nb + lane] = bt[i * nb + lane];
}

// The cut of every tree of a forest on `keys` (survivor order): its kept rows' indices,
// the trees' kept offsets and the level's totals (the select state initialised first
// unless `inited`).
int forest_cut(const pyl_ep_budget* b, const pyl_ep_forest* F, int64_t bud, int64_t* tab,
               int64_t* tab_next, const float* keys, int64_t n_cap, int64_t* rec, bool inited,
               hipStream_t s) {
  const int T = F->ntree;
  const int64_t* so = tab + PYL_EP_TAB_SO * (T + 1);
  const int64_t tree_cap = std::min<int64_t>(n_cap, F->max_sugg);
  if (!inited) fsel_init_kernel<<<(unsigned)T, 256, 0, s>>>(rec, so, bud, F->sel);
  const dim3 gsel(forest_grid(tree_cap, T), (unsigned)T);
  for (int shift = 24; shift >= 0; shift -= 8)
    sel_hist_kernel<<<gsel, kT, 0, s>>>(keys, nullptr, F->sel, shift, so);
  chunked_scan_seg(FKeepSrc{keys, F->sel}, FKeepSink{b->kept, F->sel, so, F->kcount}, so, T,
                   F->part, s);
  fkeep_offsets_kernel<<<1, 64, 0, s>>>(rec, tab, tab_next, T, F->kcount, 0, bud);
  PYL_CHECK_LAUNCH("fkeep_offsets_kernel");
  return PYL_OK;
}

}  // namespace

int pyl::ep_budget_forest_enqueue(const pyl_ep_budget* b, const pyl_ep_forest* F, int P,
                                  int level, int64_t* tab, int64_t* tab_next,
                                  const double* leaves, const float* folds,
                                  const float* scores, const int32_t* bt, int64_t n_cap,
                                  int64_t* rec, double* o_leaves, float* o_folds,
                                  float* o_scores, int32_t* o_hist, hipStream_t s) {
  PYL_REQUIRE(b != nullptr && F != nullptr, "null budget");
  PYL_REQUIRE(level >= 0 && level <= b->nlevels, "level %d outside the search", level);
  PYL_REQUIRE(b->kind == PYL_BUDGET_SCORE || b->kind == PYL_BUDGET_CNN ||
                  b->kind == PYL_BUDGET_HYBRID,
              "unknown budget kind %d", b->kind);
  const int kind = kind_of(b, level);
  PYL_REQUIRE(kind == PYL_BUDGET_SCORE || kind == b->kind,
              "level %d ranks by kind %d, the budget by %d", level, kind, b->kind);
  const int64_t bud = budget_of(b, level);
  PYL_REQUIRE(bud >= 1, "budget must be >= 1");
  PYL_REQUIRE(n_cap >= 0 && n_cap < 0x7fffffffLL, "too many survivors for the budget");
  const int T = F->ntree;
  const int row_floats = b->complex_folds ? 4 * (b->nbins / 2 + 1) : 2 * b->nbins;
  const int64_t* so = tab + PYL_EP_TAB_SO * (T + 1);
  const int64_t rows = std::min<int64_t>(n_cap, (int64_t)T * bud);
  const unsigned g_rows = (unsigned)std::max<int64_t>(1, pyl::ceil_div(rows, kT / 64));
  if (b->level_mask != nullptr && b->level_mask[level] == 0) {
    fkeep_offsets_kernel<<<1, 64, 0, s>>>(rec, tab, tab_next, T, F->kcount, 1, bud);
    fgather_kernel<<<g_rows, kT, 0, s>>>(tab, T, nullptr, rec, P, row_floats, leaves, folds,
                                         scores, bt, o_leaves, o_folds, o_scores, o_hist);
    PYL_CHECK_LAUNCH("fgather_kernel");
    return PYL_OK;
  }
  const bool net = kind != PYL_BUDGET_SCORE;
  PYL_REQUIRE(b->kept && F->sel && F->part && F->kcount &&
                  (!net || (b->params && b->keys && b->feats && F->stats && F->shist)),
              "null budget buffer");
  const int64_t* r_dev = rec + PYL_EP_REC_NRANK;  // 0 when every tree's survivors fit
  const int64_t tree_cap = std::min<int64_t>(n_cap, F->max_sugg);
  const float* keys = scores;
  if (net) {
    fsel_init_kernel<<<(unsigned)T, 256, 0, s>>>(rec, so, bud, F->sel);
    int rc = features_enqueue(leaves, scores, tree_cap, r_dev, P, level, b->nlevels, F->stats,
                              F->shist, b->feats, s, so, T);
    if (rc != PYL_OK) return rc;
    rc = pyl::cnn_forward_count(b->params, folds, n_cap, r_dev, b->nbins, b->complex_folds,
                                b->feats, b->keys, b->precision == PYL_BUDGET_BF16, s);
    if (rc != PYL_OK) return rc;
    keys = b->keys;
    if (kind == PYL_BUDGET_HYBRID) {
      const dim3 gs(kScanBlocks, (unsigned)T), gr(forest_grid(tree_cap, T), (unsigned)T);
      feat_stats_part_kernel<<<gs, kT, 0, s>>>(b->keys, tree_cap, r_dev, F->stats, F->shist,
                                               so);
      feat_hist_kernel<<<gr, kT, 0, s>>>(b->keys, tree_cap, r_dev, F->stats, F->shist, so);
      hybrid_key_kernel<<<gr, kT, 0, s>>>(b->keys, b->feats, tree_cap, r_dev, F->stats,
                                          F->shist, so);
      PYL_CHECK_LAUNCH("hybrid_key_kernel");
    }
  }
  int rc = forest_cut(b, F, bud, tab, tab_next, keys, n_cap, rec, net, s);
  if (rc != PYL_OK) return rc;
  fgather_kernel<<<g_rows, kT, 0, s>>>(tab, T, b->kept, rec, P, row_floats, leaves, folds,
                                       scores, bt, o_leaves, o_folds, o_scores, o_hist);
  PYL_CHECK_LAUNCH("fgather_kernel");
  return PYL_OK;
}

int pyl::ep_budget_forest_select(const pyl_ep_budget* b, const pyl_ep_forest* F, int level,
                                 int64_t* tab, int64_t* tab_next, const float* keys,
                                 int64_t n_cap, int64_t* rec, hipStream_t s) {
  PYL_REQUIRE(b != nullptr && F != nullptr, "null budget");
  PYL_REQUIRE(level >= 0 && level <= b->nlevels, "level %d outside the search", level);
  PYL_REQUIRE(kind_of(b, level) == PYL_BUDGET_SCORE, "the select-first cut ranks by S/N only");
  const int64_t bud = budget_of(b, level);
  PYL_REQUIRE(bud >= 1, "budget must be >= 1");
  PYL_REQUIRE(n_cap >= 0 && n_cap < 0x7fffffffLL, "too many survivors for the budget");
  const int T = F->ntree;
  if (b->level_mask != nullptr && b->level_mask[level] == 0) {
    fkeep_offsets_kernel<<<1, 64, 0, s>>>(rec, tab, tab_next, T, F->kcount, 1, bud);
    PYL_CHECK_LAUNCH("fkeep_offsets_kernel");
    return PYL_OK;
  }
  PYL_REQUIRE(b->kept && keys && F->sel && F->part && F->kcount, "null budget buffer");
  return forest_cut(b, F, bud, tab, tab_next, keys, n_cap, rec, false, s);
}

int pyl::ep_budget_enqueue(const pyl_ep_budget* b, int P, int level, const double* leaves,
                           const float* folds, const float* scores, const int32_t* bt,
                           int64_t n_cap, int64_t* rec, double* o_leaves, float* o_folds,
                           float* o_scores, int32_t* o_hist, hipStream_t s) {
  PYL_REQUIRE(b != nullptr, "null budget");
  PYL_REQUIRE(b->kind == PYL_BUDGET_SCORE || b->kind == PYL_BUDGET_CNN ||
                  b->kind == PYL_BUDGET_HYBRID,
              "unknown budget kind %d", b->kind);
  PYL_REQUIRE(level >= 0 && level <= b->nlevels, "level %d outside the search", level);
  const int kind = kind_of(b, level);
  PYL_REQUIRE(kind == PYL_BUDGET_SCORE || kind == b->kind,
              "level %d ranks by kind %d, the budget by %d", level, kind, b->kind);
  const int64_t bud = budget_of(b, level);
  PYL_REQUIRE(bud >= 1, "budget must be >= 1");
  PYL_REQUIRE(n_cap >= 0 && n_cap < 0x7fffffffLL, "too many survivors for the budget");
  const bool net = kind != PYL_BUDGET_SCORE;
  PYL_REQUIRE(b->kept && b->sel && b->part &&
                  (!net || (b->params && b->keys && b->feats && b->stats && b->shist)),
              "null budget buffer");
  const int64_t* m_dev = rec + kRecNsurv;
  const int64_t* r_dev = rec + PYL_EP_REC_NRANK;  // 0 when every survivor is kept
  sel_init_kernel<<<1, 256, 0, s>>>(rec, bud, b->sel);
  const float* keys = scores;
  if (net) {
    int rc = features_enqueue(leaves, scores, n_cap, r_dev, P, level, b->nlevels, b->stats,
                              b->shist, b->feats, s);
    if (rc != PYL_OK) return rc;
    rc = pyl::cnn_forward_count(b->params, folds, n_cap, r_dev, b->nbins, b->complex_folds,
                                b->feats, b->keys, b->precision == PYL_BUDGET_BF16, s);
    if (rc != PYL_OK) return rc;
    keys = b->keys;
    if (kind == PYL_BUDGET_HYBRID) {  // the logits' histogram, then the hybrid key
      feat_stats_part_kernel<<<kScanBlocks, kT, 0, s>>>(b->keys, n_cap, r_dev, b->stats,
                                                         b->shist);
      feat_hist_kernel<<<stride_grid(n_cap), kT, 0, s>>>(b->keys, n_cap, r_dev, b->stats,
                                                         b->shist);
      hybrid_key_kernel<<<stride_grid(n_cap), kT, 0, s>>>(b->keys, b->feats, n_cap, r_dev,
                                                          b->stats, b->shist);
      PYL_CHECK_LAUNCH("hybrid_key_kernel");
    }
  }
  for (int shift = 24; shift >= 0; shift -= 8) {
    sel_hist_kernel<<<stride_grid(n_cap), kT, 0, s>>>(keys, m_dev, b->sel, shift);
  }
  chunked_scan(KeepSrc{keys, b->sel}, KeepSink{b->kept, b->sel, rec}, m_dev, b->part, s);
  const int row_floats = b->complex_folds ? 4 * (b->nbins / 2 + 1) : 2 * b->nbins;
  const int64_t rows = std::min<int64_t>(n_cap, bud);
  budget_gather_kernel<<<(unsigned)std::max<int64_t>(1, pyl::ceil_div(rows, kT / 64)), kT, 0,
                         s>>>(b->kept, rec, P, row_floats, leaves, folds, scores, bt,
                              o_leaves, o_folds, o_scores, o_hist);
  PYL_CHECK_LAUNCH("budget_gather_kernel");
  return PYL_OK;
}

int pyl::ep_budget_level_enqueue(const pyl_ep_budget* b, int P, int level,
                                 const double* leaves, const float* folds,
                                 const float* scores, const int32_t* bt, int64_t n_cap,
                                 int64_t* rec, double* o_leaves, float* o_folds,
                                 float* o_scores, int32_t* o_hist, hipStream_t s) {
  PYL_REQUIRE(b != nullptr, "null budget");
  PYL_REQUIRE(level >= 0 && level <= b->nlevels, "level %d outside the search", level);
  if (b->level_mask == nullptr || b->level_mask[level] != 0)
    return ep_budget_enqueue(b, P, level, leaves, folds, scores, bt, n_cap, rec, o_leaves,
                             o_folds, o_scores, o_hist, s);
  const int64_t bud = budget_of(b, level);
  PYL_REQUIRE(bud >= 1, "budget must be >= 1");
  const int row_floats = b->complex_folds ? 4 * (b->nbins / 2 + 1) : 2 * b->nbins;
  const int64_t rows = std::min<int64_t>(n_cap, bud);
  budget_pass_kernel<<<(unsigned)std::max<int64_t>(1, pyl::ceil_div(rows, kT / 64)), kT, 0,
                       s>>>(rec, bud, P, row_floats, leaves, folds, scores, bt,
                            o_leaves, o_folds, o_scores, o_hist);
  PYL_CHECK_LAUNCH("budget_pass_kernel");
  return PYL_OK;
}

extern "C" int pyl_ep_budget_apply(const pyl_ep_budget* b, int32_t P, int32_t level,
                                   const double* leaves, const float* folds,
                                   const float* scores, const int32_t* bt, int64_t n,
                                   int64_t* rec, double* out_leaves, float* out_folds,
                                   float* out_scores, int32_t* out_bt, void* stream) {
  PYL_REQUIRE(b != nullptr, "null budget");
  PYL_REQUIRE(P >= 2 && P <= 5, "P must be in [2, 5]");
  PYL_REQUIRE(n >= 0, "negative row count");
  PYL_REQUIRE(leaves && folds && scores && bt && rec && out_leaves && out_folds && out_scores &&
                  out_bt,
              "null device pointer");
  const hipStream_t s = pyl::as_stream(stream);
  budget_record_kernel<<<1, 64, 0, s>>>(rec, n);
  PYL_CHECK_LAUNCH("budget_record_kernel");
  return pyl::ep_budget_enqueue(b, P, level, leaves, folds, scores, bt, n, rec, out_leaves,
                                out_folds, out_scores, out_bt, s);
}

extern "C" int pyl_cnn_features(const double* leaves, const float* scores, int64_t n, int32_t P,
                                int32_t level, int32_t nlevels, double* stats, uint32_t* shist,
                                float* feats, void* stream) {
  PYL_REQUIRE(P >= 2 && P <= 5, "P must be in [2, 5]");
  PYL_REQUIRE(n >= 0, "negative row count");
  if (n == 0) return PYL_OK;
  PYL_REQUIRE(leaves && scores && stats && shist && feats, "null device pointer");
  return features_enqueue(leaves, scores, n, nullptr, P, level, nlevels, stats, shist, feats,
                          pyl::as_stream(stream));
}
