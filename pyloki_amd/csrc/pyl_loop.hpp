// Device-side pieces of the sync-free EP level loop: the per-level record words, and the
// chunked scans that turn per-item counts or a threshold test into offsets / indices with
// the item count read from device memory.
//
// A chunked scan is two launches of kScanBlocks blocks: block b owns the contiguous chunk
// [b*ceil(n/G), (b+1)*ceil(n/G)) of the n items; the first launch writes each chunk's sum,
// the second re-scans each chunk from the sum of the chunks before it. Results are exact
// integers, so they do not depend on the chunking; the second launch's last block also
// hands the grand total to a "finish" step that updates the level record. Sums are taken
// modulo 2^64 (unsigned arithmetic): a packed source whose grand total wraps still gives
// exact differences of prefix sums wherever those differences fit.
#pragma once

#include <cstdint>
#include <vector>

#include "pyl_common.hpp"

// Phase timer of the level loops (opaque in the C ABI): event pairs on the loop's stream.
struct pyl_ep_timer {
  int32_t capacity = 0;
  std::vector<hipEvent_t> ev;  // 2 per record
  std::vector<int32_t> phase;
  int32_t n = 0;
  bool open = false;
  void begin(int32_t ph, hipStream_t s) {
    if (n >= capacity) return;
    phase[n] = ph;
    (void)hipEventRecord(ev[2 * n], s);
    open = true;
  }
  void end(hipStream_t s) {
    if (!open) return;
    (void)hipEventRecord(ev[2 * n + 1], s);
    ++n;
    open = false;
  }
};

namespace pyl_ep {
namespace {  // internal linkage: every translation unit gets its own copy

constexpr int kScanBlocks = PYL_EP_SCAN_BLOCKS;
constexpr int kScanThreads = 256;

// Level record words (int64), one record per level plus one for the level after the run.
constexpr int kRecNpar = PYL_EP_REC_NPAR;
constexpr int kRecNchild = PYL_EP_REC_NCHILD;
constexpr int kRecNsurv = PYL_EP_REC_NSURV;
constexpr int kRecStatus = PYL_EP_REC_STATUS;
constexpr int kRecErr = PYL_EP_REC_ERR;
constexpr int kRecHist = PYL_EP_REC_HIST;
constexpr int kRecNvalid = PYL_EP_REC_NVALID;
constexpr int kRecWords = PYL_EP_REC_WORDS;

// Wrapping int64 arithmetic (through uint64).
__device__ __forceinline__ int64_t wadd(int64_t a, int64_t b) {
  return (int64_t)((uint64_t)a + (uint64_t)b);
}
__device__ __forceinline__ int64_t wsub(int64_t a, int64_t b) {
  return (int64_t)((uint64_t)a - (uint64_t)b);
}

// Sum over the block (all threads get it).
__device__ __forceinline__ int64_t block_sum_i64(int64_t x0) {
  __shared__ uint64_t s_red[kScanThreads / 64];
  uint64_t x = (uint64_t)x0;
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) s_red[w] = x;
  __syncthreads();
  uint64_t t = 0;
  for (int i = 0; i < kScanThreads / 64; ++i) t += s_red[i];
  __syncthreads();
  return (int64_t)t;
}

// Inclusive scan over the block; *total receives the block's sum.
__device__ __forceinline__ int64_t block_scan_i64(int64_t x0, int64_t* total) {
  __shared__ uint64_t s_wave[kScanThreads / 64];
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  uint64_t x = (uint64_t)x0;
  for (int o = 1; o < 64; o <<= 1) {
    const uint64_t y = __shfl_up(x, o);
    if (lane >= o) x += y;
  }
  if (lane == 63) s_wave[w] = x;
  __syncthreads();
  uint64_t pre = 0, tot = 0;
  for (int i = 0; i < kScanThreads / 64; ++i) {
    const uint64_t v = s_wave[i];
    pre += i < w ? v : 0;
    tot += v;
  }
  __syncthreads();
  *total = (int64_t)tot;
  return (int64_t)(x + pre);
}

__device__ __forceinline__ void chunk_bounds(int64_t n, int64_t* b0, int64_t* b1) {
  const int64_t chunk = (n + gridDim.x - 1) / gridDim.x;
  *b0 = std::min<int64_t>(n, (int64_t)blockIdx.x * chunk);
  *b1 = std::min<int64_t>(n, *b0 + chunk);
}

// Item sources.
struct CountSrc {  // x[i]
  const int64_t* x;
  __device__ int64_t operator()(int64_t i) const { return x[i]; }
};
struct PassSrc {  // score[i] (widened to FP64) >= threshold: the survivor test
  const float* score;
  double thr;
  __device__ int64_t operator()(int64_t i) const { return (double)score[i] >= thr ? 1 : 0; }
};
struct FlagSrc {  // flag[i] != 0
  const uint8_t* flag;
  __device__ int64_t operator()(int64_t i) const { return flag[i] != 0 ? 1 : 0; }
};

template <class Src>
__global__ void __launch_bounds__(kScanThreads)
chunk_sum_kernel(Src src, const int64_t* __restrict__ n_dev, int64_t* __restrict__ part) {
  int64_t b0, b1;
  chunk_bounds(*n_dev, &b0, &b1);
  int64_t acc = 0;
  for (int64_t i = b0 + threadIdx.x; i < b1; i += kScanThreads) acc = wadd(acc, src(i));
  acc = block_sum_i64(acc);
  if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

// sink(i, exclusive prefix, value) for every item; sink.finish(total) once, by the last
// block, after its chunk.
template <class Src, class Sink>
__global__ void __launch_bounds__(kScanThreads)
chunk_scan_kernel(Src src, Sink sink, const int64_t* __restrict__ n_dev,
                  const int64_t* __restrict__ part) {
  int64_t b0, b1;
  chunk_bounds(*n_dev, &b0, &b1);
  int64_t pre = 0, tot = 0;
  for (int i = threadIdx.x; i < (int)gridDim.x; i += kScanThreads) {
    const int64_t v = part[i];
    pre = wadd(pre, i < (int)blockIdx.x ? v : 0);
    tot = wadd(tot, v);
  }
  pre = block_sum_i64(pre);
  tot = block_sum_i64(tot);
  for (int64_t base = b0; base < b1; base += kScanThreads) {
    const int64_t i = base + threadIdx.x;
    const int64_t x = i < b1 ? src(i) : 0;
    int64_t t;
    const int64_t inc = block_scan_i64(x, &t);
    if (i < b1) sink(i, wsub(wadd(pre, inc), x), x);
    pre = wadd(pre, t);
  }
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) sink.finish(tot);
}

template <class Src, class Sink>
void chunked_scan(Src src, Sink sink, const int64_t* n_dev, int64_t* part, hipStream_t s) {
  chunk_sum_kernel<Src><<<kScanBlocks, kScanThreads, 0, s>>>(src, n_dev, part);
  chunk_scan_kernel<Src, Sink><<<kScanBlocks, kScanThreads, 0, s>>>(src, sink, n_dev, part);
}

// Chunked scans over the trees of a forest at once: tree y = blockIdx.y holds the items
// [off[y], off[y + 1]) and its chunk sums in part + y * kScanBlocks; src(i, y), then
// sink(i, exclusive prefix within the tree, value, y), and sink.finish(y, total) once per
// tree. kSegScanBlocks chunks per tree.
constexpr int kSegScanBlocks = 64;

__device__ __forceinline__ void seg_chunk_bounds(const int64_t* off, int64_t* b0, int64_t* b1) {
  const int64_t lo = off[blockIdx.y], n = off[blockIdx.y + 1] - lo;
  const int64_t chunk = (n + gridDim.x - 1) / gridDim.x;
  *b0 = lo + std::min<int64_t>(n, (int64_t)blockIdx.x * chunk);
  *b1 = lo + std::min<int64_t>(n, (int64_t)(blockIdx.x + 1) * chunk);
}

template <class Src>
__global__ void __launch_bounds__(kScanThreads)
chunk_sum_seg_kernel(Src src, const int64_t* __restrict__ off, int64_t* __restrict__ part) {
  int64_t b0, b1;
  seg_chunk_bounds(off, &b0, &b1);
  int64_t acc = 0;
  for (int64_t i = b0 + threadIdx.x; i < b1; i += kScanThreads) acc += src(i, blockIdx.y);
  acc = block_sum_i64(acc);
  if (threadIdx.x == 0) part[(int64_t)blockIdx.y * kScanBlocks + blockIdx.x] = acc;
}

template <class Src, class Sink>
__global__ void __launch_bounds__(kScanThreads)
chunk_scan_seg_kernel(Src src, Sink sink, const int64_t* __restrict__ off,
                      const int64_t* __restrict__ part) {
  const int y = blockIdx.y;
  int64_t b0, b1;
  seg_chunk_bounds(off, &b0, &b1);
  const int64_t* pt = part + (int64_t)y * kScanBlocks;
  int64_t pre = 0, tot = 0;
  for (int i = threadIdx.x; i < (int)gridDim.x; i += kScanThreads) {
    const int64_t v = pt[i];
    pre += i < (int)blockIdx.x ? v : 0;
    tot += v;
  }
  pre = block_sum_i64(pre);
  tot = block_sum_i64(tot);
  for (int64_t base = b0; base < b1; base += kScanThreads) {
    const int64_t i = base + threadIdx.x;
    const int64_t x = i < b1 ? src(i, y) : 0;
    int64_t t;
    const int64_t inc = block_scan_i64(x, &t);
    if (i < b1) sink(i, pre + inc - x, x, y);
    pre += t;
  }
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) sink.finish(y, tot);
}

template <class Src, class Sink>
void chunked_scan_seg(Src src, Sink sink, const int64_t* off, int ntree, int64_t* part,
                      hipStream_t s) {
  static_assert(kSegScanBlocks <= kScanBlocks, "chunk sums per tree");
  const dim3 g(kSegScanBlocks, (unsigned)ntree);
  chunk_sum_seg_kernel<Src><<<g, kScanThreads, 0, s>>>(src, off, part);
  chunk_scan_seg_kernel<Src, Sink><<<g, kScanThreads, 0, s>>>(src, sink, off, part);
}

// Children of a level: off[i + 1] = inclusive count; the total becomes the level's child
// count unless the branch step failed, the children exceed the buffers or the history
// could not take a full survivor buffer (checked here, before any child is scored).
struct ChildOffSink {
  int64_t* off;
  int64_t* rec;  // this level's record
  int64_t cap_child;
  int64_t max_sugg;
  int64_t cap_hist;
  __device__ void operator()(int64_t i, int64_t excl, int64_t x) const { off[i + 1] = excl + x; }
  __device__ void finish(int64_t total) const {
    off[0] = 0;
    int64_t st = rec[kRecStatus];
    if (st == 0 && rec[kRecErr] != INT64_MAX) st = PYL_EP_ST_BRANCH;
    if (st == 0 && total > cap_child) st = PYL_EP_ST_CHILD_CAP;
    if (st == 0 && rec[kRecHist] + max_sugg > cap_hist) st = PYL_EP_ST_HIST_FULL;
    rec[kRecStatus] = st;
    rec[kRecNchild] = st ? 0 : total;
    rec[kRecNvalid] = st ? 0 : total;
  }
};

// Survivors of a level: the indices of the passing children in increasing order. They
// fit the survivor buffer, else the level stops (the overflow rule runs on the host).
// The next level's parent count and backtrack row follow.
struct SurvSink {
  int64_t* surv;
  int64_t* rec;  // this level's record; rec + kRecWords is the next level's
  int64_t max_sugg;
  int64_t cap_hist;
  const float* score = nullptr;  // with keys: keys[j] = score of survivor j
  float* keys = nullptr;
  __device__ void operator()(int64_t i, int64_t excl, int64_t x) const {
    if (x && excl < max_sugg) {
      surv[excl] = i;
      if (keys) keys[excl] = score[i];
    }
  }
  __device__ void finish(int64_t npass) const {
    int64_t st = rec[kRecStatus];
    if (st == 0 && npass > max_sugg) st = PYL_EP_ST_OVERFLOW;
    if (st == 0 && rec[kRecHist] + npass > cap_hist) st = PYL_EP_ST_HIST_FULL;
    const int64_t m = st ? 0 : npass;
    rec[kRecStatus] = st;
    rec[kRecNsurv] = m;
    rec[kRecWords + kRecNpar] = m;
    rec[kRecWords + kRecHist] = rec[kRecHist] + m;
  }
};

// Records of a run: level 0 gets the input tree's size and the first backtrack row.
__global__ void ep_records_init_kernel(int64_t* __restrict__ rec, float* __restrict__ minmax,
                                       int nlev, int64_t n_in, int64_t hist_row0) {
  const int l = blockIdx.x * blockDim.x + threadIdx.x;
  if (l > nlev) return;
  int64_t* r = rec + (int64_t)l * kRecWords;
  for (int w = 0; w < kRecWords; ++w) r[w] = 0;
  r[kRecErr] = INT64_MAX;
  if (l == 0) {
    r[kRecNpar] = n_in;
    r[kRecHist] = hist_row0;
  }
  if (l < nlev) {
    minmax[2 * l] = INFINITY;
    minmax[2 * l + 1] = -INFINITY;
  }
}

// Blocks of a launch whose threads stride over at most `cap` items (capped: blocks that
// find no item cost a dispatch each, and a level's count is unknown at launch).
inline unsigned strided_grid(int64_t cap, int per_block) {
  const int64_t g = (cap + per_block - 1) / per_block;
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>(g, 2048));
}

// Blocks of a launch that covers `cap` items, `per_block` per block (at most 2^31 - 1).
inline unsigned capacity_grid(int64_t cap, int per_block) {
  const int64_t g = (cap + per_block - 1) / per_block;
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>(g, 0x7fffffff));
}

}  // namespace
}  // namespace pyl_ep
