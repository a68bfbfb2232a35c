// Leaf brute fold (core/fold.py:108-171 brutefold_start).
//
// The segment-relative phase of sample i at trial frequency f does not depend on the
// segment, so a plan resolves it once in FP64 (psr_utils.py:49-56) and stores, per
// (frequency, bin), the ordered list of sample indices that land in that bin: u16 indices
// when a segment has fewer than 65536 samples, else u32. Layout [freq][bin][kpad]: every
// list of a frequency is padded to kpad = its longest list rounded up to 8 entries with the
// index seg_len, which points at a zero sample, so a lane reads its list with 16-byte loads
// and every lane of a wavefront runs the same number of entries.
//
// The fold kernel stages S segments of one series in LDS as interleaved (e, v) pairs (plus
// one zero pair per segment for the padding) and gives each wavefront one frequency at a
// time: lane b walks bin b's list in sample order and, for each entry, adds the (e, v) pair
// of every staged segment to that segment's FP32 accumulators. The list arrives 8 entries
// per load, four loads ahead of use, so the wavefront waits on LDS rather than on L2. The
// sums run in sample order per bin, which reproduces the reference's sequential
// `fold[...] += x` rounding exactly (a padding entry adds +0.0f, which leaves an FP32 sum
// that starts at +0.0f unchanged); one list load serves S segments, and one 8-byte LDS
// read brings both components.
//
// A launch folds a batch of series (up to kMaxBatch per launch, longest first or in any
// order) that share the plan: blocks are dealt over the series' segment groups, and each
// series' rows land at its own segment offset of the output, so many series of different
// lengths take one launch.
#include <algorithm>
#include <atomic>
#include <vector>

#include "pyl_common.hpp"

struct pyl_bf_plan {
  int64_t nfreq = 0;
  int64_t seg_len = 0;
  int32_t nbins = 0;
  int32_t idx16 = 0;          // 1: list holds u16 indices, else u32
  int32_t* cnt = nullptr;     // [nfreq][nbins]
  int64_t* off = nullptr;     // [nfreq] element offset of a frequency's list block
  int32_t* kpad = nullptr;    // [nfreq] padded list length of a frequency (multiple of 8)
  void* list = nullptr;       // [nfreq][nbins][kpad[f]] sample indices, padded with seg_len
  uint16_t* pmap = nullptr;   // [nfreq][seg_len] phase-bin map (plan-time scratch)
  int32_t* maxcnt = nullptr;  // [nfreq]
  int64_t list_len = 0;
  int64_t boundary = 0;       // phase bins that a +-1 ulp change of the FP64 phase flips
};

namespace {

constexpr int kBlock = 256;
constexpr int kLdsThreads = 512;
constexpr int kMaxBatch = 32;
constexpr int kLdsBudget = 80 * 1024;
constexpr int kLdsMax = 160 * 1024;

// Segments staged per block: 0 = the default (as many as fit kLdsBudget, at most 4).
std::atomic<int> g_bf_segs{0};

__global__ void bf_phase_kernel(const double* __restrict__ freqs, int64_t nfreq,
                                int64_t seg_len, int nbins, double tsamp, double t_ref,
                                uint16_t* __restrict__ pmap,
                                unsigned long long* __restrict__ boundary) {
  int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= nfreq * seg_len) return;
  int64_t f = gid / seg_len;
  int64_t i = gid - f * seg_len;
  // fold.py:150: proper_time = i * tsamp - t_ref (two roundings, no contraction).
  double pt = (double)i * tsamp;
  pt = pt - t_ref;
  const double x = pyl::phase_idx(pt, freqs[f], nbins, 0.0);
  const int b = pyl::shift_int(x, nbins);
  pmap[gid] = (uint16_t)b;
  // The count of bins an ulp of the FP64 phase would move (the reference's fastmath
  // evaluation may differ from this one by that much).
  if (pyl::shift_int(nextafter(x, -INFINITY), nbins) != b ||
      pyl::shift_int(nextafter(x, INFINITY), nbins) != b)
    atomicAdd(boundary, 1ull);
}

// One block per frequency: per-bin sample counts and their maximum.
__global__ void bf_count_kernel(const uint16_t* __restrict__ pmap, int64_t seg_len, int nbins,
                                int32_t* __restrict__ cnt, int32_t* __restrict__ maxcnt) {
  extern __shared__ int32_t s_cnt[];
  __shared__ int32_t s_max;
  const int64_t f = blockIdx.x;
  for (int b = threadIdx.x; b < nbins; b += blockDim.x) s_cnt[b] = 0;
  if (threadIdx.x == 0) s_max = 0;
  __syncthreads();
  const uint16_t* m = pmap + f * seg_len;
  for (int64_t i = threadIdx.x; i < seg_len; i += blockDim.x) atomicAdd(&s_cnt[m[i]], 1);
  __syncthreads();
  for (int b = threadIdx.x; b < nbins; b += blockDim.x) {
    cnt[f * nbins + b] = s_cnt[b];
    atomicMax(&s_max, s_cnt[b]);
  }
  __syncthreads();
  if (threadIdx.x == 0) maxcnt[f] = s_max;
}

// One block per frequency: thread b scans the map in sample order and appends the
// samples of bin b, so every list is in increasing sample order (stable), then pads it
// with seg_len.
template <typename Idx>
__global__ void bf_fill_kernel(const uint16_t* __restrict__ pmap, int64_t seg_len, int nbins,
                               const int64_t* __restrict__ off, const int32_t* __restrict__ kpad,
                               Idx* __restrict__ list) {
  constexpr int kChunk = 4096;
  __shared__ uint16_t s_map[kChunk];
  const int64_t f = blockIdx.x;
  const uint16_t* m = pmap + f * seg_len;
  const int64_t kp = kpad[f];
  for (int bg = 0; bg < nbins; bg += blockDim.x) {
    const int b = bg + threadIdx.x;
    Idx* L = list + off[f] + (int64_t)b * kp;
    int64_t k = 0;
    for (int64_t c0 = 0; c0 < seg_len; c0 += kChunk) {
      const int64_t clen = std::min<int64_t>(kChunk, seg_len - c0);
      __syncthreads();
      for (int64_t i = threadIdx.x; i < clen; i += blockDim.x) s_map[i] = m[c0 + i];
      __syncthreads();
      if (b < nbins) {
        for (int64_t i = 0; i < clen; ++i) {
          if (s_map[i] == b) {
            L[k] = (Idx)(c0 + i);
            ++k;
          }
        }
      }
    }
    if (b < nbins)
      for (; k < kp; ++k) L[k] = (Idx)seg_len;
  }
}

// The series of one launch, passed by value.
struct BatchArgs {
  int32_t nseries;
  int32_t spb;                       // segments per block (segment group size)
  const float* e[kMaxBatch];
  const float* v[kMaxBatch];
  int64_t nsamps[kMaxBatch];
  int64_t seg0[kMaxBatch];           // first output segment of the series
  int64_t grp0[kMaxBatch + 1];       // first segment group of the series (prefix sums)
};

__device__ __forceinline__ int find_series(const BatchArgs& a, int64_t grp) {
  int s = 0;
  while (s + 1 < a.nseries && grp >= a.grp0[s + 1]) ++s;
  return s;
}

__device__ __forceinline__ uint4 load_list8(const uint16_t* __restrict__ p) {
  return *reinterpret_cast<const uint4*>(p);
}

template <int S>
__device__ __forceinline__ void add_entry(const float2* __restrict__ s_ev, int64_t stride,
                                          uint32_t idx, float (&ae)[S], float (&av)[S]) {
#pragma unroll
  for (int j = 0; j < S; ++j) {
    const float2 x = s_ev[j * stride + idx];
    ae[j] += x.x;
    av[j] += x.y;
  }
}

template <int S>
__device__ __forceinline__ void add_chunk(const float2* __restrict__ s_ev, int64_t stride,
                                          const uint4& c, float (&ae)[S], float (&av)[S]) {
  add_entry<S>(s_ev, stride, c.x & 0xffffu, ae, av);
  add_entry<S>(s_ev, stride, c.x >> 16, ae, av);
  add_entry<S>(s_ev, stride, c.y & 0xffffu, ae, av);
  add_entry<S>(s_ev, stride, c.y >> 16, ae, av);
  add_entry<S>(s_ev, stride, c.z & 0xffffu, ae, av);
  add_entry<S>(s_ev, stride, c.z >> 16, ae, av);
  add_entry<S>(s_ev, stride, c.w & 0xffffu, ae, av);
  add_entry<S>(s_ev, stride, c.w >> 16, ae, av);
}

// LDS path (u16 lists): S segments of one series per block, kLdsThreads threads.
template <int S, typename T>
__global__ void __launch_bounds__(kLdsThreads)
bf_fold_lds_kernel(BatchArgs a, int64_t seg_len, int64_t nfreq, int nbins, int fpb,
                   const int64_t* __restrict__ off, const int32_t* __restrict__ kpad,
                   const uint16_t* __restrict__ list, T* __restrict__ fold) {
  extern __shared__ float2 s_ev[];  // [S][seg_len + 1] (e, v); entry seg_len is zero
  const int64_t stride = seg_len + 1;
  const int64_t grp = blockIdx.x;
  const int sr = find_series(a, grp);
  const int64_t lseg0 = (grp - a.grp0[sr]) * S;    // first segment of the group in the series
  const int64_t nseg_s = (a.nsamps[sr] + seg_len - 1) / seg_len;
  const int nvalid = (int)std::min<int64_t>(S, nseg_s - lseg0);
  const float* __restrict__ se = a.e[sr];
  const float* __restrict__ sv = a.v[sr];
  const int64_t nsamps = a.nsamps[sr];
  for (int j = 0; j < S; ++j) {
    const int64_t start = (lseg0 + j) * seg_len;
    for (int64_t i = threadIdx.x; i < stride; i += blockDim.x) {
      // Samples past the end (and the padding entry) are +0.0f.
      const int64_t g = start + i;
      const bool in = j < nvalid && i < seg_len && g < nsamps;
      s_ev[j * stride + i] = in ? make_float2(se[g], sv[g]) : make_float2(0.0f, 0.0f);
    }
  }
  __syncthreads();
  const int64_t f0 = (int64_t)blockIdx.y * fpb;
  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const int nwaves = blockDim.x >> 6;
  const int nbg = (nbins + 63) >> 6;
  const uint4 zero = make_uint4(0, 0, 0, 0);
  for (int w = wave; w < fpb * nbg; w += nwaves) {
    const int64_t f = f0 + w / nbg;
    if (f >= nfreq) break;
    const int bb = (w % nbg) * 64 + lane;
    const bool live = bb < nbins;
    const int b = live ? bb : nbins - 1;  // idle lanes read a valid list and store nothing
    const int nch = kpad[f] >> 3;         // wave-uniform
    const uint16_t* L = list + off[f] + (int64_t)b * kpad[f];
    float ae[S], av[S];
#pragma unroll
    for (int j = 0; j < S; ++j) ae[j] = av[j] = 0.0f;
    uint4 c0 = nch > 0 ? load_list8(L) : zero;
    uint4 c1 = nch > 1 ? load_list8(L + 8) : zero;
    uint4 c2 = nch > 2 ? load_list8(L + 16) : zero;
    uint4 c3 = nch > 3 ? load_list8(L + 24) : zero;
    for (int ch = 0; ch < nch; ++ch) {
      const uint4 cur = c0;
      c0 = c1;
      c1 = c2;
      c2 = c3;
      c3 = ch + 4 < nch ? load_list8(L + 8 * (ch + 4)) : zero;
      add_chunk<S>(s_ev, stride, cur, ae, av);
    }
    if (!live) continue;
#pragma unroll
    for (int j = 0; j < S; ++j) {
      if (j < nvalid) {
        T* o = fold + (((a.seg0[sr] + lseg0 + j) * nfreq + f) * 2) * nbins;
        o[b] = pyl::Store<T>::store(ae[j]);
        o[nbins + b] = pyl::Store<T>::store(av[j]);
      }
    }
  }
}

// Segments too long to stage in LDS: one segment per block, samples read from HBM/L2.
template <typename T, typename Idx>
__global__ void __launch_bounds__(kBlock)
bf_fold_global_kernel(BatchArgs a, int64_t seg_len, int64_t nfreq, int nbins, int fpb,
                      const int32_t* __restrict__ cnt, const int64_t* __restrict__ off,
                      const int32_t* __restrict__ kpad, const Idx* __restrict__ list,
                      T* __restrict__ fold) {
  const int64_t grp = blockIdx.x;
  const int sr = find_series(a, grp);
  const int64_t lseg = grp - a.grp0[sr];
  const int64_t start = lseg * seg_len;
  const int64_t len = std::min<int64_t>(seg_len, a.nsamps[sr] - start);
  const float* __restrict__ se = a.e[sr] + start;
  const float* __restrict__ sv = a.v[sr] + start;
  const int64_t f0 = (int64_t)blockIdx.y * fpb;
  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const int nwaves = blockDim.x >> 6;
  const int nbg = (nbins + 63) >> 6;
  for (int w = wave; w < fpb * nbg; w += nwaves) {
    const int64_t f = f0 + w / nbg;
    if (f >= nfreq) break;
    const int b = (w % nbg) * 64 + lane;
    if (b >= nbins) continue;
    const int c = cnt[f * nbins + b];
    const Idx* L = list + off[f] + (int64_t)b * kpad[f];
    float ae = 0.0f, av = 0.0f;
    for (int k = 0; k < c; ++k) {
      const int64_t idx = L[k];
      if (idx >= len) break;
      ae += se[idx];
      av += sv[idx];
    }
    T* o = fold + (((a.seg0[sr] + lseg) * nfreq + f) * 2) * nbins;
    o[b] = pyl::Store<T>::store(ae);
    o[nbins + b] = pyl::Store<T>::store(av);
  }
}

// The default stages as many segments as fit kLdsBudget (at most 4); a shape set with
// pyl_brutefold_set_segments_per_block may use up to the 160 KB a gfx950 workgroup holds.
int segs_per_block(int64_t seg_len) {
  if (seg_len >= 65536) return 0;  // u32 lists: the global path
  const int64_t per = (seg_len + 1) * (int64_t)sizeof(float2);
  int s = g_bf_segs.load(std::memory_order_relaxed);
  const int64_t budget = s > 0 ? kLdsMax : kLdsBudget;
  if (s <= 0) s = 4;
  while (s > 1 && s * per > budget) s >>= 1;
  return s * per <= budget ? s : 0;  // 0: the global path
}

}  // namespace

extern "C" int pyl_brutefold_plan_create(const double* freqs, int64_t nfreq, int64_t seg_len,
                                         int32_t nbins, double tsamp, double t_ref,
                                         void* stream, pyl_bf_plan** plan_out) {
  PYL_REQUIRE(plan_out != nullptr, "plan_out must not be NULL");
  PYL_REQUIRE(freqs != nullptr && nfreq > 0, "freq_arr must not be empty");
  PYL_REQUIRE(seg_len > 0 && seg_len <= (int64_t)1 << 31, "invalid segment_len %lld",
              (long long)seg_len);
  PYL_REQUIRE(nbins > 0 && nbins <= 65535, "Number of bins must be positive (<= 65535).");
  PYL_REQUIRE(tsamp > 0, "tsamp must be positive");
  for (int64_t f = 0; f < nfreq; ++f)
    PYL_REQUIRE(freqs[f] > 0, "Frequency must be positive.");
  hipStream_t s = pyl::as_stream(stream);
  auto* p = new pyl_bf_plan();
  p->nfreq = nfreq;
  p->seg_len = seg_len;
  p->nbins = nbins;
  p->idx16 = seg_len < 65536 ? 1 : 0;
  double* d_freqs = nullptr;
  unsigned long long* d_bnd = nullptr;
  auto cleanup = [&]() {
    if (d_freqs) (void)hipFree(d_freqs);
    if (d_bnd) (void)hipFree(d_bnd);
    (void)pyl_brutefold_plan_destroy(p);
  };
#define PLAN_HIP(call)                                                               \
  do {                                                                               \
    hipError_t e_ = (call);                                                          \
    if (e_ != hipSuccess) {                                                          \
      cleanup();                                                                     \
      return pyl::fail(PYL_ERR_HIP, "%s failed: %s", #call, hipGetErrorString(e_));   \
    }                                                                                \
  } while (0)
  PLAN_HIP(hipMalloc(&d_freqs, sizeof(double) * nfreq));
  PLAN_HIP(hipMalloc(&d_bnd, sizeof(unsigned long long)));
  PLAN_HIP(hipMalloc(&p->pmap, sizeof(uint16_t) * nfreq * seg_len));
  PLAN_HIP(hipMalloc(&p->cnt, sizeof(int32_t) * nfreq * nbins));
  PLAN_HIP(hipMalloc(&p->maxcnt, sizeof(int32_t) * nfreq));
  PLAN_HIP(hipMalloc(&p->off, sizeof(int64_t) * nfreq));
  PLAN_HIP(hipMalloc(&p->kpad, sizeof(int32_t) * nfreq));
  PLAN_HIP(hipMemcpyAsync(d_freqs, freqs, sizeof(double) * nfreq, hipMemcpyHostToDevice, s));
  PLAN_HIP(hipMemsetAsync(d_bnd, 0, sizeof(unsigned long long), s));
  {
    const int64_t total = nfreq * seg_len;
    bf_phase_kernel<<<(unsigned)pyl::ceil_div(total, kBlock), kBlock, 0, s>>>(
        d_freqs, nfreq, seg_len, nbins, tsamp, t_ref, p->pmap, d_bnd);
    PLAN_HIP(hipGetLastError());
    bf_count_kernel<<<(unsigned)nfreq, kBlock, sizeof(int32_t) * nbins, s>>>(
        p->pmap, seg_len, nbins, p->cnt, p->maxcnt);
    PLAN_HIP(hipGetLastError());
  }
  std::vector<int32_t> h_max(nfreq);
  unsigned long long h_bnd = 0;
  PLAN_HIP(hipMemcpyAsync(h_max.data(), p->maxcnt, sizeof(int32_t) * nfreq,
                          hipMemcpyDeviceToHost, s));
  PLAN_HIP(hipMemcpyAsync(&h_bnd, d_bnd, sizeof(h_bnd), hipMemcpyDeviceToHost, s));
  PLAN_HIP(hipStreamSynchronize(s));
  p->boundary = (int64_t)h_bnd;
  std::vector<int64_t> h_off(nfreq);
  std::vector<int32_t> h_kpad(nfreq);
  int64_t acc = 0;
  for (int64_t f = 0; f < nfreq; ++f) {
    h_off[f] = acc;
    h_kpad[f] = (h_max[f] + 7) & ~7;
    acc += (int64_t)h_kpad[f] * nbins;
  }
  p->list_len = acc;
  const size_t isz = p->idx16 ? sizeof(uint16_t) : sizeof(uint32_t);
  PLAN_HIP(hipMalloc(&p->list, isz * std::max<int64_t>(acc, 1)));
  PLAN_HIP(hipMemcpyAsync(p->off, h_off.data(), sizeof(int64_t) * nfreq, hipMemcpyHostToDevice,
                          s));
  PLAN_HIP(hipMemcpyAsync(p->kpad, h_kpad.data(), sizeof(int32_t) * nfreq,
                          hipMemcpyHostToDevice, s));
  if (p->idx16)
    bf_fill_kernel<uint16_t><<<(unsigned)nfreq, kBlock, 0, s>>>(
        p->pmap, seg_len, nbins, p->off, p->kpad, static_cast<uint16_t*>(p->list));
  else
    bf_fill_kernel<uint32_t><<<(unsigned)nfreq, kBlock, 0, s>>>(
        p->pmap, seg_len, nbins, p->off, p->kpad, static_cast<uint32_t*>(p->list));
  PLAN_HIP(hipGetLastError());
  PLAN_HIP(hipStreamSynchronize(s));
  PLAN_HIP(hipFree(d_freqs));
  d_freqs = nullptr;
  PLAN_HIP(hipFree(d_bnd));
  d_bnd = nullptr;
  PLAN_HIP(hipFree(p->pmap));
  p->pmap = nullptr;
#undef PLAN_HIP
  *plan_out = p;
  return PYL_OK;
}

extern "C" int pyl_brutefold_plan_destroy(pyl_bf_plan* p) {
  if (!p) return PYL_OK;
  if (p->cnt) (void)hipFree(p->cnt);
  if (p->off) (void)hipFree(p->off);
  if (p->kpad) (void)hipFree(p->kpad);
  if (p->list) (void)hipFree(p->list);
  if (p->pmap) (void)hipFree(p->pmap);
  if (p->maxcnt) (void)hipFree(p->maxcnt);
  delete p;
  return PYL_OK;
}

extern "C" int pyl_brutefold_plan_boundary_count(const pyl_bf_plan* p, int64_t* count_out) {
  PYL_REQUIRE(p != nullptr && count_out != nullptr, "null pointer argument");
  *count_out = p->boundary;
  return PYL_OK;
}

extern "C" int pyl_brutefold_set_segments_per_block(int32_t segs) {
  PYL_REQUIRE(segs == 0 || segs == 1 || segs == 2 || segs == 4,
              "segments per block must be 0 (default), 1, 2 or 4 (got %d)", segs);
  g_bf_segs.store(segs, std::memory_order_relaxed);
  return PYL_OK;
}

template <typename T, typename Idx>
static int launch_batch(const pyl_bf_plan* p, BatchArgs& a, T* fold, hipStream_t s) {
  const int spb = segs_per_block(p->seg_len);
  a.spb = spb > 0 ? spb : 1;
  int64_t g = 0;
  for (int i = 0; i < a.nseries; ++i) {
    a.grp0[i] = g;
    g += pyl::ceil_div(pyl::ceil_div(a.nsamps[i], p->seg_len), a.spb);
  }
  a.grp0[a.nseries] = g;
  PYL_REQUIRE(g <= 0x7fffffff, "too many segments");
  // Frequencies per block: enough blocks to fill 256 CUs several times over, while each
  // block amortises one LDS staging of its segments over several frequencies.
  int fpb = (int)std::min<int64_t>(p->nfreq, 128);
  const int64_t target_blocks = 2048;
  while (fpb > 8 && g * pyl::ceil_div(p->nfreq, fpb) < target_blocks) fpb >>= 1;
  dim3 grid((unsigned)g, (unsigned)pyl::ceil_div(p->nfreq, fpb));
  if (spb > 0) {
    const uint16_t* list = static_cast<const uint16_t*>(p->list);
    const size_t lds = sizeof(float2) * (size_t)spb * (size_t)(p->seg_len + 1);
    switch (spb) {
      case 4:
        bf_fold_lds_kernel<4, T><<<grid, kLdsThreads, lds, s>>>(
            a, p->seg_len, p->nfreq, p->nbins, fpb, p->off, p->kpad, list, fold);
        break;
      case 2:
        bf_fold_lds_kernel<2, T><<<grid, kLdsThreads, lds, s>>>(
            a, p->seg_len, p->nfreq, p->nbins, fpb, p->off, p->kpad, list, fold);
        break;
      default:
        bf_fold_lds_kernel<1, T><<<grid, kLdsThreads, lds, s>>>(
            a, p->seg_len, p->nfreq, p->nbins, fpb, p->off, p->kpad, list, fold);
        break;
    }
    PYL_CHECK_LAUNCH("bf_fold_lds_kernel");
  } else {
    bf_fold_global_kernel<T, Idx><<<grid, kBlock, 0, s>>>(a, p->seg_len, p->nfreq, p->nbins, fpb,
                                                         p->cnt, p->off, p->kpad,
                                                         static_cast<const Idx*>(p->list), fold);
    PYL_CHECK_LAUNCH("bf_fold_global_kernel");
  }
  return PYL_OK;
}

template <typename T>
static int brutefold_batch(const pyl_bf_plan* p, const float* const* ts_e,
                           const float* const* ts_v, const int64_t* nsamps, int32_t nseries,
                           T* fold, void* stream) {
  PYL_REQUIRE(p != nullptr, "plan must not be NULL");
  PYL_REQUIRE(ts_e && ts_v && nsamps && fold, "null pointer argument");
  PYL_REQUIRE(nseries > 0, "empty batch");
  hipStream_t s = pyl::as_stream(stream);
  int64_t seg = 0;
  for (int32_t i0 = 0; i0 < nseries; i0 += kMaxBatch) {
    BatchArgs a{};
    a.nseries = std::min<int32_t>(kMaxBatch, nseries - i0);
    for (int i = 0; i < a.nseries; ++i) {
      PYL_REQUIRE(ts_e[i0 + i] && ts_v[i0 + i], "null device pointer (series %d)", i0 + i);
      PYL_REQUIRE(nsamps[i0 + i] > 0, "empty time series (series %d)", i0 + i);
      a.e[i] = ts_e[i0 + i];
      a.v[i] = ts_v[i0 + i];
      a.nsamps[i] = nsamps[i0 + i];
      a.seg0[i] = seg;
      seg += pyl::ceil_div(nsamps[i0 + i], p->seg_len);
    }
    const int rc = p->idx16 ? launch_batch<T, uint16_t>(p, a, fold, s)
                            : launch_batch<T, uint32_t>(p, a, fold, s);
    if (rc != PYL_OK) return rc;
  }
  return PYL_OK;
}

extern "C" int pyl_brutefold_execute_batch(const pyl_bf_plan* p, const float* const* ts_e,
                                           const float* const* ts_v, const int64_t* nsamps,
                                           int32_t nseries, float* fold, void* stream) {
  return brutefold_batch<float>(p, ts_e, ts_v, nsamps, nseries, fold, stream);
}

extern "C" int pyl_brutefold_execute_batch_bf16(const pyl_bf_plan* p, const float* const* ts_e,
                                                const float* const* ts_v, const int64_t* nsamps,
                                                int32_t nseries, uint16_t* fold, void* stream) {
  return brutefold_batch<uint16_t>(p, ts_e, ts_v, nsamps, nseries, fold, stream);
}

extern "C" int pyl_brutefold_execute(const pyl_bf_plan* p, const float* ts_e,
                                     const float* ts_v, int64_t nsamps, float* fold,
                                     void* stream) {
  return brutefold_batch<float>(p, &ts_e, &ts_v, &nsamps, 1, fold, stream);
}

extern "C" int pyl_brutefold_execute_bf16(const pyl_bf_plan* p, const float* ts_e,
                                          const float* ts_v, int64_t nsamps, uint16_t* fold,
                                          void* stream) {
  return brutefold_batch<uint16_t>(p, &ts_e, &ts_v, &nsamps, 1, fold, stream);
}
