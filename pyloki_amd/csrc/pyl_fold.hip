// Leaf brute fold (core/fold.py:108-171 brutefold_start).
//
// The segment-relative phase of sample i at trial frequency f does not depend on the
// segment, so a plan resolves it once in FP64 (psr_utils.py:49-56) and stores, per
// (frequency, bin), the ordered list of sample indices that land in that bin. The lists
// are interleaved ([freq][k][bin]) so that lane b of a wavefront reads entry k of bin b
// and the 64 lanes read 64 consecutive words.
//
// The fold kernel stages one segment of ts_e/ts_v in LDS and gives each wavefront one
// (segment, frequency) pair: lane b walks bin b's list in sample order and accumulates in
// FP32, which reproduces the reference's sequential `fold[...] += x` rounding exactly.
#include <algorithm>
#include <vector>

#include "pyl_common.hpp"

struct pyl_bf_plan {
  int64_t nfreq = 0;
  int64_t seg_len = 0;
  int32_t nbins = 0;
  int32_t* cnt = nullptr;     // [nfreq][nbins]
  int64_t* off = nullptr;     // [nfreq] element offset of a frequency's list block
  uint32_t* list = nullptr;   // [sum_f maxcnt[f] * nbins], interleaved [k][bin]
  uint16_t* pmap = nullptr;   // [nfreq][seg_len] phase-bin map (plan-time scratch)
  int32_t* maxcnt = nullptr;  // [nfreq]
  int64_t list_len = 0;
};

namespace {

constexpr int kBlock = 256;

__global__ void bf_phase_kernel(const double* __restrict__ freqs, int64_t nfreq,
                                int64_t seg_len, int nbins, double tsamp, double t_ref,
                                uint16_t* __restrict__ pmap) {
  int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= nfreq * seg_len) return;
  int64_t f = gid / seg_len;
  int64_t i = gid - f * seg_len;
  // fold.py:150: proper_time = i * tsamp - t_ref (two roundings, no contraction).
  double pt = (double)i * tsamp;
  pt = pt - t_ref;
  pmap[gid] = (uint16_t)pyl::shift_int(pyl::phase_idx(pt, freqs[f], nbins, 0.0), nbins);
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
// samples of bin b, so every list is in increasing sample order (stable).
__global__ void bf_fill_kernel(const uint16_t* __restrict__ pmap, int64_t seg_len, int nbins,
                               const int64_t* __restrict__ off, uint32_t* __restrict__ list) {
  constexpr int kChunk = 4096;
  __shared__ uint16_t s_map[kChunk];
  const int64_t f = blockIdx.x;
  const uint16_t* m = pmap + f * seg_len;
  uint32_t* L = list + off[f];
  for (int bg = 0; bg < nbins; bg += blockDim.x) {
    const int b = bg + threadIdx.x;
    int64_t k = 0;
    for (int64_t c0 = 0; c0 < seg_len; c0 += kChunk) {
      const int64_t clen = std::min<int64_t>(kChunk, seg_len - c0);
      __syncthreads();
      for (int64_t i = threadIdx.x; i < clen; i += blockDim.x) s_map[i] = m[c0 + i];
      __syncthreads();
      if (b < nbins) {
        for (int64_t i = 0; i < clen; ++i) {
          if (s_map[i] == b) {
            L[k * nbins + b] = (uint32_t)(c0 + i);
            ++k;
          }
        }
      }
    }
  }
}

template <bool kUseLds, typename T>
__global__ void __launch_bounds__(kBlock)
bf_fold_kernel(const float* __restrict__ ts_e, const float* __restrict__ ts_v, int64_t nsamps,
               int64_t seg_len, int64_t nfreq, int nbins, int fpb,
               const int32_t* __restrict__ cnt, const int64_t* __restrict__ off,
               const uint32_t* __restrict__ list, T* __restrict__ fold) {
  extern __shared__ float s_seg[];  // [2][seg_len] when kUseLds
  const int64_t seg = blockIdx.x;
  const int64_t f0 = (int64_t)blockIdx.y * fpb;
  const int64_t start = seg * seg_len;
  const int64_t len = std::min<int64_t>(seg_len, nsamps - start);
  const float* se = ts_e + start;
  const float* sv = ts_v + start;
  if (kUseLds) {
    for (int64_t i = threadIdx.x; i < seg_len; i += blockDim.x) {
      // Padding samples past the end are +0.0f: adding +0.0f leaves an FP32 sum unchanged.
      s_seg[i] = i < len ? se[i] : 0.0f;
      s_seg[seg_len + i] = i < len ? sv[i] : 0.0f;
    }
    __syncthreads();
    se = s_seg;
    sv = s_seg + seg_len;
  }
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
    const uint32_t* L = list + off[f] + b;
    float ae = 0.0f, av = 0.0f;
    for (int k = 0; k < c; ++k) {
      const uint32_t idx = L[(int64_t)k * nbins];
      if (!kUseLds && (int64_t)idx >= len) break;
      ae += se[idx];
      av += sv[idx];
    }
    T* o = fold + ((seg * nfreq + f) * 2) * nbins;
    o[b] = pyl::Store<T>::store(ae);
    o[nbins + b] = pyl::Store<T>::store(av);
  }
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
  double* d_freqs = nullptr;
  auto cleanup = [&]() {
    if (d_freqs) (void)hipFree(d_freqs);
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
  PLAN_HIP(hipMalloc(&p->pmap, sizeof(uint16_t) * nfreq * seg_len));
  PLAN_HIP(hipMalloc(&p->cnt, sizeof(int32_t) * nfreq * nbins));
  PLAN_HIP(hipMalloc(&p->maxcnt, sizeof(int32_t) * nfreq));
  PLAN_HIP(hipMalloc(&p->off, sizeof(int64_t) * nfreq));
  PLAN_HIP(hipMemcpyAsync(d_freqs, freqs, sizeof(double) * nfreq, hipMemcpyHostToDevice, s));
  {
    const int64_t total = nfreq * seg_len;
    bf_phase_kernel<<<(unsigned)pyl::ceil_div(total, kBlock), kBlock, 0, s>>>(
        d_freqs, nfreq, seg_len, nbins, tsamp, t_ref, p->pmap);
    PLAN_HIP(hipGetLastError());
    bf_count_kernel<<<(unsigned)nfreq, kBlock, sizeof(int32_t) * nbins, s>>>(
        p->pmap, seg_len, nbins, p->cnt, p->maxcnt);
    PLAN_HIP(hipGetLastError());
  }
  std::vector<int32_t> h_max(nfreq);
  PLAN_HIP(hipMemcpyAsync(h_max.data(), p->maxcnt, sizeof(int32_t) * nfreq,
                          hipMemcpyDeviceToHost, s));
  PLAN_HIP(hipStreamSynchronize(s));
  std::vector<int64_t> h_off(nfreq);
  int64_t acc = 0;
  for (int64_t f = 0; f < nfreq; ++f) {
    h_off[f] = acc;
    acc += (int64_t)h_max[f] * nbins;
  }
  p->list_len = acc;
  PLAN_HIP(hipMalloc(&p->list, sizeof(uint32_t) * std::max<int64_t>(acc, 1)));
  PLAN_HIP(hipMemcpyAsync(p->off, h_off.data(), sizeof(int64_t) * nfreq, hipMemcpyHostToDevice,
                          s));
  bf_fill_kernel<<<(unsigned)nfreq, kBlock, 0, s>>>(p->pmap, seg_len, nbins, p->off, p->list);
  PLAN_HIP(hipGetLastError());
  PLAN_HIP(hipStreamSynchronize(s));
  PLAN_HIP(hipFree(d_freqs));
  d_freqs = nullptr;
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
  if (p->list) (void)hipFree(p->list);
  if (p->pmap) (void)hipFree(p->pmap);
  if (p->maxcnt) (void)hipFree(p->maxcnt);
  delete p;
  return PYL_OK;
}

template <typename T>
static int brutefold_execute(const pyl_bf_plan* p, const float* ts_e, const float* ts_v,
                             int64_t nsamps, T* fold, void* stream) {
  PYL_REQUIRE(p != nullptr, "plan must not be NULL");
  PYL_REQUIRE(ts_e && ts_v && fold, "null device pointer");
  PYL_REQUIRE(nsamps > 0, "empty time series");
  hipStream_t s = pyl::as_stream(stream);
  const int64_t nseg = pyl::ceil_div(nsamps, p->seg_len);
  PYL_REQUIRE(nseg <= 0x7fffffff, "too many segments");
  const size_t lds = sizeof(float) * 2 * (size_t)p->seg_len;
  const bool use_lds = lds <= 64 * 1024;
  // Frequencies per block: enough blocks to fill 256 CUs several times over, while each
  // block amortises one LDS staging of its segment over several frequencies.
  int fpb = (int)std::min<int64_t>(p->nfreq, 32);
  const int64_t target_blocks = 2048;
  while (fpb > 4 && nseg * pyl::ceil_div(p->nfreq, fpb) < target_blocks) fpb >>= 1;
  dim3 grid((unsigned)nseg, (unsigned)pyl::ceil_div(p->nfreq, fpb));
  if (use_lds) {
    bf_fold_kernel<true, T><<<grid, kBlock, lds, s>>>(ts_e, ts_v, nsamps, p->seg_len, p->nfreq,
                                                      p->nbins, fpb, p->cnt, p->off, p->list,
                                                      fold);
  } else {
    bf_fold_kernel<false, T><<<grid, kBlock, 0, s>>>(ts_e, ts_v, nsamps, p->seg_len, p->nfreq,
                                                     p->nbins, fpb, p->cnt, p->off, p->list,
                                                     fold);
  }
  PYL_CHECK_LAUNCH("bf_fold_kernel");
  return PYL_OK;
}

extern "C" int pyl_brutefold_execute(const pyl_bf_plan* p, const float* ts_e,
                                     const float* ts_v, int64_t nsamps, float* fold,
                                     void* stream) {
  return brutefold_execute<float>(p, ts_e, ts_v, nsamps, fold, stream);
}

extern "C" int pyl_brutefold_execute_bf16(const pyl_bf_plan* p, const float* ts_e,
                                          const float* ts_v, int64_t nsamps, uint16_t* fold,
                                          void* stream) {
  return brutefold_execute<uint16_t>(p, ts_e, ts_v, nsamps, fold, stream);
}
