// Complex (Fourier-domain) FFA: leaf fold, level shift-add and scoring
// (use_fourier=True: fold.py:395-481,572-614; common.py:344-372; ffa.py:400-411).
//
// Leaf fold. X[s, f, c, m] = sum_t x_c[s, t] e^{-2 pi i f m t tsamp}, m = 0..B/2, is a
// dense contraction: the segments (rows of length seg_len) times a basis of n_comb =
// nfreq * (B/2 + 1) complex exponentials. The basis is data-independent, so the plan
// builds it once in HBM (FP32 cos / sin of the FP64 angle f32(f m) * t * f32(-2 pi tsamp),
// as the reference does). Each execute is one launch of cfold_mfma_kernel, an FP32 product
// on the matrix cores (v_mfma_f32_32x32x2f32): a block takes 64 segments and 64 comb
// entries; its A tile holds the e and the v rows of the segments (so both components share
// every basis tile) and its B tile the cos and the sin rows of the comb entries (so a lane
// ends with the real and the imaginary part of the same output). The epilogue applies the
// e^{+2 pi i f m t_ref} correction in complex64 and writes the fold layout
// [nseg][nfreq][2][B/2 + 1]; nothing is staged in HBM between the product and the fold.
//
// Level shift-add. out[k] = tail[k] phi_t^k + head[k] phi_h^k with the reference's
// complex64 recurrence phi^{k+1} = phi^k * delta, delta = (cos, sin)(f32(-2 pi f32(shift)/B)).
// A block owns 64 output rows; 128 threads first build the 64 x 2 phasor tables in LDS
// (one recurrence per thread), then lane = (row, component, harmonic) forms the outputs
// with coalesced complex loads.
//
// Scoring. irfft to B real bins (x[n] = (X_0 + (-1)^n X_{B/2} + 2 sum Re(X_k e^{2 pi i k n/B}))
// / B, FP32 with an FP32 twiddle table), e / sqrt(v), then the boxcar scorer of
// pyl_score.hpp (one thread per profile).
#include <algorithm>
#include <cmath>
#include <vector>

#include "pyl_score.hpp"

struct pyl_cfold_plan {
  int64_t nfreq = 0;
  int64_t seg_len = 0;
  int32_t nbins = 0;
  int32_t nb_f = 0;
  int64_t n_comb = 0;
  float* basis = nullptr;   // [2][n_comb][seg_len]: cos rows, then sin rows
  float2* corr = nullptr;   // [n_comb] complex64 t_ref correction
};

namespace {

constexpr int kRowsPerBlock = 64;

__global__ void cfold_basis_kernel(const float* __restrict__ target, int64_t n_comb,
                                   int64_t seg_len, float time_factor, float* __restrict__ basis) {
  const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= n_comb * seg_len) return;
  const int64_t i = gid / seg_len;
  const int64_t t = gid - i * seg_len;
  // fold.py:449-456: time_val = t * f32(-2 pi tsamp) (FP64), ang = f32(f m) * time_val.
  const double ang = (double)target[i] * ((double)t * (double)time_factor);
  basis[gid] = (float)cos(ang);
  basis[n_comb * seg_len + gid] = (float)sin(ang);
}

using pyl::cmul;

typedef float cf_f32x16 __attribute__((ext_vector_type(16)));

constexpr int kCfSeg = 64;            // segments per block (A rows: their e, then v rows)
constexpr int kCfComb = 64;           // comb entries per block (B rows: cos, then sin)
constexpr int kCfK = 16;              // K per LDS stage
constexpr int kCfLd = kCfK + 1;       // padded LDS row (floats): conflict-free column reads
constexpr int kCfThreads = 256;

// One stage of the A and B tiles (128 rows x 16 K each) as 2 + 2 float4 per thread, zero
// beyond the segments, the comb entries and seg_len. kVec: rows start 16-byte aligned.
template <bool kVec>
__device__ inline void cf_load(const float* __restrict__ xe, const float* __restrict__ xv,
                               const float* __restrict__ basis, int64_t nseg, int64_t seg_len,
                               int64_t n_comb, int64_t s0, int64_t i0, int64_t k0, int tid,
                               float4 ra[2], float4 rb[2]) {
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    const int idx = tid + kCfThreads * j;  // 0..511
    const int r = idx >> 2, q = idx & 3;   // tile row 0..127, K quad 0..3
    const int64_t k = k0 + 4 * q;
    const int64_t sg = s0 + (r & 63);
    const int64_t ci = i0 + (r & 63);
    const float* a = (r < 64 ? xe : xv) + sg * seg_len + k;
    const float* b = basis + ((r < 64 ? 0 : n_comb) + ci) * seg_len + k;
    const bool av = sg < nseg, bv = ci < n_comb;
    if (kVec) {
      const bool kin = k < seg_len;
      ra[j] = av && kin ? *reinterpret_cast<const float4*>(a) : make_float4(0.f, 0.f, 0.f, 0.f);
      rb[j] = bv && kin ? *reinterpret_cast<const float4*>(b) : make_float4(0.f, 0.f, 0.f, 0.f);
    } else {
      float va[4], vb[4];
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const bool kin = k + c < seg_len;
        va[c] = av && kin ? a[c] : 0.0f;
        vb[c] = bv && kin ? b[c] : 0.0f;
      }
      ra[j] = make_float4(va[0], va[1], va[2], va[3]);
      rb[j] = make_float4(vb[0], vb[1], vb[2], vb[3]);
    }
  }
}

// fold[s][f][c][m] = (sum_t x_c[s][t] cos_t, sum_t x_c[s][t] sin_t) * corr for the block's
// 64 segments x 64 comb entries; wavefront w takes component w & 1 and comb half w >> 1,
// i.e. two 32 x 32 tiles of segments times the cos and the sin tile of 32 comb entries.
template <bool kVec>
__global__ void __launch_bounds__(kCfThreads)
cfold_mfma_kernel(const float* __restrict__ xe, const float* __restrict__ xv,
                  const float* __restrict__ basis, const float2* __restrict__ corr,
                  int64_t nseg, int64_t seg_len, int64_t n_comb, int nb_f, int m_tiles,
                  int nblk, float2* __restrict__ fold) {
  __shared__ float As[2 * kCfSeg * kCfLd];
  __shared__ float Bs[2 * kCfComb * kCfLd];
  // The grid is padded to a multiple of 8 and block p runs on XCD p % 8: logical blocks
  // x * per .. x * per + per - 1 go to XCD x, so the blocks that share a basis tile (M
  // fastest) share that XCD's L2.
  const int per = (int)(gridDim.x / 8);
  const int L = (int)(blockIdx.x % 8) * per + (int)(blockIdx.x / 8);
  if (L >= nblk) return;
  const int64_t s0 = (int64_t)(L % m_tiles) * kCfSeg;
  const int64_t i0 = (int64_t)(L / m_tiles) * kCfComb;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  const int row = lane & 31, half = lane >> 5;
  cf_f32x16 acc_c[2], acc_s[2];
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    acc_c[0][i] = 0.0f;
    acc_c[1][i] = 0.0f;
    acc_s[0][i] = 0.0f;
    acc_s[1][i] = 0.0f;
  }
  float4 ra[2], rb[2];
  cf_load<kVec>(xe, xv, basis, nseg, seg_len, n_comb, s0, i0, 0, tid, ra, rb);
  const float* a0p = As + (wm * 64 + row) * kCfLd + half;
  const float* a1p = a0p + 32 * kCfLd;
  const float* bcp = Bs + (wn * 32 + row) * kCfLd + half;
  const float* bsp = bcp + 64 * kCfLd;
  for (int64_t k0 = 0; k0 < seg_len; k0 += kCfK) {
    __syncthreads();  // the previous stage's reads are done
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int idx = tid + kCfThreads * j;
      const int r = idx >> 2, q = idx & 3;
      float* da = As + r * kCfLd + 4 * q;
      float* db = Bs + r * kCfLd + 4 * q;
      da[0] = ra[j].x; da[1] = ra[j].y; da[2] = ra[j].z; da[3] = ra[j].w;
      db[0] = rb[j].x; db[1] = rb[j].y; db[2] = rb[j].z; db[3] = rb[j].w;
    }
    __syncthreads();
    if (k0 + kCfK < seg_len)  // the next stage's loads overlap this stage's products
      cf_load<kVec>(xe, xv, basis, nseg, seg_len, n_comb, s0, i0, k0 + kCfK, tid, ra, rb);
#pragma unroll
    for (int kk = 0; kk < kCfK; kk += 2) {
      const float a0 = a0p[kk], a1 = a1p[kk], bc = bcp[kk], bs = bsp[kk];
      acc_c[0] = __builtin_amdgcn_mfma_f32_32x32x2f32(a0, bc, acc_c[0], 0, 0, 0);
      acc_s[0] = __builtin_amdgcn_mfma_f32_32x32x2f32(a0, bs, acc_s[0], 0, 0, 0);
      acc_c[1] = __builtin_amdgcn_mfma_f32_32x32x2f32(a1, bc, acc_c[1], 0, 0, 0);
      acc_s[1] = __builtin_amdgcn_mfma_f32_32x32x2f32(a1, bs, acc_s[1], 0, 0, 0);
    }
  }
  // acc[i] of lane l: segment row 8 (i / 4) + 4 half + i % 4 of the tile, comb column l & 31.
  const int64_t ci = i0 + wn * 32 + row;
  if (ci >= n_comb) return;
  const float2 cr = corr[ci];
  const int64_t nfreq = n_comb / nb_f;
  const int64_t f = ci / nb_f;
  const int m = (int)(ci - f * nb_f);
#pragma unroll
  for (int mi = 0; mi < 2; ++mi) {
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const int64_t sg = s0 + mi * 32 + 8 * (i / 4) + 4 * half + (i % 4);
      if (sg < nseg) {
        const float2 x = make_float2(acc_c[mi][i], acc_s[mi][i]);
        fold[((sg * nfreq + f) * 2 + wm) * nb_f + m] = cmul(x, cr);
      }
    }
  }
}

// common.py:344-372 with the rows of one level (segment pair q, set j).
template <int NBF>
__global__ void __launch_bounds__(128)
cshift_add_kernel(const float2* __restrict__ in, int64_t ncoords_prev,
                  const int4* __restrict__ table, const double2* __restrict__ phases,
                  int64_t ncoords_cur, int nbins, int64_t nblk_j, float2* __restrict__ out) {
  __shared__ float2 s_ph[kRowsPerBlock][2][NBF];
  __shared__ int s_row[kRowsPerBlock][2];
  const int tid = threadIdx.x;
  const int64_t q = blockIdx.x / nblk_j;
  const int64_t j0 = (blockIdx.x - q * nblk_j) * kRowsPerBlock;
  {
    const int r = tid >> 1;
    const int which = tid & 1;
    const int64_t j = j0 + r;
    if (j < ncoords_cur) {
      const double2 ph = phases[j];
      const float s32 = (float)(which == 0 ? ph.x : ph.y);
      const float step = (float)((-2.0 * M_PI * (double)s32) / (double)nbins);
      const float2 delta = make_float2(cosf(step), sinf(step));
      float2 p = make_float2(1.0f, 0.0f);
      for (int k = 0; k < NBF; ++k) {
        s_ph[r][which][k] = p;
        p = cmul(p, delta);
      }
      const int4 tb = table[j];
      s_row[r][which] = which == 0 ? tb.x : tb.z;
    }
  }
  __syncthreads();
  constexpr int E = 2 * NBF;  // complex elements per row
  const int nrows = (int)std::min<int64_t>(kRowsPerBlock, ncoords_cur - j0);
  for (int e = tid; e < nrows * E; e += 128) {
    const int r = e / E;
    const int rem = e - r * E;
    const int comp = rem / NBF;
    const int k = rem - comp * NBF;
    const float2 t = in[((2 * q) * ncoords_prev + s_row[r][0]) * E + comp * NBF + k];
    const float2 h = in[((2 * q + 1) * ncoords_prev + s_row[r][1]) * E + comp * NBF + k];
    const float2 a = cmul(t, s_ph[r][0][k]);
    const float2 b = cmul(h, s_ph[r][1][k]);
    out[(q * ncoords_cur + j0 + r) * E + comp * NBF + k] = make_float2(a.x + b.x, a.y + b.y);
  }
}

// irfft (FP32) + normalisation + boxcar S/N of complex rows [n][2][nb_f].
__global__ void __launch_bounds__(pyl::kScoreBlock)
cscore_kernel(const float2* __restrict__ folds, int64_t n, int nbins, pyl::ScoreConst k,
              int mode, float* __restrict__ out) {
  extern __shared__ float S[];  // [(B + wmax)][kScoreStride], then twiddles [2][B]
  const int tid = threadIdx.x;
  const int nb_f = nbins / 2 + 1;
  const int64_t p0 = (int64_t)blockIdx.x * pyl::kScoreBlock;
  const int nprof = (int)std::min<int64_t>(pyl::kScoreBlock, n - p0);
  float* tw = S + (nbins + k.wmax) * pyl::kScoreStride;  // cos, then sin of 2 pi j / B
  pyl::fill_twiddles(tw, nbins, tid, pyl::kScoreBlock);
  __syncthreads();
  for (int r = 0; r < nprof; ++r) {
    const float2* row = folds + (p0 + r) * 2 * nb_f;
    for (int b = tid; b < nbins; b += pyl::kScoreBlock) {
      const float e = pyl::irfft_bin(row, nbins, b, tw);
      const float v = pyl::irfft_bin(row + nb_f, nbins, b, tw);
      S[b * pyl::kScoreStride + r] = pyl::norm_ev(e, v);
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

extern "C" int pyl_cfold_plan_destroy(pyl_cfold_plan* plan) {
  if (!plan) return PYL_OK;
  if (plan->basis) (void)hipFree(plan->basis);
  if (plan->corr) (void)hipFree(plan->corr);
  delete plan;
  return PYL_OK;
}

extern "C" int pyl_cfold_plan_create(const double* freqs, int64_t nfreq, int64_t seg_len,
                                     int32_t nbins, double tsamp, double t_ref, void* stream,
                                     pyl_cfold_plan** plan_out) {
  PYL_REQUIRE(plan_out != nullptr, "plan_out must not be NULL");
  PYL_REQUIRE(freqs != nullptr && nfreq > 0, "freq_arr must not be empty");
  PYL_REQUIRE(seg_len > 0 && seg_len <= (int64_t)1 << 30, "invalid segment_len %lld",
              (long long)seg_len);
  PYL_REQUIRE(nbins >= 4 && (nbins & (nbins - 1)) == 0,
              "the complex fold needs a power-of-two nbins >= 4, got %d", nbins);
  PYL_REQUIRE(tsamp > 0, "tsamp must be positive");
  *plan_out = nullptr;
  hipStream_t s = pyl::as_stream(stream);
  auto* p = new pyl_cfold_plan();
  p->nfreq = nfreq;
  p->seg_len = seg_len;
  p->nbins = nbins;
  p->nb_f = nbins / 2 + 1;
  p->n_comb = nfreq * p->nb_f;
  // fold.py:441-446 and 474-479 on the host: FP32 comb frequencies and correction phasors.
  std::vector<float> target(p->n_comb);
  std::vector<float2> corr(p->n_comb);
  const float corr_factor = (float)(2.0 * M_PI * t_ref);
  for (int64_t f = 0; f < nfreq; ++f) {
    for (int m = 0; m < p->nb_f; ++m) {
      const float tf = (float)(freqs[f] * (double)m);
      target[f * p->nb_f + m] = tf;
      const double a = (double)corr_factor * (double)tf;
      corr[f * p->nb_f + m] = make_float2((float)cos(a), (float)sin(a));
    }
  }
  float* d_target = nullptr;
  auto fail_hip = [&](hipError_t e, const char* what) {
    if (d_target) (void)hipFree(d_target);
    (void)pyl_cfold_plan_destroy(p);
    return pyl::fail(PYL_ERR_HIP, "%s failed: %s", what, hipGetErrorString(e));
  };
  hipError_t e;
  if ((e = hipMalloc(&p->basis, sizeof(float) * 2 * p->n_comb * seg_len)) != hipSuccess)
    return fail_hip(e, "hipMalloc(basis)");
  if ((e = hipMalloc(&p->corr, sizeof(float2) * p->n_comb)) != hipSuccess)
    return fail_hip(e, "hipMalloc(corr)");
  if ((e = hipMalloc(&d_target, sizeof(float) * p->n_comb)) != hipSuccess)
    return fail_hip(e, "hipMalloc(target)");
  if ((e = hipMemcpyAsync(d_target, target.data(), sizeof(float) * p->n_comb,
                          hipMemcpyHostToDevice, s)) != hipSuccess)
    return fail_hip(e, "hipMemcpyAsync");
  if ((e = hipMemcpyAsync(p->corr, corr.data(), sizeof(float2) * p->n_comb,
                          hipMemcpyHostToDevice, s)) != hipSuccess)
    return fail_hip(e, "hipMemcpyAsync");
  const float time_factor = (float)(-2.0 * M_PI * tsamp);
  const int64_t total = p->n_comb * seg_len;
  cfold_basis_kernel<<<(unsigned)pyl::ceil_div(total, 256), 256, 0, s>>>(
      d_target, p->n_comb, seg_len, time_factor, p->basis);
  if ((e = hipGetLastError()) != hipSuccess) return fail_hip(e, "cfold_basis_kernel");
  if ((e = hipStreamSynchronize(s)) != hipSuccess) return fail_hip(e, "hipStreamSynchronize");
  (void)hipFree(d_target);
  d_target = nullptr;
  *plan_out = p;
  return PYL_OK;
}

extern "C" int pyl_cfold_execute(pyl_cfold_plan* p, const float* ts_e, const float* ts_v,
                                 int64_t nsamps, float* fold_out, void* stream) {
  PYL_REQUIRE(p != nullptr, "null plan");
  PYL_REQUIRE(ts_e && ts_v && fold_out, "null device pointer");
  PYL_REQUIRE(nsamps > 0 && nsamps % p->seg_len == 0,
              "the complex fold needs nsamps to be a multiple of segment_len");
  hipStream_t s = pyl::as_stream(stream);
  const int64_t nseg = nsamps / p->seg_len;
  const int64_t m_tiles = pyl::ceil_div(nseg, kCfSeg);
  const int64_t n_tiles = pyl::ceil_div(p->n_comb, kCfComb);
  PYL_REQUIRE(m_tiles * n_tiles < ((int64_t)1 << 31) - 8, "complex fold grid too large");
  const int nblk = (int)(m_tiles * n_tiles);
  const unsigned grid = (unsigned)(pyl::ceil_div(nblk, 8) * 8);
  const bool vec = p->seg_len % 4 == 0 && reinterpret_cast<uintptr_t>(ts_e) % 16 == 0 &&
                   reinterpret_cast<uintptr_t>(ts_v) % 16 == 0;
  auto* out = reinterpret_cast<float2*>(fold_out);
  if (vec) {
    cfold_mfma_kernel<true><<<grid, kCfThreads, 0, s>>>(ts_e, ts_v, p->basis, p->corr, nseg,
                                                        p->seg_len, p->n_comb, p->nb_f,
                                                        (int)m_tiles, nblk, out);
  } else {
    cfold_mfma_kernel<false><<<grid, kCfThreads, 0, s>>>(ts_e, ts_v, p->basis, p->corr, nseg,
                                                         p->seg_len, p->n_comb, p->nb_f,
                                                         (int)m_tiles, nblk, out);
  }
  PYL_CHECK_LAUNCH("cfold_mfma_kernel");
  return PYL_OK;
}

template <int NBF>
static void launch_cshift(const float2* in, int64_t nprev, const int4* tb, const double2* ph,
                          int64_t ncur, int nbins, int64_t nseg_out, float2* out,
                          hipStream_t s) {
  const int64_t nblk_j = pyl::ceil_div(ncur, kRowsPerBlock);
  cshift_add_kernel<NBF><<<(unsigned)(nblk_j * nseg_out), 128, 0, s>>>(
      in, nprev, tb, ph, ncur, nbins, nblk_j, out);
}

extern "C" int pyl_ffa_shift_add_complex(const float* fold_in, int64_t nseg_in,
                                         int64_t ncoords_prev, const int32_t* table_dev,
                                         const double* phases_dev, int64_t ncoords_cur,
                                         int32_t nbins, float* fold_out, void* stream) {
  PYL_REQUIRE(fold_in && table_dev && phases_dev && fold_out, "null device pointer");
  PYL_REQUIRE(nseg_in >= 2 && nseg_in % 2 == 0, "nsegments must be even and >= 2, got %lld",
              (long long)nseg_in);
  PYL_REQUIRE(ncoords_prev > 0 && ncoords_cur > 0, "empty parameter grid");
  PYL_REQUIRE(nbins == 32 || nbins == 64 || nbins == 128,
              "the complex level supports nbins in {32, 64, 128}, got %d", nbins);
  hipStream_t s = pyl::as_stream(stream);
  const int64_t nseg_out = nseg_in / 2;
  const auto* in = reinterpret_cast<const float2*>(fold_in);
  auto* out = reinterpret_cast<float2*>(fold_out);
  const auto* tb = reinterpret_cast<const int4*>(table_dev);
  const auto* ph = reinterpret_cast<const double2*>(phases_dev);
  switch (nbins) {
    case 32: launch_cshift<17>(in, ncoords_prev, tb, ph, ncoords_cur, nbins, nseg_out, out, s); break;
    case 64: launch_cshift<33>(in, ncoords_prev, tb, ph, ncoords_cur, nbins, nseg_out, out, s); break;
    default: launch_cshift<65>(in, ncoords_prev, tb, ph, ncoords_cur, nbins, nseg_out, out, s); break;
  }
  PYL_CHECK_LAUNCH("cshift_add_kernel");
  return PYL_OK;
}

extern "C" int pyl_score_boxcar_complex(const float* folds, int64_t n, int32_t nbins,
                                        const int64_t* widths, int32_t nw, double stdnoise,
                                        int32_t mode, float* out, void* stream) {
  PYL_REQUIRE(widths, "null widths");
  PYL_REQUIRE(n >= 0, "negative profile count");
  PYL_REQUIRE(nbins >= 4 && (nbins & (nbins - 1)) == 0,
              "the complex scorer needs a power-of-two nbins >= 4, got %d", nbins);
  PYL_REQUIRE(mode == PYL_SCORE_PER_WIDTH || mode == PYL_SCORE_MAX, "invalid mode %d", mode);
  pyl::ScoreConst k;
  int rc = pyl::build_score_const(nbins, widths, nw, stdnoise, &k);
  if (rc != PYL_OK) return rc;
  if (n == 0) return PYL_OK;
  PYL_REQUIRE(folds && out, "null device pointer");
  const size_t lds = sizeof(float) * ((size_t)(nbins + k.wmax) * pyl::kScoreStride + 2 * nbins);
  PYL_REQUIRE(lds <= 64 * 1024, "nbins too large for the complex scorer (%d)", nbins);
  cscore_kernel<<<(unsigned)pyl::ceil_div(n, pyl::kScoreBlock), pyl::kScoreBlock, lds,
                  pyl::as_stream(stream)>>>(reinterpret_cast<const float2*>(folds), n, nbins, k,
                                            mode, out);
  PYL_CHECK_LAUNCH("cscore_kernel");
  return PYL_OK;
}
