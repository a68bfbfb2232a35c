// Complex (Fourier-domain) FFA: leaf fold, level shift-add and scoring
// (use_fourier=True: fold.py:395-481,572-614; common.py:344-372; ffa.py:400-411).
//
// Leaf fold. X[s, f, c, m] = sum_t x_c[s, t] e^{-2 pi i f m t tsamp}, m = 0..B/2, is a
// dense contraction: the segments (rows of length seg_len) times a basis of n_comb =
// nfreq * (B/2 + 1) complex exponentials. The basis is data-independent, so the plan
// builds it once in HBM (FP32 cos / sin of the FP64 angle f32(f m) * t * f32(-2 pi tsamp),
// as the reference does); each execute is one FP32 GEMM per component through rocBLAS
// (MFMA on gfx950) followed by a kernel that applies the e^{+2 pi i f m t_ref} correction in
// complex64 and writes the fold layout [nseg][nfreq][2][B/2 + 1].
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
#include <rocblas/rocblas.h>

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
  rocblas_handle blas = nullptr;
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

// C_comp[s][i] (real part, i < n_comb) and C_comp[s][n_comb + i] (imaginary part) times the
// correction phasor, into fold[s][f][comp][m].
__global__ void cfold_finish_kernel(const float* __restrict__ ce, const float* __restrict__ cv,
                                    const float2* __restrict__ corr, int64_t nseg,
                                    int64_t n_comb, int nb_f, float2* __restrict__ fold) {
  const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= nseg * 2 * n_comb) return;
  const int64_t s = gid / (2 * n_comb);
  const int64_t r = gid - s * 2 * n_comb;
  const int comp = (int)(r / n_comb);
  const int64_t i = r - (int64_t)comp * n_comb;
  const float* c = comp == 0 ? ce : cv;
  const float2 x = make_float2(c[s * 2 * n_comb + i], c[s * 2 * n_comb + n_comb + i]);
  const int64_t f = i / nb_f;
  const int m = (int)(i - f * nb_f);
  fold[((s * (n_comb / nb_f) + f) * 2 + comp) * nb_f + m] = cmul(x, corr[i]);
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
  if (plan->blas) (void)rocblas_destroy_handle(plan->blas);
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
  if (rocblas_create_handle(&p->blas) != rocblas_status_success) {
    (void)pyl_cfold_plan_destroy(p);
    return pyl::fail(PYL_ERR_HIP, "rocblas_create_handle failed");
  }
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
  const int64_t ncols = 2 * p->n_comb;
  float* ce = nullptr;
  PYL_CHECK_HIP(hipMallocAsync(reinterpret_cast<void**>(&ce),
                               sizeof(float) * 2 * nseg * ncols, s));
  float* cv = ce + nseg * ncols;
  int rc = PYL_OK;
  if (rocblas_set_stream(p->blas, s) != rocblas_status_success) {
    rc = pyl::fail(PYL_ERR_HIP, "rocblas_set_stream failed");
  } else {
    // Row-major C[nseg][2 n_comb] = X[nseg][seg_len] * W[2 n_comb][seg_len]^T, i.e.
    // column-major C^T = W^T' * X^T with W^T' the transpose of the stored W.
    const float one = 1.0f, zero = 0.0f;
    for (int c = 0; c < 2 && rc == PYL_OK; ++c) {
      const float* x = c == 0 ? ts_e : ts_v;
      float* out = c == 0 ? ce : cv;
      const rocblas_status st = rocblas_sgemm(
          p->blas, rocblas_operation_transpose, rocblas_operation_none, (rocblas_int)ncols,
          (rocblas_int)nseg, (rocblas_int)p->seg_len, &one, p->basis, (rocblas_int)p->seg_len,
          x, (rocblas_int)p->seg_len, &zero, out, (rocblas_int)ncols);
      if (st != rocblas_status_success)
        rc = pyl::fail(PYL_ERR_HIP, "rocblas_sgemm failed (status %d)", (int)st);
    }
  }
  if (rc == PYL_OK) {
    const int64_t total = nseg * 2 * p->n_comb;
    cfold_finish_kernel<<<(unsigned)pyl::ceil_div(total, 256), 256, 0, s>>>(
        ce, cv, p->corr, nseg, p->n_comb, p->nb_f, reinterpret_cast<float2*>(fold_out));
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
      rc = pyl::fail(PYL_ERR_HIP, "cfold_finish_kernel launch failed: %s", hipGetErrorString(e));
  }
  PYL_CHECK_HIP(hipFreeAsync(ce, s));
  return rc;
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
