// Branching patterns on the device (core/taylor.py:414-499 generate_bp_poly_taylor,
// core/circular.py:905-1010 generate_bp_circ_taylor): the expected children per parent of
// every pruning level. Each final-grid frequency evolves its own cell widths [d_k .. d_1]
// level by level; one thread owns one frequency and writes its branch count per level.
// The per-level constants (the target step dpf, the shift factor fac, the moving grid's
// Taylor shift matrix) come from the host, and every FP64 operation follows the host
// generator's order, so the counts are the host's bit for bit; the weighted averages over
// frequencies are then formed on the host.
#include <algorithm>
#include <cmath>

#include "pyl_common.hpp"

namespace {

constexpr int kMaxP = 5;

__global__ void bp_counts_kernel(int kind, int np, int nlev, int64_t nf,
                                 const double* __restrict__ f0, const double* __restrict__ dinit,
                                 const double* __restrict__ dpf, const double* __restrict__ fac,
                                 const double* __restrict__ tmat, int moving, int tiling,
                                 double eta, double* __restrict__ nbr_out) {
  const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  const double fr = f0[f];
  double dcur[kMaxP];
  for (int j = 0; j < np; ++j) dcur[j] = dinit[j];
  dcur[np - 1] = dcur[np - 1] * (pyl::kCVal / fr);
  bool snap_seen = false;
  const int n1 = np + 1;
  for (int l = 0; l < nlev; ++l) {
    const double* dp = dpf + (int64_t)l * np;
    const double* fc = fac + (int64_t)l * np;
    double nbr = 1.0;
    bool snap_split = false;
    double dnext[kMaxP];
    for (int j = 0; j < np; ++j) {
      dnext[j] = dcur[j];
      if (kind == 1 && j == 0) continue;  // circular: crackle never splits
      const double dnew = dp[j] * pyl::kCVal / fr;
      const double shift = fabs(dcur[j] - dnew) * (fc[j] * (fr / pyl::kCVal));
      if (!(shift >= eta - pyl::kFloatEps)) continue;
      const double num = fmax(1.0, ceil(dcur[j] / dnew - pyl::kFloatEps));
      nbr = nbr * num;
      dnext[j] = dcur[j] / num;
      if (kind == 1 && j == 1 && num > 1.0) snap_split = true;
    }
    if (kind == 1 && snap_split && !snap_seen) {
      nbr = nbr * 0.5;
      snap_seen = true;
    }
    nbr_out[(int64_t)l * nf + f] = nbr;
    if (moving) {
      const double* T = tmat + (int64_t)l * n1 * n1;
      double v[kMaxP + 1];
      for (int j = 0; j < np; ++j) v[j] = dnext[j];
      v[np] = 0.0;
      for (int i = 0; i < np; ++i) {
        double o;
        if (tiling == 0) {
          o = v[i] * fabs(T[i * n1 + i]);
        } else if (tiling == 1) {
          o = 0.0;
          for (int j = 0; j < n1; ++j) o = o + v[j] * fabs(T[i * n1 + j]);
        } else {
          o = 0.0;
          for (int j = 0; j < n1; ++j) o = o + (v[j] * v[j]) * (T[i * n1 + j] * T[i * n1 + j]);
          o = sqrt(o);
        }
        dcur[i] = o;
      }
    } else {
      for (int j = 0; j < np; ++j) dcur[j] = dnext[j];
    }
  }
}

}  // namespace

extern "C" int pyl_bp_counts(int32_t kind, int32_t n_params, int32_t nlev, int64_t nf,
                             const double* f0, const double* dinit, const double* dpf,
                             const double* fac, const double* tmat, int32_t moving,
                             int32_t tiling, double eta, double* nbr_out, void* stream) {
  PYL_REQUIRE(kind == 0 || kind == 1, "pattern kind must be 0 (Taylor) or 1 (circular)");
  PYL_REQUIRE(n_params >= 1 && n_params <= kMaxP, "n_params must be in [1, %d]", kMaxP);
  PYL_REQUIRE(kind == 0 || n_params == 5,
              "Circular branching pattern requires exactly 5 parameters.");
  PYL_REQUIRE(nlev >= 0 && nf >= 1, "empty pattern");
  PYL_REQUIRE(tiling >= 0 && tiling <= 2, "unknown tiling strategy %d", tiling);
  PYL_REQUIRE(f0 && dinit && dpf && fac && nbr_out && (!moving || tmat), "null argument");
  if (nlev == 0) return PYL_OK;
  hipStream_t s = pyl::as_stream(stream);
  const int n1 = n_params + 1;
  const size_t nd = (size_t)nf + n_params + 2 * (size_t)nlev * n_params +
                    (moving ? (size_t)nlev * n1 * n1 : 0) + (size_t)nlev * nf;
  double* buf = nullptr;
  PYL_CHECK_HIP(hipMallocAsync(reinterpret_cast<void**>(&buf), sizeof(double) * nd, s));
  double* d_f0 = buf;
  double* d_init = d_f0 + nf;
  double* d_dpf = d_init + n_params;
  double* d_fac = d_dpf + (size_t)nlev * n_params;
  double* d_t = d_fac + (size_t)nlev * n_params;
  double* d_out = d_t + (moving ? (size_t)nlev * n1 * n1 : 0);
  int rc = PYL_OK;
  auto up = [&](double* dst, const double* src, size_t n) {
    if (rc != PYL_OK || n == 0) return;
    const hipError_t e = hipMemcpyAsync(dst, src, sizeof(double) * n, hipMemcpyHostToDevice, s);
    if (e != hipSuccess) rc = pyl::fail(PYL_ERR_HIP, "hipMemcpyAsync: %s", hipGetErrorString(e));
  };
  up(d_f0, f0, nf);
  up(d_init, dinit, n_params);
  up(d_dpf, dpf, (size_t)nlev * n_params);
  up(d_fac, fac, (size_t)nlev * n_params);
  if (moving) up(d_t, tmat, (size_t)nlev * n1 * n1);
  if (rc == PYL_OK) {
    bp_counts_kernel<<<(unsigned)pyl::ceil_div(nf, 64), 64, 0, s>>>(
        kind, n_params, nlev, nf, d_f0, d_init, d_dpf, d_fac, d_t, moving, tiling, eta, d_out);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
      e = hipMemcpyAsync(nbr_out, d_out, sizeof(double) * (size_t)nlev * nf,
                         hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) rc = pyl::fail(PYL_ERR_HIP, "bp_counts: %s", hipGetErrorString(e));
  }
  (void)hipFreeAsync(buf, s);
  return rc;
}
