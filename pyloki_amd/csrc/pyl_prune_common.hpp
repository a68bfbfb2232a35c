// Shared pieces of the EP level kernels (Taylor and circular bases): per-level constants,
// grid index truncation, the parent search, odometer child decode, the FP64 Taylor
// translation product and the boxcar scorer of a profile held in an LDS column.
#pragma once

#include <algorithm>
#include <cfloat>
#include <climits>
#include <cmath>

#include "pyl_score.hpp"

namespace pyl_ep {
namespace {  // internal linkage: every translation unit gets its own copy

constexpr int kMaxP = 5;
constexpr int kMaxN = kMaxP + 1;
constexpr int kBlk = pyl::kScoreBlock;
constexpr int kStride = pyl::kScoreStride;

struct LevelConst {
  int32_t P, nbins, branch_max, tiling;
  int32_t fixed;  // fixed grid: resolve from coord_init without the init offsets, no transform
  int32_t nparams_ffa, n_accel, n_freq;
  double eta;
  double dpf[kMaxP];          // poly_taylor_step_f(P, scale_cur) row order d_P..d_1
  double fac[kMaxP];          // shift factors, same order
  double tadd[kMaxN * kMaxN];
  double tinit[kMaxN * kMaxN];
  double tnext[kMaxN * kMaxN];
  double dt_add_init;         // t0_add - t0_init
  double amin, fmin, sa, sf;  // accel/freq grid origin and inverse step
  pyl::ScoreConst sc;         // boxcar widths and template constants
  // Chebyshev basis: tadd / tinit / tnext hold transposed domain-transfer matrices
  // (cur -> add, cur -> init, cur -> next), tprev the one of prev -> cur; rmat, cfact and
  // tspow convert coefficients on the current domain to Taylor coefficients.
  double tprev[kMaxN * kMaxN];
  double rmat[kMaxN * kMaxN];  // R[k][m]: T_k(x) = sum_m R[k][m] x^m
  double cfact[kMaxN];         // m!
  double tspow[kMaxN];         // scale_cur^m
};

int build_const(const pyl_prune_params* p, const pyl_prune_coords* c, LevelConst* k) {
  PYL_REQUIRE(p != nullptr, "null pyl_prune_params");
  PYL_REQUIRE(p->poly_order >= 2 && p->poly_order <= kMaxP, "poly_order must be in [2, %d]",
              kMaxP);
  PYL_REQUIRE(p->nbins > 1, "nbins must be > 1");
  PYL_REQUIRE(p->nparams_ffa >= 2 && p->nparams_ffa <= p->poly_order,
              "FFA grid must have 2..poly_order parameters");
  PYL_REQUIRE(p->n_accel >= 1 && p->n_freq >= 1, "empty FFA grid");
  PYL_REQUIRE(p->tiling >= 0 && p->tiling <= 2, "invalid tiling strategy");
  *k = LevelConst{};
  const int P = p->poly_order;
  k->P = P;
  k->nbins = p->nbins;
  k->branch_max = p->branch_max;
  k->tiling = p->tiling;
  k->nparams_ffa = p->nparams_ffa;
  k->n_accel = p->n_accel;
  k->n_freq = p->n_freq;
  k->eta = p->eta;
  PYL_REQUIRE(p->fixed_grid == 0 || p->fixed_grid == 1, "fixed_grid must be 0 or 1");
  k->fixed = p->fixed_grid;
  if (c) {
    // psr_utils.py:59-96 (t_ref = 0, Chebyshev coarsening), reversed to d_P..d_1.
    const double dphi = p->eta / p->nbins;
    for (int kk = 0; kk < P; ++kk) {
      const double v = dphi * pyl::fact_d(kk + 1) / pyl::powi(c->scale_cur, kk + 1);
      k->dpf[P - 1 - kk] = (double)(1LL << kk) * v;
    }
    // psr_utils.py:227-248: (tobs)^(k+1) * nbins / (k+1)! / 2^k, k = P-1..0.
    for (int idx = 0; idx < P; ++idx) {
      const int kk = P - 1 - idx;
      const double f = pyl::powi(c->scale_cur, kk + 1) * p->nbins / pyl::fact_d(kk + 1);
      k->fac[idx] = f / (double)(1LL << kk);
    }
    pyl::taylor_matrix(P + 1, c->t0_add - c->t0_cur, k->tadd);
    pyl::taylor_matrix(P + 1, c->t0_init - c->t0_cur, k->tinit);
    pyl::taylor_matrix(P + 1, c->t0_next - c->t0_cur, k->tnext);
    k->dt_add_init = c->t0_add - c->t0_init;
  }
  k->amin = p->accel_lim[0];
  k->fmin = p->freq_lim[0];
  k->sa = p->accel_lim[1] > p->accel_lim[0] ? p->n_accel / (p->accel_lim[1] - p->accel_lim[0])
                                             : 0.0;
  k->sf = p->freq_lim[1] > p->freq_lim[0] ? p->n_freq / (p->freq_lim[1] - p->freq_lim[0]) : 0.0;
  int64_t widths[PYL_MAX_WIDTHS];
  PYL_REQUIRE(p->nw >= 1 && p->nw <= PYL_MAX_WIDTHS, "number of widths must be in [1, %d]",
              PYL_MAX_WIDTHS);
  for (int w = 0; w < p->nw; ++w) widths[w] = p->widths[w];
  int rc = pyl::build_score_const(p->nbins, widths, p->nw, 1.0, &k->sc);
  if (rc != PYL_OK) return rc;
  return PYL_OK;
}

// ---- Chebyshev tables (maths.py:399-557), host FP64 in the reference's evaluation order;
// float-to-integer powers by binary powering.
inline double nbinom_d(int n, int k) {
  if (k < 0 || k > n) return 0.0;
  int64_t r = 1;
  for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;  // exact: each prefix is binomial
  return (double)r;
}

inline double cheb_conn_s(int k, int m) {  // x^k = sum_m S[k][m] T_m(x)
  if (k < 0 || m < 0 || m > k || (k - m) % 2 != 0) return 0.0;
  const int n = (k - m) / 2;
  const int d = m == 0 ? 1 : 0;
  return std::ldexp(1.0, 1 - k - d) * nbinom_d(k, n);
}

inline double cheb_conn_r(int k, int m) {  // T_k(x) = sum_m R[k][m] x^m
  if (k < 0 || m < 0 || m > k || (k - m) % 2 != 0) return 0.0;
  if (m == 0 && k % 2 == 0) return (k / 2) % 2 == 0 ? 1.0 : -1.0;
  const int n = (k - m) / 2;
  const int r = (k + m) / 2;
  const double sign = n % 2 == 0 ? 1.0 : -1.0;
  return sign * pyl::powi(2.0, m - 1) * (2.0 * k / (k + m)) * nbinom_d(r, n);
}

inline double cheb_trans_c(int n, int k, double p, double q) {  // T_n(px+q) = sum_k C T_k
  if (k > n) return 0.0;
  double res = 0.0;
  for (int m = k; m <= n; ++m) {
    const double r_nm = cheb_conn_r(n, m);
    double inner = 0.0;
    for (int i = k; i <= m; ++i)
      inner += nbinom_d(m, i) * pyl::powi(p, i) * pyl::powi(q, m - i) * cheb_conn_s(i, k);
    res += r_nm * inner;
  }
  return res;
}

// Transposed descending domain-transfer matrix (poly_chebyshev_transform_matrix with
// coeff_order 1): out[j * n + i] = C[i][j], so that mat_row(out, alpha, n, j) is column j
// of alpha @ C summed over i in ascending order.
inline void cheb_matrix_t(int K, double tc1, double ts1, double tc2, double ts2, double* out) {
  const int n = K + 1;
  const double p = ts2 / ts1;
  const double q = (tc2 - tc1) / ts1;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) out[j * n + i] = cheb_trans_c(K - i, K - j, p, q);
}

// The Chebyshev tables of one level on top of the Taylor ones: step eta/nbins * C / f0 and
// shift factor nbins for every coefficient (psr_utils.py:258-278), transfer matrices and
// the conversion to Taylor coefficients on the current domain.
[[maybe_unused]] int build_cheb_const(const pyl_prune_params* p, const pyl_prune_coords* c, LevelConst* k) {
  int rc = build_const(p, c, k);
  if (rc != PYL_OK) return rc;
  const int P = k->P;
  const int n = P + 1;
  if (c) {
    PYL_REQUIRE(c->scale_cur > 0 && c->scale_prev > 0 && c->scale_next > 0,
                "ts1 and ts2 must be positive.");
    const double dphi = p->eta / p->nbins;
    for (int j = 0; j < P; ++j) {
      k->dpf[j] = dphi;
      k->fac[j] = (double)p->nbins;
    }
    cheb_matrix_t(P, c->t0_cur, c->scale_cur, c->t0_add, c->scale_cur, k->tadd);
    cheb_matrix_t(P, c->t0_cur, c->scale_cur, c->t0_init, c->scale_cur, k->tinit);
    cheb_matrix_t(P, c->t0_cur, c->scale_cur, c->t0_next, c->scale_next, k->tnext);
    cheb_matrix_t(P, c->t0_prev, c->scale_prev, c->t0_cur, c->scale_cur, k->tprev);
    for (int m = 0; m < n; ++m) k->tspow[m] = pyl::powi(c->scale_cur, m);
  }
  for (int kk = 0; kk < n; ++kk)
    for (int m = 0; m < n; ++m) k->rmat[kk * n + m] = cheb_conn_r(kk, m);
  for (int m = 0; m < n; ++m) k->cfact[m] = pyl::fact_d(m);
  return PYL_OK;
}

// Per-leaf arrays hold at most kMaxN values. Their loops run over all kMaxN slots with a
// guard (j < n), fully unrolled, so that every index is a compile-time constant and the
// arrays stay in registers instead of per-lane scratch memory; the arithmetic and its order
// are those of the plain loops. pick / put read and write slot i (run-time) by selects.
__device__ __forceinline__ double pick(const double* a, int i) {
  double r = a[0];
#pragma unroll
  for (int t = 1; t < kMaxN; ++t) r = i == t ? a[t] : r;
  return r;
}
__device__ __forceinline__ void put(double* a, int i, double v) {
#pragma unroll
  for (int t = 0; t < kMaxN; ++t)
    if (i == t) a[t] = v;
}

// transforms.py:532-560 cheby_to_taylor of descending coefficients a on the current domain:
// d_m = ((sum_k a_k R[k][m]) * m!) / scale^m, k ascending, written descending.
__device__ __forceinline__ void cheb_to_taylor(const LevelConst& k, const double* a,
                                               double* d) {
  const int n = k.P + 1;
  double asc[kMaxN];  // a in ascending order: asc[kk] = a[n - 1 - kk]
#pragma unroll
  for (int kk = 0; kk < kMaxN; ++kk) asc[kk] = pick(a, n - 1 - kk);
#pragma unroll
  for (int m = 0; m < kMaxN; ++m) {
    if (m < n) {
      double acc = 0.0;
#pragma unroll
      for (int kk = 0; kk < kMaxN; ++kk)
        if (kk < n) acc += asc[kk] * k.rmat[kk * n + m];
      put(d, n - 1 - m, (acc * k.cfact[m]) / k.tspow[m]);
    }
  }
}

// psr_utils.py:517-554: (accel, freq) grid indices by truncation with a 1e-6 guard.
__device__ __forceinline__ int64_t trunc_idx(double v, double vmin, double sinv, int n) {
  if (!(sinv > 0.0)) return 0;
  int64_t i = (int64_t)((v - vmin) * sinv + pyl::kFloatEps);
  if (i < 0) i = 0;
  else if (i >= n) i = n - 1;
  return i;
}

__device__ void atomic_min_f32(float* addr, float v) {
  int* ia = reinterpret_cast<int*>(addr);
  int old = __atomic_load_n(ia, __ATOMIC_RELAXED);
  while (__int_as_float(old) > v) {
    int assumed = old;
    old = atomicCAS(ia, assumed, __float_as_int(v));
    if (old == assumed) break;
  }
}

__device__ void atomic_max_f32(float* addr, float v) {
  int* ia = reinterpret_cast<int*>(addr);
  int old = __atomic_load_n(ia, __ATOMIC_RELAXED);
  while (__int_as_float(old) < v) {
    int assumed = old;
    old = atomicCAS(ia, assumed, __float_as_int(v));
    if (old == assumed) break;
  }
}

// Children (or survivors) per block of the children and materialize kernels: fewer on
// small levels, so that more blocks share a level whose blocks would otherwise run one
// long dependent sequence each on a few CUs. A launch for n items needs
// items_grid(n) blocks; per_block depends only on n, so a capacity-sized launch and an
// exact one partition a level alike, and a block never sees the other's rows.
__host__ __device__ inline int per_block(int64_t n) {
  return n < (1 << 15) ? 16 : (n < (1 << 17) ? 32 : kBlk);
}
__host__ __device__ inline int64_t items_grid(int64_t n) {
  return (n + per_block(n) - 1) / per_block(n);
}
// Blocks that cover any count up to cap.
inline unsigned items_grid_cap(int64_t cap) {
  const int64_t g = std::max<int64_t>(items_grid(cap), std::max<int64_t>(
      items_grid(std::min<int64_t>(cap, (1 << 15) - 1)),
      items_grid(std::min<int64_t>(cap, (1 << 17) - 1))));
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>(g, 0x7fffffff));
}

struct ChildLeaf {
  double val[kMaxN];  // d_P..d_1, d_0
  double err[kMaxN];
  double f0, flag;
};

// Parent of child c: the last p with off[p] <= c.
__device__ __forceinline__ int64_t find_parent(const int64_t* __restrict__ off, int64_t n_par,
                                               int64_t c) {
  int64_t lo = 0, hi = n_par;  // off[lo] <= c < off[hi]
  while (hi - lo > 1) {
    const int64_t mid = (lo + hi) >> 1;
    if (off[mid] <= c) lo = mid;
    else hi = mid;
  }
  return lo;
}

// The parents of the consecutive children c_first..c_last of a block: wavefront 0 finds
// the parent of c_first, wavefront 1 that of c_last, each by a 64-ary search (each step
// probes 64 evenly spaced offsets and keeps the interval where off <= c stops holding:
// about 4 dependent loads for 4 M parents instead of 22). Every thread of the block calls
// this; it returns after a barrier with s_range = {parent of c_first, parent of c_last}.
__device__ inline void block_parent_range(const int64_t* __restrict__ off, int64_t n_par,
                                          int64_t c_first, int64_t c_last, int64_t* s_range) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (wave < 2) {
    const int64_t c = wave == 0 ? c_first : c_last;
    int64_t lo = 0, hi = n_par;  // off[lo] <= c < off[hi]
    while (hi - lo > 1) {
      const int64_t step = (hi - lo + 63) / 64;
      const int64_t q = lo + (int64_t)lane * step;
      const unsigned long long m = __ballot(q < hi && off[q] <= c);  // a prefix of lanes
      const int64_t nlo = lo + (int64_t)(__popcll(m) - 1) * step;
      hi = std::min<int64_t>(nlo + step, hi);
      lo = nlo;
    }
    if (lane == 0) s_range[wave] = lo;
  }
  __syncthreads();
}

// find_parent within [lo, hi) where off[lo] <= c < off[hi].
__device__ __forceinline__ int64_t find_parent_in(const int64_t* __restrict__ off, int64_t lo,
                                                  int64_t hi, int64_t c) {
  while (hi - lo > 1) {
    const int64_t mid = (lo + hi) >> 1;
    if (off[mid] <= c) lo = mid;
    else hi = mid;
  }
  return lo;
}

__device__ __forceinline__ void decode_child(const LevelConst& k, const double* __restrict__ lf,
                                             const double* __restrict__ br,
                                             const int32_t* __restrict__ cn, int64_t local,
                                             ChildLeaf& out) {
  const int P = k.P;
#pragma unroll
  for (int j = kMaxP - 1; j >= 0; --j) {
    if (j < P) {
      const int32_t cnt = cn[j];
      const int64_t n = cnt < 0 ? 1 : cnt;
      const int64_t ij = local % n;
      local /= n;
      out.val[j] = cnt < 0 ? br[2 * j] : br[2 * j] + (double)ij * br[2 * j + 1];
      out.err[j] = br[2 * j + 1];
    }
  }
  put(out.val, P, lf[P * 2]);  // d_0 copied from the parent; its error starts at 0
  put(out.err, P, 0.0);
  out.f0 = lf[(P + 1) * 2];
  out.flag = lf[(P + 1) * 2 + 1];
}

__device__ __forceinline__ double mat_row(const double* __restrict__ T, const double* v, int n,
                                          int i) {
  double acc = 0.0;
#pragma unroll
  for (int j = 0; j < kMaxN; ++j)
    if (j < n) acc += v[j] * T[i * n + j];
  return acc;
}

// Sequential scan + boxcar S/N (max over widths) of the profile in column `tid` of S
// (scoring.py:249-282).
__device__ __forceinline__ float score_column(const LevelConst& k, float* S, int tid) {
  return (float)pyl::score_column<false>(k.sc, k.nbins, S, tid, nullptr);
}

// The fixed grid's transform (dyn_poly_taylor.py:554-568, dyn_poly_cheby.py:698-712,
// dyn_circular_taylor.py:600-615): the child leaf as branched.
__device__ __forceinline__ void keep_leaf(const LevelConst& k, const ChildLeaf& ch, double* ol) {
  const int n = k.P + 1;
#pragma unroll
  for (int i = 0; i < kMaxN; ++i) {
    if (i < n) {
      ol[2 * i] = ch.val[i];
      ol[2 * i + 1] = ch.err[i];
    }
  }
}

// transforms.py:93-137 shift_taylor_full: the Taylor translation T of the values and the
// error propagation of the tiling strategy.
__device__ __forceinline__ void taylor_full(const LevelConst& k, const double* T,
                                            const ChildLeaf& ch, double* ol) {
  const int n = k.P + 1;
#pragma unroll
  for (int i = 0; i < kMaxN; ++i) {
    if (i < n) {
      ol[2 * i] = mat_row(T, ch.val, n, i);
      double e;
      if (k.tiling == PYL_TILING_AGGRESSIVE) {
        e = ch.err[i] * fabs(T[i * n + i]);
      } else if (k.tiling == PYL_TILING_CONSERVATIVE) {
        e = 0.0;
#pragma unroll
        for (int j = 0; j < kMaxN; ++j)
          if (j < n) e += ch.err[j] * fabs(T[i * n + j]);
      } else {
        e = 0.0;
#pragma unroll
        for (int j = 0; j < kMaxN; ++j) {
          if (j < n) {
            const double t = T[i * n + j];
            e += (ch.err[j] * ch.err[j]) * (t * t);
          }
        }
        e = sqrt(e);
      }
      ol[2 * i + 1] = e;
    }
  }
}

// Block-wide min/max of the scores of the block's valid threads, one atomic each.
__device__ __forceinline__ void block_minmax(float sc, bool valid, float* minmax) {
  float mn = valid ? sc : FLT_MAX;
  float mx = valid ? sc : -FLT_MAX;
  for (int o = 32; o > 0; o >>= 1) {
    mn = fminf(mn, __shfl_xor(mn, o));
    mx = fmaxf(mx, __shfl_xor(mx, o));
  }
  if ((threadIdx.x & 63) == 0 && mn <= mx) {
    atomic_min_f32(&minmax[0], mn);
    atomic_max_f32(&minmax[1], mx);
  }
}

// ---- Complex folds (use_fourier=True) ------------------------------------------------
// common.py:403-429 shift_add_complex_batch for one child: X[c][k] = seg[c][k] phi^k +
// parent[c][k], phi^k by the complex64 recurrence of delta = (cos, sin)(f32(-2 f32(pi)
// f32(phase) / B)). X is [2][B/2 + 1].
__device__ inline void cfold_child(const float2* __restrict__ seg, const float2* __restrict__ par,
                                   double phase, int nbins, float2* X) {
  const int nb_f = nbins / 2 + 1;
  const float s32 = (float)phase;
  const float ang = (float)((-2.0 * (double)(float)M_PI * (double)s32) / (double)nbins);
  const float2 d = make_float2(cosf(ang), sinf(ang));
  float2 p = make_float2(1.0f, 0.0f);
  for (int kk = 0; kk < nb_f; ++kk) {
    for (int c = 0; c < 2; ++c) {
      const float2 a = pyl::cmul(seg[c * nb_f + kk], p);
      const float2 f = par[c * nb_f + kk];
      X[c * nb_f + kk] = make_float2(a.x + f.x, a.y + f.y);
    }
    p = pyl::cmul(p, d);
  }
}

// The same child fold for a block of children, in two steps whose arithmetic is
// cfold_child's: every child's phasor powers phi^0..phi^(B/2) (its thread, the same
// recurrence) into ph[child][k], then the rows formed with lane = (component, harmonic),
// one child after another, so that segment and parent rows are read coalesced. Xs (or an
// output) row r = child r of the block.
__device__ inline void cphasors(double phase, int nbins, float2* ph) {
  const int nb_f = nbins / 2 + 1;
  const float s32 = (float)phase;
  const float ang = (float)((-2.0 * (double)(float)M_PI * (double)s32) / (double)nbins);
  const float2 d = make_float2(cosf(ang), sinf(ang));
  float2 p = make_float2(1.0f, 0.0f);
  for (int kk = 0; kk < nb_f; ++kk) {
    ph[kk] = p;
    p = pyl::cmul(p, d);
  }
}

__device__ inline void cfold_rows(const float2* __restrict__ seg,
                                  const float2* __restrict__ folds,
                                  const int32_t* s_row, const int32_t* s_parent,
                                  const float2* ph, int nvalid, int nbins, float2* X,
                                  int tid, int nthreads) {
  const int nb_f = nbins / 2 + 1;
#pragma unroll 8
  for (int r = 0; r < nvalid; ++r) {
    const float2* sr = seg + (int64_t)s_row[r] * 2 * nb_f;
    const float2* fr = folds + (int64_t)s_parent[r] * 2 * nb_f;
    for (int j = tid; j < 2 * nb_f; j += nthreads) {
      const int kk = j < nb_f ? j : j - nb_f;
      const float2 a = pyl::cmul(sr[j], ph[r * nb_f + kk]);
      const float2 f = fr[j];
      X[(int64_t)r * 2 * nb_f + j] = make_float2(a.x + f.x, a.y + f.y);
    }
  }
}

// common.py:463-490: X += row[c][k] phi^k for one segment of the ascend.
__device__ inline void cascend_add(const float2* __restrict__ row, double phase, int nbins,
                                   float2* X) {
  const int nb_f = nbins / 2 + 1;
  const float s32 = (float)phase;
  const float ang = (float)((-2.0 * (double)(float)M_PI * (double)s32) / (double)nbins);
  const float2 d = make_float2(cosf(ang), sinf(ang));
  float2 p = make_float2(1.0f, 0.0f);
  for (int kk = 0; kk < nb_f; ++kk) {
    for (int c = 0; c < 2; ++c) {
      const float2 a = pyl::cmul(row[c * nb_f + kk], p);
      const float2 x = X[c * nb_f + kk];
      X[c * nb_f + kk] = make_float2(x.x + a.x, x.y + a.y);
    }
    p = pyl::cmul(p, d);
  }
}

// LDS of the complex block scorer: rows Xs [kBlk][2][B/2 + 1] complex, the scorer's columns
// and the irfft twiddles. The children's phasor powers [kBlk][B/2 + 1] are needed only
// before the columns are written, so they share the columns' space when it is large
// enough, else they follow the twiddles (16-byte aligned).
__host__ __device__ inline size_t complex_ph_offset(int nbins, int wmax) {
  const size_t xs = sizeof(float2) * (size_t)kBlk * 2 * (nbins / 2 + 1);
  const size_t cols = sizeof(float) * (size_t)(nbins + wmax) * kStride;
  const size_t ph = sizeof(float2) * (size_t)kBlk * (nbins / 2 + 1);
  if (ph <= cols) return xs;
  return (xs + cols + sizeof(float) * 2 * (size_t)nbins + 15) & ~(size_t)15;
}
__host__ __device__ inline size_t complex_block_lds(int nbins, int wmax) {
  const size_t base = sizeof(float2) * (size_t)kBlk * 2 * (nbins / 2 + 1) +
                      sizeof(float) * ((size_t)(nbins + wmax) * kStride + 2 * nbins);
  const size_t ph_end =
      complex_ph_offset(nbins, wmax) + sizeof(float2) * (size_t)kBlk * (nbins / 2 + 1);
  return base > ph_end ? base : ph_end;
}

// Threads of an EP children block: four wavefronts share the block's (at most kBlk)
// children in the fold, the inverse transforms (complex folds) and the widths of the
// scorer, so that a small level's few blocks still occupy every SIMD of their CUs.
constexpr int kCBlk = 4 * kBlk;

typedef float f32x16_t __attribute__((ext_vector_type(16)));

// irfft + e/sqrt(v) of the block's rows into the scorer's columns, as a real product on the
// matrix cores: Y[m][b] = sum_j A[m][j] W[j][b] over the 2 (B/2 + 1) real coefficients
// j = (harmonic s, real or imaginary part) of row m = 2 r + component of child r, with
//   W[(s, re)][b] = w_s cos(2 pi s b / B), w_s = 1 at s = 0 and s = B/2, else 2,
//   W[(s, im)][b] = -2 sin(2 pi s b / B), 0 at s = 0 and s = B/2
// (pyl::irfft_bin's terms: the imaginary parts of the DC and Nyquist harmonics are
// ignored, as a real inverse transform does), then times 1/B. The block's wavefronts take
// the 32 x 32 tiles of Y (v_mfma_f32_32x32x2f32, one harmonic per step: lanes 0-31 carry
// real parts, lanes 32-63 imaginary ones); a lane ends holding bin b of 8 children's e and
// v. FP32 products and sums in the matrix unit's order: within a few ulps of irfft_bin,
// not bit-equal to it. The caller has filled Xs [r][component][B/2 + 1] and synchronised;
// returns after a barrier, the columns then scored by block_score_columns (or
// score_column by thread tid < nvalid).
__device__ inline float* complex_block_columns(const LevelConst& k, char* lds, int nvalid,
                                               int tid, int nthreads) {
  const int B = k.nbins;
  const int nb_f = B / 2 + 1;
  const float* Xf = reinterpret_cast<const float*>(lds);  // row m: 2 nb_f floats
  float* S = reinterpret_cast<float*>(lds + sizeof(float2) * (size_t)kBlk * 2 * nb_f);
  float* tw = S + (B + k.sc.wmax) * kStride;
  pyl::fill_twiddles(tw, B, tid, nthreads);
  __syncthreads();
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int nwaves = nthreads >> 6;
  const int row = lane & 31, half = lane >> 5;
  const int m_tiles = (2 * nvalid + 31) / 32;
  const int n_tiles = (B + 31) / 32;
  const float inv = 1.0f / (float)B;
  for (int t = wave; t < m_tiles * n_tiles; t += nwaves) {
    const int mt = t / n_tiles, nt = t - mt * n_tiles;
    const int m = 32 * mt + row;
    const bool mvalid = m < 2 * nvalid;
    const float* xa = Xf + (size_t)(mvalid ? m : 0) * 2 * nb_f + half;
    const int b = 32 * nt + row;
    const bool bvalid = b < B;
    const float* twp = tw + (half ? B : 0);
    f32x16_t acc;
#pragma unroll
    for (int i = 0; i < 16; ++i) acc[i] = 0.0f;
#pragma unroll 4
    for (int sh = 0; sh < nb_f; ++sh) {
      const float a = mvalid ? xa[2 * sh] : 0.0f;
      const bool edge = sh == 0 || sh == nb_f - 1;
      const float wt = edge ? (half ? 0.0f : 1.0f) : (half ? -2.0f : 2.0f);
      const float w = bvalid ? wt * twp[(sh * b) & (B - 1)] : 0.0f;
      acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a, w, acc, 0, 0, 0);
    }
    // acc[i] is row 8 (i / 4) + 4 half + i % 4 of the tile: rows (2 r, 2 r + 1) = (e, v)
    // of child r sit in acc[i], acc[i + 1] for even i.
    if (bvalid) {
#pragma unroll
      for (int i = 0; i < 16; i += 2) {
        const int r = (32 * mt + 8 * (i / 4) + 4 * half + (i % 4)) / 2;
        if (r < nvalid) S[b * kStride + r] = pyl::norm_ev(acc[i] * inv, acc[i + 1] * inv);
      }
    }
  }
  __syncthreads();
  return S;
}

// score_column of the block's columns (nvalid <= kBlk) with the widths spread over the
// wavefronts: the columns' owners (lanes of wavefront 0) scan, every wavefront takes every
// nwaves-th width, and the owner keeps the largest S/N. A maximum does not depend on the
// order of its terms, so thread tid < nvalid gets score_column's value; every thread of
// the block calls this.
__device__ inline float block_score_columns(const LevelConst& k, float* S, int nvalid,
                                              int tid, int nthreads) {
  __shared__ double s_best[kCBlk];
  const int B = k.nbins;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int nwaves = nthreads >> 6;
  if (tid < nvalid) {
    float run = S[tid];
    for (int i = 1; i < B; ++i) {
      run = run + S[i * kStride + tid];
      S[i * kStride + tid] = run;
    }
    for (int j = 0; j < k.sc.wmax; ++j)
      S[(B + j) * kStride + tid] = S[j * kStride + tid] + run;
  }
  __syncthreads();
  double best = -(double)FLT_MAX;
  if (lane < nvalid) {
    const float total = S[(B - 1) * kStride + lane];
    for (int w = wave; w < k.sc.nw; w += nwaves) {
      const int wd = k.sc.widths[w];
      float dmax = -FLT_MAX;
      for (int i = 0; i < B; ++i) {
        const float d = S[(i + wd) * kStride + lane] - S[i * kStride + lane];
        dmax = d > dmax ? d : dmax;
      }
      const double t1 = k.sc.hpb[w] * (double)dmax;
      const double t2 = k.sc.bb[w] * (double)total;
      const double snr = t1 - t2;
      best = snr > best ? snr : best;
    }
  }
  s_best[tid] = best;
  __syncthreads();
  if (tid < nvalid) {
    for (int v = 1; v < nwaves; ++v) {
      const double o = s_best[v * 64 + tid];
      best = o > best ? o : best;
    }
  }
  return (float)best;
}

// Dynamic LDS of the complex materialize blocks: phasors [kBlk][B/2 + 1], rows and parents.
inline size_t materialize_lds(bool complex_folds, int nbins) {
  return complex_folds ? sizeof(float2) * (size_t)kBlk * (nbins / 2 + 1) + 2 * sizeof(int32_t) * kBlk
                       : 0;
}

// Dynamic LDS of the EP scoring blocks: the real scorer's columns, or the complex rows, the
// columns and the twiddles. gfx950 gives a workgroup up to 160 KB.
constexpr size_t kMaxEpLds = 160 * 1024;
template <bool kComplex>
inline size_t ep_block_lds(const LevelConst& k) {
  if constexpr (kComplex) return complex_block_lds(k.nbins, k.sc.wmax);
  return sizeof(float) * (size_t)(k.nbins + k.sc.wmax) * kStride;
}

}  // namespace
}  // namespace pyl_ep
