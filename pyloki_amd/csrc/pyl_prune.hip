// EP (enumerate-and-prune) Taylor search on the GPU: one pruning level of
// prune.py:165-332 as three kernels, plus the final re-integration (ascend).
//
// Work decomposition. The reference walks parents in batches of 1024 and, per batch,
// branches, resolves, shift-adds and scores every child in parent-major, odometer order.
// Here every child of the level is one GPU thread:
//   prune_branch_kernel    thread per parent: per-parameter child counts, first centre and
//                          step (FP64, psr_utils.py:347-379 + taylor.py:118-142);
//   (caller)               exclusive prefix sum of the child counts -> child offsets, so
//                          child c of the level is the reference's c-th child;
//   prune_children_kernel  thread per child: parent by binary search, odometer decode,
//                          FP64 resolve (taylor.py:172-241); then the 64 children of a
//                          block are folded (parent row + rotated segment row, coalesced
//                          lane = bin loads) and scored by the sequential per-thread scan;
//   prune_materialize_kernel thread per survivor: transform (taylor.py:323-339), backtrack
//                          (prune.py:302-314); survivor rows re-folded coalesced.
// The child folds are never stored: a survivor's fold is recomputed from its parent row and
// segment row (one extra 1 KB read per survivor instead of a 512 B write per child).
//
// Taylor translation matrices T[i][j] = dt**(i-j)/(i-j)! are per-level constants; they are
// built on the host with the binary-powering power (powi) the reference's numba code uses,
// and the GPU only does the per-leaf FP64 matrix-vector products, in ascending column order.
#include <vector>

#include "pyl_budget.hpp"
#include "pyl_loop.hpp"
#include "pyl_prune_common.hpp"

namespace {

using namespace pyl_ep;

// The kernels below take their item count either by value or, when the pointer is not null,
// from device memory (the sync-free level loop). The light thread-per-parent kernels
// stride over their items, so the loop launches them with a capped grid. The children and
// materialize kernels have no such loop (a loop makes the compiler keep the level
// constants live in scalar registers across iterations, which spills them and costs the
// children kernel a fifth of its speed): the loop sizes their launches for the buffers'
// capacity and blocks past the count exit at once.
__device__ __forceinline__ int64_t dev_count(int64_t n, const int64_t* n_dev) {
  return n_dev ? *n_dev : n;
}

// Branching of parent i: per-parameter child count, first centre and step (FP64).
__device__ __forceinline__ void branch_one(const LevelConst& k, const double* __restrict__ leaves,
                                           int64_t i, double* __restrict__ branch,
                                           int32_t* __restrict__ counts,
                                           int64_t* __restrict__ nchild,
                                           unsigned long long* __restrict__ err) {
  const int P = k.P;
  const double* lf = leaves + i * (P + 2) * 2;
  const double f0 = lf[(P + 1) * 2];
  int64_t total = 1;
  for (int j = 0; j < P; ++j) {
    const double pc = lf[2 * j];
    const double dc = lf[2 * j + 1];
    const double dn = k.dpf[j] * pyl::kCVal / f0;
    const double shift = fabs(dc - dn) * (k.fac[j] * (f0 / pyl::kCVal));
    int32_t cnt = 1;
    double first = pc, step = dc;
    if (dc <= 0.0 || dn <= 0.0) {
      atomicMin(err, (unsigned long long)((i * P + j) * 4 + 2));
    } else {
      const double ratio = dc / dn;
      double nump = ceil(ratio - pyl::kFloatEps);
      if (nump < 1.0) nump = 1.0;
      if (nump > (double)k.branch_max) {
        atomicMin(err, (unsigned long long)((i * P + j) * 4 + 1));
        nump = 1.0;
      }
      cnt = (int32_t)nump;
      step = dc / (double)cnt;
      first = (pc - dc / 2.0) + step / 2.0;
    }
    if (shift < (k.eta - pyl::kFloatEps)) {
      cnt = -1;  // not split: the child keeps the parent value and step
      first = pc;
      step = dc;
    }
    branch[(i * P + j) * 2] = first;
    branch[(i * P + j) * 2 + 1] = step;
    counts[i * P + j] = cnt;
    total *= cnt < 0 ? 1 : cnt;
  }
  nchild[i] = total;
}

// chebyshev.py:72-177, first step: parent i moved from coord_prev to coord_cur
// (shift_cheby_full, values and widths); f0 and flag copied.
__device__ __forceinline__ void cheb_parent_one(const LevelConst& k,
                                                const double* __restrict__ leaves, int64_t i,
                                                double* __restrict__ leaves_cur) {
  const int P = k.P;
  const double* lf = leaves + i * (P + 2) * 2;
  ChildLeaf ch;
#pragma unroll
  for (int r = 0; r < kMaxN; ++r) {
    if (r <= P) {
      ch.val[r] = lf[2 * r];
      ch.err[r] = lf[2 * r + 1];
    }
  }
  double* ol = leaves_cur + i * (P + 2) * 2;
  taylor_full(k, k.tprev, ch, ol);
  ol[2 * (P + 1)] = lf[2 * (P + 1)];
  ol[2 * (P + 1) + 1] = lf[2 * (P + 1) + 1];
}

// Thread per parent.
__global__ void prune_branch_kernel(LevelConst k, const double* __restrict__ leaves,
                                    int64_t n_arg, const int64_t* __restrict__ n_dev,
                                    double* __restrict__ branch, int32_t* __restrict__ counts,
                                    int64_t* __restrict__ nchild,
                                    unsigned long long* __restrict__ err) {
  const int64_t n = dev_count(n_arg, n_dev);
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    branch_one(k, leaves, i, branch, counts, nchild, err);
}

__global__ void cheb_parents_kernel(LevelConst k, const double* __restrict__ leaves,
                                    int64_t n_arg, const int64_t* __restrict__ n_dev,
                                    double* __restrict__ leaves_cur) {
  const int64_t n = dev_count(n_arg, n_dev);
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    cheb_parent_one(k, leaves, i, leaves_cur);
}

// Folds are real [2][B] rows or, with kComplex, complex64 [2][B/2 + 1] rows (as float2).
// kCheby: the leaves are Chebyshev coefficients on the current domain (chebyshev.py:180-258).
template <bool kComplex, bool kCheby>
__global__ void __launch_bounds__(kCBlk)
prune_children_kernel(LevelConst k, const double* __restrict__ leaves,
                      const float* __restrict__ folds, int64_t n_par_arg,
                      const int64_t* __restrict__ n_par_dev,
                      const double* __restrict__ branch, const int32_t* __restrict__ counts,
                      const int64_t* __restrict__ off, int64_t n_child_arg,
                      const int64_t* __restrict__ n_child_dev,
                      const float* __restrict__ seg, int32_t* __restrict__ c_parent,
                      int32_t* __restrict__ c_row, int32_t* __restrict__ c_shift,
                      double* __restrict__ c_phase, float* __restrict__ c_score,
                      float* __restrict__ minmax) {
  extern __shared__ float S[];  // real: [(B + wmax)][kStride]; complex: complex_block_lds
  __shared__ int32_t s_parent[kBlk], s_row[kBlk], s_shift[kBlk];
  const int tid = threadIdx.x;
  const int64_t n_child = dev_count(n_child_arg, n_child_dev);
  const int per = per_block(n_child);
  const int64_t c0 = (int64_t)blockIdx.x * per;
  if (c0 >= n_child) return;  // the whole block
  const int64_t n_par = dev_count(n_par_arg, n_par_dev);
  const int P = k.P;
  const int B = k.nbins;
  const int nb_f = B / 2 + 1;
  const int nvalid = (int)std::min<int64_t>(per, n_child - c0);
  __shared__ int64_t s_range[2];
  block_parent_range(off, n_par, c0, c0 + nvalid - 1, s_range);
  if (tid < nvalid) {
    const int64_t c = c0 + tid;
    const int64_t par = find_parent_in(off, s_range[0], s_range[1] + 1, c);
    const double* lf = leaves + par * (P + 2) * 2;
    ChildLeaf ch;
    decode_child(k, lf, branch + par * P * 2, counts + par * P, c - off[par], ch);
    const int n = P + 1;
    double accel, v_add, d_add, v_init, d_init;
    if constexpr (kCheby) {
      // cheby_to_taylor_param_shift to t_add and t_init (transforms.py:703-737).
      double a[kMaxN], da[kMaxN], di[kMaxN];
#pragma unroll
      for (int j = 0; j < kMaxN; ++j) a[j] = j < n ? mat_row(k.tadd, ch.val, n, j) : 0.0;
      cheb_to_taylor(k, a, da);
#pragma unroll
      for (int j = 0; j < kMaxN; ++j) a[j] = j < n ? mat_row(k.tinit, ch.val, n, j) : 0.0;
      cheb_to_taylor(k, a, di);
      accel = pick(da, P - 2);
      v_add = pick(da, P - 1);
      d_add = pick(da, P);
      v_init = pick(di, P - 1);
      d_init = pick(di, P);
    } else {
      accel = mat_row(k.tadd, ch.val, n, P - 2);
      v_add = mat_row(k.tadd, ch.val, n, P - 1);
      d_add = mat_row(k.tadd, ch.val, n, P);
      v_init = mat_row(k.tinit, ch.val, n, P - 1);
      d_init = mat_row(k.tinit, ch.val, n, P);
    }
    if (k.fixed) {
      // taylor.py:244-277 / chebyshev.py:261-300: the leaf lives at t_init, so the
      // velocity and delay at t_add are taken whole (nothing subtracted).
      v_init = 0.0;
      d_init = 0.0;
    }
    const double vel = v_add - v_init;
    const double freq = ch.f0 * (1.0 - vel / pyl::kCVal);
    const double delay = (d_add - d_init) / pyl::kCVal;
    const double phase = pyl::phase_idx(k.dt_add_init, ch.f0, B, delay);
    const int64_t ia = trunc_idx(accel, k.amin, k.sa, k.n_accel);
    const int64_t ifr = trunc_idx(freq, k.fmin, k.sf, k.n_freq);
    const int32_t row = (int32_t)(ia * k.n_freq + ifr);
    const int32_t sh = pyl::shift_int(phase, B);
    c_parent[c] = (int32_t)par;
    c_row[c] = row;
    c_shift[c] = sh;
    c_phase[c] = phase;
    s_parent[tid] = (int32_t)par;
    s_row[tid] = row;
    s_shift[tid] = sh;
    if constexpr (kComplex)
      cphasors(phase, B,
               reinterpret_cast<float2*>(reinterpret_cast<char*>(S) +
                                         complex_ph_offset(B, k.sc.wmax)) + (size_t)tid * nb_f);
  }
  __syncthreads();
  float* Sc = S;
  if constexpr (kComplex) {
    char* lds = reinterpret_cast<char*>(S);
    cfold_rows(reinterpret_cast<const float2*>(seg), reinterpret_cast<const float2*>(folds),
               s_row, s_parent,
               reinterpret_cast<const float2*>(lds + complex_ph_offset(B, k.sc.wmax)), nvalid,
               B, reinterpret_cast<float2*>(lds), tid, kCBlk);
    __syncthreads();
    Sc = complex_block_columns(k, lds, nvalid, tid, kCBlk);
  } else {
    // Fold (common.py:389-399: parent row + segment row rotated by the shift) and
    // normalise, rows over the block's wavefronts, lane = bin, transposed into S.
#pragma unroll 4
    for (int r = tid >> 6; r < nvalid; r += kCBlk / 64) {
      const float* fr = folds + (int64_t)s_parent[r] * 2 * B;
      const float* sr = seg + (int64_t)s_row[r] * 2 * B;
      const int sh = s_shift[r];
      for (int b = tid & 63; b < B; b += 64) {
        int src = b - sh;
        if (src < 0) src += B;
        const float e = fr[b] + sr[src];
        const float v = fr[B + b] + sr[B + src];
        S[b * kStride + r] = pyl::norm_ev(e, v);
      }
    }
    __syncthreads();
  }
  const float sc = block_score_columns(k, Sc, nvalid, tid, kCBlk);
  if (tid < nvalid) c_score[c0 + tid] = sc;
  // Block min/max of the scores, then one atomic per wavefront.
  block_minmax(sc, tid < nvalid, minmax);
}

template <bool kComplex>
__global__ void __launch_bounds__(kCBlk)
prune_materialize_kernel(LevelConst k, const double* __restrict__ leaves,
                         const float* __restrict__ folds,
                         const double* __restrict__ branch, const int32_t* __restrict__ counts,
                         const int64_t* __restrict__ off, const float* __restrict__ seg,
                         const int32_t* __restrict__ c_parent, const int32_t* __restrict__ c_row,
                         const int32_t* __restrict__ c_shift,
                         const double* __restrict__ c_phase, const float* __restrict__ c_score,
                         const int64_t* __restrict__ surv, int64_t n_surv_arg,
                         const int64_t* __restrict__ n_surv_dev,
                         double* __restrict__ o_leaves, float* __restrict__ o_folds,
                         float* __restrict__ o_scores, int32_t* __restrict__ o_bt_base,
                         const int64_t* __restrict__ bt_row0) {
  __shared__ int64_t s_child[kBlk];
  __shared__ int32_t s_fpar[kBlk], s_frow[kBlk], s_fsh[kBlk];  // real folds
  extern __shared__ float2 s_ph[];  // complex: [kBlk][B/2 + 1] phasors, then rows, parents
  const int tid = threadIdx.x;
  const int64_t n_surv = dev_count(n_surv_arg, n_surv_dev);
  const int P = k.P;
  const int B = k.nbins;
  // Backtracks go to rows bt_row0[0].. of o_bt_base when bt_row0 is given (the level
  // loop's history), else to rows 0..
  const int per = per_block(n_surv);
  const int64_t s0 = (int64_t)blockIdx.x * per;
  if (s0 >= n_surv) return;  // the whole block
  int32_t* o_bt = o_bt_base + (bt_row0 ? *bt_row0 : 0) * (P + 2);
  const int nvalid = (int)std::min<int64_t>(per, n_surv - s0);
  if (tid < nvalid) {
    const int64_t s = s0 + tid;
    const int64_t c = surv[s];
    s_child[tid] = c;
    const int64_t par = c_parent[c];
    const double* lf = leaves + par * (P + 2) * 2;
    ChildLeaf ch;
    decode_child(k, lf, branch + par * P * 2, counts + par * P, c - off[par], ch);
    double* ol = o_leaves + s * (P + 2) * 2;
    // transforms.py:93-137 shift_taylor_full to coord_next (fixed grid: unchanged).
    if (k.fixed)
      keep_leaf(k, ch, ol);
    else
      taylor_full(k, k.tnext, ch, ol);
    ol[2 * (P + 1)] = ch.f0;
    ol[2 * (P + 1) + 1] = ch.flag;
    o_scores[s] = c_score[c];
    // prune.py:309-314: [global parent, grid indices, int32(phase) (truncated)].
    int32_t* bt = o_bt + s * (P + 2);
    bt[0] = (int32_t)par;
    const int32_t row = c_row[c];
    for (int q = 1; q <= k.nparams_ffa; ++q) bt[q] = 0;
    bt[k.nparams_ffa - 1] = row / k.n_freq;
    bt[k.nparams_ffa] = row % k.n_freq;
    for (int q = k.nparams_ffa + 1; q < P + 1; ++q) bt[q] = 0;
    bt[P + 1] = (int32_t)c_phase[c];
    if constexpr (kComplex) cphasors(c_phase[c], B, s_ph + (size_t)tid * (B / 2 + 1));
  }
  if constexpr (kComplex) {
    __syncthreads();
    int32_t* s_rw = reinterpret_cast<int32_t*>(s_ph + (size_t)kBlk * (B / 2 + 1));
    int32_t* s_pa = s_rw + kBlk;
    if (tid < nvalid) {
      s_rw[tid] = c_row[s_child[tid]];
      s_pa[tid] = c_parent[s_child[tid]];
    }
    __syncthreads();
    cfold_rows(reinterpret_cast<const float2*>(seg), reinterpret_cast<const float2*>(folds),
               s_rw, s_pa, s_ph, nvalid, B,
               reinterpret_cast<float2*>(o_folds) + s0 * 2 * (B / 2 + 1), tid, kCBlk);
  } else {
    if (tid < nvalid) {
      const int64_t c = s_child[tid];
      s_fpar[tid] = c_parent[c];
      s_frow[tid] = c_row[c];
      s_fsh[tid] = c_shift[c];
    }
    __syncthreads();
    // Rows over the block's wavefronts, lane = bin.
#pragma unroll 4
    for (int r = tid >> 6; r < nvalid; r += kCBlk / 64) {
      const float* fr = folds + (int64_t)s_fpar[r] * 2 * B;
      const float* sr = seg + (int64_t)s_frow[r] * 2 * B;
      const int sh = s_fsh[r];
      float* of = o_folds + (s0 + r) * 2 * B;
      for (int b = tid & 63; b < B; b += 64) {
        int src = b - sh;
        if (src < 0) src += B;
        of[b] = fr[b] + sr[src];
        of[B + b] = fr[B + b] + sr[B + src];
      }
    }
  }
}

// kCheby: cseg holds per segment the transposed transfer matrix from the survivors' domain
// to the segment's reference time (chebyshev.py:261-349); unused otherwise.
template <bool kComplex, bool kCheby>
__global__ void __launch_bounds__(kCBlk)
prune_ascend_kernel(LevelConst k, const double* __restrict__ leaves, int64_t n,
                    const float* __restrict__ dyp, int64_t ncoords, const int64_t* __restrict__ segi,
                    const double* __restrict__ segt, const double* __restrict__ cseg,
                    int32_t nseg, double t0_cur,
                    int32_t* __restrict__ scratch_row, double* __restrict__ scratch_phase,
                    float* __restrict__ o_folds, float* __restrict__ o_scores) {
  extern __shared__ float S[];
  const int tid = threadIdx.x;
  const int64_t s0 = (int64_t)blockIdx.x * kBlk;
  const int nvalid = (int)std::min<int64_t>(kBlk, n - s0);
  const int P = k.P;
  const int B = k.nbins;
  const int nb_f = B / 2 + 1;
  const int nn = P + 1;
  // taylor.py:280-320: resolve every segment so far for this survivor.
  if (tid < nvalid) {
    const double* lf = leaves + (s0 + tid) * (P + 2) * 2;
    double v[kMaxN];
    for (int i = 0; i < nn; ++i) v[i] = lf[2 * i];
    const double f0 = lf[(P + 1) * 2];
    for (int s = 0; s < nseg; ++s) {
      const double dt = segt[s] - t0_cur;
      double accel, vel, d0;
      if constexpr (kCheby) {
        const double* Ct = cseg + (size_t)s * nn * nn;
        double a[kMaxN], d[kMaxN];
        for (int j = 0; j < nn; ++j) a[j] = mat_row(Ct, v, nn, j);
        cheb_to_taylor(k, a, d);
        accel = d[P - 2];
        vel = d[P - 1];
        d0 = d[P];
      } else {
        double T[kMaxN * kMaxN];
        for (int i = 0; i < nn; ++i)
          for (int j = 0; j < nn; ++j)
            T[i * nn + j] = j <= i ? pyl::powi(dt, i - j) / pyl::fact_d(i - j) : 0.0;
        accel = mat_row(T, v, nn, P - 2);
        vel = mat_row(T, v, nn, P - 1);
        d0 = mat_row(T, v, nn, P);
      }
      const double freq = f0 * (1.0 - vel / pyl::kCVal);
      const double delay = d0 / pyl::kCVal;
      const double phase = pyl::phase_idx(dt, f0, B, delay);
      const int64_t ia = trunc_idx(accel, k.amin, k.sa, k.n_accel);
      const int64_t ifr = trunc_idx(freq, k.fmin, k.sf, k.n_freq);
      scratch_row[(s0 + tid) * nseg + s] = (int32_t)(ia * k.n_freq + ifr);
      scratch_phase[(s0 + tid) * nseg + s] = phase;
    }
    if constexpr (kComplex) {
      // common.py:463-490: sum of the rotated complex rows in segment order, from zero.
      float2* X = reinterpret_cast<float2*>(S) + (size_t)tid * 2 * nb_f;
      for (int i = 0; i < 2 * nb_f; ++i) X[i] = make_float2(0.0f, 0.0f);
      const float2* dc = reinterpret_cast<const float2*>(dyp);
      for (int s = 0; s < nseg; ++s) {
        const int64_t rr = (int64_t)segi[s] * ncoords + scratch_row[(s0 + tid) * nseg + s];
        cascend_add(dc + rr * 2 * nb_f, scratch_phase[(s0 + tid) * nseg + s], B, X);
      }
      float2* of = reinterpret_cast<float2*>(o_folds) + (s0 + tid) * 2 * nb_f;
      for (int i = 0; i < 2 * nb_f; ++i) of[i] = X[i];
    }
  }
  __syncthreads();
  float* Sc = S;
  if constexpr (kComplex) {
    Sc = complex_block_columns(k, reinterpret_cast<char*>(S), nvalid, tid, kCBlk);
  } else {
    // common.py:432-459: accumulate the rotated rows in segment order, from zero (rows
    // over the block's wavefronts, lane = bin).
    for (int r = tid >> 6; r < nvalid; r += kCBlk / 64) {
      const int64_t sv = s0 + r;
      float* of = o_folds + sv * 2 * B;
      for (int b = tid & 63; b < B; b += 64) {
        float ae = 0.0f, av = 0.0f;
        for (int s = 0; s < nseg; ++s) {
          const float* rowp =
              dyp + ((int64_t)segi[s] * ncoords + scratch_row[sv * nseg + s]) * 2 * B;
          int src = b - pyl::shift_int(scratch_phase[sv * nseg + s], B);
          if (src < 0) src += B;
          ae = ae + rowp[src];
          av = av + rowp[B + src];
        }
        of[b] = ae;
        of[B + b] = av;
        S[b * kStride + r] = pyl::norm_ev(ae, av);
      }
    }
    __syncthreads();
  }
  const float sc = block_score_columns(k, Sc, nvalid, tid, kCBlk);
  if (tid < nvalid) o_scores[s0 + tid] = sc;
}

}  // namespace

extern "C" int pyl_prune_branch(const pyl_prune_params* p, const pyl_prune_coords* c,
                                const double* leaves, int64_t n_par, double* branch,
                                int32_t* counts, int64_t* nchild, int64_t* err, void* stream) {
  LevelConst k;
  int rc = build_const(p, c, &k);
  if (rc != PYL_OK) return rc;
  PYL_REQUIRE(c != nullptr, "null pyl_prune_coords");
  PYL_REQUIRE(leaves && branch && counts && nchild && err, "null device pointer");
  if (n_par == 0) return PYL_OK;
  const int threads = 128;
  prune_branch_kernel<<<(unsigned)pyl::ceil_div(n_par, threads), threads, 0,
                        pyl::as_stream(stream)>>>(k, leaves, n_par, nullptr, branch, counts,
                                                  nchild,
                                                  reinterpret_cast<unsigned long long*>(err));
  PYL_CHECK_LAUNCH("prune_branch_kernel");
  return PYL_OK;
}

template <bool kComplex, bool kCheby = false>
static int prune_children(const pyl_prune_params* p, const pyl_prune_coords* c,
                          const double* leaves, const float* folds, int64_t n_par,
                          const double* branch, const int32_t* counts, const int64_t* off,
                          int64_t n_child, const float* seg, int32_t* c_parent, int32_t* c_row,
                          int32_t* c_shift, double* c_phase, float* c_score, float* minmax,
                          void* stream) {
  LevelConst k;
  int rc = kCheby ? build_cheb_const(p, c, &k) : build_const(p, c, &k);
  if (rc != PYL_OK) return rc;
  PYL_REQUIRE(c != nullptr, "null pyl_prune_coords");
  PYL_REQUIRE(leaves && folds && branch && counts && off && seg && c_parent && c_row &&
                  c_shift && c_phase && c_score && minmax,
              "null device pointer");
  PYL_REQUIRE((int64_t)p->n_accel * p->n_freq <= 0x7fffffff, "FFA grid too large");
  if (n_child == 0) return PYL_OK;
  const size_t lds = ep_block_lds<kComplex>(k);
  PYL_REQUIRE(lds <= kMaxEpLds, "nbins too large for the EP scorer (%d)", k.nbins);
  if constexpr (kComplex)
    PYL_REQUIRE((k.nbins & (k.nbins - 1)) == 0,
                "complex folds need a power-of-two nbins (%d)", k.nbins);
  prune_children_kernel<kComplex, kCheby><<<(unsigned)items_grid(n_child),
                                            kCBlk, lds,
                                            pyl::as_stream(stream)>>>(
      k, leaves, folds, n_par, nullptr, branch, counts, off, n_child, nullptr, seg, c_parent,
      c_row, c_shift, c_phase, c_score, minmax);
  PYL_CHECK_LAUNCH("prune_children_kernel");
  return PYL_OK;
}

extern "C" int pyl_prune_children(const pyl_prune_params* p, const pyl_prune_coords* c,
                                  const double* leaves, const float* folds, int64_t n_par,
                                  const double* branch, const int32_t* counts,
                                  const int64_t* off, int64_t n_child, const float* seg,
                                  int32_t* c_parent, int32_t* c_row, int32_t* c_shift,
                                  double* c_phase, float* c_score, float* minmax,
                                  void* stream) {
  return prune_children<false>(p, c, leaves, folds, n_par, branch, counts, off, n_child, seg,
                               c_parent, c_row, c_shift, c_phase, c_score, minmax, stream);
}

extern "C" int pyl_prune_children_complex(const pyl_prune_params* p, const pyl_prune_coords* c,
                                          const double* leaves, const float* folds,
                                          int64_t n_par, const double* branch,
                                          const int32_t* counts, const int64_t* off,
                                          int64_t n_child, const float* seg, int32_t* c_parent,
                                          int32_t* c_row, int32_t* c_shift, double* c_phase,
                                          float* c_score, float* minmax, void* stream) {
  return prune_children<true>(p, c, leaves, folds, n_par, branch, counts, off, n_child, seg,
                              c_parent, c_row, c_shift, c_phase, c_score, minmax, stream);
}

// kCheby only changes the level constants: tnext is then the Chebyshev transfer matrix.
template <bool kComplex, bool kCheby = false>
static int prune_materialize(const pyl_prune_params* p, const pyl_prune_coords* c,
                             const double* leaves, const float* folds, const double* branch,
                             const int32_t* counts, const int64_t* off, const float* seg,
                             const int32_t* c_parent, const int32_t* c_row,
                             const int32_t* c_shift, const double* c_phase,
                             const float* c_score, const int64_t* surv, int64_t n_surv,
                             double* o_leaves, float* o_folds, float* o_scores, int32_t* o_bt,
                             void* stream) {
  LevelConst k;
  int rc = kCheby ? build_cheb_const(p, c, &k) : build_const(p, c, &k);
  if (rc != PYL_OK) return rc;
  PYL_REQUIRE(c != nullptr, "null pyl_prune_coords");
  if (n_surv == 0) return PYL_OK;
  PYL_REQUIRE(leaves && folds && branch && counts && off && seg && c_parent && c_row &&
                  c_shift && c_phase && c_score && surv && o_leaves && o_folds && o_scores &&
                  o_bt,
              "null device pointer");
  prune_materialize_kernel<kComplex><<<(unsigned)items_grid(n_surv), kCBlk,
                                       materialize_lds(kComplex, k.nbins),
                                       pyl::as_stream(stream)>>>(
      k, leaves, folds, branch, counts, off, seg, c_parent, c_row, c_shift, c_phase, c_score,
      surv, n_surv, nullptr, o_leaves, o_folds, o_scores, o_bt, nullptr);
  PYL_CHECK_LAUNCH("prune_materialize_kernel");
  return PYL_OK;
}

extern "C" int pyl_prune_materialize(const pyl_prune_params* p, const pyl_prune_coords* c,
                                     const double* leaves, const float* folds,
                                     const double* branch, const int32_t* counts,
                                     const int64_t* off, const float* seg,
                                     const int32_t* c_parent, const int32_t* c_row,
                                     const int32_t* c_shift, const double* c_phase,
                                     const float* c_score, const int64_t* surv, int64_t n_surv,
                                     double* o_leaves, float* o_folds, float* o_scores,
                                     int32_t* o_bt, void* stream) {
  return prune_materialize<false>(p, c, leaves, folds, branch, counts, off, seg, c_parent,
                                  c_row, c_shift, c_phase, c_score, surv, n_surv, o_leaves,
                                  o_folds, o_scores, o_bt, stream);
}

extern "C" int pyl_prune_materialize_complex(
    const pyl_prune_params* p, const pyl_prune_coords* c, const double* leaves,
    const float* folds, const double* branch, const int32_t* counts, const int64_t* off,
    const float* seg, const int32_t* c_parent, const int32_t* c_row, const int32_t* c_shift,
    const double* c_phase, const float* c_score, const int64_t* surv, int64_t n_surv,
    double* o_leaves, float* o_folds, float* o_scores, int32_t* o_bt, void* stream) {
  return prune_materialize<true>(p, c, leaves, folds, branch, counts, off, seg, c_parent,
                                 c_row, c_shift, c_phase, c_score, surv, n_surv, o_leaves,
                                 o_folds, o_scores, o_bt, stream);
}

template <bool kComplex, bool kCheby = false>
static int prune_ascend(const pyl_prune_params* p, const double* leaves, int64_t n,
                        const float* dyp, int64_t ncoords, const int64_t* seg_idx,
                        const double* seg_t0, int32_t nseg, double t0_cur, float* o_folds,
                        float* o_scores, void* stream, double scale_cur = 1.0) {
  LevelConst k;
  int rc = kCheby ? build_cheb_const(p, nullptr, &k) : build_const(p, nullptr, &k);
  if (rc != PYL_OK) return rc;
  PYL_REQUIRE(nseg >= 1, "no segments to integrate");
  PYL_REQUIRE(leaves && dyp && seg_idx && seg_t0 && o_folds && o_scores, "null pointer");
  const int nn = k.P + 1;
  std::vector<double> cseg;
  if constexpr (kCheby) {
    PYL_REQUIRE(scale_cur > 0, "ts1 and ts2 must be positive.");
    for (int m = 0; m < nn; ++m) k.tspow[m] = pyl::powi(scale_cur, m);
    cseg.resize((size_t)nseg * nn * nn);
    for (int q = 0; q < nseg; ++q)
      cheb_matrix_t(k.P, t0_cur, scale_cur, seg_t0[q], scale_cur, cseg.data() + (size_t)q * nn * nn);
  }
  if (n == 0) return PYL_OK;
  const size_t lds = ep_block_lds<kComplex>(k);
  PYL_REQUIRE(lds <= kMaxEpLds, "nbins too large for the EP scorer (%d)", k.nbins);
  if constexpr (kComplex)
    PYL_REQUIRE((k.nbins & (k.nbins - 1)) == 0,
                "complex folds need a power-of-two nbins (%d)", k.nbins);
  hipStream_t s = pyl::as_stream(stream);
  // Small per-call tables: segment indices/times and per-(survivor, segment) row/phase.
  int64_t* d_idx = nullptr;
  double* d_t0 = nullptr;
  int32_t* d_row = nullptr;
  double* d_ph = nullptr;
  PYL_CHECK_HIP(hipMallocAsync(reinterpret_cast<void**>(&d_idx), sizeof(int64_t) * nseg, s));
  PYL_CHECK_HIP(hipMallocAsync(reinterpret_cast<void**>(&d_t0), sizeof(double) * nseg, s));
  PYL_CHECK_HIP(hipMallocAsync(reinterpret_cast<void**>(&d_row), sizeof(int32_t) * n * nseg, s));
  PYL_CHECK_HIP(hipMallocAsync(reinterpret_cast<void**>(&d_ph), sizeof(double) * n * nseg, s));
  PYL_CHECK_HIP(hipMemcpyAsync(d_idx, seg_idx, sizeof(int64_t) * nseg, hipMemcpyHostToDevice, s));
  PYL_CHECK_HIP(hipMemcpyAsync(d_t0, seg_t0, sizeof(double) * nseg, hipMemcpyHostToDevice, s));
  double* d_c = nullptr;
  if constexpr (kCheby) {
    PYL_CHECK_HIP(hipMallocAsync(reinterpret_cast<void**>(&d_c), sizeof(double) * cseg.size(), s));
    PYL_CHECK_HIP(hipMemcpyAsync(d_c, cseg.data(), sizeof(double) * cseg.size(),
                                 hipMemcpyHostToDevice, s));
  }
  prune_ascend_kernel<kComplex, kCheby><<<(unsigned)pyl::ceil_div(n, kBlk), kCBlk, lds, s>>>(
      k, leaves, n, dyp, ncoords, d_idx, d_t0, d_c, nseg, t0_cur, d_row, d_ph, o_folds,
      o_scores);
  PYL_CHECK_LAUNCH("prune_ascend_kernel");
  if (d_c) PYL_CHECK_HIP(hipFreeAsync(d_c, s));
  PYL_CHECK_HIP(hipFreeAsync(d_idx, s));
  PYL_CHECK_HIP(hipFreeAsync(d_t0, s));
  PYL_CHECK_HIP(hipFreeAsync(d_row, s));
  PYL_CHECK_HIP(hipFreeAsync(d_ph, s));
  return PYL_OK;
}

extern "C" int pyl_prune_ascend(const pyl_prune_params* p, const double* leaves, int64_t n,
                                const float* dyp, int64_t ncoords, const int64_t* seg_idx,
                                const double* seg_t0, int32_t nseg, double t0_cur,
                                float* o_folds, float* o_scores, void* stream) {
  return prune_ascend<false>(p, leaves, n, dyp, ncoords, seg_idx, seg_t0, nseg, t0_cur,
                             o_folds, o_scores, stream);
}

extern "C" int pyl_prune_ascend_complex(const pyl_prune_params* p, const double* leaves,
                                        int64_t n, const float* dyp, int64_t ncoords,
                                        const int64_t* seg_idx, const double* seg_t0,
                                        int32_t nseg, double t0_cur, float* o_folds,
                                        float* o_scores, void* stream) {
  return prune_ascend<true>(p, leaves, n, dyp, ncoords, seg_idx, seg_t0, nseg, t0_cur,
                            o_folds, o_scores, stream);
}

// ---- Chebyshev basis (dyn_poly_cheby.py): the Taylor level kernels with the Chebyshev
// level constants, on the parents moved to coord_cur.
extern "C" int pyl_cheb_branch(const pyl_prune_params* p, const pyl_prune_coords* c,
                               const double* leaves, int64_t n_par, double* leaves_cur,
                               double* branch, int32_t* counts, int64_t* nchild, int64_t* err,
                               void* stream) {
  LevelConst k;
  PYL_REQUIRE(c != nullptr, "null pyl_prune_coords");
  int rc = build_cheb_const(p, c, &k);
  if (rc != PYL_OK) return rc;
  PYL_REQUIRE(leaves && leaves_cur && branch && counts && nchild && err, "null device pointer");
  if (n_par == 0) return PYL_OK;
  const int threads = 128;
  const unsigned grid = (unsigned)pyl::ceil_div(n_par, threads);
  hipStream_t s = pyl::as_stream(stream);
  cheb_parents_kernel<<<grid, threads, 0, s>>>(k, leaves, n_par, nullptr, leaves_cur);
  PYL_CHECK_LAUNCH("cheb_parents_kernel");
  prune_branch_kernel<<<grid, threads, 0, s>>>(k, leaves_cur, n_par, nullptr, branch, counts,
                                               nchild,
                                               reinterpret_cast<unsigned long long*>(err));
  PYL_CHECK_LAUNCH("prune_branch_kernel");
  return PYL_OK;
}

extern "C" int pyl_cheb_children(const pyl_prune_params* p, const pyl_prune_coords* c,
                                 const double* leaves_cur, const float* folds, int64_t n_par,
                                 const double* branch, const int32_t* counts,
                                 const int64_t* off, int64_t n_child, const float* seg,
                                 int32_t* c_parent, int32_t* c_row, int32_t* c_shift,
                                 double* c_phase, float* c_score, float* minmax, void* stream) {
  return prune_children<false, true>(p, c, leaves_cur, folds, n_par, branch, counts, off,
                                     n_child, seg, c_parent, c_row, c_shift, c_phase, c_score,
                                     minmax, stream);
}

extern "C" int pyl_cheb_children_complex(const pyl_prune_params* p, const pyl_prune_coords* c,
                                         const double* leaves_cur, const float* folds,
                                         int64_t n_par, const double* branch,
                                         const int32_t* counts, const int64_t* off,
                                         int64_t n_child, const float* seg, int32_t* c_parent,
                                         int32_t* c_row, int32_t* c_shift, double* c_phase,
                                         float* c_score, float* minmax, void* stream) {
  return prune_children<true, true>(p, c, leaves_cur, folds, n_par, branch, counts, off,
                                    n_child, seg, c_parent, c_row, c_shift, c_phase, c_score,
                                    minmax, stream);
}

extern "C" int pyl_cheb_materialize(const pyl_prune_params* p, const pyl_prune_coords* c,
                                    const double* leaves_cur, const float* folds,
                                    const double* branch, const int32_t* counts,
                                    const int64_t* off, const float* seg,
                                    const int32_t* c_parent, const int32_t* c_row,
                                    const int32_t* c_shift, const double* c_phase,
                                    const float* c_score, const int64_t* surv, int64_t n_surv,
                                    double* o_leaves, float* o_folds, float* o_scores,
                                    int32_t* o_bt, void* stream) {
  return prune_materialize<false, true>(p, c, leaves_cur, folds, branch, counts, off, seg,
                                        c_parent, c_row, c_shift, c_phase, c_score, surv,
                                        n_surv, o_leaves, o_folds, o_scores, o_bt, stream);
}

extern "C" int pyl_cheb_materialize_complex(
    const pyl_prune_params* p, const pyl_prune_coords* c, const double* leaves_cur,
    const float* folds, const double* branch, const int32_t* counts, const int64_t* off,
    const float* seg, const int32_t* c_parent, const int32_t* c_row, const int32_t* c_shift,
    const double* c_phase, const float* c_score, const int64_t* surv, int64_t n_surv,
    double* o_leaves, float* o_folds, float* o_scores, int32_t* o_bt, void* stream) {
  return prune_materialize<true, true>(p, c, leaves_cur, folds, branch, counts, off, seg,
                                       c_parent, c_row, c_shift, c_phase, c_score, surv, n_surv,
                                       o_leaves, o_folds, o_scores, o_bt, stream);
}

extern "C" int pyl_cheb_ascend(const pyl_prune_params* p, const double* leaves, int64_t n,
                               const float* dyp, int64_t ncoords, const int64_t* seg_idx,
                               const double* seg_t0, int32_t nseg, double t0_cur,
                               double scale_cur, float* o_folds, float* o_scores,
                               void* stream) {
  return prune_ascend<false, true>(p, leaves, n, dyp, ncoords, seg_idx, seg_t0, nseg, t0_cur,
                                   o_folds, o_scores, stream, scale_cur);
}

extern "C" int pyl_cheb_ascend_complex(const pyl_prune_params* p, const double* leaves,
                                       int64_t n, const float* dyp, int64_t ncoords,
                                       const int64_t* seg_idx, const double* seg_t0,
                                       int32_t nseg, double t0_cur, double scale_cur,
                                       float* o_folds, float* o_scores, void* stream) {
  return prune_ascend<true, true>(p, leaves, n, dyp, ncoords, seg_idx, seg_t0, nseg, t0_cur,
                                  o_folds, o_scores, stream, scale_cur);
}

// ---- The sync-free level loop ------------------------------------------------------
extern "C" int pyl_ep_timer_create(int32_t capacity, pyl_ep_timer** out) {
  PYL_REQUIRE(out != nullptr && capacity >= 1, "invalid timer arguments");
  auto* t = new pyl_ep_timer();
  t->capacity = capacity;
  t->ev.resize(2 * (size_t)capacity);
  t->phase.resize(capacity);
  for (auto& e : t->ev) {
    if (hipEventCreate(&e) != hipSuccess) {
      e = nullptr;
      pyl_ep_timer_destroy(t);
      return pyl::fail(PYL_ERR_HIP, "hipEventCreate failed");
    }
  }
  *out = t;
  return PYL_OK;
}

extern "C" int pyl_ep_timer_destroy(pyl_ep_timer* t) {
  if (t == nullptr) return PYL_OK;
  for (auto e : t->ev)
    if (e) (void)hipEventDestroy(e);
  delete t;
  return PYL_OK;
}

extern "C" int pyl_ep_timer_read(pyl_ep_timer* t, int32_t* phase_out, float* ms_out,
                                 int32_t cap, int32_t* n_out) {
  PYL_REQUIRE(t != nullptr && phase_out != nullptr && ms_out != nullptr && n_out != nullptr,
              "null timer argument");
  const int32_t n = std::min(t->n, cap);
  for (int32_t i = 0; i < n; ++i) {
    if (hipEventSynchronize(t->ev[2 * i + 1]) != hipSuccess)
      return pyl::fail(PYL_ERR_HIP, "hipEventSynchronize failed");
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, t->ev[2 * i], t->ev[2 * i + 1]) != hipSuccess)
      return pyl::fail(PYL_ERR_HIP, "hipEventElapsedTime failed");
    phase_out[i] = t->phase[i];
    ms_out[i] = ms;
  }
  *n_out = n;
  t->n = 0;
  return PYL_OK;
}

namespace {

template <bool kComplex, bool kCheby>
int ep_levels(const pyl_prune_params* p, const pyl_prune_coords* coords, const double* thr,
              const int64_t* seg_idx, int32_t nlev, const float* fold, int64_t seg_stride,
              const double* in_leaves, const float* in_folds, int64_t n_in, int64_t hist_row0,
              int32_t first_out, const pyl_ep_loop* L, hipStream_t s) {
  PYL_REQUIRE(L != nullptr && coords != nullptr && thr != nullptr && seg_idx != nullptr,
              "null host argument");
  PYL_REQUIRE(nlev >= 1, "nlev must be >= 1");
  PYL_REQUIRE(first_out == 0 || first_out == 1, "first_out must be 0 or 1");
  PYL_REQUIRE(L->cap_par >= 1 && L->cap_child >= 1 && L->cap_hist >= 0 && L->max_sugg >= 1,
              "invalid capacities");
  PYL_REQUIRE(L->max_sugg <= L->cap_par, "max_sugg exceeds the tree buffers");
  PYL_REQUIRE(n_in >= 0 && n_in <= L->cap_par, "input tree exceeds the tree buffers");
  PYL_REQUIRE(hist_row0 >= 0 && hist_row0 <= L->cap_hist, "hist_row0 out of range");
  PYL_REQUIRE(fold && in_leaves && in_folds && L->leaves[0] && L->leaves[1] && L->folds[0] &&
                  L->folds[1] && L->scores[0] && L->scores[1] && L->hist && L->branch &&
                  L->counts && L->nchild && L->off && L->c_parent && L->c_row &&
                  L->c_shift && L->c_phase && L->c_score && L->surv && L->part && L->rec &&
                  L->minmax && (!kCheby || L->leaves_cur),
              "null device pointer");
  const pyl_ep_budget* bud = L->budget;
  const int out0 = bud ? 0 : first_out;  // with a budget: materialise into 0, keep into 1
  PYL_REQUIRE(in_leaves != L->leaves[out0] && in_folds != L->folds[out0],
              "the input tree is the first output buffer");
  PYL_REQUIRE(!bud || bud->bt_tmp, "null budget buffer");
  PYL_REQUIRE((int64_t)p->n_accel * p->n_freq <= 0x7fffffff, "FFA grid too large");
  ep_records_init_kernel<<<(unsigned)pyl::ceil_div(nlev + 1, 64), 64, 0, s>>>(
      L->rec, L->minmax, nlev, n_in, hist_row0);
  PYL_CHECK_LAUNCH("ep_records_init_kernel");
  const int tpar = 128;
  unsigned g_par = 0, g_child = 0, g_surv = 0;
  for (int l = 0; l < nlev; ++l) {
    LevelConst k;
    int rc = kCheby ? build_cheb_const(p, &coords[l], &k) : build_const(p, &coords[l], &k);
    if (rc != PYL_OK) return rc;
    const size_t lds = ep_block_lds<kComplex>(k);
    if (l == 0) {  // grids for the capacities: blocks past a level's count exit at once
      g_par = strided_grid(L->cap_par, tpar);
      g_child = items_grid_cap(L->cap_child);
      // Survivors are children of the level: the grid covers the smaller capacity.
      g_surv = items_grid_cap(std::min(L->max_sugg, L->cap_child));
    }
    PYL_REQUIRE(lds <= kMaxEpLds, "nbins too large for the EP scorer (%d)", k.nbins);
    if constexpr (kComplex)
      PYL_REQUIRE((k.nbins & (k.nbins - 1)) == 0,
                  "complex folds need a power-of-two nbins (%d)", k.nbins);
    PYL_REQUIRE(seg_idx[l] >= 0, "negative segment index");
    int64_t* rec = L->rec + (int64_t)l * kRecWords;
    const int in_buf = bud ? 1 : (first_out + l - 1) & 1;
    const double* leaves = l == 0 ? in_leaves : L->leaves[in_buf];
    const float* folds = l == 0 ? in_folds : L->folds[in_buf];
    const int o = bud ? 0 : (first_out + l) & 1;
    const float* seg = fold + seg_idx[l] * seg_stride;
    const double* src = leaves;
    pyl_ep_timer* tm = L->timer;
    if (tm) tm->begin(PYL_EP_PH_BRANCH, s);
    if constexpr (kCheby) {
      cheb_parents_kernel<<<g_par, tpar, 0, s>>>(k, leaves, 0, rec + kRecNpar, L->leaves_cur);
      PYL_CHECK_LAUNCH("cheb_parents_kernel");
      src = L->leaves_cur;
    }
    prune_branch_kernel<<<g_par, tpar, 0, s>>>(
        k, src, 0, rec + kRecNpar, L->branch, L->counts, L->nchild,
        reinterpret_cast<unsigned long long*>(rec + kRecErr));
    PYL_CHECK_LAUNCH("prune_branch_kernel");
    chunked_scan(CountSrc{L->nchild},
                 ChildOffSink{L->off, rec, L->cap_child, L->max_sugg, L->cap_hist},
                 rec + kRecNpar, L->part, s);
    PYL_CHECK_LAUNCH("chunk_scan_kernel");
    if (tm) tm->end(s);
    if (tm) tm->begin(PYL_EP_PH_CHILDREN, s);
    prune_children_kernel<kComplex, kCheby><<<g_child, kCBlk, lds, s>>>(
        k, src, folds, 0, rec + kRecNpar, L->branch, L->counts, L->off, 0, rec + kRecNchild,
        seg, L->c_parent, L->c_row, L->c_shift, L->c_phase, L->c_score, L->minmax + 2 * l);
    PYL_CHECK_LAUNCH("prune_children_kernel");
    if (tm) tm->end(s);
    if (tm) tm->begin(PYL_EP_PH_SELECT, s);
    chunked_scan(PassSrc{L->c_score, thr[l]},
                 SurvSink{L->surv, rec, L->max_sugg, L->cap_hist}, rec + kRecNchild, L->part,
                 s);
    PYL_CHECK_LAUNCH("chunk_scan_kernel");
    if (tm) tm->end(s);
    if (tm) tm->begin(PYL_EP_PH_MATERIALIZE, s);
    prune_materialize_kernel<kComplex><<<g_surv, kCBlk, materialize_lds(kComplex, k.nbins), s>>>(
        k, src, folds, L->branch, L->counts, L->off, seg, L->c_parent, L->c_row, L->c_shift,
        L->c_phase, L->c_score, L->surv, 0, rec + kRecWords + kRecNpar, L->leaves[o],
        L->folds[o], L->scores[o], bud ? bud->bt_tmp : L->hist, bud ? nullptr : rec + kRecHist);
    PYL_CHECK_LAUNCH("prune_materialize_kernel");
    if (tm) tm->end(s);
    if (bud) {
      if (tm) tm->begin(PYL_EP_PH_BUDGET, s);
      rc = pyl::ep_budget_level_enqueue(bud, k.P, bud->level0 + l, L->leaves[0],
                                        L->folds[0], L->scores[0], bud->bt_tmp, L->max_sugg,
                                        rec, L->leaves[1], L->folds[1], L->scores[1], L->hist,
                                        s);
      if (rc != PYL_OK) return rc;
      if (tm) tm->end(s);
    }
  }
  return PYL_OK;
}

}  // namespace

extern "C" int pyl_ep_levels(const pyl_prune_params* p, int32_t flags,
                             const pyl_prune_coords* coords, const double* thresholds,
                             const int64_t* seg_idx, int32_t nlev, const float* dyp_fold,
                             int64_t seg_stride, const double* in_leaves, const float* in_folds,
                             int64_t n_in, int64_t hist_row0, int32_t first_out,
                             const pyl_ep_loop* L, void* stream) {
  PYL_REQUIRE(p != nullptr, "null pyl_prune_params");
  PYL_REQUIRE((flags & ~(PYL_EP_COMPLEX | PYL_EP_CHEBYSHEV)) == 0, "unknown flags %d", flags);
  const hipStream_t s = pyl::as_stream(stream);
  const bool cplx = flags & PYL_EP_COMPLEX;
  const bool cheb = flags & PYL_EP_CHEBYSHEV;
  auto* f = cplx ? (cheb ? ep_levels<true, true> : ep_levels<true, false>)
                 : (cheb ? ep_levels<false, true> : ep_levels<false, false>);
  return f(p, coords, thresholds, seg_idx, nlev, dyp_fold, seg_stride, in_leaves, in_folds,
           n_in, hist_row0, first_out, L, s);
}
