// EP search in the circular basis (core/circular.py) on the GPU: one pruning level of
// prune.py:165-332 with the circular plugin (dyn_circular_taylor.py:484-678), plus ascend.
//
// Leaves have 7 rows: [crackle, snap, jerk, accel, vel, d0, (f0, basis flag)].
//
// Child enumeration. The reference branches every parameter but crackle (odometer order,
// crackle held at the parent value), then, inside each batch of parents, keeps the children
// that are not in the snap-accel "hole" in that order and appends, for every hole child
// whose crackle still needs a split, its crackle children. Here:
//   circ_branch_kernel   thread per parent: counts / first centre / step of snap..vel as in
//                        the Taylor branch; crackle held (step already the split step when
//                        a split is needed) and its split kept (first centre, count);
//   circ_stage1_kernel   thread per held child: parent and the expand decision;
//   (caller)             per-batch prefix sums -> final position of every held child;
//   circ_expand_kernel   thread per held child, looping over its outputs: held child and
//                        crackle index of every final child, validation and region;
//   circ_children_kernel thread per validated child: resolve with the region's propagator,
//                        then shift-add + score as in the Taylor level;
//   circ_materialize_kernel thread per survivor: basis flag, transform, backtrack, fold.
//
// The circular propagator (transforms.py:248-313) uses the device's FP64 cos/sin; the
// other operations are the reference's, in its order, with no contraction.
#include "pyl_budget.hpp"
#include "pyl_loop.hpp"
#include "pyl_prune_common.hpp"

namespace {

using namespace pyl_ep;

constexpr int kP = 5;      // crackle, snap, jerk, accel, vel
constexpr int kN = kP + 1;  // with d0
constexpr int kRows = kP + 2;

struct CircConst {
  double prop_sig, val_sig, omega_max_sq, x_mass_const;
  int64_t batch_size;
  double dt_add, dt_init, dt_next;  // t0_add - t0_cur, t0_init - t0_cur, t0_next - t0_cur
};

int build_circ(const pyl_prune_params* p, const pyl_circ_params* q, const pyl_prune_coords* c,
               LevelConst* k, CircConst* cc) {
  PYL_REQUIRE(p != nullptr && q != nullptr, "null parameters");
  PYL_REQUIRE(p->poly_order == kP, "the circular basis needs poly_order 5, got %d",
              p->poly_order);
  PYL_REQUIRE(p->nparams_ffa == kP, "the circular basis needs a 5-parameter FFA grid");
  PYL_REQUIRE(q->batch_size >= 1, "batch_size must be >= 1");
  PYL_REQUIRE(q->prop_sig > 0 && q->val_sig > 0, "significances must be positive");
  int rc = build_const(p, c, k);
  if (rc != PYL_OK) return rc;
  cc->prop_sig = q->prop_sig;
  cc->val_sig = q->val_sig;
  cc->omega_max_sq = q->omega_max_sq;
  cc->x_mass_const = q->x_mass_const;
  cc->batch_size = q->batch_size;
  cc->dt_add = c ? c->t0_add - c->t0_cur : 0.0;
  cc->dt_init = c ? c->t0_init - c->t0_cur : 0.0;
  cc->dt_next = c ? c->t0_next - c->t0_cur : 0.0;
  return PYL_OK;
}

// circular.py:13-79: region of a cell (values v, errors e over crackle..accel).
__device__ __forceinline__ int8_t circ_region(const double* v, const double* e, double sig) {
  const bool sig_s = fabs(v[1]) > sig * e[1];
  const bool sig_a = fabs(v[3]) > sig * e[3];
  const bool sig_j = fabs(v[2]) > sig * e[2];
  if (sig_s && sig_a && (-v[1]) * v[3] > 0.0) return PYL_REGION_SNAP;
  if (sig_j && (-v[0]) * v[2] > 0.0) return PYL_REGION_CRACKLE;
  return PYL_REGION_TAYLOR;
}

// circular.py:82-124: false when validation removes the cell.
__device__ __forceinline__ bool circ_valid(const CircConst& q, const double* v,
                                           const double* e) {
  const bool sig_s = fabs(v[1]) > q.val_sig * e[1];
  const bool sig_a = fabs(v[3]) > q.val_sig * e[3];
  if (!(sig_s && sig_a)) return true;  // the hole and unresolved cells are never validated
  if (!((-v[1]) * v[3] > 0.0)) return false;
  const double osq = (-v[1]) / v[3];
  const double lim = q.x_mass_const * pow(osq, 2.0 / 3.0);
  return osq <= q.omega_max_sq && fabs(v[3]) <= lim;
}

// transforms.py:248-313: values [d5..d0] moved by dt along the circular orbit pinned by
// -d4/d2 (snap region) or -d5/d3 (hole).
__device__ __forceinline__ void circ_shift(const double* v, double dt, bool in_hole,
                                           double* o) {
  const double d5 = v[0], d4 = v[1], d3 = v[2], d2 = v[3], d1 = v[4], d0 = v[5];
  const double osq = in_hole ? (-d5) / d3 : (-d4) / d2;
  const double om = sqrt(osq);
  const double odt = om * dt;
  const double cs = cos(odt);
  const double sn = sin(odt);
  const double d2j = d2 * cs + (d3 / om) * sn;
  const double d3j = d3 * cs - (d2 * om) * sn;
  o[3] = d2j;
  o[2] = d3j;
  o[1] = (-osq) * d2j;
  o[0] = (-osq) * d3j;
  const double d1ci = (-d3) / osq;
  const double d1cj = (-d3j) / osq;
  o[4] = d1cj + (d1 - d1ci);
  const double d0cj = (-d2j) / osq;
  const double d0ci = (-d2) / osq;
  o[5] = d0cj + (d0 - d0ci) + (d1 - d1ci) * dt;
}

// Values moved by dt with the region's propagator; T = Taylor translation by dt.
__device__ __forceinline__ void region_shift(int8_t reg, const double* v, double dt,
                                             const double* T, double* o) {
  if (reg == PYL_REGION_TAYLOR) {
    for (int i = 0; i < kN; ++i) o[i] = mat_row(T, v, kN, i);
  } else {
    circ_shift(v, dt, reg == PYL_REGION_CRACKLE, o);
  }
}

// A final child: held child s1 (parent par) with crackle index sub (-1: held crackle).
__device__ __forceinline__ void circ_decode(const LevelConst& k, const double* __restrict__ leaves,
                                            const pyl_circ_level& L, int64_t par, int64_t s1,
                                            int32_t sub, ChildLeaf& ch) {
  const double* lf = leaves + par * kRows * 2;
  const double* br = L.branch + par * kP * 2;
  decode_child(k, lf, br, L.counts + par * kP, s1 - L.off1[par], ch);
  if (sub >= 0) ch.val[0] = L.cr_first[par] + (double)sub * br[1];
}

// Item counts come from the level struct or, when the pointer is not null, from device
// memory (the sync-free level loop, whose launches are sized for the buffers' capacity).
__device__ __forceinline__ int64_t dev_count(int64_t n, const int64_t* n_dev) {
  return n_dev ? *n_dev : n;
}

// ---- Forests: several trees of one configuration in one run of the level loop ---------
// A level's parents (and so its held children, batches, children and survivors) are the
// trees' ones one after the other. Per level, the tree tables hold, for tree t, the offsets
// po (parents), bo (batches), cb (children), vb (validated children), so (survivors) and
// ko (kept rows, the next level's parents), each [T + 1]. Batches never straddle trees:
// tree t's batches are the reference's batches of its own parents.
constexpr int kTabArrays = 6;
struct FTab {
  const int64_t* tab;   // this level's tables; null: a single tree
  int32_t T;
  int64_t row_stride;   // segment rows between consecutive trees' FFA folds
  float* minmax;        // this level's per-tree [T][2] (forest only)
};

// Largest t in [0, T) with off[t] <= x (off non-decreasing): the tree holding item x.
__device__ __forceinline__ int f_tree(const int64_t* off, int T, int64_t x) {
  int lo = 0, hi = T;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (off[mid] <= x) lo = mid;
    else hi = mid;
  }
  return lo;
}

// Batch of parent par, and the parent range of batch b.
__device__ __forceinline__ int64_t f_batch(const FTab& f, int64_t par, int64_t bs) {
  if (!f.tab) return par / bs;
  const int64_t* po = f.tab;
  const int64_t* bo = f.tab + (f.T + 1);
  const int t = f_tree(po, f.T, par);
  return bo[t] + (par - po[t]) / bs;
}
__device__ __forceinline__ void f_batch_parents(const FTab& f, int64_t b, int64_t bs,
                                                int64_t n_par, int64_t* p0, int64_t* p1) {
  if (!f.tab) {
    *p0 = b * bs;
    *p1 = std::min<int64_t>(*p0 + bs, n_par);
    return;
  }
  const int64_t* po = f.tab;
  const int64_t* bo = f.tab + (f.T + 1);
  const int t = f_tree(bo, f.T, b);
  *p0 = po[t] + (b - bo[t]) * bs;
  *p1 = std::min<int64_t>(*p0 + bs, po[t + 1]);
}
// Segment row of a child of parent par in the stacked trees' folds.
__device__ __forceinline__ int32_t f_row(const FTab& f, int64_t par, int32_t row) {
  if (!f.tab) return row;
  return row + (int32_t)(f_tree(f.tab, f.T, par) * f.row_stride);
}

// Rows a forest's materialisation writes when a budget's cut comes first: output row r
// is survivor kept[SO[t] + r - KO[t]] of tree t (KO[t] <= r < KO[t + 1]); without kept,
// survivor r.
struct KMap {
  const int64_t* kept;
  const int64_t* tab;
  int32_t T;
  __device__ __forceinline__ int64_t operator()(int64_t r) const {
    if (!kept) return r;
    const int64_t* so = tab + PYL_EP_TAB_SO * (T + 1);
    const int64_t* ko = tab + PYL_EP_TAB_KO * (T + 1);
    const int t = f_tree(ko, T, r);
    return kept[so[t] + r - ko[t]];
  }
};

// Per-tree score minimum / maximum of a block's children (one atomic pair per wavefront
// when its children are of one tree, else per child).
__device__ __forceinline__ void forest_minmax(float sc, bool valid, int t, float* mm) {
  int tlo = valid ? t : INT_MAX, thi = valid ? t : -1;
  for (int o = 32; o > 0; o >>= 1) {
    tlo = min(tlo, __shfl_xor(tlo, o));
    thi = max(thi, __shfl_xor(thi, o));
  }
  if (thi < 0) return;  // no child in the wavefront
  if (tlo == thi) {
    float mn = valid ? sc : FLT_MAX;
    float mx = valid ? sc : -FLT_MAX;
    for (int o = 32; o > 0; o >>= 1) {
      mn = fminf(mn, __shfl_xor(mn, o));
      mx = fmaxf(mx, __shfl_xor(mx, o));
    }
    if ((threadIdx.x & 63) == 0) {
      atomic_min_f32(&mm[2 * tlo], mn);
      atomic_max_f32(&mm[2 * tlo + 1], mx);
    }
  } else if (valid) {
    atomic_min_f32(&mm[2 * t], sc);
    atomic_max_f32(&mm[2 * t + 1], sc);
  }
}

// Per-item bodies of the thread-per-item kernels (defined below the kernels).
__device__ __forceinline__ void circ_branch_one(const LevelConst& k, const CircConst& q,
                                                const double* __restrict__ leaves,
                                                const pyl_circ_level& L, int64_t i,
                                                int64_t* __restrict__ nchild,
                                                unsigned long long* __restrict__ err);
__device__ __forceinline__ void circ_stage1_one(const LevelConst& k, const CircConst& q,
                                                const double* __restrict__ leaves,
                                                const pyl_circ_level& L, int64_t n_par,
                                                int64_t s);
__device__ __forceinline__ void circ_expand_one(const LevelConst& k, const CircConst& q,
                                                const double* __restrict__ leaves,
                                                const pyl_circ_level& L, int64_t s);

__device__ __forceinline__ void circ_branch_one(const LevelConst& k, const CircConst& q,
                                                const double* __restrict__ leaves,
                                                const pyl_circ_level& L, int64_t i,
                                                int64_t* __restrict__ nchild,
                                                unsigned long long* __restrict__ err) {
  const double* lf = leaves + i * kRows * 2;
  const double f0 = lf[(kP + 1) * 2];
  // Failures in the order the reference raises them: per batch, parameters 1..4 of every
  // parent (circular.py:255-265), then crackle (:267-277).
  const int64_t b = i / q.batch_size;
  const int64_t ib = i - b * q.batch_size;
  int64_t total = 1;
  for (int j = 0; j < kP; ++j) {
    const double pc = lf[2 * j];
    const double dc = lf[2 * j + 1];
    const double dn = k.dpf[j] * pyl::kCVal / f0;
    const double shift = fabs(dc - dn) * (k.fac[j] * (f0 / pyl::kCVal));
    const int64_t key = ((b * 2 + (j == 0 ? 1 : 0)) * q.batch_size + ib) * kP + j;
    int32_t cnt = 1;
    double first = pc, step = dc;
    if (dc <= 0.0 || dn <= 0.0) {
      atomicMin(err, (unsigned long long)(key * 4 + 2));
    } else {
      double nump = ceil(dc / dn - pyl::kFloatEps);
      if (nump < 1.0) nump = 1.0;
      if (nump > (double)k.branch_max) {
        atomicMin(err, (unsigned long long)(key * 4 + 1));
        nump = 1.0;
      }
      cnt = (int32_t)nump;
      step = dc / (double)cnt;
      first = (pc - dc / 2.0) + step / 2.0;
    }
    const bool split = !(shift < (k.eta - pyl::kFloatEps));
    if (j == 0) {
      // circular.py:267-285: held at the parent value; the step is the split step when
      // crackle needs a split, else the parent's.
      L.cr_first[i] = first;
      L.cr_num[i] = split ? cnt : 0;
      L.branch[(i * kP) * 2] = pc;
      L.branch[(i * kP) * 2 + 1] = split ? step : dc;
      L.counts[i * kP] = -1;
      continue;
    }
    if (!split) {
      cnt = -1;
      first = pc;
      step = dc;
    }
    L.branch[(i * kP + j) * 2] = first;
    L.branch[(i * kP + j) * 2 + 1] = step;
    L.counts[i * kP + j] = cnt;
    total *= cnt < 0 ? 1 : cnt;
  }
  nchild[i] = total;
}

__global__ void circ_branch_kernel(LevelConst k, CircConst q, const double* __restrict__ leaves,
                                   pyl_circ_level L, const int64_t* __restrict__ n_dev,
                                   int64_t* __restrict__ nchild,
                                   unsigned long long* __restrict__ err) {
  const int64_t n = dev_count(L.n_par, n_dev);
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    circ_branch_one(k, q, leaves, L, i, nchild, err);
}

__global__ void circ_stage1_kernel(LevelConst k, CircConst q, const double* __restrict__ leaves,
                                   pyl_circ_level L, const int64_t* __restrict__ npar_dev,
                                   const int64_t* __restrict__ n_dev) {
  const int64_t n = dev_count(L.n_s1, n_dev);
  const int64_t n_par = dev_count(L.n_par, npar_dev);
  for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < n;
       s += (int64_t)gridDim.x * blockDim.x)
    circ_stage1_one(k, q, leaves, L, n_par, s);
}

__device__ __forceinline__ void circ_stage1_one(const LevelConst& k, const CircConst& q,
                                                const double* __restrict__ leaves,
                                                const pyl_circ_level& L, int64_t n_par,
                                                int64_t s) {
  const int64_t par = find_parent(L.off1, n_par, s);
  ChildLeaf ch;
  circ_decode(k, leaves, L, par, s, -1, ch);
  // circular.py:169-201: the snap-accel system is not circular but jerk is significant.
  const bool sig_s = fabs(ch.val[1]) > q.prop_sig * ch.err[1];
  const bool sig_a = fabs(ch.val[3]) > q.prop_sig * ch.err[3];
  const bool sig_j = fabs(ch.val[2]) > q.prop_sig * ch.err[2];
  const bool snap = sig_s && sig_a && (-ch.val[1]) * ch.val[3] > 0.0;
  L.s1_parent[s] = (int32_t)par;
  L.s1_expand[s] = (!snap && sig_j && L.cr_num[par] > 0) ? 1 : 0;
}

__global__ void circ_expand_kernel(LevelConst k, CircConst q, const double* __restrict__ leaves,
                                   pyl_circ_level L, const int64_t* __restrict__ n_dev) {
  const int64_t n = dev_count(L.n_s1, n_dev);
  for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < n;
       s += (int64_t)gridDim.x * blockDim.x)
    circ_expand_one(k, q, leaves, L, s);
}

__device__ __forceinline__ void circ_expand_one(const LevelConst& k, const CircConst& q,
                                                const double* __restrict__ leaves,
                                                const pyl_circ_level& L, int64_t s) {
  const int64_t par = L.s1_parent[s];
  const bool expand = L.s1_expand[s] != 0;
  const int32_t nout = expand ? L.cr_num[par] : 1;
  const int64_t pos = L.s1_pos[s];
  for (int32_t sub = 0; sub < nout; ++sub) {
    const int64_t c = pos + sub;
    ChildLeaf ch;
    circ_decode(k, leaves, L, par, s, expand ? sub : -1, ch);
    L.c_s1[c] = s;
    L.c_sub[c] = expand ? sub : -1;
    L.c_keep[c] = circ_valid(q, ch.val, ch.err) ? 1 : 0;
    L.c_region[c] = circ_region(ch.val, ch.err, q.prop_sig);
  }
}

template <bool kComplex, bool kForest = false>
__global__ void __launch_bounds__(kCBlk)
circ_children_kernel(LevelConst k, CircConst q, const double* __restrict__ leaves,
                     const float* __restrict__ folds, pyl_circ_level L,
                     const int64_t* __restrict__ valid, int64_t n_valid_arg,
                     const int64_t* __restrict__ n_dev,
                     const float* __restrict__ seg, int32_t* __restrict__ c_parent,
                     int32_t* __restrict__ c_row, int32_t* __restrict__ c_shift,
                     double* __restrict__ c_phase, float* __restrict__ c_score,
                     float* __restrict__ minmax, uint8_t* __restrict__ has_taylor,
                     FTab f = FTab{}) {
  extern __shared__ float S[];  // [(B + wmax)][kStride]
  __shared__ int32_t s_parent[kBlk], s_row[kBlk], s_shift[kBlk];
  const int tid = threadIdx.x;
  const int64_t n_valid = dev_count(n_valid_arg, n_dev);
  const int per = per_block(n_valid);
  const int64_t v0 = (int64_t)blockIdx.x * per;
  if (v0 >= n_valid) return;  // the whole block
  const int nvalid = (int)std::min<int64_t>(per, n_valid - v0);
  const int B = k.nbins;
  if (tid < nvalid) {
    const int64_t v = v0 + tid;
    const int64_t c = valid[v];
    const int64_t s1 = L.c_s1[c];
    const int64_t par = L.s1_parent[s1];
    ChildLeaf ch;
    circ_decode(k, leaves, L, par, s1, L.c_sub[c], ch);
    const int8_t reg = L.c_region[c];
    // The level loop's batch flags: a batch holds a Taylor-region child (cleared per level).
    if (has_taylor && reg == PYL_REGION_TAYLOR) {
      if constexpr (kForest) has_taylor[f_batch(f, par, q.batch_size)] = 1;
      else has_taylor[par / q.batch_size] = 1;
    }
    double da[kN], di[kN];
    region_shift(reg, ch.val, q.dt_add, k.tadd, da);
    region_shift(reg, ch.val, q.dt_init, k.tinit, di);
    if (k.fixed) {
      // circular.py:531-590: velocity and delay at t_add taken whole.
      di[kN - 2] = 0.0;
      di[kN - 1] = 0.0;
    }
    const double accel = da[kN - 3];
    const double vel = da[kN - 2] - di[kN - 2];
    const double freq = ch.f0 * (1.0 - vel / pyl::kCVal);
    const double delay = (da[kN - 1] - di[kN - 1]) / pyl::kCVal;
    const double phase = pyl::phase_idx(k.dt_add_init, ch.f0, B, delay);
    const int64_t ia = trunc_idx(accel, k.amin, k.sa, k.n_accel);
    const int64_t ifr = trunc_idx(freq, k.fmin, k.sf, k.n_freq);
    const int32_t row = (int32_t)(ia * k.n_freq + ifr);
    const int32_t sh = pyl::shift_int(phase, B);
    c_parent[v] = (int32_t)par;
    c_row[v] = row;
    c_shift[v] = sh;
    c_phase[v] = phase;
    s_parent[tid] = (int32_t)par;
    if constexpr (kForest) s_row[tid] = f_row(f, par, row);  // the tree's own fold
    else s_row[tid] = row;
    s_shift[tid] = sh;
    if constexpr (kComplex)
      cphasors(phase, B,
               reinterpret_cast<float2*>(reinterpret_cast<char*>(S) +
                                         complex_ph_offset(B, k.sc.wmax)) +
                   (size_t)tid * (B / 2 + 1));
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
    // Rows over the block's wavefronts, lane = bin.
#pragma unroll 4
    for (int r = tid >> 6; r < nvalid; r += kCBlk / 64) {
      const float* fr = folds + (int64_t)s_parent[r] * 2 * B;
      const float* sr = seg + (int64_t)s_row[r] * 2 * B;
      const int sh = s_shift[r];
      for (int b = tid & 63; b < B; b += 64) {
        int src = b - sh;
        if (src < 0) src += B;
        const float e = fr[b] + sr[src];
        const float vv = fr[B + b] + sr[B + src];
        S[b * kStride + r] = pyl::norm_ev(e, vv);
      }
    }
    __syncthreads();
  }
  const float sc = block_score_columns(k, Sc, nvalid, tid, kCBlk);
  if (tid < nvalid) c_score[v0 + tid] = sc;
  if constexpr (kForest) {
    const int t = tid < nvalid ? f_tree(f.tab, f.T, s_parent[tid]) : 0;
    forest_minmax(sc, tid < nvalid, t, f.minmax);
  } else {
    block_minmax(sc, tid < nvalid, minmax);
  }
}

template <bool kComplex, bool kForest = false>
__global__ void __launch_bounds__(kCBlk)
circ_materialize_kernel(LevelConst k, CircConst q, const double* __restrict__ leaves,
                        const float* __restrict__ folds, pyl_circ_level L,
                        const int64_t* __restrict__ valid, const float* __restrict__ seg,
                        const int32_t* __restrict__ c_parent, const int32_t* __restrict__ c_row,
                        const int32_t* __restrict__ c_shift, const double* __restrict__ c_phase,
                        const float* __restrict__ c_score,
                        const uint8_t* __restrict__ has_taylor, const int64_t* __restrict__ surv,
                        int64_t n_surv_arg, const int64_t* __restrict__ n_dev,
                        double* __restrict__ o_leaves,
                        float* __restrict__ o_folds, float* __restrict__ o_scores,
                        int32_t* __restrict__ o_bt_base, const int64_t* __restrict__ bt_row0,
                        FTab f = FTab{}, KMap km = KMap{}) {
  __shared__ int64_t s_v[kBlk];
  __shared__ int32_t s_fpar[kBlk], s_frow[kBlk], s_fsh[kBlk];  // real folds
  extern __shared__ float2 s_ph[];  // complex: [kBlk][B/2 + 1] phasors, then rows, parents
  const int tid = threadIdx.x;
  const int64_t n_surv = dev_count(n_surv_arg, n_dev);
  const int per = per_block(n_surv);
  const int64_t s0 = (int64_t)blockIdx.x * per;
  if (s0 >= n_surv) return;  // the whole block
  int32_t* o_bt = o_bt_base + (bt_row0 ? *bt_row0 : 0) * kRows;
  const int nvalid = (int)std::min<int64_t>(per, n_surv - s0);
  const int B = k.nbins;
  if (tid < nvalid) {
    const int64_t s = s0 + tid;
    const int64_t v = surv[km(s)];
    s_v[tid] = v;
    const int64_t c = valid[v];
    const int64_t s1 = L.c_s1[c];
    const int64_t par = c_parent[v];
    ChildLeaf ch;
    circ_decode(k, leaves, L, par, s1, L.c_sub[c], ch);
    const int8_t reg = L.c_region[c];
    double* ol = o_leaves + s * kRows * 2;
    if (k.fixed) {
      keep_leaf(k, ch, ol);  // no transform on the fixed grid
    } else if (reg == PYL_REGION_TAYLOR) {
      taylor_full(k, k.tnext, ch, ol);  // transforms.py:93-137
    } else {
      // transforms.py:384-432: values moved along the orbit, errors unchanged.
      double o[kN];
      circ_shift(ch.val, q.dt_next, reg == PYL_REGION_CRACKLE, o);
      for (int i = 0; i < kN; ++i) {
        ol[2 * i] = o[i];
        ol[2 * i + 1] = ch.err[i];
      }
    }
    // circular.py:478-511: snap -> 1, crackle -> 2, crackle reset to 0 when the batch
    // holds a Taylor cell; Taylor cells keep the parent's flag.
    // The fixed-grid resolve (circular.py:531-590) leaves the flag as it is.
    double flag = ch.flag;
    if (!k.fixed) {
      if (reg == PYL_REGION_SNAP) {
        flag = 1.0;
      } else if (reg == PYL_REGION_CRACKLE) {
        int64_t b;
        if constexpr (kForest) b = f_batch(f, par, q.batch_size);
        else b = par / q.batch_size;
        flag = has_taylor[b] ? 0.0 : 2.0;
      }
    }
    ol[2 * (kP + 1)] = ch.f0;
    ol[2 * (kP + 1) + 1] = flag;
    o_scores[s] = c_score[v];
    int32_t* bt = o_bt + s * kRows;
    if constexpr (kForest) bt[0] = (int32_t)(par - f.tab[f_tree(f.tab, f.T, par)]);  // in its tree
    else bt[0] = (int32_t)par;
    const int32_t row = c_row[v];
    for (int qq = 1; qq <= kP; ++qq) bt[qq] = 0;
    bt[kP - 1] = row / k.n_freq;
    bt[kP] = row % k.n_freq;
    bt[kP + 1] = (int32_t)c_phase[v];
    if constexpr (kComplex) cphasors(c_phase[v], B, s_ph + (size_t)tid * (B / 2 + 1));
  }
  if constexpr (kComplex) {
    __syncthreads();
    int32_t* s_rw = reinterpret_cast<int32_t*>(s_ph + (size_t)kBlk * (B / 2 + 1));
    int32_t* s_pa = s_rw + kBlk;
    if (tid < nvalid) {
      if constexpr (kForest) s_rw[tid] = f_row(f, c_parent[s_v[tid]], c_row[s_v[tid]]);
      else s_rw[tid] = c_row[s_v[tid]];
      s_pa[tid] = c_parent[s_v[tid]];
    }
    __syncthreads();
    cfold_rows(reinterpret_cast<const float2*>(seg), reinterpret_cast<const float2*>(folds),
               s_rw, s_pa, s_ph, nvalid, B,
               reinterpret_cast<float2*>(o_folds) + s0 * 2 * (B / 2 + 1), tid, kCBlk);
  } else {
    if (tid < nvalid) {
      const int64_t v = s_v[tid];
      s_fpar[tid] = c_parent[v];
      if constexpr (kForest) s_frow[tid] = f_row(f, c_parent[v], c_row[v]);
      else s_frow[tid] = c_row[v];
      s_fsh[tid] = c_shift[v];
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

template <bool kComplex>
__global__ void __launch_bounds__(kCBlk)
circ_ascend_kernel(LevelConst k, CircConst q, const double* __restrict__ leaves, int64_t n,
                   const float* __restrict__ dyp, int64_t ncoords,
                   const int64_t* __restrict__ segi, const double* __restrict__ segt,
                   int32_t nseg, double t0_cur, int32_t* __restrict__ scratch_row,
                   double* __restrict__ scratch_phase, float* __restrict__ o_folds,
                   float* __restrict__ o_scores, const int64_t* __restrict__ toff = nullptr,
                   int T = 1, int64_t tree_stride = 0) {
  // A forest (toff not null): row r belongs to tree t with toff[t] <= r < toff[t + 1],
  // whose FFA fold starts tree_stride * t floats into dyp.
  extern __shared__ float S[];
  const int tid = threadIdx.x;
  const int64_t s0 = (int64_t)blockIdx.x * kBlk;
  const int nvalid = (int)std::min<int64_t>(kBlk, n - s0);
  const int B = k.nbins;
  // circular.py:605-666: every segment so far, with the survivor's region propagator.
  if (tid < nvalid) {
    const double* lf = leaves + (s0 + tid) * kRows * 2;
    double v[kN], e[kN];
    for (int i = 0; i < kN; ++i) {
      v[i] = lf[2 * i];
      e[i] = lf[2 * i + 1];
    }
    const double f0 = lf[(kP + 1) * 2];
    const int8_t reg = circ_region(v, e, q.prop_sig);
    for (int s = 0; s < nseg; ++s) {
      const double dt = segt[s] - t0_cur;
      double T[kN * kN];
      for (int i = 0; i < kN; ++i)
        for (int j = 0; j < kN; ++j)
          T[i * kN + j] = j <= i ? pyl::powi(dt, i - j) / pyl::fact_d(i - j) : 0.0;
      double d[kN];
      region_shift(reg, v, dt, T, d);
      const double freq = f0 * (1.0 - d[kN - 2] / pyl::kCVal);
      const double delay = d[kN - 1] / pyl::kCVal;
      const double phase = pyl::phase_idx(dt, f0, B, delay);
      const int64_t ia = trunc_idx(d[kN - 3], k.amin, k.sa, k.n_accel);
      const int64_t ifr = trunc_idx(freq, k.fmin, k.sf, k.n_freq);
      scratch_row[(s0 + tid) * nseg + s] = (int32_t)(ia * k.n_freq + ifr);
      scratch_phase[(s0 + tid) * nseg + s] = phase;
    }
    if constexpr (kComplex) {
      // common.py:463-490: sum of the rotated complex rows in segment order, from zero.
      const int nb_f = B / 2 + 1;
      float2* X = reinterpret_cast<float2*>(S) + (size_t)tid * 2 * nb_f;
      for (int i = 0; i < 2 * nb_f; ++i) X[i] = make_float2(0.0f, 0.0f);
      const float2* dc = reinterpret_cast<const float2*>(
          dyp + (toff ? f_tree(toff, T, s0 + tid) * tree_stride : 0));
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
      const float* dr = dyp + (toff ? f_tree(toff, T, sv) * tree_stride : 0);
      for (int b = tid & 63; b < B; b += 64) {
        float ae = 0.0f, av = 0.0f;
        for (int s = 0; s < nseg; ++s) {
          const float* rowp =
              dr + ((int64_t)segi[s] * ncoords + scratch_row[sv * nseg + s]) * 2 * B;
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

extern "C" int pyl_circ_branch(const pyl_prune_params* p, const pyl_circ_params* q,
                               const pyl_prune_coords* c, const double* leaves,
                               const pyl_circ_level* L, int64_t* nchild1, int64_t* err,
                               void* stream) {
  PYL_REQUIRE(c != nullptr && L != nullptr, "null coords or level");
  LevelConst k;
  CircConst cc;
  int rc = build_circ(p, q, c, &k, &cc);
  if (rc != PYL_OK) return rc;
  PYL_REQUIRE(leaves && L->branch && L->counts && L->cr_first && L->cr_num && nchild1 && err,
              "null device pointer");
  if (L->n_par == 0) return PYL_OK;
  const int threads = 128;
  circ_branch_kernel<<<(unsigned)pyl::ceil_div(L->n_par, threads), threads, 0,
                       pyl::as_stream(stream)>>>(k, cc, leaves, *L, nullptr, nchild1,
                                                 reinterpret_cast<unsigned long long*>(err));
  PYL_CHECK_LAUNCH("circ_branch_kernel");
  return PYL_OK;
}

extern "C" int pyl_circ_stage1(const pyl_prune_params* p, const pyl_circ_params* q,
                               const pyl_prune_coords* c, const double* leaves,
                               const pyl_circ_level* L, void* stream) {
  PYL_REQUIRE(c != nullptr && L != nullptr, "null coords or level");
  LevelConst k;
  CircConst cc;
  int rc = build_circ(p, q, c, &k, &cc);
  if (rc != PYL_OK) return rc;
  if (L->n_s1 == 0) return PYL_OK;
  PYL_REQUIRE(leaves && L->branch && L->counts && L->cr_num && L->off1 && L->s1_parent &&
                  L->s1_expand,
              "null device pointer");
  const int threads = 128;
  circ_stage1_kernel<<<(unsigned)pyl::ceil_div(L->n_s1, threads), threads, 0,
                       pyl::as_stream(stream)>>>(k, cc, leaves, *L, nullptr, nullptr);
  PYL_CHECK_LAUNCH("circ_stage1_kernel");
  return PYL_OK;
}

extern "C" int pyl_circ_expand(const pyl_prune_params* p, const pyl_circ_params* q,
                               const pyl_prune_coords* c, const double* leaves,
                               const pyl_circ_level* L, void* stream) {
  PYL_REQUIRE(c != nullptr && L != nullptr, "null coords or level");
  LevelConst k;
  CircConst cc;
  int rc = build_circ(p, q, c, &k, &cc);
  if (rc != PYL_OK) return rc;
  if (L->n_s1 == 0) return PYL_OK;
  PYL_REQUIRE(leaves && L->branch && L->counts && L->cr_first && L->cr_num && L->off1 &&
                  L->s1_parent && L->s1_expand && L->s1_pos && L->c_s1 && L->c_sub &&
                  L->c_keep && L->c_region,
              "null device pointer");
  const int threads = 128;
  circ_expand_kernel<<<(unsigned)pyl::ceil_div(L->n_s1, threads), threads, 0,
                       pyl::as_stream(stream)>>>(k, cc, leaves, *L, nullptr);
  PYL_CHECK_LAUNCH("circ_expand_kernel");
  return PYL_OK;
}

template <bool kComplex>
static int circ_children(const pyl_prune_params* p, const pyl_circ_params* q,
                         const pyl_prune_coords* c, const double* leaves, const float* folds,
                         const pyl_circ_level* L, const int64_t* valid, int64_t n_valid,
                         const float* seg, int32_t* c_parent, int32_t* c_row, int32_t* c_shift,
                         double* c_phase, float* c_score, float* minmax, void* stream) {
  PYL_REQUIRE(c != nullptr && L != nullptr, "null coords or level");
  LevelConst k;
  CircConst cc;
  int rc = build_circ(p, q, c, &k, &cc);
  if (rc != PYL_OK) return rc;
  PYL_REQUIRE((int64_t)p->n_accel * p->n_freq <= 0x7fffffff, "FFA grid too large");
  if (n_valid == 0) return PYL_OK;
  PYL_REQUIRE(leaves && folds && valid && seg && c_parent && c_row && c_shift && c_phase &&
                  c_score && minmax && L->c_s1 && L->c_sub && L->c_region && L->s1_parent,
              "null device pointer");
  const size_t lds = ep_block_lds<kComplex>(k);
  PYL_REQUIRE(lds <= kMaxEpLds, "nbins too large for the EP scorer (%d)", k.nbins);
  if constexpr (kComplex)
    PYL_REQUIRE((k.nbins & (k.nbins - 1)) == 0,
                "complex folds need a power-of-two nbins (%d)", k.nbins);
  circ_children_kernel<kComplex><<<(unsigned)items_grid(n_valid), kCBlk,
                                   lds, pyl::as_stream(stream)>>>(
      k, cc, leaves, folds, *L, valid, n_valid, nullptr, seg, c_parent, c_row, c_shift, c_phase,
      c_score, minmax, nullptr);
  PYL_CHECK_LAUNCH("circ_children_kernel");
  return PYL_OK;
}

extern "C" int pyl_circ_children(const pyl_prune_params* p, const pyl_circ_params* q,
                                 const pyl_prune_coords* c, const double* leaves,
                                 const float* folds, const pyl_circ_level* L,
                                 const int64_t* valid, int64_t n_valid, const float* seg,
                                 int32_t* c_parent, int32_t* c_row, int32_t* c_shift,
                                 double* c_phase, float* c_score, float* minmax, void* stream) {
  return circ_children<false>(p, q, c, leaves, folds, L, valid, n_valid, seg, c_parent, c_row,
                              c_shift, c_phase, c_score, minmax, stream);
}

extern "C" int pyl_circ_children_complex(const pyl_prune_params* p, const pyl_circ_params* q,
                                         const pyl_prune_coords* c, const double* leaves,
                                         const float* folds, const pyl_circ_level* L,
                                         const int64_t* valid, int64_t n_valid,
                                         const float* seg, int32_t* c_parent, int32_t* c_row,
                                         int32_t* c_shift, double* c_phase, float* c_score,
                                         float* minmax, void* stream) {
  return circ_children<true>(p, q, c, leaves, folds, L, valid, n_valid, seg, c_parent, c_row,
                             c_shift, c_phase, c_score, minmax, stream);
}

template <bool kComplex>
static int circ_materialize(const pyl_prune_params* p, const pyl_circ_params* q,
                            const pyl_prune_coords* c, const double* leaves, const float* folds,
                            const pyl_circ_level* L, const int64_t* valid, const float* seg,
                            const int32_t* c_parent, const int32_t* c_row,
                            const int32_t* c_shift, const double* c_phase,
                            const float* c_score, const uint8_t* has_taylor,
                            const int64_t* surv, int64_t n_surv, double* o_leaves,
                            float* o_folds, float* o_scores, int32_t* o_bt, void* stream) {
  PYL_REQUIRE(c != nullptr && L != nullptr, "null coords or level");
  LevelConst k;
  CircConst cc;
  int rc = build_circ(p, q, c, &k, &cc);
  if (rc != PYL_OK) return rc;
  if (n_surv == 0) return PYL_OK;
  PYL_REQUIRE(leaves && folds && valid && seg && c_parent && c_row && c_shift && c_phase &&
                  c_score && has_taylor && surv && o_leaves && o_folds && o_scores && o_bt,
              "null device pointer");
  circ_materialize_kernel<kComplex><<<(unsigned)items_grid(n_surv), kCBlk,
                                      materialize_lds(kComplex, k.nbins),
                                      pyl::as_stream(stream)>>>(
      k, cc, leaves, folds, *L, valid, seg, c_parent, c_row, c_shift, c_phase, c_score,
      has_taylor, surv, n_surv, nullptr, o_leaves, o_folds, o_scores, o_bt, nullptr);
  PYL_CHECK_LAUNCH("circ_materialize_kernel");
  return PYL_OK;
}

extern "C" int pyl_circ_materialize(const pyl_prune_params* p, const pyl_circ_params* q,
                                    const pyl_prune_coords* c, const double* leaves,
                                    const float* folds, const pyl_circ_level* L,
                                    const int64_t* valid, const float* seg,
                                    const int32_t* c_parent, const int32_t* c_row,
                                    const int32_t* c_shift, const double* c_phase,
                                    const float* c_score, const uint8_t* has_taylor,
                                    const int64_t* surv, int64_t n_surv, double* o_leaves,
                                    float* o_folds, float* o_scores, int32_t* o_bt,
                                    void* stream) {
  return circ_materialize<false>(p, q, c, leaves, folds, L, valid, seg, c_parent, c_row,
                                 c_shift, c_phase, c_score, has_taylor, surv, n_surv, o_leaves,
                                 o_folds, o_scores, o_bt, stream);
}

extern "C" int pyl_circ_materialize_complex(
    const pyl_prune_params* p, const pyl_circ_params* q, const pyl_prune_coords* c,
    const double* leaves, const float* folds, const pyl_circ_level* L, const int64_t* valid,
    const float* seg, const int32_t* c_parent, const int32_t* c_row, const int32_t* c_shift,
    const double* c_phase, const float* c_score, const uint8_t* has_taylor, const int64_t* surv,
    int64_t n_surv, double* o_leaves, float* o_folds, float* o_scores, int32_t* o_bt,
    void* stream) {
  return circ_materialize<true>(p, q, c, leaves, folds, L, valid, seg, c_parent, c_row,
                                c_shift, c_phase, c_score, has_taylor, surv, n_surv, o_leaves,
                                o_folds, o_scores, o_bt, stream);
}

// tree_off (host [T + 1]) and tree_stride: a forest's trees (null: one tree).
template <bool kComplex>
static int circ_ascend(const pyl_prune_params* p, const pyl_circ_params* q,
                       const double* leaves, int64_t n, const float* dyp, int64_t ncoords,
                       const int64_t* seg_idx, const double* seg_t0, int32_t nseg,
                       double t0_cur, float* o_folds, float* o_scores, void* stream,
                       const int64_t* tree_off = nullptr, int T = 1,
                       int64_t tree_stride = 0) {
  LevelConst k;
  CircConst cc;
  int rc = build_circ(p, q, nullptr, &k, &cc);
  if (rc != PYL_OK) return rc;
  PYL_REQUIRE(nseg >= 1, "no segments to integrate");
  PYL_REQUIRE(leaves && dyp && seg_idx && seg_t0 && o_folds && o_scores, "null pointer");
  if (n == 0) return PYL_OK;
  const size_t lds = ep_block_lds<kComplex>(k);
  PYL_REQUIRE(lds <= kMaxEpLds, "nbins too large for the EP scorer (%d)", k.nbins);
  if constexpr (kComplex)
    PYL_REQUIRE((k.nbins & (k.nbins - 1)) == 0,
                "complex folds need a power-of-two nbins (%d)", k.nbins);
  hipStream_t s = pyl::as_stream(stream);
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
  int64_t* d_off = nullptr;
  if (tree_off) {
    PYL_CHECK_HIP(hipMallocAsync(reinterpret_cast<void**>(&d_off), sizeof(int64_t) * (T + 1), s));
    PYL_CHECK_HIP(
        hipMemcpyAsync(d_off, tree_off, sizeof(int64_t) * (T + 1), hipMemcpyHostToDevice, s));
  }
  circ_ascend_kernel<kComplex><<<(unsigned)pyl::ceil_div(n, kBlk), kCBlk, lds, s>>>(
      k, cc, leaves, n, dyp, ncoords, d_idx, d_t0, nseg, t0_cur, d_row, d_ph, o_folds,
      o_scores, d_off, T, tree_stride);
  PYL_CHECK_LAUNCH("circ_ascend_kernel");
  if (d_off) PYL_CHECK_HIP(hipFreeAsync(d_off, s));
  PYL_CHECK_HIP(hipFreeAsync(d_idx, s));
  PYL_CHECK_HIP(hipFreeAsync(d_t0, s));
  PYL_CHECK_HIP(hipFreeAsync(d_row, s));
  PYL_CHECK_HIP(hipFreeAsync(d_ph, s));
  return PYL_OK;
}

extern "C" int pyl_circ_ascend(const pyl_prune_params* p, const pyl_circ_params* q,
                               const double* leaves, int64_t n, const float* dyp,
                               int64_t ncoords, const int64_t* seg_idx, const double* seg_t0,
                               int32_t nseg, double t0_cur, float* o_folds, float* o_scores,
                               void* stream) {
  return circ_ascend<false>(p, q, leaves, n, dyp, ncoords, seg_idx, seg_t0, nseg, t0_cur,
                            o_folds, o_scores, stream);
}

extern "C" int pyl_circ_ascend_complex(const pyl_prune_params* p, const pyl_circ_params* q,
                                       const double* leaves, int64_t n, const float* dyp,
                                       int64_t ncoords, const int64_t* seg_idx,
                                       const double* seg_t0, int32_t nseg, double t0_cur,
                                       float* o_folds, float* o_scores, void* stream) {
  return circ_ascend<true>(p, q, leaves, n, dyp, ncoords, seg_idx, seg_t0, nseg, t0_cur,
                           o_folds, o_scores, stream);
}

extern "C" int pyl_circ_forest_ascend(const pyl_prune_params* p, const pyl_circ_params* q,
                                      int32_t flags, const double* leaves,
                                      const int64_t* tree_off, int32_t ntree,
                                      const float* dyp, int64_t tree_stride, int64_t ncoords,
                                      const int64_t* seg_idx, const double* seg_t0,
                                      int32_t nseg, double t0_cur, float* o_folds,
                                      float* o_scores, void* stream) {
  PYL_REQUIRE(p != nullptr && q != nullptr && tree_off != nullptr, "null parameters");
  PYL_REQUIRE((flags & ~PYL_EP_COMPLEX) == 0, "unknown flags %d", flags);
  PYL_REQUIRE(ntree >= 1 && ntree <= PYL_EP_FOREST_MAX, "ntree must be in [1, %d]",
              PYL_EP_FOREST_MAX);
  PYL_REQUIRE(tree_off[0] == 0, "tree offsets must start at 0");
  for (int t = 0; t < ntree; ++t)
    PYL_REQUIRE(tree_off[t + 1] >= tree_off[t], "tree offsets must not decrease");
  PYL_REQUIRE(tree_stride >= 0, "negative tree_stride");
  const int64_t n = tree_off[ntree];
  auto* f = (flags & PYL_EP_COMPLEX) ? circ_ascend<true> : circ_ascend<false>;
  return f(p, q, leaves, n, dyp, ncoords, seg_idx, seg_t0, nseg, t0_cur, o_folds, o_scores,
           stream, tree_off, ntree, tree_stride);
}

// ---- The sync-free level loop (circular basis) ---------------------------------------
namespace {

// Held children: the level goes on only if they fit the buffers and every crackle
// expansion's outputs stay below 2^31 in the packed prefix sums below.
struct HeldOffSink {
  int64_t* off1;
  int64_t* rec;
  int64_t cap_child, max_sugg, cap_hist, batch_size, branch_max;
  const int64_t* nb_forest;  // a forest's batch count (its bo[T]), else null
  __device__ void operator()(int64_t i, int64_t excl, int64_t x) const { off1[i + 1] = excl + x; }
  __device__ void finish(int64_t n_s1) const {
    off1[0] = 0;
    int64_t st = rec[kRecStatus];
    if (st == 0 && rec[kRecErr] != INT64_MAX) st = PYL_EP_ST_BRANCH;
    if (st == 0 && (n_s1 > cap_child || n_s1 * branch_max > 0x7fffffffLL))
      st = PYL_EP_ST_CHILD_CAP;
    if (st == 0 && rec[kRecHist] + max_sugg > cap_hist) st = PYL_EP_ST_HIST_FULL;
    rec[kRecStatus] = st;
    rec[PYL_EP_REC_NHELD] = st ? 0 : n_s1;
    rec[PYL_EP_REC_NBATCH] =
        st ? 0 : (nb_forest ? *nb_forest : (rec[kRecNpar] + batch_size - 1) / batch_size);
  }
};

// Per held child, packed: 1 in the low half when it does not expand, its crackle outputs
// in the high half when it does (both halves stay below 2^31).
struct HeldSrc {
  const uint8_t* expand;
  const int32_t* s1_parent;
  const int32_t* cr_num;
  __device__ int64_t operator()(int64_t s) const {
    return expand[s] ? ((int64_t)cr_num[s1_parent[s]] << 32) : 1;
  }
};
struct PackedOffSink {  // prefix sums modulo 2^64: only their differences are used
  int64_t* off;
  __device__ void operator()(int64_t i, int64_t excl, int64_t x) const {
    off[i + 1] = wadd(excl, x);
  }
  __device__ void finish(int64_t) const { off[0] = 0; }
};

// Children of every batch of parents (circular.py:359-394): its held children that do not
// expand, then the crackle children of those that do.
__global__ void circ_batch_totals_kernel(const int64_t* __restrict__ off1,
                                         const int64_t* __restrict__ held_off,
                                         const int64_t* __restrict__ rec, int64_t batch_size,
                                         int64_t* __restrict__ batch_tot,
                                         int64_t* __restrict__ batch_nk,
                                         uint8_t* __restrict__ has_taylor, FTab f) {
  const int64_t nb = rec[PYL_EP_REC_NBATCH];
  const int64_t n_par = rec[kRecNpar];
  for (int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; b < nb;
       b += (int64_t)gridDim.x * blockDim.x) {
    has_taylor[b] = 0;  // set again after the children are scored
    int64_t p0, p1;
    f_batch_parents(f, b, batch_size, n_par, &p0, &p1);
    const uint64_t d = (uint64_t)wsub(held_off[off1[p1]], held_off[off1[p0]]);
    const int64_t nk = (int64_t)(d & 0xffffffffull);
    batch_nk[b] = nk;
    batch_tot[b] = nk + (int64_t)(d >> 32);
  }
}

// First child of every batch; the level's children must fit the buffers.
struct BatchOffSink {
  int64_t* batch_off;
  int64_t* rec;
  int64_t cap_child;
  __device__ void operator()(int64_t i, int64_t excl, int64_t) const { batch_off[i] = excl; }
  __device__ void finish(int64_t total) const {
    int64_t st = rec[kRecStatus];
    if (st == 0 && total > cap_child) st = PYL_EP_ST_CHILD_CAP;
    rec[kRecStatus] = st;
    rec[kRecNchild] = st ? 0 : total;
    rec[PYL_EP_REC_NEXPAND] = st ? 0 : rec[PYL_EP_REC_NHELD];
  }
};

__global__ void circ_positions_kernel(const int64_t* __restrict__ off1,
                                      const int64_t* __restrict__ held_off,
                                      const int32_t* __restrict__ s1_parent,
                                      const uint8_t* __restrict__ s1_expand,
                                      const int64_t* __restrict__ batch_off,
                                      const int64_t* __restrict__ batch_nk,
                                      const int64_t* __restrict__ n_dev, int64_t batch_size,
                                      int64_t* __restrict__ s1_pos, FTab f) {
  const int64_t n = *n_dev;
  for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < n;
       s += (int64_t)gridDim.x * blockDim.x) {
    const int64_t b = f_batch(f, s1_parent[s], batch_size);
    int64_t p0, p1;
    f_batch_parents(f, b, batch_size, INT64_MAX, &p0, &p1);
    const uint64_t d = (uint64_t)wsub(held_off[s], held_off[off1[p0]]);
    s1_pos[s] = batch_off[b] + (s1_expand[s] ? batch_nk[b] + (int64_t)(d >> 32)
                                              : (int64_t)(d & 0xffffffffull));
  }
}

struct ValidSink {
  int64_t* valid;
  int64_t* rec;
  __device__ void operator()(int64_t i, int64_t excl, int64_t x) const {
    if (x) valid[excl] = i;
  }
  __device__ void finish(int64_t nv) const { rec[kRecNvalid] = nv; }
};

// ---- Forest tables -------------------------------------------------------------------
struct TreeSizes {
  int64_t n[PYL_EP_FOREST_MAX];
};

// One block: every level's tables cleared, per-tree min/max reset, and level 0's parent
// offsets from the input trees' sizes.
__global__ void forest_init_kernel(int64_t* __restrict__ tab, float* __restrict__ minmax,
                                   int T, int nlev, TreeSizes n_in) {
  const int64_t words = (int64_t)(nlev + 1) * PYL_EP_FOREST_TAB * (T + 1);
  for (int64_t i = threadIdx.x; i < words; i += blockDim.x) tab[i] = 0;
  for (int64_t i = threadIdx.x; i < (int64_t)nlev * T; i += blockDim.x) {
    minmax[2 * i] = INFINITY;
    minmax[2 * i + 1] = -INFINITY;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int64_t acc = 0;
    for (int t = 0; t < T; ++t) {
      tab[t] = acc;
      acc += n_in.n[t];
    }
    tab[T] = acc;
  }
}

// A level's batch offsets: each tree's parents in batches of bs (circular.py:359-394).
__global__ void forest_begin_kernel(int64_t* __restrict__ tab, int T, int64_t bs) {
  if (threadIdx.x != 0) return;
  const int64_t* po = tab + PYL_EP_TAB_PO * (T + 1);
  int64_t* bo = tab + PYL_EP_TAB_BO * (T + 1);
  int64_t acc = 0;
  for (int t = 0; t < T; ++t) {
    bo[t] = acc;
    acc += (po[t + 1] - po[t] + bs - 1) / bs;
  }
  bo[T] = acc;
}

// Every tree's held children times branch_max below 2^31 (the bound of the packed
// per-batch sums, batches never straddling trees), else the level stops with
// PYL_EP_ST_TREE_OVERFLOW (the forest's trees then go on alone).
__global__ void forest_held_check_kernel(int64_t* __restrict__ rec,
                                         const int64_t* __restrict__ tab, int T,
                                         const int64_t* __restrict__ off1, int64_t branch_max) {
  __shared__ int s_over;
  const int t = threadIdx.x;
  if (t == 0) s_over = 0;
  __syncthreads();
  const int64_t* po = tab + PYL_EP_TAB_PO * (T + 1);
  if (rec[kRecStatus] == 0 && t < T && (off1[po[t + 1]] - off1[po[t]]) * branch_max > 0x7fffffffLL)
    s_over = 1;
  __syncthreads();
  if (t == 0 && s_over) {
    rec[kRecStatus] = PYL_EP_ST_TREE_OVERFLOW;
    rec[PYL_EP_REC_NHELD] = 0;
    rec[PYL_EP_REC_NBATCH] = 0;
  }
}

// First index i in [0, n) with a[i] >= x (a increasing), else n.
__device__ __forceinline__ int64_t lower_bound_i64(const int64_t* a, int64_t n, int64_t x) {
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (a[mid] < x) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// After the survivor scan: every tree's children (its batches' first child), scored
// children and survivors; a tree with more survivors than tree_max stops the level
// (PYL_EP_ST_TREE_OVERFLOW: nothing is materialised). Without a budget, the survivors are the
// next level's parents. One block of at least T + 1 threads.
__global__ void forest_split_kernel(int64_t* __restrict__ rec, int64_t* __restrict__ tab,
                                    int64_t* __restrict__ tab_next, int T,
                                    const int64_t* __restrict__ batch_off,
                                    const int64_t* __restrict__ valid,
                                    const int64_t* __restrict__ surv, int64_t tree_max,
                                    int budget) {
  __shared__ int s_over;
  const int t = threadIdx.x;
  const int64_t* bo = tab + PYL_EP_TAB_BO * (T + 1);
  int64_t* cb = tab + PYL_EP_TAB_CB * (T + 1);
  int64_t* vb = tab + PYL_EP_TAB_VB * (T + 1);
  int64_t* so = tab + PYL_EP_TAB_SO * (T + 1);
  int64_t* ko = tab + PYL_EP_TAB_KO * (T + 1);
  if (t == 0) s_over = 0;
  __syncthreads();
  const bool ok = rec[kRecStatus] == 0;
  if (t <= T) {
    int64_t c = 0, v = 0, sv = 0;
    if (ok) {
      const int64_t nb = rec[PYL_EP_REC_NBATCH], nc = rec[kRecNchild];
      c = (t < T && bo[t] < nb) ? batch_off[bo[t]] : nc;
      v = lower_bound_i64(valid, rec[kRecNvalid], c);
      sv = lower_bound_i64(surv, rec[kRecNsurv], v);
    }
    cb[t] = c;
    vb[t] = v;
    so[t] = sv;
  }
  __syncthreads();
  if (t < T && so[t + 1] - so[t] > tree_max) s_over = 1;
  __syncthreads();
  if (s_over) {
    if (t <= T) so[t] = 0;
    if (t == 0) {
      rec[kRecStatus] = PYL_EP_ST_TREE_OVERFLOW;
      rec[kRecNsurv] = 0;
      rec[kRecWords + kRecNpar] = 0;
      rec[kRecWords + kRecHist] = rec[kRecHist];
    }
  }
  __syncthreads();
  if (!budget && t <= T) {
    ko[t] = so[t];
    tab_next[t] = so[t];  // PO of the next level
  }
}


// One tree (F null), or a forest of F->ntree trees whose input sizes are n_tree.
template <bool kComplex, bool kForest>
int circ_levels(const pyl_prune_params* p, const pyl_circ_params* q,
                const pyl_prune_coords* coords, const double* thr, const int64_t* seg_idx,
                int32_t nlev, const float* fold, int64_t seg_stride, const double* in_leaves,
                const float* in_folds, int64_t n_in, int64_t hist_row0, int32_t first_out,
                const pyl_ep_loop* L, const pyl_circ_loop* C, const pyl_ep_forest* F,
                const int64_t* n_tree, hipStream_t s) {
  PYL_REQUIRE(L != nullptr && C != nullptr && coords != nullptr && thr != nullptr &&
                  seg_idx != nullptr,
              "null host argument");
  PYL_REQUIRE(nlev >= 1, "nlev must be >= 1");
  PYL_REQUIRE(first_out == 0 || first_out == 1, "first_out must be 0 or 1");
  PYL_REQUIRE(L->cap_par >= 1 && L->cap_child >= 1 && L->cap_child < 0x7fffffffLL &&
                  L->max_sugg >= 1 && L->max_sugg <= L->cap_par,
              "invalid capacities");
  PYL_REQUIRE(n_in >= 0 && n_in <= L->cap_par, "input tree exceeds the tree buffers");
  PYL_REQUIRE(hist_row0 >= 0 && hist_row0 <= L->cap_hist, "hist_row0 out of range");
  PYL_REQUIRE(fold && in_leaves && in_folds && L->leaves[0] && L->leaves[1] && L->folds[0] &&
                  L->folds[1] && L->scores[0] && L->scores[1] && L->hist && L->branch &&
                  L->counts && L->nchild && L->c_parent && L->c_row && L->c_shift &&
                  L->c_phase && L->c_score && L->surv && L->part && L->rec && L->minmax &&
                  C->cr_first && C->cr_num && C->off1 && C->s1_parent && C->s1_expand &&
                  C->s1_pos && C->held_off && C->c_s1 && C->c_sub && C->c_keep &&
                  C->c_region && C->valid && C->batch_tot && C->batch_nk && C->batch_off &&
                  C->has_taylor,
              "null device pointer");
  const pyl_ep_budget* bud = L->budget;
  // Tree buffers, level by level. A level without a budget, or a forest level ranked by
  // S/N (cut before anything is materialised, so only the kept survivors are), writes the
  // buffer its input is not in (the first level: first_out). A level ranked by a network
  // materialises every survivor into the buffer its input is not in, then gathers the
  // kept rows into the other one (the input's, or 1 for an outside input). Host code
  // follows the same rule to find a level's tree.
  const auto cuts_first = [&](int l) {
    return kForest && bud && pyl::kind_of(bud, bud->level0 + l) == PYL_BUDGET_SCORE;
  };
  int prev = in_leaves == L->leaves[0] ? 0 : (in_leaves == L->leaves[1] ? 1 : -1);
  PYL_REQUIRE((prev == 0) == (in_folds == L->folds[0]) &&
                  (prev == 1) == (in_folds == L->folds[1]),
              "the input leaves and folds are in different tree buffers");
  PYL_REQUIRE(prev < 0 || !(!bud || cuts_first(0)) || first_out == 1 - prev,
              "the input tree is the first output buffer");
  PYL_REQUIRE(kForest || !bud || prev != 0, "the input tree is the first output buffer");
  PYL_REQUIRE(!bud || bud->bt_tmp, "null budget buffer");
  PYL_REQUIRE(!bud || !kForest || (bud->keys && bud->kept), "null budget buffer");
  PYL_REQUIRE((int64_t)p->n_accel * p->n_freq <= 0x7fffffff, "FFA grid too large");
  int T = 1;
  int64_t row_stride = 0;
  if constexpr (kForest) {
    PYL_REQUIRE(F != nullptr && n_tree != nullptr, "null forest");
    T = F->ntree;
    PYL_REQUIRE(T >= 1 && T <= PYL_EP_FOREST_MAX, "ntree must be in [1, %d] (got %d)",
                PYL_EP_FOREST_MAX, T);
    PYL_REQUIRE(F->tab && F->minmax, "null forest table");
    PYL_REQUIRE(F->max_sugg >= 1, "invalid per-tree max_sugg");
    PYL_REQUIRE(!bud || (F->sel && F->part && F->kcount), "null forest budget state");
    const int64_t row_floats = kComplex ? 4 * (p->nbins / 2 + 1) : 2 * (int64_t)p->nbins;
    PYL_REQUIRE(F->tree_stride >= 0 && F->tree_stride % row_floats == 0,
                "tree_stride must be a multiple of a fold row (%lld floats)",
                (long long)row_floats);
    row_stride = F->tree_stride / row_floats;
    PYL_REQUIRE((T - 1) * row_stride + (int64_t)p->n_accel * p->n_freq <= 0x7fffffff,
                "stacked folds too large for 32-bit row indices");
    TreeSizes sz{};
    int64_t tot = 0;
    for (int t = 0; t < T; ++t) {
      PYL_REQUIRE(n_tree[t] >= 0, "negative tree size");
      sz.n[t] = n_tree[t];
      tot += n_tree[t];
    }
    PYL_REQUIRE(tot == n_in, "tree sizes do not add up to the input rows");
    // every tree's batches together: at most cap_par (the per-batch arrays)
    PYL_REQUIRE(q->batch_size == 1 || L->cap_par >= 2 * (int64_t)T,
                "cap_par must be at least twice the number of trees");
    forest_init_kernel<<<1, 256, 0, s>>>(F->tab, F->minmax, T, nlev, sz);
    PYL_CHECK_LAUNCH("forest_init_kernel");
  }
  ep_records_init_kernel<<<(unsigned)pyl::ceil_div(nlev + 1, 64), 64, 0, s>>>(
      L->rec, L->minmax, nlev, n_in, hist_row0);
  PYL_CHECK_LAUNCH("ep_records_init_kernel");
  const int64_t tabw = (int64_t)PYL_EP_FOREST_TAB * (T + 1);
  pyl_circ_level cl{};
  cl.branch = L->branch;
  cl.counts = L->counts;
  cl.cr_first = C->cr_first;
  cl.cr_num = C->cr_num;
  cl.off1 = C->off1;
  cl.s1_parent = C->s1_parent;
  cl.s1_expand = C->s1_expand;
  cl.s1_pos = C->s1_pos;
  cl.c_s1 = C->c_s1;
  cl.c_sub = C->c_sub;
  cl.c_keep = C->c_keep;
  cl.c_region = C->c_region;
  const int tpar = 128;
  const unsigned g_par = strided_grid(L->cap_par, tpar);
  const unsigned g_held = strided_grid(L->cap_child, tpar);
  const unsigned g_child = items_grid_cap(L->cap_child);
  // Survivors are children of the level: the materialize grid covers the smaller capacity.
  const unsigned g_surv = items_grid_cap(std::min(L->max_sugg, L->cap_child));
  for (int l = 0; l < nlev; ++l) {
    LevelConst k;
    CircConst cc;
    int rc = build_circ(p, q, &coords[l], &k, &cc);
    if (rc != PYL_OK) return rc;
    const size_t lds = ep_block_lds<kComplex>(k);
    PYL_REQUIRE(lds <= kMaxEpLds, "nbins too large for the EP scorer (%d)", k.nbins);
    if constexpr (kComplex)
      PYL_REQUIRE((k.nbins & (k.nbins - 1)) == 0,
                  "complex folds need a power-of-two nbins (%d)", k.nbins);
    PYL_REQUIRE(seg_idx[l] >= 0, "negative segment index");
    int64_t* rec = L->rec + (int64_t)l * kRecWords;
    const double* leaves = l == 0 ? in_leaves : L->leaves[prev];
    const float* folds = l == 0 ? in_folds : L->folds[prev];
    const bool cut_first = cuts_first(l);
    const bool pingpong = !bud || cut_first;
    // o: the buffer the materialize kernel writes; kb: a network level's kept rows.
    const int o = pingpong ? (prev < 0 ? first_out : 1 - prev) : (prev == 0 ? 1 : 0);
    const int kb = 1 - o;
    prev = pingpong ? o : kb;
    const float* seg = fold + seg_idx[l] * seg_stride;
    const int64_t bs = cc.batch_size;
    FTab ft{};
    int64_t* tab = nullptr;
    if constexpr (kForest) {
      tab = F->tab + l * tabw;
      ft = FTab{tab, T, row_stride, F->minmax + (int64_t)l * T * 2};
    }
    pyl_ep_timer* tm = L->timer;
    if (tm) tm->begin(PYL_EP_PH_BRANCH, s);
    if constexpr (kForest) forest_begin_kernel<<<1, 64, 0, s>>>(tab, T, bs);
    circ_branch_kernel<<<g_par, tpar, 0, s>>>(
        k, cc, leaves, cl, rec + kRecNpar, L->nchild,
        reinterpret_cast<unsigned long long*>(rec + kRecErr));
    const int64_t bmax = std::max(p->branch_max, 1);
    chunked_scan(CountSrc{L->nchild},
                 HeldOffSink{C->off1, rec, L->cap_child, L->max_sugg, L->cap_hist, bs,
                             kForest ? 0 : bmax,  // a forest checks its trees one by one
                             kForest ? tab + PYL_EP_TAB_BO * (T + 1) + T : nullptr},
                 rec + kRecNpar, L->part, s);
    if constexpr (kForest) {
      forest_held_check_kernel<<<1, 64, 0, s>>>(rec, tab, T, C->off1, bmax);
      PYL_CHECK_LAUNCH("forest_held_check_kernel");
    }
    circ_stage1_kernel<<<g_held, tpar, 0, s>>>(k, cc, leaves, cl, rec + kRecNpar,
                                              rec + PYL_EP_REC_NHELD);
    chunked_scan(HeldSrc{C->s1_expand, C->s1_parent, C->cr_num}, PackedOffSink{C->held_off},
                 rec + PYL_EP_REC_NHELD, L->part, s);
    circ_batch_totals_kernel<<<g_par, tpar, 0, s>>>(C->off1, C->held_off, rec, bs, C->batch_tot,
                                                   C->batch_nk, C->has_taylor, ft);
    chunked_scan(CountSrc{C->batch_tot}, BatchOffSink{C->batch_off, rec, L->cap_child},
                 rec + PYL_EP_REC_NBATCH, L->part, s);
    circ_positions_kernel<<<g_held, tpar, 0, s>>>(C->off1, C->held_off, C->s1_parent,
                                                 C->s1_expand, C->batch_off, C->batch_nk,
                                                 rec + PYL_EP_REC_NEXPAND, bs, C->s1_pos, ft);
    circ_expand_kernel<<<g_held, tpar, 0, s>>>(k, cc, leaves, cl, rec + PYL_EP_REC_NEXPAND);
    chunked_scan(FlagSrc{C->c_keep}, ValidSink{C->valid, rec}, rec + kRecNchild, L->part, s);
    PYL_CHECK_LAUNCH("circ_expand_kernel");
    if (tm) tm->end(s);
    if (tm) tm->begin(PYL_EP_PH_CHILDREN, s);
    circ_children_kernel<kComplex, kForest><<<g_child, kCBlk, lds, s>>>(
        k, cc, leaves, folds, cl, C->valid, 0, rec + kRecNvalid, seg, L->c_parent, L->c_row,
        L->c_shift, L->c_phase, L->c_score, L->minmax + 2 * l, C->has_taylor, ft);
    PYL_CHECK_LAUNCH("circ_children_kernel");
    if (tm) tm->end(s);
    if (tm) tm->begin(PYL_EP_PH_SELECT, s);
    chunked_scan(PassSrc{L->c_score, thr[l]},
                 SurvSink{L->surv, rec, L->max_sugg, L->cap_hist, L->c_score,
                          cut_first ? bud->keys : nullptr},
                 rec + kRecNvalid, L->part, s);
    PYL_CHECK_LAUNCH("chunk_scan_kernel");
    KMap km{};
    if constexpr (kForest) {
      forest_split_kernel<<<1, 128, 0, s>>>(rec, tab, tab + tabw, T, C->batch_off, C->valid,
                                            L->surv, F->max_sugg, bud ? 1 : 0);
      PYL_CHECK_LAUNCH("forest_split_kernel");
      if (cut_first) {
        rc = pyl::ep_budget_forest_select(bud, F, bud->level0 + l, tab, tab + tabw, bud->keys,
                                          L->max_sugg, rec, s);
        if (rc != PYL_OK) return rc;
        const bool cut = bud->level_mask == nullptr || bud->level_mask[bud->level0 + l] != 0;
        km = KMap{cut ? bud->kept : nullptr, tab, T};
      }
    }
    if (tm) tm->end(s);
    if (tm) tm->begin(PYL_EP_PH_MATERIALIZE, s);
    circ_materialize_kernel<kComplex, kForest>
        <<<g_surv, kCBlk, materialize_lds(kComplex, k.nbins), s>>>(
            k, cc, leaves, folds, cl, C->valid, seg, L->c_parent, L->c_row, L->c_shift,
            L->c_phase, L->c_score, C->has_taylor, L->surv, 0, rec + kRecWords + kRecNpar,
            L->leaves[o], L->folds[o], L->scores[o], pingpong ? L->hist : bud->bt_tmp,
            pingpong ? rec + kRecHist : nullptr, ft, km);
    PYL_CHECK_LAUNCH("circ_materialize_kernel");
    if (tm) tm->end(s);
    if (bud && !cut_first) {
      if (tm) tm->begin(PYL_EP_PH_BUDGET, s);
      if constexpr (kForest)
        rc = pyl::ep_budget_forest_enqueue(bud, F, kP, bud->level0 + l, tab, tab + tabw,
                                           L->leaves[o], L->folds[o], L->scores[o],
                                           bud->bt_tmp, L->max_sugg, rec, L->leaves[kb],
                                           L->folds[kb], L->scores[kb], L->hist, s);
      else
        rc = pyl::ep_budget_level_enqueue(bud, kP, bud->level0 + l, L->leaves[o],
                                          L->folds[o], L->scores[o], bud->bt_tmp, L->max_sugg,
                                          rec, L->leaves[kb], L->folds[kb], L->scores[kb],
                                          L->hist, s);
      if (rc != PYL_OK) return rc;
      if (tm) tm->end(s);
    }
  }
  return PYL_OK;
}

}  // namespace

extern "C" int pyl_circ_levels(const pyl_prune_params* p, const pyl_circ_params* q,
                               int32_t flags, const pyl_prune_coords* coords,
                               const double* thresholds, const int64_t* seg_idx, int32_t nlev,
                               const float* dyp_fold, int64_t seg_stride,
                               const double* in_leaves, const float* in_folds, int64_t n_in,
                               int64_t hist_row0, int32_t first_out, const pyl_ep_loop* L,
                               const pyl_circ_loop* C, void* stream) {
  PYL_REQUIRE(p != nullptr && q != nullptr, "null parameters");
  PYL_REQUIRE((flags & ~PYL_EP_COMPLEX) == 0, "unknown flags %d", flags);
  auto* f = (flags & PYL_EP_COMPLEX) ? circ_levels<true, false> : circ_levels<false, false>;
  return f(p, q, coords, thresholds, seg_idx, nlev, dyp_fold, seg_stride, in_leaves, in_folds,
           n_in, hist_row0, first_out, L, C, nullptr, nullptr, pyl::as_stream(stream));
}

extern "C" int pyl_circ_forest_levels(const pyl_prune_params* p, const pyl_circ_params* q,
                                      int32_t flags, const pyl_prune_coords* coords,
                                      const double* thresholds, const int64_t* seg_idx,
                                      int32_t nlev, const float* dyp_fold, int64_t seg_stride,
                                      const double* in_leaves, const float* in_folds,
                                      const int64_t* n_in, int64_t hist_row0,
                                      int32_t first_out, const pyl_ep_loop* L,
                                      const pyl_circ_loop* C, const pyl_ep_forest* F,
                                      void* stream) {
  PYL_REQUIRE(p != nullptr && q != nullptr && F != nullptr && n_in != nullptr,
              "null parameters");
  PYL_REQUIRE((flags & ~PYL_EP_COMPLEX) == 0, "unknown flags %d", flags);
  PYL_REQUIRE(F->ntree >= 1 && F->ntree <= PYL_EP_FOREST_MAX, "ntree must be in [1, %d]",
              PYL_EP_FOREST_MAX);
  int64_t tot = 0;
  for (int t = 0; t < F->ntree; ++t) tot += n_in[t];
  auto* f = (flags & PYL_EP_COMPLEX) ? circ_levels<true, true> : circ_levels<false, true>;
  return f(p, q, coords, thresholds, seg_idx, nlev, dyp_fold, seg_stride, in_leaves, in_folds,
           tot, hist_row0, first_out, L, C, F, n_in, pyl::as_stream(stream));
}
