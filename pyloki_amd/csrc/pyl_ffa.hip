// One FFA level (dyn_pffa.py:435-489 unify_fold) as two kernels.
//
// ffa_resolve_kernel: one thread per output parameter set j. It decodes j into grid
//   indices (cartesian product, last parameter fastest), then for the tail (latter=0)
//   and head (latter=1) half resolves in FP64 exactly as core/fold.py:617-671 does:
//   P == 1: delta_t = latter * 2^(L-1) * tseg_brute, parameters unchanged;
//   P  > 1: delta_t = (latter - 1/2) * 2^(L-1) * tseg_brute, the Taylor vector
//           [d_{P-1}..d_2, 0, 0] is translated by delta_t (transforms.py:10-47,140-170),
//           f' = f (1 - v/c), delay = d/c.
//   It writes the flattened previous-level row of each half and the integer bin shift.
//
// ffa_shift_add_kernel: the memory-bound part. Each output row (segment pair q, set j)
//   is 2 x nbins floats = 512 B at nbins = 64; a row is read from the tail and head
//   rows with a circular rotation. Rows are processed as float4 chunks: a lane loads the
//   two aligned 16 B chunks of the source row that hold its rotated elements and selects
//   -- every load instruction of a wavefront covers whole 64 B segments, so the rotation
//   costs no extra HBM traffic.
//   Rows of consecutive j share their source rows (the previous level's grid is about
//   half as fine), so each block takes a contiguous run of j and the runs are dealt to
//   the 8 XCDs in contiguous slabs: neighbours' rows are reused from the same L2.
#include <algorithm>
#include <atomic>
#include <mutex>

#include "pyl_score.hpp"

namespace {

constexpr int kMaxP = 5;
constexpr int kRoundsMaxNW = 16;  // widths handled by the round-based fused kernel

struct ResolveArgs {
  int32_t P;
  int32_t nbins;
  int32_t level;
  int32_t pad;
  int64_t ncoords;
  int64_t counts_cur[kMaxP];
  int64_t counts_prev[kMaxP];
  int64_t offs[kMaxP];
  double lim_lo[kMaxP];
  double lim_hi[kMaxP];
  double tseg_brute;
  double delta_t[2];                       // tail, head
  double tmat[2][(kMaxP + 1) * (kMaxP + 1)];  // Taylor shift matrix per half (row-major)
};

// With `boundary` set the kernel writes no table and instead counts the decisions a +-1 ulp
// change of their FP64 input would flip: boundary[0] grid indices, boundary[1] shifts.
__global__ void ffa_resolve_kernel(const double* __restrict__ params, ResolveArgs a,
                                   int4* __restrict__ table, double2* __restrict__ phases,
                                   unsigned long long* __restrict__ boundary = nullptr) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= a.ncoords) return;
  const int P = a.P;
  double p[kMaxP];
  int64_t r = j;
  for (int k = P - 1; k >= 0; --k) {
    const int64_t ik = r % a.counts_cur[k];
    r /= a.counts_cur[k];
    p[k] = params[a.offs[k] + ik];
  }
  int out[4];
  double ph[2];
  unsigned n_idx = 0, n_shift = 0;
  for (int latter = 0; latter < 2; ++latter) {
    double pnew[kMaxP];
    double delay;
    const double delta_t = a.delta_t[latter];
    if (P == 1) {
      pnew[0] = p[0];
      delay = 0.0;
    } else {
      double tv[kMaxP + 1], tn[kMaxP + 1];
      const int n = P + 1;
      const double* T = a.tmat[latter];
      for (int k = 0; k < n; ++k) tv[k] = 0.0;
      for (int k = 0; k < P - 1; ++k) tv[k] = p[k];
      for (int i = 0; i < n; ++i) {
        double acc = 0.0;
        for (int jj = 0; jj < n; ++jj) acc += tv[jj] * T[i * n + jj];
        tn[i] = acc;
      }
      for (int k = 0; k < P - 1; ++k) pnew[k] = tn[k];
      pnew[P - 1] = p[P - 1] * (1.0 - tn[P - 1] / pyl::kCVal);
      delay = tn[P] / pyl::kCVal;
    }
    const double phase = pyl::phase_idx(delta_t, p[P - 1], a.nbins, delay);
    int64_t flat = 0;
    for (int k = 0; k < P; ++k)
      flat = flat * a.counts_prev[k] +
             pyl::nearest_idx(pnew[k], a.counts_prev[k], a.lim_lo[k], a.lim_hi[k]);
    out[2 * latter] = (int)flat;
    out[2 * latter + 1] = pyl::shift_int(phase, a.nbins);
    ph[latter] = phase;
    if (boundary) {
      for (int k = 0; k < P; ++k)
        if (a.counts_prev[k] > 1)
          n_idx += pyl::idx_flips(
              pyl::nearest_raw(pnew[k], a.counts_prev[k], a.lim_lo[k], a.lim_hi[k]),
              a.counts_prev[k]);
      n_shift += pyl::shift_flips(phase, a.nbins);
    }
  }
  if (boundary) {
    if (n_idx) atomicAdd(&boundary[0], (unsigned long long)n_idx);
    if (n_shift) atomicAdd(&boundary[1], (unsigned long long)n_shift);
    return;
  }
  table[j] = make_int4(out[0], out[1], out[2], out[3]);
  if (phases) phases[j] = make_double2(ph[0], ph[1]);
}

template <typename T>
struct Vec4;
template <>
struct Vec4<float> {
  using type = float4;
};
template <>
struct Vec4<uint16_t> {
  using type = ushort4;
};

template <typename T>
__device__ __forceinline__ float4 to_f32x4(const typename Vec4<T>::type& v) {
  return make_float4(pyl::Store<T>::load(v.x), pyl::Store<T>::load(v.y),
                     pyl::Store<T>::load(v.z), pyl::Store<T>::load(v.w));
}

template <typename T>
__device__ __forceinline__ void store4(T* __restrict__ dst, int c4, const float4& r) {
  typename Vec4<T>::type v;
  v.x = pyl::Store<T>::store(r.x);
  v.y = pyl::Store<T>::store(r.y);
  v.z = pyl::Store<T>::store(r.z);
  v.w = pyl::Store<T>::store(r.w);
  reinterpret_cast<typename Vec4<T>::type*>(dst)[c4] = v;
}

// Rotated load of the 4-element chunk `c4` (elements 4*c4 .. 4*c4+3) of a row component of
// B elements (a power of two), widened to FP32: element e of the result is
// src[(4*c4 + e - s) mod B]. Branch-free: the two aligned source chunks are always loaded
// (the second one is usually in the same cache line) and the rotation is a select, so lanes
// with different shifts never diverge.
template <int B, typename T = float>
__device__ __forceinline__ float4 rotated_chunk(const T* __restrict__ src, int c4, int s) {
  static_assert((B & (B - 1)) == 0 && B >= 4, "B must be a power of two");
  constexpr int C = B / 4;
  const int m = (4 * c4 - s) & (B - 1);       // first source element
  const int q = m >> 2;                        // source chunk
  const int o = m & 3;                         // offset inside the chunk
  const typename Vec4<T>::type* s4 = reinterpret_cast<const typename Vec4<T>::type*>(src);
  const float4 a = to_f32x4<T>(s4[q]);
  const float4 b = to_f32x4<T>(s4[(q + 1) & (C - 1)]);
  float4 r;
  r.x = o == 0 ? a.x : (o == 1 ? a.y : (o == 2 ? a.z : a.w));
  r.y = o == 0 ? a.y : (o == 1 ? a.z : (o == 2 ? a.w : b.x));
  r.z = o == 0 ? a.z : (o == 1 ? a.w : (o == 2 ? b.x : b.y));
  r.w = o == 0 ? a.w : (o == 1 ? b.x : (o == 2 ? b.y : b.z));
  return r;
}

// Row layouts of the FP32 level kernels. Dense: [2][B] per row (the reference's layout).
// Wrap: [2][B + 4] per row, each component followed by a copy of its first 4 bins, so that
// the 4 consecutive bins (4*c4 - s) mod B .. +3 of any rotation are one contiguous,
// dword-aligned 16-byte load (no select, no second load).
template <int B, bool kWrap>
struct RowLayout {
  static constexpr int64_t kComp = kWrap ? B + 4 : B;  // elements per component
  static constexpr int64_t kRow = 2 * kComp;            // elements per row
};

// Rotated chunk c4 of one component (element e = comp[(4*c4 + e - s) mod B]) in either
// layout.
template <int B, typename T, bool kWrap>
__device__ __forceinline__ float4 row_chunk(const T* __restrict__ comp, int c4, int s) {
  if constexpr (kWrap) {
    static_assert(sizeof(T) == 4, "the wrap layout is FP32 only");
    float4 v;
    __builtin_memcpy(&v, comp + ((4 * c4 - s) & (B - 1)), sizeof(float4));
    return v;
  } else {
    return rotated_chunk<B, T>(comp, c4, s);
  }
}

// XCD-aware block order: hardware deals block i to XCD i % 8. Returns the virtual block
// index such that the blocks one XCD receives form a contiguous slab of the grid, so
// neighbouring blocks (which share source rows) hit the same L2.
__device__ __forceinline__ int64_t xcd_slab_block(int64_t bid, int64_t nblocks) {
  const int64_t per_xcd = (nblocks + 7) / 8;
  const int64_t xcd = bid % 8;
  const int64_t slot = bid / 8;
  const int64_t full = nblocks % 8 == 0 ? 8 : nblocks % 8;  // XCDs with per_xcd blocks
  return xcd < full ? xcd * per_xcd + slot
                    : full * per_xcd + (xcd - full) * (per_xcd - 1) + slot;
}

// Rows per block = kThreads / (B/2) (each row = 2 components x B/4 chunks).
template <int B, typename T, bool kWrapIn, bool kWrapOut>
__global__ void __launch_bounds__(256)
ffa_shift_add_kernel(const T* __restrict__ in, int64_t ncoords_prev,
                     const int4* __restrict__ table, int64_t ncoords_cur, int64_t nseg_out,
                     int64_t nblk_j, T* __restrict__ out) {
  constexpr int CPR = B / 4;          // chunks per component
  constexpr int TPR = 2 * CPR;        // threads per row
  constexpr int RPB = 256 / TPR;      // rows per block
  using Li = RowLayout<B, kWrapIn>;
  using Lo = RowLayout<B, kWrapOut>;
  const int64_t vb = xcd_slab_block(blockIdx.x, gridDim.x);
  const int64_t q = vb / nblk_j;
  const int64_t jb = vb - q * nblk_j;
  if (q >= nseg_out) return;
  const int t = threadIdx.x;
  const int64_t j = jb * RPB + t / TPR;
  if (j >= ncoords_cur) return;
  const int e4 = t % TPR;
  const int comp = e4 / CPR;
  const int c4 = e4 % CPR;
  const int4 tb = table[j];
  const T* tail = in + ((2 * q) * ncoords_prev + tb.x) * Li::kRow + comp * Li::kComp;
  const T* head = in + ((2 * q + 1) * ncoords_prev + tb.z) * Li::kRow + comp * Li::kComp;
  const float4 a = row_chunk<B, T, kWrapIn>(tail, c4, tb.y);
  const float4 b = row_chunk<B, T, kWrapIn>(head, c4, tb.w);
  float4 r;
  r.x = a.x + b.x;
  r.y = a.y + b.y;
  r.z = a.z + b.z;
  r.w = a.w + b.w;
  T* o = out + (q * ncoords_cur + j) * Lo::kRow + comp * Lo::kComp;
  store4<T>(o, c4, r);
  if (kWrapOut && c4 == 0) store4<T>(o + B, 0, r);  // the wrap copy of bins 0..3
}

// Two levels in one pass (FP32, dense or wrap input): output row (q, j) of level L+2 from the
// four rows of segments 4q .. 4q+3 at level L that feed it. table_b[j] names the level-L+1
// rows (tail of segment 2q, head of segment 2q+1) and their shifts; table_a names, for
// each of those, its two level-L rows and shifts. The rotations compose (rotating a sum is
// the sum of the rotated rows), so the tail part is the pair sum the intermediate level
// would store, rounded once, at the composed shifts; likewise the head part; then
// tail + head. Equal bit for bit to two ffa_shift_add_kernel launches; the intermediate
// fold never reaches HBM.
template <int B, bool kWrapIn, bool kWrapOut>
__global__ void __launch_bounds__(256)
ffa_shift_add2_kernel(const float* __restrict__ in, int64_t ncoords_prev2,
                      const int4* __restrict__ table_a, const int4* __restrict__ table_b,
                      int64_t ncoords_cur, int64_t nseg_out, int64_t nblk_j,
                      float* __restrict__ out) {
  constexpr int CPR = B / 4;
  constexpr int TPR = 2 * CPR;
  constexpr int RPB = 256 / TPR;
  using Li = RowLayout<B, kWrapIn>;
  using Lo = RowLayout<B, kWrapOut>;
  constexpr int64_t kRow = Li::kRow;
  const int64_t vb = xcd_slab_block(blockIdx.x, gridDim.x);
  const int64_t q = vb / nblk_j;
  const int64_t jb = vb - q * nblk_j;
  if (q >= nseg_out) return;
  const int t = threadIdx.x;
  const int64_t j = jb * RPB + t / TPR;
  if (j >= ncoords_cur) return;
  const int e4 = t % TPR;
  const int comp = e4 / CPR;
  const int c4 = e4 % CPR;
  const int4 tb = table_b[j];
  const int4 ta = table_a[tb.x];
  const int4 ha = table_a[tb.z];
  const int64_t seg = ncoords_prev2 * kRow;
  const float* base = in + (4 * q) * seg + comp * Li::kComp;
  const float* t1 = base + (int64_t)ta.x * kRow;
  const float* t2 = base + seg + (int64_t)ta.z * kRow;
  const float* h1 = base + 2 * seg + (int64_t)ha.x * kRow;
  const float* h2 = base + 3 * seg + (int64_t)ha.z * kRow;
  const float4 a1 = row_chunk<B, float, kWrapIn>(t1, c4, (tb.y + ta.y) & (B - 1));
  const float4 a2 = row_chunk<B, float, kWrapIn>(t2, c4, (tb.y + ta.w) & (B - 1));
  const float4 b1 = row_chunk<B, float, kWrapIn>(h1, c4, (tb.w + ha.y) & (B - 1));
  const float4 b2 = row_chunk<B, float, kWrapIn>(h2, c4, (tb.w + ha.w) & (B - 1));
  float4 r;
  r.x = (a1.x + a2.x) + (b1.x + b2.x);
  r.y = (a1.y + a2.y) + (b1.y + b2.y);
  r.z = (a1.z + a2.z) + (b1.z + b2.z);
  r.w = (a1.w + a2.w) + (b1.w + b2.w);
  float* o = out + (q * ncoords_cur + j) * Lo::kRow + comp * Lo::kComp;
  store4<float>(o, c4, r);
  if (kWrapOut && c4 == 0) store4<float>(o + B, 0, r);  // the wrap copy of bins 0..3
}

// Generic nbins: one thread per output element.
__global__ void ffa_shift_add_generic_kernel(const float* __restrict__ in, int64_t ncoords_prev,
                                             const int4* __restrict__ table, int64_t ncoords_cur,
                                             int nbins, int64_t nseg_out,
                                             float* __restrict__ out) {
  const int64_t E = 2 * (int64_t)nbins;
  const int64_t total = nseg_out * ncoords_cur * E;
  for (int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; gid < total;
       gid += (int64_t)gridDim.x * blockDim.x) {
    const int64_t row = gid / E;
    const int e = (int)(gid - row * E);
    const int64_t q = row / ncoords_cur;
    const int64_t j = row - q * ncoords_cur;
    const int comp = e / nbins;
    const int b = e - comp * nbins;
    const int4 tb = table[j];
    const float* tail = in + ((2 * q) * ncoords_prev + tb.x) * E + comp * nbins;
    const float* head = in + ((2 * q + 1) * ncoords_prev + tb.z) * E + comp * nbins;
    const int it = ((b - tb.y) % nbins + nbins) % nbins;
    const int ih = ((b - tb.w) % nbins + nbins) % nbins;
    out[gid] = tail[it] + head[ih];
  }
}


// Last FFA level fused with scoring: the output row of (segment pair q, set j) is formed
// in registers, normalised (e / sqrt(v), ffa.py:411) and scored (scoring.py:99-165)
// without being stored. A block owns 64 output rows of one pair; 16 lanes cover one row
// (4 bins x 2 components per lane, float4 loads with the rotation of rotated_chunk), so a
// wavefront forms 4 rows per step; the normalised rows are transposed into LDS and every
// thread then scans and scores one row.
template <int B, typename T, bool kWrapIn>
__global__ void __launch_bounds__(pyl::kScoreBlock)
ffa_shift_add_score_kernel(const T* __restrict__ in, int64_t ncoords_prev,
                           const int4* __restrict__ table, int64_t ncoords_cur,
                           int64_t nblk_j, pyl::ScoreConst k, float* __restrict__ out) {
  constexpr int CPR = B / 4;                 // float4 chunks per component
  constexpr int RPS = pyl::kScoreBlock / CPR;  // rows formed per step
  extern __shared__ float S[];
  __shared__ int4 s_tab[pyl::kScoreBlock];
  const int tid = threadIdx.x;
  const int64_t q = blockIdx.x / nblk_j;
  const int64_t j0 = (blockIdx.x - q * nblk_j) * pyl::kScoreBlock;
  const int nvalid = (int)std::min<int64_t>(pyl::kScoreBlock, ncoords_cur - j0);
  if (tid < nvalid) s_tab[tid] = table[j0 + tid];
  __syncthreads();
  using Li = RowLayout<B, kWrapIn>;
  const int c4 = tid % CPR;
  const int rsub = tid / CPR;
#pragma unroll 2
  for (int r0 = 0; r0 < nvalid; r0 += RPS) {
    const int r = r0 + rsub;
    if (r < nvalid) {
      const int4 tb = s_tab[r];
      const T* tail = in + ((2 * q) * ncoords_prev + tb.x) * Li::kRow;
      const T* head = in + ((2 * q + 1) * ncoords_prev + tb.z) * Li::kRow;
      const float4 te = row_chunk<B, T, kWrapIn>(tail, c4, tb.y);
      const float4 he = row_chunk<B, T, kWrapIn>(head, c4, tb.w);
      const float4 tv = row_chunk<B, T, kWrapIn>(tail + Li::kComp, c4, tb.y);
      const float4 hv = row_chunk<B, T, kWrapIn>(head + Li::kComp, c4, tb.w);
      const int b = 4 * c4;
      // The level's fold element as it would be stored (identity for FP32, BF16 rounding
      // for BF16 storage), then e / sqrt(v).
      auto st = [](float x) { return pyl::Store<T>::load(pyl::Store<T>::store(x)); };
      const float4 x = pyl::norm_ev4(
          make_float4(st(te.x + he.x), st(te.y + he.y), st(te.z + he.z), st(te.w + he.w)),
          make_float4(st(tv.x + hv.x), st(tv.y + hv.y), st(tv.z + hv.z), st(tv.w + hv.w)));
      S[(b + 0) * pyl::kScoreStride + r] = x.x;
      S[(b + 1) * pyl::kScoreStride + r] = x.y;
      S[(b + 2) * pyl::kScoreStride + r] = x.z;
      S[(b + 3) * pyl::kScoreStride + r] = x.w;
    }
  }
  __syncthreads();
  if (tid < nvalid)
    pyl::score_column<true>(k, B, S, tid, out + (q * ncoords_cur + j0 + tid) * k.nw);
}

// Round-based variant of the fused last level (nbins 32 or 64, at most kRoundsMaxNW widths,
// widths <= WC). A block owns a tile of 256 output rows of one segment pair: tile_o_rows
// (up to 16) "outer" x 256/tile_o_rows "inner" rows, inner = index along the last
// (frequency) axis, outer = the product of the other axes. Rows at neighbouring outer
// indices (e.g. accelerations) resolve to the same source rows, so the tile reuses each
// source row from L2 instead of re-reading it from HBM; blocks are dealt to the 8 XCDs in
// contiguous slabs so that neighbouring tiles (which share boundary rows) meet in one L2.
//
// The row gather is latency-bound, so the kernel keeps its LDS footprint small (occupancy
// is what hides the gather latency): rows are formed in B/32 rounds of 32 bins. In a round
// 8 lanes cover one row's 32 bins (rotated float4 loads, 8 rows per wavefront step), write
// the normalised values transposed into a 32 x 257-word LDS buffer, and after a barrier
// every thread extends its own row's sequential FP32 prefix sum in registers. Scoring then
// slides a window of 16 + WC prefix values per row through the same LDS buffer (only the
// owning thread touches its column, so no further barriers): for window r the buffer holds
// S[16r + j], j < 16 + WC (S[B + t] = S[t] + S[B-1], the one-wrap extension), and every
// width w updates its running maximum of S[i + w] - S[i] over the window's 16 i, with S[i]
// from registers and S[i + w] read at a runtime offset. The per-width maxima sit in a small
// register array indexed by the (uniform) width counter. The i order, and so every FP32
// operation, is that of the reference's sequential loop.
//
// kTwo: the last two levels in one pass. ``in`` then holds the fold two levels before the
// last (segments 4q .. 4q+3 feed output segment q) and ``table2`` is the resolve table of
// the level before the last: an output row's tail row is the sum of two such rows (its
// table2 entry), rotated by the tail shift, and likewise its head row. Each pair is added
// and stored exactly as the intermediate level would store it, so the result is that of
// the two levels run one after the other; the intermediate fold is never written.
//
// kPack (kTwo, input segments of fewer than 2^32 bytes): a row's four source rows and their
// composed shifts sit in LDS as one uint4 of (row << 8 | 4 * shift) words, 16 B per row
// instead of the 48 B of its table and table2 entries. The block then needs 37 KB of LDS,
// so four blocks fit a CU (the register limit) instead of three.
template <int B, typename T, int WC, bool kWrapIn, bool kTwo = false, bool kPack = false>
__global__ void __launch_bounds__(256)
ffa_shift_add_score_rounds_kernel(const T* __restrict__ in, int64_t ncoords_prev,
                                  const int4* __restrict__ table, int64_t ncoords_cur,
                                  int64_t n_inner, int tile_o_rows, int64_t tiles_inner,
                                  int64_t tiles_per_pair, pyl::ScoreConst k,
                                  float* __restrict__ out,
                                  const int4* __restrict__ table2 = nullptr,
                                  int64_t ncoords_prev2 = 0) {
  static_assert(!kPack || kTwo, "packed descriptors are for two fused levels");
  constexpr int kRows = 256;
  constexpr int kLds = kRows + 1;
  constexpr int kSeg = 16;            // scoring window step
  constexpr int kRounds = B / kSeg;    // scoring windows
  constexpr int kWin = kSeg + WC;      // LDS rows of a scoring window
  constexpr int kSeg1 = 32;            // bins formed per gather round
  constexpr int kRounds1 = B / kSeg1;  // gather rounds
  constexpr int kLpr = kSeg1 / 4;      // lanes per row in a gather round
  constexpr int kRps = 64 / kLpr;      // rows per wavefront step
  constexpr int kSteps = 64 / kRps;    // steps per wavefront and round
  constexpr int kLdsRows = kWin > kSeg1 ? kWin : kSeg1;
  static_assert(B % kSeg1 == 0 && WC < B, "unsupported round layout");
  extern __shared__ float S[];  // [kLdsRows][kLds]
  __shared__ int4 s_tab[kPack ? 1 : kRows];
  const int tid = threadIdx.x;
  const int64_t vb = xcd_slab_block(blockIdx.x, gridDim.x);
  const int64_t q = vb / tiles_per_pair;
  const int64_t tile = vb - q * tiles_per_pair;
  const int64_t tile_o = tile / tiles_inner;
  const int64_t tile_i = tile - tile_o * tiles_inner;
  const int64_t n_outer = ncoords_cur / n_inner;
  const int tile_i_rows = kRows / tile_o_rows;
  const int64_t outer = tile_o * tile_o_rows + tid / tile_i_rows;
  const int64_t inner = tile_i * tile_i_rows + tid % tile_i_rows;
  const bool ok = outer < n_outer && inner < n_inner;
  const int64_t j = outer * n_inner + inner;
  const int4 t0 = ok ? table[j] : make_int4(-1, 0, 0, 0);
  __shared__ int4 s_tab2[kTwo && !kPack ? 2 * kRows : 1];
  __shared__ uint4 s_pk[kPack ? kRows : 1];
  constexpr uint32_t kNoRow = 0xffffffffu;  // above every packed word when rows < 2^24
  if constexpr (kPack) {
    uint4 pk = make_uint4(kNoRow, 0u, 0u, 0u);
    if (ok) {
      const int4 ta = table2[t0.x];
      const int4 ha = table2[t0.z];
      pk.x = ((uint32_t)ta.x << 8) | (uint32_t)(4 * ((t0.y + ta.y) & (B - 1)));
      pk.y = ((uint32_t)ta.z << 8) | (uint32_t)(4 * ((t0.y + ta.w) & (B - 1)));
      pk.z = ((uint32_t)ha.x << 8) | (uint32_t)(4 * ((t0.w + ha.y) & (B - 1)));
      pk.w = ((uint32_t)ha.z << 8) | (uint32_t)(4 * ((t0.w + ha.w) & (B - 1)));
    }
    s_pk[tid] = pk;
  } else {
    s_tab[tid] = t0;
    if constexpr (kTwo) {
      if (ok) {
        s_tab2[2 * tid] = table2[t0.x];
        s_tab2[2 * tid + 1] = table2[t0.z];
      }
    }
  }
  using Li = RowLayout<B, kWrapIn>;
  const int wave = tid >> 6;
  const int lane = tid & 63;
  const int c4 = lane % kLpr;   // float4 chunk inside the round's bins
  const int rsub = lane / kLpr;  // row inside the wavefront step
  const int64_t seg_rows = (kTwo ? ncoords_prev2 : ncoords_prev) * Li::kRow;
  const T* base_t = in + (kTwo ? 4 * q : 2 * q) * seg_rows;
  const T* base_h = in + (kTwo ? 4 * q + 2 : 2 * q + 1) * seg_rows;
  float x[B];  // this thread's row: prefix sums S[0..B-1]
#pragma unroll
  for (int rd = 0; rd < kRounds1; ++rd) {
    __syncthreads();  // s_tab ready (rd == 0) / previous round consumed
#pragma unroll
    for (int step = 0; step < kSteps; ++step) {
      const int r = wave * 64 + step * kRps + rsub;
      int4 tb;
      uint4 pk;
      if constexpr (kPack) pk = s_pk[r];
      else tb = s_tab[r];
      if (kPack ? pk.x != kNoRow : tb.x >= 0) {
        const int cc = rd * kLpr + c4;
        // The level's fold element as it would be stored (identity for FP32, BF16 rounding
        // for BF16 storage), then e / sqrt(v).
        auto st = [](float v) { return pyl::Store<T>::load(pyl::Store<T>::store(v)); };
        float4 te, he, tv, hv;
        if constexpr (kPack) {
          auto add = [&](float4 a, float4 b) {
            return make_float4(st(a.x + b.x), st(a.y + b.y), st(a.z + b.z), st(a.w + b.w));
          };
          if constexpr (kWrapIn) {
            // 32-bit byte offsets from the four segments' (uniform) bases: a segment holds
            // fewer than 2^32 bytes here, so the loads take an SGPR base and a VGPR offset
            // and the row address is one 24-bit multiply-add. The byte offset of the
            // rotated chunk, 4 ((4 cc - s) mod B), is (16 cc - w) mod 4B masked to whole
            // words, since w = 256 row + 4 s and 4B divides 256.
            constexpr uint32_t kRowB = (uint32_t)(Li::kRow * sizeof(float));
            constexpr uint32_t kCompB = (uint32_t)(Li::kComp * sizeof(float));
            const char* bt1 = reinterpret_cast<const char*>(base_t);
            const char* bt2 = reinterpret_cast<const char*>(base_t + seg_rows);
            const char* bh1 = reinterpret_cast<const char*>(base_h);
            const char* bh2 = reinterpret_cast<const char*>(base_h + seg_rows);
            auto off = [&](uint32_t w) {
              return __umul24(w >> 8, kRowB) + ((16u * (uint32_t)cc - w) & (4u * (B - 1)));
            };
            auto ld = [](const char* b, uint32_t o) {
              float4 v;
              __builtin_memcpy(&v, b + o, sizeof(float4));
              return v;
            };
            const uint32_t o1 = off(pk.x), o2 = off(pk.y), o3 = off(pk.z), o4 = off(pk.w);
            te = add(ld(bt1, o1), ld(bt2, o2));
            he = add(ld(bh1, o3), ld(bh2, o4));
            tv = add(ld(bt1 + kCompB, o1), ld(bt2 + kCompB, o2));
            hv = add(ld(bh1 + kCompB, o3), ld(bh2 + kCompB, o4));
          } else {
            const T* t1 = base_t + (int64_t)(pk.x >> 8) * Li::kRow;
            const T* t2 = base_t + seg_rows + (int64_t)(pk.y >> 8) * Li::kRow;
            const T* h1 = base_h + (int64_t)(pk.z >> 8) * Li::kRow;
            const T* h2 = base_h + seg_rows + (int64_t)(pk.w >> 8) * Li::kRow;
            const int st1 = (int)((pk.x >> 2) & 63u), st2 = (int)((pk.y >> 2) & 63u);
            const int sh1 = (int)((pk.z >> 2) & 63u), sh2 = (int)((pk.w >> 2) & 63u);
            te = add(row_chunk<B, T, kWrapIn>(t1, cc, st1), row_chunk<B, T, kWrapIn>(t2, cc, st2));
            he = add(row_chunk<B, T, kWrapIn>(h1, cc, sh1), row_chunk<B, T, kWrapIn>(h2, cc, sh2));
            tv = add(row_chunk<B, T, kWrapIn>(t1 + Li::kComp, cc, st1),
                     row_chunk<B, T, kWrapIn>(t2 + Li::kComp, cc, st2));
            hv = add(row_chunk<B, T, kWrapIn>(h1 + Li::kComp, cc, sh1),
                     row_chunk<B, T, kWrapIn>(h2 + Li::kComp, cc, sh2));
          }
        } else if constexpr (kTwo) {
          const int4 ta = s_tab2[2 * r];
          const int4 ha = s_tab2[2 * r + 1];
          const T* t1 = base_t + (int64_t)ta.x * Li::kRow;
          const T* t2 = base_t + seg_rows + (int64_t)ta.z * Li::kRow;
          const T* h1 = base_h + (int64_t)ha.x * Li::kRow;
          const T* h2 = base_h + seg_rows + (int64_t)ha.z * Li::kRow;
          const int st1 = (tb.y + ta.y) & (B - 1), st2 = (tb.y + ta.w) & (B - 1);
          const int sh1 = (tb.w + ha.y) & (B - 1), sh2 = (tb.w + ha.w) & (B - 1);
          auto add = [&](float4 a, float4 b) {
            return make_float4(st(a.x + b.x), st(a.y + b.y), st(a.z + b.z), st(a.w + b.w));
          };
          te = add(row_chunk<B, T, kWrapIn>(t1, cc, st1), row_chunk<B, T, kWrapIn>(t2, cc, st2));
          he = add(row_chunk<B, T, kWrapIn>(h1, cc, sh1), row_chunk<B, T, kWrapIn>(h2, cc, sh2));
          tv = add(row_chunk<B, T, kWrapIn>(t1 + Li::kComp, cc, st1),
                   row_chunk<B, T, kWrapIn>(t2 + Li::kComp, cc, st2));
          hv = add(row_chunk<B, T, kWrapIn>(h1 + Li::kComp, cc, sh1),
                   row_chunk<B, T, kWrapIn>(h2 + Li::kComp, cc, sh2));
        } else {
          const T* tail = base_t + (int64_t)tb.x * Li::kRow;
          const T* head = base_h + (int64_t)tb.z * Li::kRow;
          te = row_chunk<B, T, kWrapIn>(tail, cc, tb.y);
          he = row_chunk<B, T, kWrapIn>(head, cc, tb.w);
          tv = row_chunk<B, T, kWrapIn>(tail + Li::kComp, cc, tb.y);
          hv = row_chunk<B, T, kWrapIn>(head + Li::kComp, cc, tb.w);
        }
        const int b = 4 * c4;
        const float4 x = pyl::norm_ev4(
            make_float4(st(te.x + he.x), st(te.y + he.y), st(te.z + he.z), st(te.w + he.w)),
            make_float4(st(tv.x + hv.x), st(tv.y + hv.y), st(tv.z + hv.z), st(tv.w + hv.w)));
        S[(b + 0) * kLds + r] = x.x;
        S[(b + 1) * kLds + r] = x.y;
        S[(b + 2) * kLds + r] = x.z;
        S[(b + 3) * kLds + r] = x.w;
      }
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < kSeg1; ++b) {
      const float v = S[b * kLds + tid];
      const int i = rd * kSeg1 + b;
      if (i == 0) x[0] = v;
      else x[i] = x[i - 1] + v;  // scoring.py:53-55, sequential
    }
  }
  if (!ok) return;  // no barrier below
  const float total = x[B - 1];
  const int nw = k.nw;
  float dmax[kRoundsMaxNW];
#pragma unroll
  for (int m = 0; m < kRoundsMaxNW; ++m) dmax[m] = -FLT_MAX;
#pragma unroll
  for (int rd = 0; rd < kRounds; ++rd) {
#pragma unroll
    for (int jw = 1; jw < kWin; ++jw) {
      const int g = rd * kSeg + jw;
      S[jw * kLds + tid] = g < B ? x[g < B ? g : 0] : (x[g >= B ? g - B : 0] + total);
    }
#pragma unroll 1
    for (int m = 0; m < nw; ++m) {
      const float* Sw = S + k.widths[m] * kLds + tid;
      float dm = dmax[m];
#pragma unroll
      for (int b = 0; b < kSeg; b += 2) {
        const float d0 = Sw[b * kLds] - x[rd * kSeg + b];
        const float d1 = Sw[(b + 1) * kLds] - x[rd * kSeg + b + 1];
        dm = pyl::max3_seq(dm, d0, d1);
      }
      dmax[m] = dm;
    }
  }
  // A zero maximum takes the sign of the first zero difference in the reference's order
  // (later zeros never replace it): a second pass over the windows finds it. Real data
  // reach this only for all-zero stretches of a profile.
  uint32_t zero = 0;
#pragma unroll 1
  for (int m = 0; m < nw; ++m) zero |= (dmax[m] == 0.0f ? 1u : 0u) << m;
  if (zero) {
    uint32_t neg = 0;
#pragma unroll
    for (int rd = 0; rd < kRounds; ++rd) {
#pragma unroll
      for (int jw = 1; jw < kWin; ++jw) {
        const int g = rd * kSeg + jw;
        S[jw * kLds + tid] = g < B ? x[g < B ? g : 0] : (x[g >= B ? g - B : 0] + total);
      }
      S[tid] = x[rd * kSeg];
#pragma unroll 1
      for (int m = 0; m < nw; ++m) {
        if (!((zero >> m) & 1u)) continue;
        const float* Sw = S + k.widths[m] * kLds + tid;
        for (int b = 0; b < kSeg; ++b) {
          const float d = Sw[b * kLds] - S[b * kLds + tid];
          if (d == 0.0f) {
            neg |= (__float_as_uint(d) >> 31) << m;
            zero &= ~(1u << m);
            break;
          }
        }
      }
    }
#pragma unroll 1
    for (int m = 0; m < nw; ++m)
      if (dmax[m] == 0.0f) dmax[m] = ((neg >> m) & 1u) ? -0.0f : 0.0f;
  }
  float* o = out + (q * ncoords_cur + j) * nw;
#pragma unroll 1
  for (int m = 0; m < nw; ++m) {
    const double t1 = k.hpb[m] * (double)dmax[m];
    const double t2 = k.bb[m] * (double)total;
    o[m] = (float)((t1 - t2) * k.inv_std);
  }
}

}  // namespace

static int ffa_resolve_impl(int32_t nparams, const double* params_dev,
                            const int64_t* counts_cur, const int64_t* counts_prev,
                            const double* param_limits, int32_t ffa_level, double tseg_brute,
                            int32_t nbins, int32_t* table_dev, double* phases_dev,
                            unsigned long long* boundary_dev, void* stream) {
  PYL_REQUIRE(nparams >= 1 && nparams <= kMaxP, "nparams must be in [1, 5], got %d", nparams);
  PYL_REQUIRE(params_dev && counts_cur && counts_prev && param_limits &&
                  (table_dev || boundary_dev),
              "null pointer argument");
  PYL_REQUIRE(ffa_level >= 1 && ffa_level < 62, "ffa_level must be >= 1, got %d", ffa_level);
  PYL_REQUIRE(nbins > 0, "Number of bins must be positive.");
  ResolveArgs a{};
  a.P = nparams;
  a.nbins = nbins;
  a.level = ffa_level;
  a.tseg_brute = tseg_brute;
  int64_t ncur = 1, nprev = 1, off = 0;
  for (int k = 0; k < nparams; ++k) {
    PYL_REQUIRE(counts_cur[k] > 0 && counts_prev[k] > 0, "grid counts must be positive");
    PYL_REQUIRE(param_limits[2 * k] < param_limits[2 * k + 1],
                "param_limits must have min < max");
    a.counts_cur[k] = counts_cur[k];
    a.counts_prev[k] = counts_prev[k];
    a.offs[k] = off;
    off += counts_cur[k];
    a.lim_lo[k] = param_limits[2 * k];
    a.lim_hi[k] = param_limits[2 * k + 1];
    ncur *= counts_cur[k];
    nprev *= counts_prev[k];
  }
  PYL_REQUIRE(nprev <= 0x7fffffff, "previous level has too many coordinates (%lld)",
              (long long)nprev);
  a.ncoords = ncur;
  // fold.py:660-664: delta_t of each half; transforms.py:41-43: the matrix dt**p / p! is a
  // per-level constant, built once on the host.
  for (int latter = 0; latter < 2; ++latter) {
    double dt;
    if (nparams == 1)
      dt = (double)((int64_t)latter * ((int64_t)1 << (ffa_level - 1))) * tseg_brute;
    else
      dt = ((latter - 0.5) * (double)((int64_t)1 << (ffa_level - 1))) * tseg_brute;
    a.delta_t[latter] = dt;
    pyl::taylor_matrix(nparams + 1, dt, a.tmat[latter]);
  }
  hipStream_t s = pyl::as_stream(stream);
  const int threads = 128;
  ffa_resolve_kernel<<<(unsigned)pyl::ceil_div(ncur, threads), threads, 0, s>>>(
      params_dev, a, reinterpret_cast<int4*>(table_dev),
      reinterpret_cast<double2*>(phases_dev), boundary_dev);
  PYL_CHECK_LAUNCH("ffa_resolve_kernel");
  return PYL_OK;
}

extern "C" int pyl_ffa_resolve(int32_t nparams, const double* params_dev,
                               const int64_t* counts_cur, const int64_t* counts_prev,
                               const double* param_limits, int32_t ffa_level, double tseg_brute,
                               int32_t nbins, int32_t* table_dev, double* phases_dev,
                               void* stream) {
  return ffa_resolve_impl(nparams, params_dev, counts_cur, counts_prev, param_limits, ffa_level,
                          tseg_brute, nbins, table_dev, phases_dev, nullptr, stream);
}

extern "C" int pyl_ffa_resolve_boundary_count(int32_t nparams, const double* params_dev,
                                              const int64_t* counts_cur,
                                              const int64_t* counts_prev,
                                              const double* param_limits, int32_t ffa_level,
                                              double tseg_brute, int32_t nbins,
                                              unsigned long long* counts_dev, void* stream) {
  return ffa_resolve_impl(nparams, params_dev, counts_cur, counts_prev, param_limits, ffa_level,
                          tseg_brute, nbins, nullptr, nullptr, counts_dev, stream);
}

template <int B, typename T, bool kWrapIn, bool kWrapOut>
static void launch_shift_add(const T* in, int64_t nprev, const int4* tb, int64_t ncur,
                             int64_t nseg_out, T* out, hipStream_t s) {
  constexpr int RPB = 256 / (B / 2);
  const int64_t nblk_j = pyl::ceil_div(ncur, RPB);
  const int64_t nblocks = nblk_j * nseg_out;
  ffa_shift_add_kernel<B, T, kWrapIn, kWrapOut><<<(unsigned)nblocks, 256, 0, s>>>(
      in, nprev, tb, ncur, nseg_out, nblk_j, out);
}

template <int B>
static void launch_shift_add_f32(const float* in, bool wrap_in, int64_t nprev, const int4* tb,
                                 int64_t ncur, int64_t nseg_out, float* out, bool wrap_out,
                                 hipStream_t s) {
  if (wrap_in) {
    if (wrap_out) launch_shift_add<B, float, true, true>(in, nprev, tb, ncur, nseg_out, out, s);
    else launch_shift_add<B, float, true, false>(in, nprev, tb, ncur, nseg_out, out, s);
  } else {
    if (wrap_out) launch_shift_add<B, float, false, true>(in, nprev, tb, ncur, nseg_out, out, s);
    else launch_shift_add<B, float, false, false>(in, nprev, tb, ncur, nseg_out, out, s);
  }
}

static bool valid_layout(int32_t layout) {
  return layout == PYL_FOLD_DENSE || layout == PYL_FOLD_WRAP;
}

extern "C" int pyl_ffa_shift_add_layout(const float* fold_in, int32_t in_layout, int64_t nseg_in,
                                        int64_t ncoords_prev, const int32_t* table_dev,
                                        int64_t ncoords_cur, int32_t nbins, float* fold_out,
                                        int32_t out_layout, void* stream) {
  PYL_REQUIRE(fold_in && table_dev && fold_out, "null device pointer");
  PYL_REQUIRE(nseg_in >= 2 && nseg_in % 2 == 0, "nsegments must be even and >= 2, got %lld",
              (long long)nseg_in);
  PYL_REQUIRE(ncoords_prev > 0 && ncoords_cur > 0, "empty parameter grid");
  PYL_REQUIRE(nbins > 0, "Number of bins must be positive.");
  PYL_REQUIRE(valid_layout(in_layout) && valid_layout(out_layout), "invalid fold layout");
  const bool wi = in_layout == PYL_FOLD_WRAP, wo = out_layout == PYL_FOLD_WRAP;
  PYL_REQUIRE(!(wi || wo) || nbins == 32 || nbins == 64 || nbins == 128 || nbins == 256,
              "the wrap layout supports nbins in {32, 64, 128, 256}, got %d", nbins);
  hipStream_t s = pyl::as_stream(stream);
  const int64_t nseg_out = nseg_in / 2;
  const int4* tb = reinterpret_cast<const int4*>(table_dev);
  switch (nbins) {
    case 32: launch_shift_add_f32<32>(fold_in, wi, ncoords_prev, tb, ncoords_cur, nseg_out, fold_out, wo, s); break;
    case 64: launch_shift_add_f32<64>(fold_in, wi, ncoords_prev, tb, ncoords_cur, nseg_out, fold_out, wo, s); break;
    case 128: launch_shift_add_f32<128>(fold_in, wi, ncoords_prev, tb, ncoords_cur, nseg_out, fold_out, wo, s); break;
    case 256: launch_shift_add_f32<256>(fold_in, wi, ncoords_prev, tb, ncoords_cur, nseg_out, fold_out, wo, s); break;
    default: {
      const int64_t total = nseg_out * ncoords_cur * 2 * (int64_t)nbins;
      const int64_t blocks = std::min<int64_t>(pyl::ceil_div(total, 256), 1 << 20);
      ffa_shift_add_generic_kernel<<<(unsigned)blocks, 256, 0, s>>>(
          fold_in, ncoords_prev, tb, ncoords_cur, nbins, nseg_out, fold_out);
    }
  }
  PYL_CHECK_LAUNCH("ffa_shift_add_kernel");
  return PYL_OK;
}

extern "C" int pyl_ffa_shift_add(const float* fold_in, int64_t nseg_in, int64_t ncoords_prev,
                                 const int32_t* table_dev, int64_t ncoords_cur, int32_t nbins,
                                 float* fold_out, void* stream) {
  return pyl_ffa_shift_add_layout(fold_in, PYL_FOLD_DENSE, nseg_in, ncoords_prev, table_dev,
                                  ncoords_cur, nbins, fold_out, PYL_FOLD_DENSE, stream);
}

// Outer rows per 256-row tile of the round-based fused kernels: 0 = the default (16 x 16
// tiles), else the value the autotuner measured best (pyl_ffa_set_tile_rows). Fewer outer
// rows on short outer axes. The rows a tile holds never change a result.
static std::atomic<int> g_tile_outer{0};

// Packed gather descriptors in the two-level kernel (kPack): on by default wherever the
// row indices fit; pyl_ffa_set_packed_tables(0) selects the unpacked tables (a check that
// both variants agree). Neither changes a result.
static std::atomic<int> g_packed{1};

static int tile_outer_rows(int64_t n_outer) {
  int to = g_tile_outer.load(std::memory_order_relaxed);
  if (to <= 0) to = 16;
  while (to > 1 && n_outer < to) to /= 2;
  return to;
}

template <int B, typename T, int WC, bool kWrapIn>
static void launch_rounds(const T* in, int64_t nprev, const int4* tb, int64_t ncur,
                          int64_t n_inner, int64_t nseg_out, const pyl::ScoreConst& k,
                          float* out, hipStream_t s) {
  const int64_t n_outer = ncur / n_inner;
  const int to = tile_outer_rows(n_outer);
  const int64_t tiles_inner = pyl::ceil_div(n_inner, 256 / to);
  const int64_t tiles_outer = pyl::ceil_div(n_outer, to);
  const int64_t per_pair = tiles_inner * tiles_outer;
  const size_t lds = sizeof(float) * (size_t)std::max(16 + WC, 32) * 257;
  ffa_shift_add_score_rounds_kernel<B, T, WC, kWrapIn>
      <<<(unsigned)(per_pair * nseg_out), 256, lds, s>>>(in, nprev, tb, ncur, n_inner, to,
                                                          tiles_inner, per_pair, k, out);
}

template <int B, int WC, bool kWrapIn>
static void launch_rounds2(const float* in, int64_t nprev2, const int4* tb2, int64_t nprev,
                           const int4* tb, int64_t ncur, int64_t n_inner, int64_t nseg_out,
                           const pyl::ScoreConst& k, float* out, hipStream_t s) {
  const int64_t n_outer = ncur / n_inner;
  const int to = tile_outer_rows(n_outer);
  const int64_t tiles_inner = pyl::ceil_div(n_inner, 256 / to);
  const int64_t tiles_outer = pyl::ceil_div(n_outer, to);
  const int64_t per_pair = tiles_inner * tiles_outer;
  const size_t lds = sizeof(float) * (size_t)std::max(16 + WC, 32) * 257;
  const unsigned grid = (unsigned)(per_pair * nseg_out);
  // Row indices fit the packed (row << 8 | 4 shift) words, and a segment's bytes the 32-bit
  // offsets of the wrap-layout gather (so a row index also fits the 24-bit multiply).
  const bool fits = nprev2 < ((int64_t)1 << 24) &&
                    nprev2 * RowLayout<B, kWrapIn>::kRow * (int64_t)sizeof(float) <
                        ((int64_t)1 << 32);
  if (fits && g_packed.load(std::memory_order_relaxed))
    ffa_shift_add_score_rounds_kernel<B, float, WC, kWrapIn, true, true>
        <<<grid, 256, lds, s>>>(in, nprev, tb, ncur, n_inner, to, tiles_inner, per_pair, k,
                                out, tb2, nprev2);
  else
    ffa_shift_add_score_rounds_kernel<B, float, WC, kWrapIn, true, false>
        <<<grid, 256, lds, s>>>(in, nprev, tb, ncur, n_inner, to, tiles_inner, per_pair, k,
                                out, tb2, nprev2);
}

template <int B, bool kWrapIn>
static void launch_two(const float* in, int64_t nprev2, const int4* tb2, int64_t nprev,
                       const int4* tb, int64_t ncur, int64_t n_inner, int64_t nseg_out,
                       const pyl::ScoreConst& k, float* out, hipStream_t s) {
  if (k.wmax <= 16)
    return launch_rounds2<B, 16, kWrapIn>(in, nprev2, tb2, nprev, tb, ncur, n_inner, nseg_out, k, out, s);
  if (k.wmax <= B / 2)
    return launch_rounds2<B, B / 2, kWrapIn>(in, nprev2, tb2, nprev, tb, ncur, n_inner, nseg_out, k, out, s);
  launch_rounds2<B, B - 1, kWrapIn>(in, nprev2, tb2, nprev, tb, ncur, n_inner, nseg_out, k, out, s);
}

template <int B, typename T, bool kWrapIn>
static void launch_shift_add_score(const T* in, int64_t nprev, const int4* tb, int64_t ncur,
                                   int64_t n_inner, int64_t nseg_out, const pyl::ScoreConst& k,
                                   float* out, hipStream_t s) {
  if constexpr (B <= 64) {
    // Round-based kernel: the scoring window is 16 + WC prefix values, WC >= widest width.
    if (k.nw <= kRoundsMaxNW) {
      if (k.wmax <= 16)
        return launch_rounds<B, T, 16, kWrapIn>(in, nprev, tb, ncur, n_inner, nseg_out, k, out, s);
      if (k.wmax <= B / 2)
        return launch_rounds<B, T, B / 2, kWrapIn>(in, nprev, tb, ncur, n_inner, nseg_out, k, out, s);
      return launch_rounds<B, T, B - 1, kWrapIn>(in, nprev, tb, ncur, n_inner, nseg_out, k, out, s);
    }
  }
  const int64_t nblk_j = pyl::ceil_div(ncur, pyl::kScoreBlock);
  const size_t lds = sizeof(float) * (size_t)(B + k.wmax) * pyl::kScoreStride;
  ffa_shift_add_score_kernel<B, T, kWrapIn>
      <<<(unsigned)(nblk_j * nseg_out), pyl::kScoreBlock, lds, s>>>(in, nprev, tb, ncur,
                                                                      nblk_j, k, out);
}

template <int B, typename T>
static void launch_score_any(const T* in, bool wrap_in, int64_t nprev, const int4* tb,
                             int64_t ncur, int64_t n_inner, int64_t nseg_out,
                             const pyl::ScoreConst& k, float* out, hipStream_t s) {
  if constexpr (sizeof(T) == 4) {
    if (wrap_in)
      return launch_shift_add_score<B, T, true>(in, nprev, tb, ncur, n_inner, nseg_out, k, out, s);
  }
  launch_shift_add_score<B, T, false>(in, nprev, tb, ncur, n_inner, nseg_out, k, out, s);
}

template <typename T>
static int shift_add_score(const T* fold_in, int32_t in_layout, int64_t nseg_in,
                           int64_t ncoords_prev, const int32_t* table_dev, int64_t ncoords_cur,
                           int64_t n_inner, int32_t nbins, const int64_t* widths, int32_t nw,
                           double stdnoise, float* snr_out, void* stream) {
  PYL_REQUIRE(fold_in && table_dev && snr_out, "null device pointer");
  PYL_REQUIRE(nseg_in >= 2 && nseg_in % 2 == 0, "nsegments must be even and >= 2, got %lld",
              (long long)nseg_in);
  PYL_REQUIRE(ncoords_prev > 0 && ncoords_cur > 0, "empty parameter grid");
  PYL_REQUIRE(nbins == 32 || nbins == 64 || nbins == 128,
              "fused level + score supports nbins in {32, 64, 128}, got %d", nbins);
  PYL_REQUIRE(n_inner >= 1 && ncoords_cur % n_inner == 0,
              "n_inner must divide ncoords_cur (last-axis grid count)");
  PYL_REQUIRE(valid_layout(in_layout), "invalid fold layout");
  PYL_REQUIRE(in_layout == PYL_FOLD_DENSE || sizeof(T) == 4, "the wrap layout is FP32 only");
  pyl::ScoreConst k;
  int rc = pyl::build_score_const(nbins, widths, nw, stdnoise, &k);
  if (rc != PYL_OK) return rc;
  hipStream_t s = pyl::as_stream(stream);
  const int64_t nseg_out = nseg_in / 2;
  const int4* tb = reinterpret_cast<const int4*>(table_dev);
  const bool wi = in_layout == PYL_FOLD_WRAP;
  switch (nbins) {
    case 32: launch_score_any<32, T>(fold_in, wi, ncoords_prev, tb, ncoords_cur, n_inner, nseg_out, k, snr_out, s); break;
    case 64: launch_score_any<64, T>(fold_in, wi, ncoords_prev, tb, ncoords_cur, n_inner, nseg_out, k, snr_out, s); break;
    default: launch_score_any<128, T>(fold_in, wi, ncoords_prev, tb, ncoords_cur, n_inner, nseg_out, k, snr_out, s); break;
  }
  PYL_CHECK_LAUNCH("ffa_shift_add_score_kernel");
  return PYL_OK;
}

extern "C" int pyl_ffa_shift_add_score_layout(const float* fold_in, int32_t in_layout,
                                              int64_t nseg_in, int64_t ncoords_prev,
                                              const int32_t* table_dev, int64_t ncoords_cur,
                                              int64_t n_inner, int32_t nbins,
                                              const int64_t* widths, int32_t nw,
                                              double stdnoise, float* snr_out, void* stream) {
  return shift_add_score<float>(fold_in, in_layout, nseg_in, ncoords_prev, table_dev,
                                ncoords_cur, n_inner, nbins, widths, nw, stdnoise, snr_out,
                                stream);
}

extern "C" int pyl_ffa_shift_add2_score_layout(
    const float* fold_in, int32_t in_layout, int64_t nseg_in, int64_t ncoords_prev2,
    const int32_t* table_prev, int64_t ncoords_prev, const int32_t* table_dev,
    int64_t ncoords_cur, int64_t n_inner, int32_t nbins, const int64_t* widths, int32_t nw,
    double stdnoise, float* snr_out, void* stream) {
  PYL_REQUIRE(fold_in && table_prev && table_dev && snr_out, "null device pointer");
  PYL_REQUIRE(nseg_in >= 4 && nseg_in % 4 == 0,
              "two fused levels need a multiple of 4 segments, got %lld", (long long)nseg_in);
  PYL_REQUIRE(ncoords_prev2 > 0 && ncoords_prev > 0 && ncoords_cur > 0, "empty parameter grid");
  PYL_REQUIRE(nbins == 32 || nbins == 64, "two fused levels support nbins 32 or 64, got %d",
              nbins);
  PYL_REQUIRE(n_inner >= 1 && ncoords_cur % n_inner == 0,
              "n_inner must divide ncoords_cur (last-axis grid count)");
  PYL_REQUIRE(valid_layout(in_layout), "invalid fold layout");
  pyl::ScoreConst k;
  int rc = pyl::build_score_const(nbins, widths, nw, stdnoise, &k);
  if (rc != PYL_OK) return rc;
  PYL_REQUIRE(k.nw <= kRoundsMaxNW, "two fused levels support at most %d widths",
              kRoundsMaxNW);
  hipStream_t s = pyl::as_stream(stream);
  const int64_t nseg_out = nseg_in / 4;
  const int4* tb = reinterpret_cast<const int4*>(table_dev);
  const int4* tb2 = reinterpret_cast<const int4*>(table_prev);
  const bool wi = in_layout == PYL_FOLD_WRAP;
  if (nbins == 32) {
    if (wi) launch_two<32, true>(fold_in, ncoords_prev2, tb2, ncoords_prev, tb, ncoords_cur, n_inner, nseg_out, k, snr_out, s);
    else launch_two<32, false>(fold_in, ncoords_prev2, tb2, ncoords_prev, tb, ncoords_cur, n_inner, nseg_out, k, snr_out, s);
  } else {
    if (wi) launch_two<64, true>(fold_in, ncoords_prev2, tb2, ncoords_prev, tb, ncoords_cur, n_inner, nseg_out, k, snr_out, s);
    else launch_two<64, false>(fold_in, ncoords_prev2, tb2, ncoords_prev, tb, ncoords_cur, n_inner, nseg_out, k, snr_out, s);
  }
  PYL_CHECK_LAUNCH("ffa_shift_add_score_rounds_kernel<two levels>");
  return PYL_OK;
}

extern "C" int pyl_ffa_shift_add_score(const float* fold_in, int64_t nseg_in,
                                       int64_t ncoords_prev, const int32_t* table_dev,
                                       int64_t ncoords_cur, int64_t n_inner, int32_t nbins,
                                       const int64_t* widths, int32_t nw, double stdnoise,
                                       float* snr_out, void* stream) {
  return shift_add_score<float>(fold_in, PYL_FOLD_DENSE, nseg_in, ncoords_prev, table_dev,
                                ncoords_cur, n_inner, nbins, widths, nw, stdnoise, snr_out,
                                stream);
}

extern "C" int pyl_ffa_shift_add_score_bf16(const uint16_t* fold_in, int64_t nseg_in,
                                            int64_t ncoords_prev, const int32_t* table_dev,
                                            int64_t ncoords_cur, int64_t n_inner, int32_t nbins,
                                            const int64_t* widths, int32_t nw, double stdnoise,
                                            float* snr_out, void* stream) {
  return shift_add_score<uint16_t>(fold_in, PYL_FOLD_DENSE, nseg_in, ncoords_prev, table_dev,
                                   ncoords_cur, n_inner, nbins, widths, nw, stdnoise, snr_out,
                                   stream);
}

extern "C" int pyl_ffa_shift_add_bf16(const uint16_t* fold_in, int64_t nseg_in,
                                      int64_t ncoords_prev, const int32_t* table_dev,
                                      int64_t ncoords_cur, int32_t nbins, uint16_t* fold_out,
                                      void* stream) {
  PYL_REQUIRE(fold_in && table_dev && fold_out, "null device pointer");
  PYL_REQUIRE(nseg_in >= 2 && nseg_in % 2 == 0, "nsegments must be even and >= 2, got %lld",
              (long long)nseg_in);
  PYL_REQUIRE(ncoords_prev > 0 && ncoords_cur > 0, "empty parameter grid");
  PYL_REQUIRE(nbins == 32 || nbins == 64 || nbins == 128 || nbins == 256,
              "BF16 storage supports nbins in {32, 64, 128, 256}, got %d", nbins);
  hipStream_t s = pyl::as_stream(stream);
  const int64_t nseg_out = nseg_in / 2;
  const int4* tb = reinterpret_cast<const int4*>(table_dev);
  switch (nbins) {
    case 32: launch_shift_add<32, uint16_t, false, false>(fold_in, ncoords_prev, tb, ncoords_cur, nseg_out, fold_out, s); break;
    case 64: launch_shift_add<64, uint16_t, false, false>(fold_in, ncoords_prev, tb, ncoords_cur, nseg_out, fold_out, s); break;
    case 128: launch_shift_add<128, uint16_t, false, false>(fold_in, ncoords_prev, tb, ncoords_cur, nseg_out, fold_out, s); break;
    default: launch_shift_add<256, uint16_t, false, false>(fold_in, ncoords_prev, tb, ncoords_cur, nseg_out, fold_out, s); break;
  }
  PYL_CHECK_LAUNCH("ffa_shift_add_kernel<bf16>");
  return PYL_OK;
}

template <int B>
static void launch_shift_add2(const float* in, bool wrap_in, int64_t nprev2, const int4* ta,
                              const int4* tb, int64_t ncur, int64_t nseg_out, float* out,
                              bool wrap_out, hipStream_t s) {
  constexpr int RPB = 256 / (B / 2);
  const int64_t nblk_j = pyl::ceil_div(ncur, RPB);
  const unsigned nblocks = (unsigned)(nblk_j * nseg_out);
#define PYL_SA2(WI, WO)                                                                 \
  ffa_shift_add2_kernel<B, WI, WO><<<nblocks, 256, 0, s>>>(in, nprev2, ta, tb, ncur,     \
                                                          nseg_out, nblk_j, out)
  if (wrap_in) {
    if (wrap_out) PYL_SA2(true, true);
    else PYL_SA2(true, false);
  } else {
    if (wrap_out) PYL_SA2(false, true);
    else PYL_SA2(false, false);
  }
#undef PYL_SA2
}

extern "C" int pyl_ffa_shift_add2_layout(const float* fold_in, int32_t in_layout, int64_t nseg_in,
                                         int64_t ncoords_prev2, const int32_t* table_prev,
                                         int64_t ncoords_prev, const int32_t* table_dev,
                                         int64_t ncoords_cur, int32_t nbins, float* fold_out,
                                         int32_t out_layout, void* stream) {
  PYL_REQUIRE(fold_in && table_prev && table_dev && fold_out, "null device pointer");
  PYL_REQUIRE(nseg_in >= 4 && nseg_in % 4 == 0,
              "two fused levels need a multiple of 4 segments, got %lld", (long long)nseg_in);
  PYL_REQUIRE(ncoords_prev2 > 0 && ncoords_prev > 0 && ncoords_cur > 0, "empty parameter grid");
  PYL_REQUIRE(nbins == 32 || nbins == 64 || nbins == 128 || nbins == 256,
              "two fused levels support nbins in {32, 64, 128, 256}, got %d", nbins);
  PYL_REQUIRE(valid_layout(in_layout) && valid_layout(out_layout), "invalid fold layout");
  PYL_REQUIRE(ncoords_prev2 * nseg_in <= ((int64_t)1 << 40), "fold too large");
  const int64_t nseg_out = nseg_in / 4;
  PYL_REQUIRE(pyl::ceil_div(ncoords_cur, 256 / (nbins / 2)) * nseg_out <= 0x7fffffff,
              "too many output rows for one launch");
  hipStream_t s = pyl::as_stream(stream);
  const int4* ta = reinterpret_cast<const int4*>(table_prev);
  const int4* tb = reinterpret_cast<const int4*>(table_dev);
  const bool wi = in_layout == PYL_FOLD_WRAP, wo = out_layout == PYL_FOLD_WRAP;
  switch (nbins) {
    case 32: launch_shift_add2<32>(fold_in, wi, ncoords_prev2, ta, tb, ncoords_cur, nseg_out, fold_out, wo, s); break;
    case 64: launch_shift_add2<64>(fold_in, wi, ncoords_prev2, ta, tb, ncoords_cur, nseg_out, fold_out, wo, s); break;
    case 128: launch_shift_add2<128>(fold_in, wi, ncoords_prev2, ta, tb, ncoords_cur, nseg_out, fold_out, wo, s); break;
    default: launch_shift_add2<256>(fold_in, wi, ncoords_prev2, ta, tb, ncoords_cur, nseg_out, fold_out, wo, s); break;
  }
  PYL_CHECK_LAUNCH("ffa_shift_add2_kernel");
  return PYL_OK;
}

extern "C" int pyl_ffa_level(int32_t nparams, const double* params_dev, const int64_t* counts_cur,
                             const int64_t* counts_prev, const double* param_limits,
                             int32_t ffa_level, double tseg_brute, int32_t nbins,
                             const float* fold_in_dev, int64_t nseg_in, int32_t* table_dev,
                             float* fold_out_dev, void* stream) {
  int rc = pyl_ffa_resolve(nparams, params_dev, counts_cur, counts_prev, param_limits, ffa_level,
                           tseg_brute, nbins, table_dev, nullptr, stream);
  if (rc != PYL_OK) return rc;
  int64_t ncur = 1, nprev = 1;
  for (int k = 0; k < nparams; ++k) {
    ncur *= counts_cur[k];
    nprev *= counts_prev[k];
  }
  return pyl_ffa_shift_add(fold_in_dev, nseg_in, nprev, table_dev, ncur, nbins, fold_out_dev,
                           stream);
}

extern "C" int pyl_ffa_set_tile_rows(int32_t rows) {
  PYL_REQUIRE(rows == 0 || (rows >= 1 && rows <= 256 && (rows & (rows - 1)) == 0),
              "tile rows must be 0 or a power of two in [1, 256] (got %d)", rows);
  g_tile_outer.store(rows, std::memory_order_relaxed);
  return PYL_OK;
}

extern "C" int pyl_ffa_get_tile_rows(void) { return g_tile_outer.load(std::memory_order_relaxed); }

extern "C" int pyl_ffa_set_packed_tables(int32_t enable) {
  PYL_REQUIRE(enable == 0 || enable == 1, "enable must be 0 or 1 (got %d)", enable);
  g_packed.store(enable, std::memory_order_relaxed);
  return PYL_OK;
}

// ---- One FFA level without any allocation -------------------------------------------
// The resolve table of pyl_ffa_level lives in a file-scope device array instead of a
// caller workspace: at most kNoAllocMaxCoords parameter sets (16 B each, 32 MiB), held by
// every GPU the library runs on until the process ends. Host code never passes the array's
// name to a kernel (that name is a host-side stand-in); it passes the device address that
// hipGetSymbolAddress returns for the current device, looked up once per device on the
// first call there. The resolve kernel writes every entry the shift-add reads, so nothing
// needs clearing; two calls that overlap on one GPU race on the table.
constexpr int64_t kNoAllocMaxCoords = (int64_t)1 << 21;
constexpr int kNoAllocMaxDevices = 64;

static __device__ __align__(256) int4 scratch[kNoAllocMaxCoords];

static std::mutex g_scratch_mu;
static int4* g_scratch_dev[kNoAllocMaxDevices] = {};

static int scratch_table(int4** out) {
  int dev = -1;
  PYL_CHECK_HIP(hipGetDevice(&dev));
  PYL_REQUIRE(dev >= 0 && dev < kNoAllocMaxDevices, "device %d is beyond the %d supported", dev,
              kNoAllocMaxDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_dev[dev] == nullptr) {
    void* p = nullptr;
    PYL_CHECK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    PYL_REQUIRE(p != nullptr, "no device address for the level table on device %d", dev);
    g_scratch_dev[dev] = static_cast<int4*>(p);
  }
  *out = g_scratch_dev[dev];
  return PYL_OK;
}

extern "C" int64_t pyl_ffa_level_noalloc_max_coords(void) { return kNoAllocMaxCoords; }

extern "C" int pyl_ffa_level_noalloc(int32_t nparams, const double* params_dev,
                                     const int64_t* counts_cur, const int64_t* counts_prev,
                                     const double* param_limits, int32_t ffa_level,
                                     double tseg_brute, int32_t nbins, const float* fold_in_dev,
                                     int64_t nseg_in, float* fold_out_dev, void* stream) {
  PYL_REQUIRE(nparams >= 1 && nparams <= kMaxP, "nparams must be in [1, 5], got %d", nparams);
  PYL_REQUIRE(counts_cur != nullptr, "null pointer argument");
  int64_t ncur = 1;
  for (int k = 0; k < nparams; ++k) {
    PYL_REQUIRE(counts_cur[k] > 0, "grid counts must be positive");
    PYL_REQUIRE(counts_cur[k] <= kNoAllocMaxCoords && ncur <= kNoAllocMaxCoords / counts_cur[k],
                "the level has more than %lld parameter sets, the maximum without allocation",
                (long long)kNoAllocMaxCoords);
    ncur *= counts_cur[k];
  }
  int4* table = nullptr;
  const int rc = scratch_table(&table);
  if (rc != PYL_OK) return rc;
  return pyl_ffa_level(nparams, params_dev, counts_cur, counts_prev, param_limits, ffa_level,
                       tseg_brute, nbins, fold_in_dev, nseg_in, reinterpret_cast<int32_t*>(table),
                       fold_out_dev, stream);
}
