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
//   aligned 16 B chunk of the source row that holds its rotated elements (two chunks
//   when the shift is not a multiple of 4) -- every load instruction of a wavefront
//   covers whole 64 B segments, so the rotation costs no extra HBM traffic.
//   Rows of consecutive j share their source rows (the previous level's grid is about
//   half as fine), so each block takes a contiguous run of j and the runs are dealt to
//   the 8 XCDs in contiguous slabs: neighbours' rows are reused from the same L2.
#include <algorithm>

#include "pyl_score.hpp"

namespace {

constexpr int kMaxP = 5;

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

__global__ void ffa_resolve_kernel(const double* __restrict__ params, ResolveArgs a,
                                   int4* __restrict__ table, double2* __restrict__ phases) {
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
// B elements, widened to FP32: element e of the result is src[(4*c4 + e - s) mod B].
template <int B, typename T = float>
__device__ __forceinline__ float4 rotated_chunk(const T* __restrict__ src, int c4, int s) {
  constexpr int C = B / 4;
  const int base = 4 * c4 - s;                 // may be negative
  const int m = (base % B + B) % B;            // first source element
  const int q = m >> 2;                        // source chunk
  const int o = m & 3;                         // offset inside the chunk
  const typename Vec4<T>::type* s4 = reinterpret_cast<const typename Vec4<T>::type*>(src);
  const float4 a = to_f32x4<T>(s4[q]);
  if (o == 0) return a;
  const float4 b = to_f32x4<T>(s4[(q + 1) % C]);
  float4 r;
  if (o == 1) {
    r.x = a.y; r.y = a.z; r.z = a.w; r.w = b.x;
  } else if (o == 2) {
    r.x = a.z; r.y = a.w; r.z = b.x; r.w = b.y;
  } else {
    r.x = a.w; r.y = b.x; r.z = b.y; r.w = b.z;
  }
  return r;
}

// Rows per block = kThreads / (B/2) (each row = 2 components x B/4 chunks).
template <int B, typename T>
__global__ void __launch_bounds__(256)
ffa_shift_add_kernel(const T* __restrict__ in, int64_t ncoords_prev,
                     const int4* __restrict__ table, int64_t ncoords_cur, int64_t nseg_out,
                     int64_t nblk_j, T* __restrict__ out) {
  constexpr int CPR = B / 4;          // chunks per component
  constexpr int TPR = 2 * CPR;        // threads per row
  constexpr int RPB = 256 / TPR;      // rows per block
  // XCD-aware block order: hardware deals block i to XCD i % 8. Remap so that the
  // blocks one XCD receives form a contiguous slab of (q, j) space.
  const int64_t nblocks = (int64_t)gridDim.x;
  const int64_t bid = blockIdx.x;
  const int64_t per_xcd = (nblocks + 7) / 8;
  const int64_t xcd = bid % 8;
  const int64_t slot = bid / 8;
  const int64_t full = nblocks % 8 == 0 ? 8 : nblocks % 8;  // XCDs with per_xcd blocks
  int64_t vb;
  if (xcd < full) vb = xcd * per_xcd + slot;
  else vb = full * per_xcd + (xcd - full) * (per_xcd - 1) + slot;
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
  constexpr int64_t E = 2 * B;
  const T* tail = in + ((2 * q) * ncoords_prev + tb.x) * E + comp * B;
  const T* head = in + ((2 * q + 1) * ncoords_prev + tb.z) * E + comp * B;
  const float4 a = rotated_chunk<B, T>(tail, c4, tb.y);
  const float4 b = rotated_chunk<B, T>(head, c4, tb.w);
  float4 r;
  r.x = a.x + b.x;
  r.y = a.y + b.y;
  r.z = a.z + b.z;
  r.w = a.w + b.w;
  store4<T>(out + (q * ncoords_cur + j) * E + comp * B, c4, r);
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
template <int B, typename T>
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
  constexpr int64_t E = 2 * B;
  const int c4 = tid % CPR;
  const int rsub = tid / CPR;
#pragma unroll 2
  for (int r0 = 0; r0 < nvalid; r0 += RPS) {
    const int r = r0 + rsub;
    if (r < nvalid) {
      const int4 tb = s_tab[r];
      const T* tail = in + ((2 * q) * ncoords_prev + tb.x) * E;
      const T* head = in + ((2 * q + 1) * ncoords_prev + tb.z) * E;
      const float4 te = rotated_chunk<B, T>(tail, c4, tb.y);
      const float4 he = rotated_chunk<B, T>(head, c4, tb.w);
      const float4 tv = rotated_chunk<B, T>(tail + B, c4, tb.y);
      const float4 hv = rotated_chunk<B, T>(head + B, c4, tb.w);
      const int b = 4 * c4;
      // The level's fold element as it would be stored (identity for FP32, BF16 rounding
      // for BF16 storage), then e / sqrt(v).
      auto st = [](float x) { return pyl::Store<T>::load(pyl::Store<T>::store(x)); };
      S[(b + 0) * pyl::kScoreStride + r] = st(te.x + he.x) / sqrtf(st(tv.x + hv.x));
      S[(b + 1) * pyl::kScoreStride + r] = st(te.y + he.y) / sqrtf(st(tv.y + hv.y));
      S[(b + 2) * pyl::kScoreStride + r] = st(te.z + he.z) / sqrtf(st(tv.z + hv.z));
      S[(b + 3) * pyl::kScoreStride + r] = st(te.w + he.w) / sqrtf(st(tv.w + hv.w));
    }
  }
  __syncthreads();
  if (tid < nvalid)
    pyl::score_column<true>(k, B, S, tid, out + (q * ncoords_cur + j0 + tid) * k.nw);
}

// Register-scoring variant of the fused last level (nbins 32 or 64). A block owns a tile of
// 256 output rows of one segment pair: tile_o_rows (1, 2 or 4) "outer" x 256/tile_o_rows
// "inner" rows, inner = index along the last (frequency) axis, outer = the product of the
// other axes. Rows at neighbouring outer
// indices (e.g. accelerations) resolve to the same source rows, so the tile reuses each
// source row from L1/L2 several times instead of re-reading it from HBM.
// Phase 1 forms the rows with coalesced, rotated float4 loads (16 lanes per row at B = 64)
// and writes the normalised values transposed into LDS (row stride 257 words:
// conflict-free both ways). Phase 2: every thread pulls its row into registers, scans it
// sequentially and scores it with the unrolled max-difference loops of pyl_score.hpp.
template <int B, typename T>
__global__ void __launch_bounds__(256)
ffa_shift_add_score_regs_kernel(const T* __restrict__ in, int64_t ncoords_prev,
                                const int4* __restrict__ table, int64_t ncoords_cur,
                                int64_t n_inner, int tile_o_rows, int64_t tiles_inner,
                                int64_t tiles_per_pair, pyl::ScoreConst k,
                                float* __restrict__ out) {
  constexpr int kRows = 256;
  constexpr int kLds = kRows + 1;
  constexpr int CPR = B / 4;       // float4 chunks per component
  constexpr int RPS = 64 / CPR;    // rows per wavefront step
  __shared__ float S[B * kLds];
  __shared__ int4 s_tab[kRows];
  __shared__ int64_t s_j[kRows];
  const int tid = threadIdx.x;
  const int64_t q = blockIdx.x / tiles_per_pair;
  const int64_t tile = blockIdx.x - q * tiles_per_pair;
  const int64_t tile_o = tile / tiles_inner;
  const int64_t tile_i = tile - tile_o * tiles_inner;
  const int64_t n_outer = ncoords_cur / n_inner;
  {
    const int tile_i_rows = kRows / tile_o_rows;
    const int64_t outer = tile_o * tile_o_rows + tid / tile_i_rows;
    const int64_t inner = tile_i * tile_i_rows + tid % tile_i_rows;
    const bool ok = outer < n_outer && inner < n_inner;
    const int64_t j = ok ? outer * n_inner + inner : -1;
    s_j[tid] = j;
    if (ok) s_tab[tid] = table[j];
  }
  __syncthreads();
  constexpr int64_t E = 2 * B;
  const int wave = tid >> 6;
  const int lane = tid & 63;
  const int c4 = lane % CPR;
  const int rsub = lane / CPR;
  const T* base_t = in + (2 * q) * ncoords_prev * E;
  const T* base_h = in + (2 * q + 1) * ncoords_prev * E;
#pragma unroll 4
  for (int step = 0; step < 64 / RPS; ++step) {
    const int r = wave * 64 + step * RPS + rsub;
    if (s_j[r] >= 0) {
      const int4 tb = s_tab[r];
      const T* tail = base_t + (int64_t)tb.x * E;
      const T* head = base_h + (int64_t)tb.z * E;
      const float4 te = rotated_chunk<B, T>(tail, c4, tb.y);
      const float4 he = rotated_chunk<B, T>(head, c4, tb.w);
      const float4 tv = rotated_chunk<B, T>(tail + B, c4, tb.y);
      const float4 hv = rotated_chunk<B, T>(head + B, c4, tb.w);
      auto st = [](float x) { return pyl::Store<T>::load(pyl::Store<T>::store(x)); };
      const int b = 4 * c4;
      S[(b + 0) * kLds + r] = st(te.x + he.x) / sqrtf(st(tv.x + hv.x));
      S[(b + 1) * kLds + r] = st(te.y + he.y) / sqrtf(st(tv.y + hv.y));
      S[(b + 2) * kLds + r] = st(te.z + he.z) / sqrtf(st(tv.z + hv.z));
      S[(b + 3) * kLds + r] = st(te.w + he.w) / sqrtf(st(tv.w + hv.w));
    }
  }
  __syncthreads();
  const int64_t j = s_j[tid];
  if (j >= 0) {
    float x[B];
#pragma unroll
    for (int i = 0; i < B; ++i) x[i] = S[i * kLds + tid];
    pyl::scan_regs<B>(x);
    pyl::score_regs<B, true>(k, x, out + (q * ncoords_cur + j) * k.nw);
  }
}

}  // namespace

extern "C" int pyl_ffa_resolve(int32_t nparams, const double* params_dev,
                               const int64_t* counts_cur, const int64_t* counts_prev,
                               const double* param_limits, int32_t ffa_level, double tseg_brute,
                               int32_t nbins, int32_t* table_dev, double* phases_dev,
                               void* stream) {
  PYL_REQUIRE(nparams >= 1 && nparams <= kMaxP, "nparams must be in [1, 5], got %d", nparams);
  PYL_REQUIRE(params_dev && counts_cur && counts_prev && param_limits && table_dev,
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
      reinterpret_cast<double2*>(phases_dev));
  PYL_CHECK_LAUNCH("ffa_resolve_kernel");
  return PYL_OK;
}

template <int B, typename T>
static void launch_shift_add(const T* in, int64_t nprev, const int4* tb, int64_t ncur,
                             int64_t nseg_out, T* out, hipStream_t s) {
  constexpr int RPB = 256 / (B / 2);
  const int64_t nblk_j = pyl::ceil_div(ncur, RPB);
  const int64_t nblocks = nblk_j * nseg_out;
  ffa_shift_add_kernel<B, T><<<(unsigned)nblocks, 256, 0, s>>>(in, nprev, tb, ncur, nseg_out,
                                                               nblk_j, out);
}

extern "C" int pyl_ffa_shift_add(const float* fold_in, int64_t nseg_in, int64_t ncoords_prev,
                                 const int32_t* table_dev, int64_t ncoords_cur, int32_t nbins,
                                 float* fold_out, void* stream) {
  PYL_REQUIRE(fold_in && table_dev && fold_out, "null device pointer");
  PYL_REQUIRE(nseg_in >= 2 && nseg_in % 2 == 0, "nsegments must be even and >= 2, got %lld",
              (long long)nseg_in);
  PYL_REQUIRE(ncoords_prev > 0 && ncoords_cur > 0, "empty parameter grid");
  PYL_REQUIRE(nbins > 0, "Number of bins must be positive.");
  hipStream_t s = pyl::as_stream(stream);
  const int64_t nseg_out = nseg_in / 2;
  const int4* tb = reinterpret_cast<const int4*>(table_dev);
  switch (nbins) {
    case 32: launch_shift_add<32, float>(fold_in, ncoords_prev, tb, ncoords_cur, nseg_out, fold_out, s); break;
    case 64: launch_shift_add<64, float>(fold_in, ncoords_prev, tb, ncoords_cur, nseg_out, fold_out, s); break;
    case 128: launch_shift_add<128, float>(fold_in, ncoords_prev, tb, ncoords_cur, nseg_out, fold_out, s); break;
    case 256: launch_shift_add<256, float>(fold_in, ncoords_prev, tb, ncoords_cur, nseg_out, fold_out, s); break;
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


template <int B, typename T>
static void launch_shift_add_score(const T* in, int64_t nprev, const int4* tb, int64_t ncur,
                                   int64_t n_inner, int64_t nseg_out, const pyl::ScoreConst& k,
                                   float* out, hipStream_t s) {
  (void)n_inner;
  if constexpr (B <= 64) {
    const int64_t n_outer = ncur / n_inner;
    const int to = n_outer >= 4 ? 4 : (n_outer >= 2 ? 2 : 1);
    const int64_t tiles_inner = pyl::ceil_div(n_inner, 256 / to);
    const int64_t tiles_outer = pyl::ceil_div(n_outer, to);
    const int64_t per_pair = tiles_inner * tiles_outer;
    ffa_shift_add_score_regs_kernel<B, T><<<(unsigned)(per_pair * nseg_out), 256, 0, s>>>(
        in, nprev, tb, ncur, n_inner, to, tiles_inner, per_pair, k, out);
    return;
  }
  const int64_t nblk_j = pyl::ceil_div(ncur, pyl::kScoreBlock);
  const size_t lds = sizeof(float) * (size_t)(B + k.wmax) * pyl::kScoreStride;
  ffa_shift_add_score_kernel<B, T><<<(unsigned)(nblk_j * nseg_out), pyl::kScoreBlock, lds, s>>>(
      in, nprev, tb, ncur, nblk_j, k, out);
}

template <typename T>
static int shift_add_score(const T* fold_in, int64_t nseg_in, int64_t ncoords_prev,
                           const int32_t* table_dev, int64_t ncoords_cur, int64_t n_inner,
                           int32_t nbins, const int64_t* widths, int32_t nw, double stdnoise,
                           float* snr_out, void* stream) {
  PYL_REQUIRE(fold_in && table_dev && snr_out, "null device pointer");
  PYL_REQUIRE(nseg_in >= 2 && nseg_in % 2 == 0, "nsegments must be even and >= 2, got %lld",
              (long long)nseg_in);
  PYL_REQUIRE(ncoords_prev > 0 && ncoords_cur > 0, "empty parameter grid");
  PYL_REQUIRE(nbins == 32 || nbins == 64 || nbins == 128,
              "fused level + score supports nbins in {32, 64, 128}, got %d", nbins);
  PYL_REQUIRE(n_inner >= 1 && ncoords_cur % n_inner == 0,
              "n_inner must divide ncoords_cur (last-axis grid count)");
  pyl::ScoreConst k;
  int rc = pyl::build_score_const(nbins, widths, nw, stdnoise, &k);
  if (rc != PYL_OK) return rc;
  hipStream_t s = pyl::as_stream(stream);
  const int64_t nseg_out = nseg_in / 2;
  const int4* tb = reinterpret_cast<const int4*>(table_dev);
  switch (nbins) {
    case 32: launch_shift_add_score<32, T>(fold_in, ncoords_prev, tb, ncoords_cur, n_inner, nseg_out, k, snr_out, s); break;
    case 64: launch_shift_add_score<64, T>(fold_in, ncoords_prev, tb, ncoords_cur, n_inner, nseg_out, k, snr_out, s); break;
    default: launch_shift_add_score<128, T>(fold_in, ncoords_prev, tb, ncoords_cur, n_inner, nseg_out, k, snr_out, s); break;
  }
  PYL_CHECK_LAUNCH("ffa_shift_add_score_kernel");
  return PYL_OK;
}

extern "C" int pyl_ffa_shift_add_score(const float* fold_in, int64_t nseg_in,
                                       int64_t ncoords_prev, const int32_t* table_dev,
                                       int64_t ncoords_cur, int64_t n_inner, int32_t nbins,
                                       const int64_t* widths, int32_t nw, double stdnoise,
                                       float* snr_out, void* stream) {
  return shift_add_score<float>(fold_in, nseg_in, ncoords_prev, table_dev, ncoords_cur, n_inner,
                                nbins, widths, nw, stdnoise, snr_out, stream);
}

extern "C" int pyl_ffa_shift_add_score_bf16(const uint16_t* fold_in, int64_t nseg_in,
                                            int64_t ncoords_prev, const int32_t* table_dev,
                                            int64_t ncoords_cur, int64_t n_inner, int32_t nbins,
                                            const int64_t* widths, int32_t nw, double stdnoise,
                                            float* snr_out, void* stream) {
  return shift_add_score<uint16_t>(fold_in, nseg_in, ncoords_prev, table_dev, ncoords_cur,
                                   n_inner, nbins, widths, nw, stdnoise, snr_out, stream);
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
    case 32: launch_shift_add<32, uint16_t>(fold_in, ncoords_prev, tb, ncoords_cur, nseg_out, fold_out, s); break;
    case 64: launch_shift_add<64, uint16_t>(fold_in, ncoords_prev, tb, ncoords_cur, nseg_out, fold_out, s); break;
    case 128: launch_shift_add<128, uint16_t>(fold_in, ncoords_prev, tb, ncoords_cur, nseg_out, fold_out, s); break;
    default: launch_shift_add<256, uint16_t>(fold_in, ncoords_prev, tb, ncoords_cur, nseg_out, fold_out, s); break;
  }
  PYL_CHECK_LAUNCH("ffa_shift_add_kernel<bf16>");
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
