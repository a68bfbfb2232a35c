// Learned branch ranking for the EP search: a small 1-D CNN over a survivor's folded
// profile plus 8 parameter features, trained on the GPU with these kernels. It has no
// counterpart in the reference; the budgeted search keeps the survivors it ranks highest.
//
// Network (one sample; B bins, B in {32, 64, 128}):
//   x  = (p - mean p) / (std p + 1e-6), p = e / sqrt(v) (irfft of complex rows first);
//   a1 = relu(conv(x;  W1[32][1][5], b1)), a2 = relu(conv(a1; W2[32][32][5], b2)),
//   a3 = relu(conv(a2; W3[32][32][5], b3)), conv circular: y[c][p] = b[c] +
//        sum_{i,t} W[c][i][t] x[i][(p + t - 2) mod B];
//   h[c] = max_p a3[c][p] (first maximum); z = (h, features[8]);
//   u = relu(Wf1[32][40] z + bf1); logit = Wf2[32] . u + bf2.
// Loss: sum_n w_pos y softplus(-logit) + w_neg (1 - y) softplus(logit).
//
// Work decomposition: a block of 256 threads holds 256 / B samples, one thread per
// (sample, bin) for the profile and the first conv layer (32 output channels per thread
// in registers, weights read with block-uniform indices); the two 32 -> 32 layers are FP32
// products on the matrix cores (fwd_mfma) over the block's activations in LDS (33 KB).
// The backward kernel walks tiles of samples
// grid-stride; the weight gradients of the two 32 -> 32 layers are products on the matrix
// cores (FP32, wgrad_mfma), every wavefront owning five 16 x 16 output tiles accumulated
// over the block's tiles; the other gradients are owned by fixed threads. Each block writes
// its partial sum once; a reduction adds the partials in groups of 32 blocks and then the
// group sums, each in a fixed order, so the gradient is deterministic.
#include <algorithm>
#include <cfloat>
#include <cmath>

#include "pyl_budget.hpp"
#include "pyl_common.hpp"
#include "pyl_score.hpp"

namespace {

constexpr int kC = 32, kK = 5, kF = 8, kH = 32, kZ = kC + kF;
constexpr int kThreads = 256;
constexpr int oW1 = 0, oB1 = oW1 + kC * kK, oW2 = oB1 + kC, oB2 = oW2 + kC * kC * kK;
constexpr int oW3 = oB2 + kC, oB3 = oW3 + kC * kC * kK, oF1 = oB3 + kC, oBF1 = oF1 + kH * kZ;
constexpr int oF2 = oBF1 + kH, oBF2 = oF2 + kH, kNParams = oBF2 + 1;
constexpr int kF1PerThread = (kH * kZ + kThreads - 1) / kThreads;  // 5

// Per-sample training cache (floats): a1 [kC][B], a2 [kC][B], x0 [B], argmax [kC] (int32
// bits), z [kZ], q [kH].
__host__ __device__ constexpr int64_t cache_stride(int B) {
  return 2 * (int64_t)kC * B + B + kC + kZ + kH;
}
__host__ __device__ constexpr int64_t off_a2(int B) { return (int64_t)kC * B; }
__host__ __device__ constexpr int64_t off_x0(int B) { return 2 * (int64_t)kC * B; }
__host__ __device__ constexpr int64_t off_arg(int B) { return off_x0(B) + B; }
__host__ __device__ constexpr int64_t off_z(int B) { return off_arg(B) + kC; }
__host__ __device__ constexpr int64_t off_q(int B) { return off_z(B) + kZ; }

// One circular conv layer for thread (s, p): acc[co] = b[co] + sum_ci sum_t
// W[co][ci][t] in[s][ci][(p + t - 2) & (B - 1)], ci and t ascending.
template <int B, int CIN>
__device__ __forceinline__ void conv_layer(const float* __restrict__ W,
                                           const float* __restrict__ bias,
                                           const float* in, int p, float* acc) {
#pragma unroll
  for (int co = 0; co < kC; ++co) acc[co] = bias[co];
#pragma unroll 1
  for (int ci = 0; ci < CIN; ++ci) {
    float v[kK];
#pragma unroll
    for (int t = 0; t < kK; ++t) v[t] = in[ci * B + ((p + t - 2) & (B - 1))];
#pragma unroll
    for (int co = 0; co < kC; ++co) {
      const float* w = W + (co * CIN + ci) * kK;
      float a = acc[co];
#pragma unroll
      for (int t = 0; t < kK; ++t) a += w[t] * v[t];
      acc[co] = a;
    }
  }
}

typedef float cnn_f32x4 __attribute__((ext_vector_type(4)));

// Forward 32 -> 32 conv layer on the matrix cores: y[ss][co][p] = b[co] + sum over
// k = (ci, t) of W[co][ci][t] in[ss][ci][(p + t - 2) mod B], per sample a (32 x 160) x
// (160 x B) product. The block's S samples x 2 channel tiles x B/16 bin tiles (32 tiles of
// 16 x 16 for every B) are split over the four wavefronts, eight each (tiles wave + 4 j);
// v_mfma_f32_16x16x4f32, lane l: A = W[16 mt + l % 16][ci][t], B = in[ss][ci][bin of
// column 16 nt + l % 16 shifted by t], (ci, t) = k0 + l / 16; acc[j][r] starts at the bias
// and ends as channel 16 mt + 4 (l / 16) + r at bin 16 nt + l % 16. FP32 products and sums
// in the matrix unit's order.
template <int B, int S>
__device__ __forceinline__ void fwd_mfma(const float* __restrict__ W,
                                         const float* __restrict__ bias,
                                         const float (*In)[kC][B + 1], cnn_f32x4 (&acc)[8],
                                         int wave, int lane) {
  constexpr int NT = B / 16;
  const int m = lane & 15, kq = lane >> 4;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int mt = ((wave + 4 * j) % (2 * NT)) / NT;
#pragma unroll
    for (int r = 0; r < 4; ++r) acc[j][r] = bias[16 * mt + 4 * kq + r];
  }
#pragma unroll 2
  for (int k0 = 0; k0 < kC * kK; k0 += 4) {
    const int k = k0 + kq;
    const int ci = k / kK, t = k - ci * kK;
    const float w0 = W[(m * kC + ci) * kK + t], w1 = W[((16 + m) * kC + ci) * kK + t];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int T = wave + 4 * j;
      const int ss = T / (2 * NT), rem = T - ss * 2 * NT;
      const int mt = rem / NT, nt = rem - mt * NT;
      const float b = In[ss][ci][(nt * 16 + m + t - 2) & (B - 1)];
      acc[j] = __builtin_amdgcn_mfma_f32_16x16x4f32(mt == 0 ? w0 : w1, b, acc[j], 0, 0, 0);
    }
  }
}

// relu of fwd_mfma's tiles into A (after every wavefront has finished reading it); with
// cache, also into each valid sample's training cache at offset coff.
template <int B, int S>
__device__ __forceinline__ void fwd_store(float (*A)[kC][B + 1], const cnn_f32x4 (&acc)[8],
                                          int wave, int lane, float* __restrict__ cache,
                                          int64_t i0, int64_t n, int64_t coff) {
  constexpr int NT = B / 16;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int T = wave + 4 * j;
    const int ss = T / (2 * NT), rem = T - ss * 2 * NT;
    const int mt = rem / NT, nt = rem - mt * NT;
    const int pb = nt * 16 + (lane & 15);
    const bool keep = cache != nullptr && i0 + ss < n;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int co = mt * 16 + 4 * (lane >> 4) + r;
      const float a = fmaxf(acc[j][r], 0.0f);
      A[ss][co][pb] = a;
      if (keep) cache[(i0 + ss) * cache_stride(B) + coff + co * B + pb] = a;
    }
  }
}

template <int B, bool kComplex, bool kTrain>
__global__ void __launch_bounds__(kThreads)
cnn_forward_kernel(const float* __restrict__ th, const float* __restrict__ folds, int64_t n_arg,
                   const int64_t* __restrict__ n_dev, const float* __restrict__ feats,
                   float* __restrict__ logits, float* __restrict__ cache) {
  constexpr int S = kThreads / B;
  const int64_t n = n_dev ? *n_dev : n_arg;  // device count: the EP level loop
  if ((int64_t)blockIdx.x * S >= n) return;  // the whole block
  __shared__ float A[S][kC][B + 1];
  __shared__ float X[S][B];
  __shared__ float stat[S][2];
  __shared__ float Z[S][kZ];
  __shared__ float U[S][kH];
  __shared__ float tw[kComplex ? 2 * B : 1];
  const int tid = threadIdx.x;
  const int s = tid / B, p = tid % B;
  const int64_t i = (int64_t)blockIdx.x * S + s;
  const bool valid = i < n;
  if constexpr (kComplex) {
    pyl::fill_twiddles(tw, B, tid, kThreads);
    __syncthreads();
  }
  float pv = 0.0f;
  if (valid) {
    float e, v;
    if constexpr (kComplex) {
      const float2* row = reinterpret_cast<const float2*>(folds) + i * 2 * (B / 2 + 1);
      e = pyl::irfft_bin(row, B, p, tw);
      v = pyl::irfft_bin(row + (B / 2 + 1), B, p, tw);
    } else {
      e = folds[i * 2 * B + p];
      v = folds[i * 2 * B + B + p];
    }
    pv = pyl::norm_ev(e, v);
  }
  X[s][p] = pv;
  __syncthreads();
  if (p == 0) {
    float sum = 0.0f;
#pragma unroll 1
    for (int q = 0; q < B; ++q) sum += X[s][q];
    const float mu = sum / (float)B;
    float ss = 0.0f;
    for (int q = 0; q < B; ++q) {
      const float d = X[s][q] - mu;
      ss += d * d;
    }
    stat[s][0] = mu;
    stat[s][1] = sqrtf(ss / (float)B);
  }
  __syncthreads();
  const float x0 = valid ? (pv - stat[s][0]) / (stat[s][1] + 1e-6f) : 0.0f;
  __syncthreads();
  X[s][p] = x0;
  float* cs = cache + (kTrain && valid ? i * cache_stride(B) : 0);
  if constexpr (kTrain) {
    if (valid) cs[off_x0(B) + p] = x0;
  }
  __syncthreads();
  float acc[kC];
  conv_layer<B, 1>(th + oW1, th + oB1, &X[s][0], p, acc);
#pragma unroll
  for (int co = 0; co < kC; ++co) {
    const float a = fmaxf(acc[co], 0.0f);
    A[s][co][p] = a;
    if constexpr (kTrain) {
      if (valid) cs[co * B + p] = a;
    }
  }
  __syncthreads();
  {
    // The two 32 -> 32 layers on the matrix cores.
    const int wave = tid >> 6, lane = tid & 63;
    const int64_t i0 = (int64_t)blockIdx.x * S;
    cnn_f32x4 acc8[8];
    fwd_mfma<B, S>(th + oW2, th + oB2, A, acc8, wave, lane);
    __syncthreads();
    fwd_store<B, S>(A, acc8, wave, lane, kTrain ? cache : nullptr, i0, n, off_a2(B));
    __syncthreads();
    fwd_mfma<B, S>(th + oW3, th + oB3, A, acc8, wave, lane);
    __syncthreads();
    fwd_store<B, S>(A, acc8, wave, lane, nullptr, i0, n, 0);
    __syncthreads();
  }
  // Max-pool over bins (first maximum), then the features.
  if (p < kC) {
    float best = A[s][p][0];
    int arg = 0;
#pragma unroll 1
    for (int q = 1; q < B; ++q) {
      const float a = A[s][p][q];
      if (a > best) {
        best = a;
        arg = q;
      }
    }
    Z[s][p] = best;
    if constexpr (kTrain) {
      if (valid) reinterpret_cast<int32_t*>(cs + off_arg(B))[p] = arg;
    }
  }
  if (p < kF) Z[s][kC + p] = valid ? feats[i * kF + p] : 0.0f;
  __syncthreads();
  if (p < kH) {
    float q = th[oBF1 + p];
    for (int z = 0; z < kZ; ++z) q += th[oF1 + p * kZ + z] * Z[s][z];
    U[s][p] = fmaxf(q, 0.0f);
    if constexpr (kTrain) {
      if (valid) cs[off_q(B) + p] = q;
    }
  }
  if constexpr (kTrain) {
    for (int z = p; z < kZ; z += B)
      if (valid) cs[off_z(B) + z] = Z[s][z];
  }
  __syncthreads();
  if (p == 0 && valid) {
    float lg = th[oBF2];
    for (int j = 0; j < kH; ++j) lg += th[oF2 + j] * U[s][j];
    logits[i] = lg;
  }
}

// Weight gradient of a 32 -> 32 conv layer on the matrix cores: dW[co][n] (n = ci kK + t)
// += sum over k = (sample ss, bin q) of dz[ss][co][q] in[ss][ci][(q + t - 2) mod B], a
// (32 x K) x (K x 160) product with K = the tile's S B = 256 values. The 2 x 10 output
// tiles of 16 x 16 are split over the four wavefronts, five each (tiles wave + 4 j);
// v_mfma_f32_16x16x4f32, lane l: A = dz[..][16 mt + l % 16][k], B = in[..] at column
// 16 nt + l % 16, k = k0 + l / 16; acc[j][r] is row 16 mt + 4 (l / 16) + r, column
// 16 nt + l % 16. FP32 products and sums in the matrix unit's order.
template <int B, int S>
__device__ __forceinline__ void wgrad_mfma(const float (*D)[kC][B + 1], const float (*A)[kC][B + 1],
                                           cnn_f32x4 (&acc)[5], int wave, int lane) {
  const int m = lane & 15, kq = lane >> 4;
  int mt[5], ci[5], tt[5];
#pragma unroll
  for (int j = 0; j < 5; ++j) {
    const int T = wave + 4 * j;
    mt[j] = T / 10;
    const int n = (T % 10) * 16 + m;
    ci[j] = n / kK;
    tt[j] = n % kK;
  }
#pragma unroll 2
  for (int k0 = 0; k0 < S * B; k0 += 4) {
    const int k = k0 + kq;
    const int ss = k / B, q = k % B;
    const float a0 = D[ss][m][q], a1 = D[ss][16 + m][q];
#pragma unroll
    for (int j = 0; j < 5; ++j) {
      const float b = A[ss][ci[j]][(q + tt[j] - 2) & (B - 1)];
      acc[j] = __builtin_amdgcn_mfma_f32_16x16x4f32(mt[j] == 0 ? a0 : a1, b, acc[j], 0, 0, 0);
    }
  }
}

// Data gradient of a 32 -> 32 conv layer on the matrix cores: dz[ss][ci][p] = sum over
// k = (co, t) of W[co][ci][t] d[ss][co][(p - t + 2) mod B], per sample a (32 x 160) x
// (160 x B) product. The tile's S samples x 2 channel tiles x B/16 bin tiles (32 tiles of
// 16 x 16 for every B) are split over the four wavefronts, eight each (tiles wave + 4 j);
// lane l: A = W[co][16 mt + l % 16][t], B = d[ss][co][bin of column 16 nt + l % 16], with
// (co, t) = k0 + l / 16; acc[j][r] is channel 16 mt + 4 (l / 16) + r at bin 16 nt + l % 16.
template <int B, int S>
__device__ __forceinline__ void dgrad_mfma(const float* __restrict__ W,
                                           const float (*D)[kC][B + 1], cnn_f32x4 (&acc)[8],
                                           int wave, int lane) {
  constexpr int NT = B / 16;  // bin tiles
  const int m = lane & 15, kq = lane >> 4;
#pragma unroll
  for (int j = 0; j < 8; ++j) acc[j] = cnn_f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll 2
  for (int k0 = 0; k0 < kC * kK; k0 += 4) {
    const int k = k0 + kq;
    const int co = k / kK, t = k - co * kK;
    const float w0 = W[(co * kC + m) * kK + t], w1 = W[(co * kC + 16 + m) * kK + t];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int T = wave + 4 * j;
      const int ss = T / (2 * NT), rem = T - ss * 2 * NT;
      const int mt = rem / NT, nt = rem - mt * NT;
      const float b = D[ss][co][(nt * 16 + m - t + 2) & (B - 1)];
      acc[j] = __builtin_amdgcn_mfma_f32_16x16x4f32(mt == 0 ? w0 : w1, b, acc[j], 0, 0, 0);
    }
  }
}

// dz masked by the layer input's relu (act > 0), stored over D (after every wavefront has
// finished reading D).
template <int B>
__device__ __forceinline__ void dgrad_store(float (*D)[kC][B + 1], const float (*A)[kC][B + 1],
                                            const cnn_f32x4 (&acc)[8], int wave, int lane) {
  constexpr int NT = B / 16;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int T = wave + 4 * j;
    const int ss = T / (2 * NT), rem = T - ss * 2 * NT;
    const int mt = rem / NT, nt = rem - mt * NT;
    const int pb = nt * 16 + (lane & 15);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int ci = mt * 16 + 4 * (lane >> 4) + r;
      D[ss][ci][pb] = A[ss][ci][pb] > 0.0f ? acc[j][r] : 0.0f;
    }
  }
}

template <int B>
__device__ __forceinline__ void wgrad_store(float* __restrict__ out, const cnn_f32x4 (&acc)[5],
                                            int wave, int lane) {
#pragma unroll
  for (int j = 0; j < 5; ++j) {
    const int T = wave + 4 * j;
    const int n = (T % 10) * 16 + (lane & 15);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int co = (T / 10) * 16 + 4 * (lane >> 4) + r;
      out[co * kC * kK + n] = acc[j][r];
    }
  }
}

template <int B>
__global__ void __launch_bounds__(kThreads)
cnn_backward_kernel(const float* __restrict__ th, const float* __restrict__ cache,
                    const float* __restrict__ logits, const float* __restrict__ labels,
                    int64_t n, float w_pos, float w_neg, float* __restrict__ partial) {
  constexpr int S = kThreads / B;
  __shared__ float Aa[S][kC][B + 1];  // a2, then a1 (rows padded: the products' reads)
  __shared__ float D[S][kC][B + 1];   // dz3, dz2, dz1
  __shared__ float X[S][B];
  __shared__ float DQ[S][kH], ZZ[S][kZ], UU[S][kH], DH[S][kC], DL[S];
  __shared__ int ARG[S][kC];
  const int tid = threadIdx.x;
  const int s = tid / B, p = tid % B;
  const int wave = tid >> 6, lane = tid & 63;
  cnn_f32x4 g3[5], g2[5];
  float gf1[kF1PerThread];
#pragma unroll
  for (int j = 0; j < 5; ++j) g3[j] = g2[j] = cnn_f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int r = 0; r < kF1PerThread; ++r) gf1[r] = 0.0f;
  float g1 = 0.0f, gb1 = 0.0f, gb2 = 0.0f, gb3 = 0.0f, gbf1 = 0.0f, gf2 = 0.0f, gbf2 = 0.0f;
  const int64_t ntiles = (n + S - 1) / S;
  for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const int64_t i = tile * S + s;
    const bool valid = i < n;
    const float* cs = cache + (valid ? i * cache_stride(B) : 0);
    if (p < kH) UU[s][p] = valid ? fmaxf(cs[off_q(B) + p], 0.0f) : 0.0f;
    for (int z = p; z < kZ; z += B) ZZ[s][z] = valid ? cs[off_z(B) + z] : 0.0f;
    if (p < kC) ARG[s][p] = valid ? reinterpret_cast<const int32_t*>(cs + off_arg(B))[p] : 0;
    if (p == 0) {
      float dl = 0.0f;
      if (valid) {
        const float lg = logits[i];
        const float y = labels[i];
        const float sig = 1.0f / (1.0f + expf(-lg));
        dl = w_pos * y * (sig - 1.0f) + w_neg * (1.0f - y) * sig;
      }
      DL[s] = dl;
    }
    __syncthreads();
    if (p < kH) {
      const float q = valid ? cs[off_q(B) + p] : 0.0f;
      DQ[s][p] = q > 0.0f ? DL[s] * th[oF2 + p] : 0.0f;
    }
    __syncthreads();
    if (p < kC) {
      float dh = 0.0f;
      for (int j = 0; j < kH; ++j) dh += DQ[s][j] * th[oF1 + j * kZ + p];
      DH[s][p] = ZZ[s][p] > 0.0f ? dh : 0.0f;  // relu mask of a3 at its maximum
    }
    __syncthreads();
    for (int c = 0; c < kC; ++c) {
      D[s][c][p] = (p == ARG[s][c]) ? DH[s][c] : 0.0f;
      Aa[s][c][p] = valid ? cs[off_a2(B) + c * B + p] : 0.0f;
    }
    __syncthreads();
    // Head gradients.
    if (tid < kH) {
      for (int ss = 0; ss < S; ++ss) {
        gf2 += DL[ss] * UU[ss][tid];
        gbf1 += DQ[ss][tid];
      }
    }
    if (tid == 0)
      for (int ss = 0; ss < S; ++ss) gbf2 += DL[ss];
#pragma unroll
    for (int r = 0; r < kF1PerThread; ++r) {
      const int o = tid + kThreads * r;
      if (o < kH * kZ) {
        const int j = o / kZ, z = o % kZ;
        for (int ss = 0; ss < S; ++ss) gf1[r] += DQ[ss][j] * ZZ[ss][z];
      }
    }
    // conv3: weight gradients from dz3 and a2, then da2 -> dz2 (kept in registers).
    wgrad_mfma<B, S>(D, Aa, g3, wave, lane);
    if (tid < kC)
      for (int ss = 0; ss < S; ++ss)
#pragma unroll 1
        for (int q = 0; q < B; ++q) gb3 += D[ss][tid][q];
    cnn_f32x4 dz[8];
    dgrad_mfma<B, S>(th + oW3, D, dz, wave, lane);
    __syncthreads();
    dgrad_store<B>(D, Aa, dz, wave, lane);  // relu mask of a2
    __syncthreads();
#pragma unroll
    for (int c = 0; c < kC; ++c) Aa[s][c][p] = valid ? cs[c * B + p] : 0.0f;  // a1
    __syncthreads();
    // conv2: weight gradients from dz2 and a1, then da1 -> dz1.
    wgrad_mfma<B, S>(D, Aa, g2, wave, lane);
    if (tid < kC)
      for (int ss = 0; ss < S; ++ss)
#pragma unroll 1
        for (int q = 0; q < B; ++q) gb2 += D[ss][tid][q];
    dgrad_mfma<B, S>(th + oW2, D, dz, wave, lane);
    __syncthreads();
    dgrad_store<B>(D, Aa, dz, wave, lane);  // relu mask of a1
    X[s][p] = valid ? cs[off_x0(B) + p] : 0.0f;
    __syncthreads();
    // conv1: weight gradients from dz1 and x0.
    if (tid < kC * kK) {
      const int co = tid / kK, t = tid % kK;
      for (int ss = 0; ss < S; ++ss)
#pragma unroll 1
        for (int q = 0; q < B; ++q) g1 += D[ss][co][q] * X[ss][(q + t - 2) & (B - 1)];
    }
    if (tid < kC)
      for (int ss = 0; ss < S; ++ss)
#pragma unroll 1
        for (int q = 0; q < B; ++q) gb1 += D[ss][tid][q];
    __syncthreads();
  }
  float* out = partial + (int64_t)blockIdx.x * kNParams;
  wgrad_store<B>(out + oW3, g3, wave, lane);
  wgrad_store<B>(out + oW2, g2, wave, lane);
#pragma unroll
  for (int r = 0; r < kF1PerThread; ++r) {
    const int o = tid + kThreads * r;
    if (o < kH * kZ) out[oF1 + o] = gf1[r];
  }
  if (tid < kC * kK) out[oW1 + tid] = g1;
  if (tid < kC) {
    out[oB1 + tid] = gb1;
    out[oB2 + tid] = gb2;
    out[oB3 + tid] = gb3;
  }
  if (tid < kH) {
    out[oBF1 + tid] = gbf1;
    out[oF2 + tid] = gf2;
  }
  if (tid == 0) out[oBF2] = gbf2;
}

// Sums of the blocks' partial gradients in fixed order: blockIdx.y = g adds the partials
// [g per, (g + 1) per) (of npart) in order into out[g]. The gradient takes one launch of
// groups of kGroup partials and one over the group sums (one launch when npart <= kGroup),
// so it depends only on the number of partials.
constexpr int kGroup = 32;
__global__ void cnn_reduce_kernel(const float* __restrict__ partial, int npart, int per,
                                  float* __restrict__ out) {
  const int o = blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= kNParams) return;
  const int b0 = (int)blockIdx.y * per;
  const int b1 = std::min(npart, b0 + per);
  float a = 0.0f;
  for (int b = b0; b < b1; ++b) a += partial[(int64_t)b * kNParams + o];
  out[(int64_t)blockIdx.y * kNParams + o] = a;
}

__global__ void cnn_adam_kernel(float* __restrict__ th, const float* __restrict__ g,
                                float* __restrict__ m, float* __restrict__ v, int64_t np,
                                float lr, float b1, float b2, float eps, float c1, float c2) {
  const int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= np) return;
  const float gi = g[o];
  const float mi = b1 * m[o] + (1.0f - b1) * gi;
  const float vi = b2 * v[o] + (1.0f - b2) * gi * gi;
  m[o] = mi;
  v[o] = vi;
  th[o] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// ---- Inference on MFMA (BF16 operands, FP32 accumulation) -----------------------------
// The two 32 -> 32 convolutions are dense contractions, out[co][p] = sum over (t, ci) of
// W[co][t][ci] act[(p + t - 2) mod B][ci]: a 32 x B x 160 product per sample, done with
// v_mfma_f32_32x32x8_bf16 (K ordered tap-major, so that a lane's 4 K values are 4
// consecutive channels at one position: one 8-byte LDS read per operand). One wavefront
// owns one sample at a time; its activations live in wave-private LDS as bf16 [p][ci];
// the weights are staged once per block as bf16 [co][t][ci]. The profile, the first
// convolution (1 -> 32), the max-pool and the head stay FP32. Used for ranking only: the
// logits differ from the FP32 path by the bf16 rounding of weights and activations.
typedef short bf16x4_t __attribute__((ext_vector_type(4)));
typedef float f32x16_t __attribute__((ext_vector_type(16)));

__device__ __forceinline__ uint16_t to_bf16(float x) {  // round to nearest even
  uint32_t u = __float_as_uint(x);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

__device__ __forceinline__ void wave_sync_lds() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// pyl::irfft_bin's arithmetic, in its order, over a row staged in LDS (partly unrolled:
// full unrolling keeps every harmonic in registers).
template <int B>
__device__ __forceinline__ float irfft_bin_lds(const float2* X, int b, const float* tw) {
  constexpr int nb_f = B / 2 + 1;
  float acc = X[0].x;
  const float ny = X[nb_f - 1].x;
  acc += (b & 1) ? -ny : ny;
#pragma unroll 4
  for (int kk = 1; kk < nb_f - 1; ++kk) {
    const int idx = (kk * b) & (B - 1);
    acc += 2.0f * (X[kk].x * tw[idx] - X[kk].y * tw[B + idx]);
  }
  return acc / (float)B;
}

// One 32 -> 32 convolution of wave-private act_in into the 16 accumulators per N-tile.
// Lane l: column p = 32 nt + l % 32; acc[r] is row co = 8 (r / 4) + 4 (l / 32) + r % 4.
template <int B>
__device__ __forceinline__ void conv_mfma(const uint16_t* __restrict__ sw,
                                          const uint16_t* __restrict__ act, int nt, int lane,
                                          f32x16_t& acc) {
  const int row = lane & 31, kg = lane >> 5;
  for (int r = 0; r < 16; ++r) acc[r] = 0.0f;
#pragma unroll
  for (int t = 0; t < kK; ++t) {
    const int p = (32 * nt + row + t - 2) & (B - 1);
#pragma unroll
    for (int cs = 0; cs < kC / 8; ++cs) {
      const bf16x4_t a =
          *reinterpret_cast<const bf16x4_t*>(sw + (row * kK + t) * kC + 8 * cs + 4 * kg);
      const bf16x4_t b = *reinterpret_cast<const bf16x4_t*>(act + p * kC + 8 * cs + 4 * kg);
      acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, acc, 0, 0, 0);
    }
  }
}

template <int B, bool kComplex>
__global__ void __launch_bounds__(kThreads)
cnn_forward_mfma_kernel(const float* __restrict__ th, const float* __restrict__ folds,
                        int64_t n_arg, const int64_t* __restrict__ n_dev,
                        const float* __restrict__ feats, float* __restrict__ logits) {
  constexpr int kW = kThreads / 64;
  __shared__ __align__(16) uint16_t sW2[kC * kK * kC];
  __shared__ __align__(16) uint16_t sW3[kC * kK * kC];
  __shared__ __align__(16) uint16_t sAct[kW][2][B * kC];
  __shared__ float sX[kW][B];
  __shared__ float sH[kW][kC];
  __shared__ float tw[kComplex ? 2 * B : 1];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int64_t n = n_dev ? *n_dev : n_arg;
  if ((int64_t)blockIdx.x * kW >= n) return;  // the whole block
  for (int j = tid; j < kC * kC * kK; j += kThreads) {  // W[co][ci][t] -> [co][t][ci]
    const int co = j / (kC * kK), rem = j % (kC * kK), ci = rem / kK, t = rem % kK;
    sW2[(co * kK + t) * kC + ci] = to_bf16(th[oW2 + j]);
    sW3[(co * kK + t) * kC + ci] = to_bf16(th[oW3 + j]);
  }
  if constexpr (kComplex) pyl::fill_twiddles(tw, B, tid, kThreads);
  __syncthreads();
  float* x = sX[wave];
  uint16_t* a1 = sAct[wave][0];
  uint16_t* a2 = sAct[wave][1];
  for (int64_t i = (int64_t)blockIdx.x * kW + wave; i < n; i += (int64_t)gridDim.x * kW) {
    // Profile e / sqrt(v), standardised over the bins (FP32). A complex row is staged in
    // the wave's activation buffer first (one coalesced load), then inverse-transformed.
    float sum = 0.0f;
    float2* xs = reinterpret_cast<float2*>(a1);
    if constexpr (kComplex) {
      const float2* rowp = reinterpret_cast<const float2*>(folds) + i * 2 * (B / 2 + 1);
      for (int j = lane; j < 2 * (B / 2 + 1); j += 64) xs[j] = rowp[j];
      wave_sync_lds();
    }
    for (int p = lane; p < B; p += 64) {
      float e, v;
      if constexpr (kComplex) {
        e = irfft_bin_lds<B>(xs, p, tw);
        v = irfft_bin_lds<B>(xs + (B / 2 + 1), p, tw);
      } else {
        e = folds[i * 2 * B + p];
        v = folds[i * 2 * B + B + p];
      }
      const float pv = pyl::norm_ev(e, v);
      x[p] = pv;
      sum += pv;
    }
    for (int o = 32; o > 0; o >>= 1) sum += __shfl_xor(sum, o);
    const float mu = sum / (float)B;
    float ss = 0.0f;
    for (int p = lane; p < B; p += 64) {
      const float d = x[p] - mu;
      ss += d * d;
    }
    for (int o = 32; o > 0; o >>= 1) ss += __shfl_xor(ss, o);
    const float sd = sqrtf(ss / (float)B);
    for (int p = lane; p < B; p += 64) x[p] = (x[p] - mu) / (sd + 1e-6f);
    wave_sync_lds();
    // conv1 (1 -> 32), FP32, relu, stored bf16 [p][co].
    for (int p = lane; p < B; p += 64) {
      float xv[kK];
#pragma unroll
      for (int t = 0; t < kK; ++t) xv[t] = x[(p + t - 2) & (B - 1)];
#pragma unroll 4
      for (int co = 0; co < kC; ++co) {
        float acc = th[oB1 + co];
#pragma unroll
        for (int t = 0; t < kK; ++t) acc += th[oW1 + co * kK + t] * xv[t];
        a1[p * kC + co] = to_bf16(fmaxf(acc, 0.0f));
      }
    }
    wave_sync_lds();
    // conv2 on MFMA, bias + relu, stored bf16.
    const int kg = lane >> 5;
    for (int nt = 0; nt < B / 32; ++nt) {
      f32x16_t acc;
      conv_mfma<B>(sW2, a1, nt, lane, acc);
      const int p = 32 * nt + (lane & 31);
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int co = 8 * (r >> 2) + 4 * kg + (r & 3);
        a2[p * kC + co] = to_bf16(fmaxf(acc[r] + th[oB2 + co], 0.0f));
      }
    }
    wave_sync_lds();
    // conv3 on MFMA, max over positions (relu and bias commute with it).
    float mx[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) mx[r] = -FLT_MAX;
    for (int nt = 0; nt < B / 32; ++nt) {
      f32x16_t acc;
      conv_mfma<B>(sW3, a2, nt, lane, acc);
#pragma unroll
      for (int r = 0; r < 16; ++r) mx[r] = fmaxf(mx[r], acc[r]);
    }
#pragma unroll
    for (int r = 0; r < 16; ++r)
      for (int o = 16; o > 0; o >>= 1) mx[r] = fmaxf(mx[r], __shfl_xor(mx[r], o));
    if ((lane & 31) == 0) {
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int co = 8 * (r >> 2) + 4 * kg + (r & 3);
        sH[wave][co] = fmaxf(mx[r] + th[oB3 + co], 0.0f);
      }
    }
    wave_sync_lds();
    // Head (FP32): z = (h, features); u = relu(Wf1 z + bf1); logit = Wf2 . u + bf2.
    float part = 0.0f;
    if (lane < kH) {
      float q = th[oBF1 + lane];
      for (int z = 0; z < kC; ++z) q += th[oF1 + lane * kZ + z] * sH[wave][z];
      for (int f = 0; f < kF; ++f) q += th[oF1 + lane * kZ + kC + f] * feats[i * kF + f];
      part = th[oF2 + lane] * fmaxf(q, 0.0f);
    }
    for (int o = 32; o > 0; o >>= 1) part += __shfl_xor(part, o);
    if (lane == 0) logits[i] = th[oBF2] + part;
    wave_sync_lds();
  }
}

template <int B>
int launch_forward_mfma(const float* th, const float* folds, int64_t n, const int64_t* n_dev,
                        const float* feats, float* logits, int complex_folds, hipStream_t s) {
  constexpr int kW = kThreads / 64;
  // At most two blocks per CU: a block stages the weights once, then its wavefronts
  // stride over the samples.
  const unsigned grid = (unsigned)std::max<int64_t>(
      1, std::min<int64_t>(pyl::ceil_div(n, kW), 512));
  if (complex_folds)
    cnn_forward_mfma_kernel<B, true><<<grid, kThreads, 0, s>>>(th, folds, n, n_dev, feats, logits);
  else
    cnn_forward_mfma_kernel<B, false><<<grid, kThreads, 0, s>>>(th, folds, n, n_dev, feats, logits);
  PYL_CHECK_LAUNCH("cnn_forward_mfma_kernel");
  return PYL_OK;
}

// n samples, or at most n with the count read from n_dev.
template <int B>
int launch_forward(const float* th, const float* folds, int64_t n, const int64_t* n_dev,
                   const float* feats, float* logits, float* cache, int complex_folds,
                   hipStream_t s) {
  constexpr int S = kThreads / B;
  const unsigned grid = (unsigned)pyl::ceil_div(n, S);
  if (complex_folds) {
    if (cache) cnn_forward_kernel<B, true, true><<<grid, kThreads, 0, s>>>(th, folds, n, n_dev, feats, logits, cache);
    else cnn_forward_kernel<B, true, false><<<grid, kThreads, 0, s>>>(th, folds, n, n_dev, feats, logits, cache);
  } else {
    if (cache) cnn_forward_kernel<B, false, true><<<grid, kThreads, 0, s>>>(th, folds, n, n_dev, feats, logits, cache);
    else cnn_forward_kernel<B, false, false><<<grid, kThreads, 0, s>>>(th, folds, n, n_dev, feats, logits, cache);
  }
  PYL_CHECK_LAUNCH("cnn_forward_kernel");
  return PYL_OK;
}

bool valid_bins(int nbins) { return nbins == 32 || nbins == 64 || nbins == 128; }

}  // namespace

extern "C" int pyl_cnn_num_params(void) { return kNParams; }

extern "C" int64_t pyl_cnn_cache_floats(int nbins) {
  return valid_bins(nbins) ? cache_stride(nbins) : -1;
}

extern "C" int pyl_cnn_forward(const float* params, const float* folds, int64_t n, int nbins,
                               int complex_folds, const float* feats, float* logits,
                               float* cache, void* stream) {
  PYL_REQUIRE(valid_bins(nbins), "the ranking network takes 32, 64 or 128 bins (got %d)", nbins);
  PYL_REQUIRE(n >= 0, "negative sample count");
  if (n == 0) return PYL_OK;
  PYL_REQUIRE(params && folds && feats && logits, "null device pointer");
  hipStream_t s = pyl::as_stream(stream);
  switch (nbins) {
    case 32: return launch_forward<32>(params, folds, n, nullptr, feats, logits, cache, complex_folds, s);
    case 64: return launch_forward<64>(params, folds, n, nullptr, feats, logits, cache, complex_folds, s);
    default: return launch_forward<128>(params, folds, n, nullptr, feats, logits, cache, complex_folds, s);
  }
}

extern "C" int pyl_cnn_forward_bf16(const float* params, const float* folds, int64_t n,
                                    int nbins, int complex_folds, const float* feats,
                                    float* logits, void* stream) {
  PYL_REQUIRE(valid_bins(nbins), "the ranking network takes 32, 64 or 128 bins (got %d)", nbins);
  PYL_REQUIRE(n >= 0, "negative sample count");
  if (n == 0) return PYL_OK;
  PYL_REQUIRE(params && folds && feats && logits, "null device pointer");
  return pyl::cnn_forward_count(params, folds, n, nullptr, nbins, complex_folds, feats, logits,
                                1, pyl::as_stream(stream));
}

// The forward pass over at most n_cap samples, the count read from n_dev when not null
// (no cache); bf16: the MFMA inference kernel.
int pyl::cnn_forward_count(const float* params, const float* folds, int64_t n_cap,
                           const int64_t* n_dev, int nbins, int complex_folds,
                           const float* feats, float* logits, int bf16, hipStream_t s) {
  PYL_REQUIRE(valid_bins(nbins), "the ranking network takes 32, 64 or 128 bins (got %d)", nbins);
  if (n_cap <= 0) return PYL_OK;
  if (bf16) {
    switch (nbins) {
      case 32: return launch_forward_mfma<32>(params, folds, n_cap, n_dev, feats, logits, complex_folds, s);
      case 64: return launch_forward_mfma<64>(params, folds, n_cap, n_dev, feats, logits, complex_folds, s);
      default: return launch_forward_mfma<128>(params, folds, n_cap, n_dev, feats, logits, complex_folds, s);
    }
  }
  switch (nbins) {
    case 32: return launch_forward<32>(params, folds, n_cap, n_dev, feats, logits, nullptr, complex_folds, s);
    case 64: return launch_forward<64>(params, folds, n_cap, n_dev, feats, logits, nullptr, complex_folds, s);
    default: return launch_forward<128>(params, folds, n_cap, n_dev, feats, logits, nullptr, complex_folds, s);
  }
}

extern "C" int pyl_cnn_backward(const float* params, const float* cache, const float* logits,
                                const float* labels, int64_t n, int nbins, float w_pos,
                                float w_neg, float* work, int64_t work_floats, float* grad,
                                void* stream) {
  PYL_REQUIRE(valid_bins(nbins), "the ranking network takes 32, 64 or 128 bins (got %d)", nbins);
  PYL_REQUIRE(params && cache && logits && labels && work && grad, "null device pointer");
  PYL_REQUIRE(work_floats >= kNParams, "workspace must hold at least %d floats", kNParams);
  hipStream_t s = pyl::as_stream(stream);
  const int S = kThreads / nbins;
  const int64_t tiles = std::max<int64_t>(1, (n + S - 1) / S);
  // Blocks: one per tile up to 2048 and the workspace's room, which also holds the group
  // sums of the two-level reduction (one per kGroup blocks) after the partials.
  const int64_t room = work_floats / kNParams;
  int nblk = (int)std::min<int64_t>({tiles, (int64_t)2048, room});
  while (nblk > kGroup && nblk + (nblk + kGroup - 1) / kGroup > room) --nblk;
  const int ngroup = nblk > kGroup ? (nblk + kGroup - 1) / kGroup : 1;
  switch (nbins) {
    case 32: cnn_backward_kernel<32><<<nblk, kThreads, 0, s>>>(params, cache, logits, labels, n, w_pos, w_neg, work); break;
    case 64: cnn_backward_kernel<64><<<nblk, kThreads, 0, s>>>(params, cache, logits, labels, n, w_pos, w_neg, work); break;
    default: cnn_backward_kernel<128><<<nblk, kThreads, 0, s>>>(params, cache, logits, labels, n, w_pos, w_neg, work); break;
  }
  PYL_CHECK_LAUNCH("cnn_backward_kernel");
  const unsigned pblk = (unsigned)pyl::ceil_div(kNParams, 256);
  if (ngroup == 1) {
    cnn_reduce_kernel<<<dim3(pblk, 1), 256, 0, s>>>(work, nblk, nblk, grad);
  } else {
    float* groups = work + (int64_t)nblk * kNParams;
    cnn_reduce_kernel<<<dim3(pblk, (unsigned)ngroup), 256, 0, s>>>(work, nblk, kGroup, groups);
    PYL_CHECK_LAUNCH("cnn_reduce_kernel");
    cnn_reduce_kernel<<<dim3(pblk, 1), 256, 0, s>>>(groups, ngroup, ngroup, grad);
  }
  PYL_CHECK_LAUNCH("cnn_reduce_kernel");
  return PYL_OK;
}

extern "C" int pyl_cnn_adam(float* params, const float* grad, float* m, float* v,
                            int64_t nparams, float lr, float beta1, float beta2, float eps,
                            int step, void* stream) {
  PYL_REQUIRE(params && grad && m && v, "null device pointer");
  PYL_REQUIRE(step >= 1, "Adam step counts from 1");
  if (nparams <= 0) return PYL_OK;
  const float c1 = 1.0f - (float)std::pow((double)beta1, step);
  const float c2 = 1.0f - (float)std::pow((double)beta2, step);
  cnn_adam_kernel<<<(unsigned)pyl::ceil_div(nparams, 256), 256, 0, pyl::as_stream(stream)>>>(
      params, grad, m, v, nparams, lr, beta1, beta2, eps, c1, c2);
  PYL_CHECK_LAUNCH("cnn_adam_kernel");
  return PYL_OK;
}
