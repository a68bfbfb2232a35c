// Bandwidth probes of known size, for the roofline of the search kernels: a pure read
// (every element of a buffer loaded once, one float written per block) and a plain copy.
// Timed on the device they give the attainable HBM bandwidth; profiled with the
// FETCH_SIZE / WRITE_SIZE counters they give the factor that turns those counters into
// bytes.
#include "pyl_common.hpp"

namespace {

__global__ void __launch_bounds__(256)
hbm_read_kernel(const float4* __restrict__ src, int64_t n, float* __restrict__ partial) {
  float acc = 0.0f;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const float4 v = src[i];
    acc += (v.x + v.y) + (v.z + v.w);
  }
  __shared__ float s[256];
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = s[0];
}

__global__ void __launch_bounds__(256)
hbm_copy_kernel(const float4* __restrict__ src, float4* __restrict__ dst, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

}  // namespace

extern "C" int pyl_hbm_read(const float* src, int64_t n_float4, float* partial,
                            int32_t nblocks, void* stream) {
  PYL_REQUIRE(src && partial, "null device pointer");
  PYL_REQUIRE(n_float4 > 0 && nblocks > 0, "empty read");
  hbm_read_kernel<<<(unsigned)nblocks, 256, 0, pyl::as_stream(stream)>>>(
      reinterpret_cast<const float4*>(src), n_float4, partial);
  PYL_CHECK_LAUNCH("hbm_read_kernel");
  return PYL_OK;
}

extern "C" int pyl_hbm_copy(const float* src, float* dst, int64_t n_float4, int32_t nblocks,
                            void* stream) {
  PYL_REQUIRE(src && dst, "null device pointer");
  PYL_REQUIRE(n_float4 > 0 && nblocks > 0, "empty copy");
  hbm_copy_kernel<<<(unsigned)nblocks, 256, 0, pyl::as_stream(stream)>>>(
      reinterpret_cast<const float4*>(src), reinterpret_cast<float4*>(dst), n_float4);
  PYL_CHECK_LAUNCH("hbm_copy_kernel");
  return PYL_OK;
}
