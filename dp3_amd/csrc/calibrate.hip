// Bandwidth calibration: a plain device copy and a plain streaming read of
// known sizes. Their rates put the solver's HBM figures in context (the
// achievable fraction of the 8 TB/s peak), and their FETCH_SIZE/WRITE_SIZE
// counter readings scale the counters of the solver kernels.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "../../include/ddecal_hip.h"
#include "common.h"

namespace ddh {
namespace {

typedef float float4_v __attribute__((ext_vector_type(4)));

__global__ void __launch_bounds__(256) CopyKernel(const float4_v* __restrict__ src,
                                                  float4_v* __restrict__ dst, size_t n) {
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += size_t(gridDim.x) * blockDim.x)
    dst[i] = __builtin_nontemporal_load(src + i);
}

__global__ void __launch_bounds__(256) ReadKernel(const float4_v* __restrict__ src, size_t n,
                                                  float* __restrict__ out) {
  float acc = 0.0f;
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += size_t(gridDim.x) * blockDim.x) {
    const float4_v v = __builtin_nontemporal_load(src + i);
    acc += v.x + v.y + v.z + v.w;
  }
  out[size_t(blockIdx.x) * blockDim.x + threadIdx.x] = acc;
}

__global__ void FillKernel(float4_v* dst, size_t n) {
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += size_t(gridDim.x) * blockDim.x)
    dst[i] = float4_v{1.0f, 2.0f, 3.0f, 4.0f};
}

}  // namespace
}  // namespace ddh

using namespace ddh;

extern "C" int ddh_calibrate_bandwidth(int64_t bytes, int32_t repeats, double* copy_gbs,
                                       double* read_gbs) {
  return Guard([&] {
    DDH_REQUIRE(bytes >= (1 << 20) && repeats >= 1, -1, "bytes must be >= 1 MiB");
    const size_t n = size_t(bytes) / sizeof(float4_v);
    float4_v* a = nullptr;
    float4_v* b = nullptr;
    float* out = nullptr;
    int dev = 0, n_cu = 0;
    DDH_HIP_CHECK(hipGetDevice(&dev));
    DDH_HIP_CHECK(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev));
    const int blocks = std::max(1, n_cu) * 8;
    DDH_HIP_CHECK(hipMalloc(&a, n * sizeof(float4_v)));
    DDH_HIP_CHECK(hipMalloc(&b, n * sizeof(float4_v)));
    DDH_HIP_CHECK(hipMalloc(&out, size_t(blocks) * 256 * sizeof(float)));
    hipEvent_t e0, e1, e2;
    DDH_HIP_CHECK(hipEventCreate(&e0));
    DDH_HIP_CHECK(hipEventCreate(&e1));
    DDH_HIP_CHECK(hipEventCreate(&e2));
    hipLaunchKernelGGL(FillKernel, dim3(blocks), dim3(256), 0, nullptr, a, n);
    hipLaunchKernelGGL(CopyKernel, dim3(blocks), dim3(256), 0, nullptr, a, b, n);  // warm-up
    hipLaunchKernelGGL(ReadKernel, dim3(blocks), dim3(256), 0, nullptr, a, n, out);
    DDH_HIP_CHECK(hipEventRecord(e0, nullptr));
    for (int r = 0; r < repeats; ++r)
      hipLaunchKernelGGL(CopyKernel, dim3(blocks), dim3(256), 0, nullptr, a, b, n);
    DDH_HIP_CHECK(hipEventRecord(e1, nullptr));
    for (int r = 0; r < repeats; ++r)
      hipLaunchKernelGGL(ReadKernel, dim3(blocks), dim3(256), 0, nullptr, a, n, out);
    DDH_HIP_CHECK(hipEventRecord(e2, nullptr));
    DDH_HIP_CHECK(hipEventSynchronize(e2));
    float ms_copy = 0.0f, ms_read = 0.0f;
    DDH_HIP_CHECK(hipEventElapsedTime(&ms_copy, e0, e1));
    DDH_HIP_CHECK(hipEventElapsedTime(&ms_read, e1, e2));
    const double moved = double(n) * sizeof(float4_v);
    if (copy_gbs) *copy_gbs = 2.0 * moved * repeats / (ms_copy * 1e-3) / 1e9;
    if (read_gbs) *read_gbs = moved * repeats / (ms_read * 1e-3) / 1e9;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipEventDestroy(e2);
    DDH_HIP_CHECK(hipFree(a));
    DDH_HIP_CHECK(hipFree(b));
    DDH_HIP_CHECK(hipFree(out));
  });
}
