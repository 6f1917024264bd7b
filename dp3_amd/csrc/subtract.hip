// Applying the solutions to the model data and subtracting them from the data
// after a solve: DDECal::CorrectAndSubtractModels (steps/DDECal.cc:1074-1124)
// with ApplySolutionAndSubtract (:1031-1072) and the 2x2 products of
// ApplyFullJonesSolution / ApplyDiagonalSolution / ApplyScalarSolution
// (:992-1020), which the reference evaluates in double (aocommon::MC2x2).
//
// Buffers are in DPBuffer order for the time steps of one solution interval:
// data [n_t][n_bl][n_ch][4] and model [n_dir][n_t][n_bl][n_ch][4] complex64,
// solutions [n_cb][n_ant][n_dir][n_pol] complex128. One thread per
// visibility; HBM bound: per visibility 32 B of data read and written and
// n_dir x 32 B of model read (and written with keep_model).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "../../include/ddecal_hip.h"
#include "common.h"

namespace ddh {
namespace {

typedef float float4_v __attribute__((ext_vector_type(4)));

struct C2 {
  double re, im;
};
__device__ __forceinline__ C2 Mul(C2 a, C2 b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
// a * conj(b)
__device__ __forceinline__ C2 MulC(C2 a, C2 b) { return {a.re * b.re + a.im * b.im, a.im * b.re - a.re * b.im}; }
__device__ __forceinline__ C2 Add(C2 a, C2 b) { return {a.re + b.re, a.im + b.im}; }

constexpr int kDirBatch = 4;
constexpr int kMaxSubtractBlocks = 1024;

template <int NPOL>
__global__ void __launch_bounds__(256)
    CorrectSubtractKernel(float4_v* __restrict__ data, float4_v* __restrict__ model,
                          const double2* __restrict__ sol, const int32_t* __restrict__ ant1,
                          const int32_t* __restrict__ ant2, const int32_t* __restrict__ cb_start,
                          int n_t, int n_bl, int n_ch, int n_cb, int n_dir, int n_ant,
                          int subtract, int keep_model) {
  const int64_t n_vis = int64_t(n_t) * n_bl * n_ch;
  // the block starts in LDS for the bisection below (the host checks the size)
  __shared__ int32_t s_start[kMaxSubtractBlocks + 1];
  for (int i = threadIdx.x; i <= n_cb; i += blockDim.x) s_start[i] = cb_start[i];
  __syncthreads();
  const bool small = n_vis < (int64_t(1) << 32);
  for (int64_t v = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; v < n_vis;
       v += int64_t(gridDim.x) * blockDim.x) {
    int ch, bl;
    if (small) {  // 32-bit division where the index fits
      const uint32_t u = uint32_t(v), q = u / uint32_t(n_ch);
      ch = int(u - q * uint32_t(n_ch));
      bl = int(q % uint32_t(n_bl));
    } else {
      ch = int(v % n_ch);
      bl = int((v / n_ch) % n_bl);
    }
    const int a1 = ant1[bl], a2 = ant2[bl];
    // The block of this channel (DDECal.cc:1106-1108): the first cb with
    // cb_start[cb + 1] > ch, by bisection over the increasing starts
    // (cb_start[n_cb] = n_ch ends it).
    int lo = 0, hi = n_cb - 1;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (s_start[mid + 1] > ch)
        hi = mid;
      else
        lo = mid + 1;
    }
    const int cb = lo;
    const double2* s = sol + size_t(cb) * n_ant * n_dir * NPOL;
    // the data first (read before the models, so its latency overlaps theirs)
    float4_v* dp = data + size_t(v) * 2;
    float4_v d0 = {0, 0, 0, 0}, d1 = {0, 0, 0, 0};
    if (subtract) {
      d0 = __builtin_nontemporal_load(dp);
      d1 = __builtin_nontemporal_load(dp + 1);
    }
    C2 sum[4] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
    // directions kDirBatch at a time: their model loads issued together, then
    // the products added in direction order
    for (int d0i = 0; d0i < n_dir; d0i += kDirBatch) {
      float4_v mv[kDirBatch][2];
#pragma unroll
      for (int j = 0; j < kDirBatch; ++j) {
        const int d = d0i + j < n_dir ? d0i + j : d0i;
        const float4_v* mp = model + (size_t(d) * n_vis + size_t(v)) * 2;
        mv[j][0] = __builtin_nontemporal_load(mp);
        mv[j][1] = __builtin_nontemporal_load(mp + 1);
      }
#pragma unroll
      for (int j = 0; j < kDirBatch; ++j) {
        const int d = d0i + j;
        if (d >= n_dir) break;
        float4_v* mp = model + (size_t(d) * n_vis + size_t(v)) * 2;
        const float4_v m0 = mv[j][0], m1 = mv[j][1];
        const C2 m[4] = {{m0.x, m0.y}, {m0.z, m0.w}, {m1.x, m1.y}, {m1.z, m1.w}};
        const double2* p1 = s + (size_t(a1) * n_dir + d) * NPOL;
        const double2* p2 = s + (size_t(a2) * n_dir + d) * NPOL;
        C2 out[4];
        if constexpr (NPOL == 1) {
          const C2 f = MulC({p1[0].x, p1[0].y}, {p2[0].x, p2[0].y});
          for (int c = 0; c < 4; ++c) out[c] = Mul(m[c], f);
        } else if constexpr (NPOL == 2) {
          const C2 s10 = {p1[0].x, p1[0].y}, s11 = {p1[1].x, p1[1].y};
          const C2 s20 = {p2[0].x, p2[0].y}, s21 = {p2[1].x, p2[1].y};
          out[0] = MulC(Mul(s10, m[0]), s20);
          out[1] = MulC(Mul(s10, m[1]), s21);
          out[2] = MulC(Mul(s11, m[2]), s20);
          out[3] = MulC(Mul(s11, m[3]), s21);
        } else {
          const C2 a[4] = {{p1[0].x, p1[0].y}, {p1[1].x, p1[1].y}, {p1[2].x, p1[2].y},
                           {p1[3].x, p1[3].y}};
          const C2 b[4] = {{p2[0].x, p2[0].y}, {p2[1].x, p2[1].y}, {p2[2].x, p2[2].y},
                           {p2[3].x, p2[3].y}};
          // t = S1 M, out = t S2^H
          const C2 t[4] = {Add(Mul(a[0], m[0]), Mul(a[1], m[2])),
                           Add(Mul(a[0], m[1]), Mul(a[1], m[3])),
                           Add(Mul(a[2], m[0]), Mul(a[3], m[2])),
                           Add(Mul(a[2], m[1]), Mul(a[3], m[3]))};
          out[0] = Add(MulC(t[0], b[0]), MulC(t[1], b[1]));
          out[1] = Add(MulC(t[0], b[2]), MulC(t[1], b[3]));
          out[2] = Add(MulC(t[2], b[0]), MulC(t[3], b[1]));
          out[3] = Add(MulC(t[2], b[2]), MulC(t[3], b[3]));
        }
#pragma unroll
        for (int c = 0; c < 4; ++c) sum[c] = Add(sum[c], out[c]);
        if (keep_model) {
          mp[0] = float4_v{float(out[0].re), float(out[0].im), float(out[1].re), float(out[1].im)};
          mp[1] = float4_v{float(out[2].re), float(out[2].im), float(out[3].re), float(out[3].im)};
        }
      }
    }
    if (subtract) {
      // complex<float> -= complex<double> (DDECal.cc:1070): the difference is
      // taken in double and rounded once
      d0.x = float(double(d0.x) - sum[0].re);
      d0.y = float(double(d0.y) - sum[0].im);
      d0.z = float(double(d0.z) - sum[1].re);
      d0.w = float(double(d0.w) - sum[1].im);
      d1.x = float(double(d1.x) - sum[2].re);
      d1.y = float(double(d1.y) - sum[2].im);
      d1.z = float(double(d1.z) - sum[3].re);
      d1.w = float(double(d1.w) - sum[3].im);
      dp[0] = d0;
      dp[1] = d1;
    }
  }
}

template <typename T>
struct Scratch {
  T* ptr = nullptr;
  ~Scratch() {
    if (ptr) (void)hipFree(ptr);
  }
  T* Get(size_t n) {
    DDH_HIP_CHECK(hipMalloc(reinterpret_cast<void**>(&ptr), std::max<size_t>(n, 1) * sizeof(T)));
    return ptr;
  }
};

void CheckShape(int32_t n_t, int32_t n_bl, int32_t n_ch, int32_t n_dir, int32_t n_ant,
                int32_t n_cb, int32_t n_pol) {
  DDH_REQUIRE(n_t >= 1 && n_bl >= 1 && n_ch >= 1 && n_dir >= 1 && n_ant >= 1 && n_cb >= 1, -1,
              "empty dimension");
  DDH_REQUIRE(n_pol == 1 || n_pol == 2 || n_pol == 4, -1,
              "solution polarizations must be 1, 2 or 4");
  DDH_REQUIRE(n_cb <= kMaxSubtractBlocks, -1, "at most 1024 channel blocks");
}

void Launch(int32_t n_t, int32_t n_bl, int32_t n_ch, int32_t n_cb, int32_t n_dir, int32_t n_ant,
            int32_t n_pol,
            const int32_t* d_ant1, const int32_t* d_ant2, const int32_t* d_cb_start,
            const double* d_sol, void* d_data, void* d_model, int subtract, int keep_model,
            hipStream_t stream) {
  const int64_t n_vis = int64_t(n_t) * n_bl * n_ch;
  const unsigned blocks = unsigned(std::min<int64_t>((n_vis + 255) / 256, 1 << 20));
  auto* data = reinterpret_cast<float4_v*>(d_data);
  auto* model = reinterpret_cast<float4_v*>(d_model);
  auto* sol = reinterpret_cast<const double2*>(d_sol);
  if (n_pol == 1)
    hipLaunchKernelGGL(CorrectSubtractKernel<1>, dim3(blocks), dim3(256), 0, stream, data, model,
                       sol, d_ant1, d_ant2, d_cb_start, n_t, n_bl, n_ch, n_cb, n_dir, n_ant, subtract,
                       keep_model);
  else if (n_pol == 2)
    hipLaunchKernelGGL(CorrectSubtractKernel<2>, dim3(blocks), dim3(256), 0, stream, data, model,
                       sol, d_ant1, d_ant2, d_cb_start, n_t, n_bl, n_ch, n_cb, n_dir, n_ant, subtract,
                       keep_model);
  else
    hipLaunchKernelGGL(CorrectSubtractKernel<4>, dim3(blocks), dim3(256), 0, stream, data, model,
                       sol, d_ant1, d_ant2, d_cb_start, n_t, n_bl, n_ch, n_cb, n_dir, n_ant, subtract,
                       keep_model);
  DDH_HIP_CHECK(hipGetLastError());
}

}  // namespace
}  // namespace ddh

using namespace ddh;

extern "C" {

int ddh_correct_and_subtract_device(int32_t n_times, int32_t n_baselines, int32_t n_channels,
                                    int32_t n_directions, int32_t n_antennas,
                                    int32_t n_channel_blocks, int32_t n_pol,
                                    const int32_t* antenna1, const int32_t* antenna2,
                                    const int32_t* channel_block_start, const double* solutions,
                                    float* data, float* model, int32_t subtract,
                                    int32_t keep_model, void* stream) {
  return Guard([&] {
    DDH_REQUIRE(antenna1 && antenna2 && channel_block_start && solutions && data && model, -1,
                "NULL argument");
    CheckShape(n_times, n_baselines, n_channels, n_directions, n_antennas, n_channel_blocks,
               n_pol);
    Launch(n_times, n_baselines, n_channels, n_channel_blocks, n_directions, n_antennas, n_pol,
           antenna1, antenna2,
           channel_block_start, solutions, data, model, subtract, keep_model,
           reinterpret_cast<hipStream_t>(stream));
  });
}

int ddh_correct_and_subtract(int32_t n_times, int32_t n_baselines, int32_t n_channels,
                             int32_t n_directions, int32_t n_antennas, int32_t n_channel_blocks,
                             int32_t n_pol, const int32_t* antenna1, const int32_t* antenna2,
                             const int32_t* channel_block_start, const double* solutions,
                             float* data, float* model, int32_t subtract, int32_t keep_model,
                             double* kernel_ms) {
  return Guard([&] {
    DDH_REQUIRE(antenna1 && antenna2 && channel_block_start && solutions && data && model, -1,
                "NULL argument");
    CheckShape(n_times, n_baselines, n_channels, n_directions, n_antennas, n_channel_blocks,
               n_pol);
    for (int cb = 0; cb < n_channel_blocks; ++cb)
      DDH_REQUIRE(channel_block_start[cb] < channel_block_start[cb + 1], -1,
                  "channel blocks must be ordered and hold >= 1 channel");
    DDH_REQUIRE(channel_block_start[0] == 0 && channel_block_start[n_channel_blocks] == n_channels,
                -1, "channel blocks must cover the channels");
    for (int b = 0; b < n_baselines; ++b)
      DDH_REQUIRE(antenna1[b] >= 0 && antenna1[b] < n_antennas && antenna2[b] >= 0 &&
                      antenna2[b] < n_antennas,
                  -1, "antenna index out of range");
    const size_t n_vis = size_t(n_times) * n_baselines * n_channels;
    const size_t n_sol = size_t(n_channel_blocks) * n_antennas * n_directions * n_pol;
    Scratch<int32_t> a1, a2, cbs;
    Scratch<double> sol;
    Scratch<float> d, m;
    hipStream_t stream = nullptr;
    DDH_HIP_CHECK(hipMemcpy(a1.Get(n_baselines), antenna1, n_baselines * 4, hipMemcpyHostToDevice));
    DDH_HIP_CHECK(hipMemcpy(a2.Get(n_baselines), antenna2, n_baselines * 4, hipMemcpyHostToDevice));
    DDH_HIP_CHECK(hipMemcpy(cbs.Get(n_channel_blocks + 1), channel_block_start,
                            (n_channel_blocks + 1) * 4, hipMemcpyHostToDevice));
    DDH_HIP_CHECK(hipMemcpy(sol.Get(n_sol * 2), solutions, n_sol * 16, hipMemcpyHostToDevice));
    DDH_HIP_CHECK(hipMemcpy(d.Get(n_vis * 8), data, n_vis * 32, hipMemcpyHostToDevice));
    DDH_HIP_CHECK(hipMemcpy(m.Get(n_vis * 8 * n_directions), model, n_vis * 32 * n_directions,
                            hipMemcpyHostToDevice));
    hipEvent_t e0, e1;
    DDH_HIP_CHECK(hipEventCreate(&e0));
    DDH_HIP_CHECK(hipEventCreate(&e1));
    DDH_HIP_CHECK(hipEventRecord(e0, stream));
    Launch(n_times, n_baselines, n_channels, n_channel_blocks, n_directions, n_antennas, n_pol,
           a1.ptr, a2.ptr,
           cbs.ptr, sol.ptr, d.ptr, m.ptr, subtract, keep_model, stream);
    DDH_HIP_CHECK(hipEventRecord(e1, stream));
    DDH_HIP_CHECK(hipEventSynchronize(e1));
    float ms = 0.0f;
    DDH_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (kernel_ms) *kernel_ms = ms;
    if (subtract) DDH_HIP_CHECK(hipMemcpy(data, d.ptr, n_vis * 32, hipMemcpyDeviceToHost));
    if (keep_model)
      DDH_HIP_CHECK(hipMemcpy(model, m.ptr, n_vis * 32 * n_directions, hipMemcpyDeviceToHost));
  });
}

}  // extern "C"
