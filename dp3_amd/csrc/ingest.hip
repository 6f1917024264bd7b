// Weighting and packing of an interval's raw buffers into a device problem:
// AssignAndWeight (ddecal/gain_solvers/SolverTools.cc:81-167, with Weigh
// :48-79) followed by the regular SolveData constructor
// (ddecal/gain_solvers/SolveData.cc:77-188), in one pass over the buffers.
//
// Input, in DPBuffer order over the n_t time steps of the interval:
//   data    [n_t][n_bl][n_ch][4] complex64
//   flags   [n_t][n_bl][n_ch][4] bytes (or none)
//   weights [n_t][n_bl][n_ch][4] float (or none: unit weights)
//   model   [n_dir][n_t][n_bl][n_ch][4] complex64
// where n_bl counts every baseline of the buffer, auto-correlations included.
//
// Per visibility (baseline, channel) one flag covers every correlation and
// direction: a set input flag or a non-finite value in any correlation of the
// data or of any model (SolverTools.cc:115-128). Data and models are then
// multiplied by sqrt(weight) of their correlation (by the weight itself in
// linear weighting mode, :131-141), and flagged visibilities are zeroed
// (:150, :162). Packing drops auto-correlations (SolveData.cc:96-100, 137),
// keeps correlations 0 and 3 for Duo (:25-27) or all four for Full, and
// writes the device layout [time][channel-in-block][baseline_stride] of each
// block's data plane and model planes.
//
// FP16 problems take a per-(block, direction) scale: the largest |re|, |im|
// of the finite weighted model values of the block maps to 2^14 (the same
// rule as uploads of packed blocks), found by a first pass over the buffers.
//
// Threads walk the raw visibilities in buffer order, so reads are contiguous;
// each write lands in the row of its channel. HBM bound: per visibility 32 B data, 4 B flags, 16 B
// weights and n_dir x 32 B models read; 16 B (Duo) or 32 B (Full) written for
// the data and n_dir times 16 B / 32 B (complex64) or 8 B / 16 B (FP16) for
// the models, auto-correlations excepted.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "common.h"
#include "solver_kernels.h"

namespace ddh {
namespace {

typedef float float4_v __attribute__((ext_vector_type(4)));
typedef _Float16 half2_v __attribute__((ext_vector_type(2)));
typedef unsigned int uint2_v __attribute__((ext_vector_type(2)));
typedef unsigned int uint4_v __attribute__((ext_vector_type(4)));

constexpr int kIngestThreads = 256;
constexpr int kIngestMaxDirections = 16;
// LDS of the FP16 scale pass: one word per (block, direction)
constexpr size_t kIngestMaxLds = 32 * 1024;

struct Vis4 {
  float4_v lo;  // correlations 0, 1
  float4_v hi;  // correlations 2, 3
};

__device__ __forceinline__ Vis4 LoadVis4(const float4_v* p, size_t v) {
  return {p[2 * v], p[2 * v + 1]};
}

__device__ __forceinline__ bool Finite4(float4_v x) {
  return isfinite(x.x) && isfinite(x.y) && isfinite(x.z) && isfinite(x.w);
}

__device__ __forceinline__ uint32_t HalfBits(float re, float im) {
  half2_v h;
  h.x = static_cast<_Float16>(re);
  h.y = static_cast<_Float16>(im);
  return __builtin_bit_cast(uint32_t, h);
}

// Weighted, flag-zeroed visibility: correlation c scaled by w[c].
__device__ __forceinline__ Vis4 Weigh(Vis4 x, float4_v w, bool flagged) {
  const float4_v wl = {w.x, w.x, w.y, w.y};
  const float4_v wh = {w.z, w.z, w.w, w.w};
  const float4_v zero = {0.0f, 0.0f, 0.0f, 0.0f};
  return {flagged ? zero : x.lo * wl, flagged ? zero : x.hi * wh};
}

// Largest finite |re|, |im| over the correlations a plane keeps.
template <int NC>
__device__ __forceinline__ float MaxAbs(Vis4 x, float m) {
  float v[8] = {x.lo.x, x.lo.y, x.lo.z, x.lo.w, x.hi.x, x.hi.y, x.hi.z, x.hi.w};
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    if (NC == 2 && i >= 2 && i < 6) continue;  // Duo keeps correlations 0 and 3
    const float a = fabsf(v[i]);
    if (isfinite(a) && a > m) m = a;
  }
  return m;
}

// Stores one packed visibility at element `idx` of a plane.
template <int NC, bool F16>
__device__ __forceinline__ void StoreVis(void* base, size_t idx, Vis4 x, float inv) {
  if constexpr (!F16) {
    float4_v* o = reinterpret_cast<float4_v*>(base);
    if constexpr (NC == 2) {
      o[idx] = float4_v{x.lo.x, x.lo.y, x.hi.z, x.hi.w};
    } else {
      o[2 * idx] = x.lo;
      o[2 * idx + 1] = x.hi;
    }
  } else {
    if constexpr (NC == 2) {
      reinterpret_cast<uint2_v*>(base)[idx] =
          uint2_v{HalfBits(x.lo.x * inv, x.lo.y * inv), HalfBits(x.hi.z * inv, x.hi.w * inv)};
    } else {
      reinterpret_cast<uint4_v*>(base)[idx] =
          uint4_v{HalfBits(x.lo.x * inv, x.lo.y * inv), HalfBits(x.lo.z * inv, x.lo.w * inv),
                  HalfBits(x.hi.x * inv, x.hi.y * inv), HalfBits(x.hi.z * inv, x.hi.w * inv)};
    }
  }
}

// MAX: per-(block, direction) maxima of the weighted models only.
// Otherwise: writes the data plane and model planes of each block.
// Threads walk the visibilities in DPBuffer order (contiguous reads; each
// write lands in the row of its channel). NDB >= n_dir bounds the directions
// held in registers: every load of a visibility is issued before the first is
// used. MAX keeps per-workgroup maxima in LDS (s_max, n_cb * n_dir words),
// merged into max_bits with one atomic per entry and workgroup.
template <int NC, bool F16, bool MAX, int NDB>
__global__ void __launch_bounds__(kIngestThreads) AssignWeighKernel(IngestArgs g) {
  extern __shared__ uint32_t s_max[];
  const int n_max = MAX ? g.n_cb * g.n_dir : 0;
  for (int e = threadIdx.x; e < n_max; e += kIngestThreads) s_max[e] = 0u;
  if constexpr (MAX) __syncthreads();
  const float4_v* data = reinterpret_cast<const float4_v*>(g.data);
  const float4_v* model = reinterpret_cast<const float4_v*>(g.model);
  const int64_t n_vis_all = int64_t(g.n_t) * g.n_bl_all * g.n_ch;
  const float4_v zero = {0.0f, 0.0f, 0.0f, 0.0f};
  for (int64_t v = int64_t(blockIdx.x) * kIngestThreads + threadIdx.x; v < n_vis_all;
       v += int64_t(gridDim.x) * kIngestThreads) {
    const int ch = int(v % g.n_ch);
    const int64_t tb = v / g.n_ch;
    const int bl = int(tb % g.n_bl_all);
    const int t = int(tb / g.n_bl_all);
    const int cross = g.cross[bl];
    if (cross < 0) continue;  // auto-correlation: not solve data
    const Vis4 x = LoadVis4(data, size_t(v));
    const uint32_t fl = g.flags ? reinterpret_cast<const uint32_t*>(g.flags)[v] : 0u;
    float4_v w = {1.0f, 1.0f, 1.0f, 1.0f};
    if (g.weights) w = reinterpret_cast<const float4_v*>(g.weights)[v];
    Vis4 m[NDB];
#pragma unroll
    for (int d = 0; d < NDB; ++d)
      m[d] = d < g.n_dir ? LoadVis4(model, size_t(d) * n_vis_all + v) : Vis4{zero, zero};
    bool flagged = fl != 0u || !Finite4(x.lo) || !Finite4(x.hi);
#pragma unroll
    for (int d = 0; d < NDB; ++d) flagged = flagged || !Finite4(m[d].lo) || !Finite4(m[d].hi);
    if (g.weights && !g.linear) w = float4_v{sqrtf(w.x), sqrtf(w.y), sqrtf(w.z), sqrtf(w.w)};
    const int cb = g.ch_cb[ch];
    if constexpr (MAX) {
#pragma unroll
      for (int d = 0; d < NDB; ++d) {
        if (d < g.n_dir) {
          const float mv = MaxAbs<NC>(Weigh(m[d], w, flagged), 0.0f);
          if (mv > 0.0f) atomicMax(&s_max[cb * g.n_dir + d], __float_as_uint(mv));
        }
      }
    } else {
      const int c0 = g.ch_start[cb];
      const int n_ch_cb = g.ch_start[cb + 1] - c0;
      const CbInfo info = g.cbinfo[cb];
      const size_t out = (size_t(t) * n_ch_cb + (ch - c0)) * g.bl_stride + cross;
      const size_t plane = size_t(g.n_t) * n_ch_cb * g.bl_stride;
      // data planes are complex64 in every storage mode
      StoreVis<NC, false>(g.out_data, size_t(info.data_off) + out, Weigh(x, w, flagged), 1.0f);
#pragma unroll
      for (int d = 0; d < NDB; ++d) {
        if (d < g.n_dir) {
          const float inv = F16 ? 1.0f / g.scale[cb * g.n_dir + d] : 1.0f;
          StoreVis<NC, F16>(g.out_model, size_t(info.model_off) + d * plane + out,
                            Weigh(m[d], w, flagged), inv);
        }
      }
    }
  }
  if constexpr (MAX) {
    __syncthreads();
    for (int e = threadIdx.x; e < n_max; e += kIngestThreads)
      if (s_max[e] != 0u) atomicMax(&g.max_bits[e], s_max[e]);
  }
}

__global__ void FillU32Kernel(uint32_t* p, int n, uint32_t value) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = value;
}

using IngestFn = void (*)(IngestArgs);

template <bool MAX, int NDB>
IngestFn PickIngest(int n_corr, bool f16) {
  if (n_corr == 2)
    return f16 ? AssignWeighKernel<2, true, MAX, NDB> : AssignWeighKernel<2, false, MAX, NDB>;
  return f16 ? AssignWeighKernel<4, true, MAX, NDB> : AssignWeighKernel<4, false, MAX, NDB>;
}

template <bool MAX>
void LaunchAssignWeighPass(const IngestArgs& g, int n_corr, bool f16, hipStream_t stream) {
  const int64_t n_vis = int64_t(g.n_t) * g.n_bl_all * g.n_ch;
  // grid-stride: about eight workgroups per compute unit keep the loads in
  // flight; fewer workgroups also mean fewer merges of the LDS maxima
  const int64_t blocks = std::min<int64_t>((n_vis + kIngestThreads - 1) / kIngestThreads, 2048);
  const size_t lds = MAX ? size_t(g.n_cb) * g.n_dir * sizeof(uint32_t) : 0;
  const IngestFn kernel = g.n_dir <= 4   ? PickIngest<MAX, 4>(n_corr, f16)
                          : g.n_dir <= 8 ? PickIngest<MAX, 8>(n_corr, f16)
                                         : PickIngest<MAX, kIngestMaxDirections>(n_corr, f16);
  hipLaunchKernelGGL(kernel, dim3(unsigned(blocks)), dim3(kIngestThreads), lds, stream, g);
  DDH_HIP_CHECK(hipGetLastError());
}

}  // namespace

void LaunchAssignAndWeigh(const IngestArgs& g, int n_corr, bool f16, hipStream_t stream) {
  DDH_REQUIRE(g.n_dir >= 1 && g.n_dir <= kIngestMaxDirections, -1,
              "ingest supports 1 to 16 directions");
  DDH_REQUIRE(n_corr == 2 || n_corr == 4, -1, "n_correlations must be 2 or 4");
  if (g.n_t == 0 || g.n_bl_all == 0 || g.n_ch == 0) return;
  if (f16) {
    const int n = g.n_cb * g.n_dir;
    DDH_REQUIRE(size_t(n) * sizeof(uint32_t) <= kIngestMaxLds, -1,
                "too many channel blocks x directions for the FP16 scale pass");
    hipLaunchKernelGGL(FillU32Kernel, dim3((n + 255) / 256), dim3(256), 0, stream, g.max_bits, n,
                       0u);
    DDH_HIP_CHECK(hipGetLastError());
    LaunchAssignWeighPass<true>(g, n_corr, f16, stream);
    LaunchScaleFromMax(g.max_bits, const_cast<float*>(g.scale), n, stream);
  }
  LaunchAssignWeighPass<false>(g, n_corr, f16, stream);
}

}  // namespace ddh
