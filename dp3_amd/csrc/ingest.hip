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
// A workgroup takes a tile of 16 baselines x 32 channels of one time step:
// it reads runs of each baseline's row and, through LDS, stores runs of each
// channel's row. HBM bound: per cross-correlation visibility 32 B data, 4 B
// flags, 16 B weights and n_dir x 32 B models read (auto-correlations are
// not read); 16 B (Duo) or 32 B (Full) written for the data and n_dir times
// 16 B / 32 B (complex64) or 8 B / 16 B (FP16) for the models.
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

constexpr int kIngestThreads = 512;
// (block, direction) entries of the FP16 scale pass held in LDS
constexpr int kIngestMaxScale = 2048;
constexpr int kIngestMaxDirections = 16;

__device__ __forceinline__ bool Finite4(float4_v x) {
  return isfinite(x.x) && isfinite(x.y) && isfinite(x.z) && isfinite(x.w);
}

__device__ __forceinline__ uint32_t HalfBits(float re, float im) {
  half2_v h;
  h.x = static_cast<_Float16>(re);
  h.y = static_cast<_Float16>(im);
  return __builtin_bit_cast(uint32_t, h);
}

// Visibility x (correlations 0..3) weighted by w (per correlation) and
// zeroed when flagged. For single-correlation problems (uni) correlations 0
// and 3 then both hold 0.5f * (XX + YY) (SolveData.cc:29-31), which the
// problem keeps as a Duo pair.
__device__ __forceinline__ void Weigh(const float4_v (&x)[2], float4_v w, bool flagged, bool uni,
                                      float4_v (&out)[2]) {
  const float4_v z = {0.0f, 0.0f, 0.0f, 0.0f};
  out[0] = flagged ? z : x[0] * float4_v{w.x, w.x, w.y, w.y};
  out[1] = flagged ? z : x[1] * float4_v{w.z, w.z, w.w, w.w};
  if (uni) {
    const float re = 0.5f * (out[0].x + out[1].z), im = 0.5f * (out[0].y + out[1].w);
    out[0].x = out[1].z = re;
    out[0].y = out[1].w = im;
  }
}

// Largest finite |re|, |im| over the correlations a plane keeps.
template <int NC>
__device__ __forceinline__ float MaxAbs(const float4_v (&x)[2], float m) {
  const float v[8] = {x[0].x, x[0].y, x[0].z, x[0].w, x[1].x, x[1].y, x[1].z, x[1].w};
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    if (NC == 2 && i >= 2 && i < 6) continue;  // Duo keeps correlations 0 and 3
    const float a = fabsf(v[i]);
    if (isfinite(a) && a > m) m = a;
  }
  return m;
}

// Packed form of one visibility in a plane of the problem.
template <int NC, bool F16>
struct Packed;
template <>
struct Packed<2, false> {
  float4_v v;  // correlations 0 and 3
};
template <>
struct Packed<4, false> {
  float4_v v[2];
};
template <>
struct Packed<2, true> {
  uint2_v v;
};
template <>
struct Packed<4, true> {
  uint4_v v;
};

template <int NC, bool F16>
__device__ __forceinline__ Packed<NC, F16> Pack(const float4_v (&x)[2], float inv) {
  Packed<NC, F16> p;
  if constexpr (!F16 && NC == 2) {
    p.v = float4_v{x[0].x, x[0].y, x[1].z, x[1].w};
  } else if constexpr (!F16) {
    p.v[0] = x[0];
    p.v[1] = x[1];
  } else if constexpr (NC == 2) {
    p.v = uint2_v{HalfBits(x[0].x * inv, x[0].y * inv), HalfBits(x[1].z * inv, x[1].w * inv)};
  } else {
    p.v = uint4_v{HalfBits(x[0].x * inv, x[0].y * inv), HalfBits(x[0].z * inv, x[0].w * inv),
                  HalfBits(x[1].x * inv, x[1].y * inv), HalfBits(x[1].z * inv, x[1].w * inv)};
  }
  return p;
}

// Tile of one workgroup: kTileBl consecutive buffer baselines (time step t)
// by kTileCh consecutive channels, one visibility per thread.
constexpr int kTileCh = 32;
constexpr int kTileBl = kIngestThreads / kTileCh;

// MAX: per-(block, direction) maxima of the weighted models only.
// Otherwise: writes the data plane and model planes of each block.
//
// Read phase: thread i takes baseline i / kTileCh and channel i % kTileCh of
// the tile, so each wave reads runs of kTileCh x 32 B of consecutive
// baselines' rows. NDB >= n_dir bounds the directions held in registers: every
// load of a visibility is issued before the first is used. Write phase, per
// plane: the packed values pass through LDS, and thread j stores channel
// j / kTileBl, baseline j % kTileBl, so the stores of a channel row are
// contiguous. The grid is a multiple of the channel-group count, so every
// thread keeps one channel (one block) for the whole launch; MAX keeps its
// maxima in registers and merges them through LDS (s_max, n_cb * n_dir words)
// into max_bits once per workgroup.
template <int NC, bool F16, bool MAX, int NDB>
__global__ void __launch_bounds__(kIngestThreads) AssignWeighKernel(IngestArgs g) {
  using P = Packed<NC, F16>;
  using PD = Packed<NC, false>;  // data planes are complex64 in every storage mode
  __shared__ P s_model[2][kIngestThreads];
  __shared__ PD s_data[kIngestThreads];
  __shared__ uint32_t s_max[kIngestMaxScale];
  const int n_max = MAX ? g.n_cb * g.n_dir : 0;
  for (int e = threadIdx.x; e < n_max; e += kIngestThreads) s_max[e] = 0u;

  const int n_cg = (g.n_ch + kTileCh - 1) / kTileCh;
  const int n_blg = (g.n_bl_all + kTileBl - 1) / kTileBl;
  const int cg = int(blockIdx.x % n_cg);
  // read-phase coordinates
  const int r_bl = threadIdx.x / kTileCh;
  const int r_ch = cg * kTileCh + int(threadIdx.x % kTileCh);
  // write-phase coordinates
  const int w_ch_l = threadIdx.x / kTileBl;
  const int w_bl = threadIdx.x % kTileBl;
  const int w_ch = cg * kTileCh + w_ch_l;
  const bool w_ch_ok = w_ch < g.n_ch;
  const int w_cb = w_ch_ok ? g.ch_cb[w_ch] : 0;
  const int w_c0 = g.ch_start[w_cb];
  const int w_nch = g.ch_start[w_cb + 1] - w_c0;
  const CbInfo w_info = g.cbinfo[w_cb];
  const size_t w_plane = size_t(g.n_t) * w_nch * g.bl_stride;
  const int r_cb = r_ch < g.n_ch ? g.ch_cb[r_ch] : 0;

  const int64_t n_units = int64_t(g.n_t) * n_blg * n_cg;
  const int64_t n_vis_all = int64_t(g.n_t) * g.n_bl_all * g.n_ch;
  const float4_v* data = reinterpret_cast<const float4_v*>(g.data);
  const float4_v* model = reinterpret_cast<const float4_v*>(g.model);
  const float4_v zero = {0.0f, 0.0f, 0.0f, 0.0f};
  float mx[NDB];
#pragma unroll
  for (int d = 0; d < NDB; ++d) mx[d] = 0.0f;
  if constexpr (MAX) __syncthreads();

  for (int64_t u = blockIdx.x; u < n_units; u += gridDim.x) {
    const int64_t tg = u / n_cg;
    const int t = int(tg / n_blg);
    const int bl0 = int(tg % n_blg) * kTileBl;
    // ---- read phase
    const int bl = bl0 + r_bl;
    const bool r_ok = bl < g.n_bl_all && r_ch < g.n_ch && g.cross[bl] >= 0;
    float4_v x[2] = {zero, zero};
    float4_v w = {1.0f, 1.0f, 1.0f, 1.0f};
    float4_v m[NDB][2];
    bool flagged = false;
    if (r_ok) {
      const int64_t v = (int64_t(t) * g.n_bl_all + bl) * g.n_ch + r_ch;
      x[0] = data[2 * v];
      x[1] = data[2 * v + 1];
      const uint32_t fl = g.flags ? reinterpret_cast<const uint32_t*>(g.flags)[v] : 0u;
      if (g.weights) w = reinterpret_cast<const float4_v*>(g.weights)[v];
#pragma unroll
      for (int d = 0; d < NDB; ++d) {
        const size_t mv = size_t(d) * n_vis_all + v;
        m[d][0] = d < g.n_dir ? model[2 * mv] : zero;
        m[d][1] = d < g.n_dir ? model[2 * mv + 1] : zero;
      }
      flagged = fl != 0u || !Finite4(x[0]) || !Finite4(x[1]);
#pragma unroll
      for (int d = 0; d < NDB; ++d) flagged = flagged || !Finite4(m[d][0]) || !Finite4(m[d][1]);
      if (g.weights && !g.linear) w = float4_v{sqrtf(w.x), sqrtf(w.y), sqrtf(w.z), sqrtf(w.w)};
    } else {
#pragma unroll
      for (int d = 0; d < NDB; ++d) m[d][0] = m[d][1] = zero;
    }
    float4_v o[2];
    if constexpr (MAX) {
      if (r_ok) {
#pragma unroll
        for (int d = 0; d < NDB; ++d) {
          if (d < g.n_dir) {
            Weigh(m[d], w, flagged, g.uni != 0, o);
            mx[d] = MaxAbs<NC>(o, mx[d]);
          }
        }
      }
    } else {
      // ---- write phase through LDS (every thread reaches every barrier)
      const int s_idx = (threadIdx.x % kTileCh) * kTileBl + r_bl;  // [channel][baseline]
      const int wbl = bl0 + w_bl;
      const int w_cross = (wbl < g.n_bl_all && w_ch_ok) ? g.cross[wbl] : -1;
      const size_t out = (size_t(t) * w_nch + (w_ch - w_c0)) * g.bl_stride + w_cross;
      Weigh(x, w, flagged, g.uni != 0, o);
      s_data[s_idx] = Pack<NC, false>(o, 1.0f);
      __syncthreads();
      if (w_cross >= 0)
        reinterpret_cast<PD*>(g.out_data)[size_t(w_info.data_off) + out] = s_data[threadIdx.x];
#pragma unroll
      for (int d = 0; d < NDB; ++d) {
        if (d < g.n_dir) {
          const float inv = F16 ? 1.0f / g.scale[r_cb * g.n_dir + d] : 1.0f;
          Weigh(m[d], w, flagged, g.uni != 0, o);
          s_model[d & 1][s_idx] = Pack<NC, F16>(o, inv);
          __syncthreads();
          if (w_cross >= 0)
            reinterpret_cast<P*>(g.out_model)[size_t(w_info.model_off) + d * w_plane + out] =
                s_model[d & 1][threadIdx.x];
        }
      }
      __syncthreads();  // the next unit reuses the LDS tiles
    }
  }
  if constexpr (MAX) {
    if (r_ch < g.n_ch) {
#pragma unroll
      for (int d = 0; d < NDB; ++d)
        if (d < g.n_dir && mx[d] > 0.0f)
          atomicMax(&s_max[r_cb * g.n_dir + d], __float_as_uint(mx[d]));
    }
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
  const int64_t n_cg = (g.n_ch + kTileCh - 1) / kTileCh;
  const int64_t n_units = int64_t(g.n_t) * ((g.n_bl_all + kTileBl - 1) / kTileBl) * n_cg;
  // grid-stride over tiles: about four workgroups per compute unit keep the
  // loads in flight; a multiple of the channel-group count fixes each
  // thread's channel
  const int64_t cap = std::max<int64_t>(1, 1024 / n_cg) * n_cg;
  const int64_t blocks = std::min<int64_t>(n_units, cap);
  const IngestFn kernel = g.n_dir <= 4   ? PickIngest<MAX, 4>(n_corr, f16)
                          : g.n_dir <= 8 ? PickIngest<MAX, 8>(n_corr, f16)
                                         : PickIngest<MAX, kIngestMaxDirections>(n_corr, f16);
  hipLaunchKernelGGL(kernel, dim3(unsigned(blocks)), dim3(kIngestThreads), 0, stream, g);
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
    DDH_REQUIRE(n <= kIngestMaxScale, -1,
                "FP16 ingest supports at most 2048 channel blocks x directions");
    hipLaunchKernelGGL(FillU32Kernel, dim3((n + 255) / 256), dim3(256), 0, stream, g.max_bits, n,
                       0u);
    DDH_HIP_CHECK(hipGetLastError());
    LaunchAssignWeighPass<true>(g, n_corr, f16, stream);
    LaunchScaleFromMax(g.max_bits, const_cast<float*>(g.scale), n, stream);
  }
  LaunchAssignWeighPass<false>(g, n_corr, f16, stream);
}

}  // namespace ddh
