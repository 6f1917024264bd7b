// Device kernels of the iterative diagonal solver (see solver_kernels.h for
// the decomposition and the reference lines each kernel restates).
#include "solver_kernels.h"

#include <hip/hip_fp16.h>

#include <algorithm>
#include <array>
#include <limits>
#include <utility>

namespace ddh {
namespace {

constexpr int kWave = 64;
constexpr int kMaxDirections = 16;
constexpr int kCorrelateMaxThreads = 256;
constexpr int kApplyThreads = 2 * kWave;  // Gram apply: one wave per Duo correlation
constexpr int kBlockThreads = 1024;

// Correlation k of the visibility matrix maps to gain polarizations (i, j):
// V_k = g_a,i * M_k * conj(g_b,j). Duo keeps the diagonal: k -> (k, k).
template <int NC>
__device__ __forceinline__ int PolI(int k) { return NC == 2 ? k : (k >> 1); }
template <int NC>
__device__ __forceinline__ int PolJ(int k) { return NC == 2 ? k : (k & 1); }

typedef _Float16 half2_v __attribute__((ext_vector_type(2)));
typedef float float4_v __attribute__((ext_vector_type(4)));
typedef unsigned int uint2_v __attribute__((ext_vector_type(2)));
typedef unsigned int uint4_v __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float2 HalfBitsToFloat2(uint32_t bits) {
  const half2_v h = __builtin_bit_cast(half2_v, bits);
  return make_float2(float(h.x), float(h.y));
}

// Round-to-nearest-even conversion of a complex value to two FP16 halves.
__device__ __forceinline__ uint32_t Float2ToHalfBits(float re, float im) {
  half2_v h;
  h.x = static_cast<_Float16>(re);
  h.y = static_cast<_Float16>(im);
  return __builtin_bit_cast(uint32_t, h);
}

__device__ __forceinline__ double2 FixedSolution(const double2* sol, size_t pair, int pol,
                                                 double mean);

// Loads the NC correlations of visibility `idx` from a plane.
template <int NC, bool F16>
__device__ __forceinline__ void LoadVis(const void* base, size_t idx, float2 (&v)[NC]) {
  if constexpr (!F16) {
    // A Full visibility is two 16-byte halves, so each half's load covers
    // every other 16 bytes of the wave's lines: the first is a plain load,
    // which keeps the lines in L2 for the second (non-temporal; as the first
    // it would let them go and the second would fetch them again).
    const float4_v* p = reinterpret_cast<const float4_v*>(base) + idx * (NC / 2);
#pragma unroll
    for (int h = 0; h < NC / 2; ++h) {
      const float4_v x = (NC == 4 && h == 0) ? p[h] : __builtin_nontemporal_load(p + h);
      v[2 * h] = make_float2(x.x, x.y);
      v[2 * h + 1] = make_float2(x.z, x.w);
    }
  } else if constexpr (NC == 2) {
    const uint2_v u = __builtin_nontemporal_load(reinterpret_cast<const uint2_v*>(base) + idx);
    v[0] = HalfBitsToFloat2(u.x);
    v[1] = HalfBitsToFloat2(u.y);
  } else {
    const uint4_v u = __builtin_nontemporal_load(reinterpret_cast<const uint4_v*>(base) + idx);
    v[0] = HalfBitsToFloat2(u.x);
    v[1] = HalfBitsToFloat2(u.y);
    v[2] = HalfBitsToFloat2(u.z);
    v[3] = HalfBitsToFloat2(u.w);
  }
}

// One visibility of a plane as loaded: complex64 planes as float4 vectors,
// FP16 planes as packed half2 words, converted to float when used (fewer
// registers per row in flight).
template <int NC, bool F16>
struct RawVis {
  float4_v v[NC / 2];
};
template <>
struct RawVis<2, true> {
  uint2_v v;
};
template <>
struct RawVis<4, true> {
  uint4_v v;
};

template <int NC, bool F16>
constexpr uint32_t VisBytes() {
  return F16 ? uint32_t(NC) * 4u : uint32_t(NC) * 8u;
}

// Plane reads go through buffer resources (base, size) in scalar registers:
// one 32-bit offset register per lane serves every plane, and a read at or
// past the size returns zeros, so rows beyond the end of a segment need no
// branch. 0x00020000 is the gfx9 resource word 3 (32-bit float data format);
// cache policy 2 marks the reads non-temporal.
constexpr int kResourceWord3 = 0x00020000;
constexpr int kNonTemporal = 2;

__device__ __forceinline__ __amdgpu_buffer_rsrc_t PlaneResource(const void* base, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), short(0), int(bytes),
                                           kResourceWord3);
}

template <int NC, bool F16>
__device__ __forceinline__ void LoadRaw(__amdgpu_buffer_rsrc_t rsrc, uint32_t byte_off,
                                        RawVis<NC, F16>& r) {
  if constexpr (!F16) {
#pragma unroll
    for (int h = 0; h < NC / 2; ++h)
      r.v[h] = __builtin_bit_cast(
          float4_v, __builtin_amdgcn_raw_buffer_load_b128(rsrc, byte_off + 16u * h, 0, kNonTemporal));
  } else if constexpr (NC == 2) {
    r.v = __builtin_bit_cast(uint2_v,
                             __builtin_amdgcn_raw_buffer_load_b64(rsrc, byte_off, 0, kNonTemporal));
  } else {
    r.v = __builtin_bit_cast(uint4_v,
                             __builtin_amdgcn_raw_buffer_load_b128(rsrc, byte_off, 0, kNonTemporal));
  }
}

// Correlation k of a loaded visibility as float2.
template <int NC, bool F16>
__device__ __forceinline__ float2 RawGet(const RawVis<NC, F16>& r, int k) {
  if constexpr (!F16) {
    const float4_v x = r.v[k / 2];
    return (k & 1) ? make_float2(x.z, x.w) : make_float2(x.x, x.y);
  } else if constexpr (NC == 2) {
    return HalfBitsToFloat2(k ? r.v.y : r.v.x);
  } else {
    return HalfBitsToFloat2(r.v[k]);
  }
}

// ---------------------------------------------------------------------------
// correlate: grid (bl tiles, segments, channel blocks), block = waves x 64.
// Lane l of every wave owns baseline tile*64+l; the waves of a block split the
// (time, channel) rows of the segment and are combined through LDS in a fixed
// order, so results do not depend on scheduling. Each wave keeps ROWS rows of
// loads in flight (a ring of row registers refilled as each row is used).
// ---------------------------------------------------------------------------
template <int ND, int NC, bool F16, int ROWS>
__device__ __forceinline__ void CorrelateTile(const IterArgs& a, const int tile, const int seg,
                                              const int cb) {
  constexpr uint32_t kDataBytes = VisBytes<NC, false>();
  constexpr uint32_t kModelBytes = VisBytes<NC, F16>();
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = __builtin_amdgcn_readfirstlane(int(threadIdx.x) / kWave);
  const int n_waves = blockDim.x / kWave;
  const CbInfo info = LoadConst(a.cbinfo + cb);
  const int n_ch = info.n_ch;
  const size_t plane = size_t(a.n_t) * n_ch * a.bl_stride;
  const int bl = tile * kWave + lane;
  const char* dbase = reinterpret_cast<const char*>(a.data) + size_t(info.data_off) * kDataBytes;
  const char* mbase = reinterpret_cast<const char*>(a.model) + size_t(info.model_off) * kModelBytes;
  const size_t mplane = plane * kModelBytes;

  extern __shared__ float smem[];
  float2* gg_s = reinterpret_cast<float2*>(smem);  // [ND][64][NC]
  float* red_s = smem + 2 * ND * NC * kWave;       // [3*ND*NC][64]

  float ur[ND][NC], ui[ND][NC], sm[ND][NC];
#pragma unroll
  for (int d = 0; d < ND; ++d)
#pragma unroll
    for (int k = 0; k < NC; ++k) {
      ur[d][k] = 0.0f;
      ui[d][k] = 0.0f;
      sm[d][k] = 0.0f;
    }

  const int r0 = AsConst(a.seg_t0)[seg] * n_ch;
  const int r1 = AsConst(a.seg_t0)[seg + 1] * n_ch;
  // Resources end at the segment's last row: a row past it reads as zeros,
  // which adds exactly nothing.
  const uint32_t row_vis = uint32_t(r1) * uint32_t(a.bl_stride);
  const __amdgpu_buffer_rsrc_t drsrc = PlaneResource(dbase, row_vis * kDataBytes);
  __amdgpu_buffer_rsrc_t mrsrc[ND];
#pragma unroll
  for (int d = 0; d < ND; ++d) mrsrc[d] = PlaneResource(mbase + d * mplane, row_vis * kModelBytes);
  struct Row {
    RawVis<NC, false> d;
    RawVis<NC, F16> m[ND];
  };
  auto load_row = [&](int row, Row& r) {
    const uint32_t e = uint32_t(row) * uint32_t(a.bl_stride) + uint32_t(bl);
    LoadRaw<NC, false>(drsrc, e * kDataBytes, r.d);
#pragma unroll
    for (int d = 0; d < ND; ++d) LoadRaw<NC, F16>(mrsrc[d], e * kModelBytes, r.m[d]);
  };
  auto accumulate = [&](const Row& row) {
    float2 r[NC];
#pragma unroll
    for (int k = 0; k < NC; ++k) r[k] = RawGet<NC, false>(row.d, k);
    // The gain products are read from LDS for every row: an index the
    // compiler cannot see through keeps it from hoisting all ND * NC of them
    // into registers for the whole loop, which would cost occupancy.
    int gi = lane * NC;
    asm volatile("" : "+v"(gi));
    // Full residual, directions subtracted in order (IterativeDiagonalSolver.cc:108-111).
#pragma unroll
    for (int d = 0; d < ND; ++d)
#pragma unroll
      for (int k = 0; k < NC; ++k) {
        const float2 c = cmul(gg_s[d * kWave * NC + gi + k], RawGet<NC, F16>(row.m[d], k));
        r[k].x -= c.x;
        r[k].y -= c.y;
      }
#pragma unroll
    for (int d = 0; d < ND; ++d)
#pragma unroll
      for (int k = 0; k < NC; ++k) {
        const float2 mm = RawGet<NC, F16>(row.m[d], k);
        ur[d][k] += r[k].x * mm.x + r[k].y * mm.y;
        ui[d][k] += r[k].y * mm.x - r[k].x * mm.y;
        sm[d][k] += mm.x * mm.x + mm.y * mm.y;
      }
  };
  // The first rows' loads do not depend on the gains: issue them before the
  // staging so that its memory round trips overlap with them.
  Row ring[ROWS];
#pragma unroll
  for (int u = 0; u < ROWS; ++u) load_row(r0 + wave + u * n_waves, ring[u]);
  // Stage G_a,i * conj(G_b,j) of every (direction, correlation) for the 64
  // baselines of the tile (solutions repaired, then cast to float, as
  // SolveData.h:230-235). Wave 0 does it, lane = baseline; the loads of all
  // directions are independent and issued together, so staging costs about
  // two memory round trips instead of one per loop trip.
  if (wave == 0) {
    const size_t gbase = size_t(cb) * a.n_ant * a.n_sub;
    const double2* sol = (AsConst(a.state)->iteration & 1) ? a.sol1 : a.sol0;
    const double mean = AsConst(a.mean)[cb];
    const bool valid = bl < a.n_bl;
    const int a1 = valid ? AsGlobal(a.ant1)[bl] : 0;
    const int a2 = valid ? AsGlobal(a.ant2)[bl] : 0;
    // Directions in groups: each group's loads are issued together (one
    // memory round trip per group) and a compiler barrier between groups
    // bounds the registers they occupy.
    constexpr int kGroup = F16 ? 4 : 2;
#pragma unroll
    for (int d0 = 0; d0 < ND; d0 += kGroup) {
      double2 g1[kGroup][2], g2[kGroup][2];
#pragma unroll
      for (int dd = 0; dd < kGroup; ++dd) {
        if (d0 + dd >= ND) break;
        const int s = AsConst(a.seg_sol)[seg * ND + d0 + dd];
        const size_t p1 = gbase + size_t(a1) * a.n_sub + s;
        const size_t p2 = gbase + size_t(a2) * a.n_sub + s;
        g1[dd][0] = LoadGlobal(sol + p1 * 2);
        g1[dd][1] = LoadGlobal(sol + p1 * 2 + 1);
        g2[dd][0] = LoadGlobal(sol + p2 * 2);
        g2[dd][1] = LoadGlobal(sol + p2 * 2 + 1);
      }
#pragma unroll
      for (int dd = 0; dd < kGroup; ++dd) {
        if (d0 + dd >= ND) break;
        const int d = d0 + dd;
        // MakeSolutionsFinite2Pol: a non-finite pair reads as (mean, mean).
        if (!(cfinite(g1[dd][0]) && cfinite(g1[dd][1])))
          g1[dd][0] = g1[dd][1] = make_double2(mean, 0.0);
        if (!(cfinite(g2[dd][0]) && cfinite(g2[dd][1])))
          g2[dd][0] = g2[dd][1] = make_double2(mean, 0.0);
        float sc = 1.0f;
        if constexpr (F16) sc = AsConst(a.model_scale)[cb * ND + d];
#pragma unroll
        for (int k = 0; k < NC; ++k) {
          const double2 ga = g1[dd][PolI<NC>(k)];
          const double2 gb = g2[dd][PolJ<NC>(k)];
          float2 g = cmulc(make_float2(float(ga.x), float(ga.y)),
                           make_float2(float(gb.x), float(gb.y)));
          g.x *= sc;
          g.y *= sc;
          gg_s[(d * kWave + lane) * NC + k] = valid ? g : make_float2(0.0f, 0.0f);
        }
      }
      asm volatile("" ::: "memory");
    }
  }
  __syncthreads();

  // The scheduling barrier keeps each refill behind the use of its slot, so a
  // wave holds exactly ROWS rows of registers (hoisting the refills above the
  // arithmetic would need fresh registers for every slot).
  for (int row = r0 + wave; row < r1; row += ROWS * n_waves) {
#pragma unroll
    for (int u = 0; u < ROWS; ++u) {
      accumulate(ring[u]);
      __builtin_amdgcn_sched_barrier(0);
      load_row(row + (u + ROWS) * n_waves, ring[u]);
    }
  }

  // Combine the waves: wave w-1 += wave w, for w = n_waves-1 .. 1.
  for (int w = n_waves - 1; w >= 1; --w) {
    if (wave == w) {
#pragma unroll
      for (int d = 0; d < ND; ++d)
#pragma unroll
        for (int k = 0; k < NC; ++k) {
          const int v = (d * NC + k) * 3;
          red_s[(v + 0) * kWave + lane] = ur[d][k];
          red_s[(v + 1) * kWave + lane] = ui[d][k];
          red_s[(v + 2) * kWave + lane] = sm[d][k];
        }
    }
    __syncthreads();
    if (wave == w - 1) {
#pragma unroll
      for (int d = 0; d < ND; ++d)
#pragma unroll
        for (int k = 0; k < NC; ++k) {
          const int v = (d * NC + k) * 3;
          ur[d][k] += red_s[(v + 0) * kWave + lane];
          ui[d][k] += red_s[(v + 1) * kWave + lane];
          sm[d][k] += red_s[(v + 2) * kWave + lane];
        }
    }
    __syncthreads();
  }
  if (wave == 0) {
    GlobalT<float>* out =
        AsGlobal(a.partials) + (size_t(cb) * a.n_seg + seg) * ND * NC * 3 * a.bl_stride;
#pragma unroll
    for (int d = 0; d < ND; ++d)
#pragma unroll
      for (int k = 0; k < NC; ++k) {
        const size_t v = size_t(d * NC + k) * 3;
        out[(v + 0) * a.bl_stride + bl] = ur[d][k];
        out[(v + 1) * a.bl_stride + bl] = ui[d][k];
        out[(v + 2) * a.bl_stride + bl] = sm[d][k];
      }
  }
}

template <int ND, int NC, bool F16, int ROWS>
__global__ void __launch_bounds__(kCorrelateMaxThreads)
    CorrelateKernel(const IterArgs a) {
  if (a.state->done) return;
  CorrelateTile<ND, NC, F16, ROWS>(a, blockIdx.x, blockIdx.y, blockIdx.z);
}

// ---------------------------------------------------------------------------
// Gram iteration (Duo, n_dir <= kMaxGramDirections). Within a solution
// segment the gain products gg_e of a baseline are constant, so the
// correlate's sums factor once more:
//   U_d,k = sum_rows R_all,k conj(M_d,k) = C_d,k - sum_e gg_e,k G_e,d,k
//   S_d,k = G_d,d,k,  C_d,k = sum_rows D_k conj(M_d,k),
//   G_e,d,k = sum_rows M_e,k conj(M_d,k) = conj(G_d,e,k).
// C and G do not depend on the gains: gram makes them once per solve (one
// pass over the data and model planes) and apply evaluates U and S from them
// and the current gains every iteration, without visibility traffic, into
// the partials layout of correlate (antenna_solve is unchanged). For FP16
// planes both hold the unscaled half values and apply scales gg, as
// correlate does. Layout [cb][seg][k][q][bl_stride], q < ND: C_d,
// q = ND + d(d+1)/2 + e: G_e,d (e <= d).
// ---------------------------------------------------------------------------
constexpr int kGramChunk = 8;

template <int ND>
constexpr int GramPair(int e, int d) {
  return ND + d * (d + 1) / 2 + e;
}

template <int ND, bool F16>
__device__ __forceinline__ void GramTile(const IterArgs& a, const int tile, const int seg,
                                         const int cb) {
  constexpr int NC = 2;
  constexpr int NQ = ND + ND * (ND + 1) / 2;
  constexpr uint32_t kDataBytes = VisBytes<NC, false>();
  constexpr uint32_t kModelBytes = VisBytes<NC, F16>();
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = __builtin_amdgcn_readfirstlane(int(threadIdx.x) / kWave);
  const int n_waves = blockDim.x / kWave;
  const CbInfo info = LoadConst(a.cbinfo + cb);
  const int n_ch = info.n_ch;
  const size_t plane = size_t(a.n_t) * n_ch * a.bl_stride;
  const int bl = tile * kWave + lane;
  const char* dbase = reinterpret_cast<const char*>(a.data) + size_t(info.data_off) * kDataBytes;
  const char* mbase = reinterpret_cast<const char*>(a.model) + size_t(info.model_off) * kModelBytes;
  const size_t mplane = plane * kModelBytes;
  const int r0 = AsConst(a.seg_t0)[seg] * n_ch;
  const int r1 = AsConst(a.seg_t0)[seg + 1] * n_ch;
  const uint32_t row_vis = uint32_t(r1) * uint32_t(a.bl_stride);
  const __amdgpu_buffer_rsrc_t drsrc = PlaneResource(dbase, row_vis * kDataBytes);
  __amdgpu_buffer_rsrc_t mrsrc[ND];
#pragma unroll
  for (int d = 0; d < ND; ++d) mrsrc[d] = PlaneResource(mbase + d * mplane, row_vis * kModelBytes);

  float acc[NQ][NC][2];
#pragma unroll
  for (int q = 0; q < NQ; ++q)
#pragma unroll
    for (int k = 0; k < NC; ++k) acc[q][k][0] = acc[q][k][1] = 0.0f;

  for (int row = r0 + wave; row < r1; row += n_waves) {
    const uint32_t e = uint32_t(row) * uint32_t(a.bl_stride) + uint32_t(bl);
    RawVis<NC, false> dr;
    RawVis<NC, F16> mr[ND];
    LoadRaw<NC, false>(drsrc, e * kDataBytes, dr);
#pragma unroll
    for (int d = 0; d < ND; ++d) LoadRaw<NC, F16>(mrsrc[d], e * kModelBytes, mr[d]);
#pragma unroll
    for (int k = 0; k < NC; ++k) {
      const float2 dv = RawGet<NC, false>(dr, k);
      float2 m[ND];
#pragma unroll
      for (int d = 0; d < ND; ++d) m[d] = RawGet<NC, F16>(mr[d], k);
#pragma unroll
      for (int d = 0; d < ND; ++d) {
        const float2 c = cmulc(dv, m[d]);
        acc[d][k][0] += c.x;
        acc[d][k][1] += c.y;
#pragma unroll
        for (int f = 0; f <= d; ++f) {
          const float2 g = cmulc(m[f], m[d]);
          acc[GramPair<ND>(f, d)][k][0] += g.x;
          acc[GramPair<ND>(f, d)][k][1] += g.y;
        }
      }
    }
  }

  // Combine the waves in a fixed order (wave w-1 += wave w), as correlate,
  // kGramChunk terms at a time to bound the LDS.
  extern __shared__ float smem[];
#pragma unroll
  for (int q0 = 0; q0 < NQ; q0 += kGramChunk) {
    for (int w = n_waves - 1; w >= 1; --w) {
      if (wave == w) {
#pragma unroll
        for (int q = q0; q < q0 + kGramChunk && q < NQ; ++q)
#pragma unroll
          for (int k = 0; k < NC; ++k) {
            smem[(((q - q0) * NC + k) * 2 + 0) * kWave + lane] = acc[q][k][0];
            smem[(((q - q0) * NC + k) * 2 + 1) * kWave + lane] = acc[q][k][1];
          }
      }
      __syncthreads();
      if (wave == w - 1) {
#pragma unroll
        for (int q = q0; q < q0 + kGramChunk && q < NQ; ++q)
#pragma unroll
          for (int k = 0; k < NC; ++k) {
            acc[q][k][0] += smem[(((q - q0) * NC + k) * 2 + 0) * kWave + lane];
            acc[q][k][1] += smem[(((q - q0) * NC + k) * 2 + 1) * kWave + lane];
          }
      }
      __syncthreads();
    }
  }
  if (wave == 0) {
    GlobalT<float>* out = (GlobalT<float>*)AsGlobal(
        a.gram + size_t(cb * a.n_seg + seg) * NC * NQ * a.bl_stride + bl);
#pragma unroll
    for (int k = 0; k < NC; ++k)
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
        out[2 * size_t(k * NQ + q) * a.bl_stride + 0] = acc[q][k][0];
        out[2 * size_t(k * NQ + q) * a.bl_stride + 1] = acc[q][k][1];
      }
    // S_d,k = G_d,d,k does not change over the solve: its slot of the
    // correlate partials is written here, once, and the apply writes U only.
    GlobalT<float>* part =
        AsGlobal(a.partials) + (size_t(cb) * a.n_seg + seg) * ND * NC * 3 * a.bl_stride + bl;
#pragma unroll
    for (int d = 0; d < ND; ++d)
#pragma unroll
      for (int k = 0; k < NC; ++k)
        part[(size_t(d * NC + k) * 3 + 2) * a.bl_stride] = acc[GramPair<ND>(d, d)][k][0];
  }
}

template <int ND, bool F16>
__device__ __forceinline__ void ApplyGramTile(const IterArgs& a, const int tile, const int seg,
                                              const int cb) {
  // One wave per correlation k: each wave issues its cross-product loads
  // (8-byte vector loads) at once; the segment's gains are staged in LDS
  // once per workgroup.
  constexpr int NC = 2;
  constexpr int NQ = ND + ND * (ND + 1) / 2;
  typedef float f2v __attribute__((ext_vector_type(2)));
  // [n_ant][ND * NC + 1], repaired, cast to float. Lanes of a tile read
  // antennas a2 = a1 + 1, a1 + 2, ...: the one-value pad per antenna makes the
  // row stride 2 * ND * NC + 2 dwords (an odd multiple of 2), so the 64-bit
  // reads of 32 consecutive antennas fall in distinct bank pairs.
  extern __shared__ float2 s_gain[];
  constexpr int kGainRow = ND * NC + 1;
  const int lane = threadIdx.x & (kWave - 1);
  const int k = int(threadIdx.x >> 6);  // kApplyThreads = NC waves
  const int bl = tile * kWave + lane;
  const bool valid = bl < a.n_bl;
  // The segment's gains of every antenna and direction, repaired (a pair with
  // a non-finite entry reads as the block mean) and cast to float, as
  // correlate stages them. Vector loads complete in order, so the first
  // kStageUnroll entries per thread are loaded before the cross products:
  // writing them to LDS then waits for these loads only, while the cross
  // products stream in. Further entries (more than kStageUnroll x
  // kApplyThreads) are staged after.
  constexpr int kStageUnroll = 4;
  const int n_entries = a.n_ant * ND;
  const size_t gbase = size_t(cb) * a.n_ant * a.n_sub;
  const double2* sol = (AsConst(a.state)->iteration & 1) ? a.sol1 : a.sol0;
  // entry e: antenna e / ND, direction e % ND
  auto pair_of = [&](int e, int sub) { return gbase + size_t(e / ND) * a.n_sub + sub; };
  // Loads without branches (an entry past the end loads entry 0 and is not
  // written), so that they issue back to back.
  int sub[kStageUnroll];
#pragma unroll
  for (int u = 0; u < kStageUnroll; ++u) {
    const int e = threadIdx.x + u * kApplyThreads;
    sub[u] = AsConst(a.seg_sol)[seg * ND + (e < n_entries ? e : 0) % ND];
  }
  double2 sg[kStageUnroll][2];
#pragma unroll
  for (int u = 0; u < kStageUnroll; ++u) {
    const int e = threadIdx.x + u * kApplyThreads;
    const size_t p = pair_of(e < n_entries ? e : 0, sub[u]);
    sg[u][0] = LoadGlobal(sol + p * 2);
    sg[u][1] = LoadGlobal(sol + p * 2 + 1);
  }
  // keep the scheduler from moving the cross-product loads ahead of these
  __builtin_amdgcn_sched_barrier(0);
  const GlobalT<f2v>* in = (const GlobalT<f2v>*)AsGlobal(
      a.gram + (size_t(cb * a.n_seg + seg) * NC * NQ + size_t(k) * NQ) * a.bl_stride + bl);
  f2v q[NQ];
#pragma unroll
  for (int i = 0; i < NQ; ++i) q[i] = in[size_t(i) * a.bl_stride];
  const int bl_c = valid ? bl : 0;
  const int a1v = AsGlobal(a.ant1)[bl_c];
  const int a2v = AsGlobal(a.ant2)[bl_c];
  {
    const double mean = AsConst(a.mean)[cb];
    auto put = [&](int e, double2 g0, double2 g1) {
      if (!(cfinite(g0) && cfinite(g1))) g0 = g1 = make_double2(mean, 0.0);
      s_gain[e * NC + e / ND + 0] = make_float2(float(g0.x), float(g0.y));
      s_gain[e * NC + e / ND + 1] = make_float2(float(g1.x), float(g1.y));
    };
#pragma unroll
    for (int u = 0; u < kStageUnroll; ++u) {
      const int e = threadIdx.x + u * kApplyThreads;
      if (e < n_entries) put(e, sg[u][0], sg[u][1]);
    }
    for (int e = threadIdx.x + kStageUnroll * kApplyThreads; e < n_entries; e += kApplyThreads) {
      const size_t p = pair_of(e, AsConst(a.seg_sol)[seg * ND + e % ND]);
      put(e, LoadGlobal(sol + p * 2), LoadGlobal(sol + p * 2 + 1));
    }
  }
  const int a1 = valid ? a1v : 0;
  const int a2 = valid ? a2v : 0;
  __syncthreads();
  // gg_d,k with the FP16 scale folded in
  float2 gg[ND];
#pragma unroll
  for (int d = 0; d < ND; ++d) {
    float sc = 1.0f;
    if constexpr (F16) sc = AsConst(a.model_scale)[cb * ND + d];
    const float2 g = cmulc(s_gain[a1 * kGainRow + d * NC + k], s_gain[a2 * kGainRow + d * NC + k]);
    gg[d] = valid ? make_float2(g.x * sc, g.y * sc) : make_float2(0.0f, 0.0f);
  }
  GlobalT<float>* out =
      AsGlobal(a.partials) + (size_t(cb) * a.n_seg + seg) * ND * NC * 3 * a.bl_stride + bl;
#pragma unroll
  for (int d = 0; d < ND; ++d) {
    float2 u = make_float2(q[d].x, q[d].y);
    // directions subtracted in order, as the residual is formed
#pragma unroll
    for (int e = 0; e < ND; ++e) {
      const f2v ge = e <= d ? q[GramPair<ND>(e, d)] : q[GramPair<ND>(d, e)];
      const float2 g = make_float2(ge.x, e <= d ? ge.y : -ge.y);
      const float2 c = cmul(gg[e], g);
      u.x -= c.x;
      u.y -= c.y;
    }
    const size_t v = size_t(d * NC + k) * 3;
    out[(v + 0) * a.bl_stride] = u.x;
    out[(v + 1) * a.bl_stride] = u.y;
    // S (slot v + 2) was written by the cross-products pass
  }
}

template <int ND, bool F16>
__global__ void __launch_bounds__(kCorrelateMaxThreads) GramKernel(const IterArgs a) {
  GramTile<ND, F16>(a, blockIdx.x, blockIdx.y, blockIdx.z);
}

template <int ND, bool F16>
__global__ void __launch_bounds__(kApplyThreads) ApplyGramKernel(const IterArgs a) {
  if (AsConst(a.state)->done) return;
  ApplyGramTile<ND, F16>(a, blockIdx.x, blockIdx.y, blockIdx.z);
}

template <int ND, bool F16>
__global__ void __launch_bounds__(kCorrelateMaxThreads)
    BatchGramKernel(const IterArgs* __restrict__ args, const BatchTile* __restrict__ tiles) {
  const BatchTile t = LoadConst(tiles + blockIdx.x);
  const IterArgs a = LoadConst(args + t.interval);
  GramTile<ND, F16>(a, t.tile, t.seg, t.cb);
}

// Workgroups go to the 8 XCDs round-robin (workgroup w to XCD w % 8), and
// each XCD has its own L2. Workgroup w takes tile-list entry
// (w % 8) * per + w / 8 (per = n / 8), so every XCD works through one
// contiguous run of the list, in which consecutive tiles of an interval's
// block share that block's gains in its L2; the n % 8 tail keeps its order.
// A permutation of the list: results do not depend on it.
constexpr int kXcds = 8;
__device__ __forceinline__ int XcdContiguous(int w, int n) {
  const int per = n / kXcds;
  if (w >= per * kXcds) return w;
  return (w % kXcds) * per + w / kXcds;
}

template <int ND, bool F16>
__global__ void __launch_bounds__(kApplyThreads)
    BatchApplyGramKernel(const IterArgs* __restrict__ args, const BatchTile* __restrict__ tiles) {
  const BatchTile t = LoadConst(tiles + XcdContiguous(blockIdx.x, gridDim.x));
  const IterArgs a = LoadConst(args + t.interval);
  if (AsConst(a.state)->done) return;
  ApplyGramTile<ND, F16>(a, t.tile, t.seg, t.cb);
}

// Solutions of MakeSolutionsFinite2Pol (SolverBase.cc:342-371), applied on
// the fly: a pair with a non-finite entry reads as (mean, mean), where mean is
// the block's average |pol-0 value| over finite pairs (:354), kept per block.
__device__ __forceinline__ double2 FixedSolution(const double2* sol, size_t pair, int pol,
                                                 double mean) {
  const double2 g0 = sol[pair * 2];
  const double2 g1 = sol[pair * 2 + 1];
  if (!(cfinite(g0) && cfinite(g1))) return make_double2(mean, 0.0);
  return pol ? g1 : g0;
}

// Copies n 4-byte words from global memory to LDS with 16-byte loads, four
// independent loads in flight per thread (src and dst 16-byte aligned).
template <typename T>
__device__ void StageWords(const T* __restrict__ src, T* dst, int n) {
  static_assert(sizeof(T) == 4, "4-byte words");
  typedef unsigned int u4 __attribute__((ext_vector_type(4)));
  const int n4 = n / 4;
  const u4* s4 = reinterpret_cast<const u4*>(src);
  u4* d4 = reinterpret_cast<u4*>(dst);
  const int bd = blockDim.x;
  int i = threadIdx.x;
  for (; i + 3 * bd < n4; i += 4 * bd) {
    const u4 v0 = s4[i], v1 = s4[i + bd], v2 = s4[i + 2 * bd], v3 = s4[i + 3 * bd];
    d4[i] = v0;
    d4[i + bd] = v1;
    d4[i + 2 * bd] = v2;
    d4[i + 3 * bd] = v3;
  }
  for (; i < n4; i += bd) d4[i] = s4[i];
  for (int j = n4 * 4 + threadIdx.x; j < n; j += bd) dst[j] = src[j];
}

// Block-wide deterministic sums of two (double, int) pairs: a fixed xor
// shuffle tree inside each wave, then the per-wave results in wave order.
// `red` needs 2 * (blockDim.x / 64) doubles, `redc` as many ints, used by
// nothing else in the workgroup (one barrier, between their writes and reads).
__device__ void BlockSum2(double& v, int& c, double& w, int& e, double* red, int* redc) {
  for (int off = 32; off > 0; off >>= 1) {
    v += __shfl_xor(v, off);
    c += __shfl_xor(c, off);
    w += __shfl_xor(w, off);
    e += __shfl_xor(e, off);
  }
  const int n_waves = blockDim.x / 64;
  const int wave = threadIdx.x / 64;
  if ((threadIdx.x & 63) == 0) {
    red[wave] = v;
    redc[wave] = c;
    red[n_waves + wave] = w;
    redc[n_waves + wave] = e;
  }
  __syncthreads();
  v = 0.0;
  c = 0;
  w = 0.0;
  e = 0;
  for (int i = 0; i < n_waves; ++i) {
    v += red[i];
    c += redc[i];
    w += red[n_waves + i];
    e += redc[n_waves + i];
  }
}

// Block-wide deterministic sum of one double and one int: a fixed shuffle
// tree in each wave, then the wave sums in wave order after one barrier.
__device__ void BlockSum(double& v, int& c, double* red, int* redc) {
  for (int off = 32; off > 0; off >>= 1) {
    v += __shfl_xor(v, off);
    c += __shfl_xor(c, off);
  }
  const int n_waves = blockDim.x / 64;
  if ((threadIdx.x & 63) == 0) {
    red[threadIdx.x / 64] = v;
    redc[threadIdx.x / 64] = c;
  }
  __syncthreads();
  v = 0.0;
  c = 0;
  for (int i = 0; i < n_waves; ++i) {
    v += red[i];
    c += redc[i];
  }
}

// The FP64 finishing arithmetic below is written without fused multiply-adds
// (contraction off), so every kernel that inlines it (single, batched, fused)
// rounds the same way and their solutions agree bit for bit.

// a / b, scaled by the larger part of b (Smith's algorithm).
__device__ __forceinline__ double2 CDiv(double2 a, double2 b) {
#pragma clang fp contract(off)
  if (fabs(b.y) <= fabs(b.x)) {
    const double r = b.y / b.x, den = b.x + b.y * r;
    return make_double2((a.x + a.y * r) / den, (a.y - a.x * r) / den);
  }
  const double r = b.x / b.y, den = b.x * r + b.y;
  return make_double2((a.x * r + a.y) / den, (a.y * r - a.x) / den);
}

// Convergence term of one solution pair (SolverBase.cc:254-268) between its
// repaired old value and its next value, and the next value's |pol 0| for the
// repair mean (SolverBase.cc:342-371), added to the running sums. A scalar
// solver's pair holds one gain twice: its term is that of one solution,
// |(next - old) / old| unless old is 0 (:244-253), and its mean is over the
// finite solutions (:308-333).
__device__ __forceinline__ void AddPairTerms(const double2 (&old)[2], const double2 (&next)[2],
                                             bool scalar, double& conv_sum, int& conv_n,
                                             double& mean_sum, int& mean_n) {
#pragma clang fp contract(off)
  if (scalar) {
    if (old[0].x != 0.0 || old[0].y != 0.0) {
      const double2 q =
          CDiv(make_double2(next[0].x - old[0].x, next[0].y - old[0].y), old[0]);
      const double a = hypot(q.x, q.y);
      if (isfinite(a)) {
        conv_sum += a;
        conv_n += 1;
      }
    }
    if (cfinite(next[0])) {
      mean_sum += hypot(next[0].x, next[0].y);
      mean_n += 1;
    }
    return;
  }
  // (the complex products written out here, so that they are not contracted)
  const double n0 = old[0].x * old[0].x + old[0].y * old[0].y;
  const double n1 = old[1].x * old[1].x + old[1].y * old[1].y;
  const double2 inv0 = make_double2(old[0].x / n0, -old[0].y / n0);
  const double2 inv1 = make_double2(old[1].x / n1, -old[1].y / n1);
  if (cfinite(inv0) && cfinite(inv1)) {
    const double2 e0 = make_double2(next[0].x - old[0].x, next[0].y - old[0].y);
    const double2 e1 = make_double2(next[1].x - old[1].x, next[1].y - old[1].y);
    const double2 d0 = make_double2(e0.x * inv0.x - e0.y * inv0.y, e0.x * inv0.y + e0.y * inv0.x);
    const double2 d1 = make_double2(e1.x * inv1.x - e1.y * inv1.y, e1.x * inv1.y + e1.y * inv1.x);
    const double sumabs = hypot(d0.x, d0.y) + hypot(d1.x, d1.y);
    if (isfinite(sumabs)) {
      conv_sum += sumabs;
      conv_n += 2;
    }
  }
  if (cfinite(next[0]) && cfinite(next[1])) {
    mean_sum += hypot(next[0].x, next[0].y);
    mean_n += 1;
  }
}

// The end of SolveDirection for one solution pair (antenna, sub-solution) of
// a block, from its stored current values cv and its FP64 numerator and
// denominator sums per polarization: the
// division with NaN for a zero denominator (IterativeDiagonalSolver.cc:
// 206-221), the step (SolverBase.cc:88-158, the polar step for phase-only
// solves), the optional denominators, and the pair's convergence and repair-
// mean terms added to the running sums. The scalar solver fits one gain to
// both polarizations: the sums of both (IterativeScalarSolver.cc:142-173).
__device__ __forceinline__ void FinishSolution(const IterArgs& a, const double2 (&cv)[2],
                                               double2* nxt, const size_t pair, const double mean,
                                               double (&nr)[2], double (&ni)[2], double (&dn)[2],
                                               double& conv_sum, int& conv_n, double& mean_sum,
                                               int& mean_n, double2 (&out)[2]) {
#pragma clang fp contract(off)
  if (a.scalar) {
    nr[0] = nr[1] = nr[0] + nr[1];
    ni[0] = ni[1] = ni[0] + ni[1];
    dn[0] = dn[1] = dn[0] + dn[1];
  }
  const double nan = std::numeric_limits<double>::quiet_NaN();
  double2 stepped[2];
  double2 old[2];
#pragma unroll
  for (int p = 0; p < 2; ++p) {
    const double2 nx = dn[p] == 0.0 ? make_double2(nan, 0.0)
                                    : make_double2(nr[p] / dn[p], ni[p] / dn[p]);
    // the repaired current value (FixedSolution on the pair's stored values)
    old[p] = (cfinite(cv[0]) && cfinite(cv[1])) ? cv[p] : make_double2(mean, 0.0);
    if (!a.phase_only) {
      stepped[p] = make_double2(old[p].x * (1.0 - a.step_size) + nx.x * a.step_size,
                                old[p].y * (1.0 - a.step_size) + nx.y * a.step_size);
    } else {
      // Step along the complex circle (SolverBase.cc:147-155).
      const double phase_from = atan2(old[p].y, old[p].x);
      double distance = atan2(nx.y, nx.x) - phase_from;
      if (distance > M_PI)
        distance -= 2.0 * M_PI;
      else if (distance < -M_PI)
        distance += 2.0 * M_PI;
      double sn, cs;
      sincos(phase_from + a.step_size * distance, &sn, &cs);
      stepped[p] = make_double2(cs, sn);
    }
    if (nxt) {  // null: the caller keeps the values (the fused iteration's other workgroups)
      nxt[pair * 2 + p] = stepped[p];
      if (a.den_out) a.den_out[pair * 2 + p] = dn[p];
    }
    out[p] = stepped[p];
  }
  AddPairTerms(old, stepped, a.scalar != 0, conv_sum, conv_n, mean_sum, mean_n);
}

// ---------------------------------------------------------------------------
// antenna_solve: grid (n_dir, n_cb); `split` threads per (antenna, solution
// of d) share that antenna's baseline list and are combined by a fixed
// shuffle tree. With T_k = U_k + G_a,i conj(G_b,j) S_k = sum_v R_d,k conj(M_k)
// (the residual with direction d added back), the reference's sums become
//   antenna a as first antenna, pol p:  num += G_b,j T_k,  den += |G_b,j|^2 S_k
//                                       over k with i(k) = p
//   antenna b as second antenna, pol p: num += G_a,i conj(T_k),
//                                       den += |G_a,i|^2 S_k over k with j(k) = p
// (IterativeDiagonalSolver.cc:172-197, AddNormToDenominator :20-35). Then the
// FP64 division (:206-221), the step (SolverBase.cc:88-158) and the
// convergence term (:254-268).
// Solutions ping-pong between sol0 and sol1 by iteration parity.
// ---------------------------------------------------------------------------
// CHUNKED: the partials staged `solve_chunk` baselines at a time (arrays
// whose block does not fit one stage); otherwise all of a block's at once.
template <int NC, bool F16, bool CHUNKED>
__device__ __forceinline__ void AntennaSolveBlock(const IterArgs& a, const int d, const int cb) {
  __shared__ double red[2 * kBlockThreads / 64];
  __shared__ int redc[2 * kBlockThreads / 64];
  const int iteration = a.state->iteration;
  const double2* cur = (iteration & 1) ? a.sol1 : a.sol0;
  double2* nxt = (iteration & 1) ? a.sol0 : a.sol1;
  const int nsd = a.n_sol_dir[d];
  const int s0 = a.sol_start[d];
  const int n_items = a.n_ant * nsd;
  const int split = a.split;
  const int item = threadIdx.x / split;
  const int q = threadIdx.x - item * split;
  const bool active = item < n_items;
  const int ant = active ? item / nsd : 0;
  const int s_abs = s0 + (active ? item - ant * nsd : 0);
  const size_t gbase = size_t(cb) * a.n_ant * a.n_sub;
  const double mean = a.mean[cb];
  float scale = 1.0f;
  if constexpr (F16) scale = a.model_scale[cb * a.n_dir + d];
  const double sc = scale;
  const double sc2 = sc * sc;
  const int bs = a.bl_stride;
  // LDS: the gains of this direction's solutions as the correlate pass used
  // them (repaired, then cast to float) and the partials of one segment and
  // correlation, `chunk` baselines at a time (all of them where they fit).
  // The per-antenna baseline lists are read from global memory (cached):
  // each entry holds (baseline << 1 | side, other antenna), in baseline
  // order, so a thread's walk over its list is cut into chunks by a cursor
  // and visits its entries in the same order whatever the chunk size.
  extern __shared__ double smem_d[];
  double2* g_s = reinterpret_cast<double2*>(smem_d);             // [n_ant][nsd][2]
  float* part_s = reinterpret_cast<float*>(g_s + size_t(a.n_ant) * nsd * 2);  // [3][chunk]
  [[maybe_unused]] const int chunk =
      (CHUNKED && a.solve_chunk > 0 && a.solve_chunk < bs) ? a.solve_chunk : bs;
  for (int i = threadIdx.x; i < a.n_ant * nsd; i += blockDim.x) {
    const int an = i / nsd;
    const size_t pair = gbase + size_t(an) * a.n_sub + s0 + (i - an * nsd);
    double2 g0 = LoadGlobal(cur + pair * 2), g1 = LoadGlobal(cur + pair * 2 + 1);
    if (!(cfinite(g0) && cfinite(g1))) g0 = g1 = make_double2(mean, 0.0);
    g_s[2 * i] = make_double2(double(float(g0.x)), double(float(g0.y)));
    g_s[2 * i + 1] = make_double2(double(float(g1.x)), double(float(g1.y)));
  }
  int e_begin = 0, e_end = 0;
  if (active) {
    e_begin = AsGlobal(a.csr_off)[ant];
    e_end = AsGlobal(a.csr_off)[ant + 1];
  }
  const GlobalT<int32_t>* csr = AsGlobal(a.csr);
  const int rel = s_abs - s0;
  auto gain = [&](int an, int pol) { return g_s[(an * nsd + rel) * 2 + pol]; };

  double nr[2] = {0.0, 0.0}, ni[2] = {0.0, 0.0}, dn[2] = {0.0, 0.0};
  if constexpr (!CHUNKED) {
    // the first segment's first correlation is staged under the gains' barrier
    StageWords(a.partials + size_t(cb) * a.n_seg * a.n_dir * NC * 3 * size_t(bs) +
                   size_t(d) * NC * 3 * size_t(bs),
               part_s, 3 * bs);
    __syncthreads();
    const double2 gant[2] = {active ? gain(ant, 0) : make_double2(0, 0),
                             active ? gain(ant, 1) : make_double2(0, 0)};
    // The partials are staged one correlation at a time (3 x bl_stride floats),
    // which keeps the LDS per workgroup small enough for six workgroups per CU.
    for (int seg = 0; seg < a.n_seg; ++seg) {
      const int s_seg = AsConst(a.seg_sol)[seg * a.n_dir + d];
      const float* src =
          a.partials + ((size_t(cb) * a.n_seg + seg) * a.n_dir + d) * NC * 3 * size_t(bs);
  #pragma unroll
      for (int k = 0; k < NC; ++k) {
        if (seg > 0 || k > 0) {
          __syncthreads();  // the previous stage has been read
          StageWords(src + size_t(k) * 3 * bs, part_s, 3 * bs);
          __syncthreads();
        }
        if (!active || s_seg != s_abs) continue;
        const int i = PolI<NC>(k);
        const int j = PolJ<NC>(k);
        for (int e = e_begin + q; e < e_end; e += split) {
          const int code = csr[2 * e];
          const int other = csr[2 * e + 1];
          const int b = code >> 1;
          const int side = code & 1;
          const double2 u = make_double2(sc * part_s[b], sc * part_s[bs + b]);
          const double s = sc2 * part_s[2 * bs + b];
          if (side == 0) {
            // this antenna is a (first); the other is b
            const double2 go = gain(other, j);
            const double2 gg = cmulc(gant[i], go);
            const double2 t = make_double2(u.x + gg.x * s, u.y + gg.y * s);
            const double2 c = cmul(go, t);
            nr[i] += c.x;
            ni[i] += c.y;
            dn[i] += cnorm(go) * s;
          } else {
            // this antenna is b (second); the other is a
            const double2 go = gain(other, i);
            const double2 gg = cmulc(go, gant[j]);
            const double2 t = make_double2(u.x + gg.x * s, u.y + gg.y * s);
            const double2 c = cmulc(go, t);  // G_a,i * conj(T)
            nr[j] += c.x;
            ni[j] += c.y;
            dn[j] += cnorm(go) * s;
          }
        }
      }
    }
  } else {
    // one correlation's partials [3][bs] (from src), baselines [c0, c0 + chunk)
    auto stage = [&](const float* src, int c0) {
      if (chunk == bs) {
        StageWords(src, part_s, 3 * bs);
      } else {
        const int len = min(chunk, bs - c0);
  #pragma unroll
        for (int r = 0; r < 3; ++r) StageWords(src + size_t(r) * bs + c0, part_s + r * chunk, len);
      }
    };
    // the first segment's first correlation (first chunk) is staged under the
    // gains' barrier
    stage(a.partials + size_t(cb) * a.n_seg * a.n_dir * NC * 3 * size_t(bs) +
              size_t(d) * NC * 3 * size_t(bs),
          0);
    __syncthreads();
    const double2 gant[2] = {active ? gain(ant, 0) : make_double2(0, 0),
                             active ? gain(ant, 1) : make_double2(0, 0)};
    // The partials are staged one correlation at a time (3 x bl_stride floats),
    // which keeps the LDS per workgroup small enough for six workgroups per CU.
    for (int seg = 0; seg < a.n_seg; ++seg) {
      const int s_seg = AsConst(a.seg_sol)[seg * a.n_dir + d];
      const float* src =
          a.partials + ((size_t(cb) * a.n_seg + seg) * a.n_dir + d) * NC * 3 * size_t(bs);
  #pragma unroll
      for (int k = 0; k < NC; ++k) {
        int e = e_begin + q;  // the thread's cursor in its baseline list
        for (int c0 = 0; c0 < bs; c0 += chunk) {
          if (seg > 0 || k > 0 || c0 > 0) {
            __syncthreads();  // the previous stage has been read
            stage(src + size_t(k) * 3 * bs, c0);
            __syncthreads();
          }
          if (!active || s_seg != s_abs) continue;
          const int i = PolI<NC>(k);
          const int j = PolJ<NC>(k);
          // one entry of the list: its baseline b (relative to the chunk)
          auto add_edge = [&](int ei, int code) {
            const int b = (code >> 1) - c0;
            const int other = csr[2 * ei + 1];
            const int side = code & 1;
            const double2 u = make_double2(sc * part_s[b], sc * part_s[chunk + b]);
            const double s = sc2 * part_s[2 * chunk + b];
            if (side == 0) {
              // this antenna is a (first); the other is b
              const double2 go = gain(other, j);
              const double2 gg = cmulc(gant[i], go);
              const double2 t = make_double2(u.x + gg.x * s, u.y + gg.y * s);
              const double2 c = cmul(go, t);
              nr[i] += c.x;
              ni[i] += c.y;
              dn[i] += cnorm(go) * s;
            } else {
              // this antenna is b (second); the other is a
              const double2 go = gain(other, i);
              const double2 gg = cmulc(go, gant[j]);
              const double2 t = make_double2(u.x + gg.x * s, u.y + gg.y * s);
              const double2 c = cmulc(go, t);  // G_a,i * conj(T)
              nr[j] += c.x;
              ni[j] += c.y;
              dn[j] += cnorm(go) * s;
            }
          };
          if (chunk == bs) {
            // one stage: the whole list, a loop without a data-dependent exit
            for (; e < e_end; e += split) add_edge(e, csr[2 * e]);
          } else {
            for (; e < e_end; e += split) {
              const int code = csr[2 * e];
              if ((code >> 1) - c0 >= chunk) break;  // in a later chunk
              add_edge(e, code);
            }
          }
        }
      }
    }
  }
  // Combine the `split` lanes of an item (adjacent lanes, fixed tree).
  for (int m = 1; m < split; m <<= 1) {
#pragma unroll
    for (int p = 0; p < 2; ++p) {
      nr[p] += __shfl_xor(nr[p], m);
      ni[p] += __shfl_xor(ni[p], m);
      dn[p] += __shfl_xor(dn[p], m);
    }
  }

  double conv_sum = 0.0, mean_sum = 0.0;
  int conv_n = 0, mean_n = 0;
  if (active && q == 0) {
    const size_t pair = gbase + size_t(ant) * a.n_sub + s_abs;
    const double2 cv[2] = {LoadGlobal(cur + pair * 2), LoadGlobal(cur + pair * 2 + 1)};
    double2 unused[2];
    FinishSolution(a, cv, nxt, pair, mean, nr, ni, dn, conv_sum, conv_n, mean_sum, mean_n,
                   unused);
  }
  BlockSum2(conv_sum, conv_n, mean_sum, mean_n, red, redc);
  const int wg = cb * a.n_dir + d;
  if (threadIdx.x == 0) {
    a.wg_conv[wg] = conv_sum;
    a.wg_conv_n[wg] = conv_n;
    a.wg_mean[wg] = mean_sum;
    a.wg_mean_n[wg] = mean_n;
  }

}

template <int NC, bool F16, bool CHUNKED>
__global__ void __launch_bounds__(kBlockThreads) AntennaSolveKernel(const IterArgs a) {
  if (a.state->done) return;
  AntennaSolveBlock<NC, F16, CHUNKED>(a, blockIdx.x, blockIdx.y);
}

// ---------------------------------------------------------------------------
// Fused Gram iteration: the apply and the antenna solve of a channel block
// with the per-antenna sums formed on chip (no per-baseline partials in HBM).
// For problems of Duo data with at most 64 antennas and one solution per
// direction in one time segment.
//
// Every wave is a ring of the block's n antennas, lane a = antenna a. At
// walk step s (1 <= s <= n/2) lane a takes the baseline {a, (a + s) mod n}
// (for even n at s = n/2 only lanes a < n/2, which covers each such pair
// once), so the steps visit every baseline exactly once. The cross products
// are stored in this walk order, row by row: [cb][k][step row][term][n]
// (zero where a pair has no baseline), so a wave's row is one contiguous
// run of NQ * n * 8 bytes and a workgroup's rows follow each other. Lanes
// without a pair (a >= n, and a >= n/2 in the last row for even n) load
// nothing: their offset is past the buffer resource's range, which returns
// zeros without a memory access.
// The terms are C_d (d < ND), the off-diagonal G_e,d (e < d) and the real
// diagonals G_d,d two per 8-byte slot (they are sums of |M_d|^2).
//
// For its baseline (A, B) and every direction d the lane forms
//   U_d = C_d - sum_e gg_e G_e,d,  S_d = G_d,d,  T_d = U_d + gg_d S_d
// (gg_e = g_A conj(g_B), the FP16 scale folded in) and the two terms of the
// reference's SolveDirection (IterativeDiagonalSolver.cc:172-197):
//   antenna A: num += g_B T, den += |g_B|^2 S;  B: num += g_A conj(T), den += |g_A|^2 S.
// U is FP32 as in the apply; T, the two terms and the sums are FP64, as the
// antenna solve forms them. The lane's own antenna's term goes to its
// accumulators; the partner's term is pulled by the partner's lane (lane b
// reads lane (b - s) mod n, whose partner is b).
//
// Work split: a block's walk rows of one polarization are cut into `groups`
// row groups; one workgroup of four waves per (block, polarization, group),
// its waves taking every fourth row. The workgroup adds its waves' sums in
// LDS in wave order and stores them (per antenna and direction, FP64) in
// the scratch `gs_part`; a second kernel (GramFinishBlock, one workgroup per
// block) adds the block's 2 x groups sums in a fixed order and finishes every
// solution. (Finishing in the last workgroup to arrive needs device-scope
// fences, which write back and invalidate an XCD's L2 per workgroup; the
// kernel boundary is cheaper.) Results differ from apply + antenna solve by
// the summation order only, and do not depend on the order in which
// workgroups run.
// ---------------------------------------------------------------------------
constexpr int kGramSolveThreads = 256;
constexpr int kGramSolveWaves = kGramSolveThreads / kWave;

// Walk-order term slots: C_d at d, G_e,d (e < d) after them, then the
// diagonal pairs (G_2j,2j, G_2j+1,2j+1).
template <int ND>
constexpr int WalkOffPair(int e, int d) {
  return ND + d * (d - 1) / 2 + e;
}
template <int ND>
constexpr int WalkDiag(int d) {
  return ND + ND * (ND - 1) / 2 + d / 2;
}

// The decision after an iteration from its convergence sums (sum, count over
// every (block, direction) in a fixed order) and the state before it: the
// next state and the iteration's step magnitude. AssignSolutions
// (SolverBase.cc:296-306); stall detection (:63-86); the stopping rule
// (SolverBase.h:274-285). The only constraint the device applies is the
// smoothness constraint, which is always satisfied (Constraint::Satisfied),
// so the plain accuracy applies; the solve time is unused because smoothing
// does not depend on it. Shared by the decision kernel and the merged fused
// iteration, so both decide with the same arithmetic.
__device__ __forceinline__ DeviceState DecideState(const IterArgs& a, DeviceState st, double sum,
                                                   int count, double& step_magnitude) {
#pragma clang fp contract(off)
  const int it = st.iteration + 1;
  step_magnitude = count == 0 ? 0.0 : sum / a.step_size / count;
  const bool converged = step_magnitude <= a.accuracy;
  bool stalled = false;
  if (a.detect_stalling) {
    if (it == 1) {
      st.n_var_count = 1.0;
      st.step_mean = step_magnitude;
      st.step_var = 0.0;
    } else {
      st.n_var_count += 1.0;
      const double d1 = step_magnitude - st.step_mean;
      st.step_mean += d1 / st.n_var_count;
      const double d2 = step_magnitude - st.step_mean;
      st.step_var += d1 * d2;
    }
    if (it >= 30) stalled = st.step_mean < a.step_diff_sigma * sqrt(st.step_var / st.n_var_count);
  }
  const bool ready = it >= a.max_iterations || converged || stalled;
  const bool stop = it >= a.min_iterations && ready;
  st.iteration = it;
  st.has_converged = converged ? 1 : 0;
  st.last_step = step_magnitude;
  st.done = stop ? 1 : 0;
  return st;
}

// FP64 add to an LDS word (ds_add_f64, no value returned).
__device__ __forceinline__ void LdsAdd(double* p, double v) {
  __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// The fused iteration's per-workgroup sums of iteration parity `par`: two
// sets, so that a kernel can finish one iteration from its sums while it
// stores those of the next.
__device__ __forceinline__ double* GramParts(const IterArgs& a, int par) {
  return a.gs_part + size_t(par) * a.n_cb * 2 * kGramSolveMaxGroups * a.n_dir * 3 * kWave;
}

// Sums of row groups g0 and g0 + 1 (those below `groups`) of both
// polarizations for the thread's (direction, antenna) items: direction
// wave + u * waves, antenna = lane.
template <int ND, int NE>
__device__ __forceinline__ void LoadGroupSums(const GlobalT<double>* all, const int groups,
                                              const int g0, const int wave, const int lane,
                                              double (&v)[NE][2][2][3]) {
  const size_t part_stride = size_t(ND) * 3 * kWave;
#pragma unroll
  for (int u = 0; u < NE; ++u) {
    const int d = wave + u * kGramSolveWaves;
    if (d >= ND) continue;
#pragma unroll
    for (int p = 0; p < 2; ++p)
#pragma unroll
      for (int j = 0; j < 2; ++j)
        if (g0 + j < groups) {
          const GlobalT<double>* x =
              all + size_t(p * groups + g0 + j) * part_stride + size_t(d) * 3 * kWave + lane;
#pragma unroll
          for (int c = 0; c < 3; ++c) v[u][p][j][c] = x[c * kWave];
        }
  }
}

template <int NE>
__device__ __forceinline__ void AddGroupSums(const double (&v)[NE][2][2][3], const int groups,
                                             const int g0, double (&nr)[NE][2],
                                             double (&ni)[NE][2], double (&dn)[NE][2]) {
#pragma unroll
  for (int u = 0; u < NE; ++u)
#pragma unroll
    for (int p = 0; p < 2; ++p)
#pragma unroll
      for (int j = 0; j < 2; ++j)
        if (g0 + j < groups) {
          nr[u][p] += v[u][p][j][0];
          ni[u][p] += v[u][p][j][1];
          dn[u][p] += v[u][p][j][2];
        }
}

// Fused Gram iteration kernel body: the sums of the iteration the state
// names (FIN false; a finishing kernel and the decision follow), or (FIN
// true) one launch per iteration that decides, finishes and forms sums:
//  1. the decision of the iteration finished by the previous launch (state
//     `pending`), from its convergence partials (the same sums in the same
//     order, and the same DecideState, as DecideBlock); every workgroup of
//     the interval takes it, and the last workgroup to have read the state
//     (an arrival count; each arrives after its own read) stores the next
//     state. A stopping decision ends the launch: the current solutions are
//     those the previous launch stored.
//  2. the end of the next iteration f: every workgroup of a block finishes
//     all of the block's solutions from the stored sums (the same sums in the
//     same order, and the same FinishSolution, as GramFinishBlock) and keeps
//     its share in registers; the block's first workgroup stores them with
//     the convergence and repair-mean partials of f (their own parity half).
//     The repair means of f's current and new solutions are formed as the
//     decision forms them (the directions' |pol 0| sums in direction order).
//  3. the sums of iteration f + 1 from the new gains.
// So the values are those of sums, finishing kernel and decision launched in
// turn, bit for bit; an interval that stops has formed one more iteration's
// sums, which nothing reads.
template <int ND, bool F16, bool FIN>
__device__ __forceinline__ void GramSolveBlock(const IterArgs& a, const int cb, const int part) {
  constexpr int NC = 2;
  constexpr int NQ = ND + ND * (ND - 1) / 2 + (ND + 1) / 2;
  constexpr int kGainEntries = (ND * kWave + kGramSolveThreads - 1) / kGramSolveThreads;
  typedef float f2v __attribute__((ext_vector_type(2)));
  __shared__ float2 s_gain[ND][kWave];
  // each wave's per-antenna sums (numerator re, im, denominator per
  // direction): FP64 adds in LDS, which need no registers (a wave's region
  // is its own, so every address is added to in program order)
  __shared__ double s_acc[kGramSolveWaves][ND][3][kWave];
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = __builtin_amdgcn_readfirstlane(int(threadIdx.x) / kWave);
  const int groups = a.gs_groups;
  const int k = part / groups;  // polarization
  const int group = part - k * groups;
  const int n = a.n_ant;
  const DeviceState st = LoadConst(a.state);
  // FIN: the iteration finished here (the decided count, plus the pending one)
  const int f = st.iteration + (FIN ? st.pending : 0);
  const int iteration = FIN ? f + 1 : st.iteration;  // the iteration whose sums are formed
  const double2* cur = (iteration & 1) ? a.sol1 : a.sol0;
  const size_t gbase = size_t(cb) * n * a.n_sub;
  [[maybe_unused]] const int n_wg = a.n_cb * ND;  // (block, direction) partials per half

  const int rows = a.walk_rows;
  const int per = (rows + groups - 1) / groups;
  const int r0 = group * per;
  const int r1 = min(rows, r0 + per);
  const uint32_t row_bytes = uint32_t(NQ) * uint32_t(n) * 8u;
  // the block's terms of polarization k through one buffer resource: a lane's
  // byte offset in a VGPR, the term's offset in an SGPR
  const __amdgpu_buffer_rsrc_t grsrc = PlaneResource(
      a.gram_walk + (size_t(cb) * NC + k) * size_t(rows) * NQ * n, uint32_t(rows) * row_bytes);
  const GlobalT<int32_t>* codes = AsGlobal(a.walk_code);

  // Gains of polarization k of every antenna and direction (one solution
  // each). Their loads are issued first and the first row's cross products
  // right after them: vector loads complete in order, so writing the gains
  // to LDS waits for the gain loads only while the row streams in. With FIN
  // the stored sums the gains are finished from are loaded and added first
  // (the row then streams in during the finishing step).
  double2 gl[kGainEntries][2];
  [[maybe_unused]] double nr[kGainEntries][2] = {}, ni[kGainEntries][2] = {},
                          dn[kGainEntries][2] = {};
  // FIN: the stored current values of the thread's (direction, antenna) pairs
  [[maybe_unused]] double2 fcv[kGainEntries][2];
  [[maybe_unused]] double dsum = 0.0;
  [[maybe_unused]] int dcount = 0;
  if constexpr (FIN) {
    const double2* fcur = (f & 1) ? a.sol1 : a.sol0;
#pragma unroll
    for (int u = 0; u < kGainEntries; ++u) {
      const int d = min(wave + u * kGramSolveWaves, ND - 1);
      const size_t pair = gbase + size_t(lane < n ? lane : 0) * a.n_sub + AsConst(a.seg_sol)[d];
      fcv[u][0] = LoadGlobal(fcur + pair * 2);
      fcv[u][1] = LoadGlobal(fcur + pair * 2 + 1);
    }
    const GlobalT<double>* fsums =
        AsGlobal(GramParts(a, f & 1)) + size_t(cb) * 2 * groups * ND * 3 * kWave;
    double fv[kGainEntries][2][2][3];
    // groups >= 1: the first pass unconditional, so that no path reaches the
    // row loads with these loads (or the current values) still in flight
    LoadGroupSums<ND, kGainEntries>(fsums, groups, 0, wave, lane, fv);
    AddGroupSums<kGainEntries>(fv, groups, 0, nr, ni, dn);
    for (int g0 = 2; g0 < groups; g0 += 2) {
      LoadGroupSums<ND, kGainEntries>(fsums, groups, g0, wave, lane, fv);
      AddGroupSums<kGainEntries>(fv, groups, g0, nr, ni, dn);
    }
    // the pending iteration's convergence partials (half st.iteration & 1),
    // strided over the threads as the decision kernel's 256 threads take them
    if (st.pending) {
      const size_t h = size_t(st.iteration & 1) * n_wg;
      for (int i = threadIdx.x; i < n_wg; i += kGramSolveThreads) {
        dsum += AsGlobal(a.wg_conv)[h + i];
        dcount += AsGlobal(a.wg_conv_n)[h + i];
      }
    }
  } else {
#pragma unroll
    for (int u = 0; u < kGainEntries; ++u) {
      const int e = threadIdx.x + u * kGramSolveThreads;
      const int ant = e & (kWave - 1);
      const int d = (e < ND * kWave ? e : 0) / kWave;
      const size_t pair =
          gbase + size_t(ant < n ? ant : 0) * a.n_sub + AsConst(a.seg_sol)[d];
      gl[u][0] = LoadGlobal(cur + pair * 2);
      gl[u][1] = LoadGlobal(cur + pair * 2 + 1);
    }
  }
  __builtin_amdgcn_sched_barrier(0);
  f2v q[NQ];
  // lanes with a pair in row r: n, or n/2 in the last row for even n
  auto load_row = [&](int r) {
    const int lanes = (2 * (r + 1) == n) ? n / 2 : n;
    const uint32_t off =
        lane < lanes ? uint32_t(r) * row_bytes + uint32_t(lane) * 8u : 0x40000000u;
#pragma unroll
    for (int i = 0; i < NQ; ++i)
      q[i] = __builtin_bit_cast(f2v, __builtin_amdgcn_raw_buffer_load_b64(
                                         grsrc, off, uint32_t(i) * uint32_t(n) * 8u, 0));
  };
  int r = r0 + wave;
  if (r < r1) load_row(r);
  __builtin_amdgcn_sched_barrier(0);
  double mean;
  if constexpr (FIN) {
    __shared__ double s_msum[ND];
    __shared__ int s_mn[ND];
    __shared__ double s_red[kGramSolveWaves];
    __shared__ int s_redc[kGramSolveWaves];
    if (st.pending) {
      BlockSum(dsum, dcount, s_red, s_redc);
      double step_magnitude;
      DeviceState next = DecideState(a, st, dsum, dcount, step_magnitude);
      const bool stop = next.done != 0;
      next.pending = stop ? 0 : 1;  // f is finished below unless the solve stops
      if (threadIdx.x == 0) {
        const int arrived = atomicAdd(a.arrive, 1);
        if (arrived == a.n_cb * 2 * groups - 1) {
          *a.arrive = 0;
          a.step_mags[st.iteration] = step_magnitude;
          *a.state = next;
        }
      }
      if (stop) return;
    } else if (threadIdx.x == 0) {
      // the first launch after the sums of iteration 0: nothing to decide
      const int arrived = atomicAdd(a.arrive, 1);
      if (arrived == a.n_cb * 2 * groups - 1) {
        *a.arrive = 0;
        DeviceState next = st;
        next.pending = 1;
        *a.state = next;
      }
    }
    double2* fnxt = part == 0 ? ((f & 1) ? a.sol0 : a.sol1) : nullptr;
    // the repair mean of f's current solutions: the prologue's for f = 0, else
    // from iteration f - 1's partials as the decision forms it
    double old_mean = AsConst(a.mean)[cb];
    if (f > 0) {
      const size_t h = size_t((f - 1) & 1) * n_wg + size_t(cb) * ND;
      double ms = 0.0;
      int mn = 0;
#pragma unroll
      for (int d = 0; d < ND; ++d) {
        ms += AsConst(a.wg_mean)[h + d];
        mn += AsConst(a.wg_mean_n)[h + d];
      }
      old_mean = mn == 0 ? 1.0 : ms / mn;
    }
#pragma unroll
    for (int u = 0; u < kGainEntries; ++u) {
      const int d = wave + u * kGramSolveWaves;
      if (d >= ND) continue;
      double conv_sum = 0.0, mean_sum = 0.0;
      int conv_n = 0, mean_n = 0;
      gl[u][0] = gl[u][1] = make_double2(0.0, 0.0);
      if (lane < n)
        FinishSolution(a, fcv[u], fnxt, gbase + size_t(lane) * a.n_sub + AsConst(a.seg_sol)[d],
                       old_mean, nr[u], ni[u], dn[u], conv_sum, conv_n, mean_sum, mean_n, gl[u]);
      for (int off = 32; off > 0; off >>= 1) {
        conv_sum += __shfl_xor(conv_sum, off);
        conv_n += __shfl_xor(conv_n, off);
        mean_sum += __shfl_xor(mean_sum, off);
        mean_n += __shfl_xor(mean_n, off);
      }
      if (lane == 0) {
        if (part == 0) {
          const size_t wg = size_t(f & 1) * n_wg + size_t(cb) * ND + d;
          a.wg_conv[wg] = conv_sum;
          a.wg_conv_n[wg] = conv_n;
          a.wg_mean[wg] = mean_sum;
          a.wg_mean_n[wg] = mean_n;
        }
        s_msum[d] = mean_sum;
        s_mn[d] = mean_n;
      }
    }
    __syncthreads();
    // the block's repair mean, as the decision forms it (DecideBlock)
    double ms = 0.0;
    int mn = 0;
#pragma unroll
    for (int d = 0; d < ND; ++d) {
      ms += s_msum[d];
      mn += s_mn[d];
    }
    mean = mn == 0 ? 1.0 : ms / mn;
  } else {
    mean = AsConst(a.mean)[cb];
  }
  // repaired (a pair with a non-finite entry reads as the block mean) and
  // cast to float, as the correlate uses them (SolveData.h:230-235)
  {
#pragma unroll
    for (int u = 0; u < kGainEntries; ++u) {
      const int e = threadIdx.x + u * kGramSolveThreads;
      if (e < ND * kWave) {
        const int ant = e & (kWave - 1);
        double2 g0 = gl[u][0], g1 = gl[u][1];
        if (!(cfinite(g0) && cfinite(g1))) g0 = g1 = make_double2(mean, 0.0);
        const double2 g = k ? g1 : g0;
        s_gain[e / kWave][ant] =
            ant < n ? make_float2(float(g.x), float(g.y)) : make_float2(0.0f, 0.0f);
      }
    }
  }
  float sc[ND];
#pragma unroll
  for (int d = 0; d < ND; ++d) sc[d] = F16 ? AsConst(a.model_scale)[cb * ND + d] : 1.0f;
  __syncthreads();

#pragma unroll
  for (int d = 0; d < ND; ++d)
#pragma unroll
    for (int c = 0; c < 3; ++c) s_acc[wave][d][c][lane] = 0.0;
  // One row at a time (its terms loaded at the top of the loop, the first
  // while the gains were staged): without a second row in flight a wave
  // needs at most 168 registers, so three workgroups share a compute unit
  // and the other waves cover a row's load latency.
  for (; r < r1; r += kGramSolveWaves) {
    if (r != r0 + wave) load_row(r);
    const int step = r + 1;
    const uint32_t slot = uint32_t(r) * kWave + uint32_t(lane);
    const bool first = codes[slot] != 2;  // 1: lane's antenna is A; 0 (no baseline): zeros
    // The own gains are the same every row; reading them through an index the
    // compiler cannot see through keeps them (and their FP64 norms) from being
    // hoisted out of the loop into registers the next row's loads need.
    int own_lane = lane;
    asm volatile("" : "+v"(own_lane));
    const int pa = lane < n ? (lane + step < n ? lane + step : lane + step - n) : lane;
    const int src = lane < n ? (lane >= step ? lane - step : lane - step + n) : lane;
    float2 ggs[ND];
#pragma unroll
    for (int d = 0; d < ND; ++d) {
      const float2 go = s_gain[d][own_lane];
      const float2 gp = s_gain[d][pa];
      // g_A conj(g_B) with the FP16 scale of d: g_own conj(g_partner) for
      // antenna A's lane, its conjugate for B's (selects, no branch)
      const float2 x = cmulc(go, gp);
      ggs[d] = make_float2(x.x * sc[d], (first ? x.y : -x.y) * sc[d]);
    }
    float ss[ND];
#pragma unroll
    for (int d = 0; d < ND; ++d) ss[d] = (d & 1) ? q[WalkDiag<ND>(d)].y : q[WalkDiag<ND>(d)].x;
    // U_d = C_d - sum_e ggs_e G_e,d (FP32, as the apply; G_e,d = conj(G_d,e)
    // for e > d, real for e = d), four FMAs per complex term
    float2 tt[ND];
#pragma unroll
    for (int d = 0; d < ND; ++d) {
      float ux = q[d].x, uy = q[d].y;
#pragma unroll
      for (int e = 0; e < ND; ++e) {
        if (e == d) {
          ux = fmaf(-ggs[e].x, ss[d], ux);
          uy = fmaf(-ggs[e].y, ss[d], uy);
          continue;
        }
        const f2v ge = e < d ? q[WalkOffPair<ND>(e, d)] : q[WalkOffPair<ND>(d, e)];
        const float gi = e < d ? ge.y : -ge.y;
        ux = fmaf(-ggs[e].x, ge.x, ux);
        ux = fmaf(ggs[e].y, gi, ux);
        uy = fmaf(-ggs[e].x, gi, uy);
        uy = fmaf(-ggs[e].y, ge.x, uy);
      }
      tt[d] = make_float2(ux, uy);
    }
    // In units of the FP16 scale: T = sc_d T' with T' = U_d + ggs_d G_d,d and
    // S = sc_d^2 G_d,d, so the numerator sums carry sc_d and the denominator
    // sums sc_d^2, applied once when the sums are stored.
#pragma unroll
    for (int d = 0; d < ND; ++d) {
      const double sd = double(ss[d]);
      double2 t = make_double2(double(tt[d].x) + double(ggs[d].x) * sd,
                               double(tt[d].y) + double(ggs[d].y) * sd);
      // T with the lane's orientation: T for antenna A's lane, conj(T) for B's
      t.y = first ? t.y : -t.y;
      // the gains again from LDS (kept out of registers while the next row loads)
      const float2 go = s_gain[d][own_lane];
      const float2 gp = s_gain[d][pa];
      const double2 o = make_double2(go.x, go.y);
      const double2 p = make_double2(gp.x, gp.y);
      const double2 own = cmul(p, t);    // A: g_B T; B: g_A conj(T)
      const double2 prt = cmulc(o, t);   // partner B: g_A conj(T); partner A: g_B T
      LdsAdd(&s_acc[wave][d][0][lane], own.x + __shfl(prt.x, src));
      LdsAdd(&s_acc[wave][d][1][lane], own.y + __shfl(prt.y, src));
      LdsAdd(&s_acc[wave][d][2][lane], cnorm(p) * sd + __shfl(cnorm(o) * sd, src));
    }
  }

  // the four waves' sums, added in wave order
  __syncthreads();
  const size_t part_stride = size_t(ND) * 3 * kWave;  // doubles per workgroup's sums
  double* mine = GramParts(a, iteration & 1) + (size_t(cb) * 2 * groups + part) * part_stride;
  for (int d = wave; d < ND; d += kGramSolveWaves) {
    const double scd = F16 ? double(AsConst(a.model_scale)[cb * ND + d]) : 1.0;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      double v = s_acc[0][d][c][lane];
#pragma unroll
      for (int w = 1; w < kGramSolveWaves; ++w) v += s_acc[w][d][c][lane];
      mine[(d * 3 + c) * kWave + lane] = v * (c == 2 ? scd * scd : scd);
    }
  }
}

// The end of the fused iteration for one block, after every workgroup of the
// block stored its sums (the kernel boundary makes them visible): per
// direction (one wave each, lane = antenna) the 2 x groups sums are added in
// a fixed order and every solution is finished (FinishSolution), with the
// direction's convergence partials written as the antenna solve does.
__device__ __forceinline__ void GramFinishBlock(const IterArgs& a, const int cb) {
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = int(threadIdx.x) / kWave;
  const int n_waves = int(blockDim.x) / kWave;
  const int groups = a.gs_groups;
  const int n = a.n_ant;
  const int iteration = AsConst(a.state)->iteration;
  const double2* cur = (iteration & 1) ? a.sol1 : a.sol0;
  double2* nxt = (iteration & 1) ? a.sol0 : a.sol1;
  const size_t gbase = size_t(cb) * n * a.n_sub;
  const double mean = AsConst(a.mean)[cb];
  const size_t part_stride = size_t(a.n_dir) * 3 * kWave;
  const GlobalT<double>* all =
      AsGlobal(GramParts(a, iteration & 1)) + size_t(cb) * 2 * groups * part_stride;
  for (int d = wave; d < a.n_dir; d += n_waves) {
    // every sum loaded before the first is added (one round trip, not one per
    // row group), then added in row-group order
    double v[2][kGramSolveMaxGroups][3];
#pragma unroll
    for (int p = 0; p < 2; ++p)
#pragma unroll
      for (int g = 0; g < kGramSolveMaxGroups; ++g)
        if (g < groups) {
          const GlobalT<double>* x =
              all + size_t(p * groups + g) * part_stride + size_t(d) * 3 * kWave + lane;
          v[p][g][0] = x[0];
          v[p][g][1] = x[kWave];
          v[p][g][2] = x[2 * kWave];
        }
    double nr[2] = {0.0, 0.0}, ni[2] = {0.0, 0.0}, dn[2] = {0.0, 0.0};
#pragma unroll
    for (int p = 0; p < 2; ++p)
#pragma unroll
      for (int g = 0; g < kGramSolveMaxGroups; ++g)
        if (g < groups) {
          nr[p] += v[p][g][0];
          ni[p] += v[p][g][1];
          dn[p] += v[p][g][2];
        }
    double conv_sum = 0.0, mean_sum = 0.0;
    int conv_n = 0, mean_n = 0;
    if (lane < n) {
      const size_t pair = gbase + size_t(lane) * a.n_sub + AsConst(a.seg_sol)[d];
      const double2 cv[2] = {LoadGlobal(cur + pair * 2), LoadGlobal(cur + pair * 2 + 1)};
      double2 unused[2];
      FinishSolution(a, cv, nxt, pair, mean, nr, ni, dn, conv_sum, conv_n, mean_sum, mean_n,
                     unused);
    }
    for (int off = 32; off > 0; off >>= 1) {
      conv_sum += __shfl_xor(conv_sum, off);
      conv_n += __shfl_xor(conv_n, off);
      mean_sum += __shfl_xor(mean_sum, off);
      mean_n += __shfl_xor(mean_n, off);
    }
    if (lane == 0) {
      const int wg = cb * a.n_dir + d;
      a.wg_conv[wg] = conv_sum;
      a.wg_conv_n[wg] = conv_n;
      a.wg_mean[wg] = mean_sum;
      a.wg_mean_n[wg] = mean_n;
    }
  }
}

constexpr int kGramFinishThreads = 512;

__global__ void __launch_bounds__(kGramFinishThreads) GramFinishKernel(const IterArgs a) {
  if (AsConst(a.state)->done) return;
  GramFinishBlock(a, blockIdx.x);
}

// grid (2 x groups, n_cb); FIN: the end of the state's iteration, then the
// next iteration's sums
template <int ND, bool F16, bool FIN>
__global__ void __launch_bounds__(kGramSolveThreads, 3) GramSolveKernel(const IterArgs a) {
  if (AsConst(a.state)->done) return;
  GramSolveBlock<ND, F16, FIN>(a, blockIdx.y, blockIdx.x);
}

// Cross products from baseline order [cb][seg 0][k][q][bl_stride] into the
// walk order [cb][k][step row][qw][n] of the fused iteration (zeros where a
// pair has no baseline): walk term qw < n_dir is C_qw, then the
// off-diagonal G_e,d (e < d), then the real parts of the diagonal G_d,d two
// per slot. One workgroup per (block, correlation, walk term), so a term's
// baseline row is read into one XCD's L2 once. Once per solve.
__device__ __forceinline__ void GramPermuteRow(const IterArgs& a, const int cb, const int k,
                                               const int qw) {
  typedef float f2v __attribute__((ext_vector_type(2)));
  const int nd = a.n_dir;
  const int nq = GramTerms(nd);
  const int n = a.n_ant;
  const int nqw = GramWalkTerms(nd);
  const int rows = a.walk_rows;
  const int slots = rows * n;
  const GlobalT<f2v>* src = (const GlobalT<f2v>*)AsGlobal(
      a.gram + (size_t(cb) * a.n_seg * 2 + k) * nq * a.bl_stride);
  GlobalT<f2v>* dst = (GlobalT<f2v>*)AsGlobal(
      a.gram_walk + (size_t(cb) * 2 + k) * size_t(rows) * nqw * n + size_t(qw) * n);
  // source terms (baseline-order index q = nd + d(d+1)/2 + e for G_e,d)
  int q0 = qw, q1 = -1;  // q1 >= 0: a diagonal pair
  const int n_off = nd * (nd - 1) / 2;
  if (qw >= nd + n_off) {
    const int d = 2 * (qw - nd - n_off);
    q0 = nd + d * (d + 1) / 2 + d;
    q1 = d + 1 < nd ? nd + (d + 1) * (d + 2) / 2 + d + 1 : -2;
  } else if (qw >= nd) {
    int d = 1, j = qw - nd;
    while (j >= d) j -= d++;
    q0 = nd + d * (d + 1) / 2 + j;  // G_j,d
  }
  const GlobalT<int32_t>* bl = AsGlobal(a.walk_bl);
  for (int i = threadIdx.x; i < slots; i += blockDim.x) {
    const int r = i / n;
    const int b = bl[r * kWave + (i - r * n)];  // -1 past the last row's pairs
    f2v v = {0.0f, 0.0f};
    if (b >= 0) {
      if (q1 == -1) {
        v = src[size_t(q0) * a.bl_stride + b];
      } else {
        v.x = src[size_t(q0) * a.bl_stride + b].x;
        if (q1 >= 0) v.y = src[size_t(q1) * a.bl_stride + b].x;
      }
    }
    dst[size_t(r) * nqw * n + (i - r * n)] = v;
  }
}

__global__ void __launch_bounds__(256) GramPermuteKernel(const IterArgs a) {
  GramPermuteRow(a, blockIdx.z, blockIdx.y, blockIdx.x);
}

// The repaired value of 2x2 solution `m` (MakeSolutionsFinite4Pol,
// SolverBase.cc:373-405): a matrix with a non-finite entry reads as the
// block's mean |J_00| in all four entries, or as the identity when the block
// had no finite matrix (mean < 0 marks that case).
__device__ __forceinline__ void FixedMatrix(const double2* sol, size_t m, double mean,
                                            double2 (&out)[4]) {
  bool finite = true;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    out[k] = LoadGlobal(sol + m * 4 + k);
    finite = finite && cfinite(out[k]);
  }
  if (finite) return;
  if (mean < 0.0) {
    out[0] = out[3] = make_double2(1.0, 0.0);
    out[1] = out[2] = make_double2(0.0, 0.0);
  } else {
#pragma unroll
    for (int k = 0; k < 4; ++k) out[k] = make_double2(mean, 0.0);
  }
}

// acc += a * b, acc += conj(a) * b and acc -= conj(a) * b (complex64), each
// component as two fused multiply-adds into the accumulator (four VALU
// operations per complex product; the form acc += (p - q) costs six).
__device__ __forceinline__ void CMac(float2& acc, float2 a, float2 b) {
  acc.x = fmaf(a.x, b.x, acc.x);
  acc.x = fmaf(-a.y, b.y, acc.x);
  acc.y = fmaf(a.x, b.y, acc.y);
  acc.y = fmaf(a.y, b.x, acc.y);
}
__device__ __forceinline__ void CMacC(float2& acc, float2 a, float2 b) {
  acc.x = fmaf(a.x, b.x, acc.x);
  acc.x = fmaf(a.y, b.y, acc.x);
  acc.y = fmaf(a.x, b.y, acc.y);
  acc.y = fmaf(-a.y, b.x, acc.y);
}
__device__ __forceinline__ void CMsubC(float2& acc, float2 a, float2 b) {
  acc.x = fmaf(-a.x, b.x, acc.x);
  acc.x = fmaf(-a.y, b.y, acc.x);
  acc.y = fmaf(-a.x, b.y, acc.y);
  acc.y = fmaf(a.y, b.x, acc.y);
}

// 2x2 inverse in FP64 (adjugate over determinant); false when the
// determinant is exactly zero.
__device__ __forceinline__ bool Invert2x2(const double2 (&m)[4], double2 (&inv)[4]) {
  const double2 p = cmul(m[0], m[3]), q = cmul(m[1], m[2]);
  const double2 det = make_double2(p.x - q.x, p.y - q.y);
  if (det.x == 0.0 && det.y == 0.0) return false;
  const double n = cnorm(det);
  const double2 r = make_double2(det.x / n, -det.y / n);
  inv[0] = cmul(m[3], r);
  inv[1] = cmul(make_double2(-m[1].x, -m[1].y), r);
  inv[2] = cmul(make_double2(-m[2].x, -m[2].y), r);
  inv[3] = cmul(m[0], r);
  return true;
}

__device__ __forceinline__ void Mul2x2(const double2 (&a)[4], const double2 (&b)[4],
                                       double2 (&c)[4]) {
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const double2 x = cmul(a[2 * i], b[j]), y = cmul(a[2 * i + 1], b[2 + j]);
      c[2 * i + j] = make_double2(x.x + y.x, x.y + y.y);
    }
}

// Convergence term of one 2x2 solution (SolverBase.cc:269-283): the entries
// of (next - old) old^-1, skipped when old is singular; and |next_00| for the
// next repair mean when next is finite.
__device__ __forceinline__ void AddMatrixTerms(const double2 (&old)[4], const double2 (&next)[4],
                                               double& conv_sum, int& conv_n, double& mean_sum,
                                               int& mean_n) {
  double2 inv[4];
  if (Invert2x2(old, inv)) {
    double2 diff[4], ns[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) diff[k] = make_double2(next[k].x - old[k].x, next[k].y - old[k].y);
    Mul2x2(diff, inv, ns);
    const double sumabs =
        hypot(ns[0].x, ns[0].y) + hypot(ns[1].x, ns[1].y) + hypot(ns[2].x, ns[2].y) +
        hypot(ns[3].x, ns[3].y);
    if (isfinite(sumabs)) {
      conv_sum += sumabs;
      conv_n += 4;
    }
  }
  if (cfinite(next[0]) && cfinite(next[1]) && cfinite(next[2]) && cfinite(next[3])) {
    mean_sum += hypot(next[0].x, next[0].y);
    mean_n += 1;
  }
}

// ---------------------------------------------------------------------------
// Smoothness constraint (SmoothnessConstraint.cc:50-97 with the Gaussian
// kernel of KernelSmoother.h): each (antenna, sub-solution, polarization)
// series of the stepped solutions over the channel blocks is replaced by its
// weighted kernel mean; non-finite values enter with weight 0, and a block
// whose weights sum to 0 becomes NaN. The kernel windows [first, last) and
// the frequency corrections of each output block depend only on the
// frequencies; the host prepares them. The size factor of a series is
// (1 / dd_factor) / antenna_factor. One thread per (output block, series);
// the results go to sm_out, and a second pass per (direction, block) writes
// them back with the convergence partials, which replace antenna_solve's.
// ---------------------------------------------------------------------------
// Solution values the smoothing pass covers (pairs, or 2x2 matrices).
__host__ __device__ __forceinline__ int64_t SmoothElements(const IterArgs& a) {
  return int64_t(a.n_cb) * a.n_ant * a.n_sub * (a.full_jones ? 4 : 2);
}

__device__ __forceinline__ void SmoothElement(const IterArgs& a, const int64_t e) {
  const int n_pol = a.full_jones ? 4 : 2;
  const int n_series = a.n_ant * a.n_sub * n_pol;
  const int i = int(e / n_series);
  const int idx = int(e - int64_t(i) * n_series);
  const int sub = (idx / n_pol) % a.n_sub;
  const int ant = idx / (a.n_sub * n_pol);
  const double2* nxt = (a.state->iteration & 1) ? a.sol0 : a.sol1;
  const double* w = a.sm_weights + (a.sm_sub_weights ? size_t(sub) * a.n_ant * a.n_cb : 0) +
                    size_t(ant) * a.n_cb;
  const double correction = a.sm_corr[i] * ((1.0 / a.sm_dd_factor[sub]) / a.sm_ant_factor[ant]);
  const double fi = a.sm_freq[i];
  double sr = 0.0, si = 0.0, ws = 0.0;
  for (int j = a.sm_window[2 * i]; j < a.sm_window[2 * i + 1]; ++j) {
    const double2 v = nxt[size_t(j) * n_series + idx];
    const bool finite = cfinite(v);
    const double x = ((fi - a.sm_freq[j]) * correction) / a.sm_bandwidth;
    const double k = (a.sm_truncate && (x < -1.0 || x > 1.0)) ? 0.0 : exp(-x * x * 9.0);
    const double wt = k * (finite ? w[j] : 0.0);
    sr += (finite ? v.x : 0.0) * wt;
    si += (finite ? v.y : 0.0) * wt;
    ws += wt;
  }
  const double nan = std::numeric_limits<double>::quiet_NaN();
  a.sm_out[e] = ws == 0.0 ? make_double2(nan, nan) : make_double2(sr / ws, si / ws);
}

__device__ __forceinline__ void SmoothFinishBlock(const IterArgs& a, const int d, const int cb) {
  __shared__ double red[2 * kBlockThreads / 64];
  __shared__ int redc[2 * kBlockThreads / 64];
  const int iteration = a.state->iteration;
  const double2* cur = (iteration & 1) ? a.sol1 : a.sol0;
  double2* nxt = (iteration & 1) ? a.sol0 : a.sol1;
  const int nsd = a.n_sol_dir[d];
  const int s0 = a.sol_start[d];
  const size_t gbase = size_t(cb) * a.n_ant * a.n_sub;
  const double mean = a.mean[cb];
  double conv_sum = 0.0, mean_sum = 0.0;
  int conv_n = 0, mean_n = 0;
  for (int item = threadIdx.x; item < a.n_ant * nsd; item += blockDim.x) {
    const int ant = item / nsd;
    const size_t pair = gbase + size_t(ant) * a.n_sub + s0 + (item - ant * nsd);
    if (a.full_jones) {
      double2 next4[4], old4[4];
      FixedMatrix(cur, pair, mean, old4);
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        next4[k] = a.sm_out[pair * 4 + k];
        nxt[pair * 4 + k] = next4[k];
      }
      AddMatrixTerms(old4, next4, conv_sum, conv_n, mean_sum, mean_n);
      continue;
    }
    double2 next[2], old[2];
#pragma unroll
    for (int p = 0; p < 2; ++p) {
      next[p] = a.sm_out[pair * 2 + p];
      nxt[pair * 2 + p] = next[p];
      old[p] = FixedSolution(cur, pair, p, mean);
    }
    AddPairTerms(old, next, a.scalar != 0, conv_sum, conv_n, mean_sum, mean_n);
  }
  BlockSum2(conv_sum, conv_n, mean_sum, mean_n, red, redc);
  const int wg = cb * a.n_dir + d;
  if (threadIdx.x == 0) {
    a.wg_conv[wg] = conv_sum;
    a.wg_conv_n[wg] = conv_n;
    a.wg_mean[wg] = mean_sum;
    a.wg_mean_n[wg] = mean_n;
  }
}

constexpr int kSmoothThreads = 256;

__global__ void __launch_bounds__(kSmoothThreads) SmoothKernel(const IterArgs a) {
  if (a.state->done) return;
  const int64_t n = SmoothElements(a);
  for (int64_t e = int64_t(blockIdx.x) * kSmoothThreads + threadIdx.x; e < n;
       e += int64_t(gridDim.x) * kSmoothThreads)
    SmoothElement(a, e);
}

__global__ void __launch_bounds__(kBlockThreads) SmoothFinishKernel(const IterArgs a) {
  if (a.state->done) return;
  SmoothFinishBlock(a, blockIdx.x, blockIdx.y);
}

// Batched forms: blockIdx.y selects the interval (SmoothKernel), the
// antenna-solve items select (interval, block, direction) (finish).
__global__ void __launch_bounds__(kSmoothThreads)
    BatchSmoothKernel(const IterArgs* __restrict__ args) {
  const IterArgs& a = args[blockIdx.y];
  if (!a.sm_freq || a.state->done) return;
  const int64_t n = SmoothElements(a);
  for (int64_t e = int64_t(blockIdx.x) * kSmoothThreads + threadIdx.x; e < n;
       e += int64_t(gridDim.x) * kSmoothThreads)
    SmoothElement(a, e);
}

__global__ void __launch_bounds__(kBlockThreads)
    BatchSmoothFinishKernel(const IterArgs* __restrict__ args, const BatchTile* __restrict__ items) {
  const BatchTile t = items[blockIdx.x];
  const IterArgs& a = args[t.interval];
  if (!a.sm_freq || a.state->done) return;
  SmoothFinishBlock(a, t.seg, t.cb);  // seg holds the direction
}

// One workgroup: the iteration's decision from the per-workgroup partials of
// antenna_solve (the kernel boundary makes them visible without fences).
__device__ __forceinline__ void DecideBlock(const IterArgs& a) {
  __shared__ double red[kBlockThreads / 64];
  __shared__ int redc[kBlockThreads / 64];
  const int n_wg = a.n_cb * a.n_dir;
  // The state, the first convergence terms and the first block's mean terms
  // are loaded together (one round trip for the common case of one block per
  // thread and at most 8 directions); the terms are added in the same order
  // in any case.
  DeviceState st;
  if (threadIdx.x == 0) st = *a.state;
  constexpr int kMeanBatch = 8;
  const int c0 = threadIdx.x;
  double mv0[kMeanBatch];
  int nv0[kMeanBatch];
  if (c0 < a.n_cb) {
#pragma unroll
    for (int j = 0; j < kMeanBatch; ++j)
      if (j < a.n_dir) {
        mv0[j] = a.wg_mean[c0 * a.n_dir + j];
        nv0[j] = a.wg_mean_n[c0 * a.n_dir + j];
      }
  }
  double sum = 0.0;
  int count = 0;
  for (int i = threadIdx.x; i < n_wg; i += blockDim.x) {
    sum += a.wg_conv[i];
    count += a.wg_conv_n[i];
  }
  BlockSum(sum, count, red, redc);
  for (int c = c0; c < a.n_cb; c += blockDim.x) {
    double ms = 0.0;
    int mn = 0;
    // the block's terms loaded kMeanBatch at a time, added in direction order
    for (int d0 = 0; d0 < a.n_dir; d0 += kMeanBatch) {
      double mv[kMeanBatch];
      int nv[kMeanBatch];
#pragma unroll
      for (int j = 0; j < kMeanBatch; ++j)
        if (d0 + j < a.n_dir) {
          if (c == c0 && d0 == 0) {
            mv[j] = mv0[j];
            nv[j] = nv0[j];
          } else {
            mv[j] = a.wg_mean[c * a.n_dir + d0 + j];
            nv[j] = a.wg_mean_n[c * a.n_dir + d0 + j];
          }
        }
#pragma unroll
      for (int j = 0; j < kMeanBatch; ++j)
        if (d0 + j < a.n_dir) {
          ms += mv[j];
          mn += nv[j];
        }
    }
    // no finite solution: 1 for both polarizations (SolverBase.cc:356-358),
    // the identity for 2x2 solutions (:390-394), marked by -1
    a.mean[c] = mn == 0 ? (a.full_jones ? -1.0 : 1.0) : ms / mn;
  }
  if (threadIdx.x == 0) {
    double step_magnitude;
    st = DecideState(a, st, sum, count, step_magnitude);
    a.step_mags[st.iteration - 1] = step_magnitude;
    *a.state = st;
  }
}

__global__ void __launch_bounds__(kBlockThreads) DecideKernel(const IterArgs a) {
  if (a.state->done) return;
  DecideBlock(a);
}

// Resets the solve state and computes the repair means of the initial
// solutions (the first MakeSolutionsFinite2Pol of the loop); one workgroup
// per channel block.
__device__ __forceinline__ void PrologueBlock(const IterArgs& a, const int cb) {
  __shared__ double red[2 * kBlockThreads / 64];
  __shared__ int redc[2 * kBlockThreads / 64];
  if (cb == 0 && threadIdx.x == 0) {
    DeviceState st = {};
    *a.state = st;
    if (a.arrive) *a.arrive = 0;
  }
  const int n_pairs = a.n_ant * a.n_sub;
  const int n_pol = a.full_jones ? 4 : 2;
  const double2* s = a.sol0 + size_t(cb) * n_pairs * n_pol;
  double sum = 0.0, unused = 0.0;
  int count = 0, unused_n = 0;
  for (int i = threadIdx.x; i < n_pairs; i += blockDim.x) {
    bool finite = true;
    for (int k = 0; k < n_pol; ++k) finite = finite && cfinite(s[n_pol * i + k]);
    if (finite) {
      sum += hypot(s[n_pol * i].x, s[n_pol * i].y);
      ++count;
    }
  }
  BlockSum2(sum, count, unused, unused_n, red, redc);
  if (threadIdx.x == 0) a.mean[cb] = count == 0 ? (a.full_jones ? -1.0 : 1.0) : sum / count;
}

__global__ void __launch_bounds__(kBlockThreads) PrologueKernel(const IterArgs a) {
  PrologueBlock(a, blockIdx.x);
}

// chi^2 = sum over blocks, visibilities and correlations of
// |D - sum_d Diag(g_a) M_d Diag(g_b)^H|^2 in FP64 (gains unrepaired, as
// returned by the solver). Grid (bl tiles, blocks); one partial per block.
template <int NC, bool F16>
__global__ void __launch_bounds__(256) Chi2Kernel(const IterArgs a, const double2* sol,
                                                  double* partial) {
  __shared__ double red[256 / 64];
  const int tile = blockIdx.x;
  const int cb = blockIdx.y;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x / 64;
  const int n_waves = blockDim.x / 64;
  const int bl = tile * 64 + lane;
  const CbInfo info = a.cbinfo[cb];
  const size_t plane = size_t(a.n_t) * info.n_ch * a.bl_stride;
  double acc = 0.0;
  if (bl < a.n_bl) {
    const size_t gbase = size_t(cb) * a.n_ant * a.n_sub;
    const int a1 = a.ant1[bl], a2 = a.ant2[bl];
    for (int seg = 0; seg < a.n_seg; ++seg) {
      const int r0 = a.seg_t0[seg] * info.n_ch, r1 = a.seg_t0[seg + 1] * info.n_ch;
      for (int row = r0 + wave; row < r1; row += n_waves) {
        const size_t e = size_t(row) * a.bl_stride + bl;
        float2 df[NC];
        LoadVis<NC, false>(a.data, info.data_off + e, df);
        double2 r[NC];
#pragma unroll
        for (int k = 0; k < NC; ++k) r[k] = make_double2(df[k].x, df[k].y);
        for (int d = 0; d < a.n_dir; ++d) {
          const int s = a.seg_sol[seg * a.n_dir + d];
          float2 mf[NC];
          LoadVis<NC, F16>(a.model, info.model_off + d * plane + e, mf);
          const double sc = F16 ? double(a.model_scale[cb * a.n_dir + d]) : 1.0;
#pragma unroll
          for (int k = 0; k < NC; ++k) {
            const double2 ga = sol[(gbase + size_t(a1) * a.n_sub + s) * 2 + PolI<NC>(k)];
            const double2 gb = sol[(gbase + size_t(a2) * a.n_sub + s) * 2 + PolJ<NC>(k)];
            const double2 m = make_double2(sc * mf[k].x, sc * mf[k].y);
            const double2 c = cmulc(cmul(ga, m), gb);
            r[k].x -= c.x;
            r[k].y -= c.y;
          }
        }
#pragma unroll
        for (int k = 0; k < NC; ++k) acc += cnorm(r[k]);
      }
    }
  }
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
  if (lane == 0) red[wave] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < n_waves; ++w) t += red[w];
    partial[size_t(cb) * gridDim.x + tile] = t;
  }
}

__global__ void SumPartialsKernel(const double* partial, int n, double factor, double* out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    double t = 0.0;
    for (int i = 0; i < n; ++i) t += partial[i];
    *out = t * factor;
  }
}

// ---------------------------------------------------------------------------
// Full-Jones solver (IterativeFullJonesSolver.cc:18-224): 2x2 solutions
// [n_cb][n_ant][n_sub][4], row-major (00, 01, 10, 11), Full data. For
// direction d of baseline (a, b), with A = J_a M, B = J_b M^H and R_all the
// residual with every direction subtracted, the reference's per-visibility
// sums (:143-176) over the residual R_d = R_all + A J_b^H become
//   antenna a: numerator = sum R_d B   = X + J_a D_a,  denominator = D_a
//   antenna b: numerator = sum R_d^H A = Y + J_b D_b,  denominator = D_b
// with X = sum R_all B, Y = sum R_all^H A, D_a = sum B^H B, D_b = sum A^H A
// (M J_b^H = B^H). The gains are constant over a baseline's rows within a
// segment, so the correlate sums X and Y per baseline (16 floats), and the
// denominators follow from model products Q made once per solve:
//   D_a,ij = sum_kl (J_b^H J_b)_kl Q[2i+k][2j+l],
//   D_b,ij = sum_kl (J_a^H J_a)_kl Q[2l+j][2k+i],   Q[p][q] = sum M_p conj(M_q)
// (p, q index the four entries of M). The antenna solve adds them over the
// antenna's baselines in FP64 and forms J + X D^-1 = numerator *
// denominator^-1 (:178-185).
// ---------------------------------------------------------------------------
constexpr int kFullJonesGroup = kFullJonesGroupSize;  // directions accumulated per workgroup

// The four correlations of visibility `idx` of a Full plane, FP16 planes
// scaled back.
template <bool F16>
__device__ __forceinline__ void LoadFull(const void* base, size_t idx, float sc, float2 (&v)[4]) {
  LoadVis<4, F16>(base, idx, v);
  if constexpr (F16) {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      v[k].x *= sc;
      v[k].y *= sc;
    }
  }
}

// One lane's 2x2 gain in the correlate's LDS, entries kWave words apart: the
// layout [direction][J_a, J_b entry][lane] has the 64 lanes of a read on
// consecutive 8-byte words (with the entries of a lane adjacent, lanes 64 B
// apart, a read touched two bank groups).
struct LaneGains {
  const float2* p;
  __device__ __forceinline__ float2 operator[](int k) const { return p[k * kWave]; }
};

// R -= (J_a M) J_b^H; returns A = J_a M.
__device__ __forceinline__ void SubtractDirection(const LaneGains ja, const LaneGains jb,
                                                  const float2 (&m)[4], float2 (&r)[4],
                                                  float2 (&a)[4]) {
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      a[2 * i + j] = make_float2(0.0f, 0.0f);
      CMac(a[2 * i + j], ja[2 * i], m[j]);
      CMac(a[2 * i + j], ja[2 * i + 1], m[2 + j]);
    }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      // (A J_b^H)_ij = A_i0 conj(Jb_j0) + A_i1 conj(Jb_j1) = sum conj(Jb_jk) A_ik
      CMsubC(r[2 * i + j], jb[2 * j], a[2 * i]);
      CMsubC(r[2 * i + j], jb[2 * j + 1], a[2 * i + 1]);
    }
}

// correlate: grid (bl tiles x direction groups, segments, channel blocks),
// block = waves x 64, lane = baseline. Every workgroup forms the full
// residual over all directions and accumulates X and Y of its group's (up to
// NG) directions, combined across waves through LDS in a fixed order.
template <int NG, bool F16>
__device__ __forceinline__ void FullJonesCorrelateTile(const IterArgs& a, const int tile,
                                                       const int group, const int seg,
                                                       const int cb) {
  const int g0 = group * NG;
  const int ng = min(NG, a.n_dir - g0);
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = threadIdx.x / kWave;
  const int n_waves = blockDim.x / kWave;
  const CbInfo info = LoadConst(a.cbinfo + cb);
  const size_t plane = size_t(a.n_t) * info.n_ch * a.bl_stride;
  const int bl = tile * kWave + lane;

  extern __shared__ float smem[];
  float2* g_s = reinterpret_cast<float2*>(smem);  // [n_dir][8][64]: J_a, then J_b
  float* red_s = smem + size_t(a.n_dir) * kWave * 16;  // [16 * NG][64]

  // Stage the repaired, float-cast gains of every direction (SolveData.h
  // casts the solutions to float); waves split the directions.
  {
    const size_t gbase = size_t(cb) * a.n_ant * a.n_sub;
    const double2* sol = (AsConst(a.state)->iteration & 1) ? a.sol1 : a.sol0;
    const double mean = AsConst(a.mean)[cb];
    const bool valid = bl < a.n_bl;
    const int a1 = valid ? a.ant1[bl] : 0;
    const int a2 = valid ? a.ant2[bl] : 0;
    for (int d = wave; d < a.n_dir; d += n_waves) {
      const int s = AsConst(a.seg_sol)[seg * a.n_dir + d];
      double2 j1[4], j2[4];
      FixedMatrix(sol, gbase + size_t(a1) * a.n_sub + s, mean, j1);
      FixedMatrix(sol, gbase + size_t(a2) * a.n_sub + s, mean, j2);
      float2* dst = g_s + size_t(d) * 8 * kWave + lane;
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        dst[k * kWave] =
            valid ? make_float2(float(j1[k].x), float(j1[k].y)) : make_float2(0.0f, 0.0f);
        dst[(4 + k) * kWave] =
            valid ? make_float2(float(j2[k].x), float(j2[k].y)) : make_float2(0.0f, 0.0f);
      }
    }
  }
  __syncthreads();

  float2 x[NG][4], y[NG][4];
#pragma unroll
  for (int j = 0; j < NG; ++j)
#pragma unroll
    for (int k = 0; k < 4; ++k) x[j][k] = y[j][k] = make_float2(0.0f, 0.0f);
  float sc[NG];
#pragma unroll
  for (int j = 0; j < NG; ++j)
    sc[j] = (F16 && j < ng) ? AsConst(a.model_scale)[cb * a.n_dir + g0 + j] : 1.0f;

  const int r0 = AsConst(a.seg_t0)[seg] * info.n_ch;
  const int r1 = AsConst(a.seg_t0)[seg + 1] * info.n_ch;
  for (int row = r0 + wave; row < r1; row += n_waves) {
    const size_t e = size_t(row) * a.bl_stride + bl;
    float2 r[4];
    LoadVis<4, false>(a.data, info.data_off + e, r);
    float2 m[NG][4];
#pragma unroll
    for (int j = 0; j < NG; ++j)
      if (j < ng) LoadFull<F16>(a.model, info.model_off + (g0 + j) * plane + e, sc[j], m[j]);
    // The group's directions, then the others (IterativeFullJonesSolver.cc:
    // 88-89 subtracts every direction from the data).
#pragma unroll
    for (int j = 0; j < NG; ++j) {
      if (j < ng) {
        const LaneGains gj{g_s + size_t(g0 + j) * 8 * kWave + lane};
        float2 av[4];
        SubtractDirection(gj, LaneGains{gj.p + 4 * kWave}, m[j], r, av);
      }
    }
    for (int d = 0; d < a.n_dir; ++d) {
      if (d >= g0 && d < g0 + ng) continue;
      const float s = F16 ? AsConst(a.model_scale)[cb * a.n_dir + d] : 1.0f;
      float2 mo[4];
      LoadFull<F16>(a.model, info.model_off + d * plane + e, s, mo);
      const LaneGains gd{g_s + size_t(d) * 8 * kWave + lane};
      float2 av[4];
      SubtractDirection(gd, LaneGains{gd.p + 4 * kWave}, mo, r, av);
    }
#pragma unroll
    for (int j = 0; j < NG; ++j) {
      if (j >= ng) continue;
      const LaneGains ja{g_s + size_t(g0 + j) * 8 * kWave + lane};
      const LaneGains jb{ja.p + 4 * kWave};
      float2 av[4], bv[4];
      // A = J_a M
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int k = 0; k < 2; ++k) {
          av[2 * i + k] = make_float2(0.0f, 0.0f);
          CMac(av[2 * i + k], ja[2 * i], m[j][k]);
          CMac(av[2 * i + k], ja[2 * i + 1], m[j][2 + k]);
        }
      // B = J_b M^H: B_ik = Jb_i0 conj(M_k0) + Jb_i1 conj(M_k1)
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int k = 0; k < 2; ++k) {
          float2 t = make_float2(0.0f, 0.0f);
          CMacC(t, m[j][2 * k], jb[2 * i]);
          CMacC(t, m[j][2 * k + 1], jb[2 * i + 1]);
          bv[2 * i + k] = t;
        }
      // X += R B; Y += R^H A
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int k = 0; k < 2; ++k) {
          CMac(x[j][2 * i + k], r[2 * i], bv[k]);
          CMac(x[j][2 * i + k], r[2 * i + 1], bv[2 + k]);
          CMacC(y[j][2 * i + k], r[i], av[k]);
          CMacC(y[j][2 * i + k], r[2 + i], av[2 + k]);
        }
    }
  }

  auto term = [&](int j, int v) -> float& {
    if (v < 8) return (v & 1) ? x[j][v >> 1].y : x[j][v >> 1].x;
    return (v & 1) ? y[j][(v - 8) >> 1].y : y[j][(v - 8) >> 1].x;
  };
  // Combine the waves: wave w-1 += wave w, for w = n_waves-1 .. 1.
  for (int w = n_waves - 1; w >= 1; --w) {
    if (wave == w) {
#pragma unroll
      for (int j = 0; j < NG; ++j)
#pragma unroll
        for (int v = 0; v < kFullJonesTerms; ++v)
          red_s[(j * kFullJonesTerms + v) * kWave + lane] = term(j, v);
    }
    __syncthreads();
    if (wave == w - 1) {
#pragma unroll
      for (int j = 0; j < NG; ++j)
#pragma unroll
        for (int v = 0; v < kFullJonesTerms; ++v)
          term(j, v) += red_s[(j * kFullJonesTerms + v) * kWave + lane];
    }
    __syncthreads();
  }
  if (wave == 0) {
#pragma unroll
    for (int j = 0; j < NG; ++j) {
      if (j >= ng) continue;
      float* out = a.partials +
                   ((size_t(cb) * a.n_seg + seg) * a.n_dir + g0 + j) * kFullJonesTerms * a.bl_stride;
#pragma unroll
      for (int v = 0; v < kFullJonesTerms; ++v) out[size_t(v) * a.bl_stride + bl] = term(j, v);
    }
  }
}

// Single solve: grid (bl tiles x direction groups, segments, blocks).
template <int NG, bool F16>
__global__ void __launch_bounds__(kCorrelateMaxThreads) FullJonesCorrelateKernel(const IterArgs a) {
  if (AsConst(a.state)->done) return;
  const int n_tiles = a.bl_stride / kWave;
  const int group = blockIdx.x / n_tiles;
  FullJonesCorrelateTile<NG, F16>(a, blockIdx.x - group * n_tiles, group, blockIdx.y, blockIdx.z);
}

// Batch: one workgroup per tile-list entry, whose `tile` field holds
// tile + group * (bl_stride / 64) of its interval.
template <int NG, bool F16>
__global__ void __launch_bounds__(kCorrelateMaxThreads)
    BatchFullJonesCorrelateKernel(const IterArgs* __restrict__ args,
                                  const BatchTile* __restrict__ tiles) {
  const BatchTile t = LoadConst(tiles + blockIdx.x);
  const IterArgs a = LoadConst(args + t.interval);
  if (AsConst(a.state)->done) return;
  const int n_tiles = a.bl_stride / kWave;
  const int group = t.tile / n_tiles;
  FullJonesCorrelateTile<NG, F16>(a, t.tile - group * n_tiles, group, t.seg, t.cb);
}

// Model products, once per solve: Q[p][q] = sum over a segment's rows of
// M_p conj(M_q) per (block, segment, direction, baseline), Hermitian in
// (p, q): the four real diagonal terms, then (p < q) in the order (0,1),
// (0,2), (0,3), (1,2), (1,3), (2,3) as (re, im). Grid (bl tiles, segments,
// blocks x directions), lane = baseline; waves split the rows and are
// combined through LDS in a fixed order.
template <bool F16>
__device__ __forceinline__ void FullJonesModelProductsTile(const IterArgs& a, const int tile,
                                                           const int seg, const int cb,
                                                           const int d) {
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = threadIdx.x / kWave;
  const int n_waves = blockDim.x / kWave;
  const CbInfo info = LoadConst(a.cbinfo + cb);
  const size_t plane = size_t(a.n_t) * info.n_ch * a.bl_stride;
  const int bl = tile * kWave + lane;
  const float sc = F16 ? AsConst(a.model_scale)[cb * a.n_dir + d] : 1.0f;
  __shared__ float red_s[kFullJonesModelTerms * kWave];
  float q[kFullJonesModelTerms];
#pragma unroll
  for (int v = 0; v < kFullJonesModelTerms; ++v) q[v] = 0.0f;
  const int r0 = AsConst(a.seg_t0)[seg] * info.n_ch;
  const int r1 = AsConst(a.seg_t0)[seg + 1] * info.n_ch;
  for (int row = r0 + wave; row < r1; row += n_waves) {
    float2 m[4];
    LoadFull<F16>(a.model, info.model_off + d * plane + size_t(row) * a.bl_stride + bl, sc, m);
#pragma unroll
    for (int p = 0; p < 4; ++p) q[p] += m[p].x * m[p].x + m[p].y * m[p].y;
    int v = 4;
#pragma unroll
    for (int p = 0; p < 4; ++p)
#pragma unroll
      for (int r = p + 1; r < 4; ++r) {
        // M_p conj(M_r)
        q[v] += m[p].x * m[r].x + m[p].y * m[r].y;
        q[v + 1] += m[p].y * m[r].x - m[p].x * m[r].y;
        v += 2;
      }
  }
  for (int w = n_waves - 1; w >= 1; --w) {
    if (wave == w) {
#pragma unroll
      for (int v = 0; v < kFullJonesModelTerms; ++v) red_s[v * kWave + lane] = q[v];
    }
    __syncthreads();
    if (wave == w - 1) {
#pragma unroll
      for (int v = 0; v < kFullJonesModelTerms; ++v) q[v] += red_s[v * kWave + lane];
    }
    __syncthreads();
  }
  if (wave == 0) {
    float* out = a.fj_products +
                 ((size_t(cb) * a.n_seg + seg) * a.n_dir + d) * kFullJonesModelTerms * a.bl_stride;
#pragma unroll
    for (int v = 0; v < kFullJonesModelTerms; ++v) out[size_t(v) * a.bl_stride + bl] = q[v];
  }
}

template <bool F16>
__global__ void __launch_bounds__(kCorrelateMaxThreads)
    FullJonesModelProductsKernel(const IterArgs a) {
  const int cb = blockIdx.z / a.n_dir;
  FullJonesModelProductsTile<F16>(a, blockIdx.x, blockIdx.y, cb, blockIdx.z - cb * a.n_dir);
}

// Batch: one workgroup per (interval, block, segment, tile) entry, over every
// direction of its interval in turn.
template <bool F16>
__global__ void __launch_bounds__(kCorrelateMaxThreads)
    BatchFullJonesModelProductsKernel(const IterArgs* __restrict__ args,
                                      const BatchTile* __restrict__ tiles) {
  const BatchTile t = LoadConst(tiles + blockIdx.x);
  const IterArgs a = LoadConst(args + t.interval);
  for (int d = 0; d < a.n_dir; ++d) {
    FullJonesModelProductsTile<F16>(a, t.tile, t.seg, t.cb, d);
    __syncthreads();  // the next direction reuses the reduction LDS
  }
}

// ---------------------------------------------------------------------------
// Full-Jones Gram iteration. The gains are constant over a baseline's rows in
// a segment, so with P_d[ij][km] = sum_v R_all,ij conj(M_d,km) (R_all the
// residual with every direction subtracted; entry index 2i+j of the 2x2
// matrices) the correlate's sums are
//   X_d,ik = sum_jm Jb_d,jm P_d[ij][km],   Y_d,ik = sum_jm Ja_d,jm conj(P_d[ji][mk])
// and P_d follows from cross products made once per solve:
//   P_d[ij][km] = C_d[ij][km] - sum_e sum_no Ja_e,in conj(Jb_e,jo) G_e,d[no][km],
//   C_d[p][q] = sum_v D_p conj(M_d,q),   G_e,d[p][q] = sum_v M_e,p conj(M_d,q)
// (G_d,d is the model product Q_d; G_e,d = conj(G_d,e) transposed). With the
// gains contracted first, per pair (e, d):
//   X_d,ik -= sum_n Ja_e,in V[n][k],   V[n][k] = sum_om (Jb_e^H Jb_d)_om G_e,d[no][km]
//   Y_d,ik -= sum_o Jb_e,io W[o][k],   W[o][k] = sum_nm (Ja_e^H Ja_d)_nm conj(G_e,d[no][mk])
// and the data term adds sum_jm Jb_d,jm C_d[ij][km] to X and
// sum_jm Ja_d,jm conj(C_d[ji][mk]) to Y. The apply writes X and Y in the
// correlate's partials layout, so the full-Jones antenna solve and the
// decision are those of the streaming iteration.
// Cross products [n_cb][n_seg][FullJonesGramTerms(n_dir)][bl_stride]
// complex64 in blocks of 16 terms (term 4p + q: entry p of the first factor,
// entry q of the conjugated model): C_d at block d, then G_e,d (e < d) at
// block n_dir + d(d-1)/2 + e. FP16 planes are scaled back as they are read.
// ---------------------------------------------------------------------------
// One visibility of a Full plane through L2 (plain loads: several waves of the
// workgroup read the same rows), FP16 planes scaled back.
template <bool F16>
__device__ __forceinline__ void LoadFullShared(const void* base, size_t idx, float sc,
                                               float2 (&v)[4]) {
  if constexpr (!F16) {
    const float4_v* p = reinterpret_cast<const float4_v*>(base) + idx * 2;
    const float4_v x0 = p[0], x1 = p[1];
    v[0] = make_float2(x0.x, x0.y);
    v[1] = make_float2(x0.z, x0.w);
    v[2] = make_float2(x1.x, x1.y);
    v[3] = make_float2(x1.z, x1.w);
  } else {
    const uint4_v u = reinterpret_cast<const uint4_v*>(base)[idx];
    v[0] = HalfBitsToFloat2(u.x);
    v[1] = HalfBitsToFloat2(u.y);
    v[2] = HalfBitsToFloat2(u.z);
    v[3] = HalfBitsToFloat2(u.w);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      v[k].x *= sc;
      v[k].y *= sc;
    }
  }
}

// Cross products of one (block, segment, 64-baseline tile), lane = baseline.
// The n_dir(n_dir+1)/2 product blocks in the order (target d, source) =
// (0, D), (1, D), (1, M_0), (2, D), (2, M_0), (2, M_1), ... are dealt out
// three per wave; every wave runs over all of the segment's rows (the rows of
// the planes it shares with other waves come from L2).
template <bool F16>
__device__ __forceinline__ void FullJonesGramTile(const IterArgs& a, const int tile, const int seg,
                                                  const int cb) {
  constexpr int NB = kFjGramBlocksPerWave;
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = __builtin_amdgcn_readfirstlane(int(threadIdx.x) / kWave);
  const int nd = a.n_dir;
  const int n_blocks = nd * (nd + 1) / 2;
  const CbInfo info = LoadConst(a.cbinfo + cb);
  const size_t plane = size_t(a.n_t) * info.n_ch * a.bl_stride;
  const int bl = tile * kWave + lane;
  int tgt[NB], src[NB], pos[NB];
  float sct[NB], scs[NB];
#pragma unroll
  for (int k = 0; k < NB; ++k) {
    const int b = min(wave * NB + k, n_blocks - 1);  // past the last block: unused
    int d = 0;
    while ((d + 1) * (d + 2) / 2 <= b) ++d;
    tgt[k] = d;
    src[k] = b - d * (d + 1) / 2 - 1;  // -1: the data
    pos[k] = src[k] < 0 ? d : nd + d * (d - 1) / 2 + src[k];
    sct[k] = F16 ? AsConst(a.model_scale)[cb * nd + d] : 1.0f;
    scs[k] = (F16 && src[k] >= 0) ? AsConst(a.model_scale)[cb * nd + src[k]] : 1.0f;
  }
  float2 acc[NB][16];
#pragma unroll
  for (int k = 0; k < NB; ++k)
#pragma unroll
    for (int t = 0; t < 16; ++t) acc[k][t] = make_float2(0.0f, 0.0f);
  const int r0 = AsConst(a.seg_t0)[seg] * info.n_ch;
  const int r1 = AsConst(a.seg_t0)[seg + 1] * info.n_ch;
  for (int row = r0; row < r1; ++row) {
    const size_t e = size_t(row) * a.bl_stride + bl;
#pragma unroll
    for (int k = 0; k < NB; ++k) {
      if (wave * NB + k >= n_blocks) continue;
      float2 mt[4], ms[4];
      LoadFullShared<F16>(a.model, info.model_off + tgt[k] * plane + e, sct[k], mt);
      if (src[k] < 0)
        LoadFullShared<false>(a.data, info.data_off + e, 1.0f, ms);
      else
        LoadFullShared<F16>(a.model, info.model_off + src[k] * plane + e, scs[k], ms);
#pragma unroll
      for (int p = 0; p < 4; ++p)
#pragma unroll
        for (int q = 0; q < 4; ++q) CMacC(acc[k][4 * p + q], mt[q], ms[p]);
    }
  }
  const size_t base = (size_t(cb) * a.n_seg + seg) * FullJonesGramTerms(nd);
#pragma unroll
  for (int k = 0; k < NB; ++k) {
    if (wave * NB + k >= n_blocks) continue;
    float2* out = a.gram + (base + size_t(pos[k]) * 16) * a.bl_stride + bl;
#pragma unroll
    for (int t = 0; t < 16; ++t) out[size_t(t) * a.bl_stride] = acc[k][t];
  }
}

// At most FjGramWaves(8) = 12 waves per workgroup.
constexpr int kFjGramMaxThreads = 12 * kWave;

// Single solve: grid (bl tiles, segments, blocks), FjGramWaves(n_dir) waves.
template <bool F16>
__global__ void __launch_bounds__(kFjGramMaxThreads) FullJonesGramKernel(const IterArgs a) {
  FullJonesGramTile<F16>(a, blockIdx.x, blockIdx.y, blockIdx.z);
}

// Batch: one workgroup per (interval, block, segment, tile) entry.
template <bool F16>
__global__ void __launch_bounds__(kFjGramMaxThreads)
    BatchFullJonesGramKernel(const IterArgs* __restrict__ args, const BatchTile* __restrict__ tiles) {
  const BatchTile t = LoadConst(tiles + blockIdx.x);
  const IterArgs a = LoadConst(args + t.interval);
  if (int(threadIdx.x) >= FjGramWaves(a.n_dir) * kWave) return;
  FullJonesGramTile<F16>(a, t.tile, t.seg, t.cb);
}

// A complex64 value through a global load (8 bytes).
__device__ __forceinline__ float2 LoadGlobalF2(const float2* p) {
  typedef float f2v __attribute__((ext_vector_type(2)));
  const f2v v = *(const GlobalT<f2v>*)p;
  return make_float2(v.x, v.y);
}

// The apply: X and Y of one (block, segment, 64-baseline tile), lane =
// baseline; one wave per target direction d (n_dir waves). Each stored block
// G_e,d (e < d) serves two waves, d's and e's: the waves take their sources e
// in the rounds of a round-robin pairing (round r pairs a and b with
// a + b = 2r mod (m - 1) for the m = n_dir rounded up to even players, the
// last player with r), so both readers of a block load it in the same round
// and the second load hits in cache. The gains are staged in LDS as the
// correlate stages them.
__device__ __forceinline__ void FullJonesApplyGramTile(const IterArgs& a, const int tile,
                                                       const int seg, const int cb) {
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = threadIdx.x / kWave;
  const int n_waves = blockDim.x / kWave;
  const int nd = a.n_dir;
  const int bl = tile * kWave + lane;
  extern __shared__ float smem[];
  float2* g_s = reinterpret_cast<float2*>(smem);  // [n_dir][8][64]: J_a, then J_b
  {
    const size_t gbase = size_t(cb) * a.n_ant * a.n_sub;
    const double2* sol = (AsConst(a.state)->iteration & 1) ? a.sol1 : a.sol0;
    const double mean = AsConst(a.mean)[cb];
    const bool valid = bl < a.n_bl;
    const int a1 = valid ? a.ant1[bl] : 0;
    const int a2 = valid ? a.ant2[bl] : 0;
    for (int d = wave; d < nd; d += n_waves) {
      const int sidx = AsConst(a.seg_sol)[seg * nd + d];
      double2 j1[4], j2[4];
      FixedMatrix(sol, gbase + size_t(a1) * a.n_sub + sidx, mean, j1);
      FixedMatrix(sol, gbase + size_t(a2) * a.n_sub + sidx, mean, j2);
      float2* dst = g_s + size_t(d) * 8 * kWave + lane;
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        dst[k * kWave] =
            valid ? make_float2(float(j1[k].x), float(j1[k].y)) : make_float2(0.0f, 0.0f);
        dst[(4 + k) * kWave] =
            valid ? make_float2(float(j2[k].x), float(j2[k].y)) : make_float2(0.0f, 0.0f);
      }
    }
  }
  __syncthreads();
  const size_t base = (size_t(cb) * a.n_seg + seg) * FullJonesGramTerms(nd);
  const float2* gram = a.gram + bl;
  const GlobalT<float>* prod = AsGlobal(static_cast<const float*>(a.fj_products)) + bl;
  for (int d = wave; d < nd; d += n_waves) {  // one pass (n_waves = n_dir)
    const LaneGains ja{g_s + size_t(d) * 8 * kWave + lane};
    const LaneGains jb{ja.p + 4 * kWave};
    float2 x[4], y[4];
    {
      float2 c[16];
      const float2* cd = gram + (base + size_t(d) * 16) * a.bl_stride;
#pragma unroll
      for (int t = 0; t < 16; ++t) c[t] = LoadGlobalF2(cd + size_t(t) * a.bl_stride);
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int k = 0; k < 2; ++k) {
          float2 xv = make_float2(0.0f, 0.0f), yv = make_float2(0.0f, 0.0f);
#pragma unroll
          for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int m = 0; m < 2; ++m) {
              // X += Jb_d,jm C[ij][km];  Y += Ja_d,jm conj(C[ji][mk])
              CMac(xv, jb[2 * j + m], c[4 * (2 * i + j) + 2 * k + m]);
              const float2 cc = c[4 * (2 * j + i) + 2 * m + k];
              CMac(yv, ja[2 * j + m], make_float2(cc.x, -cc.y));
            }
          x[2 * i + k] = xv;
          y[2 * i + k] = yv;
        }
    }
    const int m = (nd + 1) & ~1;  // players of the pairing (a bye for odd n_dir)
    for (int round = -1; round < m - 1; ++round) {
      // round -1: e = d (Q_d); then d's partner of the round
      int e = d;
      if (round >= 0) {
        if (d == m - 1)
          e = round;
        else if (d == round)
          e = m - 1;
        else
          e = (2 * round - d + 2 * (m - 1)) % (m - 1);
        if (e >= nd) continue;  // the bye
      }
      // G_e,d[p][q] for p, q = 0..3
      float2 g[16];
      if (e < d) {
        const float2* src = gram + (base + size_t(nd + d * (d - 1) / 2 + e) * 16) * a.bl_stride;
#pragma unroll
        for (int t = 0; t < 16; ++t) g[t] = LoadGlobalF2(src + size_t(t) * a.bl_stride);
      } else if (e > d) {
        // G_e,d[p][q] = conj(G_d,e[q][p])
        const float2* src = gram + (base + size_t(nd + e * (e - 1) / 2 + d) * 16) * a.bl_stride;
#pragma unroll
        for (int p = 0; p < 4; ++p)
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const float2 v = LoadGlobalF2(src + size_t(4 * q + p) * a.bl_stride);
            g[4 * p + q] = make_float2(v.x, -v.y);
          }
      } else {
        // Q_d: the real diagonal, then (p < q) in the order (0,1), (0,2),
        // (0,3), (1,2), (1,3), (2,3) as (re, im)
        const GlobalT<float>* q16 =
            prod + ((size_t(cb) * a.n_seg + seg) * nd + d) * kFullJonesModelTerms * a.bl_stride;
        float qv[kFullJonesModelTerms];
#pragma unroll
        for (int t = 0; t < kFullJonesModelTerms; ++t) qv[t] = q16[size_t(t) * a.bl_stride];
#pragma unroll
        for (int p = 0; p < 4; ++p) g[5 * p] = make_float2(qv[p], 0.0f);
        int v = 4;
#pragma unroll
        for (int p = 0; p < 4; ++p)
#pragma unroll
          for (int q = p + 1; q < 4; ++q) {
            g[4 * p + q] = make_float2(qv[v], qv[v + 1]);
            g[4 * q + p] = make_float2(qv[v], -qv[v + 1]);
            v += 2;
          }
      }
      const LaneGains jae{g_s + size_t(e) * 8 * kWave + lane};
      const LaneGains jbe{jae.p + 4 * kWave};
      // H = Jb_e^H Jb_d, H2 = Ja_e^H Ja_d: H[o][m] = sum_j conj(Jb_e,jo) Jb_d,jm
      float2 h[4], h2[4];
#pragma unroll
      for (int o = 0; o < 2; ++o)
#pragma unroll
        for (int m = 0; m < 2; ++m) {
          float2 t = make_float2(0.0f, 0.0f), t2 = make_float2(0.0f, 0.0f);
#pragma unroll
          for (int j = 0; j < 2; ++j) {
            CMacC(t, jbe[2 * j + o], jb[2 * j + m]);
            CMacC(t2, jae[2 * j + o], ja[2 * j + m]);
          }
          h[2 * o + m] = t;
          h2[2 * o + m] = t2;
        }
#pragma unroll
      for (int n = 0; n < 2; ++n)
#pragma unroll
        for (int k = 0; k < 2; ++k) {
          // V[n][k] = sum_om H[o][m] G[no][km];  W[n][k] = sum_om H2[o][m] conj(G[on][mk])
          float2 vv = make_float2(0.0f, 0.0f), ww = make_float2(0.0f, 0.0f);
#pragma unroll
          for (int o = 0; o < 2; ++o)
#pragma unroll
            for (int m = 0; m < 2; ++m) {
              CMac(vv, h[2 * o + m], g[4 * (2 * n + o) + 2 * k + m]);
              const float2 gc = g[4 * (2 * o + n) + 2 * m + k];
              CMac(ww, h2[2 * o + m], make_float2(gc.x, -gc.y));
            }
          // X_ik -= Ja_e,in V[n][k];  Y_ik -= Jb_e,in W[n][k]
#pragma unroll
          for (int i = 0; i < 2; ++i) {
            const float2 ta = jae[2 * i + n];
            const float2 tb = jbe[2 * i + n];
            CMac(x[2 * i + k], make_float2(-ta.x, -ta.y), vv);
            CMac(y[2 * i + k], make_float2(-tb.x, -tb.y), ww);
          }
        }
    }
    float* out = a.partials + ((size_t(cb) * a.n_seg + seg) * nd + d) * kFullJonesTerms * a.bl_stride;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      out[size_t(2 * v) * a.bl_stride + bl] = x[v].x;
      out[size_t(2 * v + 1) * a.bl_stride + bl] = x[v].y;
      out[size_t(8 + 2 * v) * a.bl_stride + bl] = y[v].x;
      out[size_t(9 + 2 * v) * a.bl_stride + bl] = y[v].y;
    }
  }
}

// Single solve: grid (bl tiles, segments, blocks), n_dir waves.
__global__ void __launch_bounds__(kMaxGramDirections * kWave)
    FullJonesApplyGramKernel(const IterArgs a) {
  if (AsConst(a.state)->done) return;
  FullJonesApplyGramTile(a, blockIdx.x, blockIdx.y, blockIdx.z);
}

// Batch: one workgroup per (interval, block, segment, tile) entry, launched
// with the batch's largest n_dir waves; the waves past an interval's
// direction count stage gains and wait at the barrier only.
__global__ void __launch_bounds__(kMaxGramDirections * kWave)
    BatchFullJonesApplyGramKernel(const IterArgs* __restrict__ args,
                                  const BatchTile* __restrict__ tiles) {
  const BatchTile t = LoadConst(tiles + blockIdx.x);
  const IterArgs a = LoadConst(args + t.interval);
  if (AsConst(a.state)->done) return;
  FullJonesApplyGramTile(a, t.tile, t.seg, t.cb);
}

// Q[p][q] of the model products of one baseline (FP64).
struct ModelProducts {
  double v[kFullJonesModelTerms];
  __device__ __forceinline__ double2 at(int p, int q) const {
    if (p == q) return make_double2(v[p], 0.0);
    const int lo = p < q ? p : q, hi = p < q ? q : p;
    // offset of pair (lo, hi) among (0,1), (0,2), (0,3), (1,2), (1,3), (2,3)
    const int idx = 4 + 2 * (lo == 0 ? hi - 1 : lo == 1 ? hi + 1 : 5);
    return p < q ? make_double2(v[idx], v[idx + 1]) : make_double2(v[idx], -v[idx + 1]);
  }
};

// Adds the denominator of one baseline to sd (00, 11, re 01, im 01):
// side 0 (this antenna first):  D_ij = sum_kl H_kl Q[2i+k][2j+l]
// side 1 (this antenna second): D_ij = sum_kl H_kl Q[2l+j][2k+i]
// with H = J^H J of the other antenna's (float-cast) gain.
__device__ __forceinline__ void AddDenominator(const ModelProducts& q, const double2 (&h)[4],
                                               int side, double (&sd)[4]) {
  double2 dv[3] = {make_double2(0.0, 0.0), make_double2(0.0, 0.0), make_double2(0.0, 0.0)};
  const int ij[3][2] = {{0, 0}, {1, 1}, {0, 1}};
#pragma unroll
  for (int e = 0; e < 3; ++e) {
    const int i = ij[e][0], j = ij[e][1];
#pragma unroll
    for (int k = 0; k < 2; ++k)
#pragma unroll
      for (int l = 0; l < 2; ++l) {
        const double2 qq = side == 0 ? q.at(2 * i + k, 2 * j + l) : q.at(2 * l + j, 2 * k + i);
        const double2 t = cmul(h[2 * k + l], qq);
        dv[e].x += t.x;
        dv[e].y += t.y;
      }
  }
  sd[0] += dv[0].x;
  sd[1] += dv[1].x;
  sd[2] += dv[2].x;
  sd[3] += dv[2].y;
}

// antenna solve: grid (n_dir, n_cb); `split` threads per (antenna, solution
// of d) add the antenna's baseline sums (X and D_a where it is the first
// antenna, Y and D_b where it is the second) in FP64 and are combined by a
// fixed shuffle tree; then numerator * denominator^-1 with NaN for a singular
// denominator (IterativeFullJonesSolver.cc:178-185), the step
// (SolverBase.cc:88-158, entry by entry) and the convergence partials.
__device__ __forceinline__ void FullJonesSolveBlock(const IterArgs& a, const int d, const int cb) {
  __shared__ double red[2 * kBlockThreads / 64];
  __shared__ int redc[2 * kBlockThreads / 64];
  const int iteration = a.state->iteration;
  const double2* cur = (iteration & 1) ? a.sol1 : a.sol0;
  double2* nxt = (iteration & 1) ? a.sol0 : a.sol1;
  const int nsd = a.n_sol_dir[d];
  const int s0 = a.sol_start[d];
  const int n_items = a.n_ant * nsd;
  const int split = a.split;
  const int item = threadIdx.x / split;
  const int q = threadIdx.x - item * split;
  const bool active = item < n_items;
  const int ant = active ? item / nsd : 0;
  const int s_abs = s0 + (active ? item - ant * nsd : 0);
  const size_t gbase = size_t(cb) * a.n_ant * a.n_sub;
  const double mean = a.mean[cb];
  const size_t bs = size_t(a.bl_stride);
  int e_begin = 0, e_end = 0;
  if (active) {
    e_begin = AsGlobal(a.csr_off)[ant];
    e_end = AsGlobal(a.csr_off)[ant + 1];
  }
  const GlobalT<int32_t>* csr = AsGlobal(a.csr);
  double sx[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  double sd[4] = {0.0, 0.0, 0.0, 0.0};
  for (int seg = 0; seg < a.n_seg; ++seg) {
    if (!active || AsConst(a.seg_sol)[seg * a.n_dir + d] != s_abs) continue;
    const size_t off = (size_t(cb) * a.n_seg + seg) * a.n_dir + d;
    const GlobalT<float>* src = AsGlobal(a.partials) + off * kFullJonesTerms * bs;
    const GlobalT<float>* qsrc = AsGlobal(a.fj_products) + off * kFullJonesModelTerms * bs;
    for (int e = e_begin + q; e < e_end; e += split) {
      const int code = csr[2 * e];
      const int other = csr[2 * e + 1];
      const int b = code >> 1;
      const int side = code & 1;
      const GlobalT<float>* xs = src + (side ? 8 : 0) * bs + b;
#pragma unroll
      for (int k = 0; k < 8; ++k) sx[k] += double(xs[k * bs]);
      ModelProducts mp;
#pragma unroll
      for (int k = 0; k < kFullJonesModelTerms; ++k) mp.v[k] = double(qsrc[k * bs + b]);
      // H = J^H J of the other antenna's gain as the correlate used it
      double2 jo[4], h[4];
      FixedMatrix(cur, gbase + size_t(other) * a.n_sub + s_abs, mean, jo);
#pragma unroll
      for (int k = 0; k < 4; ++k) jo[k] = make_double2(double(float(jo[k].x)), double(float(jo[k].y)));
#pragma unroll
      for (int k = 0; k < 2; ++k)
#pragma unroll
        for (int l = 0; l < 2; ++l) {
          const double2 t0 = cmulc(jo[l], jo[k]), t1 = cmulc(jo[2 + l], jo[2 + k]);
          h[2 * k + l] = make_double2(t0.x + t1.x, t0.y + t1.y);  // sum_m conj(J_mk) J_ml
        }
      AddDenominator(mp, h, side, sd);
    }
  }
  for (int m = 1; m < split; m <<= 1) {
#pragma unroll
    for (int k = 0; k < 8; ++k) sx[k] += __shfl_xor(sx[k], m);
#pragma unroll
    for (int k = 0; k < 4; ++k) sd[k] += __shfl_xor(sd[k], m);
  }

  double conv_sum = 0.0, mean_sum = 0.0;
  int conv_n = 0, mean_n = 0;
  if (active && q == 0) {
    const size_t sidx = gbase + size_t(ant) * a.n_sub + s_abs;
    double2 old[4];
    FixedMatrix(cur, sidx, mean, old);
    // the gain the correlate used (cast to float)
    double2 jf[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) jf[k] = make_double2(double(float(old[k].x)), double(float(old[k].y)));
    const double2 den[4] = {make_double2(sd[0], 0.0), make_double2(sd[2], sd[3]),
                            make_double2(sd[2], -sd[3]), make_double2(sd[1], 0.0)};
    double2 jd[4], num[4], inv[4], nx[4];
    Mul2x2(jf, den, jd);
#pragma unroll
    for (int k = 0; k < 4; ++k) num[k] = make_double2(sx[2 * k] + jd[k].x, sx[2 * k + 1] + jd[k].y);
    const double nan = std::numeric_limits<double>::quiet_NaN();
    if (Invert2x2(den, inv)) {
      Mul2x2(num, inv, nx);
    } else {
#pragma unroll
      for (int k = 0; k < 4; ++k) nx[k] = make_double2(nan, nan);
    }
    double2 stepped[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      if (!a.phase_only) {
        stepped[k] = make_double2(old[k].x * (1.0 - a.step_size) + nx[k].x * a.step_size,
                                  old[k].y * (1.0 - a.step_size) + nx[k].y * a.step_size);
      } else {
        const double phase_from = atan2(old[k].y, old[k].x);
        double distance = atan2(nx[k].y, nx[k].x) - phase_from;
        if (distance > M_PI)
          distance -= 2.0 * M_PI;
        else if (distance < -M_PI)
          distance += 2.0 * M_PI;
        double sn, cs;
        sincos(phase_from + a.step_size * distance, &sn, &cs);
        stepped[k] = make_double2(cs, sn);
      }
      nxt[sidx * 4 + k] = stepped[k];
    }
    // the optional denominators: the diagonal of the 2x2 denominator, two
    // per solution as for a pair (the warm start's features)
    if (a.den_out) {
      a.den_out[sidx * 2] = sd[0];
      a.den_out[sidx * 2 + 1] = sd[1];
    }
    AddMatrixTerms(old, stepped, conv_sum, conv_n, mean_sum, mean_n);
  }
  BlockSum2(conv_sum, conv_n, mean_sum, mean_n, red, redc);
  const int wg = cb * a.n_dir + d;
  if (threadIdx.x == 0) {
    a.wg_conv[wg] = conv_sum;
    a.wg_conv_n[wg] = conv_n;
    a.wg_mean[wg] = mean_sum;
    a.wg_mean_n[wg] = mean_n;
  }
}

__global__ void __launch_bounds__(kBlockThreads) FullJonesSolveKernel(const IterArgs a) {
  if (a.state->done) return;
  FullJonesSolveBlock(a, blockIdx.x, blockIdx.y);
}

__global__ void __launch_bounds__(kBlockThreads)
    BatchFullJonesSolveKernel(const IterArgs* __restrict__ args,
                              const BatchTile* __restrict__ items) {
  const BatchTile t = items[blockIdx.x];
  const IterArgs& a = args[t.interval];
  if (a.state->done) return;
  FullJonesSolveBlock(a, t.seg, t.cb);  // seg holds the direction
}

// chi^2 with 2x2 solutions: |D - sum_d J_a M_d J_b^H|^2 in FP64 (gains
// unrepaired). Grid (bl tiles, blocks); one partial per workgroup.
template <bool F16>
__global__ void __launch_bounds__(256) FullJonesChi2Kernel(const IterArgs a, const double2* sol,
                                                           double* partial) {
  __shared__ double red[256 / 64];
  const int tile = blockIdx.x;
  const int cb = blockIdx.y;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x / 64;
  const int n_waves = blockDim.x / 64;
  const int bl = tile * 64 + lane;
  const CbInfo info = a.cbinfo[cb];
  const size_t plane = size_t(a.n_t) * info.n_ch * a.bl_stride;
  double acc = 0.0;
  if (bl < a.n_bl) {
    const size_t gbase = size_t(cb) * a.n_ant * a.n_sub;
    const int a1 = a.ant1[bl], a2 = a.ant2[bl];
    for (int seg = 0; seg < a.n_seg; ++seg) {
      const int r0 = a.seg_t0[seg] * info.n_ch, r1 = a.seg_t0[seg + 1] * info.n_ch;
      for (int row = r0 + wave; row < r1; row += n_waves) {
        const size_t e = size_t(row) * a.bl_stride + bl;
        float2 df[4];
        LoadVis<4, false>(a.data, info.data_off + e, df);
        double2 r[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) r[k] = make_double2(df[k].x, df[k].y);
        for (int d = 0; d < a.n_dir; ++d) {
          const int s = a.seg_sol[seg * a.n_dir + d];
          float2 mf[4];
          LoadVis<4, F16>(a.model, info.model_off + d * plane + e, mf);
          const double sc = F16 ? double(a.model_scale[cb * a.n_dir + d]) : 1.0;
          double2 ja[4], jb[4], m[4], am[4];
#pragma unroll
          for (int k = 0; k < 4; ++k) {
            ja[k] = sol[(gbase + size_t(a1) * a.n_sub + s) * 4 + k];
            jb[k] = sol[(gbase + size_t(a2) * a.n_sub + s) * 4 + k];
            m[k] = make_double2(sc * mf[k].x, sc * mf[k].y);
          }
          Mul2x2(ja, m, am);
#pragma unroll
          for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) {
              const double2 p = cmulc(am[2 * i], jb[2 * j]);
              const double2 t = cmulc(am[2 * i + 1], jb[2 * j + 1]);
              r[2 * i + j].x -= p.x + t.x;
              r[2 * i + j].y -= p.y + t.y;
            }
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) acc += cnorm(r[k]);
      }
    }
  }
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
  if (lane == 0) red[wave] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < n_waves; ++w) t += red[w];
    partial[size_t(cb) * gridDim.x + tile] = t;
  }
}

// ---------------------------------------------------------------------------
// Batched kernels: many solution intervals (each with its own IterArgs,
// shapes and device state) in one launch per kernel. A tile / item list maps
// blockIdx.x to (interval, block, segment, tile). The correlate is compiled
// per direction count (its registers scale with it), so a batch launches it
// once per direction count present, over that count's tiles.
// ---------------------------------------------------------------------------
template <int ND, int NC, bool F16, int ROWS>
__global__ void __launch_bounds__(kCorrelateMaxThreads)
    BatchCorrelateKernel(const IterArgs* __restrict__ args, const BatchTile* __restrict__ tiles) {
  const BatchTile t = LoadConst(tiles + blockIdx.x);
  const IterArgs a = LoadConst(args + t.interval);
  if (AsConst(a.state)->done) return;
  CorrelateTile<ND, NC, F16, ROWS>(a, t.tile, t.seg, t.cb);
}

template <int NC, bool F16, bool CHUNKED>
__global__ void __launch_bounds__(kBlockThreads)
    BatchAntennaSolveKernel(const IterArgs* __restrict__ args, const BatchTile* __restrict__ items) {
  const BatchTile t = items[blockIdx.x];
  const IterArgs& a = args[t.interval];
  if (a.state->done) return;
  AntennaSolveBlock<NC, F16, CHUNKED>(a, t.seg, t.cb);  // seg holds the direction
}

using SolveKernelFn = void (*)(IterArgs);
using BatchSolveKernelFn = void (*)(const IterArgs*, const BatchTile*);
// [NC == 4][F16][CHUNKED]
const SolveKernelFn kAntennaSolve[2][2][2] = {
    {{AntennaSolveKernel<2, false, false>, AntennaSolveKernel<2, false, true>},
     {AntennaSolveKernel<2, true, false>, AntennaSolveKernel<2, true, true>}},
    {{AntennaSolveKernel<4, false, false>, AntennaSolveKernel<4, false, true>},
     {AntennaSolveKernel<4, true, false>, AntennaSolveKernel<4, true, true>}}};
const BatchSolveKernelFn kBatchAntennaSolve[2][2][2] = {
    {{BatchAntennaSolveKernel<2, false, false>, BatchAntennaSolveKernel<2, false, true>},
     {BatchAntennaSolveKernel<2, true, false>, BatchAntennaSolveKernel<2, true, true>}},
    {{BatchAntennaSolveKernel<4, false, false>, BatchAntennaSolveKernel<4, false, true>},
     {BatchAntennaSolveKernel<4, true, false>, BatchAntennaSolveKernel<4, true, true>}}};

// items (interval, cb, part); one workgroup per item
template <int ND, bool F16, bool FIN>
__global__ void __launch_bounds__(kGramSolveThreads, 3)
    BatchGramSolveKernel(const IterArgs* __restrict__ args, const BatchTile* __restrict__ items) {
  const BatchTile t = LoadConst(items + XcdContiguous(blockIdx.x, gridDim.x));
  const IterArgs a = LoadConst(args + t.interval);
  if (AsConst(a.state)->done) return;
  GramSolveBlock<ND, F16, FIN>(a, t.cb, t.seg);
}

// items (interval, cb): the end of the fused iteration of each block
__global__ void __launch_bounds__(kGramFinishThreads)
    BatchGramFinishKernel(const IterArgs* __restrict__ args, const BatchTile* __restrict__ items) {
  const BatchTile t = LoadConst(items + blockIdx.x);
  const IterArgs a = LoadConst(args + t.interval);
  if (AsConst(a.state)->done) return;
  GramFinishBlock(a, t.cb);
}

// items: (interval, cb); grid (terms, 2, n_items)
__global__ void __launch_bounds__(256)
    BatchGramPermuteKernel(const IterArgs* __restrict__ args, const BatchTile* __restrict__ items) {
  const BatchTile t = LoadConst(items + blockIdx.z);
  const IterArgs a = LoadConst(args + t.interval);
  if (int(blockIdx.x) >= GramWalkTerms(a.n_dir)) return;
  GramPermuteRow(a, t.cb, blockIdx.y, blockIdx.x);
}

__global__ void __launch_bounds__(256) BatchDecideKernel(const IterArgs* __restrict__ args) {
  const IterArgs& a = args[blockIdx.x];
  if (a.state->done) return;
  DecideBlock(a);
}

__global__ void __launch_bounds__(kBlockThreads)
    BatchPrologueKernel(const IterArgs* __restrict__ args, const BatchTile* __restrict__ items) {
  const BatchTile t = items[blockIdx.x];
  PrologueBlock(args[t.interval], t.cb);
}

// Copies the current solutions of one batched interval (the buffer its
// iteration count selects) to dst.
// grid (blocks, intervals)
__global__ void BatchMoveKernel(const IterArgs* __restrict__ args, const int64_t* __restrict__ off,
                                int mode, double2* __restrict__ buf) {
  const IterArgs& a = args[blockIdx.y];
  const int64_t o = off[blockIdx.y];
  const int64_t count = off[blockIdx.y + 1] - o;
  double2* sol = mode == kMoveToSol0     ? a.sol0
                 : mode == kMoveFromSol1 ? a.sol1
                                         : ((a.state->iteration & 1) ? a.sol1 : a.sol0);
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < count;
       i += int64_t(gridDim.x) * blockDim.x) {
    if (mode == kMoveToSol0)
      sol[i] = buf[o + i];
    else
      buf[o + i] = sol[i];
  }
}

// ---------------------------------------------------------------------------
// Layout conversion for host-provided SolveData.
// ---------------------------------------------------------------------------
template <int NC, bool F16, bool UNI = false>
__global__ void RepackKernel(const float* __restrict__ src, void* __restrict__ dst,
                             int n_planes, int n_t, int n_bl, int n_ch, int bl_stride,
                             const float* plane_scale) {
  const int64_t total = int64_t(n_planes) * n_t * n_ch * bl_stride;
  for (int64_t gid = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; gid < total;
       gid += int64_t(gridDim.x) * blockDim.x) {
    const int bl = int(gid % bl_stride);
    int64_t rest = gid / bl_stride;
    const int ch = int(rest % n_ch);
    rest /= n_ch;
    const int t = int(rest % n_t);
    const int plane = int(rest / n_t);
    float v[2 * NC];
    if (bl < n_bl && UNI) {
      // one value per visibility, into both correlations of the pair
      const float* s = src + (((int64_t(plane) * n_t + t) * n_bl + bl) * n_ch + ch) * 2;
      v[0] = v[2] = s[0];
      v[1] = v[3] = s[1];
    } else if (bl < n_bl) {
      const float* s = src + ((((int64_t(plane) * n_t + t) * n_bl + bl) * n_ch + ch) * NC) * 2;
#pragma unroll
      for (int i = 0; i < 2 * NC; ++i) v[i] = s[i];
    } else {
#pragma unroll
      for (int i = 0; i < 2 * NC; ++i) v[i] = 0.0f;
    }
    if constexpr (!F16) {
      float4* o = reinterpret_cast<float4*>(dst) + gid * (NC / 2);
#pragma unroll
      for (int h = 0; h < NC / 2; ++h)
        o[h] = make_float4(v[4 * h], v[4 * h + 1], v[4 * h + 2], v[4 * h + 3]);
    } else {
      const float inv = 1.0f / plane_scale[plane];
      uint32_t bits[NC];
#pragma unroll
      for (int k = 0; k < NC; ++k) {
        bits[k] = Float2ToHalfBits(v[2 * k] * inv, v[2 * k + 1] * inv);
      }
      if constexpr (NC == 2) {
        reinterpret_cast<uint2*>(dst)[gid] = make_uint2(bits[0], bits[1]);
      } else {
        reinterpret_cast<uint4*>(dst)[gid] = make_uint4(bits[0], bits[1], bits[2], bits[3]);
      }
    }
  }
}

__global__ void PlaneMaxAbsKernel(const float* __restrict__ src, int64_t plane_floats,
                                  uint32_t* max_bits) {
  const int plane = blockIdx.y;
  const float* s = src + int64_t(plane) * plane_floats;
  float m = 0.0f;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < plane_floats;
       i += int64_t(gridDim.x) * blockDim.x) {
    const float x = fabsf(s[i]);
    if (x > m) m = x;  // NaN compares false and is ignored
  }
  // wave reduction then one atomic per wave (non-negative floats order as ints)
  for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_down(m, off));
  if ((threadIdx.x & 63) == 0 && m > 0.0f && isfinite(m))
    atomicMax(&max_bits[plane], __float_as_uint(m));
}

__global__ void ScaleFromMaxKernel(const uint32_t* max_bits, float* scale, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    const float m = __uint_as_float(max_bits[i]);
    // Largest model component maps to 2^14, far from FP16 overflow (65504).
    scale[i] = m > 0.0f ? m / 16384.0f : 1.0f;
  }
}

using KernelFn = void (*)(const IterArgs);
using BatchKernelFn = void (*)(const IterArgs*, const BatchTile*);
constexpr int kMaxRows = 2;

template <int NC, bool F16, int ROWS, int... ND>
constexpr auto MakeCorrelateTable(std::integer_sequence<int, ND...>) {
  return std::array<KernelFn, sizeof...(ND)>{&CorrelateKernel<ND + 1, NC, F16, ROWS>...};
}
template <int NC, bool F16, int ROWS, int... ND>
constexpr auto MakeBatchCorrelateTable(std::integer_sequence<int, ND...>) {
  return std::array<BatchKernelFn, sizeof...(ND)>{&BatchCorrelateKernel<ND + 1, NC, F16, ROWS>...};
}

using Dirs = std::make_integer_sequence<int, kMaxDirections>;
// [n_corr == 4][f16][rows - 1][n_dir - 1]
const std::array<KernelFn, kMaxDirections> kCorrelate[2][2][kMaxRows] = {
    {{MakeCorrelateTable<2, false, 1>(Dirs{}), MakeCorrelateTable<2, false, 2>(Dirs{})},
     {MakeCorrelateTable<2, true, 1>(Dirs{}), MakeCorrelateTable<2, true, 2>(Dirs{})}},
    {{MakeCorrelateTable<4, false, 1>(Dirs{}), MakeCorrelateTable<4, false, 2>(Dirs{})},
     {MakeCorrelateTable<4, true, 1>(Dirs{}), MakeCorrelateTable<4, true, 2>(Dirs{})}}};
const std::array<BatchKernelFn, kMaxDirections> kBatchCorrelate[2][2][kMaxRows] = {
    {{MakeBatchCorrelateTable<2, false, 1>(Dirs{}), MakeBatchCorrelateTable<2, false, 2>(Dirs{})},
     {MakeBatchCorrelateTable<2, true, 1>(Dirs{}), MakeBatchCorrelateTable<2, true, 2>(Dirs{})}},
    {{MakeBatchCorrelateTable<4, false, 1>(Dirs{}), MakeBatchCorrelateTable<4, false, 2>(Dirs{})},
     {MakeBatchCorrelateTable<4, true, 1>(Dirs{}), MakeBatchCorrelateTable<4, true, 2>(Dirs{})}}};

void CheckCorrelateVariant(int n_dir, int n_corr, int rows) {
  DDH_REQUIRE(n_dir >= 1 && n_dir <= kMaxDirections, -4,
              "number of directions must be 1.." + std::to_string(kMaxDirections));
  DDH_REQUIRE(n_corr == 2 || n_corr == 4, -1, "n_correlations must be 2 or 4");
  DDH_REQUIRE(rows >= 1 && rows <= kMaxRows, -1, "rows per trip must be 1 or 2");
}

KernelFn CorrelateFor(int n_dir, int n_corr, bool f16, int rows) {
  CheckCorrelateVariant(n_dir, n_corr, rows);
  return kCorrelate[n_corr == 4][f16][rows - 1][n_dir - 1];
}

BatchKernelFn BatchCorrelateFor(int n_dir, int n_corr, bool f16, int rows) {
  CheckCorrelateVariant(n_dir, n_corr, rows);
  return kBatchCorrelate[n_corr == 4][f16][rows - 1][n_dir - 1];
}

template <bool F16, int... ND>
constexpr auto MakeGramTable(std::integer_sequence<int, ND...>) {
  return std::array<KernelFn, sizeof...(ND)>{&GramKernel<ND + 1, F16>...};
}
template <bool F16, int... ND>
constexpr auto MakeApplyGramTable(std::integer_sequence<int, ND...>) {
  return std::array<KernelFn, sizeof...(ND)>{&ApplyGramKernel<ND + 1, F16>...};
}
template <bool F16, int... ND>
constexpr auto MakeBatchGramTable(std::integer_sequence<int, ND...>) {
  return std::array<BatchKernelFn, sizeof...(ND)>{&BatchGramKernel<ND + 1, F16>...};
}
template <bool F16, int... ND>
constexpr auto MakeBatchApplyGramTable(std::integer_sequence<int, ND...>) {
  return std::array<BatchKernelFn, sizeof...(ND)>{&BatchApplyGramKernel<ND + 1, F16>...};
}
using GramDirs = std::make_integer_sequence<int, kMaxGramDirections>;
// [f16][n_dir - 1]
const std::array<KernelFn, kMaxGramDirections> kGram[2] = {MakeGramTable<false>(GramDirs{}),
                                                          MakeGramTable<true>(GramDirs{})};
const std::array<KernelFn, kMaxGramDirections> kApplyGram[2] = {
    MakeApplyGramTable<false>(GramDirs{}), MakeApplyGramTable<true>(GramDirs{})};
const std::array<BatchKernelFn, kMaxGramDirections> kBatchGram[2] = {
    MakeBatchGramTable<false>(GramDirs{}), MakeBatchGramTable<true>(GramDirs{})};
const std::array<BatchKernelFn, kMaxGramDirections> kBatchApplyGram[2] = {
    MakeBatchApplyGramTable<false>(GramDirs{}), MakeBatchApplyGramTable<true>(GramDirs{})};

void CheckGramDirections(int n_dir) {
  DDH_REQUIRE(n_dir >= 1 && n_dir <= kMaxGramDirections, -4,
              "the Gram iteration supports 1.." + std::to_string(kMaxGramDirections) +
                  " directions");
}

}  // namespace

size_t GramSharedBytes(int n_dir, int waves) {
  (void)n_dir;
  return waves > 1 ? size_t(kGramChunk) * 2 * 2 * kWave * sizeof(float) : 0;
}

void LaunchGram(const IterArgs& a, bool f16, int waves, hipStream_t stream) {
  CheckGramDirections(a.n_dir);
  DDH_REQUIRE(waves >= 1 && waves * kWave <= kCorrelateMaxThreads, -1, "bad waves per tile");
  const dim3 grid(a.bl_stride / kWave, a.n_seg, a.n_cb);
  hipLaunchKernelGGL(kGram[f16][a.n_dir - 1], grid, dim3(waves * kWave),
                     GramSharedBytes(a.n_dir, waves), stream, a);
  DDH_HIP_CHECK(hipGetLastError());
}

size_t ApplyGramSharedBytes(int n_ant, int n_dir) {
  return size_t(n_ant) * (n_dir * 2 + 1) * sizeof(float2);  // padded rows
}

void LaunchApplyGram(const IterArgs& a, bool f16, hipStream_t stream) {
  CheckGramDirections(a.n_dir);
  const dim3 grid(a.bl_stride / kWave, a.n_seg, a.n_cb);
  hipLaunchKernelGGL(kApplyGram[f16][a.n_dir - 1], grid, dim3(kApplyThreads),
                     ApplyGramSharedBytes(a.n_ant, a.n_dir), stream, a);
  DDH_HIP_CHECK(hipGetLastError());
}

void LaunchBatchGram(const IterArgs* d_args, const BatchTile* d_tiles, const TileGroup* groups,
                     int n_groups, bool f16, int waves, hipStream_t stream) {
  DDH_REQUIRE(waves >= 1 && waves * kWave <= kCorrelateMaxThreads, -1, "bad waves per tile");
  for (int g = 0; g < n_groups; ++g) {
    if (groups[g].count == 0) continue;
    CheckGramDirections(groups[g].n_dir);
    hipLaunchKernelGGL(kBatchGram[f16][groups[g].n_dir - 1], dim3(groups[g].count),
                       dim3(waves * kWave), GramSharedBytes(groups[g].n_dir, waves), stream,
                       d_args, d_tiles + groups[g].first);
  }
  DDH_HIP_CHECK(hipGetLastError());
}

void LaunchBatchApply(const IterArgs* d_args, const BatchTile* d_tiles, const TileGroup* groups,
                      int n_groups, int apply_ant, const BatchTile* d_items, int n_items,
                      int n_intervals, size_t solve_lds, int solve_threads, bool f16,
                      const SmoothLaunch& smooth, hipStream_t stream, bool solve_chunked) {
  for (int g = 0; g < n_groups; ++g) {
    if (groups[g].count == 0) continue;
    CheckGramDirections(groups[g].n_dir);
    hipLaunchKernelGGL(kBatchApplyGram[f16][groups[g].n_dir - 1], dim3(groups[g].count),
                       dim3(kApplyThreads), ApplyGramSharedBytes(apply_ant, groups[g].n_dir),
                       stream, d_args, d_tiles + groups[g].first);
  }
  DDH_HIP_CHECK(hipGetLastError());
  LaunchBatchTail(d_args, d_items, n_items, n_intervals, solve_lds, solve_threads, 2, f16, smooth,
                  stream, solve_chunked);
}

namespace {
template <bool F16, bool FIN, int... ND>
constexpr auto MakeGramSolveTable(std::integer_sequence<int, ND...>) {
  return std::array<KernelFn, sizeof...(ND)>{&GramSolveKernel<ND + 1, F16, FIN>...};
}
template <bool F16, bool FIN, int... ND>
constexpr auto MakeBatchGramSolveTable(std::integer_sequence<int, ND...>) {
  return std::array<BatchKernelFn, sizeof...(ND)>{&BatchGramSolveKernel<ND + 1, F16, FIN>...};
}
// [FIN][F16]
const std::array<KernelFn, kMaxGramDirections> kGramSolve[2][2] = {
    {MakeGramSolveTable<false, false>(GramDirs{}), MakeGramSolveTable<true, false>(GramDirs{})},
    {MakeGramSolveTable<false, true>(GramDirs{}), MakeGramSolveTable<true, true>(GramDirs{})}};
const std::array<BatchKernelFn, kMaxGramDirections> kBatchGramSolve[2][2] = {
    {MakeBatchGramSolveTable<false, false>(GramDirs{}),
     MakeBatchGramSolveTable<true, false>(GramDirs{})},
    {MakeBatchGramSolveTable<false, true>(GramDirs{}),
     MakeBatchGramSolveTable<true, true>(GramDirs{})}};
}  // namespace

void LaunchGramPermute(const IterArgs& a, hipStream_t stream) {
  DDH_REQUIRE(a.walk_rows > 0 && a.n_seg == 1, -1, "no walk order for this problem");
  hipLaunchKernelGGL(GramPermuteKernel, dim3(GramWalkTerms(a.n_dir), 2, a.n_cb), dim3(256), 0, stream,
                     a);
  DDH_HIP_CHECK(hipGetLastError());
}

void LaunchGramSolve(const IterArgs& a, bool f16, hipStream_t stream, GramSolveMode mode) {
  CheckGramDirections(a.n_dir);
  DDH_REQUIRE(a.walk_rows > 0 && a.n_ant <= kWave, -1, "no walk order for this problem");
  DDH_REQUIRE(a.gs_groups >= 1 && a.gs_groups <= kGramSolveMaxGroups, -1, "bad row groups");
  hipLaunchKernelGGL(kGramSolve[mode == kGramFinishThenSolve][f16][a.n_dir - 1],
                     dim3(2 * a.gs_groups, a.n_cb), dim3(kGramSolveThreads), 0, stream, a);
  if (mode == kGramSolveAndFinish)
    hipLaunchKernelGGL(GramFinishKernel, dim3(a.n_cb), dim3(kGramFinishThreads), 0, stream, a);
  DDH_HIP_CHECK(hipGetLastError());
}

void LaunchBatchGramPermute(const IterArgs* d_args, const BatchTile* d_items, int n_items,
                            int max_dir, hipStream_t stream) {
  if (n_items == 0) return;
  hipLaunchKernelGGL(BatchGramPermuteKernel, dim3(GramWalkTerms(max_dir), 2, n_items), dim3(256), 0,
                     stream, d_args, d_items);
  DDH_HIP_CHECK(hipGetLastError());
}

void LaunchBatchGramSolve(const IterArgs* d_args, const BatchTile* d_items,
                          const TileGroup* groups, int n_groups, bool f16,
                          const BatchTile* d_blocks, int n_blocks, hipStream_t stream,
                          bool finish_first) {
  for (int g = 0; g < n_groups; ++g) {
    if (groups[g].count == 0) continue;
    CheckGramDirections(groups[g].n_dir);
    hipLaunchKernelGGL(kBatchGramSolve[finish_first][f16][groups[g].n_dir - 1],
                       dim3(groups[g].count), dim3(kGramSolveThreads), 0, stream, d_args,
                       d_items + groups[g].first);
  }
  if (n_blocks > 0)
    hipLaunchKernelGGL(BatchGramFinishKernel, dim3(n_blocks), dim3(kGramFinishThreads), 0, stream,
                       d_args, d_blocks);
  DDH_HIP_CHECK(hipGetLastError());
}

size_t CorrelateSharedBytes(int n_dir, int n_corr, int waves) {
  size_t bytes = size_t(n_dir) * n_corr * kWave * sizeof(float2);
  if (waves > 1) bytes += size_t(3) * n_dir * n_corr * kWave * sizeof(float);
  return bytes;
}

namespace {
using FullJonesKernelFn = void (*)(IterArgs);
using BatchFullJonesKernelFn = void (*)(const IterArgs*, const BatchTile*);
const BatchFullJonesKernelFn kBatchFullJonesCorrelate[2][kFullJonesGroup] = {
    {BatchFullJonesCorrelateKernel<1, false>, BatchFullJonesCorrelateKernel<2, false>,
     BatchFullJonesCorrelateKernel<3, false>, BatchFullJonesCorrelateKernel<4, false>,
     BatchFullJonesCorrelateKernel<5, false>, BatchFullJonesCorrelateKernel<6, false>,
     BatchFullJonesCorrelateKernel<7, false>, BatchFullJonesCorrelateKernel<8, false>},
    {BatchFullJonesCorrelateKernel<1, true>, BatchFullJonesCorrelateKernel<2, true>,
     BatchFullJonesCorrelateKernel<3, true>, BatchFullJonesCorrelateKernel<4, true>,
     BatchFullJonesCorrelateKernel<5, true>, BatchFullJonesCorrelateKernel<6, true>,
     BatchFullJonesCorrelateKernel<7, true>, BatchFullJonesCorrelateKernel<8, true>}};
const FullJonesKernelFn kFullJonesCorrelate[2][kFullJonesGroup] = {
    {FullJonesCorrelateKernel<1, false>, FullJonesCorrelateKernel<2, false>,
     FullJonesCorrelateKernel<3, false>, FullJonesCorrelateKernel<4, false>,
     FullJonesCorrelateKernel<5, false>, FullJonesCorrelateKernel<6, false>,
     FullJonesCorrelateKernel<7, false>, FullJonesCorrelateKernel<8, false>},
    {FullJonesCorrelateKernel<1, true>, FullJonesCorrelateKernel<2, true>,
     FullJonesCorrelateKernel<3, true>, FullJonesCorrelateKernel<4, true>,
     FullJonesCorrelateKernel<5, true>, FullJonesCorrelateKernel<6, true>,
     FullJonesCorrelateKernel<7, true>, FullJonesCorrelateKernel<8, true>}};

int g_max_lds = 64 * 1024;
int g_max_lds_solve = 160 * 1024;
}

size_t MaxSharedBytes() { return size_t(g_max_lds); }

void PrepareKernelAttributes() {
  static bool done = false;
  if (done) return;
  int dev = 0;
  DDH_HIP_CHECK(hipGetDevice(&dev));
  int limit = 0;
  DDH_HIP_CHECK(hipDeviceGetAttribute(&limit, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
  g_max_lds = std::min(limit, 160 * 1024);
  const int max_lds = g_max_lds;
  for (int c = 0; c < 2; ++c)
    for (int h = 0; h < 2; ++h)
      for (int r = 0; r < kMaxRows; ++r)
        for (int d = 0; d < kMaxDirections; ++d) {
          DDH_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(kCorrelate[c][h][r][d]),
                                            hipFuncAttributeMaxDynamicSharedMemorySize, max_lds));
          DDH_HIP_CHECK(hipFuncSetAttribute(
              reinterpret_cast<const void*>(kBatchCorrelate[c][h][r][d]),
              hipFuncAttributeMaxDynamicSharedMemorySize, max_lds));
        }
  for (int h = 0; h < 2; ++h)
    for (int d = 0; d < kMaxGramDirections; ++d) {
      DDH_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(kApplyGram[h][d]),
                                        hipFuncAttributeMaxDynamicSharedMemorySize,
                                        int(kApplyGramMaxShared)));
      DDH_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(kBatchApplyGram[h][d]),
                                        hipFuncAttributeMaxDynamicSharedMemorySize,
                                        int(kApplyGramMaxShared)));
    }
  for (int h = 0; h < 2; ++h)
    for (int g = 0; g < kFullJonesGroup; ++g) {
      DDH_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(kFullJonesCorrelate[h][g]),
                                        hipFuncAttributeMaxDynamicSharedMemorySize, max_lds));
      DDH_HIP_CHECK(
          hipFuncSetAttribute(reinterpret_cast<const void*>(kBatchFullJonesCorrelate[h][g]),
                              hipFuncAttributeMaxDynamicSharedMemorySize, max_lds));
    }
  // antenna_solve also holds static LDS for its reductions
  for (int c = 0; c < 2; ++c)
    for (int h = 0; h < 2; ++h)
      for (int k = 0; k < 2; ++k)
        for (const void* f : {reinterpret_cast<const void*>(kAntennaSolve[c][h][k]),
                              reinterpret_cast<const void*>(kBatchAntennaSolve[c][h][k])}) {
          hipFuncAttributes attr{};
          DDH_HIP_CHECK(hipFuncGetAttributes(&attr, f));
          const int dyn = max_lds - int(attr.sharedSizeBytes);
          g_max_lds_solve = std::min(g_max_lds_solve, dyn);
          DDH_HIP_CHECK(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, dyn));
        }
  done = true;
}

void LaunchBatch(const IterArgs* d_args, const BatchTile* d_tiles, const TileGroup* groups,
                 int n_groups, const BatchTile* d_items, int n_items, int n_intervals,
                 size_t solve_lds, int solve_threads, int n_corr, bool f16, int waves, int rows,
                 const SmoothLaunch& smooth, hipStream_t stream, bool solve_chunked) {
  DDH_REQUIRE(waves >= 1 && waves * kWave <= kCorrelateMaxThreads, -1, "bad waves per tile");
  for (int g = 0; g < n_groups; ++g) {
    const TileGroup& grp = groups[g];
    if (grp.count == 0) continue;
    const size_t lds = CorrelateSharedBytes(grp.n_dir, n_corr, waves);
    DDH_REQUIRE(lds <= MaxSharedBytes(), -4, "batch correlate LDS exceeds the device");
    hipLaunchKernelGGL(BatchCorrelateFor(grp.n_dir, n_corr, f16, rows), dim3(grp.count),
                       dim3(waves * kWave), lds, stream, d_args, d_tiles + grp.first);
  }
  DDH_HIP_CHECK(hipGetLastError());
  LaunchBatchTail(d_args, d_items, n_items, n_intervals, solve_lds, solve_threads, n_corr, f16,
                  smooth, stream, solve_chunked);
}

void LaunchBatchTail(const IterArgs* d_args, const BatchTile* d_items, int n_items,
                     int n_intervals, size_t solve_lds, int solve_threads, int n_corr, bool f16,
                     const SmoothLaunch& smooth, hipStream_t stream, bool solve_chunked) {
  DDH_REQUIRE(solve_threads >= kWave && solve_threads <= kBlockThreads, -1,
              "bad antenna solve block size");
  // antenna solve + decide only
  DDH_REQUIRE(solve_lds <= size_t(g_max_lds_solve), -4, "batch antenna solve LDS exceeds the device");
  if (n_items > 0)
    hipLaunchKernelGGL(kBatchAntennaSolve[n_corr == 4][f16][solve_chunked], dim3(n_items),
                       dim3(solve_threads), solve_lds, stream, d_args, d_items);
  LaunchBatchSmoothDecide(d_args, d_items, n_items, n_intervals, smooth, stream);
}

void LaunchBatchSmoothDecide(const IterArgs* d_args, const BatchTile* d_items, int n_items,
                             int n_intervals, const SmoothLaunch& smooth, hipStream_t stream) {
  if (n_items > 0 && smooth.n_intervals > 0) {
    DDH_REQUIRE(smooth.finish_threads >= kWave && smooth.finish_threads <= kBlockThreads, -1,
                "bad smoothing block size");
    const int64_t blocks = std::min<int64_t>((smooth.max_elements + kSmoothThreads - 1) /
                                                 kSmoothThreads, 1024);
    hipLaunchKernelGGL(BatchSmoothKernel, dim3(unsigned(std::max<int64_t>(blocks, 1)),
                                               smooth.n_intervals),
                       dim3(kSmoothThreads), 0, stream, d_args);
    hipLaunchKernelGGL(BatchSmoothFinishKernel, dim3(n_items), dim3(smooth.finish_threads), 0,
                       stream, d_args, d_items);
  }
  if (n_intervals > 0)
    hipLaunchKernelGGL(BatchDecideKernel, dim3(n_intervals), dim3(256), 0, stream, d_args);
  DDH_HIP_CHECK(hipGetLastError());
}

void LaunchBatchMove(const IterArgs* d_args, const int64_t* d_off, int n_intervals,
                     int64_t max_count, BatchMove mode, double2* buf, hipStream_t stream) {
  if (n_intervals == 0) return;
  const int64_t blocks = std::min<int64_t>((max_count + 255) / 256, 1024);
  hipLaunchKernelGGL(BatchMoveKernel, dim3(unsigned(std::max<int64_t>(blocks, 1)), n_intervals),
                     dim3(256), 0, stream, d_args, d_off, int(mode), buf);
  DDH_HIP_CHECK(hipGetLastError());
}

void LaunchBatchPrologue(const IterArgs* d_args, const BatchTile* d_items, int n_items,
                         hipStream_t stream) {
  hipLaunchKernelGGL(BatchPrologueKernel, dim3(n_items), dim3(kBlockThreads), 0, stream, d_args,
                     d_items);
  DDH_HIP_CHECK(hipGetLastError());
}

size_t FullJonesCorrelateSharedBytes(int n_dir) {
  const int ng = std::min(n_dir, kFullJonesGroup);
  return size_t(n_dir) * kWave * 8 * sizeof(float2) +
         size_t(kFullJonesTerms) * ng * kWave * sizeof(float);
}

// Full-Jones correlate: directions in groups of up to kFullJonesGroup, one
// workgroup per (tile, group, segment, block).
void LaunchFullJonesCorrelate(const IterArgs& a, bool f16, int waves, hipStream_t stream) {
  const int ng = std::min(a.n_dir, kFullJonesGroup);
  const int n_groups = (a.n_dir + ng - 1) / ng;
  const dim3 grid((a.bl_stride / kWave) * n_groups, a.n_seg, a.n_cb);
  const size_t lds = FullJonesCorrelateSharedBytes(a.n_dir);
  DDH_REQUIRE(lds <= MaxSharedBytes(), -4, "full-Jones correlate needs more LDS than the device has");
  hipLaunchKernelGGL(kFullJonesCorrelate[f16][ng - 1], grid, dim3(waves * kWave), lds, stream, a);
  DDH_HIP_CHECK(hipGetLastError());
}

void LaunchFullJonesModelProducts(const IterArgs& a, bool f16, int waves, hipStream_t stream) {
  DDH_REQUIRE(waves >= 1 && waves * kWave <= kCorrelateMaxThreads, -1, "bad waves per tile");
  const dim3 grid(a.bl_stride / kWave, a.n_seg, a.n_cb * a.n_dir);
  if (f16)
    hipLaunchKernelGGL(FullJonesModelProductsKernel<true>, grid, dim3(waves * kWave), 0, stream, a);
  else
    hipLaunchKernelGGL(FullJonesModelProductsKernel<false>, grid, dim3(waves * kWave), 0, stream,
                       a);
  DDH_HIP_CHECK(hipGetLastError());
}

namespace {
// The full-Jones Gram apply's workgroup: one wave per target direction, the
// gains of every direction in LDS.
int FjApplyWaves(int n_dir) { return n_dir; }
size_t FjApplySharedBytes(int n_dir) { return size_t(n_dir) * 8 * kWave * sizeof(float2); }
}  // namespace

void LaunchFullJonesGram(const IterArgs& a, bool f16, hipStream_t stream) {
  CheckGramDirections(a.n_dir);
  DDH_REQUIRE(a.full_jones && a.gram && a.fj_products, -1, "no full-Jones Gram buffers");
  const dim3 grid(a.bl_stride / kWave, a.n_seg, a.n_cb);
  const dim3 block(FjGramWaves(a.n_dir) * kWave);
  DDH_REQUIRE(int(block.x) <= kFjGramMaxThreads, -1, "too many full-Jones Gram waves");
  if (f16)
    hipLaunchKernelGGL(FullJonesGramKernel<true>, grid, block, 0, stream, a);
  else
    hipLaunchKernelGGL(FullJonesGramKernel<false>, grid, block, 0, stream, a);
  DDH_HIP_CHECK(hipGetLastError());
}

void LaunchFullJonesApplyGram(const IterArgs& a, hipStream_t stream) {
  CheckGramDirections(a.n_dir);
  hipLaunchKernelGGL(FullJonesApplyGramKernel, dim3(a.bl_stride / kWave, a.n_seg, a.n_cb),
                     dim3(FjApplyWaves(a.n_dir) * kWave), FjApplySharedBytes(a.n_dir), stream, a);
  DDH_HIP_CHECK(hipGetLastError());
}

void LaunchBatchFullJonesGram(const IterArgs* d_args, const BatchTile* d_tiles, int n_tiles,
                              int max_dir, bool f16, hipStream_t stream) {
  if (n_tiles == 0) return;
  CheckGramDirections(max_dir);
  const dim3 block(FjGramWaves(max_dir) * kWave);
  DDH_REQUIRE(int(block.x) <= kFjGramMaxThreads, -1, "too many full-Jones Gram waves");
  if (f16)
    hipLaunchKernelGGL(BatchFullJonesGramKernel<true>, dim3(n_tiles), block, 0, stream, d_args,
                       d_tiles);
  else
    hipLaunchKernelGGL(BatchFullJonesGramKernel<false>, dim3(n_tiles), block, 0, stream, d_args,
                       d_tiles);
  DDH_HIP_CHECK(hipGetLastError());
}

void LaunchBatchFullJones(const IterArgs* d_args, const BatchTile* d_tiles, const TileGroup* groups,
                          int n_groups, const BatchTile* d_items, int n_items, int n_intervals,
                          int solve_threads, bool f16, int waves, const SmoothLaunch& smooth,
                          hipStream_t stream, const BatchTile* d_apply_tiles, int n_apply_tiles,
                          int apply_dir) {
  DDH_REQUIRE(waves >= 1 && waves * kWave <= kCorrelateMaxThreads, -1, "bad waves per tile");
  if (d_apply_tiles && n_apply_tiles > 0) {
    // Gram iteration: X and Y from the cross products (every direction of a
    // tile in one workgroup) instead of the correlate
    CheckGramDirections(apply_dir);
    hipLaunchKernelGGL(BatchFullJonesApplyGramKernel, dim3(n_apply_tiles),
                       dim3(FjApplyWaves(apply_dir) * kWave), FjApplySharedBytes(apply_dir),
                       stream, d_args, d_apply_tiles);
    n_groups = 0;
  }
  for (int g = 0; g < n_groups; ++g) {
    const TileGroup& grp = groups[g];
    if (grp.count == 0) continue;
    const int ng = std::min(grp.n_dir, kFullJonesGroup);
    const size_t lds = FullJonesCorrelateSharedBytes(grp.n_dir);
    DDH_REQUIRE(lds <= MaxSharedBytes(), -4, "full-Jones correlate needs more LDS than the device has");
    hipLaunchKernelGGL(kBatchFullJonesCorrelate[f16][ng - 1], dim3(grp.count),
                       dim3(waves * kWave), lds, stream, d_args, d_tiles + grp.first);
  }
  DDH_HIP_CHECK(hipGetLastError());
  DDH_REQUIRE(solve_threads >= kWave && solve_threads <= kBlockThreads, -1,
              "bad antenna solve block size");
  if (n_items > 0)
    hipLaunchKernelGGL(BatchFullJonesSolveKernel, dim3(n_items), dim3(solve_threads), 0, stream,
                       d_args, d_items);
  LaunchBatchSmoothDecide(d_args, d_items, n_items, n_intervals, smooth, stream);
}

void LaunchBatchFullJonesModelProducts(const IterArgs* d_args, const BatchTile* d_tiles,
                                       int n_tiles, bool f16, int waves, hipStream_t stream) {
  DDH_REQUIRE(waves >= 1 && waves * kWave <= kCorrelateMaxThreads, -1, "bad waves per tile");
  if (n_tiles == 0) return;
  if (f16)
    hipLaunchKernelGGL(BatchFullJonesModelProductsKernel<true>, dim3(n_tiles), dim3(waves * kWave),
                       0, stream, d_args, d_tiles);
  else
    hipLaunchKernelGGL(BatchFullJonesModelProductsKernel<false>, dim3(n_tiles),
                       dim3(waves * kWave), 0, stream, d_args, d_tiles);
  DDH_HIP_CHECK(hipGetLastError());
}

void LaunchCorrelate(const IterArgs& a, int n_corr, bool f16, int waves, int rows,
                     hipStream_t stream) {
  DDH_REQUIRE(waves >= 1 && waves * kWave <= kCorrelateMaxThreads, -1, "bad waves per tile");
  if (a.full_jones) {
    DDH_REQUIRE(n_corr == 4, -1, "the full-Jones solver needs Full data");
    LaunchFullJonesCorrelate(a, f16, waves, stream);
    return;
  }
  const KernelFn f = CorrelateFor(a.n_dir, n_corr, f16, rows);
  const dim3 grid(a.bl_stride / kWave, a.n_seg, a.n_cb);
  const size_t lds = CorrelateSharedBytes(a.n_dir, n_corr, waves);
  DDH_REQUIRE(lds <= MaxSharedBytes(), -4, "correlate tile needs more LDS than the device has");
  hipLaunchKernelGGL(f, grid, dim3(waves * kWave), lds, stream, a);
  DDH_HIP_CHECK(hipGetLastError());
}

size_t AntennaSolveSharedBytes(const IterArgs& a, int n_corr) {
  (void)n_corr;  // one correlation's partials are staged at a time
  const int max_nsd = std::max(1, a.max_nsd);
  const int chunk = (a.solve_chunk > 0 && a.solve_chunk < a.bl_stride) ? a.solve_chunk
                                                                      : a.bl_stride;
  return size_t(a.n_ant) * max_nsd * 2 * sizeof(double2) + size_t(3) * chunk * sizeof(float);
}

int AntennaSolveChunk(const IterArgs& a) {
  // every baseline of a block in one stage where that fits the workgroup's
  // LDS; else stages of the largest multiple of 64 baselines that fits in
  // 64 KB with the gains (two or more workgroups per compute unit), or in
  // the whole LDS where the gains alone take most of that
  const size_t gains = size_t(a.n_ant) * std::max(1, a.max_nsd) * 2 * sizeof(double2);
  const size_t limit = size_t(g_max_lds_solve);
  const size_t row = 3 * sizeof(float);  // one baseline's three partials
  if (gains + a.bl_stride * row <= limit) return 0;
  DDH_REQUIRE(gains + kWave * row <= limit, -4,
              "too many antennas x solutions for the antenna solve LDS");
  const size_t budget = gains + 1024 * row <= size_t(64) * 1024 ? size_t(64) * 1024 : limit;
  return int((budget - gains) / row / kWave * kWave);
}

int AntennaSolveThreads(int n_ant, int max_nsd) {
  // about four threads per (antenna, solution), whole waves, at most 1024
  const int items = n_ant * std::max(1, max_nsd);
  const int t = ((items * 4 + kWave - 1) / kWave) * kWave;
  return std::min(kBlockThreads, std::max(kWave, t));
}

int DefaultSolveSplit(int n_ant, int max_nsd) {
  const int threads = AntennaSolveThreads(n_ant, max_nsd);
  const int items = n_ant * std::max(1, max_nsd);
  int split = 1;
  while (split < 8 && items * split * 2 <= threads) split *= 2;
  return split;
}

int SolveThreads(const IterArgs& a, int requested) {
  if (a.full_jones) return AntennaSolveThreads(a.n_ant, a.max_nsd);
  const int items = a.n_ant * std::max(1, a.max_nsd);
  const int need = std::max(kWave, ((items * a.split + kWave - 1) / kWave) * kWave);
  DDH_REQUIRE(need <= kBlockThreads, -4,
              "antenna solve: threads per (antenna, solution) x solutions exceed 1024");
  if (requested <= 0) return std::max(need, AntennaSolveThreads(a.n_ant, a.max_nsd));
  DDH_REQUIRE(requested % kWave == 0 && requested <= kBlockThreads && requested >= need, -1,
              "antenna solve block: a multiple of 64 up to 1024 that covers every (antenna, "
              "solution) x split");
  return requested;
}

void LaunchAntennaSolve(const IterArgs& a, int n_corr, bool f16, int threads,
                        hipStream_t stream) {
  const dim3 block(SolveThreads(a, threads));
  if (a.full_jones) {
    hipLaunchKernelGGL(FullJonesSolveKernel, dim3(a.n_dir, a.n_cb), block, 0, stream, a);
    DDH_HIP_CHECK(hipGetLastError());
    return;
  }
  const size_t lds = AntennaSolveSharedBytes(a, n_corr);
  DDH_REQUIRE(lds <= size_t(g_max_lds_solve), -4, "too many baselines for the antenna solve LDS");
  const dim3 grid(a.n_dir, a.n_cb);
  hipLaunchKernelGGL(kAntennaSolve[n_corr == 4][f16][a.solve_chunk > 0], grid, block, lds, stream,
                     a);
  DDH_HIP_CHECK(hipGetLastError());
}

int Chi2Partials(const IterArgs& a) { return (a.bl_stride / kWave) * a.n_cb; }

void LaunchChi2(const IterArgs& a, int n_corr, bool f16, const double2* sol, double* partial,
                int n_partial, double* out, hipStream_t stream, double factor) {
  DDH_REQUIRE(n_partial >= Chi2Partials(a), -1, "chi2 partial buffer too small");
  const dim3 grid(a.bl_stride / kWave, a.n_cb);
  if (a.full_jones) {
    if (f16)
      hipLaunchKernelGGL((FullJonesChi2Kernel<true>), grid, dim3(256), 0, stream, a, sol, partial);
    else
      hipLaunchKernelGGL((FullJonesChi2Kernel<false>), grid, dim3(256), 0, stream, a, sol,
                         partial);
  } else if (n_corr == 2) {
    if (f16)
      hipLaunchKernelGGL((Chi2Kernel<2, true>), grid, dim3(256), 0, stream, a, sol, partial);
    else
      hipLaunchKernelGGL((Chi2Kernel<2, false>), grid, dim3(256), 0, stream, a, sol, partial);
  } else {
    if (f16)
      hipLaunchKernelGGL((Chi2Kernel<4, true>), grid, dim3(256), 0, stream, a, sol, partial);
    else
      hipLaunchKernelGGL((Chi2Kernel<4, false>), grid, dim3(256), 0, stream, a, sol, partial);
  }
  hipLaunchKernelGGL(SumPartialsKernel, dim3(1), dim3(64), 0, stream, partial,
                     Chi2Partials(a), factor, out);
  DDH_HIP_CHECK(hipGetLastError());
}

int SmoothFinishThreads(int n_ant, int max_nsd) {
  const int items = n_ant * std::max(1, max_nsd);
  return std::min(kBlockThreads, std::max(kWave, (items + kWave - 1) / kWave * kWave));
}

void LaunchSmoothing(const IterArgs& a, hipStream_t stream) {
  if (!a.sm_freq) return;
  const int64_t n = SmoothElements(a);
  const int64_t blocks = std::min<int64_t>((n + kSmoothThreads - 1) / kSmoothThreads, 1024);
  hipLaunchKernelGGL(SmoothKernel, dim3(unsigned(std::max<int64_t>(blocks, 1))),
                     dim3(kSmoothThreads), 0, stream, a);
  hipLaunchKernelGGL(SmoothFinishKernel, dim3(a.n_dir, a.n_cb),
                     dim3(SmoothFinishThreads(a.n_ant, a.max_nsd)), 0, stream, a);
  DDH_HIP_CHECK(hipGetLastError());
}

void LaunchDecide(const IterArgs& a, hipStream_t stream) {
  hipLaunchKernelGGL(DecideKernel, dim3(1), dim3(256), 0, stream, a);
  DDH_HIP_CHECK(hipGetLastError());
}

void LaunchPrologue(const IterArgs& a, hipStream_t stream) {
  hipLaunchKernelGGL(PrologueKernel, dim3(a.n_cb), dim3(kBlockThreads), 0, stream, a);
  DDH_HIP_CHECK(hipGetLastError());
}

void LaunchRepack(const float* src, void* dst, int n_planes, int n_t, int n_bl, int n_ch,
                  int bl_stride, int n_corr, bool f16, const float* plane_scale,
                  hipStream_t stream, bool uni) {
  const int64_t total = int64_t(n_planes) * n_t * n_ch * bl_stride;
  if (total == 0) return;
  const int threads = 256;
  const int64_t blocks64 = std::min<int64_t>((total + threads - 1) / threads, 65536);
  const dim3 grid(static_cast<unsigned>(blocks64));
  DDH_REQUIRE(!uni || n_corr == 2, -1, "single-correlation planes are kept as Duo pairs");
  if (uni) {
    if (f16)
      hipLaunchKernelGGL((RepackKernel<2, true, true>), grid, dim3(threads), 0, stream, src, dst,
                         n_planes, n_t, n_bl, n_ch, bl_stride, plane_scale);
    else
      hipLaunchKernelGGL((RepackKernel<2, false, true>), grid, dim3(threads), 0, stream, src,
                         dst, n_planes, n_t, n_bl, n_ch, bl_stride, plane_scale);
  } else if (n_corr == 2) {
    if (f16)
      hipLaunchKernelGGL((RepackKernel<2, true>), grid, dim3(threads), 0, stream, src, dst,
                         n_planes, n_t, n_bl, n_ch, bl_stride, plane_scale);
    else
      hipLaunchKernelGGL((RepackKernel<2, false>), grid, dim3(threads), 0, stream, src, dst,
                         n_planes, n_t, n_bl, n_ch, bl_stride, plane_scale);
  } else {
    if (f16)
      hipLaunchKernelGGL((RepackKernel<4, true>), grid, dim3(threads), 0, stream, src, dst,
                         n_planes, n_t, n_bl, n_ch, bl_stride, plane_scale);
    else
      hipLaunchKernelGGL((RepackKernel<4, false>), grid, dim3(threads), 0, stream, src, dst,
                         n_planes, n_t, n_bl, n_ch, bl_stride, plane_scale);
  }
  DDH_HIP_CHECK(hipGetLastError());
}

void LaunchPlaneMaxAbs(const float* src, int64_t plane_floats, int n_planes,
                       uint32_t* max_bits, hipStream_t stream) {
  if (plane_floats == 0 || n_planes == 0) return;
  const int threads = 256;
  const int64_t blocks = std::min<int64_t>((plane_floats + threads - 1) / threads, 1024);
  hipLaunchKernelGGL(PlaneMaxAbsKernel, dim3(unsigned(blocks), n_planes), dim3(threads), 0,
                     stream, src, plane_floats, max_bits);
  DDH_HIP_CHECK(hipGetLastError());
}

void LaunchScaleFromMax(const uint32_t* max_bits, float* scale, int n, hipStream_t stream) {
  if (n == 0) return;
  hipLaunchKernelGGL(ScaleFromMaxKernel, dim3((n + 255) / 256), dim3(256), 0, stream,
                     max_bits, scale, n);
  DDH_HIP_CHECK(hipGetLastError());
}

}  // namespace ddh
