// Applying the solutions to the model data and subtracting them from the data
// after a solve: DDECal::CorrectAndSubtractModels (steps/DDECal.cc:1074-1124)
// with ApplySolutionAndSubtract (:1031-1072) and the 2x2 products of
// ApplyFullJonesSolution / ApplyDiagonalSolution / ApplyScalarSolution
// (:992-1020), which the reference evaluates in double (aocommon::MC2x2).
//
// Buffers are in DPBuffer order for the time steps of one solution interval:
// data [n_t][n_bl][n_ch][4] and model [n_dir][n_t][n_bl][n_ch][4] complex64,
// solutions [n_cb][n_ant][n_dir][n_pol] complex128. Two threads per
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

// Two lanes per visibility: lane h (0, 1) of a pair holds row h of the 2x2
// matrices (correlations (h, 0) and (h, 1), one 16-byte word), so each load
// and store instruction of a wave covers 1 KiB of consecutive memory. A
// full-Jones product needs the other row of the model, which the partner lane
// passes over (S1 M is formed row by row, then (S1 M) S2^H, the order of the
// 2x2 products above).
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
  const bool small = n_vis < (int64_t(1) << 31);
  const int h = threadIdx.x & 1;
  // the stride is even, so a pair's lanes stay together
  for (int64_t g = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; g < 2 * n_vis;
       g += int64_t(gridDim.x) * blockDim.x) {
    const int64_t v = g >> 1;
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
    float4_v* dp = data + size_t(g);
    float4_v dv = {0, 0, 0, 0};
    if (subtract) dv = __builtin_nontemporal_load(dp);
    C2 sum[2] = {{0, 0}, {0, 0}};
    // directions kDirBatch at a time: their model loads issued together, then
    // the products added in direction order
    for (int d0i = 0; d0i < n_dir; d0i += kDirBatch) {
      float4_v mv[kDirBatch];
#pragma unroll
      for (int j = 0; j < kDirBatch; ++j) {
        const int d = d0i + j < n_dir ? d0i + j : d0i;
        mv[j] = __builtin_nontemporal_load(model + size_t(d) * n_vis * 2 + size_t(g));
      }
#pragma unroll
      for (int j = 0; j < kDirBatch; ++j) {
        const int d = d0i + j;
        if (d >= n_dir) break;
        const float4_v m = mv[j];
        // this lane's row of M: (h, 0), (h, 1)
        const C2 mr[2] = {{m.x, m.y}, {m.z, m.w}};
        const double2* p1 = s + (size_t(a1) * n_dir + d) * NPOL;
        const double2* p2 = s + (size_t(a2) * n_dir + d) * NPOL;
        C2 out[2];
        if constexpr (NPOL == 1) {
          const C2 f = MulC({p1[0].x, p1[0].y}, {p2[0].x, p2[0].y});
          out[0] = Mul(mr[0], f);
          out[1] = Mul(mr[1], f);
        } else if constexpr (NPOL == 2) {
          const C2 s1 = {p1[h].x, p1[h].y};
          const C2 s20 = {p2[0].x, p2[0].y}, s21 = {p2[1].x, p2[1].y};
          out[0] = MulC(Mul(s1, mr[0]), s20);
          out[1] = MulC(Mul(s1, mr[1]), s21);
        } else {
          // the partner's row of M
          float4_v o;
          o.x = __shfl_xor(m.x, 1);
          o.y = __shfl_xor(m.y, 1);
          o.z = __shfl_xor(m.z, 1);
          o.w = __shfl_xor(m.w, 1);
          const C2 orow[2] = {{o.x, o.y}, {o.z, o.w}};
          const C2* r0 = h == 0 ? mr : orow;  // row 0 of M
          const C2* r1 = h == 0 ? orow : mr;  // row 1
          const C2 a0 = {p1[2 * h].x, p1[2 * h].y}, a1v = {p1[2 * h + 1].x, p1[2 * h + 1].y};
          const C2 b[4] = {{p2[0].x, p2[0].y}, {p2[1].x, p2[1].y}, {p2[2].x, p2[2].y},
                           {p2[3].x, p2[3].y}};
          // row h of t = S1 M, then row h of t S2^H
          const C2 t0 = Add(Mul(a0, r0[0]), Mul(a1v, r1[0]));
          const C2 t1 = Add(Mul(a0, r0[1]), Mul(a1v, r1[1]));
          out[0] = Add(MulC(t0, b[0]), MulC(t1, b[1]));
          out[1] = Add(MulC(t0, b[2]), MulC(t1, b[3]));
        }
        sum[0] = Add(sum[0], out[0]);
        sum[1] = Add(sum[1], out[1]);
        if (keep_model)
          model[size_t(d) * n_vis * 2 + size_t(g)] =
              float4_v{float(out[0].re), float(out[0].im), float(out[1].re), float(out[1].im)};
      }
    }
    if (subtract) {
      // complex<float> -= complex<double> (DDECal.cc:1070): the difference is
      // taken in double and rounded once
      dv.x = float(double(dv.x) - sum[0].re);
      dv.y = float(double(dv.y) - sum[0].im);
      dv.z = float(double(dv.z) - sum[1].re);
      dv.w = float(double(dv.w) - sum[1].im);
      *dp = dv;
    }
  }
}

// The solution factors of one (row, block, direction) in LDS, in double2
// values: scalar s1 conj(s2); diagonal f_hk = s1_h conj(s2_k) (row-major);
// full-Jones S1 then S2 (row-major 2x2 each).
template <int NPOL>
constexpr int RowEntry() {
  return NPOL == 1 ? 1 : NPOL == 2 ? 4 : 8;
}
constexpr int kRowTableBytes = 24576;  // factor table per workgroup
constexpr int kMaxRowChannels = 8192;  // channel -> block table (int16)

// The same products over whole rows (time, baseline) of the buffers: a
// workgroup takes `rows` consecutive rows (rows * n_ch * 64 contiguous bytes
// of data and of each direction's model), first forms every (row, block,
// direction) factor of the solutions in LDS, then streams the rows with two
// lanes per visibility as above. Each lane reads its factors from LDS (two
// 16-byte words per direction for diagonal solutions) instead of the
// solution entries of both antennas from global memory, and the per-channel
// block comes from a table built once per workgroup.
template <int NPOL>
__global__ void __launch_bounds__(256)
    CorrectSubtractRowsKernel(float4_v* __restrict__ data, float4_v* __restrict__ model,
                              const double2* __restrict__ sol, const int32_t* __restrict__ ant1,
                              const int32_t* __restrict__ ant2,
                              const int32_t* __restrict__ cb_start, int n_t, int n_bl, int n_ch,
                              int n_cb, int n_dir, int n_ant, int rows, int subtract,
                              int keep_model) {
  constexpr int E = RowEntry<NPOL>();
  extern __shared__ double2 s_dyn[];
  double2* s_tab = s_dyn;  // [rows][n_cb][n_dir][E]
  int16_t* s_cb = reinterpret_cast<int16_t*>(s_dyn + size_t(rows) * n_cb * n_dir * E);
  // channel -> block: the first cb with cb_start[cb + 1] > ch
  for (int c = threadIdx.x; c < n_ch; c += blockDim.x) {
    int lo = 0, hi = n_cb - 1;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (cb_start[mid + 1] > c)
        hi = mid;
      else
        lo = mid + 1;
    }
    s_cb[c] = int16_t(lo);
  }
  const int64_t n_rows = int64_t(n_t) * n_bl;
  const int64_t n_vis = n_rows * n_ch;
  const int64_t n_groups = (n_rows + rows - 1) / rows;
  const int h = threadIdx.x & 1;
  const int per_row = n_cb * n_dir;
  for (int64_t grp = blockIdx.x; grp < n_groups; grp += gridDim.x) {
    const int64_t r0 = grp * rows;
    const int nr = int(std::min<int64_t>(rows, n_rows - r0));
    __syncthreads();  // the previous group's readers are done (and s_cb is written)
    for (int e = threadIdx.x; e < nr * per_row; e += blockDim.x) {
      const int d = e % n_dir;
      const int cb = (e / n_dir) % n_cb;
      const int bl = int((r0 + e / per_row) % n_bl);
      const double2* p1 = sol + ((size_t(cb) * n_ant + ant1[bl]) * n_dir + d) * NPOL;
      const double2* p2 = sol + ((size_t(cb) * n_ant + ant2[bl]) * n_dir + d) * NPOL;
      double2* t = s_tab + size_t(e) * E;
      if constexpr (NPOL == 1) {
        const C2 f = MulC({p1[0].x, p1[0].y}, {p2[0].x, p2[0].y});
        t[0] = make_double2(f.re, f.im);
      } else if constexpr (NPOL == 2) {
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int k = 0; k < 2; ++k) {
            const C2 f = MulC({p1[i].x, p1[i].y}, {p2[k].x, p2[k].y});
            t[2 * i + k] = make_double2(f.re, f.im);
          }
      } else {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          t[k] = p1[k];
          t[4 + k] = p2[k];
        }
      }
    }
    __syncthreads();
    const int64_t g0 = 2 * r0 * n_ch;
    const int span = 2 * nr * n_ch;  // lanes of this group (even: pairs stay together)
    for (int l = threadIdx.x; l < span; l += blockDim.x) {
      const int64_t g = g0 + l;
      const uint32_t v = uint32_t(l) >> 1;
      const uint32_t rr = v / uint32_t(n_ch);
      const int ch = int(v - rr * uint32_t(n_ch));
      const double2* tab = s_tab + (size_t(rr) * n_cb + s_cb[ch]) * n_dir * E;
      float4_v* dp = data + size_t(g);
      float4_v dv = {0, 0, 0, 0};
      if (subtract) dv = __builtin_nontemporal_load(dp);
      C2 sum[2] = {{0, 0}, {0, 0}};
      for (int d0i = 0; d0i < n_dir; d0i += kDirBatch) {
        float4_v mv[kDirBatch];
#pragma unroll
        for (int j = 0; j < kDirBatch; ++j) {
          const int d = d0i + j < n_dir ? d0i + j : d0i;
          mv[j] = __builtin_nontemporal_load(model + size_t(d) * n_vis * 2 + size_t(g));
        }
#pragma unroll
        for (int j = 0; j < kDirBatch; ++j) {
          const int d = d0i + j;
          if (d >= n_dir) break;
          const float4_v m = mv[j];
          const C2 mr[2] = {{m.x, m.y}, {m.z, m.w}};
          const double2* f = tab + d * E;
          C2 out[2];
          if constexpr (NPOL == 1) {
            const C2 fs = {f[0].x, f[0].y};
            out[0] = Mul(mr[0], fs);
            out[1] = Mul(mr[1], fs);
          } else if constexpr (NPOL == 2) {
            const double2 f0 = f[2 * h], f1 = f[2 * h + 1];
            out[0] = Mul(mr[0], {f0.x, f0.y});
            out[1] = Mul(mr[1], {f1.x, f1.y});
          } else {
            float4_v o;
            o.x = __shfl_xor(m.x, 1);
            o.y = __shfl_xor(m.y, 1);
            o.z = __shfl_xor(m.z, 1);
            o.w = __shfl_xor(m.w, 1);
            const C2 orow[2] = {{o.x, o.y}, {o.z, o.w}};
            const C2* r0m = h == 0 ? mr : orow;  // row 0 of M
            const C2* r1m = h == 0 ? orow : mr;  // row 1
            const double2 x0 = f[2 * h], x1 = f[2 * h + 1];
            const C2 a0 = {x0.x, x0.y}, a1v = {x1.x, x1.y};
            const C2 b[4] = {{f[4].x, f[4].y}, {f[5].x, f[5].y}, {f[6].x, f[6].y},
                             {f[7].x, f[7].y}};
            const C2 t0 = Add(Mul(a0, r0m[0]), Mul(a1v, r1m[0]));
            const C2 t1 = Add(Mul(a0, r0m[1]), Mul(a1v, r1m[1]));
            out[0] = Add(MulC(t0, b[0]), MulC(t1, b[1]));
            out[1] = Add(MulC(t0, b[2]), MulC(t1, b[3]));
          }
          sum[0] = Add(sum[0], out[0]);
          sum[1] = Add(sum[1], out[1]);
          if (keep_model)
            __builtin_nontemporal_store(
                float4_v{float(out[0].re), float(out[0].im), float(out[1].re), float(out[1].im)},
                model + size_t(d) * n_vis * 2 + size_t(g));
        }
      }
      if (subtract) {
        dv.x = float(double(dv.x) - sum[0].re);
        dv.y = float(double(dv.y) - sum[0].im);
        dv.z = float(double(dv.z) - sum[1].re);
        dv.w = float(double(dv.w) - sum[1].im);
        __builtin_nontemporal_store(dv, dp);  // streamed out, not read again
      }
    }
  }
}

template <typename T>
struct Scratch {
  T* ptr = nullptr;
  size_t n = 0;
  ~Scratch() {
    if (ptr) (void)hipFree(ptr);
  }
  // grows only: a later call of the same or a smaller shape reuses the memory
  T* Get(size_t count) {
    if (count > n) {
      if (ptr) DDH_HIP_CHECK(hipFree(ptr));
      ptr = nullptr;
      n = 0;
      DDH_HIP_CHECK(hipMalloc(reinterpret_cast<void**>(&ptr), std::max<size_t>(count, 1) * sizeof(T)));
      n = count;
    }
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
  auto* data = reinterpret_cast<float4_v*>(d_data);
  auto* model = reinterpret_cast<float4_v*>(d_model);
  auto* sol = reinterpret_cast<const double2*>(d_sol);
  // whole rows with the solution factors in LDS where one row's factors fit
  const int entry = n_pol == 1 ? RowEntry<1>() : n_pol == 2 ? RowEntry<2>() : RowEntry<4>();
  const size_t row_bytes = size_t(n_cb) * n_dir * entry * sizeof(double2);
  if (row_bytes <= size_t(kRowTableBytes) && n_ch <= kMaxRowChannels && n_cb <= INT16_MAX &&
      int64_t(2) * n_ch * (kRowTableBytes / row_bytes) < INT32_MAX) {
    // rows per workgroup: as many as the table holds, up to about two passes
    // of the workgroup's lanes (at 240 channels one row: 480 lanes; more rows
    // per group, or one pass of 512 lanes, measured slower)
    const int rows = int(std::max<size_t>(
        1, std::min<size_t>(kRowTableBytes / row_bytes, size_t(2 * 256) / (2 * size_t(n_ch)))));
    const int64_t n_rows = int64_t(n_t) * n_bl;
    const unsigned groups = unsigned(std::min<int64_t>((n_rows + rows - 1) / rows, 1 << 20));
    const size_t lds = size_t(rows) * row_bytes + ((size_t(n_ch) * 2 + 15) & ~size_t(15));
#define DDH_ROWS_LAUNCH(NP)                                                                      \
  hipLaunchKernelGGL(CorrectSubtractRowsKernel<NP>, dim3(groups), dim3(256), lds, stream, data, \
                     model, sol, d_ant1, d_ant2, d_cb_start, n_t, n_bl, n_ch, n_cb, n_dir, n_ant, \
                     rows, subtract, keep_model)
    if (n_pol == 1)
      DDH_ROWS_LAUNCH(1);
    else if (n_pol == 2)
      DDH_ROWS_LAUNCH(2);
    else
      DDH_ROWS_LAUNCH(4);
#undef DDH_ROWS_LAUNCH
    DDH_HIP_CHECK(hipGetLastError());
    return;
  }
  const unsigned blocks = unsigned(std::min<int64_t>((2 * n_vis + 255) / 256, 1 << 20));
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
    // device copies of the inputs, kept per host thread for the next call
    static thread_local Scratch<int32_t> a1, a2, cbs;
    static thread_local Scratch<double> sol;
    static thread_local Scratch<float> d, m;
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
