// Warm-start network: a 3-layer perceptron (n_in -> hidden -> hidden -> n_out,
// ReLU) trained and evaluated on the GPU with bf16 MFMA GEMMs (FP32
// accumulation, FP32 master weights, Adam). It maps, per (channel block,
// antenna, solution), the previous interval's solution and summaries of the
// current interval's data/model correlation to a multiplicative correction of
// the initial gains; the plain solver then runs to its stopping criterion
// from that start.
//
// No counterpart in the reference: DP3's only warm starts are
// propagatesolutions and H5Parm initial solutions (steps/DDECal.cc:483-531).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <random>
#include <utility>
#include <vector>

#include "../../include/ddecal_hip.h"
#include "common.h"
#include "warm.h"

namespace ddh {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float float16_v __attribute__((ext_vector_type(16)));

// FP32 -> bf16 bits, round to nearest even (NaN kept quiet).
__host__ __device__ inline uint16_t ToBf16(float f) {
  const uint32_t u = __builtin_bit_cast(uint32_t, f);
  if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu)) return uint16_t((u >> 16) | 0x40);
  const uint32_t rounding = 0x7fffu + ((u >> 16) & 1u);
  return uint16_t((u + rounding) >> 16);
}
__host__ __device__ inline float FromBf16(uint16_t b) {
  return __builtin_bit_cast(float, uint32_t(b) << 16);
}

// C[M x N] (fp32, row-major, ldc) = opA(A)[M x K] * opB(B)[K x N] (+ bias[N])
// (ReLU) for bf16 operands with arbitrary strides:
//   opA(A)(m, k) = trans_a ? A[k * lda + m] : A[m * lda + k]
//   opB(B)(k, n) = trans_b ? B[n * ldb + k] : B[k * ldb + n]
// One 256-thread workgroup per 64 x 64 tile of C, one wave per 32 x 32
// quarter; K in chunks of kGemmK staged in LDS as [row][k] (k contiguous) for
// both operands, read by each lane as one 16-byte fragment per
// v_mfma_f32_32x32x16_bf16 step. Operand rows that are 16-byte aligned and
// whole are loaded as 16-byte vectors; the rest element by element, zero
// outside the matrices. Operand layout of the instruction (wave64): lane l
// supplies row (A) or column (B) l % 32 and the 8 k values 8 * (l / 32) .. + 7;
// the 16 results of lane l are C(8 * (i / 4) + 4 * (l / 32) + i % 4, l % 32).
// K ascends 16 per step, so an output's sum is the same sequence of MFMA steps
// as the one-kernel forward pass makes. With `partial` set, the workgroups of
// grid.y split K into chunks of k_len and store raw sums to
// partial[blockIdx.y][M][N], which GemmReduceKernel adds in split order.
constexpr int kGemmTile = 64;
constexpr int kGemmK = 32;
constexpr int kGemmRow = kGemmK + 8;  // LDS row in bf16 (80 bytes: 16-byte aligned rows)

// The 8 bf16 at columns c .. c + 7 (c a multiple of 8) of row r of a matrix
// with `rows` rows, `cols` columns and leading dimension ld; zero outside.
__device__ __forceinline__ void LoadRow8(const uint16_t* __restrict__ base, int ld, int r,
                                         int rows, int c, int cols, uint16_t (&v)[8]) {
  if (r < rows && c + 7 < cols && (ld & 7) == 0 &&
      (reinterpret_cast<uintptr_t>(base) & 15) == 0) {  // c is a multiple of 8
    const uint4 q = *reinterpret_cast<const uint4*>(base + size_t(r) * ld + c);
    v[0] = q.x & 0xffff; v[1] = q.x >> 16; v[2] = q.y & 0xffff; v[3] = q.y >> 16;
    v[4] = q.z & 0xffff; v[5] = q.z >> 16; v[6] = q.w & 0xffff; v[7] = q.w >> 16;
    return;
  }
#pragma unroll
  for (int j = 0; j < 8; ++j) v[j] = (r < rows && c + j < cols) ? base[size_t(r) * ld + c + j] : 0;
}

// One operand's 64 x kGemmK chunk into LDS rows [64][kGemmRow] (row = m for A,
// n for B). `kmajor`: the matrix is stored with k along its rows' contiguous
// axis ([row][k], ld between rows); otherwise [k][row] and the chunk is
// transposed on the way into LDS.
__device__ __forceinline__ void StageChunk(uint16_t* __restrict__ s, const uint16_t* __restrict__ g,
                                           int ld, bool kmajor, int r0, int rows, int k0,
                                           int K) {
  const int t = threadIdx.x;
  uint16_t v[8];
  if (kmajor) {
    const int r = t >> 2, kk = (t & 3) * 8;  // 64 rows x 4 groups of 8 k
    LoadRow8(g, ld, r0 + r, rows, k0 + kk, K, v);
    uint4 q;
    q.x = v[0] | (uint32_t(v[1]) << 16); q.y = v[2] | (uint32_t(v[3]) << 16);
    q.z = v[4] | (uint32_t(v[5]) << 16); q.w = v[6] | (uint32_t(v[7]) << 16);
    *reinterpret_cast<uint4*>(s + r * kGemmRow + kk) = q;
  } else {
    const int k = t >> 3, rr = (t & 7) * 8;  // 32 k x 8 groups of 8 rows
    LoadRow8(g, ld, k0 + k, K, r0 + rr, rows, v);
#pragma unroll
    for (int j = 0; j < 8; ++j) s[(rr + j) * kGemmRow + k] = v[j];
  }
}

__global__ void __launch_bounds__(256) GemmBf16Kernel(const uint16_t* __restrict__ A, int lda,
                                                      int trans_a,
                                                      const uint16_t* __restrict__ B, int ldb,
                                                      int trans_b, float* __restrict__ C,
                                                      int ldc, int M, int N, int K,
                                                      const float* __restrict__ bias, int relu,
                                                      int accumulate, int k_len,
                                                      float* __restrict__ partial) {
  __shared__ __align__(16) uint16_t sA[kGemmTile * kGemmRow];
  __shared__ __align__(16) uint16_t sB[kGemmTile * kGemmRow];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int tiles_n = (N + kGemmTile - 1) / kGemmTile;
  const int tm = blockIdx.x / tiles_n, tn = blockIdx.x - tm * tiles_n;
  const int m0 = tm * kGemmTile, n0 = tn * kGemmTile;
  const int wm = wave >> 1, wn = wave & 1;
  const int k_begin = blockIdx.y * k_len, k_end = min(K, k_begin + k_len);
  float16_v acc;
#pragma unroll
  for (int i = 0; i < 16; ++i) acc[i] = 0.0f;
  for (int k0 = k_begin; k0 < k_end; k0 += kGemmK) {
    __syncthreads();  // the previous chunk's fragments are read
    StageChunk(sA, A, lda, !trans_a, m0, M, k0, k_end);
    StageChunk(sB, B, ldb, trans_b != 0, n0, N, k0, k_end);
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < kGemmK / 16; ++ks) {
      const int kk = ks * 16 + (lane >> 5) * 8;
      const bf16x8 a = *reinterpret_cast<const bf16x8*>(sA + (wm * 32 + (lane & 31)) * kGemmRow + kk);
      const bf16x8 b = *reinterpret_cast<const bf16x8*>(sB + (wn * 32 + (lane & 31)) * kGemmRow + kk);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
    }
  }
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    const int row = m0 + wm * 32 + 8 * (i / 4) + 4 * (lane >> 5) + (i % 4);
    const int col = n0 + wn * 32 + (lane & 31);
    if (row < M && col < N) {
      if (partial) {
        partial[(size_t(blockIdx.y) * M + row) * N + col] = acc[i];
        continue;
      }
      float v = acc[i];
      if (bias) v += bias[col];
      if (relu) v = fmaxf(v, 0.0f);
      float* out = C + size_t(row) * ldc + col;
      *out = accumulate ? *out + v : v;
    }
  }
}

// C = sum over the splits (in split order) of the partial sums, then the
// epilogue of GemmBf16Kernel.
__global__ void GemmReduceKernel(const float* __restrict__ partial, int splits, int M, int N,
                                 float* __restrict__ C, int ldc, const float* __restrict__ bias,
                                 int relu, int accumulate) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= int64_t(M) * N) return;
  const int row = int(i / N), col = int(i - int64_t(row) * N);
  float v = partial[i];
  for (int s = 1; s < splits; ++s) v += partial[size_t(s) * M * N + i];
  if (bias) v += bias[col];
  if (relu) v = fmaxf(v, 0.0f);
  float* out = C + size_t(row) * ldc + col;
  *out = accumulate ? *out + v : v;
}

// W [K x N] (bf16, row-major, leading dimension N) in the B-operand fragment
// order of v_mfma_f32_32x32x16_bf16: fragment (tn, ks) holds, for lane l, the
// 8 values W[16 ks + 8 (l / 32) + j][32 tn + l % 32], j = 0..7 (zero outside
// K x N), so a wave loads a fragment as one contiguous 1 KiB read.
__global__ void PackFragmentsKernel(const uint16_t* __restrict__ w, int K, int N, int n_tn,
                                    int n_ks, uint16_t* __restrict__ packed) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= int64_t(n_tn) * n_ks * 64 * 8) return;
  const int j = int(i % 8);
  const int lane = int((i / 8) % 64);
  const int frag = int(i / (8 * 64));
  const int tn = frag / n_ks, ks = frag % n_ks;
  const int k = 16 * ks + 8 * (lane / 32) + j;
  const int n = 32 * tn + lane % 32;
  packed[i] = (k < K && n < N) ? w[size_t(k) * N + n] : uint16_t(0);
}

// The warm-start network's forward pass in one kernel, for kWarmFeatures (16)
// inputs, kWarmOutputs (4) outputs and 32 * NT hidden units: one wave per 32
// rows runs the three layers, its hidden activations (rounded to bf16, as the
// next layer's operand) staged in its own LDS rows, the weights read in
// fragment order. Every output is the same sequence of
// v_mfma_f32_32x32x16_bf16 steps (k ascending) over the same bf16 operands,
// with the same bias and ReLU, as the layer-by-layer GemmBf16Kernel path, so
// the two give the same bits; only the inputs and outputs touch HBM.
constexpr int kForwardRows = 32;

template <int NT>
__device__ __forceinline__ void StoreHidden(uint16_t* hs, int stride, const float16_v& acc,
                                            int tn, const float* __restrict__ bias, int lane) {
  const int col = 32 * tn + (lane & 31);
  const float b = bias[col];
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    const int row = 8 * (i / 4) + 4 * (lane >> 5) + (i % 4);
    hs[row * stride + col] = ToBf16(fmaxf(acc[i] + b, 0.0f));
  }
}

template <int NT>
__global__ void __launch_bounds__(64) MlpForwardKernel(
    const float* __restrict__ x, int64_t n, const bf16x8* __restrict__ w1p,
    const float* __restrict__ b1, const bf16x8* __restrict__ w2p, const float* __restrict__ b2,
    const bf16x8* __restrict__ w3p, const float* __restrict__ b3, float* __restrict__ y) {
  constexpr int kHidden = 32 * NT;
  constexpr int kSteps = kHidden / 16;  // k steps of layers 2 and 3
  constexpr int kStride = kHidden + 8;  // bf16 per LDS row: 16 bytes of padding
  __shared__ __attribute__((aligned(16))) uint16_t hs[kForwardRows * kStride];
  const int lane = threadIdx.x;
  const int64_t row0 = int64_t(blockIdx.x) * kForwardRows;
  const int r = lane & 31;
  const int kg = 8 * (lane >> 5);
  // layer 1: the row's 16 features as the A operand (row r, k = kg..kg+7)
  bf16x8 a;
  {
    const int64_t m = row0 + r;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const uint16_t v = m < n ? ToBf16(x[m * kWarmFeatures + kg + j]) : uint16_t(0);
      a[j] = __builtin_bit_cast(__bf16, v);
    }
  }
#pragma unroll
  for (int tn = 0; tn < NT; ++tn) {
    float16_v acc;
#pragma unroll
    for (int i = 0; i < 16; ++i) acc[i] = 0.0f;
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, w1p[tn * 64 + lane], acc, 0, 0, 0);
    StoreHidden<NT>(hs, kStride, acc, tn, b1, lane);
  }
  __syncthreads();
  // layer 2: every output tile advances over k together; H1 read once
  float16_v acc2[NT];
#pragma unroll
  for (int tn = 0; tn < NT; ++tn)
#pragma unroll
    for (int i = 0; i < 16; ++i) acc2[tn][i] = 0.0f;
  for (int ks = 0; ks < kSteps; ++ks) {
    const bf16x8 a2 = *reinterpret_cast<const bf16x8*>(hs + r * kStride + 16 * ks + kg);
#pragma unroll
    for (int tn = 0; tn < NT; ++tn)
      acc2[tn] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a2, w2p[(tn * kSteps + ks) * 64 + lane],
                                                         acc2[tn], 0, 0, 0);
  }
  __syncthreads();  // every read of H1 is done before H2 overwrites it
#pragma unroll
  for (int tn = 0; tn < NT; ++tn) StoreHidden<NT>(hs, kStride, acc2[tn], tn, b2, lane);
  __syncthreads();
  // layer 3: outputs in columns 0..3 of one tile
  float16_v acc3;
#pragma unroll
  for (int i = 0; i < 16; ++i) acc3[i] = 0.0f;
  for (int ks = 0; ks < kSteps; ++ks) {
    const bf16x8 a3 = *reinterpret_cast<const bf16x8*>(hs + r * kStride + 16 * ks + kg);
    acc3 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a3, w3p[ks * 64 + lane], acc3, 0, 0, 0);
  }
  if (r < kWarmOutputs) {
    const float b = b3[r];
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const int64_t m = row0 + 8 * (i / 4) + 4 * (lane >> 5) + (i % 4);
      if (m < n) y[m * kWarmOutputs + r] = acc3[i] + b;
    }
  }
}

__global__ void ToBf16Kernel(const float* __restrict__ x, uint16_t* __restrict__ y, int64_t n) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) y[i] = ToBf16(x[i]);
}

// dY = 2 (Y - T) / (n * n_out); also per-row squared error for the loss.
__global__ void MseGradKernel(const float* __restrict__ y, const float* __restrict__ t,
                              float* __restrict__ dy, float* __restrict__ sq, int64_t n_rows,
                              int n_out) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n_rows * n_out) return;
  const float e = y[i] - t[i];
  dy[i] = 2.0f * e / float(n_rows * n_out);
  sq[i] = e * e;
}

// g *= (h > 0): ReLU backward through the stored post-activation h.
__global__ void ReluBackwardKernel(float* __restrict__ g, const float* __restrict__ h,
                                   int64_t n) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n && h[i] <= 0.0f) g[i] = 0.0f;
}

// Column sums (bias gradients), fixed order per column.
__global__ void ColumnSumKernel(const float* __restrict__ g, int64_t n_rows, int n_cols,
                                float* __restrict__ out) {
  const int c = blockIdx.x;
  __shared__ float red[256];
  float acc = 0.0f;
  for (int64_t r = threadIdx.x; r < n_rows; r += blockDim.x) acc += g[r * n_cols + c];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[c] = red[0];
}

__global__ void SumKernel(const float* __restrict__ x, int64_t n, float* __restrict__ out) {
  __shared__ float red[256];
  float acc = 0.0f;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) acc += x[i];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0];
}

// Adam on FP32 master weights; refreshes the bf16 copy.
__global__ void AdamKernel(float* __restrict__ w, const float* __restrict__ g,
                           float* __restrict__ m, float* __restrict__ v,
                           uint16_t* __restrict__ w_bf16, int64_t n, float lr, float beta1,
                           float beta2, float eps, float bias1, float bias2) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = beta1 * m[i] + (1.0f - beta1) * gi;
  const float vi = beta2 * v[i] + (1.0f - beta2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float upd = lr * (mi / bias1) / (sqrtf(vi / bias2) + eps);
  w[i] -= upd;
  if (w_bf16) w_bf16[i] = ToBf16(w[i]);
}

__global__ void GatherRowsKernel(const float* __restrict__ src, const int32_t* __restrict__ idx,
                                 int64_t n_rows, int n_cols, float* __restrict__ dst) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n_rows * n_cols) return;
  const int64_t r = i / n_cols;
  const int c = int(i - r * n_cols);
  dst[i] = src[int64_t(idx[r]) * n_cols + c];
}

// ---- warm-start features -------------------------------------------------
// Per solution pair (cb, antenna, sub) and polarization p, with g0 the
// previous interval's solution on the current interval:
//   r1 = T(g0) / g0    the undamped one-step update (data/model correlation)
//   ra = g_a / g0      after the first stretch of damped solver iterations
//   rb = g_b / g0      after the second stretch
//   s  = log10(den / mean den over the antennas of this (cb, sub))
// Features (kWarmFeatures, zero padded): for p = 0, 1:
//   [log|r1|, arg r1, log|ra|, arg ra, log|rb|, arg rb, s].
// A 2x2 solution (span ne = 4) takes its diagonal entries J_00, J_11 as the
// two polarizations; its denominators are the diagonal of the 2x2
// denominator, two per solution as for a pair.

__device__ inline void LogRatio(double2 a, double2 b, float& lmag, float& ang) {
  // log(a / b) = log|a| - log|b| + i (arg a - arg b), wrapped to (-pi, pi]
  const double na = hypot(a.x, a.y), nb = hypot(b.x, b.y);
  if (!(na > 0.0 && nb > 0.0 && isfinite(na) && isfinite(nb))) {
    lmag = 0.0f;
    ang = 0.0f;
    return;
  }
  lmag = float(log(na / nb));
  // arg(a conj(b))
  const double re = a.x * b.x + a.y * b.y, im = a.y * b.x - a.x * b.y;
  ang = float(atan2(im, re));
}

// One launch serves many intervals: blockIdx.y selects the interval's span
// (its first solution pair in the concatenated buffers and its shape).
__global__ void WarmFeaturesKernel(const double2* __restrict__ g0_all,
                                   const double2* __restrict__ next1_all,
                                   const double2* __restrict__ ga_all,
                                   const double2* __restrict__ gb_all,
                                   const double* __restrict__ den_all,
                                   const WarmSpan* __restrict__ spans, float* __restrict__ feat_all) {
  const WarmSpan sp = spans[blockIdx.y];
  const int n_ant = sp.n_ant, n_sub = sp.n_sub;
  const int64_t n_pairs = int64_t(sp.n_cb) * n_ant * n_sub;
  const int ne = sp.ne;  // complex values per solution
  const double2* g0 = g0_all + ne * sp.pair0;
  const double2* next1 = next1_all + ne * sp.pair0;
  const double2* ga = ga_all + ne * sp.pair0;
  const double2* gb = gb_all + ne * sp.pair0;
  const double* den = den_all + 2 * sp.pair0;
  float* feat = feat_all + sp.pair0 * kWarmFeatures;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n_pairs;
       i += int64_t(gridDim.x) * blockDim.x) {
    const int s = int(i % n_sub);
    const int cb = int(i / (int64_t(n_ant) * n_sub));
    float f[kWarmFeatures] = {};
    for (int p = 0; p < 2; ++p) {
      float* fp = f + 7 * p;
      const int64_t v = i * ne + p * (ne - 1);  // value p of the pair, entry pp of a matrix
      const double2 base = g0[v];
      LogRatio(next1[v], base, fp[0], fp[1]);
      LogRatio(ga[v], base, fp[2], fp[3]);
      LogRatio(gb[v], base, fp[4], fp[5]);
      double mean = 0.0;
      int cnt = 0;
      for (int a = 0; a < n_ant; ++a) {
        const double dv = den[((int64_t(cb) * n_ant + a) * n_sub + s) * 2 + p];
        if (dv > 0.0 && isfinite(dv)) {
          mean += dv;
          ++cnt;
        }
      }
      const double dv = den[i * 2 + p];
      fp[6] = (cnt > 0 && dv > 0.0 && isfinite(dv)) ? float(log10(dv / (mean / cnt))) : -3.0f;
    }
    for (int k = 0; k < kWarmFeatures; ++k)
      feat[i * kWarmFeatures + k] = isfinite(f[k]) ? f[k] : 0.0f;
  }
}

// init = prev * exp(d_log_mag + i d_arg) per polarization; non-finite prev
// falls back to 1. A 2x2 solution (ne = 4): row p of prev times the factor
// of polarization p (the diagonal entries take the pair's rule, the leakage
// terms scale with their row); a matrix with a non-finite entry falls back
// to the identity. Spans as for the features.
__global__ void WarmApplyKernel(const double2* __restrict__ prev_all,
                                const float* __restrict__ pred_all,
                                const WarmSpan* __restrict__ spans, double2* __restrict__ init_all) {
  const WarmSpan sp = spans[blockIdx.y];
  const int64_t n_pairs = int64_t(sp.n_cb) * sp.n_ant * sp.n_sub;
  const int ne = sp.ne;
  const double2* prev = prev_all + ne * sp.pair0;
  const float* pred = pred_all + sp.pair0 * kWarmOutputs;
  double2* init = init_all + ne * sp.pair0;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n_pairs;
       i += int64_t(gridDim.x) * blockDim.x) {
    if (ne == 4) {
      double2 m[4];
      bool finite = true;
      for (int k = 0; k < 4; ++k) {
        m[k] = prev[i * 4 + k];
        finite = finite && isfinite(m[k].x) && isfinite(m[k].y);
      }
      if (!finite) {
        m[0] = m[3] = make_double2(1.0, 0.0);
        m[1] = m[2] = make_double2(0.0, 0.0);
      }
      for (int p = 0; p < 2; ++p) {
        const double lm = pred[i * kWarmOutputs + 2 * p];
        const double an = pred[i * kWarmOutputs + 2 * p + 1];
        double sn, cs;
        sincos(an, &sn, &cs);
        const double sc = exp(lm);
        for (int q = 0; q < 2; ++q) {
          const double2 g = m[2 * p + q];
          init[i * 4 + 2 * p + q] =
              make_double2(sc * (g.x * cs - g.y * sn), sc * (g.x * sn + g.y * cs));
        }
      }
      continue;
    }
    const float* pr = pred + i * kWarmOutputs;
    // a scalar solution: both values with the outputs' mean (float, as the
    // host helpers form it)
    const float sym[2] = {0.5f * (pr[0] + pr[2]), 0.5f * (pr[1] + pr[3])};
    for (int p = 0; p < 2; ++p) {
      double2 g = prev[i * 2 + p];
      if (!(isfinite(g.x) && isfinite(g.y))) g = make_double2(1.0, 0.0);
      const double lm = sp.scalar ? sym[0] : pr[2 * p];
      const double an = sp.scalar ? sym[1] : pr[2 * p + 1];
      double sn, cs;
      sincos(an, &sn, &cs);
      const double sc = exp(lm);
      init[i * 2 + p] = make_double2(sc * (g.x * cs - g.y * sn), sc * (g.x * sn + g.y * cs));
    }
  }
}

template <typename T>
struct Buf {
  T* ptr = nullptr;
  size_t n = 0;
  ~Buf() {
    if (ptr) (void)hipFree(ptr);
  }
  T* Get(size_t count) {
    if (count > n) {
      if (ptr) DDH_HIP_CHECK(hipFree(ptr));
      ptr = nullptr;
      DDH_HIP_CHECK(hipMalloc(reinterpret_cast<void**>(&ptr), std::max<size_t>(count, 1) * sizeof(T)));
      n = count;
    }
    return ptr;
  }
};

inline unsigned Blocks(int64_t n, int t = 256) { return unsigned((n + t - 1) / t); }

// Split-K partial sums of GEMMs issued from this host thread (one stream at a
// time per thread).
Buf<float>& GemmWorkspace() {
  static thread_local Buf<float> ws;
  return ws;
}

void Gemm(const uint16_t* A, int lda, bool ta, const uint16_t* B, int ldb, bool tb, float* C,
          int ldc, int M, int N, int K, const float* bias, bool relu, bool accumulate,
          hipStream_t s) {
  const int tiles = ((M + kGemmTile - 1) / kGemmTile) * ((N + kGemmTile - 1) / kGemmTile);
  // Long sums over few tiles (the weight gradients, K = rows of the batch)
  // are split over workgroups: about 256 workgroups in all, chunks of at
  // least 512 k.
  int splits = 1;
  if (K >= 1024 && tiles < 128)
    splits = std::max(1, std::min({(256 + tiles - 1) / tiles, K / 512, 64}));
  const int k_len = ((K + splits - 1) / splits + kGemmK - 1) / kGemmK * kGemmK;
  splits = (K + k_len - 1) / k_len;
  float* partial = nullptr;
  if (splits > 1) partial = GemmWorkspace().Get(size_t(splits) * M * N);
  hipLaunchKernelGGL(GemmBf16Kernel, dim3(tiles, splits), dim3(256), 0, s, A, lda, int(ta), B,
                     ldb, int(tb), C, ldc, M, N, K, bias, int(relu), int(accumulate), k_len,
                     partial);
  if (splits > 1)
    hipLaunchKernelGGL(GemmReduceKernel, dim3(Blocks(int64_t(M) * N)), dim3(256), 0, s, partial,
                       splits, M, N, C, ldc, bias, int(relu), int(accumulate));
  DDH_HIP_CHECK(hipGetLastError());
}

void ToBf16(const float* x, uint16_t* y, int64_t n, hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(ToBf16Kernel, dim3(Blocks(n)), dim3(256), 0, s, x, y, n);
}

}  // namespace
}  // namespace ddh

struct ddh_mlp {
  int n_in = 0, n_hidden = 0, n_out = 0;
  hipStream_t stream = nullptr;
  // parameter block: W1 [n_in x h], b1 [h], W2 [h x h], b2 [h], W3 [h x n_out], b3 [n_out]
  int64_t off[7] = {};
  ddh::Buf<float> w, g, m, v;
  ddh::Buf<uint16_t> w16;
  int64_t step = 0;
  // activations
  ddh::Buf<float> x, t, h1, h2, y, dy, dh1, dh2, sq, loss;
  ddh::Buf<uint16_t> x16, h1_16, h2_16, dy16, dh1_16, dh2_16;
  ddh::Buf<float> data_x, data_t;
  ddh::Buf<int32_t> perm;
  // bf16 weights in MFMA fragment order for the one-kernel forward pass,
  // repacked when the weights change (weights_version)
  ddh::Buf<uint16_t> packed;
  int64_t weights_version = 0, packed_version = -1;
  bool fused_forward = true;
  // serializes the packing of the weights (and the layered forward pass,
  // which uses the activation buffers above) between batches that share the
  // network from different host threads
  std::mutex mu;

  int64_t n_params() const { return off[6]; }
};

namespace ddh {
namespace {

void Forward(ddh_mlp& net, int64_t n, hipStream_t s) {
  const int h = net.n_hidden;
  ToBf16(net.x.ptr, net.x16.ptr, n * net.n_in, s);
  // H1 = relu(X W1 + b1)
  Gemm(net.x16.ptr, net.n_in, false, net.w16.ptr + net.off[0], h, false, net.h1.ptr, h, int(n), h,
       net.n_in, net.w.ptr + net.off[1], true, false, s);
  ToBf16(net.h1.ptr, net.h1_16.ptr, n * h, s);
  Gemm(net.h1_16.ptr, h, false, net.w16.ptr + net.off[2], h, false, net.h2.ptr, h, int(n), h, h,
       net.w.ptr + net.off[3], true, false, s);
  ToBf16(net.h2.ptr, net.h2_16.ptr, n * h, s);
  Gemm(net.h2_16.ptr, h, false, net.w16.ptr + net.off[4], net.n_out, false, net.y.ptr, net.n_out,
       int(n), net.n_out, h, net.w.ptr + net.off[5], false, false, s);
}

// The one-kernel forward pass serves 16 -> 32 * NT -> 32 * NT -> 4 networks,
// NT = 1..8; other shapes run layer by layer.
constexpr int kForwardMaxTiles = 8;

bool UsesFusedForward(const ddh_mlp& net) {
  return net.fused_forward && net.n_in == kWarmFeatures && net.n_out == kWarmOutputs &&
         net.n_hidden % 32 == 0 && net.n_hidden / 32 <= kForwardMaxTiles;
}

// Offsets (in bf16 values) of the packed W1, W2 and W3 fragments.
void PackedOffsets(const ddh_mlp& net, int64_t (&o)[4]) {
  const int nt = net.n_hidden / 32, ks = net.n_hidden / 16;
  o[0] = 0;
  o[1] = o[0] + int64_t(nt) * 1 * 64 * 8;
  o[2] = o[1] + int64_t(nt) * ks * 64 * 8;
  o[3] = o[2] + int64_t(1) * ks * 64 * 8;
}

void PackWeights(ddh_mlp& net, hipStream_t s) {
  if (net.packed_version == net.weights_version) return;
  const int h = net.n_hidden, nt = h / 32, ks = h / 16;
  int64_t o[4];
  PackedOffsets(net, o);
  uint16_t* p = net.packed.Get(o[3]);
  hipLaunchKernelGGL(PackFragmentsKernel, dim3(Blocks(o[1] - o[0])), dim3(256), 0, s,
                     net.w16.ptr + net.off[0], kWarmFeatures, h, nt, 1, p + o[0]);
  hipLaunchKernelGGL(PackFragmentsKernel, dim3(Blocks(o[2] - o[1])), dim3(256), 0, s,
                     net.w16.ptr + net.off[2], h, h, nt, ks, p + o[1]);
  hipLaunchKernelGGL(PackFragmentsKernel, dim3(Blocks(o[3] - o[2])), dim3(256), 0, s,
                     net.w16.ptr + net.off[4], h, kWarmOutputs, 1, ks, p + o[2]);
  DDH_HIP_CHECK(hipGetLastError());
  net.packed_version = net.weights_version;
}

typedef void (*ForwardKernelFn)(const float*, int64_t, const bf16x8*, const float*, const bf16x8*,
                                const float*, const bf16x8*, const float*, float*);
template <int... NT>
constexpr std::array<ForwardKernelFn, sizeof...(NT)> MakeForwardTable(
    std::integer_sequence<int, NT...>) {
  return {&MlpForwardKernel<NT + 1>...};
}
constexpr auto kForwardKernels =
    MakeForwardTable(std::make_integer_sequence<int, kForwardMaxTiles>{});

// Network outputs of the n rows of net.x into net.y on stream s.
void Infer(ddh_mlp& net, int64_t n, hipStream_t s) {
  if (!UsesFusedForward(net)) {
    Forward(net, n, s);
    return;
  }
  PackWeights(net, s);
  int64_t o[4];
  PackedOffsets(net, o);
  const bf16x8* p = reinterpret_cast<const bf16x8*>(net.packed.ptr);
  const int64_t blocks = (n + kForwardRows - 1) / kForwardRows;
  DDH_REQUIRE(blocks <= INT32_MAX, -1, "too many rows for one forward pass");
  hipLaunchKernelGGL(kForwardKernels[net.n_hidden / 32 - 1], dim3(unsigned(blocks)), dim3(64), 0,
                     s, net.x.ptr, n, p + o[0] / 8, net.w.ptr + net.off[1], p + o[1] / 8,
                     net.w.ptr + net.off[3], p + o[2] / 8, net.w.ptr + net.off[5], net.y.ptr);
  DDH_HIP_CHECK(hipGetLastError());
}

// Gradients of the mean squared error over n rows into net.g; loss into net.loss[0].
void Backward(ddh_mlp& net, int64_t n) {
  hipStream_t s = net.stream;
  const int h = net.n_hidden, no = net.n_out, ni = net.n_in;
  hipLaunchKernelGGL(MseGradKernel, dim3(Blocks(n * no)), dim3(256), 0, s, net.y.ptr, net.t.ptr,
                     net.dy.ptr, net.sq.ptr, n, no);
  hipLaunchKernelGGL(SumKernel, dim3(1), dim3(256), 0, s, net.sq.ptr, n * no, net.loss.ptr);
  ToBf16(net.dy.ptr, net.dy16.ptr, n * no, s);
  // dW3 = H2^T dY ; db3 = colsum(dY)
  Gemm(net.h2_16.ptr, h, true, net.dy16.ptr, no, false, net.g.ptr + net.off[4], no, h, no, int(n),
       nullptr, false, false, s);
  hipLaunchKernelGGL(ColumnSumKernel, dim3(no), dim3(256), 0, s, net.dy.ptr, n, no,
                     net.g.ptr + net.off[5]);
  // dH2 = dY W3^T * relu'(H2)
  Gemm(net.dy16.ptr, no, false, net.w16.ptr + net.off[4], no, true, net.dh2.ptr, h, int(n), h, no,
       nullptr, false, false, s);
  hipLaunchKernelGGL(ReluBackwardKernel, dim3(Blocks(n * h)), dim3(256), 0, s, net.dh2.ptr,
                     net.h2.ptr, n * h);
  ToBf16(net.dh2.ptr, net.dh2_16.ptr, n * h, s);
  Gemm(net.h1_16.ptr, h, true, net.dh2_16.ptr, h, false, net.g.ptr + net.off[2], h, h, h, int(n),
       nullptr, false, false, s);
  hipLaunchKernelGGL(ColumnSumKernel, dim3(h), dim3(256), 0, s, net.dh2.ptr, n, h,
                     net.g.ptr + net.off[3]);
  // dH1 = dH2 W2^T * relu'(H1)
  Gemm(net.dh2_16.ptr, h, false, net.w16.ptr + net.off[2], h, true, net.dh1.ptr, h, int(n), h, h,
       nullptr, false, false, s);
  hipLaunchKernelGGL(ReluBackwardKernel, dim3(Blocks(n * h)), dim3(256), 0, s, net.dh1.ptr,
                     net.h1.ptr, n * h);
  ToBf16(net.dh1.ptr, net.dh1_16.ptr, n * h, s);
  Gemm(net.x16.ptr, ni, true, net.dh1_16.ptr, h, false, net.g.ptr + net.off[0], h, ni, h, int(n),
       nullptr, false, false, s);
  hipLaunchKernelGGL(ColumnSumKernel, dim3(h), dim3(256), 0, s, net.dh1.ptr, n, h,
                     net.g.ptr + net.off[1]);
  DDH_HIP_CHECK(hipGetLastError());
}

void Reserve(ddh_mlp& net, int64_t n) {
  const int h = net.n_hidden;
  net.x.Get(n * net.n_in);
  net.x16.Get(n * net.n_in);
  net.t.Get(n * net.n_out);
  net.h1.Get(n * h);
  net.h2.Get(n * h);
  net.h1_16.Get(n * h);
  net.h2_16.Get(n * h);
  net.y.Get(n * net.n_out);
  net.dy.Get(n * net.n_out);
  net.dy16.Get(n * net.n_out);
  net.sq.Get(n * net.n_out);
  net.dh1.Get(n * h);
  net.dh2.Get(n * h);
  net.dh1_16.Get(n * h);
  net.dh2_16.Get(n * h);
  net.loss.Get(1);
}

}  // namespace

void CheckWarmNetwork(const ddh_mlp& net) {
  DDH_REQUIRE(net.n_in == kWarmFeatures && net.n_out == kWarmOutputs, -1,
              "the warm-start network maps 16 features to 4 outputs");
}

void WaitForNetwork(ddh_mlp& net) { DDH_HIP_CHECK(hipStreamSynchronize(net.stream)); }

void MlpForwardRows(ddh_mlp& net, const float* x, int64_t n, float* y, hipStream_t stream) {
  if (UsesFusedForward(net)) {
    {
      // the fragment-ordered weights, packed once per weight update on the
      // network's stream
      std::lock_guard<std::mutex> lock(net.mu);
      if (net.packed_version != net.weights_version) {
        PackWeights(net, net.stream);
        DDH_HIP_CHECK(hipStreamSynchronize(net.stream));
      }
    }
    int64_t o[4];
    PackedOffsets(net, o);
    const bf16x8* p = reinterpret_cast<const bf16x8*>(net.packed.ptr);
    const int64_t blocks = (n + kForwardRows - 1) / kForwardRows;
    DDH_REQUIRE(blocks <= INT32_MAX, -1, "too many rows for one forward pass");
    hipLaunchKernelGGL(kForwardKernels[net.n_hidden / 32 - 1], dim3(unsigned(blocks)), dim3(64),
                       0, stream, x, n, p + o[0] / 8, net.w.ptr + net.off[1], p + o[1] / 8,
                       net.w.ptr + net.off[3], p + o[2] / 8, net.w.ptr + net.off[5], y);
    DDH_HIP_CHECK(hipGetLastError());
    return;
  }
  // layered forward through the network's own activation buffers: one
  // caller at a time
  std::lock_guard<std::mutex> lock(net.mu);
  Reserve(net, n);
  DDH_HIP_CHECK(hipMemcpyAsync(net.x.ptr, x, size_t(n) * net.n_in * sizeof(float),
                               hipMemcpyDeviceToDevice, stream));
  Forward(net, n, stream);
  DDH_HIP_CHECK(hipMemcpyAsync(y, net.y.ptr, size_t(n) * net.n_out * sizeof(float),
                               hipMemcpyDeviceToDevice, stream));
  DDH_HIP_CHECK(hipStreamSynchronize(stream));
}

dim3 WarmGrid(int64_t max_pairs, int n_spans) {
  return dim3(unsigned(std::min<int64_t>(std::max<int64_t>(Blocks(max_pairs), 1), 4096)),
              unsigned(n_spans));
}

void LaunchWarmFeatures(const double2* g0, const double2* next1, const double2* ga,
                        const double2* gb, const double* den, const WarmSpan* spans, int n_spans,
                        int64_t max_pairs, float* features, hipStream_t stream) {
  hipLaunchKernelGGL(WarmFeaturesKernel, WarmGrid(max_pairs, n_spans), dim3(256), 0, stream, g0,
                     next1, ga, gb, den, spans, features);
  DDH_HIP_CHECK(hipGetLastError());
}

void LaunchWarmApply(const double2* g0, const float* pred, const WarmSpan* spans, int n_spans,
                     int64_t max_pairs, double2* init, hipStream_t stream) {
  hipLaunchKernelGGL(WarmApplyKernel, WarmGrid(max_pairs, n_spans), dim3(256), 0, stream, g0, pred,
                     spans, init);
  DDH_HIP_CHECK(hipGetLastError());
}

}  // namespace ddh

namespace ddh {
namespace {
// Host arrays through the warm-start kernels: ne complex values per solution
// (2: a diagonal pair, 4: a 2x2 matrix), two denominators per solution.
void WarmFeaturesHost(const double* g0, const double* next1, const double* ga, const double* gb,
                      const double* den, int32_t n_cb, int32_t n_ant, int32_t n_sub, int ne,
                      float* features) {
  DDH_REQUIRE(g0 && next1 && ga && gb && den && features, -1, "NULL argument");
  DDH_REQUIRE(n_cb >= 1 && n_ant >= 1 && n_sub >= 1, -1, "empty dimension");
  const int64_t n_pairs = int64_t(n_cb) * n_ant * n_sub;
  Buf<double2> d0, d1, da, db;
  Buf<double> dd;
  Buf<float> df;
  const size_t bytes = n_pairs * ne * sizeof(double2);
  DDH_HIP_CHECK(hipMemcpy(d0.Get(n_pairs * ne), g0, bytes, hipMemcpyHostToDevice));
  DDH_HIP_CHECK(hipMemcpy(d1.Get(n_pairs * ne), next1, bytes, hipMemcpyHostToDevice));
  DDH_HIP_CHECK(hipMemcpy(da.Get(n_pairs * ne), ga, bytes, hipMemcpyHostToDevice));
  DDH_HIP_CHECK(hipMemcpy(db.Get(n_pairs * ne), gb, bytes, hipMemcpyHostToDevice));
  DDH_HIP_CHECK(hipMemcpy(dd.Get(n_pairs * 2), den, n_pairs * 2 * sizeof(double),
                          hipMemcpyHostToDevice));
  Buf<WarmSpan> span;
  const WarmSpan one{0, n_cb, n_ant, n_sub, ne, 0};
  DDH_HIP_CHECK(hipMemcpy(span.Get(1), &one, sizeof(one), hipMemcpyHostToDevice));
  LaunchWarmFeatures(d0.ptr, d1.ptr, da.ptr, db.ptr, dd.ptr, span.ptr, 1, n_pairs,
                     df.Get(n_pairs * kWarmFeatures), nullptr);
  DDH_HIP_CHECK(hipMemcpy(features, df.ptr, n_pairs * kWarmFeatures * sizeof(float),
                          hipMemcpyDeviceToHost));
}

void WarmApplyHost(const double* prev, const float* pred, int64_t n_pairs, int ne, double* init) {
  DDH_REQUIRE(prev && pred && init, -1, "NULL argument");
  DDH_REQUIRE(n_pairs >= 1 && n_pairs <= INT32_MAX, -1, "bad pair count");
  Buf<double2> dp, di;
  Buf<float> dpred;
  DDH_HIP_CHECK(hipMemcpy(dp.Get(n_pairs * ne), prev, n_pairs * ne * sizeof(double2),
                          hipMemcpyHostToDevice));
  DDH_HIP_CHECK(hipMemcpy(dpred.Get(n_pairs * kWarmOutputs), pred,
                          n_pairs * kWarmOutputs * sizeof(float), hipMemcpyHostToDevice));
  Buf<WarmSpan> span;
  const WarmSpan one{0, 1, 1, int32_t(n_pairs), ne, 0};
  DDH_HIP_CHECK(hipMemcpy(span.Get(1), &one, sizeof(one), hipMemcpyHostToDevice));
  LaunchWarmApply(dp.ptr, dpred.ptr, span.ptr, 1, n_pairs, di.Get(n_pairs * ne), nullptr);
  DDH_HIP_CHECK(hipMemcpy(init, di.ptr, n_pairs * ne * sizeof(double2), hipMemcpyDeviceToHost));
}
}  // namespace
}  // namespace ddh

using namespace ddh;

extern "C" {

ddh_mlp* ddh_mlp_create(int32_t n_in, int32_t n_hidden, int32_t n_out, uint32_t seed) {
  ddh_mlp* out = nullptr;
  const int rc = Guard([&] {
    DDH_REQUIRE(n_in >= 1 && n_hidden >= 1 && n_out >= 1, -1, "bad layer sizes");
    auto net = std::make_unique<ddh_mlp>();
    net->n_in = n_in;
    net->n_hidden = n_hidden;
    net->n_out = n_out;
    const int64_t h = n_hidden;
    const int64_t sizes[6] = {n_in * h, h, h * h, h, h * n_out, n_out};
    net->off[0] = 0;
    for (int i = 0; i < 6; ++i) net->off[i + 1] = net->off[i] + sizes[i];
    const int64_t np = net->n_params();
    // He-normal weights, zero biases (host RNG, seeded)
    std::vector<float> w(np, 0.0f);
    std::mt19937 rng(seed);
    const int fan_in[3] = {n_in, n_hidden, n_hidden};
    for (int layer = 0; layer < 3; ++layer) {
      std::normal_distribution<float> nd(0.0f, std::sqrt(2.0f / fan_in[layer]));
      for (int64_t i = net->off[2 * layer]; i < net->off[2 * layer + 1]; ++i) w[i] = nd(rng);
    }
    // the output layer starts small: the initial correction is near zero
    for (int64_t i = net->off[4]; i < net->off[5]; ++i) w[i] *= 0.01f;
    DDH_HIP_CHECK(hipStreamCreateWithFlags(&net->stream, hipStreamNonBlocking));
    net->w.Get(np);
    net->g.Get(np);
    net->m.Get(np);
    net->v.Get(np);
    net->w16.Get(np);
    DDH_HIP_CHECK(hipMemcpy(net->w.ptr, w.data(), np * sizeof(float), hipMemcpyHostToDevice));
    DDH_HIP_CHECK(hipMemset(net->m.ptr, 0, np * sizeof(float)));
    DDH_HIP_CHECK(hipMemset(net->v.ptr, 0, np * sizeof(float)));
    ToBf16(net->w.ptr, net->w16.ptr, np, net->stream);
    DDH_HIP_CHECK(hipStreamSynchronize(net->stream));
    out = net.release();
  });
  return rc == 0 ? out : nullptr;
}

void ddh_mlp_destroy(ddh_mlp* net) {
  if (!net) return;
  (void)hipStreamSynchronize(net->stream);
  (void)hipStreamDestroy(net->stream);
  delete net;
}

int64_t ddh_mlp_num_params(const ddh_mlp* net) { return net ? net->n_params() : 0; }

int ddh_mlp_get_params(const ddh_mlp* net, float* params) {
  return Guard([&] {
    DDH_REQUIRE(net && params, -1, "NULL argument");
    DDH_HIP_CHECK(hipMemcpy(params, net->w.ptr, net->n_params() * sizeof(float),
                            hipMemcpyDeviceToHost));
  });
}

int ddh_mlp_set_params(ddh_mlp* net, const float* params) {
  return Guard([&] {
    DDH_REQUIRE(net && params, -1, "NULL argument");
    DDH_HIP_CHECK(hipMemcpy(net->w.ptr, params, net->n_params() * sizeof(float),
                            hipMemcpyHostToDevice));
    ToBf16(net->w.ptr, net->w16.ptr, net->n_params(), net->stream);
    ++net->weights_version;
    DDH_HIP_CHECK(hipStreamSynchronize(net->stream));
  });
}

int ddh_mlp_set_fused_forward(ddh_mlp* net, int32_t enable) {
  return Guard([&] {
    DDH_REQUIRE(net, -1, "NULL argument");
    net->fused_forward = enable != 0;
  });
}

int ddh_mlp_forward(ddh_mlp* net, const float* x, int64_t n, float* y) {
  return Guard([&] {
    DDH_REQUIRE(net && x && y && n >= 1, -1, "bad argument");
    Reserve(*net, n);
    DDH_HIP_CHECK(hipMemcpyAsync(net->x.ptr, x, n * net->n_in * sizeof(float),
                                 hipMemcpyHostToDevice, net->stream));
    Infer(*net, n, net->stream);
    DDH_HIP_CHECK(hipMemcpyAsync(y, net->y.ptr, n * net->n_out * sizeof(float),
                                 hipMemcpyDeviceToHost, net->stream));
    DDH_HIP_CHECK(hipStreamSynchronize(net->stream));
  });
}

int ddh_mlp_loss_and_grads(ddh_mlp* net, const float* x, const float* t, int64_t n,
                           float* loss, float* grads) {
  return Guard([&] {
    DDH_REQUIRE(net && x && t && n >= 1, -1, "bad argument");
    Reserve(*net, n);
    DDH_HIP_CHECK(hipMemcpyAsync(net->x.ptr, x, n * net->n_in * sizeof(float),
                                 hipMemcpyHostToDevice, net->stream));
    DDH_HIP_CHECK(hipMemcpyAsync(net->t.ptr, t, n * net->n_out * sizeof(float),
                                 hipMemcpyHostToDevice, net->stream));
    Forward(*net, n, net->stream);
    Backward(*net, n);
    float sq = 0.0f;
    DDH_HIP_CHECK(hipMemcpyAsync(&sq, net->loss.ptr, sizeof(float), hipMemcpyDeviceToHost,
                                 net->stream));
    if (grads)
      DDH_HIP_CHECK(hipMemcpyAsync(grads, net->g.ptr, net->n_params() * sizeof(float),
                                   hipMemcpyDeviceToHost, net->stream));
    DDH_HIP_CHECK(hipStreamSynchronize(net->stream));
    if (loss) *loss = sq / float(n * net->n_out);
  });
}

// Minibatch Adam over (x [n x n_in], t [n x n_out]) for `epochs` passes with a
// seeded shuffle. Returns the mean loss of the last epoch.
int ddh_mlp_train(ddh_mlp* net, const float* x, const float* t, int64_t n, int32_t epochs,
                  int32_t batch, float lr, uint32_t seed, float* last_epoch_loss) {
  return Guard([&] {
    DDH_REQUIRE(net && x && t && n >= 1 && epochs >= 1 && batch >= 1, -1, "bad argument");
    hipStream_t s = net->stream;
    float* dx = net->data_x.Get(n * net->n_in);
    float* dt = net->data_t.Get(n * net->n_out);
    DDH_HIP_CHECK(hipMemcpyAsync(dx, x, n * net->n_in * sizeof(float), hipMemcpyHostToDevice, s));
    DDH_HIP_CHECK(hipMemcpyAsync(dt, t, n * net->n_out * sizeof(float), hipMemcpyHostToDevice, s));
    Reserve(*net, batch);
    int32_t* dperm = net->perm.Get(n);
    std::vector<int32_t> perm(n);
    std::mt19937 rng(seed);
    const float beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
    double epoch_loss = 0.0;
    for (int e = 0; e < epochs; ++e) {
      for (int64_t i = 0; i < n; ++i) perm[i] = int32_t(i);
      std::shuffle(perm.begin(), perm.end(), rng);
      DDH_HIP_CHECK(hipMemcpyAsync(dperm, perm.data(), n * sizeof(int32_t), hipMemcpyHostToDevice, s));
      DDH_HIP_CHECK(hipStreamSynchronize(s));  // perm is reused next epoch
      epoch_loss = 0.0;
      int64_t batches = 0;
      for (int64_t b0 = 0; b0 + batch <= n; b0 += batch) {
        hipLaunchKernelGGL(GatherRowsKernel, dim3(Blocks(int64_t(batch) * net->n_in)), dim3(256), 0,
                           s, dx, dperm + b0, int64_t(batch), net->n_in, net->x.ptr);
        hipLaunchKernelGGL(GatherRowsKernel, dim3(Blocks(int64_t(batch) * net->n_out)), dim3(256),
                           0, s, dt, dperm + b0, int64_t(batch), net->n_out, net->t.ptr);
        Forward(*net, batch, net->stream);
        Backward(*net, batch);
        net->step += 1;
        const float bias1 = 1.0f - std::pow(beta1, float(net->step));
        const float bias2 = 1.0f - std::pow(beta2, float(net->step));
        hipLaunchKernelGGL(AdamKernel, dim3(Blocks(net->n_params())), dim3(256), 0, s, net->w.ptr,
                           net->g.ptr, net->m.ptr, net->v.ptr, net->w16.ptr, net->n_params(), lr,
                           beta1, beta2, eps, bias1, bias2);
        ++net->weights_version;
        float sq = 0.0f;
        if (e == epochs - 1) {
          DDH_HIP_CHECK(hipMemcpyAsync(&sq, net->loss.ptr, sizeof(float), hipMemcpyDeviceToHost, s));
          DDH_HIP_CHECK(hipStreamSynchronize(s));
          epoch_loss += sq / float(int64_t(batch) * net->n_out);
        }
        ++batches;
      }
      if (batches > 0) epoch_loss /= batches;
    }
    DDH_HIP_CHECK(hipStreamSynchronize(s));
    if (last_epoch_loss) *last_epoch_loss = float(epoch_loss);
  });
}

// C = A B with bf16 operands (given as FP32, rounded to bf16 on the device)
// and FP32 accumulation; A [M x K], B [K x N] row-major. Exposes the MFMA
// GEMM for testing.
int ddh_gemm_bf16(const float* a, const float* b, float* c, int32_t M, int32_t N, int32_t K,
                  int32_t trans_a, int32_t trans_b) {
  return Guard([&] {
    DDH_REQUIRE(a && b && c && M >= 1 && N >= 1 && K >= 1, -1, "bad argument");
    Buf<float> fa, fb, fc;
    Buf<uint16_t> ha, hb;
    const size_t na = size_t(M) * K, nb = size_t(K) * N;
    DDH_HIP_CHECK(hipMemcpy(fa.Get(na), a, na * sizeof(float), hipMemcpyHostToDevice));
    DDH_HIP_CHECK(hipMemcpy(fb.Get(nb), b, nb * sizeof(float), hipMemcpyHostToDevice));
    ToBf16(fa.ptr, ha.Get(na), na, nullptr);
    ToBf16(fb.ptr, hb.Get(nb), nb, nullptr);
    const int lda = trans_a ? M : K;
    const int ldb = trans_b ? K : N;
    Gemm(ha.ptr, lda, trans_a, hb.ptr, ldb, trans_b, fc.Get(size_t(M) * N), N, M, N, K, nullptr,
         false, false, nullptr);
    DDH_HIP_CHECK(hipMemcpy(c, fc.ptr, size_t(M) * N * sizeof(float), hipMemcpyDeviceToHost));
  });
}

int ddh_warmstart_features(const double* g0, const double* next1, const double* ga,
                           const double* gb, const double* den, int32_t n_cb, int32_t n_ant,
                           int32_t n_sub, float* features) {
  return Guard([&] { WarmFeaturesHost(g0, next1, ga, gb, den, n_cb, n_ant, n_sub, 2, features); });
}

int ddh_warmstart_apply(const double* prev, const float* pred, int64_t n_pairs, double* init) {
  return Guard([&] { WarmApplyHost(prev, pred, n_pairs, 2, init); });
}

int ddh_warmstart_features_2x2(const double* g0, const double* next1, const double* ga,
                               const double* gb, const double* den, int32_t n_cb, int32_t n_ant,
                               int32_t n_sub, float* features) {
  return Guard([&] { WarmFeaturesHost(g0, next1, ga, gb, den, n_cb, n_ant, n_sub, 4, features); });
}

int ddh_warmstart_apply_2x2(const double* prev, const float* pred, int64_t n_solutions,
                            double* init) {
  return Guard([&] { WarmApplyHost(prev, pred, n_solutions, 4, init); });
}

}  // extern "C"
