// Host side of the solver library: device-resident problems in the
// baseline-fastest layout, the iteration driver, and the C-ABI of
// include/ddecal_hip.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/ddecal_hip.h"
#include "common.h"
#include "solver_kernels.h"
#include "warm.h"

namespace ddh {

namespace {
thread_local std::string g_last_error;

// Scratch of ddh_chi2_device, the entry point that allocates nothing: the
// per-workgroup partial sums of the chi^2 kernels (one double per
// 64-baseline tile and channel block), at most kNoAllocChi2Partials of them
// (512 KiB). Each GPU holds its own copy of the array for the life of the
// process. The array's name is a host-side stand-in, not a device address:
// host code uses only the per-device address that hipGetSymbolAddress
// returns (NoAllocScratch).
constexpr int kNoAllocChi2Partials = 1 << 16;
constexpr int kNoAllocMaxDevices = 64;
double* g_scratch_ptr[kNoAllocMaxDevices] = {};
std::mutex g_scratch_mutex;
}  // namespace

static __device__ __align__(256) double scratch[kNoAllocChi2Partials];

// The address of the current device's copy of `scratch`, looked up once per
// device (at solver creation, or the first ddh_chi2_device call on a device).
double* NoAllocScratch() {
  int dev = 0;
  DDH_HIP_CHECK(hipGetDevice(&dev));
  DDH_REQUIRE(dev >= 0 && dev < kNoAllocMaxDevices, -3, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    DDH_HIP_CHECK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    DDH_REQUIRE(ptr != nullptr, -3, "no device address for the chi^2 scratch");
    g_scratch_ptr[dev] = static_cast<double*>(ptr);
  }
  return g_scratch_ptr[dev];
}

void SetLastError(const std::string& message) { g_last_error = message; }
const char* LastError() { return g_last_error.c_str(); }

namespace {

constexpr int kWave = 64;

template <typename T>
struct DeviceBuffer {
  T* ptr = nullptr;
  size_t count = 0;
  DeviceBuffer() = default;
  DeviceBuffer(const DeviceBuffer&) = delete;
  DeviceBuffer& operator=(const DeviceBuffer&) = delete;
  ~DeviceBuffer() { Free(); }
  void Free() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    count = 0;
  }
  void Resize(size_t n) {
    if (n <= count && ptr) return;
    Free();
    if (n == 0) return;
    DDH_HIP_CHECK(hipMalloc(reinterpret_cast<void**>(&ptr), n * sizeof(T)));
    count = n;
  }
  void Upload(const std::vector<T>& v, hipStream_t s) {
    Resize(std::max<size_t>(v.size(), 1));
    if (!v.empty())
      DDH_HIP_CHECK(hipMemcpyAsync(ptr, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
  }
};

// Pinned host memory for staging solution transfers: one DMA per interval
// instead of one pageable copy per channel block.
struct PinnedBuffer {
  unsigned char* ptr = nullptr;
  size_t bytes = 0;
  PinnedBuffer() = default;
  PinnedBuffer(const PinnedBuffer&) = delete;
  PinnedBuffer& operator=(const PinnedBuffer&) = delete;
  ~PinnedBuffer() {
    if (ptr) (void)hipHostFree(ptr);
  }
  template <typename T>
  T* Get(size_t count) {
    const size_t n = count * sizeof(T);
    if (n > bytes || !ptr) {
      if (ptr) (void)hipHostFree(ptr);
      ptr = nullptr;
      bytes = 0;
      DDH_HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&ptr), std::max<size_t>(n, 1),
                                  hipHostMallocDefault));
      bytes = n;
    }
    return reinterpret_cast<T*>(ptr);
  }
};

struct SolverConfig {
  ddh_config c;
  std::vector<int32_t> n_sol_dir;
  std::vector<int32_t> sol_start;
  int n_sub = 0;
};

}  // namespace
}  // namespace ddh

struct ddh_problem {
  // structure
  int n_t = 0, n_bl = 0, bl_stride = 0, n_cb = 0, n_dir = 0, n_corr = 0, n_ant = 0;
  int in_corr = 0;  // correlations per visibility of host input (1: single, kept as Duo)
  int n_seg = 0;
  bool f16 = false;
  std::vector<int32_t> ant1, ant2, ch_per_cb, seg_t0, seg_sol;
  std::vector<ddh::CbInfo> cbinfo;
  size_t data_vis = 0, model_vis = 0;  // visibility counts of the device planes
  // device
  ddh::DeviceBuffer<unsigned char> data, model;
  ddh::DeviceBuffer<float> scale;
  ddh::DeviceBuffer<int32_t> d_ant1, d_ant2, d_csr_off, d_csr, d_seg_t0, d_seg_sol;
  ddh::DeviceBuffer<ddh::CbInfo> d_cbinfo;
  // Walk order of the fused Gram iteration (at most 64 antennas, each
  // antenna pair at most once): walk_rows = floor(n_ant / 2) rows of 64
  // slots, slot (row, lane) = the baseline {lane, (lane + row + 1) mod n_ant};
  // walk_rows = 0 when the baselines admit no walk.
  int walk_rows = 0;
  std::vector<int32_t> walk_code, walk_bl;
  ddh::DeviceBuffer<int32_t> d_walk_code, d_walk_bl;
  ddh::DeviceBuffer<float> staging;
  ddh::DeviceBuffer<uint32_t> max_bits;
  // raw-buffer ingest: the DPBuffer baseline list the map was built for,
  // DPBuffer baseline -> solve-data baseline (-1: auto-correlation), first
  // channel per block, and staging of host buffers
  std::vector<int32_t> ingest_a1, ingest_a2, ingest_cross, ingest_ch_start, ingest_ch_cb;
  ddh::DeviceBuffer<int32_t> d_ingest_cross, d_ingest_ch_start, d_ingest_ch_cb;
  ddh::DeviceBuffer<unsigned char> ingest_staging;
  hipStream_t stream = nullptr;

  size_t DataVisBytes() const { return size_t(n_corr) * 8; }
  size_t ModelVisBytes() const { return f16 ? size_t(n_corr) * 4 : size_t(n_corr) * 8; }
  size_t Rows(int cb) const { return size_t(n_t) * ch_per_cb[cb]; }
};

struct ddh_solver {
  ddh::SolverConfig cfg;
  hipStream_t stream = nullptr;
  // Bumped whenever something a batch copies out of the solver at build time
  // changes: a device buffer's address (reallocation for another problem
  // shape, constraint weights of another size) or the smoothness settings.
  // A batch compares it with the version it was built from and rebuilds its
  // arguments, launch plan and graph when they differ.
  uint64_t version = 0;
  int waves = 4;
  int unroll = 1;
  int use_graph = 0;
  int poll = 8;
  int fuse = 1;       // fused Gram iteration where the problem allows it (1: merged
                      // finishing step, 2: a finishing kernel per iteration)
  int gs_groups = 2;  // its row groups per polarization (ddh_set_gram_solve_groups)
  // antenna solve schedule (ddh_set_antenna_solve_schedule; 0: the default)
  int solve_split = 0;
  int solve_threads = 0;
  // per-solve device buffers
  ddh::DeviceBuffer<double2> sol0, sol1;
  ddh::DeviceBuffer<float2> gram;       // Gram iteration cross products
  ddh::DeviceBuffer<float2> gram_walk;  // the same in walk order (fused iteration)
  ddh::DeviceBuffer<double> gs_part;     // fused iteration: per-workgroup sums
  ddh::DeviceBuffer<double> mean, wg_conv, wg_mean, step_mags;
  ddh::DeviceBuffer<int32_t> wg_conv_n, wg_mean_n;
  ddh::DeviceBuffer<int32_t> arrive;  // merged fused iteration: arrival count per launch
  ddh::DeviceBuffer<float> partials;
  ddh::DeviceBuffer<float> fj_products;  // full-Jones model products (once per solve)
  ddh::DeviceBuffer<ddh::DeviceState> state;
  ddh::DeviceBuffer<int32_t> n_sol_dir, sol_start;
  ddh::DeviceBuffer<double> den, chi2;  // one-step denominators; chi^2 partials + total
  // captured iteration body (poll iterations), valid for graph_key
  hipGraphExec_t graph_exec = nullptr;
  std::vector<const void*> graph_key;
  ddh::DeviceState* host_state = nullptr;  // pinned, two slots (pipelined polling)
  hipEvent_t poll_ev[2] = {nullptr, nullptr};
  ddh::PinnedBuffer stage;                 // pinned staging of solutions
  // timing of the last solve
  std::vector<hipEvent_t> events;
  int profile = 0;
  double last_correlate_ms = 0.0;
  int64_t last_correlate_launches = 0;
  double last_solve_ms = 0.0;
  // cached problem of the drop-in path
  std::unique_ptr<ddh_problem, void (*)(ddh_problem*)> cached{nullptr, nullptr};
  // smoothness constraint (on when sm_freq is set): the kernel windows and
  // frequency corrections per block, weights, factors, smoothed solutions
  ddh::DeviceBuffer<double> sm_freq, sm_corr, sm_weights, sm_ant_factor, sm_dd_factor;
  ddh::DeviceBuffer<int32_t> sm_window;
  ddh::DeviceBuffer<double2> sm_out;
  bool sm_on = false;
  bool sm_sub_weights = false;
  double sm_bandwidth = 0.0;
  int sm_truncate = 1;
};

namespace ddh {
namespace {

// Correlations of the device planes: single-correlation data (one value,
// 0.5 (XX + YY), per visibility) is kept as Duo pairs of equal values, which a
// scalar solve treats exactly as the single value (both terms of its sums are
// equal, so they double exactly and the quotient is unchanged).
int DeviceCorrelations(const ddh_config& c) { return c.n_correlations == 1 ? 2 : c.n_correlations; }

void ValidateConfig(const ddh_config& c) {
  DDH_REQUIRE(c.iteration == DDH_ITERATION_STREAM || c.iteration == DDH_ITERATION_GRAM, -1,
              "unknown iteration kind");
  DDH_REQUIRE(c.iteration != DDH_ITERATION_GRAM ||
                  ((DeviceCorrelations(c) == 2 || c.n_solution_polarizations == 4) &&
                   c.n_directions <= kMaxGramDirections),
              -4, "the Gram iteration is built for Duo data (diagonal and scalar solvers) and the "
                  "full-Jones solver, with at most 8 directions");
  DDH_REQUIRE(c.iteration != DDH_ITERATION_GRAM || c.n_solution_polarizations == 4 ||
                  ApplyGramSharedBytes(c.n_antennas, c.n_directions) <= kApplyGramMaxShared,
              -4, "the Gram iteration stages every antenna's gains in 64 KB of LDS "
                  "(n_antennas x (2 n_directions + 1) x 8 bytes)");
  DDH_REQUIRE(c.n_antennas > 1, -1, "n_antennas must be > 1");
  DDH_REQUIRE(c.n_directions >= 1 && c.n_directions <= 16, -4,
              "n_directions must be 1..16");
  DDH_REQUIRE(c.n_channel_blocks >= 1, -1, "n_channel_blocks must be >= 1");
  DDH_REQUIRE(c.n_correlations == 1 || c.n_correlations == 2 || c.n_correlations == 4, -1,
              "n_correlations must be 1 (single), 2 (Duo) or 4 (Full)");
  DDH_REQUIRE(c.n_correlations != 1 || c.n_solution_polarizations == 1, -1,
              "single-correlation data needs the scalar solver (n_solution_polarizations 1)");
  DDH_REQUIRE(c.n_solution_polarizations == 1 || c.n_solution_polarizations == 2 ||
                  c.n_solution_polarizations == 4,
              -1, "n_solution_polarizations must be 1 (scalar), 2 (diagonal) or 4 (full-Jones)");
  DDH_REQUIRE(c.n_solution_polarizations != 4 || c.n_correlations == 4, -1,
              "the full-Jones solver needs Full data (n_correlations 4)");
  DDH_REQUIRE(c.n_solutions_per_direction != nullptr, -1,
              "n_solutions_per_direction is required");
  DDH_REQUIRE(c.max_iterations >= 1 && c.max_iterations <= 100000, -1,
              "max_iterations must be 1..100000");
  DDH_REQUIRE(c.min_iterations >= 0, -1, "min_iterations must be >= 0");
  DDH_REQUIRE(c.step_size > 0.0, -1, "step_size must be > 0");
  DDH_REQUIRE(c.model_storage == DDH_MODEL_C64 || c.model_storage == DDH_MODEL_F16, -1,
              "model_storage must be DDH_MODEL_C64 or DDH_MODEL_F16");
  for (int d = 0; d < c.n_directions; ++d) {
    DDH_REQUIRE(c.n_solutions_per_direction[d] >= 1, -1,
                "n_solutions_per_direction entries must be >= 1");
    DDH_REQUIRE(c.n_antennas * c.n_solutions_per_direction[d] <= 1024, -4,
                "n_antennas * solutions of one direction must be <= 1024");
  }
}

// Builds segments: maximal time ranges in which every direction keeps one
// solution index.
void BuildSegments(ddh_problem& p, const uint32_t* sol_index, const SolverConfig& cfg) {
  p.seg_t0.clear();
  p.seg_sol.clear();
  for (int t = 0; t < p.n_t; ++t) {
    bool boundary = (t == 0);
    for (int d = 0; d < p.n_dir && !boundary; ++d)
      boundary = sol_index[d * p.n_t + t] != sol_index[d * p.n_t + t - 1];
    if (boundary) {
      p.seg_t0.push_back(t);
      for (int d = 0; d < p.n_dir; ++d) {
        const uint32_t s = sol_index[d * p.n_t + t];
        DDH_REQUIRE(int(s) >= cfg.sol_start[d] && int(s) < cfg.sol_start[d] + cfg.n_sol_dir[d],
                    -2, "solution index out of range for direction " + std::to_string(d));
        p.seg_sol.push_back(int32_t(s));
      }
    }
  }
  p.seg_t0.push_back(p.n_t);
  p.n_seg = int(p.seg_t0.size()) - 1;
}

// The ring walk of the fused Gram iteration over the problem's baselines:
// step s = 1 .. n/2 pairs lane a with antenna (a + s) mod n (for even n at
// s = n/2 only lanes a < n/2), which visits every antenna pair once. A slot
// records the baseline of its pair and whether the lane's antenna is the
// baseline's first (1) or second (2). Left empty (walk_rows 0) for more than
// 64 antennas or when a pair has two baselines.
void BuildWalk(ddh_problem& p) {
  p.walk_rows = 0;
  p.walk_code.clear();
  p.walk_bl.clear();
  const int n = p.n_ant;
  if (n < 2 || n > kWave) return;
  std::vector<int32_t> pair_bl(size_t(n) * n, -1);
  for (int b = 0; b < p.n_bl; ++b) {
    const int lo = std::min(p.ant1[b], p.ant2[b]), hi = std::max(p.ant1[b], p.ant2[b]);
    int32_t& e = pair_bl[size_t(lo) * n + hi];
    if (e >= 0) return;  // two baselines of one pair
    e = b;
  }
  const int rows = n / 2;
  p.walk_code.assign(size_t(rows) * kWave, 0);
  p.walk_bl.assign(size_t(rows) * kWave, -1);
  for (int r = 0; r < rows; ++r) {
    const int step = r + 1;
    for (int a = 0; a < n; ++a) {
      if (2 * step == n && a >= n / 2) continue;
      const int q = (a + step) % n;
      const int b = pair_bl[size_t(std::min(a, q)) * n + std::max(a, q)];
      if (b < 0) continue;
      p.walk_bl[size_t(r) * kWave + a] = b;
      p.walk_code[size_t(r) * kWave + a] = p.ant1[b] == a ? 1 : 2;
    }
  }
  p.walk_rows = rows;
}

ddh_problem* CreateProblem(const ddh_solver& s, int n_times, int n_baselines,
                           const int32_t* antenna1, const int32_t* antenna2,
                           const int32_t* channels_per_block, const uint32_t* sol_index) {
  const ddh_config& c = s.cfg.c;
  DDH_REQUIRE(n_times >= 1 && n_baselines >= 1, -1, "empty problem");
  auto p = std::make_unique<ddh_problem>();
  p->n_t = n_times;
  p->n_bl = n_baselines;
  p->bl_stride = (n_baselines + kWave - 1) / kWave * kWave;
  p->n_cb = c.n_channel_blocks;
  p->n_dir = c.n_directions;
  p->n_corr = DeviceCorrelations(c);
  p->in_corr = c.n_correlations;
  p->n_ant = c.n_antennas;
  p->f16 = c.model_storage == DDH_MODEL_F16;
  p->stream = s.stream;
  p->ant1.assign(antenna1, antenna1 + n_baselines);
  p->ant2.assign(antenna2, antenna2 + n_baselines);
  for (int b = 0; b < n_baselines; ++b) {
    DDH_REQUIRE(p->ant1[b] >= 0 && p->ant1[b] < p->n_ant && p->ant2[b] >= 0 &&
                    p->ant2[b] < p->n_ant,
                -1, "antenna index out of range");
    DDH_REQUIRE(p->ant1[b] != p->ant2[b], -1,
                "auto-correlations are not solve data (SolveData drops them)");
  }
  p->ch_per_cb.assign(channels_per_block, channels_per_block + p->n_cb);
  BuildSegments(*p, sol_index, s.cfg);
  // Planes: per block one data plane, then n_dir model planes.
  size_t data_off = 0, model_off = 0;
  for (int cb = 0; cb < p->n_cb; ++cb) {
    DDH_REQUIRE(p->ch_per_cb[cb] >= 1, -1, "channel blocks must hold >= 1 channel");
    const size_t plane = p->Rows(cb) * p->bl_stride;
    // Kernels address one plane with 32-bit byte offsets (with room for the
    // rows a wave reads past a segment's end).
    DDH_REQUIRE((plane + 16 * size_t(p->bl_stride)) * p->DataVisBytes() < (size_t(1) << 32), -4,
                "a channel block plane must stay below 4 GiB");
    ddh::CbInfo info{};
    info.n_ch = p->ch_per_cb[cb];
    info.data_off = int64_t(data_off);
    info.model_off = int64_t(model_off);
    p->cbinfo.push_back(info);
    data_off += plane;
    model_off += plane * p->n_dir;
  }
  p->data_vis = data_off;
  p->model_vis = model_off;
  p->data.Resize(data_off * p->DataVisBytes());
  p->model.Resize(model_off * p->ModelVisBytes());
  // Padding lanes must read zeros.
  DDH_HIP_CHECK(hipMemsetAsync(p->data.ptr, 0, data_off * p->DataVisBytes(), s.stream));
  DDH_HIP_CHECK(hipMemsetAsync(p->model.ptr, 0, model_off * p->ModelVisBytes(), s.stream));
  p->scale.Upload(std::vector<float>(size_t(p->n_cb) * p->n_dir, 1.0f), s.stream);
  // Baseline tables padded to the stride.
  std::vector<int32_t> a1(p->bl_stride, 0), a2(p->bl_stride, 0);
  std::copy(p->ant1.begin(), p->ant1.end(), a1.begin());
  std::copy(p->ant2.begin(), p->ant2.end(), a2.begin());
  p->d_ant1.Upload(a1, s.stream);
  p->d_ant2.Upload(a2, s.stream);
  // Per-antenna baseline lists, two words per entry: (baseline << 1) | side
  // (side 1 = second antenna) and the baseline's other antenna.
  std::vector<int32_t> off(p->n_ant + 1, 0), csr(size_t(4) * n_baselines);
  for (int b = 0; b < n_baselines; ++b) {
    off[p->ant1[b] + 1]++;
    off[p->ant2[b] + 1]++;
  }
  for (int a = 0; a < p->n_ant; ++a) off[a + 1] += off[a];
  std::vector<int32_t> fill(off.begin(), off.end() - 1);
  for (int b = 0; b < n_baselines; ++b) {
    const int e1 = fill[p->ant1[b]]++, e2 = fill[p->ant2[b]]++;
    csr[2 * e1] = (b << 1);
    csr[2 * e1 + 1] = p->ant2[b];
    csr[2 * e2] = (b << 1) | 1;
    csr[2 * e2 + 1] = p->ant1[b];
  }
  p->d_csr_off.Upload(off, s.stream);
  p->d_csr.Upload(csr, s.stream);
  BuildWalk(*p);
  if (p->walk_rows > 0) {
    p->d_walk_code.Upload(p->walk_code, s.stream);
    p->d_walk_bl.Upload(p->walk_bl, s.stream);
  }
  p->d_seg_t0.Upload(p->seg_t0, s.stream);
  p->d_seg_sol.Upload(p->seg_sol, s.stream);
  p->d_cbinfo.Upload(p->cbinfo, s.stream);
  DDH_HIP_CHECK(hipStreamSynchronize(s.stream));
  return p.release();
}

// Uploads one channel block given in reference order, [t][bl][ch][n_corr]
// (data) and [d][t][bl][ch][n_corr] (model), converting to the device layout.
void UploadBlock(ddh_problem& p, int cb, const float* data, const float* model) {
  const int n_ch = p.ch_per_cb[cb];
  const size_t n_vis = size_t(p.n_t) * p.n_bl * n_ch;
  const size_t vis_floats = size_t(p.in_corr) * 2;  // host input
  const size_t need = std::max(n_vis, n_vis * p.n_dir) * vis_floats;
  p.staging.Resize(need);
  // data
  DDH_HIP_CHECK(hipMemcpyAsync(p.staging.ptr, data, n_vis * vis_floats * sizeof(float),
                               hipMemcpyHostToDevice, p.stream));
  LaunchRepack(p.staging.ptr, p.data.ptr + p.cbinfo[cb].data_off * p.DataVisBytes(), 1, p.n_t,
               p.n_bl, n_ch, p.bl_stride, p.n_corr, false, nullptr, p.stream, p.in_corr == 1);
  // model
  DDH_HIP_CHECK(hipMemcpyAsync(p.staging.ptr, model,
                               n_vis * p.n_dir * vis_floats * sizeof(float),
                               hipMemcpyHostToDevice, p.stream));
  float* scale = nullptr;
  if (p.f16) {
    p.max_bits.Resize(p.n_dir);
    DDH_HIP_CHECK(hipMemsetAsync(p.max_bits.ptr, 0, p.n_dir * sizeof(uint32_t), p.stream));
    LaunchPlaneMaxAbs(p.staging.ptr, int64_t(n_vis * vis_floats), p.n_dir, p.max_bits.ptr,
                      p.stream);
    scale = p.scale.ptr + size_t(cb) * p.n_dir;
    LaunchScaleFromMax(p.max_bits.ptr, scale, p.n_dir, p.stream);
  }
  LaunchRepack(p.staging.ptr, p.model.ptr + p.cbinfo[cb].model_off * p.ModelVisBytes(),
               p.n_dir, p.n_t, p.n_bl, n_ch, p.bl_stride, p.n_corr, p.f16, scale, p.stream,
               p.in_corr == 1);
  // The staging buffer is reused by the next block.
  DDH_HIP_CHECK(hipStreamSynchronize(p.stream));
}

// Structure of host SolveData channel blocks: n_t, baseline list, channels per
// block and per-(direction, time) solution index. Verifies the
// time -> baseline -> channel order of SolveData.cc:129-181.
struct Structure {
  int n_t = 0, n_bl = 0;
  std::vector<int32_t> a1, a2, ch;
  std::vector<uint32_t> sol_index;
};

Structure DetectStructure(const ddh_cb_view* cbs, int n_cb, int n_dir) {
  Structure st;
  for (int cb = 0; cb < n_cb; ++cb) {
    const ddh_cb_view& v = cbs[cb];
    DDH_REQUIRE(v.n_visibilities > 0 && v.data && v.model && v.antenna_pairs && v.solution_map,
                -1, "channel block " + std::to_string(cb) + " is empty or incomplete");
    const uint32_t* pr = v.antenna_pairs;
    const int64_t nv = v.n_visibilities;
    // Channels in block: run length of the first pair. Baselines: runs until
    // the first pair comes back (the next time step) or the block ends.
    int64_t run = 1;
    while (run < nv && pr[2 * run] == pr[0] && pr[2 * run + 1] == pr[1]) ++run;
    int64_t n_bl = 1;
    while (n_bl * run < nv &&
           !(pr[2 * n_bl * run] == pr[0] && pr[2 * n_bl * run + 1] == pr[1]))
      ++n_bl;
    int64_t n_t = nv / (n_bl * run);
    if (n_bl == 1) {
      // A single baseline cannot separate time from channel; treat every
      // visibility as its own row.
      run = 1;
      n_t = nv;
    }
    DDH_REQUIRE(n_t * n_bl * run == nv, -2, "visibility count does not factor as time x baseline x channel");
    if (cb == 0) {
      st.n_t = int(n_t);
      st.n_bl = int(n_bl);
      for (int64_t b = 0; b < n_bl; ++b) {
        st.a1.push_back(int32_t(pr[2 * b * run]));
        st.a2.push_back(int32_t(pr[2 * b * run + 1]));
      }
      st.sol_index.resize(size_t(n_dir) * n_t);
      for (int d = 0; d < n_dir; ++d)
        for (int64_t t = 0; t < n_t; ++t)
          st.sol_index[d * n_t + t] = v.solution_map[d * nv + t * n_bl * run];
    } else {
      DDH_REQUIRE(n_t == st.n_t && n_bl == st.n_bl, -2,
                  "channel blocks differ in time steps or baselines");
    }
    st.ch.push_back(int32_t(run));
    // verify pairs and the time-only dependence of the solution map
    for (int64_t t = 0; t < n_t; ++t)
      for (int64_t b = 0; b < n_bl; ++b)
        for (int64_t c = 0; c < run; ++c) {
          const int64_t i = (t * n_bl + b) * run + c;
          DDH_REQUIRE(int32_t(pr[2 * i]) == st.a1[b] && int32_t(pr[2 * i + 1]) == st.a2[b], -2,
                      "antenna pairs do not follow the time -> baseline -> channel order");
        }
    for (int d = 0; d < n_dir; ++d)
      for (int64_t t = 0; t < n_t; ++t) {
        const uint32_t s = st.sol_index[d * n_t + t];
        const uint32_t* m = v.solution_map + d * nv + t * n_bl * run;
        DDH_REQUIRE(m[0] == s && m[n_bl * run - 1] == s, -2,
                    "solution map varies within a time step");
      }
  }
  return st;
}

// Complex values per (antenna, sub-solution) on the device: a 2x2 matrix for
// the full-Jones solver, else a pair (equal for the scalar solver).
int DevicePols(const ddh_config& c) { return c.n_solution_polarizations == 4 ? 4 : 2; }

// The fused Gram iteration (apply and antenna solve in one kernel) serves
// Gram solves of Duo data whose problem has a walk (at most 64 antennas, each
// pair once) and one solution per direction in one time segment.
bool FusedGram(const ddh_solver& s, const ddh_problem& p) {
  const ddh_config& c = s.cfg.c;
  if (!s.fuse || c.iteration != DDH_ITERATION_GRAM || c.n_solution_polarizations == 4) return false;
  if (p.n_corr != 2 || p.walk_rows == 0 || p.n_seg != 1 || p.n_dir > kMaxGramDirections)
    return false;
  for (int v : s.cfg.n_sol_dir)
    if (v != 1) return false;
  return true;
}

// The merged fused iteration (the finishing step of one iteration in the
// kernel of the next) needs the stepped solutions unchanged until the next
// iteration reads them: not with the smoothness constraint, which replaces
// them between the finishing step and the decision.
bool MergedGram(const ddh_solver& s, const IterArgs& a) {
  return a.walk_rows > 0 && s.fuse == 1 && a.sm_freq == nullptr;
}

// fused: -1 decides by FusedGram; 0 / 1 forces (a batch uses one path for all
// of its intervals).
IterArgs MakeArgs(ddh_solver& s, ddh_problem& p, int fused = -1) {
  const ddh_config& c = s.cfg.c;
  IterArgs a{};
  if (fused < 0) fused = FusedGram(s, p) ? 1 : 0;
  if (fused) {
    a.gram_walk = s.gram_walk.ptr;
    a.walk_code = p.d_walk_code.ptr;
    a.walk_bl = p.d_walk_bl.ptr;
    a.walk_rows = p.walk_rows;
    a.gs_groups = s.gs_groups;
    a.gs_part = s.gs_part.ptr;
  }
  a.data = p.data.ptr;
  a.model = p.model.ptr;
  a.model_scale = p.scale.ptr;
  a.ant1 = p.d_ant1.ptr;
  a.ant2 = p.d_ant2.ptr;
  a.csr_off = p.d_csr_off.ptr;
  a.csr = p.d_csr.ptr;
  a.cbinfo = p.d_cbinfo.ptr;
  a.seg_t0 = p.d_seg_t0.ptr;
  a.seg_sol = p.d_seg_sol.ptr;
  a.n_sol_dir = s.n_sol_dir.ptr;
  a.sol_start = s.sol_start.ptr;
  a.sol0 = s.sol0.ptr;
  a.sol1 = s.sol1.ptr;
  a.mean = s.mean.ptr;
  a.wg_conv = s.wg_conv.ptr;
  a.wg_conv_n = s.wg_conv_n.ptr;
  a.wg_mean = s.wg_mean.ptr;
  a.wg_mean_n = s.wg_mean_n.ptr;
  a.arrive = s.arrive.ptr;
  a.partials = s.partials.ptr;
  a.state = s.state.ptr;
  a.step_mags = s.step_mags.ptr;
  a.n_t = p.n_t;
  a.n_bl = p.n_bl;
  a.bl_stride = p.bl_stride;
  a.n_cb = p.n_cb;
  a.n_dir = p.n_dir;
  a.n_ant = p.n_ant;
  a.n_sub = s.cfg.n_sub;
  a.n_seg = p.n_seg;
  a.max_iterations = c.max_iterations;
  a.min_iterations = c.min_iterations;
  a.accuracy = c.accuracy;
  a.step_size = c.step_size;
  a.step_diff_sigma = c.step_diff_sigma;
  a.detect_stalling = c.detect_stalling;
  a.phase_only = c.phase_only;
  a.den_out = nullptr;
  a.gram = s.gram.ptr;
  int max_nsd = 1;
  for (int v : s.cfg.n_sol_dir) max_nsd = std::max(max_nsd, v);
  a.split = DefaultSolveSplit(c.n_antennas, max_nsd);
  if (s.solve_split > 0 && c.n_solution_polarizations != 4) a.split = s.solve_split;
  a.max_nsd = max_nsd;
  a.scalar = c.n_solution_polarizations == 1 ? 1 : 0;
  a.full_jones = c.n_solution_polarizations == 4 ? 1 : 0;
  a.solve_chunk = a.full_jones ? 0 : AntennaSolveChunk(a);
  a.fj_products = s.fj_products.ptr;
  if (s.sm_on) {
    a.sm_freq = s.sm_freq.ptr;
    a.sm_corr = s.sm_corr.ptr;
    a.sm_window = s.sm_window.ptr;
    a.sm_weights = s.sm_weights.ptr;
    a.sm_ant_factor = s.sm_ant_factor.ptr;
    a.sm_dd_factor = s.sm_dd_factor.ptr;
    a.sm_out = s.sm_out.ptr;
    a.sm_bandwidth = s.sm_bandwidth;
    a.sm_truncate = s.sm_truncate;
    a.sm_sub_weights = s.sm_sub_weights ? 1 : 0;
  }
  return a;
}

void DestroyGraph(ddh_solver& s) {
  if (s.graph_exec) (void)hipGraphExecDestroy(s.graph_exec);
  s.graph_exec = nullptr;
  s.graph_key.clear();
}

bool GramMode(const ddh_solver& s) { return s.cfg.c.iteration == DDH_ITERATION_GRAM; }

void CheckSolveSchedule(int split, int threads) {
  DDH_REQUIRE(split >= 0 && split <= kWave && (split & (split - 1)) == 0, -1,
              "antenna solve split must be 0 (default) or a power of two up to 64");
  DDH_REQUIRE(threads >= 0 && threads <= kSolveMaxThreads && threads % kWave == 0, -1,
              "antenna solve block must be 0 (default) or a multiple of 64 up to 1024");
}

int ConfigMaxNsd(const ddh_solver& s) {
  int m = 1;
  for (int v : s.cfg.n_sol_dir) m = std::max(m, v);
  return m;
}

// The correlate partials of one iteration, by the configured iteration kind.
void LaunchCorrelation(const ddh_solver& s, const IterArgs& a, const ddh_problem& p) {
  if (GramMode(s) && a.full_jones)
    LaunchFullJonesApplyGram(a, s.stream);
  else if (GramMode(s))
    LaunchApplyGram(a, p.f16, s.stream);
  else
    LaunchCorrelate(a, p.n_corr, p.f16, s.waves, s.unroll, s.stream);
}

// Gram iteration: the cross products of the problem, before the first iteration.
void LaunchSolvePrologue(const ddh_solver& s, const IterArgs& a, const ddh_problem& p) {
  LaunchPrologue(a, s.stream);
  if (GramMode(s) && a.full_jones)
    LaunchFullJonesGram(a, p.f16, s.stream);
  else if (GramMode(s))
    LaunchGram(a, p.f16, s.waves, s.stream);
  if (a.walk_rows > 0) LaunchGramPermute(a, s.stream);
  if (MergedGram(s, a)) LaunchGramSolve(a, p.f16, s.stream, kGramSolveOnly);
  if (a.full_jones) LaunchFullJonesModelProducts(a, p.f16, s.waves, s.stream);
}

// Correlate + antenna solve of one iteration (the fused Gram iteration: one
// kernel for both; merged: the one kernel decides the previous iteration,
// finishes the next one and forms the following sums, after the solve
// prologue formed the first sums).
void LaunchCorrelateAndSolve(ddh_solver& s, const IterArgs& a, const ddh_problem& p) {
  if (a.walk_rows > 0) {
    LaunchGramSolve(a, p.f16, s.stream,
                    MergedGram(s, a) ? kGramFinishThenSolve : kGramSolveAndFinish);
    return;
  }
  LaunchCorrelation(s, a, p);
  LaunchAntennaSolve(a, p.n_corr, p.f16, s.solve_threads, s.stream);
}

void LaunchIteration(ddh_solver& s, const IterArgs& a, const ddh_problem& p) {
  LaunchCorrelateAndSolve(s, a, p);
  if (MergedGram(s, a)) return;  // the merged kernel decides
  LaunchSmoothing(a, s.stream);
  LaunchDecide(a, s.stream);
}

// Launches of the solve loop: the iteration limit, plus one for the merged
// fused iteration, whose launch decides the iteration the previous one
// finished (the last decision needs a launch of its own).
int LoopLaunches(const ddh_solver& s, const IterArgs& a, int limit) {
  return limit + (MergedGram(s, a) ? 1 : 0);
}

// Captures `poll` iterations into one graph; reused while the problem and
// the solver buffers stay the same.
void EnsureGraph(ddh_solver& s, const IterArgs& a, const ddh_problem& p) {
  const std::vector<const void*> key = {&p, p.data.ptr, p.model.ptr, s.sol0.ptr, s.partials.ptr,
                                        s.gram.ptr, s.gram_walk.ptr,
                                        reinterpret_cast<const void*>(intptr_t(a.walk_rows)),
                                        reinterpret_cast<const void*>(intptr_t(a.gs_groups)),
                                        s.fj_products.ptr, s.sm_out.ptr,
                                        s.sm_weights.ptr,
                                        reinterpret_cast<const void*>(intptr_t(s.poll)),
                                        reinterpret_cast<const void*>(intptr_t(s.waves)),
                                        reinterpret_cast<const void*>(intptr_t(s.unroll)),
                                        reinterpret_cast<const void*>(intptr_t(a.split)),
                                        reinterpret_cast<const void*>(intptr_t(s.solve_threads))};
  if (s.graph_exec && key == s.graph_key) return;
  DestroyGraph(s);
  hipGraph_t graph = nullptr;
  DDH_HIP_CHECK(hipStreamBeginCapture(s.stream, hipStreamCaptureModeThreadLocal));
  for (int i = 0; i < s.poll; ++i) LaunchIteration(s, a, p);
  DDH_HIP_CHECK(hipStreamEndCapture(s.stream, &graph));
  const hipError_t e = hipGraphInstantiate(&s.graph_exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  DDH_HIP_CHECK(e);
  s.graph_key = key;
}

void CheckMatch(const ddh_solver& s, const ddh_problem& p) {
  const ddh_config& c = s.cfg.c;
  DDH_REQUIRE(p.n_cb == c.n_channel_blocks && p.n_dir == c.n_directions &&
                  p.in_corr == c.n_correlations && p.n_ant == c.n_antennas,
              -1, "problem does not match the solver configuration");
  DDH_REQUIRE(p.f16 == (c.model_storage == DDH_MODEL_F16), -1,
              "problem model storage does not match the solver");
}

// Sizes the solver's per-solve device buffers for problem p.
void PrepareBuffers(ddh_solver& s, const ddh_problem& p) {
  const ddh_config& c = s.cfg.c;
  const size_t sol_per_cb = size_t(c.n_antennas) * s.cfg.n_sub * DevicePols(c);
  const size_t n_sol = sol_per_cb * c.n_channel_blocks;
  const size_t n_wg = size_t(c.n_channel_blocks) * c.n_directions;
  auto addresses = [&] {
    return std::vector<const void*>{s.sol0.ptr,      s.sol1.ptr,      s.mean.ptr,
                                    s.wg_conv.ptr,   s.wg_conv_n.ptr, s.wg_mean.ptr,
                                    s.wg_mean_n.ptr, s.step_mags.ptr, s.partials.ptr,
                                    s.state.ptr,     s.gram.ptr,      s.fj_products.ptr,
                                    s.gram_walk.ptr, s.gs_part.ptr,  s.arrive.ptr};
  };
  const std::vector<const void*> before = addresses();
  s.sol0.Resize(n_sol);
  s.sol1.Resize(n_sol);
  s.mean.Resize(c.n_channel_blocks);
  // two halves by iteration parity (the merged fused iteration decides one
  // iteration while it finishes the next)
  s.wg_conv.Resize(2 * n_wg);
  s.wg_conv_n.Resize(2 * n_wg);
  s.wg_mean.Resize(2 * n_wg);
  s.wg_mean_n.Resize(2 * n_wg);
  s.arrive.Resize(1);
  s.step_mags.Resize(size_t(std::max(c.max_iterations, c.min_iterations)) + 2);
  const size_t terms = c.n_solution_polarizations == 4 ? size_t(kFullJonesTerms)
                                                      : size_t(p.n_corr) * 3;
  s.partials.Resize(size_t(p.n_cb) * p.n_seg * p.n_dir * terms * p.bl_stride);
  s.state.Resize(1);
  if (c.iteration == DDH_ITERATION_GRAM)
    s.gram.Resize(size_t(p.n_cb) * p.n_seg * p.bl_stride *
                  (c.n_solution_polarizations == 4 ? FullJonesGramTerms(p.n_dir)
                                                   : 2 * GramTerms(p.n_dir)));
  if (FusedGram(s, p)) {
    s.gram_walk.Resize(size_t(p.n_cb) * 2 * GramWalkTerms(p.n_dir) * p.walk_rows * p.n_ant);
    s.gs_part.Resize(size_t(2) * p.n_cb * 2 * kGramSolveMaxGroups * p.n_dir * 3 * kWave);
  }
  if (c.n_solution_polarizations == 4)
    s.fj_products.Resize(size_t(p.n_cb) * p.n_seg * p.n_dir * kFullJonesModelTerms * p.bl_stride);
  if (addresses() != before) {
    DestroyGraph(s);
    ++s.version;
  }
}

// Solutions cross the C-ABI in the caller's layout: per (antenna,
// sub-solution) two values (diagonal solver), one (scalar solver) or four
// (full-Jones solver, a row-major 2x2 matrix). The device holds two per pair
// (equal for a scalar solver) or the four of a matrix.
void SolutionsIn(const ddh_solver& s, double2* dev, const double* host, size_t n_pairs,
                 hipStream_t stream) {
  const int np = s.cfg.c.n_solution_polarizations;
  if (np == 2 || np == 4) {
    DDH_HIP_CHECK(hipMemcpyAsync(dev, host, n_pairs * np * sizeof(double2),
                                 hipMemcpyHostToDevice, stream));
    return;
  }
  std::vector<double2> tmp(n_pairs * 2);
  for (size_t i = 0; i < n_pairs; ++i)
    tmp[2 * i] = tmp[2 * i + 1] = make_double2(host[2 * i], host[2 * i + 1]);
  DDH_HIP_CHECK(hipMemcpyAsync(dev, tmp.data(), tmp.size() * sizeof(double2),
                               hipMemcpyHostToDevice, stream));
  DDH_HIP_CHECK(hipStreamSynchronize(stream));
}

// Device pairs to the caller's layout.
void SolutionsOut(const ddh_solver& s, double* host, const void* dev, size_t n_pairs,
                  hipStream_t stream) {
  const int np = s.cfg.c.n_solution_polarizations;
  if (np == 2 || np == 4) {
    DDH_HIP_CHECK(hipMemcpyAsync(host, dev, n_pairs * np * sizeof(double2),
                                 hipMemcpyDeviceToHost, stream));
    return;
  }
  std::vector<double2> tmp(n_pairs * 2);
  DDH_HIP_CHECK(hipMemcpyAsync(tmp.data(), dev, n_pairs * 2 * sizeof(double2),
                               hipMemcpyDeviceToHost, stream));
  DDH_HIP_CHECK(hipStreamSynchronize(stream));
  for (size_t i = 0; i < n_pairs; ++i) {
    host[2 * i] = tmp[2 * i].x;
    host[2 * i + 1] = tmp[2 * i].y;
  }
}

// All channel blocks' solutions of one solver between the caller's per-block
// arrays and the device buffer `dev` ([n_cb][n_ant][n_sub][n_pol]), through
// pinned staging: the blocks are gathered (or scattered) on the host and move
// in one DMA. Diagonal and full-Jones layouts only (the scalar layout is
// converted per block by SolutionsIn / SolutionsOut).
void BulkIn(const ddh_solver& s, PinnedBuffer& stage, double2* dev, double* const* host,
            hipStream_t stream) {
  const ddh_config& c = s.cfg.c;
  const size_t per_cb = size_t(c.n_antennas) * s.cfg.n_sub * DevicePols(c);
  double2* st = stage.Get<double2>(per_cb * c.n_channel_blocks);
  DDH_HIP_CHECK(hipStreamSynchronize(stream));  // the stage may still feed an earlier copy
  for (int cb = 0; cb < c.n_channel_blocks; ++cb)
    std::memcpy(st + cb * per_cb, host[cb], per_cb * sizeof(double2));
  DDH_HIP_CHECK(hipMemcpyAsync(dev, st, per_cb * c.n_channel_blocks * sizeof(double2),
                               hipMemcpyHostToDevice, stream));
}

void BulkOut(const ddh_solver& s, PinnedBuffer& stage, double* const* host, const double2* dev,
             hipStream_t stream) {
  const ddh_config& c = s.cfg.c;
  const size_t per_cb = size_t(c.n_antennas) * s.cfg.n_sub * DevicePols(c);
  double2* st = stage.Get<double2>(per_cb * c.n_channel_blocks);
  DDH_HIP_CHECK(hipMemcpyAsync(st, dev, per_cb * c.n_channel_blocks * sizeof(double2),
                               hipMemcpyDeviceToHost, stream));
  DDH_HIP_CHECK(hipStreamSynchronize(stream));
  for (int cb = 0; cb < c.n_channel_blocks; ++cb)
    std::memcpy(host[cb], st + cb * per_cb, per_cb * sizeof(double2));
}

void Solve(ddh_solver& s, ddh_problem& p, double* const* solutions, ddh_result* result) {
  const ddh_config& c = s.cfg.c;
  CheckMatch(s, p);
  const size_t n_sols = size_t(c.n_antennas) * s.cfg.n_sub;
  const size_t sol_per_cb = n_sols * DevicePols(c);
  PrepareBuffers(s, p);
  if (c.n_solution_polarizations == 1) {
    for (int cb = 0; cb < c.n_channel_blocks; ++cb)
      SolutionsIn(s, s.sol0.ptr + cb * sol_per_cb, solutions[cb], n_sols, s.stream);
  } else {
    BulkIn(s, s.stage, s.sol0.ptr, solutions, s.stream);
  }
  const IterArgs a = MakeArgs(s, p);
  const int limit = LoopLaunches(s, a, std::max(c.max_iterations, c.min_iterations));
  const bool graph = s.use_graph && !s.profile;

  hipEvent_t t0 = nullptr, t1 = nullptr;
  DDH_HIP_CHECK(hipEventCreate(&t0));
  DDH_HIP_CHECK(hipEventCreate(&t1));
  if (s.profile && int(s.events.size()) < 2 * limit) {
    for (hipEvent_t e : s.events) (void)hipEventDestroy(e);
    s.events.assign(2 * limit, nullptr);
    for (hipEvent_t& e : s.events) DDH_HIP_CHECK(hipEventCreate(&e));
  }
  if (graph) EnsureGraph(s, a, p);
  DDH_HIP_CHECK(hipEventRecord(t0, s.stream));
  LaunchSolvePrologue(s, a, p);
  for (hipEvent_t& e : s.poll_ev)
    if (!e) DDH_HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  // Pipelined polling: the next chunk of `poll` iterations is queued before
  // the host waits for the state after the current one, so the device does
  // not idle while the host reads it; a chunk queued after the solve stopped
  // returns at once (every kernel reads the done flag first).
  int launched = 0;
  auto enqueue = [&](int slot) {
    if (graph) {
      DDH_HIP_CHECK(hipGraphLaunch(s.graph_exec, s.stream));
      launched += s.poll;
    } else {
      const int k = std::min(s.poll, limit - launched);
      for (int i = 0; i < k; ++i) {
        const int it = launched + i;
        // events bracket the correlate (fused Gram iteration: apply and
        // antenna solve in one kernel)
        if (s.profile) DDH_HIP_CHECK(hipEventRecord(s.events[2 * it], s.stream));
        if (a.walk_rows > 0) {
          LaunchGramSolve(a, p.f16, s.stream,
                          MergedGram(s, a) ? kGramFinishThenSolve : kGramSolveAndFinish);
          if (s.profile) DDH_HIP_CHECK(hipEventRecord(s.events[2 * it + 1], s.stream));
        } else {
          LaunchCorrelation(s, a, p);
          if (s.profile) DDH_HIP_CHECK(hipEventRecord(s.events[2 * it + 1], s.stream));
          LaunchAntennaSolve(a, p.n_corr, p.f16, s.solve_threads, s.stream);
        }
        if (!MergedGram(s, a)) {  // the merged kernel decides
          LaunchSmoothing(a, s.stream);
          LaunchDecide(a, s.stream);
        }
      }
      launched += k;
    }
    DDH_HIP_CHECK(hipMemcpyAsync(s.host_state + slot, s.state.ptr, sizeof(DeviceState),
                                 hipMemcpyDeviceToHost, s.stream));
    DDH_HIP_CHECK(hipEventRecord(s.poll_ev[slot], s.stream));
  };
  int slot = 0;
  enqueue(slot);
  for (;;) {
    const bool more = launched < limit;
    if (more) enqueue(slot ^ 1);
    DDH_HIP_CHECK(hipEventSynchronize(s.poll_ev[slot]));
    if (s.host_state[slot].done || !more) break;
    slot ^= 1;
  }
  DDH_HIP_CHECK(hipEventRecord(t1, s.stream));
  DDH_HIP_CHECK(hipStreamSynchronize(s.stream));
  const DeviceState st = s.host_state[slot];
  DDH_REQUIRE(st.done, -3, "solver did not reach its stopping criterion");
  const double2* final_sol = (st.iteration & 1) ? s.sol1.ptr : s.sol0.ptr;
  if (c.n_solution_polarizations == 1) {
    for (int cb = 0; cb < c.n_channel_blocks; ++cb)
      SolutionsOut(s, solutions[cb], final_sol + cb * sol_per_cb, n_sols, s.stream);
    DDH_HIP_CHECK(hipStreamSynchronize(s.stream));
  } else {
    BulkOut(s, s.stage, solutions, final_sol, s.stream);
  }
  float ms = 0.0f;
  DDH_HIP_CHECK(hipEventElapsedTime(&ms, t0, t1));
  s.last_solve_ms = ms;
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
  s.last_correlate_ms = 0.0;
  s.last_correlate_launches = st.iteration;
  if (s.profile) {
    for (int it = 0; it < st.iteration; ++it) {
      float e = 0.0f;
      DDH_HIP_CHECK(hipEventElapsedTime(&e, s.events[2 * it], s.events[2 * it + 1]));
      s.last_correlate_ms += e;
    }
  }
  if (result) {
    result->iterations = st.has_converged ? st.iteration : st.iteration + 1;
    result->constraint_iterations = 0;
    result->has_converged = st.has_converged;
    result->n_step_magnitudes = st.iteration;
    result->last_step_magnitude = st.last_step;
  }
}

}  // namespace
}  // namespace ddh

namespace ddh {

// Maps the DPBuffer baseline list (auto-correlations included) of raw
// interval buffers onto the problem's baselines: the cross-correlations must
// be the problem's baselines in order (SolveData.cc:96-100, 129-137).
void PrepareIngest(ddh_problem& p, int n_bl_all, const int32_t* a1, const int32_t* a2) {
  DDH_REQUIRE(n_bl_all >= 1 && a1 && a2, -1, "empty baseline list");
  if (p.ingest_a1.size() == size_t(n_bl_all) && std::equal(a1, a1 + n_bl_all, p.ingest_a1.begin()) &&
      std::equal(a2, a2 + n_bl_all, p.ingest_a2.begin()))
    return;
  std::vector<int32_t> cross(n_bl_all, -1);
  int k = 0;
  for (int b = 0; b < n_bl_all; ++b) {
    DDH_REQUIRE(a1[b] >= 0 && a1[b] < p.n_ant && a2[b] >= 0 && a2[b] < p.n_ant, -1,
                "antenna index out of range");
    if (a1[b] == a2[b]) continue;
    DDH_REQUIRE(k < p.n_bl && p.ant1[k] == a1[b] && p.ant2[k] == a2[b], -2,
                "the cross-correlations of the buffers differ from the problem's baselines");
    cross[b] = k++;
  }
  DDH_REQUIRE(k == p.n_bl, -2, "the buffers hold fewer cross-correlations than the problem");
  std::vector<int32_t> ch_start(p.n_cb + 1, 0), ch_cb;
  for (int cb = 0; cb < p.n_cb; ++cb) {
    ch_start[cb + 1] = ch_start[cb] + p.ch_per_cb[cb];
    ch_cb.insert(ch_cb.end(), p.ch_per_cb[cb], cb);
  }
  p.ingest_cross = std::move(cross);
  p.ingest_ch_start = std::move(ch_start);
  p.ingest_ch_cb = std::move(ch_cb);
  p.d_ingest_cross.Upload(p.ingest_cross, p.stream);
  p.d_ingest_ch_start.Upload(p.ingest_ch_start, p.stream);
  p.d_ingest_ch_cb.Upload(p.ingest_ch_cb, p.stream);
  p.max_bits.Resize(size_t(p.n_cb) * p.n_dir);
  DDH_HIP_CHECK(hipStreamSynchronize(p.stream));
  p.ingest_a1.assign(a1, a1 + n_bl_all);
  p.ingest_a2.assign(a2, a2 + n_bl_all);
}

void AssignAndWeighInto(ddh_problem& p, int n_bl_all, const int32_t* a1, const int32_t* a2,
                        const float* data, const uint8_t* flags, const float* weights,
                        const float* model, int linear, hipStream_t stream) {
  DDH_REQUIRE(data && model, -1, "NULL data or model");
  auto aligned = [](const void* ptr, uintptr_t a) { return (reinterpret_cast<uintptr_t>(ptr) % a) == 0; };
  DDH_REQUIRE(aligned(data, 16) && aligned(model, 16) && aligned(weights, 16) && aligned(flags, 4),
              -1, "data, model and weights must be 16-byte aligned, flags 4-byte aligned");
  PrepareIngest(p, n_bl_all, a1, a2);
  IngestArgs g{};
  g.data = data;
  g.flags = flags;
  g.weights = weights;
  g.model = model;
  g.cross = p.d_ingest_cross.ptr;
  g.ch_start = p.d_ingest_ch_start.ptr;
  g.ch_cb = p.d_ingest_ch_cb.ptr;
  g.cbinfo = p.d_cbinfo.ptr;
  g.out_data = p.data.ptr;
  g.out_model = p.model.ptr;
  g.scale = p.scale.ptr;
  g.max_bits = p.max_bits.ptr;
  g.n_t = p.n_t;
  g.n_bl_all = n_bl_all;
  g.n_ch = p.ingest_ch_start.back();
  g.n_cb = p.n_cb;
  g.n_dir = p.n_dir;
  g.bl_stride = p.bl_stride;
  g.linear = linear ? 1 : 0;
  g.uni = p.in_corr == 1 ? 1 : 0;
  LaunchAssignAndWeigh(g, p.n_corr, p.f16, stream);
}

ProblemView GetProblemView(ddh_problem* p) {
  ProblemView v;
  v.model = p->model.ptr;
  v.data = p->data.ptr;
  v.scale = p->scale.ptr;
  v.ant1 = p->d_ant1.ptr;
  v.ant2 = p->d_ant2.ptr;
  v.cbinfo = p->d_cbinfo.ptr;
  v.n_t = p->n_t;
  v.n_bl = p->n_bl;
  v.bl_stride = p->bl_stride;
  v.n_cb = p->n_cb;
  v.n_dir = p->n_dir;
  v.n_corr = p->n_corr;
  v.uni = p->in_corr == 1;
  v.n_ant = p->n_ant;
  v.f16 = p->f16;
  v.ch_per_cb = p->ch_per_cb;
  v.cb_host = p->cbinfo;
  v.stream = p->stream;
  return v;
}
}  // namespace ddh

// ---------------------------------------------------------------------------
// C-ABI
// ---------------------------------------------------------------------------
using namespace ddh;

extern "C" {

const char* ddh_last_error(void) { return LastError(); }

int ddh_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

void ddh_config_defaults(ddh_config* c) {
  if (!c) return;
  std::memset(c, 0, sizeof(*c));
  c->n_correlations = 2;
  c->max_iterations = 100;
  c->min_iterations = 0;
  c->accuracy = 1.0e-5;
  c->constraint_accuracy = 1.0e-4;
  c->step_size = 0.2;
  c->detect_stalling = 1;
  c->step_diff_sigma = 0.1;
  c->phase_only = 0;
  c->model_storage = DDH_MODEL_C64;
  c->iteration = DDH_ITERATION_STREAM;
  c->n_solution_polarizations = 2;
}

ddh_solver* ddh_solver_create(const ddh_config* config) {
  ddh_solver* out = nullptr;
  const int rc = Guard([&] {
    DDH_REQUIRE(config != nullptr, -1, "config is NULL");
    ValidateConfig(*config);
    int n_dev = 0;
    DDH_REQUIRE(hipGetDeviceCount(&n_dev) == hipSuccess && n_dev > 0, -3,
                "no HIP device available");
    auto s = std::make_unique<ddh_solver>();
    s->cfg.c = *config;
    s->cfg.n_sol_dir.assign(config->n_solutions_per_direction,
                            config->n_solutions_per_direction + config->n_directions);
    s->cfg.c.n_solutions_per_direction = s->cfg.n_sol_dir.data();
    int acc = 0;
    for (int v : s->cfg.n_sol_dir) {
      s->cfg.sol_start.push_back(acc);
      acc += v;
    }
    s->cfg.n_sub = acc;
    PrepareKernelAttributes();
    (void)NoAllocScratch();  // this device's scratch address, before any caller needs it
    DDH_HIP_CHECK(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
    DDH_HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&s->host_state), 2 * sizeof(DeviceState),
                                hipHostMallocDefault));
    s->n_sol_dir.Upload(s->cfg.n_sol_dir, s->stream);
    s->sol_start.Upload(s->cfg.sol_start, s->stream);
    DDH_HIP_CHECK(hipStreamSynchronize(s->stream));
    out = s.release();
  });
  return rc == 0 ? out : nullptr;
}

void ddh_problem_destroy(ddh_problem* p) {
  if (!p) return;
  delete p;
}

void ddh_solver_destroy(ddh_solver* s) {
  if (!s) return;
  (void)hipStreamSynchronize(s->stream);
  s->cached.reset();
  DestroyGraph(*s);
  for (hipEvent_t e : s->events) (void)hipEventDestroy(e);
  if (s->host_state) (void)hipHostFree(s->host_state);
  for (hipEvent_t e : s->poll_ev)
    if (e) (void)hipEventDestroy(e);
  s->sol0.Free();
  s->sol1.Free();
  s->mean.Free();
  s->wg_conv.Free();
  s->wg_conv_n.Free();
  s->wg_mean.Free();
  s->wg_mean_n.Free();
  s->step_mags.Free();
  s->partials.Free();
  s->state.Free();
  s->n_sol_dir.Free();
  s->sol_start.Free();
  s->den.Free();
  s->chi2.Free();
  (void)hipStreamDestroy(s->stream);
  delete s;
}

ddh_problem* ddh_problem_create(const ddh_solver* s, int32_t n_times, int32_t n_baselines,
                                const int32_t* antenna1, const int32_t* antenna2,
                                const int32_t* channels_per_block, const uint32_t* sol_index) {
  ddh_problem* out = nullptr;
  const int rc = Guard([&] {
    DDH_REQUIRE(s && antenna1 && antenna2 && channels_per_block && sol_index, -1,
                "NULL argument");
    out = CreateProblem(*s, n_times, n_baselines, antenna1, antenna2, channels_per_block,
                        sol_index);
  });
  return rc == 0 ? out : nullptr;
}

int ddh_problem_baseline_stride(const ddh_problem* p) { return p ? p->bl_stride : 0; }

int ddh_problem_assign_and_weight_device(ddh_problem* p, int32_t n_baselines,
                                         const int32_t* antenna1, const int32_t* antenna2,
                                         const float* data, const uint8_t* flags,
                                         const float* weights, const float* model,
                                         int32_t linear_weighting, void* stream) {
  return Guard([&] {
    DDH_REQUIRE(p, -1, "NULL problem");
    hipStream_t st = stream ? static_cast<hipStream_t>(stream) : p->stream;
    AssignAndWeighInto(*p, n_baselines, antenna1, antenna2, data, flags, weights, model,
                       linear_weighting, st);
  });
}

int ddh_problem_assign_and_weight(ddh_problem* p, int32_t n_baselines, const int32_t* antenna1,
                                  const int32_t* antenna2, const float* data,
                                  const uint8_t* flags, const float* weights, const float* model,
                                  int32_t linear_weighting, double* kernel_ms) {
  return Guard([&] {
    DDH_REQUIRE(p && data && model, -1, "NULL argument");
    DDH_REQUIRE(n_baselines >= 1, -1, "empty baseline list");
    const int n_ch = std::accumulate(p->ch_per_cb.begin(), p->ch_per_cb.end(), 0);
    const size_t n_vis = size_t(p->n_t) * n_baselines * n_ch;
    // staging: data, models, weights, flags (each 256-byte aligned)
    auto up = [](size_t b) { return (b + 255) / 256 * 256; };
    const size_t b_data = up(n_vis * 32), b_model = up(n_vis * 32 * p->n_dir);
    const size_t b_w = weights ? up(n_vis * 16) : 0, b_f = flags ? up(n_vis * 4) : 0;
    p->ingest_staging.Resize(b_data + b_model + b_w + b_f);
    unsigned char* base = p->ingest_staging.ptr;
    float* d_data = reinterpret_cast<float*>(base);
    float* d_model = reinterpret_cast<float*>(base + b_data);
    float* d_w = weights ? reinterpret_cast<float*>(base + b_data + b_model) : nullptr;
    uint8_t* d_f = flags ? base + b_data + b_model + b_w : nullptr;
    DDH_HIP_CHECK(hipMemcpyAsync(d_data, data, n_vis * 32, hipMemcpyHostToDevice, p->stream));
    DDH_HIP_CHECK(hipMemcpyAsync(d_model, model, n_vis * 32 * p->n_dir, hipMemcpyHostToDevice,
                                 p->stream));
    if (weights)
      DDH_HIP_CHECK(hipMemcpyAsync(d_w, weights, n_vis * 16, hipMemcpyHostToDevice, p->stream));
    if (flags)
      DDH_HIP_CHECK(hipMemcpyAsync(d_f, flags, n_vis * 4, hipMemcpyHostToDevice, p->stream));
    PrepareIngest(*p, n_baselines, antenna1, antenna2);
    hipEvent_t e0, e1;
    DDH_HIP_CHECK(hipEventCreate(&e0));
    DDH_HIP_CHECK(hipEventCreate(&e1));
    DDH_HIP_CHECK(hipEventRecord(e0, p->stream));
    AssignAndWeighInto(*p, n_baselines, antenna1, antenna2, d_data, d_f, d_w, d_model,
                       linear_weighting, p->stream);
    DDH_HIP_CHECK(hipEventRecord(e1, p->stream));
    DDH_HIP_CHECK(hipStreamSynchronize(p->stream));
    float ms = 0.0f;
    DDH_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (kernel_ms) *kernel_ms = ms;
  });
}

int ddh_problem_upload(ddh_problem* p, const ddh_cb_view* cbs) {
  return Guard([&] {
    DDH_REQUIRE(p && cbs, -1, "NULL argument");
    for (int cb = 0; cb < p->n_cb; ++cb) {
      const int64_t nv = int64_t(p->n_t) * p->n_bl * p->ch_per_cb[cb];
      DDH_REQUIRE(cbs[cb].n_visibilities == nv, -2,
                  "channel block " + std::to_string(cb) + " has the wrong visibility count");
      UploadBlock(*p, cb, cbs[cb].data, cbs[cb].model);
    }
  });
}

void* ddh_problem_data_ptr(ddh_problem* p, int32_t cb) {
  if (!p || cb < 0 || cb >= p->n_cb) return nullptr;
  return p->data.ptr + p->cbinfo[cb].data_off * p->DataVisBytes();
}

void* ddh_problem_model_ptr(ddh_problem* p, int32_t cb, int32_t d) {
  if (!p || cb < 0 || cb >= p->n_cb || d < 0 || d >= p->n_dir) return nullptr;
  const size_t plane = p->Rows(cb) * p->bl_stride;
  return p->model.ptr + (p->cbinfo[cb].model_off + d * plane) * p->ModelVisBytes();
}

int ddh_problem_set_model_scale(ddh_problem* p, const float* scale) {
  return Guard([&] {
    DDH_REQUIRE(p && scale, -1, "NULL argument");
    DDH_HIP_CHECK(hipMemcpy(p->scale.ptr, scale, sizeof(float) * p->n_cb * p->n_dir,
                            hipMemcpyHostToDevice));
  });
}

int ddh_problem_download_plane(ddh_problem* p, int32_t cb, int32_t plane, void* host) {
  return Guard([&] {
    DDH_REQUIRE(p && host && cb >= 0 && cb < p->n_cb && plane >= 0 && plane <= p->n_dir, -1,
                "bad argument");
    const size_t n = p->Rows(cb) * p->bl_stride;
    const void* src = plane == 0 ? ddh_problem_data_ptr(p, cb) : ddh_problem_model_ptr(p, cb, plane - 1);
    const size_t bytes = n * (plane == 0 ? p->DataVisBytes() : p->ModelVisBytes());
    DDH_HIP_CHECK(hipStreamSynchronize(p->stream));
    DDH_HIP_CHECK(hipMemcpy(host, src, bytes, hipMemcpyDeviceToHost));
  });
}

int ddh_solve_problem(ddh_solver* s, ddh_problem* p, double* const* solutions,
                      ddh_result* result) {
  return Guard([&] {
    DDH_REQUIRE(s && p && solutions, -1, "NULL argument");
    Solve(*s, *p, solutions, result);
  });
}

int ddh_solve(ddh_solver* s, const ddh_cb_view* cbs, int32_t n_cb, double* const* solutions,
              double time, ddh_result* result) {
  (void)time;
  return Guard([&] {
    DDH_REQUIRE(s && cbs && solutions, -1, "NULL argument");
    DDH_REQUIRE(n_cb == s->cfg.c.n_channel_blocks, -1,
                "number of channel blocks differs from Initialize");
    const Structure st = DetectStructure(cbs, n_cb, s->cfg.c.n_directions);
    ddh_problem* cached = s->cached.get();
    bool reuse = cached && cached->n_t == st.n_t && cached->n_bl == st.n_bl &&
                 cached->ant1 == st.a1 && cached->ant2 == st.a2 && cached->ch_per_cb == st.ch;
    if (reuse) {
      // the segmentation depends on the solution map
      std::vector<int32_t> seg_t0, seg_sol;
      ddh_problem probe;
      probe.n_t = st.n_t;
      probe.n_dir = s->cfg.c.n_directions;
      BuildSegments(probe, st.sol_index.data(), s->cfg);
      reuse = probe.seg_t0 == cached->seg_t0 && probe.seg_sol == cached->seg_sol;
    }
    if (!reuse) {
      s->cached = std::unique_ptr<ddh_problem, void (*)(ddh_problem*)>(
          CreateProblem(*s, st.n_t, st.n_bl, st.a1.data(), st.a2.data(), st.ch.data(),
                        st.sol_index.data()),
          &ddh_problem_destroy);
    }
    ddh_problem& p = *s->cached;
    for (int cb = 0; cb < n_cb; ++cb) UploadBlock(p, cb, cbs[cb].data, cbs[cb].model);
    Solve(*s, p, solutions, result);
  });
}

int ddh_solver_last_timing(const ddh_solver* s, double* correlate_ms, int64_t* launches,
                           double* solve_ms) {
  if (!s) return -1;
  if (correlate_ms) *correlate_ms = s->last_correlate_ms;
  if (launches) *launches = s->last_correlate_launches;
  if (solve_ms) *solve_ms = s->last_solve_ms;
  return 0;
}

// New constraint weights. A captured iteration graph holds the weights
// address and kind by value: it is kept while both stay the same (DDECal sets
// the weights once per interval), and rebuilt otherwise. Batches take the
// arguments at creation or rebind.
static void UploadConstraintWeights(ddh_solver& s, const std::vector<double>& w, bool per_sub) {
  const double* before = s.sm_weights.ptr;
  s.sm_weights.Upload(w, s.stream);
  if (s.sm_weights.ptr != before || s.sm_sub_weights != per_sub) {
    DestroyGraph(s);
    ++s.version;
  }
  s.sm_sub_weights = per_sub;
  DDH_HIP_CHECK(hipStreamSynchronize(s.stream));
}

int ddh_solver_set_smoothness(ddh_solver* s, const double* channel_block_frequencies,
                              double bandwidth_hz, double bandwidth_ref_frequency_hz,
                              double spectral_exponent, int32_t kernel_truncation,
                              const double* antenna_factors, const double* dd_factors) {
  return Guard([&] {
    DDH_REQUIRE(s, -1, "NULL solver");
    DestroyGraph(*s);
    ++s->version;
    if (!channel_block_frequencies) {
      s->sm_on = false;
      return;
    }
    const ddh_config& c = s->cfg.c;
    const int n_cb = c.n_channel_blocks, n_ant = c.n_antennas, n_sub = s->cfg.n_sub;
    DDH_REQUIRE(bandwidth_hz > 0.0, -1, "the smoothing bandwidth must be > 0");
    DDH_REQUIRE(c.n_correlations == 1 || c.n_correlations == 2 || c.n_correlations == 4, -1,
                "bad correlations");
    std::vector<double> f(channel_block_frequencies, channel_block_frequencies + n_cb);
    for (int i = 1; i < n_cb; ++i)
      DDH_REQUIRE(f[i] >= f[i - 1], -1, "channel-block frequencies must ascend");
    // KernelSmoother's truncation windows move up the grid with the output
    // block; they and the frequency correction of the kernel size depend only
    // on the grid (KernelSmoother.h, Smooth).
    std::vector<int32_t> window(size_t(2) * n_cb);
    std::vector<double> corr(n_cb);
    size_t left = 0, right = 0;
    const size_t n = size_t(n_cb);
    while (right != n && f[right] < f[0] + bandwidth_hz * 0.5) ++right;
    right = std::min(right + 1, n);
    for (size_t i = 0; i != n; ++i) {
      const double local_bw = bandwidth_ref_frequency_hz == 0.0
                                  ? bandwidth_hz
                                  : bandwidth_hz * f[i] / bandwidth_ref_frequency_hz;
      size_t first = 0, last = n;
      if (kernel_truncation) {
        while (f[left] < f[i] - local_bw * 0.5) ++left;
        while (right != n && f[right] < f[i] + local_bw * 0.5) ++right;
        first = left > 0 ? left - 1 : 0;
        last = right < n ? right + 1 : n;
      }
      window[2 * i] = int32_t(first);
      window[2 * i + 1] = int32_t(last);
      corr[i] = std::pow(local_bw / bandwidth_hz, spectral_exponent);
    }
    std::vector<double> ant_f(n_ant, 1.0), dd_f(n_sub, 1.0);
    if (antenna_factors) ant_f.assign(antenna_factors, antenna_factors + n_ant);
    if (dd_factors) dd_f.assign(dd_factors, dd_factors + n_sub);
    s->sm_freq.Upload(f, s->stream);
    s->sm_corr.Upload(corr, s->stream);
    s->sm_window.Upload(window, s->stream);
    s->sm_ant_factor.Upload(ant_f, s->stream);
    s->sm_dd_factor.Upload(dd_f, s->stream);
    if (!s->sm_on) {
      // unit weights until the caller sets them
      s->sm_weights.Upload(std::vector<double>(size_t(n_ant) * n_cb, 1.0), s->stream);
      s->sm_sub_weights = false;
    }
    s->sm_out.Resize(size_t(n_cb) * n_ant * n_sub * DevicePols(s->cfg.c));
    s->sm_bandwidth = bandwidth_hz;
    s->sm_truncate = kernel_truncation ? 1 : 0;
    DDH_HIP_CHECK(hipStreamSynchronize(s->stream));
    s->sm_on = true;
  });
}

int ddh_solver_set_constraint_weights(ddh_solver* s, const double* weights) {
  return Guard([&] {
    DDH_REQUIRE(s && weights, -1, "NULL argument");
    const size_t n = size_t(s->cfg.c.n_antennas) * s->cfg.c.n_channel_blocks;
    UploadConstraintWeights(*s, std::vector<double>(weights, weights + n), false);
  });
}

int ddh_solver_set_sub_solution_weights(ddh_solver* s, const double* weights) {
  return Guard([&] {
    DDH_REQUIRE(s && weights, -1, "NULL argument");
    const size_t n = size_t(s->cfg.n_sub) * s->cfg.c.n_antennas * s->cfg.c.n_channel_blocks;
    UploadConstraintWeights(*s, std::vector<double>(weights, weights + n), true);
  });
}

int ddh_set_schedule(ddh_solver* s, int32_t waves_per_tile, int32_t rows_per_trip,
                     int32_t use_graph, int32_t poll_interval) {
  return Guard([&] {
    DDH_REQUIRE(s, -1, "NULL solver");
    DDH_REQUIRE(waves_per_tile >= 1 && waves_per_tile <= 4, -1, "waves_per_tile must be 1..4");
    DDH_REQUIRE(rows_per_trip == 1 || rows_per_trip == 2, -1, "rows_per_trip must be 1 or 2");
    DDH_REQUIRE(poll_interval >= 1, -1, "poll_interval must be >= 1");
    s->unroll = rows_per_trip;
    s->waves = waves_per_tile;
    s->use_graph = use_graph;
    s->poll = poll_interval;
  });
}

int ddh_set_gram_fusion(ddh_solver* s, int32_t enable) {
  return Guard([&] {
    DDH_REQUIRE(s, -1, "NULL solver");
    DDH_REQUIRE(enable >= 0 && enable <= 2, -1, "fusion must be 0, 1 or 2");
    if (enable != s->fuse) {
      s->fuse = enable;
      DestroyGraph(*s);
      ++s->version;
    }
  });
}

int ddh_set_gram_solve_groups(ddh_solver* s, int32_t groups) {
  return Guard([&] {
    DDH_REQUIRE(s, -1, "NULL solver");
    DDH_REQUIRE(groups >= 1 && groups <= kGramSolveMaxGroups, -1, "row groups must be 1..8");
    if (groups != s->gs_groups) {
      s->gs_groups = groups;
      DestroyGraph(*s);
    }
  });
}

int ddh_set_antenna_solve_schedule(ddh_solver* s, int32_t split, int32_t threads) {
  return Guard([&] {
    DDH_REQUIRE(s, -1, "NULL solver");
    CheckSolveSchedule(split, threads);
    if (split != s->solve_split || threads != s->solve_threads) {
      s->solve_split = split;
      s->solve_threads = threads;
      DestroyGraph(*s);
      ++s->version;  // batches built from this solver copy its split
    }
  });
}

int ddh_antenna_solve_schedule(const ddh_solver* s, int32_t* split, int32_t* threads) {
  return Guard([&] {
    DDH_REQUIRE(s && split && threads, -1, "NULL argument");
    const ddh_config& c = s->cfg.c;
    IterArgs a{};
    a.n_ant = c.n_antennas;
    a.max_nsd = ConfigMaxNsd(*s);
    a.full_jones = c.n_solution_polarizations == 4 ? 1 : 0;
    a.split = DefaultSolveSplit(a.n_ant, a.max_nsd);
    if (s->solve_split > 0 && !a.full_jones) a.split = s->solve_split;
    *split = a.split;
    *threads = SolveThreads(a, s->solve_threads);
  });
}

int ddh_set_profiling(ddh_solver* s, int32_t enable) {
  if (!s) return -1;
  s->profile = enable;
  return 0;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Batched solves: many solution intervals per launch.
// ---------------------------------------------------------------------------
struct ddh_batch {
  std::vector<ddh_solver*> solvers;
  std::vector<ddh_problem*> problems;
  // Shape of each interval's problem at build time: a rebind checks new
  // problems against these copies, so the batch never reads a problem it has
  // been moved away from (the caller may release it).
  struct Shape {
    int n_t, n_bl, n_cb, n_dir, n_corr, bl_stride;
    bool f16;
    std::vector<int32_t> ch_per_cb, seg_t0, seg_sol, ant1, ant2;
    bool operator==(const Shape& o) const {
      return n_t == o.n_t && n_bl == o.n_bl && n_cb == o.n_cb && n_dir == o.n_dir &&
             n_corr == o.n_corr && bl_stride == o.bl_stride && f16 == o.f16 &&
             ch_per_cb == o.ch_per_cb && seg_t0 == o.seg_t0 && seg_sol == o.seg_sol &&
             ant1 == o.ant1 && ant2 == o.ant2;
    }
  };
  std::vector<Shape> shapes;
  hipStream_t stream = nullptr;
  int n_corr = 2;
  bool f16 = false;
  bool gram = false;  // Gram iteration
  bool full_jones = false;  // full-Jones intervals (all or none)
  bool fused = false;       // fused Gram iteration for every interval
  bool merged = false;      // its finishing step merged into the next iteration's kernel
  bool solve_chunked = false;  // an interval's antenna solve stages its partials in chunks
  int gs_groups = 2;        // its row groups per polarization
  int max_dir = 0;
  // fused Gram iteration: items (interval, cb) grouped by direction count
  std::vector<ddh::BatchTile> gs_items;
  std::vector<ddh::TileGroup> gs_launch;  // gs_items split by direction count
  ddh::DeviceBuffer<ddh::BatchTile> d_gs_items;
  int waves = 2, unroll = 1, use_graph = 1, poll = 8;
  int limit = 0;  // largest iteration limit over the intervals
  size_t solve_lds = 0;
  int solve_threads = 0;
  // antenna solve schedule requested for every interval (0: the default)
  int solve_split_req = 0, solve_threads_req = 0;
  std::vector<ddh::IterArgs> args;
  std::vector<uint64_t> versions;      // member solver versions the arguments were built from
  std::vector<ddh::TileGroup> groups;  // tiles of one direction count each
  ddh::SmoothLaunch smooth{};          // smoothness constraint launches (none: 0 intervals)
  std::vector<ddh::BatchTile> tiles, items, pro_items;
  std::vector<ddh::BatchTile> prod_tiles;  // full-Jones model products: plain tiles
  ddh::DeviceBuffer<ddh::IterArgs> d_args;
  ddh::DeviceBuffer<ddh::BatchTile> d_tiles, d_items, d_pro_items, d_prod_tiles;
  ddh::DeviceBuffer<ddh::DeviceState> d_states;
  ddh::DeviceState* host_states = nullptr;  // pinned, two slots of n (pipelined polling)
  hipEvent_t poll_ev[2] = {nullptr, nullptr};
  ddh::PinnedBuffer stage;  // pinned staging of every interval's solutions
  hipGraphExec_t graph_exec = nullptr;
  int graph_poll = 0, graph_waves = 0, graph_unroll = 0;
  double last_ms = 0.0;
  int profile = 0;
  std::vector<hipEvent_t> events;
  double last_correlate_ms = 0.0;
  double last_solve_ms = 0.0;  // antenna solve + decision
  int64_t last_correlate_launches = 0;
  // learned warm start: argument variants and per-interval solution stages
  ddh::DeviceBuffer<ddh::IterArgs> d_args_step, d_args_aux;
  ddh::DeviceBuffer<double2> warm_g0, warm_next, warm_a, warm_b;
  ddh::DeviceBuffer<double> warm_den;
  ddh::DeviceBuffer<float> warm_x, warm_y;  // the network's input and output rows
  // every interval's place in the warm buffers (as last uploaded)
  std::vector<int64_t> warm_off;
  std::vector<ddh::WarmSpan> warm_spans;
  ddh::DeviceBuffer<int64_t> d_warm_off;
  ddh::DeviceBuffer<ddh::WarmSpan> d_warm_spans;
  hipEvent_t ev_warm0 = nullptr, ev_warm1 = nullptr;
  double last_warm_ms = 0.0;
};

namespace ddh {
namespace {

void DestroyBatchGraph(ddh_batch& b) {
  if (b.graph_exec) (void)hipGraphExecDestroy(b.graph_exec);
  b.graph_exec = nullptr;
}

ddh_batch::Shape ShapeOf(const ddh_problem& p) {
  return ddh_batch::Shape{p.n_t,   p.n_bl,     p.n_cb,   p.n_dir,   p.n_corr, p.bl_stride,
                          p.f16,   p.ch_per_cb, p.seg_t0, p.seg_sol, p.ant1,   p.ant2};
}

void BuildBatch(ddh_batch& b) {
  const int n = int(b.solvers.size());
  b.shapes.clear();
  for (const ddh_problem* p : b.problems) b.shapes.push_back(ShapeOf(*p));
  b.args.clear();
  b.tiles.clear();
  b.items.clear();
  b.pro_items.clear();
  b.prod_tiles.clear();
  b.d_states.Resize(n);
  b.limit = 0;
  b.solve_lds = 0;
  b.solve_threads = 0;
  b.solve_chunked = false;
  struct Costed {
    BatchTile t;
    int n_dir;
    double cost;
  };
  std::vector<Costed> costed;
  b.fused = b.gram;
  b.max_dir = 0;
  for (int i = 0; i < n; ++i) {
    CheckMatch(*b.solvers[i], *b.problems[i]);
    PrepareBuffers(*b.solvers[i], *b.problems[i]);
    b.fused = b.fused && FusedGram(*b.solvers[i], *b.problems[i]);
    b.max_dir = std::max(b.max_dir, b.problems[i]->n_dir);
  }
  b.gs_items.clear();
  b.gs_launch.clear();
  for (int i = 0; i < n; ++i) {
    ddh_solver& s = *b.solvers[i];
    ddh_problem& p = *b.problems[i];
    IterArgs a = MakeArgs(s, p, b.fused ? 1 : 0);
    a.gs_groups = b.gs_groups;
    a.state = b.d_states.ptr + i;
    if (b.solve_split_req > 0 && !a.full_jones) a.split = b.solve_split_req;
    b.args.push_back(a);
    b.limit = std::max(b.limit, std::max(s.cfg.c.max_iterations, s.cfg.c.min_iterations));
    b.solve_lds = std::max(b.solve_lds, AntennaSolveSharedBytes(a, p.n_corr));
    if (a.solve_chunk > 0) b.solve_chunked = true;
    b.solve_threads = std::max(b.solve_threads, SolveThreads(a, b.solve_threads_req));
    const double vis_bytes = double(p.DataVisBytes() + p.n_dir * p.ModelVisBytes());
    for (int cb = 0; cb < p.n_cb; ++cb) {
      for (int seg = 0; seg < p.n_seg; ++seg) {
        const double rows = double(p.seg_t0[seg + 1] - p.seg_t0[seg]) * p.ch_per_cb[cb];
        const int n_tiles = p.bl_stride / kWave;
        // full-Jones: one entry per (tile, direction group), tile + group * n_tiles
        const int n_groups = b.full_jones ? (p.n_dir + kFullJonesGroupSize - 1) / kFullJonesGroupSize
                                          : 1;
        for (int tile = 0; tile < n_tiles; ++tile) {
          if (b.full_jones) b.prod_tiles.push_back(BatchTile{i, cb, seg, tile});
          for (int g = 0; g < n_groups; ++g)
            costed.push_back({BatchTile{i, cb, seg, tile + g * n_tiles}, p.n_dir, rows * vis_bytes});
        }
      }
      for (int d = 0; d < p.n_dir; ++d) b.items.push_back(BatchTile{i, cb, d, 0});
      b.pro_items.push_back(BatchTile{i, cb, 0, 0});
    }
  }
  // One correlate launch per direction count; inside it longest tiles first:
  // workgroups are dispatched in order, so this is a greedy
  // longest-processing-time schedule over the compute units.
  std::stable_sort(costed.begin(), costed.end(), [](const Costed& x, const Costed& y) {
    return x.n_dir != y.n_dir ? x.n_dir > y.n_dir : x.cost > y.cost;
  });
  b.groups.clear();
  for (const Costed& c : costed) {
    if (b.groups.empty() || b.groups.back().n_dir != c.n_dir)
      b.groups.push_back(TileGroup{c.n_dir, int32_t(b.tiles.size()), 0});
    ++b.groups.back().count;
    b.tiles.push_back(c.t);
  }
  if (b.fused) {
    // fused workgroups (interval, block, polarization x row group), one launch
    // per direction count; a block's workgroups are adjacent in the list
    std::vector<std::pair<int, BatchTile>> gs;
    for (int i = 0; i < n; ++i)
      for (int cb = 0; cb < b.problems[i]->n_cb; ++cb)
        for (int part = 0; part < 2 * b.gs_groups; ++part)
          gs.push_back({b.problems[i]->n_dir, BatchTile{i, cb, part, 0}});
    std::stable_sort(gs.begin(), gs.end(),
                     [](const auto& x, const auto& y) { return x.first > y.first; });
    for (const auto& e : gs) {
      if (b.gs_launch.empty() || b.gs_launch.back().n_dir != e.first)
        b.gs_launch.push_back(TileGroup{e.first, int32_t(b.gs_items.size()), 0});
      ++b.gs_launch.back().count;
      b.gs_items.push_back(e.second);
    }
    b.d_gs_items.Upload(b.gs_items, b.stream);
  }
  b.smooth = SmoothLaunch{};
  for (const IterArgs& a : b.args) {
    if (!a.sm_freq) continue;
    b.smooth.n_intervals = int(b.args.size());
    b.smooth.max_elements =
        std::max(b.smooth.max_elements, int64_t(a.n_cb) * a.n_ant * a.n_sub * (a.full_jones ? 4 : 2));
    b.smooth.finish_threads =
        std::max(b.smooth.finish_threads, SmoothFinishThreads(a.n_ant, a.max_nsd));
  }
  // merged fused iteration: every interval asks for it and none smooths (the
  // smoothing replaces the stepped solutions before the decision)
  b.merged = b.fused && b.smooth.n_intervals == 0;
  for (const ddh_solver* s : b.solvers) b.merged = b.merged && s->fuse == 1;
  b.versions.clear();
  for (const ddh_solver* s : b.solvers) b.versions.push_back(s->version);
  b.d_args.Upload(b.args, b.stream);
  b.d_tiles.Upload(b.tiles, b.stream);
  b.d_items.Upload(b.items, b.stream);
  b.d_pro_items.Upload(b.pro_items, b.stream);
  if (b.full_jones) b.d_prod_tiles.Upload(b.prod_tiles, b.stream);
  DDH_HIP_CHECK(hipStreamSynchronize(b.stream));
  DestroyBatchGraph(b);
}

// Rebuilds the batch when a member solver changed since the build (new
// buffers, constraint settings or weights of another size): the arguments,
// launch plan and captured graph hold those by value.
void RefreshBatch(ddh_batch& b) {
  for (size_t i = 0; i < b.solvers.size(); ++i)
    if (b.solvers[i]->version != b.versions[i]) {
      BuildBatch(b);
      return;
    }
}

// One batched iteration over `args` (the batch's arguments or a variant of
// them): the correlate (Gram iteration: apply) of every tile group, then the
// antenna solve over n_items items and the decision for n_intervals intervals.
void LaunchBatchStep(ddh_batch& b, const IterArgs* args, int n_items, int n_intervals) {
  if (b.full_jones) {
    // Gram iteration: the apply over (interval, block, segment, tile) entries
    LaunchBatchFullJones(args, b.d_tiles.ptr, b.groups.data(), int(b.groups.size()),
                         b.d_items.ptr, n_items, n_intervals, b.solve_threads, b.f16, b.waves,
                         b.smooth, b.stream, b.gram ? b.d_prod_tiles.ptr : nullptr,
                         b.gram ? int(b.prod_tiles.size()) : 0, b.max_dir);
  } else if (b.fused) {
    if (n_items > 0)
      LaunchBatchGramSolve(args, b.d_gs_items.ptr, b.gs_launch.data(), int(b.gs_launch.size()),
                           b.f16, b.d_pro_items.ptr, int(b.pro_items.size()), b.stream);
    LaunchBatchSmoothDecide(args, b.d_items.ptr, n_items, n_intervals, b.smooth, b.stream);
  } else if (b.gram) {
    int apply_ant = 0;
    for (const IterArgs& a : b.args) apply_ant = std::max(apply_ant, a.n_ant);
    LaunchBatchApply(args, b.d_tiles.ptr, b.groups.data(), int(b.groups.size()), apply_ant,
                     b.d_items.ptr, n_items, n_intervals, b.solve_lds, b.solve_threads, b.f16,
                     b.smooth, b.stream, b.solve_chunked);
  } else {
    LaunchBatch(args, b.d_tiles.ptr, b.groups.data(), int(b.groups.size()), b.d_items.ptr,
                n_items, n_intervals, b.solve_lds, b.solve_threads, b.n_corr, b.f16, b.waves,
                b.unroll, b.smooth, b.stream, b.solve_chunked);
  }
}

// Gram iteration: the cross products of every interval (shared by all
// argument variants, which point at the same buffers).
void LaunchBatchCrossProducts(ddh_batch& b) {
  if (b.full_jones)
    LaunchBatchFullJonesModelProducts(b.d_args.ptr, b.d_prod_tiles.ptr, int(b.prod_tiles.size()),
                                      b.f16, b.waves, b.stream);
  if (b.gram && b.full_jones)
    LaunchBatchFullJonesGram(b.d_args.ptr, b.d_prod_tiles.ptr, int(b.prod_tiles.size()),
                             b.max_dir, b.f16, b.stream);
  else if (b.gram)
    LaunchBatchGram(b.d_args.ptr, b.d_tiles.ptr, b.groups.data(), int(b.groups.size()), b.f16,
                    b.waves, b.stream);
  if (b.fused)  // once per (interval, block)
    LaunchBatchGramPermute(b.d_args.ptr, b.d_pro_items.ptr, int(b.pro_items.size()), b.max_dir,
                           b.stream);
}

// One iteration of the solve loop. Merged fused iteration: one kernel that
// decides the iteration the previous launch finished, finishes the next one
// and forms the following sums (the loop's first sums come from
// LaunchBatchFirstSums).
void LaunchBatchIteration(ddh_batch& b) {
  if (b.merged) {
    LaunchBatchGramSolve(b.d_args.ptr, b.d_gs_items.ptr, b.gs_launch.data(),
                         int(b.gs_launch.size()), b.f16, b.d_pro_items.ptr, 0, b.stream, true);
    return;
  }
  LaunchBatchStep(b, b.d_args.ptr, int(b.items.size()), int(b.solvers.size()));
}

// Merged fused iteration: the sums of the state's iteration (the first of
// the loop), without finishing it.
void LaunchBatchFirstSums(ddh_batch& b) {
  if (b.merged)
    LaunchBatchGramSolve(b.d_args.ptr, b.d_gs_items.ptr, b.gs_launch.data(),
                         int(b.gs_launch.size()), b.f16, b.d_pro_items.ptr, 0, b.stream);
}

void EnsureBatchGraph(ddh_batch& b) {
  if (b.graph_exec && b.graph_poll == b.poll && b.graph_waves == b.waves &&
      b.graph_unroll == b.unroll)
    return;
  DestroyBatchGraph(b);
  hipGraph_t graph = nullptr;
  DDH_HIP_CHECK(hipStreamBeginCapture(b.stream, hipStreamCaptureModeThreadLocal));
  for (int i = 0; i < b.poll; ++i) LaunchBatchIteration(b);
  DDH_HIP_CHECK(hipStreamEndCapture(b.stream, &graph));
  const hipError_t e = hipGraphInstantiate(&b.graph_exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  DDH_HIP_CHECK(e);
  b.graph_poll = b.poll;
  b.graph_waves = b.waves;
  b.graph_unroll = b.unroll;
}

// Complex solution values of interval i over all its channel blocks.
size_t SolutionCount(const ddh_batch& b, int i) {
  const ddh_solver& s = *b.solvers[i];
  return size_t(s.cfg.c.n_channel_blocks) * s.cfg.c.n_antennas * s.cfg.n_sub * DevicePols(s.cfg.c);
}

// Host solutions[i][cb] -> dst(i) on the batch stream: every interval's blocks
// gathered into one pinned stage and copied with one DMA per interval
// (scalar layouts: per block, converted).
template <typename Dst>
void UploadSolutions(ddh_batch& b, double* const* const* solutions, Dst&& dst) {
  const int n = int(b.solvers.size());
  size_t total = 0;
  for (int i = 0; i < n; ++i) total += SolutionCount(b, i);
  double2* st = b.stage.Get<double2>(total);
  DDH_HIP_CHECK(hipStreamSynchronize(b.stream));  // the stage may still feed an earlier copy
  size_t off = 0;
  for (int i = 0; i < n; ++i) {
    const ddh_solver& s = *b.solvers[i];
    const size_t n_sols = size_t(s.cfg.c.n_antennas) * s.cfg.n_sub;
    const size_t per_cb = n_sols * DevicePols(s.cfg.c);
    if (s.cfg.c.n_solution_polarizations == 1) {
      for (int cb = 0; cb < s.cfg.c.n_channel_blocks; ++cb)
        SolutionsIn(s, dst(i) + cb * per_cb, solutions[i][cb], n_sols, b.stream);
    } else {
      for (int cb = 0; cb < s.cfg.c.n_channel_blocks; ++cb)
        std::memcpy(st + off + cb * per_cb, solutions[i][cb], per_cb * sizeof(double2));
      DDH_HIP_CHECK(hipMemcpyAsync(dst(i), st + off, SolutionCount(b, i) * sizeof(double2),
                                   hipMemcpyHostToDevice, b.stream));
    }
    off += SolutionCount(b, i);
  }
}

// Runs the iterations from the solutions in each interval's sol0 to every
// interval's stopping criterion and downloads the results.
void RunBatch(ddh_batch& b, double* const* const* solutions, ddh_result* results,
              bool cross_products);

void SolveBatch(ddh_batch& b, double* const* const* solutions, ddh_result* results) {
  RefreshBatch(b);
  UploadSolutions(b, solutions, [&](int i) { return b.args[i].sol0; });
  b.last_warm_ms = 0.0;
  RunBatch(b, solutions, results, true);
}

// The learned warm start on the device, for every interval at once:
//   next1 = one undamped update of g0 (with its denominators),
//   ga    = `stretch` damped iterations from g0, gb = `stretch` more from ga,
//   x     = features(g0, next1, ga, gb, den), pred = network(x),
//   start = gb * exp(pred) (the network predicts the correction left after
//           the stretches, so their progress is kept),
// then the solve from `start`. Matches the per-interval host sequence
// ddh_one_step, two ddh_solve_problem calls of `stretch` iterations at
// accuracy 0 without stall detection, ddh_warmstart_features,
// ddh_mlp_forward and ddh_warmstart_apply.
void SolveBatchWarm(ddh_batch& b, ddh_mlp& net, int stretch, double* const* const* solutions,
                    ddh_result* results) {
  CheckWarmNetwork(net);
  RefreshBatch(b);
  const int n = int(b.solvers.size());
  std::vector<size_t> off(n + 1, 0);
  for (int i = 0; i < n; ++i) {
    DDH_REQUIRE(stretch >= 1 && stretch <= std::max(b.solvers[i]->cfg.c.max_iterations,
                                                    b.solvers[i]->cfg.c.min_iterations),
                -1, "stretch must be 1..max_iterations");
    off[i + 1] = off[i] + SolutionCount(b, i);
  }
  const size_t total = off[n];
  b.warm_g0.Resize(total);
  b.warm_next.Resize(total);
  b.warm_a.Resize(total);
  b.warm_b.Resize(total);
  // complex values per solution: a diagonal pair, or a 2x2 matrix (whose
  // diagonal entries the features and the start use); two denominators each
  const int ne = b.full_jones ? 4 : 2;
  b.warm_den.Resize(total / ne * 2);
  // every interval's place in the concatenated buffers, for one launch each
  std::vector<int64_t> off64(off.begin(), off.end());
  std::vector<WarmSpan> spans(n);
  int64_t max_count = 0;
  for (int i = 0; i < n; ++i) {
    const ddh_config& c = b.solvers[i]->cfg.c;
    spans[i] = WarmSpan{int64_t(off[i] / ne), c.n_channel_blocks, c.n_antennas,
                        b.solvers[i]->cfg.n_sub, ne, c.n_solution_polarizations == 1 ? 1 : 0};
    max_count = std::max<int64_t>(max_count, int64_t(off[i + 1] - off[i]));
  }
  const bool same_spans =
      off64 == b.warm_off &&
      std::equal(spans.begin(), spans.end(), b.warm_spans.begin(), b.warm_spans.end(),
                 [](const WarmSpan& x, const WarmSpan& y) {
                   return x.pair0 == y.pair0 && x.n_cb == y.n_cb && x.n_ant == y.n_ant &&
                          x.n_sub == y.n_sub && x.ne == y.ne && x.scalar == y.scalar;
                 });
  if (!same_spans) {
    b.d_warm_off.Upload(off64, b.stream);
    b.d_warm_spans.Upload(spans, b.stream);
    b.warm_off = off64;
    b.warm_spans = spans;
  }
  // undamped single step with denominators; fixed-length damped stretches
  std::vector<IterArgs> step(b.args), aux(b.args);
  for (int i = 0; i < n; ++i) {
    step[i].step_size = 1.0;
    step[i].phase_only = 0;
    step[i].den_out = b.warm_den.ptr + off[i] / ne * 2;
    aux[i].max_iterations = stretch;
    aux[i].min_iterations = 0;
    aux[i].accuracy = 0.0;
    aux[i].detect_stalling = 0;
  }
  b.d_args_step.Upload(step, b.stream);
  b.d_args_aux.Upload(aux, b.stream);
  if (!b.ev_warm0) DDH_HIP_CHECK(hipEventCreate(&b.ev_warm0));
  if (!b.ev_warm1) DDH_HIP_CHECK(hipEventCreate(&b.ev_warm1));
  WaitForNetwork(net);  // weights of any pending training
  DDH_HIP_CHECK(hipEventRecord(b.ev_warm0, b.stream));
  UploadSolutions(b, solutions, [&](int i) { return b.warm_g0.ptr + off[i]; });
  const int n_pro = int(b.pro_items.size());
  // the solver buffers of every interval <-> the concatenated buffers; the
  // argument variants share the solution buffers and states
  auto move = [&](BatchMove mode, double2* buf) {
    LaunchBatchMove(b.d_args.ptr, b.d_warm_off.ptr, n, max_count, mode, buf, b.stream);
  };
  auto run_aux = [&](double2* out) {
    LaunchBatchPrologue(b.d_args_aux.ptr, b.d_pro_items.ptr, n_pro, b.stream);
    for (int k = 0; k < stretch; ++k)
      LaunchBatchStep(b, b.d_args_aux.ptr, int(b.items.size()), n);
    move(kMoveFromCurrent, out);
  };
  // next1: iteration 0 reads sol0 and writes sol1; no decision is taken
  move(kMoveToSol0, b.warm_g0.ptr);
  LaunchBatchPrologue(b.d_args_step.ptr, b.d_pro_items.ptr, n_pro, b.stream);
  LaunchBatchCrossProducts(b);  // once for the whole warm start and solve
  LaunchBatchStep(b, b.d_args_step.ptr, int(b.items.size()), 0);
  move(kMoveFromSol1, b.warm_next.ptr);
  run_aux(b.warm_a.ptr);  // sol0 still holds g0
  move(kMoveToSol0, b.warm_a.ptr);
  run_aux(b.warm_b.ptr);
  // features of all intervals as consecutive rows, one forward pass
  const int64_t rows = int64_t(total / ne);
  b.warm_x.Resize(size_t(rows) * kWarmFeatures);
  b.warm_y.Resize(size_t(rows) * kWarmOutputs);
  float* x = b.warm_x.ptr;
  LaunchWarmFeatures(b.warm_g0.ptr, b.warm_next.ptr, b.warm_a.ptr, b.warm_b.ptr, b.warm_den.ptr,
                     b.d_warm_spans.ptr, n, max_count / ne, x, b.stream);
  MlpForwardRows(net, x, rows, b.warm_y.ptr, b.stream);
  const float* pred = b.warm_y.ptr;
  // the starts from the second stretch's iterates, through warm_a (no longer
  // needed), into every sol0
  LaunchWarmApply(b.warm_b.ptr, pred, b.d_warm_spans.ptr, n, max_count / ne, b.warm_a.ptr,
                  b.stream);
  move(kMoveToSol0, b.warm_a.ptr);
  DDH_HIP_CHECK(hipEventRecord(b.ev_warm1, b.stream));
  RunBatch(b, solutions, results, false);
  float ms = 0.0f;
  DDH_HIP_CHECK(hipEventElapsedTime(&ms, b.ev_warm0, b.ev_warm1));
  b.last_warm_ms = ms;
}

void RunBatch(ddh_batch& b, double* const* const* solutions, ddh_result* results,
              bool cross_products) {
  const int n = int(b.solvers.size());
  const bool graph = b.use_graph != 0 && !b.profile;
  // merged fused iteration: one more launch, whose decision ends the solve at
  // the iteration limit
  const int limit = b.limit + (b.merged ? 1 : 0);
  if (graph) EnsureBatchGraph(b);
  if (b.profile && int(b.events.size()) < 3 * limit) {
    for (hipEvent_t e : b.events) (void)hipEventDestroy(e);
    b.events.assign(3 * limit, nullptr);
    for (hipEvent_t& e : b.events) DDH_HIP_CHECK(hipEventCreate(&e));
  }
  hipEvent_t t0, t1;
  DDH_HIP_CHECK(hipEventCreate(&t0));
  DDH_HIP_CHECK(hipEventCreate(&t1));
  DDH_HIP_CHECK(hipEventRecord(t0, b.stream));
  LaunchBatchPrologue(b.d_args.ptr, b.d_pro_items.ptr, int(b.pro_items.size()), b.stream);
  if (cross_products) LaunchBatchCrossProducts(b);
  LaunchBatchFirstSums(b);
  for (hipEvent_t& e : b.poll_ev)
    if (!e) DDH_HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  // Pipelined polling: the next chunk of `poll` iterations is queued before
  // the host waits for the states after the current one, so the device does
  // not idle while the host reads them; a chunk queued after every interval
  // stopped returns at once (every kernel reads the done flags first).
  int launched = 0;
  auto enqueue = [&](int slot) {
    if (graph) {
      DDH_HIP_CHECK(hipGraphLaunch(b.graph_exec, b.stream));
      launched += b.poll;
    } else {
      const int k = std::min(b.poll, limit - launched);
      for (int i = 0; i < k; ++i) {
        const int it = launched + i;
        if (b.profile) {
          // correlate alone between the first two events, antenna solve and
          // decision between the last two
          DDH_HIP_CHECK(hipEventRecord(b.events[3 * it], b.stream));
          if (b.fused)  // apply and antenna solve in one kernel (its sums; merged:
                        // after finishing the previous iteration)
            LaunchBatchGramSolve(b.d_args.ptr, b.d_gs_items.ptr, b.gs_launch.data(),
                                 int(b.gs_launch.size()), b.f16, b.d_pro_items.ptr, 0, b.stream,
                                 b.merged);
          else
            LaunchBatchStep(b, b.d_args.ptr, 0, 0);
          DDH_HIP_CHECK(hipEventRecord(b.events[3 * it + 1], b.stream));
          if (b.merged) {
            // nothing: the merged kernel decides
          } else if (b.fused) {
            // the finishing kernel alone (no sums groups), then the decision
            LaunchBatchGramSolve(b.d_args.ptr, b.d_gs_items.ptr, b.gs_launch.data(), 0, b.f16,
                                 b.d_pro_items.ptr, int(b.pro_items.size()), b.stream);
            LaunchBatchSmoothDecide(b.d_args.ptr, b.d_items.ptr, int(b.items.size()),
                                    int(b.solvers.size()), b.smooth, b.stream);
          }
          else if (b.full_jones)
            LaunchBatchFullJones(b.d_args.ptr, b.d_tiles.ptr, b.groups.data(), 0, b.d_items.ptr,
                                 int(b.items.size()), int(b.solvers.size()), b.solve_threads, b.f16,
                                 b.waves, b.smooth, b.stream);
          else
            LaunchBatchTail(b.d_args.ptr, b.d_items.ptr, int(b.items.size()),
                            int(b.solvers.size()), b.solve_lds, b.solve_threads, b.n_corr, b.f16,
                            b.smooth, b.stream, b.solve_chunked);
          DDH_HIP_CHECK(hipEventRecord(b.events[3 * it + 2], b.stream));
        } else {
          LaunchBatchIteration(b);
        }
      }
      launched += k;
    }
    DDH_HIP_CHECK(hipMemcpyAsync(b.host_states + slot * n, b.d_states.ptr, n * sizeof(DeviceState),
                                 hipMemcpyDeviceToHost, b.stream));
    DDH_HIP_CHECK(hipEventRecord(b.poll_ev[slot], b.stream));
  };
  int slot = 0;
  bool all_done = false;
  enqueue(slot);
  for (;;) {
    const bool more = launched < limit;
    if (more) enqueue(slot ^ 1);
    DDH_HIP_CHECK(hipEventSynchronize(b.poll_ev[slot]));
    all_done = true;
    for (int i = 0; i < n; ++i) all_done = all_done && b.host_states[slot * n + i].done;
    if (all_done || !more) break;
    slot ^= 1;
  }
  DDH_HIP_CHECK(hipEventRecord(t1, b.stream));
  DDH_HIP_CHECK(hipStreamSynchronize(b.stream));
  DDH_REQUIRE(all_done, -3, "a batched interval did not reach its stopping criterion");
  const DeviceState* final_states = b.host_states + slot * n;
  // final solutions: one DMA per interval into the pinned stage, then
  // scattered to the caller's blocks
  size_t total = 0;
  for (int i = 0; i < n; ++i) total += SolutionCount(b, i);
  double2* stage = b.stage.Get<double2>(total);
  size_t off = 0;
  for (int i = 0; i < n; ++i) {
    const ddh_solver& s = *b.solvers[i];
    const DeviceState st = final_states[i];
    const size_t n_sols = size_t(s.cfg.c.n_antennas) * s.cfg.n_sub;
    const size_t per_cb = n_sols * DevicePols(s.cfg.c);
    const double2* fin = (st.iteration & 1) ? b.args[i].sol1 : b.args[i].sol0;
    if (s.cfg.c.n_solution_polarizations == 1) {
      for (int cb = 0; cb < s.cfg.c.n_channel_blocks; ++cb)
        SolutionsOut(s, solutions[i][cb], fin + cb * per_cb, n_sols, b.stream);
    } else {
      DDH_HIP_CHECK(hipMemcpyAsync(stage + off, fin, SolutionCount(b, i) * sizeof(double2),
                                   hipMemcpyDeviceToHost, b.stream));
    }
    off += SolutionCount(b, i);
  }
  DDH_HIP_CHECK(hipStreamSynchronize(b.stream));
  off = 0;
  for (int i = 0; i < n; ++i) {
    const ddh_solver& s = *b.solvers[i];
    const DeviceState st = final_states[i];
    const size_t per_cb = size_t(s.cfg.c.n_antennas) * s.cfg.n_sub * DevicePols(s.cfg.c);
    if (s.cfg.c.n_solution_polarizations != 1)
      for (int cb = 0; cb < s.cfg.c.n_channel_blocks; ++cb)
        std::memcpy(solutions[i][cb], stage + off + cb * per_cb, per_cb * sizeof(double2));
    off += SolutionCount(b, i);
    if (results) {
      results[i].iterations = st.has_converged ? st.iteration : st.iteration + 1;
      results[i].constraint_iterations = 0;
      results[i].has_converged = st.has_converged;
      results[i].n_step_magnitudes = st.iteration;
      results[i].last_step_magnitude = st.last_step;
    }
  }
  DDH_HIP_CHECK(hipStreamSynchronize(b.stream));
  float ms = 0.0f;
  DDH_HIP_CHECK(hipEventElapsedTime(&ms, t0, t1));
  b.last_ms = ms;
  b.last_correlate_ms = 0.0;
  b.last_solve_ms = 0.0;
  b.last_correlate_launches = 0;
  if (b.profile) {
    // launches after the last interval stopped do no work: count those that did
    int worked = 0;
    for (int i = 0; i < n; ++i) worked = std::max(worked, final_states[i].iteration);
    for (int it = 0; it < worked; ++it) {
      float e = 0.0f, f = 0.0f;
      DDH_HIP_CHECK(hipEventElapsedTime(&e, b.events[3 * it], b.events[3 * it + 1]));
      DDH_HIP_CHECK(hipEventElapsedTime(&f, b.events[3 * it + 1], b.events[3 * it + 2]));
      b.last_correlate_ms += e;
      b.last_solve_ms += f;
    }
    b.last_correlate_launches = worked;
  }
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
}

// Mean duration of `reps` correlate launches under the current schedule.
template <typename LaunchFn>
double TimeLaunches(hipStream_t stream, int reps, LaunchFn&& launch) {
  hipEvent_t e0, e1;
  DDH_HIP_CHECK(hipEventCreate(&e0));
  DDH_HIP_CHECK(hipEventCreate(&e1));
  launch();  // warm-up
  DDH_HIP_CHECK(hipEventRecord(e0, stream));
  for (int r = 0; r < reps; ++r) launch();
  DDH_HIP_CHECK(hipEventRecord(e1, stream));
  DDH_HIP_CHECK(hipEventSynchronize(e1));
  float ms = 0.0f;
  DDH_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return ms / reps;
}

// Candidate correlate schedules: (waves per tile, rows per trip).
std::vector<std::pair<int, int>> ScheduleCandidates(int n_corr) {
  std::vector<std::pair<int, int>> c;
  (void)n_corr;
  for (int w : {1, 2, 3, 4})
    for (int u : {1, 2}) c.push_back({w, u});
  return c;
}

// Antenna solve schedules (threads per (antenna, solution), block threads)
// for `items` (antenna, solution) items: each split with blocks of one, two
// and four times the threads it needs; (0, 0) is the default.
std::vector<std::pair<int, int>> SolveScheduleCandidates(int items) {
  std::vector<std::pair<int, int>> c{{0, 0}};
  for (int split : {1, 2, 4, 8, 16}) {
    const int need = std::max(kWave, ((items * split + kWave - 1) / kWave) * kWave);
    if (need > kSolveMaxThreads) break;
    for (int t = need; t <= std::min(kSolveMaxThreads, 4 * need); t *= 2) c.push_back({split, t});
  }
  return c;
}

}  // namespace
}  // namespace ddh

extern "C" {

ddh_batch* ddh_batch_create(ddh_solver* const* solvers, ddh_problem* const* problems,
                            int32_t n_intervals) {
  ddh_batch* out = nullptr;
  const int rc = Guard([&] {
    DDH_REQUIRE(solvers && problems && n_intervals >= 1, -1, "empty batch");
    auto b = std::make_unique<ddh_batch>();
    b->solvers.assign(solvers, solvers + n_intervals);
    b->problems.assign(problems, problems + n_intervals);
    for (int i = 0; i < n_intervals; ++i)
      DDH_REQUIRE(b->solvers[i] && b->problems[i], -1, "NULL interval in batch");
    b->n_corr = b->problems[0]->n_corr;
    b->f16 = b->problems[0]->f16;
    b->gram = GramMode(*b->solvers[0]);
    b->full_jones = b->solvers[0]->cfg.c.n_solution_polarizations == 4;
    for (int i = 0; i < n_intervals; ++i) {
      DDH_REQUIRE(b->problems[i]->n_corr == b->n_corr && b->problems[i]->f16 == b->f16, -4,
                  "all intervals of a batch share correlations and model storage");
      DDH_REQUIRE(GramMode(*b->solvers[i]) == b->gram, -4,
                  "all intervals of a batch share the iteration kind");
      DDH_REQUIRE((b->solvers[i]->cfg.c.n_solution_polarizations == 4) ==
                      (b->solvers[0]->cfg.c.n_solution_polarizations == 4),
                  -4, "the intervals of a batch are all full-Jones or none");
    }
    DDH_HIP_CHECK(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    DDH_HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&b->host_states),
                                2 * sizeof(DeviceState) * n_intervals, hipHostMallocDefault));
    // solvers' buffers are used from the batch stream
    for (ddh_solver* s : b->solvers) DDH_HIP_CHECK(hipStreamSynchronize(s->stream));
    BuildBatch(*b);
    out = b.release();
  });
  return rc == 0 ? out : nullptr;
}

void ddh_batch_destroy(ddh_batch* b) {
  if (!b) return;
  (void)hipStreamSynchronize(b->stream);
  DestroyBatchGraph(*b);
  for (hipEvent_t e : b->events) (void)hipEventDestroy(e);
  if (b->ev_warm0) (void)hipEventDestroy(b->ev_warm0);
  if (b->ev_warm1) (void)hipEventDestroy(b->ev_warm1);
  if (b->host_states) (void)hipHostFree(b->host_states);
  for (hipEvent_t e : b->poll_ev)
    if (e) (void)hipEventDestroy(e);
  b->d_args.Free();
  b->d_tiles.Free();
  b->d_items.Free();
  b->d_pro_items.Free();
  b->d_states.Free();
  (void)hipStreamDestroy(b->stream);
  delete b;
}

int ddh_batch_set_schedule(ddh_batch* b, int32_t waves_per_tile, int32_t rows_per_trip,
                           int32_t use_graph, int32_t poll_interval) {
  return Guard([&] {
    DDH_REQUIRE(b, -1, "NULL batch");
    DDH_REQUIRE(waves_per_tile >= 1 && waves_per_tile <= 4, -1, "waves_per_tile must be 1..4");
    DDH_REQUIRE(rows_per_trip == 1 || rows_per_trip == 2, -1, "rows_per_trip must be 1 or 2");
    DDH_REQUIRE(poll_interval >= 1, -1, "poll_interval must be >= 1");
    b->waves = waves_per_tile;
    b->unroll = rows_per_trip;
    b->use_graph = use_graph;
    b->poll = poll_interval;
    BuildBatch(*b);
  });
}

int ddh_batch_set_gram_solve_groups(ddh_batch* b, int32_t groups) {
  return Guard([&] {
    DDH_REQUIRE(b, -1, "NULL batch");
    DDH_REQUIRE(groups >= 1 && groups <= kGramSolveMaxGroups, -1, "row groups must be 1..8");
    b->gs_groups = groups;
    BuildBatch(*b);
  });
}

int ddh_batch_gram_solve_groups(const ddh_batch* b) { return b ? b->gs_groups : -1; }

int ddh_batch_set_antenna_solve_schedule(ddh_batch* b, int32_t split, int32_t threads) {
  return Guard([&] {
    DDH_REQUIRE(b, -1, "NULL batch");
    CheckSolveSchedule(split, threads);
    const int old_split = b->solve_split_req, old_threads = b->solve_threads_req;
    b->solve_split_req = split;
    b->solve_threads_req = threads;
    try {
      BuildBatch(*b);
    } catch (...) {
      b->solve_split_req = old_split;
      b->solve_threads_req = old_threads;
      BuildBatch(*b);
      throw;
    }
  });
}

int ddh_batch_antenna_solve_schedule(const ddh_batch* b, int32_t* split, int32_t* threads) {
  return Guard([&] {
    DDH_REQUIRE(b && split && threads, -1, "NULL argument");
    DDH_REQUIRE(!b->args.empty(), -1, "empty batch");
    *split = b->args[0].split;
    *threads = b->solve_threads;
  });
}

int ddh_batch_rebind(ddh_batch* b, ddh_problem* const* problems) {
  return Guard([&] {
    DDH_REQUIRE(b && problems, -1, "NULL argument");
    const int n = int(b->solvers.size());
    for (int i = 0; i < n; ++i) {
      DDH_REQUIRE(problems[i], -1, "NULL problem");
      DDH_REQUIRE(ShapeOf(*problems[i]) == b->shapes[i], -2,
                  "rebinding needs problems of the same shapes (baselines, channels per "
                  "block, time steps, directions, solution intervals, storage)");
    }
    bool changed = false;
    for (int i = 0; i < n; ++i) {
      CheckMatch(*b->solvers[i], *problems[i]);
      PrepareBuffers(*b->solvers[i], *problems[i]);
      changed = changed || b->solvers[i]->version != b->versions[i];
    }
    if (changed) {
      // a member solver changed since the build: full rebuild (new plan and graph)
      for (int i = 0; i < n; ++i) b->problems[i] = problems[i];
      BuildBatch(*b);
      return;
    }
    std::vector<IterArgs> args(n);
    for (int i = 0; i < n; ++i) {
      ddh_solver& s = *b->solvers[i];
      args[i] = MakeArgs(s, *problems[i], b->fused ? 1 : 0);
      args[i].gs_groups = b->gs_groups;
      args[i].state = b->args[i].state;
      DDH_REQUIRE(AntennaSolveSharedBytes(args[i], b->n_corr) <= b->solve_lds, -2,
                  "rebinding needs problems of the same shapes");
    }
    for (int i = 0; i < n; ++i) b->problems[i] = problems[i];
    b->args = args;
    // Same shapes: same tiles, same buffer sizes, so the captured graph (which
    // reads the interval arguments from d_args) stays valid.
    DDH_HIP_CHECK(hipMemcpyAsync(b->d_args.ptr, b->args.data(), n * sizeof(IterArgs),
                                 hipMemcpyHostToDevice, b->stream));
    DDH_HIP_CHECK(hipStreamSynchronize(b->stream));
  });
}

int ddh_batch_solve(ddh_batch* b, double* const* const* solutions, ddh_result* results) {
  return Guard([&] {
    DDH_REQUIRE(b && solutions, -1, "NULL argument");
    SolveBatch(*b, solutions, results);
  });
}

int ddh_batch_solve_warm(ddh_batch* b, ddh_mlp* net, int32_t stretch,
                         double* const* const* solutions, ddh_result* results) {
  return Guard([&] {
    DDH_REQUIRE(b && net && solutions, -1, "NULL argument");
    SolveBatchWarm(*b, *net, stretch, solutions, results);
  });
}

int ddh_batch_last_warm_time(const ddh_batch* b, double* ms) {
  if (!b || !ms) return -1;
  *ms = b->last_warm_ms;
  return 0;
}

int ddh_batch_last_time(const ddh_batch* b, double* ms) {
  if (!b || !ms) return -1;
  *ms = b->last_ms;
  return 0;
}

int ddh_batch_set_profiling(ddh_batch* b, int32_t enable) {
  if (!b) return -1;
  b->profile = enable;
  return 0;
}

int ddh_batch_last_timing(const ddh_batch* b, double* correlate_ms, double* solve_ms,
                          int64_t* launches) {
  if (!b) return -1;
  if (correlate_ms) *correlate_ms = b->last_correlate_ms;
  if (solve_ms) *solve_ms = b->last_solve_ms;
  if (launches) *launches = b->last_correlate_launches;
  return 0;
}

int ddh_batch_autotune(ddh_batch* b, int32_t* waves_out, int32_t* rows_out, double* best_ms) {
  return Guard([&] {
    DDH_REQUIRE(b, -1, "NULL batch");
    RefreshBatch(*b);
    // unit gains and a reset state make every interval live for the timing
    for (size_t i = 0; i < b->solvers.size(); ++i) {
      const ddh_solver& s = *b->solvers[i];
      const size_t n = size_t(s.cfg.c.n_antennas) * s.cfg.n_sub * 2 * s.cfg.c.n_channel_blocks;
      std::vector<double2> ones(n, make_double2(1.0, 0.0));
      DDH_HIP_CHECK(hipMemcpyAsync(b->args[i].sol0, ones.data(), n * sizeof(double2),
                                   hipMemcpyHostToDevice, b->stream));
      DDH_HIP_CHECK(hipStreamSynchronize(b->stream));
    }
    LaunchBatchPrologue(b->d_args.ptr, b->d_pro_items.ptr, int(b->pro_items.size()), b->stream);
    double best = 1e30;
    int bw = b->waves, bu = b->unroll;
    for (const auto& cand : ScheduleCandidates(b->n_corr)) {
      b->waves = cand.first;
      b->unroll = cand.second;
      BuildBatch(*b);
      const double ms = TimeLaunches(b->stream, 5, [&] {
        // correlate only (Gram iteration: the cross products, which the
        // schedule shapes): neither changes the solve state
        if (b->gram)
          LaunchBatchCrossProducts(*b);
        else
          LaunchBatchStep(*b, b->d_args.ptr, 0, 0);
      });
      if (ms < best) {
        best = ms;
        bw = cand.first;
        bu = cand.second;
      }
    }
    b->waves = bw;
    b->unroll = bu;
    BuildBatch(*b);
    if (b->fused) {
      // Row groups of the fused iteration: the walk rows of a block and
      // polarization split over 1..8 workgroups. Timed on the cross products of
      // the unit-gain state; with separate finishing kernels a launch only
      // rewrites the next solutions and the partials of the same iteration, so
      // repeats leave the state unchanged (the merged kernel advances it from a
      // fresh state per candidate).
      LaunchBatchCrossProducts(*b);
      double best_gs = 1e30;
      int bg = b->gs_groups;
      for (int groups : {1, 2, 4, 8}) {
        b->gs_groups = groups;
        BuildBatch(*b);
        // merged: a fresh state and the first sums in this layout, then the
        // kernel that decides, finishes and forms the next iteration's sums
        // (repeats run a few iterations of the unit-gain solve)
        if (b->merged) {
          LaunchBatchPrologue(b->d_args.ptr, b->d_pro_items.ptr, int(b->pro_items.size()),
                              b->stream);
          LaunchBatchFirstSums(*b);
        }
        const double ms = TimeLaunches(b->stream, 5, [&] {
          if (b->merged)
            LaunchBatchGramSolve(b->d_args.ptr, b->d_gs_items.ptr, b->gs_launch.data(),
                                 int(b->gs_launch.size()), b->f16, b->d_pro_items.ptr, 0,
                                 b->stream, true);
          else
            LaunchBatchGramSolve(b->d_args.ptr, b->d_gs_items.ptr, b->gs_launch.data(),
                                 int(b->gs_launch.size()), b->f16, b->d_pro_items.ptr,
                                 int(b->pro_items.size()), b->stream);
        });
        if (ms < best_gs) {
          best_gs = ms;
          bg = groups;
        }
      }
      b->gs_groups = bg;
      BuildBatch(*b);
    } else if (!b->full_jones) {
      // Antenna solve schedule, timed on the partials of one correlate (or
      // apply) of the unit-gain state; the solve alone, without the decision,
      // leaves the state unchanged.
      if (b->gram) LaunchBatchCrossProducts(*b);
      LaunchBatchStep(*b, b->d_args.ptr, 0, 0);
      int items = 0;
      for (const IterArgs& a : b->args) items = std::max(items, a.n_ant * std::max(1, a.max_nsd));
      double best_s = 1e30;
      std::pair<int, int> bs{b->solve_split_req, b->solve_threads_req};
      for (const auto& cand : SolveScheduleCandidates(items)) {
        b->solve_split_req = cand.first;
        b->solve_threads_req = cand.second;
        BuildBatch(*b);
        const double ms = TimeLaunches(b->stream, 5, [&] {
          LaunchBatchTail(b->d_args.ptr, b->d_items.ptr, int(b->items.size()), 0, b->solve_lds,
                          b->solve_threads, b->n_corr, b->f16, SmoothLaunch{}, b->stream,
                          b->solve_chunked);
        });
        if (ms < best_s) {
          best_s = ms;
          bs = cand;
        }
      }
      b->solve_split_req = bs.first;
      b->solve_threads_req = bs.second;
      BuildBatch(*b);
    }
    if (waves_out) *waves_out = bw;
    if (rows_out) *rows_out = bu;
    if (best_ms) *best_ms = best;
  });
}

int ddh_autotune(ddh_solver* s, ddh_problem* p, int32_t* waves_out, int32_t* rows_out,
                 double* best_ms) {
  return Guard([&] {
    DDH_REQUIRE(s && p, -1, "NULL argument");
    CheckMatch(*s, *p);
    PrepareBuffers(*s, *p);
    const size_t n = size_t(s->cfg.c.n_antennas) * s->cfg.n_sub * DevicePols(s->cfg.c) *
                     s->cfg.c.n_channel_blocks;
    std::vector<double2> ones(n, make_double2(1.0, 0.0));
    DDH_HIP_CHECK(hipMemcpyAsync(s->sol0.ptr, ones.data(), n * sizeof(double2),
                                 hipMemcpyHostToDevice, s->stream));
    IterArgs a = MakeArgs(*s, *p);
    LaunchPrologue(a, s->stream);
    double best = 1e30;
    int bw = s->waves, bu = s->unroll;
    for (const auto& cand : ScheduleCandidates(p->n_corr)) {
      // Gram iteration: the schedule shapes the once-per-solve cross products
      const double ms = TimeLaunches(s->stream, 5, [&] {
        if (GramMode(*s) && a.full_jones)
          LaunchFullJonesGram(a, p->f16, s->stream);
        else if (GramMode(*s))
          LaunchGram(a, p->f16, cand.first, s->stream);
        else
          LaunchCorrelate(a, p->n_corr, p->f16, cand.first, cand.second, s->stream);
      });
      if (ms < best) {
        best = ms;
        bw = cand.first;
        bu = cand.second;
      }
    }
    s->waves = bw;
    s->unroll = bu;
    DestroyGraph(*s);
    // The fused iteration's row groups, or the antenna solve schedule, timed
    // on the cross products (partials) of the unit-gain state: a launch
    // rewrites only the next solutions and the per-workgroup sums of the same
    // iteration, so repeats leave the state unchanged (the merged fused
    // kernel advances it from a fresh state per candidate).
    IterArgs t = MakeArgs(*s, *p);
    if (t.walk_rows > 0) {
      LaunchSolvePrologue(*s, t, *p);
      const bool merged = MergedGram(*s, t);
      double best_g = 1e30;
      int bg = s->gs_groups;
      for (int groups : {1, 2, 4, 8}) {
        IterArgs u = t;
        u.gs_groups = groups;
        // merged: a fresh state and the first iteration's sums in this
        // row-group layout, then the kernel that decides, finishes and forms
        // the next sums (repeats run a few iterations of the unit-gain solve)
        if (merged) {
          LaunchPrologue(u, s->stream);
          LaunchGramSolve(u, p->f16, s->stream, kGramSolveOnly);
        }
        const double ms = TimeLaunches(s->stream, 5, [&] {
          LaunchGramSolve(u, p->f16, s->stream,
                          merged ? kGramFinishThenSolve : kGramSolveAndFinish);
        });
        if (ms < best_g) {
          best_g = ms;
          bg = groups;
        }
      }
      s->gs_groups = bg;
    } else if (!t.full_jones) {
      LaunchSolvePrologue(*s, t, *p);
      LaunchCorrelation(*s, t, *p);
      double best_s = 1e30;
      std::pair<int, int> bs{s->solve_split, s->solve_threads};
      for (const auto& cand : SolveScheduleCandidates(t.n_ant * std::max(1, t.max_nsd))) {
        IterArgs u = t;
        if (cand.first > 0) u.split = cand.first;
        else u.split = DefaultSolveSplit(t.n_ant, t.max_nsd);
        const double ms = TimeLaunches(s->stream, 5, [&] {
          LaunchAntennaSolve(u, p->n_corr, p->f16, cand.second, s->stream);
        });
        if (ms < best_s) {
          best_s = ms;
          bs = cand;
        }
      }
      if (bs.first != s->solve_split || bs.second != s->solve_threads) {
        s->solve_split = bs.first;
        s->solve_threads = bs.second;
        ++s->version;
      }
    }
    DestroyGraph(*s);
    if (waves_out) *waves_out = bw;
    if (rows_out) *rows_out = bu;
    if (best_ms) *best_ms = best;
  });
}

}  // extern "C"

extern "C" {

int ddh_one_step(ddh_solver* s, ddh_problem* p, double* const* solutions_in,
                 double* const* next_out, double* const* den_out) {
  return Guard([&] {
    DDH_REQUIRE(s && p && solutions_in && next_out, -1, "NULL argument");
    CheckMatch(*s, *p);
    PrepareBuffers(*s, *p);
    const ddh_config& c = s->cfg.c;
    const size_t n_sols = size_t(c.n_antennas) * s->cfg.n_sub;  // solutions per block
    const size_t per_cb = n_sols * DevicePols(c);                // device values per block
    s->den.Resize(n_sols * 2 * c.n_channel_blocks);              // two denominators each
    for (int cb = 0; cb < c.n_channel_blocks; ++cb)
      SolutionsIn(*s, s->sol0.ptr + cb * per_cb, solutions_in[cb], n_sols, s->stream);
    IterArgs a = MakeArgs(*s, *p);
    a.step_size = 1.0;  // stepped == next: the undamped update
    a.phase_only = 0;
    a.den_out = s->den.ptr;
    LaunchSolvePrologue(*s, a, *p);
    LaunchCorrelateAndSolve(*s, a, *p);
    for (int cb = 0; cb < c.n_channel_blocks; ++cb) {
      SolutionsOut(*s, next_out[cb], s->sol1.ptr + cb * per_cb, n_sols, s->stream);
      // two per solution in every layout: a pair's, a scalar solution's
      // (both its polarizations' summed denominator), the diagonal of a 2x2 one
      if (den_out)
        DDH_HIP_CHECK(hipMemcpyAsync(den_out[cb], s->den.ptr + cb * n_sols * 2,
                                     n_sols * 2 * sizeof(double), hipMemcpyDeviceToHost,
                                     s->stream));
    }
    DDH_HIP_CHECK(hipStreamSynchronize(s->stream));
  });
}

int ddh_chi2(ddh_solver* s, ddh_problem* p, double* const* solutions, double* chi2) {
  return Guard([&] {
    DDH_REQUIRE(s && p && solutions && chi2, -1, "NULL argument");
    CheckMatch(*s, *p);
    PrepareBuffers(*s, *p);
    const ddh_config& c = s->cfg.c;
    const size_t n_sols = size_t(c.n_antennas) * s->cfg.n_sub;
    const size_t per_cb = n_sols * DevicePols(c);
    for (int cb = 0; cb < c.n_channel_blocks; ++cb)
      SolutionsIn(*s, s->sol1.ptr + cb * per_cb, solutions[cb], n_sols, s->stream);
    const IterArgs a = MakeArgs(*s, *p);
    const int n_partial = Chi2Partials(a);
    s->chi2.Resize(size_t(n_partial) + 1);
    LaunchChi2(a, p->n_corr, p->f16, s->sol1.ptr, s->chi2.ptr, n_partial,
               s->chi2.ptr + n_partial, s->stream);
    DDH_HIP_CHECK(hipMemcpyAsync(chi2, s->chi2.ptr + n_partial, sizeof(double),
                                 hipMemcpyDeviceToHost, s->stream));
    DDH_HIP_CHECK(hipStreamSynchronize(s->stream));
    // single-correlation problems hold each value in both correlations
    if (c.n_correlations == 1) *chi2 *= 0.5;
  });
}

int ddh_chi2_device(ddh_solver* s, ddh_problem* p, const double* solutions, double* chi2,
                    void* stream) {
  return Guard([&] {
    DDH_REQUIRE(s && p && solutions && chi2, -1, "NULL argument");
    CheckMatch(*s, *p);
    const ddh_config& c = s->cfg.c;
    DDH_REQUIRE(c.n_solution_polarizations != 1, -4,
                "ddh_chi2_device takes pairs or 2x2 matrices (device layout); the scalar "
                "layout needs ddh_chi2");
    const IterArgs a = MakeArgs(*s, *p);  // reads the problem's buffers, allocates nothing
    const int n_partial = Chi2Partials(a);
    DDH_REQUIRE(n_partial <= kNoAllocChi2Partials, -4,
                "ddh_chi2_device: the problem needs " + std::to_string(n_partial) +
                    " partial sums (64-baseline tiles x channel blocks); its fixed scratch holds " +
                    std::to_string(kNoAllocChi2Partials));
    double* partial = NoAllocScratch();
    // every partial is written before it is read: nothing to clear
    LaunchChi2(a, p->n_corr, p->f16, reinterpret_cast<const double2*>(solutions), partial,
               n_partial, chi2, reinterpret_cast<hipStream_t>(stream),
               c.n_correlations == 1 ? 0.5 : 1.0);
  });
}

}  // extern "C"
