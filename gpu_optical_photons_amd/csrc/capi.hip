// capi.hip -- host runtime and C-ABI of liboptphot_hip.so (include/optphot_hip.h).
//
// One context per process and device: a non-blocking HIP stream at the
// default priority, timing events, and grow-only device buffers that are
// reused across calls (their first allocation is reported as alloc_ms, as the
// reference reports its per-call cudaMalloc, cuda_util.cuh:88-96). Every
// entry point validates its arguments before touching the GPU, runs
// synchronously on the context's stream, and converts exceptions into a
// return code plus a thread-local message.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <functional>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "../../include/optphot_hip.h"
#include "batch_queue.hpp"
#include "kernels.hpp"

// Device scratch of optphot_run_gpu_tally_wavefront_noalloc, the entry point
// that allocates nothing: the queue head, the block sums and the per-block
// partial histograms of its launch. Each device holds its own copy for the
// life of the process. Host code must never pass `scratch` itself to a
// kernel (in host code the name is a stand-in, not a device address): the
// context looks its device address up once at set-up (hipGetSymbolAddress,
// Context::init) and only that address is used.
namespace noalloc_store {
constexpr size_t kScratchBytes = size_t{16} << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
}  // namespace noalloc_store

namespace {

using namespace ophip;

static_assert(sizeof(optphot_sim_params) == 72, "SimParams layout (params.hpp:40-63)");
static_assert(sizeof(optphot_tally_config) == 12, "TallyConfig layout (transport.hpp:164-168)");

thread_local std::string g_last_error;

struct InvalidArgument : std::invalid_argument {
  using std::invalid_argument::invalid_argument;
};

void hip_check(hipError_t e, const char* what, const char* file, int line) {
  if (e != hipSuccess) {
    throw std::runtime_error(std::string("HIP error: ") + hipGetErrorString(e) + " at " + file +
                             ":" + std::to_string(line) + " (" + what + ")");
  }
}
#define HIPCHK(call) hip_check((call), #call, __FILE__, __LINE__)

double ms_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

// Grow-only device buffer.
struct DevBuf {
  void* ptr = nullptr;
  size_t bytes = 0;
  // Returns true if it had to (re)allocate.
  bool ensure(size_t want) {
    if (want <= bytes && ptr) return false;
    if (ptr) HIPCHK(hipFree(ptr));
    ptr = nullptr;
    bytes = 0;
    want = std::max<size_t>(want, 256);
    HIPCHK(hipMalloc(&ptr, want));
    bytes = want;
    return true;
  }
  template <class T>
  T* as() const {
    return static_cast<T*>(ptr);
  }
};

// Grow-only pinned host buffer (hipHostMalloc), the staging ring of the
// records copy-out.
struct PinnedBuf {
  void* ptr = nullptr;
  size_t bytes = 0;
  bool ensure(size_t want) {
    if (want <= bytes && ptr) return false;
    if (ptr) HIPCHK(hipHostFree(ptr));
    ptr = nullptr;
    bytes = 0;
    HIPCHK(hipHostMalloc(&ptr, want, hipHostMallocDefault));
    bytes = want;
    return true;
  }
};

// A few host threads that copy staged records into the caller's arrays.
class CopyPool {
 public:
  explicit CopyPool(unsigned n) {
    for (unsigned i = 0; i < n; ++i) workers_.emplace_back([this] { loop(); });
  }
  ~CopyPool() {
    {
      std::lock_guard<std::mutex> l(m_);
      stop_ = true;
    }
    cv_.notify_all();
    for (auto& t : workers_) t.join();
  }
  size_t size() const { return workers_.size(); }
  // Runs every task (on the workers and the calling thread) and returns when
  // all are done.
  void run(std::vector<std::function<void()>>& tasks) {
    if (tasks.empty()) return;
    {
      std::lock_guard<std::mutex> l(m_);
      tasks_ = &tasks;
      next_ = 0;
      pending_ = tasks.size();
    }
    cv_.notify_all();
    drain();
    std::unique_lock<std::mutex> l(m_);
    done_cv_.wait(l, [this] { return pending_ == 0; });
    tasks_ = nullptr;
  }

 private:
  void drain() {
    for (;;) {
      std::function<void()>* t = nullptr;
      {
        std::lock_guard<std::mutex> l(m_);
        if (!tasks_ || next_ >= tasks_->size()) return;
        t = &(*tasks_)[next_++];
      }
      (*t)();
      std::lock_guard<std::mutex> l(m_);
      if (--pending_ == 0) done_cv_.notify_all();
    }
  }
  void loop() {
    for (;;) {
      {
        std::unique_lock<std::mutex> l(m_);
        cv_.wait(l, [this] { return stop_ || (tasks_ && next_ < tasks_->size()); });
        if (stop_) return;
      }
      drain();
    }
  }
  std::vector<std::thread> workers_;
  std::mutex m_;
  std::condition_variable cv_, done_cv_;
  std::vector<std::function<void()>>* tasks_ = nullptr;
  size_t next_ = 0, pending_ = 0;
  bool stop_ = false;
};

struct Context {
  std::mutex mu;
  bool ready = false;
  int device = 0;
  int cu_count = 0;
  size_t max_lds = 0;
  std::string name;
  hipStream_t stream = nullptr;
  hipStream_t copy_stream = nullptr;  // records copy-out, overlapping the kernels
  hipEvent_t ev[4] = {};
  DevBuf records[8];  // fate, nb, ns, t, x, y, z, path
  static constexpr int kStageSlots = 3;
  PinnedBuf stage[kStageSlots];
  std::vector<hipEvent_t> piece_ev;  // records runs: launch start / end, copy done
  std::unique_ptr<CopyPool> pool;
  DevBuf partials, scratch, tally;
  DevBuf batch_table, batch_offsets, batch_counts, batch_slots;
  DevBuf rayleigh;  // 2^24 Rayleigh cosines (DevParams::rayleigh), built at init
  DevBuf adj_weights, adj_grad, adj_queue;
  // Captured graph of the batched launch and the arguments it was built for.
  hipGraphExec_t batch_exec = nullptr;
  std::string batch_key;
  // Calibration step (forward + chi-square weights + adjoint) graph and buffers.
  DevBuf step_target, step_inv_nt, step_hist, step_loss;
  DevBuf gs_data, gs_slots;
  hipGraphExec_t step_exec = nullptr;
  std::string step_key;
  // Per-interaction path: state, queues and rows in one buffer; the captured
  // group of iterations; the pinned status words the host polls.
  DevBuf sorted;
  PinnedBuf sorted_status;
  PinnedBuf tally_host;  // pinned landing buffer of a tally's device-to-host copy
  hipGraphExec_t sorted_exec = nullptr;
  std::string sorted_key;
  // Device address of noalloc_store::scratch on each device, looked up at
  // set-up with that device current (only index `device` is used).
  static constexpr int kMaxDevices = 64;
  void* static_scratch[kMaxDevices] = {};

  // Every captured graph refers to the buffers it was built with.
  void drop_graphs() {
    if (batch_exec) (void)hipGraphExecDestroy(batch_exec);
    if (step_exec) (void)hipGraphExecDestroy(step_exec);
    if (sorted_exec) (void)hipGraphExecDestroy(sorted_exec);
    batch_exec = step_exec = sorted_exec = nullptr;
    batch_key.clear();
    step_key.clear();
    sorted_key.clear();
  }

  void init() {
    if (ready) return;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
      throw std::runtime_error("no HIP device available");
    HIPCHK(hipGetDevice(&device));
    hipDeviceProp_t prop{};
    HIPCHK(hipGetDeviceProperties(&prop, device));
    cu_count = prop.multiProcessorCount;
    max_lds = prop.sharedMemPerBlock;
    name = std::string(prop.name) + " (" + prop.gcnArchName + ", " + std::to_string(cu_count) +
           " CUs)";
    // The ring kernels address their LDS ring by OR-ing slot bits into
    // addresses that start at LDS 0; static LDS would move that start.
    if (!ring_kernels_start_at_lds_zero())
      throw std::runtime_error("LDS-ring kernels were built with static LDS; their ring "
                               "addressing needs the dynamic LDS to start at address 0");
    HIPCHK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    HIPCHK(hipStreamCreateWithFlags(&copy_stream, hipStreamNonBlocking));
    for (auto& e : ev) HIPCHK(hipEventCreate(&e));
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    pool = std::make_unique<CopyPool>(std::min(8u, std::max(1u, hw / 2)) - 1);
    rayleigh.ensure(kRayleighTableEntries * sizeof(float));
    launch_rayleigh_table(rayleigh.as<float>(), stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(stream));
    if (device < 0 || device >= kMaxDevices)
      throw std::runtime_error("device index " + std::to_string(device) + " above " +
                               std::to_string(kMaxDevices - 1));
    void* sp = nullptr;
    HIPCHK(hipGetSymbolAddress(&sp, HIP_SYMBOL(noalloc_store::scratch)));
    if (!sp) throw std::runtime_error("hipGetSymbolAddress returned NULL for the static scratch");
    static_scratch[device] = sp;
    ready = true;
  }

  // Events of a records run with up to `pieces` pieces: a start / end pair
  // per launch, then one copy-done event per piece.
  void ensure_piece_events(size_t pieces) {
    while (piece_ev.size() < 3 * pieces) {
      hipEvent_t e;
      HIPCHK(hipEventCreate(&e));
      piece_ev.push_back(e);
    }
  }

  float elapsed(int a, int b) {
    HIPCHK(hipEventSynchronize(ev[b]));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, ev[a], ev[b]));
    return ms;
  }
};

Context& ctx() {
  static Context c;
  return c;
}

// Captures what `enqueue` launches on c.stream into an executable graph. The
// capture is ended and the partial graph freed on every path, so a failure
// leaves the stream usable.
template <class F>
hipGraphExec_t capture_graph(Context& c, F&& enqueue) {
  HIPCHK(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
  hipGraph_t g = nullptr;
  try {
    enqueue();
  } catch (...) {
    (void)hipStreamEndCapture(c.stream, &g);
    if (g) (void)hipGraphDestroy(g);
    throw;
  }
  HIPCHK(hipStreamEndCapture(c.stream, &g));
  hipGraphExec_t exec = nullptr;
  const hipError_t e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  HIPCHK(e);
  return exec;
}

// ---- parameter conversion -------------------------------------------------

DevParams to_dev(const optphot_sim_params& p) {
  DevParams d{};
  d.hx = p.hx;
  d.hy = p.hy;
  d.hz = p.hz;
  d.inv_speed = p.n_scint / 299.792458f;  // transport.hpp:74
  d.abs_length = p.abs_length;
  d.scat_length = p.scat_length;
  d.n_scint = p.n_scint;
  d.n_det = p.n_det;
  d.n_out = p.n_out;
  const float eta_det = p.n_scint / p.n_det;  // fresnel.hpp:37-38
  const float eta_out = p.n_scint / p.n_out;
  d.eta2_det = eta_det * eta_det;
  d.eta2_out = eta_out * eta_out;
  d.det_matched = p.n_scint == p.n_det;
  d.out_matched = p.n_scint == p.n_out;
  d.reflectivity = p.reflectivity;
  d.src_x = p.src_x;
  d.src_y = p.src_y;
  d.src_z = p.src_z;
  d.tau = p.tau;
  d.surface = p.surface;
  d.source = p.source;
  d.max_steps = p.max_steps;
  auto in = [](float v, float lo, float hi) { return v >= lo && v <= hi; };
  const bool box = in(p.hx, 0x1p-60f, 0x1p40f) && in(p.hy, 0x1p-60f, 0x1p40f) &&
                   in(p.hz, 0x1p-60f, 0x1p40f);
  // an unmatched pair must also have eta^2 != 1 (two distinct floats can
  // have a ratio that rounds to 1), so that |1 - eta^2| >= 2^-24
  const bool idx = in(p.n_scint, 0x1p-20f, 0x1p20f) && in(p.n_det, 0x1p-20f, 0x1p20f) &&
                   in(p.n_out, 0x1p-20f, 0x1p20f) && (d.det_matched || d.eta2_det != 1.0f) &&
                   (d.out_matched || d.eta2_out != 1.0f);
  d.fast_div = (box ? 1u : 0u) | (idx ? 2u : 0u);
  d.rayleigh = ctx().rayleigh.as<const float>();
  return d;
}

PhiloxKey key_of(const optphot_sim_params& p) {
  return PhiloxKey{static_cast<uint32_t>(p.seed), static_cast<uint32_t>(p.seed >> 32)};
}

void validate_params(const optphot_sim_params* p) {
  if (!p) throw InvalidArgument("params is NULL");
  if (p->surface > 3u) throw InvalidArgument("surface must be 0..3 (polished|black|specular|lambertian)");
  if (p->source > 1u) throw InvalidArgument("source must be 0 (point) or 1 (volume)");
  if (!(p->hx > 0.0f && p->hy > 0.0f && p->hz > 0.0f))
    throw InvalidArgument("box half sizes must be positive");
}

void validate_block(int block_size, const char* who) {
  if (block_size <= 0 || block_size % 64 != 0 || block_size > 1024)
    throw InvalidArgument(std::string(who) + ": block_size must be a multiple of 64 <= 1024");
}

DevTally to_dev_tally(const optphot_tally_config& c, const optphot_sim_params& p) {
  if (c.time_bins == 0 || c.xy_bins == 0) throw InvalidArgument("tally bins must be >= 1");
  if (!(c.time_max > 0.0f)) throw InvalidArgument("time_max must be > 0");
  if (c.xy_bins > 4096u || c.time_bins > (1u << 24))
    throw InvalidArgument("tally binning too large");
  DevTally t{};
  t.time_bins = c.time_bins;
  t.xy_bins = c.xy_bins;
  t.time_scale = static_cast<float>(c.time_bins) / c.time_max;  // transport.hpp:171
  t.two_hx = 2.0f * p.hx;                                       // transport.hpp:178-179
  t.two_hy = 2.0f * p.hy;
  t.words = 5u + c.time_bins + 1u + c.xy_bins * c.xy_bins;
  return t;
}

void check_lds(Context& c, const DevTally& t) {
  const size_t bytes = static_cast<size_t>(t.words) * sizeof(uint32_t);
  if (bytes > c.max_lds) {
    throw std::runtime_error("tally histograms need " + std::to_string(bytes) +
                             " bytes of LDS; reduce xy_bins / time_bins");
  }
}

// ---- records helpers --------------------------------------------------------

DevRecords ensure_records(Context& c, uint64_t cap, double& alloc_ms) {
  const auto t0 = std::chrono::steady_clock::now();
  const size_t sz[8] = {1, 4, 4, 4, 4, 4, 4, 4};
  bool grew = false;
  for (int i = 0; i < 8; ++i) grew |= c.records[i].ensure(cap * sz[i]);
  if (grew) alloc_ms += ms_since(t0);
  return DevRecords{c.records[0].as<uint8_t>(),  c.records[1].as<uint32_t>(),
                    c.records[2].as<uint32_t>(), c.records[3].as<float>(),
                    c.records[4].as<float>(),    c.records[5].as<float>(),
                    c.records[6].as<float>(),    c.records[7].as<float>()};
}

void validate_records(const optphot_records* out) {
  if (!out || !out->fate || !out->n_boundary || !out->n_scatter || !out->t || !out->x || !out->y ||
      !out->z || !out->path)
    throw InvalidArgument("records: every output array must be non-NULL");
}

// Scratch words: [0] queue head, [1 .. 16] counters, [17 ..] block sums.
constexpr size_t kScratchQueue = 0, kScratchCounters = 1, kScratchSums = 1 + kCntSlots;

unsigned long long* ensure_scratch(Context& c, int grid, double& alloc_ms) {
  const auto t0 = std::chrono::steady_clock::now();
  if (c.scratch.ensure((kScratchSums + 2 * static_cast<size_t>(grid)) * 8)) {
    alloc_ms += ms_since(t0);
    c.drop_graphs();  // captured graphs hold the old scratch address
  }
  return c.scratch.as<unsigned long long>();
}

void read_counters(Context& c, optphot_step_counters* counters) {
  if (!counters) return;
  HIPCHK(hipMemcpyAsync(counters->v, c.scratch.as<unsigned long long>() + kScratchCounters,
                        sizeof(counters->v), hipMemcpyDeviceToHost, c.stream));
  HIPCHK(hipStreamSynchronize(c.stream));
}

// ---- the runs ---------------------------------------------------------------

// Whether a history launch (tally of `tally_words`, 0 for records) keeps the
// Fresnel cell bounds: whenever they fit beside the tally.
bool history_bounds(const Context& c, uint32_t tally_words, int surface) {
  return history_lds_bytes(tally_words, true, surface) <= c.max_lds;
}

int history_grid_for_tally(Context& c, int block, uint32_t tally_words, int surface, uint64_t n) {
  const bool bounds = history_bounds(c, tally_words, surface);
  const int per_cu =
      std::max(1, history_blocks_per_cu(true, block,
                                        history_lds_bytes(tally_words, bounds, surface)));
  const uint64_t needed = (n + block - 1) / block;
  return static_cast<int>(std::max<uint64_t>(
      1, std::min<uint64_t>(needed, static_cast<uint64_t>(per_cu) * c.cu_count)));
}

void run_tally_common(Context& c, const optphot_sim_params& p, const optphot_tally_config& tc,
                      uint64_t n, uint64_t* tally_out, uint64_t first_id, bool wavefront,
                      int block, int grid, const WavefrontShape& shape,
                      optphot_run_timing& tm, optphot_event_stats* stats,
                      optphot_step_counters* counters) {
  const auto t0 = std::chrono::steady_clock::now();
  const DevTally T = to_dev_tally(tc, p);
  check_lds(c, T);
  if (n / static_cast<uint64_t>(grid) >= 0xffffffffull)
    throw std::runtime_error("too many photons for one tally launch; split the run");
  unsigned long long* scratch = ensure_scratch(c, grid, tm.alloc_ms);
  const auto ta = std::chrono::steady_clock::now();
  const size_t tally_bytes = (static_cast<size_t>(T.words) + 2) * 8;
  bool grew = c.partials.ensure(static_cast<size_t>(grid) * T.words * sizeof(uint32_t));
  grew |= c.tally.ensure(tally_bytes);
  grew |= c.tally_host.ensure(tally_bytes);
  if (grew) tm.alloc_ms += ms_since(ta);

  LaunchOut out{};
  out.tally = T;
  out.partials = c.partials.as<uint32_t>();
  out.sums = scratch + kScratchSums;
  // Branch-event counting is a separate kernel variant, used only on request.
  out.counters = (counters || stats) ? scratch + kScratchCounters : nullptr;
  out.tally_on = true;
  out.records = false;
  out.fresnel_bounds = wavefront ? shape.fresnel_bounds
                                 : history_bounds(c, T.words, static_cast<int>(p.surface));

  launch_zero_u64_pair(scratch, kScratchSums + 2 * static_cast<size_t>(grid),
                       c.tally.as<unsigned long long>(), static_cast<size_t>(T.words) + 2,
                       c.stream);
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(c.ev[0], c.stream));
  if (wavefront)
    launch_wavefront(to_dev(p), key_of(p), first_id, n, shape, scratch + kScratchQueue, out,
                     c.stream);
  else
    launch_history(to_dev(p), key_of(p), first_id, n, block, grid, out, c.stream);
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(c.ev[1], c.stream));
  launch_tally_reduce(T, grid, out.partials, out.sums, c.tally.as<unsigned long long>(), c.stream);
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(c.ev[2], c.stream));
  // the copy is queued behind the reduction, so the host waits once; it
  // lands in pinned memory (a pageable destination costs a staged copy)
  HIPCHK(hipMemcpyAsync(c.tally_host.ptr, c.tally.ptr, tally_bytes, hipMemcpyDeviceToHost,
                        c.stream));
  HIPCHK(hipEventRecord(c.ev[3], c.stream));
  HIPCHK(hipStreamSynchronize(c.stream));
  std::memcpy(tally_out, c.tally_host.ptr, tally_bytes);
  const float k_ms = c.elapsed(0, 1);
  tm.kernel_ms = c.elapsed(0, 2);
  tm.d2h_ms = c.elapsed(2, 3);
  read_counters(c, counters);
  if (stats) {  // counted by the same launch (out.counters is set when stats is)
    optphot_step_counters cnt{};
    read_counters(c, &cnt);
    stats->launches = 1;
    stats->useful_lane_steps = cnt.v[kCntUseful];
    stats->issued_lane_steps = cnt.v[kCntIssued];
    stats->kernels_ms = k_ms;
  }
  tm.h2d_ms = 0.0;  // photons are generated on the device (gpu_backend.cu:151-153)
  tm.total_ms = ms_since(t0);
}

WavefrontShape wavefront_shape(Context& c, const optphot_wavefront_tuning* tu, bool records,
                               bool tally, size_t lds, uint64_t n, int surface = -1) {
  optphot_wavefront_tuning t{256, 4, 256, 0, 0, 0};
  if (tu) {
    if (tu->block_size) t.block_size = tu->block_size;
    if (tu->steps_per_refill) t.steps_per_refill = tu->steps_per_refill;
    if (tu->queue_chunk) t.queue_chunk = tu->queue_chunk;
    if (tu->blocks_per_cu) t.blocks_per_cu = tu->blocks_per_cu;
    if (tally && !records) {
      t.successor_threshold = tu->successor_threshold;
      t.stream_in_lds = tu->stream_in_lds;
      t.sorted_phases = tu->sorted_phases;
    }
  }
  if (t.successor_threshold > 64)
    throw InvalidArgument("wavefront: successor_threshold must be in [0, 64]");
  if (t.stream_in_lds &&
      wavefront_ring_lds_bytes(static_cast<uint32_t>(lds / 4), t.block_size, surface, false) >
          c.max_lds)
    throw InvalidArgument("wavefront: tally plus LDS streams exceed the LDS of a block; use a "
                          "smaller block_size");
  if (t.sorted_phases &&
      phase_lds_bytes(static_cast<uint32_t>(lds / 4), t.block_size) > c.max_lds)
    throw InvalidArgument("wavefront: tally, LDS streams and phase queues exceed the LDS of a "
                          "block; use a smaller block_size");
  validate_block(t.block_size, "wavefront");
  if (t.queue_chunk < 64 || t.queue_chunk % 64 != 0 || t.queue_chunk > (1u << 24))
    throw InvalidArgument("wavefront: queue_chunk must be a multiple of 64 in [64, 2^24]");
  if (t.steps_per_refill == 0)
    throw InvalidArgument("wavefront: steps_per_refill must be >= 1");
  // Fresnel cell bounds: kept by the ring and register-stream kernels when
  // they fit beside the tally (and ring)
  const uint32_t words = static_cast<uint32_t>(lds / 4);
  const bool bounds =
      t.sorted_phases == 0 && (t.stream_in_lds != 0 || t.successor_threshold == 0) &&
      (t.stream_in_lds != 0
           ? wavefront_ring_lds_bytes(words, t.block_size, surface, true) <= c.max_lds
           : history_lds_bytes(words, true, surface) <= c.max_lds);
  const int max_per_cu =
      std::max(1, wavefront_blocks_per_cu(records, tally, t.block_size, lds,
                                          t.successor_threshold > 0, t.stream_in_lds != 0,
                                          t.sorted_phases != 0, surface, bounds));
  const int per_cu = t.blocks_per_cu > 0 ? std::min(t.blocks_per_cu, max_per_cu) : max_per_cu;
  const uint64_t needed = (n + t.block_size - 1) / t.block_size;
  WavefrontShape s{};
  s.block = t.block_size;
  s.grid = static_cast<int>(std::max<uint64_t>(
      1, std::min<uint64_t>(needed, static_cast<uint64_t>(per_cu) * c.cu_count)));
  s.steps_per_refill = t.steps_per_refill;
  s.queue_chunk = t.queue_chunk;
  s.successor_threshold = t.successor_threshold;
  s.stream_in_lds = t.stream_in_lds;
  s.sorted_phases = t.sorted_phases;
  s.fresnel_bounds = bounds;
  return s;
}

// ---- records pipeline ------------------------------------------------------
//
// Photons first_id .. first_id+n-1 are transported in chunks of at most
// `chunk` photons (the reference's bound on device memory, backend.hpp:86-89).
// The wavefront kernel runs one launch per chunk (the reference's launch
// count, gpu_event.cu:174-244); the history kernel splits chunks above 4M
// photons into a few launches. Kernels write into the device record buffers, or
// straight into the caller's device arrays. For host output each piece is
// copied on c.copy_stream into pinned slot k % kStageSlots once the launch
// that produced it has finished, and the host copies that slot into the
// caller's arrays (several threads) while the GPU computes and copies later
// pieces. A new chunk's kernels wait for the previous chunk's copies, since
// they reuse the device buffers.

enum class RecKernel { History, Wavefront };

// Automatic records kernel: the wavefront kernel (kRecordsAutoSteps steps
// between refills) once n reaches kRecordsAutoPhotonsPerLane photons per lane
// of its persistent grid, else the history kernel.
constexpr uint64_t kRecordsAutoPhotonsPerLane = 8;
constexpr uint32_t kRecordsAutoSteps = 2;

// Launch shape of a records run: the history kernel (block) or the wavefront
// kernel (block, steps between refills, queue chunk, persistent blocks per CU;
// 0 = the occupancy limit).
struct RecordsShape {
  RecKernel kind = RecKernel::History;
  int block = 256;
  uint32_t steps = 0;
  uint32_t queue_chunk = 256;
  int blocks_per_cu = 0;
};

constexpr size_t kRecordBytes = 29;                                  // backend.hpp:49
constexpr uint64_t kStagePhotonsMax = (32ull << 20) / kRecordBytes;  // 32 MB per slot
constexpr uint64_t kPieceMin = 1ull << 16;
constexpr uint64_t kWholeLaunchMax = 4ull << 20;

struct RecordsRun {
  uint64_t launches = 0;
  double kernel_ms = 0.0;
};

DevRecords offset_records(const DevRecords& r, uint64_t k) {
  return DevRecords{r.fate + k, r.nb + k, r.ns + k, r.t + k, r.x + k, r.y + k, r.z + k,
                    r.path + k};
}

// A staged piece holds the 8 record fields of m photons, field a at byte
// offset stride x (bytes per photon of fields 0 .. a-1). Copies them into the
// caller's arrays at photon `off`, on several threads for large pieces.
void scatter_slot(Context& c, const char* src, uint64_t stride, uint64_t off, uint64_t m,
                  const optphot_records* host) {
  char* dst[8] = {reinterpret_cast<char*>(host->fate + off),
                  reinterpret_cast<char*>(host->n_boundary + off),
                  reinterpret_cast<char*>(host->n_scatter + off),
                  reinterpret_cast<char*>(host->t + off),
                  reinterpret_cast<char*>(host->x + off),
                  reinterpret_cast<char*>(host->y + off),
                  reinterpret_cast<char*>(host->z + off),
                  reinterpret_cast<char*>(host->path + off)};
  const size_t sz[8] = {1, 4, 4, 4, 4, 4, 4, 4};
  std::vector<std::function<void()>> tasks;
  const size_t split = m >= (uint64_t{1} << 18) ? c.pool->size() + 1 : 1;
  for (int a = 0; a < 8; ++a) {
    const size_t bytes = m * sz[a];
    const size_t part = (bytes + split - 1) / split;
    for (size_t b = 0; b < bytes; b += part) {
      const size_t len = std::min(part, bytes - b);
      char* d = dst[a] + b;
      const char* from = src + b;
      tasks.emplace_back([d, from, len] { std::memcpy(d, from, len); });
    }
    src += stride * sz[a];
  }
  c.pool->run(tasks);
}

// Photons per staged copy piece of a run of `cap` photons per device buffer:
// about an eighth of it, at least kPieceMin, at most one staging slot.
uint64_t stage_piece(uint64_t cap) {
  const uint64_t piece = std::max<uint64_t>(kPieceMin, (cap + 7) / 8);
  return std::min<uint64_t>({(piece + 63) / 64 * 64, kStagePhotonsMax, cap});
}

// Copies n finished records from the device buffers `buf` into the caller's
// host arrays: once `ready` (recorded on c.stream) has fired, piece k goes to
// pinned slot k % kStageSlots on c.copy_stream, and the host scatters each
// slot into the caller's arrays while later pieces are in flight. Returns
// when every piece is in the caller's arrays.
void stage_records_to_host(Context& c, const DevRecords& buf, uint64_t n,
                           const optphot_records* host, hipEvent_t ready, double& alloc_ms) {
  if (n == 0) return;
  const uint64_t piece = stage_piece(n);
  const auto ta = std::chrono::steady_clock::now();
  bool grew = false;
  for (auto& st : c.stage) grew |= st.ensure(piece * kRecordBytes);
  if (grew) alloc_ms += ms_since(ta);
  const size_t pieces = static_cast<size_t>((n + piece - 1) / piece);
  c.ensure_piece_events(pieces);
  hipEvent_t* cev = c.piece_ev.data();
  const size_t sz[8] = {1, 4, 4, 4, 4, 4, 4, 4};
  HIPCHK(hipStreamWaitEvent(c.copy_stream, ready, 0));
  size_t drained = 0;
  auto drain = [&](size_t k) {
    HIPCHK(hipEventSynchronize(cev[k]));
    const uint64_t off = k * piece;
    scatter_slot(c, static_cast<const char*>(c.stage[k % Context::kStageSlots].ptr), piece, off,
                 std::min(piece, n - off), host);
  };
  for (size_t k = 0; k < pieces; ++k) {
    // slot k % kStageSlots is free once the host has copied out its last piece
    while (drained + Context::kStageSlots <= k) drain(drained++);
    const uint64_t off = k * piece, m = std::min(piece, n - off);
    char* slot = static_cast<char*>(c.stage[k % Context::kStageSlots].ptr);
    const void* srcs[8] = {buf.fate + off, buf.nb + off, buf.ns + off, buf.t + off,
                           buf.x + off,    buf.y + off,  buf.z + off,  buf.path + off};
    for (int a = 0; a < 8; ++a) {
      HIPCHK(hipMemcpyAsync(slot, srcs[a], m * sz[a], hipMemcpyDeviceToHost, c.copy_stream));
      slot += piece * sz[a];
    }
    HIPCHK(hipEventRecord(cev[k], c.copy_stream));
  }
  while (drained < pieces) drain(drained++);
}

RecordsRun run_records(Context& c, const optphot_sim_params& p, uint64_t n,
                       const optphot_records* host, const DevRecords* dev,
                       const RecordsShape& shape, uint64_t first_id, uint64_t chunk,
                       unsigned long long* scratch, optphot_run_timing& tm) {
  const RecKernel kind = shape.kind;
  const int block = shape.block;
  RecordsRun run;
  if (n == 0) return run;
  const uint64_t limit = uint64_t{1} << 31;
  const uint64_t cap = dev ? std::min(n, limit)
                           : std::max<uint64_t>(1, std::min<uint64_t>({n, chunk, limit}));
  // copy piece; a history launch covers one piece, a wavefront launch a chunk
  const uint64_t piece = host ? stage_piece(cap) : cap;
  // Photons per launch: the wavefront kernel runs a chunk per launch; the
  // history kernel splits a large chunk into a few launches (a multiple of the
  // copy piece) so that copies start early, but each launch ends with its
  // longest history, so small chunks stay whole.
  uint64_t kpiece = cap;
  if (kind == RecKernel::History && cap > kWholeLaunchMax) {
    kpiece = std::max<uint64_t>(cap / 4, kWholeLaunchMax / 2);
    kpiece = std::min<uint64_t>((kpiece + piece - 1) / piece * piece, cap);
  }
  DevRecords buf{};
  if (!dev) buf = ensure_records(c, cap, tm.alloc_ms);
  if (host) {
    const auto ta = std::chrono::steady_clock::now();
    bool grew = false;
    for (auto& st : c.stage) grew |= st.ensure(piece * kRecordBytes);
    if (grew) tm.alloc_ms += ms_since(ta);
  }
  const uint64_t chunks = (n + cap - 1) / cap;
  const size_t max_pieces = static_cast<size_t>(chunks * ((cap + piece - 1) / piece));
  // events: [0, 2 * max_pieces) launch start / end, then one per copy piece
  c.ensure_piece_events(max_pieces);
  hipEvent_t* kev = c.piece_ev.data();
  hipEvent_t* cev = c.piece_ev.data() + 2 * max_pieces;

  const DevParams P = to_dev(p);
  const PhiloxKey key = key_of(p);
  LaunchOut lo{};
  lo.records = true;
  lo.counters = scratch ? scratch + kScratchCounters : nullptr;
  const optphot_wavefront_tuning tu{block, shape.steps, shape.queue_chunk, shape.blocks_per_cu};

  const size_t sz[8] = {1, 4, 4, 4, 4, 4, 4, 4};
  struct Piece {
    uint64_t off, m, dev_off;
    size_t launch;
  };
  std::vector<Piece> pieces;
  pieces.reserve(max_pieces);
  size_t drained = 0;
  double d2h_ms = 0.0;

  auto launch = [&](uint64_t off, uint64_t m, uint64_t dev_off) {
    const size_t k = run.launches++;
    lo.rec = dev ? offset_records(*dev, off) : offset_records(buf, dev_off);
    HIPCHK(hipEventRecord(kev[2 * k], c.stream));
    if (kind == RecKernel::History) {
      const int grid = static_cast<int>((m + block - 1) / block);
      lo.fresnel_bounds = history_bounds(c, 0, static_cast<int>(P.surface));
      launch_history(P, key, first_id + off, m, block, grid, lo, c.stream);
    } else {
      const WavefrontShape sm = wavefront_shape(c, &tu, true, false, 0, m);
      lo.fresnel_bounds = sm.fresnel_bounds;
      launch_zero_u64(scratch + kScratchQueue, 1, c.stream);
      launch_wavefront(P, key, first_id + off, m, sm, scratch + kScratchQueue, lo, c.stream);
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(kev[2 * k + 1], c.stream));
    return k;
  };

  // Host side of piece k: wait for its copy, then scatter the slot into the
  // caller's arrays.
  auto drain = [&](size_t k) {
    const auto t0 = std::chrono::steady_clock::now();
    const Piece& pc = pieces[k];
    HIPCHK(hipEventSynchronize(cev[k]));
    scatter_slot(c, static_cast<const char*>(c.stage[k % Context::kStageSlots].ptr), piece,
                 pc.off, pc.m, host);
    d2h_ms += ms_since(t0);
  };

  for (uint64_t ch = 0; ch < chunks; ++ch) {
    const uint64_t c0 = ch * cap;
    const uint64_t cm = std::min(cap, n - c0);
    const size_t first_piece = pieces.size();
    if (host && first_piece > 0)  // device buffers reused: previous chunk copied first
      HIPCHK(hipStreamWaitEvent(c.stream, cev[first_piece - 1], 0));
    size_t k = 0;
    for (uint64_t o = 0; o < cm; o += piece) {
      if (o % kpiece == 0) k = launch(c0 + o, std::min(kpiece, cm - o), o);
      pieces.push_back(Piece{c0 + o, std::min(piece, cm - o), o, k});
    }
    if (!host) continue;
    for (size_t k = first_piece; k < pieces.size(); ++k) {
      const Piece& pc = pieces[k];
      // slot k % kStageSlots is free once the host has copied out its last piece
      while (drained + Context::kStageSlots <= k) drain(drained++);
      HIPCHK(hipStreamWaitEvent(c.copy_stream, kev[2 * pc.launch + 1], 0));
      char* slot = static_cast<char*>(c.stage[k % Context::kStageSlots].ptr);
      const void* srcs[8] = {buf.fate + pc.dev_off, buf.nb + pc.dev_off, buf.ns + pc.dev_off,
                             buf.t + pc.dev_off,    buf.x + pc.dev_off,  buf.y + pc.dev_off,
                             buf.z + pc.dev_off,    buf.path + pc.dev_off};
      for (int a = 0; a < 8; ++a) {
        HIPCHK(hipMemcpyAsync(slot, srcs[a], pc.m * sz[a], hipMemcpyDeviceToHost,
                              c.copy_stream));
        slot += piece * sz[a];
      }
      HIPCHK(hipEventRecord(cev[k], c.copy_stream));
    }
  }
  while (host && drained < pieces.size()) drain(drained++);
  HIPCHK(hipStreamSynchronize(c.stream));
  for (size_t k = 0; k < run.launches; ++k) {
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, kev[2 * k], kev[2 * k + 1]));
    run.kernel_ms += ms;
  }
  tm.kernel_ms += run.kernel_ms;
  tm.d2h_ms += d2h_ms;
  return run;
}

void validate_batch(const optphot_sim_params* params, const uint64_t* photons, uint32_t n_events,
                    uint32_t time_bins, float time_max, uint32_t* out, int method) {
  if (!params || !photons || !out) throw InvalidArgument("batch: NULL argument");
  if (n_events == 0) throw InvalidArgument("batch: n_events must be >= 1");
  if (time_bins == 0 || time_bins > (1u << 20) || !(time_max > 0.0f))
    throw InvalidArgument("batch: time_bins must be in [1, 2^20] and time_max > 0");
  if (method != OPTPHOT_BATCH_WAVEFRONT && method != OPTPHOT_BATCH_PLAIN)
    throw InvalidArgument("batch: unknown method");
  for (uint32_t e = 0; e < n_events; ++e) {
    validate_params(params + e);
    if (params[e].seed != params[0].seed)
      throw InvalidArgument("batch: every event must use the run seed params[0].seed");
    if (photons[e] >= 0xffffffffull)
      throw InvalidArgument("batch: an event's photon count must fit the u32 event tally");
  }
}

// The adjoint's reflector score is hits / R per detected photon, so every
// set with reflecting faces needs R in (0, 1] (R = 0 would make the score of
// a photon that met no reflector 0 / 0).
void validate_adjoint_sets(const optphot_sim_params* params, uint32_t n_sets) {
  for (uint32_t e = 0; e < n_sets; ++e) {
    const optphot_sim_params& q = params[e];
    const bool reflector = q.surface == OPTPHOT_SURFACE_SPECULAR ||
                           q.surface == OPTPHOT_SURFACE_LAMBERTIAN;
    if (reflector && !(q.reflectivity > 0.0f && q.reflectivity <= 1.0f))
      throw InvalidArgument("adjoint: set " + std::to_string(e) +
                            " has reflecting faces and reflectivity outside (0, 1]");
  }
}

template <class F>
int guarded(F&& f) {
  try {
    g_last_error.clear();
    f();
    return OPTPHOT_OK;
  } catch (const InvalidArgument& e) {
    g_last_error = e.what();
    return OPTPHOT_ERR_INVALID_ARGUMENT;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return std::strstr(e.what(), "no HIP device") ? OPTPHOT_ERR_NO_DEVICE : OPTPHOT_ERR_RUNTIME;
  } catch (...) {
    g_last_error = "unknown error";
    return OPTPHOT_ERR_RUNTIME;
  }
}

}  // namespace

// =============================================================================
// C-ABI
// =============================================================================

extern "C" {

__attribute__((visibility("default"))) int optphot_gpu_available(void) {
  int count = 0;
  return hipGetDeviceCount(&count) == hipSuccess && count > 0 ? 1 : 0;
}

__attribute__((visibility("default"))) const char* optphot_last_error(void) {
  return g_last_error.c_str();
}

__attribute__((visibility("default"))) int optphot_gpu_device_name(char* buf, size_t len) {
  return guarded([&] {
    if (!buf || len == 0) throw InvalidArgument("buffer is NULL");
    Context& c = ctx();
    std::lock_guard<std::mutex> lock(c.mu);
    c.init();
    std::snprintf(buf, len, "%s", c.name.c_str());
  });
}

__attribute__((visibility("default"))) int optphot_rayleigh_table(float* out, size_t n) {
  return guarded([&] {
    if (!out || n != kRayleighTableEntries)
      throw InvalidArgument("rayleigh table: out must hold 2^24 floats");
    Context& c = ctx();
    std::lock_guard<std::mutex> lock(c.mu);
    c.init();
    HIPCHK(hipMemcpy(out, c.rayleigh.ptr, n * sizeof(float), hipMemcpyDeviceToHost));
  });
}

__attribute__((visibility("default"))) int optphot_device_info(char* buf, size_t len) {
  return guarded([&] {
    if (!buf || len == 0) throw InvalidArgument("buffer is NULL");
    Context& c = ctx();
    std::lock_guard<std::mutex> lock(c.mu);
    c.init();
    std::string s = "name=" + c.name + "\ncu_count=" + std::to_string(c.cu_count) +
                    "\nmax_lds_per_block=" + std::to_string(c.max_lds) + "\n";
    for (int b : {64, 128, 256, 512, 1024}) {
      s += "wavefront_tally_blocks_per_cu_" + std::to_string(b) + "=" +
           std::to_string(wavefront_blocks_per_cu(false, true, b, 17432)) + "\n";
      s += "history_records_blocks_per_cu_" + std::to_string(b) + "=" +
           std::to_string(history_blocks_per_cu(false, b, 0)) + "\n";
    }
    std::snprintf(buf, len, "%s", s.c_str());
  });
}

__attribute__((visibility("default"))) int optphot_run_gpu_records(
    const optphot_sim_params* p, uint64_t n, const optphot_records* out, int block_size,
    uint64_t first_id, uint64_t chunk, optphot_run_timing* timing) {
  return guarded([&] {
    validate_params(p);
    validate_block(block_size, "history-based backend");
    if (n > 0) validate_records(out);
    Context& c = ctx();
    std::lock_guard<std::mutex> lock(c.mu);
    c.init();
    optphot_run_timing tm{};
    const auto t0 = std::chrono::steady_clock::now();
    RecordsShape rs;
    rs.block = block_size;
    run_records(c, *p, n, out, nullptr, rs, first_id, chunk, nullptr, tm);
    tm.total_ms = ms_since(t0);
    if (timing) *timing = tm;
  });
}

__attribute__((visibility("default"))) int optphot_run_gpu_records_device(
    const optphot_sim_params* p, uint64_t n, const optphot_records* dev_out,
    uint32_t steps_per_launch, int block_size, uint64_t first_id, optphot_event_stats* stats,
    optphot_run_timing* timing) {
  return guarded([&] {
    validate_params(p);
    validate_block(block_size, "device records");
    if (n > 0) validate_records(dev_out);
    Context& c = ctx();
    std::lock_guard<std::mutex> lock(c.mu);
    c.init();
    if (n > 0) {
      hipPointerAttribute_t attr{};
      if (hipPointerGetAttributes(&attr, dev_out->t) != hipSuccess ||
          attr.type != hipMemoryTypeDevice)
        throw InvalidArgument("device records: output arrays must be device memory");
    }
    optphot_run_timing tm{};
    const auto t0 = std::chrono::steady_clock::now();
    const DevRecords d{dev_out->fate, dev_out->n_boundary, dev_out->n_scatter, dev_out->t,
                       dev_out->x,    dev_out->y,          dev_out->z,         dev_out->path};
    unsigned long long* scratch = nullptr;
    if (steps_per_launch > 0) {
      optphot_wavefront_tuning tu{block_size, steps_per_launch, 256, 0};
      const WavefrontShape s = wavefront_shape(c, &tu, true, false, 0, std::max<uint64_t>(n, 1));
      scratch = ensure_scratch(c, s.grid, tm.alloc_ms);
      launch_zero_u64(scratch, kScratchSums, c.stream);
      HIPCHK(hipGetLastError());
    }
    RecordsShape rs;
    rs.kind = steps_per_launch > 0 ? RecKernel::Wavefront : RecKernel::History;
    rs.block = block_size;
    rs.steps = steps_per_launch;
    const RecordsRun r = run_records(c, *p, n, nullptr, &d, rs, first_id, 0, scratch, tm);
    if (stats) {
      optphot_event_stats st{};
      st.launches = r.launches;
      st.kernels_ms = r.kernel_ms;
      if (scratch) {
        optphot_step_counters cnt{};
        read_counters(c, &cnt);
        st.useful_lane_steps = cnt.v[kCntUseful];
        st.issued_lane_steps = cnt.v[kCntIssued];
      }
      *stats = st;
    }
    tm.total_ms = ms_since(t0);
    if (timing) *timing = tm;
  });
}

__attribute__((visibility("default"))) int optphot_run_gpu_records_tuned(
    const optphot_sim_params* p, uint64_t n, const optphot_records* out, int out_on_device,
    const optphot_records_tuning* tuning, uint64_t first_id, optphot_records_tuning* chosen,
    optphot_event_stats* stats, optphot_run_timing* timing) {
  return guarded([&] {
    validate_params(p);
    optphot_records_tuning t{OPTPHOT_RECORDS_AUTO, 0, 0, 0, 0};
    if (tuning) t = *tuning;
    if (t.kernel < OPTPHOT_RECORDS_AUTO || t.kernel > OPTPHOT_RECORDS_WAVEFRONT)
      throw InvalidArgument("records tuning: kernel must be OPTPHOT_RECORDS_AUTO, _HISTORY or "
                            "_WAVEFRONT");
    if (t.block_size == 0) t.block_size = 256;
    if (t.steps_per_refill == 0) t.steps_per_refill = kRecordsAutoSteps;
    if (t.queue_chunk == 0) t.queue_chunk = 256;
    if (t.blocks_per_cu < 0) throw InvalidArgument("records tuning: blocks_per_cu must be >= 0");
    validate_block(t.block_size, "records");
    if (n > 0) validate_records(out);
    Context& c = ctx();
    std::lock_guard<std::mutex> lock(c.mu);
    c.init();
    if (n > 0 && out_on_device) {
      hipPointerAttribute_t attr{};
      if (hipPointerGetAttributes(&attr, out->t) != hipSuccess ||
          attr.type != hipMemoryTypeDevice)
        throw InvalidArgument("records: out_on_device set but the arrays are not device memory");
    }
    optphot_run_timing tm{};
    const auto t0 = std::chrono::steady_clock::now();
    const optphot_wavefront_tuning wt{t.block_size, t.steps_per_refill, t.queue_chunk,
                                      t.blocks_per_cu};
    const WavefrontShape ws = wavefront_shape(c, &wt, true, false, 0, std::max<uint64_t>(n, 1));
    if (t.kernel == OPTPHOT_RECORDS_AUTO) {
      // Refilling finished lanes pays once each lane of the persistent grid
      // has several photons to go through; below that every photon starts at
      // once on the history kernel, and both end with the longest history.
      const uint64_t lanes = static_cast<uint64_t>(ws.grid) * ws.block;
      t.kernel = n >= kRecordsAutoPhotonsPerLane * lanes ? OPTPHOT_RECORDS_WAVEFRONT
                                                        : OPTPHOT_RECORDS_HISTORY;
    }
    RecordsShape rs;
    rs.kind = t.kernel == OPTPHOT_RECORDS_WAVEFRONT ? RecKernel::Wavefront : RecKernel::History;
    rs.block = t.block_size;
    rs.steps = t.steps_per_refill;
    rs.queue_chunk = t.queue_chunk;
    rs.blocks_per_cu = t.blocks_per_cu;
    unsigned long long* scratch = nullptr;
    if (rs.kind == RecKernel::Wavefront) {
      scratch = ensure_scratch(c, ws.grid, tm.alloc_ms);
      launch_zero_u64(scratch, kScratchSums, c.stream);  // queue + counters
      HIPCHK(hipGetLastError());
    }
    DevRecords d{};
    if (out_on_device)
      d = DevRecords{out->fate, out->n_boundary, out->n_scatter, out->t,
                     out->x,    out->y,          out->z,         out->path};
    const RecordsRun r = run_records(c, *p, n, out_on_device ? nullptr : out,
                                     out_on_device ? &d : nullptr, rs, first_id,
                                     uint64_t{1} << 24, scratch, tm);
    if (stats) {
      optphot_event_stats st{};
      st.launches = r.launches;
      st.kernels_ms = r.kernel_ms;
      if (scratch) {
        optphot_step_counters cnt{};
        read_counters(c, &cnt);
        st.useful_lane_steps = cnt.v[kCntUseful];
        st.issued_lane_steps = cnt.v[kCntIssued];
      }
      *stats = st;
    }
    if (chosen) *chosen = t;
    tm.total_ms = ms_since(t0);
    if (timing) *timing = tm;
  });
}

__attribute__((visibility("default"))) int optphot_run_gpu_tally(
    const optphot_sim_params* p, const optphot_tally_config* tc, uint64_t n, uint64_t* tally,
    int block_size, uint64_t first_id, optphot_run_timing* timing) {
  return optphot_run_gpu_tally_counters(p, tc, n, tally, block_size, first_id, nullptr, timing);
}

__attribute__((visibility("default"))) int optphot_run_gpu_tally_counters(
    const optphot_sim_params* p, const optphot_tally_config* tc, uint64_t n, uint64_t* tally,
    int block_size, uint64_t first_id, optphot_step_counters* counters,
    optphot_run_timing* timing) {
  return guarded([&] {
    validate_params(p);
    validate_block(block_size, "history-based backend");
    if (!tc || !tally) throw InvalidArgument("tally config / output is NULL");
    Context& c = ctx();
    std::lock_guard<std::mutex> lock(c.mu);
    c.init();
    const DevTally T = to_dev_tally(*tc, *p);
    check_lds(c, T);
    const int grid =
        history_grid_for_tally(c, block_size, T.words, static_cast<int>(p->surface), n);
    optphot_run_timing tm{};
    run_tally_common(c, *p, *tc, n, tally, first_id, false, block_size, grid, WavefrontShape{},
                     tm, nullptr, counters);
    if (timing) *timing = tm;
  });
}

__attribute__((visibility("default"))) int optphot_run_gpu_tally_wavefront(
    const optphot_sim_params* p, const optphot_tally_config* tc, uint64_t n, uint64_t* tally,
    uint64_t first_id, const optphot_wavefront_tuning* tuning, optphot_event_stats* stats,
    optphot_step_counters* counters, optphot_run_timing* timing) {
  return guarded([&] {
    validate_params(p);
    if (!tc || !tally) throw InvalidArgument("tally config / output is NULL");
    Context& c = ctx();
    std::lock_guard<std::mutex> lock(c.mu);
    c.init();
    const DevTally T = to_dev_tally(*tc, *p);
    const size_t lds = static_cast<size_t>(T.words) * 4;
    check_lds(c, T);
    const WavefrontShape s = wavefront_shape(c, tuning, false, true, lds, n,
                                             static_cast<int>(p->surface));
    optphot_run_timing tm{};
    run_tally_common(c, *p, *tc, n, tally, first_id, true, s.block, s.grid, s, tm, stats,
                     counters);
    if (timing) *timing = tm;
  });
}

__attribute__((visibility("default"))) int optphot_run_gpu_tally_wavefront_noalloc(
    const optphot_sim_params* p, const optphot_tally_config* tc, uint64_t n, uint64_t* dev_tally,
    uint64_t first_id, const optphot_wavefront_tuning* tuning, optphot_run_timing* timing) {
  return guarded([&] {
    validate_params(p);
    if (!tc || !dev_tally) throw InvalidArgument("noalloc tally: config / output is NULL");
    Context& c = ctx();
    std::lock_guard<std::mutex> lock(c.mu);
    if (!c.ready)
      throw std::runtime_error(
          "noalloc tally: the library is not set up on this device; make any other call first "
          "(for example optphot_gpu_device_name), which prepares what this entry point uses");
    hipPointerAttribute_t attr{};
    if (hipPointerGetAttributes(&attr, dev_tally) != hipSuccess ||
        attr.type != hipMemoryTypeDevice)
      throw InvalidArgument("noalloc tally: the output must be device memory");
    const DevTally T = to_dev_tally(*tc, *p);
    const size_t lds = static_cast<size_t>(T.words) * 4;
    check_lds(c, T);
    WavefrontShape s = wavefront_shape(c, tuning, false, true, lds, n,
                                       static_cast<int>(p->surface));
    // Scratch layout (bytes): queue head and counters, 2 u64 sums per block,
    // then at the next 256-byte boundary the u32 partial rows of every block.
    constexpr size_t kCap = noalloc_store::kScratchBytes;
    auto rows_at = [](uint64_t g) { return (kScratchSums * 8 + 16 * g + 255) & ~size_t{255}; };
    auto bytes_for = [&](uint64_t g) { return rows_at(g) + 4ull * T.words * g; };
    const uint64_t fit = (kCap - kScratchSums * 8 - 255) / (16 + 4ull * T.words);
    const uint64_t want = static_cast<uint64_t>(s.grid);
    // fewer persistent blocks than the occupancy allows only costs speed (the
    // tally does not depend on the grid), but not fewer than one per CU
    if (want > fit) {
      if (fit < std::min<uint64_t>(want, static_cast<uint64_t>(c.cu_count)))
        throw InvalidArgument(
            "noalloc tally: a tally of " + std::to_string(T.words) + " words needs " +
            std::to_string(bytes_for(std::min<uint64_t>(want, c.cu_count))) +
            " bytes of scratch at one block per compute unit, above the static maximum of " +
            std::to_string(kCap) + " bytes; use fewer xy_bins / time_bins or "
            "optphot_run_gpu_tally_wavefront");
      s.grid = static_cast<int>(fit);
    }
    if (bytes_for(static_cast<uint64_t>(s.grid)) > kCap)
      throw std::runtime_error("noalloc tally: scratch layout exceeds its maximum");
    if (n / static_cast<uint64_t>(s.grid) >= 0xffffffffull)
      throw InvalidArgument("noalloc tally: too many photons for one tally launch; split the run");
    char* base = static_cast<char*>(c.static_scratch[c.device]);
    if (!base) throw std::runtime_error("noalloc tally: static scratch address not set up");
    auto* words64 = reinterpret_cast<unsigned long long*>(base);
    auto* out_tally = reinterpret_cast<unsigned long long*>(dev_tally);

    optphot_run_timing tm{};
    const auto t0 = std::chrono::steady_clock::now();
    LaunchOut out{};
    out.tally = T;
    out.partials = reinterpret_cast<uint32_t*>(base + rows_at(static_cast<uint64_t>(s.grid)));
    out.sums = words64 + kScratchSums;
    out.counters = nullptr;
    out.tally_on = true;
    out.records = false;
    out.fresnel_bounds = s.fresnel_bounds;
    // The array keeps the previous call's contents: clear the queue head and
    // the block sums (each block writes its whole partial row).
    launch_zero_u64(words64, kScratchSums + 2 * static_cast<size_t>(s.grid), c.stream);
    launch_zero_u64(out_tally, static_cast<size_t>(T.words) + 2, c.stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(c.ev[0], c.stream));
    launch_wavefront(to_dev(*p), key_of(*p), first_id, n, s, words64 + kScratchQueue, out,
                     c.stream);
    HIPCHK(hipGetLastError());
    launch_tally_reduce(T, s.grid, out.partials, out.sums, out_tally, c.stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(c.ev[1], c.stream));
    tm.kernel_ms = c.elapsed(0, 1);
    tm.total_ms = ms_since(t0);
    if (timing) *timing = tm;
  });
}

__attribute__((visibility("default"))) int optphot_run_gpu_records_event(
    const optphot_sim_params* p, uint64_t n, const optphot_records* out,
    uint32_t steps_per_launch, int block_size, uint64_t first_id, uint64_t chunk,
    optphot_event_stats* stats, optphot_run_timing* timing) {
  return guarded([&] {
    validate_params(p);
    validate_block(block_size, "event-based backend");
    if (steps_per_launch == 0)
      throw InvalidArgument("event-based backend: steps_per_launch must be >= 1");
    if (n > 0) validate_records(out);
    Context& c = ctx();
    std::lock_guard<std::mutex> lock(c.mu);
    c.init();
    optphot_run_timing tm{};
    const auto t0 = std::chrono::steady_clock::now();
    optphot_wavefront_tuning tu{block_size, steps_per_launch, 256, 0};
    const WavefrontShape s = wavefront_shape(c, &tu, true, false, 0, std::max<uint64_t>(n, 1));
    unsigned long long* scratch = ensure_scratch(c, s.grid, tm.alloc_ms);
    launch_zero_u64(scratch, kScratchSums, c.stream);  // queue + counters
    HIPCHK(hipGetLastError());
    RecordsShape rs;
    rs.kind = RecKernel::Wavefront;
    rs.block = block_size;
    rs.steps = steps_per_launch;
    const RecordsRun r = run_records(c, *p, n, out, nullptr, rs, first_id, chunk, scratch, tm);
    optphot_event_stats st{};
    st.launches = r.launches;
    st.kernels_ms = r.kernel_ms;
    optphot_step_counters cnt{};
    read_counters(c, &cnt);
    st.useful_lane_steps = cnt.v[kCntUseful];
    st.issued_lane_steps = cnt.v[kCntIssued];
    tm.total_ms = ms_since(t0);
    if (stats) *stats = st;
    if (timing) *timing = tm;
  });
}

__attribute__((visibility("default"))) int optphot_run_gpu_tally_sorted(
    const optphot_sim_params* p, const optphot_tally_config* tc, uint64_t n, uint64_t* tally,
    uint64_t first_id, const optphot_sorted_tuning* tuning, optphot_sorted_stats* stats,
    optphot_run_timing* timing) {
  return guarded([&] {
    validate_params(p);
    if (!tc || !tally) throw InvalidArgument("tally config / output is NULL");
    optphot_sorted_tuning tu{2048, 6, 8, 0};
    if (tuning) {
      if (tuning->slots_per_block) tu.slots_per_block = tuning->slots_per_block;
      if (tuning->blocks_per_cu) tu.blocks_per_cu = tuning->blocks_per_cu;
      if (tuning->iterations_per_poll) tu.iterations_per_poll = tuning->iterations_per_poll;
      tu.tail_threshold = tuning->tail_threshold;
    }
    if (tu.slots_per_block % kSortedBlock != 0 || tu.slots_per_block > 65536)
      throw InvalidArgument("sorted: slots_per_block must be a multiple of 256 in [256, 65536]");
    if (tu.blocks_per_cu > 64)
      throw InvalidArgument("sorted: blocks_per_cu must be in [1, 64]");
    if (tu.iterations_per_poll % 2 != 0 || tu.iterations_per_poll > 1024)
      throw InvalidArgument("sorted: iterations_per_poll must be even, in [2, 1024]");
    if (n >= 0xffffffffull) throw InvalidArgument("sorted: n must be below 2^32 - 1; split the run");
    Context& c = ctx();
    std::lock_guard<std::mutex> lock(c.mu);
    c.init();
    const auto t0 = std::chrono::steady_clock::now();
    optphot_run_timing tm{};
    const DevTally T = to_dev_tally(*tc, *p);
    check_lds(c, T);
    const uint32_t grid = tu.blocks_per_cu * static_cast<uint32_t>(c.cu_count);
    const uint32_t seg = tu.slots_per_block;
    const uint64_t W = static_cast<uint64_t>(grid) * seg;
    if (W > (1ull << 29)) throw InvalidArgument("sorted: blocks x slots_per_block must be <= 2^29");
    const uint64_t threshold = tu.tail_threshold ? tu.tail_threshold : W / 8;

    // One buffer: 12 state arrays, 3 rotate + 2 fresnel queue arrays (all
    // [W] x 4 B), then the per-block words and rows.
    const size_t q_words = 17 * W;
    const size_t off_cnt = q_words * 4;                               // 3 x [grid] u32
    const size_t off_next = (off_cnt + 16ull * grid + 255) & ~255ull;  // [grid] u64
    const size_t off_sums = off_next + 8ull * grid;                    // [2 grid] u64
    const size_t off_stats = off_sums + 16ull * grid;                  // [grid][8] u64
    const size_t off_status = off_stats + 8ull * kSortedStats * grid;  // [2] u64
    const size_t off_rows = (off_status + 16 + 255) & ~255ull;          // [grid][words] u32
    const size_t bytes = off_rows + 4ull * grid * T.words;
    const auto ta = std::chrono::steady_clock::now();
    bool grew = c.sorted.ensure(bytes);
    grew |= c.tally.ensure((static_cast<size_t>(T.words) + 2) * 8);
    grew |= c.sorted_status.ensure(16);
    if (grew) {
      tm.alloc_ms += ms_since(ta);
      if (c.sorted_exec) (void)hipGraphExecDestroy(c.sorted_exec);
      c.sorted_exec = nullptr;
      c.sorted_key.clear();
    }
    char* base = c.sorted.as<char>();
    uint32_t* q = c.sorted.as<uint32_t>();
    SortedArgs A{};
    float* f = reinterpret_cast<float*>(q);
    A.px = f;
    A.py = f + W;
    A.pz = f + 2 * W;
    A.dx = f + 3 * W;
    A.dy = f + 4 * W;
    A.dz = f + 5 * W;
    A.t = f + 6 * W;
    A.path = f + 7 * W;
    A.nb = q + 8 * W;
    A.ns = q + 9 * W;
    A.idx = q + 10 * W;
    A.draws = q + 11 * W;
    A.rot_key = q + 12 * W;
    A.rot_wc = q + 13 * W;
    A.rot_wp = q + 14 * W;
    A.fres_key = q + 15 * W;
    A.fres_w = q + 16 * W;
    uint32_t* cnt = reinterpret_cast<uint32_t*>(base + off_cnt);
    A.cnt_rot = cnt;
    A.cnt_fres = cnt + grid;
    A.occ = cnt + 2 * grid;
    A.next_local = reinterpret_cast<unsigned long long*>(base + off_next);
    A.sums = reinterpret_cast<unsigned long long*>(base + off_sums);
    A.stats = reinterpret_cast<unsigned long long*>(base + off_stats);
    A.status = reinterpret_cast<unsigned long long*>(base + off_status);
    A.partials = reinterpret_cast<uint32_t*>(base + off_rows);
    A.tally = T;
    A.first_id = first_id;
    A.n = n;
    A.seg = seg;
    A.grid = grid;
    const DevParams P = to_dev(*p);
    const PhiloxKey key = key_of(*p);

    // The captured group: iterations_per_poll iterations, then the status.
    std::string gkey(reinterpret_cast<const char*>(p), sizeof(*p));
    gkey += "/" + std::to_string(n) + "/" + std::to_string(first_id) + "/" +
            std::to_string(T.words) + "/" + std::to_string(T.time_scale) + "/" +
            std::to_string(tc->xy_bins) + "/" + std::to_string(grid) + "/" + std::to_string(seg) +
            "/" + std::to_string(tu.iterations_per_poll);
    if (!c.sorted_exec || c.sorted_key != gkey) {
      if (c.sorted_exec) HIPCHK(hipGraphExecDestroy(c.sorted_exec));
      c.sorted_exec = nullptr;
      c.sorted_exec = capture_graph(c, [&] {
        for (uint32_t i = 0; i < tu.iterations_per_poll; ++i)
          launch_sorted_iteration(P, key, A, c.stream);
        launch_sorted_status(A, c.stream);
      });
      c.sorted_key = gkey;
    }

    // Every photon takes at most max_steps steps and each occupied slot takes
    // one step per iteration, so the loop ends within this many iterations.
    const uint64_t per_block = (n + grid - 1) / grid;
    const uint64_t bound =
        ((per_block + seg - 1) / seg + 1) * (static_cast<uint64_t>(p->max_steps) + 2) +
        2ull * tu.iterations_per_poll;
    auto* st = static_cast<volatile unsigned long long*>(c.sorted_status.ptr);
    HIPCHK(hipEventRecord(c.ev[0], c.stream));
    launch_sorted_init(A, c.stream);
    HIPCHK(hipGetLastError());
    uint64_t iterations = 0;
    for (;;) {
      HIPCHK(hipGraphLaunch(c.sorted_exec, c.stream));
      iterations += tu.iterations_per_poll;
      HIPCHK(hipMemcpyAsync(c.sorted_status.ptr, A.status, 16, hipMemcpyDeviceToHost, c.stream));
      HIPCHK(hipStreamSynchronize(c.stream));
      const uint64_t occupied = st[0], left = st[1];
      if (left == 0 && (occupied <= threshold || occupied == 0)) break;
      if (iterations > bound)
        throw std::runtime_error("sorted: photons still in flight after " +
                                 std::to_string(iterations) + " iterations");
    }
    HIPCHK(hipEventRecord(c.ev[1], c.stream));
    launch_sorted_tail(P, key, A, c.stream);
    HIPCHK(hipGetLastError());
    launch_zero_u64(c.tally.as<unsigned long long>(), static_cast<size_t>(T.words) + 2, c.stream);
    launch_tally_reduce(T, static_cast<int>(grid), A.partials, A.sums,
                        c.tally.as<unsigned long long>(), c.stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(c.ev[2], c.stream));
    const float loop_ms = c.elapsed(0, 1);
    tm.kernel_ms = c.elapsed(0, 2);

    const auto td = std::chrono::steady_clock::now();
    HIPCHK(hipMemcpyAsync(tally, c.tally.ptr, (static_cast<size_t>(T.words) + 2) * 8,
                          hipMemcpyDeviceToHost, c.stream));
    HIPCHK(hipStreamSynchronize(c.stream));
    tm.d2h_ms = ms_since(td);
    if (stats) {
      std::vector<unsigned long long> rows(static_cast<size_t>(grid) * kSortedStats);
      HIPCHK(hipMemcpy(rows.data(), A.stats, rows.size() * 8, hipMemcpyDeviceToHost));
      unsigned long long sum[kSortedStats] = {};
      for (size_t i = 0; i < rows.size(); ++i) sum[i % kSortedStats] += rows[i];
      *stats = optphot_sorted_stats{};
      stats->slots = W;
      stats->iterations = iterations;
      stats->classified = sum[kSortedClassified];
      stats->fresnel = sum[kSortedFresnel];
      stats->rotated = sum[kSortedRotate];
      stats->emitted = sum[kSortedEmitted];
      stats->tail_photons = sum[kSortedTail];
      stats->loop_ms = loop_ms;
      stats->tail_ms = tm.kernel_ms - loop_ms;
    }
    tm.h2d_ms = 0.0;
    tm.total_ms = ms_since(t0);
    if (timing) *timing = tm;
  });
}

__attribute__((visibility("default"))) int optphot_run_gpu_batch(
    const optphot_sim_params* params, const uint64_t* photons, uint32_t n_events,
    uint64_t first_id, uint32_t time_bins, float time_max, uint32_t* event_tally, int method,
    const optphot_wavefront_tuning* tuning, optphot_step_counters* counters,
    optphot_run_timing* timing) {
  return guarded([&] {
    validate_batch(params, photons, n_events, time_bins, time_max, event_tally, method);
    // queue chunk 1024: a lane then takes ~16 photons of one event in a row,
    // so its per-event fate counts flush rarely (measured best for both the
    // 4096-event and the 256-set calibration batches)
    optphot_wavefront_tuning tu{256, 2, 1024, 0, 0, 0};
    if (tuning) {
      if (tuning->block_size) tu.block_size = tuning->block_size;
      if (tuning->steps_per_refill) tu.steps_per_refill = tuning->steps_per_refill;
      if (tuning->queue_chunk) tu.queue_chunk = tuning->queue_chunk;
      if (tuning->blocks_per_cu) tu.blocks_per_cu = tuning->blocks_per_cu;
      tu.stream_in_lds = tuning->stream_in_lds;
    }
    validate_block(tu.block_size, "batch");
    Context& c = ctx();
    std::lock_guard<std::mutex> lock(c.mu);
    c.init();
    optphot_run_timing tm{};
    const auto t0 = std::chrono::steady_clock::now();

    // Host tables: per-event parameters and photon-id offsets.
    std::vector<DevParams> table(n_events);
    std::vector<unsigned long long> offsets(n_events + 1);
    offsets[0] = 0;
    for (uint32_t e = 0; e < n_events; ++e) {
      table[e] = to_dev(params[e]);
      offsets[e + 1] = offsets[e] + photons[e];
    }
    const uint64_t total = offsets[n_events];
    const uint32_t words = 6u + time_bins;
    const size_t count_words = static_cast<size_t>(n_events) * words;
    const size_t zero_u64 = (count_words + 1) / 2;
    const auto ta = std::chrono::steady_clock::now();
    bool grew = c.batch_table.ensure(table.size() * sizeof(DevParams));
    grew |= c.batch_offsets.ensure(offsets.size() * sizeof(unsigned long long));
    grew |= c.batch_counts.ensure(zero_u64 * 8);
    if (grew) {
      tm.alloc_ms += ms_since(ta);
      c.drop_graphs();  // buffers moved: captured graphs are stale
    }
    const bool ring = tu.stream_in_lds != 0 && method == OPTPHOT_BATCH_WAVEFRONT;
    if (ring && batch_ring_lds_bytes(batch_tally_lds_bytes(n_events, words, tu.block_size),
                                     tu.block_size) > c.max_lds)
      throw InvalidArgument("batch: event tallies plus LDS streams exceed the LDS of a block; use "
                            "a smaller block_size");
    WavefrontShape s = [&] {
      const int max_per_cu = std::max(
          1, batch_blocks_per_cu(tu.block_size, n_events, words, ring));
      const int per_cu = tu.blocks_per_cu > 0 ? std::min(tu.blocks_per_cu, max_per_cu) : max_per_cu;
      const uint64_t needed = (total + tu.block_size - 1) / tu.block_size;
      WavefrontShape w{};
      w.block = tu.block_size;
      w.grid = static_cast<int>(std::max<uint64_t>(
          1, std::min<uint64_t>(needed, static_cast<uint64_t>(per_cu) * c.cu_count)));
      w.steps_per_refill = tu.steps_per_refill;
      w.queue_chunk = tu.queue_chunk;
      w.stream_in_lds = ring ? 1u : 0u;
      return w;
    }();
    if (s.queue_chunk < 64 || s.queue_chunk % 64 != 0 || s.steps_per_refill == 0)
      throw InvalidArgument("batch: queue_chunk must be a multiple of 64, steps_per_refill >= 1");
    const std::vector<unsigned long long> slots = batch_slots(
        params, photons, n_events, s.queue_chunk, static_cast<uint64_t>(s.grid) * s.block);
    const auto tsl = std::chrono::steady_clock::now();
    if (c.batch_slots.ensure(std::max<size_t>(1, slots.size()) * sizeof(unsigned long long))) {
      tm.alloc_ms += ms_since(tsl);
      c.drop_graphs();
    }
    unsigned long long* scratch = ensure_scratch(c, 1, tm.alloc_ms);

    const auto th = std::chrono::steady_clock::now();
    HIPCHK(hipMemcpyAsync(c.batch_table.ptr, table.data(), table.size() * sizeof(DevParams),
                          hipMemcpyHostToDevice, c.stream));
    HIPCHK(hipMemcpyAsync(c.batch_offsets.ptr, offsets.data(),
                          offsets.size() * sizeof(unsigned long long), hipMemcpyHostToDevice,
                          c.stream));
    HIPCHK(hipMemcpyAsync(c.batch_slots.ptr, slots.data(),
                          slots.size() * sizeof(unsigned long long), hipMemcpyHostToDevice,
                          c.stream));
    HIPCHK(hipStreamSynchronize(c.stream));
    tm.h2d_ms = ms_since(th);

    LaunchOut lo{};
    lo.events = true;
    lo.counters = counters ? scratch + kScratchCounters : nullptr;
    lo.evt.counts = c.batch_counts.as<uint32_t>();
    lo.evt.words = words;
    lo.evt.time_bins = time_bins;
    lo.evt.time_scale = static_cast<float>(time_bins) / time_max;  // transport.hpp:171
    lo.evt.table = c.batch_table.as<DevParams>();
    lo.evt.offsets = c.batch_offsets.as<unsigned long long>();
    lo.evt.slots = c.batch_slots.as<unsigned long long>();
    lo.evt.n_events = n_events;
    lo.evt.wave_rows = batch_lds_bytes(n_events, words) == 0 &&
                       batch_wave_row_bytes(words, s.block) != 0;
    const PhiloxKey key = key_of(params[0]);

    const uint64_t n_slots = slots.size();
    s.slots_per_claim = slots_per_claim(n_slots, static_cast<uint64_t>(s.grid) * s.block / 64u);
    if (method == OPTPHOT_BATCH_WAVEFRONT) {
      const std::string gkey = std::to_string(n_events) + "/" + std::to_string(n_slots) + "/" +
                               std::to_string(first_id) + "/" + std::to_string(params[0].seed) +
                               "/" + std::to_string(words) + "/" + std::to_string(lo.evt.time_scale) +
                               "/" + std::to_string(s.block) + "/" + std::to_string(s.grid) + "/" +
                               std::to_string(s.steps_per_refill) + "/" +
                               std::to_string(s.queue_chunk) + "/" +
                               std::to_string(s.slots_per_claim) + (counters ? "/c" : "") +
                               (s.stream_in_lds ? "/ring" : "");
      if (!c.batch_exec || c.batch_key != gkey) {
        if (c.batch_exec) HIPCHK(hipGraphExecDestroy(c.batch_exec));
        c.batch_exec = nullptr;
        c.batch_exec = capture_graph(c, [&] {
          launch_zero_u64(scratch, kScratchSums, c.stream);  // queue head + counters
          launch_zero_u64(reinterpret_cast<unsigned long long*>(lo.evt.counts), zero_u64,
                          c.stream);
          launch_batch_wavefront(key, first_id, n_slots, s, scratch + kScratchQueue, lo,
                                 c.stream);
        });
        c.batch_key = gkey;
      }
      HIPCHK(hipEventRecord(c.ev[0], c.stream));
      HIPCHK(hipGraphLaunch(c.batch_exec, c.stream));
      HIPCHK(hipEventRecord(c.ev[1], c.stream));
    } else {
      launch_zero_u64(scratch, kScratchSums, c.stream);
      launch_zero_u64(reinterpret_cast<unsigned long long*>(lo.evt.counts), zero_u64, c.stream);
      HIPCHK(hipGetLastError());
      HIPCHK(hipEventRecord(c.ev[0], c.stream));
      for (uint32_t e = 0; e < n_events; ++e) {
        if (photons[e] == 0) continue;
        lo.evt.event = e;
        launch_history_event(table[e], key, first_id + offsets[e], photons[e], tu.block_size, lo,
                             c.stream);
      }
      HIPCHK(hipGetLastError());
      HIPCHK(hipEventRecord(c.ev[1], c.stream));
    }
    tm.kernel_ms = c.elapsed(0, 1);
    const auto td = std::chrono::steady_clock::now();
    HIPCHK(hipMemcpyAsync(event_tally, lo.evt.counts, count_words * sizeof(uint32_t),
                          hipMemcpyDeviceToHost, c.stream));
    HIPCHK(hipStreamSynchronize(c.stream));
    tm.d2h_ms = ms_since(td);
    read_counters(c, counters);
    tm.total_ms = ms_since(t0);
    if (timing) *timing = tm;
  });
}

__attribute__((visibility("default"))) int optphot_run_gpu_gensteps(
    const optphot_sim_params* p, const optphot_genstep* gensteps, uint32_t n_gensteps,
    uint64_t first_id, const optphot_tally_config* tc, uint64_t* tally,
    const optphot_records* records, const optphot_wavefront_tuning* tuning,
    optphot_step_counters* counters, optphot_run_timing* timing) {
  return guarded([&] {
    validate_params(p);
    if (!tc || !tally) throw InvalidArgument("gensteps: tally config / output is NULL");
    if (n_gensteps > 0 && !gensteps) throw InvalidArgument("gensteps: gensteps is NULL");
    if (records) validate_records(records);
    uint64_t total = 0;
    std::vector<DevGenstep> dev(n_gensteps);
    std::vector<uint2> slots;
    optphot_wavefront_tuning tu{256, 2, 256, 0, 0, 0};
    if (tuning) {
      if (tuning->block_size) tu.block_size = tuning->block_size;
      if (tuning->steps_per_refill) tu.steps_per_refill = tuning->steps_per_refill;
      if (tuning->queue_chunk) tu.queue_chunk = tuning->queue_chunk;
      if (tuning->blocks_per_cu) tu.blocks_per_cu = tuning->blocks_per_cu;
      tu.stream_in_lds = tuning->stream_in_lds;
    }
    validate_block(tu.block_size, "gensteps");
    if (tu.queue_chunk < 64 || tu.queue_chunk % 64 != 0 || tu.steps_per_refill == 0)
      throw InvalidArgument("gensteps: queue_chunk must be a multiple of 64, steps_per_refill >= 1");
    auto inside = [&](float x, float y, float z) {
      return x >= -p->hx && x <= p->hx && y >= -p->hy && y <= p->hy && z >= -p->hz && z <= p->hz;
    };
    for (uint32_t k = 0; k < n_gensteps; ++k) {
      const optphot_genstep& g = gensteps[k];
      if (!inside(g.x0, g.y0, g.z0) || !inside(g.x1, g.y1, g.z1))
        throw InvalidArgument("gensteps: genstep " + std::to_string(k) +
                              " has an end point outside the box");
      if (!(g.t0 >= 0.0f) || !(g.t1 >= 0.0f) || std::isinf(g.t0) || std::isinf(g.t1))
        throw InvalidArgument("gensteps: genstep times must be finite and >= 0");
      if (!(g.cone_cos >= 0.0f && g.cone_cos <= 1.0f))
        throw InvalidArgument("gensteps: genstep " + std::to_string(k) +
                              " has cone_cos outside [0, 1]");
      if (g.cone_cos > 0.0f) {
        // the axis (x1 - x0) / |x1 - x0| of a Cherenkov cone needs a squared
        // length that is a normal float (the device computes it the same way)
        const float dx = g.x1 - g.x0, dy = g.y1 - g.y0, dz = g.z1 - g.z0;
        if (!(dx * dx + dy * dy + dz * dz >= 0x1p-100f))
          throw InvalidArgument("gensteps: Cherenkov genstep " + std::to_string(k) +
                                " needs a segment with a direction (|x1 - x0|^2 >= 2^-100)");
      }
      dev[k] = DevGenstep{g.x0, g.y0, g.z0, g.t0, g.x1, g.y1, g.z1, g.t1, g.n_photons, g.cone_cos,
                          total};
      for (uint64_t c0 = 0; c0 < g.n_photons; c0 += tu.queue_chunk)
        slots.push_back(make_uint2(k, static_cast<uint32_t>(c0 / tu.queue_chunk)));
      total += g.n_photons;
    }
    Context& c = ctx();
    std::lock_guard<std::mutex> lock(c.mu);
    c.init();
    optphot_run_timing tm{};
    const auto t0 = std::chrono::steady_clock::now();
    const DevTally T = to_dev_tally(*tc, *p);
    check_lds(c, T);
    const size_t lds = static_cast<size_t>(T.words) * 4;
    // LDS rings for the Philox windows: tally runs only
    const bool ring = tu.stream_in_lds != 0 && records == nullptr;
    if (ring && ring_lds_bytes(T.words, tu.block_size) > c.max_lds)
      throw InvalidArgument("gensteps: tally plus LDS streams exceed the LDS of a block; use a "
                            "smaller block_size");
    // Fresnel cell bounds in LDS whenever they fit
    const int surf = static_cast<int>(p->surface);
    const bool bounds =
        (ring ? wavefront_ring_lds_bytes(T.words, tu.block_size, surf, true)
              : history_lds_bytes(T.words, true, surf)) <= c.max_lds;
    const int max_per_cu = std::max(
        1, genstep_blocks_per_cu(records != nullptr, tu.block_size, lds, ring, surf, bounds));
    const int per_cu = tu.blocks_per_cu > 0 ? std::min(tu.blocks_per_cu, max_per_cu) : max_per_cu;
    WavefrontShape s{};
    s.block = tu.block_size;
    s.grid = static_cast<int>(std::max<uint64_t>(
        1, std::min<uint64_t>((total + tu.block_size - 1) / tu.block_size,
                              static_cast<uint64_t>(per_cu) * c.cu_count)));
    s.steps_per_refill = tu.steps_per_refill;
    s.queue_chunk = tu.queue_chunk;
    s.stream_in_lds = ring ? 1u : 0u;
    s.fresnel_bounds = bounds;
    s.slots_per_claim = slots_per_claim(slots.size(), static_cast<uint64_t>(s.grid) * s.block / 64u);
    if (total / static_cast<uint64_t>(s.grid) >= 0xffffffffull)
      throw std::runtime_error("too many photons for one tally launch; split the run");

    unsigned long long* scratch = ensure_scratch(c, s.grid, tm.alloc_ms);
    const auto ta = std::chrono::steady_clock::now();
    bool grew = c.gs_data.ensure(std::max<size_t>(1, dev.size()) * sizeof(DevGenstep));
    grew |= c.gs_slots.ensure(std::max<size_t>(1, slots.size()) * sizeof(uint2));
    grew |= c.partials.ensure(static_cast<size_t>(s.grid) * T.words * sizeof(uint32_t));
    grew |= c.tally.ensure((static_cast<size_t>(T.words) + 2) * 8);
    DevRecords rec{};
    if (records && total > 0) rec = ensure_records(c, total, tm.alloc_ms);
    if (grew) tm.alloc_ms += ms_since(ta);
    const auto th = std::chrono::steady_clock::now();
    if (!dev.empty())
      HIPCHK(hipMemcpyAsync(c.gs_data.ptr, dev.data(), dev.size() * sizeof(DevGenstep),
                            hipMemcpyHostToDevice, c.stream));
    if (!slots.empty())
      HIPCHK(hipMemcpyAsync(c.gs_slots.ptr, slots.data(), slots.size() * sizeof(uint2),
                            hipMemcpyHostToDevice, c.stream));
    HIPCHK(hipStreamSynchronize(c.stream));
    tm.h2d_ms = ms_since(th);

    LaunchOut lo{};
    lo.tally = T;
    lo.tally_on = true;
    lo.fresnel_bounds = s.fresnel_bounds;
    lo.records = records != nullptr && total > 0;
    lo.rec = rec;
    lo.partials = c.partials.as<uint32_t>();
    lo.sums = scratch + kScratchSums;
    lo.counters = counters ? scratch + kScratchCounters : nullptr;
    GenstepArgs ga{c.gs_data.as<DevGenstep>(), c.gs_slots.as<uint2>(), slots.size()};
    launch_zero_u64(scratch, kScratchSums + 2 * static_cast<size_t>(s.grid), c.stream);
    launch_zero_u64(c.tally.as<unsigned long long>(), T.words + 2u, c.stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(c.ev[0], c.stream));
    if (total > 0) {
      launch_gensteps(to_dev(*p), key_of(*p), first_id, s, ga, scratch + kScratchQueue, lo,
                      c.stream);
      HIPCHK(hipGetLastError());
      launch_tally_reduce(T, s.grid, lo.partials, lo.sums, c.tally.as<unsigned long long>(),
                          c.stream);
      HIPCHK(hipGetLastError());
    }
    HIPCHK(hipEventRecord(c.ev[1], c.stream));
    tm.kernel_ms = c.elapsed(0, 1);
    const auto td = std::chrono::steady_clock::now();
    HIPCHK(hipMemcpyAsync(tally, c.tally.ptr, (static_cast<size_t>(T.words) + 2) * 8,
                          hipMemcpyDeviceToHost, c.stream));
    if (lo.records) stage_records_to_host(c, rec, total, records, c.ev[1], tm.alloc_ms);
    HIPCHK(hipStreamSynchronize(c.stream));
    tm.d2h_ms = ms_since(td);
    read_counters(c, counters);
    tm.total_ms = ms_since(t0);
    if (timing) *timing = tm;
  });
}

__attribute__((visibility("default"))) int optphot_run_gpu_chi2_step(
    const optphot_sim_params* params, const uint64_t* photons, uint32_t n_sets,
    uint64_t first_id, uint32_t time_bins, float time_max, const double* target,
    const double* target_photons, double* hist, double* loss, double* grad,
    const optphot_wavefront_tuning* forward_tuning, const optphot_wavefront_tuning* adjoint_tuning,
    optphot_run_timing* timing) {
  return guarded([&] {
    if (!target || !target_photons || !hist || !loss || !grad)
      throw InvalidArgument("chi2 step: NULL argument");
    validate_batch(params, photons, n_sets, time_bins, time_max,
                   reinterpret_cast<uint32_t*>(hist), OPTPHOT_BATCH_WAVEFRONT);
    validate_adjoint_sets(params, n_sets);
    for (uint32_t e = 0; e < n_sets; ++e)
      if (!(target_photons[e] > 0.0))
        throw InvalidArgument("chi2 step: target_photons must be > 0");
    auto merge = [](const optphot_wavefront_tuning* t, optphot_wavefront_tuning d) {
      if (t) {
        if (t->block_size) d.block_size = t->block_size;
        if (t->steps_per_refill) d.steps_per_refill = t->steps_per_refill;
        if (t->queue_chunk) d.queue_chunk = t->queue_chunk;
        if (t->blocks_per_cu) d.blocks_per_cu = t->blocks_per_cu;
        d.stream_in_lds = t->stream_in_lds;
      }
      validate_block(d.block_size, "chi2 step");
      if (d.queue_chunk < 64 || d.queue_chunk % 64 != 0 || d.steps_per_refill == 0)
        throw InvalidArgument("chi2 step: queue_chunk must be a multiple of 64, "
                              "steps_per_refill >= 1");
      return d;
    };
    const optphot_wavefront_tuning tf = merge(forward_tuning, {256, 2, 1024, 0, 0, 0});
    const optphot_wavefront_tuning tb = merge(adjoint_tuning, {256, 2, 256, 0, 0, 0});
    Context& c = ctx();
    std::lock_guard<std::mutex> lock(c.mu);
    c.init();
    optphot_run_timing tm{};
    const auto t0 = std::chrono::steady_clock::now();

    std::vector<DevParams> table(n_sets);
    std::vector<unsigned long long> offsets(n_sets + 1);
    std::vector<double> inv_nt(n_sets);
    offsets[0] = 0;
    for (uint32_t e = 0; e < n_sets; ++e) {
      table[e] = to_dev(params[e]);
      offsets[e + 1] = offsets[e] + photons[e];
      inv_nt[e] = 1.0 / target_photons[e];
    }
    const uint64_t total = offsets[n_sets];
    const uint32_t words = 6u + time_bins;
    const uint32_t bins1 = time_bins + 1u;
    const size_t count_words = static_cast<size_t>(n_sets) * words;
    const size_t zero_u64 = (count_words + 1) / 2;
    const size_t wcount = static_cast<size_t>(n_sets) * bins1;

    const auto ta = std::chrono::steady_clock::now();
    bool grew = c.batch_table.ensure(table.size() * sizeof(DevParams));
    grew |= c.batch_offsets.ensure(offsets.size() * 8);
    grew |= c.batch_counts.ensure(zero_u64 * 8);
    grew |= c.adj_weights.ensure(wcount * 8);
    grew |= c.adj_grad.ensure(static_cast<size_t>(n_sets) * 5 * 8);
    grew |= c.adj_queue.ensure((static_cast<size_t>(n_sets) + 1) * 8);
    grew |= c.step_target.ensure(wcount * 8);
    grew |= c.step_inv_nt.ensure(static_cast<size_t>(n_sets) * 8);
    grew |= c.step_hist.ensure(wcount * 8);
    grew |= c.step_loss.ensure(static_cast<size_t>(n_sets) * 8);
    if (grew) {
      tm.alloc_ms += ms_since(ta);
      c.drop_graphs();
    }
    unsigned long long* scratch = ensure_scratch(c, 1, tm.alloc_ms);

    const auto th = std::chrono::steady_clock::now();
    HIPCHK(hipMemcpyAsync(c.batch_table.ptr, table.data(), table.size() * sizeof(DevParams),
                          hipMemcpyHostToDevice, c.stream));
    HIPCHK(hipMemcpyAsync(c.batch_offsets.ptr, offsets.data(), offsets.size() * 8,
                          hipMemcpyHostToDevice, c.stream));
    HIPCHK(hipMemcpyAsync(c.step_target.ptr, target, wcount * 8, hipMemcpyHostToDevice,
                          c.stream));
    HIPCHK(hipMemcpyAsync(c.step_inv_nt.ptr, inv_nt.data(), inv_nt.size() * 8,
                          hipMemcpyHostToDevice, c.stream));
    HIPCHK(hipStreamSynchronize(c.stream));
    tm.h2d_ms = ms_since(th);

    // forward: the batched launch (as optphot_run_gpu_batch, wavefront method)
    WavefrontShape sf{};
    {
      const bool ring = tf.stream_in_lds != 0;
      if (ring && batch_ring_lds_bytes(batch_tally_lds_bytes(n_sets, words, tf.block_size),
                                       tf.block_size) > c.max_lds)
        throw InvalidArgument("chi2 step: event tallies plus LDS streams exceed the LDS of a "
                              "block; use a smaller block_size");
      const int max_per_cu = std::max(
          1, batch_blocks_per_cu(tf.block_size, n_sets, words, ring));
      const int per_cu =
          tf.blocks_per_cu > 0 ? std::min(tf.blocks_per_cu, max_per_cu) : max_per_cu;
      const uint64_t needed = (total + tf.block_size - 1) / tf.block_size;
      sf.block = tf.block_size;
      sf.grid = static_cast<int>(std::max<uint64_t>(
          1, std::min<uint64_t>(needed, static_cast<uint64_t>(per_cu) * c.cu_count)));
      sf.steps_per_refill = tf.steps_per_refill;
      sf.queue_chunk = tf.queue_chunk;
      sf.stream_in_lds = ring ? 1u : 0u;
    }
    const std::vector<unsigned long long> slots = batch_slots(
        params, photons, n_sets, sf.queue_chunk, static_cast<uint64_t>(sf.grid) * sf.block);
    const auto tsl = std::chrono::steady_clock::now();
    if (c.batch_slots.ensure(std::max<size_t>(1, slots.size()) * sizeof(unsigned long long))) {
      tm.alloc_ms += ms_since(tsl);
      c.drop_graphs();
    }
    HIPCHK(hipMemcpyAsync(c.batch_slots.ptr, slots.data(),
                          slots.size() * sizeof(unsigned long long), hipMemcpyHostToDevice,
                          c.stream));
    LaunchOut lo{};
    lo.events = true;
    lo.evt.counts = c.batch_counts.as<uint32_t>();
    lo.evt.words = words;
    lo.evt.time_bins = time_bins;
    lo.evt.time_scale = static_cast<float>(time_bins) / time_max;  // transport.hpp:171
    lo.evt.table = c.batch_table.as<DevParams>();
    lo.evt.offsets = c.batch_offsets.as<unsigned long long>();
    lo.evt.slots = c.batch_slots.as<unsigned long long>();
    lo.evt.n_events = n_sets;
    lo.evt.wave_rows = batch_lds_bytes(n_sets, words) == 0 &&
                       batch_wave_row_bytes(words, sf.block) != 0;
    const PhiloxKey key = key_of(params[0]);
    const uint64_t n_slots = slots.size();
    sf.slots_per_claim = slots_per_claim(n_slots, static_cast<uint64_t>(sf.grid) * sf.block / 64u);

    // adjoint shape and arguments
    WavefrontShape sb{};
    {
      const int max_per_cu = std::max(1, adjoint_blocks_per_cu(tb.block_size));
      const int per_cu =
          tb.blocks_per_cu > 0 ? std::min(tb.blocks_per_cu, max_per_cu) : max_per_cu;
      const uint64_t needed = (total + tb.block_size - 1) / tb.block_size;
      sb.block = tb.block_size;
      sb.grid = static_cast<int>(std::max<uint64_t>(
          1, std::min<uint64_t>(needed, static_cast<uint64_t>(per_cu) * c.cu_count)));
      sb.steps_per_refill = tb.steps_per_refill;
      sb.queue_chunk = tb.queue_chunk;
    }
    AdjointArgs a{};
    a.table = c.batch_table.as<DevParams>();
    a.offsets = c.batch_offsets.as<unsigned long long>();
    a.qheads = c.adj_queue.as<unsigned long long>();
    a.closed = c.adj_queue.as<unsigned long long>() + n_sets;
    a.n_events = n_sets;
    a.time_bins = time_bins;
    a.time_scale = lo.evt.time_scale;
    a.weights = c.adj_weights.as<double>();
    a.grad = c.adj_grad.as<double>();
    int surface = static_cast<int>(params[0].surface);
    for (uint32_t e = 1; e < n_sets; ++e)
      if (static_cast<int>(params[e].surface) != surface) surface = -1;

    // One graph: forward, chi-square weights, adjoint. The parameters, photon
    // counts' offsets and targets are read from device buffers, so a replay
    // with new values uploaded above is the same computation.
    const std::string gkey =
        std::to_string(n_sets) + "/" + std::to_string(n_slots) + "/" + std::to_string(first_id) +
        "/" + std::to_string(params[0].seed) + "/" + std::to_string(words) + "/" +
        std::to_string(lo.evt.time_scale) + "/" + std::to_string(surface) + "/f" +
        std::to_string(sf.block) + "," + std::to_string(sf.grid) + "," +
        std::to_string(sf.steps_per_refill) + "," + std::to_string(sf.queue_chunk) + "," +
        std::to_string(sf.slots_per_claim) + (sf.stream_in_lds ? ",ring" : "") + "/b" +
        std::to_string(sb.block) + "," + std::to_string(sb.grid) + "," +
        std::to_string(sb.steps_per_refill) + "," + std::to_string(sb.queue_chunk);
    if (!c.step_exec || c.step_key != gkey) {
      if (c.step_exec) HIPCHK(hipGraphExecDestroy(c.step_exec));
      c.step_exec = nullptr;
      c.step_exec = capture_graph(c, [&] {
        launch_zero_u64(scratch, kScratchSums, c.stream);  // queue head + counters
        launch_zero_u64(reinterpret_cast<unsigned long long*>(lo.evt.counts), zero_u64,
                        c.stream);
        launch_batch_wavefront(key, first_id, n_slots, sf, scratch + kScratchQueue, lo,
                               c.stream);
        launch_chi2_weights(lo.evt.counts, words, bins1, a.offsets, c.step_target.as<double>(),
                            c.step_inv_nt.as<double>(), c.step_hist.as<double>(),
                            c.step_loss.as<double>(), c.adj_weights.as<double>(), n_sets,
                            c.stream);
        launch_zero_u64(a.qheads, static_cast<uint64_t>(n_sets) + 1, c.stream);
        launch_zero_u64(reinterpret_cast<unsigned long long*>(a.grad),
                        static_cast<uint64_t>(n_sets) * 5, c.stream);
        launch_adjoint(table[0], surface, key, first_id, sb, a, c.stream, false);
      });
      c.step_key = gkey;
    }
    HIPCHK(hipEventRecord(c.ev[0], c.stream));
    HIPCHK(hipGraphLaunch(c.step_exec, c.stream));
    HIPCHK(hipEventRecord(c.ev[1], c.stream));
    tm.kernel_ms = c.elapsed(0, 1);
    const auto td = std::chrono::steady_clock::now();
    HIPCHK(hipMemcpyAsync(hist, c.step_hist.ptr, wcount * 8, hipMemcpyDeviceToHost, c.stream));
    HIPCHK(hipMemcpyAsync(loss, c.step_loss.ptr, static_cast<size_t>(n_sets) * 8,
                          hipMemcpyDeviceToHost, c.stream));
    HIPCHK(hipMemcpyAsync(grad, a.grad, static_cast<size_t>(n_sets) * 5 * 8,
                          hipMemcpyDeviceToHost, c.stream));
    HIPCHK(hipStreamSynchronize(c.stream));
    tm.d2h_ms = ms_since(td);
    for (uint32_t e = 0; e < n_sets; ++e) {
      const double inv = photons[e] ? 1.0 / static_cast<double>(photons[e]) : 0.0;
      for (int j = 0; j < 5; ++j) grad[5 * e + j] *= inv;
    }
    tm.total_ms = ms_since(t0);
    if (timing) *timing = tm;
  });
}

__attribute__((visibility("default"))) int optphot_run_gpu_adjoint(
    const optphot_sim_params* params, const uint64_t* photons, uint32_t n_events,
    uint64_t first_id, uint32_t time_bins, float time_max, const double* weights, double* grad,
    const optphot_wavefront_tuning* tuning, optphot_run_timing* timing) {
  return guarded([&] {
    if (!weights || !grad) throw InvalidArgument("adjoint: NULL weights / grad");
    validate_batch(params, photons, n_events, time_bins, time_max,
                   reinterpret_cast<uint32_t*>(grad), OPTPHOT_BATCH_WAVEFRONT);
    validate_adjoint_sets(params, n_events);
    optphot_wavefront_tuning tu{256, 2, 256, 0, 0, 0};
    if (tuning) {
      if (tuning->block_size) tu.block_size = tuning->block_size;
      if (tuning->steps_per_refill) tu.steps_per_refill = tuning->steps_per_refill;
      if (tuning->queue_chunk) tu.queue_chunk = tuning->queue_chunk;
      if (tuning->blocks_per_cu) tu.blocks_per_cu = tuning->blocks_per_cu;
    }
    validate_block(tu.block_size, "adjoint");
    if (tu.queue_chunk < 64 || tu.queue_chunk % 64 != 0 || tu.steps_per_refill == 0)
      throw InvalidArgument("adjoint: queue_chunk must be a multiple of 64, steps_per_refill >= 1");
    Context& c = ctx();
    std::lock_guard<std::mutex> lock(c.mu);
    c.init();
    optphot_run_timing tm{};
    const auto t0 = std::chrono::steady_clock::now();
    std::vector<DevParams> table(n_events);
    std::vector<unsigned long long> offsets(n_events + 1);
    offsets[0] = 0;
    for (uint32_t e = 0; e < n_events; ++e) {
      table[e] = to_dev(params[e]);
      offsets[e + 1] = offsets[e] + photons[e];
    }
    const uint64_t total = offsets[n_events];
    const size_t wcount = static_cast<size_t>(n_events) * (time_bins + 1u);
    const auto ta = std::chrono::steady_clock::now();
    bool grew = c.batch_table.ensure(table.size() * sizeof(DevParams));
    grew |= c.batch_offsets.ensure(offsets.size() * 8);
    grew |= c.adj_weights.ensure(wcount * 8);
    grew |= c.adj_grad.ensure(static_cast<size_t>(n_events) * 5 * 8);
    grew |= c.adj_queue.ensure((static_cast<size_t>(n_events) + 1) * 8);
    if (grew) {
      tm.alloc_ms += ms_since(ta);
      c.drop_graphs();  // captured graphs refer to the old buffers
    }
    const auto th = std::chrono::steady_clock::now();
    HIPCHK(hipMemcpyAsync(c.batch_table.ptr, table.data(), table.size() * sizeof(DevParams),
                          hipMemcpyHostToDevice, c.stream));
    HIPCHK(hipMemcpyAsync(c.batch_offsets.ptr, offsets.data(), offsets.size() * 8,
                          hipMemcpyHostToDevice, c.stream));
    HIPCHK(hipMemcpyAsync(c.adj_weights.ptr, weights, wcount * 8, hipMemcpyHostToDevice,
                          c.stream));
    HIPCHK(hipStreamSynchronize(c.stream));
    tm.h2d_ms = ms_since(th);

    const int max_per_cu = std::max(1, adjoint_blocks_per_cu(tu.block_size));
    const int per_cu = tu.blocks_per_cu > 0 ? std::min(tu.blocks_per_cu, max_per_cu) : max_per_cu;
    const uint64_t needed = (total + tu.block_size - 1) / tu.block_size;
    WavefrontShape s{};
    s.block = tu.block_size;
    s.grid = static_cast<int>(std::max<uint64_t>(
        1, std::min<uint64_t>(needed, static_cast<uint64_t>(per_cu) * c.cu_count)));
    s.steps_per_refill = tu.steps_per_refill;
    s.queue_chunk = tu.queue_chunk;

    AdjointArgs a{};
    a.table = c.batch_table.as<DevParams>();
    a.offsets = c.batch_offsets.as<unsigned long long>();
    a.qheads = c.adj_queue.as<unsigned long long>();
    a.closed = c.adj_queue.as<unsigned long long>() + n_events;
    a.n_events = n_events;
    a.time_bins = time_bins;
    a.time_scale = static_cast<float>(time_bins) / time_max;  // transport.hpp:171
    a.weights = c.adj_weights.as<double>();
    a.grad = c.adj_grad.as<double>();
    launch_zero_u64(a.qheads, static_cast<uint64_t>(n_events) + 1, c.stream);
    launch_zero_u64(reinterpret_cast<unsigned long long*>(a.grad),
                    static_cast<uint64_t>(n_events) * 5, c.stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(c.ev[0], c.stream));
    int surface = static_cast<int>(params[0].surface);
    for (uint32_t e = 1; e < n_events; ++e)
      if (static_cast<int>(params[e].surface) != surface) surface = -1;
    launch_adjoint(table[0], surface, key_of(params[0]), first_id, s, a, c.stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(c.ev[1], c.stream));
    tm.kernel_ms = c.elapsed(0, 1);
    const auto td = std::chrono::steady_clock::now();
    HIPCHK(hipMemcpyAsync(grad, a.grad, static_cast<size_t>(n_events) * 5 * 8,
                          hipMemcpyDeviceToHost, c.stream));
    HIPCHK(hipStreamSynchronize(c.stream));
    tm.d2h_ms = ms_since(td);
    for (uint32_t e = 0; e < n_events; ++e) {
      const double inv = photons[e] ? 1.0 / static_cast<double>(photons[e]) : 0.0;
      for (int j = 0; j < 5; ++j) grad[5 * e + j] *= inv;
    }
    tm.total_ms = ms_since(t0);
    if (timing) *timing = tm;
  });
}

}  // extern "C"
