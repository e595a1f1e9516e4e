// backend.hpp -- the reference's GPU backend API (include/optphot/backend.hpp:
// 26-115) on top of liboptphot_hip's C-ABI: same types, names, default
// arguments, synchronous semantics and exception types, so that code calling
// run_gpu_records / run_gpu_tally / run_gpu_records_event compiles and links
// against this header and -loptphot_hip unchanged. Errors of the C-ABI come
// back as std::invalid_argument (bad arguments) or std::runtime_error
// (device / runtime failures), with the library's message.
//
// Differences a caller can see: block sizes must be multiples of 64 (the
// wavefront width of CDNA) instead of 32, and EventStats counts issued lane
// steps per 64-lane wavefront. PhotonRecords::set(i, PhotonResult) of the
// reference's CPU backend has no counterpart (there is no host transport here).
//
// Header only: include it and link liboptphot_hip.so.
#pragma once

#include <array>
#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../optphot_hip.h"
#include "params.hpp"

namespace optphot {

// Structure-of-arrays photon records, indexed by photon number
// (backend.hpp:31-50): 29 bytes per photon.
struct PhotonRecords {
  std::vector<uint8_t> fate;
  std::vector<uint32_t> n_boundary;
  std::vector<uint32_t> n_scatter;
  std::vector<float> t, x, y, z, path;

  void resize(std::size_t n) {
    fate.resize(n);
    n_boundary.resize(n);
    n_scatter.resize(n);
    t.resize(n);
    x.resize(n);
    y.resize(n);
    z.resize(n);
    path.resize(n);
  }
  std::size_t size() const { return fate.size(); }
  static constexpr std::size_t kBytesPerPhoton = 1 + 4 + 4 + 5 * 4;
};

// Aggregated results (backend.hpp:52-62).
struct Tally {
  std::array<uint64_t, kNumFates> fate_counts{};
  std::vector<uint64_t> time_hist;  // time_bins + 1 (last = overflow)
  std::vector<uint64_t> xy_hist;    // xy_bins * xy_bins, row-major [iy][ix]
  uint64_t boundary_sum = 0;
  uint64_t scatter_sum = 0;

  void reset(const TallyConfig& c) {
    fate_counts.fill(0);
    time_hist.assign(static_cast<std::size_t>(c.time_bins) + 1, 0);
    xy_hist.assign(static_cast<std::size_t>(c.xy_bins) * c.xy_bins, 0);
    boundary_sum = scatter_sum = 0;
  }
  void add(const Tally& o) {
    for (int i = 0; i < kNumFates; ++i) fate_counts[i] += o.fate_counts[i];
    for (std::size_t i = 0; i < time_hist.size(); ++i) time_hist[i] += o.time_hist[i];
    for (std::size_t i = 0; i < xy_hist.size(); ++i) xy_hist[i] += o.xy_hist[i];
    boundary_sum += o.boundary_sum;
    scatter_sum += o.scatter_sum;
  }
  uint64_t total() const {
    uint64_t s = 0;
    for (uint64_t c : fate_counts) s += c;
    return s;
  }
};

// Milliseconds (backend.hpp:65-71).
struct RunTiming {
  double total_ms = 0.0;
  double kernel_ms = 0.0;
  double h2d_ms = 0.0;
  double d2h_ms = 0.0;
  double alloc_ms = 0.0;
};

// Diagnostics of a wavefront run (backend.hpp:95-103); lanes are 64 wide.
struct EventStats {
  uint64_t launches = 0;
  uint64_t useful_lane_steps = 0;
  uint64_t issued_lane_steps = 0;
  double kernels_ms = 0.0;
  double simt_efficiency() const {
    return issued_lane_steps ? static_cast<double>(useful_lane_steps) / issued_lane_steps : 0.0;
  }
};

static_assert(sizeof(RunTiming) == sizeof(optphot_run_timing), "RunTiming layout");
static_assert(sizeof(EventStats) == sizeof(optphot_event_stats), "EventStats layout");

namespace detail {

inline void check(int rc) {
  if (rc == OPTPHOT_OK) return;
  const std::string msg = optphot_last_error();
  if (rc == OPTPHOT_ERR_INVALID_ARGUMENT) throw std::invalid_argument(msg);
  throw std::runtime_error(msg);
}

inline optphot_records view(PhotonRecords& r) {
  return optphot_records{r.fate.data(), r.n_boundary.data(), r.n_scatter.data(), r.t.data(),
                         r.x.data(),    r.y.data(),          r.z.data(),         r.path.data()};
}

inline void unpack(const std::vector<uint64_t>& w, const TallyConfig& c, Tally& out) {
  out.reset(c);
  std::size_t k = 0;
  for (int f = 0; f < kNumFates; ++f) out.fate_counts[f] = w[k++];
  for (auto& v : out.time_hist) v = w[k++];
  for (auto& v : out.xy_hist) v = w[k++];
  out.boundary_sum = w[k++];
  out.scatter_sum = w[k++];
}

inline std::size_t tally_words(const TallyConfig& c) {
  return kNumFates + static_cast<std::size_t>(c.time_bins) + 1 +
         static_cast<std::size_t>(c.xy_bins) * c.xy_bins + 2;
}

}  // namespace detail

// ---- GPU backend (backend.hpp:84-91) ---------------------------------------

inline bool gpu_available() { return optphot_gpu_available() != 0; }

inline std::string gpu_device_name() {
  char buf[256];
  detail::check(optphot_gpu_device_name(buf, sizeof(buf)));
  return buf;
}

// History-based kernel, one lane per photon; photons first_id .. first_id+n-1,
// produced in chunks of at most `chunk` photons (gpu_backend.cu:119-154).
inline RunTiming run_gpu_records(const SimParams& p, uint64_t n, PhotonRecords& out,
                                 int block_size = 256, uint64_t first_id = 0,
                                 uint64_t chunk = 1ull << 24) {
  out.resize(n);
  const optphot_records v = detail::view(out);
  RunTiming tm;
  detail::check(optphot_run_gpu_records(c_params(p), n, &v, block_size, first_id, chunk,
                                        reinterpret_cast<optphot_run_timing*>(&tm)));
  return tm;
}

// History-based tally (gpu_backend.cu:158-212): fate counts of every photon,
// histograms of detected photons, boundary / scatter sums of every photon.
inline RunTiming run_gpu_tally(const SimParams& p, const TallyConfig& c, uint64_t n, Tally& out,
                               int block_size = 256, uint64_t first_id = 0) {
  std::vector<uint64_t> w(detail::tally_words(c));
  RunTiming tm;
  detail::check(optphot_run_gpu_tally(c_params(p), c_tally_config(c), n, w.data(), block_size,
                                      first_id, reinterpret_cast<optphot_run_timing*>(&tm)));
  detail::unpack(w, c, out);
  return tm;
}

// ---- Wavefront backend (backend.hpp:105-115) --------------------------------

// The same records as run_gpu_records, bit for bit, from the persistent
// wavefront kernel: finished lanes refill from a device work queue every
// steps_per_launch steps.
inline RunTiming run_gpu_records_event(const SimParams& p, uint64_t n, PhotonRecords& out,
                                       uint32_t steps_per_launch, int block_size = 256,
                                       uint64_t first_id = 0, uint64_t chunk = 1ull << 24,
                                       EventStats* stats = nullptr) {
  out.resize(n);
  const optphot_records v = detail::view(out);
  RunTiming tm;
  detail::check(optphot_run_gpu_records_event(
      c_params(p), n, &v, steps_per_launch, block_size, first_id, chunk,
      reinterpret_cast<optphot_event_stats*>(stats), reinterpret_cast<optphot_run_timing*>(&tm)));
  return tm;
}

// ---- Extensions (no counterpart in the reference) ----------------------------

using WavefrontTuning = optphot_wavefront_tuning;

// The same tally as run_gpu_tally, word for word, from the persistent
// wavefront kernel. A zero tuning field takes the tuned / default value.
inline RunTiming run_gpu_tally_wavefront(const SimParams& p, const TallyConfig& c, uint64_t n,
                                         Tally& out, uint64_t first_id = 0,
                                         const WavefrontTuning* tuning = nullptr,
                                         EventStats* stats = nullptr) {
  std::vector<uint64_t> w(detail::tally_words(c));
  RunTiming tm;
  detail::check(optphot_run_gpu_tally_wavefront(
      c_params(p), c_tally_config(c), n, w.data(), first_id, tuning,
      reinterpret_cast<optphot_event_stats*>(stats), nullptr,
      reinterpret_cast<optphot_run_timing*>(&tm)));
  detail::unpack(w, c, out);
  return tm;
}

// The same tally from the per-interaction path (per-type queues and one
// kernel per interaction; the design of the reference's README.md:717-720).
using SortedTuning = optphot_sorted_tuning;
using SortedStats = optphot_sorted_stats;
inline RunTiming run_gpu_tally_sorted(const SimParams& p, const TallyConfig& c, uint64_t n,
                                      Tally& out, uint64_t first_id = 0,
                                      const SortedTuning* tuning = nullptr,
                                      SortedStats* stats = nullptr) {
  std::vector<uint64_t> w(detail::tally_words(c));
  RunTiming tm;
  detail::check(optphot_run_gpu_tally_sorted(c_params(p), c_tally_config(c), n, w.data(),
                                             first_id, tuning, stats,
                                             reinterpret_cast<optphot_run_timing*>(&tm)));
  detail::unpack(w, c, out);
  return tm;
}

// Records written straight into caller-owned device arrays (n elements each,
// e.g. hipMalloc'ed or a framework's tensors): no copy to the host.
// steps_per_launch 0 = history kernel, >= 1 = wavefront kernel.
inline RunTiming run_gpu_records_device(const SimParams& p, uint64_t n,
                                        const optphot_records& device_out,
                                        uint32_t steps_per_launch = 0, int block_size = 256,
                                        uint64_t first_id = 0, EventStats* stats = nullptr) {
  RunTiming tm;
  detail::check(optphot_run_gpu_records_device(
      c_params(p), n, &device_out, steps_per_launch, block_size, first_id,
      reinterpret_cast<optphot_event_stats*>(stats), reinterpret_cast<optphot_run_timing*>(&tm)));
  return tm;
}

// Photons generated on the device from gensteps (see optphot_run_gpu_gensteps):
// the run's tally, and with `records` the records of all photons at their
// global index.
using Genstep = optphot_genstep;
inline RunTiming run_gpu_gensteps(const SimParams& p, const TallyConfig& c,
                                  const std::vector<Genstep>& gensteps, Tally& out,
                                  uint64_t first_id = 0, PhotonRecords* records = nullptr,
                                  const WavefrontTuning* tuning = nullptr) {
  std::vector<uint64_t> w(detail::tally_words(c));
  uint64_t total = 0;
  for (const Genstep& g : gensteps) total += g.n_photons;
  optphot_records v{};
  if (records) {
    records->resize(total);
    v = detail::view(*records);
  }
  RunTiming tm;
  detail::check(optphot_run_gpu_gensteps(
      c_params(p), gensteps.data(), static_cast<uint32_t>(gensteps.size()), first_id,
      c_tally_config(c), w.data(), records ? &v : nullptr, tuning, nullptr,
      reinterpret_cast<optphot_run_timing*>(&tm)));
  detail::unpack(w, c, out);
  return tm;
}

// Per-event results of a batch: 5 fate counts, then the arrival-time
// histogram of detected photons (time_bins + 1 bins, last = overflow).
struct EventTally {
  uint32_t time_bins = 0;
  std::vector<uint32_t> counts;  // [n_events][kNumFates + time_bins + 1]
  std::size_t row() const { return kNumFates + static_cast<std::size_t>(time_bins) + 1; }
  const uint32_t* event(std::size_t e) const { return counts.data() + e * row(); }
};

// Batched events in one persistent launch (method OPTPHOT_BATCH_WAVEFRONT) or
// one history launch per event (OPTPHOT_BATCH_PLAIN). Event e uses photon ids
// first_id + photons[0] + ... + photons[e-1] onwards and the run seed
// params[0].seed, so it equals run_gpu_tally(params[e], ..., photons[e], ...,
// that first id) on the fate counts and arrival-time histogram.
inline RunTiming run_gpu_batch(const std::vector<SimParams>& params,
                               const std::vector<uint64_t>& photons, uint32_t time_bins,
                               float time_max, EventTally& out, uint64_t first_id = 0,
                               int method = OPTPHOT_BATCH_WAVEFRONT,
                               const WavefrontTuning* tuning = nullptr) {
  if (params.size() != photons.size())
    throw std::invalid_argument("batch: params and photons differ in length");
  out.time_bins = time_bins;
  out.counts.assign(params.size() * out.row(), 0);
  RunTiming tm;
  detail::check(optphot_run_gpu_batch(
      reinterpret_cast<const optphot_sim_params*>(params.data()), photons.data(),
      static_cast<uint32_t>(params.size()), first_id, time_bins, time_max, out.counts.data(),
      method, tuning, nullptr, reinterpret_cast<optphot_run_timing*>(&tm)));
  return tm;
}

// Vector-Jacobian product of the per-set arrival-time histograms (fractions
// of each set's photons, overflow bin included) with weights
// [n_sets][time_bins + 1]; returns grad [n_sets][5] with respect to
// (abs_length, scat_length, reflectivity, n_det, n_scint).
inline RunTiming run_gpu_adjoint(const std::vector<SimParams>& params,
                                 const std::vector<uint64_t>& photons, uint32_t time_bins,
                                 float time_max, const std::vector<double>& weights,
                                 std::vector<double>& grad, uint64_t first_id = 0,
                                 const WavefrontTuning* tuning = nullptr) {
  if (params.size() != photons.size() ||
      weights.size() != params.size() * (static_cast<std::size_t>(time_bins) + 1))
    throw std::invalid_argument("adjoint: params / photons / weights sizes disagree");
  grad.assign(params.size() * 5, 0.0);
  RunTiming tm;
  detail::check(optphot_run_gpu_adjoint(
      reinterpret_cast<const optphot_sim_params*>(params.data()), photons.data(),
      static_cast<uint32_t>(params.size()), first_id, time_bins, time_max, weights.data(),
      grad.data(), tuning, reinterpret_cast<optphot_run_timing*>(&tm)));
  return tm;
}

}  // namespace optphot
