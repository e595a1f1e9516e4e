// optphot -- run driver of liboptphot_hip.
//
// Takes the reference's run description (.cfg files and --key value
// overrides, the keys of src/core/config.cpp) and its output layout (.npy
// arrays + meta.json under --out), runs the transport on the GPU and prints a
// run report: which kernel ran and with what launch shape, the fate shares,
// the photon-steps taken (n_boundary + n_scatter + [absorbed in the bulk],
// the reference's definition in apps/optphot_event_bench.cpp:39-41), the
// SIMT efficiency where the kernel counts it, and the time split.
//
// Options of this driver (besides the shared run settings):
//   --kernel history|wavefront   history: one lane per photon (the reference's
//                                plain method); wavefront: the persistent
//                                kernel with a device work queue. Default:
//                                history for records, wavefront for tallies.
//   --steps_per_refill K         wavefront steps between refills (default:
//                                the library's choice)
//   --count 0|1                  tally runs: use the branch-counting kernel
//                                variant, which reports SIMT efficiency
//                                (slower; results identical)
//   --report text|json           report format (json: one line)
// Every kernel choice gives the same records and tallies, bit for bit.
#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <exception>
#include <iostream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "optphot_hip/backend.hpp"
#include "optphot_hip/config.hpp"

using namespace optphot;

namespace {

struct DriverOptions {
  std::string kernel;  // "" = by mode
  uint32_t steps_per_refill = 0;
  bool count = false;
  bool json = false;
};

// What a run produced, independent of the kernel that produced it.
struct Report {
  std::string kernel, shape;
  uint64_t photons = 0;
  std::array<uint64_t, kNumFates> fates{};
  uint64_t photon_steps = 0;
  double simt = -1.0;  // < 0: not measured
  RunTiming timing{};
};

uint32_t parse_u32(const std::string& key, const std::string& v, uint32_t lo) {
  char* end = nullptr;
  const unsigned long x = std::strtoul(v.c_str(), &end, 10);
  if (v.empty() || *end != '\0' || x < lo || x > 0xffffffffUL)
    throw std::runtime_error(key + " must be an integer >= " + std::to_string(lo));
  return static_cast<uint32_t>(x);
}

// Splits argv into this driver's options and the shared run settings.
std::vector<char*> take_driver_options(int argc, char** argv, DriverOptions& d) {
  const std::map<std::string, int> own = {
      {"--kernel", 0}, {"--steps_per_refill", 1}, {"--count", 2}, {"--report", 3}};
  std::vector<char*> rest{argv[0]};
  for (int i = 1; i < argc; ++i) {
    std::string a = argv[i], v;
    const std::size_t eq = a.find('=');
    const auto it = own.find(a.substr(0, eq));
    if (it == own.end()) {
      rest.push_back(argv[i]);
      continue;
    }
    if (eq != std::string::npos) {
      v = a.substr(eq + 1);
    } else {
      if (i + 1 >= argc) throw std::runtime_error(it->first + " needs a value");
      v = argv[++i];
    }
    switch (it->second) {
      case 0:
        if (v != "history" && v != "wavefront")
          throw std::runtime_error("--kernel takes history or wavefront");
        d.kernel = v;
        break;
      case 1:
        d.steps_per_refill = parse_u32("--steps_per_refill", v, 1);
        break;
      case 2:
        d.count = parse_u32("--count", v, 0) != 0;
        break;
      default:
        if (v != "text" && v != "json") throw std::runtime_error("--report takes text or json");
        d.json = v == "json";
    }
  }
  return rest;
}

Report run_records_mode(const SimParams& p, const RunOptions& o, const DriverOptions& d,
                        PhotonRecords& rec) {
  Report r;
  r.photons = o.n_photons;
  if (d.kernel == "wavefront") {
    const uint32_t k = d.steps_per_refill ? d.steps_per_refill : 2u;
    EventStats st;
    r.timing = run_gpu_records_event(p, o.n_photons, rec, k, o.block_size, 0, 1ull << 24, &st);
    r.kernel = "wavefront";
    r.shape = "block " + std::to_string(o.block_size) + ", " + std::to_string(k) +
              " steps per refill";
    r.simt = st.issued_lane_steps ? static_cast<double>(st.useful_lane_steps) /
                                        static_cast<double>(st.issued_lane_steps)
                                  : -1.0;
  } else {
    r.timing = run_gpu_records(p, o.n_photons, rec, o.block_size);
    r.kernel = "history";
    r.shape = "block " + std::to_string(o.block_size) + ", one lane per photon";
  }
  for (std::size_t i = 0; i < rec.size(); ++i) {
    ++r.fates[rec.fate[i]];
    r.photon_steps += static_cast<uint64_t>(rec.n_boundary[i]) + rec.n_scatter[i] +
                      (rec.fate[i] == static_cast<uint8_t>(Fate::AbsorbedBulk) ? 1u : 0u);
  }
  return r;
}

Report run_tally_mode(const SimParams& p, const RunOptions& o, const DriverOptions& d,
                      Tally& tally) {
  Report r;
  r.photons = o.n_photons;
  if (d.kernel == "history") {
    std::vector<uint64_t> w(OPTPHOT_TALLY_WORDS(o.tally));
    optphot_step_counters cnt{};
    detail::check(optphot_run_gpu_tally_counters(
        c_params(p), c_tally_config(o.tally), o.n_photons, w.data(), o.block_size, 0,
        d.count ? &cnt : nullptr, reinterpret_cast<optphot_run_timing*>(&r.timing)));
    detail::unpack(w, o.tally, tally);
    r.kernel = "history";
    r.shape = "block " + std::to_string(o.block_size) + ", grid-stride over photons";
    if (d.count && cnt.v[OPTPHOT_CNT_ISSUED_LANE_STEPS])
      r.simt = static_cast<double>(cnt.v[OPTPHOT_CNT_USEFUL_LANE_STEPS]) /
               static_cast<double>(cnt.v[OPTPHOT_CNT_ISSUED_LANE_STEPS]);
  } else {
    WavefrontTuning tu{};
    tu.steps_per_refill = d.steps_per_refill;
    EventStats st;
    r.timing = run_gpu_tally_wavefront(p, o.tally, o.n_photons, tally, 0, &tu,
                                       d.count ? &st : nullptr);
    r.kernel = "wavefront";
    r.shape = d.steps_per_refill ? std::to_string(d.steps_per_refill) + " steps per refill"
                                 : "library default shape";
    if (d.count && st.issued_lane_steps)
      r.simt = static_cast<double>(st.useful_lane_steps) /
               static_cast<double>(st.issued_lane_steps);
  }
  r.fates = tally.fate_counts;
  r.photon_steps = tally.boundary_sum + tally.scatter_sum +
                   tally.fate_counts[static_cast<int>(Fate::AbsorbedBulk)];
  return r;
}

double per_second(double count, double ms) { return ms > 0.0 ? count / (ms * 1e-3) : 0.0; }

void print_text(const Report& r, const SimParams& p, const RunOptions& o) {
  const double n = static_cast<double>(r.photons);
  std::printf("device    %s\n", gpu_device_name().c_str());
  std::printf("run       %s of %llu photons, seed %llu; %s walls, %s source\n", o.mode.c_str(),
              static_cast<unsigned long long>(r.photons), static_cast<unsigned long long>(p.seed),
              surface_name(p.surface).c_str(), source_name(p.source).c_str());
  std::printf("kernel    %s (%s)\n", r.kernel.c_str(), r.shape.c_str());
  std::printf("fates\n");
  for (int f = 0; f < kNumFates; ++f) {
    const double share = n > 0 ? static_cast<double>(r.fates[f]) / n : 0.0;
    std::printf("  %-17s %14llu   %.6f +- %.6f\n", fate_name(static_cast<Fate>(f)).c_str(),
                static_cast<unsigned long long>(r.fates[f]), share,
                n > 0 ? std::sqrt(share * (1.0 - share) / n) : 0.0);
  }
  std::printf("steps     %llu photon-steps, %.4f per photon\n",
              static_cast<unsigned long long>(r.photon_steps),
              n > 0 ? static_cast<double>(r.photon_steps) / n : 0.0);
  if (r.simt >= 0.0) std::printf("SIMT      %.4f of issued lane-steps useful (wave64)\n", r.simt);
  std::printf("time      kernels %.3f ms; call %.3f ms (device to host %.3f, allocation %.3f)\n",
              r.timing.kernel_ms, r.timing.total_ms, r.timing.d2h_ms, r.timing.alloc_ms);
  std::printf("rate      %.4g photons/s, %.4g photon-steps/s over kernel time\n",
              per_second(n, r.timing.kernel_ms),
              per_second(static_cast<double>(r.photon_steps), r.timing.kernel_ms));
}

void print_json(const Report& r, const RunOptions& o) {
  std::ostringstream s;
  s.precision(10);
  s << "{\"mode\": \"" << o.mode << "\", \"kernel\": \"" << r.kernel << "\", \"photons\": "
    << r.photons << ", \"fates\": [";
  for (int f = 0; f < kNumFates; ++f) s << (f ? ", " : "") << r.fates[f];
  s << "], \"photon_steps\": " << r.photon_steps << ", \"simt_efficiency\": ";
  if (r.simt >= 0.0)
    s << r.simt;
  else
    s << "null";
  s << ", \"kernel_ms\": " << r.timing.kernel_ms << ", \"total_ms\": " << r.timing.total_ms
    << ", \"photons_per_s\": " << per_second(static_cast<double>(r.photons), r.timing.kernel_ms)
    << ", \"photon_steps_per_s\": "
    << per_second(static_cast<double>(r.photon_steps), r.timing.kernel_ms) << "}";
  std::printf("%s\n", s.str().c_str());
}

}  // namespace

int main(int argc, char** argv) {
  try {
    DriverOptions d;
    std::vector<char*> rest = take_driver_options(argc, argv, d);
    SimParams p;
    RunOptions o;
    o.backend = "gpu";
    if (!parse_command_line(static_cast<int>(rest.size()), rest.data(), p, o)) {
      std::cout << usage()
                << "\nDriver options: --kernel history|wavefront, --steps_per_refill K,\n"
                   "  --count 0|1 (SIMT efficiency of tally runs), --report text|json.\n";
      return 0;
    }
    if (o.backend != "gpu") {
      std::cerr << "error: this library runs the GPU backend only (--backend gpu)\n";
      return 2;
    }
    if (!gpu_available()) {
      std::cerr << "error: no HIP device available\n";
      return 2;
    }
    Report r;
    const std::string meta = params_to_json(p, o);
    if (o.mode == "records") {
      PhotonRecords rec;
      r = run_records_mode(p, o, d, rec);
      if (!o.out.empty()) write_records(o.out, rec, meta);
    } else {
      Tally tally;
      r = run_tally_mode(p, o, d, tally);
      if (!o.out.empty()) write_tally(o.out, tally, meta);
    }
    if (d.json)
      print_json(r, o);
    else
      print_text(r, p, o);
    if (!o.out.empty() && !d.json) std::printf("files     %s\n", o.out.c_str());
  } catch (const std::exception& e) {
    std::cerr << "error: " << e.what() << "\n";
    return 1;
  }
  return 0;
}
