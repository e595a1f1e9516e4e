// optphot -- command-line driver of the MI355X library with the reference's
// options and output (apps/optphot_main.cpp): run N photons, print the fate
// summary and timing, optionally write .npy arrays + meta.json.
//
// The backend defaults to gpu here (the reference's CPU backend is not part of
// this library). Two options of this driver select the kernel:
//   --kernel history|wavefront   (default history, as run_gpu_records / _tally)
//   --steps_per_launch K         wavefront steps between refills (default 4)
// Both kernels give bit-identical records and tallies.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <exception>
#include <iostream>
#include <string>
#include <vector>

#include "optphot_hip/backend.hpp"
#include "optphot_hip/config.hpp"

using namespace optphot;

namespace {

void print_summary(const std::array<uint64_t, kNumFates>& counts, uint64_t n, const RunTiming& tm,
                   double mean_boundaries) {
  std::printf("%-18s %12s %10s %10s\n", "fate", "count", "fraction", "stat.err");
  for (int f = 0; f < kNumFates; ++f) {
    const double frac = n ? static_cast<double>(counts[f]) / n : 0.0;
    std::printf("%-18s %12llu %10.6f %10.6f\n", fate_name(static_cast<Fate>(f)).c_str(),
                static_cast<unsigned long long>(counts[f]), frac,
                n ? std::sqrt(frac * (1 - frac) / n) : 0.0);
  }
  std::printf("mean boundary interactions / photon: %.4f\n", mean_boundaries);
  std::printf("time: total %.2f ms, transport %.2f ms", tm.total_ms, tm.kernel_ms);
  if (tm.h2d_ms + tm.d2h_ms + tm.alloc_ms > 0)
    std::printf(" (alloc %.2f ms, H2D %.3f ms, D2H %.2f ms)", tm.alloc_ms, tm.h2d_ms, tm.d2h_ms);
  std::printf("\nthroughput: %.3e photons/s (transport only)\n",
              tm.kernel_ms > 0 ? n / (tm.kernel_ms * 1e-3) : 0.0);
}

}  // namespace

int main(int argc, char** argv) {
  SimParams p;
  RunOptions o;
  o.backend = "gpu";
  std::string kernel = "history";
  uint32_t steps_per_launch = 4;
  try {
    // This driver's own options are taken out before the shared parser runs.
    std::vector<char*> rest{argv[0]};
    for (int i = 1; i < argc; ++i) {
      const std::string a = argv[i];
      const std::size_t eq = a.find('=');
      const std::string key = a.substr(0, eq);
      if (key != "--kernel" && key != "--steps_per_launch") {
        rest.push_back(argv[i]);
        continue;
      }
      std::string v;
      if (eq != std::string::npos)
        v = a.substr(eq + 1);
      else if (i + 1 < argc)
        v = argv[++i];
      else
        throw std::runtime_error("missing value for " + key);
      if (key == "--kernel") {
        if (v != "history" && v != "wavefront")
          throw std::runtime_error("kernel must be history|wavefront");
        kernel = v;
      } else {
        char* end = nullptr;
        const unsigned long k = std::strtoul(v.c_str(), &end, 10);
        if (v.empty() || *end != '\0' || k < 1 || k > 0xffffffffUL)
          throw std::runtime_error("steps_per_launch must be an integer >= 1");
        steps_per_launch = static_cast<uint32_t>(k);
      }
    }
    if (!parse_command_line(static_cast<int>(rest.size()), rest.data(), p, o)) {
      std::cout << usage()
                << "\nThis driver also takes --kernel history|wavefront and --steps_per_launch K.\n";
      return 0;
    }
    if (o.backend != "gpu") {
      std::cerr << "error: this build runs the GPU backend only (--backend gpu)\n";
      return 2;
    }
    if (!gpu_available()) {
      std::cerr << "error: no HIP device available\n";
      return 2;
    }
    std::printf("optphot: %llu photons, backend %s, mode %s, surface %s, source %s\n",
                static_cast<unsigned long long>(o.n_photons), o.backend.c_str(), o.mode.c_str(),
                surface_name(p.surface).c_str(), source_name(p.source).c_str());
    std::printf("GPU: %s, kernel %s\n", gpu_device_name().c_str(), kernel.c_str());

    const std::string meta = params_to_json(p, o);
    const bool wave = kernel == "wavefront";
    if (o.mode == "records") {
      PhotonRecords rec;
      const RunTiming tm = wave ? run_gpu_records_event(p, o.n_photons, rec, steps_per_launch,
                                                        o.block_size)
                                : run_gpu_records(p, o.n_photons, rec, o.block_size);
      std::array<uint64_t, kNumFates> counts{};
      double nb = 0;
      for (std::size_t i = 0; i < rec.size(); ++i) {
        ++counts[rec.fate[i]];
        nb += rec.n_boundary[i];
      }
      print_summary(counts, o.n_photons, tm, o.n_photons ? nb / o.n_photons : 0.0);
      if (!o.out.empty()) write_records(o.out, rec, meta);
    } else {
      Tally tally;
      RunTiming tm;
      if (wave) {
        WavefrontTuning tu{o.block_size, steps_per_launch, 0, 0, 0, 0, 0};
        tm = run_gpu_tally_wavefront(p, o.tally, o.n_photons, tally, 0, &tu);
      } else {
        tm = run_gpu_tally(p, o.tally, o.n_photons, tally, o.block_size);
      }
      print_summary(tally.fate_counts, o.n_photons, tm,
                    o.n_photons ? static_cast<double>(tally.boundary_sum) / o.n_photons : 0.0);
      if (!o.out.empty()) write_tally(o.out, tally, meta);
    }
    if (!o.out.empty()) std::printf("output written to %s\n", o.out.c_str());
  } catch (const std::exception& e) {
    std::cerr << "error: " << e.what() << "\n";
    return 1;
  }
  return 0;
}
