// optphot_compare -- compares two runs' photon records on disk (the .npy +
// meta.json layout of optphot --out, this library's or the reference's):
// photon-by-photon agreement (meaningful for the same seed and photon ids,
// e.g. the reference's CPU backend against this GPU backend) and distribution
// tests (meaningful for independent seeds), as the reference's
// apps/optphot_compare.cpp reports them for its in-process runs.
//
//   optphot_compare DIR_A DIR_B [--out FILE] [--time_bins N] [--t_max NS] [--xy_bins N]
//
// The readout face's size is read from DIR_A/meta.json (size_x, size_y).
// Exit status 0 on success (the agreement itself is reported, not judged).
#include <cstdio>
#include <cstdlib>
#include <exception>
#include <iostream>
#include <string>

#include "optphot_hip/compare.hpp"
#include "optphot_hip/config.hpp"

using namespace optphot;

int main(int argc, char** argv) {
  try {
    std::string a_dir, b_dir, out;
    int time_bins = 100, xy_bins = 20;
    float t_max = 0.0f;
    for (int i = 1; i < argc; ++i) {
      const std::string arg = argv[i];
      auto value = [&]() -> std::string {
        if (i + 1 >= argc) throw std::runtime_error("missing value for " + arg);
        return argv[++i];
      };
      if (arg == "-h" || arg == "--help") {
        std::cout << "usage: optphot_compare DIR_A DIR_B [--out FILE] [--time_bins N] "
                     "[--t_max NS] [--xy_bins N]\n";
        return 0;
      } else if (arg == "--out") {
        out = value();
      } else if (arg == "--time_bins") {
        time_bins = std::stoi(value());
      } else if (arg == "--t_max") {
        t_max = std::stof(value());
      } else if (arg == "--xy_bins") {
        xy_bins = std::stoi(value());
      } else if (arg.rfind("--", 0) == 0) {
        throw std::runtime_error("unknown option '" + arg + "'");
      } else if (a_dir.empty()) {
        a_dir = arg;
      } else if (b_dir.empty()) {
        b_dir = arg;
      } else {
        throw std::runtime_error("unexpected argument '" + arg + "'");
      }
    }
    if (a_dir.empty() || b_dir.empty()) throw std::runtime_error("two record directories needed");
    if (time_bins < 1 || xy_bins < 1) throw std::runtime_error("bins must be >= 1");
    const PhotonRecords a = read_records(a_dir), b = read_records(b_dir);
    const float hx = 0.5f * read_meta_float(a_dir + "/meta.json", "size_x");
    const float hy = 0.5f * read_meta_float(a_dir + "/meta.json", "size_y");
    std::printf("optphot_compare: %s (%zu photons) vs %s (%zu photons)\n", a_dir.c_str(), a.size(),
                b_dir.c_str(), b.size());
    std::string json = "{\n";
    if (a.size() == b.size()) {
      const HistoryAgreement h = compare_histories(a, b);
      const double n = h.n ? static_cast<double>(h.n) : 1.0;
      std::printf("photon by photon:\n");
      std::printf("  bitwise identical records : %llu / %llu (%.6f)\n",
                  static_cast<unsigned long long>(h.bitwise_identical),
                  static_cast<unsigned long long>(h.n), h.bitwise_identical / n);
      std::printf("  same discrete history     : %llu / %llu (%.6f)\n",
                  static_cast<unsigned long long>(h.same_discrete),
                  static_cast<unsigned long long>(h.n), h.same_discrete / n);
      std::printf("  different fate            : %llu\n",
                  static_cast<unsigned long long>(h.fate_differs));
      std::printf("  max |dt| = %.3g ns, max |dpos| = %.3g mm (same discrete history)\n",
                  h.max_abs_dt, h.max_abs_dpos);
      json += "\"history\": " + to_json(h) + ",\n";
    } else {
      std::printf("photon by photon: skipped (different photon counts)\n");
    }
    const DistributionComparison d =
        compare_distributions(a, b, hx, hy, time_bins, t_max, xy_bins);
    std::printf("distributions:\n");
    std::printf("  efficiency A %.6f +- %.6f, B %.6f +- %.6f, z = %.2f\n", d.eff_a, d.err_a,
                d.eff_b, d.err_b, d.eff_z);
    std::printf("  arrival time: chi2/ndf = %.1f/%d (p = %.3f), KS D = %.2e (p = %.3f)\n",
                d.time_chi2.statistic, d.time_chi2.ndf, d.time_chi2.pvalue, d.time_ks.statistic,
                d.time_ks.pvalue);
    std::printf("  hit position: chi2/ndf = %.1f/%d (p = %.3f)\n", d.xy_chi2.statistic,
                d.xy_chi2.ndf, d.xy_chi2.pvalue);
    json += "\"distributions\": " + to_json(d) + "\n}";
    if (!out.empty()) {
      write_text_file(out, json);
      std::printf("written to %s\n", out.c_str());
    }
  } catch (const std::exception& e) {
    std::cerr << "error: " << e.what() << "\n";
    return 1;
  }
  return 0;
}
