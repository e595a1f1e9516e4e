// compare.hpp -- comparison of two sets of photon records on disk or in
// memory, with the reference's measures (include/optphot/compare.hpp): photon
// by photon agreement for runs with the same seed, and distribution tests
// for independent runs. Implemented in gpu_optical_photons_amd/host/compare.cpp.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "backend.hpp"

namespace optphot {

struct HistoryAgreement {
  uint64_t n = 0;
  uint64_t bitwise_identical = 0;  // every field equal to the last bit
  uint64_t same_discrete = 0;      // same fate, n_boundary, n_scatter (includes the above)
  uint64_t fate_differs = 0;
  double max_abs_dt = 0.0;    // ns, over the same-discrete histories
  double max_abs_dpos = 0.0;  // mm, idem
  int64_t first_divergent = -1;
};

struct TestResult {
  double statistic = 0.0;  // chi-square, or the KS distance D
  int ndf = 0;             // chi-square degrees of freedom
  double pvalue = 1.0;
};

struct DistributionComparison {
  uint64_t n_a = 0, n_b = 0, det_a = 0, det_b = 0;
  double eff_a = 0, eff_b = 0, err_a = 0, err_b = 0;
  double eff_z = 0;  // (eff_a - eff_b) / sqrt(err_a^2 + err_b^2)
  TestResult time_chi2, xy_chi2, time_ks;
};

HistoryAgreement compare_histories(const PhotonRecords& a, const PhotonRecords& b);

// Detected photons' arrival times on [0, t_max) in time_bins bins plus an
// overflow bin (t_max <= 0: the latest detected arrival in either sample), hit
// positions on the readout face ([-hx, hx] x [-hy, hy]) in xy_bins^2 bins.
DistributionComparison compare_distributions(const PhotonRecords& a, const PhotonRecords& b,
                                             float hx, float hy, int time_bins = 100,
                                             float t_max = 0.0f, int xy_bins = 20);

namespace stats {
// Q(a, x) = Gamma(a, x) / Gamma(a), the regularised upper incomplete gamma.
double gamma_q(double a, double x);
double chi2_pvalue(double chi2, int ndf);
// Two-sample chi-square of histograms with different totals; bins empty in
// both are skipped; ndf = used bins - 1.
TestResult chi2_two_sample(const std::vector<double>& a, const std::vector<double>& b);
// P(D > d) for the Kolmogorov distribution at lambda.
double kolmogorov_q(double lambda);
TestResult ks_two_sample(std::vector<double> a, std::vector<double> b);
}  // namespace stats

// Reads <dir>/{fate,n_boundary,n_scatter,t,x,y,z,path}.npy as written by
// write_records (or the reference's optphot --out).
PhotonRecords read_records(const std::string& dir);
// A float field of a meta.json written by params_to_json ("inf" -> infinity).
float read_meta_float(const std::string& meta_path, const std::string& key);

std::string to_json(const HistoryAgreement& h);
std::string to_json(const DistributionComparison& d);

}  // namespace optphot
