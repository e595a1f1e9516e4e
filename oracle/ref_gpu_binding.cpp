// ref_gpu_binding.cpp -- the reference-side binding of liboptphot_hip: the
// reference's declared GPU backend functions (include/optphot/backend.hpp:
// 84-115) implemented over this library's C-ABI with the reference's own
// types. Dropped into the reference in place of src/gpu/*.cu, it lets the
// reference's unmodified apps run on the MI355X library. Here it is compiled
// only by `make ref-apps` (oracle/Makefile) against the reference's sources
// where they lie, to check that drop-in on the GPU; it is never part of the
// product library.
#include <stdexcept>
#include "optphot/backend.hpp"
#include "optphot_hip.h"

namespace optphot {
namespace {
void check(int rc) {
  if (rc == OPTPHOT_OK) return;
  if (rc == OPTPHOT_ERR_INVALID_ARGUMENT) throw std::invalid_argument(optphot_last_error());
  throw std::runtime_error(optphot_last_error());
}
optphot_records view(PhotonRecords& r) {
  return {r.fate.data(), r.n_boundary.data(), r.n_scatter.data(), r.t.data(),
          r.x.data(), r.y.data(), r.z.data(), r.path.data()};
}
const optphot_sim_params* cp(const SimParams& p) {
  static_assert(sizeof(SimParams) == sizeof(optphot_sim_params));
  return reinterpret_cast<const optphot_sim_params*>(&p);
}
}  // namespace

bool gpu_available() { return optphot_gpu_available() != 0; }
std::string gpu_device_name() {
  char b[256];
  check(optphot_gpu_device_name(b, sizeof b));
  return b;
}
RunTiming run_gpu_records(const SimParams& p, uint64_t n, PhotonRecords& out, int bs,
                          uint64_t first_id, uint64_t chunk) {
  out.resize(n);
  const optphot_records v = view(out);
  RunTiming t;
  check(optphot_run_gpu_records(cp(p), n, &v, bs, first_id, chunk,
                                reinterpret_cast<optphot_run_timing*>(&t)));
  return t;
}
RunTiming run_gpu_tally(const SimParams& p, const TallyConfig& c, uint64_t n, Tally& out,
                        int bs, uint64_t first_id) {
  const optphot_tally_config cc{c.time_bins, c.time_max, c.xy_bins};
  std::vector<uint64_t> w(OPTPHOT_TALLY_WORDS(cc));
  RunTiming t;
  check(optphot_run_gpu_tally(cp(p), &cc, n, w.data(), bs, first_id,
                              reinterpret_cast<optphot_run_timing*>(&t)));
  out.reset(c);
  size_t k = 0;
  for (auto& f : out.fate_counts) f = w[k++];
  for (auto& h : out.time_hist) h = w[k++];
  for (auto& h : out.xy_hist) h = w[k++];
  out.boundary_sum = w[k++];
  out.scatter_sum = w[k++];
  return t;
}
RunTiming run_gpu_records_event(const SimParams& p, uint64_t n, PhotonRecords& out,
                                uint32_t k, int bs, uint64_t first_id, uint64_t chunk,
                                EventStats* stats) {
  out.resize(n);
  const optphot_records v = view(out);
  RunTiming t;
  check(optphot_run_gpu_records_event(cp(p), n, &v, k, bs, first_id, chunk,
                                      reinterpret_cast<optphot_event_stats*>(stats),
                                      reinterpret_cast<optphot_run_timing*>(&t)));
  return t;
}
}  // namespace optphot
