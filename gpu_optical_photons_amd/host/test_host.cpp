// Unit tests of the C++ host layer: configuration parsing and validation,
// config files with command-line overrides, the .npy writer (after the
// reference's tests/test_config.cpp), the backend.hpp value types and error
// mapping, and the slot table of batched launches. With the argument "gpu" it
// also runs the GPU entry points through backend.hpp and checks the history
// and wavefront kernels against each other.
// Exit status 0 when every check passes.
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <iterator>
#include <limits>
#include <vector>
#include <stdexcept>
#include <string>

#include "optphot_hip/backend.hpp"
#include "optphot_hip/config.hpp"
#include "../csrc/batch_queue.hpp"

using namespace optphot;

namespace {

int g_checks = 0, g_failed = 0;

void check(bool ok, const char* what, int line) {
  ++g_checks;
  if (!ok) {
    ++g_failed;
    std::printf("FAILED line %d: %s\n", line, what);
  }
}
#define CHECK(x) check((x), #x, __LINE__)

template <class E, class F>
bool throws(F&& f) {
  try {
    f();
  } catch (const E&) {
    return true;
  } catch (...) {
    return false;
  }
  return false;
}
#define CHECK_THROWS_AS(expr, E) check(throws<E>([&] { expr; }), #expr " throws " #E, __LINE__)

std::string slurp(const std::string& path) {
  std::ifstream f(path, std::ios::binary);
  return std::string((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
}

void test_settings() {
  SimParams p;
  RunOptions o;
  apply_setting(p, o, "size_x", "30");
  apply_setting(p, o, "abs_length", "inf");
  apply_setting(p, o, "surface", "lambertian");
  apply_setting(p, o, "n_photons", "1e7");
  apply_setting(p, o, "seed", "18446744073709551615");
  apply_setting(p, o, "source", "volume");
  apply_setting(p, o, "time_max", "100");
  CHECK(p.hx == 15.0f);
  CHECK(is_inf(p.abs_length));
  CHECK(p.surface == Surface::Lambertian);
  CHECK(p.source == SourceType::UniformVolume);
  CHECK(o.n_photons == 10000000u);
  CHECK(p.seed == 18446744073709551615ull);
  CHECK(o.tally.time_max == 100.0f);
  CHECK_THROWS_AS(apply_setting(p, o, "nonsense", "1"), std::runtime_error);
  CHECK_THROWS_AS(apply_setting(p, o, "n_scint", "1.6x"), std::runtime_error);
  CHECK_THROWS_AS(apply_setting(p, o, "n_scint", ""), std::runtime_error);
  CHECK_THROWS_AS(apply_setting(p, o, "n_scint", "1e60"), std::runtime_error);
  CHECK_THROWS_AS(apply_setting(p, o, "surface", "shiny"), std::runtime_error);
  CHECK_THROWS_AS(apply_setting(p, o, "source", "line"), std::runtime_error);
  CHECK_THROWS_AS(apply_setting(p, o, "n_photons", "-5"), std::runtime_error);
  CHECK_THROWS_AS(apply_setting(p, o, "seed", "18446744073709551616"), std::runtime_error);
  CHECK_THROWS_AS(apply_setting(p, o, "mode", "histogram"), std::runtime_error);
  CHECK_THROWS_AS(apply_setting(p, o, "backend", "tpu"), std::runtime_error);
  // defaults of an untouched description (params.hpp:42-62, transport.hpp:165-167)
  const SimParams d;
  CHECK(d.hx == 25.0f && d.n_scint == 1.63f && d.n_det == 1.63f && d.n_out == 1.0f);
  CHECK(d.abs_length == 1000.0f && is_inf(d.scat_length) && d.reflectivity == 0.95f);
  CHECK(d.tau == 4.0f && d.max_steps == 10000u && d.seed == 12345u);
  const TallyConfig tc;
  CHECK(tc.time_bins == 100u && tc.time_max == 50.0f && tc.xy_bins == 50u);
}

void test_files_and_overrides() {
  const std::string path = "test_host_tmp.cfg";
  {
    std::ofstream f(path);
    f << "# comment\nsurface = black   # trailing comment\n\n n_det=1.5\nn_photons = 1000\n";
  }
  SimParams p;
  RunOptions o;
  const char* argv[] = {"optphot", "--config", path.c_str(), "--n_det", "1.4", "--tau=0"};
  CHECK(parse_command_line(6, const_cast<char**>(argv), p, o));
  CHECK(p.surface == Surface::Black);
  CHECK(p.n_det == 1.4f);
  CHECK(p.tau == 0.0f);
  CHECK(o.n_photons == 1000u);
  const char* help[] = {"optphot", "--help"};
  CHECK(!parse_command_line(2, const_cast<char**>(help), p, o));
  const char* stray[] = {"optphot", "n_det", "1.4"};
  CHECK_THROWS_AS(parse_command_line(3, const_cast<char**>(stray), p, o), std::runtime_error);
  const char* dangling[] = {"optphot", "--n_det"};
  CHECK_THROWS_AS(parse_command_line(2, const_cast<char**>(dangling), p, o), std::runtime_error);
  {
    std::ofstream f(path);
    f << "n_det 1.5\n";
  }
  CHECK_THROWS_AS(load_config_file(path, p, o), std::runtime_error);
  std::filesystem::remove(path);
  CHECK_THROWS_AS(load_config_file("does_not_exist.cfg", p, o), std::runtime_error);
}

void test_npy() {
  const std::string path = "test_host_tmp.npy";
  const float data[3] = {1.0f, 2.0f, 3.0f};
  write_npy(path, data, "<f4", {3}, 4);
  const std::string b = slurp(path);
  std::filesystem::remove(path);
  CHECK(b.size() >= 10);
  CHECK(b.substr(1, 5) == "NUMPY");
  CHECK(static_cast<unsigned char>(b[0]) == 0x93 && b[6] == 1 && b[7] == 0);
  const std::size_t hlen =
      static_cast<unsigned char>(b[8]) | (static_cast<unsigned char>(b[9]) << 8);
  CHECK((10 + hlen) % 64 == 0);
  CHECK(b.size() == 10 + hlen + sizeof(data));
  CHECK(b.find("'shape': (3,)") != std::string::npos);
  CHECK(b[10 + hlen - 1] == '\n');
  CHECK(std::memcmp(b.data() + 10 + hlen, data, sizeof(data)) == 0);

  const uint64_t grid[6] = {1, 2, 3, 4, 5, 6};
  write_npy(path, grid, "<u8", {2, 3}, 8);
  const std::string g = slurp(path);
  std::filesystem::remove(path);
  CHECK(g.find("'shape': (2, 3)") != std::string::npos);
  CHECK(g.find("'descr': '<u8'") != std::string::npos);
}

void test_meta_json() {
  SimParams p;
  RunOptions o;
  const std::string j = params_to_json(p, o);
  CHECK(j.find("\"size_x\": 50, \"size_y\": 50, \"size_z\": 50,") != std::string::npos);
  CHECK(j.find("\"scat_length\": \"inf\"") != std::string::npos);
  CHECK(j.find("\"n_scint\": 1.63, \"n_det\": 1.63, \"n_out\": 1,") != std::string::npos);
  CHECK(j.find("\"surface\": \"polished\"") != std::string::npos);
  CHECK(j.find("\"seed\": 12345,") != std::string::npos);
  CHECK(j.front() == '{' && j.back() == '}');
}

void test_value_types() {
  PhotonRecords r;
  r.resize(7);
  CHECK(r.size() == 7 && r.path.size() == 7 && r.n_scatter.size() == 7);
  CHECK(PhotonRecords::kBytesPerPhoton == 29);
  TallyConfig c;
  c.time_bins = 4;
  c.xy_bins = 3;
  Tally a, b;
  a.reset(c);
  b.reset(c);
  CHECK(a.time_hist.size() == 5 && a.xy_hist.size() == 9);
  a.fate_counts[0] = 3;
  b.fate_counts[2] = 4;
  b.time_hist[4] = 2;
  b.boundary_sum = 9;
  a.add(b);
  CHECK(a.total() == 7 && a.time_hist[4] == 2 && a.boundary_sum == 9);
  EventStats s;
  CHECK(s.simt_efficiency() == 0.0);
  s.useful_lane_steps = 3;
  s.issued_lane_steps = 4;
  CHECK(s.simt_efficiency() == 0.75);
}

void test_errors() {
  // Argument checks come before any device work, so they hold without a GPU.
  SimParams p;
  PhotonRecords r;
  CHECK_THROWS_AS(run_gpu_records(p, 10, r, 100), std::invalid_argument);
  CHECK_THROWS_AS(run_gpu_records_event(p, 10, r, 0), std::invalid_argument);
  CHECK_THROWS_AS(run_gpu_records_event(p, 10, r, 4, 2048), std::invalid_argument);
  SimParams bad = p;
  bad.surface = static_cast<Surface>(7);
  CHECK_THROWS_AS(run_gpu_records(bad, 10, r), std::invalid_argument);
  if (!gpu_available()) CHECK_THROWS_AS(run_gpu_records(p, 10, r), std::runtime_error);
}

void test_gpu() {
  CHECK(gpu_available());
  CHECK(!gpu_device_name().empty());
  SimParams p;
  p.hx = p.hy = 50.0f;
  p.scat_length = 300.0f;
  p.surface = Surface::Lambertian;
  p.reflectivity = 0.97f;
  p.source = SourceType::UniformVolume;
  p.n_det = 1.5f;
  p.seed = 1;
  const uint64_t n = 100000;
  PhotonRecords h, w;
  run_gpu_records(p, n, h, 256, 7, 30000);
  EventStats st;
  run_gpu_records_event(p, n, w, 3, 128, 7, 1 << 24, &st);
  CHECK(h.size() == n && w.size() == n);
  CHECK(h.fate == w.fate && h.n_boundary == w.n_boundary && h.n_scatter == w.n_scatter);
  CHECK(std::memcmp(h.t.data(), w.t.data(), n * 4) == 0);
  CHECK(std::memcmp(h.path.data(), w.path.data(), n * 4) == 0);
  CHECK(st.launches == 1 && st.useful_lane_steps > 0 &&
        st.issued_lane_steps >= st.useful_lane_steps);
  TallyConfig c{256, 100.0f, 64};
  Tally th, tw;
  run_gpu_tally(p, c, n, th, 256, 7);
  run_gpu_tally_wavefront(p, c, n, tw, 7);
  CHECK(th.fate_counts == tw.fate_counts && th.time_hist == tw.time_hist &&
        th.xy_hist == tw.xy_hist && th.boundary_sum == tw.boundary_sum &&
        th.scatter_sum == tw.scatter_sum);
  CHECK(th.total() == n);
  // the per-interaction path gives the same tally
  Tally ts;
  SortedStats ss{};
  run_gpu_tally_sorted(p, c, n, ts, 7, nullptr, &ss);
  CHECK(th.fate_counts == ts.fate_counts && th.time_hist == ts.time_hist &&
        th.xy_hist == ts.xy_hist && th.boundary_sum == ts.boundary_sum &&
        th.scatter_sum == ts.scatter_sum);
  CHECK(ss.emitted == n && ss.iterations >= 2);
  uint64_t det = 0, nb = 0;
  for (std::size_t i = 0; i < n; ++i) {
    det += h.fate[i] == 0;
    nb += h.n_boundary[i];
  }
  CHECK(th.fate_counts[0] == det && th.boundary_sum == nb);
  // batch of one event equals the single-event tally on its words
  EventTally et;
  run_gpu_batch({p}, {n}, 256, 100.0f, et, 7);
  bool same = true;
  for (int f = 0; f < kNumFates; ++f) same &= et.event(0)[f] == th.fate_counts[f];
  for (uint32_t k = 0; k <= 256; ++k) same &= et.event(0)[kNumFates + k] == th.time_hist[k];
  CHECK(same);
}

// The batch queue's slot table: every photon of every event in exactly one
// slot, the events expected to have long histories first (and only when such
// a history is long against the launch), and chunk j of the events of a part
// before chunk j + 1.
void test_batch_slots() {
  optphot_sim_params base{};
  base.hx = base.hy = base.hz = 25.0f;
  base.n_scint = 1.63f;
  base.n_det = 1.5f;
  base.n_out = 1.0f;
  base.abs_length = 1000.0f;
  base.scat_length = std::numeric_limits<float>::infinity();
  base.reflectivity = 0.9f;
  base.max_steps = 10000;
  base.seed = 1;
  std::vector<optphot_sim_params> ev;
  std::vector<uint64_t> ph;
  uint64_t state = 12345;
  auto next = [&] { state = state * 6364136223846793005ull + 1442695040888963407ull;
                    return state >> 33; };
  for (int e = 0; e < 300; ++e) {
    optphot_sim_params q = base;
    q.surface = static_cast<uint32_t>(next() % 4);
    q.hx = 1.0f + static_cast<float>(next() % 50);
    q.abs_length = (next() % 5 == 0) ? std::numeric_limits<float>::infinity()
                                     : 100.0f + static_cast<float>(next() % 3000);
    ev.push_back(q);
    ph.push_back(next() % 7 == 0 ? 0 : next() % 5000);
  }
  // one heavy event: polished, tiny box, no absorption
  ev[17].surface = OPTPHOT_SURFACE_POLISHED;
  ev[17].hx = ev[17].hy = ev[17].hz = 1.0f;
  ev[17].abs_length = std::numeric_limits<float>::infinity();
  ph[17] = 3000;
  for (const uint32_t chunk : {64u, 512u, 1024u}) {
    for (const uint64_t lanes : {uint64_t{64}, uint64_t{4096}, uint64_t{393216}}) {
      const auto slots = ophip::batch_slots(ev.data(), ph.data(), 300, chunk, lanes);
      std::vector<std::vector<int>> seen(300);
      for (int e = 0; e < 300; ++e) seen[e].assign(ph[e], 0);
      bool ok = true;
      for (const auto sl : slots) {
        const uint32_t e = static_cast<uint32_t>(sl >> 32);
        const uint64_t start = static_cast<uint32_t>(sl) * uint64_t{chunk};
        ok &= e < 300 && start < ph[e];
        if (!ok) break;
        for (uint64_t p = start; p < std::min<uint64_t>(start + chunk, ph[e]); ++p) ++seen[e][p];
      }
      for (int e = 0; ok && e < 300; ++e)
        for (const int c : seen[e]) ok &= c == 1;
      CHECK(ok);
      // with few lanes every history is short against the launch: the plain
      // layout, chunk 0 of all non-empty events first; with many lanes all of
      // the heavy event comes before any black-walled event
      if (lanes == 393216) {
        size_t last17 = 0, first_black = slots.size();
        for (size_t i = 0; i < slots.size(); ++i) {
          const uint32_t e = static_cast<uint32_t>(slots[i] >> 32);
          if (e == 17u) last17 = i;
          if (ev[e].surface == OPTPHOT_SURFACE_BLACK) first_black = std::min(first_black, i);
        }
        CHECK(last17 < first_black);
      }
      if (lanes == 64) {
        size_t nonempty = 0;
        for (int e = 0; e < 300; ++e) nonempty += ph[e] > 0;
        bool first_round = true;
        for (size_t i = 0; i < nonempty; ++i) first_round &= static_cast<uint32_t>(slots[i]) == 0;
        CHECK(first_round);
      }
    }
  }
  const auto h = ophip::history_estimate(ev[17], ph[17]);
  CHECK(h.longest == 10000.0);  // capped at max_steps
  optphot_sim_params black = base;
  black.surface = OPTPHOT_SURFACE_BLACK;
  CHECK(ophip::history_estimate(black, 1000).mean == 1.0);
  CHECK(ophip::batch_slots(ev.data(), std::vector<uint64_t>(300, 0).data(), 300, 64, 64).empty());
}

}  // namespace

int main(int argc, char** argv) {
  const bool gpu = argc > 1 && std::string(argv[1]) == "gpu";
  try {
    test_settings();
    test_files_and_overrides();
    test_npy();
    test_meta_json();
    test_value_types();
    test_errors();
    test_batch_slots();
    if (gpu) test_gpu();
  } catch (const std::exception& e) {
    std::printf("FAILED: unexpected exception: %s\n", e.what());
    return 1;
  }
  std::printf("%d checks, %d failed\n", g_checks, g_failed);
  return g_failed ? 1 : 0;
}
