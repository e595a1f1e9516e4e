// params.hpp -- C++ run description for liboptphot_hip, with the names,
// enum values and defaults of the reference's include/optphot/params.hpp:15-63
// and TallyConfig of include/optphot/transport.hpp:164-168, so that code
// written against the reference compiles unchanged. Each struct is
// layout-identical to its C counterpart in optphot_hip.h (checked below), and
// the C++ backend passes it to the C-ABI by pointer.
#pragma once

#include <cstddef>
#include <cstdint>
#include <limits>

#include "../optphot_hip.h"

namespace optphot {

constexpr float kSpeedOfLight = 299.792458f;  // mm / ns
constexpr float kInfinity = std::numeric_limits<float>::infinity();
inline bool is_inf(float x) { return x == kInfinity; }

enum class Surface : uint32_t { Polished = 0, Black = 1, Specular = 2, Lambertian = 3 };
enum class SourceType : uint32_t { Point = 0, UniformVolume = 1 };
enum class Fate : uint8_t {
  Detected = 0,
  AbsorbedBulk = 1,
  AbsorbedSurface = 2,
  Escaped = 3,
  MaxSteps = 4,
};
constexpr int kNumFates = 5;
constexpr int kReadoutFace = 5;  // faces 0:-x 1:+x 2:-y 3:+y 4:-z 5:+z

struct SimParams {
  float hx = 25.0f, hy = 25.0f, hz = 25.0f;  // half sizes, mm
  float n_scint = 1.63f;
  float n_det = 1.63f;
  float n_out = 1.0f;
  float abs_length = 1000.0f;     // mm, kInfinity disables absorption
  float scat_length = kInfinity;  // mm, kInfinity disables Rayleigh scattering
  Surface surface = Surface::Polished;
  float reflectivity = 0.95f;
  SourceType source = SourceType::Point;
  float src_x = 0.0f, src_y = 0.0f, src_z = 0.0f;
  float tau = 4.0f;  // ns, 0 = prompt emission
  uint32_t max_steps = 10000;
  uint64_t seed = 12345;
};

struct TallyConfig {
  uint32_t time_bins = 100;  // arrival-time histogram on [0, time_max) + overflow bin
  float time_max = 50.0f;    // ns
  uint32_t xy_bins = 50;     // xy_bins x xy_bins hit map on the readout face
};

static_assert(sizeof(SimParams) == 72 && sizeof(SimParams) == sizeof(optphot_sim_params),
              "SimParams must match the C-ABI struct");
static_assert(offsetof(SimParams, surface) == offsetof(optphot_sim_params, surface) &&
                  offsetof(SimParams, max_steps) == offsetof(optphot_sim_params, max_steps) &&
                  offsetof(SimParams, seed) == offsetof(optphot_sim_params, seed),
              "SimParams field offsets must match the C-ABI struct");
static_assert(sizeof(TallyConfig) == sizeof(optphot_tally_config) &&
                  offsetof(TallyConfig, xy_bins) == offsetof(optphot_tally_config, xy_bins),
              "TallyConfig must match the C-ABI struct");

inline const optphot_sim_params* c_params(const SimParams& p) {
  return reinterpret_cast<const optphot_sim_params*>(&p);
}
inline const optphot_tally_config* c_tally_config(const TallyConfig& c) {
  return reinterpret_cast<const optphot_tally_config*>(&c);
}

}  // namespace optphot
