// batch_queue.hpp -- the slot table of a batched-events launch (host code).
//
// A batch (many events, one persistent launch) hands out its photons through
// a device queue of slots; each slot is a run of photons of one event. This
// header builds that table on the host. The order of the slots changes only
// the speed of the launch: every event's photons keep their ids, so every
// event's tally row is the same whatever the order.
#pragma once

#include <stdint.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "../../include/optphot_hip.h"

namespace ophip {

// Expected photon history length of an event, in steps up to a constant
// factor: the mean (`mean`) and the longest of its `n` photons (`longest`).
// They only order the batch queue. A photon survives a wall with probability
// w (the reflectivity of specular and Lambertian walls; 1 for polished walls,
// where total internal reflection keeps the photons that make long
// histories; 0 for black walls) and the bulk from wall to wall with
// probability exp(-m / L), m = 4 V / S the mean chord of the box, L the
// absorption length (and, for polished walls, the scattering length, since a
// scatter lets a trapped photon out). A history then has about 1 / (1 - s)
// steps, s = w exp(-m / L), and the longest of n about ln(n) times that,
// capped at max_steps.
struct HistoryEstimate {
  double mean, longest;
};
inline HistoryEstimate history_estimate(const optphot_sim_params& p, uint64_t n) {
  const double hx = p.hx, hy = p.hy, hz = p.hz;
  const double area = hx * hy + hy * hz + hz * hx;
  const double chord = area > 0.0 ? 4.0 * hx * hy * hz / area : 0.0;
  double att = 0.0;
  if (std::isfinite(p.abs_length) && p.abs_length > 0.0f) att += chord / p.abs_length;
  if (p.surface == OPTPHOT_SURFACE_POLISHED && std::isfinite(p.scat_length) && p.scat_length > 0.0f)
    att += chord / p.scat_length;
  const double wall = p.surface == OPTPHOT_SURFACE_BLACK      ? 0.0
                      : p.surface == OPTPHOT_SURFACE_POLISHED ? 1.0
                                               : std::min(1.0, std::max(0.0, double(p.reflectivity)));
  const double s = std::min(wall * std::exp(-att), 1.0 - 1e-9);
  const double cap = static_cast<double>(p.max_steps);
  HistoryEstimate h{std::min(1.0 / (1.0 - s), cap),
                    std::min(std::log(static_cast<double>(n) + 1.0) / (1.0 - s), cap)};
  if (!std::isfinite(h.mean)) h.mean = 0.0;
  if (!std::isfinite(h.longest)) h.longest = 0.0;
  return h;
}

// The batch queue: one slot (event << 32 | chunk index) per chunk of `chunk`
// photons of every event, for a launch of `lanes` lanes. Chunk j of every
// event precedes chunk j + 1, so the waves in flight work on many events at
// once and the per-event tally atomics spread over many rows; the events are
// grouped by branch structure (surface model, scattering on, absorption on),
// so the consecutive slots a wave claims mostly share it.
// A persistent launch lasts at least as long as its longest history. When an
// event's longest history is expected to reach a fifth of the steps each lane
// makes on average, that history would stretch the end of the launch if it
// started late: such events then form a first part of the queue, laid out
// the same way with the longest first in each group, so that their long
// histories run alongside the bulk.
inline std::vector<unsigned long long> batch_slots(const optphot_sim_params* params,
                                            const uint64_t* photons, uint32_t n,
                                            uint32_t chunk, uint64_t lanes) {
  std::vector<HistoryEstimate> est(n);
  double steps = 0.0;
  for (uint32_t e = 0; e < n; ++e) {
    est[e] = history_estimate(params[e], photons[e]);
    steps += est[e].mean * static_cast<double>(photons[e]);
  }
  const double per_lane = steps / static_cast<double>(std::max<uint64_t>(1, lanes));
  auto long_first = [&](uint32_t e) { return est[e].longest >= 0.2 * per_lane; };
  auto group = [&](uint32_t e) {
    return params[e].surface * 4u + (!std::isinf(params[e].scat_length) ? 2u : 0u) +
           (!std::isinf(params[e].abs_length) ? 1u : 0u);
  };
  std::vector<uint32_t> order(n);
  for (uint32_t e = 0; e < n; ++e) order[e] = e;
  std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
    if (long_first(a) != long_first(b)) return long_first(a);
    if (group(a) != group(b)) return group(a) < group(b);
    // the first part longest first
    return long_first(a) && est[a].longest > est[b].longest;
  });
  uint64_t total = 0;
  for (uint32_t e = 0; e < n; ++e) total += (photons[e] + chunk - 1) / chunk;
  std::vector<unsigned long long> slots;
  slots.reserve(total);
  // round j of a part: chunk j of every event of the part that has one
  // (events drop out of `live` once their chunks are used up)
  auto emit_part = [&](size_t lo, size_t hi) {
    std::vector<uint32_t> live(order.begin() + lo, order.begin() + hi);
    for (uint64_t j = 0; !live.empty(); ++j) {
      size_t keep = 0;
      for (const uint32_t e : live) {
        const uint64_t chunks = (photons[e] + chunk - 1) / chunk;
        if (j < chunks) {
          slots.push_back((static_cast<unsigned long long>(e) << 32) | j);
          if (j + 1 < chunks) live[keep++] = e;
        }
      }
      live.resize(keep);
    }
  };
  size_t split = 0;
  while (split < n && long_first(order[split])) ++split;
  emit_part(0, split);
  emit_part(split, n);
  return slots;
}

}  // namespace ophip
