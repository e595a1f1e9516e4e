// kernels.hpp -- launch interface between the host runtime (capi.hip) and the
// transport kernels (transport_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "photon.hpp"

namespace ophip {

// Device pointers of the structure-of-arrays photon records (29 B/photon).
struct DevRecords {
  uint8_t* fate;
  uint32_t* nb;
  uint32_t* ns;
  float *t, *x, *y, *z, *path;
};

// Tally binning constants derived once on the host (transport.hpp:170-181).
struct DevTally {
  uint32_t time_bins, xy_bins;
  float time_scale;      // time_bins / time_max
  float two_hx, two_hy;  // 2 hx, 2 hy
  uint32_t words;        // u32 words of one block-private histogram
                         // = 5 + time_bins + 1 + xy_bins^2
};

// Branch-event counter slots (same numbering as optphot_hip.h's OPTPHOT_CNT_*).
enum : int {
  kCntUseful = 0,
  kCntIssued = 1,
  kCntAbsorb = 2,
  kCntScatter = 3,
  kCntReadout = 4,
  kCntPolished = 5,
  kCntBlack = 6,
  kCntReflector = 7,
  kCntLambert = 8,
  kCntEmitted = 9,
  kCntDetected = 10,
  kCntAbsOn = 11,   // photon-steps that sampled an absorption distance
  kCntScatOn = 12,  // photon-steps that sampled a scattering distance
  kCntSlots = 16
};

// Per-event tallies of a batch: for event e, u32 words
// [e * words .. e * words + 5) fate counts and
// [e * words + 5 .. e * words + 5 + time_bins] arrival-time histogram with an
// overflow bin (the reference's time_bin(), transport.hpp:170-173).
struct DevEventTally {
  uint32_t* counts;                 // [n_events][words]
  uint32_t words;                   // 5 + time_bins + 1
  uint32_t time_bins;
  float time_scale;                 // time_bins / time_max
  const DevParams* table;           // [n_events] per-event parameters (batched kernel)
  const unsigned long long* offsets;  // [n_events + 1] prefix sums of photons per event
  const unsigned long long* slots;  // [n_slots] batch queue: slot c is chunk
                                    // (slots[c] & 0xffffffff) of event slots[c] >> 32
  uint32_t n_events;
  uint32_t event;                   // the single event of a plain per-event launch
  uint32_t wave_rows;               // batch launches with rows in global memory: each
                                    // wave keeps its current event's row in LDS
};

// Output selection of a transport launch.
struct LaunchOut {
  DevRecords rec;              // used when `records`
  DevTally tally;              // used when `tally`
  uint32_t* partials;          // [grid][tally.words] block histograms
  unsigned long long* sums;    // [grid][2] block boundary / scatter sums
  unsigned long long* counters;  // [kCntSlots], accumulated
  DevEventTally evt;           // used when `events`
  bool records, tally_on, events;
  // History, register-stream and ring wavefront launches of one parameter
  // set: keep the Fresnel cell bounds in LDS (set when they fit).
  bool fresnel_bounds;
};

// Genstep launch: photons come from gensteps (device array), handed out as
// queue slots; slot c is chunk `slots[c].y` (photons slots[c].y * chunk ..)
// of genstep slots[c].x.
struct GenstepArgs {
  const DevGenstep* gensteps;
  const uint2* slots;
  uint64_t n_slots;
};

// Adjoint launch: parameter sets (events) e = 0..n_events-1 with counts[e]
// photons at ids first_id + offsets[e] ..; per-set weights w[e][k] over the
// time bins k = 0..time_bins (overflow included); output
// grad[e][j] += sum over detected photons of the set of
//   w[e][bin(t)] * S_j + (j == n_scint) * time term,
// j = 0..4 for abs_length, scat_length, reflectivity, n_det, n_scint.
struct AdjointArgs {
  const DevParams* table;
  const unsigned long long* offsets;  // [n_events + 1]
  unsigned long long* qheads;         // [n_events] per-set queue heads, zeroed
  unsigned long long* closed;         // [1] number of sets whose queue ran out, zeroed
  uint32_t n_events;
  uint32_t time_bins;
  float time_scale;                   // time_bins / time_max
  const double* weights;              // [n_events][time_bins + 1]
  double* grad;                       // [n_events][5], zeroed
};

struct WavefrontShape {
  int block, grid;
  uint32_t steps_per_refill, queue_chunk;
  // Tally launches: > 0 selects the kernel whose lanes keep an emitted
  // successor photon, emitting when this many lanes of a wave lack one.
  uint32_t successor_threshold = 0;
  // Tally launches: keep each lane's Philox window in an LDS ring (takes
  // precedence over successor_threshold).
  uint32_t stream_in_lds = 0;
  // Batch / genstep launches: queue slots a wave claims per atomic.
  uint32_t slots_per_claim = 1;
  // Tally launches: ring streams plus each step's Fresnel and rotation work
  // sorted onto block-wide LDS queues (takes precedence over both above).
  uint32_t sorted_phases = 0;
  // The launch keeps the Fresnel cell bounds in LDS (LaunchOut::fresnel_bounds).
  bool fresnel_bounds = false;
};

// Slots per claim for a slot queue of n_slots drained by `waves` waves: up to
// 16 (fewer atomics on the one queue counter), but small enough that every
// wave gets at least 4 claims, so a small batch still reaches every wave.
inline uint32_t slots_per_claim(uint64_t n_slots, uint64_t waves) {
  const uint64_t k = n_slots / (4u * (waves ? waves : 1u));
  return static_cast<uint32_t>(k < 1u ? 1u : (k > 16u ? 16u : k));
}

// Per-interaction ("sorted") wavefront launches. Photon state lives in HBM
// as structure-of-arrays over W = grid * seg slots; block b of every kernel
// owns slots [b * seg, (b + 1) * seg) and photon indices [lo_b, hi_b) with
// lo_b = n * b / grid, so every queue below is a set of per-block segments
// whose counters only block b touches (no contended atomics). One iteration
// is three launches:
//   classify: for every slot, in slot order (coalesced): an empty slot takes
//             the block's next photon (emission); then one step's common
//             part; the photon then goes on one of the per-type queues below,
//             stays ready (specular reflection), or finishes;
//   fresnel:  readout and polished-face decisions from their queue;
//   rotate:   Rayleigh and Lambertian directions from their queue.
struct SortedArgs {
  // photon state [W]; idx = photon index - first_id, kSortedIdle if the
  // slot holds no photon
  float *px, *py, *pz, *dx, *dy, *dz, *t, *path;
  uint32_t *nb, *ns, *idx, *draws;
  // queues: entries [W] (segment b at b * seg), counts [grid]
  uint32_t *rot_key, *rot_wc, *rot_wp, *cnt_rot;  // key = slot | code << 29
  uint32_t *fres_key, *fres_w, *cnt_fres;         // key = slot | face << 29
  unsigned long long* next_local;  // [grid] next photon index of block b
  uint32_t* occ;                   // [grid] occupied slots of block b
  uint32_t* partials;              // [grid][tally.words]
  unsigned long long* sums;        // [grid][2] boundary / scatter sums
  unsigned long long* stats;       // [grid][kSortedStats]
  unsigned long long* status;      // [2]: occupied slots, blocks with photons left
  DevTally tally;
  uint64_t first_id, n;
  uint32_t seg, grid;
};

constexpr uint32_t kSortedIdle = 0xffffffffu;

enum : int {
  kSortedClassified = 0,  // photon-steps of classify
  kSortedFresnel = 1,
  kSortedRotate = 2,
  kSortedEmitted = 3,
  kSortedTail = 4,  // photons finished by the tail kernel
  kSortedStats = 8
};

constexpr int kSortedBlock = 256;

// Empties every slot; resets the counters, partial rows and stats.
void launch_sorted_init(const SortedArgs& A, hipStream_t stream);
// One iteration: classify, fresnel, rotate.
void launch_sorted_iteration(const DevParams& P, PhiloxKey key, const SortedArgs& A,
                             hipStream_t stream);
// status[0] = occupied slots, status[1] = blocks with photon indices left.
void launch_sorted_status(const SortedArgs& A, hipStream_t stream);
// Finishes the photons in the occupied slots, one lane per photon (history style).
void launch_sorted_tail(const DevParams& P, PhiloxKey key, const SortedArgs& A,
                        hipStream_t stream);

// History-based launches: one lane per photon (records), or a grid-stride
// loop over photons (tally). first_id offsets the Philox photon ids; records
// are written at index i - first_id.
void launch_history(const DevParams& P, PhiloxKey key, uint64_t first_id, uint64_t n, int block,
                    int grid, const LaunchOut& out, hipStream_t stream);

// Persistent wavefront launch: `queue` (one u64, zeroed by the caller) is the
// device work-queue head.
void launch_wavefront(const DevParams& P, PhiloxKey key, uint64_t first_id, uint64_t n,
                      const WavefrontShape& shape, unsigned long long* queue, const LaunchOut& out,
                      hipStream_t stream);

// Batched events: one persistent launch over all events (event e owns ids
// first_id + offsets[e] .. first_id + offsets[e+1] - 1), per-lane event
// parameters from out.evt.table, per-event tallies. The queue hands out the
// n_slots slots of out.evt.slots in order (one per chunk of queue_chunk
// photons of an event).
void launch_batch_wavefront(PhiloxKey key, uint64_t first_id, uint64_t n_slots,
                            const WavefrontShape& shape, unsigned long long* queue,
                            const LaunchOut& out, hipStream_t stream);

// Plain per-event launch: the history kernel over one event's photons with its
// parameters as kernel arguments, tallied into out.evt.counts row out.evt.event.
void launch_history_event(const DevParams& P, PhiloxKey key, uint64_t first_id, uint64_t n,
                          int block, const LaunchOut& out, hipStream_t stream);

// Per-block LDS copies of the event tallies are used when they take at most
// kBatchLdsMax bytes; batch_lds_bytes() is that size, or 0 for global atomics.
constexpr size_t kBatchLdsMax = 32768;
size_t batch_lds_bytes(uint32_t n_events, uint32_t words);
// LDS bytes of a batch launch with ring streams: the ring region
// (ring_words), then the event tallies (if any).
size_t batch_ring_lds_bytes(size_t tally_lds, int block);
// Otherwise each wave of a block keeps the row of the event it is drawing
// photons from in LDS: batch_wave_row_bytes() is that LDS for a block, or 0
// when it exceeds kBatchLdsMax (the rows then take global atomics only).
size_t batch_wave_row_bytes(uint32_t words, int block);
// The tally LDS of a batch launch: the block's copy of every event row, else
// the waves' rows (if they fit), else 0.
size_t batch_tally_lds_bytes(uint32_t n_events, uint32_t words, int block);
int batch_blocks_per_cu(int block, uint32_t n_events, uint32_t words, bool ring = false);

// Adjoint (likelihood-ratio + pathwise) pass over the histories of a set of
// parameter sets, recomputed from their Philox counters.
// surface: the model shared by every set (0..3), or -1 if they differ. With
// a single set and params_by_value its parameters P0 are passed as kernel
// arguments; otherwise lanes read them from a.table (as a captured graph
// that is replayed with new parameters must).
void launch_adjoint(const DevParams& P0, int surface, PhiloxKey key, uint64_t first_id,
                    const WavefrontShape& shape, const AdjointArgs& a, hipStream_t stream,
                    bool params_by_value = true);
int adjoint_blocks_per_cu(int block);

// Sums the per-block histograms and sums into the flat u64 tally layout.
void launch_tally_reduce(const DevTally& t, int blocks, const uint32_t* partials,
                         const unsigned long long* sums, unsigned long long* tally,
                         hipStream_t stream);

// Fills n u64 words with zero (a kernel, so it may sit inside captured graphs).
void launch_zero_u64(unsigned long long* p, uint64_t n, hipStream_t stream);
// The same for two ranges, in one launch.
void launch_zero_u64_pair(unsigned long long* a, uint64_t na, unsigned long long* b, uint64_t nb,
                          hipStream_t stream);

// Fills the 2^24-entry Rayleigh-cosine table (DevParams::rayleigh).
constexpr size_t kRayleighTableEntries = size_t{1} << 24;
void launch_rayleigh_table(float* table, hipStream_t stream);

// Persistent transport of genstep photons into the block tallies (and
// records at each photon's global index when out.records).
void launch_gensteps(const DevParams& P, PhiloxKey key, uint64_t first_id,
                     const WavefrontShape& shape, const GenstepArgs& g,
                     unsigned long long* queue, const LaunchOut& out, hipStream_t stream);
int genstep_blocks_per_cu(bool records, int block, size_t lds, bool ring = false,
                          int surface = -1, bool bounds = false);

// Per-set chi-square loss of the batch tally's arrival-time histograms
// (rows of `words` u32: 5 fates, then bins1 time bins) against target
// fractions, and its gradient with respect to the histogram fractions
// (weights for the adjoint). One block per set.
void launch_chi2_weights(const uint32_t* counts, uint32_t words, uint32_t bins1,
                         const unsigned long long* offsets, const double* target,
                         const double* inv_nt, double* hist, double* loss, double* weights,
                         uint32_t n_sets, hipStream_t stream);

// Occupancy queries (blocks per CU for a block size and dynamic LDS bytes).
int history_blocks_per_cu(bool tally, int block, size_t lds_bytes);  // all dynamic LDS
// Words of the LDS ring region of a block: 16 slots of the block size rounded
// up to a power of two (the device side's ring_region_words()).
inline size_t ring_words(int block) {
  size_t p = 1;
  while (p < static_cast<size_t>(block)) p <<= 1;
  return 16u * p;
}
// LDS bytes of a wavefront tally launch with the streams in LDS: the ring
// region and the tally words.
size_t ring_lds_bytes(uint32_t tally_words, int block);
// Fresnel cell-bound tables a launch keeps in LDS: the readout's, and the
// outside medium's when the side faces are polished (or the surface model is
// not fixed at compile time).
constexpr uint32_t fresnel_tables(int surface) {
  return (surface == static_cast<int>(kPolished) || surface < 0) ? 2u : 1u;
}
// Bytes of those tables (0 when the launch does not keep them).
constexpr size_t fresnel_lds_bytes(bool on, int surface) {
  return on ? static_cast<size_t>(fresnel_tables(surface)) * kFresnelCells * 4u : 0u;
}
// LDS of a history or register-stream wavefront launch: the tally (0 words
// for records), then the Fresnel cell bounds when `bounds`.
size_t history_lds_bytes(uint32_t tally_words, bool bounds, int surface);
// LDS of a wavefront ring launch: ring region, tally, Fresnel cell bounds
// when `bounds`.
size_t wavefront_ring_lds_bytes(uint32_t tally_words, int block, int surface, bool bounds);
// Whether every kernel that places its ring region at LDS address 0 has no
// static LDS, so that its dynamic LDS (and the region) starts at address 0.
bool ring_kernels_start_at_lds_zero();
// LDS bytes of the sorted-phase kernel: the ring layout, then the per-parity
// queue counters and 9 words per lane of queues and exchange slots.
size_t phase_lds_bytes(uint32_t tally_words, int block);
// `lds`: the tally bytes; the function adds the ring region, phase queues and
// Fresnel bounds (`bounds`, for `surface`) of the kernel selected.
int wavefront_blocks_per_cu(bool records, bool tally, int block, size_t lds,
                            bool successor = false, bool ring = false, bool phases = false,
                            int surface = -1, bool bounds = false);

}  // namespace ophip
