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
  kCntSlots = 16
};

// Output selection of a transport launch.
struct LaunchOut {
  DevRecords rec;              // used when `records`
  DevTally tally;              // used when `tally`
  uint32_t* partials;          // [grid][tally.words] block histograms
  unsigned long long* sums;    // [grid][2] block boundary / scatter sums
  unsigned long long* counters;  // [kCntSlots], accumulated
  bool records, tally_on;
};

struct WavefrontShape {
  int block, grid;
  uint32_t steps_per_refill, queue_chunk;
};

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

// Sums the per-block histograms and sums into the flat u64 tally layout.
void launch_tally_reduce(const DevTally& t, int blocks, const uint32_t* partials,
                         const unsigned long long* sums, unsigned long long* tally,
                         hipStream_t stream);

// Fills n u64 words with zero (a kernel, so it may sit inside captured graphs).
void launch_zero_u64(unsigned long long* p, uint64_t n, hipStream_t stream);

// Occupancy queries (blocks per CU for a block size and dynamic LDS bytes).
int history_blocks_per_cu(bool tally, int block, size_t lds_bytes);
int wavefront_blocks_per_cu(bool records, bool tally, int block, size_t lds_bytes);

}  // namespace ophip
