// transport_kernels.hip -- optical-photon transport kernels for gfx950.
//
// Two drivers of the same per-lane physics (photon.hpp):
//
//  * history_kernel: one lane per photon for the photon's whole life (the
//    reference's history-based kernels, gpu_backend.cu:47-100). A wave runs
//    as long as its longest history.
//
//  * wavefront_kernel: a persistent grid. Every lane holds one photon in
//    registers; each wave claims photon ids from a device work queue in
//    chunks (one u64 atomic per chunk) and, every `steps_per_refill` steps,
//    re-fills the lanes whose photon has finished. The refill is a wave64
//    stream compaction of the work queue onto the dead lanes: ballot of the
//    dead lanes, popcount for the claim, mbcnt for each lane's rank. Photon
//    state never leaves the registers, so the wave stays full until the queue
//    drains and no state traffic goes to HBM. This replaces the reference's
//    relaunch-and-compact loop (gpu_event.cu:125-170) and its per-launch host
//    round trip; results are identical because a photon's history depends
//    only on (seed, photon id).
//
// Loops whose trip count differs between lanes are written as wave-uniform
// loops driven by ballots, so the per-wave counters stay in SGPRs and no
// value is observed at a lane-dependent loop exit.
//
// Tallies: each block accumulates u32 histograms [fates | time | xy] in LDS,
// writes them once to its row of a [grid][words] buffer, and a reduction
// kernel sums the rows into the u64 tally. Integer sums: the result does not
// depend on block size, grid size or scheduling.
#include <stdexcept>

#include "kernels.hpp"
#include "wave.hpp"

namespace ophip {
namespace {

// Batched events and gensteps: a wave claims `per_claim` consecutive queue
// slots per atomic (slots_per_claim() on the host: at most 16, and few enough
// that every wave gets at least 4 claims). A slot is one chunk of one event
// or genstep, from a host-built slot table.

// Steps into a history beyond which a batch wave raises its issue priority.
constexpr uint32_t kLongHistory = 256u;

// Wave issue priority 2 while any lane's `is_long` holds, 0 otherwise (the
// priority only orders this kernel's waves against each other on a SIMD).
__device__ __forceinline__ void set_priority_for_long_history(bool is_long) {
  if (__ballot(is_long) != 0)
    __builtin_amdgcn_s_setprio(2);
  else
    __builtin_amdgcn_s_setprio(0);
}

// Wave-uniform branch-event counters (kept in SGPRs). Compiled in only for
// the launches that report them: 16 live 64-bit scalars otherwise push the
// kernel's scalar registers into VGPR lanes.
template <bool kOn>
struct WaveCounters {
  uint64_t v[kCntSlots];
  __device__ __forceinline__ void clear() {
#pragma unroll
    for (int i = 0; i < kCntSlots; ++i) v[i] = 0;
  }
  __device__ __forceinline__ void add(int slot, uint64_t x) {
    if (kOn) v[slot] += x;
  }
  __device__ __forceinline__ void add_step(uint64_t active, const StepInfo& s) {
    if (!kOn) return;
    v[kCntUseful] += popc(active);
    v[kCntIssued] += 64;
    v[kCntAbsorb] += popc(__ballot(s.absorb));
    v[kCntScatter] += popc(__ballot(s.scatter));
    v[kCntReadout] += popc(__ballot(s.readout));
    v[kCntPolished] += popc(__ballot(s.polished));
    v[kCntBlack] += popc(__ballot(s.black));
    v[kCntReflector] += popc(__ballot(s.refl));
    v[kCntLambert] += popc(__ballot(s.lambert_reflect));
    v[kCntAbsOn] += popc(__ballot(s.abs_on));
    v[kCntScatOn] += popc(__ballot(s.scat_on));
  }
  __device__ __forceinline__ void flush(unsigned long long* dst) const {
    if (kOn && lane_id() == 0) {
#pragma unroll
      for (int i = 0; i < kCntSlots; ++i)
        if (v[i]) atomicAdd(dst + i, static_cast<unsigned long long>(v[i]));
    }
  }
};

__device__ __forceinline__ void zero_lds(uint32_t* lds, uint32_t words) {
  for (uint32_t w = threadIdx.x; w < words; w += blockDim.x) lds[w] = 0u;
}

// Output of one finished photon: its record at `idx`, its block-tally entries
// and/or its event-tally entries (event `e`, global u32 atomics).
template <bool kRec, bool kTally, bool kEvt = false>
__device__ __forceinline__ void finish(const DevParams& P, const LaunchOut& out, uint32_t* lds,
                                       uint64_t idx, const Photon& ph, uint32_t fate,
                                       uint64_t& snb, uint64_t& sns, uint32_t e = 0) {
  if (kEvt) {
    uint32_t* row = out.evt.counts + static_cast<size_t>(e) * out.evt.words;
    atomicAdd(row + fate, 1u);
    if (fate == kDetected)
      atomicAdd(row + 5u + time_bin(ph.t, out.evt.time_scale, out.evt.time_bins), 1u);
  }
  if (kRec) {
    out.rec.fate[idx] = static_cast<uint8_t>(fate);
    out.rec.nb[idx] = ph.nb;
    out.rec.ns[idx] = ph.ns;
    out.rec.t[idx] = ph.t;
    out.rec.x[idx] = ph.p.x;
    out.rec.y[idx] = ph.p.y;
    out.rec.z[idx] = ph.p.z;
    out.rec.path[idx] = ph.path;
  }
  if (kTally) {
    const DevTally& T = out.tally;
    atomicAdd(&lds[fate], 1u);
    if (fate == kDetected) {
      atomicAdd(&lds[5u + time_bin(ph.t, T.time_scale, T.time_bins)], 1u);
      const uint32_t xy = xy_bin(ph.p.x, ph.p.y, P.hx, P.hy, T.two_hx, T.two_hy, T.xy_bins,
                                 P.fast_div & 1u);
      atomicAdd(&lds[6u + T.time_bins + xy], 1u);
    }
    snb += ph.nb;
    sns += ph.ns;
  }
}

// Per-lane fate counts of one event of a batch. Every finished photon adds to
// one of its event's five fate words; with few events (a calibration batch)
// those words take a global atomic per photon and the atomics queue on them.
// A lane's consecutive photons mostly come from one queue slot, i.e. one
// event, so the lane counts in registers and adds its counts to the event's
// row (global or LDS) when its photons move to another event, when a count
// reaches 2^30, and at the end. Integer sums: the rows are unchanged.
struct EventFates {
  uint32_t e = 0xffffffffu;
  uint32_t c0 = 0u, c1 = 0u, c2 = 0u, c3 = 0u, c4 = 0u;

  __device__ __forceinline__ void flush(const DevEventTally& T) {
    if (e == 0xffffffffu) return;
    uint32_t* row = T.counts + static_cast<size_t>(e) * T.words;
    if (c0) atomicAdd(row + 0, c0);
    if (c1) atomicAdd(row + 1, c1);
    if (c2) atomicAdd(row + 2, c2);
    if (c3) atomicAdd(row + 3, c3);
    if (c4) atomicAdd(row + 4, c4);
    c0 = c1 = c2 = c3 = c4 = 0u;
  }
  __device__ __forceinline__ void add(const DevEventTally& T, uint32_t event, uint32_t fate) {
    if (event != e || c0 + c1 + c2 + c3 + c4 >= (1u << 30)) {
      flush(T);
      e = event;
    }
    c0 += fate == 0u;
    c1 += fate == 1u;
    c2 += fate == 2u;
    c3 += fate == 3u;
    c4 += fate == 4u;
  }
};

// The row of the event a wave is drawing photons from, kept in LDS by that
// wave (batches whose event rows do not all fit in a block's LDS). A wave
// refills its lanes from one slot (one chunk of one event) at a time, so most
// of its finished photons belong to the event of its current slot: their
// fates and arrival-time bins add to the wave's row with LDS atomics, and the
// row goes to the event's global row, one atomic per nonzero word, when the
// wave opens a slot of another event and at the end. Photons of an earlier
// slot's event still in flight add to the global rows directly. Integer sums:
// the rows are unchanged.
struct WaveRow {
  uint32_t* row;             // `words` u32 of this wave in LDS (used when `on`)
  uint32_t e = 0xffffffffu;  // event the row holds (wave-uniform; none yet)
  bool on = false;

  // All lanes of the wave, in wave-uniform control flow.
  __device__ __forceinline__ void clear(uint32_t words) {
    if (!on) return;
    for (uint32_t w = lane_id(); w < words; w += 64u) row[w] = 0u;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  }
  __device__ __forceinline__ void flush(const DevEventTally& T) {
    if (!on || e == 0xffffffffu) return;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    uint32_t* dst = T.counts + static_cast<size_t>(e) * T.words;
    for (uint32_t w = lane_id(); w < T.words; w += 64u) {
      const uint32_t v = row[w];
      if (v) {
        atomicAdd(dst + w, v);
        row[w] = 0u;
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  }
  __device__ __forceinline__ void take(const DevEventTally& T, uint32_t event) {
    if (!on || event == e) return;
    flush(T);
    e = event;
  }
};

// A finished photon of event e of a batch: its fate and, if detected, its
// arrival-time bin into the event's row. kDirect: straight into the row (rows
// in LDS, where atomics are cheap); else into the wave's LDS row when it holds
// event e, otherwise the fate into the lane's counts and the time bin into
// the global row.
template <bool kDirect>
__device__ __forceinline__ void finish_event(const DevEventTally& T, EventFates& fates,
                                             WaveRow& wr, uint32_t e, const Photon& ph,
                                             uint32_t fate) {
  const uint32_t tb = fate == kDetected ? 5u + time_bin(ph.t, T.time_scale, T.time_bins) : 0u;
  if (!kDirect && e == wr.e) {  // wave-scope LDS adds (only this wave uses the row)
    __hip_atomic_fetch_add(wr.row + fate, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
    if (tb) __hip_atomic_fetch_add(wr.row + tb, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
    return;
  }
  uint32_t* row = T.counts + static_cast<size_t>(e) * T.words;
  if (kDirect)
    atomicAdd(row + fate, 1u);
  else
    fates.add(T, e, fate);
  if (tb) atomicAdd(row + tb, 1u);
}

template <bool kTally, bool kCount>
__device__ __forceinline__ void flush_block(const LaunchOut& out, uint32_t* lds, uint64_t snb,
                                            uint64_t sns, const WaveCounters<kCount>& wc) {
  wc.flush(out.counters);
  if (kTally) {
    const uint64_t bnb = wave_sum_u64(snb), bns = wave_sum_u64(sns);
    if (lane_id() == 0) {
      if (bnb) atomicAdd(out.sums + 2u * blockIdx.x, static_cast<unsigned long long>(bnb));
      if (bns) atomicAdd(out.sums + 2u * blockIdx.x + 1u, static_cast<unsigned long long>(bns));
    }
    __syncthreads();
    uint32_t* dst = out.partials + static_cast<size_t>(blockIdx.x) * out.tally.words;
    for (uint32_t w = threadIdx.x; w < out.tally.words; w += blockDim.x) dst[w] = lds[w];
  }
}

// Fresnel cell bounds of the run (fresnel_cell_bounds) into LDS, filled by
// the block: the readout's table, then the outside medium's when `tables` is 2.
__device__ __forceinline__ void fill_fresnel_bounds(const DevParams& P, uint32_t* fres,
                                                    uint32_t tables) {
  for (uint32_t j = threadIdx.x; j < tables * kFresnelCells; j += blockDim.x) {
    const bool ro = j < kFresnelCells;
    fres[j] = fresnel_cell_bounds(P.n_scint, ro ? P.n_det : P.n_out, ro ? P.eta2_det : P.eta2_out,
                                  ro ? P.det_matched != 0u : P.out_matched != 0u,
                                  j & (kFresnelCells - 1u));
  }
}

// A run's common case, fixed at compile time for the kernels instantiated
// with it: the short divisions in range (fast_div == 3), neither interface
// index-matched, and the Rayleigh table present. The compiler then drops the
// IEEE-division and matched-interface paths and the scalar registers that
// select them (fewer spills of loop-invariant values). The host launches such
// an instantiation only when common_case() holds.
__device__ __forceinline__ void assume_common_case(const DevParams& P) {
  __builtin_assume(P.rayleigh != nullptr);
  __builtin_assume(P.fast_div == 3u);
  __builtin_assume(P.det_matched == 0u);
  __builtin_assume(P.out_matched == 0u);
}

bool common_case(const DevParams& P) {
  return P.rayleigh != nullptr && P.fast_div == 3u && P.det_matched == 0u &&
         P.out_matched == 0u;
}

// The ring kernels' common-case instantiations also fix the block size, so
// the slot stride of the LDS ring is a compile-time constant: 4 bytes x 512
// lanes, 2^11.
constexpr int kCommonRingBlock = 512;
constexpr uint32_t kCommonRingShift = 11u;
static_assert((4u << 9) == (1u << kCommonRingShift) && kCommonRingBlock == (1 << 9),
              "ring stride of the common block size");

// ---------------------------------------------------------------------------
// History-based kernel. Records: the grid covers n, lane i <-> photon i.
// Tally: waves stride over the photons in steps of the whole grid.
template <bool kRec, bool kTally, bool kEvt, bool kCount, bool kCommon = false>
__global__ __launch_bounds__(1024) void history_kernel(DevParams P, PhiloxKey key,
                                                       uint64_t first_id, uint64_t n,
                                                       LaunchOut out) {
  if (kCommon) assume_common_case(P);
  extern __shared__ uint32_t lds[];
  // LDS: the tally (if any), then the Fresnel cell bounds (when the launch
  // keeps them; never for the short per-event launches of a batch)
  uint32_t* const fres = (!kEvt && out.fresnel_bounds)
                             ? lds + (kTally ? (out.tally.words + 3u) & ~3u : 0u)
                             : nullptr;
  if (kTally) zero_lds(lds, out.tally.words);
  if (fres) fill_fresnel_bounds(P, fres, fresnel_tables(static_cast<int>(P.surface)));
  if (kTally || fres) __syncthreads();
  const uint32_t lane = lane_id();
  const uint64_t wave0 = static_cast<uint64_t>(blockIdx.x) * blockDim.x + (threadIdx.x - lane);
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  WaveCounters<kCount> wc;
  wc.clear();
  uint64_t snb = 0, sns = 0;

  for (uint64_t base = wave0; base < n; base += stride) {  // uniform per wave
    const uint64_t i = base + lane;
    bool alive = i < n;
    Stream rng;
    Photon ph;
    uint32_t fate = kMaxSteps;
    if (alive) {
      rng.start(first_id + i, key, 0u);
      emit(P, rng, ph);
    }
    if (kCount) wc.add(kCntEmitted, popc(__ballot(alive)));
    const bool valid = alive;
    alive = alive && P.max_steps > 0u;
    uint32_t done = 0;
    for (;;) {
      const uint64_t act = __ballot(alive);
      if (act == 0) break;
      StepInfo info{};
      if (alive) {
        rng.refill(key);
        const bool still = step(P, rng, ph, fate, info, nullptr, fres);
        ++done;
        alive = still && done < P.max_steps;
        if (still && !alive) fate = kMaxSteps;
      }
      wc.add_step(act, info);
    }
    if (kCount) wc.add(kCntDetected, popc(__ballot(valid && fate == kDetected)));
    if (valid) finish<kRec, kTally, kEvt>(P, out, lds, i, ph, fate, snb, sns, out.evt.event);
  }
  if (kTally) __syncthreads();
  flush_block<kTally, kCount>(out, lds, snb, sns, wc);
}

// ---------------------------------------------------------------------------
// Persistent wavefront kernel (see the file comment).
template <bool kRec, bool kTally, bool kCount, int kSurf = -1, bool kCommon = false>
__global__ __launch_bounds__(1024) void wavefront_kernel(DevParams P0, PhiloxKey key,
                                                         uint64_t first_id, uint64_t n,
                                                         uint32_t steps_per_refill, uint32_t chunk,
                                                         unsigned long long* queue,
                                                         LaunchOut out) {
  const DevParams& P = P0;
  if (kCommon) assume_common_case(P0);
  extern __shared__ uint32_t lds[];
  // LDS: the tally (if any), then the Fresnel cell bounds (if kept)
  uint32_t* const fres =
      out.fresnel_bounds ? lds + (kTally ? (out.tally.words + 3u) & ~3u : 0u) : nullptr;
  if (kTally) zero_lds(lds, out.tally.words);
  if (fres)
    fill_fresnel_bounds(P, fres, fresnel_tables(kSurf >= 0 ? kSurf : static_cast<int>(P.surface)));
  if (kTally || fres) __syncthreads();
  const uint32_t lane = lane_id();
  WaveCounters<kCount> wc;
  wc.clear();
  uint64_t snb = 0, sns = 0;

  uint64_t q_next = 0, q_end = 0;  // photons claimed by this wave, not yet started
  bool q_done = false;             // the global queue is exhausted
  bool alive = false;
  uint64_t idx = 0;
  uint32_t done = 0, fate = kMaxSteps;
  Stream rng;
  Photon ph;

  for (;;) {
    // ---- refill: compact queued photons onto the dead lanes ----------------
    const uint64_t dead = __ballot(!alive);
    if (dead != 0 && !q_done) {
      const uint32_t need = popc(dead);
      const uint64_t have = q_end - q_next;
      const uint32_t t1 = have < need ? static_cast<uint32_t>(have) : need;
      const uint64_t b1 = q_next;
      q_next += t1;
      uint64_t b2 = 0;
      uint32_t t2 = 0;
      if (t1 < need) {
        unsigned long long got = 0;
        if (lane == 0) got = atomicAdd(queue, static_cast<unsigned long long>(chunk));
        got = first_lane_u64(got);
        if (got >= n) {
          q_done = true;
        } else {
          q_end = (got + chunk < n) ? got + chunk : n;
          const uint64_t avail = q_end - got;
          t2 = (need - t1 < avail) ? need - t1 : static_cast<uint32_t>(avail);
          b2 = got;
          q_next = got + t2;
        }
      }
      const uint32_t r = rank_in(dead);
      bool fresh = false;
      if (!alive && r < t1 + t2) {
        idx = r < t1 ? b1 + r : b2 + (r - t1);
        rng.start(first_id + idx, key, 0u);
        emit(P, rng, ph);
        done = 0;
        fate = kMaxSteps;
        alive = P.max_steps > 0u;
        fresh = true;
      }
      wc.add(kCntEmitted, t1 + t2);
      if (fresh && !alive) {  // max_steps == 0: finished at emission
        finish<kRec, kTally>(P, out, lds, idx, ph, fate, snb, sns);
      }
    }
    if (__ballot(alive) == 0) {
      if (q_done) break;
      continue;
    }

    // ---- up to steps_per_refill transport steps -----------------------------
    for (uint32_t k = 0; k < steps_per_refill; ++k) {
      const uint64_t act = __ballot(alive);
      if (act == 0) break;
      StepInfo info{};
      bool finished = false;
      if (alive) {
        rng.refill(key);
        const bool still = step<kSurf>(P, rng, ph, fate, info, nullptr, fres);
        ++done;
        if (!still || done >= P.max_steps) {
          if (still) fate = kMaxSteps;
          finished = true;
          alive = false;
        }
      }
      wc.add_step(act, info);
      if (kCount) wc.add(kCntDetected, popc(__ballot(finished && fate == kDetected)));
      if (finished) finish<kRec, kTally>(P, out, lds, idx, ph, fate, snb, sns);
    }
  }
  if (kTally) __syncthreads();
  flush_block<kTally, kCount>(out, lds, snb, sns, wc);
}

// ---------------------------------------------------------------------------
// Persistent wavefront kernel with successors (tally). Besides its current
// photon every lane may hold an already emitted successor (position,
// direction, emission time and Philox window). Emission runs for the lanes of
// a wave that lack a successor only once at least `threshold` of them do (or
// when nothing else is left), so the emission code runs with many active
// lanes; a lane whose photon finishes swaps its successor in at once and
// keeps stepping. Per photon the computation is the one of the other kernels.
template <bool kCount, int kSurf>
__global__ __launch_bounds__(1024) void wavefront_next_kernel(DevParams P0, PhiloxKey key,
                                                              uint64_t first_id, uint64_t n,
                                                              uint32_t threshold, uint32_t chunk,
                                                              unsigned long long* queue,
                                                              LaunchOut out) {
  const DevParams& P = P0;
  extern __shared__ uint32_t lds[];
  zero_lds(lds, out.tally.words);
  __syncthreads();
  const uint32_t lane = lane_id();
  WaveCounters<kCount> wc;
  wc.clear();
  uint64_t snb = 0, sns = 0;

  uint64_t q_next = 0, q_end = 0;  // photons claimed by this wave, not yet emitted
  bool q_done = false;             // the global queue is exhausted
  bool alive = false, has_next = false;
  uint32_t done = 0, fate = kMaxSteps;
  Stream rng;
  Photon ph;
  // the successor: what emission produced (path and counts start at zero;
  // the Philox window holds blocks 0 and 1, so the next block is 2)
  V3 nx_p{}, nx_d{};
  float nx_t = 0.0f;
  uint32_t nw0 = 0, nw1 = 0, nw2 = 0, nw3 = 0, nw4 = 0, nw5 = 0, nw6 = 0, nw7 = 0;
  uint32_t npos = 0, nid_lo = 0, nid_hi = 0;

  for (;;) {
    // ---- emit successors for the lanes that lack one ------------------------
    // Emit when `threshold` lanes lack a successor, when a quarter of that
    // many lanes have neither photon nor successor, or when no lane is alive.
    const uint64_t need = __ballot(!has_next);
    const uint32_t n_need = popc(need);
    const uint32_t n_idle = popc(__ballot(!alive && !has_next));
    const bool go = n_need >= threshold || 4u * n_idle >= threshold || __ballot(alive) == 0;
    if (!q_done && n_need != 0 && go) {
      const uint64_t have = q_end - q_next;
      const uint32_t t1 = have < n_need ? static_cast<uint32_t>(have) : n_need;
      const uint64_t b1 = q_next;
      q_next += t1;
      uint64_t b2 = 0;
      uint32_t t2 = 0;
      if (t1 < n_need) {
        unsigned long long got = 0;
        if (lane == 0) got = atomicAdd(queue, static_cast<unsigned long long>(chunk));
        got = first_lane_u64(got);
        if (got >= n) {
          q_done = true;
        } else {
          q_end = (got + chunk < n) ? got + chunk : n;
          const uint64_t avail = q_end - got;
          t2 = (n_need - t1 < avail) ? n_need - t1 : static_cast<uint32_t>(avail);
          b2 = got;
          q_next = got + t2;
        }
      }
      const uint32_t r = rank_in(need);
      if (!has_next && r < t1 + t2) {
        const uint64_t idx = r < t1 ? b1 + r : b2 + (r - t1);
        Stream g;
        Photon e;
        g.start(first_id + idx, key, 0u);
        emit(P, g, e);
        nx_p = e.p;
        nx_d = e.d;
        nx_t = e.t;
        nw0 = g.w0, nw1 = g.w1, nw2 = g.w2, nw3 = g.w3;
        nw4 = g.w4, nw5 = g.w5, nw6 = g.w6, nw7 = g.w7;
        npos = g.pos, nid_lo = g.id_lo, nid_hi = g.id_hi;
        has_next = true;
      }
      wc.add(kCntEmitted, t1 + t2);
    }
    // ---- lanes without a photon take their successor ------------------------
    if (!alive && has_next) {
      ph.p = nx_p;
      ph.d = nx_d;
      ph.t = nx_t;
      ph.path = 0.0f;
      ph.nb = 0u;
      ph.ns = 0u;
      rng.w0 = nw0, rng.w1 = nw1, rng.w2 = nw2, rng.w3 = nw3;
      rng.w4 = nw4, rng.w5 = nw5, rng.w6 = nw6, rng.w7 = nw7;
      rng.pos = npos;
      rng.blk = 2u;
      rng.id_lo = nid_lo;
      rng.id_hi = nid_hi;
      has_next = false;
      done = 0;
      fate = kMaxSteps;
      alive = P.max_steps > 0u;
      if (!alive) {  // max_steps == 0: finished at emission
        finish<false, true>(P, out, lds, 0, ph, fate, snb, sns);
      }
    }
    const uint64_t act = __ballot(alive);
    if (act == 0) {
      if (q_done && __ballot(has_next) == 0) break;
      continue;
    }

    // ---- one transport step ----------------------------------------------------
    StepInfo info{};
    bool finished = false;
    if (alive) {
      rng.refill(key);
      const bool still = step<kSurf>(P, rng, ph, fate, info);
      ++done;
      if (!still || done >= P.max_steps) {
        if (still) fate = kMaxSteps;
        finished = true;
        alive = false;
      }
    }
    wc.add_step(act, info);
    if (kCount) wc.add(kCntDetected, popc(__ballot(finished && fate == kDetected)));
    if (finished) finish<false, true>(P, out, lds, 0, ph, fate, snb, sns);
  }
  __syncthreads();
  flush_block<true, kCount>(out, lds, snb, sns, wc);
}

// ---------------------------------------------------------------------------
// Persistent wavefront kernel with the photons' random streams in LDS (tally).
// Each lane's Philox window is a 16-word ring in LDS (RingStream). Before a
// step the wave runs generation passes until every lane about to draw holds
// enough words (5 for a step, plus 3 or 6 for an emission); in each pass
// every such lane with room appends a block, so the lanes stay topped up and
// most steps need one pass or none. Emission draws from the same ring, so a
// refill costs no Philox code of its own. Finished photons are tallied at the
// next refill point, once per `steps_per_refill` steps instead of every step.
// Per photon the words drawn and the arithmetic are those of the other
// kernels.
template <bool kCount, int kSurf, bool kCommon>
__global__ __launch_bounds__(1024) void wavefront_ring_kernel(DevParams P0, PhiloxKey key,
                                                              uint64_t first_id, uint64_t n,
                                                              uint32_t steps_per_refill,
                                                              uint32_t chunk,
                                                              unsigned long long* queue,
                                                              LaunchOut out) {
  const DevParams& P = P0;
  if (kCommon) assume_common_case(P0);
  extern __shared__ uint32_t lds[];
  // LDS: the ring region first (at address 0, a multiple of its size), then
  // the block's tally, then the Fresnel cell bounds if kept (readout, and the
  // outside medium for polished faces)
  uint32_t* const tal = lds + ring_region_words();
  uint32_t* const fres = out.fresnel_bounds ? tal + ((out.tally.words + 3u) & ~3u) : nullptr;
  zero_lds(tal, out.tally.words);
  if (fres) fill_fresnel_bounds(P, fres, fresnel_tables(kSurf));
  __syncthreads();
  const uint32_t lane = lane_id();
  WaveCounters<kCount> wc;
  wc.clear();
  uint64_t snb = 0, sns = 0;

  uint64_t q_next = 0, q_end = 0;  // photons claimed by this wave, not yet started
  bool q_done = false;             // the global queue is exhausted
  bool alive = false;    // a photon in transport
  bool fresh = false;    // an id assigned, emission pending
  bool pending = false;  // a finished photon not yet tallied
  uint32_t done = 0, fate = kMaxSteps;
  RingStream rng;
  if (kCommon)  // launched with kCommonRingBlock lanes per block only
    rng.place(lds, kCommonRingShift);
  else
    rng.place(lds);
  Photon ph;
  const uint32_t emit_words = P.source == kVolume ? 6u : 3u;

  for (uint32_t it = 0;; ++it) {
    // ---- refill point: tally the finished photons, assign ids to free lanes --
    const bool all_dead = __ballot(alive) == 0;
    if (all_dead || it % steps_per_refill == 0) {
      if (pending) {
        finish<false, true>(P, out, tal, 0, ph, fate, snb, sns);
        pending = false;
      }
      const uint64_t dead = __ballot(!alive);
      if (dead != 0 && !q_done) {
        const uint32_t need = popc(dead);
        const uint64_t have = q_end - q_next;
        const uint32_t t1 = have < need ? static_cast<uint32_t>(have) : need;
        const uint64_t b1 = q_next;
        q_next += t1;
        uint64_t b2 = 0;
        uint32_t t2 = 0;
        if (t1 < need) {
          unsigned long long got = 0;
          if (lane == 0) got = atomicAdd(queue, static_cast<unsigned long long>(chunk));
          got = first_lane_u64(got);
          if (got >= n) {
            q_done = true;
          } else {
            q_end = (got + chunk < n) ? got + chunk : n;
            const uint64_t avail = q_end - got;
            t2 = (need - t1 < avail) ? need - t1 : static_cast<uint32_t>(avail);
            b2 = got;
            q_next = got + t2;
          }
        }
        const uint32_t r = rank_in(dead);
        if (!alive && r < t1 + t2) {
          rng.reset(first_id + (r < t1 ? b1 + r : b2 + (r - t1)));
          fresh = true;
        }
        wc.add(kCntEmitted, t1 + t2);
      }
      if (all_dead && __ballot(fresh) == 0) {
        if (q_done) break;
        continue;
      }
    }

    // ---- generation passes -------------------------------------------------
    const bool drawing = alive || fresh;
    // 0 for a lane that does not draw: its level is never below it, so the
    // loop's ballot is of one comparison
    const uint32_t need = drawing ? (fresh ? emit_words + 5u : 5u) : 0u;
    while (__ballot(rng.level() < need) != 0) {
      if (drawing && rng.room()) rng.generate(key);
    }

    // ---- emission of the newly assigned photons ----------------------------
    if (fresh) {
      emit(P, rng, ph);
      fresh = false;
      done = 0;
      fate = kMaxSteps;
      alive = P.max_steps > 0u;
      if (!alive) pending = true;  // max_steps == 0: finished at emission
    }

    // ---- one transport step --------------------------------------------------
    const uint64_t act = __ballot(alive);
    if (act == 0) continue;
    StepInfo info{};
    bool finished = false;
    if (alive) {
      const bool still = step<kSurf>(P, rng, ph, fate, info, nullptr, fres);
      ++done;
      if (!still || done >= P.max_steps) {
        if (still) fate = kMaxSteps;
        finished = true;
        alive = false;
      }
    }
    wc.add_step(act, info);
    if (kCount) wc.add(kCntDetected, popc(__ballot(finished && fate == kDetected)));
    if (finished) pending = true;
  }
  __syncthreads();
  flush_block<true, kCount>(out, tal, snb, sns, wc);
}

// ---------------------------------------------------------------------------
// Persistent wavefront kernel with each step's branch work sorted by
// interaction type (tally). The ring kernel above runs the Fresnel decision
// and the new-direction sampling inside the step, so a wave pays for them
// with only the lanes that need them active (Fresnel: readout and polished
// hits; rotation: Rayleigh scatters and Lambertian reflections, with the
// Cardano cosine for the scatters only). Here each step is split at those
// branches, the per-interaction design of the reference's next steps
// (README.md:717-720) kept inside one launch:
//   1. every lane runs the common part (photon.hpp advance()) on its own
//      photon in registers and, if it needs a branch body, appends a work
//      item to a block-wide LDS queue: Rayleigh items from the front of the
//      rotation queue, Lambertian items from its back, Fresnel items to
//      their own queue;
//   2. barrier; the block's threads process the queues lane-dense, Fresnel
//      items first, then Rayleigh, then Lambertian, each type starting on a
//      wave boundary, and write each result (new direction, or the Fresnel
//      decision) into the owner's exchange slot;
//   3. barrier; every owner takes its result and completes the step.
// Queue counters are kept per step parity, so each step's counters are
// cleared in the other step's phase 2 and no third barrier is needed. The
// loop is block-uniform: every thread reaches both barriers every step, and
// the block leaves once no lane holds a photon and no wave can claim one.
// Per photon the words drawn and the arithmetic are those of the fused step.
template <bool kCount, int kSurf>
__global__ __launch_bounds__(1024) void wavefront_phase_kernel(DevParams P0, PhiloxKey key,
                                                               uint64_t first_id, uint64_t n,
                                                               uint32_t steps_per_refill,
                                                               uint32_t chunk,
                                                               unsigned long long* queue,
                                                               LaunchOut out) {
  const DevParams& P = P0;
  extern __shared__ uint32_t lds[];
  const uint32_t B = blockDim.x;
  const uint32_t tid = threadIdx.x;
  const uint32_t ring_off = (out.tally.words + 3u) & ~3u;
  uint32_t* cnt = lds + ring_off + ring_region_words();  // [2 parities][4]
  uint32_t* rot_k = cnt + 8;  // owner | code << 16
  uint32_t* rot_c = rot_k + B;
  uint32_t* rot_p = rot_c + B;
  uint32_t* fr_k = rot_p + B;  // owner | face << 16
  uint32_t* fr_w = fr_k + B;
  uint32_t* fr_c = fr_w + B;   // cos_i bits
  float* ex = reinterpret_cast<float*>(fr_c + B);  // exchange slot of each owner
  float* ey = ex + B;
  float* ez = ey + B;
  zero_lds(lds, out.tally.words);
  if (tid < 8) cnt[tid] = 0u;
  __syncthreads();
  const uint32_t lane = lane_id();
  WaveCounters<kCount> wc;
  wc.clear();
  uint64_t snb = 0, sns = 0;

  uint64_t q_next = 0, q_end = 0;
  bool q_done = false;
  bool alive = false, fresh = false, pending = false;
  uint32_t done = 0, fate = kMaxSteps;
  // the ring follows the tally, at an address that is not a multiple of its
  // size: the unaligned addressing
  RingStreamT<false> rng;
  rng.place(lds + ring_off);
  Photon ph;
  const uint32_t emit_words = P.source == kVolume ? 6u : 3u;
  const uint32_t surface = kSurf >= 0 ? static_cast<uint32_t>(kSurf) : P.surface;
  constexpr uint32_t kScatterCode = 6u;

  for (uint32_t it = 0;; ++it) {
    // ---- refill point (per wave, as in the ring kernel) ----------------------
    const bool all_dead = __ballot(alive) == 0;
    if (all_dead || it % steps_per_refill == 0) {
      if (pending) {
        finish<false, true>(P, out, lds, 0, ph, fate, snb, sns);
        pending = false;
      }
      const uint64_t dead = __ballot(!alive);
      if (dead != 0 && !q_done) {
        const uint32_t need = popc(dead);
        const uint64_t have = q_end - q_next;
        const uint32_t t1 = have < need ? static_cast<uint32_t>(have) : need;
        const uint64_t b1 = q_next;
        q_next += t1;
        uint64_t b2 = 0;
        uint32_t t2 = 0;
        if (t1 < need) {
          unsigned long long got = 0;
          if (lane == 0) got = atomicAdd(queue, static_cast<unsigned long long>(chunk));
          got = first_lane_u64(got);
          if (got >= n) {
            q_done = true;
          } else {
            q_end = (got + chunk < n) ? got + chunk : n;
            const uint64_t avail = q_end - got;
            t2 = (need - t1 < avail) ? need - t1 : static_cast<uint32_t>(avail);
            b2 = got;
            q_next = got + t2;
          }
        }
        const uint32_t r = rank_in(dead);
        if (!alive && r < t1 + t2) {
          rng.reset(first_id + (r < t1 ? b1 + r : b2 + (r - t1)));
          fresh = true;
        }
        wc.add(kCntEmitted, t1 + t2);
      }
    }

    // ---- generation passes and emission (per wave) ----------------------------
    const bool drawing = alive || fresh;
    // 0 for a lane that does not draw: its level is never below it, so the
    // loop's ballot is of one comparison
    const uint32_t need = drawing ? (fresh ? emit_words + 5u : 5u) : 0u;
    while (__ballot(rng.level() < need) != 0) {
      if (drawing && rng.room()) rng.generate(key);
    }
    if (fresh) {
      emit(P, rng, ph);
      fresh = false;
      done = 0;
      fate = kMaxSteps;
      alive = P.max_steps > 0u;
      if (!alive) pending = true;  // max_steps == 0: finished at emission
    }

    // ---- phase 1: the common part; branch work onto the queues ---------------
    const uint64_t act = __ballot(alive);
    StepInfo info{};
    Advance a{};
    int kind = -1;  // 0 Rayleigh, 1 Lambertian, 2 Fresnel
    bool finished = false;
    uint32_t wc_word = 0, wp_word = 0;
    if (alive) {
      a = advance(P, rng, ph);
      const bool readout = a.boundary && a.face == kReadoutFace;
      const bool side = a.boundary && a.face != kReadoutFace;
      info.absorb = a.absorb;
      info.scatter = a.scatter;
      info.readout = readout;
      info.polished = side && surface == kPolished;
      info.black = side && surface == kBlack;
      info.refl = side && surface >= kSpecular;
      info.abs_on = P.abs_length != kInf;
      info.scat_on = P.scat_length != kInf;
      uint32_t used = 2u;
      if (a.absorb) {
        finished = true;
        fate = kAbsorbedBulk;
      } else if (a.scatter) {
        kind = 0;
        wc_word = rng.template peek<2>();
        wp_word = rng.template peek<3>();
        used = 4u;
      } else if (readout || info.polished) {
        kind = 2;
        wc_word = rng.template peek<2>();
        used = 3u;
      } else if (info.black) {
        finished = true;
        fate = kAbsorbedSurface;
      } else {  // reflector: u <= R survives (transport.hpp:134-137)
        used = 3u;
        info.survived = to_unit(rng.template peek<2>()) <= P.reflectivity;
        if (!info.survived) {
          finished = true;
          fate = kAbsorbedSurface;
        } else if (surface == kLambertian) {
          kind = 1;
          info.lambert_reflect = true;
          wc_word = rng.template peek<3>();
          wp_word = rng.template peek<4>();
          used = 5u;
        } else {
          flip_axis(ph.d, a.axis);
        }
      }
      rng.pos += used;
    }
    uint32_t* c = cnt + 4u * (it & 1u);
    const uint32_t p_s = reserve(&c[0], kind == 0);
    const uint32_t p_l = reserve(&c[1], kind == 1);
    const uint32_t p_f = reserve(&c[2], kind == 2);
    if (kind == 0) {
      rot_k[p_s] = tid | (kScatterCode << 16);
      rot_c[p_s] = wc_word;
      rot_p[p_s] = wp_word;
      ex[tid] = ph.d.x;
      ey[tid] = ph.d.y;
      ez[tid] = ph.d.z;
    } else if (kind == 1) {
      const uint32_t j = B - 1u - p_l;
      rot_k[j] = tid | (static_cast<uint32_t>(a.face) << 16);
      rot_c[j] = wc_word;
      rot_p[j] = wp_word;
    } else if (kind == 2) {
      fr_k[p_f] = tid | (static_cast<uint32_t>(a.face) << 16);
      fr_w[p_f] = wc_word;
      fr_c[p_f] = __float_as_uint(a.cos_i);
    }
    // barrier 1, and the block's exit test: no photon in any lane and no wave
    // that can still claim one
    if (__syncthreads_or(alive || !q_done) == 0) break;

    // ---- phase 2: the queues, lane-dense --------------------------------------
    const uint32_t n_s = c[0], n_l = c[1], n_f = c[2];
    if (tid < 4) cnt[4u * ((it + 1u) & 1u) + tid] = 0u;  // the next step's counters
    const uint32_t f_pad = (n_f + 63u) & ~63u;
    const uint32_t s_pad = (n_s + 63u) & ~63u;
    const uint32_t total = f_pad + s_pad + n_l;
    for (uint32_t v = tid; v < total; v += B) {
      if (v < f_pad) {
        if (v < n_f) {
          const uint32_t k = fr_k[v];
          const uint32_t owner = k & 0xffffu;
          const bool readout = (k >> 16) == static_cast<uint32_t>(kReadoutFace);
          bool scored;
          const float prob = fresnel_prob(P, readout, __uint_as_float(fr_c[v]), scored);
          ex[owner] = to_unit(fr_w[v]) <= prob ? 1.0f : 0.0f;
        }
      } else if (v < f_pad + s_pad) {
        const uint32_t j = v - f_pad;
        if (j < n_s) {
          const uint32_t owner = rot_k[j] & 0xffffu;
          const V3 d = redirect(P, true, V3{ex[owner], ey[owner], ez[owner]}, 0, false, rot_c[j],
                                rot_p[j]);
          ex[owner] = d.x;
          ey[owner] = d.y;
          ez[owner] = d.z;
        }
      } else {
        const uint32_t j = B - 1u - (v - f_pad - s_pad);
        const uint32_t k = rot_k[j];
        const uint32_t owner = k & 0xffffu;
        const int face = static_cast<int>(k >> 16);
        const V3 d = redirect(P, false, V3{0.0f, 0.0f, 0.0f}, face >> 1, face & 1, rot_c[j],
                              rot_p[j]);
        ex[owner] = d.x;
        ey[owner] = d.y;
        ez[owner] = d.z;
      }
    }
    __syncthreads();  // barrier 2

    // ---- phase 3: every owner completes its step -------------------------------
    if (kind == 0 || kind == 1) {
      ph.d = V3{ex[tid], ey[tid], ez[tid]};
    } else if (kind == 2) {
      info.survived = ex[tid] != 0.0f;
      if (info.survived) {
        flip_axis(ph.d, a.axis);
      } else {
        finished = true;
        fate = info.readout ? kDetected : kEscaped;
      }
    }
    if (alive) {
      ++done;
      if (!finished && done >= P.max_steps) {
        finished = true;
        fate = kMaxSteps;
      }
      if (finished) {
        alive = false;
        pending = true;
      }
    }
    wc.add_step(act, info);
    if (kCount) wc.add(kCntDetected, popc(__ballot(finished && fate == kDetected)));
  }
  __syncthreads();
  flush_block<true, kCount>(out, lds, snb, sns, wc);
}

// ---------------------------------------------------------------------------
// Batched events with the photons' random streams in LDS rings (see
// wavefront_ring_kernel): the slot queue and per-event parameters of
// batch_kernel, the ring streams, generation passes and deferred tallying of
// wavefront_ring_kernel. With kLds the block's event tallies come first in
// LDS, the rings after them.
template <bool kCount, bool kLds>
__global__ __launch_bounds__(1024) void batch_ring_kernel(PhiloxKey key, uint64_t first_id,
                                                          uint32_t steps_per_refill,
                                                          uint32_t chunk, uint64_t n_slots,
                                                          uint32_t per_claim,
                                                          unsigned long long* queue,
                                                          LaunchOut out) {
  extern __shared__ uint32_t lds[];
  // LDS: the ring region first (at address 0, a multiple of its size), then
  // the block's event tallies (kLds)
  uint32_t* const evl = lds + ring_region_words();
  const uint32_t lds_words = kLds ? out.evt.n_events * out.evt.words : 0u;
  if (kLds) {
    zero_lds(evl, lds_words);
    __syncthreads();
  }
  LaunchOut loc = out;
  if (kLds) loc.evt.counts = evl;
  const uint32_t lane = lane_id();
  const unsigned long long* off = out.evt.offsets;
  const unsigned long long* slots = out.evt.slots;
  // every event carries the same table pointer; read it once, as a uniform
  const float* const rtab = out.evt.table[0].rayleigh;
  WaveCounters<kCount> wc;
  wc.clear();
  EventFates fates;
  WaveRow wr;
  wr.on = !kLds && out.evt.wave_rows != 0u;
  wr.row = evl + (threadIdx.x / 64u) * out.evt.words;
  wr.clear(out.evt.words);

  uint32_t we = 0;                 // event of the wave's claimed range (uniform)
  uint64_t we_off = 0;             // its first photon index, off[we] (uniform)
  uint64_t q_next = 0, q_end = 0;  // claimed, not yet started (indices within event we)
  uint64_t s_next = 0, s_end = 0;  // slots claimed by the wave, not yet opened
  bool q_done = false;
  bool alive = false, fresh = false, pending = false;
  uint32_t pset = 0;
  uint32_t done = 0, fate = kMaxSteps;
  DevParams P{};
  RingStream rng;
  rng.place(lds);
  Photon ph;

  for (uint32_t it = 0;; ++it) {
    set_priority_for_long_history(alive && done >= kLongHistory);
    const bool all_dead = __ballot(alive) == 0;
    if (all_dead || it % steps_per_refill == 0) {
      if (pending) {
        finish_event<kLds>(loc.evt, fates, wr, pset, ph, fate);
        pending = false;
      }
      const uint64_t dead = __ballot(!alive);
      if (dead != 0 && !q_done) {
        const uint32_t need = popc(dead);
        const uint64_t have = q_end - q_next;
        const uint32_t t1 = have < need ? static_cast<uint32_t>(have) : need;
        const uint64_t b1 = q_next;
        const uint32_t s1 = we;
        const uint64_t s1_off = we_off;
        q_next += t1;
        uint64_t b2 = 0;
        uint32_t t2 = 0;
        while (t1 < need && !q_done) {  // uniform: take the next non-empty slot
          if (s_next >= s_end) {  // the wave's slot range is used up: claim the next one
            unsigned long long c = 0;
            if (lane == 0) c = atomicAdd(queue, static_cast<unsigned long long>(per_claim));
            c = first_lane_u64(c);
            if (c >= n_slots) {
              q_done = true;
              break;
            }
            s_next = c;
            s_end = (c + per_claim < n_slots) ? c + per_claim : n_slots;
          }
          const uint64_t c = s_next++;
          const unsigned long long sl = slots[c];
          const uint32_t e = static_cast<uint32_t>(sl >> 32);
          const uint64_t start = static_cast<uint64_t>(static_cast<uint32_t>(sl)) * chunk;
          const uint64_t e_off = off[e];
          const uint64_t ne = off[e + 1] - e_off;
          if (start >= ne) continue;
          wr.take(loc.evt, e);
          we = e;
          we_off = e_off;
          q_next = start;
          q_end = (start + chunk < ne) ? start + chunk : ne;
          t2 = (need - t1 < q_end - q_next) ? need - t1 : static_cast<uint32_t>(q_end - q_next);
          b2 = q_next;
          q_next += t2;
          break;  // one chunk per refill
        }
        const uint32_t r = rank_in(dead);
        if (!alive && r < t1 + t2) {
          pset = r < t1 ? s1 : we;
          const uint64_t idx = r < t1 ? b1 + r : b2 + (r - t1);
          P = out.evt.table[pset];
          P.rayleigh = rtab;
          rng.reset(first_id + (r < t1 ? s1_off : we_off) + idx);
          fresh = true;
        }
        wc.add(kCntEmitted, t1 + t2);
      }
      if (all_dead && __ballot(fresh) == 0) {
        if (q_done) break;
        continue;
      }
    }

    const bool drawing = alive || fresh;
    // 0 for a lane that does not draw: its level is never below it, so the
    // loop's ballot is of one comparison
    const uint32_t need = drawing ? (fresh ? (P.source == kVolume ? 6u : 3u) + 5u : 5u) : 0u;
    while (__ballot(rng.level() < need) != 0) {
      if (drawing && rng.room()) rng.generate(key);
    }
    if (fresh) {
      emit(P, rng, ph);
      fresh = false;
      done = 0;
      fate = kMaxSteps;
      alive = P.max_steps > 0u;
      if (!alive) pending = true;
    }

    const uint64_t act = __ballot(alive);
    if (act == 0) continue;
    StepInfo info{};
    bool finished = false;
    if (alive) {
      const bool still = step(P, rng, ph, fate, info);
      ++done;
      if (!still || done >= P.max_steps) {
        if (still) fate = kMaxSteps;
        finished = true;
        alive = false;
      }
    }
    wc.add_step(act, info);
    if (kCount) wc.add(kCntDetected, popc(__ballot(finished && fate == kDetected)));
    if (finished) pending = true;
  }
  fates.flush(loc.evt);
  wr.flush(loc.evt);
  wc.flush(out.counters);
  if (kLds) {
    __syncthreads();
    for (uint32_t w = threadIdx.x; w < lds_words; w += blockDim.x)
      if (evl[w]) atomicAdd(out.evt.counts + w, evl[w]);
  }
}

// ---------------------------------------------------------------------------
// Photons from gensteps. The queue and refill are those of batch_kernel
// (slots claimed `per_claim` per atomic, each slot one chunk of one
// genstep); the geometry and tally are the run's (LDS histograms, as
// wavefront_kernel), and each lane carries its genstep's segment.
template <bool kRec, bool kCount, int kSurf, bool kCommon = false>
__global__ __launch_bounds__(1024) void genstep_kernel(DevParams P0, PhiloxKey key,
                                                       uint64_t first_id,
                                                       uint32_t steps_per_refill, uint32_t chunk,
                                                       uint32_t per_claim, GenstepArgs g,
                                                       unsigned long long* queue,
                                                       LaunchOut out) {
  const DevParams& P = P0;
  if (kCommon) assume_common_case(P0);
  extern __shared__ uint32_t lds[];
  // LDS: the tally, then the Fresnel cell bounds (if kept)
  uint32_t* const fres = out.fresnel_bounds ? lds + ((out.tally.words + 3u) & ~3u) : nullptr;
  zero_lds(lds, out.tally.words);
  if (fres) fill_fresnel_bounds(P, fres, fresnel_tables(kSurf));
  __syncthreads();
  const uint32_t lane = lane_id();
  WaveCounters<kCount> wc;
  wc.clear();
  uint64_t snb = 0, sns = 0;

  uint32_t wg = 0;                 // genstep of the wave's open range (uniform)
  uint64_t q_next = 0, q_end = 0;  // open range, indices within genstep wg
  uint64_t s_next = 0, s_end = 0;  // slots claimed by the wave, not yet opened
  bool q_done = false;
  bool alive = false;
  uint64_t idx = 0;  // global photon index
  uint32_t done = 0, fate = kMaxSteps;
  Stream rng;
  Photon ph;

  for (;;) {
    const uint64_t dead = __ballot(!alive);
    if (dead != 0 && !q_done) {
      const uint32_t need = popc(dead);
      const uint64_t have = q_end - q_next;
      const uint32_t t1 = have < need ? static_cast<uint32_t>(have) : need;
      const uint64_t b1 = q_next;
      const uint32_t g1 = wg;
      q_next += t1;
      uint64_t b2 = 0;
      uint32_t t2 = 0;
      if (t1 < need) {
        if (s_next >= s_end) {
          unsigned long long c = 0;
          if (lane == 0) c = atomicAdd(queue, static_cast<unsigned long long>(per_claim));
          c = first_lane_u64(c);
          if (c >= g.n_slots) {
            q_done = true;
          } else {
            s_next = c;
            s_end = (c + per_claim < g.n_slots) ? c + per_claim : g.n_slots;
          }
        }
        if (!q_done) {
          const uint2 sl = g.slots[s_next++];
          wg = sl.x;
          const uint64_t ng = g.gensteps[wg].n_photons;
          q_next = static_cast<uint64_t>(sl.y) * chunk;
          q_end = (q_next + chunk < ng) ? q_next + chunk : ng;
          t2 = (need - t1 < q_end - q_next) ? need - t1 : static_cast<uint32_t>(q_end - q_next);
          b2 = q_next;
          q_next += t2;
        }
      }
      const uint32_t r = rank_in(dead);
      bool fresh = false;
      if (!alive && r < t1 + t2) {
        const DevGenstep G = g.gensteps[r < t1 ? g1 : wg];
        idx = G.offset + (r < t1 ? b1 + r : b2 + (r - t1));
        rng.start(first_id + idx, key, 0u);
        emit_genstep(P, G, rng, ph);
        done = 0;
        fate = kMaxSteps;
        alive = P.max_steps > 0u;
        fresh = true;
      }
      wc.add(kCntEmitted, t1 + t2);
      if (fresh && !alive) finish<kRec, true>(P, out, lds, idx, ph, fate, snb, sns);
    }
    if (__ballot(alive) == 0) {
      if (q_done) break;
      continue;
    }
    for (uint32_t k = 0; k < steps_per_refill; ++k) {
      const uint64_t act = __ballot(alive);
      if (act == 0) break;
      StepInfo info{};
      bool finished = false;
      if (alive) {
        rng.refill(key);
        const bool still = step<kSurf>(P, rng, ph, fate, info, nullptr, fres);
        ++done;
        if (!still || done >= P.max_steps) {
          if (still) fate = kMaxSteps;
          finished = true;
          alive = false;
        }
      }
      wc.add_step(act, info);
      if (kCount) wc.add(kCntDetected, popc(__ballot(finished && fate == kDetected)));
      if (finished) finish<kRec, true>(P, out, lds, idx, ph, fate, snb, sns);
    }
  }
  __syncthreads();
  flush_block<true, kCount>(out, lds, snb, sns, wc);
}

// Photons from gensteps with the Philox windows in LDS rings (tally): the
// genstep slot queue of genstep_kernel, and the ring streams, generation
// passes and deferred tallying of wavefront_ring_kernel. A lane keeps the
// index of its genstep from the refill point to its emission, which draws
// from the ring like the step.
template <bool kCount, int kSurf, bool kCommon = false>
__global__ __launch_bounds__(1024) void genstep_ring_kernel(DevParams P0, PhiloxKey key,
                                                            uint64_t first_id,
                                                            uint32_t steps_per_refill,
                                                            uint32_t chunk, uint32_t per_claim,
                                                            GenstepArgs g,
                                                            unsigned long long* queue,
                                                            LaunchOut out) {
  const DevParams& P = P0;
  if (kCommon) assume_common_case(P0);
  extern __shared__ uint32_t lds[];
  // LDS: the ring region first (at address 0, a multiple of its size), then
  // the block's tally, then the Fresnel cell bounds if kept
  uint32_t* const tal = lds + ring_region_words();
  uint32_t* const fres = out.fresnel_bounds ? tal + ((out.tally.words + 3u) & ~3u) : nullptr;
  zero_lds(tal, out.tally.words);
  if (fres) fill_fresnel_bounds(P, fres, fresnel_tables(kSurf));
  __syncthreads();
  const uint32_t lane = lane_id();
  WaveCounters<kCount> wc;
  wc.clear();
  uint64_t snb = 0, sns = 0;

  uint32_t wg = 0;                 // genstep of the wave's open range (uniform)
  uint64_t q_next = 0, q_end = 0;  // open range, indices within genstep wg
  uint64_t s_next = 0, s_end = 0;  // slots claimed by the wave, not yet opened
  bool q_done = false;
  bool alive = false, fresh = false, pending = false;
  uint32_t lg = 0;  // the lane's genstep, from its refill to its emission
  uint32_t done = 0, fate = kMaxSteps;
  RingStream rng;
  if (kCommon)  // launched with kCommonRingBlock lanes per block only
    rng.place(lds, kCommonRingShift);
  else
    rng.place(lds);
  Photon ph;
  constexpr uint32_t kEmitWords = 4u;  // u along the segment (if any), cos, phi, delay

  for (uint32_t it = 0;; ++it) {
    // ---- refill point: tally the finished photons, assign ids to free lanes --
    const bool all_dead = __ballot(alive) == 0;
    if (all_dead || it % steps_per_refill == 0) {
      if (pending) {
        finish<false, true>(P, out, tal, 0, ph, fate, snb, sns);
        pending = false;
      }
      const uint64_t dead = __ballot(!alive);
      if (dead != 0 && !q_done) {
        const uint32_t need = popc(dead);
        const uint64_t have = q_end - q_next;
        const uint32_t t1 = have < need ? static_cast<uint32_t>(have) : need;
        const uint64_t b1 = q_next;
        const uint32_t g1 = wg;
        q_next += t1;
        uint64_t b2 = 0;
        uint32_t t2 = 0;
        if (t1 < need) {
          if (s_next >= s_end) {
            unsigned long long c = 0;
            if (lane == 0) c = atomicAdd(queue, static_cast<unsigned long long>(per_claim));
            c = first_lane_u64(c);
            if (c >= g.n_slots) {
              q_done = true;
            } else {
              s_next = c;
              s_end = (c + per_claim < g.n_slots) ? c + per_claim : g.n_slots;
            }
          }
          if (!q_done) {
            const uint2 sl = g.slots[s_next++];
            wg = sl.x;
            const uint64_t ng = g.gensteps[wg].n_photons;
            q_next = static_cast<uint64_t>(sl.y) * chunk;
            q_end = (q_next + chunk < ng) ? q_next + chunk : ng;
            t2 = (need - t1 < q_end - q_next) ? need - t1
                                               : static_cast<uint32_t>(q_end - q_next);
            b2 = q_next;
            q_next += t2;
          }
        }
        const uint32_t r = rank_in(dead);
        if (!alive && r < t1 + t2) {
          lg = r < t1 ? g1 : wg;
          const uint64_t idx = g.gensteps[lg].offset + (r < t1 ? b1 + r : b2 + (r - t1));
          rng.reset(first_id + idx);
          fresh = true;
        }
        wc.add(kCntEmitted, t1 + t2);
      }
      if (all_dead && __ballot(fresh) == 0) {
        if (q_done) break;
        continue;
      }
    }

    // ---- generation passes -------------------------------------------------
    const bool drawing = alive || fresh;
    // 0 for a lane that does not draw: its level is never below it, so the
    // loop's ballot is of one comparison
    const uint32_t need = drawing ? (fresh ? kEmitWords + 5u : 5u) : 0u;
    while (__ballot(rng.level() < need) != 0) {
      if (drawing && rng.room()) rng.generate(key);
    }

    // ---- emission of the newly assigned photons ----------------------------
    if (fresh) {
      const DevGenstep G = g.gensteps[lg];
      emit_genstep(P, G, rng, ph);
      fresh = false;
      done = 0;
      fate = kMaxSteps;
      alive = P.max_steps > 0u;
      if (!alive) pending = true;  // max_steps == 0: finished at emission
    }

    // ---- one transport step --------------------------------------------------
    const uint64_t act = __ballot(alive);
    if (act == 0) continue;
    StepInfo info{};
    bool finished = false;
    if (alive) {
      const bool still = step<kSurf>(P, rng, ph, fate, info, nullptr, fres);
      ++done;
      if (!still || done >= P.max_steps) {
        if (still) fate = kMaxSteps;
        finished = true;
        alive = false;
      }
    }
    wc.add_step(act, info);
    if (kCount) wc.add(kCntDetected, popc(__ballot(finished && fate == kDetected)));
    if (finished) pending = true;
  }
  __syncthreads();
  flush_block<true, kCount>(out, tal, snb, sns, wc);
}

// Column sums of the [blocks][words] partial histograms (and the [blocks][2]
// sums) into the u64 tally, which must be zero on entry. blockIdx.y takes
// every gridDim.y-th block row, so enough waves are in flight to stream the
// partials at HBM rate; each column slice adds its sum with one atomic.
// ---------------------------------------------------------------------------
// Batched events. Lanes take photons of many events and carry their event's
// parameters in registers (loaded when the lane is refilled). The queue hands
// out slots c = 0, 1, ...; slot c is one chunk of one event (out.evt.slots).
// The host lays the slots out so that
//  * the events expected to have the longest histories come first: a
//    persistent launch lasts at least as long as its longest photon history,
//    and that history then runs alongside the bulk instead of after it;
//  * within that order, chunk j of every event precedes chunk j + 1, so the
//    waves in flight at any moment work on many different events and the
//    per-event tally atomics that remain (WaveRow) spread over many rows
//    instead of queueing on one;
//  * events with the same branch structure (surface model, scattering on,
//    absorption on) are adjacent, so the consecutive slots a wave claims
//    mostly share it: branch bodies that none of the wave's events need (the
//    rotation when no event scatters or reflects diffusely) are then skipped.
// A wave that holds a photon more than kLongHistory steps into its history
// raises its issue priority until it no longer does, so that the few long
// histories, which bound the launch, are not slowed by the waves they share a
// SIMD with.
//
// With kLds (few events: n_events * words u32 fit in LDS) each block keeps
// its own copy of all event tallies in LDS and adds it to the global rows
// once at the end; otherwise each wave keeps the row of its current slot's
// event in LDS (WaveRow, when out.evt.wave_rows) and photons of other events
// add to the global rows directly.
template <bool kCount, bool kLds>
__global__ __launch_bounds__(1024) void batch_kernel(PhiloxKey key, uint64_t first_id,
                                                     uint32_t steps_per_refill, uint32_t chunk,
                                                     uint64_t n_slots, uint32_t per_claim,
                                                     unsigned long long* queue, LaunchOut out) {
  extern __shared__ uint32_t lds[];
  const uint32_t lds_words = kLds ? out.evt.n_events * out.evt.words : 0u;
  if (kLds) {
    zero_lds(lds, lds_words);
    __syncthreads();
  }
  LaunchOut loc = out;
  if (kLds) loc.evt.counts = lds;  // finish() adds into the block's copy
  const uint32_t lane = lane_id();
  const unsigned long long* off = out.evt.offsets;
  const unsigned long long* slots = out.evt.slots;
  // every event carries the same table pointer; read it once, as a uniform
  const float* const rtab = out.evt.table[0].rayleigh;
  WaveCounters<kCount> wc;
  wc.clear();
  EventFates fates;
  WaveRow wr;
  wr.on = !kLds && out.evt.wave_rows != 0u;
  wr.row = lds + (threadIdx.x / 64u) * out.evt.words;
  wr.clear(out.evt.words);

  uint32_t we = 0;                 // event of the wave's claimed range (uniform)
  uint64_t we_off = 0;             // its first photon index, off[we] (uniform)
  uint64_t q_next = 0, q_end = 0;  // claimed, not yet started (indices within event we)
  uint64_t s_next = 0, s_end = 0;  // slots claimed by the wave, not yet opened
  bool q_done = false;
  bool alive = false;
  uint32_t pset = 0;
  uint64_t idx = 0;
  uint32_t done = 0, fate = kMaxSteps;
  DevParams P{};
  Stream rng;
  Photon ph;

  for (;;) {
    set_priority_for_long_history(alive && done >= kLongHistory);
    const uint64_t dead = __ballot(!alive);
    if (dead != 0 && !q_done) {
      const uint32_t need = popc(dead);
      const uint64_t have = q_end - q_next;
      const uint32_t t1 = have < need ? static_cast<uint32_t>(have) : need;
      const uint64_t b1 = q_next;
      const uint32_t s1 = we;
      const uint64_t s1_off = we_off;
      q_next += t1;
      uint64_t b2 = 0;
      uint32_t t2 = 0;
      while (t1 < need && !q_done) {  // uniform: take the next non-empty slot
        if (s_next >= s_end) {  // the wave's slot range is used up: claim the next one
          unsigned long long c = 0;
          if (lane == 0) c = atomicAdd(queue, static_cast<unsigned long long>(per_claim));
          c = first_lane_u64(c);
          if (c >= n_slots) {
            q_done = true;
            break;
          }
          s_next = c;
          s_end = (c + per_claim < n_slots) ? c + per_claim : n_slots;
        }
        const uint64_t c = s_next++;
        const unsigned long long sl = slots[c];
        const uint32_t e = static_cast<uint32_t>(sl >> 32);
        const uint64_t start = static_cast<uint64_t>(static_cast<uint32_t>(sl)) * chunk;
        const uint64_t e_off = off[e];
        const uint64_t ne = off[e + 1] - e_off;
        if (start >= ne) continue;
        wr.take(loc.evt, e);
        we = e;
        we_off = e_off;
        q_next = start;
        q_end = (start + chunk < ne) ? start + chunk : ne;
        t2 = (need - t1 < q_end - q_next) ? need - t1 : static_cast<uint32_t>(q_end - q_next);
        b2 = q_next;
        q_next += t2;
        break;  // one chunk per refill
      }
      const uint32_t r = rank_in(dead);
      bool fresh = false;
      if (!alive && r < t1 + t2) {
        pset = r < t1 ? s1 : we;
        idx = r < t1 ? b1 + r : b2 + (r - t1);
        P = out.evt.table[pset];
        P.rayleigh = rtab;
        rng.start(first_id + (r < t1 ? s1_off : we_off) + idx, key, 0u);
        emit(P, rng, ph);
        done = 0;
        fate = kMaxSteps;
        alive = P.max_steps > 0u;
        fresh = true;
      }
      wc.add(kCntEmitted, t1 + t2);
      if (fresh && !alive) finish_event<kLds>(loc.evt, fates, wr, pset, ph, fate);
    }
    if (__ballot(alive) == 0) {
      if (q_done) break;
      continue;
    }
    for (uint32_t k = 0; k < steps_per_refill; ++k) {
      const uint64_t act = __ballot(alive);
      if (act == 0) break;
      StepInfo info{};
      bool finished = false;
      if (alive) {
        rng.refill(key);
        const bool still = step(P, rng, ph, fate, info);
        ++done;
        if (!still || done >= P.max_steps) {
          if (still) fate = kMaxSteps;
          finished = true;
          alive = false;
        }
      }
      wc.add_step(act, info);
      if (kCount) wc.add(kCntDetected, popc(__ballot(finished && fate == kDetected)));
      if (finished) finish_event<kLds>(loc.evt, fates, wr, pset, ph, fate);
    }
  }
  fates.flush(loc.evt);
  wr.flush(loc.evt);
  wc.flush(out.counters);
  if (kLds) {
    __syncthreads();
    for (uint32_t w = threadIdx.x; w < lds_words; w += blockDim.x)
      if (lds[w]) atomicAdd(out.evt.counts + w, lds[w]);
  }
}

// ---------------------------------------------------------------------------
// Adjoint kernel. Recomputes every history from its Philox counter (no tape)
// and accumulates, per detected photon, the weight of its arrival-time bin
// times its likelihood-ratio scores:
//   abs_length:  path / L_a^2 - [absorbed in the bulk] / L_a
//                (exponential free flights competing with the other events)
//   scat_length: path / L_s^2 - n_scatter / L_s
//   reflectivity: (reflector hits - a) / R - a / (1 - R), a = [absorbed there]
//   n_det, n_scint: sum over Fresnel decisions of d log R or d log (1 - R)
// plus, for n_scint, the pathwise time term: t = t_emit + path n_scint / c
// with t_emit ~ Exp(tau), so moving t by dx shifts t_emit, and
//   d/dx P(t in bin k) = (1/tau) (P(t in k) - [x in k]),  x = path n_scint / c,
// which per photon is (path / c / tau) (w[bin(t)] - w[bin(x)]).
// Scores are f32 per photon; the weighted sums are f64.
// Work: per-set queues; a wave stays on one set until it runs out, so a lane
// changes set rarely and flushes its f64 accumulators with few atomics.
//
// kSurf >= 0: every set has that surface model; kUni: a single set whose
// parameters come as the kernel argument P0 (uniform, scalar registers).
template <int kSurf, bool kUni>
__global__ __launch_bounds__(1024) void adjoint_kernel(DevParams P0, PhiloxKey key,
                                                       uint64_t first_id,
                                                       uint32_t steps_per_refill, uint32_t chunk,
                                                       AdjointArgs a) {
  const uint32_t lane = lane_id();
  const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint32_t E = a.n_events;
  uint32_t ws = wave % E;  // the wave's current set (uniform)
  uint32_t visited = 0;
  uint64_t q_next = 0, q_end = 0;
  bool w_done = false;

  bool alive = false;
  uint32_t lset = 0xffffffffu, pset = 0;  // accumulator set, photon's set
  double g0 = 0, g1 = 0, g2 = 0, g3 = 0, g4 = 0;
  DevParams Pv{};
  const DevParams& P = kUni ? P0 : Pv;
  Stream rng;
  Photon ph;
  AdjState adj{};
  uint32_t done = 0, fate = kMaxSteps;

  // Adds the accumulators of the lanes with `want` set to grad[lset], one
  // wave-wide sum and one atomic per distinct set among them (uniform call).
  auto flush = [&](bool want) {
    want = want && lset != 0xffffffffu;
    uint64_t pend = __ballot(want);
    while (pend) {
      const uint32_t leader = static_cast<uint32_t>(__builtin_ctzll(pend));
      const uint32_t s = __builtin_amdgcn_readlane(lset, leader);
      const bool mine = want && lset == s;
      double v[5] = {mine ? g0 : 0.0, mine ? g1 : 0.0, mine ? g2 : 0.0, mine ? g3 : 0.0,
                     mine ? g4 : 0.0};
#pragma unroll
      for (int j = 0; j < 5; ++j) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v[j] += __shfl_xor(v[j], off, 64);
      }
      if (lane == 0) {
#pragma unroll
        for (int j = 0; j < 5; ++j)
          if (v[j] != 0.0) atomicAdd(a.grad + 5u * s + j, v[j]);
      }
      if (mine) {
        g0 = g1 = g2 = g3 = g4 = 0.0;
        want = false;
      }
      pend = __ballot(want);
    }
  };

  for (;;) {
    const uint64_t dead = __ballot(!alive);
    if (dead != 0 && !w_done) {
      const uint32_t need = popc(dead);
      const uint64_t have = q_end - q_next;
      const uint32_t t1 = have < need ? static_cast<uint32_t>(have) : need;
      const uint64_t b1 = q_next;
      const uint32_t s1 = ws;
      q_next += t1;
      uint64_t b2 = 0;
      uint32_t t2 = 0;
      while (t1 < need && !w_done) {  // uniform: claim a chunk, moving on past exhausted sets
        const uint64_t cnt = a.offsets[ws + 1] - a.offsets[ws];
        unsigned long long got = 0;
        if (lane == 0) got = atomicAdd(a.qheads + ws, static_cast<unsigned long long>(chunk));
        got = first_lane_u64(got);
        if (got < cnt) {
          if (got + chunk >= cnt && lane == 0) atomicAdd(a.closed, 1ull);
          q_next = got;
          q_end = (got + chunk < cnt) ? got + chunk : cnt;
          t2 = (need - t1 < q_end - q_next) ? need - t1 : static_cast<uint32_t>(q_end - q_next);
          b2 = q_next;
          q_next += t2;
          break;
        }
        ++visited;
        unsigned long long closed = 0;
        if (lane == 0) closed = __atomic_load_n(a.closed, __ATOMIC_RELAXED);
        closed = first_lane_u64(closed);
        if (visited >= E || closed >= E) {
          w_done = true;
        } else {
          ws = (ws + 1u == E) ? 0u : ws + 1u;
          q_next = q_end = 0;
        }
      }
      const uint32_t r = rank_in(dead);
      const bool gets = !alive && r < t1 + t2;
      const uint32_t nset = r < t1 ? s1 : ws;
      // A lane moving to another set hands in its sums for the old one.
      flush(gets && nset != lset);
      if (gets && nset != lset) lset = nset;
      if (gets) {
        pset = nset;
        const uint64_t local = r < t1 ? b1 + r : b2 + (r - t1);
        if (!kUni) {
          Pv = a.table[pset];
          Pv.rayleigh = P0.rayleigh;  // the same for every set; uniform
        }
        rng.start(first_id + a.offsets[pset] + local, key, 0u);
        emit(P, rng, ph);
        done = 0;
        fate = kMaxSteps;
        adj = AdjState{0.0f, 0.0f, 0u, 0.0, 0.0,
                       a.weights + static_cast<size_t>(pset) * (a.time_bins + 1u), a.time_scale,
                       a.time_bins};
        alive = P.max_steps > 0u;
      }
    }
    if (__ballot(alive) == 0) {
      if (w_done) break;
      continue;
    }
    for (uint32_t k = 0; k < steps_per_refill; ++k) {
      if (__ballot(alive) == 0) break;
      bool finished = false;
      if (alive) {
        StepInfo info{};
        rng.refill(key);
        const bool still = step<kSurf, true>(P, rng, ph, fate, info, &adj);
        ++done;
        if (!still || done >= P.max_steps) {
          if (still) fate = kMaxSteps;
          finished = true;
          alive = false;
        }
      }
      if (finished) {
        // Readout baselines (see AdjState) apply to every finished photon;
        // the other terms only to detected ones (undetected photons score 0).
        g3 -= adj.base_ndet;
        g4 -= adj.base_nscint;
      }
      if (finished && fate == kDetected) {
        const float inf = __builtin_huge_valf();
        float s0 = 0.0f, s1f = 0.0f, s2 = 0.0f;
        if (P.abs_length != inf) s0 = ph.path / (P.abs_length * P.abs_length);
        if (P.scat_length != inf)
          s1f = ph.path / (P.scat_length * P.scat_length) -
                static_cast<float>(ph.ns) / P.scat_length;
        const uint32_t surf = kSurf >= 0 ? static_cast<uint32_t>(kSurf) : P.surface;
        if (surf >= kSpecular) s2 = static_cast<float>(adj.refl_hits) / P.reflectivity;
        const uint32_t bt = time_bin(ph.t, a.time_scale, a.time_bins);
        const double* w = adj.w;
        const double wt = w[bt];
        double tterm = 0.0;
        if (P.tau > 0.0f) {
          const float x = ph.path * P.inv_speed;
          const uint32_t bx = time_bin(x, a.time_scale, a.time_bins);
          tterm = (wt - w[bx]) * static_cast<double>(ph.path / 299.792458f / P.tau);
        }
        g0 += wt * static_cast<double>(s0);
        g1 += wt * static_cast<double>(s1f);
        g2 += wt * static_cast<double>(s2);
        g3 += wt * static_cast<double>(adj.s_ndet);
        g4 += wt * static_cast<double>(adj.s_nscint) + tterm;
      }
    }
  }
  flush(true);
}

__global__ void tally_reduce_kernel(uint32_t words, int blocks, const uint32_t* partials,
                                    const unsigned long long* sums, unsigned long long* tally) {
  const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long s = 0;
  if (w < words) {
    for (int b = blockIdx.y; b < blocks; b += gridDim.y)
      s += partials[static_cast<size_t>(b) * words + w];
  } else if (w < words + 2u) {
    const uint32_t j = w - words;
    for (int b = blockIdx.y; b < blocks; b += gridDim.y) s += sums[2 * b + j];
  }
  if (w < words + 2u && s != 0) atomicAdd(tally + w, s);
}

// Entry k = rayleigh_cos of u = (k + 1) 2^-24, the uniform of every word whose
// top 24 bits are k (philox.hpp:82-84).
__global__ void rayleigh_table_kernel(float* table) {
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < (1u << 24); k += stride)
    table[k] = rayleigh_cos(to_unit(k << 8));
}

// Clears two ranges in one launch (a tally run's scratch and its tally).
__global__ void zero_u64_pair_kernel(unsigned long long* a, uint64_t na, unsigned long long* b,
                                     uint64_t nb) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < na + nb;
       i += stride) {
    if (i < na)
      a[i] = 0ull;
    else
      b[i - na] = 0ull;
  }
}

__global__ void zero_u64_kernel(unsigned long long* p, uint64_t n) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += stride)
    p[i] = 0ull;
}

// Chi-square loss of each parameter set's arrival-time histogram against its
// target, and the loss gradient that weights the adjoint (one block per set).
// h_k = counts_k / N (detected photons per bin, overflow included), f = sum h_k,
// F = sum T_k; with 1/n_t the target's per-photon regulariser:
//   loss = sum_k (h_k - T_k)^2 w_k + (f - F)^2 w_f,
//   w_k = 1 / (T_k + 1/n_t),  w_f = 1 / (F (1 - F) + 1/n_t),
//   dloss/dh_k = 2 (h_k - T_k) w_k + 2 (f - F) w_f.
// All in FP64 (the cross-photon reduction).
__global__ __launch_bounds__(256) void chi2_weights_kernel(
    const uint32_t* counts, uint32_t words, uint32_t bins1, const unsigned long long* offsets,
    const double* target, const double* inv_nt, double* hist, double* loss, double* weights) {
  const uint32_t e = blockIdx.x;
  const uint32_t* row = counts + static_cast<size_t>(e) * words + 5u;  // past the fate counts
  const double n = static_cast<double>(offsets[e + 1] - offsets[e]);
  const double* T = target + static_cast<size_t>(e) * bins1;
  __shared__ double red_f[256], red_F[256], red_l[256];
  double f = 0.0, F = 0.0, l = 0.0;
  for (uint32_t k = threadIdx.x; k < bins1; k += blockDim.x) {
    const double h = n > 0.0 ? static_cast<double>(row[k]) / n : 0.0;
    hist[static_cast<size_t>(e) * bins1 + k] = h;
    const double d = h - T[k];
    f += h;
    F += T[k];
    l += d * d * (1.0 / (T[k] + inv_nt[e]));
  }
  red_f[threadIdx.x] = f;
  red_F[threadIdx.x] = F;
  red_l[threadIdx.x] = l;
  __syncthreads();
  for (uint32_t w = blockDim.x / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      red_f[threadIdx.x] += red_f[threadIdx.x + w];
      red_F[threadIdx.x] += red_F[threadIdx.x + w];
      red_l[threadIdx.x] += red_l[threadIdx.x + w];
    }
    __syncthreads();
  }
  const double ft = red_f[0], Ft = red_F[0];
  const double wf = 1.0 / (Ft * (1.0 - Ft) + inv_nt[e]);
  const double df = ft - Ft;
  if (threadIdx.x == 0) loss[e] = red_l[0] + df * df * wf;
  for (uint32_t k = threadIdx.x; k < bins1; k += blockDim.x) {
    const double h = hist[static_cast<size_t>(e) * bins1 + k];
    weights[static_cast<size_t>(e) * bins1 + k] =
        2.0 * (h - T[k]) * (1.0 / (T[k] + inv_nt[e])) + 2.0 * df * wf;
  }
}

template <bool R, bool T, bool E, bool C>
void launch_history_t(const DevParams& P, PhiloxKey key, uint64_t first_id, uint64_t n, int block,
                      int grid, const LaunchOut& out, hipStream_t stream) {
  const size_t lds = history_lds_bytes(T ? out.tally.words : 0u, !E && out.fresnel_bounds,
                                       static_cast<int>(P.surface));
  if (!E && common_case(P))
    hipLaunchKernelGGL((history_kernel<R, T, E, C, true>), dim3(grid), dim3(block), lds, stream,
                       P, key, first_id, n, out);
  else
    hipLaunchKernelGGL((history_kernel<R, T, E, C>), dim3(grid), dim3(block), lds, stream, P, key,
                       first_id, n, out);
}

template <bool C, int S>
void launch_wavefront_ring_t(const DevParams& P, PhiloxKey key, uint64_t first_id, uint64_t n,
                             const WavefrontShape& s, unsigned long long* queue,
                             const LaunchOut& out, hipStream_t stream) {
  const size_t lds = wavefront_ring_lds_bytes(out.tally.words, s.block, S, out.fresnel_bounds);
  if (common_case(P) && s.block == kCommonRingBlock)
    hipLaunchKernelGGL((wavefront_ring_kernel<C, S, true>), dim3(s.grid), dim3(s.block), lds,
                       stream, P, key, first_id, n, s.steps_per_refill, s.queue_chunk, queue, out);
  else
    hipLaunchKernelGGL((wavefront_ring_kernel<C, S, false>), dim3(s.grid), dim3(s.block), lds,
                       stream, P, key, first_id, n, s.steps_per_refill, s.queue_chunk, queue, out);
}

template <bool C, int S>
void launch_wavefront_phase_t(const DevParams& P, PhiloxKey key, uint64_t first_id, uint64_t n,
                              const WavefrontShape& s, unsigned long long* queue,
                              const LaunchOut& out, hipStream_t stream) {
  hipLaunchKernelGGL((wavefront_phase_kernel<C, S>), dim3(s.grid), dim3(s.block),
                     phase_lds_bytes(out.tally.words, s.block), stream, P, key, first_id, n,
                     s.steps_per_refill, s.queue_chunk, queue, out);
}

template <bool C, int S>
void launch_wavefront_next_t(const DevParams& P, PhiloxKey key, uint64_t first_id, uint64_t n,
                             const WavefrontShape& s, unsigned long long* queue,
                             const LaunchOut& out, hipStream_t stream) {
  const size_t lds = static_cast<size_t>(out.tally.words) * sizeof(uint32_t);
  hipLaunchKernelGGL((wavefront_next_kernel<C, S>), dim3(s.grid), dim3(s.block), lds, stream, P,
                     key, first_id, n, s.successor_threshold, s.queue_chunk, queue, out);
}

template <bool R, bool T, bool C, int S = -1>
void launch_wavefront_t(const DevParams& P, PhiloxKey key, uint64_t first_id, uint64_t n,
                        const WavefrontShape& s, unsigned long long* queue, const LaunchOut& out,
                        hipStream_t stream) {
  const size_t lds = history_lds_bytes(T ? out.tally.words : 0u, out.fresnel_bounds,
                                       S >= 0 ? S : static_cast<int>(P.surface));
  if (common_case(P))
    hipLaunchKernelGGL((wavefront_kernel<R, T, C, S, true>), dim3(s.grid), dim3(s.block), lds,
                       stream, P, key, first_id, n, s.steps_per_refill, s.queue_chunk, queue, out);
  else
    hipLaunchKernelGGL((wavefront_kernel<R, T, C, S>), dim3(s.grid), dim3(s.block), lds, stream,
                       P, key, first_id, n, s.steps_per_refill, s.queue_chunk, queue, out);
}

// Tally launches specialised on the run's surface model.
template <bool C>
void launch_wavefront_tally(const DevParams& P, PhiloxKey key, uint64_t first_id, uint64_t n,
                            const WavefrontShape& s, unsigned long long* queue,
                            const LaunchOut& out, hipStream_t stream) {
  if (s.sorted_phases) {
    switch (P.surface) {
      case kPolished:
        launch_wavefront_phase_t<C, kPolished>(P, key, first_id, n, s, queue, out, stream);
        break;
      case kBlack:
        launch_wavefront_phase_t<C, kBlack>(P, key, first_id, n, s, queue, out, stream);
        break;
      case kSpecular:
        launch_wavefront_phase_t<C, kSpecular>(P, key, first_id, n, s, queue, out, stream);
        break;
      default:
        launch_wavefront_phase_t<C, kLambertian>(P, key, first_id, n, s, queue, out, stream);
        break;
    }
    return;
  }
  if (s.stream_in_lds) {
    switch (P.surface) {
      case kPolished:
        launch_wavefront_ring_t<C, kPolished>(P, key, first_id, n, s, queue, out, stream);
        break;
      case kBlack:
        launch_wavefront_ring_t<C, kBlack>(P, key, first_id, n, s, queue, out, stream);
        break;
      case kSpecular:
        launch_wavefront_ring_t<C, kSpecular>(P, key, first_id, n, s, queue, out, stream);
        break;
      default:
        launch_wavefront_ring_t<C, kLambertian>(P, key, first_id, n, s, queue, out, stream);
        break;
    }
    return;
  }
  if (s.successor_threshold > 0) {
    switch (P.surface) {
      case kPolished:
        launch_wavefront_next_t<C, kPolished>(P, key, first_id, n, s, queue, out, stream);
        break;
      case kBlack:
        launch_wavefront_next_t<C, kBlack>(P, key, first_id, n, s, queue, out, stream);
        break;
      case kSpecular:
        launch_wavefront_next_t<C, kSpecular>(P, key, first_id, n, s, queue, out, stream);
        break;
      default:
        launch_wavefront_next_t<C, kLambertian>(P, key, first_id, n, s, queue, out, stream);
        break;
    }
    return;
  }
  switch (P.surface) {
    case kPolished:
      launch_wavefront_t<false, true, C, kPolished>(P, key, first_id, n, s, queue, out,
                                                           stream);
      break;
    case kBlack:
      launch_wavefront_t<false, true, C, kBlack>(P, key, first_id, n, s, queue, out, stream);
      break;
    case kSpecular:
      launch_wavefront_t<false, true, C, kSpecular>(P, key, first_id, n, s, queue, out,
                                                           stream);
      break;
    default:
      launch_wavefront_t<false, true, C, kLambertian>(P, key, first_id, n, s, queue, out,
                                                             stream);
      break;
  }
}

template <class K>
int blocks_per_cu(K kernel, int block, size_t lds) {
  int b = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, kernel, block, lds) != hipSuccess) return 0;
  return b;
}

}  // namespace

size_t ring_lds_bytes(uint32_t tally_words, int block) {
  return (((static_cast<size_t>(tally_words) + 3u) & ~size_t{3}) + ring_words(block)) *
         sizeof(uint32_t);
}

size_t history_lds_bytes(uint32_t tally_words, bool bounds, int surface) {
  return ((static_cast<size_t>(tally_words) + 3u) & ~size_t{3}) * sizeof(uint32_t) +
         fresnel_lds_bytes(bounds, surface);
}

size_t wavefront_ring_lds_bytes(uint32_t tally_words, int block, int surface, bool bounds) {
  return ring_lds_bytes(tally_words, block) + fresnel_lds_bytes(bounds, surface);
}

bool ring_kernels_start_at_lds_zero() {
  const void* kernels[] = {
      reinterpret_cast<const void*>(&wavefront_ring_kernel<false, kPolished, false>),
      reinterpret_cast<const void*>(&wavefront_ring_kernel<false, kBlack, false>),
      reinterpret_cast<const void*>(&wavefront_ring_kernel<false, kSpecular, false>),
      reinterpret_cast<const void*>(&wavefront_ring_kernel<false, kLambertian, false>),
      reinterpret_cast<const void*>(&wavefront_ring_kernel<true, kPolished, false>),
      reinterpret_cast<const void*>(&wavefront_ring_kernel<true, kBlack, false>),
      reinterpret_cast<const void*>(&wavefront_ring_kernel<true, kSpecular, false>),
      reinterpret_cast<const void*>(&wavefront_ring_kernel<true, kLambertian, false>),
      reinterpret_cast<const void*>(&wavefront_ring_kernel<false, kPolished, true>),
      reinterpret_cast<const void*>(&wavefront_ring_kernel<false, kBlack, true>),
      reinterpret_cast<const void*>(&wavefront_ring_kernel<false, kSpecular, true>),
      reinterpret_cast<const void*>(&wavefront_ring_kernel<false, kLambertian, true>),
      reinterpret_cast<const void*>(&wavefront_ring_kernel<true, kPolished, true>),
      reinterpret_cast<const void*>(&wavefront_ring_kernel<true, kBlack, true>),
      reinterpret_cast<const void*>(&wavefront_ring_kernel<true, kSpecular, true>),
      reinterpret_cast<const void*>(&wavefront_ring_kernel<true, kLambertian, true>),
      reinterpret_cast<const void*>(&batch_ring_kernel<false, false>),
      reinterpret_cast<const void*>(&batch_ring_kernel<false, true>),
      reinterpret_cast<const void*>(&batch_ring_kernel<true, false>),
      reinterpret_cast<const void*>(&batch_ring_kernel<true, true>),
      reinterpret_cast<const void*>(&genstep_ring_kernel<false, kPolished>),
      reinterpret_cast<const void*>(&genstep_ring_kernel<false, kBlack>),
      reinterpret_cast<const void*>(&genstep_ring_kernel<false, kSpecular>),
      reinterpret_cast<const void*>(&genstep_ring_kernel<false, kLambertian>),
      reinterpret_cast<const void*>(&genstep_ring_kernel<true, kPolished>),
      reinterpret_cast<const void*>(&genstep_ring_kernel<true, kBlack>),
      reinterpret_cast<const void*>(&genstep_ring_kernel<true, kSpecular>),
      reinterpret_cast<const void*>(&genstep_ring_kernel<true, kLambertian>),
      reinterpret_cast<const void*>(&genstep_ring_kernel<false, kPolished, true>),
      reinterpret_cast<const void*>(&genstep_ring_kernel<false, kBlack, true>),
      reinterpret_cast<const void*>(&genstep_ring_kernel<false, kSpecular, true>),
      reinterpret_cast<const void*>(&genstep_ring_kernel<false, kLambertian, true>),
      reinterpret_cast<const void*>(&genstep_ring_kernel<true, kPolished, true>),
      reinterpret_cast<const void*>(&genstep_ring_kernel<true, kBlack, true>),
      reinterpret_cast<const void*>(&genstep_ring_kernel<true, kSpecular, true>),
      reinterpret_cast<const void*>(&genstep_ring_kernel<true, kLambertian, true>),
  };
  for (const void* k : kernels) {
    hipFuncAttributes a{};
    if (hipFuncGetAttributes(&a, k) != hipSuccess || a.sharedSizeBytes != 0) return false;
  }
  return true;
}

size_t phase_lds_bytes(uint32_t tally_words, int block) {
  return ring_lds_bytes(tally_words, block) + (8u + 9u * static_cast<size_t>(block)) * 4u;
}

// Instantiated variants: records (no counters), tally with / without
// counters, and the plain per-event launches of a batch.
void launch_history(const DevParams& P, PhiloxKey key, uint64_t first_id, uint64_t n, int block,
                    int grid, const LaunchOut& out, hipStream_t stream) {
  const bool count = out.counters != nullptr;
  if (out.records && !out.tally_on)
    launch_history_t<true, false, false, false>(P, key, first_id, n, block, grid, out, stream);
  else if (out.tally_on && !out.records && count)
    launch_history_t<false, true, false, true>(P, key, first_id, n, block, grid, out, stream);
  else if (out.tally_on && !out.records)
    launch_history_t<false, true, false, false>(P, key, first_id, n, block, grid, out, stream);
  else
    throw std::invalid_argument("launch_history: unsupported output combination");
}

void launch_wavefront(const DevParams& P, PhiloxKey key, uint64_t first_id, uint64_t n,
                      const WavefrontShape& shape, unsigned long long* queue, const LaunchOut& out,
                      hipStream_t stream) {
  const bool count = out.counters != nullptr;
  if (out.records && !out.tally_on)
    launch_wavefront_t<true, false, true>(P, key, first_id, n, shape, queue, out, stream);
  else if (out.tally_on && !out.records && count)
    launch_wavefront_tally<true>(P, key, first_id, n, shape, queue, out, stream);
  else if (out.tally_on && !out.records)
    launch_wavefront_tally<false>(P, key, first_id, n, shape, queue, out, stream);
  else
    throw std::invalid_argument("launch_wavefront: unsupported output combination");
}

void launch_batch_wavefront(PhiloxKey key, uint64_t first_id, uint64_t n_slots,
                            const WavefrontShape& s, unsigned long long* queue,
                            const LaunchOut& out, hipStream_t stream) {
  const size_t lds = batch_lds_bytes(out.evt.n_events, out.evt.words);
  // without the block copies of all rows: the waves' rows, if the host chose them
  const size_t rows = !lds && out.evt.wave_rows ? batch_wave_row_bytes(out.evt.words, s.block) : 0;
  if (!lds && out.evt.wave_rows && rows == 0)
    throw std::invalid_argument("launch_batch_wavefront: wave rows exceed the LDS budget");
  if (s.stream_in_lds) {
    const size_t bytes = batch_ring_lds_bytes(lds ? lds : rows, s.block);
    if (lds && out.counters)
      hipLaunchKernelGGL((batch_ring_kernel<true, true>), dim3(s.grid), dim3(s.block), bytes,
                         stream, key, first_id, s.steps_per_refill, s.queue_chunk, n_slots,
                         s.slots_per_claim, queue, out);
    else if (lds)
      hipLaunchKernelGGL((batch_ring_kernel<false, true>), dim3(s.grid), dim3(s.block), bytes,
                         stream, key, first_id, s.steps_per_refill, s.queue_chunk, n_slots,
                         s.slots_per_claim, queue, out);
    else if (out.counters)
      hipLaunchKernelGGL((batch_ring_kernel<true, false>), dim3(s.grid), dim3(s.block), bytes,
                         stream, key, first_id, s.steps_per_refill, s.queue_chunk, n_slots,
                         s.slots_per_claim, queue, out);
    else
      hipLaunchKernelGGL((batch_ring_kernel<false, false>), dim3(s.grid), dim3(s.block), bytes,
                         stream, key, first_id, s.steps_per_refill, s.queue_chunk, n_slots,
                         s.slots_per_claim, queue, out);
    return;
  }
  if (lds && out.counters)
    hipLaunchKernelGGL((batch_kernel<true, true>), dim3(s.grid), dim3(s.block), lds, stream, key,
                       first_id, s.steps_per_refill, s.queue_chunk, n_slots, s.slots_per_claim,
                       queue, out);
  else if (lds)
    hipLaunchKernelGGL((batch_kernel<false, true>), dim3(s.grid), dim3(s.block), lds, stream,
                       key, first_id, s.steps_per_refill, s.queue_chunk, n_slots,
                       s.slots_per_claim, queue, out);
  else if (out.counters)
    hipLaunchKernelGGL((batch_kernel<true, false>), dim3(s.grid), dim3(s.block), rows, stream,
                       key, first_id, s.steps_per_refill, s.queue_chunk, n_slots,
                       s.slots_per_claim, queue, out);
  else
    hipLaunchKernelGGL((batch_kernel<false, false>), dim3(s.grid), dim3(s.block), rows, stream,
                       key, first_id, s.steps_per_refill, s.queue_chunk, n_slots,
                       s.slots_per_claim, queue, out);
}

void launch_history_event(const DevParams& P, PhiloxKey key, uint64_t first_id, uint64_t n,
                          int block, const LaunchOut& out, hipStream_t stream) {
  const int grid = static_cast<int>((n + block - 1) / block);
  if (grid > 0)
    launch_history_t<false, false, true, false>(P, key, first_id, n, block, grid, out, stream);
}

void launch_tally_reduce(const DevTally& t, int blocks, const uint32_t* partials,
                         const unsigned long long* sums, unsigned long long* tally,
                         hipStream_t stream) {
  const uint32_t total = t.words + 2u;
  const int threads = 256;
  const int slices = blocks < 64 ? blocks : 64;
  hipLaunchKernelGGL(tally_reduce_kernel, dim3((total + threads - 1) / threads, slices),
                     dim3(threads), 0, stream, t.words, blocks, partials, sums, tally);
}

void launch_chi2_weights(const uint32_t* counts, uint32_t words, uint32_t bins1,
                         const unsigned long long* offsets, const double* target,
                         const double* inv_nt, double* hist, double* loss, double* weights,
                         uint32_t n_sets, hipStream_t stream) {
  hipLaunchKernelGGL(chi2_weights_kernel, dim3(n_sets), dim3(256), 0, stream, counts, words,
                     bins1, offsets, target, inv_nt, hist, loss, weights);
}

template <bool R, bool C, int S>
void launch_gensteps_t(const DevParams& P, PhiloxKey key, uint64_t first_id,
                       const WavefrontShape& s, const GenstepArgs& g, unsigned long long* queue,
                       const LaunchOut& out, hipStream_t stream) {
  const size_t lds = history_lds_bytes(out.tally.words, out.fresnel_bounds, S);
  if (common_case(P))
    hipLaunchKernelGGL((genstep_kernel<R, C, S, true>), dim3(s.grid), dim3(s.block), lds, stream,
                       P, key, first_id, s.steps_per_refill, s.queue_chunk, s.slots_per_claim, g,
                       queue, out);
  else
    hipLaunchKernelGGL((genstep_kernel<R, C, S>), dim3(s.grid), dim3(s.block), lds, stream, P,
                       key, first_id, s.steps_per_refill, s.queue_chunk, s.slots_per_claim, g,
                       queue, out);
}

template <bool C, int S>
void launch_gensteps_ring_t(const DevParams& P, PhiloxKey key, uint64_t first_id,
                            const WavefrontShape& s, const GenstepArgs& g,
                            unsigned long long* queue, const LaunchOut& out,
                            hipStream_t stream) {
  const size_t lds = wavefront_ring_lds_bytes(out.tally.words, s.block, S, out.fresnel_bounds);
  if (common_case(P) && s.block == kCommonRingBlock)
    hipLaunchKernelGGL((genstep_ring_kernel<C, S, true>), dim3(s.grid), dim3(s.block), lds,
                       stream, P, key, first_id, s.steps_per_refill, s.queue_chunk,
                       s.slots_per_claim, g, queue, out);
  else
    hipLaunchKernelGGL((genstep_ring_kernel<C, S>), dim3(s.grid), dim3(s.block), lds, stream, P,
                       key, first_id, s.steps_per_refill, s.queue_chunk, s.slots_per_claim, g,
                       queue, out);
}

template <bool R, bool C>
void launch_gensteps_surf(const DevParams& P, PhiloxKey key, uint64_t first_id,
                          const WavefrontShape& s, const GenstepArgs& g,
                          unsigned long long* queue, const LaunchOut& out, hipStream_t stream) {
  if (!R && s.stream_in_lds) {
    switch (P.surface) {
      case kPolished: launch_gensteps_ring_t<C, kPolished>(P, key, first_id, s, g, queue, out, stream); break;
      case kBlack: launch_gensteps_ring_t<C, kBlack>(P, key, first_id, s, g, queue, out, stream); break;
      case kSpecular: launch_gensteps_ring_t<C, kSpecular>(P, key, first_id, s, g, queue, out, stream); break;
      default: launch_gensteps_ring_t<C, kLambertian>(P, key, first_id, s, g, queue, out, stream); break;
    }
    return;
  }
  switch (P.surface) {
    case kPolished: launch_gensteps_t<R, C, kPolished>(P, key, first_id, s, g, queue, out, stream); break;
    case kBlack: launch_gensteps_t<R, C, kBlack>(P, key, first_id, s, g, queue, out, stream); break;
    case kSpecular: launch_gensteps_t<R, C, kSpecular>(P, key, first_id, s, g, queue, out, stream); break;
    default: launch_gensteps_t<R, C, kLambertian>(P, key, first_id, s, g, queue, out, stream); break;
  }
}

void launch_rayleigh_table(float* table, hipStream_t stream) {
  hipLaunchKernelGGL(rayleigh_table_kernel, dim3(4096), dim3(256), 0, stream, table);
}

void launch_zero_u64_pair(unsigned long long* a, uint64_t na, unsigned long long* b, uint64_t nb,
                          hipStream_t stream) {
  if (na + nb == 0) return;
  const int threads = 256;
  const uint64_t want = (na + nb + threads - 1) / threads;
  const int grid = static_cast<int>(want < 1024 ? want : 1024);
  hipLaunchKernelGGL(zero_u64_pair_kernel, dim3(grid), dim3(threads), 0, stream, a, na, b, nb);
}

void launch_zero_u64(unsigned long long* p, uint64_t n, hipStream_t stream) {
  if (n == 0) return;
  const int threads = 256;
  const uint64_t want = (n + threads - 1) / threads;
  const int grid = static_cast<int>(want < 1024 ? want : 1024);
  hipLaunchKernelGGL(zero_u64_kernel, dim3(grid), dim3(threads), 0, stream, p, n);
}

int history_blocks_per_cu(bool tally, int block, size_t lds) {
  return tally ? blocks_per_cu(history_kernel<false, true, false, false>, block, lds)
               : blocks_per_cu(history_kernel<true, false, false, false>, block, lds);
}

int wavefront_blocks_per_cu(bool records, bool tally, int block, size_t lds, bool successor,
                            bool ring, bool phases, int surface, bool bounds) {
  const size_t all =
      history_lds_bytes(static_cast<uint32_t>(lds / sizeof(uint32_t)), bounds, surface);
  if (records) return blocks_per_cu(wavefront_kernel<true, false, true>, block, all);
  if (phases)
    return blocks_per_cu(wavefront_phase_kernel<false, kLambertian>, block,
                         phase_lds_bytes(static_cast<uint32_t>(lds / sizeof(uint32_t)), block));
  if (ring)
    return blocks_per_cu(wavefront_ring_kernel<false, kLambertian, true>, block,
                         wavefront_ring_lds_bytes(static_cast<uint32_t>(lds / sizeof(uint32_t)),
                                                  block, surface, bounds));
  if (successor) return blocks_per_cu(wavefront_next_kernel<false, kLambertian>, block, lds);
  return blocks_per_cu(wavefront_kernel<false, true, false, kLambertian>, block, all);
}

template <int S, bool U>
void launch_adjoint_t(const DevParams& P0, PhiloxKey key, uint64_t first_id,
                      const WavefrontShape& s, const AdjointArgs& a, hipStream_t stream) {
  hipLaunchKernelGGL((adjoint_kernel<S, U>), dim3(s.grid), dim3(s.block), 0, stream, P0, key,
                     first_id, s.steps_per_refill, s.queue_chunk, a);
}

template <bool U>
void launch_adjoint_surf(int surface, const DevParams& P0, PhiloxKey key, uint64_t first_id,
                         const WavefrontShape& s, const AdjointArgs& a, hipStream_t stream) {
  switch (surface) {
    case kPolished: launch_adjoint_t<kPolished, U>(P0, key, first_id, s, a, stream); break;
    case kBlack: launch_adjoint_t<kBlack, U>(P0, key, first_id, s, a, stream); break;
    case kSpecular: launch_adjoint_t<kSpecular, U>(P0, key, first_id, s, a, stream); break;
    case kLambertian: launch_adjoint_t<kLambertian, U>(P0, key, first_id, s, a, stream); break;
    default: launch_adjoint_t<-1, U>(P0, key, first_id, s, a, stream); break;
  }
}

void launch_adjoint(const DevParams& P0, int surface, PhiloxKey key, uint64_t first_id,
                    const WavefrontShape& s, const AdjointArgs& a, hipStream_t stream,
                    bool params_by_value) {
  if (a.n_events == 1 && params_by_value)
    launch_adjoint_surf<true>(surface, P0, key, first_id, s, a, stream);
  else
    launch_adjoint_surf<false>(surface, P0, key, first_id, s, a, stream);
}

int adjoint_blocks_per_cu(int block) { return blocks_per_cu(adjoint_kernel<-1, false>, block, 0); }

size_t batch_lds_bytes(uint32_t n_events, uint32_t words) {
  const size_t b = static_cast<size_t>(n_events) * words * sizeof(uint32_t);
  return b <= kBatchLdsMax ? b : 0;
}

size_t batch_ring_lds_bytes(size_t tally_lds, int block) {
  return ring_words(block) * sizeof(uint32_t) + tally_lds;
}

size_t batch_wave_row_bytes(uint32_t words, int block) {
  const size_t b = static_cast<size_t>(block / 64) * words * sizeof(uint32_t);
  return b <= kBatchLdsMax ? b : 0;
}

size_t batch_tally_lds_bytes(uint32_t n_events, uint32_t words, int block) {
  const size_t all = batch_lds_bytes(n_events, words);
  return all ? all : batch_wave_row_bytes(words, block);
}

void launch_gensteps(const DevParams& P, PhiloxKey key, uint64_t first_id,
                     const WavefrontShape& shape, const GenstepArgs& g,
                     unsigned long long* queue, const LaunchOut& out, hipStream_t stream) {
  const bool count = out.counters != nullptr;
  if (out.records && count)
    launch_gensteps_surf<true, true>(P, key, first_id, shape, g, queue, out, stream);
  else if (out.records)
    launch_gensteps_surf<true, false>(P, key, first_id, shape, g, queue, out, stream);
  else if (count)
    launch_gensteps_surf<false, true>(P, key, first_id, shape, g, queue, out, stream);
  else
    launch_gensteps_surf<false, false>(P, key, first_id, shape, g, queue, out, stream);
}

int genstep_blocks_per_cu(bool records, int block, size_t lds, bool ring, int surface,
                          bool bounds) {
  const uint32_t words = static_cast<uint32_t>(lds / sizeof(uint32_t));
  if (!records && ring)
    return blocks_per_cu(genstep_ring_kernel<false, kLambertian>, block,
                         wavefront_ring_lds_bytes(words, block, surface, bounds));
  const size_t all = history_lds_bytes(words, bounds, surface);
  return records ? blocks_per_cu(genstep_kernel<true, false, kLambertian>, block, all)
                 : blocks_per_cu(genstep_kernel<false, false, kLambertian>, block, all);
}

int batch_blocks_per_cu(int block, uint32_t n_events, uint32_t words, bool ring) {
  const bool all = batch_lds_bytes(n_events, words) != 0;
  const size_t lds = batch_tally_lds_bytes(n_events, words, block);
  if (ring)
    return all ? blocks_per_cu(batch_ring_kernel<false, true>, block,
                               batch_ring_lds_bytes(lds, block))
               : blocks_per_cu(batch_ring_kernel<false, false>, block,
                               batch_ring_lds_bytes(lds, block));
  return all ? blocks_per_cu(batch_kernel<false, true>, block, lds)
             : blocks_per_cu(batch_kernel<false, false>, block, lds);
}

}  // namespace ophip
