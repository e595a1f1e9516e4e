// sorted_kernels.hip -- per-interaction transport kernels with the photons
// sorted by interaction type (the event-based design the reference lists as
// its next step, README.md:717-720: "per-interaction kernels with sorting by
// event type").
//
// The fused kernels in transport_kernels.hip keep each photon in a lane's
// registers and run the whole step there; lanes whose photons take different
// branches idle through each other's branch bodies. Here the step is split at
// its branches. Photon state lives in HBM (structure of arrays over W slots),
// and each iteration runs three kernels over it:
//
//   classify  every slot, in slot order: an empty slot first takes the
//             block's next photon (emission); then the two exponential
//             draws, the wall distance, the advance (photon.hpp advance());
//             absorbed bulk, black-wall and lost-reflector photons finish
//             here, specular reflections stay ready, and the rest go on a
//             per-type queue with the Philox words their next phase needs;
//   fresnel   the readout / polished-face queue: Fresnel probability, the
//             decision, detection or escape into the tally, or reflection;
//   rotate    the Rayleigh / Lambertian queue: the new direction.
//
// Every queue is segmented by block: block b of each kernel owns slots
// [b * seg, (b + 1) * seg) and photon indices [n b / grid, n (b + 1) / grid),
// so queue counters, tally rows and sums of block b are touched by block b
// only and need no global atomics. Within a block a wave reserves its queue
// entries with one LDS atomic (ballot + popcount + mbcnt rank).
//
// A photon's history is a function of (seed, photon id) only; the phases
// execute the fused step's arithmetic in the same order, so tallies equal
// the fused kernels' and the CPU path's exactly.
#include "kernels.hpp"
#include "wave.hpp"

namespace ophip {
namespace {

constexpr uint32_t kSlotBits = 29;
constexpr uint32_t kSlotMask = (1u << kSlotBits) - 1u;
constexpr uint32_t kCodeScatter = 6;  // rotate queue: Rayleigh; codes 0..5: Lambertian on that face

__device__ __forceinline__ uint64_t block_lo(const SortedArgs& A, uint32_t b) {
  return A.n * b / A.grid;
}

// Block-sum of the per-lane boundary / scatter sums into LDS.
__device__ __forceinline__ void add_sums(unsigned long long* sb, uint64_t snb, uint64_t sns) {
  const uint64_t a = wave_sum_u64(snb), c = wave_sum_u64(sns);
  if (lane_id() == 0) {
    if (a) atomicAdd(&sb[0], static_cast<unsigned long long>(a));
    if (c) atomicAdd(&sb[1], static_cast<unsigned long long>(c));
  }
}

// Adds this block's LDS counts (the first `words` words) and sums to its own
// partial row; the block is the row's only writer.
__device__ __forceinline__ void flush_row(const SortedArgs& A, const uint32_t* lds, uint32_t words,
                                          const unsigned long long* sb) {
  const uint32_t b = blockIdx.x;
  uint32_t* row = A.partials + static_cast<size_t>(b) * A.tally.words;
  for (uint32_t w = threadIdx.x; w < words; w += blockDim.x)
    if (lds[w]) row[w] += lds[w];
  if (threadIdx.x == 0) {
    A.sums[2u * b] += sb[0];
    A.sums[2u * b + 1u] += sb[1];
  }
}

__device__ __forceinline__ void add_stat(const SortedArgs& A, int k, uint64_t v) {
  A.stats[static_cast<size_t>(blockIdx.x) * kSortedStats + k] += v;
}

__device__ __forceinline__ void load_photon(const SortedArgs& A, uint32_t s, Photon& ph) {
  ph.p = V3{A.px[s], A.py[s], A.pz[s]};
  ph.d = V3{A.dx[s], A.dy[s], A.dz[s]};
  ph.t = A.t[s];
  ph.path = A.path[s];
  ph.nb = A.nb[s];
  ph.ns = A.ns[s];
}

// A finished photon into the block's LDS tally [fates | time | xy]
// (transport.hpp:164-181; counted as cpu_backend.cpp:28-36 counts).
__device__ __forceinline__ void tally_photon(const DevParams& P, const DevTally& T, uint32_t* lds,
                                             uint32_t fate, float t, float x, float y) {
  atomicAdd(&lds[fate], 1u);
  if (fate == kDetected) {
    atomicAdd(&lds[5u + time_bin(t, T.time_scale, T.time_bins)], 1u);
    const uint32_t xy =
        xy_bin(x, y, P.hx, P.hy, T.two_hx, T.two_hy, T.xy_bins, P.fast_div & 1u);
    atomicAdd(&lds[6u + T.time_bins + xy], 1u);
  }
}

// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kSortedBlock) void sorted_init_kernel(SortedArgs A) {
  const uint64_t W = static_cast<uint64_t>(A.grid) * A.seg;
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  const uint64_t i0 = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  for (uint64_t i = i0; i < W; i += stride) A.idx[i] = kSortedIdle;
  const uint64_t rows = static_cast<uint64_t>(A.grid) * A.tally.words;
  for (uint64_t i = i0; i < rows; i += stride) A.partials[i] = 0u;
  for (uint64_t i = i0; i < static_cast<uint64_t>(A.grid) * kSortedStats; i += stride)
    A.stats[i] = 0ull;
  for (uint64_t b = i0; b < A.grid; b += stride) {
    A.cnt_rot[b] = 0u;
    A.cnt_fres[b] = 0u;
    A.next_local[b] = block_lo(A, static_cast<uint32_t>(b));
    A.occ[b] = 0u;
    A.sums[2 * b] = 0ull;
    A.sums[2 * b + 1] = 0ull;
  }
}

// ---------------------------------------------------------------------------
// classify: the common part of one step for every occupied slot of the
// block, in slot order; an empty slot first takes the block's next photon
// index (emission, transport.hpp:55-67) while the block has indices left.
__global__ __launch_bounds__(kSortedBlock) void sorted_classify_kernel(DevParams P0, PhiloxKey key,
                                                                       SortedArgs A) {
  const DevParams& P = P0;
  __shared__ uint32_t q[2];  // rotate, fresnel
  __shared__ uint32_t fates[5];
  __shared__ unsigned long long sb[2];
  __shared__ uint32_t n_live, n_fin, n_new;
  if (threadIdx.x < 2) q[threadIdx.x] = 0u;
  if (threadIdx.x < 5) fates[threadIdx.x] = 0u;
  if (threadIdx.x < 2) sb[threadIdx.x] = 0ull;
  if (threadIdx.x == 0) n_live = n_fin = n_new = 0u;
  __syncthreads();
  const uint32_t b = blockIdx.x;
  const uint32_t base = b * A.seg;
  const uint64_t next = A.next_local[b];
  const uint64_t hi = block_lo(A, b + 1u);
  uint64_t snb = 0, sns = 0;
  uint32_t live = 0, fin = 0;
  for (uint32_t i0 = 0; i0 < A.seg; i0 += kSortedBlock) {
    const uint32_t s = base + i0 + threadIdx.x;
    uint32_t idx = A.idx[s];
    const bool idle = idx == kSortedIdle;
    const uint32_t r = next < hi ? reserve(&n_new, idle) : 0u;
    const bool fresh = idle && next + r < hi;
    if (fresh) idx = static_cast<uint32_t>(next + r);
    int where = -1;  // 0 rotate, 1 fresnel
    uint32_t key0 = 0, wa = 0, wb = 0;
    if (idx != kSortedIdle) {
      ++live;
      Photon ph;
      Stream rng;
      uint32_t draws;
      if (fresh) {
        rng.start(A.first_id + idx, key, 0u);
        emit(P, rng, ph);
        rng.refill(key);
        draws = rng.draws();
      } else {
        load_photon(A, s, ph);
        draws = A.draws[s];
        rng.start(A.first_id + idx, key, draws);
      }
      uint32_t used = 2u, fate = kMaxSteps;
      bool finished = P.max_steps == 0u, flipped = false;  // max_steps 0: done at emission
      Advance a{};
      if (!finished) {
        a = advance(P, rng, ph);
        const bool at_limit = ph.nb + ph.ns >= P.max_steps;  // if it survives this step
        if (a.absorb) {
          finished = true;
          fate = kAbsorbedBulk;
        } else if (a.scatter) {
          used = 4u;
          if (at_limit) {
            finished = true;
          } else {
            where = 0;
            key0 = s | (kCodeScatter << kSlotBits);
            wa = rng.template peek<2>();
            wb = rng.template peek<3>();
          }
        } else if (a.face == kReadoutFace || P.surface == kPolished) {
          used = 3u;
          where = 1;
          key0 = s | (static_cast<uint32_t>(a.face) << kSlotBits);
          wa = rng.template peek<2>();
        } else if (P.surface == kBlack) {
          finished = true;
          fate = kAbsorbedSurface;
        } else {  // specular or Lambertian reflector: u <= R survives (transport.hpp:134-137)
          used = 3u;
          if (!(to_unit(rng.template peek<2>()) <= P.reflectivity)) {
            finished = true;
            fate = kAbsorbedSurface;
          } else if (P.surface == kLambertian) {
            used = 5u;
            if (at_limit) {
              finished = true;
            } else {
              where = 0;
              key0 = s | (static_cast<uint32_t>(a.face) << kSlotBits);
              wa = rng.template peek<3>();
              wb = rng.template peek<4>();
            }
          } else {
            flip_axis(ph.d, a.axis);
            flipped = true;
            finished = at_limit;
          }
        }
      }
      if (finished) {
        ++fin;
        atomicAdd(&fates[fate], 1u);
        snb += ph.nb;
        sns += ph.ns;
        if (!fresh) A.idx[s] = kSortedIdle;
      } else {
        A.px[s] = ph.p.x;
        A.py[s] = ph.p.y;
        A.pz[s] = ph.p.z;
        A.t[s] = ph.t;
        A.path[s] = ph.path;
        A.draws[s] = draws + used;
        if (fresh) {
          A.idx[s] = idx;
          A.dx[s] = ph.d.x;
          A.dy[s] = ph.d.y;
          A.dz[s] = ph.d.z;
          A.nb[s] = ph.nb;
          A.ns[s] = ph.ns;
        } else {
          if (a.scatter)
            A.ns[s] = ph.ns;
          else
            A.nb[s] = ph.nb;
          if (flipped) {
            if (a.axis == 0) A.dx[s] = ph.d.x;
            if (a.axis == 1) A.dy[s] = ph.d.y;
            if (a.axis == 2) A.dz[s] = ph.d.z;
          }
        }
      }
    }
    const uint32_t pr = reserve(&q[0], where == 0);
    const uint32_t pf = reserve(&q[1], where == 1);
    if (where == 0) {
      A.rot_key[base + pr] = key0;
      A.rot_wc[base + pr] = wa;
      A.rot_wp[base + pr] = wb;
    } else if (where == 1) {
      A.fres_key[base + pf] = key0;
      A.fres_w[base + pf] = wa;
    }
  }
  add_sums(sb, snb, sns);
  const uint32_t wl = static_cast<uint32_t>(wave_sum_u64(live));
  const uint32_t wf = static_cast<uint32_t>(wave_sum_u64(fin));
  if (lane_id() == 0) {
    if (wl) atomicAdd(&n_live, wl);
    if (wf) atomicAdd(&n_fin, wf);
  }
  __syncthreads();
  flush_row(A, fates, 5u, sb);
  if (threadIdx.x == 0) {
    const uint64_t took = hi - next < n_new ? hi - next : static_cast<uint64_t>(n_new);
    A.next_local[b] = next + took;
    A.cnt_rot[b] = q[0];
    A.cnt_fres[b] = q[1];
    A.occ[b] = n_live - n_fin;
    add_stat(A, kSortedClassified, n_live - (P.max_steps == 0u ? took : 0u));
    add_stat(A, kSortedEmitted, took);
  }
}

// ---------------------------------------------------------------------------
// fresnel: readout and polished-face decisions (fresnel.hpp:30-46,
// transport.hpp:108-128). Reflected photons flip the approached component
// and stay ready; transmitted ones are detected (readout) or escape.
__global__ __launch_bounds__(kSortedBlock) void sorted_fresnel_kernel(DevParams P0, SortedArgs A) {
  const DevParams& P = P0;
  extern __shared__ uint32_t lds[];
  __shared__ uint32_t n_freed;
  __shared__ unsigned long long sb[2];
  const DevTally& T = A.tally;
  const uint32_t words = T.words;
  for (uint32_t w = threadIdx.x; w < words; w += blockDim.x) lds[w] = 0u;
  const uint32_t b = blockIdx.x;
  const uint32_t base = b * A.seg;
  if (threadIdx.x == 0) n_freed = 0u;
  if (threadIdx.x < 2) sb[threadIdx.x] = 0ull;
  __syncthreads();
  const uint32_t count = A.cnt_fres[b];
  uint64_t snb = 0, sns = 0;
  uint32_t freed_here = 0;
  for (uint32_t i = threadIdx.x; i < count; i += kSortedBlock) {
    bool freed = false;
    const uint32_t k = A.fres_key[base + i];
    const uint32_t w = A.fres_w[base + i];
    const uint32_t s = k & kSlotMask;
    const int face = static_cast<int>(k >> kSlotBits);
    const int axis = face >> 1;
    const bool plus = face & 1;
    const bool readout = face == kReadoutFace;
    float* dcomp = axis == 0 ? A.dx : (axis == 1 ? A.dy : A.dz);
    const float d_axis = dcomp[s];
    const float cos_i = plus ? d_axis : -d_axis;
    bool scored;
    const float prob = fresnel_prob(P, readout, cos_i, scored);
    const bool survived = to_unit(w) <= prob;
    const uint32_t nb = A.nb[s], ns = A.ns[s];
    uint32_t fate = kMaxSteps;
    if (survived) {
      dcomp[s] = -d_axis;
      freed = nb + ns >= P.max_steps;
    } else {
      freed = true;
      fate = readout ? kDetected : kEscaped;
    }
    if (freed) {
      float t = 0.0f, x = 0.0f, y = 0.0f;
      if (fate == kDetected) {
        t = A.t[s];
        x = A.px[s];
        y = A.py[s];
      }
      tally_photon(P, T, lds, fate, t, x, y);
      snb += nb;
      sns += ns;
      ++freed_here;
      A.idx[s] = kSortedIdle;
    }
  }
  add_sums(sb, snb, sns);
  const uint32_t wf = static_cast<uint32_t>(wave_sum_u64(freed_here));
  if (lane_id() == 0 && wf) atomicAdd(&n_freed, wf);
  __syncthreads();
  flush_row(A, lds, words, sb);
  if (threadIdx.x == 0) {
    A.occ[b] -= n_freed;
    add_stat(A, kSortedFresnel, count);
  }
}

// ---------------------------------------------------------------------------
// rotate: new directions for Rayleigh scattering and Lambertian reflection.
__global__ __launch_bounds__(kSortedBlock) void sorted_rotate_kernel(DevParams P0, SortedArgs A) {
  // The queue is lane-dense, so the Rayleigh cosines are computed here: the
  // table would only add a random gather per item to an HBM-bound kernel.
  DevParams P = P0;
  P.rayleigh = nullptr;
  const uint32_t b = blockIdx.x;
  const uint32_t base = b * A.seg;
  const uint32_t count = A.cnt_rot[b];
  for (uint32_t i = threadIdx.x; i < count; i += kSortedBlock) {
    const uint32_t k = A.rot_key[base + i];
    const uint32_t wc = A.rot_wc[base + i];
    const uint32_t wp = A.rot_wp[base + i];
    const uint32_t s = k & kSlotMask;
    const uint32_t code = k >> kSlotBits;
    const bool scatter = code == kCodeScatter;
    const int face = scatter ? 0 : static_cast<int>(code);
    V3 d{0.0f, 0.0f, 0.0f};
    if (scatter) d = V3{A.dx[s], A.dy[s], A.dz[s]};
    d = redirect(P, scatter, d, face >> 1, face & 1, wc, wp);
    A.dx[s] = d.x;
    A.dy[s] = d.y;
    A.dz[s] = d.z;
  }
  if (threadIdx.x == 0) add_stat(A, kSortedRotate, count);
}

// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kSortedBlock) void sorted_status_kernel(SortedArgs A) {
  __shared__ unsigned long long acc[2];
  if (threadIdx.x < 2) acc[threadIdx.x] = 0ull;
  __syncthreads();
  uint64_t occ = 0, left = 0;
  for (uint32_t b = threadIdx.x; b < A.grid; b += blockDim.x) {
    occ += A.occ[b];
    left += A.next_local[b] < block_lo(A, b + 1u);
  }
  occ = wave_sum_u64(occ);
  left = wave_sum_u64(left);
  if (lane_id() == 0) {
    atomicAdd(&acc[0], static_cast<unsigned long long>(occ));
    atomicAdd(&acc[1], static_cast<unsigned long long>(left));
  }
  __syncthreads();
  if (threadIdx.x < 2) A.status[threadIdx.x] = acc[threadIdx.x];
}

// ---------------------------------------------------------------------------
// tail: the photons still in flight once few remain, each run to the end of
// its history by one lane (the fused step of photon.hpp, resumed from the
// stored state and stream position).
__global__ __launch_bounds__(kSortedBlock) void sorted_tail_kernel(DevParams P0, PhiloxKey key,
                                                                   SortedArgs A) {
  const DevParams& P = P0;
  extern __shared__ uint32_t lds[];
  __shared__ unsigned long long sb[2];
  __shared__ uint32_t n_tail;
  const DevTally& T = A.tally;
  for (uint32_t w = threadIdx.x; w < T.words; w += blockDim.x) lds[w] = 0u;
  if (threadIdx.x < 2) sb[threadIdx.x] = 0ull;
  if (threadIdx.x == 0) n_tail = 0u;
  __syncthreads();
  const uint32_t b = blockIdx.x;
  const uint32_t base = b * A.seg;
  uint64_t snb = 0, sns = 0;
  uint32_t mine = 0;
  for (uint32_t i0 = 0; i0 < A.seg; i0 += kSortedBlock) {
    const uint32_t s = base + i0 + threadIdx.x;
    const uint32_t idx = A.idx[s];
    bool alive = idx != kSortedIdle;
    Photon ph;
    Stream rng;
    uint32_t done = 0, fate = kMaxSteps;
    if (alive) {
      ++mine;
      load_photon(A, s, ph);
      rng.start(A.first_id + idx, key, A.draws[s]);
      done = ph.nb + ph.ns;
      A.idx[s] = kSortedIdle;
    }
    while (__ballot(alive) != 0) {
      if (alive) {
        rng.refill(key);
        StepInfo info{};
        const bool still = step(P, rng, ph, fate, info);
        ++done;
        if (!still || done >= P.max_steps) {
          if (still) fate = kMaxSteps;
          alive = false;
          tally_photon(P, T, lds, fate, ph.t, ph.p.x, ph.p.y);
          snb += ph.nb;
          sns += ph.ns;
        }
      }
    }
  }
  add_sums(sb, snb, sns);
  const uint32_t wm = static_cast<uint32_t>(wave_sum_u64(mine));
  if (lane_id() == 0 && wm) atomicAdd(&n_tail, wm);
  __syncthreads();
  flush_row(A, lds, T.words, sb);
  if (threadIdx.x == 0) {
    A.occ[b] = 0u;
    add_stat(A, kSortedTail, n_tail);
  }
}

}  // namespace

void launch_sorted_init(const SortedArgs& A, hipStream_t stream) {
  const uint64_t W = static_cast<uint64_t>(A.grid) * A.seg;
  const uint64_t blocks = (W + kSortedBlock - 1) / kSortedBlock;
  sorted_init_kernel<<<static_cast<unsigned>(blocks < 8192 ? blocks : 8192), kSortedBlock, 0,
                       stream>>>(A);
}

void launch_sorted_iteration(const DevParams& P, PhiloxKey key, const SortedArgs& A,
                             hipStream_t stream) {
  const size_t lds = static_cast<size_t>(A.tally.words) * sizeof(uint32_t);
  sorted_classify_kernel<<<A.grid, kSortedBlock, 0, stream>>>(P, key, A);
  sorted_fresnel_kernel<<<A.grid, kSortedBlock, lds, stream>>>(P, A);
  sorted_rotate_kernel<<<A.grid, kSortedBlock, 0, stream>>>(P, A);
}

void launch_sorted_status(const SortedArgs& A, hipStream_t stream) {
  sorted_status_kernel<<<1, kSortedBlock, 0, stream>>>(A);
}

void launch_sorted_tail(const DevParams& P, PhiloxKey key, const SortedArgs& A,
                        hipStream_t stream) {
  const size_t lds = static_cast<size_t>(A.tally.words) * sizeof(uint32_t);
  sorted_tail_kernel<<<A.grid, kSortedBlock, lds, stream>>>(P, key, A);
}

}  // namespace ophip
