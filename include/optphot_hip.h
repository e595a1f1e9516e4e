/*
 * optphot_hip.h -- C-ABI of the MI355X (gfx950) optical-photon transport
 * library, liboptphot_hip.so.
 *
 * Plain C: POD structs, plain pointers and sizes. Every entry point is
 * synchronous (it returns when the results are in the caller's memory),
 * returns OPTPHOT_OK or an error code, and leaves a message for
 * optphot_last_error(). Calls are serialised per process by the library.
 *
 * The first group replaces the reference's GPU backend entry points
 * (include/optphot/backend.hpp:84-115); the C++ header optphot_hip/backend.hpp
 * re-exposes them with the reference's exact C++ signatures and exception
 * types. The second group is new: the persistent wavefront tally, batched
 * events and the adjoint.
 */
#ifndef OPTPHOT_HIP_H
#define OPTPHOT_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Return codes. INVALID_ARGUMENT corresponds to the reference's
 * std::invalid_argument (gpu_event.cu:177-182), RUNTIME to std::runtime_error
 * (cuda_util.cuh:17-24, gpu_backend.cu:171-185). */
#define OPTPHOT_OK 0
#define OPTPHOT_ERR_INVALID_ARGUMENT 1
#define OPTPHOT_ERR_RUNTIME 2
#define OPTPHOT_ERR_NO_DEVICE 3

/* Surface / source / fate codes: include/optphot/params.hpp:15-35. */
enum { OPTPHOT_SURFACE_POLISHED = 0, OPTPHOT_SURFACE_BLACK = 1, OPTPHOT_SURFACE_SPECULAR = 2,
       OPTPHOT_SURFACE_LAMBERTIAN = 3 };
enum { OPTPHOT_SOURCE_POINT = 0, OPTPHOT_SOURCE_VOLUME = 1 };
enum { OPTPHOT_FATE_DETECTED = 0, OPTPHOT_FATE_ABSORBED_BULK = 1,
       OPTPHOT_FATE_ABSORBED_SURFACE = 2, OPTPHOT_FATE_ESCAPED = 3, OPTPHOT_FATE_MAX_STEPS = 4 };

/* SimParams, include/optphot/params.hpp:40-63: 72 bytes, same field order,
 * lengths in mm, times in ns, +inf disables a bulk process. */
typedef struct optphot_sim_params {
  float hx, hy, hz;
  float n_scint, n_det, n_out;
  float abs_length, scat_length;
  uint32_t surface;
  float reflectivity;
  uint32_t source;
  float src_x, src_y, src_z;
  float tau;
  uint32_t max_steps;
  uint64_t seed;
} optphot_sim_params;

/* TallyConfig, include/optphot/transport.hpp:164-168. */
typedef struct optphot_tally_config {
  uint32_t time_bins;
  float time_max;
  uint32_t xy_bins;
} optphot_tally_config;

/* RunTiming, include/optphot/backend.hpp:65-71 (milliseconds). */
typedef struct optphot_run_timing {
  double total_ms, kernel_ms, h2d_ms, d2h_ms, alloc_ms;
} optphot_run_timing;

/* EventStats, include/optphot/backend.hpp:95-103, with wave64 lanes:
 * issued_lane_steps = sum over waves and steps of 64 x (any lane stepping). */
typedef struct optphot_event_stats {
  uint64_t launches, useful_lane_steps, issued_lane_steps;
  double kernels_ms;
} optphot_event_stats;

/* Caller-owned host arrays of n elements each: PhotonRecords' structure of
 * arrays (backend.hpp:31-50, 29 bytes per photon), indexed by photon number. */
typedef struct optphot_records {
  uint8_t* fate;
  uint32_t* n_boundary;
  uint32_t* n_scatter;
  float *t, *x, *y, *z, *path;
} optphot_records;

/* Flat tally layout (u64 words), the reference's device layout
 * (gpu_backend.cu:56-66): [fate counts 5 | time hist time_bins+1 |
 * xy hist xy_bins^2 row-major [iy][ix] | boundary_sum | scatter_sum]. */
#define OPTPHOT_TALLY_WORDS(c) (5u + (c).time_bins + 1u + (c).xy_bins * (c).xy_bins + 2u)

/* Branch-event counters of a run, for the algorithmic-work accounting. */
enum {
  OPTPHOT_CNT_USEFUL_LANE_STEPS = 0, /* photon-steps performed */
  OPTPHOT_CNT_ISSUED_LANE_STEPS = 1, /* 64 x wave-steps with any lane active */
  OPTPHOT_CNT_ABSORB = 2,
  OPTPHOT_CNT_SCATTER = 3,
  OPTPHOT_CNT_READOUT = 4,       /* readout-face hits (Fresnel to n_det) */
  OPTPHOT_CNT_POLISHED = 5,      /* polished side hits (Fresnel to n_out) */
  OPTPHOT_CNT_BLACK = 6,         /* black side hits */
  OPTPHOT_CNT_REFLECTOR = 7,     /* specular / Lambertian side hits */
  OPTPHOT_CNT_LAMBERT_REFLECT = 8,
  OPTPHOT_CNT_EMITTED = 9,
  OPTPHOT_CNT_DETECTED = 10,
  OPTPHOT_CNT_ABS_ON = 11,  /* photon-steps with finite abs_length (distance sampled) */
  OPTPHOT_CNT_SCAT_ON = 12, /* photon-steps with finite scat_length */
  OPTPHOT_CNT_SLOTS = 16
};
typedef struct optphot_step_counters {
  uint64_t v[OPTPHOT_CNT_SLOTS];
} optphot_step_counters;

/* Launch shape of the persistent wavefront kernel. Zero fields are filled
 * from the autotuner's choice for the run (or built-in defaults). */
typedef struct optphot_wavefront_tuning {
  int block_size;             /* threads per block, multiple of 64, <= 1024 */
  uint32_t steps_per_refill;  /* transport steps between work-queue refills */
  uint32_t queue_chunk;       /* photons a wave claims per queue atomic */
  int blocks_per_cu;          /* persistent blocks per compute unit */
  /* Wavefront tally only: 0 = refill finished lanes every steps_per_refill
   * steps; n in [1, 64] = every lane keeps an emitted successor photon and a
   * wave emits successors once n of its lanes lack one (steps_per_refill is
   * then unused). Results are identical either way. */
  uint32_t successor_threshold;
  /* Wavefront tally only: nonzero keeps each lane's Philox window in an LDS
   * ring (generation passes shared across steps; takes precedence over
   * successor_threshold). Results are identical either way. */
  uint32_t stream_in_lds;
  /* Wavefront tally only: nonzero selects the LDS-ring kernel whose steps
   * are split at their branches: the Fresnel decisions and the new
   * directions of a step are sorted onto block-wide LDS queues by
   * interaction type and computed lane-dense between two block barriers
   * (takes precedence over the two fields above). Results are identical. */
  uint32_t sorted_phases;
} optphot_wavefront_tuning;

/* ---- Reference GPU backend entry points ----------------------------------- */

/* gpu_available(), backend.hpp:84 */
int optphot_gpu_available(void);
/* gpu_device_name(), backend.hpp:85; writes a NUL-terminated name. */
int optphot_gpu_device_name(char* buf, size_t len);
/* Message of the last failed call on this thread ("" if none). */
const char* optphot_last_error(void);

/* run_gpu_records(p, n, out, block_size=256, first_id=0, chunk=1<<24),
 * backend.hpp:88-89 / gpu_backend.cu:119-154: history-based kernel, one lane
 * per photon, records of photons first_id .. first_id+n-1 into `out`.
 * block_size: multiple of 64 in [64, 1024]. `timing` may be NULL. The device
 * holds `chunk` photons of records; each chunk runs as several launches whose
 * records are copied out through pinned staging buffers while later launches
 * run (d2h_ms: host time spent waiting for and copying records). */
int optphot_run_gpu_records(const optphot_sim_params* p, uint64_t n, const optphot_records* out,
                            int block_size, uint64_t first_id, uint64_t chunk,
                            optphot_run_timing* timing);

/* run_gpu_tally(p, c, n, out, block_size=256, first_id=0), backend.hpp:90-91 /
 * gpu_backend.cu:158-212: history-based tally into `tally`
 * (OPTPHOT_TALLY_WORDS(*c) u64 words). */
int optphot_run_gpu_tally(const optphot_sim_params* p, const optphot_tally_config* c, uint64_t n,
                          uint64_t* tally, int block_size, uint64_t first_id,
                          optphot_run_timing* timing);

/* run_gpu_records_event(p, n, out, steps_per_launch, block_size, first_id,
 * chunk, stats), backend.hpp:112-115 / gpu_event.cu:174-244: the same records
 * as optphot_run_gpu_records, bit for bit, computed by the persistent
 * wavefront kernel with `steps_per_launch` steps between refills. One launch
 * per chunk. `stats` and `timing` may be NULL. */
int optphot_run_gpu_records_event(const optphot_sim_params* p, uint64_t n,
                                  const optphot_records* out, uint32_t steps_per_launch,
                                  int block_size, uint64_t first_id, uint64_t chunk,
                                  optphot_event_stats* stats, optphot_run_timing* timing);

/* ---- New entry points (no counterpart in the reference) -------------------- */

/* Records of photons first_id .. first_id+n-1 written straight into
 * caller-owned DEVICE arrays (8 arrays of n elements, as optphot_records),
 * for callers that consume records on the GPU (e.g. torch tensors): no PCIe
 * copy. steps_per_launch = 0 selects the history kernel, >= 1 the wavefront
 * kernel with that many steps between refills. Same records, bit for bit, as
 * optphot_run_gpu_records. stats / timing may be NULL. */
int optphot_run_gpu_records_device(const optphot_sim_params* p, uint64_t n,
                                   const optphot_records* dev_out, uint32_t steps_per_launch,
                                   int block_size, uint64_t first_id, optphot_event_stats* stats,
                                   optphot_run_timing* timing);

/* Launch shape of a records run (no counterpart in the reference). */
enum {
  OPTPHOT_RECORDS_AUTO = 0,      /* choose from n (see optphot_run_gpu_records_tuned) */
  OPTPHOT_RECORDS_HISTORY = 1,   /* one lane per photon */
  OPTPHOT_RECORDS_WAVEFRONT = 2  /* persistent grid, finished lanes refilled */
};
typedef struct optphot_records_tuning {
  int kernel;                 /* OPTPHOT_RECORDS_* */
  int block_size;             /* multiple of 64 in [64, 1024]; 0 = 256 */
  uint32_t steps_per_refill;  /* wavefront: steps between refills; 0 = 2 */
  uint32_t queue_chunk;       /* wavefront: photons a wave claims per atomic; 0 = 256 */
  int blocks_per_cu;          /* wavefront: persistent blocks per CU; 0 = occupancy limit */
} optphot_records_tuning;

/* Records of photons first_id .. first_id+n-1 (the records of
 * optphot_run_gpu_records, bit for bit) with a chosen launch shape, into host
 * arrays (out_on_device = 0; pipelined through pinned staging as
 * optphot_run_gpu_records) or caller-owned device arrays (out_on_device = 1).
 * No counterpart in the reference, whose run_gpu_records / run_gpu_records_event
 * (backend.hpp:88-89, 112-115) take the kernel's parameters from the caller:
 * this entry point serves callers that take a tuned shape, or none.
 * tuning NULL or kernel OPTPHOT_RECORDS_AUTO: the wavefront kernel (2 steps
 * between refills, queue chunk 256, block size as given or 256) once n
 * reaches 8 photons per lane of its persistent grid, else the history kernel.
 * `chosen` (may be NULL) receives the shape that ran. stats / timing may be
 * NULL. */
int optphot_run_gpu_records_tuned(const optphot_sim_params* p, uint64_t n,
                                  const optphot_records* out, int out_on_device,
                                  const optphot_records_tuning* tuning, uint64_t first_id,
                                  optphot_records_tuning* chosen, optphot_event_stats* stats,
                                  optphot_run_timing* timing);

/* Tally of photons first_id .. first_id+n-1 by the persistent wavefront kernel
 * (device work queue, refill of finished lanes, LDS-privatised histograms).
 * Same tally as optphot_run_gpu_tally, word for word. tuning / stats /
 * counters / timing may be NULL. */
int optphot_run_gpu_tally_wavefront(const optphot_sim_params* p, const optphot_tally_config* c,
                                    uint64_t n, uint64_t* tally, uint64_t first_id,
                                    const optphot_wavefront_tuning* tuning,
                                    optphot_event_stats* stats, optphot_step_counters* counters,
                                    optphot_run_timing* timing);

/* The tally of optphot_run_gpu_tally_wavefront, word for word, from a call
 * that allocates nothing. No counterpart in the reference: it is for callers
 * that cannot allocate device memory (serving loops, harnesses that forbid
 * hipMalloc). The caller passes `dev_tally`, OPTPHOT_TALLY_WORDS(*c) u64
 * words of DEVICE memory, which receive the tally. The launch's working
 * memory (queue head, block sums, per-block partial histograms) is one
 * static 16 MiB device array of the library, held by each device until the
 * process ends; its address is looked up when the library sets up a device,
 * so some other call (for example optphot_gpu_device_name) must come first
 * on the device, or this one fails with OPTPHOT_ERR_RUNTIME. The persistent
 * grid is trimmed to what the array holds (results do not depend on it); a
 * binning whose partial rows do not fit at one block per compute unit
 * (about 16 000 tally words) is rejected with OPTPHOT_ERR_INVALID_ARGUMENT.
 * Every call in the process on a device shares that device's array, which
 * keeps the last call's contents (each call clears what must start at zero).
 * Calls that overlapped on one device would race on it; the library's lock
 * serialises calls within the process and each returns only when its
 * kernels are done. tuning / timing may be NULL; no branch-event counters. */
int optphot_run_gpu_tally_wavefront_noalloc(const optphot_sim_params* p,
                                            const optphot_tally_config* c, uint64_t n,
                                            uint64_t* dev_tally, uint64_t first_id,
                                            const optphot_wavefront_tuning* tuning,
                                            optphot_run_timing* timing);

/* Branch-event counters of the history-based tally kernel (same work as
 * optphot_run_gpu_tally). */
int optphot_run_gpu_tally_counters(const optphot_sim_params* p, const optphot_tally_config* c,
                                   uint64_t n, uint64_t* tally, int block_size, uint64_t first_id,
                                   optphot_step_counters* counters, optphot_run_timing* timing);

/* Per-interaction ("sorted") wavefront path: photon state in device memory,
 * and per iteration one kernel for emission into empty slots and the common
 * part of a step (which sorts the photons onto per-interaction queues), one
 * for the readout / polished-face Fresnel decisions and one for the
 * Rayleigh / Lambertian directions. This is the design the
 * reference lists as its next step ("per-interaction kernels with sorting by
 * event type", README.md:717-720). Zero fields take the defaults. */
typedef struct optphot_sorted_tuning {
  uint32_t slots_per_block;     /* photons in flight per block: a multiple of 256
                                   in [256, 65536]; default 2048 */
  uint32_t blocks_per_cu;       /* blocks = blocks_per_cu x CUs, in [1, 64];
                                   default 6 (the classify kernel's occupancy) */
  uint32_t iterations_per_poll; /* iterations per graph replay between host
                                   checks: even, in [2, 1024]; default 8 */
  uint32_t tail_threshold;      /* once every photon is emitted and at most
                                   this many are in flight, one kernel
                                   finishes them history-style; default
                                   slots / 8 (0xffffffff: at once) */
} optphot_sorted_tuning;

typedef struct optphot_sorted_stats {
  uint64_t slots;        /* photons in flight at most (blocks x slots_per_block) */
  uint64_t iterations;   /* classify / fresnel / rotate rounds run */
  uint64_t classified;   /* photon-steps of the classify kernel */
  uint64_t fresnel;      /* Fresnel decisions */
  uint64_t rotated;      /* new directions (Rayleigh + Lambertian) */
  uint64_t emitted;      /* photons emitted */
  uint64_t tail_photons; /* photons finished by the tail kernel */
  double loop_ms;        /* init + iterations, host checks included */
  double tail_ms;        /* tail kernel and tally reduction */
} optphot_sorted_stats;

/* Tally of photons first_id .. first_id+n-1 (n < 2^32 - 1) by the
 * per-interaction path. Same tally as optphot_run_gpu_tally, word for word.
 * tuning / stats / timing may be NULL. */
int optphot_run_gpu_tally_sorted(const optphot_sim_params* p, const optphot_tally_config* c,
                                 uint64_t n, uint64_t* tally, uint64_t first_id,
                                 const optphot_sorted_tuning* tuning, optphot_sorted_stats* stats,
                                 optphot_run_timing* timing);

/* Batched events. Event e has parameters params[e] and photons[e] photons,
 * with Philox photon ids first_id + offset_e .. first_id + offset_e +
 * photons[e] - 1, offset_e = photons[0] + ... + photons[e-1]; every event
 * uses the run seed params[0].seed (all seeds must be equal). Event e's
 * tally is the u32 row event_tally[e * (6 + time_bins) ...]: 5 fate counts,
 * then the arrival-time histogram of detected photons on [0, time_max) with
 * an overflow bin. So event e equals the reference's
 * run_cpu_tally(params[e], {time_bins, time_max}, photons[e], ..., first_id +
 * offset_e) (backend.hpp:78-79) restricted to those words.
 *   OPTPHOT_BATCH_WAVEFRONT: one persistent launch for all events, captured in
 *     a hipGraph and replayed by later calls of the same shape;
 *   OPTPHOT_BATCH_PLAIN: one history-kernel launch per event, in order.
 * tuning / counters / timing may be NULL. */
#define OPTPHOT_BATCH_WAVEFRONT 0
#define OPTPHOT_BATCH_PLAIN 1
int optphot_run_gpu_batch(const optphot_sim_params* params, const uint64_t* photons,
                          uint32_t n_events, uint64_t first_id, uint32_t time_bins,
                          float time_max, uint32_t* event_tally, int method,
                          const optphot_wavefront_tuning* tuning,
                          optphot_step_counters* counters, optphot_run_timing* timing);

/* Adjoint of per-set arrival-time histograms (vector-Jacobian product).
 * Events are parameter sets, laid out as for optphot_run_gpu_batch. h_e[k]
 * is the fraction of set e's photons detected in time bin k (k = 0 ..
 * time_bins, the last being the overflow bin), i.e. the batch tally's time
 * histogram divided by photons[e]. Given weights w[e][k] (f64, row-major
 * [n_events][time_bins + 1]) the call writes
 *   grad[e][j] = d/d theta_j  sum_k w[e][k] h_e[k]
 * for theta = (abs_length, scat_length, reflectivity, n_det, n_scint)
 * (grad: f64 [n_events][5]), estimated on the same photons (same seed and
 * ids) by likelihood-ratio scores of every absorb / scatter / Fresnel /
 * reflector decision and, for n_scint, the pathwise arrival-time term
 * through the exponential emission time (requires tau > 0 for that term).
 * Histories are recomputed from their Philox counters (no stored tape).
 * Sets with specular or Lambertian faces need reflectivity in (0, 1] (the
 * reflector score is hits / R); others are rejected as invalid arguments. */
int optphot_run_gpu_adjoint(const optphot_sim_params* params, const uint64_t* photons,
                            uint32_t n_events, uint64_t first_id, uint32_t time_bins,
                            float time_max, const double* weights, double* grad,
                            const optphot_wavefront_tuning* tuning, optphot_run_timing* timing);

/* A step summary ("genstep"): photons are emitted uniformly along the
 * segment (x0,y0,z0) at t0 -> (x1,y1,z1) at t1. Both end points must lie in
 * the box. 40 bytes.
 *  - cone_cos == 0: scintillation. Isotropic directions, each photon with
 *    the run's Exp(tau) emission delay added.
 *  - cone_cos in (0, 1]: Cherenkov light of a charged particle moving along
 *    the segment. Directions on the cone of half-angle acos(cone_cos) about
 *    the segment direction (cone_cos = 1 / (beta n) for speed beta c in
 *    index n), uniform in azimuth, no delay. The segment must not be
 *    degenerate. */
typedef struct optphot_genstep {
  float x0, y0, z0, t0;
  float x1, y1, z1, t1;
  uint32_t n_photons;
  float cone_cos;
} optphot_genstep;

/* Photons generated on the device from gensteps and transported in the box
 * of `p` (its source fields are unused). Photon k of genstep g has global
 * index i = n_0 + ... + n_{g-1} + k and Philox id first_id + i. Draws per
 * photon: scintillation [u along the segment, only if the segment or time
 * span is non-degenerate], cos, phi, emission delay; Cherenkov u along the
 * segment, phi; then the transport steps. A
 * degenerate genstep with t0 = 0 is the point source at (x0, y0, z0), photon
 * for photon. Writes the tally (as optphot_run_gpu_tally) and, when `records`
 * is non-NULL, records of all photons at their global index (host arrays of
 * the total photon count). tuning->stream_in_lds selects the kernel with
 * the Philox windows in LDS rings (tally runs only). No counterpart in the
 * reference, which lists generation from gensteps as future work
 * (README.md:734-735). */
int optphot_run_gpu_gensteps(const optphot_sim_params* p, const optphot_genstep* gensteps,
                             uint32_t n_gensteps, uint64_t first_id,
                             const optphot_tally_config* c, uint64_t* tally,
                             const optphot_records* records,
                             const optphot_wavefront_tuning* tuning,
                             optphot_step_counters* counters, optphot_run_timing* timing);

/* One calibration step for a batch of parameter sets, as a single captured
 * hipGraph replayed while the shapes are unchanged: the batched forward
 * launch (as optphot_run_gpu_batch), the per-set chi-square loss of the
 * arrival-time histogram against target fractions target[e][k] (f64,
 * [n_sets][time_bins + 1], overflow included),
 *   loss_e = sum_k (h_ek - T_ek)^2 / (T_ek + 1/n_t) + (f_e - F_e)^2 / (F_e (1 - F_e) + 1/n_t),
 * h_ek = detected photons of set e in bin k / photons[e], f_e = sum_k h_ek,
 * F_e = sum_k T_ek, n_t = target_photons[e], and the adjoint pass with
 * weights dloss_e/dh_ek. Writes hist (f64 [n_sets][time_bins + 1]), loss
 * (f64 [n_sets]) and grad = dloss_e/dtheta_j (f64 [n_sets][5], theta as for
 * optphot_run_gpu_adjoint). Parameters, offsets and targets are uploaded
 * before each replay, so a replay with new parameters is the same
 * computation as a fresh capture. Reflectivity as for optphot_run_gpu_adjoint.
 * Tunings may be NULL. */
int optphot_run_gpu_chi2_step(const optphot_sim_params* params, const uint64_t* photons,
                              uint32_t n_sets, uint64_t first_id, uint32_t time_bins,
                              float time_max, const double* target, const double* target_photons,
                              double* hist, double* loss, double* grad,
                              const optphot_wavefront_tuning* forward_tuning,
                              const optphot_wavefront_tuning* adjoint_tuning,
                              optphot_run_timing* timing);

/* Copies the device's Rayleigh-cosine table into out (n must be 2^24):
 * entry k is the Rayleigh scattering cosine (sampling.hpp:39-44) of the
 * uniform (k + 1) 2^-24, i.e. of every Philox word whose top 24 bits are k.
 * The transport kernels read scattering cosines from this table
 * (diagnostics: tests compare it with the CPU function). */
int optphot_rayleigh_table(float* out, size_t n);

/* Number of compute units and the kernels' register / occupancy facts,
 * written as "key=value" lines into buf (diagnostics). */
int optphot_device_info(char* buf, size_t len);

#ifdef __cplusplus
}
#endif

#endif /* OPTPHOT_HIP_H */
