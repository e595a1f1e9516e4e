// ref_capi.cpp -- a C-ABI harness around the reference's own CPU path, so the
// oracle's tests can call the reference itself.
//
// TEST INFRASTRUCTURE ONLY. Built by oracle/Makefile (target `ref`) together
// with the reference's src/cpu/cpu_backend.cpp and its headers, read in place
// from the reference checkout; the output goes to oracle/_ref/ and is never
// committed. Only tests/ and bench.py's cpu_baseline leg load it.
#include <cstring>

#include "optphot/analytic.hpp"
#include "optphot/backend.hpp"
#include "optphot/transport.hpp"

using namespace optphot;

static_assert(sizeof(SimParams) == 72, "SimParams is the 72-byte POD of params.hpp");
static_assert(sizeof(TallyConfig) == 12, "TallyConfig is 12 bytes");

#define REF_EXPORT extern "C" __attribute__((visibility("default")))

REF_EXPORT int ref_portable_math() {
#if defined(OPTPHOT_PORTABLE_MATH)
  return 1;
#else
  return 0;
#endif
}

REF_EXPORT int ref_max_threads() { return cpu_max_threads(); }

REF_EXPORT void ref_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  const U32x4 r = philox4x32_10(U32x4{ctr[0], ctr[1], ctr[2], ctr[3]}, U32x2{key[0], key[1]});
  out[0] = r.x;
  out[1] = r.y;
  out[2] = r.z;
  out[3] = r.w;
}

REF_EXPORT void ref_stream_words(uint64_t seed, uint64_t id, uint32_t skip, uint32_t n,
                                 uint32_t* out) {
  PhiloxStream s(seed, id, skip);
  for (uint32_t i = 0; i < n; ++i) out[i] = s.next_u32();
}

REF_EXPORT float ref_log(float x) { return opt_log(x); }
REF_EXPORT void ref_sincos_turns(float u, float* s, float* c) { opt_sincos_turns(u, *s, *c); }
REF_EXPORT float ref_cbrt(float x) { return opt_cbrt(x); }
REF_EXPORT float ref_portable_log(float x) { return portable::log(x); }
REF_EXPORT void ref_portable_sincos_turns(float u, float* s, float* c) {
  portable::sincos_turns(u, *s, *c);
}
REF_EXPORT float ref_portable_cbrt(float x) { return portable::cbrt(x); }
REF_EXPORT float ref_fresnel_R(float n1, float n2, float cos_i) {
  return fresnel_unpolarized(n1, n2, cos_i).reflectance;
}
REF_EXPORT float ref_box_exit(const float p[3], const float d[3], float hx, float hy, float hz,
                              int* face) {
  const BoxHit h = distance_to_box_exit(Vec3{p[0], p[1], p[2]}, Vec3{d[0], d[1], d[2]}, hx, hy, hz);
  *face = h.face;
  return h.dist;
}
REF_EXPORT float ref_rayleigh_cos(float u) { return sample_rayleigh_cos(u); }

// One photon, plus the number of Philox words its history consumed.
REF_EXPORT void ref_transport_one(const SimParams* p, uint64_t id, float out_f[5],
                                  uint32_t out_u[4]) {
  PhiloxStream rng(p->seed, id);
  PhotonState s = start_photon(*p, rng);
  Fate fate = Fate::MaxSteps;
  for (uint32_t step = 0; step < p->max_steps; ++step) {
    if (!step_photon(*p, rng, s, fate)) break;
  }
  const PhotonResult r = finish_photon(s, fate);
  out_f[0] = r.t;
  out_f[1] = r.pos.x;
  out_f[2] = r.pos.y;
  out_f[3] = r.pos.z;
  out_f[4] = r.path;
  out_u[0] = static_cast<uint32_t>(r.fate);
  out_u[1] = r.n_boundary;
  out_u[2] = r.n_scatter;
  out_u[3] = rng.draws();
}

REF_EXPORT void ref_run_records(const SimParams* p, uint64_t n, uint64_t first_id, int threads,
                                uint8_t* fate, uint32_t* nb, uint32_t* ns, float* t, float* x,
                                float* y, float* z, float* path) {
  PhotonRecords r;
  run_cpu_records(*p, n, r, threads, first_id);
  std::memcpy(fate, r.fate.data(), n);
  std::memcpy(nb, r.n_boundary.data(), n * 4);
  std::memcpy(ns, r.n_scatter.data(), n * 4);
  std::memcpy(t, r.t.data(), n * 4);
  std::memcpy(x, r.x.data(), n * 4);
  std::memcpy(y, r.y.data(), n * 4);
  std::memcpy(z, r.z.data(), n * 4);
  std::memcpy(path, r.path.data(), n * 4);
}

// Tally in the flat u64 layout [fates | time hist | xy hist | boundary sum, scatter sum].
REF_EXPORT double ref_run_tally(const SimParams* p, const TallyConfig* c, uint64_t n,
                                uint64_t first_id, int threads, uint64_t* out) {
  Tally t;
  const RunTiming tm = run_cpu_tally(*p, *c, n, t, threads, first_id);
  std::size_t k = 0;
  for (uint64_t v : t.fate_counts) out[k++] = v;
  for (uint64_t v : t.time_hist) out[k++] = v;
  for (uint64_t v : t.xy_hist) out[k++] = v;
  out[k++] = t.boundary_sum;
  out[k++] = t.scatter_sum;
  return tm.kernel_ms;
}

REF_EXPORT uint32_t ref_time_bin(float t, const TallyConfig* c) { return time_bin(t, *c); }
REF_EXPORT uint32_t ref_xy_bin(float x, float y, const SimParams* p, const TallyConfig* c) {
  return xy_bin(x, y, *p, *c);
}

REF_EXPORT double ref_face_solid_angle_fraction(const SimParams* p, int face) {
  return analytic::face_solid_angle_fraction(*p, face);
}
REF_EXPORT double ref_face_survival_fraction(const SimParams* p, int face, double L) {
  return analytic::face_survival_fraction(*p, face, L);
}
REF_EXPORT double ref_mirror_box_escape_fraction(double n_scint, double n_det) {
  return analytic::mirror_box_escape_fraction(n_scint, n_det);
}
