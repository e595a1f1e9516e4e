/*
 * optphot_oracle.c -- CPU restatement of the reference's optical-photon
 * transport path, written in plain C.
 *
 * TEST INFRASTRUCTURE ONLY. This file is the oracle the HIP kernels are
 * checked against. Only tests/, __graft_entry__.smoke() and the cpu_baseline
 * leg of bench.py may load the library built from it; the product library
 * never links, calls or falls back to it.
 *
 * It follows, operation by operation and in the same evaluation order, the
 * reference's header-only physics:
 *   include/optphot/philox.hpp    Philox4x32-10 and the per-photon stream
 *   include/optphot/pmath.hpp     log / sincos(2 pi u) / cbrt, two modes
 *   include/optphot/vec3.hpp      vector algebra, Duff et al. basis, rotation
 *   include/optphot/sampling.hpp  inverse-CDF samplers
 *   include/optphot/fresnel.hpp   unpolarised Fresnel reflectance, reflect
 *   include/optphot/geometry.hpp  box exit distance, snap, face normal
 *   include/optphot/transport.hpp start / step / transport, tally binning
 *   src/cpu/cpu_backend.cpp       records and tally loops (OpenMP dynamic 4096)
 * Evaluation order matters: every + - * / is correctly rounded in IEEE-754,
 * so the same expression tree gives the same bits on any conforming machine.
 * Build with -ffp-contract=off and without -ffast-math (the reference's FP
 * policy, CMakeLists.txt:25-34).
 *
 * Math mode: compiled with ORACLE_PORTABLE_MATH the transcendental functions
 * are the reference's portable ones (pmath.hpp:62-130, only correctly rounded
 * operations => bit-reproducible); without it they are the C library's
 * logf/sinf/cosf/cbrtf (pmath.hpp:134-158), as in the reference's default
 * build.
 */
#include <math.h>
#include <stdlib.h>
#include <stdint.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "optphot_oracle.h"

#define ORACLE_EXPORT __attribute__((visibility("default")))

/* hd.hpp:35-40 */
static const float kSpeedOfLight = 299.792458f; /* mm / ns */

static inline int is_inf(float x) { return x == INFINITY; } /* hd.hpp:44 */

/* ------------------------------------------------------------------------ */
/* Philox4x32-10 (philox.hpp:37-76)                                          */

static inline void philox_block(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                                uint32_t k1, uint32_t out[4]) {
  for (int r = 0; r < 10; ++r) {
    if (r > 0) { /* key schedule, philox.hpp:69-72 */
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    /* one round, philox.hpp:54-59: products with M0 = 0xD2511F53 (word 0)
       and M1 = 0xCD9E8D57 (word 2) */
    const uint64_t p0 = (uint64_t)0xD2511F53u * (uint64_t)c0;
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * (uint64_t)c2;
    const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
    const uint32_t n1 = (uint32_t)p1;
    const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
    const uint32_t n3 = (uint32_t)p0;
    c0 = n0;
    c1 = n1;
    c2 = n2;
    c3 = n3;
  }
  out[0] = c0;
  out[1] = c1;
  out[2] = c2;
  out[3] = c3;
}

/* philox.hpp:82-84: top 24 bits -> k * 2^-24, k in [1, 2^24] */
static inline float u32_to_uniform(uint32_t bits) {
  return (float)((bits >> 8) + 1u) * (1.0f / 16777216.0f);
}

/* Per-photon stream (philox.hpp:97-149). Counter {block, 0, id_lo, id_hi},
   key {seed_lo, seed_hi}. The buffered block is read front to back, which is
   the same sequence as the reference's shift register. */
typedef struct {
  uint32_t k0, k1, id_lo, id_hi;
  uint32_t block, avail;
  uint32_t buf[4];
} stream_t;

static inline void stream_init(stream_t* s, uint64_t seed, uint64_t id, uint32_t draws) {
  s->k0 = (uint32_t)seed;
  s->k1 = (uint32_t)(seed >> 32);
  s->id_lo = (uint32_t)id;
  s->id_hi = (uint32_t)(id >> 32);
  /* resume, philox.hpp:122-134 */
  s->block = (draws + 3u) / 4u;
  s->avail = 4u * s->block - draws;
  if (s->avail > 0) philox_block(s->block - 1u, 0u, s->id_lo, s->id_hi, s->k0, s->k1, s->buf);
}

static inline uint32_t stream_next(stream_t* s) {
  if (s->avail == 0) {
    philox_block(s->block++, 0u, s->id_lo, s->id_hi, s->k0, s->k1, s->buf);
    s->avail = 4;
  }
  return s->buf[4u - s->avail--];
}

static inline float stream_uniform(stream_t* s) { return u32_to_uniform(stream_next(s)); }
static inline uint32_t stream_draws(const stream_t* s) { return 4u * s->block - s->avail; }

/* ------------------------------------------------------------------------ */
/* Transcendentals (pmath.hpp)                                               */

static inline uint32_t f2u(float x) {
  uint32_t u;
  memcpy(&u, &x, 4);
  return u;
}
static inline float u2f(uint32_t u) {
  float x;
  memcpy(&x, &u, 4);
  return x;
}

/* pmath.hpp:62-82: exponent split, m in [sqrt(1/2), sqrt(2)), atanh series */
static float portable_log(float x) {
  const uint32_t bits = f2u(x);
  int e = (int)((bits >> 23) & 0xffu) - 127;
  float m = u2f((bits & 0x007fffffu) | 0x3f800000u);
  if (m > 1.41421356f) {
    m = 0.5f * m;
    ++e;
  }
  const float s = (m - 1.0f) / (m + 1.0f);
  const float s2 = s * s;
  const float poly =
      1.0f + s2 * (1.0f / 3.0f +
                   s2 * (1.0f / 5.0f + s2 * (1.0f / 7.0f + s2 * (1.0f / 9.0f + s2 * (1.0f / 11.0f)))));
  const float ln2_hi = 6.9313812256e-01f;
  const float ln2_lo = 9.0580006145e-06f;
  const float fe = (float)e;
  return fe * ln2_hi + (2.0f * s * poly + fe * ln2_lo);
}

/* pmath.hpp:87-119: exact quarter-turn reduction, Taylor polynomials */
static void portable_sincos_turns(float u, float* s, float* c) {
  const float v = 4.0f * u;
  float q = floorf(v);
  float r = v - q;
  if (r > 0.5f) {
    r -= 1.0f;
    q += 1.0f;
  }
  const float x = r * 1.57079632679489662f;
  const float x2 = x * x;
  const float sp =
      x * (1.0f + x2 * (-1.0f / 6.0f +
                        x2 * (1.0f / 120.0f + x2 * (-1.0f / 5040.0f + x2 * (1.0f / 362880.0f)))));
  const float cp =
      1.0f + x2 * (-0.5f + x2 * (1.0f / 24.0f +
                                 x2 * (-1.0f / 720.0f +
                                       x2 * (1.0f / 40320.0f + x2 * (-1.0f / 3628800.0f)))));
  const int quadrant = (int)q & 3;
  if (quadrant == 0) {
    *s = sp;
    *c = cp;
  } else if (quadrant == 1) {
    *s = cp;
    *c = -sp;
  } else if (quadrant == 2) {
    *s = -sp;
    *c = -cp;
  } else {
    *s = -cp;
    *c = sp;
  }
}

/* pmath.hpp:123-130: integer exponent / 3 guess, 4 Newton steps */
static float portable_cbrt(float x) {
  float y = u2f(f2u(x) / 3u + 0x2a514067u);
  for (int i = 0; i < 4; ++i) y = y - (y * y * y - x) / (3.0f * y * y);
  return y;
}

#ifdef ORACLE_PORTABLE_MATH
static inline float m_log(float x) { return portable_log(x); }
static inline void m_sincos_turns(float u, float* s, float* c) { portable_sincos_turns(u, s, c); }
static inline float m_cbrt(float x) { return portable_cbrt(x); }
#else
/* pmath.hpp:134-158, default build: the platform math library */
static const float kTwoPi = 6.28318530717958647692f; /* hd.hpp:36 */
static inline float m_log(float x) { return logf(x); }
static inline void m_sincos_turns(float u, float* s, float* c) {
  const float phi = kTwoPi * u;
  *s = sinf(phi);
  *c = cosf(phi);
}
static inline float m_cbrt(float x) { return cbrtf(x); }
#endif

/* ------------------------------------------------------------------------ */
/* Vectors (vec3.hpp)                                                        */

typedef struct {
  float x, y, z;
} v3;

static inline v3 v3_make(float x, float y, float z) {
  v3 r = {x, y, z};
  return r;
}
static inline v3 v3_add(v3 a, v3 b) { return v3_make(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 v3_sub(v3 a, v3 b) { return v3_make(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 v3_scale(float s, v3 a) { return v3_make(s * a.x, s * a.y, s * a.z); }
static inline v3 v3_neg(v3 a) { return v3_make(-a.x, -a.y, -a.z); }
static inline float v3_dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
/* vec3.hpp:29-30: normalize(a) = (1 / |a|) * a */
static inline v3 v3_normalize(v3 a) { return v3_scale(1.0f / sqrtf(v3_dot(a, a)), a); }

/* vec3.hpp:35-41 (Duff et al. 2017, branch-free) */
static inline void onb(v3 n, v3* t1, v3* t2) {
  const float sign = copysignf(1.0f, n.z);
  const float a = -1.0f / (sign + n.z);
  const float b = n.x * n.y * a;
  *t1 = v3_make(1.0f + sign * n.x * n.x * a, sign * b, -sign * n.x);
  *t2 = v3_make(b, sign + n.y * n.y * a, -n.y);
}

/* vec3.hpp:45-53 */
static inline v3 rotate_about(v3 axis, float cos_theta, float phi_turns) {
  v3 t1, t2;
  onb(axis, &t1, &t2);
  const float sin_theta = sqrtf(fmaxf(0.0f, 1.0f - cos_theta * cos_theta));
  float sin_phi, cos_phi;
  m_sincos_turns(phi_turns, &sin_phi, &cos_phi);
  const v3 d = v3_add(v3_add(v3_scale(sin_theta * cos_phi, t1), v3_scale(sin_theta * sin_phi, t2)),
                      v3_scale(cos_theta, axis));
  return v3_normalize(d);
}

/* ------------------------------------------------------------------------ */
/* Sampling (sampling.hpp)                                                   */

static inline float sample_exponential(float length, float u) { /* :17-19 */
  return is_inf(length) ? INFINITY : -length * m_log(u);
}

static inline v3 sample_isotropic(float u1, float u2) { /* :22-28 */
  const float cos_t = 2.0f * u1 - 1.0f;
  const float sin_t = sqrtf(fmaxf(0.0f, 1.0f - cos_t * cos_t));
  float sin_phi, cos_phi;
  m_sincos_turns(u2, &sin_phi, &cos_phi);
  return v3_make(sin_t * cos_phi, sin_t * sin_phi, cos_t);
}

static inline float sample_rayleigh_cos(float u) { /* :39-44, Cardano */
  const float q = 4.0f * u - 2.0f;
  const float s = sqrtf(q * q + 1.0f);
  const float a = (q >= 0.0f) ? m_cbrt(q + s) : 1.0f / m_cbrt(s - q);
  return fminf(1.0f, fmaxf(-1.0f, a - 1.0f / a));
}

static inline float sample_emission_time(float tau, float u) { /* :58-60 */
  return tau > 0.0f ? -tau * m_log(u) : 0.0f;
}

/* ------------------------------------------------------------------------ */
/* Fresnel and geometry                                                      */

/* fresnel.hpp:30-46: returns R; n1 == n2 means no interface */
static inline float fresnel_R(float n1, float n2, float cos_i) {
  if (n1 == n2) return 0.0f;
  const float eta = n1 / n2;
  const float sin2_t = eta * eta * fmaxf(0.0f, 1.0f - cos_i * cos_i);
  if (sin2_t >= 1.0f) return 1.0f;
  const float cos_t = sqrtf(1.0f - sin2_t);
  const float rs = (n1 * cos_i - n2 * cos_t) / (n1 * cos_i + n2 * cos_t);
  const float rp = (n2 * cos_i - n1 * cos_t) / (n2 * cos_i + n1 * cos_t);
  return 0.5f * (rs * rs + rp * rp);
}

/* fresnel.hpp:49-51 */
static inline v3 reflect(v3 d, v3 n) { return v3_sub(d, v3_scale(2.0f * v3_dot(d, n), n)); }

/* geometry.hpp:25-46: per axis only the face approached; strict < keeps the
   earlier axis on ties; negative rounding residue clamps to 0 */
static inline float box_exit(v3 p, v3 d, float hx, float hy, float hz, int* face) {
  float dist = INFINITY;
  int f = -1;
  if (d.x != 0.0f) {
    const int pos = d.x > 0.0f;
    const float t = ((pos ? hx : -hx) - p.x) / d.x;
    if (t < dist) {
      dist = t;
      f = pos ? 1 : 0;
    }
  }
  if (d.y != 0.0f) {
    const int pos = d.y > 0.0f;
    const float t = ((pos ? hy : -hy) - p.y) / d.y;
    if (t < dist) {
      dist = t;
      f = pos ? 3 : 2;
    }
  }
  if (d.z != 0.0f) {
    const int pos = d.z > 0.0f;
    const float t = ((pos ? hz : -hz) - p.z) / d.z;
    if (t < dist) {
      dist = t;
      f = pos ? 5 : 4;
    }
  }
  *face = f;
  return fmaxf(dist, 0.0f);
}

static inline v3 face_normal(int face) { /* geometry.hpp:49-53 */
  const float s = (face & 1) ? 1.0f : -1.0f;
  const int axis = face >> 1;
  return v3_make(axis == 0 ? s : 0.0f, axis == 1 ? s : 0.0f, axis == 2 ? s : 0.0f);
}

static inline v3 snap_to_face(v3 p, int face, float hx, float hy, float hz) { /* :57-63 */
  const float s = (face & 1) ? 1.0f : -1.0f;
  const int axis = face >> 1;
  return v3_make(axis == 0 ? s * hx : fminf(fmaxf(p.x, -hx), hx),
                 axis == 1 ? s * hy : fminf(fmaxf(p.y, -hy), hy),
                 axis == 2 ? s * hz : fminf(fmaxf(p.z, -hz), hz));
}

/* ------------------------------------------------------------------------ */
/* Transport (transport.hpp)                                                 */

typedef struct {
  v3 pos, dir;
  float t, path;
  uint32_t nb, ns;
} pstate;

static inline int happens(float u, float prob) { return u <= prob; } /* :41 */

/* transport.hpp:55-67. Draw order: [x, y, z if volume], cos, phi, time. */
static pstate start_photon(const oracle_params* p, stream_t* rng) {
  pstate s;
  if (p->source == ORACLE_SOURCE_POINT) {
    s.pos = v3_make(p->src_x, p->src_y, p->src_z);
  } else {
    const float u1 = stream_uniform(rng);
    const float u2 = stream_uniform(rng);
    const float u3 = stream_uniform(rng);
    s.pos = v3_make(p->hx * (2.0f * u1 - 1.0f), p->hy * (2.0f * u2 - 1.0f), p->hz * (2.0f * u3 - 1.0f));
  }
  const float ua = stream_uniform(rng);
  const float ub = stream_uniform(rng);
  s.dir = sample_isotropic(ua, ub);
  s.t = sample_emission_time(p->tau, stream_uniform(rng));
  s.path = 0.0f;
  s.nb = 0u;
  s.ns = 0u;
  return s;
}

/* transport.hpp:170-173 */
static inline uint32_t time_bin(float t, const oracle_tally_config* c) {
  const float f = t * ((float)c->time_bins / c->time_max);
  return (f >= (float)c->time_bins) ? c->time_bins : (uint32_t)f;
}

/* Likelihood-ratio scores of one history, in double (adjoint reference).
   The history itself is the FP32 one above; only the scores are f64. */
typedef struct {
  double s_ndet, s_nscint;
  uint32_t refl_hits;
  /* Readout baseline: b = w[bin(t)] at the hit (the weight if transmitted
     now); reflected readout decisions add s to the score and b s here,
     transmitted ones add nothing, which equals subtracting b s from g s for
     every readout decision (unbiased: E[s | state] = 0). */
  double base_ndet, base_nscint;
  const double* w;
  const oracle_tally_config* tc;
} adjscore;

/* d log P(decision) / d n1, d n2 for a Fresnel decision at cos_i (float
   inputs of the FP32 history), evaluated in double from
   R = (rs^2 + rp^2)/2, cos_t = sqrt(1 - (n1/n2)^2 (1 - cos_i^2)).
   No score for matched interfaces or total internal reflection, as decided
   in double: a hit that float32 rounds into total internal reflection (a
   certain reflection in the history) but that has cos_t^2 > 0 in double is
   scored as reflected. Near the critical angle dR/dn grows like 1/cos_t, so
   those few hits carry a finite share of the derivative. */
static void fresnel_score(float n1f, float n2f, float cif, int reflected, double* d1, double* d2) {
  *d1 = *d2 = 0.0;
  if (n1f == n2f) return;
  const double n1 = n1f, n2 = n2f, ci = cif;
  const double s2i = fmax(0.0, 1.0 - ci * ci);
  const double s2t = (n1 / n2) * (n1 / n2) * s2i;
  if (s2t >= 1.0) return;
  const double ct = sqrt(1.0 - s2t);
  const double A = n1 * ci - n2 * ct, B = n1 * ci + n2 * ct;
  const double C = n2 * ci - n1 * ct, D = n2 * ci + n1 * ct;
  const double rs = A / B, rp = C / D;
  const double R = 0.5 * (rs * rs + rp * rp);
  const double dct1 = -n1 * s2i / (n2 * n2 * ct);
  const double dct2 = n1 * n1 * s2i / (n2 * n2 * n2 * ct);
  const double drs1 = ((ci - n2 * dct1) * B - A * (ci + n2 * dct1)) / (B * B);
  const double drs2 = ((-ct - n2 * dct2) * B - A * (ct + n2 * dct2)) / (B * B);
  const double drp1 = ((-ct - n1 * dct1) * D - C * (ct + n1 * dct1)) / (D * D);
  const double drp2 = ((ci - n1 * dct2) * D - C * (ci + n1 * dct2)) / (D * D);
  const double dR1 = rs * drs1 + rp * drp1, dR2 = rs * drs2 + rp * drp2;
  /* 1 - R without cancellation: 2 n1 n2 ci ct (1/B^2 + 1/D^2) */
  const double T = 2.0 * n1 * n2 * ci * ct * (1.0 / (B * B) + 1.0 / (D * D));
  const double inv = reflected ? 1.0 / R : -1.0 / T;
  *d1 = dR1 * inv;
  *d2 = dR2 * inv;
}

/* transport.hpp:71-145. Returns 1 while alive, else sets *fate. `adj`
   (optional) collects the adjoint scores of the decisions taken. */
/* Per-step decision trace (test instrumentation): what the step did and how
   close its deciding comparisons came to going the other way. */
typedef struct {
  uint8_t code;  /* 0 absorb, 1 scatter, 2 + face boundary; | 0x80 if a Bernoulli
                    decision at the face came out true (reflected / survived) */
  float margin;  /* smallest relative gap of the comparisons that decided */
} steptrace;

static inline float rel_gap(float a, float b) {
  if (is_inf(a) || is_inf(b)) return INFINITY;
  const float m = fmaxf(fabsf(a), fabsf(b));
  return m > 0.0f ? fabsf(a - b) / m : 0.0f;
}

/* gap between the nearest and the second-nearest face along d (face choice) */
static float face_choice_gap(v3 p, v3 d, float hx, float hy, float hz) {
  float t[3] = {INFINITY, INFINITY, INFINITY};
  const float pc[3] = {p.x, p.y, p.z}, dc[3] = {d.x, d.y, d.z}, h[3] = {hx, hy, hz};
  for (int a = 0; a < 3; ++a)
    if (dc[a] != 0.0f) t[a] = ((dc[a] > 0.0f ? h[a] : -h[a]) - pc[a]) / dc[a];
  float lo = fminf(t[0], fminf(t[1], t[2]));
  float second = INFINITY;
  int used = 0;
  for (int a = 0; a < 3; ++a) {
    if (!used && t[a] == lo) {
      used = 1;
      continue;
    }
    second = fminf(second, t[a]);
  }
  return rel_gap(lo, second);
}

static int step_photon_adj(const oracle_params* p, stream_t* rng, pstate* s, int* fate,
                           adjscore* adj, steptrace* tr);
static int step_photon(const oracle_params* p, stream_t* rng, pstate* s, int* fate) {
  return step_photon_adj(p, rng, s, fate, NULL, NULL);
}

static int step_photon_adj(const oracle_params* p, stream_t* rng, pstate* s, int* fate,
                           adjscore* adj, steptrace* tr) {
  const float inv_speed = p->n_scint / kSpeedOfLight;
  const float s_abs = sample_exponential(p->abs_length, stream_uniform(rng));
  const float s_sca = sample_exponential(p->scat_length, stream_uniform(rng));
  int face;
  const float dist = box_exit(s->pos, s->dir, p->hx, p->hy, p->hz, &face);
  if (tr) {
    float m = fminf(rel_gap(s_abs, dist), rel_gap(s_sca, dist));
    if (fminf(s_abs, s_sca) < dist) m = fminf(m, rel_gap(s_abs, s_sca));
    if (!(s_abs < dist && s_abs <= s_sca) && !(s_sca < dist)) {
      m = fminf(m, face_choice_gap(s->pos, s->dir, p->hx, p->hy, p->hz));
      tr->code = (uint8_t)(2 + face);
    } else {
      tr->code = (s_abs < dist && s_abs <= s_sca) ? 0 : 1;
    }
    tr->margin = m;
  }

  if (s_abs < dist && s_abs <= s_sca) { /* bulk absorption, :81-87 */
    s->pos = v3_add(s->pos, v3_scale(s_abs, s->dir));
    s->t += s_abs * inv_speed;
    s->path += s_abs;
    *fate = ORACLE_FATE_ABSORBED_BULK;
    return 0;
  }
  if (s_sca < dist) { /* Rayleigh scattering, :90-98 */
    s->pos = v3_add(s->pos, v3_scale(s_sca, s->dir));
    s->t += s_sca * inv_speed;
    s->path += s_sca;
    const float u1 = stream_uniform(rng);
    const float u2 = stream_uniform(rng);
    s->dir = rotate_about(s->dir, sample_rayleigh_cos(u1), u2);
    ++s->ns;
    return 1;
  }
  /* boundary, :101-104 */
  s->pos = snap_to_face(v3_add(s->pos, v3_scale(dist, s->dir)), face, p->hx, p->hy, p->hz);
  s->t += dist * inv_speed;
  s->path += dist;
  ++s->nb;
  const v3 normal = face_normal(face);
  const float cos_i = v3_dot(s->dir, normal);

  if (face == 5) { /* readout +z, :108-118 */
    const float R = fresnel_R(p->n_scint, p->n_det, cos_i);
    const float u = stream_uniform(rng);
    const int refl = happens(u, R);
    if (tr) {
      tr->margin = fminf(tr->margin, fabsf(u - R));
      if (refl) tr->code |= 0x80;
    }
    if (adj && refl) {
      double d1, d2;
      fresnel_score(p->n_scint, p->n_det, cos_i, refl, &d1, &d2);
      const double b = adj->w[time_bin(s->t, adj->tc)];
      adj->s_nscint += d1;
      adj->s_ndet += d2;
      adj->base_nscint += b * d1;
      adj->base_ndet += b * d2;
    }
    if (refl) {
      s->dir = reflect(s->dir, normal);
      return 1;
    }
    *fate = ORACLE_FATE_DETECTED;
    return 0;
  }
  if (p->surface == ORACLE_SURFACE_POLISHED) { /* :120-128 */
    const float R = fresnel_R(p->n_scint, p->n_out, cos_i);
    const float u = stream_uniform(rng);
    const int refl = happens(u, R);
    if (tr) {
      tr->margin = fminf(tr->margin, fabsf(u - R));
      if (refl) tr->code |= 0x80;
    }
    if (adj) {
      double d1, d2;
      fresnel_score(p->n_scint, p->n_out, cos_i, refl, &d1, &d2);
      adj->s_nscint += d1;
    }
    if (refl) {
      s->dir = reflect(s->dir, normal);
      return 1;
    }
    *fate = ORACLE_FATE_ESCAPED;
    return 0;
  }
  if (p->surface == ORACLE_SURFACE_BLACK) { /* :129-132 */
    *fate = ORACLE_FATE_ABSORBED_SURFACE;
    return 0;
  }
  if (adj) ++adj->refl_hits;
  const float u_r = stream_uniform(rng);
  if (tr) {
    tr->margin = fminf(tr->margin, fabsf(u_r - p->reflectivity));
    if (happens(u_r, p->reflectivity)) tr->code |= 0x80;
  }
  if (!happens(u_r, p->reflectivity)) { /* :134-137 */
    *fate = ORACLE_FATE_ABSORBED_SURFACE;
    return 0;
  }
  if (p->surface == ORACLE_SURFACE_SPECULAR) { /* :138-139 */
    s->dir = reflect(s->dir, normal);
  } else { /* Lambertian about the inward normal, :140-143, sampling.hpp:53-55 */
    const float u1 = stream_uniform(rng);
    const float u2 = stream_uniform(rng);
    s->dir = rotate_about(v3_neg(normal), sqrtf(u1), u2);
  }
  return 1;
}

/* Emission from a genstep (no counterpart in the reference, whose README.md
   lists generation from gensteps as future work). Draw order: [u along the
   segment, only if the segment or its time span is non-degenerate], cos, phi,
   emission delay. A degenerate genstep with t0 = 0 therefore draws and
   computes exactly what the reference's point source at (x0, y0, z0) does
   (transport.hpp:55-67): same words, and t = 0 + delay = delay.
   Cherenkov (cone_cos > 0): u along the segment, phi; the direction on the
   cone of cone_cos about the segment direction (rotate_about, vec3.hpp:35-53,
   as for Rayleigh scattering), no delay. */
static pstate start_genstep_photon(const oracle_params* p, const oracle_genstep* g,
                                   stream_t* rng) {
  pstate s;
  if (g->cone_cos > 0.0f) {
    const float u = stream_uniform(rng);
    const float up = stream_uniform(rng);
    const float dx = g->x1 - g->x0, dy = g->y1 - g->y0, dz = g->z1 - g->z0;
    s.pos = v3_make(g->x0 + u * dx, g->y0 + u * dy, g->z0 + u * dz);
    const float inv = 1.0f / sqrtf(dx * dx + dy * dy + dz * dz);
    s.dir = rotate_about(v3_make(inv * dx, inv * dy, inv * dz), g->cone_cos, up);
    s.t = g->t0 + u * (g->t1 - g->t0);
    s.path = 0.0f;
    s.nb = 0u;
    s.ns = 0u;
    return s;
  }
  s.pos = v3_make(g->x0, g->y0, g->z0);
  float t0 = g->t0;
  if (g->x0 != g->x1 || g->y0 != g->y1 || g->z0 != g->z1 || g->t0 != g->t1) {
    const float u = stream_uniform(rng);
    s.pos = v3_make(g->x0 + u * (g->x1 - g->x0), g->y0 + u * (g->y1 - g->y0),
                    g->z0 + u * (g->z1 - g->z0));
    t0 = g->t0 + u * (g->t1 - g->t0);
  }
  const float ua = stream_uniform(rng);
  const float ub = stream_uniform(rng);
  s.dir = sample_isotropic(ua, ub);
  s.t = t0 + sample_emission_time(p->tau, stream_uniform(rng));
  s.path = 0.0f;
  s.nb = 0u;
  s.ns = 0u;
  return s;
}

static void transport_genstep_photon(const oracle_params* p, const oracle_genstep* g,
                                     uint64_t id, oracle_result* r) {
  stream_t rng;
  stream_init(&rng, p->seed, id, 0u);
  pstate s = start_genstep_photon(p, g, &rng);
  int fate = ORACLE_FATE_MAX_STEPS;
  for (uint32_t step = 0; step < p->max_steps; ++step) {
    if (!step_photon(p, &rng, &s, &fate)) break;
  }
  r->fate = (uint8_t)fate;
  r->n_boundary = s.nb;
  r->n_scatter = s.ns;
  r->t = s.t;
  r->x = s.pos.x;
  r->y = s.pos.y;
  r->z = s.pos.z;
  r->path = s.path;
  r->draws = stream_draws(&rng);
}

/* transport.hpp:151-159 */
static void transport_photon(const oracle_params* p, uint64_t id, oracle_result* r) {
  stream_t rng;
  stream_init(&rng, p->seed, id, 0u);
  pstate s = start_photon(p, &rng);
  int fate = ORACLE_FATE_MAX_STEPS;
  for (uint32_t step = 0; step < p->max_steps; ++step) {
    if (!step_photon(p, &rng, &s, &fate)) break;
  }
  r->fate = (uint8_t)fate;
  r->n_boundary = s.nb;
  r->n_scatter = s.ns;
  r->t = s.t;
  r->x = s.pos.x;
  r->y = s.pos.y;
  r->z = s.pos.z;
  r->path = s.path;
  r->draws = stream_draws(&rng);
}

/* transport.hpp:176-181 */
static inline uint32_t xy_bin(float x, float y, const oracle_params* p, const oracle_tally_config* c) {
  const float nb = (float)c->xy_bins;
  const uint32_t ix = (uint32_t)fminf((x + p->hx) / (2.0f * p->hx) * nb, nb - 1.0f);
  const uint32_t iy = (uint32_t)fminf((y + p->hy) / (2.0f * p->hy) * nb, nb - 1.0f);
  return iy * c->xy_bins + ix;
}

/* ------------------------------------------------------------------------ */
/* Exported entry points                                                     */

ORACLE_EXPORT int oracle_portable_math(void) {
#ifdef ORACLE_PORTABLE_MATH
  return 1;
#else
  return 0;
#endif
}

ORACLE_EXPORT void oracle_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2],
                                        uint32_t out[4]) {
  philox_block(ctr[0], ctr[1], ctr[2], ctr[3], key[0], key[1], out);
}

ORACLE_EXPORT float oracle_u32_to_uniform(uint32_t bits) { return u32_to_uniform(bits); }

/* Words skip .. skip+n-1 of photon `id`'s stream, resuming from `skip`. */
ORACLE_EXPORT void oracle_stream_words(uint64_t seed, uint64_t id, uint32_t skip, uint32_t n,
                                       uint32_t* out) {
  stream_t s;
  stream_init(&s, seed, id, skip);
  for (uint32_t i = 0; i < n; ++i) out[i] = stream_next(&s);
}

ORACLE_EXPORT float oracle_log(float x) { return m_log(x); }
ORACLE_EXPORT void oracle_sincos_turns(float u, float* s, float* c) { m_sincos_turns(u, s, c); }
ORACLE_EXPORT float oracle_cbrt(float x) { return m_cbrt(x); }
ORACLE_EXPORT float oracle_portable_log(float x) { return portable_log(x); }
ORACLE_EXPORT void oracle_portable_sincos_turns(float u, float* s, float* c) {
  portable_sincos_turns(u, s, c);
}
ORACLE_EXPORT float oracle_portable_cbrt(float x) { return portable_cbrt(x); }
ORACLE_EXPORT float oracle_fresnel_R(float n1, float n2, float cos_i) {
  return fresnel_R(n1, n2, cos_i);
}
ORACLE_EXPORT float oracle_box_exit(const float p[3], const float d[3], float hx, float hy,
                                    float hz, int* face) {
  return box_exit(v3_make(p[0], p[1], p[2]), v3_make(d[0], d[1], d[2]), hx, hy, hz, face);
}
ORACLE_EXPORT float oracle_rayleigh_cos(float u) { return sample_rayleigh_cos(u); }
/* sample_rayleigh_cos of every 24-bit uniform: out[k] for u = (k + 1) 2^-24
 * (philox.hpp:82-84, sampling.hpp:39-44); out holds 2^24 floats. */
ORACLE_EXPORT void oracle_rayleigh_table(float* out) {
  for (uint32_t k = 0; k < (1u << 24); ++k)
    out[k] = sample_rayleigh_cos(u32_to_uniform(k << 8));
}
ORACLE_EXPORT void oracle_rotate_about(const float axis[3], float cos_theta, float phi_turns,
                                       float out[3]) {
  const v3 r = rotate_about(v3_make(axis[0], axis[1], axis[2]), cos_theta, phi_turns);
  out[0] = r.x;
  out[1] = r.y;
  out[2] = r.z;
}
ORACLE_EXPORT void oracle_isotropic(float u1, float u2, float out[3]) {
  const v3 r = sample_isotropic(u1, u2);
  out[0] = r.x;
  out[1] = r.y;
  out[2] = r.z;
}

static int thread_count(int threads) {
#ifdef _OPENMP
  return threads > 0 ? threads : omp_get_max_threads();
#else
  (void)threads;
  return 1;
#endif
}

ORACLE_EXPORT int oracle_max_threads(void) { return thread_count(0); }

/* cpu_backend.cpp:88-113: records for photons first_id .. first_id+n-1.
   `draws` (optional) receives each photon's total Philox word count. */
/* Step-by-step decision trace of photon `id` (test instrumentation): for each
   of the first `cap` steps, the step's outcome code and the smallest relative
   margin of the comparisons that decided it (see steptrace). Returns the
   number of steps the photon took (may exceed cap). With `state` non-NULL it
   also receives position, direction and time after each step (7 floats). Used to show that a
   photon whose discrete history differs between two math builds diverged at
   a near-tie, i.e. through last-bit rounding. */
ORACLE_EXPORT uint32_t oracle_trace(const oracle_params* p, uint64_t id, uint32_t cap,
                                    uint8_t* codes, float* margins, float* state) {
  stream_t rng;
  stream_init(&rng, p->seed, id, 0u);
  pstate s = start_photon(p, &rng);
  int fate = ORACLE_FATE_MAX_STEPS;
  uint32_t step = 0;
  while (step < p->max_steps) {
    steptrace tr = {0, INFINITY};
    const int alive = step_photon_adj(p, &rng, &s, &fate, NULL, &tr);
    if (step < cap) {
      codes[step] = tr.code;
      margins[step] = tr.margin;
      if (state) {
        float* o = state + 7 * (size_t)step;
        o[0] = s.pos.x, o[1] = s.pos.y, o[2] = s.pos.z;
        o[3] = s.dir.x, o[4] = s.dir.y, o[5] = s.dir.z;
        o[6] = s.t;
      }
    }
    ++step;
    if (!alive) break;
  }
  return step;
}

/* The emitted state of photon `id` (position, direction, time: 7 floats),
   i.e. the state before the first step of oracle_trace (test
   instrumentation: the first step's flight starts here). */
ORACLE_EXPORT void oracle_emitted(const oracle_params* p, uint64_t id, float* state) {
  stream_t rng;
  stream_init(&rng, p->seed, id, 0u);
  const pstate s = start_photon(p, &rng);
  state[0] = s.pos.x, state[1] = s.pos.y, state[2] = s.pos.z;
  state[3] = s.dir.x, state[4] = s.dir.y, state[5] = s.dir.z;
  state[6] = s.t;
}

ORACLE_EXPORT void oracle_run_records(const oracle_params* p, uint64_t n, uint64_t first_id,
                                      int threads, uint8_t* fate, uint32_t* n_boundary,
                                      uint32_t* n_scatter, float* t, float* x, float* y, float* z,
                                      float* path, uint32_t* draws) {
  const int64_t count = (int64_t)n;
  const int nt = thread_count(threads);
#pragma omp parallel for schedule(dynamic, 4096) num_threads(nt)
  for (int64_t i = 0; i < count; ++i) {
    oracle_result r;
    transport_photon(p, first_id + (uint64_t)i, &r);
    fate[i] = r.fate;
    n_boundary[i] = r.n_boundary;
    n_scatter[i] = r.n_scatter;
    t[i] = r.t;
    x[i] = r.x;
    y[i] = r.y;
    z[i] = r.z;
    path[i] = r.path;
    if (draws) draws[i] = r.draws;
  }
}

/* Size of the u64 tally buffer: [fates 5 | time T+1 | xy X^2 | boundary sum,
   scatter sum] (the layout of the reference's device tally,
   gpu_backend.cu:56-66, which also matches Tally's fields in order). */
ORACLE_EXPORT uint64_t oracle_tally_words(const oracle_tally_config* c) {
  return 5u + (uint64_t)c->time_bins + 1u + (uint64_t)c->xy_bins * c->xy_bins + 2u;
}

/* cpu_backend.cpp:28-36 + 115-147: every fate counted, histograms for
   detected photons only, integer sums over all photons. */
/* Photons of gensteps g = 0 .. n_g-1: photon k of genstep g has index
   i = n_0 + ... + n_{g-1} + k and Philox id first_id + i. Records (arrays may
   be NULL) are written at i; the tally (NULL to skip) counts all photons'
   fates and sums and the detected photons' histograms. */
ORACLE_EXPORT void oracle_run_gensteps(const oracle_params* p, const oracle_genstep* g,
                                       uint32_t n_g, uint64_t first_id,
                                       const oracle_tally_config* c, uint64_t* tally,
                                       uint8_t* fate, uint32_t* n_boundary, uint32_t* n_scatter,
                                       float* t, float* x, float* y, float* z, float* path,
                                       int threads) {
  uint64_t total = 0;
  for (uint32_t k = 0; k < n_g; ++k) total += g[k].n_photons;
  uint64_t* off = (uint64_t*)malloc(sizeof(uint64_t) * ((size_t)n_g + 1));
  off[0] = 0;
  for (uint32_t k = 0; k < n_g; ++k) off[k + 1] = off[k] + g[k].n_photons;
  uint64_t words = 0, time_off = 0, xy_off = 0, sums_off = 0;
  if (tally) {
    words = oracle_tally_words(c);
    time_off = 5u;
    xy_off = 5u + c->time_bins + 1u;
    sums_off = xy_off + (uint64_t)c->xy_bins * c->xy_bins;
    for (uint64_t i = 0; i < words; ++i) tally[i] = 0;
  }
  const int nt = thread_count(threads);
#pragma omp parallel for schedule(dynamic, 4096) num_threads(nt)
  for (int64_t i = 0; i < (int64_t)total; ++i) {
    uint32_t lo = 0, hi = n_g; /* genstep: off[lo] <= i < off[lo + 1] */
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) / 2;
      if (off[mid] <= (uint64_t)i) lo = mid; else hi = mid;
    }
    while (off[lo + 1] <= (uint64_t)i) ++lo; /* skip empty gensteps */
    oracle_result r;
    transport_genstep_photon(p, &g[lo], first_id + (uint64_t)i, &r);
    if (fate) {
      fate[i] = r.fate;
      n_boundary[i] = r.n_boundary;
      n_scatter[i] = r.n_scatter;
      t[i] = r.t;
      x[i] = r.x;
      y[i] = r.y;
      z[i] = r.z;
      path[i] = r.path;
    }
    if (tally) {
#pragma omp atomic
      tally[r.fate] += 1;
#pragma omp atomic
      tally[sums_off] += r.n_boundary;
#pragma omp atomic
      tally[sums_off + 1] += r.n_scatter;
      if (r.fate == ORACLE_FATE_DETECTED) {
#pragma omp atomic
        tally[time_off + time_bin(r.t, c)] += 1;
#pragma omp atomic
        tally[xy_off + xy_bin(r.x, r.y, p, c)] += 1;
      }
    }
  }
  free(off);
}

ORACLE_EXPORT void oracle_run_tally(const oracle_params* p, const oracle_tally_config* c,
                                    uint64_t n, uint64_t first_id, int threads, uint64_t* out) {
  const uint64_t words = oracle_tally_words(c);
  const uint64_t time_off = 5u, xy_off = 5u + c->time_bins + 1u;
  const uint64_t sums_off = xy_off + (uint64_t)c->xy_bins * c->xy_bins;
  for (uint64_t i = 0; i < words; ++i) out[i] = 0;
  const int64_t count = (int64_t)n;
  const int nt = thread_count(threads);
#pragma omp parallel num_threads(nt)
  {
    uint64_t local[8192];
    uint64_t* loc = local;
    int heap = 0;
    if (words > 8192) { /* large binnings: fall back to atomics on the shared buffer */
      heap = 1;
      loc = out;
    } else {
      for (uint64_t i = 0; i < words; ++i) local[i] = 0;
    }
#pragma omp for schedule(dynamic, 4096) nowait
    for (int64_t i = 0; i < count; ++i) {
      oracle_result r;
      transport_photon(p, first_id + (uint64_t)i, &r);
      if (heap) {
#pragma omp atomic
        loc[r.fate] += 1;
#pragma omp atomic
        loc[sums_off] += r.n_boundary;
#pragma omp atomic
        loc[sums_off + 1] += r.n_scatter;
        if (r.fate == ORACLE_FATE_DETECTED) {
#pragma omp atomic
          loc[time_off + time_bin(r.t, c)] += 1;
#pragma omp atomic
          loc[xy_off + xy_bin(r.x, r.y, p, c)] += 1;
        }
      } else {
        loc[r.fate] += 1;
        loc[sums_off] += r.n_boundary;
        loc[sums_off + 1] += r.n_scatter;
        if (r.fate == ORACLE_FATE_DETECTED) {
          loc[time_off + time_bin(r.t, c)] += 1;
          loc[xy_off + xy_bin(r.x, r.y, p, c)] += 1;
        }
      }
    }
    if (!heap) {
#pragma omp critical
      for (uint64_t i = 0; i < words; ++i) out[i] += local[i];
    }
  }
}

ORACLE_EXPORT uint32_t oracle_time_bin(float t, const oracle_tally_config* c) {
  return time_bin(t, c);
}
ORACLE_EXPORT uint32_t oracle_xy_bin(float x, float y, const oracle_params* p,
                                     const oracle_tally_config* c) {
  return xy_bin(x, y, p, c);
}

/* Adjoint reference: for parameter sets e (photons[e] photons at ids
   first_id + offset_e ..), grad[e][j] = (1/N_e) sum over detected photons of
   w[e][bin(t)] S_j (+ the arrival-time term for n_scint), with
   j = abs_length, scat_length, reflectivity, n_det, n_scint and the scores
     S_abs = path / L_a^2            (detected photons were not absorbed)
     S_sca = path / L_s^2 - n_scatter / L_s
     S_R   = reflector hits / R
     S_n   = sum of Fresnel decision scores
   and, for n_scint, (w[bin(t)] - w[bin(x)]) (path / c) / tau with
   x = path * (n_scint / c): the emission time t_e ~ Exp(tau) absorbs a shift
   of the time of flight. Histories are the FP32 ones of transport_photon;
   scores and sums are double. */
ORACLE_EXPORT void oracle_adjoint(const oracle_params* params, const uint64_t* photons,
                                  uint32_t n_events, uint64_t first_id, uint32_t time_bins,
                                  float time_max, const double* weights, double* grad, int threads) {
  const oracle_tally_config tc = {time_bins, time_max, 1u};
  uint64_t offset = 0;
  for (uint32_t e = 0; e < n_events; ++e) {
    const oracle_params* p = params + e;
    const double* w = weights + (size_t)e * (time_bins + 1u);
    const int64_t n = (int64_t)photons[e];
    double g0 = 0, g1 = 0, g2 = 0, g3 = 0, g4 = 0;
    const int nt = thread_count(threads);
#pragma omp parallel for schedule(dynamic, 4096) num_threads(nt) reduction(+ : g0, g1, g2, g3, g4)
    for (int64_t i = 0; i < n; ++i) {
      stream_t rng;
      stream_init(&rng, p->seed, first_id + offset + (uint64_t)i, 0u);
      pstate s = start_photon(p, &rng);
      int fate = ORACLE_FATE_MAX_STEPS;
      adjscore adj = {0.0, 0.0, 0u, 0.0, 0.0, w, &tc};
      for (uint32_t step = 0; step < p->max_steps; ++step) {
        if (!step_photon_adj(p, &rng, &s, &fate, &adj, NULL)) break;
      }
      g3 -= adj.base_ndet;
      g4 -= adj.base_nscint;
      if (fate != ORACLE_FATE_DETECTED) continue;
      const double path = s.path;
      const double wt = w[time_bin(s.t, &tc)];
      if (!is_inf(p->abs_length)) g0 += wt * (path / ((double)p->abs_length * p->abs_length));
      if (!is_inf(p->scat_length))
        g1 += wt * (path / ((double)p->scat_length * p->scat_length) -
                    (double)s.ns / p->scat_length);
      if (p->surface == ORACLE_SURFACE_SPECULAR || p->surface == ORACLE_SURFACE_LAMBERTIAN)
        g2 += wt * ((double)adj.refl_hits / p->reflectivity);
      g3 += wt * adj.s_ndet;
      double tterm = 0.0;
      if (p->tau > 0.0f) {
        const float x = s.path * (p->n_scint / kSpeedOfLight);
        tterm = (wt - w[time_bin(x, &tc)]) * (path / 299.792458 / p->tau);
      }
      g4 += wt * adj.s_nscint + tterm;
    }
    const double inv = n ? 1.0 / (double)n : 0.0;
    grad[5 * e + 0] = g0 * inv;
    grad[5 * e + 1] = g1 * inv;
    grad[5 * e + 2] = g2 * inv;
    grad[5 * e + 3] = g3 * inv;
    grad[5 * e + 4] = g4 * inv;
    offset += photons[e];
  }
}
