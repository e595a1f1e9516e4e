/*
 * optphot_oracle.h -- plain-C declarations of the oracle (test infrastructure
 * only; see optphot_oracle.c). The structs mirror the reference's POD types:
 * SimParams (include/optphot/params.hpp:40-63, 72 bytes, no padding) and
 * TallyConfig (include/optphot/transport.hpp:164-168).
 */
#ifndef OPTPHOT_ORACLE_H
#define OPTPHOT_ORACLE_H
#include <stdint.h>

enum { ORACLE_SURFACE_POLISHED = 0, ORACLE_SURFACE_BLACK = 1, ORACLE_SURFACE_SPECULAR = 2,
       ORACLE_SURFACE_LAMBERTIAN = 3 };
enum { ORACLE_SOURCE_POINT = 0, ORACLE_SOURCE_VOLUME = 1 };
enum { ORACLE_FATE_DETECTED = 0, ORACLE_FATE_ABSORBED_BULK = 1, ORACLE_FATE_ABSORBED_SURFACE = 2,
       ORACLE_FATE_ESCAPED = 3, ORACLE_FATE_MAX_STEPS = 4 };

typedef struct {
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
} oracle_params;

typedef struct {
  uint32_t time_bins;
  float time_max;
  uint32_t xy_bins;
} oracle_tally_config;

/* A step summary: photons are emitted uniformly along the segment
 * (x0,y0,z0,t0) -> (x1,y1,z1,t1). cone_cos == 0: scintillation, isotropic
 * with an Exp(tau) emission delay; cone_cos in (0, 1]: Cherenkov, on the cone
 * of that cosine about the segment direction, no delay. 40 bytes. */
typedef struct {
  float x0, y0, z0, t0;
  float x1, y1, z1, t1;
  uint32_t n_photons;
  float cone_cos;
} oracle_genstep;

typedef struct {
  uint8_t fate;
  uint32_t n_boundary, n_scatter;
  float t, x, y, z, path;
  uint32_t draws;
} oracle_result;

#endif
