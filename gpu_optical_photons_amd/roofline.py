"""Algorithmic work of the transport, for roofline fractions.

The count is of the reference algorithm's operations, not of the
instructions a kernel happens to issue. Rule: every IEEE f32 operation
(+ - * /, sqrt, min/max, compare, select, negation, int<->float conversion)
and every 32-bit integer operation counts 1; a 32x32->64 multiply counts 2
(low and high halves). Transcendentals count as the operations of the
reference's portable implementations. The per-branch costs below were counted
by hand from the reference's formulas and are frozen here:

  Philox4x32-10 block      philox.hpp:54-76   10 x (4 mul halves + 4 xor) + 18 key adds = 98
  uniform draw             philox.hpp:82-84   98 / 4 + 4 (shift, add, cvt, mul)        = 28.5
  log                      pmath.hpp:62-82                                            29
  sincos_turns             pmath.hpp:87-119                                           33
  cbrt                     pmath.hpp:123-130  2 + 4 Newton x 7                         30
  exponential distance     sampling.hpp:17-19 inf test + neg + mul + log               32 (1 if disabled)
  box exit                 geometry.hpp:25-46 3 axes x 9 + clamp                       28
  event decision           transport.hpp:81-90                                         4
  advance                  transport.hpp:82-84 (pos 6, t 2, path 1, n/c 1)             10
  boundary common          advance + snap 10 + count 1 + normal 5 + dot 5              31
  Fresnel                  fresnel.hpp:30-46                                          28
  reflect                  fresnel.hpp:49-51                                          12
  rotate_about             vec3.hpp:35-53 (basis 17, sin 4, sincos 33, sum 17, norm 10) 81
  Rayleigh cos             sampling.hpp:39-44 (q 2, sqrt 3, cbrt 30, select/inv 3, clamp 4) 42
  emission, point          transport.hpp:55-67 (3 draws, isotropic 42, time 32)         160
  emission, volume         + 3 draws + sample_in_box 9                                 255

Per step: common = 2 draws + 2 exponential distances + box exit + decision;
then absorb +10; scatter +10 +2 draws +42 +81 +1; every boundary hit +31;
readout and polished hits +28 (Fresnel) +28.5 (draw) +1 (test), +12 if
reflected; black +2; specular / Lambertian +28.5 +1 +2, then +12 (specular
reflect) or +57 +1 +3 +81 (Lambertian reflect).
"""
from __future__ import annotations

import math

# MI355X peaks (published): 256 CUs at 2.4 GHz, FP32 vector 157.3 TFLOP/s
# counting an FMA as 2 FLOPs. A CU executes 128 FP32 lane-operations per clock
# without packed math: the library's VALU calibration kernel (8 independent
# plain v_fma_f32 chains per lane) issues 9.0e11 wave64 instructions/s, i.e.
# 1.46 per CU per clock at 2.4 GHz, more than a 64-lane-per-clock CU could.
# So the ceiling for this code, which has one algorithmic op per instruction
# (no contraction), is 256 x 128 x 2.4e9 = 78.6e12 lane-ops/s.
FP32_VECTOR_PEAK_TFLOPS = 157.3
VALU_ISSUE_PEAK_TOPS = 256 * 128 * 2.4e9 / 1e12  # 78.6 lane-ops per second (x 1e12)
HBM_PEAK_GBS = 8000.0
HBM_PEAK_BYTES_PER_S = HBM_PEAK_GBS * 1e9

DRAW = 98.0 / 4.0 + 4.0
LOG = 29.0
EXPD = 3.0 + LOG  # inf test, negation, multiply, log
BOX = 28.0
DECIDE = 4.0
ADVANCE = 10.0
BOUNDARY = 31.0
FRESNEL = 28.0
REFLECT = 12.0
ROTATE = 81.0
RAYLEIGH_COS = 42.0
ISOTROPIC = 42.0
EMIT_POINT = 3 * DRAW + ISOTROPIC + 1.0 + 2.0 + LOG
EMIT_VOLUME = EMIT_POINT + 3 * DRAW + 9.0


def step_common(abs_length: float, scat_length: float) -> float:
    e_abs = 1.0 if math.isinf(abs_length) else EXPD
    e_sca = 1.0 if math.isinf(scat_length) else EXPD
    return 2 * DRAW + e_abs + e_sca + BOX + DECIDE


def algorithmic_ops(params, counters: dict, fate_counts, emitted_volume=None) -> float:
    """Total operations of the reference algorithm for a run, from the
    kernel's branch-event counters and the run's fate counts. For a batch of
    events with different parameters pass params=None: the counters'
    abs_on_steps / scat_on_steps give the steps that sampled each distance,
    and emitted_volume the photons of volume-source events (the rest are
    point-source emissions)."""
    c = counters
    steps = c["useful_lane_steps"]
    detected = int(fate_counts[0])
    escaped = int(fate_counts[3])
    absorbed_surface = int(fate_counts[2])
    if params is not None:
        abs_on = 0 if math.isinf(params.abs_length) else steps
        scat_on = 0 if math.isinf(params.scat_length) else steps
        if emitted_volume is None:
            emitted_volume = c["emitted"] if params.source == 1 else 0
    else:
        abs_on, scat_on = c["abs_on_steps"], c["scat_on_steps"]
        emitted_volume = emitted_volume or 0
    ops = steps * (2 * DRAW + BOX + DECIDE)
    ops += abs_on * EXPD + (steps - abs_on) * 1.0
    ops += scat_on * EXPD + (steps - scat_on) * 1.0
    ops += c["absorb"] * ADVANCE
    ops += c["scatter"] * (ADVANCE + 2 * DRAW + RAYLEIGH_COS + ROTATE + 1.0)
    hits = c["readout"] + c["polished"] + c["black"] + c["reflector"]
    ops += hits * BOUNDARY
    fres = c["readout"] + c["polished"]
    ops += fres * (FRESNEL + DRAW + 1.0)
    ops += (c["readout"] - detected + c["polished"] - escaped) * REFLECT
    ops += c["black"] * 2.0
    ops += c["reflector"] * (DRAW + 1.0 + 2.0)
    refl_survived = c["reflector"] - (absorbed_surface - c["black"])
    lam = c["lambert_reflect"]
    ops += (refl_survived - lam) * REFLECT
    ops += lam * (2 * DRAW + 1.0 + 3.0 + ROTATE)
    ops += emitted_volume * EMIT_VOLUME + (c["emitted"] - emitted_volume) * EMIT_POINT
    return float(ops)


def roofline(ops: float, kernel_ms: float, kernel: str, note: str = "") -> dict:
    """The roofline block of a launch: algorithmic operations over its time,
    against the FP32 vector peak (and the non-packed lane-op ceiling)."""
    achieved = ops / (kernel_ms * 1e-3) / 1e12
    out = {"bound": "valu", "achieved": achieved, "peak": FP32_VECTOR_PEAK_TFLOPS,
           "unit": "TFLOP/s", "frac": achieved / FP32_VECTOR_PEAK_TFLOPS, "kernel": kernel,
           "kernel_ms": kernel_ms, "algorithmic_ops_per_launch": ops,
           "frac_of_valu_issue_peak": achieved / VALU_ISSUE_PEAK_TOPS}
    if note:
        out["note"] = note
    return out


def photon_steps(counters: dict) -> int:
    return int(counters["useful_lane_steps"])


# Device-memory bytes of the per-interaction path (photon state in HBM as 12
# 4-byte arrays per slot), per event of the workload:
SORTED_SCAN = 4            # classify reads every slot's photon index each iteration
SORTED_STATE_IN = 44       # a step of a photon already in flight: p, d, t, path, nb, ns, draws
SORTED_STATE_OUT = 28      # after a step: p, t, path, draws and nb or ns
SORTED_EMIT_OUT = 20       # a new photon also writes idx, d and both counts
SORTED_ROT_ITEM = 12 + 12 + 12  # queue entry written and read, new direction written
SORTED_RAYLEIGH_DIR = 12   # the rotation kernel reads a scattered photon's direction
SORTED_FRES_ITEM = 8 + 8 + 4 + 8 + 4  # entry written and read, d component, counts, write-back
SORTED_DETECTED = 12       # t, x, y of a detected photon for the histograms


def sorted_path_bytes(counters: dict, n: int, slots: int, iterations: int,
                      detected: int) -> float:
    """Algorithmic device-memory bytes of one run of the per-interaction path,
    from the branch-event counts of the same photons (every step counted as
    if it ran in the sorted kernels; the history-style tail takes < 1 % of the
    steps)."""
    c = counters
    steps = c["useful_lane_steps"]
    rot = c["scatter"] + c["lambert_reflect"]
    fres = c["readout"] + c["polished"]
    return float(iterations * slots * SORTED_SCAN + (steps - n) * SORTED_STATE_IN
                 + steps * SORTED_STATE_OUT + n * SORTED_EMIT_OUT + rot * SORTED_ROT_ITEM
                 + c["scatter"] * SORTED_RAYLEIGH_DIR + fres * SORTED_FRES_ITEM
                 + detected * SORTED_DETECTED)
