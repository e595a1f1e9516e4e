"""Batched events: many events with different geometry and surface parameters
and Poisson photon counts, transported in one persistent launch.

Event e uses the run seed and photon ids first_id + offset_e ..
first_id + offset_e + N_e - 1 (offset_e = N_0 + ... + N_{e-1}), so each event
is exactly the reference's run_cpu_tally(params[e], ..., N_e, first_id +
offset_e) (backend.hpp:78-79; the Philox counter word for the photon id,
philox.hpp:106, carries the offset). Per event the launch returns the 5 fate
counts and an arrival-time histogram of detected photons with an overflow bin.
"""
from __future__ import annotations

import ctypes
import dataclasses
import math

import numpy as np

from . import backend
from .params import NUM_FATES, SimParams, Source, Surface

WAVEFRONT, PLAIN = 0, 1


@dataclasses.dataclass
class EventBatch:
    params: list          # SimParams per event
    photons: np.ndarray   # u64 photons per event

    @property
    def n_events(self):
        return len(self.params)

    @property
    def offsets(self):
        return np.concatenate([[0], np.cumsum(self.photons, dtype=np.uint64)]).astype(np.uint64)

    @property
    def total(self):
        return int(self.photons.sum())


def make_events(n_events: int = 4096, seed: int = 2024, run_seed: int = 1,
                photons_per_mev: float = 1e4, max_steps: int = 10000) -> EventBatch:
    """The batched-events workload: per event box dimensions ~U[10, 100] mm,
    surface model uniform over {polished to n_out 1.0, black, specular with
    R ~ U[0.85, 0.99], Lambertian with R ~ U[0.85, 0.99]}, n_det ~ U[1.4, 1.6],
    abs_length ~ U[300, 3000] mm, scat_length infinite or ~U[100, 1000] mm
    (50/50), a point source uniform in the box, N ~ Poisson(photons_per_mev x E)
    with E ~ U[0.1, 3] MeV; n_scint 1.63 and tau 4 ns. Drawn from a numpy
    generator seeded with `seed`; all events share the run seed `run_seed`."""
    rng = np.random.default_rng(seed)
    params = []
    counts = np.zeros(n_events, np.uint64)
    for e in range(n_events):
        size = rng.uniform(10.0, 100.0, 3)
        surf = int(rng.integers(0, 4))
        refl = float(rng.uniform(0.85, 0.99)) if surf in (2, 3) else 0.95
        n_det = float(rng.uniform(1.4, 1.6))
        abs_len = float(rng.uniform(300.0, 3000.0))
        scat = math.inf if rng.random() < 0.5 else float(rng.uniform(100.0, 1000.0))
        h = size / 2.0
        src = rng.uniform(-h, h)
        energy = rng.uniform(0.1, 3.0)
        counts[e] = rng.poisson(photons_per_mev * energy)
        params.append(SimParams(hx=float(h[0]), hy=float(h[1]), hz=float(h[2]), n_scint=1.63,
                                n_det=n_det, n_out=1.0, abs_length=abs_len, scat_length=scat,
                                surface=surf, reflectivity=refl, source=int(Source.POINT),
                                src_x=float(src[0]), src_y=float(src[1]), src_z=float(src[2]),
                                tau=4.0, max_steps=max_steps, seed=run_seed))
    return EventBatch(params, counts)


_fn = None


def _bind():
    global _fn
    if _fn is None:
        lib = backend.lib()
        f = lib.optphot_run_gpu_batch
        f.restype = ctypes.c_int
        f.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64,
                      ctypes.c_uint32, ctypes.c_float, ctypes.c_void_p, ctypes.c_int,
                      ctypes.POINTER(backend.WavefrontTuning),
                      ctypes.POINTER(backend.StepCounters), ctypes.POINTER(backend.RunTiming)]
        _fn = f
    return _fn


def run_gpu_batch(batch: EventBatch, time_bins: int = 128, time_max: float = 50.0,
                  first_id: int = 0, method: int = WAVEFRONT, tuning=None,
                  counters: bool = False):
    """Runs every event of the batch; returns (counts [E, 6 + time_bins] u32,
    RunTiming, StepCounters or None). Row layout: fates 0..4, then the time
    histogram with its overflow bin."""
    fn = _bind()
    arr = (SimParams * batch.n_events)(*batch.params)
    photons = np.ascontiguousarray(batch.photons, dtype=np.uint64)
    out = np.zeros((batch.n_events, NUM_FATES + time_bins + 1), np.uint32)
    tm = backend.RunTiming()
    cnt = backend.StepCounters()
    if isinstance(tuning, dict):
        tuning = backend.WavefrontTuning(**tuning)
    rc = fn(ctypes.addressof(arr), photons.ctypes.data, batch.n_events, first_id, time_bins,
            time_max, out.ctypes.data, method,
            ctypes.byref(tuning) if tuning is not None else None,
            ctypes.byref(cnt) if counters else None, ctypes.byref(tm))
    backend._check(rc)
    return out, tm, (cnt if counters else None)


__all__ = ["EventBatch", "make_events", "run_gpu_batch", "WAVEFRONT", "PLAIN", "Surface"]
