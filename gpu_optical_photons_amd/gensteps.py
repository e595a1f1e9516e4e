"""Photons generated on the device from gensteps (scintillation and
Cherenkov step summaries): optphot_run_gpu_gensteps.

A genstep emits `n_photons` photons uniformly along the segment
(x0, y0, z0) at t0 -> (x1, y1, z1) at t1. With cone_cos = 0 (scintillation)
they go isotropically, each delayed by the run's Exp(tau) emission time; with
cone_cos in (0, 1] (Cherenkov light of a particle of speed beta c moving along
the segment, cone_cos = 1 / (beta n)) they go on the cone of that cosine
about the segment direction, uniform in azimuth, without delay. Photon k of
genstep g has global index
n_0 + ... + n_{g-1} + k and Philox id first_id + that index, so any subset of
gensteps can be re-run photon for photon. The reference only lists
generation from gensteps as future work (README.md:734-735); a degenerate
genstep with t0 = 0 reproduces its point source exactly.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import backend
from .params import SimParams, TallyConfig

GENSTEP_DTYPE = np.dtype([("x0", "<f4"), ("y0", "<f4"), ("z0", "<f4"), ("t0", "<f4"),
                          ("x1", "<f4"), ("y1", "<f4"), ("z1", "<f4"), ("t1", "<f4"),
                          ("n_photons", "<u4"), ("cone_cos", "<f4")])
assert GENSTEP_DTYPE.itemsize == 40


def make_gensteps(n: int, p: SimParams, photons_per_mm: float = 400.0, seed: int = 7,
                  max_len: float = 10.0, time_per_mm: float = 0.0033,
                  cherenkov_fraction: float = 0.0) -> np.ndarray:
    """n synthetic track segments inside the box of `p`: start uniform in the
    box, direction isotropic, length ~U[0, max_len] mm clipped to the box,
    time t0 ~U[0, 2] ns and t1 = t0 + length x time_per_mm (a relativistic
    particle), photon count Poisson(photons_per_mm x length). A fraction
    `cherenkov_fraction` of the segments longer than 1 um are Cherenkov
    gensteps of a particle with beta ~U[1 / n_scint, 1], cone_cos =
    1 / (beta n_scint)."""
    rng = np.random.default_rng(seed)
    h = np.array([p.hx, p.hy, p.hz], np.float64)
    g = np.zeros(n, GENSTEP_DTYPE)
    a = rng.uniform(-h, h, (n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    length = rng.uniform(0.0, max_len, n)
    b = np.clip(a + d * length[:, None], -h, h)
    seg = np.linalg.norm(b - a, axis=1)
    t0 = rng.uniform(0.0, 2.0, n)
    for k, name in enumerate(("x", "y", "z")):
        g[name + "0"] = a[:, k]
        g[name + "1"] = b[:, k]
    g["t0"] = t0
    g["t1"] = t0 + seg * time_per_mm
    g["n_photons"] = rng.poisson(photons_per_mm * seg)
    if cherenkov_fraction > 0.0:
        n_scint = float(p.n_scint)
        beta = rng.uniform(1.0 / n_scint, 1.0, n)
        cher = (rng.random(n) < cherenkov_fraction) & (seg > 1e-3)
        g["cone_cos"] = np.where(cher, np.minimum(1.0, 1.0 / (beta * n_scint)), 0.0)
    return g


_fn = None


def _bind():
    global _fn
    if _fn is None:
        f = backend.lib().optphot_run_gpu_gensteps
        f.restype = ctypes.c_int
        f.argtypes = [ctypes.POINTER(SimParams), ctypes.c_void_p, ctypes.c_uint32,
                      ctypes.c_uint64, ctypes.POINTER(TallyConfig), ctypes.c_void_p,
                      ctypes.c_void_p, ctypes.POINTER(backend.WavefrontTuning),
                      ctypes.POINTER(backend.StepCounters), ctypes.POINTER(backend.RunTiming)]
        _fn = f
    return _fn


def run_gpu_gensteps(p: SimParams, c: TallyConfig, gensteps, first_id: int = 0,
                     records: bool = False, tuning=None, counters: bool = False):
    """Returns (Tally, PhotonRecords or None, RunTiming, StepCounters or None)."""
    g = np.ascontiguousarray(gensteps, dtype=GENSTEP_DTYPE)
    total = int(g["n_photons"].sum())
    flat = np.zeros(c.words, np.uint64)
    rec = backend.PhotonRecords.empty(total) if records else None
    recs_c = rec._c() if records else None
    tm, cnt = backend.RunTiming(), backend.StepCounters()
    if isinstance(tuning, dict):
        tuning = backend.WavefrontTuning(**tuning)
    rc = _bind()(ctypes.byref(p), g.ctypes.data, len(g), first_id, ctypes.byref(c),
                 flat.ctypes.data, ctypes.byref(recs_c) if records else None,
                 ctypes.byref(tuning) if tuning is not None else None,
                 ctypes.byref(cnt) if counters else None, ctypes.byref(tm))
    backend._check(rc)
    return backend.Tally.from_flat(flat, c), rec, tm, (cnt if counters else None)
