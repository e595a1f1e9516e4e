"""Run description types, mirroring the reference's POD structs.

SimParams follows include/optphot/params.hpp:40-63 field for field (72 bytes,
no padding, same defaults) so that a pointer to it is exactly what the C-ABI
(include/optphot_hip.h) takes. TallyConfig follows
include/optphot/transport.hpp:164-168.
"""
from __future__ import annotations

import ctypes
import enum
import math

INF = math.inf


class Surface(enum.IntEnum):
    """params.hpp:15-21 (surface model of the five non-readout faces)."""
    POLISHED = 0
    BLACK = 1
    SPECULAR = 2
    LAMBERTIAN = 3


class Source(enum.IntEnum):
    """params.hpp:23-26."""
    POINT = 0
    VOLUME = 1


class Fate(enum.IntEnum):
    """params.hpp:29-35."""
    DETECTED = 0
    ABSORBED_BULK = 1
    ABSORBED_SURFACE = 2
    ESCAPED = 3
    MAX_STEPS = 4


NUM_FATES = 5
READOUT_FACE = 5  # +z (params.hpp:38)
SPEED_OF_LIGHT = 299.792458  # mm / ns (hd.hpp:37)


class SimParams(ctypes.Structure):
    """One simulation setup (params.hpp:40-63). Lengths in mm, times in ns."""
    _fields_ = [
        ("hx", ctypes.c_float), ("hy", ctypes.c_float), ("hz", ctypes.c_float),
        ("n_scint", ctypes.c_float), ("n_det", ctypes.c_float), ("n_out", ctypes.c_float),
        ("abs_length", ctypes.c_float), ("scat_length", ctypes.c_float),
        ("surface", ctypes.c_uint32), ("reflectivity", ctypes.c_float),
        ("source", ctypes.c_uint32),
        ("src_x", ctypes.c_float), ("src_y", ctypes.c_float), ("src_z", ctypes.c_float),
        ("tau", ctypes.c_float),
        ("max_steps", ctypes.c_uint32),
        ("seed", ctypes.c_uint64),
    ]
    _defaults = dict(hx=25.0, hy=25.0, hz=25.0, n_scint=1.63, n_det=1.63, n_out=1.0,
                     abs_length=1000.0, scat_length=INF, surface=int(Surface.POLISHED),
                     reflectivity=0.95, source=int(Source.POINT), src_x=0.0, src_y=0.0,
                     src_z=0.0, tau=4.0, max_steps=10000, seed=12345)

    def __init__(self, **kw):
        super().__init__()
        vals = dict(self._defaults)
        for k, v in kw.items():
            if k not in vals:
                raise TypeError(f"SimParams has no field {k!r}")
            vals[k] = v
        for k, v in vals.items():
            setattr(self, k, int(v) if k in ("surface", "source", "max_steps", "seed") else v)

    def copy(self, **kw):
        d = self.as_dict()
        d.update(kw)
        return SimParams(**d)

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}

    def __repr__(self):
        return "SimParams(" + ", ".join(f"{k}={v!r}" for k, v in self.as_dict().items()) + ")"


assert ctypes.sizeof(SimParams) == 72


class TallyConfig(ctypes.Structure):
    """transport.hpp:164-168: arrival-time histogram on [0, time_max) with an
    overflow bin, and an xy_bins x xy_bins hit map on the readout face."""
    _fields_ = [("time_bins", ctypes.c_uint32), ("time_max", ctypes.c_float),
                ("xy_bins", ctypes.c_uint32)]

    def __init__(self, time_bins=100, time_max=50.0, xy_bins=50):
        super().__init__(time_bins, time_max, xy_bins)

    @property
    def words(self):
        """u64 words of the flat tally: fates | time | xy | 2 sums."""
        return NUM_FATES + self.time_bins + 1 + self.xy_bins * self.xy_bins + 2


assert ctypes.sizeof(TallyConfig) == 12
