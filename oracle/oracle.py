"""ctypes access to the oracle libraries (test infrastructure only).

Two libraries are wrapped:
  * liboracle_{portable,libm}.so -- our plain-C restatement (optphot_oracle.c);
  * _ref/libref_{portable,libm}.so -- the reference's own CPU path, compiled
    from its sources by oracle/Makefile (`make ref`), present only where the
    reference checkout was available at build time.
"portable" means the reference's OPTPHOT_PORTABLE_MATH build (correctly
rounded operations only, bit-reproducible), "libm" its default build
(logf/sinf/cosf/cbrtf from the C library).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline may import
this module; it is the checker, never the thing measured or shipped.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")

RECORD_FIELDS = ("fate", "n_boundary", "n_scatter", "t", "x", "y", "z", "path")
RECORD_DTYPES = (np.uint8, np.uint32, np.uint32, np.float32, np.float32, np.float32, np.float32,
                 np.float32)


class TallyConfigC(ctypes.Structure):
    """TallyConfig (include/optphot/transport.hpp:164-168)."""
    _fields_ = [("time_bins", ctypes.c_uint32), ("time_max", ctypes.c_float),
                ("xy_bins", ctypes.c_uint32)]


def _empty_records(n):
    return {f: np.zeros(n, d) for f, d in zip(RECORD_FIELDS, RECORD_DTYPES)}


def tally_words(time_bins, xy_bins):
    return 5 + time_bins + 1 + xy_bins * xy_bins + 2


def split_tally(flat, time_bins, xy_bins):
    """Flat u64 tally -> dict (fate_counts, time_hist, xy_hist, sums)."""
    t0 = 5
    x0 = t0 + time_bins + 1
    s0 = x0 + xy_bins * xy_bins
    return {
        "fate_counts": flat[:5].copy(),
        "time_hist": flat[t0:x0].copy(),
        "xy_hist": flat[x0:s0].reshape(xy_bins, xy_bins).copy(),
        "boundary_sum": int(flat[s0]),
        "scatter_sum": int(flat[s0 + 1]),
    }


class _Lib:
    prefix = ""

    def __init__(self, path):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} is not built (run `make -C oracle`)")
        self.path = path
        self.lib = ctypes.CDLL(path)
        L, p = self.lib, self.prefix
        vp = ctypes.c_void_p
        f = ctypes.c_float
        self._fn("philox4x32_10", None, [_u32p, _u32p, _u32p])
        self._fn("stream_words", None, [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint32,
                                        ctypes.c_uint32, _u32p])
        for name in ("log", "cbrt", "portable_log", "portable_cbrt", "rayleigh_cos"):
            self._fn(name, f, [f])
        for name in ("sincos_turns", "portable_sincos_turns"):
            self._fn(name, None, [f, ctypes.POINTER(f), ctypes.POINTER(f)])
        self._fn("fresnel_R", f, [f, f, f])
        self._fn("box_exit", f, [_f32p, _f32p, f, f, f, ctypes.POINTER(ctypes.c_int)])
        self._fn("portable_math", ctypes.c_int, [])
        self._fn("max_threads", ctypes.c_int, [])
        self._fn("time_bin", ctypes.c_uint32, [f, vp])
        self._fn("xy_bin", ctypes.c_uint32, [f, f, vp, vp])
        del L, p

    def _fn(self, name, restype, argtypes):
        fn = getattr(self.lib, self.prefix + name)
        fn.restype = restype
        fn.argtypes = argtypes
        setattr(self, name, fn)

    # -- convenience wrappers -------------------------------------------------
    def philox(self, ctr, key):
        out = np.zeros(4, np.uint32)
        self.philox4x32_10(np.asarray(ctr, np.uint32), np.asarray(key, np.uint32), out)
        return out

    def words(self, seed, photon_id, n, skip=0):
        out = np.zeros(n, np.uint32)
        self.stream_words(seed, photon_id, skip, n, out)
        return out

    def sincos(self, u, portable=False):
        s, c = ctypes.c_float(), ctypes.c_float()
        (self.portable_sincos_turns if portable else self.sincos_turns)(u, ctypes.byref(s),
                                                                        ctypes.byref(c))
        return s.value, c.value

    def exit(self, p, d, hx, hy, hz):
        face = ctypes.c_int()
        dist = self.box_exit(np.asarray(p, np.float32), np.asarray(d, np.float32), hx, hy, hz,
                             ctypes.byref(face))
        return dist, face.value


class Oracle(_Lib):
    """Our C restatement (optphot_oracle.c)."""
    prefix = "oracle_"

    def __init__(self, math="portable"):
        super().__init__(os.path.join(HERE, f"liboracle_{math}.so"))
        self.math = math
        self._fn("run_records", None, [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint64,
                                       ctypes.c_int, _u8p, _u32p, _u32p, _f32p, _f32p, _f32p,
                                       _f32p, _f32p, ctypes.c_void_p])
        self._fn("run_tally", None, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64,
                                     ctypes.c_uint64, ctypes.c_int, _u64p])
        self._fn("u32_to_uniform", ctypes.c_float, [ctypes.c_uint32])
        self._fn("rayleigh_table", None, [_f32p])
        self._fn("adjoint", None, [ctypes.c_void_p, _u64p, ctypes.c_uint32, ctypes.c_uint64,
                                   ctypes.c_uint32, ctypes.c_float, ctypes.c_void_p,
                                   ctypes.c_void_p, ctypes.c_int])
        self._fn("trace", ctypes.c_uint32, [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint32,
                                            _u8p, _f32p, ctypes.c_void_p])
        self._fn("emitted", None, [ctypes.c_void_p, ctypes.c_uint64, _f32p])
        self._fn("run_gensteps", None, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32,
                                        ctypes.c_uint64, ctypes.c_void_p, ctypes.c_void_p,
                                        ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                        ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                        ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int])

    def gensteps(self, params, gensteps, first_id=0, time_bins=100, time_max=50.0, xy_bins=50,
                 records=False, threads=0):
        """Tally (dict) and, with records=True, records of photons from gensteps
        (structured array with fields x0 y0 z0 t0 x1 y1 z1 t1 n_photons cone_cos)."""
        g = np.ascontiguousarray(gensteps)
        assert g.dtype.itemsize == 40
        total = int(g["n_photons"].sum())
        cfg = TallyConfigC(time_bins, time_max, xy_bins)
        flat = np.zeros(tally_words(time_bins, xy_bins), np.uint64)
        r = _empty_records(total) if records else None
        ptrs = [r[k].ctypes.data if records else None for k in RECORD_FIELDS]
        self.run_gensteps(ctypes.byref(params), g.ctypes.data, len(g), first_id,
                          ctypes.byref(cfg), flat.ctypes.data, *ptrs, threads)
        return split_tally(flat, time_bins, xy_bins), r

    def emitted_state(self, params, photon_id):
        """Position, direction and time (7 floats) of photon `photon_id` as
        emitted: the state before the first row of step_trace."""
        out = np.zeros(7, np.float32)
        self.emitted(ctypes.byref(params), photon_id, out)
        return out

    def step_trace(self, params, photon_id, cap=1 << 14, with_state=False):
        """(codes u8 [steps], margins f32 [steps]) of one photon: per step the
        outcome (0 absorb, 1 scatter, 2 + face boundary, | 0x80 if the face's
        Bernoulli decision came out true) and the smallest relative margin of
        the comparisons that decided it."""
        codes = np.zeros(cap, np.uint8)
        margins = np.zeros(cap, np.float32)
        state = np.zeros((cap, 7), np.float32) if with_state else None
        n = int(self.trace(ctypes.byref(params), photon_id, cap, codes, margins,
                           state.ctypes.data_as(ctypes.c_void_p) if with_state else None))
        n = min(n, cap)
        if with_state:
            return codes[:n], margins[:n], state[:n]
        return codes[:n], margins[:n]

    def gradient(self, params_list, photons, time_bins, time_max, weights, first_id=0,
                 threads=0):
        """FP64 reference of the adjoint: grad [E, 5] = d/dtheta sum_k w[e,k] h_e[k]."""
        E = len(params_list)
        arr = (type(params_list[0]) * E)(*params_list)
        ph = np.ascontiguousarray(photons, np.uint64)
        w = np.ascontiguousarray(weights, np.float64).reshape(E, time_bins + 1)
        g = np.zeros((E, 5), np.float64)
        self.adjoint(ctypes.addressof(arr), ph, E, first_id, time_bins, time_max,
                     w.ctypes.data, g.ctypes.data, threads)
        return g

    def records(self, params, n, first_id=0, threads=0, with_draws=False):
        r = _empty_records(n)
        draws = np.zeros(n, np.uint32) if with_draws else None
        self.run_records(ctypes.byref(params), n, first_id, threads,
                         *[r[k] for k in RECORD_FIELDS],
                         draws.ctypes.data_as(ctypes.c_void_p) if with_draws else None)
        if with_draws:
            r["draws"] = draws
        return r

    def rayleigh_cos_table(self):
        """sample_rayleigh_cos of every 24-bit uniform (2^24 floats): entry k
        for u = (k + 1) 2^-24."""
        out = np.empty(1 << 24, np.float32)
        self.rayleigh_table(out)
        return out

    def tally(self, params, n, time_bins=100, time_max=50.0, xy_bins=50, first_id=0, threads=0):
        cfg = TallyConfigC(time_bins, time_max, xy_bins)
        flat = np.zeros(tally_words(time_bins, xy_bins), np.uint64)
        self.run_tally(ctypes.byref(params), ctypes.byref(cfg), n, first_id, threads, flat)
        return split_tally(flat, time_bins, xy_bins)


class Reference(_Lib):
    """The reference's own CPU path (oracle/_ref, built from its sources)."""
    prefix = "ref_"

    def __init__(self, math="portable"):
        super().__init__(os.path.join(HERE, "_ref", f"libref_{math}.so"))
        self.math = math
        self._fn("run_records", None, [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint64,
                                       ctypes.c_int, _u8p, _u32p, _u32p, _f32p, _f32p, _f32p,
                                       _f32p, _f32p])
        self._fn("run_tally", ctypes.c_double, [ctypes.c_void_p, ctypes.c_void_p,
                                                ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int,
                                                _u64p])
        self._fn("transport_one", None, [ctypes.c_void_p, ctypes.c_uint64, _f32p, _u32p])
        self._fn("face_solid_angle_fraction", ctypes.c_double, [ctypes.c_void_p, ctypes.c_int])
        self._fn("face_survival_fraction", ctypes.c_double, [ctypes.c_void_p, ctypes.c_int,
                                                             ctypes.c_double])
        self._fn("mirror_box_escape_fraction", ctypes.c_double, [ctypes.c_double,
                                                                 ctypes.c_double])

    def records(self, params, n, first_id=0, threads=0):
        r = _empty_records(n)
        self.run_records(ctypes.byref(params), n, first_id, threads,
                         *[r[k] for k in RECORD_FIELDS])
        return r

    def tally(self, params, n, time_bins=100, time_max=50.0, xy_bins=50, first_id=0, threads=0):
        cfg = TallyConfigC(time_bins, time_max, xy_bins)
        flat = np.zeros(tally_words(time_bins, xy_bins), np.uint64)
        ms = self.run_tally(ctypes.byref(params), ctypes.byref(cfg), n, first_id, threads, flat)
        out = split_tally(flat, time_bins, xy_bins)
        out["ms"] = ms
        return out

    def one(self, params, photon_id):
        f = np.zeros(5, np.float32)
        u = np.zeros(4, np.uint32)
        self.transport_one(ctypes.byref(params), photon_id, f, u)
        return {"fate": int(u[0]), "n_boundary": int(u[1]), "n_scatter": int(u[2]),
                "draws": int(u[3]), "t": f[0], "x": f[1], "y": f[2], "z": f[3], "path": f[4]}


def reference_available(math="portable"):
    return os.path.exists(os.path.join(HERE, "_ref", f"libref_{math}.so"))
