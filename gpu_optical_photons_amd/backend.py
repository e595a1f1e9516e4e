"""Python mirror of the reference's GPU backend API over the C-ABI.

Same names, argument meaning and error behaviour as the reference's
include/optphot/backend.hpp:75-115 (run_gpu_records, run_gpu_tally,
run_gpu_records_event, gpu_available, gpu_device_name):
  * std::invalid_argument  -> ValueError
  * std::runtime_error     -> RuntimeError
The work runs in liboptphot_hip.so (HIP kernels for gfx950); there is no CPU
fallback: if the library is missing, importing this module raises.
"""
from __future__ import annotations

import ctypes
import dataclasses
import os

import numpy as np

from .params import NUM_FATES, Fate, SimParams, TallyConfig

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboptphot_hip.so")

RECORD_FIELDS = ("fate", "n_boundary", "n_scatter", "t", "x", "y", "z", "path")
RECORD_DTYPES = (np.uint8, np.uint32, np.uint32, np.float32, np.float32, np.float32, np.float32,
                 np.float32)

OK, ERR_INVALID_ARGUMENT, ERR_RUNTIME, ERR_NO_DEVICE = 0, 1, 2, 3

CNT_NAMES = ("useful_lane_steps", "issued_lane_steps", "absorb", "scatter", "readout", "polished",
             "black", "reflector", "lambert_reflect", "emitted", "detected", "abs_on_steps",
             "scat_on_steps")


class RunTiming(ctypes.Structure):
    """backend.hpp:65-71 (milliseconds)."""
    _fields_ = [("total_ms", ctypes.c_double), ("kernel_ms", ctypes.c_double),
                ("h2d_ms", ctypes.c_double), ("d2h_ms", ctypes.c_double),
                ("alloc_ms", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class EventStats(ctypes.Structure):
    """backend.hpp:95-103, wave64: issued lane-steps count 64 lanes per wave-step."""
    _fields_ = [("launches", ctypes.c_uint64), ("useful_lane_steps", ctypes.c_uint64),
                ("issued_lane_steps", ctypes.c_uint64), ("kernels_ms", ctypes.c_double)]

    def simt_efficiency(self):
        return self.useful_lane_steps / self.issued_lane_steps if self.issued_lane_steps else 0.0


class StepCounters(ctypes.Structure):
    _fields_ = [("v", ctypes.c_uint64 * 16)]

    def as_dict(self):
        return {name: int(self.v[i]) for i, name in enumerate(CNT_NAMES)}


class WavefrontTuning(ctypes.Structure):
    _fields_ = [("block_size", ctypes.c_int), ("steps_per_refill", ctypes.c_uint32),
                ("queue_chunk", ctypes.c_uint32), ("blocks_per_cu", ctypes.c_int),
                ("successor_threshold", ctypes.c_uint32), ("stream_in_lds", ctypes.c_uint32),
                ("sorted_phases", ctypes.c_uint32)]


RECORDS_AUTO, RECORDS_HISTORY, RECORDS_WAVEFRONT = 0, 1, 2


class RecordsTuning(ctypes.Structure):
    """optphot_records_tuning: the kernel (RECORDS_AUTO / _HISTORY /
    _WAVEFRONT) and its shape; zero fields take the library defaults."""
    _fields_ = [("kernel", ctypes.c_int), ("block_size", ctypes.c_int),
                ("steps_per_refill", ctypes.c_uint32), ("queue_chunk", ctypes.c_uint32),
                ("blocks_per_cu", ctypes.c_int)]

    def as_dict(self):
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


class SortedTuning(ctypes.Structure):
    """optphot_sorted_tuning: zero fields take the library defaults."""
    _fields_ = [("slots_per_block", ctypes.c_uint32), ("blocks_per_cu", ctypes.c_uint32),
                ("iterations_per_poll", ctypes.c_uint32), ("tail_threshold", ctypes.c_uint32)]


class SortedStats(ctypes.Structure):
    _fields_ = [("slots", ctypes.c_uint64), ("iterations", ctypes.c_uint64),
                ("classified", ctypes.c_uint64), ("fresnel", ctypes.c_uint64),
                ("rotated", ctypes.c_uint64), ("emitted", ctypes.c_uint64),
                ("tail_photons", ctypes.c_uint64), ("loop_ms", ctypes.c_double),
                ("tail_ms", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class _RecordsC(ctypes.Structure):
    _fields_ = [(name, ctypes.c_void_p) for name in RECORD_FIELDS]


def _share_torch_runtime():
    """PyTorch-ROCm wheels ship their own HIP and HSA runtimes under the same
    sonames (libamdhip64.so.7, libhsa-runtime64.so.1). Loading torch first
    makes this library bind to those already-loaded runtimes, so the process
    has one HIP runtime and torch tensors and this library's kernels share a
    device and its memory. Loading this library first would pull in the
    system runtime as well, and torch would then see no device. Skipped when
    torch is absent or OPTPHOT_NO_TORCH=1."""
    if os.environ.get("OPTPHOT_NO_TORCH") == "1":
        return
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is not built: the HIP library is required (there is no CPU fallback). "
            "Build it with `python -c 'import __graft_entry__ as g; g.build()'` or "
            "`make -C gpu_optical_photons_amd/csrc`.")
    _share_torch_runtime()
    lib = ctypes.CDLL(LIB_PATH)
    P = ctypes.POINTER
    sig = {
        "optphot_gpu_available": (ctypes.c_int, []),
        "optphot_gpu_device_name": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_size_t]),
        "optphot_device_info": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_size_t]),
        "optphot_last_error": (ctypes.c_char_p, []),
        "optphot_run_gpu_records": (ctypes.c_int, [P(SimParams), ctypes.c_uint64, P(_RecordsC),
                                                   ctypes.c_int, ctypes.c_uint64,
                                                   ctypes.c_uint64, P(RunTiming)]),
        "optphot_run_gpu_tally": (ctypes.c_int, [P(SimParams), P(TallyConfig), ctypes.c_uint64,
                                                 ctypes.c_void_p, ctypes.c_int, ctypes.c_uint64,
                                                 P(RunTiming)]),
        "optphot_run_gpu_tally_counters": (ctypes.c_int, [P(SimParams), P(TallyConfig),
                                                          ctypes.c_uint64, ctypes.c_void_p,
                                                          ctypes.c_int, ctypes.c_uint64,
                                                          P(StepCounters), P(RunTiming)]),
        "optphot_run_gpu_tally_wavefront": (ctypes.c_int, [P(SimParams), P(TallyConfig),
                                                           ctypes.c_uint64, ctypes.c_void_p,
                                                           ctypes.c_uint64, P(WavefrontTuning),
                                                           P(EventStats), P(StepCounters),
                                                           P(RunTiming)]),
        "optphot_run_gpu_tally_wavefront_noalloc": (ctypes.c_int, [
            P(SimParams), P(TallyConfig), ctypes.c_uint64, ctypes.c_void_p, ctypes.c_uint64,
            P(WavefrontTuning), P(RunTiming)]),
        "optphot_run_gpu_tally_sorted": (ctypes.c_int, [P(SimParams), P(TallyConfig),
                                                        ctypes.c_uint64, ctypes.c_void_p,
                                                        ctypes.c_uint64, P(SortedTuning),
                                                        P(SortedStats), P(RunTiming)]),
        "optphot_run_gpu_records_event": (ctypes.c_int, [P(SimParams), ctypes.c_uint64,
                                                         P(_RecordsC), ctypes.c_uint32,
                                                         ctypes.c_int, ctypes.c_uint64,
                                                         ctypes.c_uint64, P(EventStats),
                                                         P(RunTiming)]),
        "optphot_run_gpu_records_device": (ctypes.c_int, [P(SimParams), ctypes.c_uint64,
                                                          P(_RecordsC), ctypes.c_uint32,
                                                          ctypes.c_int, ctypes.c_uint64,
                                                          P(EventStats), P(RunTiming)]),
        "optphot_run_gpu_records_tuned": (ctypes.c_int, [P(SimParams), ctypes.c_uint64,
                                                         P(_RecordsC), ctypes.c_int,
                                                         P(RecordsTuning), ctypes.c_uint64,
                                                         P(RecordsTuning), P(EventStats),
                                                         P(RunTiming)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


_lib = _load()


def lib():
    """The loaded liboptphot_hip.so (ctypes.CDLL)."""
    return _lib


def _check(rc):
    if rc == OK:
        return
    msg = _lib.optphot_last_error().decode(errors="replace")
    if rc == ERR_INVALID_ARGUMENT:
        raise ValueError(msg)
    raise RuntimeError(msg)


def gpu_available() -> bool:
    """backend.hpp:84."""
    return bool(_lib.optphot_gpu_available())


def gpu_device_name() -> str:
    """backend.hpp:85 ("none" without a device, as gpu_backend.cu:106-115)."""
    buf = ctypes.create_string_buffer(256)
    if _lib.optphot_gpu_device_name(buf, 256) != OK:
        return "none"
    return buf.value.decode()


def device_info() -> dict:
    buf = ctypes.create_string_buffer(4096)
    _check(_lib.optphot_device_info(buf, 4096))
    out = {}
    for line in buf.value.decode().splitlines():
        if "=" in line:
            k, v = line.split("=", 1)
            out[k] = v
    return out


# ---------------------------------------------------------------------------
@dataclasses.dataclass
class PhotonRecords:
    """backend.hpp:31-50: structure of arrays indexed by photon number."""
    fate: np.ndarray
    n_boundary: np.ndarray
    n_scatter: np.ndarray
    t: np.ndarray
    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    path: np.ndarray

    BYTES_PER_PHOTON = 29

    @classmethod
    def empty(cls, n):
        return cls(*[np.zeros(n, d) for d in RECORD_DTYPES])

    def __len__(self):
        return len(self.fate)

    def as_dict(self):
        return {f: getattr(self, f) for f in RECORD_FIELDS}

    def _c(self):
        return _RecordsC(*[getattr(self, f).ctypes.data for f in RECORD_FIELDS])

    def steps(self):
        """Photon-steps per photon: n_boundary + n_scatter + [AbsorbedBulk]
        (apps/optphot_event_bench.cpp:39-41)."""
        return (self.n_boundary.astype(np.uint64) + self.n_scatter
                + (self.fate == Fate.ABSORBED_BULK))


@dataclasses.dataclass
class Tally:
    """backend.hpp:52-62."""
    fate_counts: np.ndarray
    time_hist: np.ndarray
    xy_hist: np.ndarray  # [iy][ix]
    boundary_sum: int
    scatter_sum: int

    @classmethod
    def from_flat(cls, flat, c: TallyConfig):
        t0 = NUM_FATES
        x0 = t0 + c.time_bins + 1
        s0 = x0 + c.xy_bins * c.xy_bins
        return cls(flat[:NUM_FATES].copy(), flat[t0:x0].copy(),
                   flat[x0:s0].reshape(c.xy_bins, c.xy_bins).copy(), int(flat[s0]),
                   int(flat[s0 + 1]))

    def total(self):
        return int(self.fate_counts.sum())

    def flat(self):
        return np.concatenate([self.fate_counts, self.time_hist, self.xy_hist.ravel(),
                               np.array([self.boundary_sum, self.scatter_sum], np.uint64)])


# ---------------------------------------------------------------------------
def run_gpu_records(p: SimParams, n: int, block_size: int = 256, first_id: int = 0,
                    chunk: int = 1 << 24, out: PhotonRecords | None = None):
    """run_gpu_records (backend.hpp:88-89): history-based kernel, one lane per
    photon. Returns (PhotonRecords, RunTiming)."""
    rec = out if out is not None else PhotonRecords.empty(n)
    tm = RunTiming()
    rc = _lib.optphot_run_gpu_records(ctypes.byref(p), n, ctypes.byref(rec._c()), block_size,
                                      first_id, chunk, ctypes.byref(tm))
    _check(rc)
    return rec, tm


def run_gpu_tally(p: SimParams, c: TallyConfig, n: int, block_size: int = 256,
                  first_id: int = 0, counters: bool = False):
    """run_gpu_tally (backend.hpp:90-91): history-based tally.
    Returns (Tally, RunTiming[, StepCounters])."""
    flat = np.zeros(c.words, np.uint64)
    tm = RunTiming()
    cnt = StepCounters()
    rc = _lib.optphot_run_gpu_tally_counters(ctypes.byref(p), ctypes.byref(c), n,
                                             flat.ctypes.data, block_size, first_id,
                                             ctypes.byref(cnt) if counters else None,
                                             ctypes.byref(tm))
    _check(rc)
    t = Tally.from_flat(flat, c)
    return (t, tm, cnt) if counters else (t, tm)


def run_gpu_records_event(p: SimParams, n: int, steps_per_launch: int, block_size: int = 256,
                          first_id: int = 0, chunk: int = 1 << 24):
    """run_gpu_records_event (backend.hpp:112-115): same records, bit for bit,
    from the persistent wavefront kernel. Returns (PhotonRecords, RunTiming,
    EventStats)."""
    rec = PhotonRecords.empty(n)
    tm = RunTiming()
    st = EventStats()
    rc = _lib.optphot_run_gpu_records_event(ctypes.byref(p), n, ctypes.byref(rec._c()),
                                            steps_per_launch, block_size, first_id, chunk,
                                            ctypes.byref(st), ctypes.byref(tm))
    _check(rc)
    return rec, tm, st


def run_gpu_records_device(p: SimParams, n: int, steps_per_launch: int = 0,
                           block_size: int = 256, first_id: int = 0, device: str = "cuda:0"):
    """Records written straight into device memory (new; no reference
    counterpart): returns ({field: torch tensor on `device`}, RunTiming,
    EventStats). steps_per_launch 0 = history kernel, >= 1 = wavefront kernel.
    torch only allocates the device arrays; the library's kernels fill them."""
    import torch
    dtypes = {np.dtype(np.uint8): torch.uint8, np.dtype(np.uint32): torch.uint32,
              np.dtype(np.float32): torch.float32}
    out = {f: torch.empty(n, dtype=dtypes[np.dtype(d)], device=device)
           for f, d in zip(RECORD_FIELDS, RECORD_DTYPES)}
    torch.cuda.synchronize(device)  # allocations complete before another stream writes
    c = _RecordsC(*[out[f].data_ptr() for f in RECORD_FIELDS])
    tm, st = RunTiming(), EventStats()
    rc = _lib.optphot_run_gpu_records_device(ctypes.byref(p), n, ctypes.byref(c),
                                             steps_per_launch, block_size, first_id,
                                             ctypes.byref(st), ctypes.byref(tm))
    _check(rc)
    return out, tm, st


def run_gpu_records_tuned(p: SimParams, n: int, tuning: RecordsTuning | dict | None = None,
                          first_id: int = 0, device: str | None = None,
                          out: PhotonRecords | None = None):
    """Records with a chosen launch shape (new; no reference counterpart):
    `tuning` None (or kernel RECORDS_AUTO) lets the library choose the kernel
    from n. Records go to host arrays (PhotonRecords), or, with `device` set,
    into torch tensors on that device. Returns (records, RunTiming,
    EventStats, the RecordsTuning that ran)."""
    if isinstance(tuning, dict):
        tuning = RecordsTuning(**tuning)
    if device is not None:
        import torch
        dtypes = {np.dtype(np.uint8): torch.uint8, np.dtype(np.uint32): torch.uint32,
                  np.dtype(np.float32): torch.float32}
        rec = {f: torch.empty(n, dtype=dtypes[np.dtype(d)], device=device)
               for f, d in zip(RECORD_FIELDS, RECORD_DTYPES)}
        torch.cuda.synchronize(device)  # allocations complete before another stream writes
        c = _RecordsC(*[rec[f].data_ptr() for f in RECORD_FIELDS])
    else:
        rec = out if out is not None else PhotonRecords.empty(n)
        c = rec._c()
    tm, st, chosen = RunTiming(), EventStats(), RecordsTuning()
    rc = _lib.optphot_run_gpu_records_tuned(ctypes.byref(p), n, ctypes.byref(c),
                                            1 if device is not None else 0,
                                            ctypes.byref(tuning) if tuning is not None else None,
                                            first_id, ctypes.byref(chosen), ctypes.byref(st),
                                            ctypes.byref(tm))
    _check(rc)
    return rec, tm, st, chosen


def run_gpu_tally_wavefront(p: SimParams, c: TallyConfig, n: int, first_id: int = 0,
                            tuning: WavefrontTuning | dict | None = None, counters: bool = True):
    """Tally by the persistent wavefront kernel (new; no reference counterpart).
    Returns (Tally, RunTiming, EventStats, StepCounters). With counters=False
    the launch uses the kernel variant without branch-event counting (the
    one to time); EventStats then carries only launches and kernels_ms and
    the StepCounters are zero."""
    if isinstance(tuning, dict):
        tuning = WavefrontTuning(**tuning)
    flat = np.zeros(c.words, np.uint64)
    tm, st, cnt = RunTiming(), EventStats(), StepCounters()
    rc = _lib.optphot_run_gpu_tally_wavefront(
        ctypes.byref(p), ctypes.byref(c), n, flat.ctypes.data, first_id,
        ctypes.byref(tuning) if tuning is not None else None,
        ctypes.byref(st) if counters else None, ctypes.byref(cnt) if counters else None,
        ctypes.byref(tm))
    _check(rc)
    if not counters:
        st.launches = 1
        st.kernels_ms = tm.kernel_ms
    return Tally.from_flat(flat, c), tm, st, cnt


def run_gpu_tally_wavefront_noalloc(p: SimParams, c: TallyConfig, n: int, out, first_id: int = 0,
                                    tuning: WavefrontTuning | dict | None = None):
    """The wavefront tally from the entry point that allocates nothing
    (optphot_run_gpu_tally_wavefront_noalloc; no reference counterpart): the
    tally goes into `out`, a caller-owned device tensor of c.words 64-bit
    words (torch.int64 or torch.uint64), and the launch works in the
    library's static device scratch. Returns the RunTiming. Some other call of
    the library must have set up the device first."""
    import torch
    if out.device.type != "cuda" or out.element_size() != 8 or out.numel() != c.words \
            or not out.is_contiguous():
        raise ValueError(f"out must be a contiguous device tensor of {c.words} 64-bit words")
    if isinstance(tuning, dict):
        tuning = WavefrontTuning(**tuning)
    torch.cuda.synchronize(out.device)  # torch's work on `out` is done before the library's
    tm = RunTiming()
    rc = _lib.optphot_run_gpu_tally_wavefront_noalloc(
        ctypes.byref(p), ctypes.byref(c), n, ctypes.c_void_p(out.data_ptr()), first_id,
        ctypes.byref(tuning) if tuning is not None else None, ctypes.byref(tm))
    _check(rc)
    return tm


def run_gpu_tally_sorted(p: SimParams, c: TallyConfig, n: int, first_id: int = 0,
                         tuning: SortedTuning | dict | None = None):
    """Tally by the per-interaction path (photon state in device memory; per
    iteration a classify kernel (emission and the step's common part), then a
    Fresnel and a rotation kernel over per-type queues; the design of the
    reference's README.md:717-720).
    Same tally as run_gpu_tally. Returns (Tally, RunTiming, SortedStats)."""
    if isinstance(tuning, dict):
        tuning = SortedTuning(**tuning)
    flat = np.zeros(c.words, np.uint64)
    tm, st = RunTiming(), SortedStats()
    rc = _lib.optphot_run_gpu_tally_sorted(
        ctypes.byref(p), ctypes.byref(c), n, flat.ctypes.data, first_id,
        ctypes.byref(tuning) if tuning is not None else None, ctypes.byref(st), ctypes.byref(tm))
    _check(rc)
    return Tally.from_flat(flat, c), tm, st
