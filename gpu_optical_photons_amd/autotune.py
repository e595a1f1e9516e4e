"""Autotuner for the launch shape of the persistent wavefront kernels and of
records runs.

Tunes block size, the refill policy -- steps between refills ("wavefront
length"), or the successor kernel's emission threshold (lanes lacking a
successor before a wave emits) --, the queue chunk a wave claims per atomic,
and persistent blocks per CU, by timing the real kernel on a sample of the
workload; for batches of events (one persistent launch over all events),
block size, steps between refills, queue chunk and the streams' place
(registers or LDS rings), by timing the batch itself; for records, the kernel
(history or wavefront) and its shape. Results never depend on the choice
(integer tallies, per-photon streams; checked by the sweep itself and by the
GPU tests), only the speed does. Choices persist in tuning.json next to this
file, keyed by a workload signature (kernel, surface model, source, which bulk
processes are on, tally size class).

    python -m gpu_optical_photons_amd.autotune   # sweep the standard workloads
"""
from __future__ import annotations

import itertools
import json
import math
import os
import time

HERE = os.path.dirname(os.path.abspath(__file__))
CACHE = os.path.join(HERE, "tuning.json")

DEFAULT = {"block_size": 256, "steps_per_refill": 4, "queue_chunk": 256, "blocks_per_cu": 0,
           "successor_threshold": 0, "stream_in_lds": 0, "sorted_phases": 0}

BLOCKS = (64, 128, 256, 512, 1024)
STEPS = (1, 2, 4, 8, 16)
THRESHOLDS = (16, 32, 48, 64)
RING_BLOCKS = (128, 256, 512)
RING_STEPS = (1, 2, 3, 4)
PHASE_BLOCKS = (256, 512, 1024)
PHASE_STEPS = (1, 2)
CHUNKS = (128, 512, 2048)


def signature(kernel: str, p, c=None) -> str:
    words = 0 if c is None else c.words
    size = "small" if words < 1024 else ("mid" if words < 8192 else "large")
    return (f"{kernel}|surface={int(p.surface)}|source={int(p.source)}"
            f"|abs={'on' if not math.isinf(p.abs_length) else 'off'}"
            f"|scat={'on' if not math.isinf(p.scat_length) else 'off'}|tally={size}")


def _load():
    if os.path.exists(CACHE):
        try:
            return json.load(open(CACHE))
        except Exception:
            return {}
    return {}


def choice_for(kernel: str, p, c=None) -> dict:
    entry = _load().get(signature(kernel, p, c))
    if entry:
        return dict(DEFAULT) | dict(entry["tuning"])
    return dict(DEFAULT)


def sweep(kernel: str, p, c, n: int, blocks=BLOCKS, steps=STEPS, chunks=CHUNKS,
          thresholds=THRESHOLDS, ring_blocks=RING_BLOCKS, ring_steps=RING_STEPS,
          phase_blocks=PHASE_BLOCKS, phase_steps=PHASE_STEPS, reps: int = 2,
          save: bool = True, log=print) -> dict:
    """Times every candidate on n photons; keeps the fastest. Every candidate
    must give the same tally word for word."""
    import numpy as np

    from . import backend

    if kernel != "wavefront_tally":
        raise ValueError(f"unknown kernel {kernel!r}")
    ref = None
    results = []
    candidates = [DEFAULT | {"block_size": b, "steps_per_refill": k, "queue_chunk": ch}
                  for b, k, ch in itertools.product(blocks, steps, chunks)]
    candidates += [DEFAULT | {"block_size": b, "steps_per_refill": 1, "queue_chunk": ch,
                              "successor_threshold": th}
                   for b, th, ch in itertools.product(blocks, thresholds, chunks)]
    candidates += [DEFAULT | {"block_size": b, "steps_per_refill": k, "queue_chunk": ch,
                              "stream_in_lds": 1}
                   for b, k, ch in itertools.product(ring_blocks, ring_steps, chunks)]
    candidates += [DEFAULT | {"block_size": b, "steps_per_refill": k, "queue_chunk": ch,
                              "sorted_phases": 1}
                   for b, k, ch in itertools.product(phase_blocks, phase_steps, chunks)]
    for tu in candidates:
        best = math.inf
        try:
            for _ in range(reps):
                t, tm, st, _ = backend.run_gpu_tally_wavefront(p, c, n, tuning=tu, counters=False)
                best = min(best, tm.kernel_ms)
        except ValueError as e:  # a shape this tally does not fit (LDS)
            log(f"  skipped {tu}: {e}")
            continue
        flat = t.flat()
        if ref is None:
            ref = flat
        elif not np.array_equal(ref, flat):
            raise AssertionError(f"tally differs for tuning {tu}: the kernel is not "
                                 "schedule-independent")
        results.append((best, tu))
        policy = (f"phase steps={tu['steps_per_refill']:2d}" if tu["sorted_phases"]
                  else f"ring steps={tu['steps_per_refill']:2d}" if tu["stream_in_lds"]
                  else f"successor>={tu['successor_threshold']:2d}" if tu["successor_threshold"]
                  else f"steps={tu['steps_per_refill']:2d}")
        log(f"  {kernel} block={tu['block_size']:4d} {policy:14s} chunk={tu['queue_chunk']:5d}: "
            f"{best:8.3f} ms")
    results.sort(key=lambda r: r[0])
    best_ms, best_tu = results[0]
    entry = {"tuning": best_tu, "kernel_ms": best_ms, "photons": n,
             "photons_per_s": n / (best_ms * 1e-3),
             "default_ms": next((ms for ms, tu in results if tu == DEFAULT), None),
             "swept": len(results), "when": time.strftime("%Y-%m-%d %H:%M:%S")}
    if save:
        cache = _load()
        cache[signature(kernel, p, c)] = entry
        with open(CACHE, "w") as f:
            json.dump(cache, f, indent=1, sort_keys=True)
    return entry


BATCH_BLOCKS = (256, 512, 1024)
BATCH_STEPS = (1, 2, 3)
BATCH_CHUNKS = (512, 1024, 2048)


def batch_signature(batch, time_bins: int) -> str:
    """Workload class of a batch: how many events, how many photons per event,
    which surface models occur, whether scattering is on, and whether the
    per-event rows fit the LDS of a block."""
    E = batch.n_events
    mean = float(batch.total) / max(E, 1)
    surfaces = "".join(sorted({str(int(p.surface)) for p in batch.params}))
    scat = {math.isinf(p.scat_length) for p in batch.params}
    rows_lds = E * (6 + time_bins) * 4 <= 32 * 1024
    return (f"batch|events={'few' if E < 64 else 'mid' if E < 1024 else 'many'}"
            f"|photons={'<1e4' if mean < 1e4 else '<1e5' if mean < 1e5 else '>=1e5'}"
            f"|surfaces={surfaces}"
            f"|scat={'off' if scat == {True} else 'on' if scat == {False} else 'mixed'}"
            f"|rows={'lds' if rows_lds else 'global'}")


def batch_choice_for(batch, time_bins: int):
    """The stored tuning for a batch of this class, or None (library defaults)."""
    entry = _load().get(batch_signature(batch, time_bins))
    return dict(entry["tuning"]) if entry else None


def sweep_batch(batch, time_bins: int, time_max: float, blocks=BATCH_BLOCKS,
                steps=BATCH_STEPS, chunks=BATCH_CHUNKS, rings=(0, 1), reps: int = 2,
                save: bool = True, log=print) -> dict:
    """Times every batch launch shape on the batch itself (one persistent
    launch per call); keeps the fastest. Every candidate must give the same
    per-event tallies word for word."""
    import numpy as np

    from .batch import run_gpu_batch

    ref = None
    results = []
    for b, k, ch, ring in itertools.product(blocks, steps, chunks, rings):
        tu = {"block_size": b, "steps_per_refill": k, "queue_chunk": ch, "stream_in_lds": ring}
        try:
            run_gpu_batch(batch, time_bins, time_max, tuning=tu)  # captures the graph
            best = math.inf
            for _ in range(reps):
                w, tm, _ = run_gpu_batch(batch, time_bins, time_max, tuning=tu)
                best = min(best, tm.kernel_ms)
        except ValueError as e:  # a shape this batch does not fit (LDS)
            log(f"  skipped {tu}: {e}")
            continue
        if ref is None:
            ref = w
        elif not np.array_equal(ref, w):
            raise AssertionError(f"batch tallies differ for tuning {tu}: the kernel is not "
                                 "schedule-independent")
        results.append((best, tu))
        log(f"  batch block={b:4d} steps={k} chunk={ch:5d} ring={ring}: {best:8.3f} ms")
    results.sort(key=lambda r: r[0])
    best_ms, best_tu = results[0]
    default = {"block_size": 256, "steps_per_refill": 2, "queue_chunk": 1024, "stream_in_lds": 0}
    entry = {"tuning": best_tu, "kernel_ms": best_ms, "photons": int(batch.total),
             "events": int(batch.n_events), "photons_per_s": batch.total / (best_ms * 1e-3),
             "default_ms": next((ms for ms, tu in results if tu == default), None),
             "swept": len(results), "when": time.strftime("%Y-%m-%d %H:%M:%S")}
    if save:
        cache = _load()
        cache[batch_signature(batch, time_bins)] = entry
        with open(CACHE, "w") as f:
            json.dump(cache, f, indent=1, sort_keys=True)
    return entry


RECORDS_HISTORY_BLOCKS = (64, 128, 256, 512, 1024)
RECORDS_BLOCKS = (128, 256, 512)
RECORDS_STEPS = (1, 2, 4, 8)
RECORDS_CHUNKS = (64, 256, 1024)
RECORDS_BLOCKS_PER_CU = (0, 2)


def records_signature(p, n: int) -> str:
    """Workload class of a records run: the tally signature's physics classes
    and the decade of n (how many photons each lane of a persistent grid
    gets decides between the kernels)."""
    return (f"records|surface={int(p.surface)}|source={int(p.source)}"
            f"|abs={'on' if not math.isinf(p.abs_length) else 'off'}"
            f"|scat={'on' if not math.isinf(p.scat_length) else 'off'}"
            f"|n=1e{int(math.floor(math.log10(max(n, 1))))}")


def records_choice_for(p, n: int):
    """The stored records shape (an optphot_records_tuning dict) for this
    class, or None (the library's automatic choice)."""
    entry = _load().get(records_signature(p, n))
    return dict(entry["tuning"]) if entry else None


def sweep_records(p, n: int, reps: int = 3, save: bool = True, log=print) -> dict:
    """Times the history kernel at every block size and the wavefront kernel
    over block size, steps between refills, queue chunk and blocks per CU,
    on n photons written to device memory; keeps the fastest. Every
    candidate must write the same records bit for bit."""
    import numpy as np

    from . import backend

    candidates = [{"kernel": backend.RECORDS_HISTORY, "block_size": b, "steps_per_refill": 0,
                   "queue_chunk": 0, "blocks_per_cu": 0} for b in RECORDS_HISTORY_BLOCKS]
    candidates += [{"kernel": backend.RECORDS_WAVEFRONT, "block_size": b, "steps_per_refill": k,
                    "queue_chunk": ch, "blocks_per_cu": bpc}
                   for b, k, ch, bpc in itertools.product(RECORDS_BLOCKS, RECORDS_STEPS,
                                                          RECORDS_CHUNKS, RECORDS_BLOCKS_PER_CU)]
    ref = None
    results = []
    for tu in candidates:
        ms = []
        for _ in range(reps + 1):
            rec, tm, _, _ = backend.run_gpu_records_tuned(p, n, tu, device="cuda:0")
            ms.append(tm.kernel_ms)
        best = float(np.median(ms[1:]))
        raw = np.concatenate([rec[f].cpu().numpy().view(np.uint8).ravel()
                              for f in backend.RECORD_FIELDS])
        if ref is None:
            ref = raw
        elif not np.array_equal(ref, raw):
            raise AssertionError(f"records differ for {tu}: the kernels are not "
                                 "schedule-independent")
        results.append((best, tu))
        log(f"  records {tu}: {best:8.3f} ms")
    results.sort(key=lambda r: r[0])
    best_ms, best_tu = results[0]
    hist = min(ms for ms, tu in results if tu["kernel"] == backend.RECORDS_HISTORY)
    entry = {"tuning": best_tu, "kernel_ms": best_ms, "photons": n,
             "photons_per_s": n / (best_ms * 1e-3), "best_history_ms": hist,
             "swept": len(results), "when": time.strftime("%Y-%m-%d %H:%M:%S")}
    if save:
        cache = _load()
        cache[records_signature(p, n)] = entry
        with open(CACHE, "w") as f:
            json.dump(cache, f, indent=1, sort_keys=True)
    return entry


def main():
    import sys
    sys.path.insert(0, os.path.dirname(HERE))
    from .params import SimParams, Source, Surface, TallyConfig

    c2 = SimParams(hx=50.0, hy=50.0, hz=25.0, n_det=1.5, abs_length=2000.0, scat_length=300.0,
                   surface=Surface.LAMBERTIAN, reflectivity=0.97, source=Source.VOLUME, seed=1)
    c1 = SimParams(n_det=1.5, abs_length=1000.0, seed=1)
    for name, p, c in (("long histories", c2, TallyConfig(256, 100.0, 64)),
                       ("plain small", c1, TallyConfig())):
        print(f"sweep: {name}", flush=True)
        e = sweep("wavefront_tally", p, c, 20_000_000, log=lambda s: print(s, flush=True))
        print(f"best: {e['tuning']} {e['kernel_ms']:.3f} ms", flush=True)
    for name, p, n in (("records, plain small", c1, 1_000_000),
                       ("records, long histories", c2, 10_000_000)):
        print(f"sweep: {name}", flush=True)
        e = sweep_records(p, n, log=lambda s: print(s, flush=True))
        print(f"best: {e['tuning']} {e['kernel_ms']:.3f} ms", flush=True)
    from . import calibrate
    from .batch import make_events
    for name, batch, tb, tmax in (("batched events", make_events(4096, seed=2024), 128, 50.0),
                                  ("calibration sets", calibrate.c5_batch(), 128, 100.0)):
        print(f"sweep: {name}", flush=True)
        e = sweep_batch(batch, tb, tmax, log=lambda s: print(s, flush=True))
        print(f"best: {e['tuning']} {e['kernel_ms']:.3f} ms", flush=True)


if __name__ == "__main__":
    main()
