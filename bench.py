#!/usr/bin/env python3
"""Throughput of the optical-photon transport hot path on one MI355X.

Workload (one "step" = one pass of the hot path over one batch): long
histories, event-based -- a 100 x 100 x 50 mm box, Rayleigh scattering with
scat_length 300 mm, five Lambertian faces with R = 0.97, +z readout to
n_det = 1.5, photons uniform in the volume, abs_length 2000 mm, n_scint 1.63,
tau 4 ns, seed 1; 1e8 photons per step through the persistent wavefront
kernel, tallied into a 256-bin arrival-time histogram over 0-100 ns and a
64 x 64 hit map. Parameters travel as kernel arguments and photons are
generated on the device, so nothing is read from the host in the timed
region.

Before its W warmup steps the GPU runs 0.5 s of untimed launches of the same
workload, so that the timed steps see its steady clock.

Prints one JSON line: photons/s (value), photon-steps/s, the plain
history-based kernel on the same workload, the roofline of the dominant
kernel (FP32 vector / VALU: the step is scalar f32 and i32 arithmetic with
branches, no dense contraction) and the CPU baseline (the reference's own CPU
path, OpenMP, timed on this host's cores on a bounded sample).

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

N_PHOTONS = 100_000_000
WORKLOAD = "long_histories_wavefront_c2"


def settle_clocks(run, seconds=0.5):
    """Runs `run` (one untimed launch of the workload) until `seconds` of wall
    time have passed. An idle MI355X raises its clocks over the first tens of
    milliseconds of work (a launch takes ~15% longer at first), so the warmup
    and timed steps that follow run at the steady clock."""
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        run()


def c2_params():
    from gpu_optical_photons_amd.params import SimParams, Source, Surface
    return SimParams(hx=50.0, hy=50.0, hz=25.0, n_scint=1.63, n_det=1.5, n_out=1.0,
                     abs_length=2000.0, scat_length=300.0, surface=Surface.LAMBERTIAN,
                     reflectivity=0.97, source=Source.VOLUME, tau=4.0, max_steps=10000, seed=1)


def cpu_baseline(p, c):
    """The reference's CPU path (oracle/_ref: its src/cpu/cpu_backend.cpp, default
    libm build, OpenMP over all cores) or, where that is not built, the oracle
    port, on a bounded sample of the same workload."""
    try:
        from oracle.oracle import Oracle, Reference, reference_available
    except Exception as e:  # pragma: no cover - oracle not built
        return {"error": f"oracle unavailable: {e}"}
    kind = "reference" if reference_available("libm") else "port"
    lib = Reference("libm") if kind == "reference" else Oracle("libm")
    cores = int(lib.max_threads())
    probe = 1_000_000
    t0 = time.perf_counter()
    lib.tally(p, probe, c.time_bins, c.time_max, c.xy_bins, threads=0)
    dt = time.perf_counter() - t0
    n = int(min(2_000_000_000, max(probe, 10.0 / max(dt, 1e-6) * probe)))
    t0 = time.perf_counter()
    lib.tally(p, n, c.time_bins, c.time_max, c.xy_bins, threads=0)
    dt = time.perf_counter() - t0
    # one thread, about 3 s of work
    n1 = int(max(10_000, min(n, 3.0 * n / dt / max(cores, 1))))
    t1 = time.perf_counter()
    lib.tally(p, n1, c.time_bins, c.time_max, c.xy_bins, threads=1)
    d1 = time.perf_counter() - t1
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            model = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), "")
    except OSError:
        pass
    return {"value": n / dt, "unit": "photons/s", "cores": cores, "kind": kind,
            "sample": f"photon ids 0..{n - 1} of the same workload (tally mode), "
                      f"{dt:.1f} s wall on {cores} OpenMP threads",
            "cpu_model": model, "nproc": os.cpu_count(),
            "one_thread": {"value": n1 / d1, "unit": "photons/s", "cores": 1,
                           "sample": f"photon ids 0..{n1 - 1}, {d1:.1f} s wall"}}


def _cpu_lib():
    """The reference's CPU path (oracle/_ref, its default libm build), or the
    oracle's port where that is not built. Returns (library, kind)."""
    from oracle.oracle import Oracle, Reference, reference_available
    if reference_available("libm"):
        return Reference("libm"), "reference"
    return Oracle("libm"), "port"


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            return next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), "")
    except OSError:
        return ""


def bench_plain_small():
    """Plain method, small: 50 mm cube, polished faces to n_out 1.0, readout
    n_det 1.5, abs 1000 mm, no Rayleigh, point source at the centre, tau 4 ns,
    1e6 photons, seed 1; records. Kernel rates from one launch writing device
    records (history kernel, one lane per photon; wavefront kernel); the
    end-to-end rate of records delivered into host memory (kernel pieces,
    pinned staging, D2H and host copies overlapped); all records identical
    bit for bit. Roofline from the branch counters of the same photons; the
    reference's CPU path on the same photons (records and tally, all cores
    and one thread)."""
    import numpy as np

    from gpu_optical_photons_amd import backend, roofline
    from gpu_optical_photons_amd.params import SimParams, TallyConfig
    p = SimParams(n_det=1.5, abs_length=1000.0, seed=1)
    n = 1_000_000
    backend.run_gpu_records_device(p, n, 0)
    dev, tm, _ = backend.run_gpu_records_device(p, n, 0)
    dev_w, tme, st = backend.run_gpu_records_device(p, n, 2)
    backend.run_gpu_records(p, n)
    host_ms = []
    for _ in range(3):
        rec, th = backend.run_gpu_records(p, n)
        host_ms.append((th.total_ms, th.d2h_ms))
    total_ms, d2h_ms = min(host_ms)
    same = all(np.array_equal(getattr(rec, f).view(np.uint8), dev[f].cpu().numpy().view(np.uint8))
               and np.array_equal(getattr(rec, f).view(np.uint8),
                                  dev_w[f].cpu().numpy().view(np.uint8))
               for f in ("fate", "n_boundary", "n_scatter", "t", "x", "y", "z", "path"))
    steps = int(rec.steps().sum())
    # the library's automatic kernel choice and the autotuned records shape
    # (tuning.json), into device arrays as above; medians of 5 after a warm-up
    from gpu_optical_photons_amd import autotune

    def _records_ms(tuning):
        backend.run_gpu_records_tuned(p, n, tuning, device="cuda:0")
        runs = [backend.run_gpu_records_tuned(p, n, tuning, device="cuda:0") for _ in range(5)]
        same_as_history = all(
            np.array_equal(r[0][f].cpu().numpy().view(np.uint8), dev[f].cpu().numpy().view(np.uint8))
            for r in runs[:1] for f in ("fate", "n_boundary", "n_scatter", "t", "x", "y", "z", "path"))
        return (float(np.median([r[1].kernel_ms for r in runs])), runs[0][3].as_dict(),
                runs[0][2], same_as_history)

    auto_ms, auto_choice, auto_st, auto_same = _records_ms(None)
    tuned = autotune.records_choice_for(p, n)
    tuned_ms, tuned_choice, tuned_st, tuned_same = (_records_ms(tuned) if tuned
                                                    else (None, None, None, True))
    same = same and auto_same and tuned_same
    # branch counts of the same photons (history tally, counting variant)
    c = TallyConfig()
    tally, _, cnt = backend.run_gpu_tally(p, c, n, counters=True)
    ops = roofline.algorithmic_ops(p, cnt.as_dict(), tally.fate_counts)
    # the same pipeline at 10x the photons, into fresh arrays (pages first
    # touched by the copies) and into arrays a caller reuses
    n10 = 10 * n
    backend.run_gpu_records(p, n10)
    fresh10 = min(backend.run_gpu_records(p, n10)[1].total_ms for _ in range(2))
    out10 = backend.PhotonRecords.empty(n10)
    for f in ("fate", "n_boundary", "n_scatter", "t", "x", "y", "z", "path"):
        getattr(out10, f).fill(0)
    reused10 = min(backend.run_gpu_records(p, n10, out=out10)[1].total_ms for _ in range(2))
    # CPU: the reference's records and tally runs on the same photons
    lib, kind = _cpu_lib()
    cores = int(lib.max_threads())
    cpu = {}
    for name, fn in (("records", lambda th: lib.records(p, n, threads=th)),
                     ("tally", lambda th: lib.tally(p, n, c.time_bins, c.time_max, c.xy_bins,
                                                    threads=th))):
        fn(0)
        t0 = time.perf_counter()
        fn(0)
        dt = time.perf_counter() - t0
        t1 = time.perf_counter()
        fn(1)
        d1 = time.perf_counter() - t1
        cpu[name] = {"value": n / dt, "unit": "photons/s", "cores": cores,
                     "one_thread_photons_per_s": n / d1}
    return {"photons": n, "history_kernel_ms": tm.kernel_ms,
            "history_photons_per_s": n / (tm.kernel_ms * 1e-3),
            "history_photon_steps_per_s": steps / (tm.kernel_ms * 1e-3),
            "wavefront_kernel_ms": tme.kernel_ms,
            "wavefront_photons_per_s": n / (tme.kernel_ms * 1e-3),
            "wavefront_photon_steps_per_s": steps / (tme.kernel_ms * 1e-3),
            "wavefront_simt_efficiency": st.simt_efficiency(),
            "roofline": roofline.roofline(ops, tm.kernel_ms, "history_kernel (records)"),
            "roofline_wavefront": roofline.roofline(ops, tme.kernel_ms,
                                                    "wavefront_kernel (records)"),
            "records_auto": {"chosen": auto_choice, "kernel_ms": auto_ms,
                             "photons_per_s": n / (auto_ms * 1e-3),
                             "simt_efficiency": (auto_st.simt_efficiency()
                                                 if auto_st.issued_lane_steps else None)},
            "records_autotuned": None if tuned is None else {
                "tuning": tuned_choice, "kernel_ms": tuned_ms,
                "photons_per_s": n / (tuned_ms * 1e-3),
                "simt_efficiency": (tuned_st.simt_efficiency()
                                    if tuned_st.issued_lane_steps else None)},
            "records_write_GBps": n * 29 / (tm.kernel_ms * 1e-3) / 1e9,
            "records_to_host_total_ms": total_ms, "records_to_host_d2h_ms": d2h_ms,
            "records_to_host_photons_per_s": n / (total_ms * 1e-3),
            "records_to_host_GBps": n * 29 / (total_ms * 1e-3) / 1e9,
            "records_identical": bool(same), "photon_steps": steps,
            "cpu_baseline": {"kind": kind, "cores": cores, "nproc": os.cpu_count(),
                             "cpu_model": _cpu_model(), "sample": f"photon ids 0..{n - 1}",
                             **cpu},
            "records_to_host_10x": {
                "photons": n10, "fresh_arrays_ms": fresh10, "reused_arrays_ms": reused10,
                "fresh_GBps": n10 * 29 / (fresh10 * 1e-3) / 1e9,
                "reused_GBps": n10 * 29 / (reused10 * 1e-3) / 1e9}}


def bench_gensteps():
    """Photons generated on the device from 100 000 gensteps (synthetic track
    segments <= 10 mm in the long-histories box, 200 photons/mm; generator seed
    1), tally 256 x 100 ns x 64^2: one persistent launch; then the same with
    a third of the gensteps Cherenkov, and the records of the first 10 000
    gensteps copied into host arrays."""
    from gpu_optical_photons_amd.gensteps import make_gensteps, run_gpu_gensteps
    from gpu_optical_photons_amd.params import TallyConfig
    p = c2_params()
    c = TallyConfig(256, 100.0, 64)
    g = make_gensteps(100_000, p, photons_per_mm=200.0, seed=1)
    n = int(g["n_photons"].sum())
    out = {"gensteps": len(g), "photons": n}
    for name, tu in (("registers", dict(block_size=512, steps_per_refill=2, queue_chunk=256)),
                     ("lds_ring", dict(block_size=512, steps_per_refill=2, queue_chunk=256,
                                       stream_in_lds=1))):
        run_gpu_gensteps(p, c, g, tuning=tu)
        settle_clocks(lambda: run_gpu_gensteps(p, c, g, tuning=tu), 0.3)
        ms = []
        for _ in range(3):
            _, _, tm, _ = run_gpu_gensteps(p, c, g, tuning=tu)
            ms.append(tm.kernel_ms)
        out[name] = {"kernel_ms": min(ms), "photons_per_s": n / (min(ms) * 1e-3)}
        out["upload_ms"] = tm.h2d_ms
    best = min(("registers", "lds_ring"), key=lambda k: out[k]["kernel_ms"])
    out["kernel"] = best
    out["kernel_ms"] = out[best]["kernel_ms"]
    out["photons_per_s"] = out[best]["photons_per_s"]
    # the same segments with a third of them Cherenkov gensteps (photons on the
    # cone about the segment, beta ~U[1/n, 1]), on the faster kernel
    gc = make_gensteps(100_000, p, photons_per_mm=200.0, seed=1, cherenkov_fraction=1.0 / 3.0)
    nc = int(gc["n_photons"].sum())
    tu = (dict(block_size=512, steps_per_refill=2, queue_chunk=256, stream_in_lds=1)
          if best == "lds_ring" else dict(block_size=512, steps_per_refill=2, queue_chunk=256))
    run_gpu_gensteps(p, c, gc, tuning=tu)
    ms = min(run_gpu_gensteps(p, c, gc, tuning=tu)[2].kernel_ms for _ in range(3))
    out["with_cherenkov"] = {"cherenkov_gensteps": int((gc["cone_cos"] > 0).sum()),
                             "photons": nc, "kernel_ms": ms, "photons_per_s": nc / (ms * 1e-3)}
    # records of the first 10 000 gensteps' photons into fresh host arrays
    # (29 B per photon, staged through pinned memory)
    gr = g[:10_000]
    nr = int(gr["n_photons"].sum())
    tms = [run_gpu_gensteps(p, c, gr, records=True)[2] for _ in range(3)]
    tm = min(tms, key=lambda t: t.total_ms)
    out["records_to_host"] = {"gensteps": len(gr), "photons": nr, "total_ms": tm.total_ms,
                              "kernel_ms": tm.kernel_ms, "d2h_ms": tm.d2h_ms,
                              "d2h_GBps": nr * 29 / (tm.d2h_ms * 1e6)}
    return out


def bench_sorted(p, c, n, fused_tally, counters, ring_tuning):
    """Sorting by interaction type, two forms, on the long-histories workload
    (same photons as the headline line):
      * the per-interaction path (optphot_run_gpu_tally_sorted): photon state
        in HBM, per iteration a classify kernel (emission + the step's common
        part, sorting photons onto per-block queues), a Fresnel kernel and a
        rotation kernel; bound by HBM traffic, so its roofline is bytes;
      * the ring kernel with each step's Fresnel and rotation work sorted onto
        block-wide LDS queues (wavefront tuning sorted_phases=1).
    Both must reproduce the headline tally word for word."""
    from gpu_optical_photons_amd import backend, roofline
    backend.run_gpu_tally_sorted(p, c, n)  # warm-up (and graph capture)
    ms, st, t = [], None, None
    for _ in range(3):
        t, tm, st = backend.run_gpu_tally_sorted(p, c, n)
        ms.append(tm.kernel_ms)
    best = min(ms)
    b = roofline.sorted_path_bytes(counters, n, st.slots, st.iterations,
                                   int(fused_tally.fate_counts[0]))
    out = {"path": {"kernel_ms": best, "photons_per_s": n / (best * 1e-3),
                    "identical_to_fused": bool((t.flat() == fused_tally.flat()).all()),
                    "stats": st.as_dict(), "algorithmic_bytes": b,
                    "achieved_GBps": b / (best * 1e-3) / 1e9,
                    "hbm_frac": b / (best * 1e-3) / roofline.HBM_PEAK_BYTES_PER_S}}
    tu = dict(ring_tuning, sorted_phases=1)
    backend.run_gpu_tally_wavefront(p, c, n, tuning=tu, counters=False)
    ms = []
    for _ in range(3):
        t2, tm, _, _ = backend.run_gpu_tally_wavefront(p, c, n, tuning=tu, counters=False)
        ms.append(tm.kernel_ms)
    out["lds_phases"] = {"tuning": tu, "kernel_ms": min(ms),
                         "photons_per_s": n / (min(ms) * 1e-3),
                         "identical_to_fused": bool((t2.flat() == fused_tally.flat()).all())}
    return out


def bench_reference_workloads():
    """The reference's own published workloads (README.md:540-552), 1e8
    photons in tally mode with its default binning (100 bins x 50 ns, 50^2):
    configs/default.cfg (50 mm cube, polished faces, point source) and
    configs/wrapped_lambertian.cfg (Lambertian R 0.95, n_det 1.5, abs 1000 mm,
    scat 2000 mm, volume source), seed 12345. The reference's figures are for
    its history kernel on a GTX 1660 Ti: context, not a like-for-like
    baseline."""
    import math

    from gpu_optical_photons_amd import autotune, backend
    from gpu_optical_photons_amd.params import SimParams, Source, Surface, TallyConfig
    c = TallyConfig()
    n = 100_000_000
    cases = {
        "default_cube": (SimParams(), 9.3e8),
        "wrapped_lambertian": (SimParams(surface=Surface.LAMBERTIAN, reflectivity=0.95, n_det=1.5,
                                         abs_length=1000.0, scat_length=2000.0,
                                         source=Source.VOLUME, tau=4.0), 3.3e8),
    }
    out = {}
    for name, (p, published) in cases.items():
        assert not math.isinf(p.abs_length)
        tu = autotune.choice_for("wavefront_tally", p, c)
        backend.run_gpu_tally_wavefront(p, c, n, tuning=tu, counters=False)
        ms = min(backend.run_gpu_tally_wavefront(p, c, n, tuning=tu, counters=False)[1].kernel_ms
                 for _ in range(3))
        out[name] = {"photons": n, "kernel_ms": ms, "photons_per_s": n / (ms * 1e-3),
                     "reference_published_photons_per_s": published,
                     "reference_hardware": "GTX 1660 Ti, history kernel"}
    return out


def bench_batched(cpu=True):
    """Batched events: 4096 events (box ~U[10,100] mm per axis, surface model
    uniform over polished / black / specular / Lambertian, n_det ~U[1.4,1.6],
    abs ~U[300,3000] mm, scat inf or ~U[100,1000] mm, point source uniform,
    N ~ Poisson(1e4 E), E ~U[0.1,3] MeV; generator seed 2024, run seed 1),
    per-event detected count + 128-bin time histogram 0-50 ns. One persistent
    graph-captured launch (the autotuned batch shape of this batch class, if
    tuning.json holds one) vs one history-kernel launch per event; roofline
    from the branch counters of the same photons; the CPU baseline is the
    reference's CPU tally run per event with the same photon-id offsets."""
    import numpy as np

    from gpu_optical_photons_amd import autotune, roofline
    from gpu_optical_photons_amd.batch import PLAIN, WAVEFRONT, make_events, run_gpu_batch
    b = make_events(4096, seed=2024)
    tu = autotune.batch_choice_for(b, 128)
    run_gpu_batch(b, 128, 50.0, method=WAVEFRONT, tuning=tu)  # capture
    settle_clocks(lambda: run_gpu_batch(b, 128, 50.0, method=WAVEFRONT, tuning=tu), 0.3)
    ms = []
    for _ in range(5):
        w, tw, _ = run_gpu_batch(b, 128, 50.0, method=WAVEFRONT, tuning=tu)
        ms.append(tw.kernel_ms)
    w_ms = float(np.median(ms))
    pl, tp, _ = run_gpu_batch(b, 128, 50.0, method=PLAIN)
    wc, _, cnt = run_gpu_batch(b, 128, 50.0, method=WAVEFRONT, tuning=tu, counters=True)
    counters = cnt.as_dict()
    fates = w[:, :5].sum(axis=0, dtype=np.uint64)
    volume = int(sum(int(n) for q, n in zip(b.params, b.photons) if q.source == 1))
    ops = roofline.algorithmic_ops(None, counters, fates, emitted_volume=volume)
    steps = roofline.photon_steps(counters)
    kernel = "batch_ring_kernel" if (tu or {}).get("stream_in_lds") else "batch_kernel"
    cpu_baseline = None
    if cpu:  # the reference's tally of every event, photon ids at the event's offset
        lib, kind = _cpu_lib()
        cores = int(lib.max_threads())
        offs = b.offsets
        t0 = time.perf_counter()
        for e in range(b.n_events):
            lib.tally(b.params[e], int(b.photons[e]), 128, 50.0, 1, first_id=int(offs[e]),
                      threads=0)
        dt = time.perf_counter() - t0
        cpu_baseline = {"value": b.total / dt, "unit": "photons/s", "cores": cores, "kind": kind,
                        "nproc": os.cpu_count(), "cpu_model": _cpu_model(),
                        "sample": f"all {b.n_events} events, one tally call per event "
                                  f"({dt:.1f} s wall)"}
    return {"events": b.n_events, "photons": b.total, "tuning": tu, "wavefront_ms": w_ms,
            "wavefront_ms_runs": ms, "plain_ms": tp.kernel_ms,
            "wavefront_photons_per_s": b.total / (w_ms * 1e-3),
            "wavefront_photon_steps_per_s": steps / (w_ms * 1e-3),
            "plain_photons_per_s": b.total / (tp.kernel_ms * 1e-3),
            "plain_photon_steps_per_s": steps / (tp.kernel_ms * 1e-3),
            "simt_efficiency": counters["useful_lane_steps"] / counters["issued_lane_steps"],
            "roofline": roofline.roofline(ops, w_ms, kernel),
            "identical": bool(np.array_equal(w, pl) and np.array_equal(w, wc)),
            "cpu_baseline": cpu_baseline}


def bench_differentiation():
    """Single-geometry differentiation: plain-small box with scat 300 mm,
    1e7 photons, loss = chi-square against a target at (abs 1500, scat 400,
    R 0.95, n_det 1.45, n_scint 1.60), seed 777. Forward (tally) + adjoint
    (likelihood-ratio + pathwise) timings, and the gradient against central
    finite differences (relative step 1e-3, common random numbers) of the
    linearised loss on 4e9 photons in 16 disjoint id ranges. Roofline of the
    forward launch from its branch counters (the adjoint recomputes the same
    transport; its scores are not counted, so its figure is a lower bound).
    CPU: the reference's forward tally and the oracle's FP64 gradient."""
    import math

    import numpy as np

    from gpu_optical_photons_amd import adjoint, roofline
    from gpu_optical_photons_amd import calibrate as cal
    from gpu_optical_photons_amd.batch import EventBatch, run_gpu_batch
    n = 10_000_000
    tpl, T = cal.c4_problem(n)
    theta = cal.C4_NOMINAL.copy()
    cal.loss_and_grad(tpl, theta, T, n, n)  # warm
    runs = [cal.loss_and_grad(tpl, theta, T, n, n) for _ in range(3)]
    f_ms = float(np.median([r[2] for r in runs]))
    a_ms = float(np.median([r[3] for r in runs]))
    loss, g, _, _, dl_dh = runs[0]
    p = adjoint.with_theta(tpl, theta)
    counts, _, cnt = run_gpu_batch(EventBatch([p], np.array([n], np.uint64)), cal.C4_BINS,
                                   cal.C4_TMAX, counters=True)
    counters = cnt.as_dict()
    ops = roofline.algorithmic_ops(p, counters, counts[0, :5])
    steps = roofline.photon_steps(counters)
    nc = 4_000_000_000
    lr, lr_se = cal.lr_gradient(tpl, theta, T, nc, n, parts=16, dl_dh=dl_dh)
    fd, fd_se = cal.fd_gradient(tpl, theta, T, nc, n, rel_step=1e-3, parts=16, dl_dh=dl_dh)
    names = ("abs_length", "scat_length", "reflectivity", "n_det", "n_scint")
    check = {}
    for j, nm in enumerate(names):
        se = math.hypot(lr_se[j], fd_se[j])
        check[nm] = {"adjoint": lr[j], "adjoint_se": lr_se[j], "fd": fd[j], "fd_se": fd_se[j],
                     "rel_diff": (lr[j] - fd[j]) / abs(fd[j]) if fd[j] else 0.0,
                     "pull": (lr[j] - fd[j]) / se if se else 0.0}
    # CPU: forward tally (reference) on the same photons, FP64 gradient (oracle
    # port) on a bounded sample
    from oracle.oracle import Oracle
    lib, kind = _cpu_lib()
    cores = int(lib.max_threads())
    t0 = time.perf_counter()
    lib.tally(p, n, cal.C4_BINS, cal.C4_TMAX, 1, threads=0)
    cf = time.perf_counter() - t0
    orc = Oracle("libm")
    m = 1_000_000
    t0 = time.perf_counter()
    orc.gradient([p], [m], cal.C4_BINS, cal.C4_TMAX, dl_dh[None, :])
    cg = time.perf_counter() - t0
    cpu_fwd = n / cf
    cpu_adj = m / cg
    return {"photons": n, "loss": loss, "grad": dict(zip(names, g.tolist())),
            "forward_ms": f_ms, "adjoint_ms": a_ms,
            "fwd_adj_over_fwd": (f_ms + a_ms) / f_ms,
            "forward_photons_per_s": n / (f_ms * 1e-3),
            "forward_photon_steps_per_s": steps / (f_ms * 1e-3),
            "fwd_adj_photons_per_s": n / ((f_ms + a_ms) * 1e-3),
            "roofline": roofline.roofline(ops, f_ms, "batch_kernel (forward, one set)"),
            "roofline_fwd_adj": roofline.roofline(
                2 * ops, f_ms + a_ms, "batch_kernel + adjoint_kernel",
                "transport operations of both passes; the adjoint's scores are not counted"),
            "gradient_check_photons": nc, "gradient_vs_fd": check,
            "cpu_baseline": {"value": 1.0 / (1.0 / cpu_fwd + 1.0 / cpu_adj), "unit": "photons/s",
                             "cores": cores, "kind": f"{kind} forward + port FP64 gradient",
                             "forward_photons_per_s": cpu_fwd, "gradient_photons_per_s": cpu_adj,
                             "nproc": os.cpu_count(), "cpu_model": _cpu_model(),
                             "sample": f"forward: photon ids 0..{n - 1}; gradient: ids "
                                       f"0..{m - 1}"}}


def bench_gradient_checks(photons=4_000_000_000):
    """The adjoint against central finite differences (common random
    numbers) of the linearised chi-square loss, `photons` per side in 16
    disjoint id ranges, with the error per parameter:
      * lambertian: the calibration geometry, R = 0.97 (the reflectivity
        enters through every wall decision);
      * specular: a 20 x 60 x 40 mm specular box, R = 0.9;
      * polished: the single-geometry workload with the FD step scanned over
        relative 1e-3, 3e-4 and 1e-4 for n_det and n_scint (FD curvature
        against estimator bias)."""
    from gpu_optical_photons_amd import calibrate as cal
    out = {"photons_per_side": photons}
    for name in ("lambertian", "specular"):
        tpl, theta, T, bins, tmax = cal.gradient_problem(name, 20_000_000)
        out[name] = cal.gradient_check(tpl, theta, T, photons, 20_000_000, bins, tmax)
    tpl, theta, T, bins, tmax = cal.gradient_problem("polished", 10_000_000)
    w = cal.linearised_weights(tpl, theta, T, 10_000_000, 10_000_000, bins, tmax)
    scan = {}
    for step in (1e-3, 3e-4, 1e-4):
        scan[f"{step:g}"] = cal.gradient_check(tpl, theta, T, photons, 10_000_000, bins, tmax,
                                               rel_step=step, params=(3, 4), dl_dh=w)
    out["polished_step_scan"] = scan
    # The point source's emission cosines lie on a 2^-23 grid that the
    # polished walls keep up to the readout, so the adjoint (the derivative
    # at fixed indices) depends on where that grid falls against the critical
    # cosine. One float32 ulp of n_det moves the critical cosine by 1.44
    # grid cells: the mean over 16 consecutive ulps is the smooth derivative
    # that finite differences see.
    import numpy as np
    nd = np.float32(theta[3])
    per_ulp = []
    for k in range(16):
        th_k = theta.copy()
        th_k[3] = float(nd)
        g_k, _ = cal.lr_gradient(tpl, th_k, T, 1_000_000_000, 10_000_000, parts=4, dl_dh=w,
                                 bins=bins, tmax=tmax)
        per_ulp.append([float(g_k[3]), float(g_k[4])])
        nd = np.nextafter(nd, np.float32(np.inf))
    per_ulp = np.array(per_ulp)
    fd1 = scan["0.001"]
    out["polished_emission_grid_average"] = {
        "n_det_ulps": 16, "photons_per_point": 1_000_000_000,
        "adjoint_n_det": per_ulp[:, 0].tolist(), "adjoint_n_scint": per_ulp[:, 1].tolist(),
        "mean_n_det": float(per_ulp[:, 0].mean()), "mean_n_scint": float(per_ulp[:, 1].mean()),
        "fd_n_det": fd1["n_det"]["fd"], "fd_n_det_se": fd1["n_det"]["fd_se"],
        "fd_n_scint": fd1["n_scint"]["fd"], "fd_n_scint_se": fd1["n_scint"]["fd_se"],
        "mean_rel_diff_n_det": float(per_ulp[:, 0].mean() / fd1["n_det"]["fd"] - 1.0),
        "mean_rel_diff_n_scint": float(per_ulp[:, 1].mean() / fd1["n_scint"]["fd"] - 1.0)}
    # exact case: mirror box with a Fresnel readout (float64 quadrature;
    # tests/golden/mirror_box_derivative.json)
    g = json.load(open(os.path.join(ROOT, "tests", "golden", "mirror_box_derivative.json")))
    import math

    import numpy as np

    from gpu_optical_photons_amd.params import SimParams
    p = SimParams(**g["params"], abs_length=math.inf, scat_length=math.inf, surface=2,
                  reflectivity=g["reflectivity"], tau=0.0, seed=21)
    th = np.array([math.inf, math.inf, g["reflectivity"], 1.5, 1.63])
    w = np.ones(9)
    lr, lr_se = cal.lr_gradient(p, th, None, 4 * photons, 1, parts=16, dl_dh=w, bins=8, tmax=5000.0)
    fd, fd_se = cal.fd_gradient(p, th, None, photons, 1, parts=16, dl_dh=w, bins=8, tmax=5000.0,
                                params=(3, 4))
    mb = {}
    for j, key in ((3, "ndet"), (4, "nscint")):
        smooth = g[f"dP_d{key}_smooth"]
        quant = smooth + g[f"grid_correction_{key}"]
        mb[key] = {"adjoint": float(lr[j]), "adjoint_se": float(lr_se[j]),
                   "quantised_exact": quant, "adjoint_pull": float((lr[j] - quant) / lr_se[j]),
                   "fd": float(fd[j]), "fd_se": float(fd_se[j]), "smooth_exact": smooth,
                   "fd_pull": float((fd[j] - smooth) / fd_se[j])}
    out["mirror_box_fresnel_exact"] = {"adjoint_photons": 4 * photons, "fd_photons_per_side": photons,
                                       **mb}
    return out


def bench_calibration():
    """Batched differentiable calibration: 256 parameter sets x 1e6 photons,
    long-histories geometry, (abs_length, R, n_det) fitted by 50 Adam
    iterations to targets from hidden parameters under a held-out seed; each
    iteration one captured graph (forward, chi-square weights, adjoint). The
    roofline is that of a forward batch launch at the fitted parameters, from
    its branch counters."""
    import numpy as np

    from gpu_optical_photons_amd import adjoint, roofline
    from gpu_optical_photons_amd import calibrate as cal
    from gpu_optical_photons_amd.batch import EventBatch, run_gpu_batch
    r = cal.calibrate(n_sets=256, photons=1_000_000, iters=50)
    tpl = cal.c5_template(r["seed"])
    batch = EventBatch([adjoint.with_theta(tpl, q) for q in cal._theta_rows(r["fitted"], 256)],
                       np.full(256, 1_000_000, np.uint64))
    tu = r["forward_tuning"]
    run_gpu_batch(batch, r["bins"], r["tmax"], tuning=tu)
    f_ms = float(np.median([run_gpu_batch(batch, r["bins"], r["tmax"], tuning=tu)[1].kernel_ms
                            for _ in range(3)]))
    counts, _, cnt = run_gpu_batch(batch, r["bins"], r["tmax"], tuning=tu, counters=True)
    counters = cnt.as_dict()
    fates = counts[:, :5].sum(axis=0, dtype=np.uint64)
    ops = roofline.algorithmic_ops(None, counters, fates, emitted_volume=batch.total)
    kernel = "batch_ring_kernel" if (tu or {}).get("stream_in_lds") else "batch_kernel"
    out = {k: r[k] for k in ("n_sets", "photons_per_set", "iterations", "median_rel_error",
                             "p90_rel_error", "graph_ms", "forward_only_ms", "wall_s",
                             "forward_photons_per_s", "fwd_adj_photons_per_s",
                             "fwd_adj_over_fwd", "forward_tuning")}
    out.update({
        "final_median_loss": r["losses"][-1], "first_median_loss": r["losses"][0],
        "forward_at_fit_ms": f_ms,
        "forward_photon_steps_per_s": roofline.photon_steps(counters) / (f_ms * 1e-3),
        "roofline": roofline.roofline(ops, f_ms, kernel + " (forward, 256 sets)"),
        "roofline_graph": roofline.roofline(
            2 * ops * r["iterations"], r["graph_ms"], "calibration step graph",
            "forward + adjoint transport operations at the fitted parameters x 50; the "
            "adjoint's scores are not counted"),
        "diagnostics": cal.calibration_diagnostics(r)})
    return out


def load_traffic():
    path = os.path.join(ROOT, "profiles", "wavefront_tally_c2_traffic.json")
    if not os.path.exists(path):
        return None
    try:
        return json.load(open(path))
    except Exception:
        return None


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--photons", type=int, default=N_PHOTONS)
    ap.add_argument("--only-main", action="store_true",
                    help="only the long-histories line (no other workloads)")
    ap.add_argument("--sorted-only", action="store_true",
                    help="only the sorting-by-interaction-type measurements on the "
                         "long-histories workload")
    ap.add_argument("--batched-only", action="store_true",
                    help="only the batched-events measurement (4096 events, autotuned shape, "
                         "against one history launch per event)")
    ap.add_argument("--gensteps-only", action="store_true",
                    help="only the generation-from-gensteps measurements (10^5 gensteps in "
                         "the long-histories box)")
    ap.add_argument("--gradients-only", action="store_true",
                    help="only the adjoint-against-finite-difference checks (4e9 photons per "
                         "side at the Lambertian, specular and polished geometries)")
    ap.add_argument("--calibration-only", action="store_true",
                    help="only the calibration measurement (256 parameter sets, one graph per "
                         "forward + loss + adjoint step)")
    ap.add_argument("--valu-probe", action="store_true",
                    help="only run the library's VALU calibration kernel (all lanes, then odd "
                         "lanes only): a known instruction count for reading rocprofv3 VALU "
                         "counters")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a plain 1 GiB device-to-device copy (torch), the known-size "
                         "reference for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE counters")
    args = ap.parse_args()
    if args.copy_probe:
        import torch
        a = torch.empty(1 << 28, dtype=torch.float32, device="cuda:0").fill_(1.0)
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_read": a.numel() * 4,
                          "copy_probe_bytes_written": b.numel() * 4, "copies": 3}))
        return
    if args.valu_probe:
        import ctypes

        # the calibration kernel lives in a library of its own, beside the
        # transport library (gpu_optical_photons_amd/csrc/probe.hip)
        path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "gpu_optical_photons_amd",
                            "liboptphot_probe.so")
        if not os.path.exists(path):
            raise SystemExit(f"bench --valu-probe: {path} is not built (make -C "
                             "gpu_optical_photons_amd/csrc)")
        fn = ctypes.CDLL(path).ophip_valu_probe
        fn.restype = ctypes.c_int
        fn.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int,
                       ctypes.POINTER(ctypes.c_double)]

        def check(rc):
            if rc != 0:
                raise RuntimeError(f"valu probe failed ({rc})")
        blocks, iters = 2048, 1 << 15
        out = {"blocks": blocks, "threads_per_block": 256, "iters": iters,
               "fma_per_wave": 8 * iters}
        for half in (0, 1):
            ms = ctypes.c_double()
            check(fn(blocks, iters, half, ctypes.byref(ms)))  # warm-up
            check(fn(blocks, iters, half, ctypes.byref(ms)))
            waves = blocks * 4
            key = "half_lanes" if half else "all_lanes"
            out[key] = {"kernel_ms": ms.value,
                        "wave_instructions_per_s": waves * 8 * iters / (ms.value * 1e-3),
                        "lane_fma_per_s": waves * 8 * iters * (32 if half else 64)
                        / (ms.value * 1e-3)}
        print(json.dumps(out))
        return
    if args.batched_only:
        print(json.dumps({"batched_events": bench_batched(cpu=not args.no_cpu)}))
        return
    if args.gensteps_only:
        print(json.dumps({"gensteps": bench_gensteps()}))
        return
    if args.gradients_only:
        print(json.dumps({"gradient_checks": bench_gradient_checks()}))
        return
    if args.calibration_only:
        print(json.dumps({"calibration": bench_calibration()}))
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(a multi-GPU run would shard photon-id ranges; only one GPU is used here)", file=sys.stderr)
        sys.exit(2)

    import numpy as np

    from gpu_optical_photons_amd import autotune, backend, roofline
    from gpu_optical_photons_amd.params import TallyConfig

    if not backend.gpu_available():
        print("bench.py: no HIP device", file=sys.stderr)
        sys.exit(1)
    p = c2_params()
    c = TallyConfig(256, 100.0, 64)
    n = args.photons
    tuning = autotune.choice_for("wavefront_tally", p, c)
    kernel_name = ("wavefront_phase_kernel" if tuning.get("sorted_phases")
                   else "wavefront_ring_kernel" if tuning.get("stream_in_lds")
                   else "wavefront_next_kernel" if tuning.get("successor_threshold")
                   else "wavefront_kernel")

    if args.sorted_only:
        ring = autotune.choice_for("wavefront_tally", p, c)
        fused, _, _, cnt = backend.run_gpu_tally_wavefront(p, c, n, tuning=ring, counters=True)
        print(json.dumps({"workload": WORKLOAD, "photons": n,
                          "sorted": bench_sorted(p, c, n, fused, cnt.as_dict(), ring)}))
        return
    settle_clocks(lambda: backend.run_gpu_tally_wavefront(p, c, n, tuning=tuning, counters=False))
    for _ in range(args.warmup):
        backend.run_gpu_tally_wavefront(p, c, n, tuning=tuning, counters=False)
    kernel_ms, transport_ms = [], []
    tally = None
    t0 = time.perf_counter()
    for _ in range(args.steps):
        tally, tm, st, _ = backend.run_gpu_tally_wavefront(p, c, n, tuning=tuning, counters=False)
        kernel_ms.append(tm.kernel_ms)
        transport_ms.append(st.kernels_ms)
    wall = time.perf_counter() - t0
    ms_per_step = wall * 1e3 / args.steps
    # Branch-event counts of the same photons (a property of the workload, so
    # one untimed run of the counting variant gives them exactly).
    tally_c, _, _, cnt = backend.run_gpu_tally_wavefront(p, c, n, tuning=tuning, counters=True)
    if not np.array_equal(tally_c.flat(), tally.flat()):
        raise RuntimeError("bench: counting and timed kernel variants disagree")
    counters = cnt.as_dict()
    steps_per_pass = roofline.photon_steps(counters)
    value = n / (ms_per_step * 1e-3)
    transport_avg = float(np.mean(transport_ms))
    ops = roofline.algorithmic_ops(p, counters, tally.fate_counts)
    achieved = ops / (transport_avg * 1e-3) / 1e12

    # Plain history-based kernel on the same workload (one lane per photon).
    from gpu_optical_photons_amd.backend import run_gpu_tally
    hist, htm = run_gpu_tally(p, c, n)  # warm
    hist, htm = run_gpu_tally(p, c, n)
    _, _, hcnt = run_gpu_tally(p, c, n, counters=True)
    hc = hcnt.as_dict()
    identical = bool(np.array_equal(hist.flat(), tally.flat()))

    line = {
        "metric": "photons/s (wavefront kernel, long-histories workload)",
        "value": value,
        "unit": "photons/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (photons generated on the device from Philox streams)",
        "config": {"workload": WORKLOAD, "photons_per_step": n, "box_mm": [100, 100, 50],
                   "surface": "lambertian R=0.97", "scat_length_mm": 300,
                   "abs_length_mm": 2000, "n_det": 1.5, "source": "volume", "seed": 1,
                   "time_bins": 256, "time_max_ns": 100, "xy_bins": 64,
                   "kernel": "wavefront (persistent, device work queue)",
                   "tuning": tuning},
        "photon_steps_per_s": steps_per_pass / (ms_per_step * 1e-3),
        "photon_steps_per_photon": steps_per_pass / n,
        "kernel_ms_per_step": float(np.mean(kernel_ms)),
        "simt_efficiency": counters["useful_lane_steps"] / counters["issued_lane_steps"],
        "plain_history_kernel": {
            "photons_per_s": n / (htm.kernel_ms * 1e-3),
            "kernel_ms": htm.kernel_ms,
            "simt_efficiency": hc["useful_lane_steps"] / hc["issued_lane_steps"],
            "tally_identical_to_wavefront": identical,
        },
        "roofline": {
            "bound": "valu",
            "achieved": achieved,
            "peak": roofline.FP32_VECTOR_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / roofline.FP32_VECTOR_PEAK_TFLOPS,
            "traffic": None,
            "kernel": kernel_name,
            "kernel_ms": transport_avg,
            "kernel_ms_note": "transport kernel + tally reduction, HIP events on the library stream",
            "algorithmic_ops_per_launch": ops,
            "frac_of_valu_issue_peak": achieved / roofline.VALU_ISSUE_PEAK_TOPS,
            "valu_issue_peak_tops": roofline.VALU_ISSUE_PEAK_TOPS,
        },
    }
    tr = load_traffic()
    if tr and tr.get("kernel", "wavefront_kernel") == kernel_name:
        line["roofline"]["traffic"] = tr.get("hbm_bytes_per_launch")
        line["roofline"]["traffic_source"] = tr.get("source")
    if not args.only_main:
        line["plain_small"] = bench_plain_small()
        line["batched_events"] = bench_batched()
        line["batched_vs_plain"] = (line["batched_events"]["plain_ms"]
                                    / line["batched_events"]["wavefront_ms"])
        # the same photons through both kernels at one large launch: what the
        # wavefront kernel itself gains over the history kernel, apart from
        # the launch count of the per-event baseline
        line["wavefront_over_history_same_photons"] = htm.kernel_ms / transport_avg
        line["gensteps"] = bench_gensteps()
        line["differentiation"] = bench_differentiation()
        line["calibration"] = bench_calibration()
        line["sorted_by_interaction"] = bench_sorted(p, c, n, tally, counters, tuning)
        line["reference_workloads"] = bench_reference_workloads()
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(p, c)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
