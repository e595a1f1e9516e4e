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
    n = int(min(N_PHOTONS, max(probe, 10.0 / max(dt, 1e-6) * probe)))
    t0 = time.perf_counter()
    lib.tally(p, n, c.time_bins, c.time_max, c.xy_bins, threads=0)
    dt = time.perf_counter() - t0
    return {"value": n / dt, "unit": "photons/s", "cores": cores, "kind": kind,
            "sample": f"photon ids 0..{n - 1} of the same workload (tally mode), "
                      f"{dt:.1f} s wall on {cores} OpenMP threads"}


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
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(photon-id ranges would shard across GPUs, see DESIGN.md)", file=sys.stderr)
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
            "kernel": "wavefront_kernel<records=false, tally=true>",
            "kernel_ms": transport_avg,
            "kernel_ms_note": "transport kernel + tally reduction, HIP events on the library stream",
            "algorithmic_ops_per_launch": ops,
            "frac_of_valu_issue_peak": achieved / roofline.VALU_ISSUE_PEAK_TOPS,
            "valu_issue_peak_tops": roofline.VALU_ISSUE_PEAK_TOPS,
        },
    }
    tr = load_traffic()
    if tr:
        line["roofline"]["traffic"] = tr.get("hbm_bytes_per_launch")
        line["roofline"]["traffic_source"] = tr.get("source")
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(p, c)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
