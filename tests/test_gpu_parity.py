"""GPU parity: the HIP kernels against the oracle, photon by photon.

The oracle is the plain-C restatement of the reference's CPU path
(oracle/optphot_oracle.c) in the reference's bit-reproducible math mode
(OPTPHOT_PORTABLE_MATH); the kernels compute in the same mode, so every
record must be identical to the last bit (the reference's own bar for its
reproducible build, tests/test_gpu.cpp:91-94).
"""
import numpy as np
import pytest

from tests import configs

pytestmark = pytest.mark.gpu

FIELDS = ("fate", "n_boundary", "n_scatter", "t", "x", "y", "z", "path")


def assert_records_identical(gpu, ref, label=""):
    for f in FIELDS:
        a = getattr(gpu, f) if not isinstance(gpu, dict) else gpu[f]
        b = ref[f] if isinstance(ref, dict) else getattr(ref, f)
        if not np.array_equal(a.view(np.uint8), b.view(np.uint8)):
            bad = np.nonzero(a.view(f"u{a.itemsize}") != b.view(f"u{b.itemsize}"))[0]
            i = int(bad[0])
            raise AssertionError(f"{label}: field {f} differs at {len(bad)} photons; first id {i}: "
                                 f"gpu {a[i]!r} oracle {b[i]!r}")


@pytest.mark.parametrize("name", sorted(configs.PARITY))
def test_history_records_bitwise(backend, oracle_portable, name):
    p = configs.PARITY[name]
    n = 200_003
    print(f"[parity] history records {name} n={n}", flush=True)
    gpu, tm = backend.run_gpu_records(p, n)
    ref = oracle_portable.records(p, n)
    assert_records_identical(gpu, ref, name)


def test_records_first_id_and_chunk(backend, oracle_portable):
    p = configs.every_process()
    n = 300_000
    a, _ = backend.run_gpu_records(p, n, block_size=256, first_id=0, chunk=1 << 24)
    b, _ = backend.run_gpu_records(p, n, block_size=64, first_id=0, chunk=70001)
    assert_records_identical(a, b.as_dict(), "chunk/block")
    tail, _ = backend.run_gpu_records(p, 500, first_id=1000)
    ref = oracle_portable.records(p, 500, first_id=1000)
    assert_records_identical(tail, ref, "first_id")
    assert np.array_equal(tail.t, a.t[1000:1500])


def test_records_pipeline_many_pieces(backend, oracle_portable):
    """Records copied out in pieces through the pinned staging ring, across
    chunk boundaries (device buffers reused), for both kernels."""
    p = configs.c1_plain()
    n = 3_000_000
    ref = oracle_portable.records(p, n)
    print("[parity] records pipeline history", flush=True)
    a, tm = backend.run_gpu_records(p, n, chunk=1_000_000)
    assert_records_identical(a, ref, "history pipeline")
    print("[parity] records pipeline wavefront", flush=True)
    ev, _, st = backend.run_gpu_records_event(p, n, 4, chunk=1_000_000)
    assert_records_identical(ev, ref, "wavefront pipeline")
    assert st.launches == 3
    # one chunk of 3e6 + 1: copy pieces of 375 040 photons (>= 2^18), which the
    # host splits over several threads; the odd count leaves odd byte ranges
    print("[parity] records pipeline, large pieces", flush=True)
    ref1 = oracle_portable.records(p, n + 1)
    b, _ = backend.run_gpu_records(p, n + 1)
    assert_records_identical(b, ref1, "large pieces")


@pytest.mark.parametrize("k", [0, 1, 5])
def test_device_records(backend, oracle_portable, k):
    p = configs.c2_long()
    n = 300_001
    print(f"[parity] device records k={k}", flush=True)
    out, tm, st = backend.run_gpu_records_device(p, n, steps_per_launch=k, first_id=17)
    ref = oracle_portable.records(p, n, first_id=17)
    host = {f: out[f].cpu().numpy() for f in out}
    assert_records_identical(host, ref, f"device k={k}")
    assert tm.d2h_ms == 0.0
    if k:
        assert st.useful_lane_steps > 0 and st.launches == 1


@pytest.mark.parametrize("name", ["default", "wrapped_lambertian", "capped", "c2_long"])
def test_event_records_equal_history_for_any_k(backend, name):
    p = configs.PARITY[name]
    n = 100_003
    hist, _ = backend.run_gpu_records(p, n)
    steps = int(hist.steps().sum())
    for k in (1, 3, 100_000):
        print(f"[parity] event {name} k={k}", flush=True)
        ev, tm, st = backend.run_gpu_records_event(p, n, k, block_size=64, chunk=30011)
        assert_records_identical(ev, hist.as_dict(), f"{name} k={k}")
        assert st.useful_lane_steps == steps
        assert st.issued_lane_steps >= st.useful_lane_steps
        assert st.launches == (n + 30010) // 30011


def test_single_launch_issued_lane_steps(backend):
    """With steps_per_launch >= max_steps each wave runs 64 consecutive photons
    to completion: issued = sum over 64-photon groups of 64 x longest."""
    p = configs.default()
    n = 200_000
    hist, _ = backend.run_gpu_records(p, n)
    ev, tm, st = backend.run_gpu_records_event(p, n, p.max_steps, block_size=256, chunk=n)
    steps = hist.steps()
    pad = (-n) % 64
    longest = np.concatenate([steps, np.zeros(pad, steps.dtype)]).reshape(-1, 64).max(axis=1)
    assert st.launches == 1
    assert st.issued_lane_steps == int(64 * longest.sum())


@pytest.mark.parametrize("name", ["default", "c1_plain", "c2_long", "wrapped_lambertian",
                                  "absorption"])
def test_tally_equals_oracle(backend, oracle_portable, name):
    from gpu_optical_photons_amd.params import TallyConfig
    p = configs.PARITY[name]
    n = 1_000_000
    c = TallyConfig(256, 100.0, 64) if name == "c2_long" else TallyConfig()
    ref = oracle_portable.tally(p, n, c.time_bins, c.time_max, c.xy_bins)
    for block in (256, 128):
        t, tm = backend.run_gpu_tally(p, c, n, block_size=block)
        np.testing.assert_array_equal(t.fate_counts, ref["fate_counts"])
        np.testing.assert_array_equal(t.time_hist, ref["time_hist"])
        np.testing.assert_array_equal(t.xy_hist, ref["xy_hist"])
        assert t.boundary_sum == ref["boundary_sum"]
        assert t.scatter_sum == ref["scatter_sum"]
    for tuning in (None, dict(block_size=64, steps_per_refill=1, queue_chunk=64),
                   dict(block_size=1024, steps_per_refill=16, queue_chunk=4096),
                   # successor-photon kernel
                   dict(block_size=256, queue_chunk=512, successor_threshold=32),
                   dict(block_size=64, queue_chunk=64, successor_threshold=1),
                   # Philox windows in an LDS ring, with deferred tallying
                   dict(block_size=512, steps_per_refill=2, queue_chunk=512, stream_in_lds=1),
                   dict(block_size=64, steps_per_refill=5, queue_chunk=64, stream_in_lds=1),
                   # each step's Fresnel and rotation work sorted onto LDS queues
                   dict(block_size=512, steps_per_refill=2, queue_chunk=512, sorted_phases=1),
                   dict(block_size=64, steps_per_refill=1, queue_chunk=64, sorted_phases=1),
                   dict(block_size=256, steps_per_refill=3, queue_chunk=256, sorted_phases=1)):
        t, tm, st, cnt = backend.run_gpu_tally_wavefront(p, c, n, tuning=tuning)
        np.testing.assert_array_equal(t.flat(), backend.Tally.from_flat(
            np.concatenate([ref["fate_counts"], ref["time_hist"], ref["xy_hist"].ravel(),
                            np.array([ref["boundary_sum"], ref["scatter_sum"]], np.uint64)]),
            c).flat())
        steps = ref["boundary_sum"] + ref["scatter_sum"] + int(ref["fate_counts"][1])
        assert cnt.as_dict()["useful_lane_steps"] == steps
        assert cnt.as_dict()["emitted"] == n
        # the launch that is timed: the same kernel without the branch counters
        t2, _, _, _ = backend.run_gpu_tally_wavefront(p, c, n, tuning=tuning, counters=False)
        np.testing.assert_array_equal(t2.flat(), t.flat())


@pytest.mark.parametrize("name,n", [("c1_plain", 1_000_000), ("c2_long", 1_000_000),
                                    ("wrapped_lambertian", 1_000_000)])
def test_records_against_default_build(backend, oracle_portable, oracle_libm, name, n):
    """The GPU against the reference's default (glibc) CPU build on the same
    seed: discrete histories identical for >= 99.99 % of photons; among those,
    arrival times within 1e-5 relative for >= 99.9 % and x, y, z within 1e-5
    of the box half-size for >= 99 % each. The GPU's records equal the
    portable build's bit for bit, so its differences from the default build
    are the portable build's: every same-history photon outside 1e-5, in time
    or any coordinate, is traced step by step in both builds and stays within
    the bound of rounding carried forward (tests/drift.py)."""
    from oracle.oracle import Oracle, Reference, reference_available
    from tests import drift
    ref = Reference("libm") if reference_available("libm") else Oracle("libm")
    p = configs.PARITY[name]
    gpu, _ = backend.run_gpu_records(p, n)
    assert_records_identical(gpu, oracle_portable.records(p, n), name)
    out = drift.check_agreement(gpu.as_dict(), ref.records(p, n), p)
    print(f"[{name}] {len(out)} same-history photons outside 1e-5 traced", flush=True)
    for i in out:
        drift.check_carried_rounding(oracle_portable, oracle_libm, p, int(i))


def test_rayleigh_table_equals_oracle(backend, oracle_portable):
    """The device's Rayleigh-cosine table (read by every transport kernel for
    its scattering cosines) against the CPU function, all 2^24 entries, bit
    for bit."""
    import ctypes
    fn = backend.lib().optphot_rayleigh_table
    fn.restype = ctypes.c_int
    fn.argtypes = [ctypes.c_void_p, ctypes.c_size_t]
    dev = np.empty(1 << 24, np.float32)
    backend._check(fn(dev.ctypes.data, dev.size))
    ref = oracle_portable.rayleigh_cos_table()
    np.testing.assert_array_equal(dev.view(np.uint32), ref.view(np.uint32))


@pytest.mark.parametrize("case", ["tiny_box", "huge_box", "huge_index", "near_matched"])
def test_division_paths_at_range_limits(backend, oracle_portable, case):
    """The kernels divide with a short exact sequence only where their
    operands are in range (box half sizes in [2^-60, 2^40], refractive indices
    in [2^-20, 2^20] with (n1/n2)^2 != 1 unless matched); outside, the IEEE
    divisions run. Boxes and indices beyond the limits, and indices one ulp
    apart (the edge of the Fresnel argument), must all give the oracle's
    records and tally."""
    from gpu_optical_photons_amd.params import SimParams, Source, Surface, TallyConfig
    p = {
        "tiny_box": SimParams(hx=2e-19, hy=3e-19, hz=1e-19, abs_length=4e-18, scat_length=5e-19,
                              surface=Surface.LAMBERTIAN, reflectivity=0.9, n_det=1.5,
                              source=Source.VOLUME, tau=0.0, seed=21),
        "huge_box": SimParams(hx=2e12, hy=3e12, hz=1.5e12, abs_length=3e13, scat_length=5e12,
                              surface=Surface.POLISHED, n_det=1.5, source=Source.VOLUME,
                              tau=4.0, seed=22),
        "huge_index": SimParams(n_det=3e6, n_out=1.0, abs_length=500.0, scat_length=200.0,
                                surface=Surface.POLISHED, seed=23),
        "near_matched": SimParams(n_det=float(np.nextafter(np.float32(1.63), np.float32(2))),
                                  n_out=float(np.nextafter(np.float32(1.63), np.float32(0))),
                                  abs_length=500.0, scat_length=200.0, surface=Surface.POLISHED,
                                  seed=24),
    }[case]
    n = 20_000
    ref = oracle_portable.records(p, n)
    gpu, _ = backend.run_gpu_records(p, n)
    assert_records_identical(gpu, ref, case)
    c = TallyConfig()
    rt = oracle_portable.tally(p, n, c.time_bins, c.time_max, c.xy_bins)
    t, _, _, _ = backend.run_gpu_tally_wavefront(
        p, c, n, tuning=dict(block_size=512, steps_per_refill=2, queue_chunk=512, stream_in_lds=1))
    np.testing.assert_array_equal(t.fate_counts, rt["fate_counts"])
    np.testing.assert_array_equal(t.time_hist, rt["time_hist"])
    np.testing.assert_array_equal(t.xy_hist, rt["xy_hist"])


def test_records_tuned_shapes_equal_history(backend, oracle_portable):
    """optphot_run_gpu_records_tuned: every kernel and shape, into host or
    device arrays, writes the oracle's records bit for bit; the automatic
    choice takes the history kernel for a few photons per lane and the
    wavefront kernel from 8 photons per lane of its grid."""
    from gpu_optical_photons_amd.backend import RECORDS_HISTORY, RECORDS_WAVEFRONT
    p = configs.PARITY["every_process"]
    n = 300_000
    ref = oracle_portable.records(p, n)
    for tu in (None, dict(kernel=RECORDS_HISTORY, block_size=64),
               dict(kernel=RECORDS_WAVEFRONT, block_size=128, steps_per_refill=1,
                    queue_chunk=64, blocks_per_cu=1),
               dict(kernel=RECORDS_WAVEFRONT, block_size=1024, steps_per_refill=8,
                    queue_chunk=1024, blocks_per_cu=2)):
        for device in (None, "cuda:0"):
            rec, _, _, chosen = backend.run_gpu_records_tuned(p, n, tu, device=device)
            if device is not None:
                rec = {f: rec[f].cpu().numpy() for f in FIELDS}
            assert_records_identical(rec, ref, f"{tu} {device}")
            if tu is not None:
                assert chosen.kernel == tu["kernel"]
    _, _, _, small = backend.run_gpu_records_tuned(p, 1000, None, device="cuda:0")
    assert small.kernel == RECORDS_HISTORY
    big_n = 8_000_000
    rec, _, st, big = backend.run_gpu_records_tuned(p, big_n, None, device="cuda:0")
    assert big.kernel == RECORDS_WAVEFRONT and st.issued_lane_steps > 0
    sample = np.arange(0, big_n, 997)
    part = oracle_portable.records(p, big_n)
    for f in FIELDS:
        assert np.array_equal(rec[f].cpu().numpy()[sample].view(np.uint8),
                              part[f][sample].view(np.uint8)), f
