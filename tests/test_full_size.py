"""The long-histories tally at its full size (10^8 photons), where the oracle
is too slow to run whole, through properties that do not depend on size:

* the autotuned wavefront kernel, the plain history kernel and the
  per-interaction path give the same tally word for word (photon histories
  depend only on (seed, photon id), test_gpu.cpp:122-146);
* splitting the id range (first_id, backend.hpp:75-91) and adding the parts
  gives the whole (test_transport_cpu.cpp:328-332);
* the tally's own sums close: every photon has one fate, every detected
  photon one time bin and one hit-map bin, and the counted photon-steps equal
  boundary + scatter + bulk absorptions (optphot_event_bench.cpp:39-41);
* the last 2*10^5 ids of the range, run alone, equal the oracle's tally of
  those ids, which ties the far end of the id space to the reference.
"""
import numpy as np
import pytest

from tests import configs

pytestmark = pytest.mark.gpu

N = 100_000_000


def _c2():
    from gpu_optical_photons_amd.params import TallyConfig
    return configs.PARITY["c2_long"], TallyConfig(256, 100.0, 64)


def test_full_size_kernels_agree_and_sums_close(backend):
    from gpu_optical_photons_amd import autotune
    p, c = _c2()
    tu = autotune.choice_for("wavefront_tally", p, c)
    print(f"[full size] wavefront {tu}", flush=True)
    wf, _, _, cnt = backend.run_gpu_tally_wavefront(p, c, N, tuning=tu)
    print("[full size] history kernel", flush=True)
    hist, _ = backend.run_gpu_tally(p, c, N)
    np.testing.assert_array_equal(wf.flat(), hist.flat())
    print("[full size] per-interaction path", flush=True)
    srt, _, st = backend.run_gpu_tally_sorted(p, c, N)
    np.testing.assert_array_equal(srt.flat(), wf.flat())
    assert st.emitted == N

    assert int(wf.fate_counts.sum()) == N
    detected = int(wf.fate_counts[0])
    assert int(wf.time_hist.sum()) == detected
    assert int(wf.xy_hist.sum()) == detected
    steps = int(wf.boundary_sum) + int(wf.scatter_sum) + int(wf.fate_counts[1])
    d = cnt.as_dict()
    assert d["useful_lane_steps"] == steps
    assert d["emitted"] == N and d["detected"] == detected
    assert d["scatter"] == int(wf.scatter_sum)


def test_full_size_split_ranges_add_up(backend):
    from gpu_optical_photons_amd import autotune
    p, c = _c2()
    tu = autotune.choice_for("wavefront_tally", p, c)
    whole, _, _, _ = backend.run_gpu_tally_wavefront(p, c, N, tuning=tu, counters=False)
    cuts = [0, 12_345_678, 50_000_000, 99_800_000, N]
    total = np.zeros_like(whole.flat())
    for a, b in zip(cuts[:-1], cuts[1:]):
        part, _, _, _ = backend.run_gpu_tally_wavefront(p, c, b - a, first_id=a, tuning=tu,
                                                        counters=False)
        total += part.flat()
    np.testing.assert_array_equal(total, whole.flat())


def test_full_size_far_end_equals_oracle(backend, oracle_portable):
    from gpu_optical_photons_amd import autotune
    p, c = _c2()
    tu = autotune.choice_for("wavefront_tally", p, c)
    first, n = N - 200_000, 200_000
    t, _, _, _ = backend.run_gpu_tally_wavefront(p, c, n, first_id=first, tuning=tu,
                                                 counters=False)
    ref = oracle_portable.tally(p, n, c.time_bins, c.time_max, c.xy_bins, first_id=first)
    want = backend.Tally.from_flat(
        np.concatenate([ref["fate_counts"], ref["time_hist"], ref["xy_hist"].ravel(),
                        np.array([ref["boundary_sum"], ref["scatter_sum"]], np.uint64)]), c)
    np.testing.assert_array_equal(t.flat(), want.flat())
