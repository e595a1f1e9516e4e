"""The per-interaction ("sorted") path against the oracle and the fused kernels.

The path splits every step over three kernels (classify with emission,
Fresnel, rotation) with photon state in device memory and per-type queues; the
reference names this design as its next step (README.md:717-720). A photon's
history depends only on (seed, photon id), so its tally must equal the CPU
path's word for word (the reference's bar for tallies, test_gpu.cpp:122-146)
for every tuning, as the reference requires of chunk / block / k
(test_gpu.cpp:148-189).
"""
import numpy as np
import pytest

from tests import configs

CASES = ["default", "c1_plain", "c2_long", "wrapped_lambertian", "absorption", "specular",
         "capped", "black_scatter", "mirror_trap", "every_process"]


def flat_ref(backend, ref, c):
    return backend.Tally.from_flat(
        np.concatenate([ref["fate_counts"], ref["time_hist"], ref["xy_hist"].ravel(),
                        np.array([ref["boundary_sum"], ref["scatter_sum"]], np.uint64)]),
        c).flat()


def test_sorted_tuning_rejected_before_any_device_work():
    from gpu_optical_photons_amd import backend as b
    from gpu_optical_photons_amd.params import SimParams, TallyConfig
    p, c = SimParams(), TallyConfig()
    for bad in (dict(slots_per_block=300), dict(slots_per_block=65536 + 256),
                dict(blocks_per_cu=65), dict(iterations_per_poll=3),
                dict(iterations_per_poll=2048)):
        with pytest.raises(ValueError, match="sorted"):
            b.run_gpu_tally_sorted(p, c, 10, tuning=bad)
    with pytest.raises(ValueError, match="2\\^32"):
        b.run_gpu_tally_sorted(p, c, 2**32)


@pytest.mark.gpu
@pytest.mark.parametrize("name", CASES)
def test_sorted_tally_equals_oracle(backend, oracle_portable, name):
    from gpu_optical_photons_amd.params import TallyConfig
    p = configs.PARITY[name]
    n = 300_000
    c = TallyConfig(256, 100.0, 64) if name == "c2_long" else TallyConfig()
    ref = oracle_portable.tally(p, n, c.time_bins, c.time_max, c.xy_bins)
    want = flat_ref(backend, ref, c)
    steps = ref["boundary_sum"] + ref["scatter_sum"] + int(ref["fate_counts"][1])
    for tuning in (None,
                   # tiny slot table: many iterations, the whole drain in the sorted kernels
                   dict(slots_per_block=256, blocks_per_cu=1, iterations_per_poll=2,
                        tail_threshold=1),
                   # the tail kernel takes over as soon as every photon is emitted
                   dict(slots_per_block=512, blocks_per_cu=3, iterations_per_poll=4,
                        tail_threshold=0xFFFFFFFF),
                   dict(slots_per_block=4096, blocks_per_cu=2, iterations_per_poll=16)):
        t, tm, st = backend.run_gpu_tally_sorted(p, c, n, tuning=tuning)
        np.testing.assert_array_equal(t.flat(), want, err_msg=f"{name} {tuning}")
        assert st.emitted == n
        assert st.classified <= steps
        if st.tail_photons == 0:
            assert st.classified == steps
        assert st.fresnel <= st.classified and st.rotated <= st.classified


@pytest.mark.gpu
def test_sorted_edge_cases(backend, oracle_portable):
    from gpu_optical_photons_amd.params import SimParams, TallyConfig
    c = TallyConfig()
    base = configs.PARITY["every_process"]
    small = dict(slots_per_block=256, blocks_per_cu=1, iterations_per_poll=2)
    for n, first_id, p in ((0, 0, base), (1, 0, base), (7, 5, base), (1000, 123456789, base),
                           (250_001, 3, base),
                           (5000, 0, configs.PARITY["capped"]),
                           (5000, 0, SimParams(**{**base.as_dict(), "max_steps": 0})),
                           (5000, 0, SimParams(**{**base.as_dict(), "max_steps": 1}))):
        ref = oracle_portable.tally(p, n, c.time_bins, c.time_max, c.xy_bins, first_id=first_id)
        for tuning in (None, small):
            t, _, st = backend.run_gpu_tally_sorted(p, c, n, first_id=first_id, tuning=tuning)
            np.testing.assert_array_equal(t.flat(), flat_ref(backend, ref, c),
                                          err_msg=f"n={n} first_id={first_id} {tuning}")
            assert st.emitted == n


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["c2_long", "wrapped_lambertian"])
def test_sorted_equals_fused_at_scale(backend, name):
    """At 2e7 photons (too many for the oracle in a test) the sorted path must
    give the fused wavefront kernel's tally word for word."""
    from gpu_optical_photons_amd.params import TallyConfig
    p = configs.PARITY[name]
    c = TallyConfig(256, 100.0, 64)
    n = 20_000_000
    fused, _, _, _ = backend.run_gpu_tally_wavefront(p, c, n, counters=False)
    t, _, st = backend.run_gpu_tally_sorted(p, c, n)
    np.testing.assert_array_equal(t.flat(), fused.flat())
    assert st.emitted == n and st.iterations >= 2


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["c2_long", "wrapped_lambertian", "every_process", "mirror_trap"])
def test_phase_kernel_equals_ring_kernel_at_scale(backend, name):
    """The fused kernel with LDS-sorted branch phases against the plain ring
    kernel at 2e7 photons, and at a block size whose waves hold few items."""
    from gpu_optical_photons_amd.params import TallyConfig
    p = configs.PARITY[name]
    c = TallyConfig(256, 100.0, 64)
    n = 20_000_000
    ring, _, _, _ = backend.run_gpu_tally_wavefront(
        p, c, n, counters=False,
        tuning=dict(block_size=512, steps_per_refill=2, queue_chunk=512, stream_in_lds=1))
    for tuning in (dict(block_size=512, steps_per_refill=2, queue_chunk=512, sorted_phases=1),
                   dict(block_size=128, steps_per_refill=4, queue_chunk=128, sorted_phases=1)):
        t, _, st, cnt = backend.run_gpu_tally_wavefront(p, c, n, tuning=tuning)
        np.testing.assert_array_equal(t.flat(), ring.flat(), err_msg=str(tuning))
        assert cnt.as_dict()["emitted"] == n
