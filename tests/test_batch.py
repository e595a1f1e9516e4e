"""Batched events (one persistent, graph-captured launch over many events)
against the plain per-event history launches and against the oracle, event by
event: event e with photon ids offset_e.. is the reference's
run_cpu_tally(params[e], N_e, first_id = offset_e)."""
import math

import numpy as np
import pytest

from gpu_optical_photons_amd.params import SimParams

TB, TMAX = 128, 50.0


def _oracle_row(oracle, p, n, first):
    t = oracle.tally(p, int(n), TB, TMAX, 1, first_id=int(first))
    return np.concatenate([t["fate_counts"], t["time_hist"]]).astype(np.int64)


def test_make_events_is_deterministic_and_in_range():
    from gpu_optical_photons_amd.batch import make_events
    a, b = make_events(64, seed=7), make_events(64, seed=7)
    assert np.array_equal(a.photons, b.photons)
    for p, q in zip(a.params, b.params):
        assert p.as_dict() == q.as_dict()
        assert 5.0 <= p.hx <= 50.0 and 1.4 <= p.n_det <= 1.6 and 300 <= p.abs_length <= 3000
        assert math.isinf(p.scat_length) or 100 <= p.scat_length <= 1000
        assert abs(p.src_x) <= p.hx and abs(p.src_z) <= p.hz
    assert {p.surface for p in make_events(400).params} == {0, 1, 2, 3}


def test_batch_rejects_mixed_seeds_without_gpu():
    from gpu_optical_photons_amd.batch import EventBatch, run_gpu_batch
    b = EventBatch([SimParams(seed=1), SimParams(seed=2)], np.array([10, 10], np.uint64))
    with pytest.raises(ValueError, match="run seed"):
        run_gpu_batch(b)
    with pytest.raises(ValueError, match="time_bins"):
        run_gpu_batch(EventBatch([SimParams()], np.array([1], np.uint64)), time_bins=0)


def test_batch_signature_classes():
    from gpu_optical_photons_amd import autotune, calibrate
    from gpu_optical_photons_amd.batch import make_events
    c3 = autotune.batch_signature(make_events(4096, seed=2024), 128)
    assert c3 == "batch|events=many|photons=<1e5|surfaces=0123|scat=mixed|rows=global"
    c5 = autotune.batch_signature(calibrate.c5_batch(n_sets=256, photons=10), 128)
    assert c5 == "batch|events=mid|photons=<1e4|surfaces=3|scat=on|rows=global"
    few = autotune.batch_signature(make_events(8, seed=12, photons_per_mev=3000), 128)
    assert few.endswith("|rows=lds") and "|events=few|" in few
    tu = autotune.batch_choice_for(make_events(4096, seed=2024), 128)
    assert tu is None or set(tu) == {"block_size", "steps_per_refill", "queue_chunk",
                                      "stream_in_lds"}


@pytest.mark.gpu
def test_small_batch_matches_oracle_event_by_event(backend, oracle_portable):
    from gpu_optical_photons_amd.batch import PLAIN, WAVEFRONT, make_events, run_gpu_batch
    b = make_events(96, seed=11, photons_per_mev=3000)
    b.photons[5] = 0  # an empty event
    w, _, _ = run_gpu_batch(b, TB, TMAX, method=WAVEFRONT)
    p, _, _ = run_gpu_batch(b, TB, TMAX, method=PLAIN)
    assert np.array_equal(w, p)
    r, _, _ = run_gpu_batch(b, TB, TMAX, method=WAVEFRONT, tuning=dict(stream_in_lds=1))
    assert np.array_equal(w, r)
    # few events: per-block event tallies in LDS (with and without ring streams)
    few = make_events(8, seed=12, photons_per_mev=3000)
    fw, _, _ = run_gpu_batch(few, TB, TMAX, method=WAVEFRONT)
    fr, _, _ = run_gpu_batch(few, TB, TMAX, method=WAVEFRONT, tuning=dict(stream_in_lds=1))
    fp, _, _ = run_gpu_batch(few, TB, TMAX, method=PLAIN)
    assert np.array_equal(fw, fp) and np.array_equal(fr, fp)
    off = b.offsets
    for e in range(b.n_events):
        assert np.array_equal(w[e].astype(np.int64), _oracle_row(oracle_portable, b.params[e],
                                                                  b.photons[e], off[e])), e
    assert w[:, :5].sum() == b.total
    # every event empty: an empty queue
    b.photons[:] = 0
    z, _, _ = run_gpu_batch(b, TB, TMAX, method=WAVEFRONT)
    assert not z.any()


@pytest.mark.gpu
def test_full_batch_graph_replays_and_plain(backend, oracle_portable):
    from gpu_optical_photons_amd.batch import PLAIN, WAVEFRONT, make_events, run_gpu_batch
    b = make_events(4096, seed=2024)
    print(f"[batch] 4096 events, {b.total} photons", flush=True)
    plain, tp, _ = run_gpu_batch(b, TB, TMAX, method=PLAIN)
    assert plain[:, :5].sum() == b.total
    for rep in range(3):  # captured graph: first call captures, later calls replay
        w, tw, _ = run_gpu_batch(b, TB, TMAX, method=WAVEFRONT)
        assert np.array_equal(w, plain), f"replay {rep}"
        backend.run_gpu_records(SimParams(), 1000)  # eager work between replays
    rng = np.random.default_rng(0)
    off = b.offsets
    for e in rng.choice(b.n_events, 24, replace=False):
        assert np.array_equal(plain[e].astype(np.int64),
                              _oracle_row(oracle_portable, b.params[e], b.photons[e], off[e])), e
    print(f"[batch] plain {tp.kernel_ms:.2f} ms, wavefront {tw.kernel_ms:.2f} ms", flush=True)


@pytest.mark.gpu
def test_batch_first_id_and_tuning_invariance(backend):
    from gpu_optical_photons_amd.batch import WAVEFRONT, make_events, run_gpu_batch
    b = make_events(200, seed=3, photons_per_mev=2000)
    ref, _, _ = run_gpu_batch(b, TB, TMAX, first_id=12345)
    for tu in (dict(block_size=64, steps_per_refill=1, queue_chunk=64),
               dict(block_size=1024, steps_per_refill=8, queue_chunk=2048),
               dict(block_size=512, steps_per_refill=2, queue_chunk=256, stream_in_lds=1),
               dict(block_size=64, steps_per_refill=3, queue_chunk=64, stream_in_lds=1)):
        w, _, _ = run_gpu_batch(b, TB, TMAX, first_id=12345, method=WAVEFRONT, tuning=tu)
        assert np.array_equal(w, ref)
    other, _, _ = run_gpu_batch(b, TB, TMAX, first_id=0)
    assert not np.array_equal(other, ref)


@pytest.mark.gpu
def test_batch_wide_time_rows(backend):
    """4096-bin time rows: at block size 64 each wave's LDS row (16 KB) fits the
    LDS budget, at 1024 the 16 rows of a block do not and every photon adds to
    the global rows; both equal the per-event launches."""
    from gpu_optical_photons_amd.batch import PLAIN, WAVEFRONT, make_events, run_gpu_batch
    b = make_events(300, seed=5, photons_per_mev=2000)
    plain, _, _ = run_gpu_batch(b, 4096, TMAX, method=PLAIN)
    for tu in (dict(block_size=64, steps_per_refill=2, queue_chunk=128),
               dict(block_size=1024, steps_per_refill=2, queue_chunk=128),
               dict(block_size=64, steps_per_refill=2, queue_chunk=128, stream_in_lds=1)):
        w, _, _ = run_gpu_batch(b, 4096, TMAX, method=WAVEFRONT, tuning=tu)
        assert np.array_equal(w, plain), tu
