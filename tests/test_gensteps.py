"""Photons from gensteps (optphot_run_gpu_gensteps; the reference lists
generation from gensteps as future work, README.md:734-735, so there is no
reference kernel to restate).

How the emission is pinned:
* a degenerate genstep (x0 = x1, t0 = t1 = 0) draws and computes what the
  reference's point source does: its photons equal the reference's own CPU
  path (oracle/_ref, portable math) record for record;
* along a segment, with absorption so short that photons stop where they are
  emitted, the emission point's fraction along the segment is U[0, 1]
  (Kolmogorov-Smirnov) and its time is t0 + u (t1 - t0) plus an Exp(tau)
  delay (mean and KS);
* global photon indices: re-running a tail of the genstep list with
  first_id = its offset reproduces the tail of the full run;
* GPU: the kernel equals the oracle bit for bit (tally and records), for any
  launch shape.
"""
import math

import numpy as np
import pytest
from scipy import stats

from gpu_optical_photons_amd.gensteps import GENSTEP_DTYPE, make_gensteps
from gpu_optical_photons_amd.params import SimParams
from tests.configs import PARITY

FIELDS = ("fate", "n_boundary", "n_scatter", "t", "x", "y", "z", "path")


def _point(p, n):
    g = np.zeros(1, GENSTEP_DTYPE)
    g["x0"] = g["x1"] = p.src_x
    g["y0"] = g["y1"] = p.src_y
    g["z0"] = g["z1"] = p.src_z
    g["n_photons"] = n
    return g


@pytest.mark.parametrize("name", ["c1_plain", "every_process", "specular"])
def test_degenerate_genstep_is_the_reference_point_source(oracle_portable, name):
    from oracle.oracle import Oracle, Reference, reference_available
    p = PARITY[name].copy(source=0)
    ref_lib = Reference("portable") if reference_available("portable") else oracle_portable
    n = 20_000
    _, r = oracle_portable.gensteps(p, _point(p, n), first_id=77, records=True)
    ref = ref_lib.records(p, n, first_id=77)
    for f in FIELDS:
        assert np.array_equal(r[f].view(np.uint8), ref[f].view(np.uint8)), f


def test_emission_along_segment(oracle_portable):
    p = SimParams(hx=50.0, hy=50.0, hz=50.0, abs_length=1e-6, tau=3.0, seed=3)
    g = np.zeros(1, GENSTEP_DTYPE)
    g["x0"], g["y0"], g["z0"], g["t0"] = -20.0, 5.0, -10.0, 1.0
    g["x1"], g["y1"], g["z1"], g["t1"] = 30.0, -15.0, 25.0, 1.5
    n = 100_000
    g["n_photons"] = n
    _, r = oracle_portable.gensteps(p, g, records=True)
    assert np.all(r["fate"] == 1)  # absorbed in the bulk, within ~1e-6 mm of emission
    a = np.array([-20.0, 5.0, -10.0])
    b = np.array([30.0, -15.0, 25.0])
    pos = np.stack([r["x"], r["y"], r["z"]], axis=1).astype(np.float64)
    u = (pos - a) @ (b - a) / ((b - a) @ (b - a))
    assert np.abs(np.linalg.norm((pos - a) - np.outer(u, b - a), axis=1)).max() < 1e-3
    assert stats.kstest(u, "uniform").pvalue > 0.01
    delay = r["t"].astype(np.float64) - (1.0 + u * 0.5) - r["path"] * 1.63 / 299.792458
    assert abs(delay.mean() - 3.0) < 5 * 3.0 / np.sqrt(n)
    assert stats.kstest(delay, "expon", args=(0.0, 3.0)).pvalue > 0.01


def test_global_index_and_first_id(oracle_portable):
    p = PARITY["c2_long"]
    g = make_gensteps(60, p, photons_per_mm=50.0, seed=4)
    g["n_photons"][3] = 0  # an empty genstep
    _, full = oracle_portable.gensteps(p, g, first_id=1000, records=True)
    off = int(g["n_photons"][:20].sum())
    _, tail = oracle_portable.gensteps(p, g[20:], first_id=1000 + off, records=True)
    for f in FIELDS:
        assert np.array_equal(tail[f], full[f][off:]), f


def test_gensteps_outside_the_box_rejected_before_gpu():
    from gpu_optical_photons_amd import backend
    from gpu_optical_photons_amd.gensteps import run_gpu_gensteps
    from gpu_optical_photons_amd.params import TallyConfig
    p = SimParams(hx=10.0, hy=10.0, hz=10.0)
    g = np.zeros(1, GENSTEP_DTYPE)
    g["x1"] = 11.0
    g["n_photons"] = 10
    with pytest.raises(ValueError, match="outside the box"):
        run_gpu_gensteps(p, TallyConfig(), g)
    g["x1"] = 0.0
    g["t0"] = -1.0
    with pytest.raises(ValueError, match="times"):
        run_gpu_gensteps(p, TallyConfig(), g)
    assert backend is not None


def test_cherenkov_cone_geometry(oracle_portable):
    """Cherenkov gensteps against the geometry of the cone. A short segment
    on the z axis, no absorption or scattering, black side walls and an
    index-matched readout face: every photon flies straight to z = hz and is
    detected there. Its radial offset over its path is then sin theta of the
    cone, its azimuth is uniform, and hz minus the path times cos theta gives
    the emission point, uniform along the segment, at the time the segment
    assigns it."""
    cone = 0.95
    p = SimParams(hx=50.0, hy=50.0, hz=50.0, n_scint=1.6, n_det=1.6, abs_length=math.inf,
                  scat_length=math.inf, surface=1, tau=4.0, seed=21)
    g = np.zeros(1, GENSTEP_DTYPE)
    g["x0"], g["y0"], g["z0"], g["t0"] = 0.0, 0.0, -40.0, 1.0
    g["x1"], g["y1"], g["z1"], g["t1"] = 0.0, 0.0, -38.0, 1.2
    g["cone_cos"] = cone
    n = 50_000
    g["n_photons"] = n
    _, r = oracle_portable.gensteps(p, g, records=True)
    assert np.all(r["fate"] == 0) and np.all(r["n_boundary"] == 1)
    x, y, path = (r[k].astype(np.float64) for k in ("x", "y", "path"))
    sin_t = math.sqrt(1.0 - cone * cone)
    assert np.abs(np.hypot(x, y) / path - sin_t).max() < 1e-5
    assert stats.kstest((np.arctan2(y, x) / (2 * np.pi)) % 1.0, "uniform").pvalue > 0.01
    z_emit = 50.0 - path * cone
    u = (z_emit + 40.0) / 2.0
    assert u.min() > -1e-4 and u.max() < 1.0 + 1e-4
    assert stats.kstest(u, "uniform").pvalue > 0.01
    t_emit = r["t"].astype(np.float64) - path * 1.6 / 299.792458
    assert np.abs(t_emit - (1.0 + 0.2 * u)).max() < 1e-4  # no emission delay


@pytest.mark.gpu
def test_gpu_cherenkov_edge_cases_equal_oracle(backend, oracle_portable):
    """Cherenkov gensteps at the edges of their range, GPU against the oracle
    bit for bit: cone_cos = 1 (photons along the segment), a cone of nearly
    90 degrees, the smallest cone_cos, a segment along each axis and one a
    few 1e-15 mm long, and scintillation gensteps in between."""
    from gpu_optical_photons_amd.gensteps import run_gpu_gensteps
    from gpu_optical_photons_amd.params import TallyConfig
    p = PARITY["every_process"]
    cases = [((0, 0, -5), (0, 0, 5), 1.0), ((1, 2, 3), (4, -2, 0), 1e-3),
             ((0, 0, 0), (1, 1, 1), float(np.float32(1e-38))), ((-3, 0, 0), (3, 0, 0), 0.6),
             ((0, -3, 0), (0, 3, 0), 0.9), ((1e-15, 0, 0), (3e-15, 0, 0), 0.7),
             ((0, 0, 0), (2, 0, 1), 0.0), ((2, 2, 2), (2, 2, 2), 0.0)]
    g = np.zeros(len(cases), GENSTEP_DTYPE)
    for k, (a, b, cc) in enumerate(cases):
        g[k]["x0"], g[k]["y0"], g[k]["z0"] = a
        g[k]["x1"], g[k]["y1"], g[k]["z1"] = b
        g[k]["t0"], g[k]["t1"] = 0.5 * k, 0.5 * k + 0.1
        g[k]["cone_cos"] = cc
        g[k]["n_photons"] = 3000
    c = TallyConfig(128, 60.0, 32)
    ref_t, ref_r = oracle_portable.gensteps(p, g, first_id=11, time_bins=128, time_max=60.0,
                                            xy_bins=32, records=True)
    t, r, _, _ = run_gpu_gensteps(p, c, g, first_id=11, records=True)
    for f in FIELDS:
        assert np.array_equal(getattr(r, f).view(np.uint8), ref_r[f].view(np.uint8)), f
    assert np.array_equal(t.fate_counts, ref_t["fate_counts"])
    assert np.array_equal(t.xy_hist, ref_t["xy_hist"])


def test_cherenkov_genstep_checks_before_gpu():
    from gpu_optical_photons_amd.gensteps import run_gpu_gensteps
    from gpu_optical_photons_amd.params import TallyConfig
    p = SimParams(hx=10.0, hy=10.0, hz=10.0)
    g = np.zeros(1, GENSTEP_DTYPE)
    g["z1"] = 1.0
    g["n_photons"] = 10
    g["cone_cos"] = 1.5
    with pytest.raises(ValueError, match="cone_cos"):
        run_gpu_gensteps(p, TallyConfig(), g)
    g["cone_cos"] = float("nan")
    with pytest.raises(ValueError, match="cone_cos"):
        run_gpu_gensteps(p, TallyConfig(), g)
    g["cone_cos"] = 0.7
    g["z1"] = 0.0  # no direction
    with pytest.raises(ValueError, match="direction"):
        run_gpu_gensteps(p, TallyConfig(), g)


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["c2_long", "every_process", "c1_plain"])
def test_gpu_gensteps_equal_oracle(backend, oracle_portable, name):
    from gpu_optical_photons_amd.gensteps import run_gpu_gensteps
    from gpu_optical_photons_amd.params import TallyConfig
    p = PARITY[name]
    # a third Cherenkov gensteps (cone about the segment, no delay)
    g = make_gensteps(400, p, photons_per_mm=200.0, seed=9, cherenkov_fraction=0.3)
    assert (g["cone_cos"] > 0).sum() > 50 and g[11]["cone_cos"] == 0.0
    g["n_photons"][7] = 0
    g[11]["x1"], g[11]["y1"], g[11]["z1"], g[11]["t1"] = g[11]["x0"], g[11]["y0"], g[11]["z0"], g[11]["t0"]
    c = TallyConfig(256, 100.0, 64)
    print(f"[gensteps] {name}: {len(g)} gensteps, {int(g['n_photons'].sum())} photons", flush=True)
    ref_t, ref_r = oracle_portable.gensteps(p, g, first_id=5, time_bins=256, time_max=100.0,
                                            xy_bins=64, records=True)
    for tu in (None, dict(block_size=64, steps_per_refill=1, queue_chunk=64),
               dict(block_size=1024, steps_per_refill=8, queue_chunk=1024)):
        t, r, tm, cnt = run_gpu_gensteps(p, c, g, first_id=5, records=True, tuning=tu,
                                         counters=True)
        assert np.array_equal(t.fate_counts, ref_t["fate_counts"])
        assert np.array_equal(t.time_hist, ref_t["time_hist"])
        assert np.array_equal(t.xy_hist, ref_t["xy_hist"])
        assert t.boundary_sum == ref_t["boundary_sum"] and t.scatter_sum == ref_t["scatter_sum"]
        for f in FIELDS:
            assert np.array_equal(getattr(r, f).view(np.uint8), ref_r[f].view(np.uint8)), f
        assert cnt.as_dict()["emitted"] == int(g["n_photons"].sum())
    # tally only, Philox windows in LDS rings
    for tu in (dict(block_size=512, steps_per_refill=2, queue_chunk=256, stream_in_lds=1),
               dict(block_size=64, steps_per_refill=1, queue_chunk=64, stream_in_lds=1),
               dict(block_size=256, steps_per_refill=5, queue_chunk=1024, stream_in_lds=1)):
        for counters in (True, False):
            t, r, tm, cnt = run_gpu_gensteps(p, c, g, first_id=5, records=False, tuning=tu,
                                             counters=counters)
            assert r is None
            assert np.array_equal(t.fate_counts, ref_t["fate_counts"]), tu
            assert np.array_equal(t.time_hist, ref_t["time_hist"]), tu
            assert np.array_equal(t.xy_hist, ref_t["xy_hist"]), tu
            assert t.boundary_sum == ref_t["boundary_sum"]
            assert t.scatter_sum == ref_t["scatter_sum"]
            if counters:
                assert cnt.as_dict()["emitted"] == int(g["n_photons"].sum())


@pytest.mark.gpu
def test_records_through_several_staging_pieces(backend, oracle_portable):
    """Records of ~7e5 photons leave the device in several staged pieces, more
    than there are staging slots, so slots are reused; then a smaller run
    reuses the buffers. Both equal the oracle's records byte for byte."""
    from gpu_optical_photons_amd.gensteps import run_gpu_gensteps
    from gpu_optical_photons_amd.params import TallyConfig
    p = PARITY["c2_long"]
    g = make_gensteps(700, p, photons_per_mm=200.0, seed=31, cherenkov_fraction=1.0 / 3.0)
    n = int(g["n_photons"].sum())
    assert n > 8 * (1 << 16)  # pieces of n / 8 photons at least: 8 pieces, 3 slots
    print(f"[gensteps] records of {n} photons", flush=True)
    c = TallyConfig(64, 100.0, 16)
    _, ref = oracle_portable.gensteps(p, g, first_id=5, time_bins=64, time_max=100.0,
                                      xy_bins=16, records=True)
    _, r, _, _ = run_gpu_gensteps(p, c, g, first_id=5, records=True)
    for f in FIELDS:
        assert np.array_equal(getattr(r, f).view(np.uint8), ref[f].view(np.uint8)), f
    m = int(g["n_photons"][:300].sum())
    _, r2, _, _ = run_gpu_gensteps(p, c, g[:300], first_id=5, records=True)
    for f in FIELDS:
        assert np.array_equal(getattr(r2, f).view(np.uint8), ref[f][:m].view(np.uint8)), f


@pytest.mark.gpu
def test_gpu_genstep_records_in_large_pieces(backend, oracle_portable):
    """Genstep records above 8 x 2^18 photons: the staged copy pieces then
    hold at least 2^18 photons each, which the host copies into the caller's
    arrays on several threads, each field split into byte ranges (an odd
    photon count makes the last range of the 1-byte fate field odd). Byte for
    byte against the oracle."""
    from gpu_optical_photons_amd.gensteps import run_gpu_gensteps
    from gpu_optical_photons_amd.params import TallyConfig
    p = PARITY["c2_long"]
    g = make_gensteps(3000, p, photons_per_mm=200.0, seed=21, cherenkov_fraction=0.3)
    g["n_photons"][0] += 1 - int(g["n_photons"].sum()) % 2  # odd total
    total = int(g["n_photons"].sum())
    assert total > 8 * (1 << 18) and total % 2 == 1
    print(f"[gensteps] large pieces: {total} photons", flush=True)
    c = TallyConfig(64, 100.0, 16)
    _, ref_r = oracle_portable.gensteps(p, g, first_id=3, time_bins=64, time_max=100.0,
                                        xy_bins=16, records=True)
    _, r, _, _ = run_gpu_gensteps(p, c, g, first_id=3, records=True)
    for f in FIELDS:
        assert np.array_equal(getattr(r, f).view(np.uint8), ref_r[f].view(np.uint8)), f
