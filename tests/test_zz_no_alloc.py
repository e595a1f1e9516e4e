"""The entry point that allocates nothing
(optphot_run_gpu_tally_wavefront_noalloc) against the allocating wavefront
tally (optphot_run_gpu_tally_wavefront) through the C-ABI: word for word on
two calls in a row, the second smaller, with the autotuned long-histories
shape (LDS-ring kernel) and with the default shape; a binning whose partial
rows exceed the static scratch is rejected."""
import numpy as np
import pytest

from gpu_optical_photons_amd.params import SimParams, Source, Surface, TallyConfig

pytestmark = pytest.mark.gpu


def _long_histories():
    return SimParams(hx=50.0, hy=50.0, hz=25.0, n_scint=1.63, n_det=1.5, n_out=1.0,
                     abs_length=2000.0, scat_length=300.0, surface=Surface.LAMBERTIAN,
                     reflectivity=0.97, source=Source.VOLUME, tau=4.0, max_steps=10000, seed=1)


def _noalloc(backend, p, c, n, first_id, tuning):
    import torch
    out = torch.full((c.words,), -1, dtype=torch.int64, device="cuda:0")  # stale contents
    backend.run_gpu_tally_wavefront_noalloc(p, c, n, out, first_id=first_id, tuning=tuning)
    return out.cpu().numpy().view(np.uint64)


def test_noalloc_equals_allocating_entry_point(backend):
    from gpu_optical_photons_amd import autotune
    print("[noalloc] device", backend.gpu_device_name(), flush=True)  # sets up the device
    p = _long_histories()
    c = TallyConfig(256, 100.0, 64)
    ring = autotune.choice_for("wavefront_tally", p, c)
    cases = [  # the second call of each pair is the smaller one
        (p, c, 3_000_000, 0, ring),
        (p, c, 200_003, 17, ring),
        (p.copy(surface=Surface.SPECULAR, reflectivity=0.9), TallyConfig(100, 50.0, 50),
         1_000_000, 5, None),
        (p.copy(surface=Surface.SPECULAR, reflectivity=0.9), TallyConfig(100, 50.0, 50),
         4_097, 0, None),
    ]
    for q, cfg, n, fid, tu in cases:
        print(f"[noalloc] n={n} first_id={fid} tuning={tu}", flush=True)
        ref, _, _, _ = backend.run_gpu_tally_wavefront(q, cfg, n, first_id=fid, tuning=tu,
                                                       counters=False)
        got = _noalloc(backend, q, cfg, n, fid, tu)
        assert np.array_equal(got, ref.flat()), (n, fid)
        assert int(got[:5].sum()) == n


def test_noalloc_rejects_binning_beyond_scratch(backend):
    import torch
    backend.gpu_device_name()
    c = TallyConfig(256, 100.0, 128)  # 16 646 words per block row
    out = torch.zeros(c.words, dtype=torch.int64, device="cuda:0")
    with pytest.raises(ValueError, match="static maximum"):
        backend.run_gpu_tally_wavefront_noalloc(_long_histories(), c, 10_000_000, out)


def test_noalloc_rejects_host_output(backend):
    import torch
    backend.gpu_device_name()
    c = TallyConfig()
    out = torch.zeros(c.words, dtype=torch.int64)
    with pytest.raises(ValueError, match="device tensor"):
        backend.run_gpu_tally_wavefront_noalloc(SimParams(), c, 1000, out)
