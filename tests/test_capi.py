"""The C-ABI library without a GPU: it loads, exports every entry point the
header declares, and rejects invalid arguments before touching the device
(the reference's std::invalid_argument cases, gpu_event.cu:177-182)."""
import ctypes
import os
import re

import numpy as np
import pytest

from gpu_optical_photons_amd import backend
from gpu_optical_photons_amd.params import SimParams, TallyConfig

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_entry_points():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            text = open(os.path.join(ROOT, "include", fn)).read()
            names |= set(re.findall(r"\b(optphot_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    names = declared_entry_points()
    assert "optphot_run_gpu_records" in names and len(names) >= 8
    lib = ctypes.CDLL(backend.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_struct_layouts():
    assert ctypes.sizeof(SimParams) == 72
    assert ctypes.sizeof(TallyConfig) == 12
    assert SimParams.seed.offset == 64 and SimParams.max_steps.offset == 60


def test_invalid_block_size_rejected_before_any_gpu_call():
    p = SimParams()
    for bs in (0, 32, 96, 2048, -64):
        with pytest.raises(ValueError, match="block_size"):
            backend.run_gpu_records(p, 10, block_size=bs)
        with pytest.raises(ValueError, match="block_size"):
            backend.run_gpu_records_event(p, 10, 4, block_size=bs)
        with pytest.raises(ValueError, match="block_size"):
            backend.run_gpu_tally(p, TallyConfig(), 10, block_size=bs)


def test_invalid_steps_and_params_rejected():
    with pytest.raises(ValueError, match="steps_per_launch"):
        backend.run_gpu_records_event(SimParams(), 10, 0)
    with pytest.raises(ValueError, match="surface"):
        backend.run_gpu_records(SimParams(surface=7), 10)
    with pytest.raises(ValueError, match="half sizes"):
        backend.run_gpu_records(SimParams(hx=0.0), 10)


def test_no_device_reports_cleanly():
    """Without a device the library says so instead of crashing."""
    if backend.gpu_available():
        pytest.skip("a device is present")
    assert backend.gpu_device_name() == "none"
    with pytest.raises(RuntimeError, match="no HIP device"):
        backend.run_gpu_records(SimParams(), 10)


def test_adjoint_rejects_reflectivity_outside_unit_interval():
    """The reflector score is hits / R: sets with reflecting faces need R in
    (0, 1], checked before the device is touched."""
    from gpu_optical_photons_amd import adjoint
    from gpu_optical_photons_amd.batch import EventBatch
    for surface in (2, 3):
        for R in (0.0, -0.1, 1.5, float("nan")):
            b = EventBatch([SimParams(surface=surface, reflectivity=R)], np.array([10], np.uint64))
            with pytest.raises(ValueError, match="reflectivity"):
                adjoint.run_gpu_adjoint(b, np.ones((1, 9)), 8, 50.0)
            with pytest.raises(ValueError, match="reflectivity"):
                adjoint.run_gpu_chi2_step(b, np.full((1, 9), 0.01), 100.0, 8, 50.0)
    # polished and black faces never use R
    b = EventBatch([SimParams(surface=0, reflectivity=0.0)], np.array([10], np.uint64))
    if not backend.gpu_available():
        with pytest.raises(RuntimeError, match="no HIP device"):
            adjoint.run_gpu_adjoint(b, np.ones((1, 9)), 8, 50.0)


def test_records_tuned_rejects_bad_shapes_before_the_device():
    """optphot_run_gpu_records_tuned checks the kernel choice and the shape
    before it touches the device; a valid shape without a GPU reaches the
    device check."""
    for bad, what in ((dict(kernel=3), "kernel"), (dict(kernel=-1), "kernel"),
                      (dict(kernel=2, block_size=100), "block"),
                      (dict(kernel=2, blocks_per_cu=-1), "blocks_per_cu")):
        with pytest.raises(ValueError, match=what):
            backend.run_gpu_records_tuned(SimParams(), 10, bad)
    if not backend.gpu_available():
        with pytest.raises(RuntimeError, match="no HIP device"):
            backend.run_gpu_records_tuned(SimParams(), 10, None)
