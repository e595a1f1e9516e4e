"""The drop-in check: the reference's own, unmodified apps
(apps/optphot_main.cpp, optphot_compare.cpp, optphot_event_bench.cpp) and CPU
backend (portable math), with the reference's GPU backend replaced by
oracle/ref_gpu_binding.cpp over liboptphot_hip.so. Built by
`make -C oracle ref-apps` where the reference checkout is present; the
binaries travel with the tree. The reference's own comparison logic
(src/core/compare.cpp) reports the agreement.
"""
import csv
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
APPS = os.path.join(ROOT, "oracle", "_ref", "apps")

pytestmark = pytest.mark.gpu

# SimParams of the plain-small and long-histories workloads as reference options
C1 = ["--size_x", "50", "--size_y", "50", "--size_z", "50", "--n_det", "1.5",
      "--abs_length", "1000", "--surface", "polished", "--source", "point", "--seed", "1"]
C2 = ["--size_x", "100", "--size_y", "100", "--size_z", "50", "--n_det", "1.5",
      "--abs_length", "2000", "--scat_length", "300", "--surface", "lambertian",
      "--reflectivity", "0.97", "--source", "volume", "--seed", "1"]
OTHERS = {
    "specular_scatter": ["--surface", "specular", "--reflectivity", "0.9", "--scat_length", "80",
                         "--n_det", "1.5", "--seed", "7"],
    "black_offcentre": ["--surface", "black", "--src_x", "10", "--src_y", "-7", "--src_z", "3",
                        "--seed", "11"],
    "capped": ["--surface", "lambertian", "--reflectivity", "0.99", "--scat_length", "30",
               "--max_steps", "5", "--seed", "5"],
}


def _app(name):
    path = os.path.join(APPS, name)
    if not os.path.exists(path):
        pytest.skip(f"{path} not built (needs the reference checkout at build time)")
    return path


def _run(args, cwd, timeout=300):
    r = subprocess.run(args, cwd=cwd, capture_output=True, text=True, timeout=timeout)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


@pytest.mark.parametrize("name,opts,n", [("c1_plain", C1, 1_000_000), ("c2_long", C2, 200_000)]
                         + [(k, v, 200_000) for k, v in OTHERS.items()])
def test_reference_compare_app_bitwise(tmp_path, name, opts, n):
    out = _run([_app("optphot_compare"), *opts, "--n_photons", str(n)], tmp_path)
    m = re.search(r"bitwise identical records : (\d+) / (\d+)", out)
    assert m, out
    assert int(m.group(1)) == int(m.group(2)) == n, out


def test_reference_event_bench_app(tmp_path):
    out_csv = tmp_path / "event.csv"
    r = subprocess.run([_app("optphot_event_bench"), *C2, "--n_photons", "200000", "--k",
                        "1,3,16,10000", "--repeat", "1", "--csv", str(out_csv)],
                       cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and "differs" not in r.stderr, r.stdout + r.stderr
    rows = list(csv.DictReader(open(out_csv)))
    assert len(rows) == 5
    for row in rows:
        assert int(row["bitwise_identical"]) == 200000, row


@pytest.mark.parametrize("mode", ["tally", "records"])
def test_reference_main_app_gpu_equals_cpu(tmp_path, mode):
    app = _app("optphot")
    files = {"tally": ["fate_counts", "time_hist", "xy_hist"],
             "records": ["fate", "n_boundary", "n_scatter", "t", "x", "y", "z", "path"]}[mode]
    common = [*C2, "--n_photons", "100000", "--mode", mode, "--time_bins", "256", "--time_max",
              "100", "--xy_bins", "64"]
    _run([app, *common, "--backend", "cpu", "--threads", "0", "--out", str(tmp_path / "cpu")],
         tmp_path)
    _run([app, *common, "--backend", "gpu", "--out", str(tmp_path / "gpu")], tmp_path)
    for f in files:
        a = np.load(tmp_path / "cpu" / f"{f}.npy")
        b = np.load(tmp_path / "gpu" / f"{f}.npy")
        assert a.dtype == b.dtype and a.shape == b.shape, f
        assert np.array_equal(a.view(np.uint8), b.view(np.uint8)), f


def test_compare_tool_on_reference_cpu_and_library_gpu_runs(tmp_path):
    """The reference's optphot with its CPU backend and this library's optphot
    on the GPU write records of the same photons; this library's
    optphot_compare reads both directories and finds them bitwise identical."""
    import json
    ref_app = _app("optphot")
    own = os.path.join(ROOT, "gpu_optical_photons_amd", "bin")
    common = [*C2, "--n_photons", "200000", "--mode", "records"]
    _run([ref_app, *common, "--backend", "cpu", "--threads", "0", "--out",
          str(tmp_path / "cpu")], tmp_path)
    _run([os.path.join(own, "optphot"), *common, "--kernel", "wavefront", "--out",
          str(tmp_path / "gpu")], tmp_path)
    _run([os.path.join(own, "optphot_compare"), str(tmp_path / "cpu"), str(tmp_path / "gpu"),
          "--out", str(tmp_path / "cmp.json")], tmp_path)
    h = json.loads((tmp_path / "cmp.json").read_text())["history"]
    assert h["bitwise_identical"] == h["n"] == 200000
