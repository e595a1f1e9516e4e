"""optphot_compare over record directories in the .npy + meta.json layout:
photon-by-photon agreement counts on constructed differences, and its
statistics against scipy (chi-square survival function, two-sample KS)."""
import json
import os
import subprocess

import numpy as np
import pytest
from scipy import stats

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
APP = os.path.join(ROOT, "gpu_optical_photons_amd", "bin", "optphot_compare")
DTYPES = {"fate": np.uint8, "n_boundary": np.uint32, "n_scatter": np.uint32, "t": np.float32,
          "x": np.float32, "y": np.float32, "z": np.float32, "path": np.float32}


def _write(dir_, rec, size=(50.0, 40.0, 30.0)):
    os.makedirs(dir_, exist_ok=True)
    for k, dt in DTYPES.items():
        np.save(os.path.join(dir_, f"{k}.npy"), np.ascontiguousarray(rec[k], dtype=dt))
    with open(os.path.join(dir_, "meta.json"), "w") as f:
        f.write('{\n  "size_x": %g, "size_y": %g, "size_z": %g,\n  "seed": 1\n}\n' % size)


def _records(n, seed):
    rng = np.random.default_rng(seed)
    return {"fate": rng.choice([0, 1, 2], n, p=[0.4, 0.3, 0.3]).astype(np.uint8),
            "n_boundary": rng.integers(0, 9, n).astype(np.uint32),
            "n_scatter": rng.integers(0, 3, n).astype(np.uint32),
            "t": rng.exponential(5.0, n).astype(np.float32),
            "x": rng.uniform(-25, 25, n).astype(np.float32),
            "y": rng.uniform(-20, 20, n).astype(np.float32),
            "z": np.full(n, 15.0, np.float32),
            "path": rng.uniform(0, 300, n).astype(np.float32)}


def _run(tmp_path, a, b, *extra):
    if not os.path.exists(APP):
        pytest.fail(f"{APP} is not built")
    out = tmp_path / "cmp.json"
    r = subprocess.run([APP, str(a), str(b), "--out", str(out), *extra], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    return json.loads(out.read_text()), r.stdout


def test_identical_and_perturbed_records(tmp_path):
    n = 5000
    a = _records(n, 1)
    _write(tmp_path / "a", a)
    _write(tmp_path / "b", a)
    res, out = _run(tmp_path, tmp_path / "a", tmp_path / "b")
    assert res["history"]["bitwise_identical"] == n and res["history"]["same_discrete"] == n
    assert "bitwise identical records : 5000 / 5000" in out
    b = {k: v.copy() for k, v in a.items()}
    b["t"][10] = np.nextafter(b["t"][10], np.float32(np.inf))  # one ulp: same discrete history
    b["fate"][20] = (b["fate"][20] + 1) % 3                     # a different fate
    b["n_scatter"][30] += 1                                      # a different step count
    _write(tmp_path / "b", b)
    res, _ = _run(tmp_path, tmp_path / "a", tmp_path / "b")
    h = res["history"]
    assert h["same_discrete"] == n - 2 and h["bitwise_identical"] == n - 3
    assert h["fate_differs"] == 1 and h["first_divergent"] == 20
    assert h["max_abs_dt"] == pytest.approx(float(b["t"][10]) - float(a["t"][10]))


def test_statistics_match_scipy(tmp_path):
    a, b = _records(20000, 2), _records(15000, 3)
    b["t"] *= np.float32(1.02)  # a small shift in arrival times
    _write(tmp_path / "a", a)
    _write(tmp_path / "b", b)
    res, out = _run(tmp_path, tmp_path / "a", tmp_path / "b", "--time_bins", "40", "--t_max",
                    "30", "--xy_bins", "8")
    assert "skipped" in out  # different photon counts: no per-photon comparison
    d = res["distributions"]
    ta, tb = a["t"][a["fate"] == 0].astype(np.float64), b["t"][b["fate"] == 0].astype(np.float64)
    assert d["det_a"] == len(ta) and d["det_b"] == len(tb)
    ks = stats.ks_2samp(ta, tb, method="asymp")
    assert d["time_ks"]["statistic"] == pytest.approx(ks.statistic, rel=1e-12)
    assert d["time_ks"]["pvalue"] == pytest.approx(ks.pvalue, rel=0.05, abs=1e-4)
    ha = np.histogram(np.minimum(ta / 30.0 * 40, 40).astype(int), bins=np.arange(42))[0]
    hb = np.histogram(np.minimum(tb / 30.0 * 40, 40).astype(int), bins=np.arange(42))[0]
    A, B = ha.sum(), hb.sum()
    m = (ha + hb) > 0
    chi2 = (((np.sqrt(B / A) * ha - np.sqrt(A / B) * hb) ** 2)[m] / (ha + hb)[m]).sum()
    assert d["time_chi2"]["statistic"] == pytest.approx(chi2, rel=1e-9)
    assert d["time_chi2"]["ndf"] == m.sum() - 1
    assert d["time_chi2"]["pvalue"] == pytest.approx(stats.chi2.sf(chi2, m.sum() - 1), rel=1e-6,
                                                     abs=1e-12)
    ea, eb = len(ta) / 20000, len(tb) / 15000
    z = (ea - eb) / np.sqrt(ea * (1 - ea) / 20000 + eb * (1 - eb) / 15000)
    assert d["eff_z"] == pytest.approx(z, rel=1e-9)


def test_rejects_malformed_input(tmp_path):
    a = _records(100, 4)
    _write(tmp_path / "a", a)
    np.save(tmp_path / "a" / "t.npy", a["t"].astype(np.float64))  # wrong dtype
    r = subprocess.run([APP, str(tmp_path / "a"), str(tmp_path / "a")], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 1 and "expected dtype <f4" in r.stderr
