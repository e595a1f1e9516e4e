"""Batched differentiable calibration at full size: 256 parameter sets x 1e6
photons, long-histories geometry (100 x 100 x 50 mm, Lambertian walls,
Rayleigh 300 mm, volume source), each fitting (abs_length, R, n_det) by 50
Adam iterations to its own target histogram (hidden parameters, held-out
seed), one captured graph (forward, chi-square weights, adjoint) per
iteration.

Checks: the first iteration's graph at 256 sets equals separate forward and
adjoint calls; R and n_det converge (90th percentile of the relative error
below 6% and 3%); and abs_length, which converges poorly, is limited by the
data: a fit continued to 300 iterations reaches the loss of the hidden
parameters while abs_length stays far off, and the loss Hessian in
normalised coordinates is nearly singular along abs_length."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def fit(backend):
    from gpu_optical_photons_amd import calibrate as cal
    print("[calibration] 256 sets x 1e6 photons x 50 iterations", flush=True)
    return cal.calibrate(n_sets=256, photons=1_000_000, iters=50, time_forward=False)


def test_first_step_graph_equals_separate_calls(fit):
    from gpu_optical_photons_amd import adjoint
    from gpu_optical_photons_amd import calibrate as cal
    from gpu_optical_photons_amd.batch import EventBatch
    n_sets, photons = 256, 1_000_000
    T, nt, bins, tmax = fit["targets"], fit["target_photons"], fit["bins"], fit["tmax"]
    x0 = np.full((n_sets, 3), 0.5)  # where every fit starts
    tpl = cal.c5_template(fit["seed"])
    batch = EventBatch([adjoint.with_theta(tpl, r)
                        for r in cal._theta_rows(cal.LO + (cal.HI - cal.LO) * x0, n_sets)],
                       np.full(n_sets, photons, np.uint64))
    hist, loss, grad, _ = adjoint.run_gpu_chi2_step(batch, T, nt, bins, tmax,
                                                    forward_tuning=fit["forward_tuning"])
    h, _ = adjoint.histograms(batch, bins, tmax)
    assert np.array_equal(hist, h)
    np.testing.assert_allclose(loss, adjoint.chi2_loss(h, T, photons, nt), rtol=1e-12)
    f, F = h.sum(1, keepdims=True), T.sum(1, keepdims=True)
    wk = 1.0 / (T + 1.0 / nt)
    wf = 1.0 / (F * (1.0 - F) + 1.0 / nt)
    g, _ = adjoint.run_gpu_adjoint(batch, 2.0 * (h - T) * wk + 2.0 * (f - F) * wf, bins, tmax)
    np.testing.assert_allclose(grad, g, rtol=1e-9, atol=1e-12 * np.abs(g).max())
    assert fit["losses"][0] == pytest.approx(float(np.median(loss)), rel=1e-12)


def test_reflectivity_and_n_det_converge(fit):
    p90 = fit["p90_rel_error"]
    med = fit["median_rel_error"]
    print(f"[calibration] median {med}, p90 {p90}; loss {fit['losses'][0]:.3e} -> "
          f"{fit['losses'][-1]:.3e}", flush=True)
    assert p90["reflectivity"] < 0.06
    assert p90["n_det"] < 0.03
    assert fit["losses"][-1] < 0.02 * fit["losses"][0]


def test_abs_length_error_is_a_data_limit(fit):
    """Run long enough, the fits reach the loss of the hidden parameters (the
    data's noise floor) with R and n_det tight but abs_length still far off;
    the loss is far flatter along abs_length than along R and n_det."""
    from gpu_optical_photons_amd import calibrate as cal
    d = cal.calibration_diagnostics(fit)
    print(f"[calibration] diagnostics {d}", flush=True)
    assert d["long_fit_loss_over_hidden_median"] < 1.02
    e = d["long_fit_median_rel_error"]
    assert e["abs_length"] > 0.15
    assert e["reflectivity"] < 0.02 and e["n_det"] < 0.005
    assert d["softest_direction_abs_component_median"] > 0.8
    lam = d["hessian_eigenvalues_median"]
    assert lam[0] < 0.01 * lam[2]
