"""Pins the Fresnel scores of the FP64 adjoint reference (oracle_adjoint) in
histories that meet the readout repeatedly at angles next to the critical
angle, against float64 quadrature.

Mirror box (specular walls R_w = 0.9), Fresnel readout n_scint 1.63 ->
n_det 1.5 (total internal reflection below the critical cosine 0.391), no bulk
processes, point source. The walls keep |d_z|, so every readout arrival of a
photon has the same cos_i; tests/analytic.py sums the arrivals exactly and
integrates over directions. dR/dn grows like 1/cos_t at the critical angle,
so hits within float32 rounding of it carry a finite share of dP/dn: the
score must be taken wherever the float64 geometry allows transmission, also
where the float32 history rounded the hit into total internal reflection.

The point source draws cos(theta) = 2u - 1 on a 2^-23 grid, so the derivative
of the simulated model at fixed indices is a Riemann sum over that grid;
finite differences over steps much wider than the grid see the smooth
derivative. Both values are in tests/golden/mirror_box_derivative.json
(tests/golden/make_mirror_box_derivative.py)."""
import json
import math
import os

import numpy as np
import pytest

from gpu_optical_photons_amd.params import SimParams
from tests import analytic

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                      "mirror_box_derivative.json")


def golden():
    with open(GOLDEN) as f:
        return json.load(f)


def mirror_box_fresnel(seed=21):
    g = golden()
    return SimParams(**g["params"], abs_length=math.inf, scat_length=math.inf, surface=2,
                     reflectivity=g["reflectivity"], tau=0.0, seed=seed)


def test_quadrature_derivative_equals_its_finite_difference():
    """The summed-arrival derivative against a central difference of P_det
    on one fixed direction grid (so the wall-count jumps cancel)."""
    p = mirror_box_fresnel()
    n1 = float(p.n_scint)
    h = 1e-3
    lo = math.sqrt(1.0 - (float(np.float32(1.5 + h)) / n1) ** 2)
    n_v, n_phi = 500, 500
    v = (np.arange(n_v) + 0.5) / n_v
    mu = lo + (1.0 - lo) * v * v
    jac = 2.0 * (1.0 - lo) * v / n_v

    def pdet(n_det):
        q = p.copy(n_det=n_det)
        c = math.sqrt(1.0 - (float(q.n_det) / n1) ** 2)
        above = mu > c
        F, _, _ = analytic._mirror_box_readout_density(q, 0.9, mu[above], n_phi)
        return float((jac[above] * F).sum()), float(q.n_det)

    (pp, np_), (pm, nm) = pdet(1.5 + h), pdet(1.5 - h)
    fd = (pp - pm) / (np_ - nm)
    _, d_ndet, _ = analytic.mirror_box_fresnel_readout(p, 0.9, n_v, n_phi)
    assert fd == pytest.approx(d_ndet, rel=2e-3)
    assert fd == pytest.approx(golden()["dP_dndet_smooth"], rel=2e-3)


def test_golden_reproduces_at_low_resolution():
    g = golden()
    pdet, d_ndet, d_nsc = analytic.mirror_box_fresnel_readout(mirror_box_fresnel(), 0.9, 600, 600)
    assert pdet == pytest.approx(g["P_det"], rel=2e-5)
    assert d_ndet == pytest.approx(g["dP_dndet_smooth"], rel=3e-4)
    assert d_nsc == pytest.approx(g["dP_dnscint_smooth"], rel=3e-4)
    # the emission-grid correction is a few tenths of a percent, opposite in
    # sign for the two indices (R depends on their ratio near the critical angle)
    assert 0.003 < g["grid_correction_ndet"] / d_ndet < 0.01
    assert 0.003 < g["grid_correction_nscint"] / d_nsc < 0.01


def test_oracle_scores_near_the_critical_angle(oracle_portable):
    """The oracle's n_det / n_scint scores on 3.2e8 photons against the
    derivative of the quantised model. Scoring only the hits that float32
    leaves outside total internal reflection gives 0.9370 for n_det, 4
    standard errors below."""
    g = golden()
    p = mirror_box_fresnel()
    n, parts = 320_000_000, 8
    m = n // parts
    w = np.ones((1, 9))
    gs = np.array([oracle_portable.gradient([p], [m], 8, 5000.0, w, first_id=k * m)[0]
                   for k in range(parts)])
    est, se = gs.mean(axis=0), gs.std(axis=0, ddof=1) / math.sqrt(parts)
    want_det = g["dP_dndet_smooth"] + g["grid_correction_ndet"]
    want_sc = g["dP_dnscint_smooth"] + g["grid_correction_nscint"]
    print(f"[mirror box Fresnel] n_det {est[3]:.5f} +- {se[3]:.5f} (quantised {want_det:.5f}), "
          f"n_scint {est[4]:.5f} +- {se[4]:.5f} (quantised {want_sc:.5f})")
    assert abs(est[3] - want_det) < 4.0 * se[3] + 2e-4
    assert abs(est[4] - want_sc) < 4.0 * se[4] + 2e-4
    assert se[3] < 0.003
