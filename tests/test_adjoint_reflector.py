"""Pins the reflector score of the FP64 adjoint reference (oracle_adjoint) at
nonzero reflectivity against a closed form.

Specular box, readout index matched to the scintillator (every photon that
reaches +z is detected), no bulk absorption or scattering, point source.
Unfolding the five mirrors, each direction reaches the readout after a fixed
number k of reflector hits, each survived with probability R, so
  P_det = <R^k>,   dP_det/dR = <k R^(k-1)>
over isotropic directions. tests/analytic.py sums both over the tiles of the
unfolded readout plane with exact rectangle solid angles, and a direct trace
of the reflections checks the tile counting. The estimator under test is the
likelihood-ratio score refl_hits / R of every detected photon (reflector
decisions, include/optphot/transport.hpp:133-137); a dropped 1/R, an extra
hit or a sign slip moves it by tens of standard errors."""
import math

import numpy as np
import pytest

from gpu_optical_photons_amd.params import SimParams
from tests import analytic

INF = math.inf


def mirror_box(R, seed=21):
    return SimParams(hx=10.0, hy=30.0, hz=20.0, n_scint=1.63, n_det=1.63, abs_length=INF,
                     scat_length=INF, surface=2, reflectivity=R, src_x=3.0, src_y=-7.0,
                     src_z=-4.0, tau=4.0, seed=seed)


def test_tile_counting_matches_traced_reflections():
    p = mirror_box(0.9)
    rng = np.random.default_rng(1)
    for _ in range(2000):
        mu = rng.uniform(-1.0, 1.0)
        if abs(mu) < 0.01:  # nearly horizontal: tens of thousands of hits
            continue
        phi = rng.uniform(0.0, 2.0 * math.pi)
        s = math.sqrt(1.0 - mu * mu)
        d = (s * math.cos(phi), s * math.sin(phi), mu)
        h = p.hz - p.src_z if mu > 0 else p.src_z + 3.0 * p.hz
        t = h / abs(mu)
        i = math.floor((p.src_x + d[0] * t + p.hx) / (2.0 * p.hx))
        j = math.floor((p.src_y + d[1] * t + p.hy) / (2.0 * p.hy))
        assert analytic.mirror_box_reflections(p, d) == abs(i) + abs(j) + (mu < 0), (mu, phi)


def test_closed_form_limits():
    p = mirror_box(0.9)
    k, f = analytic.mirror_box_tiles(p, 400)
    assert abs(f.sum() - 1.0) < 5e-3  # all but the nearly horizontal directions
    assert analytic.mirror_box_detection(p, 0.0)[0] == pytest.approx(
        analytic.face_solid_angle_fraction(p, 5), rel=1e-12)  # R = 0: direct hits only
    # the derivative of the closed form is its own finite difference
    h = 1e-6
    fd = (analytic.mirror_box_detection(p, 0.9 + h)[0]
          - analytic.mirror_box_detection(p, 0.9 - h)[0]) / (2 * h)
    assert analytic.mirror_box_detection(p, 0.9)[1] == pytest.approx(fd, rel=1e-7)


@pytest.mark.parametrize("R", [0.9, 0.6])
def test_reflector_score_against_mirror_box(oracle_portable, R):
    p = mirror_box(R)
    n, parts, tb, tmax = 4_000_000, 4, 8, 2000.0
    w = np.ones((1, tb + 1))
    m = n // parts
    gs = np.array([oracle_portable.gradient([p], [m], tb, tmax, w, first_id=k * m)[0]
                   for k in range(parts)])
    g, se = gs.mean(axis=0), gs.std(axis=0, ddof=1) / math.sqrt(parts)
    P, dP = analytic.mirror_box_detection(p, R)
    print(f"[mirror box R={R}] score {g[2]:.6f} +- {se[2]:.6f}, closed form {dP:.6f}")
    assert abs(g[2] - dP) < 4.0 * se[2] + 1e-4 * dP
    assert se[2] < 0.01 * dP  # a dropped 1/R (a factor R) would be many sigma away
    assert abs(g[2] * R - dP) > 10.0 * se[2]
    # the forward estimate of the same photons: binomial agreement with <R^k>
    t = oracle_portable.tally(p, n, tb, tmax, 1)
    assert analytic.binomial_p(int(t["fate_counts"][0]), n, P) > 0.01
    # no other parameter enters this box
    assert g[0] == 0.0 and g[1] == 0.0
