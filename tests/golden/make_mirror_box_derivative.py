"""Writes tests/golden/mirror_box_derivative.json: the detection probability
of the mirror box with a Fresnel readout (specular walls R_w = 0.9,
n_scint 1.63 -> n_det 1.5, no bulk processes, point source) and its index
derivatives, by float64 quadrature (tests/analytic.py), plus the correction
that the point source's 2^-23 grid of emission cosines makes to the
derivative of the simulated model at fixed indices. Pure mathematics: no
reference code is involved. Takes about two minutes.

    python tests/golden/make_mirror_box_derivative.py
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from gpu_optical_photons_amd.params import SimParams  # noqa: E402
from tests import analytic  # noqa: E402

R_W = 0.9
P = dict(hx=10.0, hy=30.0, hz=20.0, n_scint=1.63, n_det=1.5, src_x=3.0, src_y=-7.0, src_z=-4.0)


def main():
    p = SimParams(**P)
    pdet, d_ndet, d_nsc = analytic.mirror_box_fresnel_readout(p, R_W, 2000, 2000)
    c_ndet, c_nsc = analytic.mirror_box_emission_grid_correction(p, R_W)
    out = {"params": P, "reflectivity": R_W, "P_det": pdet,
           "dP_dndet_smooth": d_ndet, "dP_dnscint_smooth": d_nsc,
           "grid_correction_ndet": c_ndet, "grid_correction_nscint": c_nsc,
           "quadrature": "mirror_box_fresnel_readout(n_v=2000, n_phi=2000); "
                         "mirror_box_emission_grid_correction(window=2e-4, n_phi=2000)"}
    with open(os.path.join(HERE, "mirror_box_derivative.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out, indent=1))


if __name__ == "__main__":
    main()
