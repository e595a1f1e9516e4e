"""The adjoint kernel against the FP64 CPU reference of the same estimator on
the same photons, the torch.autograd.Function wiring, and the statistical
agreement of the gradient with central finite differences (common random
numbers) of the forward kernel."""
import math

import numpy as np
import pytest

from gpu_optical_photons_amd.params import SimParams, Source, Surface

pytestmark = pytest.mark.gpu

SETS = [
    SimParams(n_det=1.5, abs_length=1000.0, scat_length=300.0, seed=1),            # polished
    SimParams(hx=50.0, hy=50.0, hz=25.0, n_det=1.5, abs_length=2000.0, scat_length=300.0,
              surface=Surface.LAMBERTIAN, reflectivity=0.97, source=Source.VOLUME, seed=1),
    SimParams(surface=Surface.SPECULAR, reflectivity=0.9, n_det=1.45, scat_length=150.0,
              source=Source.VOLUME, hx=10.0, hy=30.0, hz=20.0, seed=1),
]


@pytest.mark.parametrize("k", range(len(SETS)))
def test_adjoint_equals_fp64_reference(backend, oracle_portable, k):
    from gpu_optical_photons_amd.adjoint import run_gpu_adjoint
    from gpu_optical_photons_amd.batch import EventBatch
    p = SETS[k]
    n, tb, tmax = 300_000, 64, 60.0
    w = np.random.default_rng(k).normal(size=(1, tb + 1))
    g, _ = run_gpu_adjoint(EventBatch([p], np.array([n], np.uint64)), w, tb, tmax)
    ref = oracle_portable.gradient([p], [n], tb, tmax, w)
    scale = np.abs(ref).max()
    print(f"[adjoint] set {k}: gpu {g[0]} ref {ref[0]}", flush=True)
    assert np.all(np.abs(g - ref) <= 1e-5 * scale + 2e-4 * np.abs(ref))


def test_adjoint_batch_of_sets_equals_reference(backend, oracle_portable):
    from gpu_optical_photons_amd.adjoint import run_gpu_adjoint
    from gpu_optical_photons_amd.batch import EventBatch
    n = np.array([120_000, 80_000, 50_000], np.uint64)
    tb, tmax = 32, 60.0
    w = np.random.default_rng(9).normal(size=(3, tb + 1))
    g, _ = run_gpu_adjoint(EventBatch(SETS, n), w, tb, tmax, first_id=7)
    ref = oracle_portable.gradient(SETS, n, tb, tmax, w, first_id=7)
    for e in range(3):
        scale = np.abs(ref[e]).max()
        assert np.all(np.abs(g[e] - ref[e]) <= 1e-5 * scale + 2e-4 * np.abs(ref[e])), e
    for tu in (dict(block_size=64, steps_per_refill=1, queue_chunk=64),
               dict(block_size=1024, steps_per_refill=8, queue_chunk=1024)):
        g2, _ = run_gpu_adjoint(EventBatch(SETS, n), w, tb, tmax, first_id=7, tuning=tu)
        assert np.allclose(g2, g, rtol=1e-9, atol=1e-12 * np.abs(g).max())


def test_autograd_function_backward(backend, oracle_portable):
    import torch

    from gpu_optical_photons_amd import adjoint
    n, tb, tmax = 200_000, 32, 50.0
    templates = [SETS[0], SETS[1]]
    theta = torch.tensor([[1000.0, 300.0, 0.97, 1.5, 1.63],
                          [2000.0, 300.0, 0.97, 1.5, 1.63]], dtype=torch.float64,
                         requires_grad=True)
    F = adjoint.make_function()
    h = F.apply(theta, templates, [n, n], tb, tmax, 0)
    target = h.detach() * 1.05
    loss = adjoint.chi2_loss(h, target, n, n).sum()
    loss.backward()
    h2 = h.detach().numpy()
    t2 = target.numpy()
    f, Fd = h2.sum(1, keepdims=True), t2.sum(1, keepdims=True)
    wk = 1.0 / (t2 + 1.0 / n)
    wf = 1.0 / (Fd * (1.0 - Fd) + 1.0 / n)
    dl_dh = 2.0 * (h2 - t2) * wk + 2.0 * (f - Fd) * wf
    params = [adjoint.with_theta(t, r) for t, r in zip(templates, theta.detach().numpy())]
    ref = oracle_portable.gradient(params, [n, n], tb, tmax, dl_dh)
    g = theta.grad.numpy()
    assert np.all(np.abs(g - ref) <= 1e-5 * np.abs(ref).max() + 2e-4 * np.abs(ref))
    assert g[0, 2] == 0.0  # polished faces: the reflectivity does not enter


def test_gradient_agrees_with_finite_differences(backend):
    """Single-geometry differentiation workload: d(loss)/d theta by the adjoint
    against central finite differences (relative step 1e-3, common random
    numbers) of the linearised loss, both with standard errors from four
    disjoint photon-id ranges. Agreement within 4 combined standard errors
    per parameter, the reflectivity exactly zero."""
    from gpu_optical_photons_amd import calibrate as cal
    n, nt = 40_000_000, 40_000_000
    tpl, T = cal.c4_problem(n, nt)
    theta = cal.C4_NOMINAL.copy()
    _, _, _, _, dl_dh = cal.loss_and_grad(tpl, theta, T, n, nt)
    lr, lr_se = cal.lr_gradient(tpl, theta, T, n, nt, dl_dh=dl_dh)
    fd, fd_se = cal.fd_gradient(tpl, theta, T, n, nt, rel_step=1e-3, dl_dh=dl_dh)
    for j, name in enumerate(("abs_length", "scat_length", "reflectivity", "n_det", "n_scint")):
        se = math.hypot(lr_se[j], fd_se[j])
        rel = (lr[j] - fd[j]) / abs(fd[j]) if fd[j] else float("nan")
        print(f"[gradcheck] {name:12s} adjoint {lr[j]: .6e} +- {lr_se[j]:.2e}  "
              f"FD {fd[j]: .6e} +- {fd_se[j]:.2e}  rel {rel: .4f}", flush=True)
        if name == "reflectivity":
            assert lr[j] == 0.0
        else:
            assert abs(lr[j] - fd[j]) <= 4.0 * se, name


@pytest.mark.parametrize("n_sets,ftu", [
    (1, None), (6, None),
    # the calibration sets' autotuned forward shape: LDS rings, long chunks
    (6, dict(block_size=512, steps_per_refill=2, queue_chunk=2048, stream_in_lds=1))])
def test_chi2_step_graph_equals_separate_calls(backend, n_sets, ftu):
    """The one-graph calibration step (forward, chi-square weights on the
    device, adjoint) against the same quantities from separate calls, over
    three replays with new parameters (the graph must read them afresh), with
    the default and a tuned forward launch shape."""
    from gpu_optical_photons_amd import adjoint
    from gpu_optical_photons_amd.batch import EventBatch
    from gpu_optical_photons_amd.calibrate import c5_template
    bins, tmax, n, nt = 96, 80.0, 150_000, 400_000.0
    tpl = c5_template(3)
    rng = np.random.default_rng(n_sets)
    base = np.array([2000.0, 300.0, 0.97, 1.5, 1.63])
    def draw(scale):
        th = base * (1 + scale * rng.normal(size=5))
        th[2] = min(th[2], 0.995)  # a reflectivity, at most 1
        return th
    targets = EventBatch([adjoint.with_theta(c5_template(99), draw(0.05)) for _ in range(n_sets)],
                         np.full(n_sets, int(nt), np.uint64))
    T, _ = adjoint.histograms(targets, bins, tmax)
    for rep in range(3):
        th = base * (1 + 0.04 * rng.normal(size=(n_sets, 5)))
        th[:, 2] = np.minimum(th[:, 2], 0.995)  # a reflectivity, at most 1
        batch = EventBatch([adjoint.with_theta(tpl, r) for r in th],
                           np.full(n_sets, n, np.uint64))
        print(f"[chi2 step] sets {n_sets} replay {rep}", flush=True)
        hist, loss, grad, _ = adjoint.run_gpu_chi2_step(batch, T, nt, bins, tmax,
                                                        forward_tuning=ftu)
        h, _ = adjoint.histograms(batch, bins, tmax)
        assert np.array_equal(hist, h)
        np.testing.assert_allclose(loss, adjoint.chi2_loss(h, T, n, nt), rtol=1e-12)
        f, F = h.sum(1, keepdims=True), T.sum(1, keepdims=True)
        wk = 1.0 / (T + 1.0 / nt)
        wf = 1.0 / (F * (1.0 - F) + 1.0 / nt)
        dl_dh = 2.0 * (h - T) * wk + 2.0 * (f - F) * wf
        g, _ = adjoint.run_gpu_adjoint(batch, dl_dh, bins, tmax)
        np.testing.assert_allclose(grad, g, rtol=1e-9, atol=1e-12 * np.abs(g).max())


def test_reflector_gradient_against_mirror_box_closed_form(backend):
    """dP_det/dR on the GPU against <k R^(k-1)> of the unfolded mirror box
    (specular walls, matched readout, no bulk processes, point source): the
    closed form of tests/analytic.py, no estimator involved."""
    from gpu_optical_photons_amd.adjoint import run_gpu_adjoint
    from gpu_optical_photons_amd.batch import EventBatch
    from tests import analytic
    from tests.test_adjoint_reflector import mirror_box
    for R in (0.97, 0.75):
        p = mirror_box(R)
        n, parts, tb, tmax = 100_000_000, 4, 8, 2000.0
        m = n // parts
        gs = np.array([run_gpu_adjoint(EventBatch([p], np.array([m], np.uint64)),
                                       np.ones((1, tb + 1)), tb, tmax, k * m)[0][0]
                       for k in range(parts)])
        g, se = gs.mean(axis=0), gs.std(axis=0, ddof=1) / math.sqrt(parts)
        _, dP = analytic.mirror_box_detection(p, R)
        print(f"[mirror box R={R}] GPU {g[2]:.6f} +- {se[2]:.6f}, closed form {dP:.6f}",
              flush=True)
        assert abs(g[2] - dP) < 4.0 * se[2] + 1e-4 * dP
        assert se[2] < 0.002 * dP


@pytest.mark.parametrize("name", ["lambertian", "specular"])
def test_gradient_fd_at_reflecting_geometries(backend, name):
    """Adjoint against central FD (relative step 1e-3, common random numbers)
    of the linearised chi-square loss for all five parameters where the
    reflectivity enters: the calibration geometry (Lambertian R = 0.97) and a
    specular box (R = 0.9). 8e8 photons per side in 8 disjoint id ranges;
    every parameter within 4 combined standard errors."""
    from gpu_optical_photons_amd import calibrate as cal
    n = 800_000_000
    tpl, theta, T, bins, tmax = cal.gradient_problem(name, 20_000_000)
    chk = cal.gradient_check(tpl, theta, T, n, 20_000_000, bins, tmax, parts=8)
    for k, v in chk.items():
        print(f"[gradcheck {name}] {k:12s} adjoint {v['adjoint']: .6e} +- {v['adjoint_se']:.1e} "
              f"FD {v['fd']: .6e} +- {v['fd_se']:.1e} pull {v['pull']: .2f}", flush=True)
        assert abs(v["pull"]) < 4.0, k
    assert chk["reflectivity"]["adjoint"] != 0.0
    assert abs(chk["reflectivity"]["rel_diff"]) < 0.02


def test_index_gradients_near_the_critical_angle(backend):
    """Mirror box with a Fresnel readout (tests/test_adjoint_critical.py):
    the adjoint at 1.6e10 photons against the derivative of the simulated
    model (point-source emission cosines on their 2^-23 grid), and central
    finite differences (step 1e-3, common random numbers, 4e9 photons per
    side) against the smooth derivative, both by float64 quadrature."""
    from gpu_optical_photons_amd import calibrate as cal
    from tests.test_adjoint_critical import golden, mirror_box_fresnel
    g = golden()
    p = mirror_box_fresnel()
    theta = np.array([math.inf, math.inf, g["reflectivity"], 1.5, 1.63])
    bins, tmax = 8, 5000.0
    w = np.ones(bins + 1)
    lr, lr_se = cal.lr_gradient(p, theta, None, 16_000_000_000, 1, parts=16, dl_dh=w, bins=bins,
                                tmax=tmax)
    fd, fd_se = cal.fd_gradient(p, theta, None, 4_000_000_000, 1, rel_step=1e-3, parts=16,
                                dl_dh=w, bins=bins, tmax=tmax, params=(3, 4))
    for j, key in ((3, "ndet"), (4, "nscint")):
        smooth = g[f"dP_d{key}_smooth"]
        quantised = smooth + g[f"grid_correction_{key}"]
        print(f"[critical angle] {key}: adjoint {lr[j]:.5f} +- {lr_se[j]:.5f} (quantised "
              f"{quantised:.5f}); FD {fd[j]:.5f} +- {fd_se[j]:.5f} (smooth {smooth:.5f})",
              flush=True)
        assert abs(lr[j] - quantised) < 4.0 * lr_se[j] + 1e-4 * abs(quantised)
        assert abs(fd[j] - smooth) < 4.0 * fd_se[j] + 1e-4 * abs(smooth)
