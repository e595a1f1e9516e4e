"""Differentiation workloads: gradient checks and batched calibration.

Single geometry: the plain-small box (50 mm cube, polished faces to n_out
1.0, +z readout, point source at the centre, tau 4 ns) at the nominal
theta = (abs 1000 mm, scat 300 mm, R 0.97, n_det 1.5, n_scint 1.63) -- a
finite scattering length so that d/d scat_length is not identically zero;
R does not enter with polished faces, so its gradient must vanish. Loss: the
chi-square of adjoint.chi2_loss between the simulated 256-bin arrival-time
histogram on [0, 50) ns (plus overflow and the detected fraction) and a
target generated at (1500, 400, 0.95, 1.45, 1.60) with a different seed.

Batched calibration: parameter sets with the long-histories geometry
(100 x 100 x 50 mm, Lambertian walls, Rayleigh, volume source), hidden
(abs_length, R, n_det) drawn from U[300, 3000] x U[0.85, 0.99] x U[1.4, 1.6],
targets generated under a held-out seed; each set fits its three parameters
by Adam on normalised coordinates. Each iteration is one captured hipGraph
(optphot_run_gpu_chi2_step): the batched forward launch, the chi-square loss
weights on the device, and the adjoint launch.
"""
from __future__ import annotations

import math
import time

import numpy as np

from . import adjoint
from .batch import EventBatch
from .params import SimParams, Source, Surface

C4_NOMINAL = np.array([1000.0, 300.0, 0.97, 1.5, 1.63])
C4_TARGET = np.array([1500.0, 400.0, 0.95, 1.45, 1.60])
C4_BINS, C4_TMAX = 256, 50.0


def c4_template(seed: int = 1) -> SimParams:
    return SimParams(hx=25.0, hy=25.0, hz=25.0, n_out=1.0, surface=Surface.POLISHED,
                     source=Source.POINT, tau=4.0, max_steps=10000, seed=seed)


def c4_problem(n: int, n_target: int | None = None, seed: int = 1, target_seed: int = 777):
    """(template, target histogram [bins+1]) for the single-geometry loss."""
    n_target = n_target or n
    tgt = adjoint.with_theta(c4_template(target_seed), C4_TARGET)
    T, _ = adjoint.histograms(EventBatch([tgt], np.array([n_target], np.uint64)), C4_BINS, C4_TMAX)
    return c4_template(seed), T[0]


def loss_and_grad(template, theta, target, n, n_target, first_id=0):
    """The single-geometry chi-square loss at theta and its adjoint gradient (same
    photons)."""
    p = adjoint.with_theta(template, theta)
    b = EventBatch([p], np.array([n], np.uint64))
    h, tf = adjoint.histograms(b, C4_BINS, C4_TMAX, first_id)
    h = h[0]
    f, F = h.sum(), target.sum()
    wk = 1.0 / (target + 1.0 / n_target)
    wf = 1.0 / (F * (1.0 - F) + 1.0 / n_target)
    loss = float(((h - target) ** 2 * wk).sum() + (f - F) ** 2 * wf)
    dl_dh = 2.0 * (h - target) * wk + 2.0 * (f - F) * wf
    g, ta = adjoint.run_gpu_adjoint(b, dl_dh[None, :], C4_BINS, C4_TMAX, first_id)
    return loss, g[0], tf.kernel_ms, ta.kernel_ms, dl_dh


def loss_at(template, theta, target, n, n_target, first_id=0, bins=C4_BINS, tmax=C4_TMAX):
    p = adjoint.with_theta(template, theta)
    h, _ = adjoint.histograms(EventBatch([p], np.array([n], np.uint64)), bins, tmax, first_id)
    if target is None:  # only the histogram is wanted
        return float("nan"), h[0]
    return float(adjoint.chi2_loss(h[0], target, n, n_target)), h[0]


def fd_gradient(template, theta, target, n, n_target, rel_step=1e-3, parts=4, dl_dh=None,
                bins=C4_BINS, tmax=C4_TMAX, params=range(5)):
    """Central finite differences with common random numbers (same seed and
    photon ids at theta +- h), in float64 from integer tallies. Each of
    `parts` disjoint photon-id ranges gives one estimate; returns mean and
    standard error (parameters not in `params` stay zero). With dl_dh given,
    differentiates the linearised loss sum_k dl_dh[k] h_k (the quantity the
    adjoint returns), which removes the O(h^2) curvature of the chi-square
    from the comparison."""
    m = n // parts
    est = np.zeros((parts, 5))
    for j in params:
        step = rel_step * abs(theta[j])
        tp, tm_ = theta.copy(), theta.copy()
        tp[j] += step
        tm_[j] -= step
        # the step actually taken, after float32 rounding of the parameters
        dp = np.float32(tp[j]) - np.float32(tm_[j])
        for k in range(parts):
            lp, hp = loss_at(template, tp, target, m, n_target, k * m, bins, tmax)
            lm, hm = loss_at(template, tm_, target, m, n_target, k * m, bins, tmax)
            if dl_dh is not None:
                est[k, j] = float(dl_dh @ (hp - hm)) / float(dp)
            else:
                est[k, j] = (lp - lm) / float(dp)
    return est.mean(axis=0), est.std(axis=0, ddof=1) / math.sqrt(parts)


def lr_gradient(template, theta, target, n, n_target, parts=4, dl_dh=None, bins=C4_BINS,
                tmax=C4_TMAX):
    """Adjoint gradient on `parts` disjoint id ranges: mean and standard error.
    With dl_dh given, that fixed weight vector is used for every part."""
    m = n // parts
    est = np.zeros((parts, 5))
    for k in range(parts):
        if dl_dh is None:
            _, g, _, _, _ = loss_and_grad(template, theta, target, m, n_target, first_id=k * m)
        else:
            p = adjoint.with_theta(template, theta)
            b = EventBatch([p], np.array([m], np.uint64))
            g, _ = adjoint.run_gpu_adjoint(b, dl_dh[None, :], bins, tmax, k * m)
            g = g[0]
        est[k] = g
    return est.mean(axis=0), est.std(axis=0, ddof=1) / math.sqrt(parts)


def linearised_weights(template, theta, target, n, n_target, bins, tmax):
    """dloss/dh of the chi-square at theta (the adjoint's weight vector)."""
    _, h = loss_at(template, theta, target, n, n_target, 0, bins, tmax)
    f, F = h.sum(), target.sum()
    wk = 1.0 / (target + 1.0 / n_target)
    wf = 1.0 / (F * (1.0 - F) + 1.0 / n_target)
    return 2.0 * (h - target) * wk + 2.0 * (f - F) * wf


def gradient_check(template, theta, target, n, n_target, bins, tmax, rel_step=1e-3, parts=16,
                   params=range(5), dl_dh=None):
    """Adjoint against central FD (common random numbers) of the linearised
    loss, per parameter: values, standard errors, relative difference and
    pull (difference over the combined standard error)."""
    if dl_dh is None:
        dl_dh = linearised_weights(template, theta, target, min(n, 10_000_000), n_target, bins,
                                   tmax)
    lr, lr_se = lr_gradient(template, theta, target, n, n_target, parts, dl_dh, bins, tmax)
    fd, fd_se = fd_gradient(template, theta, target, n, n_target, rel_step, parts, dl_dh, bins,
                            tmax, params)
    out = {}
    for j in params:
        se = math.hypot(lr_se[j], fd_se[j])
        out[adjoint.THETA[j]] = {
            "adjoint": float(lr[j]), "adjoint_se": float(lr_se[j]), "fd": float(fd[j]),
            "fd_se": float(fd_se[j]),
            "rel_diff": float((lr[j] - fd[j]) / abs(fd[j])) if fd[j] else 0.0,
            "pull": float((lr[j] - fd[j]) / se) if se else 0.0}
    return out


def gradient_problem(name: str, n_target: int, seed: int = 1, target_seed: int = 777):
    """(template, nominal theta, target histogram, bins, tmax) of a gradient
    check. "polished": the single-geometry workload above (R does not enter).
    "lambertian": the calibration geometry (100 x 100 x 50 mm, Lambertian
    walls R = 0.97, Rayleigh 300 mm, volume source), 128 bins on [0, 100) ns.
    "specular": a 20 x 60 x 40 mm box with specular walls R = 0.9, Rayleigh
    150 mm, n_det 1.45, volume source, 128 bins on [0, 60) ns. Targets come
    from other parameters under another seed."""
    if name == "polished":
        tpl, T = c4_problem(n_target, n_target, seed, target_seed)
        return tpl, C4_NOMINAL.copy(), T, C4_BINS, C4_TMAX
    if name == "lambertian":
        tpl = c5_template(seed)
        theta = np.array([2000.0, 300.0, 0.97, 1.5, 1.63])
        target = np.array([1500.0, 400.0, 0.95, 1.45, 1.60])
        bins, tmax = 128, 100.0
        tgt_tpl = c5_template(target_seed)
    elif name == "specular":
        tpl = SimParams(hx=10.0, hy=30.0, hz=20.0, n_scint=1.63, n_out=1.0,
                        surface=Surface.SPECULAR, source=Source.VOLUME, tau=4.0,
                        max_steps=10000, seed=seed)
        theta = np.array([1000.0, 150.0, 0.9, 1.45, 1.63])
        target = np.array([1300.0, 200.0, 0.85, 1.5, 1.60])
        bins, tmax = 128, 60.0
        tgt_tpl = tpl.copy(seed=target_seed)
    else:
        raise ValueError(f"unknown gradient problem {name!r}")
    T, _ = adjoint.histograms(EventBatch([adjoint.with_theta(tgt_tpl, target)],
                                         np.array([n_target], np.uint64)), bins, tmax)
    return tpl, theta, T[0], bins, tmax


# ---------------------------------------------------------------------------
# Batched calibration

LO = np.array([300.0, 0.85, 1.4])
HI = np.array([3000.0, 0.99, 1.6])
FIT = (0, 2, 3)  # theta indices fitted: abs_length, reflectivity, n_det


def c5_template(seed: int) -> SimParams:
    return SimParams(hx=50.0, hy=50.0, hz=25.0, n_scint=1.63, n_out=1.0, scat_length=300.0,
                     surface=Surface.LAMBERTIAN, source=Source.VOLUME, tau=4.0,
                     max_steps=10000, seed=seed)


def c5_batch(n_sets=256, photons=1_000_000, draw_seed=5, seed=1) -> EventBatch:
    """A forward batch of the calibration: every set at its hidden parameters
    (the draw of calibrate() with the same draw_seed), `photons` each."""
    rng = np.random.default_rng(draw_seed)
    th = np.tile(np.array([2000.0, 300.0, 0.97, 1.5, 1.63]), (n_sets, 1))
    th[:, list(FIT)] = LO + (HI - LO) * rng.random((n_sets, 3))
    tpl = c5_template(seed)
    return EventBatch([adjoint.with_theta(tpl, r) for r in th],
                      np.full(n_sets, photons, np.uint64))


def calibrate(n_sets=256, photons=1_000_000, iters=50, bins=128, tmax=100.0, lr=0.08,
              draw_seed=5, seed=1, target_seed=99, target_photons=None, log=None,
              time_forward=True):
    """Fits (abs_length, R, n_det) of every set to its own target histogram
    by Adam on normalised coordinates (each parameter linear over its range),
    starting at the middle of each range. Returns a dict with hidden / fitted parameters,
    errors and timings; with time_forward, each iteration also times a
    forward-only batch at the same parameters (outside the fit) for the
    forward+adjoint / forward ratio. The forward launches take the autotuned
    batch shape for this batch class, if any."""
    from . import autotune
    rng = np.random.default_rng(draw_seed)
    hidden = LO + (HI - LO) * rng.random((n_sets, 3))
    target_photons = target_photons or photons
    base = np.array([2000.0, 300.0, 0.97, 1.5, 1.63])

    def theta_rows(p):  # p: [n_sets, 3] (abs_length, R, n_det)
        th = np.tile(base, (n_sets, 1))
        th[:, list(FIT)] = p
        return th

    # targets at the hidden parameters, held-out seed
    tpl_t = c5_template(target_seed)
    tb = EventBatch([adjoint.with_theta(tpl_t, th) for th in theta_rows(hidden)],
                    np.full(n_sets, target_photons, np.uint64))
    T, _ = adjoint.histograms(tb, bins, tmax)

    def theta_of(x):
        return theta_rows(LO + (HI - LO) * x)

    x = np.full((n_sets, 3), 0.5)
    m = np.zeros_like(x)
    v = np.zeros_like(x)
    b1, b2, eps = 0.9, 0.999, 1e-8
    tpl = c5_template(seed)
    tuning = None
    step_ms = fwd_ms = 0.0
    wall0 = time.perf_counter()
    losses = []
    for it in range(1, iters + 1):
        th = theta_of(x)
        batch = EventBatch([adjoint.with_theta(tpl, r) for r in th],
                           np.full(n_sets, photons, np.uint64))
        if it == 1:
            tuning = autotune.batch_choice_for(batch, bins)
        # one captured graph per iteration: forward, chi-square weights, adjoint
        _, loss, g, ts = adjoint.run_gpu_chi2_step(batch, T, target_photons, bins, tmax,
                                                   forward_tuning=tuning)
        step_ms += ts.kernel_ms
        if time_forward:  # forward alone at the same parameters, for the cost ratio
            fwd_ms += adjoint.histograms(batch, bins, tmax, tuning=tuning)[1].kernel_ms
        gx = g[:, list(FIT)] * (HI - LO)  # chain rule to normalised coordinates
        m = b1 * m + (1 - b1) * gx
        v = b2 * v + (1 - b2) * gx * gx
        x = np.clip(x - lr * (m / (1 - b1 ** it)) / (np.sqrt(v / (1 - b2 ** it)) + eps), 0.0, 1.0)
        losses.append(float(np.median(loss)))
        if log:
            log(f"iter {it:3d}: median loss {losses[-1]:.4e}")
    wall = time.perf_counter() - wall0
    fitted = LO + (HI - LO) * x
    rel_err = np.abs(fitted - hidden) / hidden
    total = n_sets * photons * iters
    return {
        "n_sets": n_sets, "photons_per_set": photons, "iterations": iters,
        "hidden": hidden, "fitted": fitted,
        "median_rel_error": dict(zip(("abs_length", "reflectivity", "n_det"),
                                     np.median(rel_err, axis=0).tolist())),
        "p90_rel_error": dict(zip(("abs_length", "reflectivity", "n_det"),
                                  np.percentile(rel_err, 90, axis=0).tolist())),
        "losses": losses,
        "graph_ms": step_ms, "wall_s": wall, "forward_only_ms": fwd_ms,
        "forward_photons_per_s": total / (fwd_ms * 1e-3) if fwd_ms else None,
        "fwd_adj_photons_per_s": total / (step_ms * 1e-3),
        "fwd_adj_over_fwd": step_ms / fwd_ms if fwd_ms else None,
        "forward_tuning": tuning,
        "targets": T, "target_photons": target_photons, "bins": bins, "tmax": tmax,
        "seed": seed, "draw_seed": draw_seed, "lr": lr,
    }


def _theta_rows(p, n_sets):
    """Full theta rows with (abs_length, R, n_det) = p [E, 3]."""
    th = np.tile(np.array([2000.0, 300.0, 0.97, 1.5, 1.63]), (n_sets, 1))
    th[:, list(FIT)] = p
    return th


def step_at(x, T, target_photons, photons, bins, tmax, seed=1, tuning=None):
    """(loss [E], gradient in normalised coordinates [E, 3]) of every set at
    normalised parameters x [E, 3], from one calibration-step graph."""
    n_sets = len(x)
    tpl = c5_template(seed)
    batch = EventBatch([adjoint.with_theta(tpl, r)
                        for r in _theta_rows(LO + (HI - LO) * x, n_sets)],
                       np.full(n_sets, photons, np.uint64))
    _, loss, g, _ = adjoint.run_gpu_chi2_step(batch, T, target_photons, bins, tmax,
                                              forward_tuning=tuning)
    return loss, g[:, list(FIT)] * (HI - LO)


def hessian_at(x, T, target_photons, photons, bins, tmax, seed=1, tuning=None, h=0.02):
    """Loss Hessian [E, 3, 3] in normalised coordinates at x, by central
    differences of the adjoint gradient with common random numbers."""
    H = np.zeros((len(x), 3, 3))
    for j in range(3):
        xp, xm = x.copy(), x.copy()
        xp[:, j] += h
        xm[:, j] -= h
        _, gp = step_at(xp, T, target_photons, photons, bins, tmax, seed, tuning)
        _, gm = step_at(xm, T, target_photons, photons, bins, tmax, seed, tuning)
        H[:, :, j] = (gp - gm) / (2.0 * h)
    return 0.5 * (H + np.transpose(H, (0, 2, 1)))


def calibration_diagnostics(r, h=0.02, long_iters=300, log=None):
    """Why each fitted parameter ends where it does, from a calibrate()
    result r (medians over the sets):
      * loss at the fitted and at the hidden parameters (same simulation
        seed);
      * the same fit run for long_iters iterations: once its loss is at the
        hidden parameters' level (the noise floor of the data), its remaining
        error is what the data cannot resolve;
      * the Hessian of the loss at the hidden parameters in normalised
        coordinates, by central differences of the adjoint gradient with
        common random numbers (step h): eigenvalues and the abs_length share
        of the softest direction."""
    n_sets, photons = r["n_sets"], r["photons_per_set"]
    T, nt, bins, tmax = r["targets"], r["target_photons"], r["bins"], r["tmax"]
    tu, seed = r["forward_tuning"], r["seed"]

    def x_of(p):
        return (p - LO) / (HI - LO)

    loss_h, _ = step_at(x_of(r["hidden"]), T, nt, photons, bins, tmax, seed, tu)
    loss_f, _ = step_at(x_of(r["fitted"]), T, nt, photons, bins, tmax, seed, tu)
    H = hessian_at(x_of(r["hidden"]), T, nt, photons, bins, tmax, seed, tu, h)
    ev, evec = np.linalg.eigh(H)
    r2 = calibrate(n_sets=n_sets, photons=photons, iters=long_iters, bins=bins, tmax=tmax,
                   lr=r["lr"], draw_seed=r["draw_seed"], seed=seed, target_photons=nt,
                   time_forward=False, log=log)
    loss_l, _ = step_at(x_of(r2["fitted"]), T, nt, photons, bins, tmax, seed, tu)
    names = ("abs_length", "reflectivity", "n_det")
    return {
        "loss_fitted_over_hidden_median": float(np.median(loss_f / loss_h)),
        "long_fit_iterations": long_iters,
        "long_fit_loss_over_hidden_median": float(np.median(loss_l / loss_h)),
        "long_fit_fraction_loss_le_hidden": float(np.mean(loss_l <= loss_h)),
        "long_fit_median_rel_error": r2["median_rel_error"],
        "long_fit_p90_rel_error": r2["p90_rel_error"],
        "hessian_diag_median": dict(zip(names, np.median(np.diagonal(H, axis1=1, axis2=2),
                                                          axis=0).tolist())),
        "hessian_eigenvalues_median": np.median(ev, axis=0).tolist(),
        "hessian_condition_median": float(np.median(ev[:, 2] / np.maximum(ev[:, 0], 1e-30))),
        "softest_direction_abs_component_median": float(np.median(np.abs(evec[:, 0, 0]))),
        "hessian_step": h,
    }
