"""Differentiable transport: arrival-time histograms of parameter sets and
their adjoint, behind a torch.autograd.Function.

The forward pass is the batched tally (one persistent launch over all sets,
see batch.py); the output for set e is h_e[k] = (photons of set e detected in
time bin k) / N_e, k = 0..time_bins with the last bin the overflow, so
sum_k h_e[k] is the detected fraction. The backward pass is the adjoint
kernel (optphot_run_gpu_adjoint): given dL/dh it recomputes every history
from its Philox counter and returns dL/dtheta for
theta = (abs_length, scat_length, reflectivity, n_det, n_scint) per set,
by likelihood-ratio scores of the absorb / scatter / Fresnel / reflector
decisions and the pathwise arrival-time term for n_scint. Same photons (seed
and ids) in both passes, so the gradient is that of the realised estimator's
expectation, not of a different sample.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import backend
from .batch import EventBatch, run_gpu_batch
from .params import NUM_FATES, SimParams

THETA = ("abs_length", "scat_length", "reflectivity", "n_det", "n_scint")

_fn = None


def _bind():
    global _fn
    if _fn is None:
        f = backend.lib().optphot_run_gpu_adjoint
        f.restype = ctypes.c_int
        f.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64,
                      ctypes.c_uint32, ctypes.c_float, ctypes.c_void_p, ctypes.c_void_p,
                      ctypes.POINTER(backend.WavefrontTuning), ctypes.POINTER(backend.RunTiming)]
        _fn = f
    return _fn


def with_theta(template: SimParams, theta) -> SimParams:
    """`template` with theta = (abs_length, scat_length, reflectivity, n_det,
    n_scint) substituted (float32 rounding happens in SimParams)."""
    return template.copy(**{k: float(v) for k, v in zip(THETA, theta)})


def histograms(batch: EventBatch, time_bins: int, time_max: float, first_id: int = 0,
               tuning=None):
    """h [E, time_bins + 1] (fractions of each set's photons per arrival-time
    bin, detected photons only) and the RunTiming of the launch."""
    counts, tm, _ = run_gpu_batch(batch, time_bins, time_max, first_id=first_id, tuning=tuning)
    n = np.maximum(batch.photons.astype(np.float64), 1.0)[:, None]
    return counts[:, NUM_FATES:].astype(np.float64) / n, tm


def run_gpu_adjoint(batch: EventBatch, weights, time_bins: int, time_max: float,
                    first_id: int = 0, tuning=None):
    """grad [E, 5] = d/dtheta sum_k weights[e, k] h_e[k], and the RunTiming."""
    fn = _bind()
    arr = (SimParams * batch.n_events)(*batch.params)
    photons = np.ascontiguousarray(batch.photons, dtype=np.uint64)
    w = np.ascontiguousarray(weights, dtype=np.float64)
    if w.shape != (batch.n_events, time_bins + 1):
        raise ValueError(f"weights must have shape ({batch.n_events}, {time_bins + 1})")
    grad = np.zeros((batch.n_events, 5), np.float64)
    tm = backend.RunTiming()
    if isinstance(tuning, dict):
        tuning = backend.WavefrontTuning(**tuning)
    rc = fn(ctypes.addressof(arr), photons.ctypes.data, batch.n_events, first_id, time_bins,
            time_max, w.ctypes.data, grad.ctypes.data,
            ctypes.byref(tuning) if tuning is not None else None, ctypes.byref(tm))
    backend._check(rc)
    return grad, tm


_step_fn = None


def run_gpu_chi2_step(batch: EventBatch, target, target_photons, time_bins: int, time_max: float,
                      first_id: int = 0, forward_tuning=None, adjoint_tuning=None):
    """One calibration step in a single captured graph (optphot_run_gpu_chi2_step):
    forward histograms, the chi-square loss of chi2_loss against `target`
    ([E, time_bins + 1] fractions) and its adjoint gradient. Returns
    (hist [E, T+1], loss [E], grad [E, 5], RunTiming)."""
    global _step_fn
    if _step_fn is None:
        f = backend.lib().optphot_run_gpu_chi2_step
        f.restype = ctypes.c_int
        f.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64,
                      ctypes.c_uint32, ctypes.c_float, ctypes.c_void_p, ctypes.c_void_p,
                      ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                      ctypes.POINTER(backend.WavefrontTuning),
                      ctypes.POINTER(backend.WavefrontTuning), ctypes.POINTER(backend.RunTiming)]
        _step_fn = f
    E = batch.n_events
    arr = (SimParams * E)(*batch.params)
    photons = np.ascontiguousarray(batch.photons, dtype=np.uint64)
    T = np.ascontiguousarray(target, dtype=np.float64)
    if T.shape != (E, time_bins + 1):
        raise ValueError(f"target must have shape ({E}, {time_bins + 1})")
    nt = np.ascontiguousarray(np.broadcast_to(np.asarray(target_photons, np.float64), (E,)))
    hist = np.zeros((E, time_bins + 1), np.float64)
    loss = np.zeros(E, np.float64)
    grad = np.zeros((E, 5), np.float64)
    tm = backend.RunTiming()
    tunings = [backend.WavefrontTuning(**t) if isinstance(t, dict) else t
               for t in (forward_tuning, adjoint_tuning)]
    rc = _step_fn(ctypes.addressof(arr), photons.ctypes.data, E, first_id, time_bins, time_max,
                  T.ctypes.data, nt.ctypes.data, hist.ctypes.data, loss.ctypes.data,
                  grad.ctypes.data, *[ctypes.byref(t) if t is not None else None for t in tunings],
                  ctypes.byref(tm))
    backend._check(rc)
    return hist, loss, grad, tm


def make_function():
    """The torch.autograd.Function (torch is imported lazily)."""
    import torch

    class PhotonHistogram(torch.autograd.Function):
        """h = PhotonHistogram.apply(theta, templates, photons, time_bins,
        time_max, first_id): theta [E, 5] (float64) -> h [E, time_bins + 1]."""

        @staticmethod
        def forward(ctx, theta, templates, photons, time_bins, time_max, first_id):
            th = theta.detach().cpu().numpy()
            params = [with_theta(t, row) for t, row in zip(templates, th)]
            batch = EventBatch(params, np.asarray(photons, np.uint64))
            h, _ = histograms(batch, time_bins, time_max, first_id)
            ctx.batch = batch
            ctx.cfg = (time_bins, time_max, first_id)
            return torch.from_numpy(h).to(theta.device)

        @staticmethod
        def backward(ctx, grad_h):
            time_bins, time_max, first_id = ctx.cfg
            g, _ = run_gpu_adjoint(ctx.batch, grad_h.detach().cpu().double().numpy(), time_bins,
                                   time_max, first_id)
            return (torch.from_numpy(g).to(grad_h.device), None, None, None, None, None)

    return PhotonHistogram


def chi2_loss(h, target, n_sim, n_target):
    """Pearson chi-square per photon between simulated and target arrival-time
    histograms (fractions, overflow bin included) plus the detected fraction:
      sum_k (h_k - T_k)^2 / (T_k + 1/n_target)
        + (f - F)^2 / (F (1 - F) + 1/n_target),   f = sum_k h_k, F = sum_k T_k.
    Weights depend on the target only, so the loss is smooth in h. Works on
    numpy arrays or torch tensors ([..., bins])."""
    f = h.sum(-1)
    F = target.sum(-1)
    wk = 1.0 / (target + 1.0 / n_target)
    wf = 1.0 / (F * (1.0 - F) + 1.0 / n_target)
    return ((h - target) ** 2 * wk).sum(-1) + (f - F) ** 2 * wf
