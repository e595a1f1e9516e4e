"""Tracing the differences between two math builds of the transport
(portable math against the reference's default glibc build) to rounding.

The builds run the same photon with the same random words; they differ only
in the last-bit rounding of logf / sinf / cosf / cbrtf (pmath.hpp:134-158).
A photon whose discrete history is the same in both can still end at a
different position or time: the emitted direction differs by about one ulp
and that difference is carried, and amplified by the geometry, through every
later step. `carried_rounding_ratios` checks a photon's two step traces
(oracle step_trace, emitted state first) against a first-order bound on how
far such a difference can grow in one step, with R = 1e-6 for the new
rounding each step adds (8 ulps of 1; relative for times):

* position (max-norm over axes, in units of the largest box half-size h):
  a flight of length s carries the position difference and adds the
  direction difference times s; a flight that ends on a wall is projected
  onto it, which multiplies by at most 1 + 1/cos_i (cos_i: the direction's
  component along the wall normal):
      dp_j <= (1 + 1/cos_i) (dp_{j-1} + (dd_{j-1} + R) s / h) + R
* direction (max-norm): a specular or Fresnel reflection keeps it, and so
  does a photon's last step (absorbed or detected, its direction is not
  redrawn); a Lambertian reflection (the face's Bernoulli decision true)
  draws a fresh direction in the exact frame of an axis-aligned normal, so
  only rounding is left; a Rayleigh scatter rotates about the old direction
  (its frame carries the difference with a factor of at most 4) with a
  cosine that can differ by an ulp, which sin = sqrt(1 - cos^2) amplifies by
  cot(theta):
      dd_j <= dd_{j-1} + R                              (reflection)
      dd_j <= R                                         (Lambertian wall)
      dd_j <= 4 dd_{j-1} + R (1 + |cos t| / sin t)      (Rayleigh)
* time: t grows by the flight length times n_scint / c, and the flight
  lengths differ by at most sqrt(3) (dp_{j-1} + dp_j) h:
      dt_j <= dt_{j-1} + (n/c) h (sqrt(3) (dp_{j-1} + dp_j) + R s / h) + R t_j
* the emitted states differ only by rounding: dp, dd <= R and dt <= R t.

Every ratio measured / bound is <= 1 for a difference that is rounding
carried forward; a difference in the physics (a term, a sign, a different
formula in one build) shows up as a step whose difference jumps past what
the step's geometry allows.
"""
import numpy as np

R = 1e-6
C_MM_PER_NS = 299.792458  # speed of light, mm / ns (the reference's units.hpp)
READOUT_FACE = 5           # +z
LAMBERTIAN = 3             # Surface.LAMBERTIAN


def same_discrete(a, b):
    return ((a["fate"] == b["fate"]) & (a["n_boundary"] == b["n_boundary"])
            & (a["n_scatter"] == b["n_scatter"]))


def relative_differences(a, b, p):
    """Per photon: the arrival-time difference relative to the time, and the
    x, y and z differences relative to the largest box half-size h (a hit
    position's scale: coordinates run over [-h, h])."""
    h = max(p.hx, p.hy, p.hz)
    out = {"t": np.abs(a["t"].astype(np.float64) - b["t"]) / np.maximum(np.abs(b["t"]), 1e-30)}
    for k in "xyz":
        out[k] = np.abs(a[k].astype(np.float64) - b[k]) / h
    return out


def check_agreement(a, b, p, tol=1e-5):
    """Records of two builds: discrete histories identical for >= 99.99 % of
    photons; among those, arrival times within `tol` relative for >= 99.9 %
    and each hit coordinate within `tol` of h for >= 99 %. Returns the
    same-history photons outside `tol` in any of them (to be traced)."""
    same = same_discrete(a, b)
    assert same.mean() >= 0.9999, same.mean()
    d = relative_differences(a, b, p)
    assert (d["t"][same] <= tol).mean() >= 0.999, np.quantile(d["t"][same], 0.999)
    for k in "xyz":
        assert (d[k][same] <= tol).mean() >= 0.99, (k, np.quantile(d[k][same], 0.99))
    worst = np.maximum.reduce([d[k] for k in "txyz"])
    return np.nonzero(same & (worst > tol))[0]


def carried_rounding_ratios(p, emitted_a, emitted_b, state_a, state_b, codes):
    """Largest measured / bound ratio of one photon's two traces, for the
    emitted state and for position, direction and time over the steps (see
    the module docstring). `codes` are the (identical) step codes."""
    h = max(p.hx, p.hy, p.hz)
    A = np.vstack([emitted_a, state_a]).astype(np.float64)
    B = np.vstack([emitted_b, state_b]).astype(np.float64)
    dp = np.abs(A[:, :3] - B[:, :3]).max(axis=1) / h
    dd = np.abs(A[:, 3:6] - B[:, 3:6]).max(axis=1)
    dt = np.abs(A[:, 6] - B[:, 6])
    first = max(dp[0] / R, dd[0] / R, dt[0] / max(R * abs(A[0, 6]), 1e-30))
    if len(codes) == 0:
        return first, 0.0, 0.0, 0.0
    c = np.asarray(codes, np.int64) & 0x7F
    survived = (np.asarray(codes, np.int64) & 0x80) != 0
    d_in = A[:-1, 3:6]
    s = np.linalg.norm(A[1:, :3] - A[:-1, :3], axis=1) / h
    wall = c >= 2
    axis = np.where(wall, (c - 2) // 2, 0)
    cos_i = np.abs(d_in[np.arange(len(c)), axis])
    amp = np.where(wall, 1.0 + 1.0 / np.maximum(cos_i, 1e-12), 1.0)
    bp = amp * (dp[:-1] + (dd[:-1] + R) * s) + R
    cos_t = np.clip((d_in * A[1:, 3:6]).sum(axis=1), -1.0, 1.0)
    sin_t = np.sqrt(np.maximum(1.0 - cos_t * cos_t, 0.0))
    lambert = wall & survived & (c - 2 != READOUT_FACE) & (p.surface == LAMBERTIAN)
    bd = np.where(c == 1, 4.0 * dd[:-1] + R * (1.0 + np.abs(cos_t) / np.maximum(sin_t, 1e-12)),
                  np.where(lambert, R, dd[:-1] + R))
    v = p.n_scint / C_MM_PER_NS
    bt = (dt[:-1] + v * h * (np.sqrt(3.0) * (dp[:-1] + dp[1:]) + R * s)
          + R * np.abs(A[1:, 6]))
    return first, float((dp[1:] / bp).max()), float((dd[1:] / bd).max()), float((dt[1:] / bt).max())


def first_difference(port, libm, p, i):
    """(step, start drift, drift before it, margin) of the first step whose
    outcome differs between the builds' traces of photon i; drifts are the
    largest position (/ h) or direction difference."""
    ca, ma, sa = port.step_trace(p, i, with_state=True)
    cb, mb, sb = libm.step_trace(p, i, with_state=True)
    k = 0
    while k < min(len(ca), len(cb)) and ca[k] == cb[k]:
        k += 1
    assert k < max(len(ca), len(cb)), "step traces agree"
    h = max(p.hx, p.hy, p.hz)

    def drift(j):
        return max(np.abs(sa[j, :3] - sb[j, :3]).max() / h, np.abs(sa[j, 3:6] - sb[j, 3:6]).max())

    margin = min(ma[k] if k < len(ma) else np.inf, mb[k] if k < len(mb) else np.inf)
    return k, drift(0), (drift(k - 1) if k > 0 else 0.0), float(margin)


def check_divergence(port, libm, p, i):
    """A photon whose step outcomes differ: the builds start a rounding apart
    and the first differing decision was a near tie, within 10x the drift
    carried to it."""
    k, d0, dk, margin = first_difference(port, libm, p, i)
    assert d0 <= 1e-5, (i, d0)
    if k == 0:
        assert margin <= 1e-5, (i, margin)
    else:
        assert margin <= 10.0 * max(dk, 1e-7), (i, k, dk, margin)


def check_carried_rounding(port, libm, p, i):
    """Trace photon i in both builds: identical step outcomes and every
    difference within the carried-rounding bound, or, where the outcomes
    differ (same fate and counts, another face order), a near-tie divergence.
    Returns True when the outcomes were identical."""
    ca, _, sa = port.step_trace(p, i, with_state=True)
    cb, _, sb = libm.step_trace(p, i, with_state=True)
    if not np.array_equal(ca, cb):
        check_divergence(port, libm, p, i)
        return False
    r = carried_rounding_ratios(p, port.emitted_state(p, i), libm.emitted_state(p, i), sa, sb, ca)
    assert max(r) <= 1.0, (i, r)
    return True
