"""Closed-form and quadrature expectations for the transport (numpy, float64).

Restated from the mathematics the reference's include/optphot/analytic.hpp
uses: the rectangle solid angle by inclusion-exclusion of
F(x, y) = atan(x y / (h sqrt(x^2 + y^2 + h^2))) (analytic.hpp:22-27), the
attenuated solid angle (1/4pi) Int exp(-r/L) h dA / r^3 by tensor Gauss-Legendre
quadrature (analytic.hpp:85-119), and the mirror-box escape cone
1 - sqrt(1 - (n_det/n_scint)^2) (analytic.hpp:135-139).
"""
import math

import numpy as np


def rect_solid_angle(x1, x2, y1, y2, h):
    def F(x, y):
        return np.arctan(x * y / (h * np.sqrt(x * x + y * y + h * h)))
    return F(x2, y2) - F(x1, y2) - F(x2, y1) + F(x1, y1)


def face_frame(p, face):
    half = (p.hx, p.hy, p.hz)
    src = (p.src_x, p.src_y, p.src_z)
    a = face >> 1
    s = 1.0 if face & 1 else -1.0
    ua, va = (a + 1) % 3, (a + 2) % 3
    return (half[a] - s * src[a], -half[ua] - src[ua], half[ua] - src[ua], -half[va] - src[va],
            half[va] - src[va])


def face_solid_angle_fraction(p, face):
    h, u1, u2, v1, v2 = face_frame(p, face)
    return rect_solid_angle(u1, u2, v1, v2, h) / (4.0 * math.pi)


def face_survival_fraction(p, face, L, panels=24, order=16):
    h, u1, u2, v1, v2 = face_frame(p, face)
    gx, gw = np.polynomial.legendre.leggauss(order)

    def edges(a, b):
        cuts = [a, 0.0, b] if a < 0.0 < b else [a, b]
        e = []
        for k in range(len(cuts) - 1):
            e.extend(cuts[k] + (cuts[k + 1] - cuts[k]) * np.arange(panels) / panels)
        e.append(b)
        return np.array(e)

    eu, ev = edges(u1, u2), edges(v1, v2)
    cu, hu = 0.5 * (eu[1:] + eu[:-1]), 0.5 * (eu[1:] - eu[:-1])
    cv, hv = 0.5 * (ev[1:] + ev[:-1]), 0.5 * (ev[1:] - ev[:-1])
    u = (cu[:, None] + hu[:, None] * gx[None, :]).ravel()
    wu = (hu[:, None] * gw[None, :]).ravel()
    v = (cv[:, None] + hv[:, None] * gx[None, :]).ravel()
    wv = (hv[:, None] * gw[None, :]).ravel()
    r = np.sqrt(u[:, None] ** 2 + v[None, :] ** 2 + h * h)
    att = 1.0 if math.isinf(L) else np.exp(-r / L)
    return float((wu[:, None] * wv[None, :] * h / r ** 3 * att).sum() / (4.0 * math.pi))


def box_survival_fraction(p, L):
    return sum(face_survival_fraction(p, f, L) for f in range(6))


def mirror_box_escape_fraction(n_scint, n_det):
    if n_det >= n_scint:
        return 1.0
    s = n_det / n_scint
    return 1.0 - math.sqrt(1.0 - s * s)


def binomial_pull(k, n, p):
    return (k - n * p) / math.sqrt(n * p * (1.0 - p))


def binomial_p(k, n, p):
    """Two-sided p-value of k successes in n trials of probability p (normal
    approximation; n p (1 - p) is in the thousands in every use)."""
    return math.erfc(abs(binomial_pull(k, n, p)) / math.sqrt(2.0))


def face_integral(p, face, integrand, panels=24, order=16):
    """(1/4pi) Int over `face` of integrand(r, cos_i) dOmega for a point source,
    dOmega = h dA / r^3 and cos_i = h / r (the angle to the face normal), by the
    same composite Gauss-Legendre rule as face_survival_fraction."""
    h, u1, u2, v1, v2 = face_frame(p, face)
    gx, gw = np.polynomial.legendre.leggauss(order)

    def edges(a, b):
        cuts = [a, 0.0, b] if a < 0.0 < b else [a, b]
        e = []
        for k in range(len(cuts) - 1):
            e.extend(cuts[k] + (cuts[k + 1] - cuts[k]) * np.arange(panels) / panels)
        e.append(b)
        return np.array(e)

    eu, ev = edges(u1, u2), edges(v1, v2)
    cu, hu = 0.5 * (eu[1:] + eu[:-1]), 0.5 * (eu[1:] - eu[:-1])
    cv, hv = 0.5 * (ev[1:] + ev[:-1]), 0.5 * (ev[1:] - ev[:-1])
    u = (cu[:, None] + hu[:, None] * gx[None, :]).ravel()
    wu = (hu[:, None] * gw[None, :]).ravel()
    v = (cv[:, None] + hv[:, None] * gx[None, :]).ravel()
    wv = (hv[:, None] * gw[None, :]).ravel()
    r = np.sqrt(u[:, None] ** 2 + v[None, :] ** 2 + h * h)
    return float((wu[:, None] * wv[None, :] * h / r ** 3 * integrand(r, h / r)).sum()
                 / (4.0 * math.pi))


def fresnel_R(n1, n2, ci):
    """Unpolarised Fresnel reflectance (float64, vectorised), angle form."""
    ci = np.asarray(ci, np.float64)
    s2t = (n1 / n2) ** 2 * (1.0 - ci * ci)
    out = np.ones_like(ci)
    ok = s2t < 1.0
    ct = np.sqrt(1.0 - s2t[ok])
    c = ci[ok]
    rs = (n1 * c - n2 * ct) / (n1 * c + n2 * ct)
    rp = (n2 * c - n1 * ct) / (n2 * c + n1 * ct)
    out[ok] = 0.5 * (rs * rs + rp * rp)
    return out


def mirror_box_tiles(p, max_tiles=400):
    """Specular box, matched readout, no bulk processes, point source: unfold
    the five mirrors. A photon going up (d_z > 0) meets the readout plane
    z = hz at unfolded (X, Y); going down it first reflects off z = -hz and
    meets the image plane at distance src_z + 3 hz. Tile (i, j) of that plane,
    X in [(2i-1) hx, (2i+1) hx] and Y likewise, is reached through |i| x-walls
    and |j| y-walls. Returns (k, solid-angle fraction) arrays over the tiles of
    both planes, k = |i| + |j| (+1 for the downward plane), each fraction the
    exact rectangle solid angle / 4 pi."""
    i = np.arange(-max_tiles, max_tiles + 1)
    xe = np.concatenate([(2 * i - 1) * p.hx, [(2 * max_tiles + 1) * p.hx]]) - p.src_x
    ye = np.concatenate([(2 * i - 1) * p.hy, [(2 * max_tiles + 1) * p.hy]]) - p.src_y
    k_xy = np.abs(i)[:, None] + np.abs(i)[None, :]
    ks, fr = [], []
    for h, kz in ((p.hz - p.src_z, 0), (p.src_z + 3.0 * p.hz, 1)):
        F = np.arctan(xe[:, None] * ye[None, :]
                      / (h * np.sqrt(xe[:, None] ** 2 + ye[None, :] ** 2 + h * h)))
        omega = F[1:, 1:] - F[:-1, 1:] - F[1:, :-1] + F[:-1, :-1]
        ks.append((k_xy + kz).ravel())
        fr.append(omega.ravel() / (4.0 * math.pi))
    return np.concatenate(ks), np.concatenate(fr)


def mirror_box_detection(p, R, max_tiles=400):
    """(P_det, dP_det/dR) = (<R^k>, <k R^(k-1)>) over directions, from
    mirror_box_tiles: every direction that reaches the readout does so after
    exactly k reflector hits, each survived with probability R."""
    k, f = mirror_box_tiles(p, max_tiles)
    return float((f * R ** k).sum()), float((f * k * R ** np.maximum(k - 1, 0)).sum())


def mirror_box_reflections(p, d, limit=100000):
    """Reflector hits before the photon from the source with direction d
    (float64) reaches z = +hz, by tracing the mirror reflections one by one."""
    pos = np.array([p.src_x, p.src_y, p.src_z], np.float64)
    d = np.array(d, np.float64)
    half = np.array([p.hx, p.hy, p.hz], np.float64)
    hits = 0
    while hits < limit:
        t = np.full(3, np.inf)
        for a in range(3):
            if d[a] > 0:
                t[a] = (half[a] - pos[a]) / d[a]
            elif d[a] < 0:
                t[a] = (-half[a] - pos[a]) / d[a]
        a = int(np.argmin(t))
        pos = pos + t[a] * d
        if a == 2 and d[2] > 0:
            return hits
        d[a] = -d[a]
        hits += 1
    return hits


def _crossings(s0, d, dist, h):
    """Walls at odd multiples of h crossed by x = s0 + d t, 0 < t < dist
    (s0 inside the box, so tile 0)."""
    x = s0 + d * dist
    return np.where(x >= s0, np.floor((x + h) / (2.0 * h)), np.floor((h - x) / (2.0 * h)))


def _fresnel_and_index_derivatives(n1, n2, mu):
    """R, T = 1 - R (cancellation-free) and dR/dn1, dR/dn2 at cos_i = mu
    (float64, mu above the critical cosine)."""
    s2i = 1.0 - mu * mu
    ct = np.sqrt(1.0 - (n1 / n2) ** 2 * s2i)
    A, B = n1 * mu - n2 * ct, n1 * mu + n2 * ct
    C, D = n2 * mu - n1 * ct, n2 * mu + n1 * ct
    rs, rp = A / B, C / D
    R = 0.5 * (rs * rs + rp * rp)
    T = 2.0 * n1 * n2 * mu * ct * (1.0 / B ** 2 + 1.0 / D ** 2)
    dct1 = -n1 * s2i / (n2 * n2 * ct)
    dct2 = n1 * n1 * s2i / (n2 ** 3 * ct)
    out = []
    for dA, dB, dC, dD in ((mu - n2 * dct1, mu + n2 * dct1, -ct - n1 * dct1, ct + n1 * dct1),
                           (-ct - n2 * dct2, ct + n2 * dct2, mu - n1 * dct2, mu + n1 * dct2)):
        out.append(rs * (dA * B - A * dB) / B ** 2 + rp * (dC * D - C * dD) / D ** 2)
    return R, T, out[0], out[1]


def _mirror_box_readout_density(p, R_w, mu, n_phi):
    """For |d_z| = mu (array above the critical cosine): the detection
    probability and its n_det, n_scint derivatives, integrated over the
    azimuth and summed over both hemispheres, per unit of |d_z| with the
    isotropic measure dOmega / 4pi = d(d_z) dphi / 4pi. The walls keep |d_z|,
    so every readout arrival of a photon has cos_i = mu: it transmits with
    T = 1 - R_F or reflects; arrival m of a photon going up lies at z-distance
    (hz - src_z) + 4 m hz after m hits of the -z wall, going down at
    (src_z + 3 hz) + 4 m hz after m + 1; side walls as the unfolded line
    crosses them. Per direction
      F = sum_m T R_F^m R_w^K_m,  dF/dn = sum_m dR_F/dn R_F^(m-1) (m T - R_F) R_w^K_m."""
    n1, n2 = float(p.n_scint), float(p.n_det)
    mu = np.asarray(mu, np.float64)
    R, T, dR1, dR2 = _fresnel_and_index_derivatives(n1, n2, mu)
    phi = (np.arange(n_phi) + 0.5) / n_phi * 2.0 * math.pi
    st = np.sqrt(1.0 - mu * mu)
    dx = st[:, None] * np.cos(phi)[None, :]
    dy = st[:, None] * np.sin(phi)[None, :]
    F = np.zeros_like(mu)
    g_det = np.zeros_like(mu)
    g_sc = np.zeros_like(mu)
    for up in (True, False):
        h0 = (p.hz - p.src_z) if up else (p.src_z + 3.0 * p.hz)
        kz0 = 0 if up else 1
        Rm, Rm1 = np.ones_like(mu), np.zeros_like(mu)
        m = 0
        while True:
            dist = ((h0 + 4.0 * m * p.hz) / mu)[:, None]
            K = (_crossings(p.src_x, dx, dist, p.hx) + _crossings(p.src_y, dy, dist, p.hy)
                 + kz0 + m)
            wall = (R_w ** K).mean(axis=1) * 0.5  # (1/4pi) * 2pi
            F += T * Rm * wall
            core = m * T * Rm1 - Rm
            g_det += dR2 * core * wall
            g_sc += dR1 * core * wall
            if float((Rm * wall).max()) < 1e-15:
                break
            Rm1, Rm = Rm, Rm * R
            m += 1
    return F, g_det, g_sc


def mirror_box_fresnel_readout(p, R_w, n_v=2000, n_phi=2000):
    """Specular walls of reflectivity R_w, a Fresnel readout (n_scint ->
    n_det), no bulk processes, point source (see _mirror_box_readout_density):
    P_det and its n_det, n_scint derivatives, integrated over |d_z| in
    (cos_c, 1] (below the critical cosine cos_c total internal reflection
    keeps the photon forever) with |d_z| = cos_c + (1 - cos_c) v^2, midpoint
    rule in v. Returns (P_det, dP/dn_det, dP/dn_scint)."""
    n1, n2 = float(p.n_scint), float(p.n_det)
    cos_c = math.sqrt(max(0.0, 1.0 - (n2 / n1) ** 2)) if n2 < n1 else 0.0
    v = (np.arange(n_v) + 0.5) / n_v
    mu = cos_c + (1.0 - cos_c) * v * v
    jac = 2.0 * (1.0 - cos_c) * v / n_v
    out = np.zeros(3)
    for i in range(0, n_v, 250):
        F, gd, gs = _mirror_box_readout_density(p, R_w, mu[i:i + 250], n_phi)
        out += [float((jac[i:i + 250] * q).sum()) for q in (F, gd, gs)]
    return tuple(out)


def mirror_box_emission_grid_correction(p, R_w, window=2e-4, n_phi=2000, n_v=4000):
    """The point source draws cos(theta) = 2u - 1 with u on the 2^-24 grid,
    so |d_z|, kept by the walls up to the readout, takes values on the grid
    k 2^-23. The derivative of the simulated (quantised) model at fixed
    indices is then a Riemann sum over those values of a density that grows
    like 1/sqrt(|d_z| - cos_c) at the critical cosine; it differs from the
    smooth derivative (the limit of finite differences over steps that move
    cos_c across many grid cells) by this amount, taken over the singular
    window [cos_c, cos_c + window] (the grid's error is O(2^-23) beyond it).
    Returns (d/dn_det, d/dn_scint) corrections: quantised minus smooth."""
    n1, n2 = float(p.n_scint), float(p.n_det)
    c = math.sqrt(1.0 - (n2 / n1) ** 2)
    d = 2.0 ** -23
    grid = np.arange(math.ceil(c / d), math.floor((c + window) / d) + 1) * d
    riem = np.zeros(2)
    for i in range(0, len(grid), 200):
        _, gd, gs = _mirror_box_readout_density(p, R_w, grid[i:i + 200], n_phi)
        riem += [float(gd.sum()) * d, float(gs.sum()) * d]
    hi = grid[-1] + d / 2  # the grid cells cover [c, hi]: the first cell starts below its point
    v = (np.arange(n_v) + 0.5) / n_v
    mu = c + (hi - c) * v * v
    integ = np.zeros(2)
    for i in range(0, n_v, 400):
        _, gd, gs = _mirror_box_readout_density(p, R_w, mu[i:i + 400], n_phi)
        w = 2.0 * (hi - c) * v[i:i + 400] / n_v
        integ += [float((w * gd).sum()), float((w * gs).sum())]
    return tuple(riem - integ)
