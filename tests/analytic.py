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
