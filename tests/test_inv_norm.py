"""The normalisation shortcut of rotate_about, checked on the CPU.

The reference normalises a rotated direction with 1 / sqrt(|d|^2), both
correctly rounded (vec3.hpp:35-53). |d|^2 of a rotated unit vector lies within
a few ulps of 1, and for s = 1 + i ulp the kernel takes the two rounded
results from integer operations on the bits of s (photon.hpp:
inv_sqrt_near_one). Here that formula is compared with numpy's float32 square
root and division (correctly rounded) for every |i| <= 2000, beyond the
|i| <= 1024 the kernel uses it for; a tie rule on the wrong side is caught;
and |d|^2 after float32 rotations of random and axis-aligned normals is shown
to stay within a few ulps of 1.
"""
import numpy as np

F = np.float32
ONE = int(np.float32(1).view(np.int32))


def ieee_inv_sqrt(i):
    s = np.array([ONE + i], np.int32).view(F)[0]
    return int(np.array([F(F(1) / np.sqrt(s))]).view(np.int32)[0])


def formula(i, tie_up=False):
    if i >= 0:
        return ONE - (i & ~1)
    return ONE + ((3 - i) >> 2) if not tie_up else ONE + ((2 - i) >> 2)


def test_formula_equals_ieee_for_every_i_within_2000():
    assert all(formula(i) == ieee_inv_sqrt(i) for i in range(-2000, 2001))


def test_a_wrong_tie_rule_is_caught():
    assert any(formula(i, tie_up=True) != ieee_inv_sqrt(i) for i in range(-2000, 0))


def test_rotated_norms_stay_near_one():
    rng = np.random.default_rng(1)
    n = rng.normal(size=(200_000, 3))
    n = (n / np.linalg.norm(n, axis=1)[:, None]).astype(F)
    ct = (rng.random(n.shape[0]) * 2 - 1).astype(F)
    phi = 2 * np.pi * rng.random(n.shape[0])
    nx, ny, nz = n[:, 0], n[:, 1], n[:, 2]
    sign = np.copysign(F(1), nz).astype(F)
    a = (F(-1) / (sign + nz)).astype(F)
    b = (nx * ny * a).astype(F)
    t1 = (F(1) + sign * nx * nx * a, sign * b, -sign * nx)
    t2 = (b, sign + ny * ny * a, -ny)
    st = np.sqrt(np.maximum(F(0), F(1) - ct * ct)).astype(F)
    c1, c2 = st * np.cos(phi).astype(F), st * np.sin(phi).astype(F)
    d = [(c1 * t1[k] + c2 * t2[k]) + ct * n[:, k] for k in range(3)]
    s = ((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]).astype(F)
    i = s.view(np.int32).astype(np.int64) - ONE
    assert np.abs(i).max() <= 16
