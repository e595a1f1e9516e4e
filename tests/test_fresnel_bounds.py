"""The Fresnel cell bounds of the wavefront ring kernel, checked on the CPU.

A forward kernel may settle a readout or polished-face decision u <= R(cos_i)
(fresnel.hpp:30-46, transport.hpp:41) from a per-cell bound on R instead of
evaluating R (photon.hpp: fresnel_cell_bounds). The bound is computed by
evaluating the same float32 operations on the end points of every operand's
range, which is sound because each correctly rounded operation is monotone in
each operand. This file restates that construction in numpy float32 (whose
+, -, *, / and sqrt are correctly rounded, with no contraction), pins the
restated R against the oracle's fresnel_R bit for bit, and then checks
exhaustively, over every float cos_i in [2^-6, 1] and a log-uniform sample
below, that every cell's bounds hold for the float32 R of every cos_i in the
cell, for the index pairs of the workloads and some edge cases. A bound one
unit too tight is caught by the same check. The GPU parity tests then show
that the kernel's decisions equal the oracle's.
"""
import numpy as np
import pytest

F = np.float32
CELLS = 512


def fresnel_prob(n1, n2, c):
    """R as fresnel_prob computes it (photon.hpp), for an array of cos_i."""
    n1, n2 = F(n1), F(n2)
    if n1 == n2:
        return np.zeros_like(c)
    eta = F(n1 / n2)
    eta2 = F(eta * eta)
    with np.errstate(invalid="ignore", divide="ignore"):
        sin2 = eta2 * np.maximum(F(0), F(1) - c * c)
        ct = np.sqrt(F(1) - sin2)
        rs = (n1 * c - n2 * ct) / (n1 * c + n2 * ct)
        rp = (n2 * c - n1 * ct) / (n2 * c + n1 * ct)
        r = F(0.5) * (rs * rs + rp * rp)
    return np.where(sin2 >= F(1), F(1), r).astype(F)


def _quot(alo, ahi, blo, bhi):
    with np.errstate(invalid="ignore", divide="ignore"):
        return (np.fmin(alo / blo, alo / bhi), np.fmax(ahi / blo, ahi / bhi))


def _square(lo, hi):
    l2, h2 = lo * lo, hi * hi
    return (np.where(lo >= 0, l2, np.where(hi <= 0, h2, F(0))),
            np.where(lo >= 0, h2, np.where(hi <= 0, l2, np.fmax(l2, h2))))


def cell_bounds(n1, n2, tight_by=0):
    """fresnel_cell_bounds for every cell, as (lo16, hi16) arrays. `tight_by`
    plants a mistake: both bounds moved inwards by that many units."""
    n1, n2 = F(n1), F(n2)
    j = np.arange(CELLS, dtype=np.uint32)
    if n1 == n2:
        return np.zeros(CELLS, np.int64), np.zeros(CELLS, np.int64)
    eta = F(n1 / n2)
    eta2 = F(eta * eta)
    cl = j.astype(F) * F(1.0 / CELLS)
    ch = (j + 1).astype(F) * F(1.0 / CELLS)
    s_hi = eta2 * np.fmax(F(0), F(1) - cl * cl)
    s_lo = eta2 * np.fmax(F(0), F(1) - ch * ch)
    sn_hi = np.fmin(s_hi, np.nextafter(F(1), F(0)))
    with np.errstate(invalid="ignore"):
        ct_lo, ct_hi = np.sqrt(F(1) - sn_hi), np.sqrt(F(1) - s_lo)
    n1c_lo, n1c_hi, n2c_lo, n2c_hi = n1 * cl, n1 * ch, n2 * cl, n2 * ch
    n1t_lo, n1t_hi, n2t_lo, n2t_hi = n1 * ct_lo, n1 * ct_hi, n2 * ct_lo, n2 * ct_hi
    rs = _quot(n1c_lo - n2t_hi, n1c_hi - n2t_lo, n1c_lo + n2t_lo, n1c_hi + n2t_hi)
    rp = _quot(n2c_lo - n1t_hi, n2c_hi - n1t_lo, n2c_lo + n1t_lo, n2c_hi + n1t_hi)
    s2, p2 = _square(*rs), _square(*rp)
    lo = F(0.5) * (s2[0] + p2[0])
    hi = F(0.5) * (s2[1] + p2[1])
    mixed = s_hi >= F(1)
    # fmin / fmax return the number when one operand is NaN, as fminf / fmaxf do
    lo = np.where(mixed, np.fmin(lo, F(1)), lo)
    hi = np.where(mixed, np.fmax(hi, F(1)), hi)
    bad = np.isnan(lo) | np.isnan(hi)
    fl = np.floor(np.fmin(np.fmax(np.nan_to_num(lo), F(0)), F(1)) * F(65536))
    fh = np.ceil(np.fmin(np.fmax(np.nan_to_num(hi), F(0)), F(1)) * F(65536))
    l16 = np.where(fl >= 65535, 0xFFFF, fl).astype(np.int64)
    h16 = np.where(fh < 1, 0, np.where(fh >= 65536, 0xFFFF, fh - 1)).astype(np.int64)
    l16, h16 = np.where(bad, 0, l16), np.where(bad, 0xFFFF, h16)
    tir = s_lo >= F(1)
    l16, h16 = np.where(tir, 0xFFFF, l16), np.where(tir, 0xFFFF, h16)
    if tight_by:
        l16 = np.minimum(l16 + tight_by, 0xFFFF)
        h16 = np.maximum(h16 - tight_by, 0)
    return l16, h16


def all_floats(lo_exp=-6):
    """Every float32 in [2^lo_exp, 1]."""
    a = np.arange(np.float32(2.0 ** lo_exp).view(np.uint32), np.float32(1).view(np.uint32) + 1,
                  dtype=np.uint32)
    return a.view(F)


def sample_below(lo_exp=-6, n=400_000, seed=3):
    rng = np.random.default_rng(seed)
    return np.exp2(rng.uniform(-40, lo_exp, n)).astype(F)


def violations(n1, n2, tight_by=0, step=1 << 22):
    """Count of cos_i whose float32 R leaves its cell's bounds."""
    l16, h16 = cell_bounds(n1, n2, tight_by)
    grid = all_floats()
    bad = 0
    for k in range(0, grid.size + 1, step):
        c = grid[k:k + step] if k < grid.size else sample_below()
        if c.size == 0:
            continue
        r = fresnel_prob(n1, n2, c).astype(np.float64)
        cell = np.minimum((c * F(CELLS)).astype(np.int64), CELLS - 1)
        lo = l16[cell] / 65536.0
        hi = (h16[cell] + 1) / 65536.0
        bad += int(np.count_nonzero((r < lo) | (r > hi)))
    return bad


PAIRS = [
    (1.63, 1.5),          # readout of the long-histories and plain workloads
    (1.63, 1.0),          # polished faces to air
    (1.63, 1.45),         # readout at the calibration targets
    (1.60, 1.40),         # steep total internal reflection
    (1.40, 1.60),         # into a denser detector: no total internal reflection
    (1.63, 1.6299999),    # one ulp from index-matched
    (1.63, 1.63),         # matched: R = 0
]


def test_restated_R_equals_the_oracle(oracle_portable):
    rng = np.random.default_rng(7)
    c = np.concatenate([rng.random(3000, dtype=np.float32),
                        np.exp2(rng.uniform(-30, 0, 1000)).astype(F)])
    for n1, n2 in PAIRS:
        r = fresnel_prob(n1, n2, c)
        ref = np.array([oracle_portable.fresnel_R(F(n1), F(n2), F(x)) for x in c], dtype=F)
        assert np.array_equal(r.view(np.uint32), ref.view(np.uint32)), (n1, n2)


@pytest.mark.parametrize("n1,n2", PAIRS)
def test_cell_bounds_hold_for_every_cos(n1, n2):
    assert violations(n1, n2) == 0


def test_batch_index_ranges():
    rng = np.random.default_rng(11)
    for n2 in rng.uniform(1.4, 1.6, 4):
        assert violations(1.63, float(F(n2)), step=1 << 23) == 0


def test_a_bound_one_unit_too_tight_is_caught():
    assert violations(1.63, 1.5, tight_by=1) > 0
    assert violations(1.63, 1.0, tight_by=1) > 0


def test_decision_rule_against_direct_comparison():
    """k8 < lo16 means u <= R and k8 > hi16 means u > R, for drawn words."""
    rng = np.random.default_rng(5)
    for n1, n2 in PAIRS[:5]:
        l16, h16 = cell_bounds(n1, n2)
        c = rng.random(200_000, dtype=np.float32)
        c = np.where(c == 0, F(1), c)
        w = rng.integers(0, 2 ** 32, c.size, dtype=np.uint64).astype(np.uint32)
        u = ((w >> 8).astype(np.float64) + 1.0) * 2.0 ** -24
        r = fresnel_prob(n1, n2, c).astype(np.float64)
        cell = np.minimum((c * F(CELLS)).astype(np.int64), CELLS - 1)
        k8 = (w >> 16).astype(np.int64)
        sure_r, sure_t = k8 < l16[cell], k8 > h16[cell]
        assert np.all(u[sure_r] <= r[sure_r])
        assert np.all(u[sure_t] > r[sure_t])
        # most decisions are settled by the bounds
        assert np.mean(sure_r | sure_t) > 0.98
