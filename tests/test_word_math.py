"""Integer forms of the logarithm split and the quarter-turn reduction,
checked on the CPU.

The reference's portable logarithm splits x = m 2^e with m in [sqrt(1/2),
sqrt(2)) (pmath.hpp:62-82), and its sincos(2 pi u) reduces v = 4 u to a
quadrant q and a remainder r in [-1/2, 1/2] (pmath.hpp:87-119). The kernels
compute both with integer operations on the bits (photon.hpp: p_log,
p_sincos_word). Here each integer form is compared with the reference's float
steps, restated in numpy float32 (correctly rounded, no contraction), over
every input the kernels feed it: every float in [2^-24, 1] for the split (the
uniforms of to_unit lie there) and every 24-bit k of a drawn word for the
reduction. A threshold or tie rule off by one is caught, and whole values are
compared with the oracle's portable log and sincos on a sample.
"""
import numpy as np

F = np.float32


def ref_split(bits):
    """(e, bits of m) as pmath.hpp:62-82 computes them."""
    e = ((bits >> 23) & 0xFF).astype(np.int64) - 127
    m = ((bits & 0x7FFFFF) | 0x3F800000).astype(np.uint32).view(F)
    big = m > F(1.41421356)
    m = np.where(big, F(0.5) * m, m).astype(F)
    return e + big, m.view(np.uint32)


def int_split(bits, c=0x3F3504F4):
    d = bits.astype(np.int64) - c
    e = d >> 23
    m = ((d & 0x7FFFFF) + c).astype(np.uint32)
    return e, m


def test_log_split_equals_the_reference_for_every_uniform():
    lo = int(np.float32(2.0 ** -24).view(np.uint32))
    hi = int(np.float32(1.0).view(np.uint32))
    for start in range(lo, hi + 1, 1 << 23):
        bits = np.arange(start, min(start + (1 << 23), hi + 1), dtype=np.uint32)
        e0, m0 = ref_split(bits)
        e1, m1 = int_split(bits)
        assert np.array_equal(e0, e1) and np.array_equal(m0, m1), hex(start)


def test_log_split_over_all_normal_exponents():
    rng = np.random.default_rng(3)
    bits = rng.integers(0x00800000, 0x7F800000, 2_000_000, dtype=np.uint32)
    e0, m0 = ref_split(bits)
    e1, m1 = int_split(bits)
    assert np.array_equal(e0, e1) and np.array_equal(m0, m1)


def test_log_split_threshold_one_unit_off_is_caught():
    bits = np.arange(0x3F800000, 0x3F800000 + (1 << 23), dtype=np.uint32)
    e0, m0 = ref_split(bits)
    e1, m1 = int_split(bits, c=0x3F3504F3)
    assert not np.array_equal(m0, m1)


def _poly_log(e, m):
    s = ((m - F(1)) / (m + F(1))).astype(F)
    s2 = s * s
    poly = F(1) + s2 * (F(1 / 3) + s2 * (F(1 / 5) + s2 * (F(1 / 7) + s2 * (F(1 / 9) + s2 * F(1 / 11)))))
    fe = e.astype(F)
    return (fe * F(6.9313812256e-01) + (F(2) * s * poly + fe * F(9.0580006145e-06))).astype(F)


def test_log_values_equal_the_oracle(oracle_portable):
    rng = np.random.default_rng(4)
    k = rng.integers(0, 1 << 24, 4000, dtype=np.uint32)
    u = ((k.astype(F) + F(1)) * F(2.0 ** -24)).astype(F)
    e, mb = int_split(u.view(np.uint32))
    got = _poly_log(e, mb.view(F))
    ref = np.array([oracle_portable.portable_log(F(x)) for x in u], dtype=F)
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))


def ref_reduce(k):
    """(quadrant, x) of pmath.hpp:87-119 for u = (k + 1) 2^-24."""
    u = ((k.astype(F) + F(1)) * F(2.0 ** -24)).astype(F)
    v = F(4) * u
    q = np.floor(v)
    r = (v - q).astype(F)
    big = r > F(0.5)
    r = np.where(big, r - F(1), r).astype(F)
    q = np.where(big, q + F(1), q)
    return q.astype(np.int64) & 3, (r * F(1.57079632679489662)).astype(F)


def int_reduce(k, tie_bias=(1 << 21) - 1):
    t = k.astype(np.int64) + 1 + tie_bias
    q = t >> 22
    ri = (t & 0x3FFFFF) - tie_bias
    return q & 3, (ri.astype(F) * (F(1.57079632679489662) * F(2.0 ** -22))).astype(F)


def test_quarter_turn_reduction_equals_the_reference_for_every_word():
    for start in range(0, 1 << 24, 1 << 22):
        k = np.arange(start, start + (1 << 22), dtype=np.uint32)
        q0, x0 = ref_reduce(k)
        q1, x1 = int_reduce(k)
        assert np.array_equal(q0, q1), start
        assert np.array_equal(x0.view(np.uint32), x1.view(np.uint32)), start


def test_quarter_turn_tie_rule_is_caught():
    k = np.arange(0, 1 << 24, 7, dtype=np.uint32)
    q0, x0 = ref_reduce(k)
    q1, x1 = int_reduce(k, tie_bias=1 << 21)  # ties up instead of down
    assert not (np.array_equal(q0, q1) and np.array_equal(x0.view(np.uint32), x1.view(np.uint32)))


def test_sincos_values_equal_the_oracle(oracle_portable):
    rng = np.random.default_rng(5)
    k = np.concatenate([rng.integers(0, 1 << 24, 3000, dtype=np.uint32),
                        np.array([0, (1 << 21) - 1, 1 << 21, (1 << 22) - 1, (1 << 24) - 1], np.uint32)])
    q, x = int_reduce(k)
    x2 = x * x
    sp = x * (F(1) + x2 * (F(-1 / 6) + x2 * (F(1 / 120) + x2 * (F(-1 / 5040) + x2 * F(1 / 362880)))))
    cp = F(1) + x2 * (F(-0.5) + x2 * (F(1 / 24) + x2 * (F(-1 / 720) + x2 * (F(1 / 40320) + x2 * F(-1 / 3628800)))))
    s = np.select([q == 0, q == 1, q == 2], [sp, cp, -sp], -cp).astype(F)
    c = np.select([q == 0, q == 1, q == 2], [cp, -sp, -cp], sp).astype(F)
    u = ((k.astype(F) + F(1)) * F(2.0 ** -24)).astype(F)
    ref = np.array([oracle_portable.sincos(F(v), portable=True) for v in u], dtype=F)
    assert np.array_equal(s.view(np.uint32), ref[:, 0].view(np.uint32))
    assert np.array_equal(c.view(np.uint32), ref[:, 1].view(np.uint32))


def test_tiny_component_test_from_bits():
    """(bits << 1) - 2 < (bits(2^-60) << 1) - 2 exactly for 0 < |d| < 2^-60
    (photon.hpp: tiny_component), over zeros, subnormals, the neighbours of
    2^-60 and a log-uniform sample of both signs."""
    rng = np.random.default_rng(6)
    edge = np.float32(2.0 ** -60)
    vals = np.concatenate([
        np.array([0.0, -0.0, 1e-45, -1e-45, 1e-39, np.inf, -np.inf, 1.0, -1.0], F),
        np.array([np.nextafter(edge, F(0)), edge, np.nextafter(edge, F(1))], F),
        (np.exp2(rng.uniform(-149, 0, 200_000)) * rng.choice([-1, 1], 200_000)).astype(F)])
    b = vals.view(np.uint32).astype(np.uint64)
    got = ((b << 1) - 2) % (1 << 32) < ((0x21800000 << 1) - 2)
    want = (np.abs(vals) > 0) & (np.abs(vals) < edge)
    assert np.array_equal(got, want)


def test_sym_unit_equals_two_u_minus_one_for_every_word():
    """One FMA of k 2^-23 and 2^-23 - 1 gives the reference's 2 u - 1 (u the
    uniform of word k, transport.hpp:55-67) for every 24-bit k (photon.hpp:
    sym_unit). The FMA's single rounding of the exact value is taken in
    float64, where k 2^-23 + (2^-23 - 1) is exact."""
    k = np.arange(0, 1 << 24, dtype=np.uint32)
    u = (k.astype(F) * F(2.0 ** -24) + F(2.0 ** -24)).astype(F)
    ref = (F(2) * u - F(1)).astype(F)
    got = (k.astype(np.float64) * 2.0 ** -23 + (2.0 ** -23 - 1.0)).astype(F)
    assert np.array_equal(ref.view(np.uint32), got.view(np.uint32))
    # the constant itself is exact in float32
    assert float(F(2.0 ** -23 - 1.0)) == 2.0 ** -23 - 1.0
