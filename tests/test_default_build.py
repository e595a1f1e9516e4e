"""Agreement with the reference's DEFAULT build (glibc logf / sinf / cosf /
cbrtf, pmath.hpp:134-158), whose last bits no source pins.

The HIP kernels reproduce the reference's portable-math build bit for bit
(test_gpu_parity.py). Against the default build, with the same seed and
photon ids, the only difference is the last-bit rounding of the
transcendental functions. These tests measure what that does, in the
quantities the reference's own comparison reports (same discrete history,
largest time and position differences: src/core/compare.cpp:40-62,
README.md:432-438):

* the discrete history (fate, n_boundary, n_scatter) is identical for
  >= 99.99 % of photons; among those, arrival times agree within 1e-5
  relative for >= 99.9 % and each hit coordinate within 1e-5 of the box
  half-size for >= 99 %;
* every photon outside 1e-5 is traced step by step in both builds (oracle
  step traces, tests/drift.py): its differences stay within a first-order
  bound of rounding carried forward (a direction an ulp apart at emission,
  carried along the path, amplified by grazing walls and small-angle
  scatters), step by step;
* every photon whose discrete history differs, or whose face order does, is
  traced the same way: the states start a rounding apart, and the first
  differing decision comes only where its margin is within 10x the drift
  carried to it. So the divergence is rounding, carried forward, and not a
  physics difference.

The default build is the reference's own CPU path compiled from its sources
(oracle/_ref, libm); the oracle's libm mode is bitwise equal to it
(test_oracle.py) and stands in where it is not built, and for the traces.
"""
import numpy as np
import pytest

from tests import drift
from tests.configs import PARITY

CASES = [("c1_plain", 1_000_000), ("c2_long", 1_000_000), ("wrapped_lambertian", 1_000_000),
         ("every_process", 1_000_000), ("specular", 1_000_000), ("black_scatter", 1_000_000)]


@pytest.fixture(scope="module")
def default_build():
    from oracle.oracle import Oracle, Reference, reference_available
    return Reference("libm") if reference_available("libm") else Oracle("libm")


@pytest.mark.parametrize("name,n", CASES)
def test_default_build_agreement(oracle_portable, oracle_libm, default_build, name, n):
    """Agreement fractions, then every same-history photon outside 1e-5 in
    time or position traced to carried rounding."""
    p = PARITY[name]
    a = oracle_portable.records(p, n)
    b = default_build.records(p, n)
    out = drift.check_agreement(a, b, p)
    for i in out:
        drift.check_carried_rounding(oracle_portable, oracle_libm, p, int(i))


@pytest.mark.parametrize("name,n", [("wrapped_lambertian", 2_000_000), ("c2_long", 2_000_000),
                                    ("specular", 2_000_000), ("c1_plain", 4_000_000)])
def test_divergent_histories_trace_to_rounding(oracle_portable, oracle_libm, name, n):
    p = PARITY[name]
    a = oracle_portable.records(p, n)
    b = oracle_libm.records(p, n)
    for i in np.nonzero(~drift.same_discrete(a, b))[0]:
        drift.check_divergence(oracle_portable, oracle_libm, p, int(i))


def test_carried_rounding_bound_flags_planted_differences(oracle_portable, oracle_libm):
    """The bound is tight enough to see a physics difference: a real traced
    pair passes it, and the same pair with one planted difference -- a hit
    position moved by 1e-4 of h at one wall, a direction turned by 1e-4 at
    one reflection, a time 1e-5 late at one step, an emitted direction 1e-5
    off -- fails it at that step."""
    p = PARITY["wrapped_lambertian"]
    a = oracle_portable.records(p, 200_000)
    b = oracle_libm.records(p, 200_000)
    out = drift.check_agreement(a, b, p)
    i = next(int(k) for k in out
             if len(oracle_portable.step_trace(p, int(k))[0]) >= 12
             and np.array_equal(oracle_portable.step_trace(p, int(k))[0],
                                oracle_libm.step_trace(p, int(k))[0]))
    ca, _, sa = oracle_portable.step_trace(p, i, with_state=True)
    _, _, sb = oracle_libm.step_trace(p, i, with_state=True)
    ea, eb = oracle_portable.emitted_state(p, i), oracle_libm.emitted_state(p, i)
    assert max(drift.carried_rounding_ratios(p, ea, eb, sa, sb, ca)) <= 1.0
    h = max(p.hx, p.hy, p.hz)
    k = 6
    assert (ca[k] & 0x7F) >= 2  # a wall step
    moved = sb.copy()
    moved[k, 0 if abs(sb[k, 0]) != np.float32(p.hx) else 1] += np.float32(1e-4 * h)
    turned = sb.copy()
    r = next(j for j in range(3, len(ca)) if (ca[j] & 0x80) and (ca[j] & 0x7F) >= 2)
    turned[r, 3] += np.float32(1e-4)  # after a reflection off a wall
    late = sb.copy()
    late[k, 6] *= np.float32(1 + 1e-5)
    tilted = eb.copy()
    tilted[3] += np.float32(1e-5)
    ratios = [drift.carried_rounding_ratios(p, ea, eb, sa, moved, ca)[1],
              drift.carried_rounding_ratios(p, ea, eb, sa, turned, ca)[2],
              drift.carried_rounding_ratios(p, ea, eb, sa, late, ca)[3],
              drift.carried_rounding_ratios(p, ea, tilted, sa, sb, ca)[0]]
    assert min(ratios) > 1.0, ratios


def test_trace_records_agree_with_transport(oracle_portable):
    """The step trace runs the same transport: its step count and final state
    match the photon's record."""
    p = PARITY["every_process"]
    rec = oracle_portable.records(p, 64)
    for i in range(64):
        codes, margins, state = oracle_portable.step_trace(p, i, with_state=True)
        nb = int(np.sum((codes & 0x7F) >= 2))
        ns = int(np.sum((codes & 0x7F) == 1))
        assert nb == rec["n_boundary"][i] and ns == rec["n_scatter"][i]
        assert state[-1, 6] == rec["t"][i] and state[-1, 0] == rec["x"][i]
        assert np.all(margins >= 0)
