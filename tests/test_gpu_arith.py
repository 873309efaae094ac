"""Bitwise checks of the FP64 divisions and square root of the strict kernels.

The strict FP64 path replaces every IEEE division by div_rn (a correctly rounded quotient from a
double-word reciprocal, Markstein's theorem), and compute_link_q's division by 6 and the IBB
momentum term's division by RT by constant reciprocal pairs.  rbtp_arith_probe evaluates those
same inline functions on the device; here they must equal the host's IEEE division and sqrt bit
for bit over 10^7 operands per operation: random operands over a wide exponent range, operands
whose quotient (or root) lies within one part in 2^53 of a rounding midpoint (the cases that
break a faithful but not correctly rounded method), the value ranges the kernels see, and
signed zeros.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N_RANDOM = 8_000_000
N_HARD = 1_000_000
N_RANGE = 1_000_000


def random_doubles(rng, n, emin, emax, signed=True):
    m = rng.uniform(1.0, 2.0, n)
    e = rng.integers(emin, emax + 1, n)
    x = np.ldexp(m, e)
    if signed:
        x *= rng.choice([-1.0, 1.0], n)
    return x


def near_midpoint_quotients(rng, b, emin, emax):
    """a such that a / b lies within 2^-53 relative of the midpoint between two doubles."""
    q = random_doubles(rng, b.size, emin, emax)
    half_ulp = (np.nextafter(q, np.inf) - q).astype(np.longdouble) / 2
    return (b.astype(np.longdouble) * (q.astype(np.longdouble) + half_ulp)).astype(np.float64)


def assert_bitwise(got, want, what):
    bad = got.view(np.uint64) != want.view(np.uint64)
    n = int(np.count_nonzero(bad))
    assert n == 0, f"{what}: {n} of {got.size} differ, first at {np.flatnonzero(bad)[:5]}"


def divisor_operands(rng):
    b = np.concatenate([
        random_doubles(rng, N_RANDOM, -100, 100),
        random_doubles(rng, N_HARD, -60, 60),
        # grad_denom = rho/3 + 2 tauc/3 and tau, tauc near 1/2 .. 2
        rng.uniform(0.3, 2.0, N_RANGE),
    ])
    a = np.concatenate([
        random_doubles(rng, N_RANDOM, -200, 200),
        near_midpoint_quotients(rng, b[N_RANDOM:N_RANDOM + N_HARD], -60, 60),
        random_doubles(rng, N_RANGE, -60, 2),
    ])
    a[:1000] = 0.0
    a[1000:2000] = -0.0
    return a, b


@pytest.mark.parametrize("op", ["div", "div_hostrecip"])
def test_division_is_ieee(rb, op):
    rng = np.random.default_rng(1 if op == "div" else 2)
    a, b = divisor_operands(rng)
    with np.errstate(all="ignore"):
        want = a / b
    assert_bitwise(rb.arith_probe(op, a, b), want, op)


@pytest.mark.parametrize("op,divisor", [("div6", 6.0), ("div_rt", 1.0 / 3.0)])
def test_constant_division_is_ieee(rb, op, divisor):
    rng = np.random.default_rng(3 if op == "div6" else 4)
    d = np.full(N_HARD, divisor)
    a = np.concatenate([
        random_doubles(rng, N_RANDOM, -1000, 1000),
        near_midpoint_quotients(rng, d, -200, 200),
        # link_q numerators -b - sqrt(disc) (|b| <= 2 sqrt 3 (R + 2)) and IBB momentum terms
        random_doubles(rng, N_RANGE, -40, 6),
    ])
    a[:1000] = 0.0
    a[1000:2000] = -0.0
    want = a / divisor
    assert_bitwise(rb.arith_probe(op, a), want, op)


def test_sqrt_is_ieee(rb):
    rng = np.random.default_rng(5)
    # roots within 2^-53 of a midpoint: squares of midpoints, rounded
    r = random_doubles(rng, N_HARD, -400, 400, signed=False)
    mid = r.astype(np.longdouble) + (np.nextafter(r, np.inf) - r).astype(np.longdouble) / 2
    a = np.concatenate([
        random_doubles(rng, N_RANDOM, -1000, 1000, signed=False),
        (mid * mid).astype(np.float64),
        # discriminants of compute_link_q and squared pair distances
        rng.uniform(0.0, 400.0, N_RANGE),
    ])
    a[:1000] = 0.0
    a[1000:2000] = -0.0
    assert_bitwise(rb.arith_probe("sqrt", a), np.sqrt(a), "sqrt")
