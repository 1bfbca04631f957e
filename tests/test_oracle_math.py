"""Pins of the scalar arithmetic the oracle and the kernels share.

* sincos: the FreeBSD-msun double-evaluated float kernels must stay within
  1 ulp of the correctly rounded sin/cos (a wrong coefficient or quadrant
  rule moves thousands of ulps).
* quotient: the kernels replace a float division by a grid or model
  constant b with RN32(RN64(a) * RN64(1/b)); for every divisor the problems
  use, that must equal IEEE a / b for all 2^32 dividends.
* heading wrap: f - floor(f/n3)*n3 must land in [0, n3] wherever the
  reference can index, which is what lets the kernels reject outside it
  without changing any in-domain result.
"""
import math

import numpy as np
import pytest

from conftest import nthreads


def ulp_dist(a, b):
    ai = a.view(np.int32).astype(np.int64)
    bi = b.view(np.int32).astype(np.int64)
    ai = np.where(ai < 0, -(ai & 0x7FFFFFFF), ai)
    bi = np.where(bi < 0, -(bi & 0x7FFFFFFF), bi)
    return np.abs(ai - bi)


@pytest.mark.parametrize("lo,hi", [(-0.8, 0.8), (-7.0, 7.0), (-100.0, 100.0), (-3.0e4, 3.0e4)])
def test_sincos_within_one_ulp(orc, lo, hi):
    rng = np.random.default_rng(11)
    x = rng.uniform(lo, hi, 2_000_000).astype(np.float32)
    s, c = orc.sincosf(x)
    rs = np.sin(x.astype(np.float64)).astype(np.float32)
    rc = np.cos(x.astype(np.float64)).astype(np.float32)
    us, uc = ulp_dist(s, rs), ulp_dist(c, rc)
    assert us.max() <= 1 and uc.max() <= 1
    assert (us == 0).mean() > 0.999 and (uc == 0).mean() > 0.999


def test_sincos_special_values(orc):
    x = np.array([0.0, -0.0, np.float32(math.pi) / 4, -np.float32(math.pi) / 4,
                  np.float32(math.pi), -np.float32(math.pi), 1e-30, 1.5707964], dtype=np.float32)
    s, c = orc.sincosf(x)
    assert np.signbit(s[1]) and s[1] == 0 and c[1] == 1
    assert s[0] == 0 and c[0] == 1
    ref_s = np.sin(x.astype(np.float64)).astype(np.float32)
    ref_c = np.cos(x.astype(np.float64)).astype(np.float32)
    assert ulp_dist(s, ref_s).max() <= 1 and ulp_dist(c, ref_c).max() <= 1
    nan_s, nan_c = orc.sincosf(np.array([np.inf, np.nan], dtype=np.float32))
    assert np.isnan(nan_s).all() and np.isnan(nan_c).all()


def problem_divisors(pg):
    out = set()
    for cfg in (1, 2, 4, 5):
        g = pg.configs.make_grid(pg.configs.SHAPES[cfg])
        out.update(float(v) for v in g.step)
        out.add(float(np.float32(g.n[2])))
    out.update([5.0, float(np.float32(0.15))])
    return sorted(out)


def test_quotient_equals_ieee_division_exhaustively(pg, orc):
    divs = problem_divisors(pg)
    assert len(divs) >= 15
    bad = {b: orc.check_quotient(b, nthreads()) for b in divs}
    assert all(v == 0 for v in bad.values()), bad


def test_quotient_check_detects_a_wrong_reciprocal(orc):
    # A divisor whose float reciprocal is inexact; the check itself must be
    # able to fail: the float-only product a * RN32(1/b) misrounds often.
    b = np.float32(0.1309)
    rb32 = np.float32(1.0) / b
    a = np.random.default_rng(3).uniform(-1e3, 1e3, 200_000).astype(np.float32)
    assert np.count_nonzero(a * rb32 != a / b) > 0
    assert orc.check_quotient(float(b), nthreads()) == 0


@pytest.mark.parametrize("n3", [8, 12, 16, 24, 48, 64])
def test_wrapped_heading_stays_in_range(n3):
    n3f = np.float32(n3)
    rng = np.random.default_rng(n3)
    f = rng.uniform(-5.0 * n3, 5.0 * n3, 4_000_000).astype(np.float32)
    # dense around every multiple of n3 in the range, where rounding bites
    ks = np.arange(-5, 6, dtype=np.float32) * n3f
    base = ks.view(np.int32)
    near = (base[:, None] + np.arange(-4096, 4097, dtype=np.int32)[None, :]).ravel().view(np.float32)
    f = np.concatenate([f, near[np.isfinite(near)], -near[np.isfinite(near)]])
    # f = (h - lo)/step with h and lo floats near pi is 0 or at least
    # ulp(pi)/step in magnitude; only subnormal negatives (|f| < n3 * 2^-149,
    # where f / n3 underflows to -0) wrap below zero, and those are rejected.
    tiny = (f < 0) & (f > -np.float32(n3) * np.float32(2.0 ** -140))
    w = f - np.floor(f / n3f) * n3f
    assert w.dtype == np.float32
    assert (w[~tiny] >= 0).all() and (w[~tiny] <= n3f).all()
