"""The time-series division of the value-table flagger path.

The device forms y = x / ts[t] as (float)((double)x * rts[t]) with
rts[t] = 1 / (double)ts[t] (flagger.hip `tsdiv`), while the oracle and the
reference (ts_correct, src/hella/src/timeseries.cu:19-74) divide in single
precision.  The two agree bit for bit: the double product is within 2^-52 of
the quotient, and a quotient of two floats is never within 2^-49 of a
rounding boundary of float.  These checks pin that claim on the CPU (numpy's
float32 division is IEEE division) over random bit patterns, flagger-like
values, quotients placed next to rounding midpoints, and the special values;
the last check shows that a single-precision reciprocal would be caught.
"""
import numpy as np

F32 = np.float32


def recip_div(x, ts):
    with np.errstate(all="ignore"):
        r = 1.0 / ts.astype(np.float64)
        return (x.astype(np.float64) * r).astype(F32)


def ieee_div(x, ts):
    with np.errstate(all="ignore"):
        return (x / ts).astype(F32)


def assert_same(a, b):
    a = np.asarray(a, F32)
    b = np.asarray(b, F32)
    na, nb = np.isnan(a), np.isnan(b)
    assert np.array_equal(na, nb)
    ia, ib = a.view(np.uint32)[~na], b.view(np.uint32)[~nb]
    bad = np.nonzero(ia != ib)[0]
    assert bad.size == 0, (a[~na][bad[:5]], b[~nb][bad[:5]])


def test_random_bit_patterns():
    rng = np.random.default_rng(11)
    n = 1 << 22
    x = rng.integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32).view(F32)
    ts = rng.integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32).view(F32)
    assert_same(recip_div(x, ts), ieee_div(x, ts))


def test_flagger_like_values():
    # table values near 1 (bandpass-divided), time series near 1
    rng = np.random.default_rng(12)
    n = 1 << 22
    x = (1.0 + 0.3 * rng.standard_normal(n)).astype(F32)
    ts = (1.0 + 0.01 * rng.standard_normal(n)).astype(F32)
    assert_same(recip_div(x, ts), ieee_div(x, ts))
    # every float in [1, 2) over a few divisors: exhaustive in x
    x = (np.arange(1 << 23, dtype=np.uint32) + np.uint32(0x3F800000)).view(F32)
    for d in (F32(1.0000001), F32(0.99999994), F32(3.0), F32(1.1), F32(0.7)):
        ts = np.full(x.shape, d, F32)
        assert_same(recip_div(x, ts), ieee_div(x, ts))


def test_quotients_next_to_midpoints():
    # x = the floats nearest m * ts for midpoints m of adjacent floats, so
    # x / ts lands as close to a rounding boundary as floats allow
    rng = np.random.default_rng(13)
    n = 1 << 20
    ts = rng.uniform(0.5, 2.0, n).astype(F32)
    q = rng.uniform(0.5, 2.0, n).astype(F32)
    m = (q.astype(np.float64) + np.nextafter(q, F32(np.inf)).astype(np.float64)) / 2
    x0 = (m * ts.astype(np.float64)).astype(F32)
    for k in range(-3, 4):
        x = x0.view(np.uint32).astype(np.int64) + k
        x = x.astype(np.uint32).view(F32)
        assert_same(recip_div(x, ts), ieee_div(x, ts))


def test_special_values():
    vals = np.array([0.0, -0.0, np.inf, -np.inf, np.nan, 1.0, -1.0,
                     np.finfo(F32).max, -np.finfo(F32).max, np.finfo(F32).tiny,
                     np.finfo(F32).smallest_subnormal, 2.0 ** -140, 3.0e38, 1.0e-38,
                     0.5, 2.0, 7.0], F32)
    x, ts = np.meshgrid(vals, vals)
    assert_same(recip_div(x.ravel(), ts.ravel()), ieee_div(x.ravel(), ts.ravel()))
    # quotients at the overflow threshold and in the subnormal range
    rng = np.random.default_rng(14)
    n = 1 << 20
    big = rng.uniform(1.0, 2.0, n).astype(F32) * F32(2.0 ** 127)
    small = rng.uniform(0.5, 1.0, n).astype(F32)
    assert_same(recip_div(big, small), ieee_div(big, small))
    tiny = (rng.integers(1, 1 << 23, n).astype(np.uint32)).view(F32)  # subnormals
    d = rng.uniform(0.5, 4.0, n).astype(F32)
    assert_same(recip_div(tiny, d), ieee_div(tiny, d))
    x = rng.uniform(1.0, 2.0, n).astype(F32) * F32(2.0 ** -120)
    d = rng.uniform(1.0, 2.0, n).astype(F32) * F32(2.0 ** 20)
    assert_same(recip_div(x, d), ieee_div(x, d))


def test_float_reciprocal_would_be_caught():
    rng = np.random.default_rng(15)
    n = 1 << 16
    x = (1.0 + 0.3 * rng.standard_normal(n)).astype(F32)
    ts = (1.0 + 0.01 * rng.standard_normal(n)).astype(F32)
    with np.errstate(all="ignore"):
        wrong = x * (F32(1.0) / ts)
    assert np.count_nonzero(wrong.view(np.uint32) != ieee_div(x, ts).view(np.uint32)) > 0
