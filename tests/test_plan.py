"""Plan (A1): delay table, DM list, max delay -- oracle pinned by closed forms
and invariants, library plan checked bit-exact against the oracle.  CPU only:
the plan is host code."""
import numpy as np
import pytest

# (nch, dt, f0, df, dm_lo, dm_hi, ti_us, tol): the two single-beam workloads
GEOMS = {
    "c1_256ch": (256, 256e-6, 500.0, -0.390625, 0.0, 500.0, 64.0, 1.25),
    "c2_1024ch": (1024, 131.072e-6, 500.0, -0.09765625, 0.0, 2000.0, 64.0, 1.25),
    "hella_3072ch": (3072, 1e-3, 468.75, -0.03075, 0.0, 100.0, 40.0, 1.3),
}


@pytest.mark.parametrize("name", list(GEOMS))
def test_delay_table_closed_form(orc, name):
    nch, dt, f0, df = GEOMS[name][:4]
    d = orc.delay_table(nch, dt, f0, df).astype(np.float64)
    f = f0 + np.arange(nch) * df
    ref = 4.15e3 / dt * (1.0 / f**2 - 1.0 / f0**2)
    assert d[0] == 0.0
    assert np.all(np.diff(d) > 0), "delay must grow towards low frequency"
    # a*a - b*b is formed in float: its absolute error is a few ulp of 1/f^2
    atol = 4.15e3 / dt * 4 * 2.0**-24 / f.min() ** 2
    np.testing.assert_allclose(d, ref, rtol=1e-6, atol=atol)


@pytest.mark.parametrize("name", list(GEOMS))
def test_dm_list_invariants(orc, name):
    nch, dt, f0, df, lo, hi, ti, tol = GEOMS[name]
    dms = orc.dm_list(lo, hi, dt, ti, f0, df, nch, tol).astype(np.float64)
    assert dms[0] == lo
    assert dms[-1] >= hi and dms[-2] < hi
    assert np.all(np.diff(dms) > 0)
    # every step solves  a^2 x^2 + b^2 (x - p)^2 = c + tol^2 a^2 p^2  (smearing tolerance)
    f = (f0 + (nch // 2 - 0.5) * df) * 1e-3
    a2 = (8.3 * df / f**3) ** 2
    b2 = a2 * nch * nch / 16.0
    c = ((dt * 1e6) ** 2 + ti**2) * (tol**2 - 1.0)
    p, x = dms[:-1], dms[1:]
    # the larger root of the tolerance quadratic, solved independently with numpy
    A, B, Cq = a2 + b2, -2.0 * b2 * p, b2 * p * p - (c + tol**2 * a2 * p * p)
    root = (-B + np.sqrt(B * B - 4 * A * Cq)) / (2 * A)
    assert np.all(np.abs(A * root**2 + B * root + Cq) <= 1e-6 * (b2 * p * p + c + 1.0))
    # each stored DM is that root rounded to float (1 ulp slack for the
    # different but equivalent algebra)
    ulp = np.spacing(root.astype(np.float32)).astype(np.float64)
    assert np.all(np.abs(x - root) <= ulp)


def test_production_scale_numbers(orc):
    """Order-of-magnitude pins quoted for the 256- and 1024-channel workloads."""
    p1 = orc.Plan(*GEOMS["c1_256ch"])
    p2 = orc.Plan(*GEOMS["c2_1024ch"])
    assert abs(p1.dms.size - 430) <= 5 and abs(p1.max_delay - 18332) <= 200
    assert abs(p2.dms.size - 1889) <= 20 and abs(p2.max_delay - 142647) <= 1500


@pytest.mark.parametrize("name", list(GEOMS))
def test_max_delay_and_delays(orc, name):
    p = orc.Plan(*GEOMS[name])
    assert p.max_delay == p.delta.max() == p.delta[-1, -1]
    assert np.all(np.diff(p.delta, axis=0) >= 0), "delay monotone in DM"
    assert np.all(p.delta[:, 0] == 0)
    # round half up of the float product
    prod = (p.dms[:, None].astype(np.float32) * p.delay[None, :]).astype(np.float64)
    np.testing.assert_array_equal(p.delta, np.floor(prod + 0.5).astype(np.int32))


@pytest.mark.parametrize("name", list(GEOMS))
def test_library_plan_matches_oracle(orc, hamd, name):
    nch, dt, f0, df, lo, hi, ti, tol = GEOMS[name]
    p = orc.Plan(*GEOMS[name])
    lp = hamd.DedispPlan(nch, dt, f0, df)
    lp.generate_dm_list(lo, hi, ti, tol)
    np.testing.assert_array_equal(lp.dm_list(), p.dms)
    assert lp.max_delay == p.max_delay
    assert lp.dm_count == p.dms.size


def test_library_plan_errors(hamd):
    with pytest.raises(hamd.DedispError):
        hamd.DedispPlan(0, 1e-3, 400.0, -0.1)
    lp = hamd.DedispPlan(16, 1e-3, 400.0, -1.0)
    with pytest.raises(hamd.DedispError):   # no DM list yet
        lp.execute_adv(np.zeros((64, 16), np.uint8))
    lp.set_dm_list([0.0, 1.0, 2.0])
    with pytest.raises(hamd.DedispError):   # nsamps <= max_delay
        lp.execute_adv(np.zeros((max(1, lp.max_delay), 16), np.uint8))
