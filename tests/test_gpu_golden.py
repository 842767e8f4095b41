"""The HIP path against the committed golden fixtures (tests/golden)."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    return np.load(os.path.join(HERE, name + ".npz"), allow_pickle=False)


def test_plan_matches_fixture(hamd):
    g = _load("plan_c1")
    lp = hamd.DedispPlan(256, 256e-6, 500.0, -0.390625)
    lp.generate_dm_list(0.0, 500.0, 64.0, 1.25)
    np.testing.assert_array_equal(lp.dm_list(), g["dms"])
    assert lp.max_delay == int(g["max_delay"])


def test_dedisp_matches_fixture(hamd):
    g = _load("dedisp_small")
    lp = hamd.DedispPlan(16, 256e-6, 500.0, -6.25)
    lp.generate_dm_list(0.0, 60.0, 64.0, 1.25)
    np.testing.assert_array_equal(lp.dm_list(), g["dms"])
    lp.set_killmask(g["killmask"])
    got = lp.execute_adv(g["data"])
    np.testing.assert_array_equal(got, g["out"].astype(np.float32))


def test_flagger_matches_fixture(hamd):
    g = _load("flag_small")
    cfg = dict(INPUT="FILE", NCHAN=32, FCH1=500.0, FOFF=-3.125, TSAMP=256e-6, GULP=2048,
               FLAG1=3, FLAG2=5, SPEC_MIN=-0.05, SPEC_MAX=0.15, SCRUNCH=[(1, 1, 5.0)],
               DM_MIN=0, DM_MAX=20, DM_TOL=1.25, PULSE_WIDTH=64, WIDTH_MIN=1, WIDTH_MAX=8, SNR=7)
    h = hamd.Hella(cfg)
    h.process_gulp(g["data"])
    np.testing.assert_array_equal(h.window(0)[:, :2048], g["out"])
    np.testing.assert_array_equal(h.flag_counts(0), g["counts"])


@pytest.mark.parametrize("mode", ["BOXCAR", "HELLA"])
def test_search_matches_fixture(hamd, mode):
    """The search stages (hella_smooth, hella_find_peaks) on the fixture's
    dedispersed plane (24 trials x 3016 samples, widths 1..16, S/N 6, at most
    1000 candidates): statistics, sigma0 and the candidate list of the
    fixture.  16 channels over 500 MHz down in 6.25 MHz steps, DM 0-5 give the
    24 trials and a 177-sample max delay, so a 3193-sample FILE gulp searches
    3000 samples."""
    g = _load("search_small")
    name = mode.lower()
    cfg = dict(INPUT="FILE", NCHAN=16, FCH1=500.0, FOFF=-6.25, TSAMP=256e-6, GULP=3193,
               DM_MIN=0, DM_MAX=5, DM_TOL=1.25, PULSE_WIDTH=64, WIDTH_MIN=1, WIDTH_MAX=16,
               SNR=6, MAX_GIANTS=1000, SEARCH=mode)
    h = hamd.Hella(cfg)
    i = h.info(0)
    assert (i["ndm"], i["ntime_dd"], i["ntime_out"]) == (24, 3016, 3000)
    h.set_dedisp(0, g["D"])
    h.smooth(1)
    h.find_peaks(0)
    st = h.search_stats(0)
    np.testing.assert_array_equal(st["mu"], g[f"{name}_mu"])
    np.testing.assert_array_equal(st["sigma"], g[f"{name}_sigma"])
    assert st["sigma0"] == g[f"{name}_sigma0"]
    p = h.peaks()
    ref = g[f"{name}_cands"]
    assert len(p) == len(ref) > 0
    for f in ("samp", "width", "dm_idx"):
        np.testing.assert_array_equal(p[f], ref[f])
    np.testing.assert_allclose(p["snr"], ref["snr"], rtol=1e-5, atol=1e-5)
    h.close()
