"""The per-stage C ABI (the drop-in for the hella:: stage functions the
reference's gulp loop calls at casm:678-736) on the GPU, the fp16
channel-major DADA input, and a dedisp plan whose tile shape changes between
executions."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GEOM = dict(NCHAN=64, FCH1=500.0, FOFF=-1.5625, TSAMP=256e-6, PULSE_WIDTH=64, DM_TOL=1.25)


def _cfg(**kw):
    c = dict(GEOM)
    c.update(INPUT="FILTERBANK", GULP=2048, NBEAM=3, DM_MIN=0, DM_MAX=80, WIDTH_MIN=1,
             WIDTH_MAX=8, SNR=6, SPEC_MIN=-0.05, SPEC_MAX=0.15, SCRUNCH=[(1, 1, 5.0)])
    c.update(kw)
    return c


NCHS, FOFFS, DMX = (64, 32, 96), (-1.5625, -1.5625, -1.0), (80.0, 50.0, 120.0)


def _ctx(hamd, cfg):
    h = hamd.Hella(cfg)
    h.set_beam(1, NCHAN=32, DM_MAX=50)
    h.set_beam(2, NCHAN=96, FOFF=-1.0, DM_MAX=120)
    km = np.ones(96, np.int32)
    km[[5, 17, 60]] = 0
    h.set_killmask(2, km)
    return h, km


def _gulp(hamd, g, W=2048):
    pulses = ([(0.7, 40.0, 1.0, 30.0), (2.2, 70.0, 1.0, 25.0)], [(1.1, 20.0, 1.0, 30.0)],
              [(1.4, 90.0, 2.0, 30.0)])
    parts = [hamd.synth(NCHS[b], W, 500.0, FOFFS[b], 256e-6, seed=1000 + b, t_offset=g * W,
                        pulses=pulses[b], burst_period=1500, burst_phase=200 * b, burst_len=5,
                        burst_amp=5.0) for b in range(3)]
    return parts


def test_stage_api_equals_gulp_and_oracle(hamd, orc):
    """clear_peaks -> fast_flagger -> per beam (dedisperse, smooth,
    find_peaks) -> end_gulp over eight gulps of three uneven beams gives the
    candidates and clusters of hella_process_gulp on the same gulps, and
    every beam's window, plane, statistics and candidates equal the
    oracle's."""
    cfg = _cfg()
    hs, km = _ctx(hamd, cfg)
    hg, _ = _ctx(hamd, cfg)
    info = [hs.info(b) for b in range(3)]
    plans = [orc.Plan(NCHS[b], 256e-6, 500.0, FOFFS[b], 0.0, DMX[b], 64.0, 1.25)
             for b in range(3)]
    kms = [None, None, km]
    flagged = [[], [], []]
    searched = 0
    for g in range(8):
        parts = _gulp(hamd, g)
        gulp = np.concatenate([p.ravel() for p in parts])
        for b in range(3):
            flagged[b].append(orc.flag_u8(parts[b], -0.05, 0.15, -1, -1,
                                       ((1, 1, 5.0),))[0])
        hg.process_gulp(gulp)
        hs.clear_peaks()
        hs.fast_flagger(gulp)
        wins = {}
        for b in range(3):
            if g < info[b]["min_gulp"]:
                continue
            i = info[b]
            a0 = (g + 1) * 2048 - i["ntime"]
            wins[b] = np.concatenate(flagged[b], axis=1)[:, a0:a0 + i["ntime"]]
            hs.dedisperse(b)
            hs.smooth(1)
            hs.find_peaks(b)
            # this beam's plane and search against the oracle, before the
            # next beam reuses the work set
            D = hs.dedisp_plane(b)
            np.testing.assert_array_equal(
                D, orc.dedisperse_ft(wins[b], plans[b].delta, i["ntime_dd"], killmask=kms[b]))
            ref = orc.search(D, i["ntime_out"], 8, 1, orc.MODE_BOXCAR, 6.0)
            st = hs.search_stats(b)
            np.testing.assert_array_equal(st["mu"], ref["mu"])
            np.testing.assert_array_equal(st["sigma"], ref["sigma"])
            assert hs.crossings(b) == ref["count"]
            mine = hs.peaks()
            mine = mine[mine["beam"] == b]
            for f in ("samp", "width", "dm_idx"):
                np.testing.assert_array_equal(mine[f], ref["cands"][f])
            np.testing.assert_allclose(mine["snr"], ref["cands"]["snr"], rtol=1e-5, atol=1e-5)
            searched += 1
        hs.end_gulp()
        # the windows (oldest sample first) once the gulp is complete
        for b, win in wins.items():
            np.testing.assert_array_equal(hs.window(b), win)
        np.testing.assert_array_equal(hs.peaks(), hg.peaks())
        np.testing.assert_array_equal(hs.clusters(), hg.clusters())
    assert searched >= 9
    # out-of-order calls are errors, not crashes
    hs.clear_peaks()
    with pytest.raises(hamd.HellaError):
        hs.find_peaks(1)          # the last smoothed beam is 2
    with pytest.raises(hamd.HellaError):
        hs.smooth(2)              # only the normalised widths exist
    hs.close()
    hg.close()


def test_dada_fp16_channel_major_input(hamd, orc):
    """INPUT DADA: gulps are fp16 channel-major blocks ([nchan][gulp] per beam,
    the PSRDADA format of transpose.cu:87-90 / casm:688-695); the streaming
    window equals the oracle's flagger on the same fp16 values, and the
    sums and candidates follow."""
    W = 2048
    cfg = dict(GEOM, INPUT="DADA", GULP=W, DM_MIN=0, DM_MAX=60, WIDTH_MIN=1, WIDTH_MAX=8,
               SNR=6, SCRUNCH=[(1, 1, 5.0), (4, 8, 3.0)])
    h = hamd.Hella(cfg)
    i = h.info(0)
    op = orc.Plan(64, 256e-6, 500.0, -1.5625, 0.0, 60.0, 64.0, 1.25)
    flagged = []
    for g in range(i["min_gulp"] + 2):
        x = hamd.synth(64, W, 500.0, -1.5625, 256e-6, seed=61, t_offset=g * W, dtype="f32",
                       mean=0.0, sigma=1.0, rfi_chan=(9,), rfi_mult=20.0,
                       pulses=[(1.0, 40.0, 1.0, 30.0)])
        ft16 = np.ascontiguousarray(x.T.astype(np.float16))        # [nchan][gulp]
        h.process_gulp(ft16)
        out, _ = orc.flag(ft16.astype(np.float32), -0.05, 0.15, -1, -1, ((1, 1, 5.0), (4, 8, 3.0)))
        flagged.append(out)
    F = np.concatenate(flagged, axis=1)
    a0 = (i["min_gulp"] + 2) * W - i["ntime"]
    win = F[:, a0:a0 + i["ntime"]]
    np.testing.assert_array_equal(h.window(0), win)
    D = h.dedisp_plane(0)
    np.testing.assert_array_equal(D, orc.dedisperse_ft(win, op.delta, i["ntime_dd"]))
    ref = orc.search(D, i["ntime_out"], 8, 1, orc.MODE_BOXCAR, 6.0)
    p = h.peaks()
    assert h.crossings(0) == ref["count"] > 0
    for f in ("samp", "width", "dm_idx"):
        np.testing.assert_array_equal(p[f], ref["cands"][f])
    h.close()


@pytest.mark.parametrize("dtype,scr", [("U8", ((1, 1, 5.0), (1, 1, 3.0))),
                                       ("F32", ((1, 1, 5.0), (4, 8, 3.0)))])
def test_flag_spectrum_and_specflags(hamd, orc, dtype, scr):
    """The flag spectrum (the reference's flagSpec: per scrunch, the mean of
    the grown mask over the statistics columns, accumulated over gulps) of
    two beams on the value-table path (u8, 1 x 1 scrunches) and on the FP32
    path (a 4 x 8 box scrunch) equals the oracle's, bit for bit, and the
    spectral-flag counts follow casm:705-716 (specflags[i] += (int)(NTIME x
    flagSpec[i]), the beams' spectra summed)."""
    W = 4000                     # statistics columns 3840: the window matters
    cfg = dict(GEOM, INPUT="FILTERBANK", INPUT_DTYPE=dtype, GULP=W, NBEAM=2, DM_MIN=0, DM_MAX=40,
               WIDTH_MIN=1, WIDTH_MAX=8, SNR=6, SCRUNCH=list(scr))
    h = hamd.Hella(cfg)
    i = h.info(0)
    specs = [np.zeros(64, np.float32), np.zeros(64, np.float32)]
    specflags = np.zeros(64, np.int32)
    want_flags = np.zeros(64, np.int32)
    for g in range(i["min_gulp"] + 3):
        parts = []
        for b in range(2):
            kw = dict(dtype="f32", mean=0.0, sigma=1.0) if dtype == "F32" else {}
            x = hamd.synth(64, W, 500.0, -1.5625, 256e-6, seed=300 + b, t_offset=g * W,
                           rfi_chan=(3 + b, 40), rfi_mult=6.0, burst_period=1700,
                           burst_phase=100 * b, burst_len=6, burst_amp=5.0,
                           pulses=[(0.4, 30.0, 1.0, 40.0)], **kw)
            parts.append(x)
            if dtype == "F32":
                orc.flag(orc.ingest_f32_tf(x), -0.05, 0.15, -1, -1, scr, spec=specs[b])
            else:
                orc.flag_u8(x, -0.05, 0.15, -1, -1, scr, spec=specs[b])
        h.process_gulp(np.concatenate([p.ravel() for p in parts]))
        for b in range(2):
            np.testing.assert_array_equal(h.flag_spectrum(b), specs[b])
        if g >= i["min_gulp"]:
            h.accumulate_specflags(specflags)
            tot = specs[0].astype(np.float64) + specs[1].astype(np.float64)
            want_flags += (i["ntime"] * tot).astype(np.int32)
            np.testing.assert_array_equal(specflags, want_flags)
    assert specs[0].max() > 0 and specflags.max() > 0
    h.close()


@pytest.mark.parametrize("dtype,scr,W", [("U8", ((1, 1, 5.0), (1, 1, 3.0)), 4000),
                                         ("U8", ((1, 1, 5.0),), 4096),
                                         ("F32", ((1, 1, 5.0), (4, 8, 3.0)), 4000)])
def test_robust_statistics_mode(hamd, orc, dtype, scr, W):
    """FLAG_STATS ROBUST (per-channel median and 1.4826 x MAD in every
    normalisation; radix selection on the device): on the value-table path
    (u8, 1 x 1 scrunches) and on the FP32 path (a 4 x 8 box scrunch), the
    streaming window equals the oracle's robust flagger bit for bit, and the
    dedispersed sums and candidates follow."""
    cfg = dict(GEOM, INPUT="FILTERBANK", INPUT_DTYPE=dtype, GULP=W, DM_MIN=0, DM_MAX=60,
               WIDTH_MIN=1, WIDTH_MAX=8, SNR=6, SCRUNCH=list(scr), FLAG_STATS="ROBUST")
    h = hamd.Hella(cfg)
    i = h.info(0)
    op = orc.Plan(64, 256e-6, 500.0, -1.5625, 0.0, 60.0, 64.0, 1.25)
    flagged = []
    for g in range(i["min_gulp"] + 2):
        kw = dict(dtype="f32", mean=0.0, sigma=1.0) if dtype == "F32" else {}
        x = hamd.synth(64, W, 500.0, -1.5625, 256e-6, seed=71, t_offset=g * W, rfi_chan=(9, 33),
                       rfi_mult=12.0, burst_period=1500, burst_phase=300, burst_len=6,
                       burst_amp=5.0, pulses=[(t, 40.0, 1.0, 30.0) for t in (0.6, 1.4, 2.2, 3.0)],
                       bp_slope=0.4, **kw)
        h.process_gulp(x)
        if dtype == "F32":
            out, _ = orc.flag(orc.ingest_f32_tf(x), -0.05, 0.15, -1, -1, scr, robust=True)
        else:
            out, _ = orc.flag_u8(x, -0.05, 0.15, -1, -1, scr, robust=True)
        flagged.append(out)
    F = np.concatenate(flagged, axis=1)
    a0 = (i["min_gulp"] + 2) * W - i["ntime"]
    win = F[:, a0:a0 + i["ntime"]]
    np.testing.assert_array_equal(h.window(0), win)
    D = h.dedisp_plane(0)
    np.testing.assert_array_equal(D, orc.dedisperse_ft(win, op.delta, i["ntime_dd"]))
    ref = orc.search(D, i["ntime_out"], 8, 1, orc.MODE_BOXCAR, 6.0)
    p = h.peaks()
    assert h.crossings(0) == ref["count"] > 0
    for f in ("samp", "width", "dm_idx"):
        np.testing.assert_array_equal(p[f], ref["cands"][f])
    h.close()


def _spread4(delta, live):
    d = delta[:, live]
    return int(max((d[j + 3] - d[j]).max() for j in range(d.shape[0] - 3)))


@pytest.mark.parametrize("change", ["killmask", "dm_list"])
def test_dedisp_tile_follows_plan_changes(hamd, orc, change):
    """A dedisp plan executed, then given a killmask or a DM list (same count,
    same max delay) that makes it choose another tile shape (one trial per
    wave with 1792-sample tiles <-> 32 lanes per trial with 8-trial blocks),
    then executed again on the same number of samples: both results equal
    the oracle's sums (the tile table is rebuilt with the shape)."""
    nch, dt, f0, df = 64, 256e-6, 500.0, -1.5625
    dense = np.linspace(0.0, 179.0, 40).astype(np.float32)
    jump = np.r_[np.arange(20), np.arange(160, 180)].astype(np.float32)
    delay = orc.delay_table(nch, dt, f0, df)
    low_killed = np.r_[np.ones(32), np.zeros(32)].astype(np.int32)
    if change == "killmask":
        steps = [(jump, None), (jump, low_killed), (jump, None)]
    else:
        steps = [(dense, None), (jump, None), (dense, None)]
    # the steps really alternate between the two shapes (the one-trial-per-
    # wave tile holds a 4-trial delay spread of at most 4316 samples)
    spreads = [_spread4(orc.delays(dms, delay), (km if km is not None else np.ones(nch)) > 0)
               for dms, km in steps]
    assert spreads[0] > 4316 >= spreads[1] if change == "killmask" else \
        spreads[1] > 4316 >= spreads[0]
    rng = np.random.default_rng(5)
    md = orc.max_delay(jump, delay)
    assert md == orc.max_delay(dense, delay)
    data = rng.integers(0, 256, size=(md + 3000, nch), dtype=np.uint8)
    lp = hamd.DedispPlan(nch, dt, f0, df)
    for dms, km in steps:
        lp.set_dm_list(dms)
        lp.set_killmask(km)
        assert lp.max_delay == md
        got = lp.execute_adv(data)
        ref = orc.dedisperse_tf(data, orc.delays(dms, delay), 3000, killmask=km)
        np.testing.assert_array_equal(got, ref.astype(np.float32))
    lp.close()
