"""Full-size parity of the production configurations on the GPU (through the
C ABI) against the oracle:

* the plain single beam at production size (1024 channels over 500->400 MHz,
  131.072 us, 65536-sample chunks plus the 142647-sample max-DM tail, DM
  0-2000, widths 1..256) with its 20 injected pulses inside the searched
  window -- every flagged byte of the window, every dedispersed sum, the
  per-width statistics, the candidate list and the clusters;
* the uneven multi-beam shape (512 ch / DM 1000, 1024 ch / DM 2000, 2048 ch /
  DM 3000, killmasks) through the batched graph, beam by beam;
* the streaming shape (1024 ch, 8192-sample chunks, DM 0-2000) through the
  server, batched and plain, every searched chunk of four beams.

Each case prints what it runs before the GPU work starts."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TSAMP = 131.072e-6


def _geom(nch, **kw):
    c = dict(INPUT="FILTERBANK", NCHAN=nch, FCH1=500.0, FOFF=-100.0 / nch, TSAMP=TSAMP,
             DM_MIN=0, DM_MAX=2000, DM_TOL=1.25, PULSE_WIDTH=64, WIDTH_MIN=1, WIDTH_MAX=256,
             SNR=7.0, SCRUNCH=[(1, 1, 6.0)])
    c.update(kw)
    return c


def _plan(orc, cfg):
    return orc.Plan(cfg["NCHAN"], cfg["TSAMP"], cfg["FCH1"], cfg["FOFF"], cfg["DM_MIN"],
                    cfg["DM_MAX"], cfg["PULSE_WIDTH"], cfg["DM_TOL"])


def _flag(orc, cfg, chunk_tf):
    out, _ = orc.flag_u8(chunk_tf, -0.05, 0.15, -1, -1, tuple(cfg.get("SCRUNCH", ())))
    return out


def _same_cands(got, ref_cands):
    assert len(got) == len(ref_cands)
    for f in ("samp", "width", "dm_idx"):
        np.testing.assert_array_equal(got[f], ref_cands[f])
    np.testing.assert_allclose(got["snr"], ref_cands["snr"], rtol=1e-5, atol=1e-5)


def _same_clusters(orc, got, cands, ct=64, cd=16, rep0=0):
    ref = orc.cluster(cands, ct, cd)
    assert [(int(c["rep"]) - rep0, int(c["count"]), int(c["samp_min"]), int(c["samp_max"]),
             int(c["dm_min"]), int(c["dm_max"])) for c in got] == \
        [(c["rep"], c["count"], c["samp_min"], c["samp_max"], c["dm_min"], c["dm_max"])
         for c in ref]


def _say(msg):
    print(msg, flush=True)


# ------------------------------------------------------------------ plain C2
def c2_pulses(seed=20261206, n=20, t_lo=7.1, t_hi=15.5):
    """the production configuration's 20 pulses (DM U(50,1800), FWHM
    log-uniform 0.2-10 ms, S/N U(8,50)), arriving between t_lo and t_hi
    seconds -- inside the output window searched after the fourth chunk"""
    rng = np.random.default_rng(seed)
    return [(float(rng.uniform(t_lo, t_hi)), float(rng.uniform(50, 1800)),
             float(np.exp(rng.uniform(np.log(0.2), np.log(10.0)))), float(rng.uniform(8, 50)))
            for _ in range(n)]


@pytest.mark.timeout(600)
def test_c2_full_size_plain(hamd, orc):
    """One production chunk plus its tail through hella_process_gulp (the
    default one-trial-per-wave kernel with entry tables): flagged window,
    dedispersed plane, statistics, candidates and clusters all equal the
    oracle's; the pulses inside the searched window are found."""
    nch, W = 1024, 65536
    cfg = _geom(nch, GULP=W)
    rng = np.random.default_rng(20261207)
    rfi = rng.choice(nch, int(0.05 * nch), replace=False)
    pulses = c2_pulses()
    h = hamd.Hella(cfg)
    i = h.info(0)
    op = _plan(orc, cfg)
    assert (i["ndm"], i["max_delay"]) == (op.dms.size, op.max_delay) == (1889, 142647)
    ng = i["min_gulp"] + 1
    flagged = []
    _say(f"c2: {ng} chunks of {nch} x {W}, ndm {i['ndm']}, ntime {i['ntime']}")
    for g in range(ng):
        chunk = hamd.synth(nch, W, 500.0, -100.0 / nch, TSAMP, seed=20261208, t_offset=g * W,
                           rfi_chan=rfi, rfi_mult=10.0, pulses=pulses,
                           burst_period=int(round(2.0 / TSAMP)), burst_phase=1000, burst_len=10,
                           burst_amp=6.0)
        _say(f"c2: process_gulp {g}")
        h.process_gulp(chunk)
        flagged.append(_flag(orc, cfg, chunk))
    a0 = ng * W - i["ntime"]
    win = np.concatenate(flagged, axis=1)[:, a0:a0 + i["ntime"]]
    del flagged
    np.testing.assert_array_equal(h.window(0), win)
    D = h.dedisp_plane(0)
    np.testing.assert_array_equal(D, orc.dedisperse_ft(win, op.delta, i["ntime_dd"]))
    del win
    ref = orc.search(D, i["ntime_out"], cfg["WIDTH_MAX"], 1, orc.MODE_BOXCAR, cfg["SNR"],
                     max_giants=10000)
    st = h.search_stats(0)
    np.testing.assert_array_equal(st["mu"], ref["mu"])
    np.testing.assert_array_equal(st["sigma"], ref["sigma"])
    assert st["sigma0"] == ref["sigma0"] and not st["gated"]
    assert h.crossings(0) == ref["count"]
    p = h.peaks()
    _same_cands(p, ref["cands"])
    _same_clusters(orc, h.clusters(), ref["cands"])
    # injected pulses inside the searched window: most are recovered near
    # their arrival sample and DM (the weakest, widest ones may not cross)
    samp0 = a0 + cfg["WIDTH_MAX"] // 2
    found = 0
    for t, dm, fwhm, snr in pulses:
        s = int(round(t / TSAMP)) - samp0
        near = p[(np.abs(p["samp"] - s) < 64 + 4 * fwhm * 1e-3 / TSAMP) &
                 (np.abs(p["dm"] - dm) < 0.05 * dm + 5)]
        found += len(near) > 0
    assert found >= 10, f"only {found} of 20 pulses recovered"
    h.close()


# ------------------------------------------------------------------ uneven beams, batched
@pytest.mark.timeout(600)
def test_c3_uneven_beams_batched(hamd, orc):
    """One beam of each kind of the uneven configuration (512 ch / DM 0-1000,
    1024 ch / DM 0-2000, 2048 ch / DM 0-3000, seeded killmasks of up to 10% of
    the channels, 65536-sample chunks) through the batched graph-captured
    chain: per beam the window equals the oracle's flagged chunks, sampled
    DM rows equal the oracle's sums, and statistics, candidates and clusters
    equal the oracle's search of the plane."""
    W = 65536
    kinds = [(512, 1000.0), (1024, 2000.0), (2048, 3000.0)]
    cfg = _geom(1024, GULP=W, NBEAM=3, MAX_GIANTS=1 << 20)
    rng = np.random.default_rng(31)
    h = hamd.Hella(cfg)
    kms, chunks, plans = [], [], []
    for b, (nch, dmx) in enumerate(kinds):
        h.set_beam(b, NCHAN=nch, FOFF=-100.0 / nch, DM_MAX=dmx)
        km = np.ones(nch, np.int32)
        km[rng.choice(nch, int(rng.uniform(0.02, 0.1) * nch), replace=False)] = 0
        h.set_killmask(b, km)
        kms.append(km)
        chunks.append(hamd.synth(nch, W, 500.0, -100.0 / nch, TSAMP, seed=900 + b,
                                 pulses=[(float(rng.uniform(1, 7)), float(rng.uniform(50, 0.9 * dmx)),
                                          2.0, 25.0)]))
        bc = dict(cfg, NCHAN=nch, FOFF=-100.0 / nch, DM_MAX=dmx)
        plans.append((bc, _plan(orc, bc)))
    info = [h.info(b) for b in range(3)]
    ng = max(i["min_gulp"] for i in info) + 1
    gulp = np.concatenate([c.ravel() for c in chunks])
    _say(f"c3: {ng} batched gulps of 3 beams ({[i['ndm'] for i in info]} trials)")
    for g in range(ng):
        h.process_gulp(gulp, batched=True)
    peaks, clusters = h.peaks(), h.clusters()
    for b in range(3):
        bc, op = plans[b]
        i = info[b]
        assert i["ndm"] == op.dms.size and i["nlive"] == int(kms[b].sum())
        fl = _flag(orc, bc, chunks[b])
        reps = -(-(ng * W) // W)
        a0 = ng * W - i["ntime"]
        win = np.tile(fl, (1, reps))[:, a0:a0 + i["ntime"]]
        _say(f"c3: beam {b} ({bc['NCHAN']} ch): window, sampled rows, search")
        np.testing.assert_array_equal(h.window(b), win)
        D = h.dedisp_plane(b)
        rows = np.unique(np.r_[0, i["ndm"] - 1, rng.choice(i["ndm"], 40, replace=False)])
        np.testing.assert_array_equal(
            D[rows], orc.dedisperse_ft(win, op.delta[rows], i["ntime_dd"], killmask=kms[b]))
        ref = orc.search(D, i["ntime_out"], bc["WIDTH_MAX"], 1, orc.MODE_BOXCAR, bc["SNR"],
                         max_giants=1 << 20)
        st = h.search_stats(b)
        np.testing.assert_array_equal(st["mu"], ref["mu"])
        np.testing.assert_array_equal(st["sigma"], ref["sigma"])
        assert st["sigma0"] == ref["sigma0"]
        assert h.crossings(b) == ref["count"]
        mine = np.nonzero(peaks["beam"] == b)[0]
        _same_cands(peaks[mine], ref["cands"])
        cl = clusters[clusters["beam"] == b]
        _same_clusters(orc, cl, ref["cands"], rep0=int(mine[0]) if len(mine) else 0)
    assert len(peaks) > 0
    h.close()


# ------------------------------------------------------------------ streaming shape, server
SERVE_W = 8192


def _serve_pool(hamd, nch, P, rfi):
    # one continuous stream cut into P chunks; a pulse arrives in each of the
    # first 11 chunks (DM 100-900, so each sweep stays inside the pool)
    rng = np.random.default_rng(41)
    cs = SERVE_W * TSAMP
    pulses = [((j + float(rng.uniform(0.2, 0.8))) * cs, float(rng.uniform(100, 900)),
               float(rng.uniform(1, 4)), 30.0) for j in range(11)]
    data = hamd.synth(nch, P * SERVE_W, 500.0, -100.0 / nch, TSAMP, seed=4242, rfi_chan=rfi,
                      rfi_mult=10.0, pulses=pulses, burst_period=int(round(2.0 / TSAMP)),
                      burst_phase=1000, burst_len=10, burst_amp=6.0, bp_slope=0.3)
    return [np.ascontiguousarray(data[k * SERVE_W:(k + 1) * SERVE_W]) for k in range(P)]


@pytest.mark.timeout(600)
@pytest.mark.parametrize("mode", ["batched", "plain"])
def test_c4_server_shape(hamd, orc, mode):
    """Four beams of the streaming configuration (1024 ch, 8192-sample
    chunks, DM 0-2000, widths 1..256) through the server until each has two
    searched chunks: every searched chunk's crossings, candidates and
    clusters equal the oracle's search of the oracle's window."""
    nch, P, nbeam = 1024, 24, 4
    cfg = _geom(nch, GULP=SERVE_W, CAND_CAP=65536)
    rng = np.random.default_rng(43)
    rfi = rng.choice(nch, int(0.05 * nch), replace=False)
    pool = _serve_pool(hamd, nch, P, rfi)
    op = _plan(orc, cfg)
    ntime = 32 * ((SERVE_W + op.max_delay + 256) // 32 + 1)
    mg = ntime // SERVE_W
    nseq = mg + 2
    order = {b: [(k + 3 * b) % P for k in range(nseq)] for b in range(nbeam)}
    ev = [(0.001 * k, b, order[b][k], k, 0) for k in range(nseq) for b in range(nbeam)]
    opts = dict(plain=True) if mode == "plain" else dict(max_batch=4, lanes=2)
    _say(f"c4 {mode}: {nbeam} beams x {nseq} chunks, window fills after {mg}")
    srv = hamd.Server(cfg, hbm_budget=48 << 30, **opts)
    ring = hamd.Ring(32, pool[0].nbytes, f"NCHAN {nch}\n")
    hamd.serve(srv, ring, ev, pool, time_scale=1.0)
    res = srv.results()
    srv.close()
    ring.close()
    _say("c4: oracle")
    flagged = [_flag(orc, cfg, c) for c in pool]
    checked, ncand = 0, 0
    for r in res:
        if not r["searched"]:
            assert r["seq"] < mg
            continue
        b, k = r["beam"], r["seq"]
        F = np.concatenate([flagged[order[b][j]] for j in range(k + 1)], axis=1)
        a0 = (k + 1) * SERVE_W - ntime
        win = F[:, a0:a0 + ntime]
        D = orc.dedisperse_ft(win, op.delta, SERVE_W + 256)
        ref = orc.search(D, SERVE_W, 256, 1, orc.MODE_BOXCAR, cfg["SNR"], max_giants=10000)
        assert r["crossings"] == ref["count"], (b, k)
        _same_cands(r["cands"], ref["cands"])
        np.testing.assert_array_equal(r["cands"]["abs_samp"], a0 + 128 + r["cands"]["samp"])
        _same_clusters(orc, r["clusters"], ref["cands"])
        checked += 1
        ncand += len(r["cands"])
    assert checked == nbeam * (nseq - mg)
    assert ncand > 0


# ------------------------------------------------------------------ stress shape, whole path
def _box_rows(D, rows, nout, widths, maxW):
    """box sums of the given rows of D for every width: [len(widths)][len(rows)][nout] u64"""
    P = np.zeros((len(rows), D.shape[1] + 1), np.uint64)
    np.cumsum(D[rows].astype(np.uint64), axis=1, out=P[:, 1:])
    c0 = maxW // 2
    out = []
    for w in widths:
        a = np.arange(nout) + c0 - w // 2
        out.append(P[:, a + w] - P[:, a])
    return np.stack(out)


@pytest.mark.timeout(900)
def test_c5_stress_whole_path(hamd, orc):
    """One beam of the stress configuration (4096 ch FP32 N(0,1) input, 20% of
    the channels at 30 sigma, broadband bursts every second, 65.536 us,
    262144-sample chunks, DM 0-5000, widths 1..1024) through the whole plain
    path, with pulses bright enough for more than 1e5 threshold crossings:
    * the last chunk's flagged bytes equal the oracle's flagger;
    * the DM-0 row equals the window's channel sums, sampled DM rows equal
      the oracle's sums;
    * every candidate's S/N, recomputed on the host from the plane with the
      search's own normalisation, agrees within 1e-5;
    * on sampled DM rows every width's crossings are exactly the samples whose
      recomputed S/N exceeds the threshold;
    * the clusters of the (uncapped) list equal the oracle's clusters."""
    nch, W, ts = 4096, 262144, 65.536e-6
    cfg = dict(INPUT="FILTERBANK", INPUT_DTYPE="F32", NCHAN=nch, FCH1=500.0, FOFF=-100.0 / nch,
               TSAMP=ts, GULP=W, DM_MIN=0, DM_MAX=5000, DM_TOL=1.25, PULSE_WIDTH=64,
               WIDTH_MIN=1, WIDTH_MAX=1024, SNR=7.0, SCRUNCH=[(1, 1, 6.0)],
               MAX_GIANTS=1 << 21, CAND_CAP=1 << 21)
    rng = np.random.default_rng(55)
    rfi = rng.choice(nch, nch // 5, replace=False)
    # bright pulses whose sweeps stay inside one chunk (the chunk repeats)
    pulses = [(float(rng.uniform(0.3, 4.0)), float(rng.uniform(100, 900)),
               float(rng.uniform(0.5, 4.0)), float(rng.uniform(40, 80))) for _ in range(24)]
    h = hamd.Hella(cfg)
    i = h.info(0)
    assert (i["ndm"], i["max_delay"]) == (7511, 712492)
    _say(f"c5: synthesising a {nch} x {W} FP32 chunk")
    chunk = hamd.synth(nch, W, 500.0, -100.0 / nch, ts, seed=5005, dtype="f32", mean=0.0,
                       sigma=1.0, rfi_chan=rfi, rfi_mult=30.0, pulses=pulses,
                       burst_period=int(round(1.0 / ts)), burst_phase=500, burst_len=10,
                       burst_amp=6.0)
    buf = hamd.DeviceBuffer(chunk.nbytes)
    buf.upload(chunk)
    ng = i["min_gulp"] + 1
    for g in range(ng):
        _say(f"c5: gulp {g}")
        h.process_gulp(buf)
    buf.free()
    crossings = h.crossings(0)
    p = h.peaks()
    _say(f"c5: {crossings} crossings, {len(p)} candidates; oracle flagger")
    assert 100000 < crossings < (1 << 21) and len(p) == crossings
    # flagger: every gulp is the same chunk, so the window is that chunk's
    # flagged bytes repeated; the GPU's last chunk against the oracle
    fl, _ = orc.flag(orc.ingest_f32_tf(chunk), -0.05, 0.15, -1, -1, ((1, 1, 6.0),))
    del chunk
    win = h.window(0)
    a0 = ng * W - i["ntime"]
    np.testing.assert_array_equal(win[:, -W:], fl)
    np.testing.assert_array_equal(win[:, :W - a0], fl[:, a0:])
    del fl
    _say("c5: plane")
    D = h.dedisp_plane(0)
    ntdd = i["ntime_dd"]
    np.testing.assert_array_equal(D[0], win[:, :ntdd].sum(axis=0, dtype=np.uint32))
    op = orc.Plan(nch, ts, 500.0, -100.0 / nch, 0.0, 5000.0, 64.0, 1.25)
    rows = np.unique(np.r_[i["ndm"] - 1, rng.choice(i["ndm"], 14, replace=False),
                           np.unique(p["dm_idx"])[:: max(1, len(np.unique(p["dm_idx"])) // 6)]])
    np.testing.assert_array_equal(D[rows], orc.dedisperse_ft(win, op.delta[rows], ntdd))
    del win
    # candidates: S/N from the plane with the search's normalisation
    _say("c5: candidates")
    mu, inv, thr = h.search_norm(0)
    st = h.search_stats(0)
    assert not st["gated"]
    widths = [1 << k for k in range(i["nwidth"])]
    nout = i["ntime_out"]
    for k, w in enumerate(widths):
        sel = p[p["width"] == k]
        if not len(sel):
            continue
        lo = sel["samp"] + 512 - w // 2
        box = np.zeros(len(sel), np.uint64)
        for j in range(w):
            box += D[sel["dm_idx"], lo + j]
        snr = ((box.astype(np.float64) - mu[k]) * inv[k]).astype(np.float32)
        np.testing.assert_allclose(sel["snr"], snr, rtol=1e-5, atol=1e-5)
    # sampled rows: the crossings of every width are exactly the host's
    srows = np.unique(np.r_[rows, np.unique(p["dm_idx"])[:8]])
    B = _box_rows(D, srows, nout, widths, 1024)
    for k in range(len(widths)):
        y = ((B[k].astype(np.float64) - mu[k]) * inv[k]).astype(np.float32)
        for j, r in enumerate(srows):
            want = np.nonzero(y[j] > thr)[0]
            got = p[(p["width"] == k) & (p["dm_idx"] == r)]["samp"]
            np.testing.assert_array_equal(got, want)
    del D, B
    _say("c5: clusters")
    ref = orc.cluster(p, 64, 16)
    got = h.clusters()
    assert [(int(c["rep"]), int(c["count"]), int(c["samp_min"]), int(c["samp_max"]),
             int(c["dm_min"]), int(c["dm_max"])) for c in got] == \
        [(c["rep"], c["count"], c["samp_min"], c["samp_max"], c["dm_min"], c["dm_max"]) for c in ref]
    h.close()
