"""Streaming server on the GPU: every beam's result is the same alone and in
the mix (continuous batching, several lanes, beams joining and leaving,
admission waits under a small HBM budget), and the same as the plain
per-gulp pipeline on that beam's chunk sequence."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GEOM = dict(NCHAN=64, FCH1=500.0, FOFF=-1.5625, TSAMP=256e-6, PULSE_WIDTH=64, DM_TOL=1.25)
CHUNK = 1024


def _cfg(**kw):
    c = dict(GEOM)
    c.update(DM_MIN=0, DM_MAX=50, WIDTH_MIN=1, WIDTH_MAX=16, SNR=6, GULP=CHUNK,
             SCRUNCH=[(1, 1, 6.0), (4, 4, 4.0)], CAND_CAP=65536)
    c.update(kw)
    return c


def _beam_stream(hamd, cfg, beam, nchunks):
    """a beam's continuous synthetic stream, cut into chunks (u8 [CHUNK][nch])"""
    rng = np.random.default_rng(1000 + beam)
    t_end = nchunks * CHUNK * cfg["TSAMP"]
    pulses = [(float(t), float(rng.uniform(5, 45)), float(rng.uniform(0.5, 3)),
               float(rng.uniform(12, 40))) for t in rng.uniform(0.2, t_end, 3)]
    data = hamd.synth(cfg["NCHAN"], nchunks * CHUNK, cfg["FCH1"], cfg["FOFF"], cfg["TSAMP"],
                      seed=77 + beam, pulses=pulses, rfi_chan=(beam % 60, 61), rfi_mult=6.0,
                      burst_period=3000, burst_phase=100 * beam, burst_len=6, burst_amp=5.0)
    return [data[k * CHUNK:(k + 1) * CHUNK] for k in range(nchunks)]


def _schedule(streams, starts, period):
    """events (t, beam, pool, seq, flags) and the pool of chunks"""
    pool, ev = [], []
    for beam, chunks in streams.items():
        for k, ch in enumerate(chunks):
            pool.append(ch)
            last = k == len(chunks) - 1
            ev.append((starts[beam] + k * period, beam, len(pool) - 1, k, 1 if last else 0))
    ev.sort(key=lambda e: (e[0], e[1]))
    return ev, pool


def _run(hamd, cfg, streams, starts, period=0.002, **opts):
    ev, pool = _schedule(streams, starts, period)
    srv = hamd.Server(cfg, **opts)
    ring = hamd.Ring(16, pool[0].nbytes, "NCHAN %d\n" % cfg["NCHAN"])
    hamd.serve(srv, ring, ev, pool, time_scale=1.0)
    res = srv.results()
    st = srv.stats()
    srv.close()
    ring.close()
    by = {}
    for r in res:
        by.setdefault(r["beam"], {})[r["seq"]] = r
    return by, st


def _same(a, b):
    assert a["searched"] == b["searched"] and a["crossings"] == b["crossings"]
    ca, cb = a["cands"], b["cands"]
    assert len(ca) == len(cb)
    for f in ("samp", "width", "dm_idx", "abs_samp", "beam"):
        np.testing.assert_array_equal(ca[f], cb[f])
    np.testing.assert_array_equal(ca["snr"], cb["snr"])   # same kernels, same order: bit-exact
    np.testing.assert_array_equal(a["clusters"], b["clusters"])


@pytest.fixture(scope="module")
def streams(hamd):
    cfg = _cfg()
    lens = {0: 9, 1: 7, 2: 8, 3: 6, 4: 9, 5: 5}
    return cfg, {b: _beam_stream(hamd, cfg, b, n) for b, n in lens.items()}


@pytest.mark.parametrize("opts", [dict(max_batch=4, lanes=2), dict(max_batch=2, lanes=1),
                                  dict(max_batch=8, lanes=3)])
def test_alone_equals_mix(hamd, streams, opts):
    cfg, st = streams
    starts = {0: 0.0, 1: 0.001, 2: 0.004, 3: 0.0, 4: 0.007, 5: 0.003}
    mix, stats = _run(hamd, cfg, st, starts, **opts)
    assert stats["chunks"] == sum(len(v) for v in st.values())
    assert stats["batches"] < stats["chunks"]           # batching happened
    for beam in st:
        alone, _ = _run(hamd, cfg, {beam: st[beam]}, {beam: 0.0}, **opts)
        assert sorted(alone[beam]) == sorted(mix[beam]) == list(range(len(st[beam])))
        for k in alone[beam]:
            _same(alone[beam][k], mix[beam][k])
    # some chunks were searched and found candidates
    assert any(r["searched"] and r["ncand"] > 0 for v in mix.values() for r in v.values())


def test_server_matches_plain_pipeline(hamd, streams):
    """chunk-by-chunk hella_process_gulp on one beam gives the server's results"""
    cfg, st = streams
    beam = 4
    srv, _ = _run(hamd, cfg, {beam: st[beam]}, {beam: 0.0}, max_batch=4, lanes=2)
    h = hamd.Hella(dict(cfg, NBEAM=1))
    for k, ch in enumerate(st[beam]):
        h.process_gulp(ch)
        got = srv[beam][k]
        ref = h.peaks()
        assert got["searched"] == int(k >= h.info(0)["min_gulp"])
        assert len(ref) == len(got["cands"])
        for f in ("snr", "samp", "width", "dm_idx", "abs_samp"):
            np.testing.assert_array_equal(ref[f], got["cands"][f])
        if got["searched"]:
            assert h.crossings(0) == got["crossings"]
    h.close()


def test_plain_mode_equals_batched(hamd, streams):
    cfg, st = streams
    sub = {b: st[b] for b in (0, 2, 5)}
    starts = {0: 0.0, 2: 0.0005, 5: 0.001}
    a, sa = _run(hamd, cfg, sub, starts, plain=True)
    b, sb = _run(hamd, cfg, sub, starts, max_batch=8, lanes=2)
    assert sa["batches"] == sa["chunks"]
    for beam in sub:
        for k in a[beam]:
            _same(a[beam][k], b[beam][k])


def test_admission_waits_for_hbm(hamd, streams):
    """a budget for two resident beams: later beams wait, results unchanged"""
    cfg, st = streams
    opts = dict(max_batch=4, lanes=2)
    beam_b, work_b = hamd.Server.footprint(cfg, **opts)
    starts = {b: 0.0005 * b for b in st}
    mix, stats = _run(hamd, cfg, st, starts, hbm_budget=work_b + 2 * beam_b + beam_b // 2, **opts)
    assert stats["max_resident"] == 2
    assert stats["beams_waited"] >= 1
    assert stats["beams_admitted"] == len(st)
    for beam in (1, 5):
        alone, _ = _run(hamd, cfg, {beam: st[beam]}, {beam: 0.0}, **opts)
        for k in alone[beam]:
            _same(alone[beam][k], mix[beam][k])


def test_latency_is_reported(hamd, streams):
    cfg, st = streams
    res, stats = _run(hamd, cfg, {0: st[0], 1: st[1]}, {0: 0.0, 1: 0.001}, period=0.01,
                      max_batch=4, lanes=2)
    lat = [(r["t_done_ns"] - r["t_arrival_ns"]) * 1e-9 for v in res.values() for r in v.values()]
    assert all(0 < x < 5.0 for x in lat)
    assert stats["wall_s"] > 0 and 0 < stats["busy_s"] <= stats["wall_s"] + 1e-3


def test_bad_budget_and_block_size(hamd, streams):
    cfg, st = streams
    with pytest.raises(hamd.HellaError):
        hamd.Server(cfg, hbm_budget=1 << 20, max_batch=4, lanes=2)
    srv = hamd.Server(cfg, max_batch=2, lanes=1)
    ring = hamd.Ring(4, 64)
    ev = [(0.0, 0, 0, 0, 1)]
    with pytest.raises(hamd.HellaError):
        hamd.serve(srv, ring, ev, [np.zeros(64, np.uint8)])
    srv.close()
    ring.close()
