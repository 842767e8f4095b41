"""Candidate clustering on the GPU (through the C ABI) against the oracle's
clustering of the same candidate lists: 8-connected cells per beam, the
highest-S/N member (lowest index on ties) as representative, clusters by
representative index; in the plain and batched context paths, with the
MAX_GIANTS cut falling inside a beam's list, and in the streaming server."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GEOM = dict(NCHAN=64, FCH1=500.0, FOFF=-1.5625, TSAMP=256e-6, PULSE_WIDTH=64, DM_TOL=1.25)


def _cfg(**kw):
    c = dict(GEOM)
    c.update(INPUT="FILTERBANK", DM_MIN=0, DM_MAX=80, WIDTH_MIN=1, WIDTH_MAX=8, SNR=5)
    c.update(kw)
    return c


def _oracle_clusters(orc, peaks, ct, cd):
    """the oracle's clusters of each beam's run of the list, in list order"""
    out = []
    beams = list(dict.fromkeys(peaks["beam"].tolist()))
    for b in beams:
        idx = np.nonzero(peaks["beam"] == b)[0]
        assert np.all(np.diff(idx) == 1)          # a beam's candidates are contiguous
        for c in orc.cluster(peaks[idx], ct, cd):
            out.append((c["rep"] + int(idx[0]), c["count"], c["samp_min"], c["samp_max"],
                        c["dm_min"], c["dm_max"], b))
    return out


def _device_clusters(cl):
    return [(int(c["rep"]), int(c["count"]), int(c["samp_min"]), int(c["samp_max"]),
             int(c["dm_min"]), int(c["dm_max"]), int(c["beam"])) for c in cl]


def _run(hamd, cfg, batched, gulps=6):
    """three uneven beams (64, 32 and 96 channels, one with a killmask), with
    dispersed pulses and broadband bursts; per gulp the list and clusters"""
    W = cfg["GULP"]
    h = hamd.Hella(cfg)
    h.set_beam(1, NCHAN=32, DM_MAX=50)
    h.set_beam(2, NCHAN=96, FOFF=-1.0, DM_MAX=120)
    km = np.ones(96, np.int32)
    km[[5, 17, 60]] = 0
    h.set_killmask(2, km)
    nchs, foffs = (64, 32, 96), (-1.5625, -1.5625, -1.0)
    pulses = ([(0.9, 40.0, 1.0, 40.0), (1.45, 10.0, 3.0, 25.0)], [(1.2, 20.0, 1.0, 35.0)],
              [(1.9, 90.0, 2.0, 40.0)])
    out = []
    for g in range(gulps):
        parts = [hamd.synth(nchs[b], W, 500.0, foffs[b], 256e-6, seed=500 + b, t_offset=g * W,
                            pulses=pulses[b], burst_period=700, burst_phase=50 * b,
                            burst_len=4, burst_amp=4.0).ravel() for b in range(3)]
        h.process_gulp(np.concatenate(parts), batched=batched)
        out.append((h.peaks().copy(), h.clusters().copy()))
    h.close()
    return out


@pytest.mark.parametrize("batched", [False, True])
@pytest.mark.parametrize("cell", [(64, 16), (8, 2), (1, 1)])
def test_clusters_match_oracle(hamd, orc, batched, cell):
    ct, cd = cell
    cfg = _cfg(GULP=2048, NBEAM=3, CLUSTER_DT=ct, CLUSTER_DDM=cd)
    total = multi = 0
    for peaks, cl in _run(hamd, cfg, batched):
        assert _device_clusters(cl) == _oracle_clusters(orc, peaks, ct, cd)
        total += len(cl)
        multi += int(np.sum(cl["count"] > 1))
    assert total > 0 and multi > 0                # clusters of several crossings occurred


@pytest.mark.parametrize("batched", [False, True])
@pytest.mark.parametrize("cap", [7, 40, 150])
def test_clusters_with_max_giants_cut(hamd, orc, batched, cap):
    """the cut of the gulp's list at MAX_GIANTS falls inside a beam: the
    clusters are those of the candidates kept"""
    cfg = _cfg(GULP=2048, NBEAM=3, MAX_GIANTS=cap, CLUSTER_DT=8, CLUSTER_DDM=2)
    cut = 0
    for peaks, cl in _run(hamd, cfg, batched, gulps=5):
        assert len(peaks) <= cap
        cut += int(len(peaks) == cap)
        assert _device_clusters(cl) == _oracle_clusters(orc, peaks, 8, 2)
    assert cut > 0


def test_batched_clusters_equal_plain(hamd):
    cfg = _cfg(GULP=2048, NBEAM=3, CLUSTER_DT=8, CLUSTER_DDM=2)
    for (pp, pc), (bp, bc) in zip(_run(hamd, cfg, False), _run(hamd, cfg, True)):
        np.testing.assert_array_equal(pp, bp)
        np.testing.assert_array_equal(pc, bc)


def test_server_clusters_match_oracle(hamd, orc):
    """every searched chunk of the streaming server (batched and plain
    modes): its clusters are the oracle's clusters of its candidates"""
    chunk = 1024
    cfg = _cfg(GULP=chunk, DM_MAX=50, SCRUNCH=[(1, 1, 6.0)], CLUSTER_DT=8, CLUSTER_DDM=2)
    streams = {}
    for b in range(3):
        data = hamd.synth(64, 9 * chunk, 500.0, -1.5625, 256e-6, seed=900 + b,
                          pulses=[(0.5 + 0.3 * b, 30.0, 1.0, 40.0), (1.6, 12.0, 2.0, 30.0)],
                          burst_period=900, burst_phase=40 * b, burst_len=4, burst_amp=4.0)
        streams[b] = [data[k * chunk:(k + 1) * chunk] for k in range(9)]
    pool, ev = [], []
    for b, chunks in streams.items():
        for k, ch in enumerate(chunks):
            pool.append(ch)
            ev.append((0.001 * b + 0.002 * k, b, len(pool) - 1, k, 1 if k == 8 else 0))
    ev.sort(key=lambda e: (e[0], e[1]))
    checked = 0
    for opts in (dict(max_batch=4, lanes=2), dict(plain=True)):
        srv = hamd.Server(cfg, **opts)
        ring = hamd.Ring(16, pool[0].nbytes, "NCHAN 64\n")
        hamd.serve(srv, ring, ev, pool, time_scale=1.0)
        for r in srv.results():
            if not r["searched"]:
                continue
            assert _device_clusters(r["clusters"]) == _oracle_clusters(orc, r["cands"], 8, 2)
            checked += len(r["clusters"])
        srv.close()
        ring.close()
    assert checked > 0


@pytest.mark.parametrize("batched", [False, True])
@pytest.mark.parametrize("snr,lo,hi,min_cl", [(2.5, 2049, 16384, 1025), (2.0, 16385, 10 ** 9, 0)])
def test_large_lists(hamd, orc, snr, lo, hi, min_cl, batched):
    """lists beyond the LDS capacities (parents in LDS and the other words
    in the work set's buffers; then everything in the buffers); at S/N 2.5 a
    beam has more clusters than the 1024 its pinned staging slot holds, so
    the rest of its list is fetched after the counts"""
    cfg = _cfg(GULP=2048, NBEAM=3, SNR=snr, MAX_GIANTS=60000, CAND_CAP=60000, CLUSTER_DT=8,
               CLUSTER_DDM=2)
    seen, most = 0, 0
    for peaks, cl in _run(hamd, cfg, batched, gulps=4):
        assert _device_clusters(cl) == _oracle_clusters(orc, peaks, 8, 2)
        for b in set(peaks["beam"].tolist()):
            seen += int(lo <= int(np.sum(peaks["beam"] == b)) <= hi)
        for b in set(cl["beam"].tolist()):
            most = max(most, int(np.sum(cl["beam"] == b)))
    assert seen > 0
    assert most >= min_cl
