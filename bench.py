"""FRB-search hot path benchmark on one MI355X.

Workload (the production single-beam configuration): 1 beam of uint8
filterbank data, 1024 channels over 500->400 MHz, tsamp 131.072 us, 65536-sample
chunks plus the max-DM overlap tail, DM 0-2000 (tolerance 1.25, pulse width
64 us), boxcar widths 1..256 in powers of two, S/N threshold 7.  The synthetic
stream carries 20 dispersed pulses (DM U(50,1800), FWHM logU(0.2,10) ms, S/N
U(8,50)), 5% persistent narrowband RFI channels at 10 sigma and 10-sample
broadband bursts every 2 s at +6 sigma.

One step = one chunk through flag -> dedisperse -> boxcar search -> peaks on the
plain per-beam path, input already resident in HBM.  value = DM-trial samples
per second of that step.  Also reported: the batched graph path on several
beams, per-stage device times, the dedispersion kernel's integer-add roofline,
the flagger's HBM roofline and the oracle (CPU) on a bounded sample.
"""
from __future__ import annotations

import argparse
import importlib.util
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
OUT_DIR = os.path.join(ROOT, "bench_out")

CHUNK = 65536
NCH = 1024
TSAMP = 131.072e-6
FOFF = -100.0 / NCH
MAXW = 256
SNR = 7.0
PEAK_HBM_GBS = 8000.0
PEAK_LDS_GBS = 128 * 256 * 2.4  # LDS read bandwidth: 128 B per CU per cycle, 256 CUs, 2.4 GHz
PEAK_ADD_T = 78.6  # packed u16 integer adds/s (256 CU x 64 lanes x 2 x 2.4 GHz), 1e12


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "plain_1024ch_dm2000_counters.md")
TUNING_DIR = os.path.join(ROOT, "casm-hella_amd", "tuning")


def tuning_record(h, name, tune):
    """The committed autotuner record of a workload (TUNING_DIR/<name>.json),
    applied to a fresh context before its first gulp.  With --tune nothing is
    applied: the tuner runs once the windows have filled and writes
    <name>.json to --out-dir.  Returns {"best": ..., "source": ...}, or the
    built-in choice when there is no record."""
    path = os.path.join(TUNING_DIR, name + ".json")
    if tune:
        return {"source": "autotuned in this run", "record": name + ".json"}
    if not os.path.exists(path):
        return {"source": "built-in choice (no record)"}
    return {"best": h.use_tuning(path), "source": os.path.relpath(path, ROOT)}


def profiled_traffic(name, key="traffic_bytes_per_launch"):
    """HBM bytes per launch of a kernel (or per chunk of a stage) from the
    committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary of this workload (or,
    with key, another per-launch figure the summary lists), or None when the
    summary does not list it."""
    try:
        with open(TRAFFIC_PROFILE) as f:
            for line in f:
                parts = line.split()
                if line.startswith("machine-readable: " + key) and parts[2] == name:
                    return float(parts[3])
    except OSError:
        pass
    return None


def lds_roof(ms):
    """LDS read throughput of the dedispersion kernel: the bytes one launch
    reads from LDS (SQ_INSTS_LDS_LOAD_BANDWIDTH x 64 B in the committed
    counter summary) over the launch time measured in this run"""
    b = profiled_traffic("k_dedisp", "lds_load_bytes_per_launch")
    if b is None or not ms:
        return None
    a = b / (ms * 1e-3) / 1e9
    return {"achieved": a, "peak": PEAK_LDS_GBS, "unit": "GB/s", "frac": a / PEAK_LDS_GBS,
            "bytes": b}


def load_pkg():
    spec = importlib.util.spec_from_file_location(
        "hella_amd", os.path.join(ROOT, "casm-hella_amd", "hella_amd.py"))
    mod = importlib.util.module_from_spec(spec)
    sys.modules["hella_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def config(nbeam=1):
    return dict(INPUT="FILTERBANK", NCHAN=NCH, FCH1=500.0, FOFF=FOFF, TSAMP=TSAMP, GULP=CHUNK,
                DM_MIN=0, DM_MAX=2000, DM_TOL=1.25, PULSE_WIDTH=64, WIDTH_MIN=1, WIDTH_MAX=MAXW,
                SNR=SNR, NBEAM=nbeam, SCRUNCH=[(1, 1, 6.0)])


CYCLE = 8  # distinct chunks; the timed stream repeats them


def stream_params(seed=20261205):
    """20 pulses per chunk (DM U(50,1800), FWHM log-uniform 0.2-10 ms, S/N
    U(8,50)) over one cycle of CYCLE chunks, so every searched window holds
    about 20 arrivals; the 5% RFI channels"""
    rng = np.random.default_rng(seed)
    period = CYCLE * CHUNK * TSAMP
    pulses = []
    for _ in range(20 * CYCLE):
        pulses.append((float(rng.uniform(0.0, period)), float(rng.uniform(50, 1800)),
                       float(np.exp(rng.uniform(np.log(0.2), np.log(10.0)))),
                       float(rng.uniform(8, 50))))
    rfi = rng.choice(NCH, int(0.05 * NCH), replace=False)
    return pulses, rfi


def make_chunk(hamd, g, pulses, rfi, seed=20261205):
    """chunk g of the cyclic stream: the pulses repeat every CYCLE chunks (the
    copies one and two cycles earlier carry the sweeps that continue into the
    first chunks), so a stream that replays chunks 0..CYCLE-1 in a loop is
    continuous, sweeps included"""
    period = CYCLE * CHUNK * TSAMP
    per = [(t - k * period,) + p[1:] for k in (0, 1, 2) for p in pulses for t in (p[0],)]
    return hamd.synth(NCH, CHUNK, 500.0, FOFF, TSAMP, seed=seed, t_offset=g * CHUNK,
                      rfi_chan=rfi, rfi_mult=10.0, pulses=per,
                      burst_period=int(round(2.0 / TSAMP)), burst_phase=1000, burst_len=10,
                      burst_amp=6.0)


def pct(xs, q):
    return float(np.percentile(np.asarray(xs), q)) if len(xs) else None


# ---------------------------------------------------------------- serving
SERVE_CHUNK = 8192           # 1.074 s chunks of the streaming configuration
SERVE_BUDGET = 48 << 30      # HBM admission budget


def serve_config():
    c = config(1)
    c.pop("NBEAM")
    c.update(GULP=SERVE_CHUNK, CAND_CAP=65536)
    return c


def serve_min_gulp(hamd):
    plan = hamd.DedispPlan(NCH, TSAMP, 500.0, FOFF)
    plan.generate_dm_list(0.0, 2000.0, 64.0, 1.25)
    ntime = 32 * ((SERVE_CHUNK + plan.max_delay + MAXW) // 32 + 1)
    plan.close()
    return ntime // SERVE_CHUNK


def serve_pool(hamd, n, rfi, seed=20261204):
    """n consecutive chunks of the streaming configuration, which the beams
    cycle through: pulses by a Poisson process at 0.05/s with the production
    single-beam distributions (DM 50-1800, FWHM log-uniform 0.2-10 ms, S/N
    8-50), and a bandpass slope that drifts chunk by chunk over a triangle
    between -0.6 and +0.6 (channel gains 0.7-1.3 at the extremes), besides the
    persistent RFI channels and the broadband bursts"""
    rng = np.random.default_rng(seed)
    span = n * SERVE_CHUNK * TSAMP
    t, pulses = 0.0, []
    while True:
        t += float(rng.exponential(1.0 / 0.05))
        if t >= span:
            break
        pulses.append((t, float(rng.uniform(50, 1800)),
                       float(np.exp(rng.uniform(np.log(0.2), np.log(10.0)))),
                       float(rng.uniform(8, 50))))
    slopes = [0.6 * (1.0 - abs(4.0 * k / n - 2.0)) for k in range(n)]
    return [hamd.synth(NCH, SERVE_CHUNK, 500.0, FOFF, TSAMP, seed=777, t_offset=k * SERVE_CHUNK,
                       rfi_chan=rfi, rfi_mult=10.0, pulses=pulses,
                       burst_period=int(round(2.0 / TSAMP)), burst_phase=1000, burst_len=10,
                       burst_amp=6.0, bp_slope=slopes[k]) for k in range(n)]


def serve_saturation(hamd, pool, mg, nbeam, nchunk, **opts):
    """Warm nbeam beams (fill their windows), then push nchunk searched chunks
    per beam as fast as the server takes them: chunks/s and the latency of
    those chunks."""
    srv = hamd.Server(serve_config(), hbm_budget=SERVE_BUDGET, **opts)
    P = len(pool)
    ring = hamd.Ring(96, pool[0].nbytes, f"NCHAN {NCH}\nNBIT 8\nTSAMP {TSAMP * 1e6}\n")
    warm = [(0.0, b, (7 * b + k) % P, k, 0) for k in range(mg) for b in range(nbeam)]
    hamd.serve(srv, ring, warm, pool, time_scale=0.0)
    ring.close()
    n0, b0 = srv.stats()["chunks"], srv.stats()["batches"]
    ring = hamd.Ring(96, pool[0].nbytes, f"NCHAN {NCH}\n")
    ev = [(0.0, b, (7 * b + mg + k) % P, mg + k, 0) for k in range(nchunk) for b in range(nbeam)]
    t0 = time.perf_counter()
    hamd.serve(srv, ring, ev, pool, time_scale=0.0)
    dt = time.perf_counter() - t0
    ring.close()
    res = srv.results(with_candidates=False)[n0:]
    st = srv.stats()
    srv.close()
    assert all(r["searched"] for r in res)
    lat = [(r["t_done_ns"] - r["t_arrival_ns"]) * 1e-6 for r in res]
    rate = len(res) / dt
    return {"chunks": len(res), "seconds": dt, "chunks_per_s": rate,
            "realtime_beams": int(rate * SERVE_CHUNK * TSAMP),
            "mean_batch": len(res) / max(1, st["batches"] - b0),
            "latency_ms_p50": pct(lat, 50), "latency_ms_p99": pct(lat, 99)}


def serve_realtime(hamd, pool, mg, nbeam, seconds, seed=4, segment_s=60.0, budget=SERVE_BUDGET,
                   dwell_s=20.0, **opts):
    """nbeam warm beams stream chunks on the real-time cadence (random phase);
    beams leave after exponential dwell (mean 20 s) and a new beam joins in
    their place, filling its window before it is searched.  Latency = chunk
    arrival to candidates on the host.  The stream is replayed in segments of
    segment_s seconds through the same server (a progress line after each)."""
    rng = np.random.default_rng(seed)
    chunk_s = SERVE_CHUNK * TSAMP
    srv = hamd.Server(serve_config(), hbm_budget=budget, **opts)
    P = len(pool)
    ring = hamd.Ring(256, pool[0].nbytes, f"NCHAN {NCH}\n")
    warm = [(0.0, b, (3 * b + k) % P, k, 0) for k in range(mg) for b in range(nbeam)]
    t_w = time.perf_counter()
    hamd.serve(srv, ring, warm, pool, time_scale=0.0)
    t_w = time.perf_counter() - t_w
    ring.close()
    n0 = srv.stats()["chunks"]
    ev = []
    next_id = nbeam
    for slot in range(nbeam):
        # a warm beam continues its stream; when it leaves, a new beam joins
        bid, seq, t = slot, mg, float(rng.uniform(0, chunk_s))
        leave = float(rng.exponential(dwell_s)) if dwell_s > 0 else float("inf")
        while t < seconds:
            last = t + chunk_s > leave
            ev.append((t, bid, (5 * bid + seq) % P, seq, hamd.RING_LAST if last else 0))
            seq += 1
            t += chunk_s
            if last:
                bid, seq = next_id, 0
                next_id += 1
                leave = t + float(rng.exponential(dwell_s))
    ev.sort(key=lambda e: (e[0], e[1]))
    t0 = time.perf_counter()
    late, seg0 = 0.0, 0.0
    while seg0 < seconds:
        seg = [(t - seg0,) + e[1:] for e in ev for t in (e[0],) if seg0 <= t < seg0 + segment_s]
        ring = hamd.Ring(256, pool[0].nbytes, f"NCHAN {NCH}\n")
        late = max(late, hamd.serve(srv, ring, seg, pool, time_scale=1.0))
        ring.close()
        seg0 += segment_s
        if seg0 < seconds:
            print(f"# serving real time: {min(seg0, seconds):.0f} of {seconds:.0f} s streamed, "
                  f"{srv.stats()['chunks'] - n0} chunks", file=sys.stderr, flush=True)
    wall = time.perf_counter() - t0
    res = srv.results(with_candidates=False)[n0:]
    st = srv.stats()
    srv.close()
    srch = [(r["t_done_ns"] - r["t_arrival_ns"]) * 1e-6 for r in res if r["searched"]]
    allc = [(r["t_done_ns"] - r["t_arrival_ns"]) * 1e-6 for r in res]
    return {"beams": nbeam, "seconds": seconds, "wall_s": wall, "warmup_s": t_w,
            "chunks": len(res), "searched_chunks": len(srch), "joins": next_id - nbeam,
            "latency_ms_p50": pct(srch, 50), "latency_ms_p99": pct(srch, 99),
            "latency_ms_p99_all_chunks": pct(allc, 99),
            "p99_below_chunk": bool(srch) and pct(srch, 99) < chunk_s * 1e3,
            "writer_max_late_s": late, "max_resident": st["max_resident"],
            "beams_waited": st["beams_waited"], "gpu_busy_frac": st["busy_s"] / max(st["wall_s"], 1e-9),
            "work_set_GB": st["work_bytes"] / 2**30, "hbm_budget_GB": budget / 2**30,
            "resident_GB": st["peak_resident_bytes"] / 2**30}


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(hamd, chunk_tf, sample_rows=0):
    """The oracle (C restatement of the reference algorithm, OpenMP over
    channels, samples and DM rows) on one whole production chunk: flagger on
    the chunk, dedispersion of all trials over the chunk plus its tail, and
    the boxcar search of the whole plane.  sample_rows > 0 times that many DM
    trials instead and scales them to all trials."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_bind as orc

    plan = orc.Plan(NCH, TSAMP, 500.0, FOFF, 0.0, 2000.0, 64.0, 1.25)
    ndm = plan.dms.size
    t0 = time.perf_counter()
    flagged, _ = orc.flag_u8(chunk_tf, -0.05, 0.15, -1, -1, ((1, 1, 6.0),))
    t_flag = time.perf_counter() - t0
    # window: the flagged chunk repeated stands in for chunk + tail (the cost
    # does not depend on the bytes)
    ntime_dd = CHUNK + MAXW
    reps = (ntime_dd + plan.max_delay + CHUNK - 1) // CHUNK
    win = np.tile(flagged, (1, reps))
    rows = np.arange(ndm) if sample_rows <= 0 else np.linspace(0, ndm - 1, sample_rows).astype(int)
    t1 = time.perf_counter()
    D = orc.dedisperse_ft(win, plan.delta[rows], ntime_dd)
    t_dd = time.perf_counter() - t1
    del win
    t2 = time.perf_counter()
    orc.search(D, CHUNK, MAXW, 1, 0, SNR)
    t_s = time.perf_counter() - t2
    scale = ndm / len(rows)
    t_chunk = t_flag + (t_dd + t_s) * scale
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    what = ("the whole chunk" if scale == 1 else
            f"{len(rows)} of {ndm} DM trials, scaled to all trials")
    return {"value": ndm * CHUNK / t_chunk, "unit": "DM-trial·samples/s", "cores": cores,
            "kind": "port", "cpu_model": cpu_model(),
            "stages_s": {"flag": t_flag, "dedisperse": t_dd * scale, "search": t_s * scale},
            "sample": (f"oracle (C, -O3 -march=x86-64-v3, OpenMP, {cores} threads) on one "
                       f"production chunk: flagger ({t_flag:.2f} s), dedispersion of "
                       f"{len(rows)} trials x {ntime_dd} samples x {NCH} channels "
                       f"({t_dd:.2f} s), boxcar search ({t_s:.2f} s); {what}: "
                       f"{t_chunk:.1f} s per chunk")}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    # untimed steps before the timed ones; at 1024 ch / DM 2000 the first
    # three gulps only fill the window, and the clocks settle over the next few
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--batch-beams", type=int, default=4)
    ap.add_argument("--cpu-rows", type=int, default=0,
                    help="time this many DM trials of the CPU baseline and scale (0: all)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--tune", action="store_true")
    ap.add_argument("--tile", type=str, default=None,
                    help="force the dedispersion tile of the plain path: lanes,chunk,order,loader")
    ap.add_argument("--plain-only", action="store_true",
                    help="skip the batched path and the CPU baseline (profiling runs)")
    ap.add_argument("--calib-copy-mb", type=int, default=0,
                    help="also run one device-to-device copy of this size (counter calibration)")
    ap.add_argument("--no-serve", action="store_true", help="skip the streaming server runs")
    ap.add_argument("--serve-only", action="store_true",
                    help="only the streaming server runs (profiling)")
    ap.add_argument("--serve-beams", type=int, default=64,
                    help="beams of the saturation runs (plain vs batched server)")
    ap.add_argument("--serve-chunks", type=int, default=8, help="timed chunks per beam")
    ap.add_argument("--serve-rt-beams", type=int, default=256)
    ap.add_argument("--serve-rt-seconds", type=float, default=12.0,
                    help="length of the real-time run (0 skips it)")
    ap.add_argument("--serve-rt-segment", type=float, default=60.0,
                    help="the real-time stream is replayed in segments of this length")
    ap.add_argument("--no-saturation", action="store_true",
                    help="skip the saturation runs of the streaming server")
    ap.add_argument("--sustain-beams", type=int, default=1216,
                    help="also stream this many beams at the real cadence (beams sustained)")
    ap.add_argument("--sustain-seconds", type=float, default=10.0)
    ap.add_argument("--sustain-budget-gb", type=int, default=240,
                    help="HBM admission budget of the sustained run")
    ap.add_argument("--serve-batch", type=int, default=32)
    ap.add_argument("--serve-lanes", type=int, default=2)
    ap.add_argument("--uneven", type=int, default=0,
                    help="only the uneven multi-beam configuration with this many beams "
                         "(split as evenly as possible over the three kinds; 128 = 43/43/42)")
    ap.add_argument("--stress", type=int, default=0,
                    help="only the 4096-channel FP32 stress configuration, this many timed chunks")
    ap.add_argument("--out-dir", type=str, default=os.path.join(ROOT, "bench_out"),
                    help="where the autotuner records are written")
    args = ap.parse_args()
    global OUT_DIR
    OUT_DIR = args.out_dir
    os.makedirs(OUT_DIR, exist_ok=True)
    if args.gpus != 1:
        print("bench.py measures one GPU only: beams shard across GPUs with no exchange, "
              "so the multi-GPU figure is the per-GPU one times N; it is not run here",
              file=sys.stderr)
        sys.exit(2)

    hamd = load_pkg()
    if hamd.device_count() < 1:
        raise SystemExit("no GPU visible")
    pulses, rfi = stream_params()
    if args.uneven:
        print(json.dumps({"uneven": uneven(hamd, args.uneven, tune=args.tune)}), flush=True)
        return
    if args.stress:
        print(json.dumps({"stress": stress(hamd, args.stress, args.tune)}), flush=True)
        return
    if args.serve_only:
        print(json.dumps({"serving": serving(hamd, args, rfi)}), flush=True)
        return
    h = hamd.Hella(config(1))
    if args.tile:
        lanes, chunk, order, loader = (int(v) for v in args.tile.split(","))
        h.set_tile(lanes, chunk, order | (loader << 1))
        tile_plain = {"source": "--tile " + args.tile}
    else:
        tile_plain = tuning_record(h, "plain_1024ch_dm2000", args.tune)
    info = h.info(0)
    need = max(args.warmup, info["min_gulp"] + 1)
    nchunks = need + args.steps
    ndistinct = CYCLE
    t_gen = time.perf_counter()
    host_chunks = [make_chunk(hamd, g, pulses, rfi) for g in range(ndistinct)]
    dev = []
    for c in host_chunks:
        b = hamd.DeviceBuffer(c.nbytes)
        b.upload(c)
        dev.append(b)
    print(f"# generated {ndistinct} chunks in {time.perf_counter() - t_gen:.1f} s; "
          f"ndm={info['ndm']} max_delay={info['max_delay']} ntime={info['ntime']}",
          file=sys.stderr, flush=True)

    # ---------------- plain path
    for g in range(need):
        h.process_gulp(dev[g % ndistinct])
    if args.tune:
        h.autotune(os.path.join(OUT_DIR, "plain_1024ch_dm2000.json"))
        h.process_gulp(dev[0])
    lat, ncands, nclust = [], [], []
    hamd.device_sync()
    t0 = time.perf_counter()
    for k in range(args.steps):
        ts = time.perf_counter()
        ncands.append(h.process_gulp(dev[(need + k) % ndistinct]))
        lat.append((time.perf_counter() - ts) * 1e3)
        nclust.append(len(h.clusters()))
    hamd.device_sync()
    t_plain = (time.perf_counter() - t0) / args.steps
    crossings = h.crossings(0)
    print(f"# plain: {t_plain * 1e3:.2f} ms/chunk, candidates per step {ncands}, "
          f"clusters per step {nclust}", file=sys.stderr, flush=True)
    # per-stage device times (one extra step with events between stages)
    h.set_timing(True)
    h.process_gulp(dev[0])
    stages = dict(zip(["flag", "dedisperse", "search", "peaks"], h.stage_times()))
    h.set_timing(False)
    # live kernel timing with HIP events on the context stream
    ms_dd = h.bench_stage(0, 1, reps=5)
    ms_flag = h.bench_stage(0, 0, reps=3)
    ms_search = h.bench_stage(0, 2, reps=3)
    adds = info["ndm"] * info["ntime_dd"] * info["nlive"]
    achieved_add = adds / (ms_dd * 1e-3) / 1e12
    flag_bytes = 3 * NCH * CHUNK * 1  # one stats read, one read+write apply (u8 equivalents)
    # two reads of the u32 plane (statistics, then the sigma0 / threshold pass)
    search_bytes = 2 * info["ndm"] * CHUNK * 4
    # PCIe-inclusive rate: the same steps from host (pageable) chunks, the
    # per-chunk host-to-device copy inside the timed region
    hamd.device_sync()
    tp = time.perf_counter()
    npc = min(args.steps, 5)
    for k in range(npc):
        h.process_gulp(host_chunks[(need + args.steps + k) % ndistinct])
    hamd.device_sync()
    t_pcie = (time.perf_counter() - tp) / npc
    pcie = {"ms_per_step": t_pcie * 1e3, "value": info["ndm"] * CHUNK / t_pcie, "steps": npc,
            "note": "host (pageable) chunk copied to the device inside each step"}
    h.close()
    if args.calib_copy_mb:
        nbytes = args.calib_copy_mb << 20
        a, b = hamd.DeviceBuffer(nbytes), hamd.DeviceBuffer(nbytes)
        hamd.lib().hella_memcpy(b.ptr, a.ptr, nbytes, 3)
        hamd.device_sync()
        a.free()
        b.free()
    if args.plain_only:
        print(json.dumps({"ms_per_step": t_plain * 1e3, "stages_ms": stages,
                          "kernels_ms": {"dedisperse": ms_dd, "flagger": ms_flag,
                                         "search": ms_search},
                          "dedisp_Tadd_s": achieved_add}), flush=True)
        return

    # ---------------- batched path (graph-captured), several beams
    nb = args.batch_beams
    hb = hamd.Hella(config(nb))
    tile_batched = tuning_record(hb, f"batched_{nb}x1024ch_dm2000", args.tune)
    gulp_dev = []
    for i in range(ndistinct):
        buf = hamd.DeviceBuffer(host_chunks[i].nbytes * nb)
        buf.upload(np.concatenate([host_chunks[(i + j) % ndistinct].ravel() for j in range(nb)]))
        gulp_dev.append(buf)
    for g in range(need):
        hb.process_gulp(gulp_dev[g % ndistinct], batched=True)
    if args.tune:
        hb.autotune(os.path.join(OUT_DIR, f"batched_{nb}x1024ch_dm2000.json"))
        hb.process_gulp(gulp_dev[0], batched=True)
    blat = []
    hamd.device_sync()
    t0 = time.perf_counter()
    for k in range(args.steps):
        ts = time.perf_counter()
        hb.process_gulp(gulp_dev[(need + k) % ndistinct], batched=True)
        blat.append((time.perf_counter() - ts) * 1e3)
    hamd.device_sync()
    t_batch = (time.perf_counter() - t0) / args.steps
    hb.close()

    chunk_s = CHUNK * TSAMP
    value = info["ndm"] * CHUNK / t_plain
    out = {
        "metric": "real-time beams/GPU; DM-trial·samples/s; p99 chunk latency, batched vs plain",
        "value": value,
        "unit": "DM-trial·samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": need,
        "ms_per_step": t_plain * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic (seeded generator: Gaussian noise, dispersed pulses, RFI)",
        "config": {"workload": "plain single beam: 1024 ch 500->400 MHz, 131.072 us, "
                               "65536-sample chunks + max-DM tail, DM 0-2000 "
                               f"({info['ndm']} trials), boxcar widths 1..256",
                   "ndm": info["ndm"], "max_delay": info["max_delay"], "chunk": CHUNK,
                   "nchan": NCH, "scrunch": "1 1 6"},
        "candidates_per_step": ncands, "clusters_per_step": nclust,
        "crossings_last_step": crossings,
        "pcie_inclusive": pcie,
        "realtime_factor": chunk_s / t_plain,
        # throughput bounds of the production chunk (chunk duration over the
        # step time), not sustained counts; the HBM bound: windows of this
        # configuration that fit the 288 GB next to one work set
        "realtime_beams": {"plain_step_bound": int(chunk_s / t_plain),
                           "batched_step_bound": int(nb * chunk_s / t_batch),
                           "batched_beams_run": nb,
                           "hbm_bound": int((288e9 - 1.0e9) // (NCH * info["ntime"]))},
        "latency_ms": {"plain": {"p50": pct(lat, 50), "p99": pct(lat, 99)},
                       "batched": {"p50": pct(blat, 50), "p99": pct(blat, 99), "beams": nb}},
        "stages_ms": stages,
        "kernels_ms": {"dedisperse": ms_dd, "flagger": ms_flag, "search": ms_search},
        "dedisp_tile": {"plain": tile_plain, "batched": tile_batched},
        "roofline": {"bound": "valu", "kernel": "k_dedisp", "achieved": achieved_add,
                     "peak": PEAK_ADD_T, "unit": "Tadd/s", "frac": achieved_add / PEAK_ADD_T,
                     "traffic": profiled_traffic("k_dedisp"),
                     "traffic_source": os.path.relpath(TRAFFIC_PROFILE, ROOT),
                     "lds": lds_roof(ms_dd)},
        "roofline_search": {"bound": "hbm", "kernels": "boxcar statistics, sigma0 pass, "
                                                       "threshold count and compaction",
                            "achieved": search_bytes / (ms_search * 1e-3) / 1e9,
                            "peak": PEAK_HBM_GBS, "unit": "GB/s",
                            "frac": search_bytes / (ms_search * 1e-3) / 1e9 / PEAK_HBM_GBS,
                            "algorithmic_bytes": search_bytes,
                            "traffic": profiled_traffic("search_chunk"),
                            "traffic_source": os.path.relpath(TRAFFIC_PROFILE, ROOT)},
        "roofline_flagger": {"bound": "hbm", "achieved": flag_bytes / (ms_flag * 1e-3) / 1e9,
                             "peak": PEAK_HBM_GBS, "unit": "GB/s",
                             "frac": flag_bytes / (ms_flag * 1e-3) / 1e9 / PEAK_HBM_GBS,
                             "traffic": profiled_traffic("flagger_chunk"),
                             "traffic_source": os.path.relpath(TRAFFIC_PROFILE, ROOT)},
    }
    if not args.no_serve:
        out["serving"] = serving(hamd, args, rfi)
        sv = out["serving"]
        if sv["saturation"]:
            out["realtime_beams"]["serving_plain_capacity"] = sv["saturation"]["plain"]["realtime_beams"]
            out["realtime_beams"]["serving_batched_capacity"] = \
                sv["saturation"]["batched"]["realtime_beams"]
        # beams sustained (the largest count streamed at the real cadence
        # with p99 chunk-to-candidate latency below the chunk duration), on
        # the streaming configuration's 8192-sample chunks
        sus = sv.get("sustained")
        if sus:
            out["realtime_beams"]["sustained"] = sus["beams"] if sus["p99_below_chunk"] else None
            out["realtime_beams"]["sustained_p99_ms"] = sus["latency_ms_p99"]
            out["realtime_beams"]["sustained_chunk_ms"] = sv["chunk_s"] * 1e3
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(hamd, host_chunks[0], args.cpu_rows)
    print(json.dumps(out), flush=True)


def stress(hamd, gulps, tune=False):
    """The 4096-channel FP32 stress configuration, one beam on the plain path:
    500->400 MHz, 65.536 us, 262144-sample chunks (17.18 s), DM 0-5000, boxcar
    widths 1..1024, N(0,1) input with 20% of the channels at 30 sigma and
    broadband bursts every second, a 1x1 scrunch at 6 sigma."""
    nch, tsamp, W = 4096, 65.536e-6, 262144
    cfg = dict(INPUT="FILTERBANK", INPUT_DTYPE="F32", NCHAN=nch, FCH1=500.0, FOFF=-100.0 / nch,
               TSAMP=tsamp, GULP=W, DM_MIN=0, DM_MAX=5000, DM_TOL=1.25, PULSE_WIDTH=64,
               WIDTH_MIN=1, WIDTH_MAX=1024, SNR=7.0, NBEAM=1, SCRUNCH=[(1, 1, 6.0)])
    h = hamd.Hella(cfg)
    tile = tuning_record(h, "stress_4096ch_dm5000", tune)
    info = h.info(0)
    rng = np.random.default_rng(5)
    rfi = rng.choice(nch, nch // 5, replace=False)
    pulses = [(float(rng.uniform(1, 16)), float(rng.uniform(100, 4500)), 2.0, 30.0) for _ in range(5)]
    t0 = time.perf_counter()
    bufs = []
    for i in range(2):
        c = hamd.synth(nch, W, 500.0, -100.0 / nch, tsamp, seed=77 + i, t_offset=i * W,
                       dtype="f32", mean=0.0, sigma=1.0, rfi_chan=rfi, rfi_mult=30.0,
                       pulses=pulses, burst_period=int(round(1.0 / tsamp)), burst_phase=500,
                       burst_len=10, burst_amp=6.0)
        b = hamd.DeviceBuffer(c.nbytes)
        b.upload(c)
        bufs.append(b)
        del c
    print(f"# stress: input ready in {time.perf_counter() - t0:.1f} s; ndm={info['ndm']} "
          f"max_delay={info['max_delay']} min_gulp={info['min_gulp']}", file=sys.stderr, flush=True)
    need = info["min_gulp"] + 1
    for g in range(need):
        h.process_gulp(bufs[g % 2])
    if tune:
        h.autotune(os.path.join(OUT_DIR, "stress_4096ch_dm5000.json"))
        h.process_gulp(bufs[need % 2])
        need += 1
    hamd.device_sync()
    times, crossings, ncand = [], [], []
    for g in range(gulps):
        ts = time.perf_counter()
        ncand.append(h.process_gulp(bufs[(need + g) % 2]))
        times.append((time.perf_counter() - ts) * 1e3)
        crossings.append(h.crossings(0))
    h.set_timing(True)
    h.process_gulp(bufs[0])
    stages = dict(zip(["flag", "dedisperse", "search", "peaks"], h.stage_times()))
    h.close()
    chunk_s = W * tsamp
    ms = float(np.mean(times))
    adds = info["ndm"] * info["ntime_dd"] * info["nlive"]
    return {"workload": "stress: 4096 ch FP32, 262144-sample chunks (17.18 s), DM 0-5000 "
                        f"({info['ndm']} trials, max delay {info['max_delay']}), widths 1..1024",
            "ms_per_chunk": ms, "realtime_factor": chunk_s / (ms * 1e-3),
            "beams_per_gpu": int(chunk_s / (ms * 1e-3)), "stages_ms": stages,
            "dedisp_Tadd_s": adds / (stages["dedisperse"] * 1e-3) / 1e12,
            "raw_crossings_per_chunk": crossings, "candidates_per_chunk": ncand,
            "dedisp_tile": tile}


def uneven(hamd, nbeam, reps=3, tune=False):
    """The uneven multi-beam configuration: nbeam beams split as evenly as
    possible (earlier kinds first) over 512 ch / DM 0-1000, 1024 ch / DM
    0-2000 and 2048 ch / DM 0-3000 (500->400 MHz, 131.072 us), 65536-sample
    chunks with each beam's own max-DM window, seeded killmasks of 0-10% of
    the channels, pulses in a quarter of the beams.  Once every window is
    full, chunks of all beams through the batched graph-captured chain (after
    the autotuner, for up to 32 beams) against the plain per-beam sequence on
    the same chunks.  Every chunk carries the same synthetic samples."""
    groups = [(512, 1000.0), (1024, 2000.0), (2048, 3000.0)]
    counts = [nbeam // 3 + (1 if k < nbeam % 3 else 0) for k in range(3)]
    beams = [g for g, n in zip(groups, counts) for _ in range(n)]
    nb = len(beams)
    cfg = dict(INPUT="FILTERBANK", NCHAN=1024, FCH1=500.0, FOFF=-100.0 / 1024, TSAMP=TSAMP,
               GULP=CHUNK, DM_MIN=0, DM_MAX=2000, DM_TOL=1.25, PULSE_WIDTH=64, WIDTH_MIN=1,
               WIDTH_MAX=256, SNR=7.0, NBEAM=nb, SCRUNCH=[(1, 1, 6.0)])
    rng = np.random.default_rng(3)
    kms = []
    for nch, _ in beams:
        km = np.ones(nch, np.int32)
        km[rng.choice(nch, int(rng.uniform(0, 0.1) * nch), replace=False)] = 0
        kms.append(km)
    pulses = [[(float(rng.uniform(1, 8)), float(rng.uniform(50, 0.9 * dmx)), 2.0, 15.0)]
              if rng.random() < 0.25 else [] for _, dmx in beams]
    t0 = time.perf_counter()
    sizes = [nch * CHUNK for nch, _ in beams]
    host = np.empty(sum(sizes), np.uint8)
    off = 0
    for b, (nch, _) in enumerate(beams):
        host[off:off + sizes[b]] = hamd.synth(nch, CHUNK, 500.0, -100.0 / nch, TSAMP, seed=900 + b,
                                              pulses=pulses[b]).ravel()
        off += sizes[b]
    buf = hamd.DeviceBuffer(host.nbytes)
    buf.upload(host)
    del host
    print(f"# uneven: {nb} beams ({counts[0]}/{counts[1]}/{counts[2]}), input ready in "
          f"{time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)

    tile = {}

    def ctx():
        # the batched form's tuned tile serves both forms (same shapes)
        h = hamd.Hella(cfg)
        for b, (nch, dmx) in enumerate(beams):
            h.set_beam(b, NCHAN=nch, FOFF=-100.0 / nch, DM_MAX=dmx)
            h.set_killmask(b, kms[b])
        tile.update(tuning_record(h, f"uneven_{nb}beams", tune))
        return h

    def run(h, batched):
        need = max(h.info(b)["min_gulp"] for b in range(nb)) + 1
        for g in range(need):
            h.process_gulp(buf, batched=batched)
        if batched and tune:
            h.autotune(os.path.join(OUT_DIR, f"uneven_{nb}beams.json"))
        h.process_gulp(buf, batched=batched)
        hamd.device_sync()
        print(f"# uneven: {'batched' if batched else 'plain'} warm after {need + 1} chunks",
              file=sys.stderr, flush=True)
        ts = time.perf_counter()
        nc = []
        for k in range(reps):
            nc.append(h.process_gulp(buf, batched=batched))
        hamd.device_sync()
        return (time.perf_counter() - ts) / reps, nc, h.peaks()

    hb = ctx()
    t_b, nc_b, pk_b = run(hb, True)
    info = [hb.info(b) for b in range(nb)]
    hb.close()
    hp = ctx()
    t_p, nc_p, pk_p = run(hp, False)
    hp.close()
    buf.free()
    adds = sum(i["ndm"] * i["ntime_dd"] * i["nlive"] for i in info)
    return {"workload": f"uneven multi-beam: {nb} beams ({counts[0]} x 512 ch/DM 1000, "
                        f"{counts[1]} x 1024 ch/DM 2000, {counts[2]} x 2048 ch/DM 3000), "
                        "killmasks 0-10%, 65536-sample chunks",
            "batched_ms_per_chunk": t_b * 1e3, "plain_ms_per_chunk": t_p * 1e3,
            "batched_over_plain": t_p / t_b, "dedisp_adds_per_chunk": adds,
            "batched_Tadd_s_end_to_end": adds / t_b / 1e12,
            "realtime_factor_batched": CHUNK * TSAMP / t_b,
            "candidates": {"batched": nc_b, "plain": nc_p},
            "last_chunk_identical": bool(np.array_equal(pk_b, pk_p)),
            "dedisp_tile": tile}


def serving(hamd, args, rfi):
    mg = serve_min_gulp(hamd)
    t0 = time.perf_counter()
    pool = serve_pool(hamd, 24, rfi)
    print(f"# serving: pool in {time.perf_counter() - t0:.1f} s, windows fill in {mg} chunks",
          file=sys.stderr, flush=True)
    sat = {}
    for mode, opts in (("plain", dict(plain=True)),
                       ("batched", dict(max_batch=args.serve_batch, lanes=args.serve_lanes))):
        if args.no_saturation:
            break
        sat[mode] = serve_saturation(hamd, pool, mg, args.serve_beams, args.serve_chunks, **opts)
        print(f"# serving {mode}: {sat[mode]['chunks_per_s']:.0f} chunks/s -> "
              f"{sat[mode]['realtime_beams']} real-time beams", file=sys.stderr, flush=True)
    rt = None
    if args.serve_rt_seconds > 0:
        rt = serve_realtime(hamd, pool, mg, args.serve_rt_beams, args.serve_rt_seconds,
                            segment_s=args.serve_rt_segment, max_batch=args.serve_batch,
                            lanes=args.serve_lanes)
        print(f"# serving real time: {rt['beams']} beams, p50 {rt['latency_ms_p50']} ms, "
              f"p99 {rt['latency_ms_p99']} ms", file=sys.stderr, flush=True)
    # beams sustained: the real-time cadence with as many beams as the HBM
    # holds and the GPU keeps up with; sustained when p99 < the chunk
    sus = None
    if args.sustain_beams > 0:
        # every beam stays (no departures): every chunk of every beam is searched
        sus = serve_realtime(hamd, pool, mg, args.sustain_beams, args.sustain_seconds,
                             segment_s=args.serve_rt_segment, budget=args.sustain_budget_gb << 30,
                             dwell_s=0.0, max_batch=args.serve_batch, lanes=args.serve_lanes)
        print(f"# serving sustained: {sus['beams']} beams, p50 {sus['latency_ms_p50']} ms, "
              f"p99 {sus['latency_ms_p99']} ms, GPU busy {sus['gpu_busy_frac']:.2f}",
              file=sys.stderr, flush=True)
    return {"workload": f"streaming: 1024 ch, {SERVE_CHUNK}-sample chunks "
                        f"({SERVE_CHUNK * TSAMP:.3f} s), DM 0-2000, widths 1..256, "
                        f"HBM budget {SERVE_BUDGET >> 30} GB, pulses 0.05/s per beam, "
                        "bandpass slope drifting between -0.6 and +0.6",
            "chunk_s": SERVE_CHUNK * TSAMP, "window_fill_chunks": mg,
            "saturation": sat,
            "batched_over_plain": (sat["batched"]["chunks_per_s"] / sat["plain"]["chunks_per_s"]
                                   if sat else None),
            "realtime": rt, "sustained": sus, "max_batch": args.serve_batch,
            "lanes": args.serve_lanes}


if __name__ == "__main__":
    main()
