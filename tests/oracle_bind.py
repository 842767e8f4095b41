"""ctypes binding of oracle/liboracle.so -- the CPU checker.

Test infrastructure: only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg use it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
LIB = os.path.join(ORACLE_DIR, "liboracle.so")

if not os.path.exists(LIB):
    subprocess.run(["make", "-C", ORACLE_DIR], check=True, capture_output=True)

L = C.CDLL(LIB)

_VP, _SZ, _F, _D, _I = C.c_void_p, C.c_size_t, C.c_float, C.c_double, C.c_int


class Scrunch(C.Structure):
    _fields_ = [("tscrunch", _I), ("fscrunch", _I), ("thresh", _F)]


class FlagCfg(C.Structure):
    _fields_ = [("nch", _I), ("W", _I), ("spec_min", _F), ("spec_max", _F),
                ("flag1", _I), ("flag2", _I), ("nscrunch", _I), ("scrunch", C.POINTER(Scrunch)),
                ("robust", _I), ("u8", _I)]


class SearchCfg(C.Structure):
    _fields_ = [("ndm", _I), ("ntime_dd", _I), ("nout", _I), ("maxW", _I), ("minW", _I),
                ("mode", _I), ("snr", _F), ("max_giants", _I)]


class Cand(C.Structure):
    _fields_ = [("snr", _F), ("samp", _I), ("width", _I), ("dm_idx", _I)]


class Cluster(C.Structure):
    _fields_ = [("rep", _I), ("count", _I), ("samp_min", _I), ("samp_max", _I),
                ("dm_min", _I), ("dm_max", _I)]


def _d(name, res, *args):
    f = getattr(L, name)
    f.restype = res
    f.argtypes = list(args)


_d("orc_csum", _D, _VP, _SZ)
_d("orc_delay_table", None, _SZ, _F, _F, _F, _VP)
_d("orc_dm_list", _SZ, _F, _F, _D, _D, _D, _D, _SZ, _D, _VP, _SZ)
_d("orc_max_delay", _SZ, _VP, _SZ, _VP, _SZ)
_d("orc_delays", None, _VP, _SZ, _VP, _SZ, _VP)
_d("orc_dedisperse", None, _VP, _SZ, _SZ, _SZ, _VP, _SZ, _VP, _SZ, _VP, _SZ)
_d("orc_stats", None, _VP, _SZ, _SZ, _SZ, C.POINTER(_D), C.POINTER(_D))
_d("orc_stats_f", None, _VP, _SZ, _SZ, _SZ, C.POINTER(_F), C.POINTER(_F))
_d("orc_bandpass", None, _VP, _I, _I, _VP)
_d("orc_divide_bp", None, _VP, _I, _I, _VP)
_d("orc_normalize", None, _VP, _I, _I)
_d("orc_normalize_robust", None, _VP, _I, _I, _VP)
_d("orc_spec_flag", None, _VP, _I, _I, _F, _F)
_d("orc_box_mean", None, _VP, _I, _I, _I, _I, _VP)
_d("orc_scrunch_flag", None, _VP, _I, _I, _I, _I, _F, _I, _I, _VP, _VP)
_d("orc_ts_correct", None, _VP, _I, _I, _VP)
_d("orc_chan_sum", _D, _VP, _I, _I, _I)
_d("orc_quantize", None, _VP, _I, _I, _VP, _SZ)
_d("orc_ingest_u8_tf", None, _VP, _I, _I, _VP)
_d("orc_ingest_f32_tf", None, _VP, _I, _I, _VP)
_d("orc_flag", None, C.POINTER(FlagCfg), _VP, _VP, _SZ, _VP, _VP)
_d("orc_nwidth", _I, C.POINTER(SearchCfg))
_d("orc_width_of", _I, C.POINTER(SearchCfg), _I)
_d("orc_hella_kernel", None, _I, _VP)
_d("orc_search", C.c_long, C.POINTER(SearchCfg), _VP, _SZ, _VP, _VP, C.POINTER(_F),
   C.POINTER(_I), C.POINTER(Cand), _VP)
_d("orc_cluster", _I, C.POINTER(Cand), _I, _I, _I, C.POINTER(Cluster))


def _p(a):
    return a.ctypes.data


# ---------------------------------------------------------------- plan
def delay_table(nch, dt, f0, df):
    out = np.zeros(nch, np.float32)
    L.orc_delay_table(nch, dt, f0, df, _p(out))
    return out


def dm_list(dm_start, dm_end, dt, ti_us, f0, df, nch, tol):
    n = L.orc_dm_list(dm_start, dm_end, dt, ti_us, f0, df, nch, tol, None, 0)
    out = np.zeros(n, np.float32)
    L.orc_dm_list(dm_start, dm_end, dt, ti_us, f0, df, nch, tol, _p(out), n)
    return out


def max_delay(dms, delay):
    return int(L.orc_max_delay(_p(dms), dms.size, _p(delay), delay.size))


def delays(dms, delay):
    out = np.zeros((dms.size, delay.size), np.int32)
    L.orc_delays(_p(dms), dms.size, _p(delay), delay.size, _p(out))
    return out


class Plan:
    """Oracle plan: delay table, DM list, max delay and integer delays."""

    def __init__(self, nch, dt, f0, df, dm_start, dm_end, ti_us, tol):
        self.nch = nch
        self.delay = delay_table(nch, dt, f0, df)
        self.dms = dm_list(dm_start, dm_end, float(np.float32(dt)), ti_us, float(np.float32(f0)),
                           float(np.float32(df)), nch, float(np.float32(tol)))
        self.max_delay = max_delay(self.dms, self.delay)
        self.delta = delays(self.dms, self.delay)


# ---------------------------------------------------------------- dedispersion
def dedisperse_tf(data_tf, delta, nout, killmask=None):
    """data_tf: u8 [nsamps][nch]; returns u32 [ndm][nout]."""
    data_tf = np.ascontiguousarray(data_tf, np.uint8)
    ns, nch = data_tf.shape
    ndm = delta.shape[0]
    out = np.zeros((ndm, nout), np.uint32)
    km = None
    if killmask is not None:
        km = np.ascontiguousarray(killmask, np.uint8)
    L.orc_dedisperse(_p(data_tf), nch, 1, nch, _p(np.ascontiguousarray(delta, np.int32)), ndm,
                     _p(km) if km is not None else None, nout, _p(out), nout)
    return out


def dedisperse_ft(data_ft, delta, nout, killmask=None):
    """data_ft: u8 [nch][ncols]; returns u32 [ndm][nout]."""
    data_ft = np.ascontiguousarray(data_ft, np.uint8)
    nch, ncol = data_ft.shape
    ndm = delta.shape[0]
    out = np.zeros((ndm, nout), np.uint32)
    km = None if killmask is None else np.ascontiguousarray(killmask, np.uint8)
    L.orc_dedisperse(_p(data_ft), 1, ncol, nch, _p(np.ascontiguousarray(delta, np.int32)), ndm,
                     _p(km) if km is not None else None, nout, _p(out), nout)
    return out


# ---------------------------------------------------------------- reductions / stats
def csum(x):
    x = np.ascontiguousarray(x, np.float64)
    return float(L.orc_csum(_p(x), x.size))


def stats(x):
    x = np.ascontiguousarray(x, np.float32)
    h, w = x.shape
    m, v = _D(), _D()
    L.orc_stats(_p(x), h, w, w, C.byref(m), C.byref(v))
    return m.value, v.value


def stats_f(x):
    x = np.ascontiguousarray(x, np.float32)
    h, w = x.shape
    m, s = _F(), _F()
    L.orc_stats_f(_p(x), h, w, w, C.byref(m), C.byref(s))
    return m.value, s.value


# ---------------------------------------------------------------- flagger
def bandpass(x):
    x = np.ascontiguousarray(x, np.float32)
    nch, W = x.shape
    bp = np.zeros(nch, np.float32)
    L.orc_bandpass(_p(x), nch, W, _p(bp))
    return bp


def divide_bp(x, bp):
    x = np.array(x, np.float32, copy=True)
    L.orc_divide_bp(_p(x), x.shape[0], x.shape[1], _p(np.ascontiguousarray(bp, np.float32)))
    return x


def normalize(x):
    x = np.array(x, np.float32, copy=True)
    L.orc_normalize(_p(x), x.shape[0], x.shape[1])
    return x


def normalize_robust(x):
    """per-channel median / MAD normalisation; returns (plane, [nch][2] m, s)"""
    x = np.array(x, np.float32, copy=True)
    ms = np.zeros((x.shape[0], 2), np.float32)
    L.orc_normalize_robust(_p(x), x.shape[0], x.shape[1], _p(ms))
    return x, ms


def spec_flag(x, smin, smax):
    x = np.array(x, np.float32, copy=True)
    L.orc_spec_flag(_p(x), x.shape[0], x.shape[1], smin, smax)
    return x


def box_mean(x, ts, fs):
    x = np.ascontiguousarray(x, np.float32)
    out = np.zeros_like(x)
    L.orc_box_mean(_p(x), x.shape[0], x.shape[1], ts, fs, _p(out))
    return out


def scrunch_flag(x, ts, fs, thresh, flag1=-1, flag2=-1, want_spec=False):
    x = np.array(x, np.float32, copy=True)
    cnt = np.zeros(x.shape[0], np.uint32)
    spec = np.zeros(x.shape[0], np.float32)
    L.orc_scrunch_flag(_p(x), x.shape[0], x.shape[1], ts, fs, thresh, flag1, flag2, _p(cnt),
                       _p(spec))
    return (x, cnt, spec) if want_spec else (x, cnt)


def chan_sum(x, t):
    x = np.ascontiguousarray(x, np.float32)
    return float(L.orc_chan_sum(_p(x), x.shape[0], x.shape[1], t))


def ts_correct(x):
    x = np.array(x, np.float32, copy=True)
    ts = np.zeros(x.shape[1], np.float32)
    L.orc_ts_correct(_p(x), x.shape[0], x.shape[1], _p(ts))
    return x, ts


def quantize(x):
    x = np.ascontiguousarray(x, np.float32)
    out = np.zeros(x.shape, np.uint8)
    L.orc_quantize(_p(x), x.shape[0], x.shape[1], _p(out), x.shape[1])
    return out


def ingest_u8_tf(data_tf):
    data_tf = np.ascontiguousarray(data_tf, np.uint8)
    W, nch = data_tf.shape
    x = np.zeros((nch, W), np.float32)
    L.orc_ingest_u8_tf(_p(data_tf), nch, W, _p(x))
    return x


def ingest_f32_tf(data_tf):
    data_tf = np.ascontiguousarray(data_tf, np.float32)
    W, nch = data_tf.shape
    x = np.zeros((nch, W), np.float32)
    L.orc_ingest_f32_tf(_p(data_tf), nch, W, _p(x))
    return x


def flag(x_ft, spec_min=-0.05, spec_max=0.15, flag1=-1, flag2=-1, scrunches=(), spec=None,
         robust=False, u8=False):
    """Whole per-beam flagger on a float FT plane; returns (u8 FT, flagcount).
    spec: a float32 [nch] array that accumulates the flag spectrum."""
    x = np.array(x_ft, np.float32, copy=True)
    nch, W = x.shape
    sc = (Scrunch * max(1, len(scrunches)))()
    for i, (ts, fs, th) in enumerate(scrunches):
        sc[i] = Scrunch(ts, fs, th)
    cfg = FlagCfg(nch, W, spec_min, spec_max, flag1, flag2, len(scrunches), sc, 1 if robust else 0,
                  1 if u8 else 0)
    out = np.zeros((nch, W), np.uint8)
    cnt = np.zeros(nch, np.uint32)
    if spec is not None:
        assert spec.dtype == np.float32 and spec.size == nch and spec.flags.c_contiguous
    L.orc_flag(C.byref(cfg), _p(x), _p(out), W, _p(cnt), _p(spec) if spec is not None else None)
    return out, cnt


def flag_u8(gulp_tf, spec_min=-0.05, spec_max=0.15, flag1=-1, flag2=-1, scrunches=(),
            spec=None, robust=False):
    """The flagger on an 8-bit time-major gulp ([W][nch] uint8), as the
    pipeline takes it: byte-histogram order where it applies.  Returns
    (u8 FT, flagcount)."""
    return flag(ingest_u8_tf(gulp_tf), spec_min, spec_max, flag1, flag2, scrunches, spec=spec,
                robust=robust, u8=True)


# ---------------------------------------------------------------- search
MODE_BOXCAR, MODE_HELLA = 0, 1


def nwidth(mode, minW, maxW):
    cfg = SearchCfg(0, 0, 0, maxW, minW, mode, 0.0, 0)
    return int(L.orc_nwidth(C.byref(cfg)))


def hella_kernel(sm):
    k = np.zeros((3, 2 * sm + 1), np.float32)
    L.orc_hella_kernel(sm, _p(k))
    return k


def search(D, nout, maxW, minW=1, mode=MODE_BOXCAR, snr=7.0, max_giants=10000,
           want_planes=False):
    """Search one dedispersed plane D (u32 [ndm][ntime_dd])."""
    D = np.ascontiguousarray(D, np.uint32)
    ndm, ntdd = D.shape
    cfg = SearchCfg(ndm, ntdd, nout, maxW, minW, mode, snr, max_giants)
    nw = int(L.orc_nwidth(C.byref(cfg)))
    mu = np.zeros(nw, np.float32)
    sg = np.zeros(nw, np.float32)
    s0 = _F()
    gated = _I()
    cands = (Cand * max(1, max_giants))()
    rows = ndm - 2 if mode == MODE_HELLA else ndm
    planes = np.zeros((nw, rows, nout), np.float32) if want_planes else None
    n = L.orc_search(C.byref(cfg), _p(D), ntdd, _p(mu), _p(sg), C.byref(s0), C.byref(gated),
                     cands, _p(planes) if want_planes else None)
    k = min(n, max_giants)
    arr = np.array([(cands[i].snr, cands[i].samp, cands[i].width, cands[i].dm_idx)
                    for i in range(k)],
                   dtype=[("snr", "f4"), ("samp", "i4"), ("width", "i4"), ("dm_idx", "i4")])
    return dict(count=int(n), cands=arr, mu=mu, sigma=sg, sigma0=s0.value,
                gated=bool(gated.value), planes=planes)


def cluster(cands, cell_t, cell_dm):
    n = len(cands)
    cs = (Cand * max(1, n))()
    for i, c in enumerate(cands):
        cs[i] = Cand(float(c["snr"]), int(c["samp"]), int(c["width"]), int(c["dm_idx"]))
    out = (Cluster * max(1, n))()
    k = L.orc_cluster(cs, n, cell_t, cell_dm, out)
    return [dict(rep=out[i].rep, count=out[i].count, samp_min=out[i].samp_min,
                 samp_max=out[i].samp_max, dm_min=out[i].dm_min, dm_max=out[i].dm_max)
            for i in range(k)]
