"""ctypes binding of libhella_amd.so (the C ABI declared in include/dedisp.h
and include/hella.h).

The library is the product: every compute call below runs hand-written HIP
kernels for gfx950.  Loading fails loudly if the library has not been built
(``python -c "import __graft_entry__ as g; g.build()"``).
"""
from __future__ import annotations

import ctypes as C
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libhella_amd.so")

_lib = None


def lib():
    """Load libhella_amd.so once; raise if it is missing."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"libhella_amd.so not built: {LIB_PATH}")
        _lib = C.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


# dedisp flags / errors (include/dedisp.h)
DEDISP_NO_ERROR = 0
DEDISP_HOST_POINTERS = 1 << 1
DEDISP_DEVICE_POINTERS = 1 << 2
DEDISP_WAIT = 1 << 3
DEDISP_ASYNC = 1 << 4

_VP = C.c_void_p
_SZ = C.c_ulong
_F = C.c_float
_I = C.c_int


def _declare(L):
    def d(name, res, *args):
        f = getattr(L, name)
        f.restype = res
        f.argtypes = list(args)

    d("dedisp_create_plan", _I, C.POINTER(_VP), _SZ, _F, _F, _F)
    d("dedisp_destroy_plan", None, _VP)
    d("dedisp_generate_dm_list", _I, _VP, _F, _F, _F, _F)
    d("dedisp_set_dm_list", _I, _VP, C.POINTER(_F), _SZ)
    d("dedisp_set_killmask", _I, _VP, C.POINTER(_I))
    d("dedisp_set_gulp_size", _I, _VP, _SZ)
    d("dedisp_get_gulp_size", _SZ, _VP)
    d("dedisp_get_dm_list", C.POINTER(_F), _VP)
    d("dedisp_get_dm_count", _SZ, _VP)
    d("dedisp_get_max_delay", _SZ, _VP)
    d("dedisp_get_dm_delay", _SZ, _VP, _I)
    d("dedisp_get_channel_count", _SZ, _VP)
    d("dedisp_get_killmask", C.POINTER(_I), _VP)
    d("dedisp_get_dt", _F, _VP)
    d("dedisp_get_f0", _F, _VP)
    d("dedisp_get_df", _F, _VP)
    d("dedisp_execute_adv", _I, _VP, _SZ, _VP, _SZ, _SZ, _VP, _SZ, _SZ, C.c_uint)
    d("dedisp_execute", _I, _VP, _SZ, _VP, _SZ, _VP, _SZ, C.c_uint)
    d("dedisp_execute_guru", _I, _VP, _SZ, _VP, _SZ, _SZ, _VP, _SZ, _SZ, _SZ, _SZ, C.c_uint)
    d("dedisp_execute_adv_noalloc", _I, _VP, _SZ, _VP, _SZ, _SZ, _VP, _SZ, _SZ, C.c_uint)
    d("dedisp_set_device", _I, _I)
    d("dedisp_sync", _I)
    d("dedisp_get_error_string", C.c_char_p, _I)
    d("dedisp_set_stream", _I, _VP, _VP)


class DedispError(RuntimeError):
    pass


def _chk(code):
    if code != DEDISP_NO_ERROR:
        msg = lib().dedisp_get_error_string(code).decode()
        raise DedispError(f"dedisp error {code}: {msg}")


class DedispPlan:
    """Python view of a dedisp_plan (create / DM list / killmask / execute)."""

    def __init__(self, nchans, dt, f0, df):
        self._p = _VP()
        _chk(lib().dedisp_create_plan(C.byref(self._p), nchans, dt, f0, df))
        self.nchans = nchans

    def close(self):
        if self._p:
            lib().dedisp_destroy_plan(self._p)
            self._p = _VP()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def generate_dm_list(self, dm_start, dm_end, ti_us, tol):
        _chk(lib().dedisp_generate_dm_list(self._p, dm_start, dm_end, ti_us, tol))

    def set_dm_list(self, dms):
        a = np.ascontiguousarray(dms, dtype=np.float32)
        _chk(lib().dedisp_set_dm_list(self._p, a.ctypes.data_as(C.POINTER(_F)), a.size))

    def set_killmask(self, mask):
        if mask is None:
            _chk(lib().dedisp_set_killmask(self._p, None))
            return
        a = np.ascontiguousarray(mask, dtype=np.int32)
        _chk(lib().dedisp_set_killmask(self._p, a.ctypes.data_as(C.POINTER(_I))))

    @property
    def dm_count(self):
        return int(lib().dedisp_get_dm_count(self._p))

    @property
    def max_delay(self):
        return int(lib().dedisp_get_max_delay(self._p))

    def dm_list(self):
        n = self.dm_count
        ptr = lib().dedisp_get_dm_list(self._p)
        return np.ctypeslib.as_array(ptr, shape=(n,)).copy() if n else np.zeros(0, np.float32)

    def execute_adv(self, data_tf, nsamps=None, in_stride=None, out=None, flags=0):
        """Host-pointer execute: data_tf is u8 [nsamps][in_stride]."""
        data_tf = np.ascontiguousarray(data_tf, dtype=np.uint8)
        if nsamps is None:
            nsamps = data_tf.shape[0]
        if in_stride is None:
            in_stride = data_tf.shape[1]
        nout = nsamps - self.max_delay
        if out is None:
            out = np.zeros((self.dm_count, nout), dtype=np.float32)
        _chk(lib().dedisp_execute_adv(self._p, nsamps, data_tf.ctypes.data, 8, in_stride,
                                      out.ctypes.data, 32, out.strides[0], flags))
        return out

    def execute_device(self, d_in, nsamps, in_stride, d_out, out_stride, noalloc=False,
                       flags=DEDISP_DEVICE_POINTERS):
        """Device-pointer execute (the call hella makes): d_in and d_out are
        device addresses (int); noalloc selects dedisp_execute_adv_noalloc.
        Returns the dedisp_error code."""
        f = lib().dedisp_execute_adv_noalloc if noalloc else lib().dedisp_execute_adv
        return int(f(self._p, nsamps, d_in, 8, in_stride, d_out, 32, out_stride, flags))

    def execute_guru(self, data_tf, first_dm, dm_count, flags=0):
        data_tf = np.ascontiguousarray(data_tf, dtype=np.uint8)
        nsamps, in_stride = data_tf.shape
        nout = nsamps - self.max_delay
        out = np.zeros((dm_count, nout), dtype=np.float32)
        _chk(lib().dedisp_execute_guru(self._p, nsamps, data_tf.ctypes.data, 8, in_stride,
                                       out.ctypes.data, 32, out.strides[0], first_dm, dm_count,
                                       flags))
        return out


# ---------------------------------------------------------------- hella pipeline
class HellaError(RuntimeError):
    pass


class _Info(C.Structure):
    _fields_ = [("nchan", _I), ("ndm", _I), ("max_delay", C.c_longlong), ("ntime", C.c_longlong),
                ("ntime_dd", _I), ("ntime_out", _I), ("nwidth", _I), ("gulp", C.c_longlong),
                ("min_gulp", _I), ("gulp_index", C.c_longlong), ("rows", _I), ("nlive", _I)]


class _Cand(C.Structure):
    _fields_ = [("snr", _F), ("beam", _I), ("samp", _I), ("width", _I), ("dm_idx", _I),
                ("dm", _F), ("abs_samp", C.c_longlong)]


class _Cluster(C.Structure):
    _fields_ = [("rep", _I), ("count", _I), ("samp_min", _I), ("samp_max", _I),
                ("dm_min", _I), ("dm_max", _I), ("beam", _I)]


class SynthParams(C.Structure):
    _fields_ = [("seed", C.c_ulonglong), ("nchan", _I), ("nsamp", C.c_longlong),
                ("t_offset", C.c_longlong), ("fch1", C.c_double), ("foff", C.c_double),
                ("tsamp", C.c_double), ("dtype", _I), ("mean", _F), ("sigma", _F),
                ("n_rfi_chan", _I), ("rfi_chan", C.POINTER(_I)), ("rfi_sigma_mult", _F),
                ("n_pulse", _I), ("pulse_t", C.POINTER(C.c_double)), ("pulse_dm", C.POINTER(_F)),
                ("pulse_fwhm_ms", C.POINTER(_F)), ("pulse_snr", C.POINTER(_F)),
                ("burst_period", C.c_longlong), ("burst_phase", C.c_longlong), ("burst_len", _I),
                ("burst_amp", _F), ("bp_slope", _F)]


class _RingBlock(C.Structure):
    _fields_ = [("index", _I), ("beam", _I), ("seq", C.c_longlong), ("flags", _I),
                ("bytes", C.c_size_t), ("t_arrival_ns", C.c_longlong), ("data", _VP)]


class _ReplayEvent(C.Structure):
    _fields_ = [("t", C.c_double), ("beam", _I), ("pool", _I), ("seq", C.c_longlong),
                ("flags", _I)]


class _ServerOpts(C.Structure):
    _fields_ = [("hbm_budget", C.c_ulonglong), ("max_batch", _I), ("lanes", _I), ("plain", _I)]


class _ChunkResult(C.Structure):
    _fields_ = [("beam", _I), ("seq", C.c_longlong), ("searched", _I), ("ncand", _I),
                ("crossings", C.c_longlong), ("nclusters", _I), ("t_arrival_ns", C.c_longlong),
                ("t_done_ns", C.c_longlong), ("batch", _I)]


class _ServerStats(C.Structure):
    _fields_ = [("chunks", C.c_longlong), ("batches", C.c_longlong), ("beams_admitted", _I),
                ("beams_waited", _I), ("max_resident", _I), ("work_bytes", C.c_ulonglong),
                ("peak_resident_bytes", C.c_ulonglong), ("busy_s", C.c_double),
                ("wall_s", C.c_double)]


RING_LAST, RING_END = 1, 2

class FilHeader(C.Structure):
    _fields_ = ([(n, _I) for n in ("machine_id", "telescope_id", "data_type", "nchans", "nbits",
                                   "nifs", "barycentric", "pulsarcentric", "ibeam", "nbeams")] +
                [(n, C.c_double) for n in ("tstart", "tsamp", "fch1", "foff", "refdm", "az_start",
                                           "za_start", "src_raj", "src_dej", "period")] +
                [("isign", C.c_char), ("npuls", C.c_longlong), ("source_name", C.c_char * 256),
                 ("raw_data_file", C.c_char * 256), ("nfreq", _I),
                 ("frequency_table", C.c_double * 4096), ("header_bytes", C.c_longlong),
                 ("nsamples", C.c_longlong)])


CAND_DTYPE = np.dtype([("snr", "f4"), ("beam", "i4"), ("samp", "i4"), ("width", "i4"),
                       ("dm_idx", "i4"), ("dm", "f4"), ("abs_samp", "i8")])
CLUSTER_DTYPE = np.dtype([("rep", "i4"), ("count", "i4"), ("samp_min", "i4"), ("samp_max", "i4"),
                          ("dm_min", "i4"), ("dm_max", "i4"), ("beam", "i4")])


def _declare_hella(L):
    def d(name, res, *args):
        f = getattr(L, name)
        f.restype = res
        f.argtypes = list(args)

    P = _VP
    d("hella_create", _I, C.c_char_p, C.POINTER(P))
    d("hella_destroy", None, P)
    d("hella_last_error", C.c_char_p)
    d("hella_set_beam", _I, P, _I, C.c_char_p)
    d("hella_set_killmask", _I, P, _I, C.POINTER(_I))
    d("hella_get_info", _I, P, _I, C.POINTER(_Info))
    d("hella_get_dm_list", _I, P, _I, C.POINTER(_F), _I, C.POINTER(_I))
    d("hella_get_stream", P, P)
    d("hella_fast_flagger", _I, P, P, _I)
    d("hella_dedisperse", _I, P, _I)
    d("hella_smooth", _I, P, _I)
    d("hella_find_peaks", _I, P, _I)
    d("hella_clear_peaks", _I, P)
    d("hella_end_gulp", _I, P)
    d("hella_process_gulp", _I, P, P, _I, C.POINTER(_I))
    d("hella_process_gulp_batched", _I, P, P, _I, C.POINTER(_I))
    d("hella_get_peaks", _I, P, C.POINTER(_Cand), _I)
    d("hella_get_clusters", _I, P, C.POINTER(_Cluster), _I)
    d("hella_last_crossings", C.c_longlong, P, _I)
    d("hella_copy_window", _I, P, _I, P, C.c_size_t)
    d("hella_copy_dedisp", _I, P, _I, P, C.c_size_t)
    d("hella_set_dedisp", _I, P, _I, P, C.c_size_t)
    d("hella_get_search_norm", _I, P, _I, C.POINTER(C.c_double), C.POINTER(C.c_double),
      C.POINTER(_F))
    d("hella_get_search_stats", _I, P, _I, C.POINTER(_F), C.POINTER(_F), C.POINTER(_F),
      C.POINTER(_I))
    d("hella_stage_times", _I, P, C.POINTER(_F))
    d("hella_set_timing", _I, P, _I)
    d("hella_autotune", _I, P, C.c_char_p)
    d("hella_bench_stage", _I, P, _I, _I, _I, C.POINTER(_F))
    d("hella_set_tile", _I, P, _I, _I, _I)
    d("hella_device_count", _I)
    d("hella_dev_alloc", P, C.c_size_t)
    d("hella_dev_free", None, P)
    d("hella_memcpy", _I, P, P, C.c_size_t, _I)
    d("hella_device_sync", _I)
    d("hella_synth_filterbank", _I, C.POINTER(SynthParams), P)
    d("hella_ring_create", _I, _I, C.c_size_t, C.c_char_p, C.POINTER(P))
    d("hella_ring_destroy", None, P)
    d("hella_ring_header", C.c_char_p, P)
    d("hella_ring_open_write", _I, P, _I, C.POINTER(P))
    d("hella_ring_close_write", _I, P, _I, C.c_longlong, _I, C.c_size_t)
    d("hella_ring_end", _I, P)
    d("hella_ring_abort", _I, P)
    d("hella_ring_open_read", _I, P, _I, C.POINTER(_RingBlock))
    d("hella_ring_release", _I, P, _I)
    d("hella_ring_replay", _I, P, C.POINTER(_ReplayEvent), _I, C.POINTER(P), _I, C.c_size_t,
      C.c_double, C.POINTER(C.c_longlong))
    d("hella_server_create", _I, C.c_char_p, C.POINTER(_ServerOpts), C.POINTER(P))
    d("hella_server_destroy", None, P)
    d("hella_server_run", _I, P, P)
    d("hella_server_result_count", _I, P)
    d("hella_server_get_result", _I, P, _I, C.POINTER(_ChunkResult))
    d("hella_server_get_candidates", _I, P, _I, C.POINTER(_Cand), _I)
    d("hella_server_get_clusters", _I, P, _I, C.POINTER(_Cluster), _I)
    d("hella_server_get_stats", _I, P, C.POINTER(_ServerStats))
    d("hella_server_footprint", _I, C.c_char_p, C.POINTER(_ServerOpts),
      C.POINTER(C.c_ulonglong), C.POINTER(C.c_ulonglong))
    d("hella_get_flag_counts", _I, P, _I, C.POINTER(C.c_uint), _I)
    d("hella_write_candidates", _I, C.c_char_p, C.POINTER(_Cand), _I, _I)
    d("hella_format_coincidencer", _I, C.c_longlong, C.c_longlong, C.c_char_p, C.POINTER(_Cand),
      _I, _I, C.c_char_p, _I)
    d("hella_write_specflags", _I, C.c_char_p, C.POINTER(_I), _I)
    d("hella_get_flag_spectrum", _I, P, _I, C.POINTER(_F), _I)
    d("hella_accumulate_specflags", _I, P, C.POINTER(_I), _I)
    d("hella_dada_header_value", _I, C.c_char_p, C.c_size_t, C.c_char_p, C.c_char_p, _I)
    d("hella_dada_coincidencer_fields", _I, C.c_char_p, C.c_size_t, C.c_char_p, _I)
    d("hella_sigproc_read_header", _I, C.c_char_p, C.POINTER(FilHeader))
    d("hella_sigproc_read", C.c_longlong, C.c_char_p, C.POINTER(FilHeader), C.c_longlong,
      C.c_longlong, P)
    d("hella_sigproc_config", _I, C.POINTER(FilHeader), C.c_char_p, _I)


_orig_declare = _declare


def _declare(L):  # noqa: F811 -- extend the dedisp declarations
    _orig_declare(L)
    _declare_hella(L)


def _hchk(code):
    if code != 0:
        raise HellaError(f"hella error {code}: {lib().hella_last_error().decode()}")


def config_text(cfg):
    """dict -> KEY VALUE text; SCRUNCH takes a list of (ts, fs, thresh)."""
    if isinstance(cfg, str):
        return cfg
    lines = []
    for k, v in cfg.items():
        if k == "SCRUNCH":
            lines.append(f"SCRUNCH {len(v)}")
            lines += [f"{ts} {fs} {th} 1" for ts, fs, th in v]
        else:
            lines.append(f"{k} {v}")
    return "\n".join(lines) + "\n"


class DeviceBuffer:
    """Device memory from the library's allocator (hipMalloc)."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        self.ptr = lib().hella_dev_alloc(self.nbytes)
        if not self.ptr:
            raise HellaError("device allocation failed")

    def upload(self, arr):
        a = np.ascontiguousarray(arr)
        assert a.nbytes <= self.nbytes
        _hchk(lib().hella_memcpy(self.ptr, a.ctypes.data, a.nbytes, 1))

    def download(self, arr):
        _hchk(lib().hella_memcpy(arr.ctypes.data, self.ptr, arr.nbytes, 2))
        return arr

    def free(self):
        if self.ptr:
            lib().hella_dev_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Hella:
    """One pipeline context (hella_create ... hella_destroy)."""

    def __init__(self, cfg):
        self._c = _VP()
        _hchk(lib().hella_create(config_text(cfg).encode(), C.byref(self._c)))

    def close(self):
        if self._c:
            lib().hella_destroy(self._c)
            self._c = _VP()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_beam(self, beam, **kw):
        _hchk(lib().hella_set_beam(self._c, beam, config_text(kw).encode()))

    def set_killmask(self, beam, mask):
        a = np.ascontiguousarray(mask, np.int32)
        _hchk(lib().hella_set_killmask(self._c, beam, a.ctypes.data_as(C.POINTER(_I))))

    def set_tile(self, lanes, chunk, order=0):
        _hchk(lib().hella_set_tile(self._c, lanes, chunk, order))

    def use_tuning(self, path):
        """Apply an autotuner record (the JSON that autotune() writes) before
        the first gulp: its best dedispersion tile is forced for every beam.
        Returns the record's "best" entry."""
        with open(path) as f:
            best = json.load(f)["best"]
        self.set_tile(*tile_args(best))
        return best

    def set_timing(self, on=True):
        _hchk(lib().hella_set_timing(self._c, 1 if on else 0))

    def info(self, beam=0):
        i = _Info()
        _hchk(lib().hella_get_info(self._c, beam, C.byref(i)))
        return {f: getattr(i, f) for f, _ in _Info._fields_}

    def dm_list(self, beam=0):
        n = _I()
        _hchk(lib().hella_get_dm_list(self._c, beam, None, 0, C.byref(n)))
        out = np.zeros(n.value, np.float32)
        _hchk(lib().hella_get_dm_list(self._c, beam, out.ctypes.data_as(C.POINTER(_F)), n.value,
                                      C.byref(n)))
        return out

    @staticmethod
    def _ptr(gulp):
        """(pointer, host flag, keep-alive) of a gulp: a DeviceBuffer, or a host
        array in the configured INPUT format -- uint8 [gulp][nchan] (FILTERBANK,
        FILE), float32 [gulp][nchan] (INPUT_DTYPE F32) or float16 [nchan][gulp]
        (DADA: the fp16 channel-major blocks of the PSRDADA ring), beams
        concatenated"""
        if isinstance(gulp, DeviceBuffer):
            return gulp.ptr, 0
        a = np.ascontiguousarray(gulp)
        return a.ctypes.data, 1, a

    def process_gulp(self, gulp, batched=False):
        keep = self._ptr(gulp)
        n = _I()
        f = lib().hella_process_gulp_batched if batched else lib().hella_process_gulp
        _hchk(f(self._c, keep[0], keep[1], C.byref(n)))
        return n.value

    def fast_flagger(self, gulp):
        keep = self._ptr(gulp)
        _hchk(lib().hella_fast_flagger(self._c, keep[0], keep[1]))

    def dedisperse(self, beam):
        _hchk(lib().hella_dedisperse(self._c, beam))

    def smooth(self, scale=1):
        _hchk(lib().hella_smooth(self._c, scale))

    def find_peaks(self, beam):
        _hchk(lib().hella_find_peaks(self._c, beam))

    def clear_peaks(self):
        _hchk(lib().hella_clear_peaks(self._c))

    def end_gulp(self):
        _hchk(lib().hella_end_gulp(self._c))

    def peaks(self):
        n = lib().hella_get_peaks(self._c, None, 0)
        out = np.zeros(n, CAND_DTYPE)
        if n:
            lib().hella_get_peaks(self._c, out.ctypes.data_as(C.POINTER(_Cand)), n)
        return out

    def clusters(self):
        n = lib().hella_get_clusters(self._c, None, 0)
        out = np.zeros(n, CLUSTER_DTYPE)
        if n:
            lib().hella_get_clusters(self._c, out.ctypes.data_as(C.POINTER(_Cluster)), n)
        return out

    def flag_counts(self, beam=0):
        """Per-channel flagged-sample counts of the beam (accumulated)."""
        n = lib().hella_get_flag_counts(self._c, beam, None, 0)
        if n < 0:
            _hchk(1)
        out = np.zeros(n, np.uint32)
        lib().hella_get_flag_counts(self._c, beam, out.ctypes.data_as(C.POINTER(C.c_uint)), n)
        return out

    def flag_spectrum(self, beam=0):
        """The beam's flag spectrum (the reference's flagSpec): per channel the
        accumulated mean of every scrunch's grown mask."""
        n = lib().hella_get_flag_spectrum(self._c, beam, None, 0)
        if n < 0:
            _hchk(1)
        out = np.zeros(n, np.float32)
        lib().hella_get_flag_spectrum(self._c, beam, out.ctypes.data_as(C.POINTER(_F)), n)
        return out

    def accumulate_specflags(self, specflags):
        """specflags[i] += (int)(NTIME * flagSpec[i]) over the beams with
        len(specflags) channels (casm:705-716); specflags: int32 array, updated"""
        assert specflags.dtype == np.int32 and specflags.flags.c_contiguous
        _hchk(lib().hella_accumulate_specflags(self._c, specflags.ctypes.data_as(C.POINTER(_I)),
                                               int(specflags.size)))
        return specflags

    def crossings(self, beam):
        return int(lib().hella_last_crossings(self._c, beam))

    def window(self, beam=0):
        i = self.info(beam)
        out = np.zeros((i["nchan"], i["ntime"]), np.uint8)
        _hchk(lib().hella_copy_window(self._c, beam, out.ctypes.data, out.nbytes))
        return out

    def dedisp_plane(self, beam=0):
        i = self.info(beam)
        out = np.zeros((i["ndm"], i["ntime_dd"]), np.uint32)
        _hchk(lib().hella_copy_dedisp(self._c, beam, out.ctypes.data, out.nbytes))
        return out

    def set_dedisp(self, beam, D):
        """Load a dedispersed plane (u32 [ndm][ntime_dd]) as the beam's current
        plane; smooth() and find_peaks() then search it."""
        a = np.ascontiguousarray(D, np.uint32)
        _hchk(lib().hella_set_dedisp(self._c, beam, a.ctypes.data, a.nbytes))

    def search_stats(self, beam=0):
        nw = self.info(beam)["nwidth"]
        mu = np.zeros(nw, np.float32)
        sg = np.zeros(nw, np.float32)
        s0 = _F()
        gated = _I()
        _hchk(lib().hella_get_search_stats(self._c, beam, mu.ctypes.data_as(C.POINTER(_F)),
                                           sg.ctypes.data_as(C.POINTER(_F)), C.byref(s0),
                                           C.byref(gated)))
        return dict(mu=mu, sigma=sg, sigma0=s0.value, gated=bool(gated.value))

    def search_norm(self, beam=0):
        """per-width mean and 1/sigma (double) and the threshold of the last search"""
        nw = self.info(beam)["nwidth"]
        mu = np.zeros(nw, np.float64)
        inv = np.zeros(nw, np.float64)
        thr = _F()
        _hchk(lib().hella_get_search_norm(self._c, beam, mu.ctypes.data_as(C.POINTER(C.c_double)),
                                          inv.ctypes.data_as(C.POINTER(C.c_double)),
                                          C.byref(thr)))
        return mu, inv, thr.value

    def stage_times(self):
        ms = (_F * 4)()
        _hchk(lib().hella_stage_times(self._c, ms))
        return list(ms)

    def bench_stage(self, beam, which, reps=5):
        """mean device ms of one stage re-run (0 flag, 1 dedisperse, 2 search)"""
        ms = _F()
        _hchk(lib().hella_bench_stage(self._c, beam, which, reps, C.byref(ms)))
        return ms.value

    def autotune(self, path=None):
        _hchk(lib().hella_autotune(self._c, path.encode() if path else None))

    @property
    def stream(self):
        return lib().hella_get_stream(self._c)


def tile_args(best):
    """(lanes_per_dm, chunk, order) for Hella.set_tile from the "best" entry of
    an autotuner record, whose "dma" field is the loader number the tuner
    reports (0 registers, 1 LDS DMA, 2 register prefetch, 3 one trial per wave,
    4 as 3 with every LDS read of a chunk issued first, 5 as 3 with the
    windows copied by LDS DMA two chunks ahead)."""
    loader = {0: 1 << 1, 1: 0, 2: 2 << 1, 3: 3 << 1, 4: (3 << 1) | 8, 5: (3 << 1) | 16}
    dma = int(best["dma"])
    if dma not in loader:
        raise ValueError(f"autotuner record: unknown loader {dma}")
    return int(best["lanes_per_dm"]), int(best["chunk"]), (int(best["order"]) & 1) | loader[dma]


def device_count():
    return int(lib().hella_device_count())


def device_sync():
    _hchk(lib().hella_device_sync())


def synth(nchan, nsamp, fch1, foff, tsamp, seed=1, t_offset=0, dtype="u8", mean=128.0,
          sigma=16.0, rfi_chan=(), rfi_mult=1.0, pulses=(), burst_period=0, burst_phase=0,
          burst_len=0, burst_amp=0.0, bp_slope=0.0):
    """Seeded synthetic filterbank [nsamp][nchan] (u8 or f32).
    pulses: iterable of (t_seconds, dm, fwhm_ms, snr)."""
    rc = np.ascontiguousarray(rfi_chan, np.int32)
    pulses = list(pulses)
    pt = np.array([p[0] for p in pulses], np.float64)
    pdm = np.array([p[1] for p in pulses], np.float32)
    pw = np.array([p[2] for p in pulses], np.float32)
    ps = np.array([p[3] for p in pulses], np.float32)
    prm = SynthParams(seed, nchan, nsamp, t_offset, fch1, foff, tsamp, 0 if dtype == "u8" else 1,
                      mean, sigma, rc.size, rc.ctypes.data_as(C.POINTER(_I)), rfi_mult,
                      len(pulses), pt.ctypes.data_as(C.POINTER(C.c_double)),
                      pdm.ctypes.data_as(C.POINTER(_F)), pw.ctypes.data_as(C.POINTER(_F)),
                      ps.ctypes.data_as(C.POINTER(_F)), burst_period, burst_phase, burst_len,
                      burst_amp, bp_slope)
    out = np.zeros((nsamp, nchan), np.uint8 if dtype == "u8" else np.float32)
    _hchk(lib().hella_synth_filterbank(C.byref(prm), out.ctypes.data))
    return out


# ---------------------------------------------------------------- streaming
class Ring:
    """PSRDADA-style block ring in pinned host memory (hella_ring_*)."""

    def __init__(self, nblocks, block_bytes, header=""):
        self._r = _VP()
        self.block_bytes = int(block_bytes)
        _hchk(lib().hella_ring_create(nblocks, self.block_bytes, header.encode(), C.byref(self._r)))

    @property
    def handle(self):
        return self._r

    def header(self):
        return lib().hella_ring_header(self._r).decode()

    def write(self, beam, seq, data, flags=0, timeout_ms=-1):
        a = np.ascontiguousarray(data)
        if a.nbytes > self.block_bytes:
            raise ValueError("chunk larger than a ring block")
        blk = _VP()
        _hchk(lib().hella_ring_open_write(self._r, timeout_ms, C.byref(blk)))
        C.memmove(blk.value, a.ctypes.data, a.nbytes)
        _hchk(lib().hella_ring_close_write(self._r, beam, seq, flags, a.nbytes))

    def end(self):
        _hchk(lib().hella_ring_end(self._r))

    def read(self, timeout_ms=0):
        """next block as a dict (data copied), or None; {'end': True} at the end"""
        b = _RingBlock()
        _hchk(lib().hella_ring_open_read(self._r, timeout_ms, C.byref(b)))
        if b.index < 0:
            return {"end": True} if b.flags & RING_END else None
        data = np.ctypeslib.as_array((C.c_uint8 * b.bytes).from_address(b.data)).copy()
        return {"index": b.index, "beam": b.beam, "seq": b.seq, "flags": b.flags,
                "t_arrival_ns": b.t_arrival_ns, "data": data}

    def release(self, index):
        _hchk(lib().hella_ring_release(self._r, index))

    def replay(self, events, pool, time_scale=1.0):
        """events: iterable of (t, beam, pool_index, seq, flags) sorted by t; pool:
        list of arrays of one chunk each.  Blocks until the stream has been written
        and returns the largest delay behind schedule in seconds."""
        ev = list(events)
        arr = (_ReplayEvent * max(1, len(ev)))()
        for i, (t, beam, pi, seq, flags) in enumerate(ev):
            arr[i] = _ReplayEvent(float(t), int(beam), int(pi), int(seq), int(flags))
        keep = [np.ascontiguousarray(p) for p in pool]
        ptrs = (_VP * len(keep))(*[k.ctypes.data for k in keep])
        late = C.c_longlong()
        _hchk(lib().hella_ring_replay(self._r, arr, len(ev), ptrs, len(keep), keep[0].nbytes,
                                      time_scale, C.byref(late)))
        return late.value * 1e-9

    def close(self):
        if self._r:
            lib().hella_ring_destroy(self._r)
            self._r = _VP()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Streaming server with continuous batching (hella_server_*)."""

    def __init__(self, cfg, hbm_budget=0, max_batch=16, lanes=2, plain=False):
        self._s = _VP()
        self.opts = _ServerOpts(int(hbm_budget), int(max_batch), int(lanes), 1 if plain else 0)
        _hchk(lib().hella_server_create(config_text(cfg).encode(), C.byref(self.opts),
                                        C.byref(self._s)))

    @staticmethod
    def footprint(cfg, max_batch=16, lanes=2, plain=False):
        """(bytes per resident beam, bytes of the work sets)"""
        o = _ServerOpts(0, int(max_batch), int(lanes), 1 if plain else 0)
        bb, wb = C.c_ulonglong(), C.c_ulonglong()
        _hchk(lib().hella_server_footprint(config_text(cfg).encode(), C.byref(o), C.byref(bb),
                                           C.byref(wb)))
        return bb.value, wb.value

    def run(self, ring):
        _hchk(lib().hella_server_run(self._s, ring.handle))

    def results(self, with_candidates=True):
        out = []
        r = _ChunkResult()
        for i in range(lib().hella_server_result_count(self._s)):
            _hchk(lib().hella_server_get_result(self._s, i, C.byref(r)))
            d = {f: getattr(r, f) for f, _ in _ChunkResult._fields_}
            if with_candidates:
                n = lib().hella_server_get_candidates(self._s, i, None, 0)
                c = np.zeros(n, CAND_DTYPE)
                if n:
                    lib().hella_server_get_candidates(self._s, i, c.ctypes.data_as(C.POINTER(_Cand)), n)
                m = lib().hella_server_get_clusters(self._s, i, None, 0)
                k = np.zeros(m, CLUSTER_DTYPE)
                if m:
                    lib().hella_server_get_clusters(self._s, i, k.ctypes.data_as(C.POINTER(_Cluster)), m)
                d["cands"], d["clusters"] = c, k
            out.append(d)
        return out

    def stats(self):
        st = _ServerStats()
        _hchk(lib().hella_server_get_stats(self._s, C.byref(st)))
        return {f: getattr(st, f) for f, _ in _ServerStats._fields_}

    def close(self):
        if self._s:
            lib().hella_server_destroy(self._s)
            self._s = _VP()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def serve(server, ring, events, pool, time_scale=1.0):
    """Replay events into the ring from a writer thread while the server runs
    in this thread; returns the writer's largest delay behind schedule (s)."""
    import threading
    res = {}

    def writer():
        try:
            res["late"] = ring.replay(events, pool, time_scale)
        except Exception as e:  # the server aborts the ring on error
            res["error"] = e

    th = threading.Thread(target=writer, daemon=True)
    th.start()
    try:
        server.run(ring)
    finally:
        th.join()
    if "error" in res:
        raise res["error"]
    return res["late"]


# ---------------------------------------------------------------- candidate emission
def _cands_arg(cands):
    c = np.ascontiguousarray(cands, CAND_DTYPE)
    return c, (c.ctypes.data_as(C.POINTER(_Cand)) if c.size else None)


def write_candidates(path, cands, beam0=0):
    """Append candidates to a text file in the reference's format (header
    line when the file is empty, then "%g %d %g %d %d %g %d" lines)."""
    c, ptr = _cands_arg(cands)
    _hchk(lib().hella_write_candidates(str(path).encode(), ptr, int(c.size), beam0))


def format_coincidencer(samp0, gulp, cands, beam0=0, dada_fields=None):
    """The coincidencer message of one gulp as text."""
    c, ptr = _cands_arg(cands)
    df = dada_fields.encode() if dada_fields else None
    n = lib().hella_format_coincidencer(samp0, gulp, df, ptr, int(c.size), beam0, None, 0)
    if n < 0:
        raise HellaError("hella_format_coincidencer: bad argument")
    buf = C.create_string_buffer(n + 1)
    lib().hella_format_coincidencer(samp0, gulp, df, ptr, int(c.size), beam0, buf, n + 1)
    return buf.value.decode()


def _header_bytes(header):
    return header if isinstance(header, bytes) else header.encode()


def dada_header_value(header, key):
    """The value of key in a PSRDADA header text (None if absent)."""
    h = _header_bytes(header)
    n = lib().hella_dada_header_value(h, len(h), key.encode(), None, 0)
    if n < 0:
        return None
    buf = C.create_string_buffer(n + 1)
    lib().hella_dada_header_value(h, len(h), key.encode(), buf, n + 1)
    return buf.value.decode()


def dada_coincidencer_fields(header):
    """"UTC_START PICOSECONDS TSAMP" of a PSRDADA header (the coincidencer's
    first-line fields); raises if one is missing."""
    h = _header_bytes(header)
    n = lib().hella_dada_coincidencer_fields(h, len(h), None, 0)
    if n < 0:
        raise HellaError(lib().hella_last_error().decode())
    buf = C.create_string_buffer(n + 1)
    lib().hella_dada_coincidencer_fields(h, len(h), buf, n + 1)
    return buf.value.decode()


def write_specflags(path, specflags):
    """Append one "%d" line per channel (the spectral-flag file of casm:756-761)."""
    a = np.ascontiguousarray(specflags, np.int32)
    _hchk(lib().hella_write_specflags(str(path).encode(), a.ctypes.data_as(C.POINTER(_I)),
                                      int(a.size)))


# ---------------------------------------------------------------- SIGPROC input
class FilterbankFile:
    """A SIGPROC filterbank file: its header and time-major uint8 samples."""

    def __init__(self, path):
        self.path = str(path)
        self.hdr = FilHeader()
        _hchk(lib().hella_sigproc_read_header(self.path.encode(), C.byref(self.hdr)))

    def __getattr__(self, name):
        if name in dict(FilHeader._fields_):
            v = getattr(self.hdr, name)
            return v.decode() if isinstance(v, bytes) else v
        raise AttributeError(name)

    def read(self, first, nsamp):
        out = np.zeros((nsamp, self.hdr.nchans), np.uint8)
        n = lib().hella_sigproc_read(self.path.encode(), C.byref(self.hdr), first, nsamp,
                                     out.ctypes.data)
        if n < 0:
            raise HellaError(lib().hella_last_error().decode())
        return out[:n]

    def config(self):
        n = lib().hella_sigproc_config(C.byref(self.hdr), None, 0)
        buf = C.create_string_buffer(n + 1)
        lib().hella_sigproc_config(C.byref(self.hdr), buf, n + 1)
        return buf.value.decode()
