/*
 * hella.h -- C ABI of the MI355X FRB-search pipeline (libhella_amd.so).
 *
 * Mirrors casm-hella's pipeline-stage interface (namespace hella, all taking
 * the pipeline state hella::pinfo_t, src/hella/definitions.h:84-173):
 *   hella_fast_flagger  <- hella::fast_flagger(p)      flagger.h:75, flagger.cu:345-411
 *   hella_dedisperse    <- hella::dedisperse(p, beam)  dedisperse.h:17, dedisperse.cpp:17-48
 *   hella_smooth        <- hella::smooth(p, scale)     smooth.h:16, smooth.cpp:21-127
 *   hella_find_peaks    <- hella::find_peaks(p, bm)    peaks.h:17, peaks.cu:31-117
 *   hella_clear_peaks   <- hella::clear_peaks(p)       peaks.h:20, peaks.cu:119-127
 *   hella_process_gulp  <- the per-gulp loop of src/apps/casm_hella_refactored.cu:659-835
 *                          (flag all beams, then per beam dedisperse -> smooth -> find_peaks)
 *   hella_create        <- initialize() (casm:54-309): the same KEY VALUE configuration
 *                          vocabulary (casm:72-169), plus the geometry keys that are
 *                          compile-time constants in the reference (definitions.h:43-56).
 * Opaque context instead of pinfo_t; status codes instead of exit(); all device
 * work is stream-ordered on the context's stream.
 */
#ifndef HELLA_AMD_HELLA_H
#define HELLA_AMD_HELLA_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hella_ctx_s *hella_ctx;

typedef enum {
  HELLA_OK = 0,
  HELLA_ERR_INVALID = 1,   /* bad argument */
  HELLA_ERR_CONFIG = 2,    /* bad or inconsistent configuration */
  HELLA_ERR_ALLOC = 3,     /* device or host allocation failed */
  HELLA_ERR_GPU = 4,       /* HIP runtime error */
  HELLA_ERR_STATE = 5,     /* call out of order (e.g. search before the ring is full) */
  HELLA_ERR_CAPACITY = 6   /* a fixed-capacity buffer overflowed (data still valid up to it) */
} hella_status;

/* one threshold crossing (peaks.cu:101-116 fields plus absolute sample) */
typedef struct {
  float snr;
  int beam;
  int samp;        /* sample within the gulp's output window */
  int width;       /* width trial index */
  int dm_idx;      /* DM trial index into the beam's DM list */
  float dm;        /* DM value of that trial */
  long long abs_samp; /* absolute sample index (stream time) */
} hella_candidate;

typedef struct {
  int rep;         /* index of the representative (max S/N) candidate in the gulp list */
  int count;       /* crossings in the cluster */
  int samp_min, samp_max, dm_min, dm_max;
  int beam;
} hella_cluster;

typedef struct {
  int nchan;
  int ndm;
  long long max_delay;
  long long ntime;     /* ring length NTIME (casm:215-224) */
  int ntime_dd;        /* dedispersed samples per gulp (gulp + maxW) */
  int ntime_out;       /* searched samples per gulp */
  int nwidth;
  long long gulp;
  int min_gulp;        /* first gulp index with a full ring (casm:632) */
  long long gulp_index;/* gulps ingested so far */
  int rows;            /* searched DM rows */
  int nlive;           /* live channels */
} hella_info;

/* ---- lifetime ------------------------------------------------------------ */
/* config: newline-separated "KEY VALUE" lines.  Keys (defaults in brackets):
 *   reference keys:  DM_MIN [0] DM_MAX [1000] DM_TOL [1.25] WIDTH_MIN [1]
 *     WIDTH_MAX [32] SNR [7] GULP [32768] FLAG1 [-1] FLAG2 [-1]
 *     SPEC_MIN [-0.05] SPEC_MAX [0.15] NBEAM [1] GPU [0]
 *     INPUT FILTERBANK|DADA|FILE [FILTERBANK]  (FILE = one-shot window sizing)
 *     SCRUNCH n, followed by n lines "ts fs thresh nits"
 *   geometry:  NCHAN [256] FCH1 [500] (MHz) FOFF [-0.390625] (MHz) TSAMP [256e-6] (s)
 *     PULSE_WIDTH [40] (us, dedisp ti)
 *   search:    SEARCH BOXCAR|HELLA [BOXCAR]  MAX_GIANTS [10000]  CAND_CAP [1048576]
 *     CLUSTER_DT [64] CLUSTER_DDM [16]  TIME_LIMIT [100.1] (s, casm:723)
 *   input:     INPUT_DTYPE U8|F32|F16 (default from INPUT)
 *   flagger:   FLAG_STATS MEAN|ROBUST [MEAN]  (ROBUST: every normalisation uses
 *     each channel's median and 1.4826 x MAD over the statistics columns
 *     instead of the plane's mean and sigma; no reference counterpart)
 * Unknown keys are ignored, as in the reference. */
hella_status hella_create(const char *config, hella_ctx *out);
void hella_destroy(hella_ctx ctx);
/* message of the last failing call on this thread */
const char *hella_last_error(void);

/* Per-beam geometry/DM overrides ("KEY VALUE" lines: NCHAN FCH1 FOFF TSAMP DM_MIN
 * DM_MAX DM_TOL) for beam < NBEAM; call before the first gulp. */
hella_status hella_set_beam(hella_ctx ctx, int beam, const char *overrides);
/* killmask[c] != 0 keeps channel c of the beam (dedisp_set_killmask semantics) */
hella_status hella_set_killmask(hella_ctx ctx, int beam, const int *killmask);

hella_status hella_get_info(hella_ctx ctx, int beam, hella_info *out);
/* copies up to cap DM values; returns the DM count in *n */
hella_status hella_get_dm_list(hella_ctx ctx, int beam, float *out, int cap, int *n);
/* the hipStream_t (as void*) all work of the context is ordered on */
void *hella_get_stream(hella_ctx ctx);

/* ---- stages (plain per-beam path) --------------------------------------- */
/* gulp: all beams' samples, beam-major ([nbeam][gulp][nchan] for U8/F32 time-major,
 * [nbeam][nchan][gulp] for F16 channel-major).  host_ptr != 0: host memory.
 * Flags every beam into its ring (casm:701). */
hella_status hella_fast_flagger(hella_ctx ctx, const void *gulp, int host_ptr);
/* dedisperse the current window of a beam into the context's plane (casm:726) */
hella_status hella_dedisperse(hella_ctx ctx, int beam);
/* width trials + per-width statistics of the last dedispersed beam (casm:731);
 * scale != 1 keeps the reference's unnormalised mode unsupported -> HELLA_ERR_INVALID */
hella_status hella_smooth(hella_ctx ctx, int scale);
/* sigma gate, threshold, ordered compaction; appends to the gulp's list (casm:736) */
hella_status hella_find_peaks(hella_ctx ctx, int beam);
hella_status hella_clear_peaks(hella_ctx ctx);
/* advance every beam's ring by one gulp (the rewind of casm:771-797) */
hella_status hella_end_gulp(hella_ctx ctx);

/* Whole gulp on the plain path: flag all beams, then per beam
 * dedisperse -> smooth -> find_peaks while the candidate cap and the time limit
 * allow, then advance.  *ncand receives the gulp's candidate count. */
hella_status hella_process_gulp(hella_ctx ctx, const void *gulp, int host_ptr, int *ncand);
/* Same result through the batched path: every beam of the gulp in one launch
 * chain (flag all, dedisperse all, search all), captured once as a hipGraph
 * and replayed. */
hella_status hella_process_gulp_batched(hella_ctx ctx, const void *gulp, int host_ptr,
                                        int *ncand);

/* the gulp's candidates (reference order: beam, width, DM row, sample) */
int hella_get_peaks(hella_ctx ctx, hella_candidate *out, int cap);
/* clusters of the gulp's candidates (8-connected cells of CLUSTER_DT samples x
 * CLUSTER_DDM trials, per beam); returns the count, fills up to cap */
int hella_get_clusters(hella_ctx ctx, hella_cluster *out, int cap);
/* total crossings of the last searched beam before any cap */
long long hella_last_crossings(hella_ctx ctx, int beam);

/* ---- inspection ---------------------------------------------------------- */
/* copy the beam's ring (u8, [nchan][ntime], oldest sample first) */
hella_status hella_copy_window(hella_ctx ctx, int beam, void *host, size_t bytes);
/* copy the last dedispersed plane of a beam ([ndm][ntime_dd] u32) */
hella_status hella_copy_dedisp(hella_ctx ctx, int beam, void *host, size_t bytes);
/* load a dedispersed plane ([ndm][ntime_dd] u32, host memory) as the beam's
 * current plane, in place of hella_dedisperse: hella_smooth and
 * hella_find_peaks then search it (the reference's smooth/find_peaks read
 * whatever p->d_dedisp holds, smooth.cpp:21-127, peaks.cu:31-117) */
hella_status hella_set_dedisp(hella_ctx ctx, int beam, const void *host, size_t bytes);
/* per-width mean/sigma (nwidth each), sigma0, gate flag of the beam's last search */
hella_status hella_get_search_stats(hella_ctx ctx, int beam, float *mu, float *sigma,
                                    float *sigma0, int *gated);
/* the exact normalisation of the beam's last search: per width the mean and
 * 1/sigma in double (nwidth each; S/N = (float)((raw - mu) * inv)), and the
 * threshold snr * sigma0 in float */
hella_status hella_get_search_norm(hella_ctx ctx, int beam, double *mu, double *inv, float *thr);
/* per-stage device time of the last gulp in ms: flag, dedisperse, search, peaks (4) */
hella_status hella_stage_times(hella_ctx ctx, float *ms);
/* enable per-stage event timing (adds syncs between stages) */
hella_status hella_set_timing(hella_ctx ctx, int on);

/* Re-run one stage of the plain path for a beam on the data of its last run
 * (same descriptors, so the call is idempotent) reps times, timed with HIP
 * events on the context stream.  which: 0 flagger, 1 dedispersion kernel,
 * 2 search (width trials, statistics, threshold, compaction).  *ms receives
 * the mean duration of one repetition. */
hella_status hella_bench_stage(hella_ctx ctx, int beam, int which, int reps, float *ms);

/* ---- tuning -------------------------------------------------------------- */
/* time the dedispersion tile shapes on the current rings and keep the fastest;
 * the choice and its timings are written as JSON to path (may be NULL) */
hella_status hella_autotune(hella_ctx ctx, const char *path);
/* force a tile shape: lanes per DM trial (8/16/32/64), channels per LDS chunk
 * (8/16/32); order bit 0: tile order (0 DM-block-major, 1 time-major); order
 * bits 1-2: window loader (0 LDS DMA, 1 through registers, 2 next chunk
 * prefetched in registers, chunk 8 or 16; 3 one DM trial per wave: lanes 64,
 * chunk 4 or 8 = DM trials per workgroup); order bit 3 (loader 3 only): issue
 * every LDS read of a chunk before its sums; order bit 4 (loader 3, chunk 4):
 * copy the windows by LDS DMA two chunks ahead */
hella_status hella_set_tile(hella_ctx ctx, int lanes_per_dm, int chunk, int order);

/* ---- SIGPROC filterbank input (src/hella/src/sigproc.cpp:26-191) ------------ */
#define HELLA_FIL_MAXFREQ 4096
/* the readable header parameters (sigproc.h:17-49), plus where the data start
 * and how many whole samples the file holds */
typedef struct {
  int machine_id, telescope_id, data_type, nchans, nbits, nifs, barycentric, pulsarcentric;
  int ibeam, nbeams;
  double tstart, tsamp, fch1, foff, refdm, az_start, za_start, src_raj, src_dej, period;
  char isign;
  long long npuls;
  char source_name[256];
  char raw_data_file[256];
  int nfreq;                  /* fchannel entries read (fch1 = foff = 0 then) */
  double frequency_table[HELLA_FIL_MAXFREQ];
  long long header_bytes;     /* 0: the file has no SIGPROC header */
  long long nsamples;         /* whole samples after the header */
} hella_fil_header;
/* Parse the keyword header of a SIGPROC file (read_header, sigproc.cpp:58-191):
 * HELLA_ERR_CONFIG for an unknown keyword or a truncated header, as the
 * reference throws. */
hella_status hella_sigproc_read_header(const char *path, hella_fil_header *hdr);
/* Read nsamp samples from sample `first` on into out ([nsamp][nchans] uint8,
 * the time-major layout the pipeline takes as FILTERBANK/FILE input).  Only
 * 8-bit, single-IF files (sigproc.cpp:184-188).  Returns the samples read
 * (fewer at the end of the file), or -1. */
long long hella_sigproc_read(const char *path, const hella_fil_header *hdr, long long first,
                             long long nsamp, void *out);
/* The geometry keys of the configuration ("NCHAN .. FCH1 .. FOFF .. TSAMP ..")
 * from a header; returns the length (as snprintf), -1 on a bad argument. */
int hella_sigproc_config(const hella_fil_header *hdr, char *buf, int cap);

/* ---- candidate emission (src/hella/src/peaks.cu:129-259) ----------------- */
/* Per-channel count of flagged samples of the beam (pixels of every 1 x 1 /
 * box scrunch's grown mask, accumulated since the context was created).
 * Returns nchan (copies min(cap, nchan)), or -1 on error. */
int hella_get_flag_counts(hella_ctx ctx, int beam, unsigned *out, int cap);
/* The beam's flag spectrum, the reference's flagSpec (flagger.cu:116-124,
 * 333-334): per channel, each scrunch adds the mean over the statistics
 * columns (256 * floor(gulp / 256)) of its grown mask G, G being the fraction
 * of masked pixels in each ts x fs box (1 x 1: the mask itself), as
 * fspec += (float)(sum / (ts * fs) / columns), accumulated since the context
 * was created.  Returns nchan (copies min(cap, nchan)), or -1 on error. */
int hella_get_flag_spectrum(hella_ctx ctx, int beam, float *out, int cap);
/* casm:705-716 after a gulp whose windows are full: for every channel i <
 * nchan, specflags[i] += (int)(NTIME * flagSpec[i]), flagSpec being the sum
 * (in beam order, double) of the flag spectra of the beams with nchan
 * channels (the reference's beams share one flagSpec array) and NTIME the
 * window length of the first of them. */
hella_status hella_accumulate_specflags(hella_ctx ctx, int *specflags, int nchan);
/* Append candidates to the text file path as output_peaks does
 * (peaks.cu:129-182): the header "SNR SAMP_START TIME_START WIDTH DM_IDX DM
 * BEAM_IDX" if the file is empty, then one "%g %d %g %d %d %g %d" line per
 * candidate: S/N, abs_samp, abs_samp * 1e-3 / 86400 (TIME_RESOLUTION and
 * SECONDS_PER_DAY, definitions.h:56,59), width index, DM index, DM,
 * beam + beam0.  The samples print in full (%lld) beyond 2^31. */
hella_status hella_write_candidates(const char *path, const hella_candidate *cands, int n,
                                    int beam0);
/* The coincidencer message of one gulp (peaks.cu:212-231): the line
 * "(samp0 / gulp) + 1" (samp0: the gulp's first output sample, C division),
 * followed by " dada_fields" (the DADA header's "UTC_START PICOSECONDS TSAMP";
 * NULL or "" for other inputs), then the candidate lines of
 * hella_write_candidates.  Writes at most cap - 1 bytes and a NUL; returns
 * the full length (as snprintf), or -1 on a bad argument.  Sending it (a TCP
 * connection per gulp in the reference) is left to the caller. */
int hella_format_coincidencer(long long samp0, long long gulp, const char *dada_fields,
                              const hella_candidate *cands, int n, int beam0, char *buf, int cap);
/* ---- PSRDADA header (src/hella/src/dada.cpp:50-79) ----------------------
 * The header text as key/value pairs, parsed as parse_dada_header does: the
 * text ends at the first NUL (or after len bytes), one "KEY VALUE" per line,
 * the key up to the first space, the value from the first non-space after it
 * to the end of the line, empty lines skipped, a line without a space ends
 * the parse, a repeated key keeps its last value.
 * hella_dada_header_value: the value of key (writes at most cap - 1 bytes and
 * a NUL; returns its full length as snprintf does), -1 if the key is absent. */
int hella_dada_header_value(const char *header, size_t len, const char *key, char *out, int cap);
/* "UTC_START PICOSECONDS TSAMP" (the three values joined by spaces), the
 * dada_fields argument of hella_format_coincidencer for DADA input
 * (peaks.cu:215-220); -1 when one of them is missing (the reference's
 * map::at throws). */
int hella_dada_coincidencer_fields(const char *header, size_t len, char *out, int cap);

/* Append one "%d" line per channel: the spectral-flag file of casm:756-761
 * (the specflags array hella_accumulate_specflags keeps). */
hella_status hella_write_specflags(const char *path, const int *specflags, int nchan);

/* ---- PSRDADA-style block ring (streaming ingest) -------------------------
 * Replaces the PSRDADA data block casm-hella reads from (src/hella/src/dada.cpp:16-88;
 * casm:659-670,812-815): a text header and fixed-size data blocks in pinned host
 * memory, one writer and one reader, blocks read in order.  Each block carries
 * the beam it belongs to, the beam's chunk number, flags and its arrival time
 * (steady clock, ns, stamped when the writer closes it).  The reader may release
 * blocks in any order; the writer reuses a block once every older one is released. */
typedef struct hella_ring_s *hella_ring;

enum {
  HELLA_RING_LAST = 1,   /* block flag: the beam's last chunk (the beam leaves after it) */
  HELLA_RING_END = 2     /* read result: the writer ended the stream and every block was read */
};

typedef struct {
  int index;             /* block index for hella_ring_release (-1: nothing read) */
  int beam;
  long long seq;
  int flags;
  size_t bytes;
  long long t_arrival_ns;
  const void *data;
} hella_ring_block;

hella_status hella_ring_create(int nblocks, size_t block_bytes, const char *header, hella_ring *out);
void hella_ring_destroy(hella_ring ring);
const char *hella_ring_header(hella_ring ring);
/* writer: wait (timeout_ms < 0: forever) for a free block, fill it, close it */
hella_status hella_ring_open_write(hella_ring ring, int timeout_ms, void **block);
hella_status hella_ring_close_write(hella_ring ring, int beam, long long seq, int flags, size_t bytes);
/* writer: no more blocks */
hella_status hella_ring_end(hella_ring ring);
/* reader: give up; waiting writers return HELLA_ERR_STATE */
hella_status hella_ring_abort(hella_ring ring);
/* reader: next block in order; info->index = -1 on timeout, and flags = HELLA_RING_END
 * once the stream has ended and been read completely */
hella_status hella_ring_open_read(hella_ring ring, int timeout_ms, hella_ring_block *info);
hella_status hella_ring_release(hella_ring ring, int index);

/* one scheduled chunk for hella_ring_replay */
typedef struct {
  double t;              /* arrival time (s) from the start of the replay */
  int beam;
  int pool;              /* index of the chunk's data in the pool */
  long long seq;
  int flags;             /* HELLA_RING_LAST on a beam's final chunk */
} hella_replay_event;

/* Writer loop for tests and benchmarks: writes the events (sorted by t) at
 * t * time_scale seconds after the call, copying pool[pool] into each block,
 * then ends the stream.  *max_late_ns: the largest delay behind schedule (a full
 * ring holds the writer back). */
hella_status hella_ring_replay(hella_ring ring, const hella_replay_event *events, int n,
                               const void *const *pool, int npool, size_t chunk_bytes,
                               double time_scale, long long *max_late_ns);

/* ---- streaming server (continuous batching) -------------------------------
 * Serves beams that join and leave a running stream: each beam keeps its
 * max-DM window and flagger state resident on the device; beams are admitted
 * while the HBM budget allows (resident windows + the fixed work sets);
 * chunks that arrive while other batches are on the GPU are launched as soon
 * as a launch chain is free, each launch chain carrying up to max_batch beams'
 * chunks (the batched path, captured as hipGraphs).  plain = 1 runs the plain
 * per-beam sequence instead, one chunk at a time.  No reference counterpart:
 * the reference processes a fixed beam set gulp by gulp (casm:659-835). */
typedef struct hella_server_s *hella_server;

typedef struct {
  unsigned long long hbm_budget;  /* bytes for resident beams + work sets (0: no limit) */
  int max_batch;                  /* beams per launch chain (>= 1) */
  int lanes;                      /* launch chains in flight, one stream each (>= 1) */
  int plain;                      /* 1: plain per-beam path (max_batch 1, one lane) */
} hella_server_opts;

typedef struct {
  int beam;               /* beam id from the ring */
  long long seq;          /* the beam's chunk number */
  int searched;           /* 0 while the beam's window was still filling */
  int ncand;              /* candidates kept (MAX_GIANTS cap per chunk) */
  long long crossings;    /* threshold crossings before caps */
  int nclusters;
  long long t_arrival_ns; /* block closed by the writer */
  long long t_done_ns;    /* candidates and clusters on the host */
  int batch;              /* chunks in the launch chain that carried it */
} hella_chunk_result;

typedef struct {
  long long chunks;        /* chunks processed */
  long long batches;       /* launch chains */
  int beams_admitted;
  int beams_waited;        /* beams that waited for HBM before admission */
  int max_resident;        /* most beams resident at once */
  unsigned long long work_bytes;      /* work sets (fixed) */
  unsigned long long peak_resident_bytes;
  double busy_s;           /* wall time with at least one launch chain in flight */
  double wall_s;           /* duration of hella_server_run */
} hella_server_stats;

/* config: the hella_create KEY VALUE text (geometry, DM range, widths, GULP =
 * chunk length, SNR, SCRUNCH, ...).  Every beam uses it. */
hella_status hella_server_create(const char *config, const hella_server_opts *opts,
                                 hella_server *out);
void hella_server_destroy(hella_server srv);
/* Serve the ring until its writer ends the stream and every chunk is done.
 * Runs in the calling thread; the writer runs in another. */
hella_status hella_server_run(hella_server srv, hella_ring ring);
/* results in completion order */
int hella_server_result_count(hella_server srv);
hella_status hella_server_get_result(hella_server srv, int i, hella_chunk_result *out);
/* candidates / clusters of result i (reference order within the chunk) */
int hella_server_get_candidates(hella_server srv, int i, hella_candidate *out, int cap);
int hella_server_get_clusters(hella_server srv, int i, hella_cluster *out, int cap);
hella_status hella_server_get_stats(hella_server srv, hella_server_stats *out);
/* bytes one beam holds while resident, and the work-set bytes of the options */
hella_status hella_server_footprint(const char *config, const hella_server_opts *opts,
                                    unsigned long long *beam_bytes,
                                    unsigned long long *work_bytes);

/* ---- device helpers (test / bench plumbing) ------------------------------ */
int hella_device_count(void);
void *hella_dev_alloc(size_t bytes);
void hella_dev_free(void *p);
/* kind: 1 host->device, 2 device->host, 3 device->device */
hella_status hella_memcpy(void *dst, const void *src, size_t bytes, int kind);
hella_status hella_device_sync(void);

/* ---- synthetic filterbank generator (seeded, host) ---------------------- */
typedef struct {
  unsigned long long seed;
  int nchan;
  long long nsamp;        /* rows to generate */
  long long t_offset;     /* absolute index of the first row (streams stay continuous) */
  double fch1, foff, tsamp;
  int dtype;              /* 0: u8 (rounded, clipped to [0,255]); 1: f32 */
  float mean, sigma;
  int n_rfi_chan;         /* channels with noise sigma * rfi_sigma_mult */
  const int *rfi_chan;
  float rfi_sigma_mult;
  int n_pulse;            /* dispersed Gaussian pulses */
  const double *pulse_t;  /* arrival time (s) at fch1 */
  const float *pulse_dm;
  const float *pulse_fwhm_ms;
  const float *pulse_snr; /* matched-filter S/N of the band-summed pulse */
  long long burst_period; /* broadband bursts: every period samples (0: none) */
  long long burst_phase;
  int burst_len;
  float burst_amp;        /* in units of sigma */
  float bp_slope;         /* gain = 1 + bp_slope * (c/(nchan-1) - 0.5) */
} hella_synth_params;

/* out: [nsamp][nchan] of dtype.  Pulses use 4.15e3 s MHz^2 and the float delay
 * of the dedisp plan, so injection and search agree. */
hella_status hella_synth_filterbank(const hella_synth_params *p, void *out);

#ifdef __cplusplus
}
#endif
#endif
