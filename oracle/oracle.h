/*
 * oracle.h -- CPU restatement of casm-hella's FRB-search hot path
 * (flag -> dedisperse -> boxcar/smooth -> peaks -> clusters).
 *
 * THIS IS TEST INFRASTRUCTURE. Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it, and only as the checker or the
 * timed CPU baseline. The product library never links or calls it.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the casm-hella repository root).  Where the reference leaves a detail open
 * (closed NPP/dedisp code, FP16 arithmetic, reduction order) this file fixes
 * it; the fixed choices are listed in the header comment of oracle.c.
 *
 * Pinning status of each function is stated beside it:
 *   pinned   -- checked by tests against closed forms / brute force / invariants
 *   unpinned -- no reference vector or closed form fixes it (see oracle.c)
 */
#ifndef HELLA_ORACLE_H
#define HELLA_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Canonical blocked reduction: blocks of ORC_CSUM_BLOCK elements, 256 lane
 * partials per block (lane l sums elements l, l+256, ... of the block in
 * order), a pairwise tree over the 256 lanes, then the block sums added in
 * ascending block order.  All in double. */
#define ORC_CSUM_BLOCK 4096
#define ORC_LANES 256
double orc_csum(const double *x, size_t n);

/* ---- A1 plan (casm:203-224; dedisp plan API) ---------------------------- */
/* delay[c] = 4.15e3/dt * (1/(f0+c*df)^2 - 1/f0^2), float, per channel. */
void orc_delay_table(size_t nch, float dt, float f0, float df, float *delay);
/* DM list by the smearing-tolerance recurrence, FP64, stored as float each
 * step.  Returns the count (also when cap is exceeded; then only cap stored). */
size_t orc_dm_list(float dm_start, float dm_end, double dt, double ti_us,
                   double f0, double df, size_t nch, double tol,
                   float *out, size_t cap);
/* max_delay = (size_t)((double)(dm_last * delay[nch-1]) + 0.5) */
size_t orc_max_delay(const float *dms, size_t ndm, const float *delay, size_t nch);
/* delta[d*nch + c] = (int)((double)(dms[d]*delay[c]) + 0.5) */
void orc_delays(const float *dms, size_t ndm, const float *delay, size_t nch,
                int32_t *delta);

/* ---- A11 dedispersion (src/dedisperse.cpp:17-48; external dedisp) ------- */
/* out[d*out_stride + t] = sum_{c live} in[(t+delta[d][c])*in_tstride + c*in_cstride]
 * for t < nout.  killmask may be NULL (all channels live). */
void orc_dedisperse(const uint8_t *in, size_t in_tstride, size_t in_cstride,
                    size_t nch, const int32_t *delta, size_t ndm,
                    const uint8_t *killmask, size_t nout,
                    uint32_t *out, size_t out_stride);

/* ---- A4 statistics (src/normalization.cu:125-162) ----------------------- */
/* combine per-row sums S1 (of x) and S2 (of x^2) over n columns into the
 * global mean and the mean of row variances about it */
void orc_stats_rows(const double *S1, const double *S2, size_t h, double n,
                    double *mu, double *var);
/* mean of row means / mean of row variances over the first 256*floor(w/256)
 * columns (all w columns when w < 256). */
void orc_stats(const float *x, size_t h, size_t w, size_t stride,
               double *mu, double *var);
/* as orc_stats, float results: mu_f = (float)mu, sig_f = (float)sqrt(var) */
void orc_stats_f(const float *x, size_t h, size_t w, size_t stride,
                 float *mu_f, float *sig_f);

/* ---- A2..A9 flagger (src/flagger.cu:198-411, transpose.cu, timeseries.cu) */
typedef struct {
  int tscrunch, fscrunch;
  float thresh;
} orc_scrunch;

typedef struct {
  int nch;          /* channels (rows) */
  int W;            /* gulp width in samples (columns) */
  float spec_min, spec_max;
  int flag1, flag2; /* channel range [flag1, flag2) zeroed by each scrunch */
  int nscrunch;
  const orc_scrunch *scrunch;
  int robust;       /* 1: every normalisation is orc_normalize_robust */
  int u8;           /* 1: the plane holds 8-bit input bytes (0..255): with only
                       1 x 1 scrunches and mean statistics the stages up to
                       ts_correct take the byte-histogram order (orc_flag) */
} orc_flag_cfg;

/* stage functions on a float FT plane x[nch][W] (row stride W) */
void orc_bandpass(const float *x, int nch, int W, float *bp);          /* calc_bandpass_new */
void orc_divide_bp(float *x, int nch, int W, const float *bp);         /* divide_by_bp */
void orc_normalize(float *x, int nch, int W);                          /* normalize_data */
/* Robust per-channel normalisation (no reference counterpart; the opt-in
 * FLAG_STATS ROBUST mode): over the statistics columns t < wp (see
 * orc_stats) of channel c, m = the element of rank (wp-1)/2 in ascending
 * order, d = the same of |x - m| (float subtraction), s = (float)(1.4826 * d);
 * every sample becomes (float)(((double)x - m) * (1.0 / (double)s)), or 0
 * when s is 0.  Ascending order is that of the floats with -0 before +0 and
 * NaNs past the infinities.  m and s per channel go to ms (may be NULL). */
void orc_normalize_robust(float *x, int nch, int W, float *ms);
void orc_spec_flag(float *x, int nch, int W, float smin, float smax);  /* calc_bandpass_new + replace_data_and_add_bandpass */
void orc_box_mean(const float *x, int nch, int W, int ts, int fs, float *out); /* npp_convolve_handler */
/* apply_scrunch flag branch; flagspec (may be NULL) accumulates the mean of
 * the grown mask G over the statistics columns (flagger.cu:333-334): G is the
 * fraction of masked pixels in each box, so per channel
 * flagspec[c] += (float)((double)n / (ts*fs) / wp), n the masked pixels of
 * the boxes of columns t < wp = 256*floor(W/256) (W when W < 256) */
void orc_scrunch_flag(float *x, int nch, int W, int ts, int fs, float thresh,
                      int flag1, int flag2, uint32_t *flagcount, float *flagspec);
/* channel sum of sample t: ORC_TS_PARTS partitions of ceil(nch/16)
 * consecutive channels, each summed in channel order, the partials combined
 * pairwise (p[q] += p[q+h], h = 8, 4, 2, 1); double */
#define ORC_TS_PARTS 16
double orc_chan_sum(const float *x, int nch, int W, int t);
void orc_ts_correct(float *x, int nch, int W, float *ts);              /* ts_correct */
void orc_quantize(const float *x, int nch, int W, uint8_t *out, size_t out_stride); /* transpose_output */
/* input conversion (transpose_input): u8 TF [W][nch] -> float FT */
void orc_ingest_u8_tf(const uint8_t *in, int nch, int W, float *x);
void orc_ingest_f32_tf(const float *in, int nch, int W, float *x);
/* whole fast_flagger for one beam; x is the FT plane (modified in place).
 * With cfg->u8, 1 x 1 scrunches and mean statistics, every channel sum
 * before ts_correct is a sum over the channel's byte histogram (count x
 * stage value of the byte, the 256 products in the canonical tree of
 * orc_csum); otherwise sums run over the plane in the canonical blocked
 * order.  out receives u8 FT rows with stride out_stride.  flagcount (may be NULL)
 * accumulates the grown-mask count per channel, flagspec (may be NULL) the
 * flag spectrum of every scrunch (orc_scrunch_flag). */
void orc_flag(const orc_flag_cfg *cfg, float *x, uint8_t *out, size_t out_stride,
              uint32_t *flagcount, float *flagspec);

/* ---- A12/A13 search: boxcar or hella smoothing, normalise, threshold ---- */
enum { ORC_MODE_BOXCAR = 0, ORC_MODE_HELLA = 1 };

typedef struct {
  int ndm;        /* rows of D */
  int ntime_dd;   /* columns of D available (gulp + maxW) */
  int nout;       /* searched samples per row (gulp) */
  int maxW;
  int minW;
  int mode;
  float snr;
  int max_giants; /* cap on returned candidates */
} orc_search_cfg;

typedef struct {
  float snr;
  int samp;
  int width;  /* width index */
  int dm_idx; /* index into the plan DM list */
} orc_cand;

/* number of width trials for the mode */
int orc_nwidth(const orc_search_cfg *cfg);
/* width (boxcar) or smoothing scale sm (hella) of trial k */
int orc_width_of(const orc_search_cfg *cfg, int k);
/* hella smoothing kernel for scale sm: 3 x (2sm+1) floats (src/smooth.cpp:45-58) */
void orc_hella_kernel(int sm, float *k);

/* Full search of one beam's dedispersed plane D[ndm][ntime_dd] (row stride
 * d_stride, exact integer sums).  Per width, the statistics follow
 * orc_stats_rows; in BOXCAR mode the row sums S1 = sum box and S2 = sum box^2
 * are exact integers rounded to double once (S2 as hi * 2^64 + lo), in HELLA
 * mode they are canonical double sums of the float plane.  Writes per-width mu/sigma (nwidth each),
 * sigma0, returns the number of threshold crossings (all, uncapped); stores
 * the first min(count, cfg->max_giants) into cands (in width, row, sample
 * order).  plane_out (may be NULL) receives the normalised planes
 * [nwidth][rows][nout] (rows = ndm, or ndm-2 in hella mode). */
long orc_search(const orc_search_cfg *cfg, const uint32_t *D, size_t d_stride,
                float *mu_w, float *sig_w, float *sigma0, int *gated,
                orc_cand *cands, float *plane_out);

/* ---- A14 clustering (no reference counterpart) -------------------------- */
typedef struct {
  int rep;        /* index of representative candidate (max S/N, lowest index on ties) */
  int count;
  int samp_min, samp_max, dm_min, dm_max;
} orc_cluster_t;
/* cells of cell_t samples x cell_dm DM trials; 8-connected occupied cells form
 * a cluster.  Returns the cluster count; clusters ordered by rep index. */
int orc_cluster(const orc_cand *c, int n, int cell_t, int cell_dm,
                orc_cluster_t *out);

#ifdef __cplusplus
}
#endif
#endif
