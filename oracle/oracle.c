/*
 * oracle.c -- CPU restatement of casm-hella's FRB-search hot path.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Compiled with -ffp-contract=off so
 * every float/double operation is one IEEE-rounded operation in the order
 * written here.
 *
 * Choices this file fixes where the reference leaves them open:
 *  - dedisp (external, un-vendored; called at casm:205-210 and
 *    src/dedisperse.cpp:32) is restated from the published dedisp algorithm:
 *    float delay table, FP64 DM-list recurrence stored as float each step,
 *    max_delay = (size_t)(dm_last*delay_last + 0.5).  The per-(DM,channel)
 *    delay uses the same round-half-up of the float product.  Output for
 *    out_nbits=32 is float(exact integer sum).
 *  - FP16 arithmetic of the flagger (src/flagger.cu) is restated in FP32;
 *    every reduction is a double-precision canonical blocked sum (orc_csum);
 *    the channel sum of each sample (timeseries.cu:19-36) is taken over 16
 *    partitions of consecutive channels (each summed in channel order) whose
 *    partials are combined pairwise, in double (orc_chan_sum).
 *    mu_f = (float)mu, sigma_f = (float)sqrt(var).
 *  - Normalisation (flagger.cu:198-214, smooth.cpp:95-124) is
 *    y = (float)(((double)x - mu) * (1.0 / (double)sigma_f)).
 *  - NPP filters (convolution.cu:40, smooth.cpp:61) use the documented anchor
 *    convention: output (x,y) covers source columns x-ax .. x-ax+kw-1 and rows
 *    y-ay .. y-ay+kh-1, border replicate; taps accumulate row-major (kernel
 *    row outer, column inner) into a float starting at 0.
 *  - Quantise (transpose.cu:65-75) maps NaN to 0.
 *  - Statistics use 256*floor(w/256) columns (normalization.cu:129); when
 *    w < 256 (where the reference divides by zero) all w columns are used.
 *  - BOXCAR width statistics take the row sums of the box values and of
 *    their squares exactly (integers), rounded to double once; every other
 *    statistic uses the canonical blocked double sum.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
double orc_csum(const double *x, size_t n)
{
  double total = 0.0;
  for (size_t b0 = 0; b0 < n; b0 += ORC_CSUM_BLOCK) {
    size_t bn = n - b0 < ORC_CSUM_BLOCK ? n - b0 : ORC_CSUM_BLOCK;
    double p[ORC_LANES];
    for (size_t l = 0; l < ORC_LANES; ++l) {
      double s = 0.0;
      for (size_t k = l; k < bn; k += ORC_LANES) s += x[b0 + k];
      p[l] = s;
    }
    for (size_t s = ORC_LANES / 2; s >= 1; s >>= 1)
      for (size_t l = 0; l < s; ++l) p[l] += p[l + s];
    total += p[0];
  }
  return total;
}

/* ---- plan: dedisp generate_delay_table / generate_dm_list ------------------
 * Called by the reference at casm:205-210. */
void orc_delay_table(size_t nch, float dt, float f0, float df, float *delay)
{
  for (size_t c = 0; c < nch; ++c) {
    float a = 1.f / (f0 + (float)c * df);
    float b = 1.f / f0;
    float ab = a * a - b * b;
    delay[c] = (float)(4.15e3 / (double)dt * (double)ab);
  }
}

size_t orc_dm_list(float dm_start, float dm_end, double dt, double ti_us,
                   double f0, double df, size_t nch, double tol,
                   float *out, size_t cap)
{
  double dtu = dt * 1e6;
  double f = (f0 + ((double)(nch / 2) - 0.5) * df) * 1e-3;
  double tol2 = tol * tol;
  double a = 8.3 * df / (f * f * f);
  double a2 = a * a;
  double b2 = a2 * ((double)(nch * nch) / 16.0);
  double c = (dtu * dtu + ti_us * ti_us) * (tol2 - 1.0);
  size_t n = 0;
  float last = dm_start;
  if (n < cap) out[n] = last;
  ++n;
  while (last < dm_end) {
    double prev = (double)last;
    double prev2 = prev * prev;
    double k = c + tol2 * a2 * prev2;
    double dm = (b2 * prev + sqrt(-a2 * b2 * prev2 + (a2 + b2) * k)) / (a2 + b2);
    last = (float)dm;
    if (n < cap) out[n] = last;
    ++n;
  }
  return n;
}

size_t orc_max_delay(const float *dms, size_t ndm, const float *delay, size_t nch)
{
  float p = dms[ndm - 1] * delay[nch - 1];
  return (size_t)((double)p + 0.5);
}

void orc_delays(const float *dms, size_t ndm, const float *delay, size_t nch,
                int32_t *delta)
{
  for (size_t d = 0; d < ndm; ++d)
    for (size_t c = 0; c < nch; ++c) {
      float p = dms[d] * delay[c];
      delta[d * nch + c] = (int32_t)((double)p + 0.5);
    }
}

/* ---- dedispersion (dedisp_execute_adv semantics, src/dedisperse.cpp:24-38) */
void orc_dedisperse(const uint8_t *in, size_t in_tstride, size_t in_cstride,
                    size_t nch, const int32_t *delta, size_t ndm,
                    const uint8_t *killmask, size_t nout,
                    uint32_t *out, size_t out_stride)
{
  /* DM rows are independent: threads only change which core computes a row */
#pragma omp parallel for schedule(dynamic, 1)
  for (size_t d = 0; d < ndm; ++d) {
    uint32_t *o = out + d * out_stride;
    for (size_t t = 0; t < nout; ++t) o[t] = 0;
    for (size_t c = 0; c < nch; ++c) {
      if (killmask && !killmask[c]) continue;
      const uint8_t *row = in + c * in_cstride + (size_t)delta[d * nch + c] * in_tstride;
      for (size_t t = 0; t < nout; ++t) o[t] += row[t * in_tstride];
    }
  }
}

/* ---- statistics (calculate_stddev, src/normalization.cu:125-162) ---------- */
static size_t stats_width(size_t w)
{
  size_t wp = (w / 256) * 256;
  return wp ? wp : w;
}

/* Per row r: S1_r = csum(x), S2_r = csum(x*x) over the statistics columns;
 * mu = csum_r(S1_r / n) / h;  v_r = ((S2_r - (2 mu) S1_r) + (n mu) mu) / n;
 * var = csum_r(v_r) / h.  This is the reference's mean of row means and mean
 * of row variances about the global mean (normalization.cu:135-151) with the
 * row sums of squared deviations expanded, so one pass over the data gives
 * both moments. */
void orc_stats_rows(const double *S1, const double *S2, size_t h, double n,
                    double *mu, double *var)
{
  double *rr = (double *)calloc(h ? h : 1, sizeof(double));
  for (size_t r = 0; r < h; ++r) rr[r] = S1[r] / n;
  double m = orc_csum(rr, h) / (double)h;
  for (size_t r = 0; r < h; ++r) rr[r] = ((S2[r] - (2.0 * m) * S1[r]) + (n * m) * m) / n;
  double v = orc_csum(rr, h) / (double)h;
  free(rr);
  *mu = m;
  *var = v;
}

void orc_stats(const float *x, size_t h, size_t w, size_t stride,
               double *mu, double *var)
{
  size_t wp = stats_width(w);
  double *S1 = (double *)calloc(h ? h : 1, sizeof(double));
  double *S2 = (double *)calloc(h ? h : 1, sizeof(double));
  /* rows are independent (each row's sums keep their canonical order);
   * threads only change which core computes a row */
#pragma omp parallel
  {
    double *t1 = (double *)malloc((wp ? wp : 1) * sizeof(double));
    double *t2 = (double *)malloc((wp ? wp : 1) * sizeof(double));
#pragma omp for schedule(static)
    for (size_t r = 0; r < h; ++r) {
      for (size_t j = 0; j < wp; ++j) {
        double v = (double)x[r * stride + j];
        t1[j] = v;
        t2[j] = v * v;
      }
      S1[r] = orc_csum(t1, wp);
      S2[r] = orc_csum(t2, wp);
    }
    free(t1);
    free(t2);
  }
  orc_stats_rows(S1, S2, h, (double)wp, mu, var);
  free(S1); free(S2);
}

void orc_stats_f(const float *x, size_t h, size_t w, size_t stride,
                 float *mu_f, float *sig_f)
{
  double m, v;
  orc_stats(x, h, w, stride, &m, &v);
  *mu_f = (float)m;
  *sig_f = (float)sqrt(v);
}

/* ---- flagger stages ------------------------------------------------------ */

/* calc_bandpass_new (src/flagger.cu:86-112): sum of the first 32*floor(W/32)
 * samples of each channel, divided by W. */
void orc_bandpass(const float *x, int nch, int W, float *bp)
{
  size_t n = (size_t)(W / 32) * 32;
#pragma omp parallel
  {
    double *tmp = (double *)malloc((n ? n : 1) * sizeof(double));
#pragma omp for schedule(static)
    for (int c = 0; c < nch; ++c) {
      for (size_t j = 0; j < n; ++j) tmp[j] = (double)x[(size_t)c * W + j];
      bp[c] = (float)(orc_csum(tmp, n) / (double)W);
    }
    free(tmp);
  }
}

/* divide_by_bp (src/flagger.cu:180-194) */
void orc_divide_bp(float *x, int nch, int W, const float *bp)
{
#pragma omp parallel for schedule(static)
  for (int c = 0; c < nch; ++c) {
    float b = bp[c];
    float *row = x + (size_t)c * W;
    for (int t = 0; t < W; ++t) row[t] = b > 0.f ? row[t] / b : 0.f;
  }
}

/* normalize_data (src/flagger.cu:198-214) */
void orc_normalize(float *x, int nch, int W)
{
  double m, v;
  orc_stats(x, (size_t)nch, (size_t)W, (size_t)W, &m, &v);
  float sf = (float)sqrt(v);
  double inv = 1.0 / (double)sf;
  size_t n = (size_t)nch * W;
#pragma omp parallel for schedule(static)
  for (size_t i = 0; i < n; ++i) x[i] = (float)(((double)x[i] - m) * inv);
}

static inline uint32_t fkey(float f)
{
  uint32_t b;
  memcpy(&b, &f, 4);
  return b ^ ((b >> 31) ? 0xffffffffu : 0x80000000u);
}

static inline float kval(uint32_t k)
{
  uint32_t b = (k >> 31) ? (k ^ 0x80000000u) : ~k;
  float f;
  memcpy(&f, &b, 4);
  return f;
}

static int cmp_u32(const void *a, const void *b)
{
  uint32_t x = *(const uint32_t *)a, y = *(const uint32_t *)b;
  return x < y ? -1 : (x > y);
}

void orc_normalize_robust(float *x, int nch, int W, float *ms)
{
  const size_t wp = stats_width((size_t)W);
  const size_t k = (wp - 1) / 2;
#pragma omp parallel
  {
    uint32_t *key = (uint32_t *)malloc((wp ? wp : 1) * sizeof(uint32_t));
#pragma omp for schedule(static)
    for (int c = 0; c < nch; ++c) {
      float *row = x + (size_t)c * W;
      for (size_t t = 0; t < wp; ++t) key[t] = fkey(row[t]);
      qsort(key, wp, sizeof(uint32_t), cmp_u32);
      const float m = kval(key[k]);
      for (size_t t = 0; t < wp; ++t) key[t] = fkey(fabsf(row[t] - m));
      qsort(key, wp, sizeof(uint32_t), cmp_u32);
      const float d = kval(key[k]);
      const float sc = (float)(1.4826 * (double)d);
      if (ms) {
        ms[2 * c] = m;
        ms[2 * c + 1] = sc;
      }
      const double inv = 1.0 / (double)sc;
      for (int t = 0; t < W; ++t)
        row[t] = sc > 0.f ? (float)(((double)row[t] - (double)m) * inv) : 0.f;
    }
    free(key);
  }
}

/* bandpass_flag tail (src/flagger.cu:226-238) */
void orc_spec_flag(float *x, int nch, int W, float smin, float smax)
{
  float *bp = (float *)malloc(sizeof(float) * (size_t)nch);
  orc_bandpass(x, nch, W, bp);
#pragma omp parallel for schedule(static)
  for (int c = 0; c < nch; ++c) {
    float *row = x + (size_t)c * W;
    int bad = bp[c] < smin || bp[c] > smax;
    for (int t = 0; t < W; ++t) row[t] = bad ? 1.f : row[t] + 1.f;
  }
  free(bp);
}

static inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

/* npp_convolve_handler (src/convolution.cu:20-55): constant kernel
 * k = (float)(1/ts/fs), anchor (ts/2, fs/2), replicate border. */
void orc_box_mean(const float *x, int nch, int W, int ts, int fs, float *out)
{
  float k = (float)(1. / ts / fs);
  int ax = ts / 2, ay = fs / 2;
#pragma omp parallel for schedule(static)
  for (int c = 0; c < nch; ++c)
    for (int t = 0; t < W; ++t) {
      float acc = 0.f;
      for (int i = 0; i < fs; ++i) {
        const float *row = x + (size_t)clampi(c + i - ay, 0, nch - 1) * W;
        for (int j = 0; j < ts; ++j) {
          float p = row[clampi(t + j - ax, 0, W - 1)] * k;
          acc = acc + p;
        }
      }
      out[(size_t)c * W + t] = acc;
    }
}

/* apply_scrunch flag branch (src/flagger.cu:302-340) on a plane that has
 * already been bandpass-corrected and normalised. */
void orc_scrunch_flag(float *x, int nch, int W, int ts, int fs, float thresh,
                      int flag1, int flag2, uint32_t *flagcount, float *flagspec)
{
  const int wp = (int)stats_width((size_t)W);
  size_t n = (size_t)nch * W;
  float *S = (float *)malloc(n * sizeof(float));
  uint8_t *M = (uint8_t *)malloc(n);
  orc_box_mean(x, nch, W, ts, fs, S);
  float thr = (float)((double)thresh / sqrt(1. * ts * fs));
  float nthr = (float)(-1. * (double)thr);
#pragma omp parallel for schedule(static)
  for (size_t i = 0; i < n; ++i) M[i] = (S[i] > thr) || (S[i] < nthr);
  int ax = ts / 2, ay = fs / 2;
#pragma omp parallel for schedule(static)
  for (int c = 0; c < nch; ++c) {
    int fixed = c >= flag1 && c < flag2;
    uint32_t cnt = 0;
    uint64_t nbox = 0; /* masked pixels of the boxes of columns t < wp */
    for (int t = 0; t < W; ++t) {
      int n = 0;
      for (int i = 0; i < fs; ++i) {
        const uint8_t *row = M + (size_t)clampi(c + i - ay, 0, nch - 1) * W;
        for (int j = 0; j < ts; ++j) n += row[clampi(t + j - ax, 0, W - 1)] ? 1 : 0;
      }
      int g = n > 0;
      if (t < wp) nbox += (uint64_t)n;
      cnt += (uint32_t)g;
      float *v = x + (size_t)c * W + t;
      if (g || fixed) *v = 0.f;
      *v = *v + 1.f;
    }
    if (flagcount) flagcount[c] += cnt;
    if (flagspec)
      flagspec[c] = flagspec[c] + (float)((double)nbox / (double)(ts * fs) / (double)wp);
  }
  free(S);
  free(M);
}

double orc_chan_sum(const float *x, int nch, int W, int t)
{
  double p[ORC_TS_PARTS];
  const int per = (nch + ORC_TS_PARTS - 1) / ORC_TS_PARTS;
  for (int q = 0; q < ORC_TS_PARTS; ++q) {
    const int c0 = q * per < nch ? q * per : nch;
    const int c1 = c0 + per < nch ? c0 + per : nch;
    double s = 0.0;
    for (int c = c0; c < c1; ++c) s += (double)x[(size_t)c * W + t];
    p[q] = s;
  }
  for (int h = ORC_TS_PARTS / 2; h >= 1; h >>= 1)
    for (int q = 0; q < h; ++q) p[q] += p[q + h];
  return p[0];
}

/* ts_correct (src/timeseries.cu:19-74) */
void orc_ts_correct(float *x, int nch, int W, float *ts)
{
  /* samples are independent: each column's sum keeps its order */
#pragma omp parallel for schedule(static, 256)
  for (int t = 0; t < W; ++t) {
    const double s = orc_chan_sum(x, nch, W, t);
    float tv = (float)(s / (double)nch);
    if (ts) ts[t] = tv;
    for (int c = 0; c < nch; ++c) {
      float *v = x + (size_t)c * W + t;
      *v = *v / tv;
    }
  }
}

/* transpose_output quantiser (src/transpose.cu:65-75) */
static inline uint8_t quant(float x)
{
  float v = x + 4.f;
  v = v * (255.f / 14.f);
  if (v < 0.f) v = 0.f;
  if (v > 255.f) v = 255.f;
  if (v != v) v = 0.f;
  return (uint8_t)v;
}

void orc_quantize(const float *x, int nch, int W, uint8_t *out, size_t out_stride)
{
#pragma omp parallel for schedule(static)
  for (int c = 0; c < nch; ++c)
    for (int t = 0; t < W; ++t) out[(size_t)c * out_stride + t] = quant(x[(size_t)c * W + t]);
}

/* transpose_input (src/transpose.cu:23-42) */
void orc_ingest_u8_tf(const uint8_t *in, int nch, int W, float *x)
{
#pragma omp parallel for schedule(static)
  for (int c = 0; c < nch; ++c)
    for (int t = 0; t < W; ++t) x[(size_t)c * W + t] = (float)in[(size_t)t * nch + c];
}

void orc_ingest_f32_tf(const float *in, int nch, int W, float *x)
{
#pragma omp parallel for schedule(static)
  for (int c = 0; c < nch; ++c)
    for (int t = 0; t < W; ++t) x[(size_t)c * W + t] = in[(size_t)t * nch + c];
}

static void bandpass_correct(float *x, int nch, int W)
{
  float *bp = (float *)malloc(sizeof(float) * (size_t)nch);
  orc_bandpass(x, nch, W, bp);
  orc_divide_bp(x, nch, W, bp);
  free(bp);
}

/* Byte-histogram order of the flagger up to the time-series correction (8-bit
 * input, 1 x 1 scrunches, mean statistics).  Every value before ts_correct is
 * a function T_c(b) of the channel and the input byte, so each channel sum of
 * the reference's stages is a sum over the 256 byte values of count x value:
 *   bandpass   bp[c] = (float)(csum_v(h32[v] * T(v)) / W)   (columns < 32*floor(W/32))
 *   statistics S1 = csum_v(hA[v] * T(v)), S2 = csum_v(hA[v] * (T(v) * T(v)))
 *              (columns < wp), combined over channels by orc_stats_rows
 * with csum_v the canonical 256-lane tree (orc_csum of the 256 products, in
 * double).  The stage values T are the same float operations as on the plane.
 * On return x holds T(x) after the last bandpass division (the input of
 * ts_correct). */
static void flag_hist(const orc_flag_cfg *cfg, float *x, uint32_t *flagcount, float *flagspec)
{
  const int nch = cfg->nch, W = cfg->W;
  const int n32 = (W / 32) * 32, wp = (int)stats_width((size_t)W);
  uint32_t *H = (uint32_t *)calloc((size_t)nch * 768, sizeof(uint32_t));
  float *T = (float *)malloc(sizeof(float) * (size_t)nch * 256);
  double *S1 = (double *)malloc(sizeof(double) * (size_t)nch);
  double *S2 = (double *)malloc(sizeof(double) * (size_t)nch);
#pragma omp parallel for schedule(static)
  for (int c = 0; c < nch; ++c) {
    uint32_t *h = H + (size_t)c * 768;
    for (int t = 0; t < W; ++t) {
      const int v = (int)x[(size_t)c * W + t];
      h[(t < wp ? 0 : (t < n32 ? 256 : 512)) + v] += 1u;
    }
    for (int v = 0; v < 256; ++v) T[(size_t)c * 256 + v] = (float)v;
  }
  /* per channel: counts over [0, wp), [0, n32) and [0, W) */
#define HA(c, v) ((double)H[(size_t)(c) * 768 + (v)])
#define H32(c, v) (HA(c, v) + (double)H[(size_t)(c) * 768 + 256 + (v)])
#define HW(c, v) (H32(c, v) + (double)H[(size_t)(c) * 768 + 512 + (v)])
  double mu = 0.0, inv = 0.0;
  /* bandpass of the channel's current values */
#define BANDSUM(c, out)                                                   \
  do {                                                                    \
    double p_[256];                                                       \
    for (int v = 0; v < 256; ++v) p_[v] = H32(c, v) * (double)T[(size_t)(c) * 256 + v]; \
    out = (float)(orc_csum(p_, 256) / (double)W);                         \
  } while (0)
  /* statistics of the channels' current values, then mu, inv */
#define STATS()                                                           \
  do {                                                                    \
    _Pragma("omp parallel for schedule(static)")                          \
    for (int c = 0; c < nch; ++c) {                                       \
      double p1[256], p2[256];                                            \
      for (int v = 0; v < 256; ++v) {                                     \
        const double d = (double)T[(size_t)c * 256 + v];                  \
        p1[v] = HA(c, v) * d;                                             \
        p2[v] = HA(c, v) * (d * d);                                       \
      }                                                                   \
      S1[c] = orc_csum(p1, 256);                                          \
      S2[c] = orc_csum(p2, 256);                                          \
    }                                                                     \
    double var_;                                                          \
    orc_stats_rows(S1, S2, (size_t)nch, (double)wp, &mu, &var_);          \
    const float sf_ = (float)sqrt(var_);                                  \
    inv = 1.0 / (double)sf_;                                              \
  } while (0)
  /* bandpass_flag (flagger.cu:216-242) */
  for (int c = 0; c < nch; ++c) {
    float bp;
    BANDSUM(c, bp);
    for (int v = 0; v < 256; ++v) {
      float *y = &T[(size_t)c * 256 + v];
      *y = bp > 0.f ? *y / bp : 0.f;
    }
  }
  STATS();
  for (int c = 0; c < nch; ++c) {
    float *y = T + (size_t)c * 256, bp2;
    for (int v = 0; v < 256; ++v) y[v] = (float)(((double)y[v] - mu) * inv);
    BANDSUM(c, bp2);
    const int bad = bp2 < cfg->spec_min || bp2 > cfg->spec_max;
    for (int v = 0; v < 256; ++v) y[v] = bad ? 1.f : y[v] + 1.f;
  }
  /* configured 1 x 1 scrunches (flagger.cu:365-384) */
  for (int s = 0; s < cfg->nscrunch; ++s) {
    for (int c = 0; c < nch; ++c) {
      float *y = T + (size_t)c * 256, bp;
      BANDSUM(c, bp);
      for (int v = 0; v < 256; ++v) y[v] = bp > 0.f ? y[v] / bp : 0.f;
    }
    STATS();
    const float thr = (float)((double)cfg->scrunch[s].thresh / sqrt(1.));
    const float nthr = (float)(-1. * (double)thr);
    for (int c = 0; c < nch; ++c) {
      float *y = T + (size_t)c * 256;
      const int fixed = c >= cfg->flag1 && c < cfg->flag2;
      uint64_t cnt = 0, nbox = 0;
      for (int v = 0; v < 256; ++v) {
        float z = (float)(((double)y[v] - mu) * inv);
        const int g = (z > thr) || (z < nthr);
        if (g) {
          cnt += (uint64_t)HW(c, v);
          nbox += (uint64_t)HA(c, v);
        }
        if (g || fixed) z = 0.f;
        y[v] = z + 1.f;
      }
      if (flagcount) flagcount[c] += (uint32_t)cnt;
      if (flagspec)
        flagspec[c] = flagspec[c] + (float)((double)nbox / (double)1 / (double)wp);
    }
  }
  /* the bandpass division of apply_scrunch(8,8,100,flag=0,ts=1) */
  for (int c = 0; c < nch; ++c) {
    float *y = T + (size_t)c * 256, bp;
    BANDSUM(c, bp);
    for (int v = 0; v < 256; ++v) y[v] = bp > 0.f ? y[v] / bp : 0.f;
  }
#undef STATS
#undef BANDSUM
#undef HW
#undef H32
#undef HA
#pragma omp parallel for schedule(static)
  for (int c = 0; c < nch; ++c)
    for (int t = 0; t < W; ++t) {
      float *p = &x[(size_t)c * W + t];
      *p = T[(size_t)c * 256 + (int)*p];
    }
  free(H);
  free(T);
  free(S1);
  free(S2);
}

/* fast_flagger for one beam (src/flagger.cu:345-411) */
void orc_flag(const orc_flag_cfg *cfg, float *x, uint8_t *out, size_t out_stride,
              uint32_t *flagcount, float *flagspec)
{
  int nch = cfg->nch, W = cfg->W;
  int hist = cfg->u8 && !cfg->robust;
  for (int s = 0; hist && s < cfg->nscrunch; ++s)
    hist = cfg->scrunch[s].tscrunch == 1 && cfg->scrunch[s].fscrunch == 1;
  if (hist) {
    flag_hist(cfg, x, flagcount, flagspec);
    orc_ts_correct(x, nch, W, NULL);
    orc_normalize(x, nch, W);
    orc_quantize(x, nch, W, out, out_stride);
    return;
  }
  /* bandpass_flag (flagger.cu:216-242) */
  bandpass_correct(x, nch, W);
  if (cfg->robust) orc_normalize_robust(x, nch, W, NULL);
  else orc_normalize(x, nch, W);
  orc_spec_flag(x, nch, W, cfg->spec_min, cfg->spec_max);
  /* configured scrunches (flagger.cu:365-384) */
  for (int s = 0; s < cfg->nscrunch; ++s) {
    bandpass_correct(x, nch, W);
    if (cfg->robust) orc_normalize_robust(x, nch, W, NULL);
    else orc_normalize(x, nch, W);
    orc_scrunch_flag(x, nch, W, cfg->scrunch[s].tscrunch, cfg->scrunch[s].fscrunch,
                     cfg->scrunch[s].thresh, cfg->flag1, cfg->flag2, flagcount, flagspec);
  }
  /* apply_scrunch(8,8,100,flag=0,ts=1) (flagger.cu:385-401) */
  bandpass_correct(x, nch, W);
  orc_ts_correct(x, nch, W, NULL);
  if (cfg->robust) orc_normalize_robust(x, nch, W, NULL);
  else orc_normalize(x, nch, W);
  /* transpose_output_handler (transpose.cu:120-145) */
  orc_quantize(x, nch, W, out, out_stride);
}

/* ---- search ------------------------------------------------------------- */
int orc_nwidth(const orc_search_cfg *cfg)
{
  int n = 0;
  if (cfg->mode == ORC_MODE_HELLA) {
    for (int sm = cfg->minW; sm < cfg->maxW; sm *= 2) ++n; /* casm:186-191 */
  } else {
    for (int w = cfg->minW; w <= cfg->maxW; w *= 2) ++n;
  }
  return n;
}

int orc_width_of(const orc_search_cfg *cfg, int k) { return cfg->minW << k; }

/* src/smooth.cpp:45-58 */
void orc_hella_kernel(int sm, float *k)
{
  const float w[3] = {0.3f, 1.f, 0.3f};
  int n = 2 * sm + 1;
  float filtSum = 0.f;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < n; ++j) {
      double u = (j - sm) / (sm / 2.355);
      float v = (float)(1. - 0.5 * u * u + 0.25 * u * u * 0.25 * u * u
                        - 0.083 * u * u * 0.083 * u * u * 0.083 * u * u);
      float kv = w[i] * v * v;
      k[i * n + j] = kv;
      filtSum += w[i] * v * v;
    }
  for (int i = 0; i < 3 * n; ++i) k[i] /= filtSum;
}

typedef struct {
  const orc_search_cfg *cfg;
  const uint32_t *D;
  size_t ds;
  int k;              /* width index */
  int rows, row0;     /* plane rows and first D row */
  const float *kern;  /* hella kernel of this width */
  int w, sm;
} plane_ctx;

/* raw (un-normalised) value of plane k at (row r, sample s) */
static double plane_raw(const plane_ctx *p, int r, int s)
{
  const orc_search_cfg *cfg = p->cfg;
  int c0 = cfg->maxW / 2;
  if (cfg->mode == ORC_MODE_BOXCAR) {
    const uint32_t *row = p->D + (size_t)(r + p->row0) * p->ds;
    int lo = p->w / 2;
    uint64_t acc = 0;
    for (int j = 0; j < p->w; ++j) acc += row[s + c0 - lo + j];
    return (double)acc;
  } else {
    int n = 2 * p->sm + 1;
    float acc = 0.f;
    for (int i = 0; i < 3; ++i) {
      const uint32_t *row = p->D + (size_t)(r + p->row0 + i - 1) * p->ds;
      for (int j = 0; j < n; ++j) {
        int x = s + c0 + j;
        if (x > cfg->ntime_dd - 1) x = cfg->ntime_dd - 1;
        float t = p->kern[i * n + j] * (float)row[x];
        acc = acc + t;
      }
    }
    return (double)acc;
  }
}

long orc_search(const orc_search_cfg *cfg, const uint32_t *D, size_t d_stride,
                float *mu_w, float *sig_w, float *sigma0, int *gated,
                orc_cand *cands, float *plane_out)
{
  int nw = orc_nwidth(cfg);
  int rows = cfg->mode == ORC_MODE_HELLA ? cfg->ndm - 2 : cfg->ndm;
  int row0 = cfg->mode == ORC_MODE_HELLA ? 1 : 0;
  int nout = cfg->nout;
  size_t wp = stats_width((size_t)nout);
  size_t plane = (size_t)rows * nout;
  float *Y = (float *)malloc(sizeof(float) * plane * (size_t)nw);
  double *tmp = (double *)malloc(sizeof(double) * (size_t)nout);
  double *rr = (double *)malloc(sizeof(double) * (size_t)(rows > 0 ? rows : 1));
  double *raw = (double *)malloc(sizeof(double) * (size_t)nout);
  float *kern = NULL;

  for (int k = 0; k < nw; ++k) {
    plane_ctx p = {cfg, D, d_stride, k, rows, row0, NULL, 0, 0};
    if (cfg->mode == ORC_MODE_HELLA) {
      p.sm = orc_width_of(cfg, k);
      kern = (float *)realloc(kern, sizeof(float) * 3 * (size_t)(2 * p.sm + 1));
      orc_hella_kernel(p.sm, kern);
      p.kern = kern;
    } else {
      p.w = orc_width_of(cfg, k);
    }
    /* per-width mean / variance (smooth.cpp:95-124 -> calculate_stddev) */
    double *rawplane = (double *)malloc(sizeof(double) * plane);
    if (cfg->mode == ORC_MODE_BOXCAR) {
      /* exact integer box sums from a u64 prefix sum of each row */
      const int c0 = cfg->maxW / 2, w = p.w, lo = w / 2;
#pragma omp parallel for schedule(static)
      for (int r = 0; r < rows; ++r) {
        const uint32_t *row = D + (size_t)(r + row0) * d_stride;
        uint64_t *P = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)(cfg->ntime_dd + 1));
        P[0] = 0;
        for (int t = 0; t < cfg->ntime_dd; ++t) P[t + 1] = P[t] + row[t];
        for (int s = 0; s < nout; ++s) {
          int a = s + c0 - lo;
          rawplane[(size_t)r * nout + s] = (double)(P[a + w] - P[a]);
        }
        free(P);
      }
    } else {
#pragma omp parallel for schedule(static)
      for (int r = 0; r < rows; ++r)
        for (int s = 0; s < nout; ++s) rawplane[(size_t)r * nout + s] = plane_raw(&p, r, s);
    }
    double m, v;
    {
      double *S1 = (double *)malloc(sizeof(double) * (size_t)(rows > 0 ? rows : 1));
      double *S2 = (double *)malloc(sizeof(double) * (size_t)(rows > 0 ? rows : 1));
      /* rows are independent: threads only change which core sums a row */
#pragma omp parallel for schedule(static)
      for (int r = 0; r < rows; ++r) {
        const double *row = rawplane + (size_t)r * nout;
        if (cfg->mode == ORC_MODE_BOXCAR) {
          /* exact integer sums of the box values and their squares, rounded
           * to double once (S2 = hi * 2^64 + lo) */
          uint64_t s1 = 0;
          unsigned __int128 s2 = 0;
          for (size_t j = 0; j < wp; ++j) {
            const uint64_t b = (uint64_t)row[j];
            s1 += b;
            s2 += (unsigned __int128)b * b;
          }
          S1[r] = (double)s1;
          S2[r] = (double)(uint64_t)(s2 >> 64) * 18446744073709551616.0 + (double)(uint64_t)s2;
        } else {
          double *sq = (double *)malloc(sizeof(double) * (wp ? wp : 1));
          for (size_t j = 0; j < wp; ++j) sq[j] = row[j] * row[j];
          S1[r] = orc_csum(row, wp);
          S2[r] = orc_csum(sq, wp);
          free(sq);
        }
      }
      orc_stats_rows(S1, S2, (size_t)rows, (double)wp, &m, &v);
      free(S1);
      free(S2);
    }
    float sf = (float)sqrt(v);
    mu_w[k] = (float)m;
    sig_w[k] = sf;
    float *Yk = Y + (size_t)k * plane;
    if (sf > 0.f) {
      double inv = 1.0 / (double)sf;
#pragma omp parallel for schedule(static)
      for (size_t i = 0; i < plane; ++i) Yk[i] = (float)((rawplane[i] - m) * inv);
    } else {
      for (size_t i = 0; i < plane; ++i) Yk[i] = (float)rawplane[i];
    }
    free(rawplane);
  }
  (void)raw;

  /* find_peaks (src/peaks.cu:31-117): sigma of the normalised first width */
  float mu0, s0;
  orc_stats_f(Y, (size_t)rows, (size_t)nout, (size_t)nout, &mu0, &s0);
  *sigma0 = s0;
  int ok = (s0 < 5.f && s0 > 0.1f);
  *gated = !ok;
  long count = 0;
  if (ok) {
    float thr = cfg->snr * s0;
    for (int k = 0; k < nw; ++k) {
      const float *Yk = Y + (size_t)k * plane;
      for (int r = 0; r < rows; ++r)
        for (int s = 0; s < nout; ++s) {
          float y = Yk[(size_t)r * nout + s];
          if (y > thr) {
            if (count < cfg->max_giants) {
              orc_cand *c = &cands[count];
              c->snr = y;
              c->samp = s;
              c->width = k;
              c->dm_idx = r + row0;
            }
            ++count;
          }
        }
    }
  }
  if (plane_out) memcpy(plane_out, Y, sizeof(float) * plane * (size_t)nw);
  free(Y);
  free(tmp);
  free(rr);
  free(raw);
  free(kern);
  return count;
}

/* ---- clustering ---------------------------------------------------------- */
typedef struct { int64_t key; int idx; } keyidx;

static int cmp_key(const void *a, const void *b)
{
  int64_t x = ((const keyidx *)a)->key, y = ((const keyidx *)b)->key;
  if (x != y) return x < y ? -1 : 1;
  int i = ((const keyidx *)a)->idx, j = ((const keyidx *)b)->idx;
  return i < j ? -1 : (i > j);
}

static int uf_find(int *par, int x)
{
  while (par[x] != x) {
    par[x] = par[par[x]];
    x = par[x];
  }
  return x;
}

static int64_t cell_key(int cdm, int ct) { return ((int64_t)cdm << 32) | (uint32_t)ct; }

int orc_cluster(const orc_cand *c, int n, int cell_t, int cell_dm, orc_cluster_t *out)
{
  if (n <= 0) return 0;
  keyidx *ki = (keyidx *)malloc(sizeof(keyidx) * (size_t)n);
  for (int i = 0; i < n; ++i) {
    ki[i].key = cell_key(c[i].dm_idx / cell_dm, c[i].samp / cell_t);
    ki[i].idx = i;
  }
  qsort(ki, (size_t)n, sizeof(keyidx), cmp_key);
  /* unique cells */
  int64_t *cells = (int64_t *)malloc(sizeof(int64_t) * (size_t)n);
  int *cell_of = (int *)malloc(sizeof(int) * (size_t)n);
  int nc = 0;
  for (int i = 0; i < n; ++i) {
    if (i == 0 || ki[i].key != ki[i - 1].key) cells[nc++] = ki[i].key;
    cell_of[ki[i].idx] = nc - 1;
  }
  int *par = (int *)malloc(sizeof(int) * (size_t)nc);
  for (int i = 0; i < nc; ++i) par[i] = i;
  for (int i = 0; i < nc; ++i) {
    int cdm = (int)(cells[i] >> 32), ct = (int)(uint32_t)cells[i];
    for (int dd = -1; dd <= 1; ++dd)
      for (int dt = -1; dt <= 1; ++dt) {
        if (!dd && !dt) continue;
        int64_t nk = cell_key(cdm + dd, ct + dt);
        /* binary search */
        int lo = 0, hi = nc - 1, f = -1;
        while (lo <= hi) {
          int mid = (lo + hi) / 2;
          if (cells[mid] == nk) { f = mid; break; }
          if (cells[mid] < nk) lo = mid + 1; else hi = mid - 1;
        }
        if (f >= 0) {
          int a = uf_find(par, i), b = uf_find(par, f);
          if (a != b) { if (a < b) par[b] = a; else par[a] = b; }
        }
      }
  }
  /* representative per root */
  int *rep = (int *)malloc(sizeof(int) * (size_t)nc);
  for (int i = 0; i < nc; ++i) rep[i] = -1;
  for (int i = 0; i < n; ++i) {
    int r = uf_find(par, cell_of[i]);
    int cur = rep[r];
    if (cur < 0 || c[i].snr > c[cur].snr) rep[r] = i;
  }
  int nclu = 0;
  int *slot = (int *)malloc(sizeof(int) * (size_t)nc);
  for (int i = 0; i < nc; ++i) slot[i] = -1;
  /* order clusters by representative index */
  int *reps = (int *)malloc(sizeof(int) * (size_t)nc);
  for (int i = 0; i < nc; ++i)
    if (rep[i] >= 0 && uf_find(par, i) == i) reps[nclu++] = rep[i];
  /* sort reps ascending (insertion via qsort on ints) */
  for (int i = 1; i < nclu; ++i) {
    int v = reps[i], j = i - 1;
    while (j >= 0 && reps[j] > v) { reps[j + 1] = reps[j]; --j; }
    reps[j + 1] = v;
  }
  for (int q = 0; q < nclu; ++q) {
    int r = uf_find(par, cell_of[reps[q]]);
    slot[r] = q;
    out[q].rep = reps[q];
    out[q].count = 0;
    out[q].samp_min = out[q].dm_min = 0x7fffffff;
    out[q].samp_max = out[q].dm_max = -1;
  }
  for (int i = 0; i < n; ++i) {
    int q = slot[uf_find(par, cell_of[i])];
    orc_cluster_t *o = &out[q];
    o->count++;
    if (c[i].samp < o->samp_min) o->samp_min = c[i].samp;
    if (c[i].samp > o->samp_max) o->samp_max = c[i].samp;
    if (c[i].dm_idx < o->dm_min) o->dm_min = c[i].dm_idx;
    if (c[i].dm_idx > o->dm_max) o->dm_max = c[i].dm_idx;
  }
  free(ki); free(cells); free(cell_of); free(par); free(rep); free(slot); free(reps);
  return nclu;
}
