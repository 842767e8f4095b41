// flagger.hip -- RFI flagger and bandpass normaliser (FP32) for gfx950.
//
// Restates fast_flagger (src/hella/src/flagger.cu:345-411) for one or many
// beams of a gulp.  Data live in a float channel-major plane X[c][t].
// Every pass is HBM-bound streaming work; reductions follow one canonical
// order so results are reproducible bit for bit:
//   * a row is cut into blocks of 4096 samples; in block b, lane l of a
//     256-thread workgroup sums samples b*4096 + l + 256k (k ascending) in
//     double, then the 256 lane sums are folded pairwise (128, 64, ..., 1);
//   * block sums of a row are added in ascending order starting from 0;
//   * row values are combined over rows by the same blocked rule.
// Elementwise passes use the same (row, block, lane) mapping, which lets the
// reduction of a pass's output ride along in the pass that writes it.
#include "flagger.hpp"
#include "reduce.hpp"

namespace hamd {

namespace {

constexpr int NT = 256;

// A load through the global address space: it counts on vmcnt alone, so the
// waits of the LDS accesses between such loads and their use (lgkmcnt) do not
// also wait for them, as they wait for a flat load.
template <class T>
__device__ __forceinline__ T gld(const T *p)
{
  typedef __attribute__((address_space(1))) const T gt;
  return *((gt *)p);
}

// the partial sums of (channel c, block b), channels innermost: a thread per
// channel reads a block's partials of consecutive channels coalesced
__device__ __forceinline__ int64_t fslot(const FlagBeamDev &B, int c, int b)
{
  return (int64_t)b * B.nch + c;
}

__device__ __forceinline__ uint8_t quant(float x)
{
  // transpose_output (src/hella/src/transpose.cu:65-75); NaN -> 0
  float v = x + 4.f;
  v = v * (255.f / 14.f);
  if (v < 0.f) v = 0.f;
  if (v > 255.f) v = 255.f;
  if (v != v) v = 0.f;
  return (uint8_t)v;
}

// The normalisation of a pass, (float)(((double)x - a) * b): the plane's
// mean and 1/sigma (normalize_data, flagger.cu:198-214), or with robust
// statistics the channel's median m and 1/s for its scale s; zero when s is
// 0 (a channel whose samples are mostly one value), and then every value
// is 0.  Both forms are the same two double operations per value, so a and
// b are formed once per channel (a per-value branch between the forms was
// compiled into both paths and a select).
struct NormP {
  double a, b;
  bool zero;
};

// read through the constant address space (scalar loads at wave-uniform c):
// the pass that reads them never writes them
__device__ __forceinline__ NormP norm_params(const FlagBeamDev &B, int c)
{
  typedef __attribute__((address_space(4))) const FScal FScalK;
  typedef __attribute__((address_space(4))) const float FK;
  const FScalK &sc = *(FScalK *)B.scal;
  if (!B.robust) return NormP{sc.mu, sc.inv, false};
  const float m = ((FK *)B.rstat)[2 * c], sd = ((FK *)B.rstat)[2 * c + 1];
  return NormP{(double)m, 1.0 / (double)sd, !(sd > 0.f)};
}

__device__ __forceinline__ float normv(float x, const NormP &n)
{
  return n.zero ? 0.f : (float)(((double)x - n.a) * n.b);
}

// ---- ingest: input gulp -> X (transpose_input, src/hella/src/transpose.cu:23-42)
__global__ __launch_bounds__(256) void k_ingest(const FlagBeamDev *__restrict__ beams, int W)
{
  const FlagBeamDev B = beams[blockIdx.z];
  const int c0 = blockIdx.y * 64;
  const int t0 = blockIdx.x * 64;
  if (c0 >= B.nch) return;
  __shared__ float tile[64][65];
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  if (B.in_fmt == FLAG_IN_F16_FT) {
    const _Float16 *in = static_cast<const _Float16 *>(B.in);
    for (int r = ty; r < 64; r += 4) {
      const int c = c0 + r, t = t0 + tx;
      if (c < B.nch && t < W) B.X[(int64_t)c * B.xs + t] = (float)in[(int64_t)c * B.in_stride + t];
    }
    return;
  }
  for (int r = ty; r < 64; r += 4) {
    const int t = t0 + r, c = c0 + tx;
    float v = 0.f;
    if (t < W && c < B.nch) {
      if (B.in_fmt == FLAG_IN_U8_TF)
        v = (float)static_cast<const uint8_t *>(B.in)[(int64_t)t * B.in_stride + c];
      else
        v = static_cast<const float *>(B.in)[(int64_t)t * B.in_stride + c];
    }
    tile[r][tx] = v;
  }
  __syncthreads();
  for (int r = ty; r < 64; r += 4) {
    const int c = c0 + r, t = t0 + tx;
    if (c < B.nch && t < W) B.X[(int64_t)c * B.xs + t] = tile[tx][r];
  }
}

// ---- generic elementwise pass + canonical row partial of its output
// a0/a1: SPEC -> channel-mean range; NORM_FLAG1 -> +/- threshold
__global__ __launch_bounds__(NT) void k_apply(const FlagBeamDev *__restrict__ beams, int W,
                                              int op, int part, int npart, float a0, float a1)
{
  const FlagBeamDev B = beams[blockIdx.z];
  const int c = blockIdx.y;
  if (c >= B.nch) return;
  const int blk = blockIdx.x;
  const int l = threadIdx.x;
  float *row = B.X + (int64_t)c * B.xs;
  const int base = blk * FLAG_BLOCK;
  __shared__ double sh[2 * 128];
  const NormP np = norm_params(B, c);
  float bpv = 0.f;
  bool bad = false;
  if (op == FOP_DIVBP || op == FOP_DIVBP_TS) bpv = B.bp[c];
  if (op == FOP_SPEC) {
    const float b2 = B.bp[c];
    bad = (b2 < a0) || (b2 > a1);
  }
  const bool fixed = c >= B.flag1 && c < B.flag2;
  uint32_t cnt = 0, cntw = 0;
  double acc[2] = {0.0, 0.0};
  // every sample of the thread (and its time-series value) loaded before the
  // first is used: the thread reads and writes only its own samples
  constexpr int KB = FLAG_BLOCK / NT;
  float xin[KB], tsv[KB];
#pragma unroll
  for (int k = 0; k < KB; ++k) {
    const int tc = min(base + l + NT * k, W - 1);
    xin[k] = gld(row + tc);
    tsv[k] = op == FOP_DIVBP_TS ? gld(B.ts + tc) : 1.f;
  }
#pragma unroll 4
  for (int k = 0; k < KB; ++k) {
    const int t = base + l + NT * k;
    if (t >= W) break;
    float x = xin[k];
    switch (op) {
      case FOP_DIVBP: x = bpv > 0.f ? x / bpv : 0.f; row[t] = x; break;
      case FOP_NORM: x = normv(x, np); row[t] = x; break;
      case FOP_SPEC: x = bad ? 1.f : x + 1.f; row[t] = x; break;
      case FOP_NORM_QUANT: {
        x = normv(x, np);
        int64_t col = B.col0 + t;
        if (col >= B.ring_len) col -= B.ring_len;
        B.ring[(int64_t)c * B.ring_stride + col] = quant(x);
        break;
      }
      case FOP_NORM_FLAG1: {
        // box of 1 x 1: the box mean is the sample itself (kernel 1.0f)
        x = normv(x, np);
        const bool g = (x > a0) || (x < a1);
        cnt += g ? 1u : 0u;
        cntw += (g && t < B.wp) ? 1u : 0u;
        if (g || fixed) x = 0.f;
        x = x + 1.f;
        row[t] = x;
        break;
      }
      case FOP_DIVBP_TS: {
        x = bpv > 0.f ? x / bpv : 0.f;
        x = x / tsv[k];
        row[t] = x;
        break;
      }
      default: break;
    }
    if (t < npart && part != FPART_NONE) {
      const double d = (double)x;
      acc[0] += d;
      acc[1] += d * d;
    }
  }
  if (cnt) atomicAdd(&B.flagcount[c], cnt);
  if (cntw) atomicAdd(&B.gsum[c], (unsigned long long)cntw);
  if (part == FPART_NONE) return;
  tree256_multi(acc, sh);
  if (l == 0) {
    B.part[fslot(B, c, blk)] = acc[0];
    if (part == FPART_SUM2) B.part2[fslot(B, c, blk)] = acc[1];
  }
}

// bandpass from partials: bp[c] = (float)(sum_b part[c][b] / W).  After a
// scrunch (box = ts * fs > 0) also the flag spectrum: the mean of the grown
// mask G over the statistics columns, G being the fraction of masked pixels
// in each box (calc_bandpass + add_bandpass, flagger.cu:333-334):
// fspec[c] += (float)(sum / box / wp), the scrunch's sum then cleared.
__global__ __launch_bounds__(NT) void k_finish_bp(const FlagBeamDev *__restrict__ beams, int W,
                                                  int box)
{
  const FlagBeamDev B = beams[blockIdx.z];
  const int c = blockIdx.x * NT + threadIdx.x;
  if (c >= B.nch) return;
  double s = 0.0;
  for (int b = 0; b < B.nblk; ++b) s += gld(B.part + fslot(B, c, b));
  B.bp[c] = (float)(s / (double)W);
  if (box > 0) {
    const double g = (double)B.gsum[c];
    B.fspec[c] = B.fspec[c] + (float)(g / (double)box / (double)B.wp);
    B.gsum[c] = 0ull;
  }
}

// Statistics of the plane from the block sums (calculate_stddev,
// normalization.cu:125-162): S1_r, S2_r = ascending sums of row r's block
// partials; mu = canonical sum over rows of S1_r / n, over the row count; the
// variance is the same combination of ((S2_r - 2 mu S1_r) + n mu mu) / n.
// The row values of a 4096-row block are formed by all 1024 threads at once
// (rowv), then lane l < 256 adds rows l + 256k in ascending k and the lanes
// are folded by the canonical tree.
constexpr int FS_NT = 1024;

__device__ double flag_rows_csum(const FlagBeamDev &B, int n, double mu, bool var, double *rowv,
                                 double *sh)
{
  const int l = threadIdx.x;
  const double dn = (double)n;
  double total = 0.0;
  for (int r0 = 0; r0 < B.nch; r0 += FLAG_BLOCK) {
    for (int i = l; i < FLAG_BLOCK; i += FS_NT) {
      const int r = r0 + i;
      double val = 0.0;
      if (r < B.nch) {
        double s1 = 0.0;
#pragma unroll 8
        for (int b = 0; b < B.nblk; ++b) s1 += gld(B.part + fslot(B, r, b));
        if (!var) {
          val = s1 / dn;
        } else {
          double s2 = 0.0;
#pragma unroll 8
          for (int b = 0; b < B.nblk; ++b) s2 += gld(B.part2 + fslot(B, r, b));
          val = ((s2 - (2.0 * mu) * s1) + (dn * mu) * mu) / dn;
        }
      }
      rowv[i] = val;
    }
    __syncthreads();
    double acc[1] = {0.0};
    if (l < NT) {
      for (int k = 0; k < FLAG_BLOCK / NT; ++k) {
        if (r0 + l + NT * k >= B.nch) break;
        acc[0] += rowv[l + NT * k];
      }
    }
    tree256_multi(acc, sh);
    if (l == 0) sh[128] = acc[0];
    __syncthreads();
    total += sh[128];
    __syncthreads();
  }
  return total / (double)B.nch;
}

// Lanes l < 256 add the row values rowv[l + 256k] of rows below nch in
// ascending k, the canonical tree folds the lanes (flag_rows_csum's last
// steps for one 4096-row group).
__device__ double flag_rows_group(int nch, const double *rowv, double *sh)
{
  const int l = threadIdx.x;
  __syncthreads();
  double acc[1] = {0.0};
  if (l < NT) {
    for (int k = 0; k < FLAG_BLOCK / NT; ++k) {
      if (l + NT * k >= nch) break;
      acc[0] += rowv[l + NT * k];
    }
  }
  tree256_multi(acc, sh);
  if (l == 0) sh[128] = acc[0];
  __syncthreads();
  double total = 0.0;
  total += sh[128];
  __syncthreads();
  return total / (double)nch;
}

__global__ __launch_bounds__(FS_NT) void k_finish_stats(const FlagBeamDev *__restrict__ beams, int n)
{
  const FlagBeamDev B = beams[blockIdx.x];
  __shared__ double rowv[FLAG_BLOCK];
  __shared__ double sh[136];
  double mu, var;
  if (B.nch <= FLAG_BLOCK) {
    // one group of rows: each thread sums its rows' S1 and S2 once, with
    // every load issued before the additions, and keeps them for both passes
    constexpr int RPT = FLAG_BLOCK / FS_NT;
    const int l = threadIdx.x;
    const double dn = (double)n;
    double s1[RPT], s2[RPT];
#pragma unroll
    for (int i = 0; i < RPT; ++i) {
      const int r = l + FS_NT * i;
      s1[i] = 0.0;
      s2[i] = 0.0;
      if (r < B.nch) {
#pragma unroll 8
        for (int b = 0; b < B.nblk; ++b) {
          s1[i] += gld(B.part + fslot(B, r, b));
          s2[i] += gld(B.part2 + fslot(B, r, b));
        }
      }
    }
#pragma unroll
    for (int i = 0; i < RPT; ++i) rowv[l + FS_NT * i] = l + FS_NT * i < B.nch ? s1[i] / dn : 0.0;
    mu = flag_rows_group(B.nch, rowv, sh);
#pragma unroll
    for (int i = 0; i < RPT; ++i)
      rowv[l + FS_NT * i] =
          l + FS_NT * i < B.nch ? ((s2[i] - (2.0 * mu) * s1[i]) + (dn * mu) * mu) / dn : 0.0;
    var = flag_rows_group(B.nch, rowv, sh);
  } else {
    mu = flag_rows_csum(B, n, 0.0, false, rowv, sh);
    var = flag_rows_csum(B, n, mu, true, rowv, sh);
  }
  if (threadIdx.x == 0) {
    const float sf = (float)sqrt(var);
    B.scal->mu = mu;
    B.scal->sigf = sf;
    B.scal->inv = 1.0 / (double)sf;
  }
}

// The plane's statistics of a table step's values (k_hstep), formed by
// each workgroup of the next step from the per-channel sums hs[c] = (S1, S2):
// the values and order of k_finish_stats over partials whose blocks past the
// first are zero (adding a zero block leaves a row sum unchanged), lane l
// forming rows l + 256k itself in ascending k.
__device__ NormP hstep_norm(const FlagBeamDev &B, const double *hs, double *sh)
{
  const int l = threadIdx.x;
  const double dn = (double)B.wp;
  double mu = 0.0, var = 0.0;
  for (int pass = 0; pass < 2; ++pass) {
    double total = 0.0;
    for (int r0 = 0; r0 < B.nch; r0 += FLAG_BLOCK) {
      double acc[1] = {0.0};
#pragma unroll 4
      for (int k = 0; k < FLAG_BLOCK / NT; ++k) {
        const int r = r0 + l + NT * k;
        // both words loaded at a clamped row, outside the branch
        const int rc = min(r, B.nch - 1);
        const double h1 = gld(hs + 2 * rc), h2 = gld(hs + 2 * rc + 1);
        if (r < B.nch) {
          const double s1 = 0.0 + h1;
          if (pass == 0) {
            acc[0] += s1 / dn;
          } else {
            const double s2 = 0.0 + h2;
            acc[0] += ((s2 - (2.0 * mu) * s1) + (dn * mu) * mu) / dn;
          }
        }
      }
      tree256_multi(acc, sh);
      if (l == 0) sh[128] = acc[0];
      __syncthreads();
      total += sh[128];
      __syncthreads();
    }
    if (pass == 0)
      mu = total / (double)B.nch;
    else
      var = total / (double)B.nch;
  }
  const float sf = (float)sqrt(var);
  return NormP{mu, 1.0 / (double)sf, false};
}

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// S = box mean over (ts x fs) with anchor (ts/2, fs/2), replicate border,
// taps row-major into a float from 0 (npp_convolve_handler,
// src/hella/src/convolution.cu:20-55); mask = |S| > thr.
__global__ __launch_bounds__(NT) void k_box_thresh(const FlagBeamDev *__restrict__ beams, int W,
                                                   int ts, int fs, float k, float thr,
                                                   float nthr)
{
  const FlagBeamDev B = beams[blockIdx.z];
  const int c = blockIdx.y;
  if (c >= B.nch) return;
  const int ax = ts / 2, ay = fs / 2;
  const int base = blockIdx.x * FLAG_BLOCK;
  for (int kk = 0; kk < FLAG_BLOCK / NT; ++kk) {
    const int t = base + threadIdx.x + NT * kk;
    if (t >= W) break;
    float acc = 0.f;
    for (int i = 0; i < fs; ++i) {
      const float *row = B.X + (int64_t)clampi(c + i - ay, 0, B.nch - 1) * B.xs;
      for (int j = 0; j < ts; ++j) {
        const float p = row[clampi(t + j - ax, 0, W - 1)] * k;
        acc = acc + p;
      }
    }
    B.mask[(int64_t)c * B.xs + t] = (uint8_t)((acc > thr) || (acc < nthr));
  }
}

// grown mask (any masked pixel inside the window), replace, +1, count, and
// the canonical row partial of the result (apply_scrunch flag branch,
// src/hella/src/flagger.cu:322-335)
__global__ __launch_bounds__(NT) void k_grow_replace(const FlagBeamDev *__restrict__ beams, int W,
                                                     int ts, int fs, int npart)
{
  const FlagBeamDev B = beams[blockIdx.z];
  const int c = blockIdx.y;
  if (c >= B.nch) return;
  const int ax = ts / 2, ay = fs / 2;
  const int base = blockIdx.x * FLAG_BLOCK;
  const int l = threadIdx.x;
  const bool fixed = c >= B.flag1 && c < B.flag2;
  __shared__ double sh[128];
  double acc[1] = {0.0};
  uint32_t cnt = 0;
  unsigned long long nbox = 0;
  float *row = B.X + (int64_t)c * B.xs;
  for (int kk = 0; kk < FLAG_BLOCK / NT; ++kk) {
    const int t = base + l + NT * kk;
    if (t >= W) break;
    bool g = false;
    if (t < B.wp) {
      // the statistics columns: every masked pixel of the box counts (G's sum)
      uint32_t n = 0;
      for (int i = 0; i < fs; ++i) {
        const uint8_t *mrow = B.mask + (int64_t)clampi(c + i - ay, 0, B.nch - 1) * B.xs;
        for (int j = 0; j < ts; ++j) n += mrow[clampi(t + j - ax, 0, W - 1)] ? 1u : 0u;
      }
      nbox += n;
      g = n > 0;
    } else {
      for (int i = 0; i < fs && !g; ++i) {
        const uint8_t *mrow = B.mask + (int64_t)clampi(c + i - ay, 0, B.nch - 1) * B.xs;
        for (int j = 0; j < ts; ++j)
          if (mrow[clampi(t + j - ax, 0, W - 1)]) { g = true; break; }
      }
    }
    cnt += g ? 1u : 0u;
    float x = row[t];
    if (g || fixed) x = 0.f;
    x = x + 1.f;
    row[t] = x;
    if (t < npart) acc[0] += (double)x;
  }
  if (cnt) atomicAdd(&B.flagcount[c], cnt);
  if (nbox) atomicAdd(&B.gsum[c], nbox);
  tree256_multi(acc, sh);
  if (l == 0) B.part[fslot(B, c, blockIdx.x)] = acc[0];
}

// ts_correct (src/hella/src/timeseries.cu:19-74) on the bandpass-corrected
// plane: ts[t] = (sum_c x / bp[c]) / nch in double, the channels summed as 16
// partitions of consecutive channels (each in channel order), the partials
// combined pairwise (p[i] += p[i + h], h = 8, 4, 2, 1).  One wave per
// partition, 64 samples per workgroup; the division by ts happens in the
// following FOP_DIVBP_TS pass.
constexpr int TS_PARTS = 16;

__global__ __launch_bounds__(64 * TS_PARTS) void k_ts_sum(const FlagBeamDev *__restrict__ beams,
                                                           int W)
{
  const FlagBeamDev B = beams[blockIdx.z];
  __shared__ double sp[TS_PARTS][64];
  const int lane = threadIdx.x & 63, q = threadIdx.x >> 6;
  const int t = blockIdx.x * 64 + lane;
  const int per = (B.nch + TS_PARTS - 1) / TS_PARTS;
  const int c0 = min(B.nch, q * per), c1 = min(B.nch, c0 + per);
  double s = 0.0;
  if (t < W) {
    const float *col = B.X + t;
    typedef __attribute__((address_space(4))) const float FK;
#pragma unroll 8
    for (int c = c0; c < c1; ++c) {
      const float b = ((FK *)B.bp)[c];  // wave-uniform: a scalar load
      const float x = gld(col + (int64_t)c * B.xs);
      s += (double)(b > 0.f ? x / b : 0.f);
    }
  }
  sp[q][lane] = s;
  __syncthreads();
  if (q == 0 && t < W) {
    double p[TS_PARTS];
#pragma unroll
    for (int i = 0; i < TS_PARTS; ++i) p[i] = sp[i][lane];
#pragma unroll
    for (int h = TS_PARTS / 2; h >= 1; h >>= 1)
#pragma unroll
      for (int i = 0; i < h; ++i) p[i] += p[i + h];
    const float tv = (float)(p[0] / (double)B.nch);
  B.ts[t] = tv;
  B.rts[t] = 1.0 / (double)tv;
  }
}

// ---- value-table path -------------------------------------------------
// The same operations as k_apply, applied to one value of channel c; the
// statistics (bp, mu, inv) are those k_apply would read in the same pass.
__device__ __forceinline__ float value_op(int op, float x, float bpv, bool bad, bool fixed,
                                          const NormP &np, float a0, float a1, bool &g)
{
  g = false;
  switch (op) {
    case FOP_DIVBP:
    case FOP_DIVBP_TS: return bpv > 0.f ? x / bpv : 0.f;
    case FOP_NORM: return normv(x, np);
    case FOP_SPEC: return bad ? 1.f : x + 1.f;
    case FOP_NORM_FLAG1: {
      x = normv(x, np);
      g = (x > a0) || (x < a1);
      if (g || fixed) x = 0.f;
      return x + 1.f;
    }
    default: return x;  // FOP_NONE; FOP_NORM_QUANT acts after the time-series division
  }
}

// u8 time-major gulp -> u8 channel-major plane X8, 64 x 64 byte tiles: a
// thread reads 16 channels of one sample (one 16-byte load when the rows
// allow it), scatters them into the transposed LDS tile and writes 16
// samples of one channel with one 16-byte store.
typedef uint32_t u32x4f __attribute__((ext_vector_type(4)));


__global__ __launch_bounds__(256) void k_ingest_u8(const FlagBeamDev *__restrict__ beams, int W)
{
  const FlagBeamDev B = beams[blockIdx.z];
  const int c0 = blockIdx.y * 64;
  const int t0 = blockIdx.x * 64;
  if (c0 >= B.nch) return;
  __shared__ __align__(16) uint8_t tile[64][80];  // [channel][sample]
  const int r = threadIdx.x >> 2, q = (threadIdx.x & 3) * 16;
  const uint8_t *in = static_cast<const uint8_t *>(B.in);
  {
    const int t = t0 + r, c = c0 + q;
    const uint8_t *src = in + (int64_t)t * B.in_stride + c;
    uint8_t v[16];
    if (t < W && c + 16 <= B.nch && ((reinterpret_cast<uintptr_t>(src) & 15) == 0)) {
      const u32x4f w = gld(reinterpret_cast<const u32x4f *>(src));
#pragma unroll
      for (int i = 0; i < 16; ++i) v[i] = (uint8_t)(w[i >> 2] >> (8 * (i & 3)));
    } else {
#pragma unroll
      for (int i = 0; i < 16; ++i) v[i] = (t < W && c + i < B.nch) ? src[i] : (uint8_t)0;
    }
#pragma unroll
    for (int i = 0; i < 16; ++i) tile[q + i][r] = v[i];
  }
  __syncthreads();
  const int c = c0 + r, t = t0 + q;
  if (c >= B.nch || t >= W) return;
  uint8_t *dst = B.X8 + (int64_t)c * B.xs + t;
  if (t + 16 <= W) {
    *reinterpret_cast<u32x4f *>(dst) = *reinterpret_cast<const u32x4f *>(&tile[r][q]);
  } else {
    for (int i = 0; i < 16 && t + i < W; ++i) dst[i] = tile[r][q + i];
  }
}

// One pass of the value-table path over NB consecutive blocks of a channel:
// the workgroup builds the channel's table for this pass from the previous
// table (lin < 0: the identity) and the current statistics, the workgroup of
// the channel's first blocks stores it to slot lout, and every sample is
// looked up.  The bytes arrive with 16-byte loads, one per thread and block,
// and are read back from LDS in the canonical lane order; the NB blocks'
// partial sums share one tree.  Samples, partial sums and flag counts are
// those of k_apply.
template <int NB>
__global__ __launch_bounds__(NT) void k_apply_lut(const FlagBeamDev *__restrict__ beams, int W,
                                                  int op, int part, int npart, float a0, float a1,
                                                  int lin, int lout)
{
  const FlagBeamDev B = beams[blockIdx.z];
  const int c = blockIdx.y;
  if (c >= B.nch) return;
  const int blk0 = blockIdx.x * NB;
  if (blk0 >= B.nblk) return;
  const int l = threadIdx.x;
  __shared__ float tab[256];
  __shared__ uint8_t flg[256];
  __shared__ __align__(16) uint8_t bytes[NB][FLAG_BLOCK];
  __shared__ double sh[2 * NB * 128];
  const uint8_t *row = B.X8 + (int64_t)c * B.xs;
  // loads first: the blocks' bytes (the row holds xs >= W rounded up to 64
  // bytes, so a 16-byte load that starts below W stays in the row) and the
  // previous table entry
  u32x4f d[NB];
#pragma unroll
  for (int j = 0; j < NB; ++j) {
    const int t = (blk0 + j) * FLAG_BLOCK + 16 * l;
    // unconditional at a clamped 16-byte column, zeroed past W
    const u32x4f v = gld(reinterpret_cast<const u32x4f *>(row + min(t, ((W - 1) & ~15))));
    d[j] = t < W ? v : u32x4f{0u, 0u, 0u, 0u};
  }
  const float xl = gld(B.lut + ((int64_t)max(lin, 0) * B.nch + c) * 256 + l);
  const float x0 = lin < 0 ? (float)l : xl;
  const NormP np = norm_params(B, c);
  const float bpv = B.bp[c];
  const bool bad = op == FOP_SPEC && ((bpv < a0) || (bpv > a1));
  const bool fixed = c >= B.flag1 && c < B.flag2;
  {
    bool g;
    const float y = value_op(op, x0, bpv, bad, fixed, np, a0, a1, g);
    tab[l] = y;
    flg[l] = g ? 1 : 0;
    if (blk0 == 0 && lout >= 0 && lout != lin) B.lut[((int64_t)lout * B.nch + c) * 256 + l] = y;
  }
#pragma unroll
  for (int j = 0; j < NB; ++j) *reinterpret_cast<u32x4f *>(&bytes[j][16 * l]) = d[j];
  __syncthreads();
  uint32_t cnt = 0, cntw = 0;
  double acc[2 * NB];
#pragma unroll
  for (int j = 0; j < 2 * NB; ++j) acc[j] = 0.0;
  const bool sums = part != FPART_NONE;
  constexpr int KB = FLAG_BLOCK / NT;
#pragma unroll
  for (int j = 0; j < NB; ++j) {
    const int base = (blk0 + j) * FLAG_BLOCK;
    if (base + FLAG_BLOCK <= W && (!sums || base + FLAG_BLOCK <= npart)) {
      // a block with every sample inside the gulp (and the statistics
      // columns): the same operations in the same order, with every LDS
      // read of the block issued before the arithmetic
      int v[KB];
      float x[KB];
#pragma unroll
      for (int k = 0; k < KB; ++k) v[k] = bytes[j][l + NT * k];
#pragma unroll
      for (int k = 0; k < KB; ++k) {
        x[k] = tab[v[k]];
        cnt += flg[v[k]];
        cntw += base + l + NT * k < B.wp ? flg[v[k]] : 0u;
      }
      if (sums) {
#pragma unroll
        for (int k = 0; k < KB; ++k) {
          const double dd = (double)x[k];
          acc[2 * j] += dd;
          acc[2 * j + 1] += dd * dd;
        }
      }
      continue;
    }
#pragma unroll 4
    for (int k = 0; k < FLAG_BLOCK / NT; ++k) {
      const int t = base + l + NT * k;
      if (t >= W) break;
      const int v = bytes[j][l + NT * k];
      const float x = tab[v];
      cnt += flg[v];
      cntw += t < B.wp ? flg[v] : 0u;
      if (t < npart && part != FPART_NONE) {
        const double dd = (double)x;
        acc[2 * j] += dd;
        acc[2 * j + 1] += dd * dd;
      }
    }
  }
  if (cnt) atomicAdd(&B.flagcount[c], cnt);
  if (cntw) atomicAdd(&B.gsum[c], (unsigned long long)cntw);
  if (part == FPART_NONE) return;
  tree256_multi(acc, sh);
  if (l == 0) {
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      if (blk0 + j >= B.nblk) break;
      B.part[fslot(B, c, blk0 + j)] = acc[2 * j];
      if (part == FPART_SUM2) B.part2[fslot(B, c, blk0 + j)] = acc[2 * j + 1];
    }
  }
}

// The time-series division of the value-table path, y = x / ts[t] with x the
// channel's table value of the byte, formed as (float)((double)x * rts[t]),
// rts[t] = 1 / (double)ts[t] rounded to double.  The double product is within
// 2^-52 (relative) of the exact quotient, and the quotient of two floats is
// never within 2^-49 (relative) of a boundary between two float roundings (a
// midpoint of adjacent floats, or the overflow threshold): the quotient minus
// such a boundary m is (x - ts m) / ts, a nonzero multiple of the last place
// of ts m over ts.  So the product rounds to the correctly rounded quotient;
// zeros, infinities and NaNs follow IEEE division as well.
__device__ __forceinline__ float tsdiv(float x, double r) { return (float)((double)x * r); }

// The statistics pass after the time-series division (FOP_TSDIV): per
// channel and block the partial sums of y and y * y over the statistics
// columns in the canonical lane order (k_apply's), TQCH channels per
// workgroup sharing the block's reciprocals and one tree.
constexpr int TQCH = 4;

__global__ __launch_bounds__(NT) void k_tsdiv_stats(const FlagBeamDev *__restrict__ beams, int W,
                                                    int lin)
{
  const FlagBeamDev B = beams[blockIdx.z];
  const int c0 = blockIdx.y * TQCH;
  if (c0 >= B.nch) return;
  const int nc = min(TQCH, B.nch - c0);
  const int blk = blockIdx.x;
  const int l = threadIdx.x;
  constexpr int KB = FLAG_BLOCK / NT;
  __shared__ float tab[TQCH][256];
  __shared__ __align__(16) uint8_t bytes[TQCH][FLAG_BLOCK];
  __shared__ double sh[2 * TQCH * 128];
  const int base = blk * FLAG_BLOCK;
  // the rows hold xs >= W rounded up to 64 bytes: a 16-byte load that
  // starts below W stays in the row
  u32x4f d[TQCH];
#pragma unroll
  for (int j = 0; j < TQCH; ++j) {
    d[j] = u32x4f{0u, 0u, 0u, 0u};
    if (j < nc && base + 16 * l < W)
      d[j] = gld(reinterpret_cast<const u32x4f *>(B.X8 + (int64_t)(c0 + j) * B.xs + base + 16 * l));
  }
  double r[KB];
#pragma unroll
  for (int k = 0; k < KB; ++k) {
    const int t = base + l + NT * k;
    r[k] = gld(B.rts + min(t, W - 1));  // used only for t < wp <= W
  }
#pragma unroll
  for (int j = 0; j < TQCH; ++j) {
    // loaded for every j (the table exists for every channel and slot), kept
    // only where it is the value
    const float tv = gld(B.lut + ((int64_t)max(lin, 0) * B.nch + min(c0 + j, B.nch - 1)) * 256 + l);
    tab[j][l] = j >= nc ? 0.f : (lin < 0 ? (float)l : tv);
  }
#pragma unroll
  for (int j = 0; j < TQCH; ++j) *reinterpret_cast<u32x4f *>(&bytes[j][16 * l]) = d[j];
  __syncthreads();
  double acc[2 * TQCH];
#pragma unroll
  for (int j = 0; j < 2 * TQCH; ++j) acc[j] = 0.0;
#pragma unroll
  for (int j = 0; j < TQCH; ++j) {
    if (j >= nc) break;
    if (base + FLAG_BLOCK <= B.wp) {
      int v[KB];
#pragma unroll
      for (int k = 0; k < KB; ++k) v[k] = bytes[j][l + NT * k];
#pragma unroll
      for (int k = 0; k < KB; ++k) {
        const double dd = (double)tsdiv(tab[j][v[k]], r[k]);
        acc[2 * j] += dd;
        acc[2 * j + 1] += dd * dd;
      }
    } else {
#pragma unroll 4
      for (int k = 0; k < KB; ++k) {
        if (base + l + NT * k >= B.wp) break;
        const double dd = (double)tsdiv(tab[j][bytes[j][l + NT * k]], r[k]);
        acc[2 * j] += dd;
        acc[2 * j + 1] += dd * dd;
      }
    }
  }
  tree256_multi(acc, sh);
  if (l == 0) {
#pragma unroll
    for (int j = 0; j < TQCH; ++j) {
      if (j >= nc) break;
      B.part[fslot(B, c0 + j, blk)] = acc[2 * j];
      B.part2[fslot(B, c0 + j, blk)] = acc[2 * j + 1];
    }
  }
}

// The quantising pass of the value-table path (FOP_NORM_QUANT): the value of
// table slot lin divided by ts[t] (tsdiv), normalised and quantised into the
// ring.  It has no sums and no flags, so a thread takes 16 consecutive
// samples, with one 16-byte load of a channel's bytes and, when the ring
// columns are 16-aligned and do not wrap, one 16-byte store; QQCH channels
// per workgroup share the samples' reciprocals.
constexpr int QQCH = 4;

__global__ __launch_bounds__(NT) void k_quant_lut(const FlagBeamDev *__restrict__ beams, int W,
                                                  int lin)
{
  typedef double f64x2 __attribute__((ext_vector_type(2)));
  const FlagBeamDev B = beams[blockIdx.z];
  const int c0 = blockIdx.y * QQCH;
  if (c0 >= B.nch) return;
  const int nc = min(QQCH, B.nch - c0);
  const int l = threadIdx.x;
  __shared__ float tab[QQCH][256];
#pragma unroll
  for (int j = 0; j < QQCH; ++j) {
    // loaded for every j (the table exists for every channel and slot), kept
    // only where it is the value
    const float tv = gld(B.lut + ((int64_t)max(lin, 0) * B.nch + min(c0 + j, B.nch - 1)) * 256 + l);
    tab[j][l] = j >= nc ? 0.f : (lin < 0 ? (float)l : tv);
  }
  __syncthreads();
  const int t0 = (blockIdx.x * NT + l) * 16;
  if (t0 >= W) return;
  // the rows hold xs >= W rounded up to 64 bytes and rts holds W rounded up
  // to 64 doubles: 16-byte loads that start below W stay inside both
  u32x4f d[QQCH];
#pragma unroll
  for (int j = 0; j < QQCH; ++j) {
    d[j] = u32x4f{0u, 0u, 0u, 0u};
    if (j < nc) d[j] = gld(reinterpret_cast<const u32x4f *>(B.X8 + (int64_t)(c0 + j) * B.xs + t0));
  }
  double r[16];
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const f64x2 v = gld(reinterpret_cast<const f64x2 *>(B.rts + t0 + 2 * j));
    r[2 * j] = v.x;
    r[2 * j + 1] = v.y;
  }
  int64_t col = B.col0 + t0;
  if (col >= B.ring_len) col -= B.ring_len;
  const bool vec = t0 + 16 <= W && (col & 15) == 0 && col + 16 <= B.ring_len;
#pragma unroll
  for (int j = 0; j < QQCH; ++j) {
    if (j >= nc) break;
    const int c = c0 + j;
    const NormP np = norm_params(B, c);
    uint32_t q[4] = {0u, 0u, 0u, 0u};
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const uint32_t b = (d[j][i >> 2] >> (8 * (i & 3))) & 0xffu;
      const float x = normv(tsdiv(tab[j][b], r[i]), np);
      q[i >> 2] |= (uint32_t)quant(x) << (8 * (i & 3));
    }
    uint8_t *dst = B.ring + (int64_t)c * B.ring_stride;
    if (vec) {
      *reinterpret_cast<u32x4f *>(dst + col) = u32x4f{q[0], q[1], q[2], q[3]};
    } else {
      for (int i = 0; i < 16 && t0 + i < W; ++i) {
        int64_t ci = col + i;
        if (ci >= B.ring_len) ci -= B.ring_len;
        dst[ci] = (uint8_t)(q[i >> 2] >> (8 * (i & 3)));
      }
    }
  }
}

// a channel's table of the next pass on its own (the pass that reads it
// needs the table before its own statistics exist)
__global__ __launch_bounds__(NT) void k_table(const FlagBeamDev *__restrict__ beams, int op,
                                              float a0, float a1, int lin, int lout)
{
  const FlagBeamDev B = beams[blockIdx.z];
  const int c = blockIdx.y;
  if (c >= B.nch) return;
  const int l = threadIdx.x;
  const float x0 = lin < 0 ? (float)l : B.lut[((int64_t)lin * B.nch + c) * 256 + l];
  const NormP np = norm_params(B, c);
  const float bpv = B.bp[c];
  const bool bad = op == FOP_SPEC && ((bpv < a0) || (bpv > a1));
  const bool fixed = c >= B.flag1 && c < B.flag2;
  bool g;
  B.lut[((int64_t)lout * B.nch + c) * 256 + l] = value_op(op, x0, bpv, bad, fixed, np, a0, a1, g);
}

// k_ts_sum on the value-table path, in two steps.  k_ts_part: workgroup
// (time block, partition q) adds the values x / bp[c] (table slot lin,
// already divided) over the partition's channels in channel order for 1024
// samples, 4 consecutive ones per thread read with one 4-byte load per
// channel; the tables of 32 channels at a time sit in LDS.
// k_ts_combine: the 16 partition sums of a sample are combined pairwise
// exactly as k_ts_sum does; it also stores the reciprocals 1 / ts[t].
constexpr int TS_GROUP = 32;
constexpr int TS_SPT = 4;                // samples per thread
constexpr int TS_BLOCK = NT * TS_SPT;    // samples per workgroup

__global__ __launch_bounds__(NT) void k_ts_part(const FlagBeamDev *__restrict__ beams, int W,
                                                int lin)
{
  const FlagBeamDev B = beams[blockIdx.z];
  const int q = blockIdx.y;
  const int l = threadIdx.x;
  const int t0 = blockIdx.x * TS_BLOCK + TS_SPT * l;
  __shared__ float tab[TS_GROUP][256];
  const int per = (B.nch + TS_PARTS - 1) / TS_PARTS;
  const int c0 = min(B.nch, q * per), c1 = min(B.nch, c0 + per);
  const float *tabs = B.lut + (int64_t)lin * B.nch * 256;
  double s[TS_SPT];
#pragma unroll
  for (int j = 0; j < TS_SPT; ++j) s[j] = 0.0;
  // the row holds xs >= W rounded up to 64 bytes: a 4-byte load that starts
  // below W stays in the row
  const bool in = t0 < W;
  for (int g0 = c0; g0 < c1; g0 += TS_GROUP) {
    const int ng = min(TS_GROUP, c1 - g0);
    __syncthreads();
    for (int i = 0; i < ng; ++i) tab[i][l] = gld(tabs + (int64_t)(g0 + i) * 256 + l);
    __syncthreads();
    if (!in) continue;
    const uint8_t *row = B.X8 + (int64_t)g0 * B.xs + t0;
#pragma unroll 8
    for (int i = 0; i < ng; ++i) {
      const uint32_t w = gld(reinterpret_cast<const uint32_t *>(row + (int64_t)i * B.xs));
#pragma unroll
      for (int j = 0; j < TS_SPT; ++j) s[j] += (double)tab[i][(w >> (8 * j)) & 0xffu];
    }
  }
  double *out = B.tsp + (int64_t)q * B.xs + t0;
#pragma unroll
  for (int j = 0; j < TS_SPT; ++j)
    if (t0 + j < W) out[j] = s[j];
}

__global__ __launch_bounds__(NT) void k_ts_combine(const FlagBeamDev *__restrict__ beams, int W)
{
  const FlagBeamDev B = beams[blockIdx.z];
  const int t = blockIdx.x * NT + threadIdx.x;
  if (t >= W) return;
  double p[TS_PARTS];
#pragma unroll
  for (int i = 0; i < TS_PARTS; ++i) p[i] = gld(B.tsp + (int64_t)i * B.xs + t);
#pragma unroll
  for (int h = TS_PARTS / 2; h >= 1; h >>= 1)
#pragma unroll
    for (int i = 0; i < h; ++i) p[i] += p[i + h];
  const float tv = (float)(p[0] / (double)B.nch);
  B.ts[t] = tv;
  B.rts[t] = 1.0 / (double)tv;
}

// ---- byte-histogram path (value-table path with mean statistics) ----------
// Before the time-series correction every value of channel c is a function
// T_c(b) of the input byte b, so every channel sum the stages need is a sum
// over the 256 byte values of (count of b) x T_c(b): the bandpass over
// [0, n32), the statistics over [0, wp), the flag counts over [0, W).  One
// pass builds the histograms; the stages then work on 256-entry tables only,
// one workgroup per channel with thread v holding byte v, each sum being the
// canonical 256-lane tree of the products in double.

// per channel the histograms of [0, wp), [wp, n32), [n32, W), 16 bytes per
// thread and load.  PACK16: 16 LDS copies (copy = lane & 15) of u16 counters,
// two bins per word; a copy counts at most W / 16 samples, so the caller
// takes this form while W / 16 < 65536.  Copy k starts k words past a
// multiple of 32, so one bin in different copies falls in different banks;
// with 16 copies, 4 lanes of a wave share one (few colliding increments).
// Otherwise one u32 copy per wave.
template <bool PACK16>
__global__ __launch_bounds__(NT) void k_hist_u8(const FlagBeamDev *__restrict__ beams, int W)
{
  const FlagBeamDev B = beams[blockIdx.z];
  const int c = blockIdx.y;
  if (c >= B.nch) return;
  const int l = threadIdx.x, w = l >> 6;
  constexpr int NCOPY = PACK16 ? 16 : NT / 64;
  constexpr int CW = PACK16 ? 385 : 768;  // words per copy (PACK16: 384 + 1)
  __shared__ uint32_t h[NCOPY * CW];
  for (int i = l; i < NCOPY * CW; i += NT) h[i] = 0u;
  __syncthreads();
  const int n32 = (W / 32) * 32, wp = B.wp;
  const uint8_t *row = B.X8 + (int64_t)c * B.xs;
  uint32_t *mine = h + (PACK16 ? (l & 15) : w) * CW;
  // the row holds xs >= W rounded up to 64 bytes: a 16-byte load that starts
  // below W stays in the row
  for (int t0 = 16 * l; t0 < W; t0 += 16 * NT) {
    const u32x4f d = gld(reinterpret_cast<const u32x4f *>(row + t0));
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const int t = t0 + i;
      if (t >= W) break;
      const uint32_t b = (d[i >> 2] >> (8 * (i & 3))) & 0xffu;
      const uint32_t bin = (t < wp ? 0u : (t < n32 ? 256u : 512u)) + b;
      if (PACK16)
        atomicAdd(&mine[bin >> 1], 1u << (16 * (bin & 1u)));
      else
        atomicAdd(&mine[bin], 1u);
    }
  }
  __syncthreads();
  uint32_t *out = B.hist + (int64_t)c * 768;
  for (int i = l; i < 768; i += NT) {
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < NCOPY; ++k)
      s += PACK16 ? (h[k * CW + (i >> 1)] >> (16 * (i & 1))) & 0xffffu : h[k * CW + i];
    out[i] = s;
  }
}

// One step of the table chain for one channel (workgroup), thread v = byte v.
// STEP 0: bandpass of the input bytes, division by it, statistics;
// STEP 1: normalisation, bandpass, spectral flag (a0, a1: SPEC_MIN/MAX),
//         bandpass, division;
// STEP 2: normalisation and the 1 x 1 scrunch flag (a0, a1: +/- threshold),
//         flag counts and spectrum, bandpass, division.
// Steps 1 and 2 normalise with the statistics of the previous step's values,
// formed from its per-channel sums in slot sin of hsum; a step with sout >= 0
// stores the sums of its own values (S1, S2 of the statistics columns) in
// slot sout.  bp[c] ends as the last bandpass.  The values are those of the
// plane passes (k_apply_lut), the sums the histogram order.
template <int STEP>
__global__ __launch_bounds__(NT) void k_hstep(const FlagBeamDev *__restrict__ beams, int W, int lin,
                                              int lout, float a0, float a1, int sin, int sout)
{
  const FlagBeamDev B = beams[blockIdx.z];
  const int c = blockIdx.y;
  if (c >= B.nch) return;
  const int v = threadIdx.x;
  __shared__ double sh[2 * 128];
  __shared__ double bc[2];
  __shared__ unsigned long long cntsh[2];
  const uint32_t *H = B.hist + (int64_t)c * 768;
  const double hA = (double)gld(H + v);
  const double h32 = hA + (double)gld(H + 256 + v);
  const double hW = h32 + (double)gld(H + 512 + v);
  const float xl = gld(B.lut + ((int64_t)max(lin, 0) * B.nch + c) * 256 + v);
  float x = lin < 0 ? (float)v : xl;
  // bandpass of the current values (k_finish_bp's (float)(sum / W))
  auto bandpass = [&]() -> float {
    double s[1] = {h32 * (double)x};
    tree256_multi(s, sh);
    if (v == 0) bc[0] = s[0];
    __syncthreads();
    const double r = bc[0];
    __syncthreads();
    return (float)(r / (double)W);
  };
  auto divide = [&](float bp) { x = bp > 0.f ? x / bp : 0.f; };
  if (STEP == 0) {
    const float bp = bandpass();
    if (v == 0) B.bp[c] = bp;
    divide(bp);
  } else {
    const NormP np = hstep_norm(B, B.hsum + (int64_t)sin * B.nch * 2, sh);
    x = normv(x, np);
    if (STEP == 1) {
      const float b2 = bandpass();
      const bool bad = (b2 < a0) || (b2 > a1);
      x = bad ? 1.f : x + 1.f;
    } else {
      const bool g = (x > a0) || (x < a1);
      const bool fixed = c >= B.flag1 && c < B.flag2;
      if (g || fixed) x = 0.f;
      x = x + 1.f;
      if (v < 2) cntsh[v] = 0ull;
      __syncthreads();
      if (g) {
        atomicAdd(&cntsh[0], (unsigned long long)hW);
        atomicAdd(&cntsh[1], (unsigned long long)hA);
      }
      __syncthreads();
      if (v == 0) {
        B.flagcount[c] += (uint32_t)cntsh[0];
        // k_finish_bp's flag spectrum of a 1 x 1 scrunch
        B.fspec[c] = B.fspec[c] + (float)((double)cntsh[1] / (double)1 / (double)B.wp);
      }
    }
    const float bp = bandpass();
    if (v == 0) B.bp[c] = bp;
    divide(bp);
  }
  B.lut[((int64_t)lout * B.nch + c) * 256 + v] = x;
  if (sout >= 0) {
    const double d = (double)x;
    double s[2] = {hA * d, hA * (d * d)};
    tree256_multi(s, sh);
    if (v == 0) {
      double *hs = B.hsum + ((int64_t)sout * B.nch + c) * 2;
      hs[0] = s[0];
      hs[1] = s[1];
    }
  }
}

// ---- robust statistics (FLAG_STATS ROBUST) ------------------------------
// Key of a float whose unsigned order is the float order (-0 before +0, NaNs
// past the infinities), and back.
__device__ __forceinline__ uint32_t fkey(float f)
{
  const uint32_t b = __float_as_uint(f);
  return b ^ ((b >> 31) ? 0xffffffffu : 0x80000000u);
}
__device__ __forceinline__ float kval(uint32_t k)
{
  return __uint_as_float((k >> 31) ? (k ^ 0x80000000u) : ~k);
}

// The element of rank k (0-based, ascending key order) among the keys of
// value(t), t < n: a radix select, 8 bits a pass from the top, each pass
// counting in an LDS histogram the keys that share the digits chosen so far
// and keeping the bucket where rank k falls (an exclusive scan of the 256
// counts by the 256 threads).
template <class F>
__device__ uint32_t rselect(int n, uint32_t k, F value, uint32_t *hist, uint32_t *pick)
{
  const int l = threadIdx.x;
  uint32_t prefix = 0, mask = 0;
  for (int shift = 24; shift >= 0; shift -= 8) {
    hist[l] = 0u;
    __syncthreads();
    // RU values per thread requested before the first is keyed (one memory
    // latency per RU samples, not per sample)
    constexpr int RU = 8;
    for (int t0 = l; t0 < n; t0 += NT * RU) {
      float v[RU];
#pragma unroll
      for (int u = 0; u < RU; ++u) v[u] = value(min(t0 + NT * u, n - 1));
#pragma unroll
      for (int u = 0; u < RU; ++u) {
        if (t0 + NT * u >= n) break;
        const uint32_t key = fkey(v[u]);
        if ((key & mask) == prefix) atomicAdd(&hist[(key >> shift) & 0xffu], 1u);
      }
    }
    __syncthreads();
    // inclusive scan of the counts (Hillis-Steele over 256 entries)
    uint32_t h = hist[l], v = h;
    for (int off = 1; off < NT; off <<= 1) {
      __syncthreads();
      const uint32_t u = l >= off ? hist[l - off] : 0u;
      __syncthreads();
      v += u;
      hist[l] = v;
    }
    __syncthreads();
    const uint32_t before = v - h;  // keys of lower buckets
    if (h > 0 && k >= before && k < v) {
      pick[0] = (uint32_t)l;
      pick[1] = k - before;
    }
    __syncthreads();
    prefix |= pick[0] << shift;
    mask |= 0xffu << shift;
    k = pick[1];
    __syncthreads();
  }
  return prefix;
}

// Per channel (one workgroup) the median m of the statistics columns and
// the scale s = 1.4826 * median(|x - m|) (x - m in float), both lower medians
// (rank (wp - 1) / 2).  SRC 0: the FP32 plane; 1: the uint8 plane through
// table slot lin; 2: the same divided by ts[t] (the final normalisation of
// the value-table path).
template <int SRC>
__global__ __launch_bounds__(NT) void k_robust(const FlagBeamDev *__restrict__ beams, int lin)
{
  const FlagBeamDev B = beams[blockIdx.z];
  const int c = blockIdx.y;
  if (c >= B.nch) return;
  const int l = threadIdx.x;
  __shared__ uint32_t hist[NT];
  __shared__ uint32_t pick[2];
  __shared__ float tab[256];
  if (SRC >= 1) {
    const float tv = gld(B.lut + ((int64_t)max(lin, 0) * B.nch + c) * 256 + l);
    tab[l] = lin < 0 ? (float)l : tv;
  }
  __syncthreads();
  const float *row = B.X + (int64_t)c * B.xs;
  const uint8_t *row8 = B.X8 + (int64_t)c * B.xs;
  auto value = [&](int t) -> float {
    if (SRC == 0) return gld(row + t);
    const float x = tab[gld(row8 + t)];
    return SRC == 2 ? x / gld(B.ts + t) : x;
  };
  const int n = B.wp;
  const uint32_t k = (uint32_t)(n - 1) / 2u;
  const float m = kval(rselect(n, k, value, hist, pick));
  const float d = kval(rselect(n, k, [&](int t) { return fabsf(value(t) - m); }, hist, pick));
  if (l == 0) {
    B.rstat[2 * c] = m;
    B.rstat[2 * c + 1] = (float)(1.4826 * (double)d);
  }
}

__global__ void k_zero_u32(uint32_t *p, int n)
{
  const int i = blockIdx.x * NT + threadIdx.x;
  if (i < n) p[i] = 0u;
}

}  // namespace

// ------------------------------------------------------------------ host
void FlagRunner::run(const FlagBeamDev *d_beams, int nbeam, int max_nch, int W,
                     const FlagParams &P, hipStream_t s, bool value_tables)
{
  if (value_tables) {
    run_tables(d_beams, nbeam, max_nch, W, P, s);
    return;
  }
  const int nblk = (W + FLAG_BLOCK - 1) / FLAG_BLOCK;
  const int n32 = (W / 32) * 32;
  int wp = (W / 256) * 256;
  if (wp == 0) wp = W;
  const dim3 gplane((unsigned)nblk, (unsigned)max_nch, (unsigned)nbeam);
  const dim3 grow((unsigned)((max_nch + NT - 1) / NT), 1, (unsigned)nbeam);
  const dim3 gbeam((unsigned)nbeam);
  auto apply = [&](int op, int part, int npart, float a0 = 0.f, float a1 = 0.f) {
    hipLaunchKernelGGL(k_apply, gplane, dim3(NT), 0, s, d_beams, W, op, part, npart, a0, a1);
  };
  auto bp = [&](int box = 0) {
    hipLaunchKernelGGL(k_finish_bp, grow, dim3(NT), 0, s, d_beams, W, box);
  };
  auto stats = [&]() { hipLaunchKernelGGL(k_finish_stats, gbeam, dim3(FS_NT), 0, s, d_beams, wp); };
  // a division pass (by the bandpass, or also by the time series) and the
  // statistics of its result: the plane's mean and sigma from the pass's
  // partial sums, or with robust statistics each channel's median and MAD
  auto divide_stats = [&](int op) {
    if (P.robust) {
      apply(op, FPART_NONE, 0);
      hipLaunchKernelGGL(k_robust<0>, dim3(1u, (unsigned)max_nch, (unsigned)nbeam), dim3(NT), 0, s,
                         d_beams, -1);
    } else {
      apply(op, FPART_SUM2, wp);
      stats();
    }
  };

  hipLaunchKernelGGL(k_ingest, dim3((unsigned)((W + 63) / 64), (unsigned)((max_nch + 63) / 64),
                                    (unsigned)nbeam),
                     dim3(256), 0, s, d_beams, W);
  // bandpass_flag (flagger.cu:216-242): bandpass_correct, normalize_data,
  // channel means of the normalised data, spectral flag
  apply(FOP_NONE, FPART_SUM, n32);
  bp();
  divide_stats(FOP_DIVBP);
  apply(FOP_NORM, FPART_SUM, n32);
  bp();
  apply(FOP_SPEC, FPART_SUM, n32, P.spec_min, P.spec_max);
  bp();
  // configured scrunches (flagger.cu:365-384)
  for (int i = 0; i < P.nscrunch; ++i) {
    const FlagScrunch &S = P.scrunch[i];
    divide_stats(FOP_DIVBP);
    const float thr = (float)((double)S.thresh / sqrt(1. * S.tscrunch * S.fscrunch));
    const float nthr = (float)(-1. * (double)thr);
    if (S.tscrunch == 1 && S.fscrunch == 1) {
      apply(FOP_NORM_FLAG1, FPART_SUM, n32, thr, nthr);
    } else {
      apply(FOP_NORM, FPART_NONE, 0);
      const float k = (float)(1. / S.tscrunch / S.fscrunch);
      hipLaunchKernelGGL(k_box_thresh, gplane, dim3(NT), 0, s, d_beams, W, S.tscrunch, S.fscrunch,
                         k, thr, nthr);
      hipLaunchKernelGGL(k_grow_replace, gplane, dim3(NT), 0, s, d_beams, W, S.tscrunch,
                         S.fscrunch, n32);
    }
    bp(S.tscrunch * S.fscrunch);
  }
  // apply_scrunch(8,8,100,flag=0,ts=1) (flagger.cu:385-401): bandpass_correct,
  // ts_correct, normalize_data
  hipLaunchKernelGGL(k_ts_sum, dim3((unsigned)((W + 63) / 64), 1, (unsigned)nbeam), dim3(64 * TS_PARTS), 0,
                     s, d_beams, W);
  divide_stats(FOP_DIVBP_TS);
  // transpose_output_handler (transpose.cu:120-145): quantise into the ring
  apply(FOP_NORM_QUANT, FPART_NONE, 0);
  HAMD_CHECK(hipGetLastError());
}

// The pass sequence of FlagRunner::run on the value-table path: the same
// passes and statistics, each pass reading the uint8 plane through the
// channel tables instead of reading and rewriting the FP32 plane.
void FlagRunner::run_tables(const FlagBeamDev *d_beams, int nbeam, int max_nch, int W,
                            const FlagParams &P, hipStream_t s)
{
  const int nblk = (W + FLAG_BLOCK - 1) / FLAG_BLOCK;
  const int n32 = (W / 32) * 32;
  const int wp = W >= 256 ? (W / 256) * 256 : W;
  const dim3 gplane((unsigned)nblk, (unsigned)max_nch, (unsigned)nbeam);
  const dim3 grow((unsigned)((max_nch + NT - 1) / NT), 1, (unsigned)nbeam);
  int cur = -1;  // table slot holding the current values (-1: the input bytes)
  const dim3 gplane4((unsigned)((nblk + 3) / 4), (unsigned)max_nch, (unsigned)nbeam);
  const dim3 gplane2((unsigned)((nblk + 1) / 2), (unsigned)max_nch, (unsigned)nbeam);
  // a pass of the robust path through the current tables (op before the
  // time-series correction); NB blocks per workgroup: 4, or all of a short
  // gulp's blocks (each block's partial sums are its own, whatever the
  // grouping)
  auto apply = [&](int op, int part, int npart, float a0 = 0.f, float a1 = 0.f) {
    const int out = op == FOP_NONE ? cur : (cur == 0 ? 1 : 0);
    if (nblk >= 4)
      hipLaunchKernelGGL(k_apply_lut<4>, gplane4, dim3(NT), 0, s, d_beams, W, op, part, npart, a0,
                         a1, cur, out);
    else if (nblk >= 2)
      hipLaunchKernelGGL(k_apply_lut<2>, gplane2, dim3(NT), 0, s, d_beams, W, op, part, npart, a0,
                         a1, cur, out);
    else
      hipLaunchKernelGGL(k_apply_lut<1>, gplane, dim3(NT), 0, s, d_beams, W, op, part, npart, a0,
                         a1, cur, out);
    cur = out;
  };
  auto bp = [&](int box = 0) {
    hipLaunchKernelGGL(k_finish_bp, grow, dim3(NT), 0, s, d_beams, W, box);
  };
  const dim3 gchan(1u, (unsigned)max_nch, (unsigned)nbeam);
  // robust statistics after a bandpass division: the division as a table
  // and each channel's median and MAD through it
  auto divide_stats = [&]() {
    const int nxt = cur == 0 ? 1 : 0;
    hipLaunchKernelGGL(k_table, gchan, dim3(NT), 0, s, d_beams, (int)FOP_DIVBP, 0.f, 0.f, cur,
                       nxt);
    cur = nxt;
    hipLaunchKernelGGL(k_robust<1>, gchan, dim3(NT), 0, s, d_beams, cur);
  };

  hipLaunchKernelGGL(k_ingest_u8, dim3((unsigned)((W + 63) / 64), (unsigned)((max_nch + 63) / 64),
                                       (unsigned)nbeam),
                     dim3(256), 0, s, d_beams, W);
  if (!P.robust) {
    // mean statistics: the stages up to the time-series correction from the
    // byte histograms, on 256-entry tables (k_hstep), no plane pass
    if (W / 16 < 65536)
      hipLaunchKernelGGL(k_hist_u8<true>, gchan, dim3(NT), 0, s, d_beams, W);
    else
      hipLaunchKernelGGL(k_hist_u8<false>, gchan, dim3(NT), 0, s, d_beams, W);
    // each step stores its values' per-channel sums (S1, S2) in a slot of
    // hsum; the next step forms the plane's statistics from them
    hipLaunchKernelGGL(k_hstep<0>, gchan, dim3(NT), 0, s, d_beams, W, -1, 0, 0.f, 0.f, -1, 0);
    int slot = 1;
    hipLaunchKernelGGL(k_hstep<1>, gchan, dim3(NT), 0, s, d_beams, W, 0, 1, P.spec_min,
                       P.spec_max, 0, P.nscrunch > 0 ? slot : -1);
    cur = 1;
    for (int i = 0; i < P.nscrunch; ++i) {
      const FlagScrunch &S = P.scrunch[i];
      const float thr = (float)((double)S.thresh / sqrt(1. * S.tscrunch * S.fscrunch));
      const float nthr = (float)(-1. * (double)thr);
      const int nxt = cur == 0 ? 1 : 0;
      hipLaunchKernelGGL(k_hstep<2>, gchan, dim3(NT), 0, s, d_beams, W, cur, nxt, thr, nthr, slot,
                         i + 1 < P.nscrunch ? 1 - slot : -1);
      slot = 1 - slot;
      cur = nxt;
    }
  } else {
    apply(FOP_NONE, FPART_SUM, n32);
    bp();
    divide_stats();
    apply(FOP_NORM, FPART_SUM, n32);
    bp();
    apply(FOP_SPEC, FPART_SUM, n32, P.spec_min, P.spec_max);
    bp();
    for (int i = 0; i < P.nscrunch; ++i) {
      const FlagScrunch &S = P.scrunch[i];
      divide_stats();
      const float thr = (float)((double)S.thresh / sqrt(1. * S.tscrunch * S.fscrunch));
      const float nthr = (float)(-1. * (double)thr);
      apply(FOP_NORM_FLAG1, FPART_SUM, n32, thr, nthr);
      bp(1);
    }
    // bandpass_correct of the final scrunch as a table, then ts_correct
    const int nxt = cur == 0 ? 1 : 0;
    hipLaunchKernelGGL(k_table, gchan, dim3(NT), 0, s, d_beams, (int)FOP_DIVBP, 0.f, 0.f, cur,
                       nxt);
    cur = nxt;
  }
  hipLaunchKernelGGL(k_ts_part, dim3((unsigned)((W + TS_BLOCK - 1) / TS_BLOCK), TS_PARTS, (unsigned)nbeam),
                     dim3(NT), 0, s, d_beams, W, cur);
  hipLaunchKernelGGL(k_ts_combine, dim3((unsigned)((W + NT - 1) / NT), 1, (unsigned)nbeam),
                     dim3(NT), 0, s, d_beams, W);
  if (P.robust) {
    hipLaunchKernelGGL(k_robust<2>, gchan, dim3(NT), 0, s, d_beams, cur);
  } else {
    hipLaunchKernelGGL(k_tsdiv_stats, dim3((unsigned)nblk, (unsigned)((max_nch + TQCH - 1) / TQCH),
                                           (unsigned)nbeam),
                       dim3(NT), 0, s, d_beams, W, cur);
    hipLaunchKernelGGL(k_finish_stats, dim3((unsigned)nbeam), dim3(FS_NT), 0, s, d_beams, wp);
  }
  hipLaunchKernelGGL(k_quant_lut, dim3((unsigned)((W + 16 * NT - 1) / (16 * NT)),
                                       (unsigned)((max_nch + QQCH - 1) / QQCH), (unsigned)nbeam),
                     dim3(NT), 0, s, d_beams, W, cur);
  HAMD_CHECK(hipGetLastError());
}

void zero_u32(uint32_t *p, int n, hipStream_t s)
{
  if (n <= 0) return;
  hipLaunchKernelGGL(k_zero_u32, dim3((unsigned)((n + NT - 1) / NT)), dim3(NT), 0, s, p, n);
  HAMD_CHECK(hipGetLastError());
}

}  // namespace hamd
