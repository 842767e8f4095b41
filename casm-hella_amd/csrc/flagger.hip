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

namespace hamd {

namespace {

constexpr int NT = 256;

__device__ __forceinline__ double lane_tree(double v, double *sh)
{
  const int l = threadIdx.x;
  sh[l] = v;
  __syncthreads();
#pragma unroll
  for (int s = NT / 2; s >= 1; s >>= 1) {
    if (l < s) sh[l] += sh[l + s];
    __syncthreads();
  }
  double r = sh[0];
  __syncthreads();
  return r;
}

__device__ __forceinline__ uint8_t quant(float x)
{
  // transpose_output (src/hella/src/transpose.cu:428-438); NaN -> 0
  float v = x + 4.f;
  v = v * (255.f / 14.f);
  if (v < 0.f) v = 0.f;
  if (v > 255.f) v = 255.f;
  if (v != v) v = 0.f;
  return (uint8_t)v;
}

// ---- ingest: input gulp -> X (transpose_input, src/hella/src/transpose.cu:23-42)
__global__ __launch_bounds__(256) void k_ingest(const FlagBeamDev *__restrict__ beams, int W)
{
  const FlagBeamDev &B = beams[blockIdx.z];
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
__global__ __launch_bounds__(NT) void k_apply(const FlagBeamDev *__restrict__ beams, int W,
                                              int op, int part, int npart, float smin,
                                              float smax)
{
  const FlagBeamDev &B = beams[blockIdx.z];
  const int c = blockIdx.y;
  if (c >= B.nch) return;
  const int blk = blockIdx.x;
  const int l = threadIdx.x;
  float *row = B.X + (int64_t)c * B.xs;
  const int base = blk * FLAG_BLOCK;
  __shared__ double sh[NT];
  const FScal sc = *B.scal;
  float bpv = 0.f;
  bool bad = false;
  if (op == FOP_DIVBP) bpv = B.bp[c];
  if (op == FOP_SPEC) {
    const float b2 = B.bp[c];
    bad = (b2 < smin) || (b2 > smax);
  }
  double acc = 0.0;
  const double mu = (part == FPART_SQDEV) ? sc.mu : 0.0;
#pragma unroll 4
  for (int k = 0; k < FLAG_BLOCK / NT; ++k) {
    const int t = base + l + NT * k;
    if (t >= W) break;
    float x = row[t];
    switch (op) {
      case FOP_DIVBP: x = bpv > 0.f ? x / bpv : 0.f; row[t] = x; break;
      case FOP_NORM: x = (float)(((double)x - sc.mu) * sc.inv); row[t] = x; break;
      case FOP_SPEC: x = bad ? 1.f : x + 1.f; row[t] = x; break;
      case FOP_NORM_QUANT: {
        x = (float)(((double)x - sc.mu) * sc.inv);
        int64_t col = B.col0 + t;
        if (col >= B.ring_len) col -= B.ring_len;
        B.ring[(int64_t)c * B.ring_stride + col] = quant(x);
        break;
      }
      default: break;
    }
    if (t < npart) {
      if (part == FPART_SUM) acc += (double)x;
      else if (part == FPART_SQDEV) {
        const double d = (double)x - mu;
        acc += d * d;
      }
    }
  }
  if (part == FPART_NONE) return;
  const double s = lane_tree(acc, sh);
  if (l == 0) B.part[(int64_t)c * B.nblk + blk] = s;
}

// bandpass from partials: bp[c] = (float)(sum_b part[c][b] / W)
__global__ __launch_bounds__(NT) void k_finish_bp(const FlagBeamDev *__restrict__ beams, int W)
{
  const FlagBeamDev &B = beams[blockIdx.z];
  const int c = blockIdx.x * NT + threadIdx.x;
  if (c >= B.nch) return;
  double s = 0.0;
  for (int b = 0; b < B.nblk; ++b) s += B.part[(int64_t)c * B.nblk + b];
  B.bp[c] = (float)(s / (double)W);
}

// global statistic over rows: value(r) = (sum_b part[r][b]) / n, combined by
// the canonical blocked rule and divided by the row count.
// which == 0: mean -> scal.mu; which == 1: variance -> scal.sigf, scal.inv
__global__ __launch_bounds__(NT) void k_finish_stats(const FlagBeamDev *__restrict__ beams,
                                                     int n, int which)
{
  const FlagBeamDev &B = beams[blockIdx.x];
  __shared__ double sh[NT];
  const int l = threadIdx.x;
  double total = 0.0;
  for (int r0 = 0; r0 < B.nch; r0 += FLAG_BLOCK) {
    double acc = 0.0;
    for (int k = 0; k < FLAG_BLOCK / NT; ++k) {
      const int r = r0 + l + NT * k;
      if (r >= B.nch) break;
      double s = 0.0;
      for (int b = 0; b < B.nblk; ++b) s += B.part[(int64_t)r * B.nblk + b];
      acc += s / (double)n;
    }
    total += lane_tree(acc, sh);
  }
  if (l == 0) {
    const double v = total / (double)B.nch;
    if (which == 0) {
      B.scal->mu = v;
    } else {
      const float sf = (float)sqrt(v);
      B.scal->sigf = sf;
      B.scal->inv = 1.0 / (double)sf;
    }
  }
}

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// S = box mean over (ts x fs) with anchor (ts/2, fs/2), replicate border,
// taps row-major into a float from 0 (npp_convolve_handler,
// src/hella/src/convolution.cu:20-55); mask = |S| > thr.
__global__ __launch_bounds__(NT) void k_box_thresh(const FlagBeamDev *__restrict__ beams, int W,
                                                   int ts, int fs, float k, float thr,
                                                   float nthr)
{
  const FlagBeamDev &B = beams[blockIdx.z];
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
  const FlagBeamDev &B = beams[blockIdx.z];
  const int c = blockIdx.y;
  if (c >= B.nch) return;
  const int ax = ts / 2, ay = fs / 2;
  const int base = blockIdx.x * FLAG_BLOCK;
  const int l = threadIdx.x;
  const bool fixed = c >= B.flag1 && c < B.flag2;
  __shared__ double sh[NT];
  double acc = 0.0;
  uint32_t cnt = 0;
  float *row = B.X + (int64_t)c * B.xs;
  for (int kk = 0; kk < FLAG_BLOCK / NT; ++kk) {
    const int t = base + l + NT * kk;
    if (t >= W) break;
    bool g = false;
    for (int i = 0; i < fs && !g; ++i) {
      const uint8_t *mrow = B.mask + (int64_t)clampi(c + i - ay, 0, B.nch - 1) * B.xs;
      for (int j = 0; j < ts; ++j)
        if (mrow[clampi(t + j - ax, 0, W - 1)]) { g = true; break; }
    }
    cnt += g ? 1u : 0u;
    float x = row[t];
    if (g || fixed) x = 0.f;
    x = x + 1.f;
    row[t] = x;
    if (t < npart) acc += (double)x;
  }
  if (cnt) atomicAdd(&B.flagcount[c], cnt);
  const double s = lane_tree(acc, sh);
  if (l == 0) B.part[(int64_t)c * B.nblk + blockIdx.x] = s;
}

// ts_correct (src/hella/src/timeseries.cu:19-74): ts[t] = sum_c x / nch in
// channel order (double), then every channel divided by ts[t]
__global__ __launch_bounds__(NT) void k_ts_correct(const FlagBeamDev *__restrict__ beams, int W)
{
  const FlagBeamDev &B = beams[blockIdx.z];
  const int t = blockIdx.x * NT + threadIdx.x;
  if (t >= W) return;
  double s = 0.0;
  for (int c = 0; c < B.nch; ++c) s += (double)B.X[(int64_t)c * B.xs + t];
  const float tv = (float)(s / (double)B.nch);
  B.ts[t] = tv;
  for (int c = 0; c < B.nch; ++c) {
    float *p = B.X + (int64_t)c * B.xs + t;
    *p = *p / tv;
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
                     const FlagParams &P, hipStream_t s)
{
  const int nblk = (W + FLAG_BLOCK - 1) / FLAG_BLOCK;
  const int n32 = (W / 32) * 32;
  int wp = (W / 256) * 256;
  if (wp == 0) wp = W;
  const dim3 gplane((unsigned)nblk, (unsigned)max_nch, (unsigned)nbeam);
  const dim3 grow((unsigned)((max_nch + NT - 1) / NT), 1, (unsigned)nbeam);
  const dim3 gbeam((unsigned)nbeam);
  auto apply = [&](int op, int part, int npart) {
    hipLaunchKernelGGL(k_apply, gplane, dim3(NT), 0, s, d_beams, W, op, part, npart, P.spec_min,
                       P.spec_max);
  };
  auto bp = [&]() { hipLaunchKernelGGL(k_finish_bp, grow, dim3(NT), 0, s, d_beams, W); };
  auto fin = [&](int which) {
    hipLaunchKernelGGL(k_finish_stats, gbeam, dim3(NT), 0, s, d_beams, wp, which);
  };
  // normalise (normalize_data, flagger.cu:198-214) on a plane whose row
  // partials of sum(x) over wp columns are already in place
  auto stats_after_sum = [&]() {
    fin(0);
    apply(FOP_NONE, FPART_SQDEV, wp);
    fin(1);
  };

  hipLaunchKernelGGL(k_ingest, dim3((unsigned)((W + 63) / 64), (unsigned)((max_nch + 63) / 64),
                                    (unsigned)nbeam),
                     dim3(256), 0, s, d_beams, W);
  // bandpass_flag (flagger.cu:216-242)
  apply(FOP_NONE, FPART_SUM, n32);
  bp();
  apply(FOP_DIVBP, FPART_SUM, wp);
  stats_after_sum();
  apply(FOP_NORM, FPART_SUM, n32);
  bp();
  apply(FOP_SPEC, FPART_SUM, n32);
  bp();
  // configured scrunches (flagger.cu:365-384)
  for (int i = 0; i < P.nscrunch; ++i) {
    const FlagScrunch &S = P.scrunch[i];
    apply(FOP_DIVBP, FPART_SUM, wp);
    stats_after_sum();
    apply(FOP_NORM, FPART_NONE, 0);
    const float k = (float)(1. / S.tscrunch / S.fscrunch);
    const float thr = (float)((double)S.thresh / sqrt(1. * S.tscrunch * S.fscrunch));
    const float nthr = (float)(-1. * (double)thr);
    hipLaunchKernelGGL(k_box_thresh, gplane, dim3(NT), 0, s, d_beams, W, S.tscrunch, S.fscrunch,
                       k, thr, nthr);
    hipLaunchKernelGGL(k_grow_replace, gplane, dim3(NT), 0, s, d_beams, W, S.tscrunch,
                       S.fscrunch, n32);
    bp();
  }
  // apply_scrunch(8,8,100,flag=0,ts=1) (flagger.cu:385-401)
  apply(FOP_DIVBP, FPART_NONE, 0);
  hipLaunchKernelGGL(k_ts_correct, dim3((unsigned)((W + NT - 1) / NT), 1, (unsigned)nbeam),
                     dim3(NT), 0, s, d_beams, W);
  apply(FOP_NONE, FPART_SUM, wp);
  stats_after_sum();
  // transpose_output_handler (transpose.cu:120-145): quantise into the ring
  apply(FOP_NORM_QUANT, FPART_NONE, 0);
  HAMD_CHECK(hipGetLastError());
}

void zero_u32(uint32_t *p, int n, hipStream_t s)
{
  if (n <= 0) return;
  hipLaunchKernelGGL(k_zero_u32, dim3((unsigned)((n + NT - 1) / NT)), dim3(NT), 0, s, p, n);
  HAMD_CHECK(hipGetLastError());
}

}  // namespace hamd
