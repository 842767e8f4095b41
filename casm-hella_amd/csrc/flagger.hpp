// flagger.hpp -- device descriptors and host runner of the FP32 flagger.
#pragma once

#include "internal.hpp"

namespace hamd {

constexpr int FLAG_BLOCK = 4096;  // canonical reduction block (samples)

enum FlagInFmt { FLAG_IN_U8_TF = 0, FLAG_IN_F32_TF = 1, FLAG_IN_F16_FT = 2 };
enum FlagOp {
  FOP_NONE = 0,
  FOP_DIVBP = 1,       // x / bp[c] (0 when bp <= 0)
  FOP_NORM = 2,        // (x - mu) / sigma
  FOP_SPEC = 3,        // 1 if the channel mean is out of range, else x + 1
  FOP_NORM_QUANT = 4,  // normalise and quantise into the ring
  FOP_NORM_FLAG1 = 5,  // normalise, then the 1 x 1 scrunch: clip |x| > thr to 0, +1
  FOP_DIVBP_TS = 6,    // (x / bp[c]) / ts[t]
  FOP_TSDIV = 7        // x / ts[t] (value-table path: the table already holds x / bp[c])
};
enum FlagPart { FPART_NONE = 0, FPART_SUM = 1, FPART_SUM2 = 2 };

// per-beam scalars kept on the device
struct FScal {
  double mu;    // mean of the last statistics pass
  double inv;   // 1 / (double)sigf
  float sigf;   // (float)sqrt(variance)
  float pad;
};

// One beam of a gulp as seen by the flagger kernels.
struct FlagBeamDev {
  float *X;            // work plane [nch][xs]
  int64_t xs;          // row stride of X and mask (elements)
  int nch;
  int nblk;            // ceil(W / FLAG_BLOCK)
  const void *in;      // input gulp on the device
  int in_fmt;          // FlagInFmt
  int flag1, flag2;    // channel range zeroed by every scrunch
  int64_t in_stride;   // TF: elements between samples; FT: elements between channels
  uint8_t *ring;       // quantised output, channel-major
  int64_t ring_stride; // bytes between channel rows
  int64_t ring_len;    // columns (multiple of 16); column = (col0 + t) mod ring_len
  int64_t col0;
  double *part;        // [nblk][nch] block partial sums of x (channels innermost)
  double *part2;       // [nblk][nch] block partial sums of x*x
  float *bp;           // [nch] bandpass
  FScal *scal;
  uint32_t *flagcount; // [nch] grown-mask counts (accumulated)
  // the reference's flagSpec (flagger.cu:116-124, 333-334): per channel, the
  // mean over the statistics columns of each scrunch's grown mask G (the box
  // mean of the threshold mask), accumulated in float over scrunches and
  // gulps.  gsum collects a scrunch's integer sum of masked pixels over the
  // boxes of the statistics columns; k_finish_bp folds it into fspec.
  float *fspec;              // [nch]
  unsigned long long *gsum;  // [nch]
  int wp;                    // statistics columns: 256 * floor(W / 256) (W if W < 256)
  // robust statistics (FLAG_STATS ROBUST): every normalisation uses the
  // channel's own median m and scale s = 1.4826 * MAD over the statistics
  // columns, rstat[2c] = m, rstat[2c + 1] = s, instead of the plane's mean
  // and standard deviation
  int robust;
  float *rstat;        // [nch][2]
  float *ts;           // [W] time series of the final pass
  double *rts;         // [W] 1 / (double)ts[t] (value-table path)
  // value-table path, mean statistics: per channel the sums (S1, S2) of a
  // table step's values over the statistics columns, two slots
  double *hsum;        // [2][nch][2]
  uint8_t *mask;       // [nch][xs] threshold mask scratch
  // value-table path (uint8 input, 1 x 1 scrunches only): the input as a
  // uint8 channel-major plane, and two [nch][256] tables holding each
  // channel's current value of every input byte
  uint8_t *X8;         // [nch][xs]
  float *lut;          // [2][nch][256]
  double *tsp;         // [16][xs] partition sums of the time series
  // per channel the input-byte histograms of columns [0, wp), [wp, n32) and
  // [n32, W) (n32 = 32 * floor(W / 32)): mean statistics on the value-table
  // path take every channel sum before the time-series correction from them
  uint32_t *hist;      // [nch][3][256]
};

struct FlagScrunch {
  int tscrunch, fscrunch;
  float thresh;
};

struct FlagParams {
  float spec_min = -0.05f, spec_max = 0.15f;
  int nscrunch = 0;
  FlagScrunch scrunch[8];
  bool robust = false;  // per-channel median / MAD normalisation (opt-in)
};

struct FlagRunner {
  // Runs the whole flagger for nbeam beams described by d_beams (device
  // array); max_nch is the largest channel count among them.  value_tables:
  // take the value-table path (every beam's input is uint8 and every
  // scrunch is 1 x 1; see flag_value_tables).
  static void run(const FlagBeamDev *d_beams, int nbeam, int max_nch, int W, const FlagParams &P,
                  hipStream_t s, bool value_tables = false);

 private:
  static void run_tables(const FlagBeamDev *d_beams, int nbeam, int max_nch, int W,
                         const FlagParams &P, hipStream_t s);
};

// Up to the time-series correction every flagger operation on uint8 input
// with 1 x 1 scrunches maps a channel's sample value through a function of
// (channel, input byte) alone, so the plane never has to be materialised:
// each pass reads the uint8 plane and a 256-entry table per channel.
inline bool flag_value_tables(int in_fmt, const FlagParams &P)
{
  if (in_fmt != FLAG_IN_U8_TF) return false;
  for (int i = 0; i < P.nscrunch; ++i)
    if (P.scrunch[i].tscrunch != 1 || P.scrunch[i].fscrunch != 1) return false;
  return true;
}
// byte offset of the tables behind the uint8 plane in a work buffer
inline size_t flag_table_offset(int nch, int64_t xs) { return ((size_t)nch * xs + 255) / 256 * 256; }
// bytes of the value-table path's work buffer: plane, two tables, the 16
// partition sums of the time series and the byte histograms
inline size_t flag_hist_offset(int nch, int64_t xs)
{
  return flag_table_offset(nch, xs) + (size_t)2 * nch * 256 * 4 + (size_t)16 * xs * 8;
}
inline size_t flag_table_bytes(int nch, int64_t xs)
{
  return flag_hist_offset(nch, xs) + (size_t)nch * 768 * 4;
}
// a box scrunch (ts x fs > 1) needs the mask plane
inline bool flag_needs_mask(const FlagParams &P)
{
  for (int i = 0; i < P.nscrunch; ++i)
    if (P.scrunch[i].tscrunch != 1 || P.scrunch[i].fscrunch != 1) return true;
  return false;
}

// A beam's resident flagger state in one buffer: grown-mask counts (u32),
// the flag spectrum (f32) and the scrunch's masked-pixel sums (u64), nch
// each, all zero at the start.
inline size_t flag_state_spec_offset(int nch) { return (size_t)nch * 4; }
inline size_t flag_state_gsum_offset(int nch) { return ((size_t)nch * 8 + 7) / 8 * 8; }
inline size_t flag_state_bytes(int nch) { return flag_state_gsum_offset(nch) + (size_t)nch * 8; }

void zero_u32(uint32_t *p, int n, hipStream_t s);

}  // namespace hamd
