// internal.hpp -- shared declarations of libhella_amd (gfx950 only).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace hamd {

// ---------------------------------------------------------------- errors
struct HipError {
  hipError_t err;
  const char *what;
  int line;
};

#define HAMD_CHECK(expr)                                   \
  do {                                                     \
    hipError_t e_ = (expr);                                \
    if (e_ != hipSuccess) throw ::hamd::HipError{e_, #expr, __LINE__}; \
  } while (0)

void set_last_error(const std::string &msg);
const char *last_error();

inline size_t round_up(size_t v, size_t m) { return (v + m - 1) / m * m; }

// Device buffer owned by the library (hipMalloc/hipFree).
struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  DevBuf(DevBuf &&o) noexcept : p(o.p), bytes(o.bytes) { o.p = nullptr; o.bytes = 0; }
  DevBuf &operator=(DevBuf &&o) noexcept {
    release();
    p = o.p; bytes = o.bytes; o.p = nullptr; o.bytes = 0;
    return *this;
  }
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  // grows (never shrinks); contents are not preserved
  void ensure(size_t b) {
    if (b <= bytes) return;
    release();
    if (b == 0) return;
    HAMD_CHECK(hipMalloc(&p, b));
    bytes = b;
  }
  template <class T> T *as() const { return static_cast<T *>(p); }
};

// ---------------------------------------------------------------- plan
// Host description of a dedispersion plan (dedisp semantics).
struct PlanHost {
  size_t nchans = 0;
  float dt = 0.f, f0 = 0.f, df = 0.f;
  std::vector<float> delay_table;   // per channel, float
  std::vector<float> dm_list;       // ascending
  std::vector<int> killmask;        // 1 = live
  size_t max_delay = 0;
  size_t gulp_size = 65536;

  void build_delay_table();
  void generate_dm_list(float dm_start, float dm_end, double ti_us, double tol);
  void set_dm_list(const float *dms, size_t n);
  void compute_max_delay();
  // delta[d*nlive + i] for live channels only; live[i] = channel index
  void delays_live(size_t d_first, size_t nd, std::vector<int32_t> &delta,
                   std::vector<int32_t> &live) const;
  int32_t delay_of(size_t d, size_t c) const {
    float p = dm_list[d] * delay_table[c];
    return (int32_t)((double)p + 0.5);
  }
};

// ---------------------------------------------------------------- dedisp kernel
// One beam as seen by the dedispersion kernel.  Source rows: channel c of the
// FT u8 store begins at base + c*row_stride; window column x is stored at
// column (start + x) mod ring (ring is a multiple of 16).
struct DDBeamDev {
  const uint8_t *base;
  int64_t row_stride;
  int64_t ring;
  int64_t start;
  const int32_t *delta;  // [ndm][nlive] delays of live channels
  const int32_t *live;   // [nlive] channel index of each live channel
  const int32_t *delta_t;  // channel-major delays: delta_t[c * dt_stride + d] = delta[d][c]
  int dt_stride;           // (from this beam's first trial; >= ndm + 64, padded)
  // one-trial-per-wave kernel: per (trial, live channel) LDS entry of the
  // trial in its block's staged window (dd_entries), for blocks of ent_nd
  // trials; null when absent (the kernel then derives the entries itself)
  const uint32_t *ent;     // [ndm][ent_stride][2]
  int ent_stride;
  int ent_nd;
  int nlive;
  int ndm;
  int nout;              // output samples per DM row
  int out_f32;           // 1: float output, 0: u32
  void *out;
  int64_t out_stride;    // elements between DM rows
};

// Per (trial d, live channel c) entry of the one-trial-per-wave kernel for
// DM blocks of nd trials (block start d0 = d - d % nd), valid when the ring
// column of every tile start is a multiple of 16: q = (bas & 15) + own - bas
// with bas = delay(d0, c), own = delay(d, c) is the trial's byte offset in the
// channel's staged window; the entry holds q rounded down to a dword and the
// v_perm selector of the even samples at q's byte alignment.  Rows are padded
// to stride >= nlive + 16 with entries of offset 0 (the zero rows staged for
// channels past the last live one).
void dd_entries(const std::vector<int32_t> &delta, int ndm, int nlive, int nd,
                std::vector<uint32_t> &out, int &stride);

// Channel-major copy of a [ndm][nlive] delay table: [nlive][stride] with
// stride = ndm + 64, the pad repeating each channel's last delay, so any
// run of up to 64 consecutive trials can be read.
void delays_transposed(const std::vector<int32_t> &delta, int ndm, int nlive,
                       std::vector<int32_t> &out, int &stride);

// One workgroup of work: beam, first DM trial, first output sample.  beam < 0
// marks a padding entry.
struct DDTileDesc {
  int beam, d0, t0, pad;
};

// The DM-major delay table [ndm][nlive] of a beam is followed by at least
// this many readable int32 (vector scalar loads may run past its end).
constexpr int DD_DELTA_PAD = 16;

// Tile shape of the dedispersion kernel.
struct DDTile {
  int lanes_per_dm;  // G: threads sharing one DM trial (time groups); ND = 256 / G
  int chunk;         // CC: channels staged in LDS per step (divides 256)
  int dma = 1;       // 0: loads through registers; 1: double-buffered global->LDS DMA;
                     // 2: next chunk prefetched in registers (chunk 8 or 16);
                     // 3: one DM trial per wave (lanes_per_dm 64): chunk = 4 or 8 waves,
                     //    i.e. trials per workgroup, scalar delay arithmetic;
                     // 4: as 3, every LDS read of a chunk issued before its sums;
                     // 5: as 3 (chunk 4), windows copied by LDS DMA two chunks ahead
};

int dd_samples_per_thread();
// LDS bytes a dedispersion workgroup may use (the per-workgroup launch limit)
constexpr size_t DD_LDS_BUDGET = 64 * 1024;
inline int dd_tile_T(const DDTile &t) { return t.lanes_per_dm * dd_samples_per_thread(); }
inline int dd_tile_ND(const DDTile &t) { return t.dma >= 3 ? t.chunk : 256 / t.lanes_per_dm; }
// largest per-channel spread of delays inside any block of nd consecutive DMs
int dd_max_spread(const int32_t *delta, int ndm, int nlive, int nd);
size_t dd_lds_bytes(const DDTile &t, int max_spread);
// LDS a workgroup of this tile kind may use
size_t dd_lds_budget(const DDTile &t);
// the tile shape can run on windows of this spread and rings of this length
bool dd_tile_fits(const DDTile &t, int max_spread, int64_t min_ring);
// Tile table over beams: entries ordered so that the workgroups that the
// dispatcher sends to one XCD cover consecutive time tiles of one DM block.
// order 0: DM-block-major (consecutive time tiles of one DM block per XCD run);
// order 1: time-major (all DM blocks of a time tile per XCD run)
void dd_build_tiles(const std::vector<DDBeamDev> &beams, const DDTile &t,
                    std::vector<DDTileDesc> &tiles, int order = 0);
// Launch over a device tile table.
// min_ring: smallest ring length among the beams of the launch
void dd_launch(const DDBeamDev *d_beams, const DDTileDesc *d_tiles, int ntiles, const DDTile &t,
               int max_spread, hipStream_t s, int64_t min_ring = INT64_MAX);

// u8 time-major [nsamps][in_stride] -> FT rows (row_stride), zero tail to ring
void transpose_tf_to_ft(const uint8_t *in, int64_t nsamps, int64_t in_stride,
                        int nch, uint8_t *out, int64_t row_stride, int64_t ring,
                        hipStream_t s);

// zero rows [rows][n] of 32-bit words with a kernel (graph-capturable)
// Per-device kernel set-up (raised LDS limits of the kernels that take more
// than 64 KiB).  Called once per device by every entry point that creates a
// context, server or plan launch, before any stream capture.
void kernels_setup();
void dd_kernel_setup();
void cluster_kernel_setup();
void fill_zero_rows(void *out, int64_t stride, int n, int rows, hipStream_t s);

DDTile choose_dd_tile(const int32_t *delta, int ndm, int nlive, int *spread_out,
                      int lanes_pref, size_t lds_budget);

}  // namespace hamd
