// pipeline.hpp -- the hella pipeline context: per-beam plans, rings, work
// buffers and the plain / batched gulp paths.
#pragma once

#include <array>
#include <map>

#include "../../include/hella.h"
#include "flagger.hpp"
#include "internal.hpp"
#include "search.hpp"

namespace hamd {

// KEY VALUE configuration (casm:72-169 vocabulary plus runtime geometry)
struct Config {
  float dm_min = 0.f, dm_max = 1000.f, dm_tol = 1.25f, snr = 7.f;
  int width_min = 1, width_max = 32;
  long long gulp = 32768;
  int flag1 = -1, flag2 = -1;
  float spec_min = -0.05f, spec_max = 0.15f;
  int nbeam = 1, gpu = 0;
  FlagParams flag;
  int nchan = 256;
  double fch1 = 500.0, foff = -0.390625, tsamp = 256e-6;
  float pulse_width_us = 40.f;  // casm:207 passes 40 us to dedisp_generate_dm_list
  int input_fmt = FLAG_IN_U8_TF;
  bool oneshot = false;         // INPUT FILE: NTIME = gulp (casm:210-212)
  int search_mode = SEARCH_BOXCAR;
  int max_giants = 10000;       // definitions.h:49
  int cand_cap = 1 << 20;
  int cluster_dt = 64, cluster_ddm = 16;
  double time_limit = 100.1;    // definitions.h:61
  bool parse(const std::string &text, std::string &err, bool beam_only = false);
};

struct BeamState {
  int id = 0;
  int nch = 0;
  double f0 = 0, df = 0, dt = 0;
  float dm_lo = 0, dm_hi = 0, dm_tol = 1.25f;
  PlanHost plan;
  std::vector<int> killmask;
  std::vector<int32_t> delta_host, live_host;
  DevBuf d_delta, d_live;
  int nlive = 0, ndm = 0;
  int64_t md = 0;
  int64_t ntime = 0;        // ring columns (NTIME)
  int64_t ring_stride = 0;  // bytes between channel rows
  int ntime_dd = 0, nout = 0, rows = 0;
  int min_gulp = 0;
  int64_t in_off = 0;       // byte offset of this beam in a gulp buffer
  int64_t in_bytes = 0;
  DevBuf ring;
  // flagger state
  DevBuf f_part, f_part2, f_bp, f_scal, f_count, f_ts;
  // search state
  DevBuf s_scal, s_part, s_part2, s_counts, s_offs, s_tiles, s_cands, s_ncand;
  // work planes owned by the beam (batched path)
  DevBuf X, mask, D, planes;
  int64_t dstride = 0;
  // per-beam tile table for the plain path
  DevBuf d_tiles;
  int ntiles = 0;
  int spread = 0;
  long long crossings = 0;
};

struct Batch {
  bool valid = false;
  std::vector<int> beams;  // searchable beams, in order
  DevBuf d_tiles;
  int ntiles = 0;
  DDTile tile{16, 32, 1};
  int spread = 0;
  hipGraphExec_t exec = nullptr;
  bool graph_ok = false;
};

}  // namespace hamd

struct hella_ctx_s {
  hamd::Config cfg;
  std::vector<hamd::BeamState> beams;
  hipStream_t stream = nullptr;
  bool built = false;
  long long gulp_index = 0;
  int W = 0, nw = 0, max_nch = 0, max_rows = 0;
  int64_t gulp_bytes = 0;
  // shared work buffers (plain path)
  hamd::DevBuf X, mask, D, planes;
  int64_t xs = 0;
  // descriptor arrays: pinned host copies, device copies
  hamd::FlagBeamDev *h_fdesc = nullptr;
  hamd::DDBeamDev *h_ddesc = nullptr;
  hamd::SearchBeamDev *h_sdesc = nullptr;
  uint32_t *h_ncand = nullptr;
  hamd::DevBuf d_fdesc, d_ddesc, d_sdesc;
  hamd::DevBuf d_in;           // staged gulp
  uint8_t *h_in = nullptr;     // pinned staging for host gulps
  hamd::DevBuf d_kern;
  std::vector<int> kofs;
  hamd::SearchCommon common{};
  hamd::DDTile tile{16, 32, 1};
  int tile_order = 0;
  bool tile_forced = false;
  int cur_beam = -1;           // beam whose plane is in the shared D
  std::vector<hella_candidate> peaks;
  std::vector<hella_cluster> clusters;
  bool timing = false;
  std::array<float, 4> stage_ms{};
  hamd::Batch batch;
  bool batch_mode = false;     // descriptors currently describe the batched layout
};

namespace hamd {
void ctx_build(hella_ctx c);
void ctx_flag_beam(hella_ctx c, int b, const void *d_gulp_base);
void ctx_dedisperse(hella_ctx c, int b);
void ctx_smooth(hella_ctx c, int b);
int ctx_find_peaks(hella_ctx c, int b);
void ctx_end_gulp(hella_ctx c);
void ctx_process_plain(hella_ctx c, const void *gulp, bool host_ptr);
void ctx_process_batched(hella_ctx c, const void *gulp, bool host_ptr);
void ctx_cluster(hella_ctx c);
void ctx_autotune(hella_ctx c, const char *path);
float ctx_bench_stage(hella_ctx c, int b, int which, int reps);
const void *ctx_stage_input(hella_ctx c, const void *gulp, bool host_ptr);
}  // namespace hamd
