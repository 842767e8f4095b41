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

// Resident state of one beam: plan, delays, circular window and the
// flagger's running per-channel state.  Everything a single pass through the
// chain needs beyond this lives in a WorkSet.
struct BeamState {
  int id = 0;
  int nch = 0;
  double f0 = 0, df = 0, dt = 0;
  float dm_lo = 0, dm_hi = 0, dm_tol = 1.25f;
  PlanHost plan;
  std::vector<int> killmask;
  std::vector<int32_t> delta_host, live_host;
  DevBuf d_delta, d_live, d_delta_t, d_ent;
  const int32_t *delta = nullptr, *live = nullptr;  // device tables (own or shared)
  const int32_t *delta_t = nullptr;                 // channel-major delays
  int dt_stride = 0;
  const uint32_t *ent = nullptr;                    // dd_entries for blocks of ent_nd trials
  int ent_stride = 0, ent_nd = 0;
  int nlive = 0, ndm = 0;
  int64_t md = 0;
  int64_t ntime = 0;        // ring columns (NTIME)
  int64_t ring_stride = 0;  // bytes between channel rows
  int ntime_dd = 0, nout = 0, rows = 0;
  int min_gulp = 0;
  int64_t dstride = 0;      // row stride of the dedispersed plane (u32)
  int64_t in_off = 0;       // byte offset of this beam in a gulp buffer
  int64_t in_bytes = 0;
  long long gidx = 0;       // gulps ingested by this beam
  DevBuf ring;
  DevBuf f_bp, f_scal, f_count;  // bandpass, flagger scalars, grown-mask counts
  DevBuf s_scal;                 // statistics of the beam's last search
  // per-beam tile table of the plain path
  DevBuf d_tiles;
  int ntiles = 0;
  int spread = 0;
  long long crossings = 0;
};

// Sizes a WorkSet must cover (maximum over the beams that use it).
struct WorkDims {
  int nch = 0, W = 0, ndm = 0, rows = 0, nout = 0, nw = 0, cand_cap = 0;
  int cl_rows = 0, cl_cols = 0, cl_cap = 0;  // clustering grid and capacity
  int64_t xs = 0, dstride = 0;
  bool mask = false, planes = false;
  bool tables = false;  // flagger value-table path: uint8 plane + tables instead of FP32
  void merge(const WorkDims &o);
};

// Scratch of one beam's pass through the chain: flagger plane and mask,
// dedispersed plane, width planes (hella mode), block partial sums, counts
// and the candidate list.
struct WorkSet {
  DevBuf X, mask, D, planes;
  DevBuf f_part, f_part2, f_ts;
  DevBuf s_part, s_part2, s_counts, s_offs, s_bmax, s_tiles, s_cands, s_ncand;
  DevBuf c_grid, c_work, c_key, c_out, c_n;
  void ensure(const WorkDims &d);
  static size_t bytes_for(const WorkDims &d);
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
  // work sets: [0] serves the plain path; the batched path gives beam b set b
  std::vector<hamd::WorkSet> ws;
  hamd::WorkDims wdims;
  int64_t xs = 0;
  // descriptor arrays: pinned host copies, device copies
  hamd::FlagBeamDev *h_fdesc = nullptr;
  hamd::DDBeamDev *h_ddesc = nullptr;
  hamd::SearchBeamDev *h_sdesc = nullptr;
  uint32_t *h_ncand = nullptr;
  // pinned staging of the plain path's candidates and first clusters: every
  // candidate the host keeps of a beam (at most min(MAX_GIANTS, CAND_CAP))
  // and the first clusters arrive in one round trip
  hamd::Cand *h_cstage = nullptr;
  hamd::ClusterRec *h_kstage = nullptr;
  size_t cstage_n = 0;
  hamd::DevBuf d_fdesc, d_ddesc, d_sdesc;
  hamd::DevBuf d_in;           // staged gulp
  uint8_t *h_in = nullptr;     // pinned staging for host gulps
  hamd::DevBuf d_kern;
  std::vector<int> kofs;
  hamd::SearchCommon common{};
  hamd::DDTile tile{16, 32, 1};
  int tile_order = 0;
  bool tile_forced = false;
  int cur_beam = -1;           // beam whose plane is in work set 0
  std::vector<hella_candidate> peaks;
  std::vector<hella_cluster> clusters;
  bool timing = false;
  std::array<float, 4> stage_ms{};
  hamd::Batch batch;
  bool batch_mode = false;     // descriptors currently describe the batched layout
};

namespace hamd {
// plan, delays, window sizing and resident buffers of one beam
void beam_setup(BeamState &B, const Config &cfg, int W, hipStream_t s);
WorkDims beam_dims(const BeamState &B, const Config &cfg, int W, int nw);
// device descriptors of one beam's pass: d_in is the beam's gulp on the device
FlagBeamDev make_fdesc(const Config &cfg, int W, BeamState &B, WorkSet &w, const void *d_in);
DDBeamDev make_ddesc(const Config &cfg, int W, BeamState &B, WorkSet &w);
SearchBeamDev make_sdesc(const Config &cfg, BeamState &B, WorkSet &w);
// absolute sample of output sample 0 of the beam's current window
long long window_abs0(const Config &cfg, int W, const BeamState &B);
bool beam_ready(const Config &cfg, const BeamState &B);
SearchCommon search_common(const Config &cfg, const BeamState &B0, int nw);
// append a beam's search results (first entries of its ordered list) to a
// candidate list holding at most cap entries
void append_candidates(const Config &cfg, int W, const BeamState &B, int beam_id, const Cand *cs,
                       size_t n, size_t cap, std::vector<hella_candidate> &out);
// clustering arguments (SearchRunner::cluster): prefix candidates already in
// the list; cumulative: the launch's beams share one list, in order
ClusterArgs cluster_args(const Config &cfg, int prefix, bool cumulative);
// append a beam's clusters (device records; rep0 = the beam's first index in
// the candidate list)
void append_clusters(const ClusterRec *cl, size_t n, int rep0, int beam_id,
                     std::vector<hella_cluster> &out);
// the first n records of a device list whose first `staged` are already in
// pinned staging: a pointer into the staging when they all are, else the
// staged ones plus the rest copied from the device into `spill` (synchronous)
template <class T>
const T *results_view(const T *staged_h, size_t staged, const T *dev, size_t n,
                      std::vector<T> &spill, hipStream_t s)
{
  if (n <= staged) return staged_h;
  spill.resize(n);
  std::copy(staged_h, staged_h + staged, spill.begin());
  HAMD_CHECK(hipMemcpyAsync(spill.data() + staged, dev + staged, (n - staged) * sizeof(T),
                            hipMemcpyDeviceToHost, s));
  HAMD_CHECK(hipStreamSynchronize(s));
  return spill.data();
}
// default dedispersion tile shape for a set of beams
void choose_default_tile(const std::vector<BeamState *> &beams, DDTile &tile, int &order);
void build_plain_tiles(BeamState &B, const DDTile &t, int order);
void ctx_build(hella_ctx c);
void ctx_flag_beam(hella_ctx c, int b, const void *d_gulp_base);
void ctx_dedisperse(hella_ctx c, int b);
void ctx_smooth(hella_ctx c, int b);
int ctx_find_peaks(hella_ctx c, int b);
void ctx_end_gulp(hella_ctx c);
void ctx_process_plain(hella_ctx c, const void *gulp, bool host_ptr);
void ctx_process_batched(hella_ctx c, const void *gulp, bool host_ptr);
void ctx_autotune(hella_ctx c, const char *path);
float ctx_bench_stage(hella_ctx c, int b, int which, int reps);
const void *ctx_stage_input(hella_ctx c, const void *gulp, bool host_ptr);
}  // namespace hamd
