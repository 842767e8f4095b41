// pipeline.cpp -- hella pipeline context: configuration, per-beam plans and
// rings, the plain per-beam gulp path and the batched gulp path.
//
// Control flow restated from src/apps/casm_hella_refactored.cu:
//   sizing            casm:182-224 (NTIME, ntime_dd, ntime_out; FILE input: one-shot)
//   per-gulp loop     casm:659-835 (flag all beams; once the window is full,
//                     per beam dedisperse -> smooth -> find_peaks while the
//                     candidate cap and the time budget allow)
//   rewind            casm:771-797 is replaced by a circular window: absolute
//                     sample a of a beam lives at ring column a mod NTIME, so
//                     no data moves between gulps.
#include "pipeline.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <sstream>

namespace hamd {

// ------------------------------------------------------------------ config
static bool parse_line_kv(const std::string &line, std::string &k, std::string &v)
{
  std::istringstream is(line);
  if (!(is >> k)) return false;
  if (!(is >> v)) v.clear();
  return true;
}

bool Config::parse(const std::string &text, std::string &err, bool beam_only)
{
  std::istringstream in(text);
  std::string line;
  while (std::getline(in, line)) {
    std::string k, v;
    if (!parse_line_kv(line, k, v) || k[0] == '#') continue;
    auto F = [&]() { return (float)std::atof(v.c_str()); };
    auto D = [&]() { return std::atof(v.c_str()); };
    auto I = [&]() { return std::atoi(v.c_str()); };
    if (k == "DM_MIN") dm_min = F();
    else if (k == "DM_MAX") dm_max = F();
    else if (k == "DM_TOL") dm_tol = F();
    else if (k == "NCHAN") nchan = I();
    else if (k == "FCH1") fch1 = D();
    else if (k == "FOFF") foff = D();
    else if (k == "TSAMP") tsamp = D();
    else if (beam_only) continue;
    else if (k == "SNR") snr = F();
    else if (k == "WIDTH_MIN") width_min = I();
    else if (k == "WIDTH_MAX") width_max = I();
    else if (k == "GULP") gulp = std::atoll(v.c_str());
    else if (k == "FLAG1") flag1 = I();
    else if (k == "FLAG2") flag2 = I();
    else if (k == "SPEC_MIN") spec_min = F();
    else if (k == "SPEC_MAX") spec_max = F();
    else if (k == "NBEAM") nbeam = I();
    else if (k == "GPU") gpu = I();
    else if (k == "PULSE_WIDTH") pulse_width_us = F();
    else if (k == "MAX_GIANTS") max_giants = I();
    else if (k == "CAND_CAP") cand_cap = I();
    else if (k == "CLUSTER_DT") cluster_dt = I();
    else if (k == "CLUSTER_DDM") cluster_ddm = I();
    else if (k == "TIME_LIMIT") time_limit = D();
    else if (k == "INPUT") {
      // casm:74-80: DADA (fp16 FT), FILE (one window), FILTERBANK (u8 TF)
      if (v == "DADA") { input_fmt = FLAG_IN_F16_FT; oneshot = false; }
      else if (v == "FILE") { input_fmt = FLAG_IN_U8_TF; oneshot = true; }
      else if (v == "FILTERBANK") { input_fmt = FLAG_IN_U8_TF; oneshot = false; }
      else { err = "unsupported INPUT " + v; return false; }
    } else if (k == "FLAG_STATS") {
      // the flagger's normalisation: MEAN (the reference's plane mean and
      // sigma) or ROBUST (each channel's median and 1.4826 x MAD)
      if (v == "MEAN") flag.robust = false;
      else if (v == "ROBUST") flag.robust = true;
      else { err = "unsupported FLAG_STATS " + v; return false; }
    } else if (k == "INPUT_DTYPE") {
      if (v == "U8") input_fmt = FLAG_IN_U8_TF;
      else if (v == "F32") input_fmt = FLAG_IN_F32_TF;
      else if (v == "F16") input_fmt = FLAG_IN_F16_FT;
      else { err = "unsupported INPUT_DTYPE " + v; return false; }
    } else if (k == "SEARCH") {
      if (v == "BOXCAR") search_mode = SEARCH_BOXCAR;
      else if (v == "HELLA") search_mode = SEARCH_HELLA;
      else { err = "unsupported SEARCH " + v; return false; }
    } else if (k == "SCRUNCH") {
      // casm:157-167: n lines "ts fs thresh nits" follow
      int n = I();
      if (n < 0 || n > 8) { err = "SCRUNCH count out of range (0..8)"; return false; }
      flag.nscrunch = n;
      for (int i = 0; i < n; ++i) {
        if (!std::getline(in, line)) { err = "SCRUNCH: missing lines"; return false; }
        int ts = 0, fs = 0, nits = 0;
        float th = 0;
        if (std::sscanf(line.c_str(), "%d %d %f %d", &ts, &fs, &th, &nits) < 3 || ts < 1 || fs < 1) {
          err = "SCRUNCH: bad line '" + line + "'";
          return false;
        }
        flag.scrunch[i] = FlagScrunch{ts, fs, th};
      }
    }
  }
  flag.spec_min = spec_min;
  flag.spec_max = spec_max;
  return true;
}

// ------------------------------------------------------------------ build
// clusters staged in pinned memory by the plain path per beam (more are
// fetched after the counts)
constexpr size_t KSTAGE_N = 1024;

static void alloc_zero(DevBuf &b, size_t bytes, hipStream_t s)
{
  b.ensure(std::max<size_t>(bytes, 16));
  zero_u32(b.as<uint32_t>(), (int)(round_up(b.bytes, 4) / 4), s);
}

static int input_esize(const Config &cfg)
{
  return cfg.input_fmt == FLAG_IN_U8_TF ? 1 : (cfg.input_fmt == FLAG_IN_F32_TF ? 4 : 2);
}

void beam_setup(BeamState &B, const Config &cfg, int W, hipStream_t s)
{
  const int maxW = cfg.width_max;
  if (B.nch <= 0) throw std::runtime_error("NCHAN must be positive");
  B.plan = PlanHost{};
  B.plan.nchans = (size_t)B.nch;
  B.plan.dt = (float)B.dt;
  B.plan.f0 = (float)B.f0;
  B.plan.df = (float)B.df;
  if (!(B.plan.df < 0.f) || !(B.plan.dt > 0.f)) throw std::runtime_error("FOFF must be < 0, TSAMP > 0");
  B.plan.build_delay_table();
  B.plan.generate_dm_list(B.dm_lo, B.dm_hi, (double)cfg.pulse_width_us, (double)B.dm_tol);
  if (B.killmask.empty()) B.killmask.assign((size_t)B.nch, 1);
  B.plan.killmask = B.killmask;
  B.plan.delays_live(0, B.plan.dm_list.size(), B.delta_host, B.live_host);
  B.nlive = (int)B.live_host.size();
  B.ndm = (int)B.plan.dm_list.size();
  B.md = (int64_t)B.plan.max_delay;
  if (cfg.oneshot) {
    B.ntime = W;
    B.ntime_dd = (int)(W - B.md);
    B.nout = B.ntime_dd - maxW;
    B.min_gulp = 0;
  } else {
    B.ntime = 32 * ((W + B.md + maxW) / 32 + 1);
    B.ntime_dd = W + maxW;
    B.nout = W;
    B.min_gulp = (int)(B.ntime / W);
  }
  if (B.nout <= 0) throw std::runtime_error("gulp too short for the maximum DM delay");
  B.rows = cfg.search_mode == SEARCH_HELLA ? B.ndm - 2 : B.ndm;
  if (B.rows <= 0) throw std::runtime_error("too few DM trials for the search mode");
  // the exact boxcar statistics keep 16 squared box sums of a thread in u64:
  // every box sum must stay below 2^30
  if (cfg.search_mode == SEARCH_BOXCAR && (int64_t)maxW * 255 * B.nlive >= (1LL << 30))
    throw std::runtime_error("WIDTH_MAX x live channels x 255 must stay below 2^30");
  const int64_t ring_len = cfg.oneshot ? (int64_t)round_up((size_t)W, 16) : B.ntime;
  B.ring_stride = (int64_t)round_up((size_t)ring_len, 64);
  B.ntime = cfg.oneshot ? ring_len : B.ntime;
  B.ring.ensure((size_t)B.nch * B.ring_stride);
  alloc_zero(B.ring, (size_t)B.nch * B.ring_stride, s);
  B.dstride = (int64_t)round_up((size_t)B.ntime_dd, 64);
  B.in_bytes = (int64_t)W * B.nch * input_esize(cfg);
  if (B.nlive > 0 && !B.delta) {
    B.d_delta.ensure((B.delta_host.size() + DD_DELTA_PAD) * 4);
    B.d_live.ensure(B.live_host.size() * 4);
    HAMD_CHECK(hipMemcpy(B.d_delta.p, B.delta_host.data(), B.delta_host.size() * 4,
                         hipMemcpyHostToDevice));
    HAMD_CHECK(hipMemcpy(B.d_live.p, B.live_host.data(), B.live_host.size() * 4,
                         hipMemcpyHostToDevice));
    B.delta = B.d_delta.as<int32_t>();
    B.live = B.d_live.as<int32_t>();
    std::vector<int32_t> dt;
    delays_transposed(B.delta_host, B.ndm, B.nlive, dt, B.dt_stride);
    B.d_delta_t.ensure(dt.size() * 4);
    HAMD_CHECK(hipMemcpy(B.d_delta_t.p, dt.data(), dt.size() * 4, hipMemcpyHostToDevice));
    B.delta_t = B.d_delta_t.as<int32_t>();
  }
  B.f_bp.ensure((size_t)B.nch * 4);
  alloc_zero(B.f_scal, sizeof(FScal), s);
  alloc_zero(B.f_count, flag_state_bytes(B.nch), s);
  alloc_zero(B.s_scal, sizeof(SScal), s);
  B.gidx = 0;
  B.crossings = 0;
}

WorkDims beam_dims(const BeamState &B, const Config &cfg, int W, int nw)
{
  WorkDims d;
  d.nch = B.nch;
  d.W = W;
  d.xs = (int64_t)round_up((size_t)W, 64);
  d.ndm = B.ndm;
  d.dstride = B.dstride;
  d.rows = B.rows;
  d.nout = B.nout;
  d.nw = nw;
  d.cand_cap = cfg.cand_cap;
  d.mask = flag_needs_mask(cfg.flag);
  d.tables = flag_value_tables(cfg.input_fmt, cfg.flag);
  d.planes = cfg.search_mode == SEARCH_HELLA;
  d.cl_rows = (B.ndm + std::max(1, cfg.cluster_ddm) - 1) / std::max(1, cfg.cluster_ddm);
  d.cl_cols = (B.nout + std::max(1, cfg.cluster_dt) - 1) / std::max(1, cfg.cluster_dt);
  d.cl_cap = std::max(1, std::min(cfg.cand_cap, cfg.max_giants));
  return d;
}

void WorkDims::merge(const WorkDims &o)
{
  nch = std::max(nch, o.nch);
  W = std::max(W, o.W);
  ndm = std::max(ndm, o.ndm);
  rows = std::max(rows, o.rows);
  nout = std::max(nout, o.nout);
  nw = std::max(nw, o.nw);
  cand_cap = std::max(cand_cap, o.cand_cap);
  xs = std::max(xs, o.xs);
  dstride = std::max(dstride, o.dstride);
  mask = mask || o.mask;
  planes = planes || o.planes;
  tables = tables || o.tables;
  cl_rows = std::max(cl_rows, o.cl_rows);
  cl_cols = std::max(cl_cols, o.cl_cols);
  cl_cap = std::max(cl_cap, o.cl_cap);
}

struct WorkSizes {
  size_t X, mask, D, planes, fpart, fts, spart, scnt, stiles, scands;
  size_t cgrid, cwork, ckey, cout;
};

static WorkSizes work_sizes(const WorkDims &d)
{
  WorkSizes z{};
  const size_t nblk = (size_t)(d.W + FLAG_BLOCK - 1) / FLAG_BLOCK;
  const size_t sblk = (size_t)(d.nout + SEARCH_BLOCK - 1) / SEARCH_BLOCK;
  const size_t ncnt = (size_t)d.nw * d.rows * sblk;
  // FP32 plane, or (value tables) the uint8 plane and two [nch][256] tables;
  // a batch mixing both paths cannot occur (one input format per context)
  z.X = d.tables ? flag_table_bytes(d.nch, d.xs) : (size_t)d.nch * d.xs * 4;
  z.mask = d.mask ? (size_t)d.nch * d.xs : 0;
  z.D = (size_t)d.ndm * d.dstride * 4;
  z.planes = d.planes ? (size_t)d.nw * d.rows * round_up((size_t)d.nout, 64) * 4 : 0;
  z.fpart = (size_t)d.nch * nblk * 8;
  // the final pass's time series, the per-channel robust statistics, the
  // time series' reciprocals, then the table steps' channel sums
  z.fts = round_up((size_t)d.W, 64) * 12 + (size_t)d.nch * 40;
  z.spart = ncnt * 8;
  z.scnt = ncnt * 4;
  z.stiles = ((ncnt + SEARCH_BLOCK - 1) / SEARCH_BLOCK + 1) * 4;
  z.scands = (size_t)d.cand_cap * sizeof(Cand);
  z.cgrid = (size_t)d.cl_rows * d.cl_cols * 4;
  z.cwork = (size_t)6 * d.cl_cap * 4;
  z.ckey = (size_t)d.cl_cap * 8;
  z.cout = (size_t)d.cl_cap * sizeof(ClusterRec);
  return z;
}

size_t WorkSet::bytes_for(const WorkDims &d)
{
  const WorkSizes z = work_sizes(d);
  return z.X + z.mask + z.D + z.planes + 2 * z.fpart + z.fts + 3 * z.spart + 6 * z.scnt +
         z.stiles + z.scands + 16 + z.cgrid + z.cwork + z.ckey + z.cout;
}

void WorkSet::ensure(const WorkDims &d)
{
  const WorkSizes z = work_sizes(d);
  X.ensure(z.X);
  if (z.mask) mask.ensure(z.mask);
  D.ensure(z.D);
  if (z.planes) planes.ensure(z.planes);
  f_part.ensure(z.fpart);
  f_part2.ensure(z.fpart);
  f_ts.ensure(z.fts);
  s_part.ensure(z.spart);
  s_part2.ensure(2 * z.spart);  // BOXCAR statistics: u128 per slot
  s_counts.ensure(z.scnt);
  s_offs.ensure(z.scnt);
  s_bmax.ensure(4 * z.scnt);    // one maximum per wave of each block
  s_tiles.ensure(z.stiles);
  s_cands.ensure(z.scands);
  s_ncand.ensure(16);  // crossings, clusters, candidates taken
  // the clustering grid starts empty (every cell 0xffffffff) and each launch
  // empties the cells it used
  const void *g0 = c_grid.p;
  c_grid.ensure(std::max<size_t>(z.cgrid, 4));
  if (c_grid.p != g0) {
    // on the null stream, which the contexts' non-blocking streams do not
    // wait for: complete it before any launch can read the grid
    HAMD_CHECK(hipMemset(c_grid.p, 0xff, c_grid.bytes));
    HAMD_CHECK(hipDeviceSynchronize());
  }
  c_work.ensure(z.cwork);
  c_key.ensure(z.ckey);
  c_out.ensure(z.cout);
}

void choose_default_tile(const std::vector<BeamState *> &beams, DDTile &tile, int &order)
{
  // measured on MI355X: one DM trial per wave in workgroups of 4 trials,
  // time-major tile order, every LDS read of a chunk issued before its sums
  // (1024 ch, DM 0-2000: 4.07 ms, the same as reading channel by channel;
  // 4096 ch, DM 0-5000: 284 ms against 328 ms); then 32 lanes per trial with
  // 8-channel chunks staged through registers; wider lane groups take over
  // when the delay spread of 8 DM trials does not fit
  int64_t min_ring = INT64_MAX;
  for (const BeamState *B : beams) min_ring = std::min(min_ring, B->ntime);
  order = 1;
  for (const DDTile t : {DDTile{64, 4, 4}, DDTile{32, 8, 0}, DDTile{64, 8, 0}}) {
    bool fits = true;
    for (const BeamState *B : beams) {
      if (B->nlive == 0) continue;
      const int sp = dd_max_spread(B->delta_host.data(), B->ndm, B->nlive, dd_tile_ND(t));
      if (!dd_tile_fits(t, sp, min_ring)) { fits = false; break; }
    }
    if (fits) {
      tile = t;
      return;
    }
  }
  tile = DDTile{64, 8, 0};
}

// entry tables of the one-trial-per-wave kernel (its blocks are tile-sized)
static void ensure_entries(BeamState &B, const DDTile &t)
{
  if (t.dma < 3 || B.nlive == 0 || B.ent_nd == dd_tile_ND(t)) return;
  std::vector<uint32_t> e;
  dd_entries(B.delta_host, B.ndm, B.nlive, dd_tile_ND(t), e, B.ent_stride);
  B.d_ent.ensure(e.size() * 4);
  HAMD_CHECK(hipMemcpy(B.d_ent.p, e.data(), e.size() * 4, hipMemcpyHostToDevice));
  B.ent = B.d_ent.as<uint32_t>();
  B.ent_nd = dd_tile_ND(t);
}

void build_plain_tiles(BeamState &B, const DDTile &t, int order)
{
  B.spread = B.nlive ? dd_max_spread(B.delta_host.data(), B.ndm, B.nlive, dd_tile_ND(t)) : 0;
  ensure_entries(B, t);
  B.ntiles = 0;
  if (B.nlive == 0) return;
  DDBeamDev dd{};
  dd.ndm = B.ndm;
  dd.nout = B.ntime_dd;
  dd.nlive = B.nlive;
  std::vector<DDTileDesc> tiles;
  dd_build_tiles(std::vector<DDBeamDev>{dd}, t, tiles, order);
  B.d_tiles.ensure(tiles.size() * sizeof(DDTileDesc));
  HAMD_CHECK(hipMemcpy(B.d_tiles.p, tiles.data(), tiles.size() * sizeof(DDTileDesc),
                       hipMemcpyHostToDevice));
  B.ntiles = (int)tiles.size();
}

SearchCommon search_common(const Config &cfg, const BeamState &B0, int nw)
{
  SearchCommon C{};
  C.nout = B0.nout;
  C.ntime_dd = B0.ntime_dd;
  C.maxW = cfg.width_max;
  C.minW = cfg.width_min;
  C.nw = nw;
  C.mode = cfg.search_mode;
  C.snr = cfg.snr;
  C.cap = cfg.cand_cap;
  C.nblk = (C.nout + SEARCH_BLOCK - 1) / SEARCH_BLOCK;
  C.wp = (C.nout / 256) * 256;
  if (C.wp == 0) C.wp = C.nout;
  C.box_max = (int64_t)cfg.width_max * 255 * B0.nch;
  return C;
}

void ctx_build(hella_ctx c)
{
  if (c->built) return;
  Config &cfg = c->cfg;
  const int W = (int)cfg.gulp;
  c->W = W;
  const int maxW = cfg.width_max, minW = cfg.width_min;
  c->nw = search_nwidth(cfg.search_mode, minW, maxW);
  if (c->nw <= 0 || c->nw > SEARCH_MAXW || maxW > 1024 || (maxW & (maxW - 1)) ||
      (minW & (minW - 1)))
    throw std::runtime_error("WIDTH_MIN/WIDTH_MAX must be powers of two, maxW <= 1024");
  hipStream_t s = c->stream;
  int64_t in_off = 0;
  c->max_nch = 0;
  c->max_rows = 0;
  c->wdims = WorkDims{};
  int ntdd = -1, nout = -1;
  for (auto &B : c->beams) {
    beam_setup(B, cfg, W, s);
    if (ntdd >= 0 && (ntdd != B.ntime_dd || nout != B.nout))
      throw std::runtime_error("one-shot sizing needs the same max delay in every beam");
    ntdd = B.ntime_dd;
    nout = B.nout;
    B.in_off = in_off;
    in_off += B.in_bytes;
    c->max_nch = std::max(c->max_nch, B.nch);
    c->max_rows = std::max(c->max_rows, B.rows);
    c->wdims.merge(beam_dims(B, cfg, W, c->nw));
  }
  c->gulp_bytes = in_off;
  // tile shape shared by all beams
  if (!c->tile_forced) {
    std::vector<BeamState *> bp;
    for (auto &B : c->beams) bp.push_back(&B);
    choose_default_tile(bp, c->tile, c->tile_order);
  }
  c->xs = c->wdims.xs;
  c->ws.resize(1);
  c->ws[0].ensure(c->wdims);
  if (cfg.search_mode == SEARCH_HELLA) {
    std::vector<float> all, k;
    c->kofs.clear();
    for (int i = 0; i < c->nw; ++i) {
      hella_kernel_host(minW << i, k);
      c->kofs.push_back((int)all.size());
      all.insert(all.end(), k.begin(), k.end());
    }
    c->d_kern.ensure(all.size() * 4);
    HAMD_CHECK(hipMemcpy(c->d_kern.p, all.data(), all.size() * 4, hipMemcpyHostToDevice));
  }
  // descriptors
  const size_t nb = c->beams.size();
  HAMD_CHECK(hipHostMalloc((void **)&c->h_fdesc, nb * sizeof(FlagBeamDev), hipHostMallocDefault));
  HAMD_CHECK(hipHostMalloc((void **)&c->h_ddesc, nb * sizeof(DDBeamDev), hipHostMallocDefault));
  HAMD_CHECK(hipHostMalloc((void **)&c->h_sdesc, nb * sizeof(SearchBeamDev), hipHostMallocDefault));
  HAMD_CHECK(hipHostMalloc((void **)&c->h_ncand, 4 * nb * sizeof(uint32_t), hipHostMallocDefault));
  c->cstage_n = std::max<size_t>(1, std::min<size_t>((size_t)std::max(cfg.max_giants, 0),
                                                      (size_t)std::max(cfg.cand_cap, 0)));
  // one staging slot per beam (the batched path harvests every beam at once)
  HAMD_CHECK(hipHostMalloc((void **)&c->h_cstage, nb * c->cstage_n * sizeof(Cand),
                           hipHostMallocDefault));
  HAMD_CHECK(hipHostMalloc((void **)&c->h_kstage, nb * KSTAGE_N * sizeof(ClusterRec),
                           hipHostMallocDefault));
  c->d_fdesc.ensure(nb * sizeof(FlagBeamDev));
  c->d_ddesc.ensure(nb * sizeof(DDBeamDev));
  c->d_sdesc.ensure(nb * sizeof(SearchBeamDev));
  c->d_in.ensure((size_t)c->gulp_bytes);
  c->common = search_common(cfg, c->beams[0], c->nw);
  c->common.box_max = (int64_t)cfg.width_max * 255 * c->max_nch;
  for (auto &B : c->beams) build_plain_tiles(B, c->tile, c->tile_order);
  HAMD_CHECK(hipStreamSynchronize(s));
  c->built = true;
}

// ------------------------------------------------------------------ descriptors
bool beam_ready(const Config &cfg, const BeamState &B)
{
  return cfg.oneshot || B.gidx >= B.min_gulp;
}

// ring column of the beam's next ingested gulp
static int64_t write_col(const Config &cfg, int W, const BeamState &B)
{
  if (cfg.oneshot) return 0;
  return (int64_t)((B.gidx * (long long)W) % B.ntime);
}

// ring column of the oldest sample of the window that ends with that gulp
static int64_t window_start(const Config &cfg, int W, const BeamState &B)
{
  if (cfg.oneshot) return 0;
  return (int64_t)(((B.gidx + 1) * (long long)W) % B.ntime);
}

long long window_abs0(const Config &cfg, int W, const BeamState &B)
{
  if (cfg.oneshot) return 0;
  return (B.gidx + 1) * (long long)W - B.ntime;
}

FlagBeamDev make_fdesc(const Config &cfg, int W, BeamState &B, WorkSet &w, const void *d_in)
{
  FlagBeamDev f{};
  f.X = w.X.as<float>();
  f.xs = (int64_t)round_up((size_t)W, 64);
  f.nch = B.nch;
  f.nblk = (W + FLAG_BLOCK - 1) / FLAG_BLOCK;
  f.in = d_in;
  f.in_fmt = cfg.input_fmt;
  f.flag1 = cfg.flag1;
  f.flag2 = cfg.flag2;
  f.in_stride = cfg.input_fmt == FLAG_IN_F16_FT ? W : B.nch;
  f.ring = B.ring.as<uint8_t>();
  f.ring_stride = B.ring_stride;
  f.ring_len = B.ntime;
  f.col0 = write_col(cfg, W, B);
  f.part = w.f_part.as<double>();
  f.part2 = w.f_part2.as<double>();
  f.bp = B.f_bp.as<float>();
  f.scal = B.f_scal.as<FScal>();
  f.flagcount = B.f_count.as<uint32_t>();
  f.fspec = reinterpret_cast<float *>(B.f_count.as<uint8_t>() + flag_state_spec_offset(B.nch));
  f.gsum = reinterpret_cast<unsigned long long *>(B.f_count.as<uint8_t>() +
                                                  flag_state_gsum_offset(B.nch));
  f.wp = W >= 256 ? (W / 256) * 256 : W;
  f.ts = w.f_ts.as<float>();
  f.robust = cfg.flag.robust ? 1 : 0;
  f.rstat = w.f_ts.as<float>() + round_up((size_t)W, 64);
  f.rts = reinterpret_cast<double *>(f.rstat + 2 * (size_t)B.nch);
  f.hsum = f.rts + round_up((size_t)W, 64);
  f.mask = w.mask.as<uint8_t>();
  if (flag_value_tables(cfg.input_fmt, cfg.flag)) {
    f.X8 = w.X.as<uint8_t>();
    f.lut = reinterpret_cast<float *>(w.X.as<uint8_t>() + flag_table_offset(B.nch, f.xs));
    f.tsp = reinterpret_cast<double *>(f.lut + (size_t)2 * B.nch * 256);
    f.hist = reinterpret_cast<uint32_t *>(w.X.as<uint8_t>() + flag_hist_offset(B.nch, f.xs));
  }
  return f;
}

DDBeamDev make_ddesc(const Config &cfg, int W, BeamState &B, WorkSet &w)
{
  DDBeamDev d{};
  d.base = B.ring.as<uint8_t>();
  d.row_stride = B.ring_stride;
  d.ring = B.ntime;
  d.start = window_start(cfg, W, B);
  d.delta = B.delta;
  d.live = B.live;
  d.delta_t = B.delta_t;
  d.dt_stride = B.dt_stride;
  // the entries assume tile starts at ring columns that are multiples of 16:
  // window starts advance by the gulp, tile starts by multiples of 1792
  const bool aligned = cfg.oneshot || (W % 16 == 0 && B.ntime % 16 == 0);
  d.ent = aligned ? B.ent : nullptr;
  d.ent_stride = B.ent_stride;
  d.ent_nd = B.ent_nd;
  d.nlive = B.nlive;
  d.ndm = B.ndm;
  d.nout = B.ntime_dd;
  d.out_f32 = 0;
  d.out = w.D.p;
  d.out_stride = B.dstride;
  return d;
}

SearchBeamDev make_sdesc(const Config &cfg, BeamState &B, WorkSet &w)
{
  SearchBeamDev s{};
  s.D = w.D.as<uint32_t>();
  s.ds = B.dstride;
  s.rows = B.rows;
  s.row0 = cfg.search_mode == SEARCH_HELLA ? 1 : 0;
  s.planes = w.planes.as<float>();
  s.ps = (int64_t)round_up((size_t)B.nout, 64);
  s.part = w.s_part.as<double>();
  s.part2 = w.s_part2.as<double>();
  s.scal = B.s_scal.as<SScal>();
  s.counts = w.s_counts.as<uint32_t>();
  s.offs = w.s_offs.as<uint32_t>();
  s.bmax = w.s_bmax.as<uint32_t>();
  s.tiles = w.s_tiles.as<uint32_t>();
  s.cands = w.s_cands.as<Cand>();
  s.ncand = w.s_ncand.as<uint32_t>();
  const WorkDims d = beam_dims(B, cfg, (int)cfg.gulp, 1);
  s.cl_grid = w.c_grid.as<uint32_t>();
  s.cl_work = w.c_work.as<uint32_t>();
  s.cl_key = w.c_key.as<unsigned long long>();
  s.clusters = w.c_out.as<ClusterRec>();
  s.cl_n = s.ncand + 1;
  s.cl_rows = d.cl_rows;
  s.cl_cols = d.cl_cols;
  s.cl_cap = d.cl_cap;
  return s;
}

ClusterArgs cluster_args(const Config &cfg, int prefix, bool cumulative)
{
  ClusterArgs a{};
  a.cand_cap = cfg.cand_cap;
  a.max_giants = cfg.max_giants;
  a.prefix = prefix;
  a.cumulative = cumulative ? 1 : 0;
  a.ct = std::max(1, cfg.cluster_dt);
  a.cd = std::max(1, cfg.cluster_ddm);
  return a;
}

void append_clusters(const ClusterRec *cl, size_t n, int rep0, int beam_id,
                     std::vector<hella_cluster> &out)
{
  static_assert(sizeof(ClusterRec) == sizeof(hella_cluster), "cluster record layout");
  out.reserve(out.size() + n);
  for (size_t i = 0; i < n; ++i) {
    const ClusterRec &x = cl[i];
    hella_cluster h;
    h.rep = rep0 + x.rep;
    h.count = x.count;
    h.samp_min = x.samp_min;
    h.samp_max = x.samp_max;
    h.dm_min = x.dm_min;
    h.dm_max = x.dm_max;
    h.beam = beam_id;
    out.push_back(h);
  }
}

// ------------------------------------------------------------------ timing
struct StageTimer {
  hella_ctx c;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int idx;
  StageTimer(hella_ctx c_, int i) : c(c_), idx(i)
  {
    if (!c->timing) return;
    HAMD_CHECK(hipEventCreate(&e0));
    HAMD_CHECK(hipEventCreate(&e1));
    HAMD_CHECK(hipEventRecord(e0, c->stream));
  }
  ~StageTimer()
  {
    if (!c->timing) return;
    float ms = 0.f;
    if (hipEventRecord(e1, c->stream) == hipSuccess && hipEventSynchronize(e1) == hipSuccess &&
        hipEventElapsedTime(&ms, e0, e1) == hipSuccess)
      c->stage_ms[(size_t)idx] += ms;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
  }
};

// ------------------------------------------------------------------ plain stages
const void *ctx_stage_input(hella_ctx c, const void *gulp, bool host_ptr)
{
  if (!host_ptr) return gulp;
  HAMD_CHECK(hipMemcpyAsync(c->d_in.p, gulp, (size_t)c->gulp_bytes, hipMemcpyHostToDevice,
                            c->stream));
  return c->d_in.p;
}

void ctx_flag_beam(hella_ctx c, int b, const void *d_gulp)
{
  BeamState &B = c->beams[(size_t)b];
  StageTimer t(c, 0);
  c->h_fdesc[b] = make_fdesc(c->cfg, c->W, B, c->ws[0],
                             static_cast<const uint8_t *>(d_gulp) + B.in_off);
  FlagBeamDev *dd = c->d_fdesc.as<FlagBeamDev>() + b;
  HAMD_CHECK(hipMemcpyAsync(dd, &c->h_fdesc[b], sizeof(FlagBeamDev), hipMemcpyHostToDevice,
                            c->stream));
  FlagRunner::run(dd, 1, B.nch, c->W, c->cfg.flag, c->stream,
                  flag_value_tables(c->cfg.input_fmt, c->cfg.flag));
}

void ctx_dedisperse(hella_ctx c, int b)
{
  BeamState &B = c->beams[(size_t)b];
  StageTimer t(c, 1);
  c->h_ddesc[b] = make_ddesc(c->cfg, c->W, B, c->ws[0]);
  if (B.nlive == 0) {
    fill_zero_rows(c->ws[0].D.p, B.dstride, B.ntime_dd, B.ndm, c->stream);
  } else {
    DDBeamDev *dd = c->d_ddesc.as<DDBeamDev>() + b;
    HAMD_CHECK(hipMemcpyAsync(dd, &c->h_ddesc[b], sizeof(DDBeamDev), hipMemcpyHostToDevice,
                              c->stream));
    dd_launch(dd, B.d_tiles.as<DDTileDesc>(), B.ntiles, c->tile, B.spread, c->stream, B.ntime);
  }
  c->cur_beam = b;
}

void ctx_smooth(hella_ctx c, int b)
{
  BeamState &B = c->beams[(size_t)b];
  StageTimer t(c, 2);
  c->h_sdesc[b] = make_sdesc(c->cfg, B, c->ws[0]);
  SearchBeamDev *sd = c->d_sdesc.as<SearchBeamDev>() + b;
  HAMD_CHECK(hipMemcpyAsync(sd, &c->h_sdesc[b], sizeof(SearchBeamDev), hipMemcpyHostToDevice,
                            c->stream));
  SearchRunner::smooth(sd, 1, B.rows, c->common, c->d_kern.as<float>(), c->kofs, c->stream);
}

// the first candidates of a beam's search under the caps of peaks.cu:77-78
// (per beam) and peaks.cu:101-104 (per gulp, applied by the caller)
static size_t beam_take(const Config &cfg, uint32_t ncand)
{
  size_t n = std::min<size_t>(ncand, (size_t)cfg.cand_cap);
  return std::min<size_t>(n, (size_t)cfg.max_giants);
}

// append beam b's candidates to a gulp list holding at most cap entries
void append_candidates(const Config &cfg, int W, const BeamState &B, int beam_id, const Cand *cs,
                       size_t n, size_t cap, std::vector<hella_candidate> &out)
{
  const long long a0 = window_abs0(cfg, W, B) + cfg.width_max / 2;
  if (out.size() < cap) out.reserve(std::min(cap, out.size() + n));
  for (size_t i = 0; i < n; ++i) {
    const Cand &x = cs[i];
    if (out.size() >= cap) break;
    hella_candidate h;
    h.snr = x.snr;
    h.beam = beam_id;
    h.samp = x.samp;
    h.width = x.width;
    h.dm_idx = x.dm_idx;
    h.dm = B.plan.dm_list[(size_t)x.dm_idx];
    h.abs_samp = a0 + x.samp;
    out.push_back(h);
  }
}

// records of a work set's lists that fit one staging slot
static size_t stage_cap(hella_ctx c, const WorkSet &w)
{
  return std::min<size_t>(c->cstage_n, (size_t)w.s_cands.bytes / sizeof(Cand));
}

static size_t kstage_cap(const WorkSet &w)
{
  return std::min<size_t>(KSTAGE_N, (size_t)w.c_out.bytes / sizeof(ClusterRec));
}

int ctx_find_peaks(hella_ctx c, int b)
{
  BeamState &B = c->beams[(size_t)b];
  WorkSet &w = c->ws[0];
  uint32_t nc = 0;
  size_t ntake = 0, ncl = 0;
  const Cand *cs = nullptr;
  const ClusterRec *cl = nullptr;
  std::vector<Cand> spill_c;
  std::vector<ClusterRec> spill_k;
  const int rep0 = (int)c->peaks.size();
  {
    StageTimer t(c, 3);
    SearchBeamDev *sd = c->d_sdesc.as<SearchBeamDev>() + b;
    SearchRunner::peaks(sd, 1, B.rows, c->common, c->stream);
    SearchRunner::cluster(sd, 1, cluster_args(c->cfg, rep0, false), c->stream);
    // the counts, the candidates the host can keep and the first KSTAGE_N
    // clusters in one round trip: the device writes them into the pinned
    // staging; longer cluster lists fetch the rest after the counts
    uint32_t *hn = c->h_ncand + 4 * (size_t)b;
    const size_t cap = stage_cap(c, w), kcap = kstage_cap(w);
    SearchRunner::harvest(sd, hn, c->h_cstage, (int)cap, c->h_kstage, (int)kcap, c->stream);
    HAMD_CHECK(hipStreamSynchronize(c->stream));
    nc = hn[0];
    ntake = beam_take(c->cfg, nc);
    ncl = hn[1];
    cs = results_view<Cand>(c->h_cstage, cap, w.s_cands.as<Cand>(), ntake, spill_c, c->stream);
    cl = results_view<ClusterRec>(c->h_kstage, kcap, w.c_out.as<ClusterRec>(), ncl, spill_k,
                                  c->stream);
  }
  B.crossings = nc;
  append_candidates(c->cfg, c->W, B, b, cs, ntake, (size_t)c->cfg.max_giants, c->peaks);
  append_clusters(cl, ncl, rep0, b, c->clusters);
  return (int)ntake;
}

void ctx_end_gulp(hella_ctx c)
{
  c->gulp_index += 1;
  for (auto &B : c->beams) B.gidx += 1;
}

void ctx_process_plain(hella_ctx c, const void *gulp, bool host_ptr)
{
  const auto t0 = std::chrono::steady_clock::now();
  c->batch_mode = false;
  c->stage_ms.fill(0.f);
  c->peaks.clear();
  c->clusters.clear();
  const void *d_gulp = ctx_stage_input(c, gulp, host_ptr);
  for (size_t b = 0; b < c->beams.size(); ++b) ctx_flag_beam(c, (int)b, d_gulp);
  for (size_t b = 0; b < c->beams.size(); ++b) {
    BeamState &B = c->beams[b];
    if (!beam_ready(c->cfg, B)) continue;
    if ((int)c->peaks.size() >= c->cfg.max_giants) break;
    const double el =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (el >= c->cfg.time_limit) break;
    ctx_dedisperse(c, (int)b);
    ctx_smooth(c, (int)b);
    ctx_find_peaks(c, (int)b);
  }
  HAMD_CHECK(hipStreamSynchronize(c->stream));
  ctx_end_gulp(c);
}

// ------------------------------------------------------------------ batched path
// work set b belongs to beam b
static void batch_alloc(hella_ctx c)
{
  if (c->ws.size() < c->beams.size()) c->ws.resize(c->beams.size());
  for (size_t b = 0; b < c->beams.size(); ++b)
    c->ws[b].ensure(beam_dims(c->beams[b], c->cfg, c->W, c->nw));
}

static void batch_destroy_graph(Batch &bt)
{
  if (bt.exec) (void)hipGraphExecDestroy(bt.exec);
  bt.exec = nullptr;
  bt.graph_ok = false;
}

// (re)build the searchable-beam set, its tile table and descriptors
static void batch_prepare(hella_ctx c)
{
  std::vector<int> ready;
  for (size_t b = 0; b < c->beams.size(); ++b)
    if (beam_ready(c->cfg, c->beams[b])) ready.push_back((int)b);
  Batch &bt = c->batch;
  if (bt.valid && ready == bt.beams) return;
  batch_destroy_graph(bt);
  bt.beams = ready;
  std::vector<DDBeamDev> dd;
  int spread = 0;
  for (int b : ready) {
    dd.push_back(make_ddesc(c->cfg, c->W, c->beams[(size_t)b], c->ws[(size_t)b]));
    spread = std::max(spread, c->beams[(size_t)b].spread);
  }
  std::vector<DDTileDesc> tiles;
  dd_build_tiles(dd, c->tile, tiles, c->tile_order);
  bt.d_tiles.ensure(std::max<size_t>(tiles.size(), 1) * sizeof(DDTileDesc));
  if (!tiles.empty())
    HAMD_CHECK(hipMemcpy(bt.d_tiles.p, tiles.data(), tiles.size() * sizeof(DDTileDesc),
                         hipMemcpyHostToDevice));
  bt.ntiles = (int)tiles.size();
  bt.tile = c->tile;
  bt.spread = spread;
  bt.valid = true;
}

// record the device work of one batched gulp on the context stream
static void batch_record(hella_ctx c)
{
  Batch &bt = c->batch;
  const size_t nb = c->beams.size();
  hipStream_t s = c->stream;
  HAMD_CHECK(hipMemcpyAsync(c->d_fdesc.p, c->h_fdesc, nb * sizeof(FlagBeamDev),
                            hipMemcpyHostToDevice, s));
  FlagRunner::run(c->d_fdesc.as<FlagBeamDev>(), (int)nb, c->max_nch, c->W, c->cfg.flag, s,
                  flag_value_tables(c->cfg.input_fmt, c->cfg.flag));
  const int nr = (int)bt.beams.size();
  if (nr == 0) return;
  HAMD_CHECK(hipMemcpyAsync(c->d_ddesc.p, c->h_ddesc, (size_t)nr * sizeof(DDBeamDev),
                            hipMemcpyHostToDevice, s));
  HAMD_CHECK(hipMemcpyAsync(c->d_sdesc.p, c->h_sdesc, (size_t)nr * sizeof(SearchBeamDev),
                            hipMemcpyHostToDevice, s));
  for (int i = 0; i < nr; ++i) {
    const int b = bt.beams[(size_t)i];
    const BeamState &B = c->beams[(size_t)b];
    if (B.nlive == 0) fill_zero_rows(c->ws[(size_t)b].D.p, B.dstride, B.ntime_dd, B.ndm, s);
  }
  int64_t min_ring = INT64_MAX;
  for (int b : bt.beams) min_ring = std::min(min_ring, c->beams[(size_t)b].ntime);
  dd_launch(c->d_ddesc.as<DDBeamDev>(), bt.d_tiles.as<DDTileDesc>(), bt.ntiles, bt.tile, bt.spread,
            s, min_ring);
  int max_rows = 0;
  for (int b : bt.beams) max_rows = std::max(max_rows, c->beams[(size_t)b].rows);
  SearchRunner::run(c->d_sdesc.as<SearchBeamDev>(), nr, max_rows, c->common,
                    c->d_kern.as<float>(), c->kofs, s);
  SearchRunner::cluster(c->d_sdesc.as<SearchBeamDev>(), nr, cluster_args(c->cfg, 0, true), s);
  // every beam's counts, candidates and first clusters into its staging slot
  for (int i = 0; i < nr; ++i) {
    const WorkSet &w = c->ws[(size_t)bt.beams[(size_t)i]];
    SearchRunner::harvest(c->d_sdesc.as<SearchBeamDev>() + i, &c->h_ncand[4 * (size_t)i],
                          c->h_cstage + (size_t)i * c->cstage_n, (int)stage_cap(c, w),
                          c->h_kstage + (size_t)i * KSTAGE_N, (int)kstage_cap(w), s);
  }
}

void ctx_process_batched(hella_ctx c, const void *gulp, bool host_ptr)
{
  c->batch_mode = true;
  c->stage_ms.fill(0.f);
  c->peaks.clear();
  c->clusters.clear();
  batch_alloc(c);
  batch_prepare(c);
  Batch &bt = c->batch;
  const void *d_gulp = ctx_stage_input(c, gulp, host_ptr);
  // refresh the per-gulp descriptor values (read by the graph's copy nodes)
  for (size_t b = 0; b < c->beams.size(); ++b)
    c->h_fdesc[b] = make_fdesc(c->cfg, c->W, c->beams[b], c->ws[b],
                               static_cast<const uint8_t *>(d_gulp) + c->beams[b].in_off);
  for (size_t i = 0; i < bt.beams.size(); ++i) {
    const size_t b = (size_t)bt.beams[i];
    c->h_ddesc[i] = make_ddesc(c->cfg, c->W, c->beams[b], c->ws[b]);
    c->h_sdesc[i] = make_sdesc(c->cfg, c->beams[b], c->ws[b]);
  }
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (c->timing) {
    HAMD_CHECK(hipEventCreate(&e0));
    HAMD_CHECK(hipEventCreate(&e1));
    HAMD_CHECK(hipEventRecord(e0, c->stream));
  }
  if (!bt.graph_ok) {
    hipGraph_t g = nullptr;
    HAMD_CHECK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    try {
      batch_record(c);
    } catch (...) {
      (void)hipStreamEndCapture(c->stream, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    HAMD_CHECK(hipStreamEndCapture(c->stream, &g));
    HAMD_CHECK(hipGraphInstantiate(&bt.exec, g, nullptr, nullptr, 0));
    (void)hipGraphDestroy(g);
    bt.graph_ok = true;
  }
  HAMD_CHECK(hipGraphLaunch(bt.exec, c->stream));
  if (c->timing) HAMD_CHECK(hipEventRecord(e1, c->stream));
  HAMD_CHECK(hipStreamSynchronize(c->stream));
  if (c->timing) {
    float ms = 0.f;
    HAMD_CHECK(hipEventElapsedTime(&ms, e0, e1));
    c->stage_ms[0] = ms;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
  }
  // candidates and clusters from the staging slots the graph filled (lists
  // longer than a slot fetch the rest)
  std::vector<Cand> spill_c;
  std::vector<ClusterRec> spill_k;
  for (size_t i = 0; i < bt.beams.size(); ++i) {
    if ((int)c->peaks.size() >= c->cfg.max_giants) break;
    BeamState &B = c->beams[(size_t)bt.beams[i]];
    const WorkSet &w = c->ws[(size_t)bt.beams[i]];
    const uint32_t *hn = c->h_ncand + 4 * i;
    B.crossings = hn[0];
    const size_t n = beam_take(c->cfg, hn[0]), ncl = hn[1];
    const Cand *cs = results_view<Cand>(c->h_cstage + i * c->cstage_n, stage_cap(c, w),
                                        w.s_cands.as<Cand>(), n, spill_c, c->stream);
    const ClusterRec *cl = results_view<ClusterRec>(c->h_kstage + i * KSTAGE_N, kstage_cap(w),
                                                    w.c_out.as<ClusterRec>(), ncl, spill_k,
                                                    c->stream);
    const int rep0 = (int)c->peaks.size();
    append_candidates(c->cfg, c->W, B, bt.beams[i], cs, n, (size_t)c->cfg.max_giants, c->peaks);
    append_clusters(cl, ncl, rep0, bt.beams[i], c->clusters);
  }
  ctx_end_gulp(c);
}

float ctx_bench_stage(hella_ctx c, int b, int which, int reps)
{
  BeamState &B = c->beams[(size_t)b];
  hipStream_t s = c->stream;
  auto launch = [&]() {
    if (which == 0) {
      FlagRunner::run(c->d_fdesc.as<FlagBeamDev>() + b, 1, B.nch, c->W, c->cfg.flag, s,
                      flag_value_tables(c->cfg.input_fmt, c->cfg.flag));
    } else if (which == 1) {
      if (B.nlive)
        dd_launch(c->d_ddesc.as<DDBeamDev>() + b, B.d_tiles.as<DDTileDesc>(), B.ntiles, c->tile,
                  B.spread, s, B.ntime);
    } else {
      SearchRunner::run(c->d_sdesc.as<SearchBeamDev>() + b, 1, B.rows, c->common,
                        c->d_kern.as<float>(), c->kofs, s);
    }
  };
  launch();  // warm
  hipEvent_t e0, e1;
  HAMD_CHECK(hipEventCreate(&e0));
  HAMD_CHECK(hipEventCreate(&e1));
  HAMD_CHECK(hipEventRecord(e0, s));
  for (int r = 0; r < reps; ++r) launch();
  HAMD_CHECK(hipEventRecord(e1, s));
  HAMD_CHECK(hipEventSynchronize(e1));
  float ms = 0.f;
  HAMD_CHECK(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return ms / (float)reps;
}

// ------------------------------------------------------------------ autotune
void ctx_autotune(hella_ctx c, const char *path)
{
  ctx_build(c);
  batch_alloc(c);
  // time the dedispersion of every beam's current window with each shape
  std::vector<DDBeamDev> dd;
  int64_t min_ring = INT64_MAX;
  for (size_t b = 0; b < c->beams.size(); ++b) {
    BeamState &B = c->beams[b];
    if (B.nlive > 0) {
      dd.push_back(make_ddesc(c->cfg, c->W, B, c->ws[b]));
      min_ring = std::min(min_ring, B.ntime);
    }
  }
  if (dd.empty()) return;
  DevBuf d_beams;
  d_beams.ensure(dd.size() * sizeof(DDBeamDev));
  HAMD_CHECK(hipMemcpy(d_beams.p, dd.data(), dd.size() * sizeof(DDBeamDev), hipMemcpyHostToDevice));
  struct Res { int G, cc, order, dma; float ms; };
  std::vector<Res> res;
  hipEvent_t e0, e1;
  HAMD_CHECK(hipEventCreate(&e0));
  HAMD_CHECK(hipEventCreate(&e1));
  DevBuf d_tiles;
  std::vector<DDTile> shapes;
  shapes.push_back(DDTile{64, 4, 5});
  for (int dma : {3, 4})
    for (int cc : {8, 4}) shapes.push_back(DDTile{64, cc, dma});
  for (int dma : {2, 1, 0})
    for (int G : {8, 16, 32, 64})
      for (int cc : {8, 16, 32})
        if (!(dma == 2 && cc == 32)) shapes.push_back(DDTile{G, cc, dma});
  for (const DDTile &t : shapes)
      for (int order : {0, 1}) {
        const int G = t.lanes_per_dm, cc = t.chunk, dma = t.dma;
        int spread = 0;
        bool fits = true;
        for (auto &B : c->beams) {
          if (!B.nlive) continue;
          const int sp = dd_max_spread(B.delta_host.data(), B.ndm, B.nlive, dd_tile_ND(t));
          spread = std::max(spread, sp);
          if (!dd_tile_fits(t, sp, min_ring)) fits = false;
        }
        if (!fits) continue;
        if (dma >= 3) {
          // this shape's entry tables
          size_t j = 0;
          for (size_t b = 0; b < c->beams.size(); ++b) {
            BeamState &B = c->beams[b];
            if (!B.nlive) continue;
            ensure_entries(B, t);
            dd[j++] = make_ddesc(c->cfg, c->W, B, c->ws[b]);
          }
          HAMD_CHECK(hipMemcpy(d_beams.p, dd.data(), dd.size() * sizeof(DDBeamDev),
                               hipMemcpyHostToDevice));
        }
        std::vector<DDTileDesc> tiles;
        dd_build_tiles(dd, t, tiles, order);
        d_tiles.ensure(tiles.size() * sizeof(DDTileDesc));
        HAMD_CHECK(hipMemcpy(d_tiles.p, tiles.data(), tiles.size() * sizeof(DDTileDesc),
                             hipMemcpyHostToDevice));
        if (dma == 2 && min_ring < (int64_t)dd_lds_bytes(DDTile{G, 1, 0}, spread)) continue;
        dd_launch(d_beams.as<DDBeamDev>(), d_tiles.as<DDTileDesc>(), (int)tiles.size(), t, spread,
                  c->stream, min_ring);
        HAMD_CHECK(hipEventRecord(e0, c->stream));
        const int reps = 3;
        for (int r = 0; r < reps; ++r)
          dd_launch(d_beams.as<DDBeamDev>(), d_tiles.as<DDTileDesc>(), (int)tiles.size(), t,
                    spread, c->stream, min_ring);
        HAMD_CHECK(hipEventRecord(e1, c->stream));
        HAMD_CHECK(hipEventSynchronize(e1));
        float ms = 0.f;
        HAMD_CHECK(hipEventElapsedTime(&ms, e0, e1));
        res.push_back(Res{G, cc, order, dma, ms / reps});
      }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (res.empty()) return;
  const Res best = *std::min_element(res.begin(), res.end(),
                                     [](const Res &a, const Res &b) { return a.ms < b.ms; });
  c->tile = DDTile{best.G, best.cc, best.dma};
  c->tile_order = best.order;
  c->tile_forced = true;
  for (auto &B : c->beams) build_plain_tiles(B, c->tile, c->tile_order);
  c->batch.valid = false;
  batch_destroy_graph(c->batch);
  if (path && *path) {
    FILE *f = std::fopen(path, "w");
    if (f) {
      std::fprintf(f, "{\"best\": {\"lanes_per_dm\": %d, \"chunk\": %d, \"order\": %d, \"dma\": %d, \"ms\": %.4f},\n",
                   best.G, best.cc, best.order, best.dma, best.ms);
      std::fprintf(f, " \"nbeam\": %zu, \"nchan_max\": %d, \"gulp\": %d,\n \"trials\": [", c->beams.size(),
                   c->max_nch, c->W);
      for (size_t i = 0; i < res.size(); ++i)
        std::fprintf(f, "%s\n  {\"lanes_per_dm\": %d, \"chunk\": %d, \"order\": %d, \"dma\": %d, \"ms\": %.4f}",
                     i ? "," : "", res[i].G, res[i].cc, res[i].order, res[i].dma, res[i].ms);
      std::fprintf(f, "]}\n");
      std::fclose(f);
    }
  }
}

}  // namespace hamd
