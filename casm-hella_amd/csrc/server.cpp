// server.cpp -- streaming server: beams join and leave a running stream of
// chunks read from the block ring (ring.cpp), each beam's window and flagger
// state stay resident on the device, and chunks are launched in continuous
// batches on a few launch chains ("lanes", one stream each).
//
// Per beam the work is the per-gulp sequence of the reference
// (src/apps/casm_hella_refactored.cu:659-835): flag the chunk into the beam's
// window, and once the window is full (casm:632) dedisperse, search and find
// peaks.  A lane carries up to max_batch chunks of different beams through one
// launch chain: a flag graph over all of them and a search graph over those
// whose window is full, both captured once per batch size and replayed with
// fresh descriptors (uploaded by the graphs' copy nodes from pinned memory).
// A beam has at most one chunk in flight, so its chunks are processed in
// order; chunks of other beams that arrive meanwhile go to the next free
// lane.  Admission: a beam becomes resident only while the HBM budget holds
// its window plus the fixed work sets; until then its chunks wait on the host.
#include <algorithm>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <new>
#include <stdexcept>
#include <vector>

#include "pipeline.hpp"

namespace hamd {
long long now_ns();
}

using namespace hamd;

namespace {

struct Chunk {
  int ring_index = -1;        // ring block held until the chunk is done
  const void *data = nullptr; // ring block, or copy below
  std::vector<uint8_t> copy;  // host copy while the beam waits for admission
  long long seq = 0, t_arr = 0;
  bool last = false;
};

struct SBeam {
  int id = -1;
  std::unique_ptr<BeamState> B;  // resident state; null while waiting
  std::deque<Chunk> q;
  bool in_flight = false;
  bool waited = false;
};

struct Job {
  SBeam *sb = nullptr;
  Chunk ck;
  bool searched = false;
};

struct Graph {
  hipGraphExec_t exec = nullptr;
  DevBuf d_tiles;
  int ntiles = 0;
};

// Results of a lane's chunks come back in one pass: the gather kernel writes
// each work set's counts, first candidates and first clusters straight into
// the lane's pinned staging (only the entries that exist; no copy-engine
// transfer).  Longer lists fetch the rest after the counts.
constexpr int SRV_KC = 256;  // candidates staged per chunk
constexpr int SRV_KK = 64;   // clusters staged per chunk

struct StageSlot {
  const uint32_t *ncand;  // crossings, clusters, candidates taken
  const Cand *cands;
  const ClusterRec *clusters;
  int cap_c, cap_k;  // entries the two lists hold
};

struct StageOut {
  uint32_t n[4];
  Cand c[SRV_KC];
  ClusterRec k[SRV_KK];
};

__global__ __launch_bounds__(256) void k_stage_results(const StageSlot *__restrict__ slots,
                                                       StageOut *__restrict__ out, int n)
{
  const int i = blockIdx.x;
  if (i >= n) return;
  const StageSlot S = slots[i];
  StageOut &o = out[i];
  const int l = threadIdx.x;
  const uint32_t nc = S.ncand[0], nk = S.ncand[1];
  if (l < 3) o.n[l] = S.ncand[l];
  const int tc = min((int)min(nc, (uint32_t)SRV_KC), S.cap_c);
  const int tk = min((int)min(nk, (uint32_t)SRV_KK), S.cap_k);
  for (int j = l; j < tc; j += 256) o.c[j] = S.cands[j];
  for (int j = l; j < tk; j += 256) o.k[j] = S.clusters[j];
}

struct Lane {
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  std::vector<WorkSet> ws;
  FlagBeamDev *h_fdesc = nullptr;
  DDBeamDev *h_ddesc = nullptr;
  SearchBeamDev *h_sdesc = nullptr;
  StageOut *h_stage = nullptr;  // pinned: one StageOut per work set
  DevBuf d_fdesc, d_ddesc, d_sdesc, d_in, d_slots;
  std::map<int, Graph> flag_g, search_g;  // by batch size
  std::vector<Job> jobs;
  bool busy = false;
  long long t_launch = 0;
};

struct Result {
  hella_chunk_result r{};
  std::vector<hella_candidate> cands;
  std::vector<hella_cluster> clusters;
};

void destroy_graphs(std::map<int, Graph> &m)
{
  for (auto &kv : m)
    if (kv.second.exec) (void)hipGraphExecDestroy(kv.second.exec);
  m.clear();
}

}  // namespace

struct hella_server_s {
  Config cfg;
  hella_server_opts opts{};
  int W = 0, nw = 0;
  size_t chunk_bytes = 0;
  BeamState geom;        // template: plan, delays (device), sizing
  WorkDims dims;
  DDTile tile{32, 8, 0};
  int tile_order = 1;
  SearchCommon common{};
  DevBuf d_kern;
  std::vector<int> kofs;
  std::vector<Lane> lanes;
  std::map<int, SBeam> beams;                       // by beam id
  std::vector<std::unique_ptr<BeamState>> free_bs;  // windows of departed beams
  int max_beams = 0;     // resident beams the budget allows
  size_t beam_bytes = 0, work_bytes = 0;
  std::vector<Result> results;
  hella_server_stats st{};
  int resident = 0;

  ~hella_server_s()
  {
    for (auto &L : lanes) {
      if (L.stream) (void)hipStreamSynchronize(L.stream);
      destroy_graphs(L.flag_g);
      destroy_graphs(L.search_g);
      if (L.h_fdesc) (void)hipHostFree(L.h_fdesc);
      if (L.h_ddesc) (void)hipHostFree(L.h_ddesc);
      if (L.h_sdesc) (void)hipHostFree(L.h_sdesc);
      if (L.h_stage) (void)hipHostFree(L.h_stage);
      if (L.done) (void)hipEventDestroy(L.done);
      if (L.stream) (void)hipStreamDestroy(L.stream);
    }
  }
};

namespace {

hella_status sfail(hella_status s, const std::string &msg)
{
  set_last_error(msg);
  return s;
}

template <class F> hella_status sguard(F &&f)
{
  try {
    f();
    return HELLA_OK;
  } catch (const HipError &e) {
    return sfail(e.err == hipErrorOutOfMemory ? HELLA_ERR_ALLOC : HELLA_ERR_GPU,
                 std::string(e.what) + ": " + hipGetErrorString(e.err));
  } catch (const std::bad_alloc &) {
    return sfail(HELLA_ERR_ALLOC, "host allocation failed");
  } catch (const std::invalid_argument &e) {
    return sfail(HELLA_ERR_INVALID, e.what());
  } catch (const std::runtime_error &e) {
    return sfail(HELLA_ERR_CONFIG, e.what());
  } catch (...) {
    return sfail(HELLA_ERR_GPU, "unknown error");
  }
}

size_t resident_bytes(const BeamState &g)
{
  const auto r = [](size_t b) { return round_up(std::max<size_t>(b, 16), 256); };
  return r((size_t)g.nch * g.ring_stride) + r((size_t)g.nch * 4) + r(flag_state_bytes(g.nch)) +
         r(sizeof(FScal)) + r(sizeof(SScal));
}

// plan, sizing and device delay tables shared by every beam of the server
void setup_geometry(hella_server s, const char *config)
{
  std::string err;
  if (!s->cfg.parse(config ? config : "", err)) throw std::runtime_error(err);
  Config &cfg = s->cfg;
  if (cfg.oneshot) throw std::runtime_error("the server streams: INPUT FILE is not supported");
  if (cfg.gulp <= 0 || cfg.gulp > (1LL << 30) || cfg.nchan <= 0 || cfg.cand_cap <= 0)
    throw std::runtime_error("GULP, NCHAN and CAND_CAP must be positive");
  s->W = (int)cfg.gulp;
  const int maxW = cfg.width_max, minW = cfg.width_min;
  s->nw = search_nwidth(cfg.search_mode, minW, maxW);
  if (s->nw <= 0 || s->nw > SEARCH_MAXW || maxW > 1024 || (maxW & (maxW - 1)) ||
      (minW & (minW - 1)))
    throw std::runtime_error("WIDTH_MIN/WIDTH_MAX must be powers of two, maxW <= 1024");
  BeamState &g = s->geom;
  g.nch = cfg.nchan;
  g.f0 = cfg.fch1;
  g.df = cfg.foff;
  g.dt = cfg.tsamp;
  g.dm_lo = cfg.dm_min;
  g.dm_hi = cfg.dm_max;
  g.dm_tol = cfg.dm_tol;
  beam_setup(g, cfg, s->W, nullptr);
  g.ring.release();  // the template holds no window
  s->chunk_bytes = (size_t)g.in_bytes;
  s->dims = beam_dims(g, cfg, s->W, s->nw);
  std::vector<BeamState *> one{&g};
  choose_default_tile(one, s->tile, s->tile_order);
  build_plain_tiles(g, s->tile, s->tile_order);
  s->common = search_common(cfg, g, s->nw);
  if (cfg.search_mode == SEARCH_HELLA) {
    std::vector<float> all, k;
    for (int i = 0; i < s->nw; ++i) {
      hella_kernel_host(minW << i, k);
      s->kofs.push_back((int)all.size());
      all.insert(all.end(), k.begin(), k.end());
    }
    s->d_kern.ensure(all.size() * 4);
    HAMD_CHECK(hipMemcpy(s->d_kern.p, all.data(), all.size() * 4, hipMemcpyHostToDevice));
  }
  s->beam_bytes = resident_bytes(g);
}

void footprint(const BeamState &g, const WorkDims &dims, size_t chunk_bytes,
               const hella_server_opts &o, size_t &beam_bytes, size_t &work_bytes)
{
  beam_bytes = resident_bytes(g);
  const size_t per = WorkSet::bytes_for(dims) + round_up(chunk_bytes, 256);
  const size_t tables = (size_t)g.ndm * g.nlive * 4 + (size_t)g.nlive * 4 +
                        (size_t)g.nlive * g.dt_stride * 4;
  work_bytes = (size_t)o.lanes * o.max_batch * per + tables;
}

void normalise_opts(hella_server_opts &o)
{
  if (o.plain) {
    o.max_batch = 1;
    o.lanes = 1;
  }
  if (o.max_batch < 1 || o.lanes < 1 || o.max_batch > 4096 || o.lanes > 8)
    throw std::invalid_argument("max_batch must be 1..4096 and lanes 1..8");
}

void setup_lanes(hella_server s)
{
  const int nb = s->opts.max_batch;
  s->lanes.resize((size_t)s->opts.lanes);
  for (auto &L : s->lanes) {
    HAMD_CHECK(hipStreamCreateWithFlags(&L.stream, hipStreamNonBlocking));
    HAMD_CHECK(hipEventCreateWithFlags(&L.done, hipEventDisableTiming));
    L.ws.resize((size_t)nb);
    for (auto &w : L.ws) w.ensure(s->dims);
    HAMD_CHECK(hipHostMalloc((void **)&L.h_fdesc, nb * sizeof(FlagBeamDev), hipHostMallocDefault));
    HAMD_CHECK(hipHostMalloc((void **)&L.h_ddesc, nb * sizeof(DDBeamDev), hipHostMallocDefault));
    HAMD_CHECK(hipHostMalloc((void **)&L.h_sdesc, nb * sizeof(SearchBeamDev), hipHostMallocDefault));
    HAMD_CHECK(hipHostMalloc((void **)&L.h_stage, nb * sizeof(StageOut), hipHostMallocDefault));
    L.d_fdesc.ensure(nb * sizeof(FlagBeamDev));
    L.d_ddesc.ensure(nb * sizeof(DDBeamDev));
    L.d_sdesc.ensure(nb * sizeof(SearchBeamDev));
    L.d_in.ensure((size_t)nb * round_up(s->chunk_bytes, 256));
    std::vector<StageSlot> slots((size_t)nb);
    for (int i = 0; i < nb; ++i) {
      const WorkSet &w = L.ws[(size_t)i];
      slots[(size_t)i] = StageSlot{w.s_ncand.as<uint32_t>(), w.s_cands.as<Cand>(),
                                   w.c_out.as<ClusterRec>(), (int)(w.s_cands.bytes / sizeof(Cand)),
                                   (int)(w.c_out.bytes / sizeof(ClusterRec))};
    }
    L.d_slots.ensure(nb * sizeof(StageSlot));
    HAMD_CHECK(hipMemcpy(L.d_slots.p, slots.data(), nb * sizeof(StageSlot), hipMemcpyHostToDevice));
  }
  HAMD_CHECK(hipDeviceSynchronize());
}

// copy the template's geometry into a beam and give it a zeroed window
std::unique_ptr<BeamState> new_resident(hella_server s)
{
  std::unique_ptr<BeamState> B;
  if (!s->free_bs.empty()) {
    B = std::move(s->free_bs.back());
    s->free_bs.pop_back();
  } else {
    B.reset(new BeamState());
    const BeamState &g = s->geom;
    B->nch = g.nch;
    B->f0 = g.f0;
    B->df = g.df;
    B->dt = g.dt;
    B->dm_lo = g.dm_lo;
    B->dm_hi = g.dm_hi;
    B->dm_tol = g.dm_tol;
    B->plan = g.plan;
    B->delta = g.delta;
    B->live = g.live;
    B->delta_t = g.delta_t;
    B->dt_stride = g.dt_stride;
    B->ent = g.ent;
    B->ent_stride = g.ent_stride;
    B->ent_nd = g.ent_nd;
    B->nlive = g.nlive;
    B->ndm = g.ndm;
    B->md = g.md;
    B->ntime = g.ntime;
    B->ring_stride = g.ring_stride;
    B->ntime_dd = g.ntime_dd;
    B->nout = g.nout;
    B->rows = g.rows;
    B->min_gulp = g.min_gulp;
    B->dstride = g.dstride;
    B->in_bytes = g.in_bytes;
    B->spread = g.spread;
    B->ring.ensure((size_t)g.nch * g.ring_stride);
    B->f_bp.ensure((size_t)g.nch * 4);
    B->f_scal.ensure(std::max<size_t>(sizeof(FScal), 16));
    B->f_count.ensure(flag_state_bytes(g.nch));
    B->s_scal.ensure(std::max<size_t>(sizeof(SScal), 16));
  }
  // the null stream: the lanes are non-blocking streams, so this does not
  // wait for chains in flight
  hipStream_t st = nullptr;
  zero_u32(B->ring.as<uint32_t>(), (int)(B->ring.bytes / 4), st);
  zero_u32(B->f_scal.as<uint32_t>(), (int)(B->f_scal.bytes / 4), st);
  zero_u32(B->f_count.as<uint32_t>(), (int)(B->f_count.bytes / 4), st);
  zero_u32(B->s_scal.as<uint32_t>(), (int)(B->s_scal.bytes / 4), st);
  HAMD_CHECK(hipStreamSynchronize(st));
  B->gidx = 0;
  B->crossings = 0;
  return B;
}

bool try_admit(hella_server s, SBeam &sb)
{
  if (sb.B) return true;
  if (s->resident >= s->max_beams) return false;
  sb.B = new_resident(s);
  s->resident += 1;
  s->st.beams_admitted += 1;
  s->st.max_resident = std::max(s->st.max_resident, s->resident);
  s->st.peak_resident_bytes =
      std::max<unsigned long long>(s->st.peak_resident_bytes, (unsigned long long)s->resident * s->beam_bytes);
  return true;
}

void route_block(hella_server s, hella_ring ring, const hella_ring_block &blk)
{
  if (blk.bytes != s->chunk_bytes) {
    (void)hella_ring_release(ring, blk.index);
    throw std::invalid_argument("ring block size differs from the chunk size of the configuration");
  }
  SBeam &sb = s->beams[blk.beam];
  if (sb.id < 0) sb.id = blk.beam;
  Chunk ck;
  ck.seq = blk.seq;
  ck.t_arr = blk.t_arrival_ns;
  ck.last = (blk.flags & HELLA_RING_LAST) != 0;
  if (!sb.B && !try_admit(s, sb)) {
    // waiting for admission: keep a host copy so the ring keeps moving
    if (!sb.waited) s->st.beams_waited += 1;
    sb.waited = true;
    ck.copy.assign(static_cast<const uint8_t *>(blk.data),
                   static_cast<const uint8_t *>(blk.data) + blk.bytes);
    ck.data = ck.copy.data();
    if (hella_ring_release(ring, blk.index) != HELLA_OK) throw std::runtime_error("ring release failed");
  } else {
    ck.ring_index = blk.index;
    ck.data = blk.data;
  }
  sb.q.push_back(std::move(ck));
}

// beams that wait for admission, oldest first
void admit_waiting(hella_server s)
{
  std::vector<SBeam *> w;
  for (auto &kv : s->beams)
    if (!kv.second.B) w.push_back(&kv.second);
  std::sort(w.begin(), w.end(), [](const SBeam *a, const SBeam *b) {
    return a->q.front().t_arr < b->q.front().t_arr;
  });
  for (SBeam *sb : w)
    if (!try_admit(s, *sb)) break;
}

void record_flag(hella_server s, Lane &L, int n)
{
  HAMD_CHECK(hipMemcpyAsync(L.d_fdesc.p, L.h_fdesc, (size_t)n * sizeof(FlagBeamDev),
                            hipMemcpyHostToDevice, L.stream));
  FlagRunner::run(L.d_fdesc.as<FlagBeamDev>(), n, s->geom.nch, s->W, s->cfg.flag, L.stream,
                  flag_value_tables(s->cfg.input_fmt, s->cfg.flag));
}

// the counts, first candidates and first clusters of work sets 0..n-1 to the
// lane's pinned staging (stream-ordered; a graph may capture it)
void stage_results(Lane &L, int n)
{
  hipLaunchKernelGGL(k_stage_results, dim3((unsigned)n), dim3(256), 0, L.stream,
                     L.d_slots.as<StageSlot>(), L.h_stage, n);
  HAMD_CHECK(hipGetLastError());
}

void record_search(hella_server s, Lane &L, int nr, const DDTileDesc *tiles, int ntiles)
{
  const BeamState &g = s->geom;
  HAMD_CHECK(hipMemcpyAsync(L.d_ddesc.p, L.h_ddesc, (size_t)nr * sizeof(DDBeamDev),
                            hipMemcpyHostToDevice, L.stream));
  HAMD_CHECK(hipMemcpyAsync(L.d_sdesc.p, L.h_sdesc, (size_t)nr * sizeof(SearchBeamDev),
                            hipMemcpyHostToDevice, L.stream));
  if (g.nlive == 0) {
    for (int i = 0; i < nr; ++i) fill_zero_rows(L.ws[(size_t)i].D.p, g.dstride, g.ntime_dd, g.ndm, L.stream);
  } else {
    dd_launch(L.d_ddesc.as<DDBeamDev>(), tiles, ntiles, s->tile, g.spread, L.stream, g.ntime);
  }
  SearchRunner::run(L.d_sdesc.as<SearchBeamDev>(), nr, g.rows, s->common, s->d_kern.as<float>(),
                    s->kofs, L.stream);
  SearchRunner::cluster(L.d_sdesc.as<SearchBeamDev>(), nr, cluster_args(s->cfg, 0, false), L.stream);
  stage_results(L, nr);
}

template <class Rec> Graph &get_graph(Lane &L, std::map<int, Graph> &m, int key, Rec &&rec)
{
  Graph &G = m[key];
  if (G.exec) return G;
  hipGraph_t g = nullptr;
  HAMD_CHECK(hipStreamBeginCapture(L.stream, hipStreamCaptureModeThreadLocal));
  try {
    rec(G);
  } catch (...) {
    (void)hipStreamEndCapture(L.stream, &g);
    if (g) (void)hipGraphDestroy(g);
    m.erase(key);
    throw;
  }
  HAMD_CHECK(hipStreamEndCapture(L.stream, &g));
  const hipError_t e = hipGraphInstantiate(&G.exec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) {
    m.erase(key);
    throw HipError{e, "hipGraphInstantiate", __LINE__};
  }
  return G;
}

void launch_batched(hella_server s, Lane &L)
{
  const int n = (int)L.jobs.size();
  int nr = 0;
  for (const Job &j : L.jobs) nr += j.searched ? 1 : 0;
  const size_t stride = round_up(s->chunk_bytes, 256);
  for (int i = 0; i < n; ++i) {
    Job &j = L.jobs[(size_t)i];
    uint8_t *din = L.d_in.as<uint8_t>() + (size_t)i * stride;
    HAMD_CHECK(hipMemcpyAsync(din, j.ck.data, s->chunk_bytes, hipMemcpyHostToDevice, L.stream));
    BeamState &B = *j.sb->B;
    L.h_fdesc[i] = make_fdesc(s->cfg, s->W, B, L.ws[(size_t)i], din);
    if (j.searched) {
      L.h_ddesc[i] = make_ddesc(s->cfg, s->W, B, L.ws[(size_t)i]);
      L.h_sdesc[i] = make_sdesc(s->cfg, B, L.ws[(size_t)i]);
    }
  }
  Graph &gf = get_graph(L, L.flag_g, n, [&](Graph &) { record_flag(s, L, n); });
  HAMD_CHECK(hipGraphLaunch(gf.exec, L.stream));
  if (nr > 0) {
    if (!L.search_g.count(nr)) {
      // the tile table of this batch size is built and uploaded before the
      // capture starts (no synchronous copies inside a capture)
      Graph &G = L.search_g[nr];
      std::vector<DDBeamDev> dd((size_t)nr);
      for (int i = 0; i < nr; ++i) dd[(size_t)i] = L.h_ddesc[i];
      std::vector<DDTileDesc> tiles;
      dd_build_tiles(dd, s->tile, tiles, s->tile_order);
      G.d_tiles.ensure(std::max<size_t>(tiles.size(), 1) * sizeof(DDTileDesc));
      if (!tiles.empty())
        HAMD_CHECK(hipMemcpy(G.d_tiles.p, tiles.data(), tiles.size() * sizeof(DDTileDesc),
                             hipMemcpyHostToDevice));
      G.ntiles = (int)tiles.size();
    }
    Graph &gs = get_graph(L, L.search_g, nr, [&](Graph &G) {
      record_search(s, L, nr, G.d_tiles.as<DDTileDesc>(), G.ntiles);
    });
    HAMD_CHECK(hipGraphLaunch(gs.exec, L.stream));
  }
  HAMD_CHECK(hipEventRecord(L.done, L.stream));
}

// the plain per-beam sequence for one chunk (descriptor uploads, stage by
// stage, host syncs to read the candidate count and list)
void launch_plain(hella_server s, Lane &L)
{
  Job &j = L.jobs[0];
  BeamState &B = *j.sb->B;
  WorkSet &w = L.ws[0];
  uint8_t *din = L.d_in.as<uint8_t>();
  HAMD_CHECK(hipMemcpyAsync(din, j.ck.data, s->chunk_bytes, hipMemcpyHostToDevice, L.stream));
  L.h_fdesc[0] = make_fdesc(s->cfg, s->W, B, w, din);
  HAMD_CHECK(hipMemcpyAsync(L.d_fdesc.p, L.h_fdesc, sizeof(FlagBeamDev), hipMemcpyHostToDevice,
                            L.stream));
  FlagRunner::run(L.d_fdesc.as<FlagBeamDev>(), 1, B.nch, s->W, s->cfg.flag, L.stream,
                  flag_value_tables(s->cfg.input_fmt, s->cfg.flag));
  if (j.searched) {
    L.h_ddesc[0] = make_ddesc(s->cfg, s->W, B, w);
    if (B.nlive == 0) {
      fill_zero_rows(w.D.p, B.dstride, B.ntime_dd, B.ndm, L.stream);
    } else {
      HAMD_CHECK(hipMemcpyAsync(L.d_ddesc.p, L.h_ddesc, sizeof(DDBeamDev), hipMemcpyHostToDevice,
                                L.stream));
      dd_launch(L.d_ddesc.as<DDBeamDev>(), s->geom.d_tiles.as<DDTileDesc>(), s->geom.ntiles,
                s->tile, B.spread, L.stream, B.ntime);
    }
    L.h_sdesc[0] = make_sdesc(s->cfg, B, w);
    HAMD_CHECK(hipMemcpyAsync(L.d_sdesc.p, L.h_sdesc, sizeof(SearchBeamDev), hipMemcpyHostToDevice,
                              L.stream));
    SearchRunner::smooth(L.d_sdesc.as<SearchBeamDev>(), 1, B.rows, s->common, s->d_kern.as<float>(),
                         s->kofs, L.stream);
    SearchRunner::peaks(L.d_sdesc.as<SearchBeamDev>(), 1, B.rows, s->common, L.stream);
    SearchRunner::cluster(L.d_sdesc.as<SearchBeamDev>(), 1, cluster_args(s->cfg, 0, false), L.stream);
    stage_results(L, 1);
  }
  HAMD_CHECK(hipEventRecord(L.done, L.stream));
}

void harvest(hella_server s, hella_ring ring, Lane &L)
{
  HAMD_CHECK(hipStreamSynchronize(L.stream));
  const long long nb = (long long)L.jobs.size();
  for (size_t i = 0; i < L.jobs.size(); ++i) {
    Job &j = L.jobs[i];
    SBeam &sb = *j.sb;
    BeamState &B = *sb.B;
    Result res;
    res.r.beam = sb.id;
    res.r.seq = j.ck.seq;
    res.r.searched = j.searched ? 1 : 0;
    res.r.batch = (int)nb;
    res.r.t_arrival_ns = j.ck.t_arr;
    if (j.searched) {
      const StageOut &so = L.h_stage[i];
      const uint32_t nc = so.n[0], ncl = so.n[1];
      size_t take = std::min<size_t>(nc, (size_t)s->cfg.cand_cap);
      take = std::min<size_t>(take, (size_t)s->cfg.max_giants);
      std::vector<Cand> spill_c;
      std::vector<ClusterRec> spill_k;
      const Cand *cs = results_view<Cand>(so.c, SRV_KC, L.ws[i].s_cands.as<Cand>(), take, spill_c,
                                          L.stream);
      const ClusterRec *cl = results_view<ClusterRec>(so.k, SRV_KK, L.ws[i].c_out.as<ClusterRec>(),
                                                      ncl, spill_k, L.stream);
      B.crossings = nc;
      res.r.crossings = nc;
      append_candidates(s->cfg, s->W, B, sb.id, cs, take, (size_t)s->cfg.max_giants, res.cands);
      append_clusters(cl, ncl, 0, sb.id, res.clusters);
    }
    res.r.ncand = (int)res.cands.size();
    res.r.nclusters = (int)res.clusters.size();
    res.r.t_done_ns = now_ns();
    s->results.push_back(std::move(res));
    if (j.ck.ring_index >= 0) (void)hella_ring_release(ring, j.ck.ring_index);
    B.gidx += 1;
    sb.in_flight = false;
    s->st.chunks += 1;
    if (j.ck.last) {
      // the beam leaves: its window returns to the pool; chunks that already
      // arrived under the same id belong to a new beam, admitted afresh
      s->free_bs.push_back(std::move(sb.B));
      s->resident -= 1;
      if (sb.q.empty()) s->beams.erase(sb.id);
    }
  }
  L.jobs.clear();
  L.busy = false;
}

// pick up to max_batch beams with a queued chunk and none in flight, oldest
// chunk first; beams whose window is full come first in the batch
bool dispatch(hella_server s, Lane &L)
{
  std::vector<SBeam *> c;
  for (auto &kv : s->beams) {
    SBeam &sb = kv.second;
    if (sb.B && !sb.in_flight && !sb.q.empty()) c.push_back(&sb);
  }
  if (c.empty()) return false;
  std::sort(c.begin(), c.end(), [](const SBeam *a, const SBeam *b) {
    if (a->q.front().t_arr != b->q.front().t_arr) return a->q.front().t_arr < b->q.front().t_arr;
    return a->id < b->id;
  });
  if ((int)c.size() > s->opts.max_batch) c.resize((size_t)s->opts.max_batch);
  std::stable_partition(c.begin(), c.end(),
                        [&](const SBeam *b) { return beam_ready(s->cfg, *b->B); });
  L.jobs.clear();
  for (SBeam *sb : c) {
    Job j;
    j.sb = sb;
    j.ck = std::move(sb->q.front());
    sb->q.pop_front();
    j.searched = beam_ready(s->cfg, *sb->B);
    sb->in_flight = true;
    L.jobs.push_back(std::move(j));
  }
  L.busy = true;
  L.t_launch = now_ns();
  if (s->opts.plain)
    launch_plain(s, L);
  else
    launch_batched(s, L);
  s->st.batches += 1;
  return true;
}

bool work_left(hella_server s)
{
  for (auto &kv : s->beams)
    if (!kv.second.q.empty() || kv.second.in_flight) return true;
  for (auto &L : s->lanes)
    if (L.busy) return true;
  return false;
}

void serve(hella_server s, hella_ring ring)
{
  bool ended = false;
  long long busy_since = -1;
  double busy_s = 0;
  const long long t0 = now_ns();
  auto busy_lanes = [&]() {
    int k = 0;
    for (auto &L : s->lanes) k += L.busy ? 1 : 0;
    return k;
  };
  auto note_busy = [&]() {
    const int k = busy_lanes();
    const long long t = now_ns();
    if (k > 0 && busy_since < 0) busy_since = t;
    if (k == 0 && busy_since >= 0) {
      busy_s += (double)(t - busy_since) * 1e-9;
      busy_since = -1;
    }
  };
  for (;;) {
    // ingest every block that has arrived
    for (;;) {
      hella_ring_block blk;
      if (hella_ring_open_read(ring, 0, &blk) != HELLA_OK) throw std::runtime_error("ring read failed");
      if (blk.index < 0) {
        if (blk.flags & HELLA_RING_END) ended = true;
        break;
      }
      route_block(s, ring, blk);
    }
    // collect finished launch chains
    for (auto &L : s->lanes)
      if (L.busy && hipEventQuery(L.done) == hipSuccess) harvest(s, ring, L);
    note_busy();
    admit_waiting(s);
    // launch on every free lane
    bool launched = false;
    for (auto &L : s->lanes)
      if (!L.busy && dispatch(s, L)) launched = true;
    note_busy();
    if (ended && !work_left(s)) break;
    if (launched) continue;
    const int k = busy_lanes();
    if (k == (int)s->lanes.size()) {
      // every lane is busy: wait for the oldest
      Lane *o = nullptr;
      for (auto &L : s->lanes)
        if (L.busy && (!o || L.t_launch < o->t_launch)) o = &L;
      HAMD_CHECK(hipEventSynchronize(o->done));
      continue;
    }
    if (k > 0) {
      // a lane is free but nothing is queued: poll the ring briefly
      hella_ring_block blk;
      if (hella_ring_open_read(ring, 1, &blk) != HELLA_OK) throw std::runtime_error("ring read failed");
      if (blk.index >= 0) route_block(s, ring, blk);
      else if (blk.flags & HELLA_RING_END) ended = true;
      continue;
    }
    if (ended) {
      if (!work_left(s)) break;
      // queued chunks of beams still waiting for admission cannot proceed
      bool any_resident_work = false;
      for (auto &kv : s->beams)
        if (kv.second.B && !kv.second.q.empty()) any_resident_work = true;
      if (!any_resident_work) throw std::runtime_error("beams wait for HBM that never frees");
      continue;
    }
    hella_ring_block blk;
    if (hella_ring_open_read(ring, 20, &blk) != HELLA_OK) throw std::runtime_error("ring read failed");
    if (blk.index >= 0) route_block(s, ring, blk);
    else if (blk.flags & HELLA_RING_END) ended = true;
  }
  note_busy();
  s->st.busy_s += busy_s;
  s->st.wall_s += (double)(now_ns() - t0) * 1e-9;
}

}  // namespace

extern "C" {

hella_status hella_server_footprint(const char *config, const hella_server_opts *opts,
                                    unsigned long long *beam_bytes, unsigned long long *work_bytes)
{
  if (!opts || !beam_bytes || !work_bytes) return sfail(HELLA_ERR_INVALID, "bad argument");
  hella_server_s tmp;
  tmp.opts = *opts;
  return sguard([&] {
    normalise_opts(tmp.opts);
    setup_geometry(&tmp, config);
    size_t bb = 0, wb = 0;
    footprint(tmp.geom, tmp.dims, tmp.chunk_bytes, tmp.opts, bb, wb);
    *beam_bytes = bb;
    *work_bytes = wb;
  });
}

hella_status hella_server_create(const char *config, const hella_server_opts *opts,
                                 hella_server *out)
{
  if (!out || !opts) return sfail(HELLA_ERR_INVALID, "bad argument");
  *out = nullptr;
  auto *s = new (std::nothrow) hella_server_s();
  if (!s) return sfail(HELLA_ERR_ALLOC, "server allocation failed");
  s->opts = *opts;
  const hella_status st = sguard([&] {
    normalise_opts(s->opts);
    int n = 0;
    HAMD_CHECK(hipGetDeviceCount(&n));
    std::string err;
    Config probe;
    if (!probe.parse(config ? config : "", err)) throw std::runtime_error(err);
    if (probe.gpu < 0 || probe.gpu >= n) throw std::runtime_error("GPU index out of range");
    HAMD_CHECK(hipSetDevice(probe.gpu));
    kernels_setup();
    setup_geometry(s, config);
    footprint(s->geom, s->dims, s->chunk_bytes, s->opts, s->beam_bytes, s->work_bytes);
    if (s->opts.hbm_budget == 0) {
      s->max_beams = 1 << 30;
    } else {
      if (s->opts.hbm_budget < s->work_bytes + s->beam_bytes)
        throw std::invalid_argument("HBM budget below the work sets plus one beam");
      s->max_beams = (int)((s->opts.hbm_budget - s->work_bytes) / s->beam_bytes);
    }
    setup_lanes(s);
    s->st.work_bytes = s->work_bytes;
  });
  if (st != HELLA_OK) {
    delete s;
    return st;
  }
  *out = s;
  return HELLA_OK;
}

void hella_server_destroy(hella_server s) { delete s; }

hella_status hella_server_run(hella_server s, hella_ring ring)
{
  if (!s || !ring) return sfail(HELLA_ERR_INVALID, "bad argument");
  const hella_status st = sguard([&] { serve(s, ring); });
  if (st != HELLA_OK) (void)hella_ring_abort(ring);
  return st;
}

int hella_server_result_count(hella_server s) { return s ? (int)s->results.size() : 0; }

hella_status hella_server_get_result(hella_server s, int i, hella_chunk_result *out)
{
  if (!s || !out || i < 0 || (size_t)i >= s->results.size()) return sfail(HELLA_ERR_INVALID, "bad index");
  *out = s->results[(size_t)i].r;
  return HELLA_OK;
}

int hella_server_get_candidates(hella_server s, int i, hella_candidate *out, int cap)
{
  if (!s || i < 0 || (size_t)i >= s->results.size()) return -1;
  const auto &v = s->results[(size_t)i].cands;
  const int n = (int)v.size();
  if (out)
    for (int k = 0; k < std::min(n, cap); ++k) out[k] = v[(size_t)k];
  return n;
}

int hella_server_get_clusters(hella_server s, int i, hella_cluster *out, int cap)
{
  if (!s || i < 0 || (size_t)i >= s->results.size()) return -1;
  const auto &v = s->results[(size_t)i].clusters;
  const int n = (int)v.size();
  if (out)
    for (int k = 0; k < std::min(n, cap); ++k) out[k] = v[(size_t)k];
  return n;
}

hella_status hella_server_get_stats(hella_server s, hella_server_stats *out)
{
  if (!s || !out) return sfail(HELLA_ERR_INVALID, "bad argument");
  *out = s->st;
  return HELLA_OK;
}

}  // extern "C"
