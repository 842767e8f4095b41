// capi.cpp -- the hella_* C ABI (include/hella.h).  Every entry point
// catches the library's exceptions and returns a status code; the message
// is kept per thread for hella_last_error().
#include <cstring>
#include <new>
#include <stdexcept>

#include "pipeline.hpp"

using namespace hamd;

namespace {

hella_status fail(hella_status s, const std::string &msg)
{
  set_last_error(msg);
  return s;
}

template <class F> hella_status guard(F &&f)
{
  try {
    f();
    return HELLA_OK;
  } catch (const HipError &e) {
    return fail(e.err == hipErrorOutOfMemory ? HELLA_ERR_ALLOC : HELLA_ERR_GPU,
                std::string(e.what) + ": " + hipGetErrorString(e.err));
  } catch (const std::bad_alloc &) {
    return fail(HELLA_ERR_ALLOC, "host allocation failed");
  } catch (const std::invalid_argument &e) {
    return fail(HELLA_ERR_INVALID, e.what());
  } catch (const std::runtime_error &e) {
    return fail(HELLA_ERR_CONFIG, e.what());
  } catch (...) {
    return fail(HELLA_ERR_GPU, "unknown error");
  }
}

bool valid_beam(hella_ctx c, int beam) { return c && beam >= 0 && (size_t)beam < c->beams.size(); }

}  // namespace

extern "C" {

const char *hella_last_error(void) { return last_error(); }

hella_status hella_create(const char *config, hella_ctx *out)
{
  if (!out) return fail(HELLA_ERR_INVALID, "out is NULL");
  *out = nullptr;
  auto *c = new (std::nothrow) hella_ctx_s();
  if (!c) return fail(HELLA_ERR_ALLOC, "context allocation failed");
  std::string err;
  if (!c->cfg.parse(config ? config : "", err)) {
    delete c;
    return fail(HELLA_ERR_CONFIG, err);
  }
  const Config &cfg = c->cfg;
  if (cfg.nbeam <= 0 || cfg.gulp <= 0 || cfg.nchan <= 0 || cfg.max_giants < 0 || cfg.cand_cap <= 0 ||
      cfg.gulp > (1LL << 30)) {
    delete c;
    return fail(HELLA_ERR_CONFIG, "NBEAM, GULP, NCHAN and CAND_CAP must be positive");
  }
  c->beams.resize((size_t)cfg.nbeam);
  for (int b = 0; b < cfg.nbeam; ++b) {
    BeamState &B = c->beams[(size_t)b];
    B.id = b;
    B.nch = cfg.nchan;
    B.f0 = cfg.fch1;
    B.df = cfg.foff;
    B.dt = cfg.tsamp;
    B.dm_lo = cfg.dm_min;
    B.dm_hi = cfg.dm_max;
    B.dm_tol = cfg.dm_tol;
  }
  hella_status st = guard([&] {
    int n = 0;
    HAMD_CHECK(hipGetDeviceCount(&n));
    if (cfg.gpu < 0 || cfg.gpu >= n) throw std::runtime_error("GPU index out of range");
    HAMD_CHECK(hipSetDevice(cfg.gpu));
    kernels_setup();
    HAMD_CHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  });
  if (st != HELLA_OK) {
    delete c;
    return st;
  }
  *out = c;
  return HELLA_OK;
}

void hella_destroy(hella_ctx c)
{
  if (!c) return;
  (void)hipStreamSynchronize(c->stream);
  if (c->batch.exec) (void)hipGraphExecDestroy(c->batch.exec);
  if (c->h_fdesc) (void)hipHostFree(c->h_fdesc);
  if (c->h_ddesc) (void)hipHostFree(c->h_ddesc);
  if (c->h_sdesc) (void)hipHostFree(c->h_sdesc);
  if (c->h_ncand) (void)hipHostFree(c->h_ncand);
  if (c->h_cstage) (void)hipHostFree(c->h_cstage);
  if (c->h_kstage) (void)hipHostFree(c->h_kstage);
  hipStream_t s = c->stream;
  delete c;
  if (s) (void)hipStreamDestroy(s);
}

hella_status hella_set_beam(hella_ctx c, int beam, const char *overrides)
{
  if (!valid_beam(c, beam)) return fail(HELLA_ERR_INVALID, "bad beam");
  if (c->built) return fail(HELLA_ERR_STATE, "beams are fixed after the first gulp");
  Config tmp = c->cfg;
  std::string err;
  if (!tmp.parse(overrides ? overrides : "", err, true)) return fail(HELLA_ERR_CONFIG, err);
  BeamState &B = c->beams[(size_t)beam];
  B.nch = tmp.nchan;
  B.f0 = tmp.fch1;
  B.df = tmp.foff;
  B.dt = tmp.tsamp;
  B.dm_lo = tmp.dm_min;
  B.dm_hi = tmp.dm_max;
  B.dm_tol = tmp.dm_tol;
  if (B.nch <= 0) return fail(HELLA_ERR_CONFIG, "NCHAN must be positive");
  B.killmask.clear();
  return HELLA_OK;
}

hella_status hella_set_killmask(hella_ctx c, int beam, const int *killmask)
{
  if (!valid_beam(c, beam)) return fail(HELLA_ERR_INVALID, "bad beam");
  if (c->built) return fail(HELLA_ERR_STATE, "killmask is fixed after the first gulp");
  BeamState &B = c->beams[(size_t)beam];
  B.killmask.assign((size_t)B.nch, 1);
  if (killmask)
    for (int i = 0; i < B.nch; ++i) B.killmask[(size_t)i] = killmask[i] != 0;
  return HELLA_OK;
}

hella_status hella_get_info(hella_ctx c, int beam, hella_info *out)
{
  if (!valid_beam(c, beam) || !out) return fail(HELLA_ERR_INVALID, "bad argument");
  return guard([&] {
    ctx_build(c);
    const BeamState &B = c->beams[(size_t)beam];
    out->nchan = B.nch;
    out->ndm = B.ndm;
    out->max_delay = B.md;
    out->ntime = B.ntime;
    out->ntime_dd = B.ntime_dd;
    out->ntime_out = B.nout;
    out->nwidth = c->nw;
    out->gulp = c->W;
    out->min_gulp = B.min_gulp;
    out->gulp_index = c->gulp_index;
    out->rows = B.rows;
    out->nlive = B.nlive;
  });
}

hella_status hella_get_dm_list(hella_ctx c, int beam, float *out, int cap, int *n)
{
  if (!valid_beam(c, beam)) return fail(HELLA_ERR_INVALID, "bad beam");
  return guard([&] {
    ctx_build(c);
    const BeamState &B = c->beams[(size_t)beam];
    if (n) *n = B.ndm;
    if (out)
      for (int i = 0; i < B.ndm && i < cap; ++i) out[i] = B.plan.dm_list[(size_t)i];
  });
}

void *hella_get_stream(hella_ctx c) { return c ? (void *)c->stream : nullptr; }

hella_status hella_fast_flagger(hella_ctx c, const void *gulp, int host_ptr)
{
  if (!c || !gulp) return fail(HELLA_ERR_INVALID, "bad argument");
  return guard([&] {
    ctx_build(c);
    c->batch_mode = false;
    c->peaks.clear();
    c->clusters.clear();
    const void *d = ctx_stage_input(c, gulp, host_ptr != 0);
    for (size_t b = 0; b < c->beams.size(); ++b) ctx_flag_beam(c, (int)b, d);
  });
}

hella_status hella_dedisperse(hella_ctx c, int beam)
{
  if (!valid_beam(c, beam)) return fail(HELLA_ERR_INVALID, "bad beam");
  if (!c->built) return fail(HELLA_ERR_STATE, "no gulp flagged yet");
  const BeamState &B = c->beams[(size_t)beam];
  if (!beam_ready(c->cfg, B)) return fail(HELLA_ERR_STATE, "the beam's window is not full yet");
  return guard([&] { ctx_dedisperse(c, beam); });
}

hella_status hella_smooth(hella_ctx c, int scale)
{
  if (!c || c->cur_beam < 0) return fail(HELLA_ERR_STATE, "nothing dedispersed");
  if (scale != 1) return fail(HELLA_ERR_INVALID, "only scale=1 (normalised widths) is supported");
  return guard([&] { ctx_smooth(c, c->cur_beam); });
}

hella_status hella_find_peaks(hella_ctx c, int beam)
{
  if (!valid_beam(c, beam) || beam != c->cur_beam) return fail(HELLA_ERR_STATE, "beam not smoothed");
  return guard([&] { ctx_find_peaks(c, beam); });
}

hella_status hella_clear_peaks(hella_ctx c)
{
  if (!c) return fail(HELLA_ERR_INVALID, "ctx is NULL");
  c->peaks.clear();
  c->clusters.clear();
  return HELLA_OK;
}

hella_status hella_end_gulp(hella_ctx c)
{
  if (!c || !c->built) return fail(HELLA_ERR_STATE, "no gulp");
  return guard([&] {
    HAMD_CHECK(hipStreamSynchronize(c->stream));
    ctx_end_gulp(c);
  });
}

hella_status hella_process_gulp(hella_ctx c, const void *gulp, int host_ptr, int *ncand)
{
  if (!c || !gulp) return fail(HELLA_ERR_INVALID, "bad argument");
  return guard([&] {
    ctx_build(c);
    ctx_process_plain(c, gulp, host_ptr != 0);
    if (ncand) *ncand = (int)c->peaks.size();
  });
}

hella_status hella_process_gulp_batched(hella_ctx c, const void *gulp, int host_ptr, int *ncand)
{
  if (!c || !gulp) return fail(HELLA_ERR_INVALID, "bad argument");
  return guard([&] {
    ctx_build(c);
    ctx_process_batched(c, gulp, host_ptr != 0);
    if (ncand) *ncand = (int)c->peaks.size();
  });
}

int hella_get_peaks(hella_ctx c, hella_candidate *out, int cap)
{
  if (!c) return 0;
  const int n = (int)c->peaks.size();
  if (out)
    for (int i = 0; i < n && i < cap; ++i) out[i] = c->peaks[(size_t)i];
  return n;
}

int hella_get_clusters(hella_ctx c, hella_cluster *out, int cap)
{
  if (!c) return 0;
  const int n = (int)c->clusters.size();
  if (out)
    for (int i = 0; i < n && i < cap; ++i) out[i] = c->clusters[(size_t)i];
  return n;
}

int hella_get_flag_counts(hella_ctx c, int beam, unsigned *out, int cap)
{
  if (!valid_beam(c, beam) || cap < 0 || (cap > 0 && !out)) {
    fail(HELLA_ERR_INVALID, "bad argument");
    return -1;
  }
  int n = -1;
  const hella_status st = guard([&] {
    ctx_build(c);
    const BeamState &B = c->beams[(size_t)beam];
    n = B.nch;
    if (cap > 0) {
      HAMD_CHECK(hipStreamSynchronize(c->stream));
      HAMD_CHECK(hipMemcpy(out, B.f_count.p, (size_t)(cap < n ? cap : n) * 4, hipMemcpyDeviceToHost));
    }
  });
  return st == HELLA_OK ? n : -1;
}

int hella_get_flag_spectrum(hella_ctx c, int beam, float *out, int cap)
{
  if (!valid_beam(c, beam) || cap < 0 || (cap > 0 && !out)) {
    fail(HELLA_ERR_INVALID, "bad argument");
    return -1;
  }
  int n = -1;
  const hella_status st = guard([&] {
    ctx_build(c);
    const BeamState &B = c->beams[(size_t)beam];
    n = B.nch;
    if (cap > 0) {
      HAMD_CHECK(hipStreamSynchronize(c->stream));
      HAMD_CHECK(hipMemcpy(out, B.f_count.as<uint8_t>() + flag_state_spec_offset(B.nch),
                           (size_t)(cap < n ? cap : n) * 4, hipMemcpyDeviceToHost));
    }
  });
  return st == HELLA_OK ? n : -1;
}

hella_status hella_accumulate_specflags(hella_ctx c, int *specflags, int nchan)
{
  if (!c || nchan <= 0 || !specflags) return fail(HELLA_ERR_INVALID, "bad argument");
  return guard([&] {
    ctx_build(c);
    std::vector<double> sum((size_t)nchan, 0.0);
    std::vector<float> one((size_t)nchan);
    long long ntime = -1;
    HAMD_CHECK(hipStreamSynchronize(c->stream));
    for (const BeamState &B : c->beams) {
      if (B.nch != nchan) continue;
      if (ntime < 0) ntime = B.ntime;
      HAMD_CHECK(hipMemcpy(one.data(), B.f_count.as<uint8_t>() + flag_state_spec_offset(B.nch),
                           (size_t)nchan * 4, hipMemcpyDeviceToHost));
      for (int i = 0; i < nchan; ++i) sum[(size_t)i] += (double)one[(size_t)i];
    }
    if (ntime < 0) throw std::invalid_argument("no beam has that many channels");
    for (int i = 0; i < nchan; ++i)
      specflags[i] += (int)((double)ntime * sum[(size_t)i]);
  });
}

long long hella_last_crossings(hella_ctx c, int beam)
{
  if (!valid_beam(c, beam)) return -1;
  return c->beams[(size_t)beam].crossings;
}

hella_status hella_copy_window(hella_ctx c, int beam, void *host, size_t bytes)
{
  if (!valid_beam(c, beam) || !host) return fail(HELLA_ERR_INVALID, "bad argument");
  return guard([&] {
    ctx_build(c);
    const BeamState &B = c->beams[(size_t)beam];
    const size_t need = (size_t)B.nch * (size_t)B.ntime;
    if (bytes < need) throw std::invalid_argument("buffer too small");
    std::vector<uint8_t> raw((size_t)B.nch * B.ring_stride);
    HAMD_CHECK(hipStreamSynchronize(c->stream));
    HAMD_CHECK(hipMemcpy(raw.data(), B.ring.p, raw.size(), hipMemcpyDeviceToHost));
    // oldest sample first: the window of the last ingested gulp
    const long long g = c->gulp_index;  // gulps ingested
    const int64_t start = c->cfg.oneshot ? 0 : (int64_t)((g * (long long)c->W) % B.ntime);
    uint8_t *o = static_cast<uint8_t *>(host);
    const size_t first = (size_t)(B.ntime - start);  // columns start .. ntime-1, then 0 ..
    for (int ch = 0; ch < B.nch; ++ch) {
      const uint8_t *r = raw.data() + (size_t)ch * B.ring_stride;
      uint8_t *d = o + (size_t)ch * B.ntime;
      std::memcpy(d, r + start, first);
      std::memcpy(d + first, r, (size_t)start);
    }
  });
}

hella_status hella_copy_dedisp(hella_ctx c, int beam, void *host, size_t bytes)
{
  if (!valid_beam(c, beam) || !host) return fail(HELLA_ERR_INVALID, "bad argument");
  return guard([&] {
    const BeamState &B = c->beams[(size_t)beam];
    const void *src = c->batch_mode ? (c->ws.size() > (size_t)beam ? c->ws[(size_t)beam].D.p : nullptr)
                                    : (c->cur_beam == beam ? c->ws[0].D.p : nullptr);
    if (!src) throw std::invalid_argument("no dedispersed plane for this beam");
    const size_t need = (size_t)B.ndm * B.ntime_dd * 4;
    if (bytes < need) throw std::invalid_argument("buffer too small");
    HAMD_CHECK(hipStreamSynchronize(c->stream));
    HAMD_CHECK(hipMemcpy2D(host, (size_t)B.ntime_dd * 4, src, (size_t)B.dstride * 4,
                           (size_t)B.ntime_dd * 4, (size_t)B.ndm, hipMemcpyDeviceToHost));
  });
}

hella_status hella_set_dedisp(hella_ctx c, int beam, const void *host, size_t bytes)
{
  if (!valid_beam(c, beam) || !host) return fail(HELLA_ERR_INVALID, "bad argument");
  return guard([&] {
    ctx_build(c);
    BeamState &B = c->beams[(size_t)beam];
    const size_t need = (size_t)B.ndm * B.ntime_dd * 4;
    if (bytes < need) throw std::invalid_argument("buffer smaller than ndm x ntime_dd u32");
    c->batch_mode = false;
    HAMD_CHECK(hipStreamSynchronize(c->stream));
    HAMD_CHECK(hipMemcpy2D(c->ws[0].D.p, (size_t)B.dstride * 4, host, (size_t)B.ntime_dd * 4,
                           (size_t)B.ntime_dd * 4, (size_t)B.ndm, hipMemcpyHostToDevice));
    c->cur_beam = beam;
  });
}

hella_status hella_get_search_stats(hella_ctx c, int beam, float *mu, float *sigma, float *sigma0,
                                    int *gated)
{
  if (!valid_beam(c, beam)) return fail(HELLA_ERR_INVALID, "bad beam");
  return guard([&] {
    const BeamState &B = c->beams[(size_t)beam];
    SScal sc;
    HAMD_CHECK(hipStreamSynchronize(c->stream));
    HAMD_CHECK(hipMemcpy(&sc, B.s_scal.p, sizeof(SScal), hipMemcpyDeviceToHost));
    for (int k = 0; k < c->nw; ++k) {
      if (mu) mu[k] = (float)sc.mu[k];
      if (sigma) sigma[k] = sc.sigf[k];
    }
    if (sigma0) *sigma0 = sc.sig0;
    if (gated) *gated = sc.gated;
  });
}

hella_status hella_get_search_norm(hella_ctx c, int beam, double *mu, double *inv, float *thr)
{
  if (!valid_beam(c, beam)) return fail(HELLA_ERR_INVALID, "bad beam");
  return guard([&] {
    const BeamState &B = c->beams[(size_t)beam];
    SScal sc;
    HAMD_CHECK(hipStreamSynchronize(c->stream));
    HAMD_CHECK(hipMemcpy(&sc, B.s_scal.p, sizeof(SScal), hipMemcpyDeviceToHost));
    for (int k = 0; k < c->nw; ++k) {
      if (mu) mu[k] = sc.mu[k];
      if (inv) inv[k] = sc.inv[k];
    }
    if (thr) *thr = sc.thr;
  });
}

hella_status hella_stage_times(hella_ctx c, float *ms)
{
  if (!c || !ms) return fail(HELLA_ERR_INVALID, "bad argument");
  for (int i = 0; i < 4; ++i) ms[i] = c->stage_ms[(size_t)i];
  return HELLA_OK;
}

hella_status hella_set_timing(hella_ctx c, int on)
{
  if (!c) return fail(HELLA_ERR_INVALID, "ctx is NULL");
  c->timing = on != 0;
  return HELLA_OK;
}

hella_status hella_bench_stage(hella_ctx c, int beam, int which, int reps, float *ms)
{
  if (!valid_beam(c, beam) || !ms || reps <= 0 || which < 0 || which > 2)
    return fail(HELLA_ERR_INVALID, "bad argument");
  if (!c->built || c->batch_mode) return fail(HELLA_ERR_STATE, "run a plain gulp first");
  return guard([&] { *ms = ctx_bench_stage(c, beam, which, reps); });
}

hella_status hella_autotune(hella_ctx c, const char *path)
{
  if (!c) return fail(HELLA_ERR_INVALID, "ctx is NULL");
  return guard([&] { ctx_autotune(c, path); });
}

hella_status hella_set_tile(hella_ctx c, int lanes_per_dm, int chunk, int order)
{
  if (!c) return fail(HELLA_ERR_INVALID, "ctx is NULL");
  if (c->built) return fail(HELLA_ERR_STATE, "set the tile before the first gulp");
  // order bits 1-2: loader (0 LDS DMA, 1 registers, 2 register prefetch,
  // 3 one DM trial per wave: 64 lanes, chunk = 4 or 8 waves per workgroup);
  // bit 3 with loader 3: a chunk's LDS reads all issued before its sums;
  // bit 4 with loader 3 and chunk 4: windows copied by LDS DMA two chunks ahead
  const int loader = (order >> 1) & 3;
  const bool pipe = ((order >> 3) & 1) != 0;
  const bool dma2 = ((order >> 4) & 1) != 0;
  const bool ok = loader == 3
                      ? lanes_per_dm == 64 && (chunk == 4 || chunk == 8)
                      : (lanes_per_dm == 8 || lanes_per_dm == 16 || lanes_per_dm == 32 ||
                         lanes_per_dm == 64) &&
                            (chunk == 8 || chunk == 16 || chunk == 32) && !(loader == 2 && chunk == 32);
  if (!ok || order < 0 || order > 31 || (pipe && loader != 3) ||
      (dma2 && (loader != 3 || pipe || chunk != 4)))
    return fail(HELLA_ERR_INVALID, "tile shape out of range");
  c->tile = DDTile{lanes_per_dm, chunk,
                   loader == 0 ? 1 : loader == 1 ? 0 : dma2 ? 5 : pipe ? 4 : loader};
  c->tile_order = order & 1;
  c->tile_forced = true;
  return HELLA_OK;
}

int hella_device_count(void)
{
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

void *hella_dev_alloc(size_t bytes)
{
  void *p = nullptr;
  if (hipMalloc(&p, bytes) != hipSuccess) return nullptr;
  return p;
}

void hella_dev_free(void *p)
{
  if (p) (void)hipFree(p);
}

hella_status hella_memcpy(void *dst, const void *src, size_t bytes, int kind)
{
  hipMemcpyKind k = kind == 1 ? hipMemcpyHostToDevice
                  : kind == 2 ? hipMemcpyDeviceToHost
                  : kind == 3 ? hipMemcpyDeviceToDevice
                              : hipMemcpyDefault;
  if (!dst || !src) return fail(HELLA_ERR_INVALID, "NULL pointer");
  return guard([&] { HAMD_CHECK(hipMemcpy(dst, src, bytes, k)); });
}

hella_status hella_device_sync(void)
{
  return guard([&] { HAMD_CHECK(hipDeviceSynchronize()); });
}

}  // extern "C"
