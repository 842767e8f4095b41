// dedisp_api.cpp -- the dedisp-compatible C ABI (include/dedisp.h).
//
// Replaces the external dedisp library at the call sites
// src/apps/casm_hella_refactored.cu:205-210 and
// src/hella/src/dedisperse.cpp:32-48.  Errors are returned as dedisp_error
// values (the reference turns them into std::runtime_error at
// dedisperse.cpp:40-46); nothing here exits the process.
#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>

#include "../../include/dedisp.h"
#include "internal.hpp"

using namespace hamd;

struct dedisp_plan_struct {
  PlanHost h;
  hipStream_t stream = nullptr;
  std::vector<dedisp_bool> killmask_pub;
  // host tables (delays of the live channels, tile shape, channel-major
  // delays, entry tables), rebuilt by the DM-list and killmask setters, and
  // their device copies, uploaded by the first execute after a change
  bool host_ready = false;
  bool dirty = true;
  DevBuf d_delta, d_live, d_delta_t, d_ent;
  int dt_stride = 0;
  int ent_stride = 0, ent_nd = 0;  // entry tables of the one-trial-per-wave kernel
  int nlive = 0;
  std::vector<int32_t> delta_host, live_host, dt_host;
  std::vector<uint32_t> ent_host;
  DDBeamDev na_desc{};  // descriptor of the last scratch-only execute (host copy)
  DDTile tile{16, 32, 1};
  int max_spread = 0;
  DevBuf scratch_in, scratch_ft, scratch_out;
  DevBuf d_beam, d_tiles;
  std::vector<DDTileDesc> tiles;
  size_t tiles_nout = 0, tiles_first = 0, tiles_count = 0;
};

namespace hamd {
// Tile choice shared with the pipeline: the largest channel chunk whose LDS
// footprint fits the budget, shrinking the DM block when even 8 channels do
// not fit.
DDTile choose_dd_tile(const int32_t *delta, int ndm, int nlive, int *spread_out,
                      int lanes_pref, size_t lds_budget)
{
  // measured on MI355X: one DM trial per wave in workgroups of 4 trials,
  // every LDS read of a chunk issued before its sums, is fastest, then 32
  // lanes per trial with 8-channel chunks staged through registers; other
  // lane counts when the spread does not fit
  const DDTile w{64, 4, 4};
  if (lanes_pref <= 64) {
    const int spread = dd_max_spread(delta, ndm, nlive, dd_tile_ND(w));
    if (dd_tile_fits(w, spread, INT64_MAX)) {
      *spread_out = spread;
      return w;
    }
  }
  static const int gs[] = {32, 64, 16, 8};
  for (int G : gs) {
    if (G < lanes_pref) continue;
    const DDTile t{G, 8, 0};
    const int spread = dd_max_spread(delta, ndm, nlive, dd_tile_ND(t));
    if (dd_lds_bytes(t, spread) <= lds_budget) {
      *spread_out = spread;
      return t;
    }
  }
  *spread_out = dd_max_spread(delta, ndm, nlive, 4);
  return DDTile{64, 8, 0};
}
}  // namespace hamd

namespace {

// Host tables of the current DM list and killmask (host memory only; the
// setters call it, so an execute finds them ready).
void prepare_host(dedisp_plan plan)
{
  // the tile table depends on the tile shape chosen here (DM block, time
  // step), which a new DM list or killmask may change: rebuild it
  plan->tiles.clear();
  plan->h.delays_live(0, plan->h.dm_list.size(), plan->delta_host, plan->live_host);
  plan->nlive = (int)plan->live_host.size();
  const int ndm = (int)plan->h.dm_list.size();
  plan->dt_host.clear();
  plan->ent_host.clear();
  plan->ent_nd = 0;
  if (plan->nlive > 0) {
    plan->tile = choose_dd_tile(plan->delta_host.data(), ndm, plan->nlive, &plan->max_spread,
                                16, DD_LDS_BUDGET);
    delays_transposed(plan->delta_host, ndm, plan->nlive, plan->dt_host, plan->dt_stride);
    if (plan->tile.dma >= 3) {
      plan->ent_nd = dd_tile_ND(plan->tile);
      dd_entries(plan->delta_host, ndm, plan->nlive, plan->ent_nd, plan->ent_host,
                 plan->ent_stride);
    }
  }
  plan->host_ready = true;
  plan->dirty = true;
}

dedisp_error refresh_device(dedisp_plan plan)
{
  if (!plan->host_ready) prepare_host(plan);
  if (!plan->dirty) return DEDISP_NO_ERROR;
  if (plan->nlive > 0) {
    plan->d_delta.ensure((plan->delta_host.size() + DD_DELTA_PAD) * 4);
    plan->d_live.ensure(plan->live_host.size() * 4);
    HAMD_CHECK(hipMemcpy(plan->d_delta.p, plan->delta_host.data(), plan->delta_host.size() * 4,
                         hipMemcpyHostToDevice));
    HAMD_CHECK(hipMemcpy(plan->d_live.p, plan->live_host.data(), plan->live_host.size() * 4,
                         hipMemcpyHostToDevice));
    plan->d_delta_t.ensure(plan->dt_host.size() * 4);
    HAMD_CHECK(hipMemcpy(plan->d_delta_t.p, plan->dt_host.data(), plan->dt_host.size() * 4,
                         hipMemcpyHostToDevice));
    if (!plan->ent_host.empty()) {
      plan->d_ent.ensure(plan->ent_host.size() * 4);
      HAMD_CHECK(hipMemcpy(plan->d_ent.p, plan->ent_host.data(), plan->ent_host.size() * 4,
                           hipMemcpyHostToDevice));
    }
  }
  plan->dirty = false;
  return DEDISP_NO_ERROR;
}

// ---- scratch of dedisp_execute_adv_noalloc ----------------------------------
// A fixed device array (each GPU the library runs on holds its own copy until
// the process ends): the call's descriptor, tile table, delay tables and the
// channel-major copy of the input live here, so the call allocates nothing.
// In host code the array's name is only a stand-in; the device address comes
// from hipGetSymbolAddress, once per device, with that device current.
constexpr size_t NOALLOC_SCRATCH_BYTES = 64u << 20;
static __device__ __align__(256) uint8_t scratch[NOALLOC_SCRATCH_BYTES];

size_t noalloc_scratch_bytes() { return NOALLOC_SCRATCH_BYTES; }

uint8_t *noalloc_scratch()
{
  constexpr int MAXDEV = 64;
  static std::mutex mu;
  static uint8_t *ptr[MAXDEV] = {};
  int dev = 0;
  HAMD_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= MAXDEV) throw HipError{hipErrorInvalidDevice, "device index", __LINE__};
  std::lock_guard<std::mutex> lk(mu);
  if (!ptr[dev]) {
    void *p = nullptr;
    HAMD_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    ptr[dev] = static_cast<uint8_t *>(p);
  }
  return ptr[dev];
}

// The tile table of dd_build_tiles for one beam in DM-block-major order,
// written on the device: list entry k is (d0, t0) = ((k / nt) * ND,
// (k % nt) * T), and slot i of the XCD-interleaved table holds entry
// k = (i % 8) * q + i / 8 (q = slots / 8); slots past the list are empty.
__global__ void k_fill_tiles(DDTileDesc *tiles, int slots, int n, int nt, int T, int ND)
{
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= slots) return;
  const int q = slots / 8;
  const int k = (i % 8) * q + i / 8;
  tiles[i] = k < n ? DDTileDesc{0, (k / nt) * ND, (k % nt) * T, 0} : DDTileDesc{-1, 0, 0, 0};
}

void fill_tiles(DDTileDesc *tiles, int slots, int n, int nt, int nd, int T, int ND,
                hipStream_t s)
{
  (void)nd;
  hipLaunchKernelGGL(k_fill_tiles, dim3((unsigned)((slots + 255) / 256)), dim3(256), 0, s, tiles,
                     slots, n, nt, T, ND);
  HAMD_CHECK(hipGetLastError());
}

dedisp_error gpu_error(const HipError &e)
{
  set_last_error(std::string(e.what) + ": " + hipGetErrorString(e.err));
  if (e.err == hipErrorOutOfMemory) return DEDISP_MEM_ALLOC_FAILED;
  return DEDISP_INTERNAL_GPU_ERROR;
}

}  // namespace

extern "C" {

dedisp_error dedisp_create_plan(dedisp_plan *plan, dedisp_size nchans, dedisp_float dt,
                                dedisp_float f0, dedisp_float df)
{
  if (!plan) return DEDISP_INVALID_POINTER;
  *plan = nullptr;
  if (nchans == 0 || !(dt > 0.f) || !(f0 > 0.f) || !(df < 0.f)) return DEDISP_UNKNOWN_ERROR;
  if (nchans > (1u << 16)) return DEDISP_NCHANS_EXCEEDS_LIMIT;
  auto *p = new (std::nothrow) dedisp_plan_struct();
  if (!p) return DEDISP_MEM_ALLOC_FAILED;
  p->h.nchans = nchans;
  p->h.dt = dt;
  p->h.f0 = f0;
  p->h.df = df;
  p->h.build_delay_table();
  p->h.killmask.assign(nchans, 1);
  p->killmask_pub.assign(nchans, 1);
  *plan = p;
  return DEDISP_NO_ERROR;
}

void dedisp_destroy_plan(dedisp_plan plan) { delete plan; }

dedisp_error dedisp_generate_dm_list(dedisp_plan plan, dedisp_float dm_start, dedisp_float dm_end,
                                     dedisp_float ti, dedisp_float tol)
{
  if (!plan) return DEDISP_INVALID_PLAN;
  if (!(dm_end >= dm_start) || !(tol > 1.f) || !(ti >= 0.f)) return DEDISP_UNKNOWN_ERROR;
  plan->h.generate_dm_list(dm_start, dm_end, (double)ti, (double)tol);
  try {
    prepare_host(plan);
  } catch (...) {
    plan->host_ready = false;
    return DEDISP_MEM_ALLOC_FAILED;
  }
  return DEDISP_NO_ERROR;
}

dedisp_error dedisp_set_dm_list(dedisp_plan plan, const dedisp_float *dm_list, dedisp_size count)
{
  if (!plan) return DEDISP_INVALID_PLAN;
  if (!dm_list || count == 0) return DEDISP_INVALID_POINTER;
  for (dedisp_size i = 0; i < count; ++i)
    if (!(dm_list[i] >= 0.f) || (i && dm_list[i] < dm_list[i - 1])) return DEDISP_UNKNOWN_ERROR;
  plan->h.set_dm_list(dm_list, count);
  try {
    prepare_host(plan);
  } catch (...) {
    plan->host_ready = false;
    return DEDISP_MEM_ALLOC_FAILED;
  }
  return DEDISP_NO_ERROR;
}

dedisp_error dedisp_set_killmask(dedisp_plan plan, const dedisp_bool *killmask)
{
  if (!plan) return DEDISP_INVALID_PLAN;
  for (size_t c = 0; c < plan->h.nchans; ++c) {
    int v = killmask ? (killmask[c] != 0) : 1;
    plan->h.killmask[c] = v;
    plan->killmask_pub[c] = v;
  }
  plan->host_ready = false;
  plan->dirty = true;
  if (plan->h.dm_list.empty()) return DEDISP_NO_ERROR;  // tables follow the DM list
  try {
    prepare_host(plan);
  } catch (...) {
    return DEDISP_MEM_ALLOC_FAILED;
  }
  return DEDISP_NO_ERROR;
}

dedisp_error dedisp_set_gulp_size(dedisp_plan plan, dedisp_size gulp_size)
{
  if (!plan) return DEDISP_INVALID_PLAN;
  if (gulp_size == 0) return DEDISP_UNKNOWN_ERROR;
  plan->h.gulp_size = gulp_size;
  return DEDISP_NO_ERROR;
}

dedisp_size dedisp_get_gulp_size(dedisp_plan plan) { return plan ? plan->h.gulp_size : 0; }
const dedisp_float *dedisp_get_dm_list(const dedisp_plan plan)
{
  return (plan && !plan->h.dm_list.empty()) ? plan->h.dm_list.data() : nullptr;
}
dedisp_size dedisp_get_dm_count(const dedisp_plan plan) { return plan ? plan->h.dm_list.size() : 0; }
dedisp_size dedisp_get_max_delay(const dedisp_plan plan) { return plan ? plan->h.max_delay : 0; }
dedisp_size dedisp_get_dm_delay(const dedisp_plan plan, int dm_trial)
{
  if (!plan || dm_trial < 0 || (size_t)dm_trial >= plan->h.dm_list.size()) return 0;
  return (dedisp_size)plan->h.delay_of((size_t)dm_trial, plan->h.nchans - 1);
}
dedisp_size dedisp_get_channel_count(const dedisp_plan plan) { return plan ? plan->h.nchans : 0; }
const dedisp_bool *dedisp_get_killmask(const dedisp_plan plan)
{
  return plan ? plan->killmask_pub.data() : nullptr;
}
dedisp_float dedisp_get_dt(const dedisp_plan plan) { return plan ? plan->h.dt : 0.f; }
dedisp_float dedisp_get_f0(const dedisp_plan plan) { return plan ? plan->h.f0 : 0.f; }
dedisp_float dedisp_get_df(const dedisp_plan plan) { return plan ? plan->h.df : 0.f; }

dedisp_error dedisp_set_stream(dedisp_plan plan, void *stream)
{
  if (!plan) return DEDISP_INVALID_PLAN;
  plan->stream = static_cast<hipStream_t>(stream);
  return DEDISP_NO_ERROR;
}

dedisp_error dedisp_execute_guru(const dedisp_plan plan, dedisp_size nsamps, const dedisp_byte *in,
                                 dedisp_size in_nbits, dedisp_size in_stride, dedisp_byte *out,
                                 dedisp_size out_nbits, dedisp_size out_stride,
                                 dedisp_size first_dm_idx, dedisp_size dm_count, unsigned flags)
{
  if (!plan) return DEDISP_INVALID_PLAN;
  if (plan->h.dm_list.empty()) return DEDISP_NO_DM_LIST_SET;
  if (!in || !out) return DEDISP_INVALID_POINTER;
  if ((flags & DEDISP_HOST_POINTERS) && (flags & DEDISP_DEVICE_POINTERS))
    return DEDISP_INVALID_FLAG_COMBINATION;
  if ((flags & DEDISP_WAIT) && (flags & DEDISP_ASYNC)) return DEDISP_INVALID_FLAG_COMBINATION;
  if (in_nbits != 8) return DEDISP_UNSUPPORTED_IN_NBITS;
  if (out_nbits != 32) return DEDISP_UNSUPPORTED_OUT_NBITS;
  const size_t nch = plan->h.nchans;
  const size_t md = plan->h.max_delay;
  if (nsamps <= md) return DEDISP_TOO_FEW_NSAMPS;
  if (first_dm_idx + dm_count > plan->h.dm_list.size() || dm_count == 0)
    return DEDISP_UNKNOWN_ERROR;
  const size_t nout = nsamps - md;
  if (in_stride < nch) return DEDISP_INVALID_STRIDE;
  if (out_stride < nout * 4 || out_stride % 4) return DEDISP_INVALID_STRIDE;
  if (nout > (size_t)INT32_MAX || dm_count > (size_t)INT32_MAX) return DEDISP_TOO_FEW_NSAMPS;
  const bool dev = (flags & DEDISP_DEVICE_POINTERS) != 0;
  try {
    kernels_setup();  // first call on a device: raised LDS limits (not a stream operation)
    refresh_device(plan);
    hipStream_t s = plan->stream;
    const uint8_t *d_in = in;
    if (!dev) {
      plan->scratch_in.ensure(nsamps * in_stride);
      HAMD_CHECK(hipMemcpyAsync(plan->scratch_in.p, in, nsamps * in_stride, hipMemcpyHostToDevice, s));
      d_in = plan->scratch_in.as<uint8_t>();
    }
    // transpose to channel-major rows long enough for every tile read
    const size_t T = (size_t)dd_tile_T(plan->tile);
    const size_t ntiles = (nout + T - 1) / T;
    const size_t ring = round_up(ntiles * T + md + (size_t)plan->max_spread + 64, 64);
    const size_t row_stride = ring;
    plan->scratch_ft.ensure(nch * row_stride);
    transpose_tf_to_ft(d_in, (int64_t)nsamps, (int64_t)in_stride, (int)nch,
                       plan->scratch_ft.as<uint8_t>(), (int64_t)row_stride, (int64_t)ring, s);
    void *d_out = out;
    if (!dev) {
      plan->scratch_out.ensure(dm_count * out_stride);
      d_out = plan->scratch_out.p;
    }
    DDBeamDev B{};
    B.base = plan->scratch_ft.as<uint8_t>();
    B.row_stride = (int64_t)row_stride;
    B.ring = (int64_t)ring;
    B.start = 0;
    B.delta = plan->d_delta.as<int32_t>() + first_dm_idx * (size_t)plan->nlive;
    B.live = plan->d_live.as<int32_t>();
    B.delta_t = plan->d_delta_t.as<int32_t>() + first_dm_idx;
    B.dt_stride = plan->dt_stride;
    // the entry tables hold blocks that start at multiples of ent_nd
    if (plan->ent_nd > 0 && first_dm_idx % (size_t)plan->ent_nd == 0) {
      B.ent = plan->d_ent.as<uint32_t>() + first_dm_idx * (size_t)plan->ent_stride * 2;
      B.ent_stride = plan->ent_stride;
      B.ent_nd = plan->ent_nd;
    }
    B.nlive = plan->nlive;
    B.ndm = (int)dm_count;
    B.nout = (int)nout;
    B.out_f32 = 1;
    B.out = d_out;
    B.out_stride = (int64_t)(out_stride / 4);
    if (plan->nlive == 0) {
      fill_zero_rows(d_out, B.out_stride, (int)nout, (int)dm_count, s);
    } else {
      if (plan->tiles.empty() || plan->tiles_nout != nout || plan->tiles_first != first_dm_idx ||
          plan->tiles_count != dm_count) {
        dd_build_tiles(std::vector<DDBeamDev>{B}, plan->tile, plan->tiles);
        plan->d_tiles.ensure(plan->tiles.size() * sizeof(DDTileDesc));
        HAMD_CHECK(hipMemcpy(plan->d_tiles.p, plan->tiles.data(),
                             plan->tiles.size() * sizeof(DDTileDesc), hipMemcpyHostToDevice));
        plan->tiles_nout = nout;
        plan->tiles_first = first_dm_idx;
        plan->tiles_count = dm_count;
      }
      plan->d_beam.ensure(sizeof(DDBeamDev));
      HAMD_CHECK(hipMemcpyAsync(plan->d_beam.p, &B, sizeof(DDBeamDev), hipMemcpyHostToDevice, s));
      dd_launch(plan->d_beam.as<DDBeamDev>(), plan->d_tiles.as<DDTileDesc>(),
                (int)plan->tiles.size(), plan->tile, plan->max_spread, s, (int64_t)ring);
    }
    if (!dev)
      HAMD_CHECK(hipMemcpyAsync(out, d_out, dm_count * out_stride, hipMemcpyDeviceToHost, s));
    if (!(flags & DEDISP_ASYNC)) HAMD_CHECK(hipStreamSynchronize(s));
  } catch (const HipError &e) {
    return gpu_error(e);
  } catch (...) {
    return DEDISP_UNKNOWN_ERROR;
  }
  return DEDISP_NO_ERROR;
}

dedisp_error dedisp_execute_adv(const dedisp_plan plan, dedisp_size nsamps, const dedisp_byte *in,
                                dedisp_size in_nbits, dedisp_size in_stride, dedisp_byte *out,
                                dedisp_size out_nbits, dedisp_size out_stride, unsigned flags)
{
  if (!plan) return DEDISP_INVALID_PLAN;
  return dedisp_execute_guru(plan, nsamps, in, in_nbits, in_stride, out, out_nbits, out_stride, 0,
                             plan->h.dm_list.size(), flags);
}

dedisp_error dedisp_execute(const dedisp_plan plan, dedisp_size nsamps, const dedisp_byte *in,
                            dedisp_size in_nbits, dedisp_byte *out, dedisp_size out_nbits,
                            unsigned flags)
{
  if (!plan) return DEDISP_INVALID_PLAN;
  if (nsamps <= plan->h.max_delay) return DEDISP_TOO_FEW_NSAMPS;
  const size_t in_stride = plan->h.nchans * in_nbits / 8;
  const size_t out_stride = (nsamps - plan->h.max_delay) * out_nbits / 8;
  return dedisp_execute_adv(plan, nsamps, in, in_nbits, in_stride, out, out_nbits, out_stride, flags);
}

dedisp_error dedisp_execute_adv_noalloc(const dedisp_plan plan, dedisp_size nsamps,
                                        const dedisp_byte *in, dedisp_size in_nbits,
                                        dedisp_size in_stride, dedisp_byte *out,
                                        dedisp_size out_nbits, dedisp_size out_stride,
                                        unsigned flags)
{
  if (!plan) return DEDISP_INVALID_PLAN;
  if (plan->h.dm_list.empty()) return DEDISP_NO_DM_LIST_SET;
  if (!in || !out) return DEDISP_INVALID_POINTER;
  // device pointers only: host buffers would need device staging memory
  if (!(flags & DEDISP_DEVICE_POINTERS) || (flags & DEDISP_HOST_POINTERS))
    return DEDISP_INVALID_FLAG_COMBINATION;
  if ((flags & DEDISP_WAIT) && (flags & DEDISP_ASYNC)) return DEDISP_INVALID_FLAG_COMBINATION;
  if (in_nbits != 8) return DEDISP_UNSUPPORTED_IN_NBITS;
  if (out_nbits != 32) return DEDISP_UNSUPPORTED_OUT_NBITS;
  const size_t nch = plan->h.nchans;
  const size_t md = plan->h.max_delay;
  if (nsamps <= md) return DEDISP_TOO_FEW_NSAMPS;
  const size_t nout = nsamps - md;
  const size_t ndm = plan->h.dm_list.size();
  if (in_stride < nch) return DEDISP_INVALID_STRIDE;
  if (out_stride < nout * 4 || out_stride % 4) return DEDISP_INVALID_STRIDE;
  if (nout > (size_t)INT32_MAX || ndm > (size_t)INT32_MAX) return DEDISP_TOO_FEW_NSAMPS;
  // the setters leave the host tables ready; a plan whose last setter failed
  // would need them rebuilt (host memory): refuse instead
  if (!plan->host_ready) {
    set_last_error("dedisp_execute_adv_noalloc: the plan's tables are not built "
                   "(set the DM list again)");
    return DEDISP_UNKNOWN_ERROR;
  }
  try {
    hipStream_t s = plan->stream;
    if (plan->nlive == 0) {
      fill_zero_rows(out, (int64_t)(out_stride / 4), (int)nout, (int)ndm, s);
    } else {
      const DDTile t = plan->tile;
      const size_t T = (size_t)dd_tile_T(t), ND = (size_t)dd_tile_ND(t);
      const size_t nt = (nout + T - 1) / T, nd = (ndm + ND - 1) / ND;
      const size_t ring = round_up(nt * T + md + (size_t)plan->max_spread + 64, 64);
      // scratch layout (256-byte aligned parts)
      const size_t n_tiles = (nt * nd + 7) / 8 * 8;
      const size_t o_tiles = 256;
      const size_t o_delta = round_up(o_tiles + n_tiles * sizeof(DDTileDesc), 256);
      const size_t o_live = round_up(o_delta + (plan->delta_host.size() + DD_DELTA_PAD) * 4, 256);
      const size_t o_dt = round_up(o_live + plan->live_host.size() * 4, 256);
      const size_t o_ent = round_up(o_dt + plan->dt_host.size() * 4, 256);
      const size_t o_win = round_up(o_ent + plan->ent_host.size() * 4, 256);
      const size_t need = o_win + nch * ring;
      if (need > noalloc_scratch_bytes()) {
        set_last_error("dedisp_execute_adv_noalloc: needs " + std::to_string(need) +
                       " bytes of scratch, the limit is " +
                       std::to_string(noalloc_scratch_bytes()));
        return DEDISP_MEM_ALLOC_FAILED;
      }
      kernels_setup();  // first call on a device: raised LDS limits (no allocation)
      uint8_t *sc = noalloc_scratch();
      auto up = [&](size_t off, const void *src, size_t bytes) {
        if (bytes) HAMD_CHECK(hipMemcpyAsync(sc + off, src, bytes, hipMemcpyHostToDevice, s));
      };
      // every call rewrites what it reads: another call (another plan) on
      // this device may have left its own tables in the shared scratch
      up(o_delta, plan->delta_host.data(), plan->delta_host.size() * 4);
      up(o_live, plan->live_host.data(), plan->live_host.size() * 4);
      up(o_dt, plan->dt_host.data(), plan->dt_host.size() * 4);
      up(o_ent, plan->ent_host.data(), plan->ent_host.size() * 4);
      transpose_tf_to_ft(in, (int64_t)nsamps, (int64_t)in_stride, (int)nch, sc + o_win,
                         (int64_t)ring, (int64_t)ring, s);
      fill_tiles(reinterpret_cast<DDTileDesc *>(sc + o_tiles), (int)n_tiles, (int)(nt * nd),
                 (int)nt, (int)nd, (int)T, (int)ND, s);
      DDBeamDev &B = plan->na_desc;
      B = DDBeamDev{};
      B.base = sc + o_win;
      B.row_stride = (int64_t)ring;
      B.ring = (int64_t)ring;
      B.start = 0;
      B.delta = reinterpret_cast<const int32_t *>(sc + o_delta);
      B.live = reinterpret_cast<const int32_t *>(sc + o_live);
      B.delta_t = reinterpret_cast<const int32_t *>(sc + o_dt);
      B.dt_stride = plan->dt_stride;
      if (plan->ent_nd > 0) {
        B.ent = reinterpret_cast<const uint32_t *>(sc + o_ent);
        B.ent_stride = plan->ent_stride;
        B.ent_nd = plan->ent_nd;
      }
      B.nlive = plan->nlive;
      B.ndm = (int)ndm;
      B.nout = (int)nout;
      B.out_f32 = 1;
      B.out = out;
      B.out_stride = (int64_t)(out_stride / 4);
      up(0, &B, sizeof(DDBeamDev));
      dd_launch(reinterpret_cast<const DDBeamDev *>(sc), reinterpret_cast<DDTileDesc *>(sc + o_tiles),
                (int)n_tiles, t, plan->max_spread, s, (int64_t)ring);
    }
    if (!(flags & DEDISP_ASYNC)) HAMD_CHECK(hipStreamSynchronize(s));
  } catch (const HipError &e) {
    return gpu_error(e);
  } catch (...) {
    return DEDISP_UNKNOWN_ERROR;
  }
  return DEDISP_NO_ERROR;
}

dedisp_error dedisp_set_device(int device_idx)
{
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return DEDISP_PRIOR_GPU_ERROR;
  if (device_idx < 0 || device_idx >= n) return DEDISP_INVALID_DEVICE_INDEX;
  if (hipSetDevice(device_idx) != hipSuccess) return DEDISP_INTERNAL_GPU_ERROR;
  return DEDISP_NO_ERROR;
}

dedisp_error dedisp_sync(void)
{
  return hipDeviceSynchronize() == hipSuccess ? DEDISP_NO_ERROR : DEDISP_PRIOR_GPU_ERROR;
}

const char *dedisp_get_error_string(dedisp_error error)
{
  switch (error) {
    case DEDISP_NO_ERROR: return "No error";
    case DEDISP_MEM_ALLOC_FAILED: return "Memory allocation failed";
    case DEDISP_MEM_COPY_FAILED: return "Memory copy failed";
    case DEDISP_NCHANS_EXCEEDS_LIMIT: return "No. channels exceeds internal limit";
    case DEDISP_INVALID_PLAN: return "Invalid plan";
    case DEDISP_INVALID_POINTER: return "Invalid pointer";
    case DEDISP_INVALID_STRIDE: return "Invalid stride";
    case DEDISP_NO_DM_LIST_SET: return "No DM list has been set";
    case DEDISP_TOO_FEW_NSAMPS: return "No. samples < maximum delay";
    case DEDISP_INVALID_FLAG_COMBINATION: return "Invalid flag combination";
    case DEDISP_UNSUPPORTED_IN_NBITS: return "Unsupported in_nbits value";
    case DEDISP_UNSUPPORTED_OUT_NBITS: return "Unsupported out_nbits value";
    case DEDISP_INVALID_DEVICE_INDEX: return "Invalid device index";
    case DEDISP_DEVICE_ALREADY_SET: return "Device is already set and cannot be changed";
    case DEDISP_PRIOR_GPU_ERROR: return "There was a previous GPU error";
    case DEDISP_INTERNAL_GPU_ERROR: return "Internal GPU error";
    case DEDISP_UNKNOWN_ERROR: return "Unknown error";
  }
  return "Invalid error code";
}

}  // extern "C"
