// plan.cpp -- host side of the dedispersion plan (dedisp semantics).
//
// Restates the published dedisp plan algorithm that casm-hella calls at
// src/apps/casm_hella_refactored.cu:205-210: a float per-channel delay table,
// the FP64 smearing-tolerance DM-list recurrence (each step stored as float),
// and max_delay = (size_t)(dm_last * delay_last + 0.5).  The integer delay of
// every (DM, channel) pair is computed here once and uploaded, so the device
// and any checker use identical delays.
#include <cmath>

#include "internal.hpp"

namespace hamd {

namespace {
thread_local std::string g_err;
}  // namespace

void set_last_error(const std::string &msg) { g_err = msg; }
const char *last_error() { return g_err.c_str(); }

void PlanHost::build_delay_table()
{
  delay_table.resize(nchans);
  for (size_t c = 0; c < nchans; ++c) {
    float a = 1.f / (f0 + (float)c * df);
    float b = 1.f / f0;
    float ab = a * a - b * b;
    delay_table[c] = (float)(4.15e3 / (double)dt * (double)ab);
  }
}

void PlanHost::generate_dm_list(float dm_start, float dm_end, double ti_us, double tol)
{
  double dtu = (double)dt * 1e6;
  double f = ((double)f0 + ((double)(nchans / 2) - 0.5) * (double)df) * 1e-3;
  double tol2 = tol * tol;
  double a = 8.3 * (double)df / (f * f * f);
  double a2 = a * a;
  double b2 = a2 * ((double)(nchans * nchans) / 16.0);
  double c = (dtu * dtu + ti_us * ti_us) * (tol2 - 1.0);
  dm_list.clear();
  dm_list.push_back(dm_start);
  while (dm_list.back() < dm_end) {
    double prev = (double)dm_list.back();
    double prev2 = prev * prev;
    double k = c + tol2 * a2 * prev2;
    double dm = (b2 * prev + std::sqrt(-a2 * b2 * prev2 + (a2 + b2) * k)) / (a2 + b2);
    dm_list.push_back((float)dm);
  }
  compute_max_delay();
}

void PlanHost::set_dm_list(const float *dms, size_t n)
{
  dm_list.assign(dms, dms + n);
  compute_max_delay();
}

void PlanHost::compute_max_delay()
{
  if (dm_list.empty() || nchans == 0) {
    max_delay = 0;
    return;
  }
  float p = dm_list.back() * delay_table[nchans - 1];
  max_delay = (size_t)((double)p + 0.5);
}

void PlanHost::delays_live(size_t d_first, size_t nd, std::vector<int32_t> &delta,
                           std::vector<int32_t> &live) const
{
  live.clear();
  for (size_t c = 0; c < nchans; ++c)
    if (killmask.empty() || killmask[c]) live.push_back((int32_t)c);
  delta.resize(nd * live.size());
  for (size_t d = 0; d < nd; ++d)
    for (size_t i = 0; i < live.size(); ++i)
      delta[d * live.size() + i] = delay_of(d_first + d, (size_t)live[i]);
}

void dd_entries(const std::vector<int32_t> &delta, int ndm, int nlive, int nd,
                std::vector<uint32_t> &out, int &stride)
{
  auto sel = [](uint32_t s) { return s | 0x0c00u | ((s + 2u) << 16) | 0x0c000000u; };
  stride = (nlive + 16 + 3) / 4 * 4;
  out.assign((size_t)ndm * stride * 2, 0);
  for (int d = 0; d < ndm; ++d) {
    const int d0 = d - d % nd;
    uint32_t *row = out.data() + (size_t)d * stride * 2;
    for (int c = 0; c < stride; ++c) {
      int q = 0;
      if (c < nlive) {
        const int32_t bas = delta[(size_t)d0 * nlive + c], own = delta[(size_t)d * nlive + c];
        q = (bas & 15) + own - bas;
      }
      row[2 * c] = (uint32_t)(q & ~3);
      row[2 * c + 1] = sel((uint32_t)(q & 3));
    }
  }
}

void delays_transposed(const std::vector<int32_t> &delta, int ndm, int nlive,
                       std::vector<int32_t> &out, int &stride)
{
  stride = ndm + 64;
  out.assign((size_t)nlive * stride, 0);
  for (int c = 0; c < nlive; ++c) {
    int32_t *row = out.data() + (size_t)c * stride;
    for (int d = 0; d < ndm; ++d) row[d] = delta[(size_t)d * nlive + c];
    for (int d = ndm; d < stride; ++d) row[d] = ndm > 0 ? row[ndm - 1] : 0;
  }
}

}  // namespace hamd
