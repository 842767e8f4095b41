// synth.cpp -- seeded synthetic filterbank generator (host).
//
// Noise: SplitMix64 seeds a xoshiro256** stream per absolute row, so any row
// range of a stream can be generated on its own and chunks stay continuous;
// Box-Muller turns pairs of uniforms into normals.  Pulses are Gaussians in
// time, dispersed with the dedisp plan's float delay table, with the per-channel
// amplitude that gives the requested band-summed matched-filter S/N:
//   A = snr * sigma / (sqrt(nchan) * sqrt(sqrt(pi) * sigma_t)),
//   sigma_t = FWHM / 2.3548 in samples.
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "../../include/hella.h"
#include "internal.hpp"

namespace {

inline uint64_t splitmix64(uint64_t &x)
{
  uint64_t z = (x += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

struct Xoshiro {
  uint64_t s[4];
  explicit Xoshiro(uint64_t seed)
  {
    for (auto &w : s) w = splitmix64(seed);
  }
  uint64_t next()
  {
    const uint64_t r = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return r;
  }
  double uniform_open() { return ((double)(next() >> 11) + 1.0) * (1.0 / 9007199254740992.0); }
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

}  // namespace

extern "C" hella_status hella_synth_filterbank(const hella_synth_params *p, void *out)
{
  if (!p || !out || p->nchan <= 0 || p->nsamp < 0 || p->tsamp <= 0) {
    hamd::set_last_error("hella_synth_filterbank: bad parameters");
    return HELLA_ERR_INVALID;
  }
  const int nch = p->nchan;
  const int64_t ns = p->nsamp;
  std::vector<float> sig((size_t)nch, p->sigma), gain((size_t)nch, 1.f);
  for (int i = 0; i < p->n_rfi_chan; ++i) {
    const int c = p->rfi_chan[i];
    if (c >= 0 && c < nch) sig[(size_t)c] = p->sigma * p->rfi_sigma_mult;
  }
  for (int c = 0; c < nch; ++c)
    gain[(size_t)c] = 1.f + p->bp_slope * ((nch > 1 ? (float)c / (float)(nch - 1) : 0.f) - 0.5f);
  std::vector<float> buf((size_t)ns * nch);
  // noise, one RNG stream per absolute row
  const unsigned nth = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  auto work = [&](int64_t r0, int64_t r1) {
    for (int64_t r = r0; r < r1; ++r) {
      const uint64_t a = (uint64_t)(p->t_offset + r);
      Xoshiro g(p->seed ^ (a * 0xD1B54A32D192ED03ull));
      float *row = buf.data() + (size_t)r * nch;
      for (int c = 0; c < nch; c += 2) {
        const double u1 = g.uniform_open(), u2 = g.uniform();
        const double rad = std::sqrt(-2.0 * std::log(u1));
        const double th = 6.283185307179586 * u2;
        row[c] = p->mean + sig[(size_t)c] * (float)(rad * std::cos(th));
        if (c + 1 < nch) row[c + 1] = p->mean + sig[(size_t)c + 1] * (float)(rad * std::sin(th));
      }
      if (p->burst_period > 0 && (int64_t)a >= p->burst_phase &&
          ((int64_t)a - p->burst_phase) % p->burst_period < p->burst_len)
        for (int c = 0; c < nch; ++c) row[c] += p->burst_amp * p->sigma;
    }
  };
  std::vector<std::thread> th;
  const int64_t per = (ns + nth - 1) / nth;
  for (unsigned i = 0; i < nth; ++i) {
    const int64_t r0 = (int64_t)i * per, r1 = std::min(ns, r0 + per);
    if (r0 < r1) th.emplace_back(work, r0, r1);
  }
  for (auto &t : th) t.join();
  // pulses: dedisp float delay table (4.15e3 / dt * (1/f^2 - 1/f0^2))
  if (p->n_pulse > 0) {
    hamd::PlanHost plan;
    plan.nchans = (size_t)nch;
    plan.dt = (float)p->tsamp;
    plan.f0 = (float)p->fch1;
    plan.df = (float)p->foff;
    plan.build_delay_table();
    for (int k = 0; k < p->n_pulse; ++k) {
      const double sg = (double)p->pulse_fwhm_ms[k] * 1e-3 / p->tsamp / 2.3548;
      const double amp = (double)p->pulse_snr[k] * p->sigma /
                         (std::sqrt((double)nch) * std::sqrt(std::sqrt(M_PI) * sg));
      const double t0 = p->pulse_t[k] / p->tsamp;
      const double half = 6.0 * sg + 2.0;
      for (int c = 0; c < nch; ++c) {
        const double tc = t0 + (double)p->pulse_dm[k] * (double)plan.delay_table[(size_t)c];
        int64_t a0 = (int64_t)std::floor(tc - half), a1 = (int64_t)std::ceil(tc + half);
        a0 = std::max<int64_t>(a0, p->t_offset);
        a1 = std::min<int64_t>(a1, p->t_offset + ns - 1);
        for (int64_t a = a0; a <= a1; ++a) {
          const double x = ((double)a - tc) / sg;
          buf[(size_t)(a - p->t_offset) * nch + c] += (float)(amp * std::exp(-0.5 * x * x));
        }
      }
    }
  }
  const size_t n = buf.size();
  if (p->dtype == 0) {
    uint8_t *o = static_cast<uint8_t *>(out);
    for (size_t i = 0; i < n; ++i) {
      const int c = (int)(i % (size_t)nch);
      float v = std::rint(buf[i] * gain[(size_t)c]);
      v = v < 0.f ? 0.f : (v > 255.f ? 255.f : v);
      o[i] = (uint8_t)v;
    }
  } else {
    float *o = static_cast<float *>(out);
    for (size_t i = 0; i < n; ++i) o[i] = buf[i] * gain[i % (size_t)nch];
  }
  return HELLA_OK;
}
