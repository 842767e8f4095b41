// emit.cpp -- candidate emission in the reference's text formats
// (src/hella/src/peaks.cu:129-259) and the spectral-flag file
// (src/apps/casm_hella_refactored.cu:756-761).  Host code only: it formats
// candidate lists the pipeline or the server already returned.
#include <cerrno>
#include <cinttypes>
#include <cstdio>
#include <string>

#include "pipeline.hpp"

using namespace hamd;

namespace {

// definitions.h:56,59: TIME_START is the start sample at a fixed 1 ms per
// sample, in days (the reference does not use TSAMP here)
constexpr double TIME_RESOLUTION_S = 1.0e-3;
constexpr double SECONDS_PER_DAY_S = 86400.0;

const char *const HEADER = "SNR SAMP_START TIME_START WIDTH DM_IDX DM BEAM_IDX";

// one candidate as "%g %d %g %d %d %g %d" (peaks.cu:170; the socket stream
// of peaks.cu:220-227 prints the same fields with a default ostream, which
// is %g for floating values)
std::string line(const hella_candidate &c, int beam0)
{
  char buf[160];
  std::snprintf(buf, sizeof buf, "%g %lld %g %d %d %g %d\n", (double)c.snr, c.abs_samp,
                TIME_RESOLUTION_S * (double)c.abs_samp / SECONDS_PER_DAY_S, c.width, c.dm_idx,
                (double)c.dm, c.beam + beam0);
  return buf;
}

}  // namespace

extern "C" {

hella_status hella_write_candidates(const char *path, const hella_candidate *cands, int n,
                                    int beam0)
{
  if (!path || n < 0 || (n > 0 && !cands)) {
    set_last_error("bad argument");
    return HELLA_ERR_INVALID;
  }
  FILE *f = std::fopen(path, "a");
  if (!f) {
    set_last_error(std::string("cannot open ") + path + " for appending");
    return HELLA_ERR_INVALID;
  }
  bool ok = std::fseek(f, 0, SEEK_END) == 0;
  const long size = ok ? std::ftell(f) : -1;
  ok = ok && size >= 0;
  if (ok && size == 0) ok = std::fprintf(f, "%s\n", HEADER) > 0;
  for (int i = 0; ok && i < n; ++i) ok = std::fputs(line(cands[i], beam0).c_str(), f) >= 0;
  ok = (std::fclose(f) == 0) && ok;
  if (!ok) {
    set_last_error(std::string("write to ") + path + " failed");
    return HELLA_ERR_INVALID;
  }
  return HELLA_OK;
}

int hella_format_coincidencer(long long samp0, long long gulp, const char *dada_fields,
                              const hella_candidate *cands, int n, int beam0, char *buf, int cap)
{
  if (gulp <= 0 || n < 0 || (n > 0 && !cands) || cap < 0 || (cap > 0 && !buf)) return -1;
  // first line: (int)(samp / gulp) + 1 with the gulp's start sample, then the
  // DADA header's UTC_START PICOSECONDS TSAMP when the input is DADA
  std::string msg = std::to_string((long long)(samp0 / gulp) + 1);
  if (dada_fields && *dada_fields) msg += std::string(" ") + dada_fields;
  msg += "\n";
  for (int i = 0; i < n; ++i) msg += line(cands[i], beam0);
  if (cap > 0) {
    const size_t k = msg.size() < (size_t)cap - 1 ? msg.size() : (size_t)cap - 1;
    msg.copy(buf, k);
    buf[k] = '\0';
  }
  return (int)msg.size();
}

hella_status hella_write_specflags(const char *path, const int *specflags, int nchan)
{
  if (!path || nchan < 0 || (nchan > 0 && !specflags)) {
    set_last_error("bad argument");
    return HELLA_ERR_INVALID;
  }
  FILE *f = std::fopen(path, "a");
  if (!f) {
    set_last_error(std::string("cannot open ") + path + " for appending");
    return HELLA_ERR_INVALID;
  }
  bool ok = true;
  for (int i = 0; ok && i < nchan; ++i) ok = std::fprintf(f, "%d\n", specflags[i]) > 0;
  ok = (std::fclose(f) == 0) && ok;
  if (!ok) {
    set_last_error(std::string("write to ") + path + " failed");
    return HELLA_ERR_INVALID;
  }
  return HELLA_OK;
}

}  // extern "C"
