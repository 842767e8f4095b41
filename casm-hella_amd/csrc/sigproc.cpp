// sigproc.cpp -- SIGPROC filterbank reader (src/hella/src/sigproc.cpp:26-191,
// src/hella/sigproc.h:17-49): the keyword header (length-prefixed strings
// followed by binary values) and the time-major uint8 samples after it.
// Host code only.
#include <cstdio>
#include <cstring>
#include <string>

#include "pipeline.hpp"

using namespace hamd;

namespace {

struct File {
  FILE *f = nullptr;
  explicit File(const char *path) : f(std::fopen(path, "rb")) {}
  ~File() {
    if (f) std::fclose(f);
  }
};

// a header string: int length, then that many characters (get_string,
// sigproc.cpp:26-41: lengths outside 1..80 do not name a keyword)
bool get_string(FILE *f, std::string &s)
{
  int32_t n = 0;
  if (std::fread(&n, 4, 1, f) != 1) return false;
  if (n < 1 || n > 80) {
    s = "ERROR";
    return true;
  }
  char buf[81];
  if (std::fread(buf, (size_t)n, 1, f) != 1) return false;
  buf[n] = '\0';
  s = buf;
  return true;
}

template <class T> bool get(FILE *f, T &v) { return std::fread(&v, sizeof(T), 1, f) == 1; }

hella_status bad(const std::string &msg)
{
  set_last_error(msg);
  return HELLA_ERR_CONFIG;
}

}  // namespace

extern "C" {

hella_status hella_sigproc_read_header(const char *path, hella_fil_header *h)
{
  if (!path || !h) {
    set_last_error("bad argument");
    return HELLA_ERR_INVALID;
  }
  std::memset(h, 0, sizeof *h);
  File F(path);
  if (!F.f) {
    set_last_error(std::string("cannot open ") + path);
    return HELLA_ERR_INVALID;
  }
  FILE *f = F.f;
  std::string s;
  if (!get_string(f, s)) return bad("file too short for a header");
  if (s != "HEADER_START") return HELLA_OK;  // not a SIGPROC header: header_bytes 0
  bool want_raw = false, want_source = false;
  int32_t nbins = 0, nsamples = 0;
  for (;;) {
    if (!get_string(f, s)) return bad("header ends before HEADER_END");
    if (s == "HEADER_END") break;
    bool ok = true;
    if (s == "rawdatafile") want_raw = true;
    else if (s == "source_name") want_source = true;
    else if (s == "FREQUENCY_START") h->nfreq = 0;
    else if (s == "FREQUENCY_END") {}
    else if (s == "az_start") ok = get(f, h->az_start);
    else if (s == "za_start") ok = get(f, h->za_start);
    else if (s == "src_raj") ok = get(f, h->src_raj);
    else if (s == "src_dej") ok = get(f, h->src_dej);
    else if (s == "tstart") ok = get(f, h->tstart);
    else if (s == "tsamp") ok = get(f, h->tsamp);
    else if (s == "period") ok = get(f, h->period);
    else if (s == "fch1") ok = get(f, h->fch1);
    else if (s == "fchannel") {
      double v = 0;
      ok = get(f, v);
      if (ok && h->nfreq < HELLA_FIL_MAXFREQ) h->frequency_table[h->nfreq++] = v;
      h->fch1 = h->foff = 0.0;  // a table is in use
    }
    else if (s == "foff") ok = get(f, h->foff);
    else if (s == "nchans") ok = get(f, h->nchans);
    else if (s == "telescope_id") ok = get(f, h->telescope_id);
    else if (s == "machine_id") ok = get(f, h->machine_id);
    else if (s == "data_type") ok = get(f, h->data_type);
    else if (s == "ibeam") ok = get(f, h->ibeam);
    else if (s == "nbeams") ok = get(f, h->nbeams);
    else if (s == "nbits") ok = get(f, h->nbits);
    else if (s == "barycentric") ok = get(f, h->barycentric);
    else if (s == "pulsarcentric") ok = get(f, h->pulsarcentric);
    else if (s == "nbins") ok = get(f, nbins);
    else if (s == "nsamples") ok = get(f, nsamples);
    else if (s == "nifs") ok = get(f, h->nifs);
    else if (s == "npuls") ok = get(f, h->npuls);
    else if (s == "refdm") ok = get(f, h->refdm);
    else if (s == "signed") ok = get(f, h->isign);
    else if (want_raw) {
      std::snprintf(h->raw_data_file, sizeof h->raw_data_file, "%s", s.c_str());
      want_raw = false;
    } else if (want_source) {
      std::snprintf(h->source_name, sizeof h->source_name, "%s", s.c_str());
      want_source = false;
    } else {
      return bad("read_header: unknown parameter: " + s);
    }
    if (!ok) return bad("header ends inside the value of " + s);
  }
  h->header_bytes = (long long)std::ftell(f);
  if (std::fseek(f, 0, SEEK_END) == 0) {
    const long long end = (long long)std::ftell(f);
    const long long per = (long long)h->nchans * (h->nifs > 0 ? h->nifs : 1) * h->nbits / 8;
    h->nsamples = per > 0 ? (end - h->header_bytes) / per : 0;
  }
  return HELLA_OK;
}

long long hella_sigproc_read(const char *path, const hella_fil_header *h, long long first,
                             long long nsamp, void *out)
{
  if (!path || !h || !out || first < 0 || nsamp < 0) {
    set_last_error("bad argument");
    return -1;
  }
  if (h->nbits != 8 || h->nchans <= 0 || (h->nifs != 0 && h->nifs != 1)) {
    set_last_error("only 8-bit single-IF filterbanks are read (sigproc.cpp:184-188)");
    return -1;
  }
  File F(path);
  if (!F.f) {
    set_last_error(std::string("cannot open ") + path);
    return -1;
  }
  const long long per = h->nchans;
  if (std::fseek(F.f, (long)(h->header_bytes + first * per), SEEK_SET) != 0) {
    set_last_error("seek failed");
    return -1;
  }
  const size_t got = std::fread(out, (size_t)per, (size_t)nsamp, F.f);
  return (long long)got;
}

int hella_sigproc_config(const hella_fil_header *h, char *buf, int cap)
{
  if (!h || cap < 0 || (cap > 0 && !buf)) return -1;
  char tmp[256];
  const int n = std::snprintf(tmp, sizeof tmp, "NCHAN %d\nFCH1 %.17g\nFOFF %.17g\nTSAMP %.17g\n",
                              h->nchans, h->fch1, h->foff, h->tsamp);
  if (cap > 0) std::snprintf(buf, (size_t)cap, "%s", tmp);
  return n;
}

}  // extern "C"
