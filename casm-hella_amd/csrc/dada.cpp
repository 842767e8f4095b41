// dada.cpp -- the PSRDADA header as key/value pairs.
//
// casm-hella copies the ring's text header (src/hella/src/dada.cpp:38-42) and
// parses it into a key -> value map (parse_dada_header, dada.cpp:50-79):
//   * the text ends at the first NUL byte;
//   * lines are separated by '\n'; empty lines are skipped;
//   * the key runs up to the first space; a line without a space ends the
//     parse (the reference logs a warning and stops);
//   * the value starts at the first non-space after the key and runs to the
//     end of the line (trailing characters kept);
//   * a key given twice keeps its last value.
// The coincidencer message then carries UTC_START, PICOSECONDS and TSAMP from
// that map (peaks.cu:215-220).
//
// One difference: a line whose key is followed by spaces only makes the
// reference's substr(npos) throw; here that key gets the empty value.
#include <cstdio>
#include <cstring>
#include <map>
#include <string>

#include "../../include/hella.h"
#include "internal.hpp"

namespace {

std::map<std::string, std::string> parse_header(const char *text, size_t len)
{
  std::map<std::string, std::string> kv;
  const char *nul = static_cast<const char *>(std::memchr(text, '\0', len));
  const std::string all(text, nul ? (size_t)(nul - text) : len);
  size_t pos = 0;
  while (pos <= all.size()) {
    size_t eol = all.find('\n', pos);
    if (eol == std::string::npos) eol = all.size();
    const std::string line = all.substr(pos, eol - pos);
    pos = eol + 1;
    if (line.empty()) {
      if (eol == all.size()) break;
      continue;
    }
    const size_t sp = line.find(' ');
    if (sp == std::string::npos) break;  // invalid line: stop, as the reference does
    const size_t v = line.find_first_not_of(' ', sp);
    kv[line.substr(0, sp)] = v == std::string::npos ? std::string() : line.substr(v);
    if (eol == all.size()) break;
  }
  return kv;
}

int put(const std::string &s, char *out, int cap)
{
  if (out && cap > 0) {
    const size_t n = s.size() < (size_t)cap - 1 ? s.size() : (size_t)cap - 1;
    std::memcpy(out, s.data(), n);
    out[n] = '\0';
  }
  return (int)s.size();
}

}  // namespace

extern "C" {

int hella_dada_header_value(const char *header, size_t len, const char *key, char *out, int cap)
{
  if (!header || !key || cap < 0 || (cap > 0 && !out)) return -1;
  const auto kv = parse_header(header, len);
  const auto it = kv.find(key);
  if (it == kv.end()) return -1;
  return put(it->second, out, cap);
}

int hella_dada_coincidencer_fields(const char *header, size_t len, char *out, int cap)
{
  if (!header || cap < 0 || (cap > 0 && !out)) return -1;
  const auto kv = parse_header(header, len);
  std::string s;
  for (const char *k : {"UTC_START", "PICOSECONDS", "TSAMP"}) {
    const auto it = kv.find(k);
    if (it == kv.end()) {
      hamd::set_last_error(std::string("DADA header has no ") + k);
      return -1;
    }
    if (!s.empty()) s += ' ';
    s += it->second;
  }
  return put(s, out, cap);
}

}  // extern "C"
