"""The drop-in boundary compiles where the reference's code calls it (CPU
only, no GPU calls):

* the reference's own dedisp call expressions -- plan creation, DM list and
  queries (casm:205-210) and dedisp_execute_adv on device pointers with the
  error-name switch (dedisperse.cpp:24-91) -- compile against
  include/dedisp.h with the argument types the reference passes;
* the hella:: shim printed in INTEGRATION.md section 2 compiles against
  include/hella.h (with a stand-in for the reference's definitions.h that
  declares the pinfo_t fields the shim touches);
* both objects link against libhella_amd.so with every symbol resolved."""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_DIR = os.path.join(ROOT, "casm-hella_amd", "lib")

pytestmark = pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not installed")

# stand-in for src/hella/definitions.h: the geometry constants and the
# pinfo_t members the call sites and the shim use, with the reference's types
DEFINITIONS_STANDIN = r"""
#pragma once
#include <map>
#include <string>
#include <vector>
#include <dedisp.h>
#define NCHAN 3072
#define FREQ_CHANNEL_WIDTH -0.03075
#define CHAN0_FREQ 468.75e6
#define TIME_RESOLUTION 1.0e-3
namespace hella {
typedef struct pinfo {
  int inp_format;
  float minDM, maxDM;
  unsigned long gulp;
  char specflags[500];
  int out_format;
  char out_path[500];
  int BEAM0;
  std::vector<char> dada_header;
  unsigned long NTIME;
  int ndms;
  float dm_tol;
  int ntime_dd, ntime_out, ntime_dedisp;
  int maxWidth;
  const float *DMs;
  dedisp_plan dedispersion_plan;
  float *d_dedispPacked;
  unsigned char *d_inputPacked;
  unsigned char *d_gulp;
} pinfo_t;
}
"""

# the reference's call expressions, with its argument types
REFERENCE_CALLS = r"""
#include <stdexcept>
#include <string>
#include "hella/definitions.h"

namespace hella {
void plan_setup(pinfo_t *p)              // casm:205-212
{
  dedisp_create_plan(&p->dedispersion_plan, NCHAN, TIME_RESOLUTION, CHAN0_FREQ / 1e6,
                     FREQ_CHANNEL_WIDTH);
  dedisp_generate_dm_list(p->dedispersion_plan, p->minDM, p->maxDM, 40, p->dm_tol);
  p->DMs = dedisp_get_dm_list(p->dedispersion_plan);
  p->ndms = dedisp_get_dm_count(p->dedispersion_plan);
  p->ntime_dd = p->NTIME - dedisp_get_max_delay(p->dedispersion_plan);
  p->ntime_out = p->ntime_dd - p->maxWidth;
}

// the 17 error names hella's error switch handles (dedisperse.cpp:50-91)
static const dedisp_error kErrors[] = {
    DEDISP_NO_ERROR, DEDISP_MEM_ALLOC_FAILED, DEDISP_MEM_COPY_FAILED,
    DEDISP_NCHANS_EXCEEDS_LIMIT, DEDISP_INVALID_PLAN, DEDISP_INVALID_POINTER,
    DEDISP_INVALID_STRIDE, DEDISP_NO_DM_LIST_SET, DEDISP_TOO_FEW_NSAMPS,
    DEDISP_INVALID_FLAG_COMBINATION, DEDISP_UNSUPPORTED_IN_NBITS, DEDISP_UNSUPPORTED_OUT_NBITS,
    DEDISP_INVALID_DEVICE_INDEX, DEDISP_DEVICE_ALREADY_SET, DEDISP_PRIOR_GPU_ERROR,
    DEDISP_INTERNAL_GPU_ERROR, DEDISP_UNKNOWN_ERROR};
static_assert(sizeof(kErrors) / sizeof(kErrors[0]) == 17, "dedisp error names");

void dedisperse_beam(pinfo_t *p)         // dedisperse.cpp:24-46
{
  // device pointers, 8-bit time-major input of NCHAN bytes per sample, 32-bit
  // output rows of ntime_dedisp samples, flag bit 2 (device pointers)
  const dedisp_size bits_in = 8, bits_out = 32;
  const dedisp_error derror = dedisp_execute_adv(
      p->dedispersion_plan, p->NTIME, reinterpret_cast<const dedisp_byte *>(p->d_inputPacked),
      bits_in, NCHAN, reinterpret_cast<dedisp_byte *>(p->d_dedispPacked), bits_out,
      p->ntime_dedisp * bits_out / 8, 1u << 2);
  if (derror != DEDISP_NO_ERROR || kErrors[0] != DEDISP_NO_ERROR)
    throw std::runtime_error(std::string("dedisp_execute_adv failed: ") +
                             dedisp_get_error_string(derror));
}
}  // namespace hella
"""

MAIN = r"""
#include "hella/definitions.h"
namespace hella {
void plan_setup(pinfo_t *);
void dedisperse_beam(pinfo_t *);
void init(pinfo_t *, const char *);
void fast_flagger(pinfo_t *);
void dedisperse(pinfo_t *, int);
void smooth(pinfo_t *, int);
void find_peaks(pinfo_t *, int);
void clear_peaks(pinfo_t *);
void rewind(pinfo_t *);
void output_peaks(pinfo_t *, int, int);
void write_specflags(pinfo_t *, int);
}
// referenced, never run: the link must resolve every call
void (*volatile keep[])() = {
    (void (*)())&hella::plan_setup, (void (*)())&hella::dedisperse_beam,
    (void (*)())&hella::init, (void (*)())&hella::fast_flagger, (void (*)())&hella::dedisperse,
    (void (*)())&hella::smooth, (void (*)())&hella::find_peaks, (void (*)())&hella::clear_peaks,
    (void (*)())&hella::rewind, (void (*)())&hella::output_peaks,
    (void (*)())&hella::write_specflags};
int main() { return keep[0] == nullptr; }
"""


def shim_source():
    """the C++ block of INTEGRATION.md section 2 (the hella_shim.cpp file)"""
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    sec = text[text.index("## 2."):text.index("## 3.")]
    blocks = re.findall(r"```cpp\n(.*?)```", sec, re.S)
    shim = [b for b in blocks if b.startswith("// hella_shim.cpp")]
    assert len(shim) == 1
    return shim[0]


def _compile(tmp_path, name, src):
    path = tmp_path / name
    path.write_text(src)
    obj = str(path) + ".o"
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-c", str(path), "-o", obj,
                        "-I", os.path.join(ROOT, "include"), "-I", str(tmp_path)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return obj


@pytest.fixture
def standin(tmp_path):
    d = tmp_path / "hella"
    d.mkdir()
    (d / "definitions.h").write_text(DEFINITIONS_STANDIN)
    return tmp_path


def test_reference_dedisp_calls_compile(standin):
    _compile(standin, "ref_calls.cpp", REFERENCE_CALLS)


def test_integration_shim_compiles(standin):
    _compile(standin, "hella_shim.cpp", shim_source())


@pytest.mark.skipif(not os.path.exists(os.path.join(LIB_DIR, "libhella_amd.so")),
                    reason="library not built")
def test_shim_and_calls_link_against_the_library(standin):
    objs = [_compile(standin, "ref_calls.cpp", REFERENCE_CALLS),
            _compile(standin, "hella_shim.cpp", shim_source()),
            _compile(standin, "main.cpp", MAIN)]
    exe = str(standin / "linked")
    r = subprocess.run(["g++"] + objs + ["-o", exe, "-L", LIB_DIR, "-lhella_amd",
                                         "-Wl,--allow-shlib-undefined",
                                         "-Wl,--no-as-needed"],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    nm = subprocess.run(["nm", "-u", exe], capture_output=True, text=True, check=True).stdout
    used = {ln.split()[-1] for ln in nm.splitlines() if ln.strip()}
    for sym in ("dedisp_create_plan", "dedisp_execute_adv", "hella_fast_flagger",
                "hella_format_coincidencer", "hella_dada_coincidencer_fields"):
        assert sym in used
