"""The C-ABI library loads without a GPU and exports every function that
include/*.h declares (CPU only: no compute calls)."""
import ctypes
import glob
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "casm-hella_amd", "lib", "libhella_amd.so")


def declared_functions():
    """Function names declared in the public headers (after the preprocessor,
    so comments and macros are gone)."""
    names = set()
    for h in sorted(glob.glob(os.path.join(ROOT, "include", "*.h"))):
        text = subprocess.run(["gcc", "-E", "-P", "-x", "c", h], check=True,
                              capture_output=True, text=True).stdout
        # join statements and look for "<type> name(" outside typedefs
        for stmt in re.sub(r"\s+", " ", text).split(";"):
            stmt = stmt.strip()
            if not stmt or stmt.startswith("typedef") or "{" in stmt:
                continue
            m = re.match(r"^[\w\s\*]+?\b(\w+)\s*\(", stmt)
            if m and not m.group(1).startswith("__"):   # system headers
                names.add(m.group(1))
    return sorted(names)


def test_headers_declare_the_boundary():
    names = declared_functions()
    for must in ("dedisp_create_plan", "dedisp_generate_dm_list", "dedisp_execute_adv",
                 "dedisp_set_killmask", "dedisp_get_max_delay", "hella_create",
                 "hella_fast_flagger", "hella_dedisperse", "hella_smooth", "hella_find_peaks",
                 "hella_clear_peaks", "hella_process_gulp", "hella_process_gulp_batched"):
        assert must in names
    assert len(names) >= 55


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_library_exports_every_declared_function():
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, f"declared but not exported: {missing}"
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], check=True, capture_output=True,
                         text=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    # no C++ symbol leaks into the C boundary's namespace
    for n in declared_functions():
        assert n in exported


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_error_paths_need_no_gpu():
    lib = ctypes.CDLL(LIB)
    lib.dedisp_get_error_string.restype = ctypes.c_char_p
    assert lib.dedisp_get_error_string(0)
    lib.hella_create.restype = ctypes.c_int
    h = ctypes.c_void_p()
    assert lib.hella_create(b"NBEAM 1\n", None) == 1                 # HELLA_ERR_INVALID
    assert lib.hella_create(b"NBEAM 0\n", ctypes.byref(h)) == 2      # HELLA_ERR_CONFIG
    assert not h.value
    lib.hella_last_error.restype = ctypes.c_char_p
    assert lib.hella_last_error()
