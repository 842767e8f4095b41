"""Candidate emission in the reference's text formats (host code, no GPU):
output_peaks' file and coincidencer message (src/hella/src/peaks.cu:129-259)
and the spectral-flag file (src/apps/casm_hella_refactored.cu:756-761)."""
import numpy as np
import pytest


def _cands(hamd):
    c = np.zeros(2, hamd.CAND_DTYPE)
    c[0] = (12.5, 2, 100, 3, 17, 250.5, 123456)
    c[1] = (7.25, 0, 4, 0, 1, 0.75, -86399)
    return c


# The expected lines are typed from the reference's format and constants:
# "%g %d %g %d %d %g %d" of (S/N, start sample, start sample * 1e-3 / 86400,
# width index, DM index, DM, beam + BEAM0).  123456 * 1e-3 / 86400 =
# 0.001428888... -> "0.00142889"; -86399 * 1e-3 / 86400 = -0.000999988...
LINE0 = "12.5 123456 0.00142889 3 17 250.5 5"
LINE1 = "7.25 -86399 -0.000999988 0 1 0.75 3"
HEADER = "SNR SAMP_START TIME_START WIDTH DM_IDX DM BEAM_IDX"


def test_candidate_file(hamd, tmp_path):
    p = tmp_path / "cands.txt"
    c = _cands(hamd)
    hamd.write_candidates(p, c, beam0=3)
    hamd.write_candidates(p, c[:1], beam0=3)
    hamd.write_candidates(p, c[:0], beam0=3)
    assert p.read_text().splitlines() == [HEADER, LINE0, LINE1, LINE0]


def test_empty_list_still_writes_header(hamd, tmp_path):
    p = tmp_path / "none.txt"
    hamd.write_candidates(p, np.zeros(0, hamd.CAND_DTYPE))
    assert p.read_text() == HEADER + "\n"


def test_coincidencer_message(hamd):
    c = _cands(hamd)
    # gulp number = (samp0 / gulp) + 1 with C division (toward zero):
    # 126608 / 65536 -> 1 -> "2";  -70000 / 65536 -> -1 -> "0"
    m = hamd.format_coincidencer(-70000 + 65536 * 3, 65536, c, beam0=3)
    assert m.splitlines() == ["2", LINE0, LINE1]
    m = hamd.format_coincidencer(-70000, 65536, c[:1], beam0=3,
                                 dada_fields="2026-12-03-10:00:00 0 0.000131072")
    assert m.splitlines() == ["0 2026-12-03-10:00:00 0 0.000131072", LINE0]


def test_coincidencer_truncates_like_snprintf(hamd):
    lib = hamd.lib()
    c = _cands(hamd)
    import ctypes as C
    buf = C.create_string_buffer(8)
    ptr = c.ctypes.data_as(C.POINTER(hamd._Cand))
    n = lib.hella_format_coincidencer(0, 10, None, ptr, 2, 3, buf, 8)
    assert n == len("1\n" + LINE0 + "\n" + LINE1 + "\n")
    assert buf.value.decode() == ("1\n" + LINE0)[:7]
    assert lib.hella_format_coincidencer(0, 0, None, ptr, 2, 3, buf, 8) == -1


def test_specflags_file(hamd, tmp_path):
    p = tmp_path / "spec.txt"
    # the reference writes its int specflags with "%d" (casm:756-761)
    hamd.write_specflags(p, [0, 5, 2147483647])
    hamd.write_specflags(p, [1])
    assert p.read_text().splitlines() == ["0", "5", "2147483647", "1"]


def test_bad_paths(hamd, tmp_path):
    with pytest.raises(hamd.HellaError):
        hamd.write_candidates(tmp_path / "no" / "such" / "dir.txt", _cands(hamd))
    with pytest.raises(hamd.HellaError):
        hamd.write_specflags(tmp_path / "no" / "dir" / "s.txt", [1, 2])
