"""PSRDADA header parsing (src/hella/src/dada.cpp:50-79 semantics) and its
use in the coincidencer message (peaks.cu:212-231); host code, no GPU."""
import pytest

HEADER = (b"HDR_VERSION 1.0\n"
          b"UTC_START 2026-12-04-08:58:00\n"
          b"\n"
          b"PICOSECONDS   123000000\n"         # value after the run of spaces
          b"TSAMP 131.072\n"
          b"NCHAN 1024\n"
          b"NCHAN 3072\n"                      # a repeated key keeps its last value
          b"SOURCE two words  \n"              # the rest of the line, trailing spaces too
          b"NOSPACE\n"                         # invalid line: the parse ends here
          b"AFTER 1\n"
          b"\x00IGNORED 7\n")                  # text after the NUL is not header


def test_header_values(hamd):
    assert hamd.dada_header_value(HEADER, "UTC_START") == "2026-12-04-08:58:00"
    assert hamd.dada_header_value(HEADER, "PICOSECONDS") == "123000000"
    assert hamd.dada_header_value(HEADER, "NCHAN") == "3072"
    assert hamd.dada_header_value(HEADER, "SOURCE") == "two words  "
    assert hamd.dada_header_value(HEADER, "HDR_VERSION") == "1.0"
    for absent in ("AFTER", "IGNORED", "NOSPACE", "TSAMP2", ""):
        assert hamd.dada_header_value(HEADER, absent) is None
    # a key followed by spaces only: the empty value
    assert hamd.dada_header_value(b"KEY   \nX 1\n", "KEY") == ""
    assert hamd.dada_header_value(b"KEY   \nX 1\n", "X") == "1"
    # no trailing newline, no NUL: the length bounds the text
    assert hamd.dada_header_value(b"A 1\nB 22", "B") == "22"
    assert hamd.dada_header_value(b"A 1\nB 22\n", "A") == "1"


def test_coincidencer_message_from_header(hamd):
    fields = hamd.dada_coincidencer_fields(HEADER)
    assert fields == "2026-12-04-08:58:00 123000000 131.072"
    with pytest.raises(hamd.HellaError):
        hamd.dada_coincidencer_fields(b"UTC_START x\nTSAMP 1\n")
    import numpy as np
    c = np.zeros(1, hamd.CAND_DTYPE)
    c[0] = (9.5, 0, 100, 2, 17, 250.5, 3 * 65536 + 100)
    msg = hamd.format_coincidencer(3 * 65536, 65536, c, beam0=4, dada_fields=fields)
    first, line = msg.splitlines()
    assert first == "4 2026-12-04-08:58:00 123000000 131.072"
    assert line.split() == ["9.5", str(3 * 65536 + 100),
                            "%g" % ((3 * 65536 + 100) * 1e-3 / 86400), "2", "17", "250.5", "4"]


def test_ring_header_feeds_the_message(hamd):
    """the ring's header (the text a PSRDADA reader copies first) gives the
    coincidencer fields"""
    r = hamd.Ring(2, 64, HEADER.split(b"\x00")[0].decode())
    try:
        assert hamd.dada_coincidencer_fields(r.header()) == "2026-12-04-08:58:00 123000000 131.072"
    finally:
        r.close()
