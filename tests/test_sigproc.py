"""SIGPROC filterbank input (src/hella/src/sigproc.cpp:26-191): the keyword
header and the uint8 samples behind it (host code, no GPU)."""
import struct

import numpy as np
import pytest


def _s(text):
    b = text.encode()
    return struct.pack("<i", len(b)) + b


def _write(path, params, data, extra=b""):
    """A SIGPROC file as the reference's tools write it (length-prefixed
    keyword strings, little-endian int/double values)."""
    out = _s("HEADER_START")
    for k, v in params:
        out += _s(k)
        if isinstance(v, str):
            out += _s(v)
        elif isinstance(v, float):
            out += struct.pack("<d", v)
        elif isinstance(v, bytes):
            out += v
        else:
            out += struct.pack("<i", v)
    out += extra + _s("HEADER_END")
    path.write_bytes(out + np.ascontiguousarray(data, np.uint8).tobytes())
    return len(out)


PARAMS = [("rawdatafile", "burst.fil"), ("source_name", "FRB20261203"), ("machine_id", 11),
          ("telescope_id", 20), ("data_type", 1), ("fch1", 500.0), ("foff", -0.390625),
          ("nchans", 256), ("nbits", 8), ("tstart", 60000.25), ("tsamp", 256e-6), ("nifs", 1),
          ("nbeams", 4), ("ibeam", 2), ("src_raj", 123456.7), ("src_dej", -12345.6),
          ("az_start", 10.5), ("za_start", 20.25), ("barycentric", 0), ("pulsarcentric", 0),
          ("refdm", 0.0), ("period", 0.0), ("signed", b"\x01"), ("npuls", struct.pack("<q", 7))]


def test_header_and_samples(hamd, tmp_path):
    rng = np.random.default_rng(3)
    data = rng.integers(0, 256, size=(1000, 256), dtype=np.uint8)
    p = tmp_path / "a.fil"
    nhdr = _write(p, PARAMS, data)
    f = hamd.FilterbankFile(p)
    assert f.header_bytes == nhdr and f.nsamples == 1000
    assert (f.nchans, f.nbits, f.nifs, f.nbeams, f.ibeam) == (256, 8, 1, 4, 2)
    assert (f.machine_id, f.telescope_id, f.data_type) == (11, 20, 1)
    assert (f.fch1, f.foff, f.tsamp, f.tstart) == (500.0, -0.390625, 256e-6, 60000.25)
    assert (f.src_raj, f.src_dej, f.az_start, f.za_start) == (123456.7, -12345.6, 10.5, 20.25)
    assert f.source_name == "FRB20261203" and f.raw_data_file == "burst.fil"
    assert f.npuls == 7 and f.isign == "\x01"
    np.testing.assert_array_equal(f.read(0, 1000), data)
    np.testing.assert_array_equal(f.read(900, 300), data[900:])     # short at the end
    # the geometry keys, printed so that they parse back to the same doubles
    kv = dict(line.split() for line in f.config().splitlines())
    assert kv["NCHAN"] == "256"
    assert (float(kv["FCH1"]), float(kv["FOFF"]), float(kv["TSAMP"])) == (500.0, -0.390625, 256e-6)


def test_frequency_table(hamd, tmp_path):
    p = tmp_path / "t.fil"
    params = [("nchans", 3), ("nbits", 8), ("tsamp", 1e-3), ("fch1", 400.0)]
    extra = (_s("FREQUENCY_START") + _s("fchannel") + struct.pack("<d", 500.0) + _s("fchannel") +
             struct.pack("<d", 450.0) + _s("fchannel") + struct.pack("<d", 400.0) +
             _s("FREQUENCY_END"))
    _write(p, params, np.zeros((5, 3), np.uint8), extra)
    f = hamd.FilterbankFile(p)
    assert f.nfreq == 3 and list(f.hdr.frequency_table[:3]) == [500.0, 450.0, 400.0]
    assert f.fch1 == 0.0 and f.foff == 0.0


def test_no_header_and_errors(hamd, tmp_path):
    raw = tmp_path / "raw.bin"
    raw.write_bytes(struct.pack("<i", 3) + b"abc" + bytes(100))
    f = hamd.FilterbankFile(raw)
    assert f.header_bytes == 0
    bad = tmp_path / "bad.fil"
    _write(bad, [("nchans", 4), ("bogus_key", 1)], np.zeros((2, 4), np.uint8))
    with pytest.raises(hamd.HellaError, match="unknown parameter: bogus_key"):
        hamd.FilterbankFile(bad)
    cut = tmp_path / "cut.fil"
    cut.write_bytes(_s("HEADER_START") + _s("nchans"))
    with pytest.raises(hamd.HellaError):
        hamd.FilterbankFile(cut)
    b16 = tmp_path / "b16.fil"
    _write(b16, [("nchans", 4), ("nbits", 16)], np.zeros((2, 8), np.uint8))
    with pytest.raises(hamd.HellaError, match="8-bit"):
        hamd.FilterbankFile(b16).read(0, 1)
