"""Committed autotuner records: each one parses, names a shape the tuner
measures, and maps onto the set_tile encoding of the C ABI (include/hella.h:
order bit 0 the tile order, bits 1-2 the loader, bit 3 the read-ahead
pipeline, bit 4 the two-ahead LDS DMA)."""
import glob
import importlib.util
import json
import os

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
RECORDS = sorted(glob.glob(os.path.join(ROOT, "casm-hella_amd", "tuning", "*.json")))


def _pkg():
    spec = importlib.util.spec_from_file_location(
        "hella_amd_tuning", os.path.join(ROOT, "casm-hella_amd", "hella_amd.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_records_present():
    names = {os.path.basename(p) for p in RECORDS}
    for want in ("plain_1024ch_dm2000.json", "batched_4x1024ch_dm2000.json",
                 "uneven_32beams.json", "stress_4096ch_dm5000.json"):
        assert want in names


@pytest.mark.parametrize("path", RECORDS, ids=os.path.basename)
def test_record_is_best_of_its_trials(path):
    d = json.load(open(path))
    best, trials = d["best"], d["trials"]
    assert trials, "a record lists the shapes it measured"
    assert best["ms"] == min(t["ms"] for t in trials)
    assert any(all(t[k] == best[k] for k in ("lanes_per_dm", "chunk", "order", "dma"))
               for t in trials)


def test_tile_args_encoding():
    m = _pkg()
    enc = lambda dma, order=0: m.tile_args({"lanes_per_dm": 64, "chunk": 4, "order": order,
                                             "dma": dma})[2]
    # loader field: 0 LDS DMA, 1 registers, 2 register prefetch, 3 one trial per wave
    assert enc(1) == 0
    assert enc(0) == 2
    assert enc(2) == 4
    assert enc(3) == 6
    assert enc(4) == 6 | 8
    assert enc(5) == 6 | 16
    assert enc(4, order=1) == 15
    with pytest.raises(ValueError):
        enc(6)


@pytest.mark.parametrize("path", RECORDS, ids=os.path.basename)
def test_record_maps_to_valid_shape(path):
    m = _pkg()
    lanes, chunk, order = m.tile_args(json.load(open(path))["best"])
    loader = (order >> 1) & 3
    if loader == 3:
        assert lanes == 64 and chunk in (4, 8)
    else:
        assert lanes in (8, 16, 32, 64) and chunk in (8, 16, 32)
