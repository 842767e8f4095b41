"""The oracle still produces the committed golden fixtures (tests/golden,
written by tests/golden/make_golden.py from seeded inputs)."""
import os
import sys

import numpy as np
import pytest

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
sys.path.insert(0, HERE)
import make_golden  # noqa: E402


@pytest.mark.parametrize("name", sorted(make_golden.FIXTURES))
def test_oracle_reproduces_fixture(name):
    want = np.load(os.path.join(HERE, name + ".npz"), allow_pickle=False)
    got = make_golden.FIXTURES[name]()
    assert sorted(want.files) == sorted(got)
    for k in want.files:
        np.testing.assert_array_equal(np.asarray(got[k]), want[k], err_msg=f"{name}:{k}")


def test_fixture_invariants():
    p = np.load(os.path.join(HERE, "plan_c1.npz"))
    # DM list ascending from 0 past 500; max delay = round(last DM * delay[last])
    assert p["dms"][0] == 0.0 and np.all(np.diff(p["dms"]) > 0) and p["dms"][-1] >= 500.0
    assert p["max_delay"] == p["delta_rows"][-1].max()
    d = np.load(os.path.join(HERE, "dedisp_small.npz"))
    # the DM-0 row is the sum of the live channels of each sample
    live = d["killmask"] != 0
    nout = d["out"].shape[1]
    np.testing.assert_array_equal(d["out"][0], d["data"][:nout, live].sum(axis=1))
