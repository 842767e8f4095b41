"""Writes the small golden fixtures of this directory from the oracle.

The reference ships no golden vectors for its hot path, so these fixtures
freeze the oracle's outputs on small seeded inputs: the CPU tests check that
the oracle still produces them, the GPU tests check the HIP path against them
directly.  Run from the repository root after building the oracle:
    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import oracle_bind as orc  # noqa: E402

# geometry of the small pipeline fixtures (FILE sizing: one window)
FLAG_GEOM = dict(NCHAN=32, FCH1=500.0, FOFF=-3.125, TSAMP=256e-6, GULP=2048, FLAG1=3, FLAG2=5,
                 SPEC_MIN=-0.05, SPEC_MAX=0.15, SCRUNCH=((1, 1, 5.0),))


def plan_c1():
    # the CPU-runnable single-beam geometry: 256 ch over 500->400 MHz, 256 us,
    # DM 0-500, pulse width 64 us, tolerance 1.25
    p = orc.Plan(256, 256e-6, 500.0, -0.390625, 0.0, 500.0, 64.0, 1.25)
    return dict(delay=p.delay, dms=p.dms, max_delay=np.int64(p.max_delay),
                delta_rows=p.delta[[0, 1, 100, p.dms.size - 1]])


def dedisp_small():
    rng = np.random.default_rng(20261203)
    p = orc.Plan(16, 256e-6, 500.0, -6.25, 0.0, 60.0, 64.0, 1.25)
    ns = p.max_delay + 600
    data = rng.integers(0, 256, size=(ns, 16), dtype=np.uint8)
    km = np.ones(16, np.int32)
    km[[2, 9]] = 0
    out = orc.dedisperse_tf(data, p.delta, ns - p.max_delay, killmask=km)
    return dict(data=data, killmask=km, delta=p.delta, dms=p.dms, out=out)


def flag_small():
    rng = np.random.default_rng(20261204)
    g = FLAG_GEOM
    nch, W = g["NCHAN"], g["GULP"]
    x = np.clip(np.rint(rng.normal(128.0, 16.0, size=(W, nch))), 0, 255)
    x[:, 7] = np.clip(np.rint(rng.normal(128.0, 160.0, size=W)), 0, 255)   # a noisy channel
    x[1000:1010, :] = np.clip(x[1000:1010, :] + 96, 0, 255)                # a broadband burst
    data = x.astype(np.uint8)
    out, cnt = orc.flag_u8(data, g["SPEC_MIN"], g["SPEC_MAX"], g["FLAG1"],
                        g["FLAG2"], g["SCRUNCH"])
    return dict(data=data, out=out, counts=cnt)


def search_small():
    rng = np.random.default_rng(20261205)
    ndm, nout, maxW = 24, 3000, 16
    D = rng.integers(900, 1100, size=(ndm, nout + maxW)).astype(np.uint32)
    D[10:14, 1500:1508] += 400
    res = {}
    for mode, name in ((orc.MODE_BOXCAR, "boxcar"), (orc.MODE_HELLA, "hella")):
        r = orc.search(D, nout, maxW, 1, mode, 6.0, 1000)
        res[f"{name}_cands"] = r["cands"]
        res[f"{name}_mu"] = r["mu"]
        res[f"{name}_sigma"] = r["sigma"]
        res[f"{name}_sigma0"] = np.float32(r["sigma0"])
    res["D"] = D
    return res


FIXTURES = {"plan_c1": plan_c1, "dedisp_small": dedisp_small, "flag_small": flag_small,
            "search_small": search_small}

if __name__ == "__main__":
    for name, fn in FIXTURES.items():
        path = os.path.join(HERE, name + ".npz")
        np.savez_compressed(path, **fn())
        print(path, os.path.getsize(path))
