"""Dedispersion (A11) on the GPU through the dedisp-compatible C ABI,
bit-exact against the oracle's brute-force sums."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _plan(hamd, orc, nch, dt, f0, df, lo, hi, ti=64.0, tol=1.25):
    lp = hamd.DedispPlan(nch, dt, f0, df)
    lp.generate_dm_list(lo, hi, ti, tol)
    op = orc.Plan(nch, dt, f0, df, lo, hi, ti, tol)
    assert lp.max_delay == op.max_delay
    return lp, op


@pytest.mark.parametrize("nch,nextra,hi", [(64, 4000, 300.0), (96, 777, 150.0), (256, 5000, 500.0)])
def test_random_bit_exact(hamd, orc, nch, nextra, hi):
    rng = np.random.default_rng(nch + nextra)
    lp, op = _plan(hamd, orc, nch, 256e-6, 500.0, -100.0 / nch, 0.0, hi)
    ns = op.max_delay + nextra
    data = rng.integers(0, 256, size=(ns, nch), dtype=np.uint8)
    got = lp.execute_adv(data)
    ref = orc.dedisperse_tf(data, op.delta, ns - op.max_delay)
    np.testing.assert_array_equal(got, ref.astype(np.float32))


def test_killmask_and_guru(hamd, orc):
    rng = np.random.default_rng(7)
    nch = 128
    lp, op = _plan(hamd, orc, nch, 256e-6, 500.0, -0.78125, 0.0, 400.0)
    ns = op.max_delay + 3001
    data = rng.integers(0, 256, size=(ns, nch), dtype=np.uint8)
    km = (rng.random(nch) > 0.1).astype(np.int32)
    lp.set_killmask(km)
    got = lp.execute_adv(data)
    ref = orc.dedisperse_tf(data, op.delta, ns - op.max_delay, killmask=km)
    np.testing.assert_array_equal(got, ref.astype(np.float32))
    sub = lp.execute_guru(data, 17, 40)
    np.testing.assert_array_equal(sub, ref[17:57].astype(np.float32))
    lp.set_killmask(np.zeros(nch, np.int32))
    assert not lp.execute_adv(data).any()


def test_closed_forms(hamd, orc):
    nch = 64
    lp, op = _plan(hamd, orc, nch, 256e-6, 500.0, -1.5625, 0.0, 200.0)
    ns = op.max_delay + 1500
    nout = ns - op.max_delay
    # constant input v -> n_live * v everywhere
    got = lp.execute_adv(np.full((ns, nch), 200, np.uint8))
    assert np.all(got == nch * 200)
    # single impulse at (t0, c) -> appears at t0 - delay(d,c) in every DM row
    t0, c = op.max_delay + 700, 37
    data = np.zeros((ns, nch), np.uint8)
    data[t0, c] = 255
    got = lp.execute_adv(data)
    for d in range(op.dms.size):
        row = got[d]
        pos = t0 - op.delta[d, c]
        if 0 <= pos < nout:
            assert row[pos] == 255 and row.sum() == 255
        else:
            assert row.sum() == 0
    # DM = 0 row is the plain channel sum of each sample
    rng = np.random.default_rng(3)
    data = rng.integers(0, 256, size=(ns, nch), dtype=np.uint8)
    got = lp.execute_adv(data)
    np.testing.assert_array_equal(got[0], data[:nout].sum(axis=1).astype(np.float32))


def test_c1_geometry(hamd, orc):
    """256 channels, 500->400 MHz, 256 us, DM 0-500: the CPU-runnable workload."""
    rng = np.random.default_rng(20261204)
    lp, op = _plan(hamd, orc, 256, 256e-6, 500.0, -0.390625, 0.0, 500.0)
    ns = 32768
    data = np.clip(np.rint(rng.normal(128, 16, size=(ns, 256))), 0, 255).astype(np.uint8)
    got = lp.execute_adv(data)
    ref = orc.dedisperse_tf(data, op.delta, ns - op.max_delay)
    np.testing.assert_array_equal(got, ref.astype(np.float32))


def test_stress_geometry_full_size_spot_checks(hamd, orc):
    """The 4096-channel stress geometry at full size through the dedisp
    interface with device pointers: 4096 ch over 500->400 MHz, 65.536 us,
    DM 0-5000 (7511 trials, max delay 712492), 262144 output samples per
    trial (8.1e12 integer adds).  Random bytes in; 2048 (trial, sample)
    entries spread over 64 trials (the first and the last among them) are
    checked against direct sums of the input at the oracle's delays, and the
    whole row of trial 0 against the channel sums."""
    nch, dt, nout = 4096, 65.536e-6, 262144
    lp, op = _plan(hamd, orc, nch, dt, 500.0, -100.0 / nch, 0.0, 5000.0)
    assert op.dms.size == 7511 and op.max_delay == 712492
    ns = op.max_delay + nout
    data = np.frombuffer(np.random.default_rng(5000).bytes(ns * nch), np.uint8).reshape(ns, nch)
    din = hamd.DeviceBuffer(data.nbytes)
    din.upload(data)
    ndm = op.dms.size
    dout = hamd.DeviceBuffer(ndm * nout * 4)
    print(f"# stress geometry: ns={ns} nch={nch} ndm={ndm}", flush=True)
    code = hamd.lib().dedisp_execute_adv(lp._p, ns, din.ptr, 8, nch, dout.ptr, 32, nout * 4, 1 << 2)
    assert code == 0
    hamd.device_sync()
    rng = np.random.default_rng(7)
    rows = np.unique(np.concatenate([[0, ndm - 1], rng.choice(ndm, 62, replace=False)]))
    chan = np.arange(nch)
    for d in rows:
        row = np.zeros(nout, np.float32)
        hamd.lib().hella_memcpy(row.ctypes.data, dout.ptr + int(d) * nout * 4, nout * 4, 2)
        ts = rng.integers(0, nout, 32)
        for t in ts:
            want = int(data[t + op.delta[d], chan].astype(np.int64).sum())
            assert row[t] == want, (d, t)
        if d == 0:
            # DM 0: every sample of the row is the channel sum of that sample
            assert not op.delta[0].any()
            want_row = np.zeros(nout, np.int64)
            for t0 in range(0, nout, 16384):
                want_row[t0:t0 + 16384] = data[t0:t0 + 16384].sum(axis=1, dtype=np.int64)
            np.testing.assert_array_equal(row, want_row.astype(np.float32))
    din.free()
    dout.free()


ERR = ["NO_ERROR", "MEM_ALLOC_FAILED", "MEM_COPY_FAILED", "NCHANS_EXCEEDS_LIMIT", "INVALID_PLAN",
       "INVALID_POINTER", "INVALID_STRIDE", "NO_DM_LIST_SET", "TOO_FEW_NSAMPS",
       "INVALID_FLAG_COMBINATION", "UNSUPPORTED_IN_NBITS", "UNSUPPORTED_OUT_NBITS",
       "INVALID_DEVICE_INDEX", "DEVICE_ALREADY_SET", "PRIOR_GPU_ERROR", "INTERNAL_GPU_ERROR",
       "UNKNOWN_ERROR"]
E = {n: i for i, n in enumerate(ERR)}


def test_error_codes(hamd, orc):
    """invalid calls return dedisp's error values (hella turns any non-zero
    value into an exception, src/hella/src/dedisperse.cpp:40-46) and leave
    the plan usable"""
    import ctypes
    L = hamd.lib()
    nch, nout = 32, 100
    lp, op = _plan(hamd, orc, nch, 256e-6, 500.0, -3.125, 0.0, 100.0)
    md, ndm = op.max_delay, op.dms.size
    ns = md + nout
    rng = np.random.default_rng(11)
    data = rng.integers(0, 256, size=(ns, nch), dtype=np.uint8)
    din = hamd.DeviceBuffer(data.nbytes)
    din.upload(data)
    dout = hamd.DeviceBuffer(ndm * nout * 4)
    DP, HP, WAIT, ASYNC = 1 << 2, 1 << 1, 1 << 3, 1 << 4

    def ex(plan=lp._p, n=ns, i=None, ib=8, ist=nch, o=None, ob=32, ost=nout * 4, fl=DP):
        return L.dedisp_execute_adv(plan, n, din.ptr if i is None else i, ib, ist,
                                    dout.ptr if o is None else o, ob, ost, fl)

    assert ex(plan=None) == E["INVALID_PLAN"]
    assert ex(i=0) == E["INVALID_POINTER"] and ex(o=0) == E["INVALID_POINTER"]
    assert ex(fl=DP | HP) == E["INVALID_FLAG_COMBINATION"]
    assert ex(fl=DP | WAIT | ASYNC) == E["INVALID_FLAG_COMBINATION"]
    assert ex(ib=4) == E["UNSUPPORTED_IN_NBITS"] and ex(ib=32) == E["UNSUPPORTED_IN_NBITS"]
    assert ex(ob=16) == E["UNSUPPORTED_OUT_NBITS"]
    assert ex(n=md) == E["TOO_FEW_NSAMPS"] and ex(n=0) == E["TOO_FEW_NSAMPS"]
    assert ex(ist=nch - 1) == E["INVALID_STRIDE"]
    assert ex(ost=nout * 4 - 4) == E["INVALID_STRIDE"] and ex(ost=nout * 4 + 2) == E["INVALID_STRIDE"]
    L.dedisp_get_error_string.restype = ctypes.c_char_p
    for i in range(len(ERR)):
        assert L.dedisp_get_error_string(i)
    # after every rejected call the plan still computes the right sums
    assert ex() == 0
    hamd.device_sync()
    got = np.zeros((ndm, nout), np.float32)
    L.hella_memcpy(got.ctypes.data, dout.ptr, got.nbytes, 2)
    ref = orc.dedisperse_tf(data, op.delta, nout)
    np.testing.assert_array_equal(got, ref.astype(np.float32))
    lp.close()
    # a plan without a DM list; invalid plans and DM lists
    fresh = hamd.DedispPlan(nch, 256e-6, 500.0, -3.125)
    assert ex(plan=fresh._p) == E["NO_DM_LIST_SET"]
    bad = (ctypes.c_float * 3)(0.0, 20.0, 10.0)
    assert L.dedisp_set_dm_list(fresh._p, bad, 3) == E["UNKNOWN_ERROR"]
    assert L.dedisp_set_dm_list(fresh._p, bad, 0) == E["INVALID_POINTER"]
    assert L.dedisp_generate_dm_list(fresh._p, ctypes.c_float(50.0), ctypes.c_float(10.0),
                                     ctypes.c_float(64.0), ctypes.c_float(1.25)) == E["UNKNOWN_ERROR"]
    fresh.close()
    p = ctypes.c_void_p()
    assert L.dedisp_create_plan(ctypes.byref(p), 0, ctypes.c_float(256e-6), ctypes.c_float(500.0),
                                ctypes.c_float(-1.0)) == E["UNKNOWN_ERROR"]
    assert L.dedisp_create_plan(ctypes.byref(p), 1 << 17, ctypes.c_float(256e-6),
                                ctypes.c_float(500.0), ctypes.c_float(-1e-4)) == E["NCHANS_EXCEEDS_LIMIT"]


def test_custom_dm_lists(hamd, orc):
    """dedisp_set_dm_list: a single zero DM (every sample's channel sum, no
    delay) and the first trials of a generated list (the oracle's delays of
    those trials)"""
    nch = 48
    lp, op = _plan(hamd, orc, nch, 256e-6, 500.0, -2.0, 0.0, 300.0)
    rng = np.random.default_rng(12)
    ns = op.max_delay + 1234
    data = rng.integers(0, 256, size=(ns, nch), dtype=np.uint8)
    sub = hamd.DedispPlan(nch, 256e-6, 500.0, -2.0)
    sub.set_dm_list(np.array([0.0], np.float32))
    assert sub.dm_count == 1 and sub.max_delay == 0
    got = sub.execute_adv(data)
    np.testing.assert_array_equal(got[0], data.sum(axis=1).astype(np.float32))
    k = 7
    sub.set_dm_list(op.dms[:k].astype(np.float32))
    md = int(op.delta[:k].max())
    assert sub.dm_count == k and sub.max_delay == md
    got = sub.execute_adv(data)
    ref = orc.dedisperse_tf(data, op.delta[:k], ns - md)
    np.testing.assert_array_equal(got, ref.astype(np.float32))
    sub.close()
    lp.close()
