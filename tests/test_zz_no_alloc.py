"""dedisp_execute_adv_noalloc (the entry point that allocates nothing: its
scratch is a fixed device array per GPU) against dedisp_execute_adv on the
same device buffers, through the C ABI: identical sums on two calls in a
row, the second on a smaller input, also after another plan has used the
shared scratch; and the oracle's sums; requests beyond the scratch are
refused before any GPU work."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DEV = 1 << 2  # DEDISP_DEVICE_POINTERS


def _plan(hamd, nch, dt, f0, df, dm_max, killmask=None):
    lp = hamd.DedispPlan(nch, dt, f0, df)
    lp.generate_dm_list(0.0, dm_max, 64.0, 1.25)
    if killmask is not None:
        lp.set_killmask(killmask)
    return lp


def _run(hamd, lp, data, noalloc):
    ns, nch = data.shape
    nout = ns - lp.max_delay
    d_in = hamd.DeviceBuffer(data.nbytes)
    d_in.upload(data)
    d_out = hamd.DeviceBuffer(lp.dm_count * nout * 4)
    d_out.upload(np.full(lp.dm_count * nout, 0x7f, np.uint32))   # not zero: every value is written
    rc = lp.execute_device(d_in.ptr, ns, nch, d_out.ptr, nout * 4, noalloc=noalloc)
    assert rc == 0, hamd.lib().dedisp_get_error_string(rc)
    out = d_out.download(np.zeros((lp.dm_count, nout), np.float32))
    d_in.free()
    d_out.free()
    return out


def test_noalloc_equals_execute_adv(hamd, orc):
    rng = np.random.default_rng(7)
    # the CPU-runnable single-beam geometry: 256 ch over 500->400 MHz, 256 us, DM 0-500
    km = np.ones(256, np.int32)
    km[rng.choice(256, 20, replace=False)] = 0
    lp = _plan(hamd, 256, 256e-6, 500.0, -0.390625, 500.0, km)
    other = _plan(hamd, 64, 256e-6, 500.0, -1.5625, 120.0)
    md = lp.max_delay
    delta = orc.delays(lp.dm_list(), orc.delay_table(256, 256e-6, 500.0, -0.390625))
    for ns in (md + 16384, md + 3000):          # two calls in a row, the second smaller
        data = rng.integers(0, 256, size=(ns, 256), dtype=np.uint8)
        a = _run(hamd, lp, data, noalloc=True)
        b = _run(hamd, lp, data, noalloc=False)
        np.testing.assert_array_equal(a, b)
        ref = orc.dedisperse_tf(data, delta, ns - md, killmask=km)
        np.testing.assert_array_equal(a, ref.astype(np.float32))
        # another plan in between overwrites the shared scratch
        d2 = rng.integers(0, 256, size=(other.max_delay + 2000, 64), dtype=np.uint8)
        np.testing.assert_array_equal(_run(hamd, other, d2, True), _run(hamd, other, d2, False))
    lp.close()
    other.close()


def test_noalloc_refuses_what_does_not_fit(hamd):
    lp = _plan(hamd, 1024, 131.072e-6, 500.0, -100.0 / 1024, 2000.0)
    ns = lp.max_delay + 65536                     # ~213 MB channel-major copy > 64 MiB
    # refused on the host before any device work: the pointers are never used
    rc = lp.execute_device(16, ns, 1024, 16, 65536 * 4, noalloc=True)
    assert rc == 1                                # DEDISP_MEM_ALLOC_FAILED
    assert b"scratch" in hamd.lib().hella_last_error()
    # host pointers would need staging memory: refused
    assert lp.execute_device(16, ns, 1024, 16, 65536 * 4, noalloc=True,
                             flags=1 << 1) == 9   # DEDISP_INVALID_FLAG_COMBINATION
    lp.close()
