"""Block ring (PSRDADA-style ingest of the streaming server): ordering, block
metadata, the writer waiting for the oldest block to be released, end of
stream and abort.  Runs on the CPU (pageable blocks without a GPU)."""
import threading
import time

import numpy as np
import pytest


def test_header_order_and_metadata(hamd):
    r = hamd.Ring(4, 32, "HDR_VERSION 1.0\nNCHAN 16\n")
    assert r.header() == "HDR_VERSION 1.0\nNCHAN 16\n"
    pool = [np.full(32, i, np.uint8) for i in range(3)]
    ev = [(0.0005 * i, 10 + i % 3, i % 3, i // 3, hamd.RING_LAST if i >= 9 else 0)
          for i in range(12)]
    got = []

    def reader():
        while True:
            b = r.read(50)
            if b is None:
                continue
            if b.get("end"):
                return
            got.append(b)
            r.release(b["index"])

    th = threading.Thread(target=reader)
    th.start()
    r.replay(ev, pool, 1.0)
    th.join(10)
    assert not th.is_alive()
    assert [(b["beam"], b["seq"], b["flags"]) for b in got] == [(e[1], e[3], e[4]) for e in ev]
    assert all(np.all(b["data"] == e[2]) for b, e in zip(got, ev))
    t = [b["t_arrival_ns"] for b in got]
    assert t == sorted(t)
    assert r.read(0) == {"end": True}


def test_writer_waits_for_oldest_release(hamd):
    r = hamd.Ring(2, 8)
    r.write(0, 0, np.zeros(8, np.uint8))
    r.write(0, 1, np.ones(8, np.uint8))
    with pytest.raises(hamd.HellaError):          # full: times out
        r.write(0, 2, np.zeros(8, np.uint8), timeout_ms=20)
    a, b = r.read(), r.read()
    assert (a["seq"], b["seq"]) == (0, 1)
    r.release(b["index"])                          # newer released first: still full
    with pytest.raises(hamd.HellaError):
        r.write(0, 2, np.zeros(8, np.uint8), timeout_ms=20)
    r.release(a["index"])                          # oldest released: both free
    r.write(0, 2, np.full(8, 2, np.uint8), timeout_ms=20)
    r.write(0, 3, np.full(8, 3, np.uint8), timeout_ms=20)
    c = r.read()
    assert c["seq"] == 2 and np.all(c["data"] == 2)


def test_abort_releases_a_blocked_writer(hamd):
    r = hamd.Ring(1, 8)
    r.write(1, 0, np.zeros(8, np.uint8))
    err = {}

    def writer():
        try:
            r.write(1, 1, np.zeros(8, np.uint8))   # blocks: the only block is held
        except hamd.HellaError as e:
            err["e"] = e

    th = threading.Thread(target=writer)
    th.start()
    time.sleep(0.05)
    assert th.is_alive()
    lib = hamd.lib()
    assert lib.hella_ring_abort(r.handle) == 0
    th.join(5)
    assert not th.is_alive() and "e" in err


def test_oversized_block_is_rejected(hamd):
    r = hamd.Ring(2, 8)
    with pytest.raises(ValueError):
        r.write(0, 0, np.zeros(16, np.uint8))
    # the C side refuses a close that claims more than a block
    blk = hamd.C.c_void_p()
    lib = hamd.lib()
    assert lib.hella_ring_open_write(r.handle, 0, hamd.C.byref(blk)) == 0
    assert lib.hella_ring_close_write(r.handle, 0, 0, 0, 16) == 1   # HELLA_ERR_INVALID
