"""The boundary's behaviour contracts, run on the GPU through libsemper_c.so
(the Path C routes through libsemper_dev.so, whose switch turns the
global-motion search off): cancel from another thread and from the progress
callback, per-engine cancel independence and the stale cancel cleared on
entry (tests/c/contract.c:300-322, test_full_field_contracts.cpp:326), the
final 100% tick on the calling thread (:592), re-entry from the callback
refused (contract.c:203-235), a blank frame unseedable (-5,
test_full_field_contracts.cpp:631), and Path C seeding a one-row grid and
trying the next node when the centre one fails
(test_seeding_contracts.cpp:482-606)."""
import ctypes
import threading
import time

import numpy as np
import pytest

from helpers import P, engine_module, gray8, ref_speckle, stretch_rotate_aff

pytestmark = pytest.mark.gpu
sa = engine_module()
W = H = 256
PRM = sa.params((0, 0, W, H), 15, 21, 31)


@pytest.fixture(scope="module")
def lib():
    dev = sa.load_dev()
    assert dev.sdk_device_count() >= 1, "no GPU visible"
    return sa.load_c()


def scene():
    """The contract suite's pair (test_full_field_contracts.cpp:55-66): the
    speckle field of seed 11 with 600 blobs, shifted by (3, 2) px."""
    ref = gray8(ref_speckle(11, W, H, 600))
    dfm = gray8(ref_speckle(11, W, H, 600, [3.0, 2.0, 0, 0, 0, 0, W / 2, H / 2]))
    return ref, dfm


def run(lib, eng, img, prm, cb=None, user=None, cap=None):
    a = np.ascontiguousarray(img)
    cap = cap if cap is not None else 8 * 4096
    out = np.zeros(max(cap, 1), np.float32)
    met = np.zeros(24, np.float32)
    rc = lib.semper_run(eng, P(a), a.size, None, 0, ctypes.byref(prm), P(out), cap, P(met), 24,
                        cb or sa.PROGRESS_CB(), user)
    return rc, out, met


def make(lib, ref):
    e = lib.semper_create()
    assert e
    assert lib.semper_set_reference(e, P(ref), ref.size, None, 0, ref.shape[1], ref.shape[0]) == 0
    return e


def test_stale_cancel_cleared_and_engines_independent(lib):
    ref, dfm = scene()
    a, b = make(lib, ref), make(lib, ref)
    lib.semper_cancel(a)                       # engine a only
    nb, _, _ = run(lib, b, dfm, PRM)
    assert nb >= 0 and nb != sa.ERR_CANCELLED
    na, _, _ = run(lib, a, dfm, PRM)           # the stale cancel is cleared on entry
    assert na >= 0 and na != sa.ERR_CANCELLED
    lib.semper_destroy(a)
    lib.semper_destroy(b)


def big_pair():
    """A 2048 x 2048 shifted speckle pair: a solve long enough that a cancel
    from another thread lands inside it."""
    wl = sa_workloads()
    dev = sa.load_dev()
    return wl.pair(dev, "c2_2048_affine")


def sa_workloads():
    import importlib.util
    import os
    from helpers import PKG
    spec = importlib.util.spec_from_file_location("workloads", os.path.join(PKG, "workloads.py"))
    wl = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(wl)
    return wl


def test_cancel_from_another_thread(lib):
    ref, dfm = big_pair()
    eng = make(lib, ref)
    prm = sa.params((0, 0, 2048, 2048), 5, 31, 31)
    done = threading.Event()

    def canceller():
        while not done.is_set():
            lib.semper_cancel(eng)
            time.sleep(0.0002)

    th = threading.Thread(target=canceller)
    th.start()
    try:
        rc, _, _ = run(lib, eng, dfm, prm, cap=8 * 200000)
    finally:
        done.set()
        th.join()
    assert rc == sa.ERR_CANCELLED, rc
    # the handle solves again once the cancel is over
    rc2, _, _ = run(lib, eng, dfm, prm, cap=8 * 200000)
    assert rc2 > 0
    lib.semper_destroy(eng)


def test_cancel_from_the_progress_callback(lib):
    ref, dfm = big_pair()
    eng = make(lib, ref)
    prm = sa.params((0, 0, 2048, 2048), 5, 31, 31)
    ticks = []

    @sa.PROGRESS_CB
    def cb(pct, user):
        ticks.append(pct)
        lib.semper_cancel(eng)

    rc, _, _ = run(lib, eng, dfm, prm, cb=cb, cap=8 * 200000)
    assert ticks and rc == sa.ERR_CANCELLED, (rc, ticks)
    assert 100 not in ticks                    # no final tick on a cancelled solve
    lib.semper_destroy(eng)


def test_final_tick_on_the_calling_thread(lib):
    ref, dfm = big_pair()
    eng = make(lib, ref)
    prm = sa.params((0, 0, 2048, 2048), 5, 31, 31)
    calls = []
    inside = [0]
    overlap = [False]

    @sa.PROGRESS_CB
    def cb(pct, user):
        inside[0] += 1
        if inside[0] != 1:
            overlap[0] = True
        calls.append((pct, threading.get_ident()))
        time.sleep(0.001)
        inside[0] -= 1

    rc, _, _ = run(lib, eng, dfm, prm, cb=cb, cap=8 * 200000)
    me = threading.get_ident()
    assert rc > 0 and calls and not overlap[0]
    assert calls[-1] == (100, me)
    assert all(t != me for _, t in calls[:-1])
    assert all(a[0] <= b[0] for a, b in zip(calls, calls[1:]))
    lib.semper_destroy(eng)


def test_reentry_from_the_callback_is_refused(lib):
    ref, dfm = scene()
    eng = make(lib, ref)
    probe = {"calls": 0, "refused": 0}

    @sa.PROGRESS_CB
    def cb(pct, user):
        probe["calls"] += 1
        if lib.semper_set_reference(eng, P(ref), ref.size, None, 0, W, H) == sa.ERR_INIT:
            probe["refused"] += 1

    n, _, _ = run(lib, eng, dfm, PRM, cb=cb)
    assert n >= 0 and probe["calls"] > 0 and probe["refused"] == probe["calls"]
    assert lib.semper_set_reference(eng, P(ref), ref.size, None, 0, W, H) == 0   # free again
    lib.semper_destroy(eng)


def test_blank_frame_is_unseedable(lib):
    ref, _ = scene()
    eng = make(lib, ref)
    blank = np.full((H, W), 128, np.uint8)
    rc, _, _ = run(lib, eng, blank, PRM)
    assert rc == sa.ERR_SEED, rc
    lib.semper_destroy(eng)


# ---- Path C (global-motion search off, so the route is the one under test) ----

def stretch_scene():
    """A 45% stretch with Poisson 0.3 about (192, 128) on 384 x 256 (the
    reference renders it with its smooth speckle; the blob field stands in)."""
    w, h = 384, 256
    aff = stretch_rotate_aff(0.45, 0.3, 0.0, 0.0, 0.0, 192.0, 128.0)
    ref = gray8(ref_speckle(101, w, h, 1200))
    dfm = gray8(ref_speckle(101, w, h, 1200, aff))
    return ref, dfm


def dev_run(ref, dfm, prm):
    dev = sa.load_dev()
    e = dev.sdk_engine_create()
    assert dev.sdk_engine_set_flag(e, 16, 0) == 0
    h, w = ref.shape
    assert dev.sdk_engine_set_reference(e, P(np.ascontiguousarray(ref)), 0, w, h, None) == 0
    pr = np.array(prm, np.int32)
    out = np.zeros(8 * 64, np.float32)
    met = np.full(24, -1.0, np.float32)
    n = dev.sdk_engine_run(e, P(np.ascontiguousarray(dfm)), 0, P(pr), P(out), out.size, P(met), 24,
                           ctypes.c_float(1e-3))
    dev.sdk_engine_destroy(e)
    return int(n), met


def test_path_c_seeds_a_one_row_grid():
    """MetricsIdentityHoldsOnPathCRoute: no model, no lattice on a one-row
    grid, so Path C finds the seed and Path B solves the row from it; the
    metrics identity 3 + 4 + 19 = 1 + 20 + 23 holds."""
    ref, dfm = stretch_scene()
    n, m = dev_run(ref, dfm, (172, 128, 40, 5, 5, 21, 11, 0))
    print(f"path C route: rc {n} quality {m[16]:.0f} lock {m[21]:.0f} 3:{m[3]:.0f} 4:{m[4]:.0f} 19:{m[19]:.0f} "
          f"1:{m[1]:.0f} 20:{m[20]:.0f} 23:{m[23]:.0f}", flush=True)
    assert n >= 0 and m[21] == 0.0 and m[16] == 0.0
    assert m[19] == 0.0 and m[4] > 0.0
    assert m[3] + m[4] + m[19] == m[1] + m[20] + m[23]


def test_path_c_tries_the_next_node():
    """PathCTriesNextNodeWhenCentreFails: the deformed frame blank where
    Path C's first candidate (192, 128) lands; the next nearest seeds."""
    ref, dfm = stretch_scene()
    dfm = dfm.copy()
    dfm[118:139, 182:203] = 128
    n, m = dev_run(ref, dfm, (132, 128, 120, 15, 15, 21, 31, 0))
    print(f"path C next node: rc {n} quality {m[16]:.0f} lock {m[21]:.0f} path B {m[4]:.0f}", flush=True)
    assert m[21] == 0.0
    assert n != sa.ERR_SEED and n >= 0
    assert m[16] == 0.0 and m[4] > 0.0
