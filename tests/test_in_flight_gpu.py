"""Frames in flight (no reference counterpart): several engines, each with
its own stream and cached reference, driven from several host threads at
once, as a caller overlapping one frame's host phases with another's
kernels would. Each engine's output must be the single engine's, bit for
bit, on the noisy 1024^2 frame (whose flood has real work) and on the
learned-seeding frame."""
import ctypes
import importlib.util
import os
import threading

import numpy as np
import pytest

from helpers import PKG, P

pytestmark = pytest.mark.gpu


def _load(name, path):
    spec = importlib.util.spec_from_file_location(name, path)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


@pytest.fixture(scope="module")
def env():
    sa = _load("semper_amd", os.path.join(PKG, "semper_amd.py"))
    dev = sa.load_dev()
    if dev.sdk_device_count() < 1:
        pytest.skip("no GPU")
    return sa, dev, _load("workloads", os.path.join(PKG, "workloads.py"))


@pytest.mark.parametrize("name,learned", [("t_1024_noisy", False), ("c5_4096_bumps", True)])
def test_engines_in_threads_match_one_engine(env, name, learned):
    sa, dev, wl = env
    cfg = wl.config(name)
    ref, dfm = wl.pair(dev, name)
    prm = np.array(wl.solve_params(name), np.int32)
    cap = (cfg["w"] // cfg["step"] + 1) * (cfg["h"] // cfg["step"] + 1) * 8
    tol = ctypes.c_float(cfg["tol"])

    def engine():
        h = dev.sdk_engine_create()
        if learned:
            assert dev.sdk_engine_load_seeder(h, sa.SEEDER_WEIGHTS.encode()) == 0
            assert dev.sdk_engine_set_flag(h, 5, 1) == 0
        assert dev.sdk_engine_set_reference(h, P(ref), 0, cfg["w"], cfg["h"], None) == 0
        return h

    def run(h, out):
        met = np.zeros(24, np.float32)
        return dev.sdk_engine_run(h, P(dfm), 0, P(prm), P(out), cap, P(met), 24, tol)

    h0 = engine()
    out0 = np.zeros(cap, np.float32)
    n0 = run(h0, out0)
    dev.sdk_engine_destroy(h0)
    assert n0 > 0, n0
    k, runs = 3, 4
    hs = [engine() for _ in range(k)]
    outs = [np.zeros(cap, np.float32) for _ in range(k)]
    got = [[] for _ in range(k)]

    def work(i):
        for _ in range(runs):
            n = run(hs[i], outs[i])
            got[i].append((n, outs[i][: max(n, 0) * 8].copy()))

    ths = [threading.Thread(target=work, args=(i,)) for i in range(k)]
    for t in ths:
        t.start()
    for t in ths:
        t.join()
    for h in hs:
        dev.sdk_engine_destroy(h)
    print(f"{name}: {n0} points; threads {[[n for n, _ in g] for g in got]}", flush=True)
    ref_bits = out0[: n0 * 8].view(np.uint32)
    for g in got:
        assert len(g) == runs
        for n, pts in g:
            assert n == n0 and np.array_equal(pts.view(np.uint32), ref_bits)
