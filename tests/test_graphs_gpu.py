"""Graph capture of the fixed per-frame kernel chains (no reference
counterpart): the end of a run -- valid flags, five median passes, VSG, the
row-major pack, the counts -- and the incremental session's strain of the
total field. A chain is captured the second time it is met with the same
grid and buffers and replayed after that, so each replay runs between the
engine's own eager launches and synchronisations. Every replay must give
exactly what the eager chain gives, frame by frame."""
import ctypes
import importlib.util
import os

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


def engine_runs(dev, cfg, ref, frames, prm, graphs):
    h = dev.sdk_engine_create()
    assert dev.sdk_engine_set_flag(h, 9, 1 if graphs else 0) == 0
    assert dev.sdk_engine_set_reference(h, P(ref), 0, cfg["w"], cfg["h"], None) == 0
    cap = (cfg["w"] // cfg["step"] + 1) * (cfg["h"] // cfg["step"] + 1) * 8
    res = []
    for f in frames:
        out = np.zeros(cap, np.float32)
        met = np.zeros(24, np.float32)
        n = dev.sdk_engine_run(h, P(f), 0, P(prm), P(out), cap, P(met), 24, ctypes.c_float(cfg["tol"]))
        assert n > 0, n
        res.append((out[: n * 8].copy(), met[[0, 1, 2, 3, 4, 19, 20, 23]].copy()))
    stats = np.zeros(2, np.int32)
    assert dev.sdk_engine_graph_stats(h, P(stats)) == 0
    dev.sdk_engine_destroy(h)
    return res, stats


def test_run_chain_replays_match_eager(env):
    sa, dev, wl = env
    name = "c3_1024_mixed"
    cfg = wl.config(name)
    frames = [wl.sequence_frame(dev, name, k) for k in (0, 9, 21, 33, 45, 63)]
    prm = np.array(wl.solve_params(name), np.int32)
    # the same frame twice in a row too: a replay right after the capture
    seq = [frames[1], frames[2], frames[2], frames[3], frames[4], frames[5], frames[1]]
    eager, s0 = engine_runs(dev, cfg, frames[0], seq, prm, False)
    graph, s1 = engine_runs(dev, cfg, frames[0], seq, prm, True)
    print(f"graph stats: eager {s0.tolist()}, graphs {s1.tolist()}", flush=True)
    assert s0.tolist() == [0, 0] and s1[0] == 1 and s1[1] == len(seq) - 1
    for k, ((o0, m0), (o1, m1)) in enumerate(zip(eager, graph)):
        assert np.array_equal(o0, o1), k
        assert np.array_equal(m0, m1), k


def session_run(sa, dev, wl, name, frames, graphs):
    cfg = wl.config(name)
    s = sa.Incremental(dev)
    assert dev.sdk_incremental_set_flag(s.h, 9, 1 if graphs else 0) == 0
    regions = wl.region_map(name)
    assert dev.sdk_incremental_set_region_map(s.h, 4, 4, P(regions)) == 0
    assert s.init(frames[0], sa.params((0, 0, cfg["w"], cfg["h"]), cfg["step"], cfg["subset"], cfg["window"]), 1) == 0
    res = []
    for f in frames[1:]:
        n, pts, met, info = s.process(f)
        assert n > 0, n
        res.append((pts, met[[0, 1, 2, 3, 4, 19, 20, 23]].copy(),
                    [info.reference_frame, info.reference_updated, info.tracked_points, info.frame0_points,
                     info.taylor_fallbacks, info.lost_this_frame]))
    stats = np.zeros(2, np.int32)
    assert dev.sdk_incremental_graph_stats(s.h, P(stats)) == 0
    s.close()
    return res, stats


def test_session_strain_chain_replays_match_eager(env):
    sa, dev, wl = env
    name = "c3_1024_mixed"
    frames = [wl.sequence_frame(dev, name, k) for k in range(12)]
    eager, s0 = session_run(sa, dev, wl, name, frames, False)
    graph, s1 = session_run(sa, dev, wl, name, frames, True)
    print(f"session graph stats: eager {s0.tolist()}, graphs {s1.tolist()}", flush=True)
    assert s0.tolist() == [0, 0] and s1[1] >= len(frames) - 4
    for k, (a, b) in enumerate(zip(eager, graph)):
        assert np.array_equal(a[0], b[0]), k
        assert np.array_equal(a[1], b[1]) and a[2] == b[2], k
