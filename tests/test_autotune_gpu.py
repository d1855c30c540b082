"""The Path A launch-shape tuner (no reference counterpart). A shape -- 1, 2 or
4 subsets per fast-path workgroup, row-major or 8 x 8-tile node order -- only
changes which wave solves which subset and when, never a subset's arithmetic,
so every shape must give the bit-identical field, point count and metrics of
the untuned launch, on a single-size grid and on a mixed-size grid (where the
tile orders are not tried). The tuner times every candidate twice on real runs,
then settles on the one of least mean time, and the field never changes."""
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
    wl = _load("workloads", os.path.join(PKG, "workloads.py"))
    return dev, wl


def run(dev, cfg, ref, dfm, prm, flags, dims=None, runs=1):
    h = dev.sdk_engine_create()
    for f, v in flags:
        assert dev.sdk_engine_set_flag(h, f, v) == 0
    if dims is not None:
        gw, gh = cfg["w"] // cfg["step"], cfg["h"] // cfg["step"]
        assert dev.sdk_engine_set_subset_map(h, gw, gh, P(dims)) == 0
    assert dev.sdk_engine_set_reference(h, P(ref), 0, cfg["w"], cfg["h"], None) == 0
    cap = (cfg["w"] // cfg["step"] + 1) * (cfg["h"] // cfg["step"] + 1) * 8
    outs, tun = [], []
    for _ in range(runs):
        out = np.zeros(cap, np.float32)
        met = np.zeros(24, np.float32)
        n = dev.sdk_engine_run(h, P(dfm), 0, P(prm), P(out), cap, P(met), 24, ctypes.c_float(cfg["tol"]))
        assert n > 0, n
        chosen = np.zeros(2, np.int32)
        ms = np.zeros(6, np.float64)
        rc = dev.sdk_engine_tuning(h, P(chosen), P(ms))
        outs.append((out[: n * 8].copy(), met))
        tun.append((rc, chosen, ms))
    dev.sdk_engine_destroy(h)
    return outs, tun


CANDIDATES = [(1, 0), (2, 0), (4, 0), (1, 8), (2, 8), (4, 8)]
COUNTS = [0, 1, 2, 3, 4, 5, 6, 7, 8, 15, 16, 19, 20, 21, 22, 23]   # the metric slots that are not timings


@pytest.mark.parametrize("name", ["t_1024_noisy", "c3_1024_mixed"])
def test_every_shape_gives_the_same_field(env, name):
    dev, wl = env
    cfg = wl.config(name)
    ref, dfm = wl.pair(dev, name)
    mixed = name.endswith("mixed")
    dims = None
    if mixed:
        dims = wl.subset_map(name, cfg["w"] // cfg["step"], cfg["h"] // cfg["step"])
        prm = np.array((0, 0, cfg["w"], cfg["h"], cfg["step"], 21, cfg["window"], 0), np.int32)
    else:
        prm = np.array(wl.solve_params(name), np.int32)
    [(base, met0)], [tun0] = run(dev, cfg, ref, dfm, prm, [(6, 0)], dims)
    assert tun0[0] == -3   # nothing timed with the tuner off
    for i, (wpb, tile) in enumerate(CANDIDATES):
        [(out, met)], _ = run(dev, cfg, ref, dfm, prm, [(7, i + 1)], dims)
        print(f"{name}: shape {wpb} x {tile}: {len(out) // 8} points", flush=True)
        assert np.array_equal(out, base), (wpb, tile)
        assert np.array_equal(met[COUNTS], met0[COUNTS]), (wpb, tile)
    nc = 3 if mixed else 6
    outs, tun = run(dev, cfg, ref, dfm, prm, [], dims, runs=2 * nc + 2)
    for out, met in outs:
        assert np.array_equal(out, base) and np.array_equal(met[COUNTS], met0[COUNTS])
    rcs = [t[0] for t in tun]
    assert rcs == [1] * (2 * nc - 1) + [0] * 3, rcs
    rc, chosen, ms = tun[-1]
    print(f"{name}: tuner chose {tuple(chosen)}, candidate ms {np.round(ms, 4).tolist()}", flush=True)
    tried = [i for i in range(6) if ms[i] >= 0]
    assert tried == list(range(nc))
    k = CANDIDATES.index(tuple(int(c) for c in chosen))
    assert ms[k] == min(ms[i] for i in tried)


@pytest.mark.parametrize("use6", [0, 1], ids=["keys4", "keys6"])
def test_fp32_deformed_plane_gives_the_same_field(env, use6):
    """The solver sampling an FP32 copy of the deformed frame (flag 13, the
    default) against sampling the u8 frame, both by tap weights (flag 17 off:
    the coefficient plane is made from the FP32 run only): the copy holds the
    same values, and both kernels gather and sum in the same order, so the
    field, the point count and the metrics are the same bit for bit. Without
    the FP32 copy the coefficient plane is not used, whatever flag 17 says."""
    dev, wl = env
    name = "t_1024_noisy"
    cfg = wl.config(name)
    ref, dfm = wl.pair(dev, name)
    prm = np.array(wl.solve_params(name), np.int32)
    prm[7] = use6
    [(a, ma)], _ = run(dev, cfg, ref, dfm, prm, [(13, 0), (17, 0)])
    [(b, mb)], _ = run(dev, cfg, ref, dfm, prm, [(13, 1), (17, 0)])
    [(c, mc)], _ = run(dev, cfg, ref, dfm, prm, [(13, 0), (17, 1)])
    print(f"use6 {use6}: {len(a) // 8} points", flush=True)
    assert np.array_equal(a, b) and np.array_equal(ma[COUNTS], mb[COUNTS])
    assert np.array_equal(a, c) and np.array_equal(ma[COUNTS], mc[COUNTS])
