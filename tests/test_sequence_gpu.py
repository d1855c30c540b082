"""The batched / incremental sequence (no reference counterpart as a whole):
64 frames of 1024^2 speckle under a uniaxial stretch ramping linearly from 0
to 20% (Poisson 0.3), each frame with its own noise, through the incremental
session at update interval 1, with the subset sizes 15 / 21 / 31 / 41 given
by a seeded 4 x 4 region map of the frame, graph capture on. Checked against
the analytic field (displacement and Green-Lagrange strain of the frame-0
points), region by region against plain sessions at each region's size, and
the FP64 composition of the totals against the float one."""
import importlib.util
import os

import numpy as np
import pytest

from helpers import PKG, P

pytestmark = pytest.mark.gpu
NAME = "c3_1024_mixed"


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
    cfg = wl.config(NAME)
    frames = [wl.sequence_frame(dev, NAME, k) for k in range(cfg["frames"])]
    return sa, dev, wl, cfg, frames


def session(sa, dev, wl, cfg, frames, subset=None, compose64=False):
    s = sa.Incremental(dev)
    if subset is None:
        regions = wl.region_map(NAME)
        assert dev.sdk_incremental_set_region_map(s.h, 4, 4, P(regions)) == 0
    if compose64:
        assert dev.sdk_incremental_set_flag(s.h, 10, 1) == 0
    prm = sa.params((0, 0, cfg["w"], cfg["h"]), cfg["step"], subset or cfg["subset"], cfg["window"])
    assert s.init(frames[0], prm, 1) == 0
    outs = []
    for f in frames[1:]:
        n, pts, _, info = s.process(f)
        assert n > 0 and info.reference_updated == 1, n
        outs.append(pts)
    s.close()
    return outs


def truth(wl, cfg, k, pts):
    exx = wl.sequence_strain(NAME, k)
    eyy = -cfg["p"][1] * exx
    u = exx * (pts[:, 0] - cfg["w"] / 2.0)
    v = eyy * (pts[:, 1] - cfg["h"] / 2.0)
    return u, v, exx + 0.5 * exx * exx, eyy + 0.5 * eyy * eyy


def region_of(cfg, pts):
    return (pts[:, 1].astype(int) * 4 // cfg["h"]) * 4 + pts[:, 0].astype(int) * 4 // cfg["w"]


def test_sequence_mixed_sizes_against_truth_and_plain_sessions(env):
    sa, dev, wl, cfg, frames = env
    mixed = session(sa, dev, wl, cfg, frames)
    first = len(mixed[0])
    for k, pts in enumerate(mixed, start=1):
        assert len(pts) >= 0.75 * first, (k, len(pts), first)
    last = mixed[-1]
    K = len(frames) - 1
    u, v, Exx, Eyy = truth(wl, cfg, K, last)
    err = np.hypot(last[:, 2] - u, last[:, 3] - v)
    print(f"mixed, frame {K}: {len(last)} points (frame 1: {first}), rms {np.sqrt(np.mean(err ** 2)):.3f} px, "
          f"median {np.median(err):.3f}, mean |dExx| {np.mean(np.abs(last[:, 4] - Exx)):.4f}, "
          f"|dEyy| {np.mean(np.abs(last[:, 5] - Eyy)):.4f}", flush=True)
    # 63 updates of IC-GN's per-increment bias; the size-15 regions carry most
    # of it (0.52 px rms, median 0.28 when measured)
    assert np.sqrt(np.mean(err ** 2)) <= 0.6 and np.median(err) <= 0.35
    assert abs(np.median(last[:, 4]) - Exx) <= 0.01 and abs(np.median(last[:, 5]) - Eyy) <= 0.01
    reg = region_of(cfg, last)
    sizes = wl.region_map(NAME).ravel()
    for s in sorted(set(sizes.tolist())):
        plain = session(sa, dev, wl, cfg, frames, subset=s)[-1]
        if s == 21:
            # the reference's long-horizon bound for update interval 1 at
            # subset 21 (test_incremental.cpp:619, 0.45 px on rms u), here on
            # the rms of |(du, dv)|
            up, vp, _, _ = truth(wl, cfg, K, plain)
            ep = np.sqrt(np.mean((plain[:, 2] - up) ** 2 + (plain[:, 3] - vp) ** 2))
            print(f"plain subset 21, frame {K}: {len(plain)} points, rms {ep:.3f} px", flush=True)
            assert ep <= 0.45, ep
        pr = region_of(cfg, plain)
        for r in np.nonzero(sizes == s)[0]:
            a, b = last[reg == r], plain[pr == r]
            ua, va, _, _ = truth(wl, cfg, K, a)
            ub, vb, _, _ = truth(wl, cfg, K, b)
            ea = np.sqrt(np.mean((a[:, 2] - ua) ** 2 + (a[:, 3] - va) ** 2))
            eb = np.sqrt(np.mean((b[:, 2] - ub) ** 2 + (b[:, 3] - vb) ** 2))
            print(f"  region {r} (size {s}): mixed {len(a)} points rms {ea:.3f}, plain {len(b)} points rms {eb:.3f}",
                  flush=True)
            assert len(a) >= 0.8 * len(b)
            assert ea <= 1.5 * eb + 0.1


def test_fp64_composition_against_float(env):
    sa, dev, wl, cfg, frames = env
    f32 = session(sa, dev, wl, cfg, frames[:33])
    f64 = session(sa, dev, wl, cfg, frames[:33], compose64=True)
    # nothing is composed on the first frame
    assert np.array_equal(f32[0], f64[0])
    for k, (a, b) in enumerate(zip(f32, f64), start=1):
        da = {(x, y): r for x, y, *r in a.tolist()}
        db = {(x, y): r for x, y, *r in b.tolist()}
        common = sorted(set(da) & set(db))
        assert len(common) >= min(len(da), len(db)) - 3, k
        d = np.array([np.abs(np.array(da[c][:2]) - np.array(db[c][:2])) for c in common])
        assert np.mean(d.max(1) <= 1e-3) >= 0.999 and d.max() <= 2e-2, (k, d.max())
    print(f"frame {len(f32)}: float and FP64 composition, max |duv| {d.max():.2e}", flush=True)
