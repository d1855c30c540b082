"""Layered Path B (no reference counterpart; the default under learned
seeding, reach 8). Each layer claims every open node within `reach` grid
steps of an accepted one at once, seeded from the accepted node a
breadth-first search in the reference flood's pop order reaches it from
first, one launch per layer, instead of the reference's rounds of 32 pops;
with reach 1 every claim is one the reference's flood could make. A node's
converged warp does not depend on which accepted node seeded it, up to the
stop rule, so
against the reference flood on the same frames it must report the same point
set (within a few nodes whose acceptance is marginal), the same field within
the reference's cross-toolchain bound, no less accuracy against the FP64 IC-GN
optimum."""
import ctypes
import importlib.util
import os

import numpy as np
import pytest

from helpers import PKG, P, fp64_polish

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
    return dev, _load("workloads", os.path.join(PKG, "workloads.py"))


def run(dev, cfg, ref, dfm, prm, reach):
    h = dev.sdk_engine_create()
    assert dev.sdk_engine_set_flag(h, 8, reach) == 0
    assert dev.sdk_engine_set_reference(h, P(ref), 0, cfg["w"], cfg["h"], None) == 0
    cap = (cfg["w"] // cfg["step"] + 1) * (cfg["h"] // cfg["step"] + 1) * 8
    out = np.zeros(cap, np.float32)
    met = np.zeros(24, np.float32)
    tm = np.zeros(14, np.float64)
    n = dev.sdk_engine_run(h, P(dfm), 0, P(prm), P(out), cap, P(met), 24, ctypes.c_float(cfg["tol"]))
    dev.sdk_engine_timings(h, P(tm))
    dev.sdk_engine_destroy(h)
    assert n > 0, n
    return out[: n * 8].reshape(-1, 8).copy(), met, tm


def fp64_error(ref, dfm, pts, subset, k=300):
    rng = np.random.default_rng(3)
    sel = pts[rng.choice(len(pts), size=min(k, len(pts)), replace=False)]
    p0 = np.zeros((len(sel), 6))
    p0[:, 0], p0[:, 1] = sel[:, 2], sel[:, 3]
    si, po, _ = fp64_polish(ref.astype(np.float32), dfm.astype(np.float32), subset,
                            [(int(x), int(y)) for x, y in sel[:, :2]], p0)
    ok = si[:, 0] == 0
    return np.maximum(np.abs(po[ok, 0] - sel[ok, 2]), np.abs(po[ok, 1] - sel[ok, 3]))


@pytest.mark.parametrize("reach", [1, 4])
def test_layered_flood_against_reference_flood(env, reach):
    dev, wl = env
    name = "t_1024_noisy"
    cfg = wl.config(name)
    ref, dfm = wl.pair(dev, name)
    prm = np.array(wl.solve_params(name), np.int32)
    p0, m0, t0 = run(dev, cfg, ref, dfm, prm, 0)
    p1, m1, t1 = run(dev, cfg, ref, dfm, prm, reach)
    print(f"reference flood: {len(p0)} points, Path B {int(m0[4])} solved in {int(t0[12])} further launches; "
          f"layered, reach {reach}: {len(p1)} points, Path B {int(m1[4])} solved in {int(t1[12])} launches",
          flush=True)
    assert m0[4] > 50   # the frame gives the flood real work
    assert abs(len(p1) - len(p0)) <= max(3, len(p0) // 100)
    assert abs(m1[4] - m0[4]) <= max(3, m0[4] // 20)
    assert m1[3] == m0[3]   # Path A is untouched
    a = {(x, y): r for x, y, *r in p0.tolist()}
    b = {(x, y): r for x, y, *r in p1.tolist()}
    common = sorted(set(a) & set(b))
    assert len(common) >= len(p0) - max(3, len(p0) // 100)
    d = np.array([np.abs(np.array(a[k][:2]) - np.array(b[k][:2])) for k in common])
    print(f"  max |duv| {d.max():.2e}, within 1e-3 {np.mean(d.max(1) <= 1e-3):.4f}", flush=True)
    assert d.max() <= 5e-3 and np.mean(d.max(1) <= 1e-3) >= 0.99
    e0, e1 = fp64_error(ref, dfm, p0, int(prm[5])), fp64_error(ref, dfm, p1, int(prm[5]))
    print(f"  vs FP64 within 1e-3 px: layered {np.mean(e1 <= 1e-3):.4f}, reference flood {np.mean(e0 <= 1e-3):.4f}",
          flush=True)
    assert np.mean(e1 <= 1e-3) >= np.mean(e0 <= 1e-3) - 2.0 / len(e0)


def test_node_report_agrees_with_the_metrics(env):
    """The per-node provenance report (flag 12) against the run's own metrics
    slots: anchors published (slot 19), Path A points (slot 3) and the later
    paths' points (slot 4) are the nodes the report gives to each path."""
    dev, wl = env
    name = "t_1024_noisy"
    cfg = wl.config(name)
    ref, dfm = wl.pair(dev, name)
    prm = np.array(wl.solve_params(name), np.int32)
    gw, gh = cfg["w"] // cfg["step"], cfg["h"] // cfg["step"]
    h = dev.sdk_engine_create()
    assert dev.sdk_engine_set_flag(h, 12, 1) == 0
    assert dev.sdk_engine_set_reference(h, P(ref), 0, cfg["w"], cfg["h"], None) == 0
    cap = (gw + 1) * (gh + 1) * 8
    out = np.zeros(cap, np.float32)
    met = np.zeros(24, np.float32)
    assert dev.sdk_engine_run(h, P(dfm), 0, P(prm), P(out), cap, P(met), 24, ctypes.c_float(cfg["tol"])) > 0
    tag = np.zeros(gw * gh, np.uint8)
    path = np.zeros(gw * gh, np.uint8)
    assert dev.sdk_engine_node_report(h, P(tag), P(path), gw * gh - 1) == -3
    assert dev.sdk_engine_node_report(h, P(tag), P(path), gw * gh) == gw * gh
    dev.sdk_engine_destroy(h)
    counts = np.bincount(path, minlength=5)
    print(f"report: {counts.tolist()}, metrics: anchors {met[19]:.0f}, A {met[3]:.0f}, B {met[4]:.0f}, "
          f"attempted {met[0]:.0f}", flush=True)
    assert counts[2] == met[19] and counts[3] == met[3] and counts[4] == met[4]
    assert counts[1:].sum() == met[0]
    assert np.all(tag[path == 2] == 3)


def test_learned_frame_repeats_bit_for_bit(env):
    """The learned-seeding workload at full size (4096^2, eight Gaussian bumps
    of up to +-30 px and a rotation of up to 5 degrees, step 8, subset 31),
    run six times on one engine: every run reports the same points bit for
    bit, and every node the report calls skipped is one the scan skipped
    (the attempted count of slot 0). Under the layered flood a node can be
    claimed twice, a failed far claim and then an adjacent one; both results
    reaching the apply kernel raced, and a node accepted by the second claim
    was now and then left with the first claim's invalid corr."""
    dev, wl = env
    sa = _load("semper_amd", os.path.join(PKG, "semper_amd.py"))
    name = "c5_4096_bumps"
    cfg = wl.config(name)
    ref, dfm = wl.pair(dev, name)
    prm = np.array(wl.solve_params(name), np.int32)
    gw, gh = cfg["w"] // cfg["step"], cfg["h"] // cfg["step"]
    cap = (gw + 1) * (gh + 1) * 8
    h = dev.sdk_engine_create()
    assert dev.sdk_engine_load_seeder(h, sa.SEEDER_WEIGHTS.encode()) == 0
    assert dev.sdk_engine_set_flag(h, 5, 1) == 0
    assert dev.sdk_engine_set_flag(h, 12, 1) == 0
    assert dev.sdk_engine_set_reference(h, P(ref), 0, cfg["w"], cfg["h"], None) == 0
    first = None
    tag = np.zeros(gw * gh, np.uint8)
    path = np.zeros(gw * gh, np.uint8)
    log = np.zeros(2 * 64, np.int32)
    for k in range(6):
        out = np.zeros(cap, np.float32)
        met = np.zeros(24, np.float32)
        n = dev.sdk_engine_run(h, P(dfm), 0, P(prm), P(out), cap, P(met), 24, ctypes.c_float(cfg["tol"]))
        assert n > 0, n
        assert dev.sdk_engine_node_report(h, P(tag), P(path), gw * gh) == gw * gh
        counts = np.bincount(path, minlength=5)
        nl = dev.sdk_engine_launch_log(h, P(log), 64)
        print(f"run {k}: {n} points, report {counts.tolist()}, attempted {met[0]:.0f}, {nl} solver launches",
              flush=True)
        assert counts[1:].sum() == met[0]
        assert nl >= 3 and log[0] > 0 and np.all(log[1:2 * min(nl, 64):2] <= log[0:2 * min(nl, 64):2])
        pts = out[: n * 8].copy()
        if first is None:
            first = pts
        else:
            assert pts.shape == first.shape and np.array_equal(pts.view(np.uint32), first.view(np.uint32))
    dev.sdk_engine_destroy(h)


def run_flags(dev, cfg, ref, dfm, prm, flags, seeder=None):
    h = dev.sdk_engine_create()
    if seeder is not None:
        assert dev.sdk_engine_load_seeder(h, seeder) == 0
    for f, v in flags:
        assert dev.sdk_engine_set_flag(h, f, v) == 0
    assert dev.sdk_engine_set_reference(h, P(ref), 0, cfg["w"], cfg["h"], None) == 0
    cap = (cfg["w"] // cfg["step"] + 1) * (cfg["h"] // cfg["step"] + 1) * 8
    out = np.zeros(cap, np.float32)
    met = np.zeros(24, np.float32)
    tm = np.zeros(14, np.float64)
    n = dev.sdk_engine_run(h, P(dfm), 0, P(prm), P(out), cap, P(met), 24, ctypes.c_float(cfg["tol"]))
    dev.sdk_engine_timings(h, P(tm))
    log = np.zeros(2 * 256, np.int32)
    nl = dev.sdk_engine_launch_log(h, P(log), 256)
    dev.sdk_engine_destroy(h)
    assert n > 0, n
    return out[: n * 8].copy(), met, tm, log[: 2 * min(nl, 256)].copy()


@pytest.mark.parametrize("reach", [1, 4, 16])
def test_device_frontier_is_the_host_frontier(env, reach):
    """The layered flood's frontier on the device (flag 18, the default: the
    sources sorted in the flood's pop order, the breadth-first search's seeds
    by the FIFO tie-break, the claims and the updates) against the same layers
    run on the host: the same field bit for bit, the same metrics, the same
    solver launches with the same work items and hand-offs, on the noisy 1024^2
    frame where the flood has real work."""
    dev, wl = env
    name = "t_1024_noisy"
    cfg = wl.config(name)
    ref, dfm = wl.pair(dev, name)
    prm = np.array(wl.solve_params(name), np.int32)
    a, ma, ta, la = run_flags(dev, cfg, ref, dfm, prm, [(8, reach), (18, 1)])
    b, mb, tb, lb = run_flags(dev, cfg, ref, dfm, prm, [(8, reach), (18, 0)])
    print(f"reach {reach}: {len(a) // 8} points, Path B {int(ma[4])} solved in {int(ta[12])} launches "
          f"({int(ta[13])} claims)", flush=True)
    assert ma[4] > 50
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    assert np.array_equal(ma[[0, 1, 2, 3, 4, 5, 6, 7, 8, 16, 19, 20, 21, 22, 23]],
                          mb[[0, 1, 2, 3, 4, 5, 6, 7, 8, 16, 19, 20, 21, 22, 23]])
    assert ta[12] == tb[12] and ta[13] == tb[13]
    assert np.array_equal(la, lb)


def test_device_frontier_on_the_learned_frame(env):
    """The learned-seeding workload at full size (4096^2 bumps, reach 16 by
    default under learned seeding): the device frontier's field equals the
    host frontier's bit for bit, with the same metrics and launches."""
    dev, wl = env
    sa = _load("semper_amd", os.path.join(PKG, "semper_amd.py"))
    name = "c5_4096_bumps"
    cfg = wl.config(name)
    ref, dfm = wl.pair(dev, name)
    prm = np.array(wl.solve_params(name), np.int32)
    w = sa.SEEDER_WEIGHTS.encode()
    a, ma, ta, la = run_flags(dev, cfg, ref, dfm, prm, [(5, 1), (18, 1)], seeder=w)
    b, mb, tb, lb = run_flags(dev, cfg, ref, dfm, prm, [(5, 1), (18, 0)], seeder=w)
    print(f"learned frame: {len(a) // 8} points, Path B {int(ma[4])} solved in {int(ta[12])} launches, "
          f"Path B ms {ta[6]:.2f} (device) {tb[6]:.2f} (host)", flush=True)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    assert np.array_equal(ma[[0, 1, 2, 3, 4, 5, 6, 7, 8]], mb[[0, 1, 2, 3, 4, 5, 6, 7, 8]])
    assert ta[12] == tb[12] and ta[13] == tb[13]
    assert np.array_equal(la, lb)
