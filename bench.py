"""Benchmark: converged subsets per second of the full-field DIC solve
(semper_run's pipeline) on one MI355X.

Workload (default): 4096 x 4096 synthetic speckle with the heterogeneous field
u = 3 sin(2 pi x / 1024) cos(2 pi y / 1536), v = 2 exp(-r^2 / (2 600^2)),
step 8, subset 31, strain window 31, Keys 4x4, FP32 IC-GN (253k attempted
subsets). One step = one full solve of the deformed frame against the cached
reference (seeding, Path A/B, strain, packing), with both frames already in
HBM. `value` = solved subsets (status 0, ZNSSD <= 0.15) per second, scaled by
the fraction of them within 1e-3 px of the FP64 oracle on a sampled subset of
nodes.

Also reported: the roofline of the dominant kernel (the Path A subset solver
launch, FP32 VALU bound: 109 flop per pixel-iteration + 8 per subset pixel for
the SD build), timed with HIP events on the engine's stream, and the CPU
baseline: the oracle (a C++/OpenMP restatement of the reference's
run_full_field) on a bounded sample of the same workload on the host cores.
"""
import argparse
import ctypes
import importlib.util
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "semper-dic-engine_amd")


def _load(name, path):
    spec = importlib.util.spec_from_file_location(name, path)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


_PINNED = []


def pinned_f32(dev, n):
    """A float32 array of n in page-locked host memory (the output buffer of the
    timed runs: the packed points reach it by a direct copy from the device);
    held until the process ends."""
    p = dev.sdk_host_alloc(4 * n)
    if not p:
        raise MemoryError("sdk_host_alloc failed")
    a = np.ctypeslib.as_array(ctypes.cast(p, ctypes.POINTER(ctypes.c_float)), shape=(n,))
    a[:] = 0.0
    _PINNED.append(p)
    return a


FLOP_PER_PIXEL_ITER = {False: 109.0, True: 227.0}
SD_FLOP_PER_PIXEL = 8.0
PEAK_FP32_TFLOPS = 157.3
# HBM bytes per Path A launch from two rocprofv3 --pmc runs (FETCH_SIZE x 2.0,
# WRITE_SIZE x 1.0, the factors a 256 MiB copy kernel reads under the same
# counters); measured per workload with the default engine switches, recorded
# with the kernel it was measured on (null for a workload or switch set not
# measured)
TRAFFIC_BYTES = {"c4_4096_heterogeneous": 214073.7 * 1024 * 2.0 + 24362.3 * 1024 * 1.0}


def fp64_accuracy(ref, dfm, subset, pts, use6, k=2000, seed=0):
    """Fraction of a node sample whose (u, v) lie within 1e-3 px of the FP64
    IC-GN optimum of the same subset (FP64 solve started from the point's
    displacement, run to ||dp|| < 1e-10)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from helpers import fp64_polish

    rng = np.random.default_rng(seed)
    k = min(k, len(pts))
    sel = pts[rng.choice(len(pts), size=k, replace=False)]
    nodes = [(int(x), int(y)) for x, y in sel[:, :2]]
    p0 = np.zeros((k, 6), np.float64)
    p0[:, 0], p0[:, 1] = sel[:, 2], sel[:, 3]
    si, po, _ = fp64_polish(ref.astype(np.float32), dfm.astype(np.float32), subset, nodes, p0, use6=use6)
    ok = si[:, 0] == 0
    du = np.abs(po[ok, 0] - sel[ok, 2])
    dv = np.abs(po[ok, 1] - sel[ok, 3])
    frac = float(np.mean((du <= 1e-3) & (dv <= 1e-3))) if ok.any() else 0.0
    return {"sampled": int(k), "fp64_converged": int(ok.sum()), "within_1e-3px": frac,
            "max_du": float(du.max()) if ok.any() else None, "max_dv": float(dv.max()) if ok.any() else None,
            "rms_du": float(np.sqrt((du ** 2).mean())) if ok.any() else None,
            "rms_dv": float(np.sqrt((dv ** 2).mean())) if ok.any() else None}


# (name, ((flag, value), ...)): flag 2 FP64 Cholesky step, flag 3 gradient
# planes in fp16 (1) or bf16 (2), flag 4 MFMA J^T J pool, flag 17 Keys 4x4
# from the fp16 row-cubic coefficient plane
PRECISION_VARIANTS = [("fp64_cholesky_step", ((2, 1),)), ("fp16_gradients", ((3, 1),)),
                      ("bf16_gradients", ((3, 2),)), ("mfma_jtj_pool", ((4, 1),)), ("fp16_coefficients", ((17, 1),)),
                      ("all_fp16", ((2, 1), (3, 1), (4, 1), (17, 1))), ("all_bf16", ((2, 1), (3, 2), (4, 1), (17, 1)))]


def sequence_study(dev, sa, wl, frames_limit=None):
    """The batched / incremental sequence: 64 frames of 1024^2 speckle under a
    uniaxial stretch ramping 0 -> 20% (Poisson 0.3), each frame with its own
    noise, through the incremental session at update interval 1. The mixed
    15 / 21 / 31 / 41 region map with graph capture on and off, and plain
    subset 21 beside them: per-frame time (all 63 frames after the first, the
    session's own captures included), output points and the last frame's
    error against the analytic field."""
    name = "c3_1024_mixed"
    cfg = wl.config(name)
    nframes = cfg["frames"] if frames_limit is None else min(cfg["frames"], frames_limit)
    frames = [wl.sequence_frame(dev, name, k) for k in range(nframes)]
    P = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    res = {"workload": f"{name}: {cfg['frames']} frames {cfg['w']}x{cfg['h']}, exx 0 -> {cfg['ramp']}, Poisson "
                       f"{cfg['p'][1]}, step {cfg['step']}, strain window {cfg['window']}, update interval 1",
           "region_map": wl.region_map(name).tolist()}
    for label, mixed, graphs in (("mixed_graphs", True, True), ("mixed_eager", True, False),
                                 ("plain_21_graphs", False, True)):
        s = sa.Incremental(dev)
        dev.sdk_incremental_set_flag(s.h, 9, 1 if graphs else 0)
        if mixed:
            regions = wl.region_map(name)
            dev.sdk_incremental_set_region_map(s.h, 4, 4, P(regions))
        assert s.init(frames[0], sa.params((0, 0, cfg["w"], cfg["h"]), cfg["step"], cfg["subset"], cfg["window"]),
                      1) == 0
        ms, npts = [], []
        pts = None
        t8s, t14s = [], []
        for f in frames[1:]:
            t0 = time.perf_counter()
            n, pts, _, _ = s.process(f)
            ms.append((time.perf_counter() - t0) * 1e3)
            npts.append(int(n))
            t8, t14 = np.zeros(8), np.zeros(14)
            dev.sdk_incremental_timings(s.h, P(t8))
            dev.sdk_incremental_engine_timings(s.h, P(t14))
            t8s.append(t8)
            t14s.append(t14)
        stats = np.zeros(2, np.int32)
        dev.sdk_incremental_graph_stats(s.h, P(stats))
        s.close()
        k = nframes - 1
        exx = wl.sequence_strain(name, k)
        t8m, t14m = np.mean(t8s, 0), np.mean(t14s, 0)
        u = exx * (pts[:, 0] - cfg["w"] / 2.0)
        v = -cfg["p"][1] * exx * (pts[:, 1] - cfg["h"] / 2.0)
        err = np.hypot(pts[:, 2] - u, pts[:, 3] - v)
        res[label] = {"ms_per_frame_mean": float(np.mean(ms)), "ms_per_frame_median": float(np.median(ms)),
                      "frames_per_s": float(1000.0 / np.mean(ms)), "points_per_s": float(np.sum(npts) / np.sum(ms) * 1e3),
                      "points_first": npts[0], "points_last": npts[-1],
                      "last_frame_rms_px": float(np.sqrt(np.mean(err ** 2))),
                      "last_frame_median_px": float(np.median(err)),
                      "graph_captures": int(stats[0]), "graph_replays": int(stats[1]),
                      "session_phases_ms": {k_: round(float(v), 4) for k_, v in zip(
                          ["total", "engine_run", "composition", "reference_update", "field_fill", "mask_warp",
                           "mask_bounds", "adopt"], t8m)},
                      "engine_phases_ms": {k_: round(float(v), 4) for k_, v in zip(
                          ["total", "prep", "seeding", "prepass", "mesh", "path_a", "path_b", "strain",
                           "path_a_kernel"], t14m[:9])}}
    return res


def settle_tuner(dev, eng, run_once, limit=16):
    """Untimed runs until the engine's launch-shape tuner has settled (it
    times each shape on the first runs of a configuration)."""
    shape = np.zeros(2, np.int32)
    shape_ms = np.zeros(6, np.float64)
    for _ in range(limit):
        if dev.sdk_engine_tuning(eng, shape.ctypes.data_as(ctypes.c_void_p),
                                 shape_ms.ctypes.data_as(ctypes.c_void_p)) in (0, -3):
            return
        run_once()


def precision_study(dev, dref, ddef, prm, cap, tol, ref, dfm, subset, use6, steps):
    """The precision variants (FP64 Cholesky Gauss-Newton step, fp16 or bf16
    gradient storage, J^T J on MFMA, the fp16 coefficient plane): per variant the first run's Hessian pre-pass (the
    pool build), the step time, the converged throughput and the error of
    every output point's (u, v) against the FP64 IC-GN optimum on the same
    2000-node sample as the plain method."""
    P = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    w = ref.shape[1]
    res = {}
    for name, flags in PRECISION_VARIANTS:
        eng = dev.sdk_engine_create()
        for f, v in flags:
            assert dev.sdk_engine_set_flag(eng, f, v) == 0
        assert dev.sdk_engine_set_reference(eng, dref, 1, w, ref.shape[0], None) == 0
        out = pinned_f32(dev, cap)
        met = np.zeros(24, np.float32)
        tm = np.zeros(14, np.float64)
        n = dev.sdk_engine_run(eng, ddef, 1, P(prm), P(out), cap, P(met), 24, tol)
        dev.sdk_engine_timings(eng, P(tm))
        prepass = float(tm[3])
        settle_tuner(dev, eng, lambda: dev.sdk_engine_run(eng, ddef, 1, P(prm), P(out), cap, P(met), 24, tol))
        t0 = time.perf_counter()
        for _ in range(steps):
            n = dev.sdk_engine_run(eng, ddef, 1, P(prm), P(out), cap, P(met), 24, tol)
        ms = (time.perf_counter() - t0) * 1000.0 / steps
        dev.sdk_engine_destroy(eng)
        pts = out[: int(n) * 8].reshape(-1, 8).copy()
        acc = fp64_accuracy(ref, dfm, subset, pts, use6)
        solved = float(met[3] + met[4] + met[19])
        res[name] = {"ms_per_step": ms, "prepass_ms": prepass, "output_points": int(n),
                     "value": solved * acc["within_1e-3px"] / (ms / 1000.0), "iterations_per_point": float(met[8]),
                     "accuracy_vs_fp64": acc}
    return res


def cpu_model():
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def timed_oracle(ref, dfm, prm, tol, threads, runs=5):
    """The oracle's set_reference, one warm-up run (it builds the Hessian
    pool, the reference's pre-pass) and the median of `runs` further runs on
    the cached reference, each timed on its own; the last run's output and
    node planes."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from helpers import OracleEngine

    oe = OracleEngine()
    t0 = time.perf_counter()
    oe.set_reference(ref)
    t_ref = time.perf_counter() - t0
    t0 = time.perf_counter()
    oe.run(dfm, list(prm), threads=threads, tol=tol)
    t_warm = time.perf_counter() - t0
    ts = []
    for _ in range(runs):
        t0 = time.perf_counter()
        n, cpts, met, planes = oe.run_planes(dfm, list(prm), threads=threads, tol=tol)
        ts.append(time.perf_counter() - t0)
    return dict(set_reference_s=t_ref, warmup_run_s=t_warm, runs_s=ts, median_s=float(np.median(ts)), n=int(n),
                pts=cpts, met=met, planes=planes)


def cpu_baseline(ref, dfm, prm, tol, use6, threads):
    """The oracle (C++/OpenMP restatement of the reference's run_full_field)
    timed on the host cores: the whole frame on `threads` cores, and a
    1024 x 1024 window of it (the central sixteenth of the nodes) on one core,
    each as set_reference, one warm-up run and the median of five runs."""
    allc = timed_oracle(ref, dfm, prm, tol, threads)
    met = allc["met"]
    solved = float(met[3] + met[4] + met[19])
    acc = fp64_accuracy(ref, dfm, int(prm[5]), allc["pts"], use6)
    h, w = ref.shape
    win = min(1024, w, h)
    sub = np.array(prm, np.int32).copy()
    sub[0], sub[1], sub[2], sub[3] = (w - win) // 2, (h - win) // 2, win, win
    one = timed_oracle(ref, dfm, sub, tol, 1)
    solved1 = float(one["met"][3] + one["met"][4] + one["met"][19])
    value = solved * acc["within_1e-3px"] / allc["median_s"]
    res = {
        "value": value,
        "unit": "converged subsets/s",
        "cores": threads,
        "kind": "port",
        "cpu_model": cpu_model(),
        "sample": f"oracle run_full_field on the whole {w}x{h} frame pair, same parameters and stop rule, "
                  f"{threads} threads: set_reference, one warm-up run (the Hessian pool), then the median of 5 "
                  f"runs ({int(met[0])} attempted, {int(solved)} solved subsets, {allc['median_s']:.2f} s), scaled "
                  f"by its own fraction within 1e-3 px of FP64 ({acc['within_1e-3px']:.4f})",
        "set_reference_s": allc["set_reference_s"],
        "warmup_run_s": allc["warmup_run_s"],
        "runs_s": allc["runs_s"],
        "one_core": {"value": solved1 * acc["within_1e-3px"] / one["median_s"], "cores": 1,
                     "sample": f"the central {win}x{win} window of the same pair (rect {sub[:4].tolist()}, "
                               f"{int(one['met'][0])} attempted, {int(solved1)} solved subsets): set_reference, "
                               f"one warm-up run, median of 5 runs ({one['median_s']:.2f} s), scaled by the "
                               f"whole frame's fraction within 1e-3 px",
                     "set_reference_s": one["set_reference_s"], "warmup_run_s": one["warmup_run_s"],
                     "runs_s": one["runs_s"]},
    }
    return res, allc


def parity_report(ref, dfm, prm, tol, gpu, oracle, use6):
    """The GPU field against the oracle's, node by node (the GPU parity
    tests' split analysis): point sets, the largest differences, and the
    nodes over 1e-5 px by cause."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from helpers import oracle_vsg, split_analysis

    n_g, pts_g, pl_g, rep = gpu
    pts_o, pl_o = oracle["pts"], oracle["planes"]
    same_set = len(pts_g) == len(pts_o) and np.array_equal(pts_g[:, :2], pts_o[:, :2])
    out = {"points_oracle": int(len(pts_o)), "points_gpu": int(n_g), "same_point_set": bool(same_set)}
    if not same_set:
        return out
    d = np.abs(pts_g - pts_o)
    res = split_analysis(ref, dfm, prm, pl_g, pl_o, int(prm[5]), tol, guess_g=rep["guess"], path_g=rep["path"],
                         tag_g=rep["tag"])
    how = res["how"]
    pl = pl_g.copy()
    for x, y in res["nodes"]:
        pl[:, y, x] = pl_o[:, y, x]
    ok, exx, eyy, exy = oracle_vsg(pl, int(prm[4]), int(prm[6]))
    ix = ((pts_o[:, 0] - prm[0]) / prm[4]).astype(int)
    iy = ((pts_o[:, 1] - prm[1]) / prm[4]).astype(int)
    ds = np.abs(np.stack([exx[iy, ix], eyy[iy, ix], exy[iy, ix]], 1) - pts_o[:, 4:7])
    out.update({"max_duv": float(d[:, 2:4].max()), "frac_duv_within_1e-3": float((d[:, 2:4].max(1) <= 1e-3).mean()),
                "max_dstrain": float(d[:, 4:7].max()), "max_dcorr": float(d[:, 7].max()),
                "points_strain_over_1e-5": int((d[:, 4:7].max(1) > 1e-5).sum()),
                "nodes_over_1e-5px": len(res["nodes"]),
                "stop_splits": how.count("gpu") + how.count("oracle"), "moved_guess": how.count("guess"),
                "converged_in_oracle_arithmetic": how.count("converged"), "unexplained": how.count("-"),
                "max_dstrain_with_those_nodes_resolved": float(ds.max())})
    return out


def copy_probe():
    """A copy of known size for the FETCH_SIZE / WRITE_SIZE counters: 64 Mi
    fp32 read and written (256 MiB each way) per dispatch."""
    import torch

    x = torch.randn(64 << 20, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        torch.mul(x, 1.0, out=y)
    torch.cuda.synchronize()
    print("copy probe: 256 MiB read, 256 MiB written per dispatch", flush=True)


def frames_in_flight(dev, sa, dref, ddef, prm, cap, tol, w, h, steps, k, expect, learned=True, flags=()):
    """One frame pair solved with k frames in flight: k engines, each with its
    own stream and the same cached reference, driven by k host threads
    (ctypes releases the GIL for the call), so that one frame's host phases
    overlap the others' kernels. Every engine must report the single
    engine's points bit for bit."""
    import threading

    P = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    engs, outs, mets = [], [], []
    for _ in range(k):
        e = dev.sdk_engine_create()
        for f, v in flags:
            assert dev.sdk_engine_set_flag(e, f, v) == 0
        if learned:
            if dev.sdk_engine_load_seeder(e, sa.SEEDER_WEIGHTS.encode()) != 0:
                raise RuntimeError("seeder weights missing: " + sa.SEEDER_WEIGHTS)
            assert dev.sdk_engine_set_flag(e, 5, 1) == 0
        assert dev.sdk_engine_set_reference(e, dref, 1, w, h, None) == 0
        out = pinned_f32(dev, cap)
        met = np.zeros(24, np.float32)
        run = lambda e=e, out=out, met=met: dev.sdk_engine_run(e, ddef, 1, P(prm), P(out), cap, P(met), 24, tol)
        run()
        settle_tuner(dev, e, run)
        run()
        engs.append(e)
        outs.append(out)
        mets.append(met)
    counts = [[] for _ in range(k)]

    def work(i):
        for _ in range(steps):
            counts[i].append(int(dev.sdk_engine_run(engs[i], ddef, 1, P(prm), P(outs[i]), cap, P(mets[i]), 24,
                                                    tol)))

    ths = [threading.Thread(target=work, args=(i,)) for i in range(k)]
    t0 = time.perf_counter()
    for t in ths:
        t.start()
    for t in ths:
        t.join()
    dt = time.perf_counter() - t0
    same = all(c == [expect[0]] * steps for c in counts) and all(
        np.array_equal(o[: expect[0] * 8].view(np.uint32), expect[1].view(np.uint32)) for o in outs)
    solved = float(mets[0][3] + mets[0][4] + mets[0][19])
    for e in engs:
        dev.sdk_engine_destroy(e)
    return {"frames_in_flight": k, "frames": k * steps, "ms_per_frame": dt * 1000.0 / (k * steps),
            "solved_per_frame": solved, "same_points_as_one_engine": bool(same)}


def learned_seed_study(dev, sa, wl, steps, threads, name="c5_4096_bumps", in_flight=True):
    """A workload solved by the plain GPU method, by the GPU method with the
    learned seeder (the committed weights) and, with threads > 0, by the CPU
    oracle (the plain method restated): converged subsets per second, each
    scaled by its fraction within 1e-3 px of the FP64 optimum on a 2000-node
    sample. The learned-seeding workload is 4096^2 with eight seeded Gaussian
    bumps of up to +-30 px plus a global rotation of up to 5 degrees, step 8,
    subset 31."""
    P = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    cfg = wl.config(name)
    w, h = cfg["w"], cfg["h"]
    ref, dfm = wl.pair(dev, name)
    dref, ddef = dev.sdk_malloc(w * h), dev.sdk_malloc(w * h)
    dev.sdk_memcpy(dref, P(ref), w * h, 0)
    dev.sdk_memcpy(ddef, P(dfm), w * h, 0)
    prm = np.array(wl.solve_params(name), np.int32)
    cap = (w // cfg["step"] + 1) * (h // cfg["step"] + 1) * 8
    tol = ctypes.c_float(cfg["tol"])
    what = ("eight Gaussian bumps (+-30 px, width 200-800 px) and a rotation of up to 5 deg"
            if name == "c5_4096_bumps" else "the workload's field")
    res = {"workload": f"{name}: {w}x{h} speckle, {what}, step {cfg['step']}, subset {cfg['subset']}",
           "weights": "seeder_weights.bin"}
    single = None
    modes = ("plain_gpu", "learned_seed_gpu")
    engines = {}
    forced = 0
    for mode in modes:
        eng = dev.sdk_engine_create()
        if mode == "learned_seed_gpu":
            if dev.sdk_engine_load_seeder(eng, sa.SEEDER_WEIGHTS.encode()) != 0:
                raise RuntimeError("seeder weights missing: " + sa.SEEDER_WEIGHTS)
            assert dev.sdk_engine_set_flag(eng, 5, 1) == 0
        assert dev.sdk_engine_set_reference(eng, dref, 1, w, h, None) == 0
        out = pinned_f32(dev, cap)
        met = np.zeros(24, np.float32)
        if mode == "learned_seed_gpu" and forced and not in_flight:
            # the plain engine's settled launch shape, so that the tuners'
            # choices do not differ between the two
            assert dev.sdk_engine_set_flag(eng, 7, forced) == 0
        dev.sdk_engine_run(eng, ddef, 1, P(prm), P(out), cap, P(met), 24, tol)   # warm (pool, pyramids)
        settle_tuner(dev, eng, lambda: dev.sdk_engine_run(eng, ddef, 1, P(prm), P(out), cap, P(met), 24, tol))
        if mode == "plain_gpu":
            shape = np.zeros(2, np.int32)
            shape_ms = np.zeros(6, np.float64)
            if dev.sdk_engine_tuning(eng, P(shape), P(shape_ms)) == 0 and int(shape[0]) in (1, 2, 4):
                forced = {1: 1, 2: 2, 4: 3}[int(shape[0])] + (3 if int(shape[1]) > 0 else 0)
        engines[mode] = (eng, out, met)
    # Where both methods do the same work (the main workload, in_flight
    # False), their runs are interleaved on one launch shape and each is the
    # median of its runs, so that drift of the clock or the host does not
    # favour either; otherwise each engine's runs are timed back to back, so
    # that neither finds the caches as the other left them.
    runs = {mode: [] for mode in modes}
    last_n = {}
    order = [modes] * max(steps, 5) if not in_flight else [(m,) * max(steps, 5) for m in modes]
    for group in order:
        for mode in group:
            eng, out, met = engines[mode]
            t0 = time.perf_counter()
            last_n[mode] = dev.sdk_engine_run(eng, ddef, 1, P(prm), P(out), cap, P(met), 24, tol)
            runs[mode].append((time.perf_counter() - t0) * 1000.0)
    for mode in modes:
        eng, out, met = engines[mode]
        n = last_n[mode]
        ms = float(np.median(runs[mode]))
        ss = np.zeros(3, np.int32)
        dev.sdk_engine_seed_stats(eng, P(ss))
        tm = np.zeros(14, np.float64)
        dev.sdk_engine_timings(eng, P(tm))
        dev.sdk_engine_destroy(eng)
        n = int(n)
        pts = out[: max(n, 0) * 8].reshape(-1, 8).copy()
        acc = fp64_accuracy(ref, dfm, cfg["subset"], pts, False) if n > 0 else None
        solved = float(met[3] + met[4] + met[19]) if n > 0 else 0.0
        frac = acc["within_1e-3px"] if acc else 0.0
        if mode == "learned_seed_gpu" and n > 0:
            single = (n, out[: n * 8].copy())
        res[mode] = {"ms_per_step": ms, "output_points": n, "value": solved * frac / (ms / 1000.0),
                     "accuracy_vs_fp64": acc, "anchors_cnn_seeded": int(ss[0]), "anchors_cnn_held": int(ss[1]),
                     "anchors_plain_fallback": int(ss[2]),
                     "phases_ms": {k: float(v) for k, v in zip(
                         ["total", "prep", "seeding", "prepass", "mesh", "path_a", "path_b", "strain"], tm[:8])},
                     "path_a_kernel_ms": float(tm[8]), "path_a_iterations": int(tm[10]),
                     "path_b_launches": int(tm[12]), "path_b_solves": int(tm[13])}
    if single is not None and in_flight:
        frac = (res["learned_seed_gpu"]["accuracy_vs_fp64"] or {}).get("within_1e-3px", 0.0)
        for k in (2, 3):
            fl = frames_in_flight(dev, sa, dref, ddef, prm, cap, tol, w, h, max(4, 2 * steps), k, single)
            fl["value"] = fl["solved_per_frame"] * frac / (fl["ms_per_frame"] / 1000.0)
            res[f"learned_seed_gpu_{k}_in_flight"] = fl
    dev.sdk_free(dref)
    dev.sdk_free(ddef)
    if threads:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from helpers import OracleEngine

        oe = OracleEngine()
        oe.set_reference(ref)
        t0 = time.perf_counter()
        n_o, _, met_o = oe.run(dfm, list(prm), threads=threads, tol=cfg["tol"])
        dt = time.perf_counter() - t0
        res["cpu_oracle"] = {"seconds": dt, "output_points": int(n_o), "cores": threads,
                             "value": float(met_o[3] + met_o[4] + met_o[19]) / dt if n_o > 0 else 0.0}
    lv, pv = res["learned_seed_gpu"]["value"], res["plain_gpu"]["value"]
    res["learned_over_plain"] = (lv / pv) if pv > 0 else None
    if pv <= 0:
        res["learned_over_plain_note"] = ("the plain GPU method reports no converged subset on this workload "
                                          "(every anchor and Path C seed fails), so the ratio is unbounded")
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=14,
                    help="untimed runs; the launch-shape tuner times its 6 shapes twice each in the first 12")
    ap.add_argument("--workload", default="c4_4096_heterogeneous")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline and FP64 check")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-precision", action="store_true", help="skip the precision-variant study")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-seeding comparison")
    ap.add_argument("--no-sequence", action="store_true", help="skip the 64-frame incremental sequence")
    ap.add_argument("--u8-deformed", action="store_true",
                    help="the solver samples the u8 deformed frame instead of its FP32 copy (same field)")
    ap.add_argument("--coef-plane", action="store_true",
                    help="Keys 4x4 from the fp16 row-cubic coefficient plane of each frame (flag 17)")
    ap.add_argument("--flag", action="append", default=[], metavar="F=V",
                    help="another engine switch for the main line's engines (sdk_engine_set_flag F V)")
    ap.add_argument("--tol", type=float, default=None, help="IC-GN stop rule ||dp|| < tol (default: the workload's)")
    ap.add_argument("--learned-seed", action="store_true",
                    help="the main line's engine seeds its anchors with the learned seeder (the committed weights)")
    ap.add_argument("--sequence-only", type=int, default=0, metavar="FRAMES",
                    help="only run the incremental sequence study on its first FRAMES frames (for a profiler) "
                         "and print its JSON")
    ap.add_argument("--reference-only", type=int, default=0, metavar="RUNS",
                    help="only set the workload's reference and solve its frame RUNS times, each after a fresh "
                         "set_reference (the reference precompute under a profiler), and print the prepass times")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a plain 256 MiB copy kernel three times (the calibration of the HBM traffic "
                         "counters) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus above 1 is not supported"}))
        sys.exit(2)
    if args.copy_probe:
        copy_probe()
        return

    sa = _load("semper_amd", os.path.join(PKG, "semper_amd.py"))
    wl = _load("workloads", os.path.join(PKG, "workloads.py"))
    dev = sa.load_dev()
    if dev.sdk_device_count() < 1:
        raise RuntimeError("no GPU visible")
    if args.sequence_only:
        print(json.dumps({"sequence": sequence_study(dev, sa, wl, frames_limit=args.sequence_only)}))
        return
    cfg = wl.config(args.workload)
    w, h = cfg["w"], cfg["h"]
    prm = np.array(wl.solve_params(args.workload), np.int32)
    if args.reference_only:
        dref, ddef = wl.pair(dev, args.workload, on_device=True)
        eng = dev.sdk_engine_create()
        cap = (w // cfg["step"]) * (h // cfg["step"]) * 8
        out = np.zeros(cap, np.float32)
        met = np.zeros(24, np.float32)
        tm = np.zeros(14, np.float64)
        P = lambda a: a.ctypes.data_as(ctypes.c_void_p)
        res = []
        for _ in range(args.reference_only):
            assert dev.sdk_engine_set_reference(eng, dref, 1, w, h, None) == 0
            n = dev.sdk_engine_run(eng, ddef, 1, P(prm), P(out), cap, P(met), 24, ctypes.c_float(cfg["tol"]))
            assert n > 0, n
            dev.sdk_engine_timings(eng, P(tm))
            res.append({"points": int(n), "prepass_ms": float(tm[3]), "total_ms": float(tm[0])})
        print(json.dumps({"reference_only": args.workload, "runs": res}))
        return
    use6 = bool(prm[7])
    t0 = time.perf_counter()
    dref, ddef = wl.pair(dev, args.workload, on_device=True)
    t_gen = time.perf_counter() - t0
    eng = dev.sdk_engine_create()
    if args.learned_seed:
        if dev.sdk_engine_load_seeder(eng, sa.SEEDER_WEIGHTS.encode()) != 0:
            raise RuntimeError("seeder weights missing: " + sa.SEEDER_WEIGHTS)
        assert dev.sdk_engine_set_flag(eng, 5, 1) == 0
    # the main line's engine switches, applied to every engine of the same solve
    engine_flags = ([(13, 0)] if args.u8_deformed else []) + ([(17, 1)] if args.coef_plane else []) + \
        [tuple(int(x) for x in f.split("=")) for f in args.flag]
    for f, v in engine_flags:
        assert dev.sdk_engine_set_flag(eng, f, v) == 0
    assert dev.sdk_engine_set_reference(eng, dref, 1, w, h, None) == 0
    gw, gh = w // cfg["step"], h // cfg["step"]
    cap = gw * gh * 8
    out = pinned_f32(dev, cap)
    met = np.zeros(24, np.float32)
    tm = np.zeros(14, np.float64)
    P = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    tol_v = args.tol if args.tol is not None else cfg["tol"]
    tol = ctypes.c_float(tol_v)
    prepass_ms = None
    for _ in range(args.warmup):
        n = dev.sdk_engine_run(eng, ddef, 1, P(prm), P(out), cap, P(met), 24, tol)
        assert n > 0, n
        if prepass_ms is None:
            dev.sdk_engine_timings(eng, P(tm))
            prepass_ms = float(tm[3])
    # the timed steps run in the launch shape the tuner settled on
    settle_tuner(dev, eng, lambda: dev.sdk_engine_run(eng, ddef, 1, P(prm), P(out), cap, P(met), 24, tol))
    kern_ms, kern_iters, kern_nodes = [], [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        n = dev.sdk_engine_run(eng, ddef, 1, P(prm), P(out), cap, P(met), 24, tol)
        dev.sdk_engine_timings(eng, P(tm))
        kern_ms.append(tm[8])
        kern_iters.append(tm[10])
        kern_nodes.append(tm[9])
    dt = time.perf_counter() - t0
    ms_per_step = dt * 1000.0 / args.steps
    # the same solve with the deformed frame handed over in host memory (PCIe
    # copy included), as semper_run receives it; reported beside, never as value
    host_def = np.zeros((h, w), np.uint8)
    dev.sdk_memcpy(P(host_def), ddef, w * h, 1)
    met_h = np.zeros(24, np.float32)
    t1 = time.perf_counter()
    nh = max(1, args.steps // 2)
    for _ in range(nh):
        dev.sdk_engine_run(eng, P(host_def), 0, P(prm), P(out), cap, P(met_h), 24, tol)
    host_ms = (time.perf_counter() - t1) * 1000.0 / nh
    # and with the points written to pageable host memory (a numpy buffer):
    # the copy then goes through the runtime's staging; reported beside value
    out_pg = np.zeros(cap, np.float32)
    t1 = time.perf_counter()
    for _ in range(nh):
        dev.sdk_engine_run(eng, ddef, 1, P(prm), P(out_pg), cap, P(met_h), 24, tol)
    pageable_ms = (time.perf_counter() - t1) * 1000.0 / nh
    # restore the device-input run's output for the checks below
    n = dev.sdk_engine_run(eng, ddef, 1, P(prm), P(out), cap, P(met), 24, tol)
    # the same frame with 2 and 3 frames in flight (engines on their own
    # streams and host threads); reported beside value, never as it
    in_flight = {}
    if n > 0:
        for k in (2, 3):
            in_flight[str(k)] = frames_in_flight(dev, sa, dref, ddef, prm, cap, tol, w, h, max(4, args.steps), k,
                                                 (int(n), out[: int(n) * 8].copy()), learned=args.learned_seed,
                                                 flags=engine_flags)
    solved = float(met[3] + met[4] + met[19])
    n = int(n)
    shape = np.zeros(2, np.int32)
    shape_ms = np.zeros(6, np.float64)
    launch_shape = None
    tune_rc = dev.sdk_engine_tuning(eng, P(shape), P(shape_ms))
    if tune_rc >= 0:
        launch_shape = {"subsets_per_workgroup": int(shape[0]), "tile": int(shape[1]), "settled": tune_rc == 0,
                        "candidates_ms": {f"{a}x{b}": round(float(t), 4) for (a, b), t in zip(
                            [(1, 0), (2, 0), (4, 0), (1, 8), (2, 8), (4, 8)], shape_ms) if t >= 0}}
    n_sub = cfg["subset"] ** 2
    # dominant kernel: the Path A solver launch over the grid
    avg_kms = float(np.mean(kern_ms))
    flops = float(np.mean(kern_iters)) * n_sub * FLOP_PER_PIXEL_ITER[use6] + float(met[3]) * n_sub * SD_FLOP_PER_PIXEL
    achieved = float(flops / (avg_kms * 1e-3) / 1e12)
    pts = out[: n * 8].reshape(-1, 8).copy()

    cpu = None
    parity = None
    acc = None
    precision = None
    learned = None
    if not args.no_cpu:
        ref = np.zeros((h, w), np.uint8)
        dfm = np.zeros((h, w), np.uint8)
        dev.sdk_memcpy(P(ref), dref, w * h, 1)
        dev.sdk_memcpy(P(dfm), ddef, w * h, 1)
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        acc = fp64_accuracy(ref, dfm, cfg["subset"], pts, use6)
        cpu, orc = cpu_baseline(ref, dfm, prm, tol_v, use6, threads)
        # the node planes of the same solve, from an engine with the node report on
        e2 = dev.sdk_engine_create()
        assert dev.sdk_engine_set_flag(e2, 12, 1) == 0
        for f, v in engine_flags:
            assert dev.sdk_engine_set_flag(e2, f, v) == 0
        if args.learned_seed:
            dev.sdk_engine_load_seeder(e2, sa.SEEDER_WEIGHTS.encode())
            dev.sdk_engine_set_flag(e2, 5, 1)
        assert dev.sdk_engine_set_reference(e2, dref, 1, w, h, None) == 0
        out2 = np.zeros(cap, np.float32)
        met2 = np.zeros(24, np.float32)
        n2 = int(dev.sdk_engine_run(e2, ddef, 1, P(prm), P(out2), cap, P(met2), 24, tol))
        ogh, ogw = orc["planes"].shape[1:]
        pl2 = np.zeros(7 * ogw * ogh, np.float32)
        gs2 = np.zeros(6 * ogw * ogh, np.float32)
        tg2 = np.zeros(ogw * ogh, np.uint8)
        pa2 = np.zeros(ogw * ogh, np.uint8)
        assert dev.sdk_engine_node_planes(e2, 0, P(pl2), pl2.size) == pl2.size
        assert dev.sdk_engine_node_planes(e2, 2, P(gs2), gs2.size) == gs2.size
        assert dev.sdk_engine_node_report(e2, P(tg2), P(pa2), tg2.size) == tg2.size
        dev.sdk_engine_destroy(e2)
        rep = {"guess": gs2.reshape(6, ogh, ogw), "tag": tg2.reshape(ogh, ogw), "path": pa2.reshape(ogh, ogw)}
        parity = parity_report(ref, dfm, prm, tol_v, (n2, out2[: n2 * 8].reshape(-1, 8), pl2.reshape(7, ogh, ogw),
                                                      rep), orc, use6)
        parity["same_points_as_timed_engine"] = bool(n2 == n and np.array_equal(out2[: n2 * 8], out[: n * 8]))
        if not args.no_precision:
            precision = precision_study(dev, dref, ddef, prm, cap, tol, ref, dfm, cfg["subset"], use6,
                                        max(2, args.steps // 2))
        if not args.no_learned:
            learned = learned_seed_study(dev, sa, wl, 3, threads)
            learned["on_the_main_workload"] = learned_seed_study(dev, sa, wl, 5, 0, name=args.workload,
                                                                 in_flight=False)
    sequence = None if args.no_sequence else sequence_study(dev, sa, wl)
    frac = acc["within_1e-3px"] if acc else 1.0
    value = float(solved * frac / (ms_per_step / 1000.0))
    for fl in in_flight.values():
        fl["value"] = fl["solved_per_frame"] * frac / (fl["ms_per_frame"] / 1000.0)
    line = {
        "metric": "converged subsets/sec at \u22641e-3 px: learned-seed vs plain GPU vs CPU, 4096\u00b2",
        "value": value,
        "unit": "converged subsets/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"{args.workload}: {w}x{h} speckle, step {cfg['step']}, subset {cfg['subset']}, "
                               f"strain window {cfg['window']}, {'Keys 6x6' if use6 else 'Keys 4x4'} IC-GN",
                   "method": "learned-seed GPU" if args.learned_seed else "plain GPU",
                   "attempted_subsets": int(met[0]), "solved_subsets": int(solved), "output_points": int(n),
                   "pair_generation_s": round(t_gen, 3), "stop_rule": tol_v,
                   "output_buffer": "page-locked host memory"},
        "roofline": {"bound": "valu", "compute_roof": "FP32 vector ALU (MI355X FP32 peak 157.3 TFLOP/s)",
                     "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32_TFLOPS,
                     "traffic": None if (engine_flags or args.learned_seed) else TRAFFIC_BYTES.get(args.workload),
                     "traffic_unit": "bytes per launch",
                     "kernel": "Path A subset solver launch (k_fast + k_slow)", "avg_launch_ms": avg_kms,
                     "iterations_per_launch": float(np.mean(kern_iters)), "launch_shape": launch_shape},
        "cpu_baseline": cpu,
        "accuracy_vs_fp64": acc,
        "parity_vs_oracle": parity,
        "host_output_pageable": {"ms_per_step": pageable_ms,
                                 "value": float(solved * (acc["within_1e-3px"] if acc else 1.0) / (pageable_ms / 1000.0)),
                                 "note": "points written to pageable host memory instead of the page-locked buffer "
                                         "of the timed steps"},
        "host_input": {"ms_per_step": host_ms, "value": float(solved * (acc["within_1e-3px"] if acc else 1.0)
                                                               / (host_ms / 1000.0)),
                       "note": "deformed frame passed in host memory (pageable), PCIe copy included"},
        "metrics": [float(x) for x in met],
        "phases_ms": {k: float(v) for k, v in zip(
            ["total", "prep", "seeding", "prepass", "mesh", "path_a", "path_b", "strain"], tm[:8])},
        "path_b": {"speculative_solves": int(tm[11]), "further_launches": int(tm[12]), "further_solves": int(tm[13])},
        "frames_in_flight": in_flight,
        "prepass_ms": prepass_ms,
        "precision_study": precision,
        "learned_seed": learned,
        "sequence": sequence,
    }
    dev.sdk_engine_destroy(eng)
    dev.sdk_free(dref)
    dev.sdk_free(ddef)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
