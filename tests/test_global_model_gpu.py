"""The engine's global-motion model (sdk_engine_global_model) against the
oracle's on the same pairs, and against the motion the benchmark's affine
workload renders: the 2048^2 pair under a 2 degree rotation, exx 1%,
eyy -0.3%, exy 0.5% and a (12.3, -7.8) px translation about the frame
centre, which is also the model's centre."""
import ctypes

import numpy as np
import pytest

from helpers import P, engine_module, model_at, oracle_global_motion
from test_global_model import CASES, CX, CY, RECT, pair

pytestmark = pytest.mark.gpu
sa = engine_module()


@pytest.fixture(scope="module")
def dev():
    lib = sa.load_dev()
    assert lib.sdk_device_count() >= 1, "no GPU visible"
    return lib


def engine_model(dev, ref, dfm, prm, tol=1e-3):
    h, w = ref.shape
    e = dev.sdk_engine_create()
    assert dev.sdk_engine_set_reference(e, P(np.ascontiguousarray(ref)), 0, w, h, None) == 0
    pr = np.array(prm, np.int32)
    cap = (prm[2] // prm[4] + 1) * (prm[3] // prm[4] + 1) * 8
    out = np.zeros(cap, np.float32)
    met = np.zeros(24, np.float32)
    n = dev.sdk_engine_run(e, P(np.ascontiguousarray(dfm)), 0, P(pr), P(out), cap, P(met), 24, ctypes.c_float(tol))
    g = np.zeros(12, np.float64)
    assert dev.sdk_engine_global_model(e, P(g)) == 0
    dev.sdk_engine_destroy(e)
    m = dict(locked=bool(g[0]), full=bool(g[1]), a=g[2:8].copy(), cx=g[8], cy=g[9], patches=int(g[10]),
             passes=int(g[11]))
    return m, int(n), out[: max(n, 0) * 8].reshape(-1, 8), met


@pytest.mark.parametrize("name", sorted(CASES))
def test_engine_model_is_the_oracles(dev, name):
    ref, dfm = pair(CASES[name])
    m_o = oracle_global_motion(ref, dfm, RECT)
    m_g, n, _, met = engine_model(dev, ref, dfm, list(RECT) + [16, 21, 49, 0])
    print(f"{name}: engine a {np.round(m_g['a'], 6).tolist()} pass {m_g['passes']}", flush=True)
    assert n > 0 and met[21] == 1.0
    assert m_g["locked"] == m_o["locked"] and m_g["full"] == m_o["full"] and m_g["patches"] == m_o["patches"]
    assert (m_g["cx"], m_g["cy"]) == (CX, CY)
    # the same patches and shifts; the two fits' double sums may round apart
    assert np.abs(m_g["a"] - m_o["a"]).max() <= 1e-9


def test_affine_workload_model_and_field(dev):
    wl = sa_workloads()
    name = "c2_2048_affine"
    cfg = wl.config(name)
    ref, dfm = wl.pair(dev, name)
    prm = list(wl.solve_params(name))
    m, n, pts, met = engine_model(dev, ref, dfm, prm, cfg["tol"])
    want = wl.affine_params(2.0, 0.01, -0.003, 0.005, 12.3, -7.8)   # u, v, ux, uy, vx, vy
    assert (m["cx"], m["cy"]) == (cfg["w"] / 2, cfg["h"] / 2)
    a = m["a"]
    centre = np.hypot(a[0] - want[0], a[3] - want[1])
    grad = np.abs(np.array([a[1], a[2], a[4], a[5]]) - np.array(want[2:])).max()
    # at the patch centres (8 x 8 on this ROI)
    g = 8
    xs = np.array([round((i + 0.5) * cfg["w"] / g) for i in range(g)], float)
    X, Y = np.meshgrid(xs, xs)
    mu, mv = model_at(m, X, Y)
    tu, tv = wl.displacement(name, X, Y)
    err = np.hypot(mu - tu, mv - tv).max()
    # the field: every output point against the analytic affine
    tu_p, tv_p = wl.displacement(name, pts[:, 0].astype(np.float64), pts[:, 1].astype(np.float64))
    rms = float(np.sqrt(np.mean((pts[:, 2] - tu_p) ** 2 + (pts[:, 3] - tv_p) ** 2)))
    print(f"affine workload: model a {np.round(a, 6).tolist()}, truth {np.round(want, 6).tolist()}; centre "
          f"{centre:.4f} px, gradients {grad:.1e}, patch centres max {err:.4f} px; field {n} points, rms {rms:.5f} px",
          flush=True)
    assert m["locked"] and m["full"]
    assert centre <= 0.25 and grad <= 3e-3 and err <= 0.5
    # the field within the reference's own end-to-end bound on synthetic
    # truth (0.05 px rms, test_seeding_contracts.cpp:244,294): the rest is the
    # Keys interpolator's bias on this 2-4 px speckle
    assert n > 150000 and rms <= 0.05, rms


def sa_workloads():
    import importlib.util
    import os
    from helpers import PKG
    spec = importlib.util.spec_from_file_location("workloads", os.path.join(PKG, "workloads.py"))
    wl = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(wl)
    return wl
