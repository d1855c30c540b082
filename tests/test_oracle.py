"""The CPU oracle pinned against the reference's own golden fixtures and
known answers (no GPU). The oracle is the checker of every GPU parity test,
so it is checked first."""
import ctypes

import numpy as np
import pytest

from helpers import (FF_SLOTS, P, CORPUS_SCENARIOS, OracleEngine, corpus_images, corpus_nodes, ff_golden_images,
                     fp64_polish, full_field_golden, golden_corpus, oracle, oracle_solve_nodes, ref_speckle)


def on_truth(x, y, u, v, s):
    sc = CORPUS_SCENARIOS[s]
    dx, dy = x - 256.0, y - 256.0
    return np.hypot(u - (sc[0] + sc[2] * dx + sc[3] * dy), v - (sc[1] + sc[4] * dx + sc[5] * dy)) <= 0.1


@pytest.mark.parametrize("kind,use6", [("bicubic", 0), ("keys6x6", 1)])
def test_oracle_golden_corpus(kind, use6):
    """tests/integration/test_golden_corpus.cpp:189-327: 554 subset solves per
    interpolator, statuses identical (bar slow-convergence flips), values
    within 1e-2 px / 1e-3 (5x at local minima)."""
    g = golden_corpus(kind)
    cap = 1000
    iv = np.zeros((cap, 3), np.int32)
    fv = np.zeros((cap, 7), np.float32)
    n = oracle().orc_golden_corpus(use6, cap, P(iv), P(fv))
    assert n == len(g)
    iv, fv = iv[:n], fv[:n]
    assert (iv[:, 0] == g["x"]).all() and (iv[:, 1] == g["y"]).all()
    # the corpus lists scenario 0's initialised nodes, then scenario 1's; both
    # scenarios share the reference image, so the same nodes initialise and
    # each scenario holds half of the list
    assert n % 2 == 0 and (iv[: n // 2, :2] == iv[n // 2:, :2]).all()
    flips = 0
    for i in range(n):
        gv = np.array([g[k][i] for k in ("u", "v", "ux", "uy", "vx", "vy")])
        s = 0 if i < n // 2 else 1
        exact = on_truth(g["x"][i], g["y"][i], g["u"][i], g["v"][i], s)
        sc = 1.0 if exact else 5.0
        agree = (np.abs(gv[:2] - fv[i, :2]) <= 1e-2 * sc).all() and (np.abs(gv[2:] - fv[i, 2:6]) <= 1e-3 * sc).all()
        if g["status"][i] != iv[i, 2]:
            assert g["corr"][i] < 2.0 and fv[i, 6] < 2.0 and agree, i
            flips += 1
            continue
        if g["status"][i] == 0:
            assert agree, (i, gv, fv[i])
    assert flips <= 6


@pytest.mark.parametrize("scenario", [0, 1])
def test_oracle_full_field_golden(scenario):
    """tests/integration/test_full_field_golden.cpp:280-356: exact point count
    and grid coordinates, u/v within 5e-3 px, strain and corr within 1e-4, and
    the counting metrics slots the replaced seeding front end cannot move."""
    code, counts, gpts = full_field_golden()[scenario]
    ref, dfm = ff_golden_images(scenario)
    oe = OracleEngine()
    assert oe.set_reference(ref) == 0
    n, pts, met = oe.run(dfm, (0, 0, 320, 320, 12, 21, 48, 0))
    assert n == code
    assert (pts[:, :2] == gpts[:, :2]).all()
    d = np.abs(pts - gpts).max(0)
    assert d[2] <= 5e-3 and d[3] <= 5e-3
    assert d[4:7].max() <= 1e-4 and d[7] <= 1e-4
    m = met[FF_SLOTS]
    for k, slot in enumerate(FF_SLOTS):
        if slot in (0, 1, 2, 3, 4, 16, 19, 20, 21, 22, 23):
            assert abs(m[k] - counts[k]) <= (1e-4 if slot == 22 else 0), (slot, m[k], counts[k])
    # accounting identity (docs/CONTRACT.md metrics): 3 + 4 + 19 == 1 + 20 + 23
    assert met[3] + met[4] + met[19] == met[1] + met[20] + met[23]


def test_interpolators_exact_at_integers_and_ramps():
    """tests/unit/test_image.cpp:42-131: every interpolator reproduces the
    samples at integer positions and a linear ramp anywhere."""
    W = H = 32
    yy, xx = np.mgrid[0:H, 0:W].astype(np.float32)
    ramp = (3.0 * xx + 2.0 * yy + 7.0).astype(np.float32)
    rng = np.random.default_rng(1)
    img = rng.uniform(0, 255, (H, W)).astype(np.float32)
    xs = rng.uniform(4, W - 6, 200).astype(np.float32)
    ys = rng.uniform(4, H - 6, 200).astype(np.float32)
    ix = np.floor(xs).astype(np.float32)
    iy = np.floor(ys).astype(np.float32)
    out = np.zeros(200, np.float32)
    for kind in (0, 1, 2):
        oracle().orc_interp(P(img), W, H, kind, 200, P(ix), P(iy), P(out))
        assert np.allclose(out, img[iy.astype(int), ix.astype(int)], atol=1e-3)
        oracle().orc_interp(P(ramp), W, H, kind, 200, P(xs), P(ys), P(out))
        assert np.allclose(out, 3.0 * xs + 2.0 * ys + 7.0, atol=2e-3)


def test_gradients_exact_on_cubics():
    """The 5-tap 4th-order stencil is exact for polynomials up to degree 4."""
    W, H = 40, 30
    yy, xx = np.mgrid[0:H, 0:W].astype(np.float64)
    f = (0.01 * xx ** 3 - 0.2 * xx ** 2 + 3 * xx + 0.02 * yy ** 3 + yy).astype(np.float32)
    gx = np.zeros(W * H, np.float32)
    gy = np.zeros(W * H, np.float32)
    oracle().orc_gradients(P(np.ascontiguousarray(f)), W, H, P(gx), P(gy))
    gx, gy = gx.reshape(H, W), gy.reshape(H, W)
    tx = 0.03 * xx ** 2 - 0.4 * xx + 3
    ty = 0.06 * yy ** 2 + 1
    assert np.abs(gx[2:-2, 2:-2] - tx[2:-2, 2:-2]).max() < 2e-3
    assert np.abs(gy[2:-2, 2:-2] - ty[2:-2, 2:-2]).max() < 2e-3
    assert (gx[:2] == 0).all() and (gx[:, -2:] == 0).all()


def test_vsg_linear_field_exact():
    """tests/unit/test_strain_calculator.cpp:63-102: a linear displacement field
    gives its Green-Lagrange strain exactly (1e-5)."""
    gw, gh, step = 20, 16, 5
    y, x = np.mgrid[0:gh, 0:gw].astype(np.float64) * step
    ux, uy, vx, vy = 0.01, 0.004, -0.003, 0.02
    u = (0.5 + ux * x + uy * y).astype(np.float32).ravel()
    v = (-0.2 + vx * x + vy * y).astype(np.float32).ravel()
    valid = np.ones(gw * gh, np.uint8)
    exx = np.zeros(gw * gh, np.float32)
    eyy = np.zeros_like(exx)
    exy = np.zeros_like(exx)
    rej = oracle().orc_outliers_and_vsg(gw, gh, step, 21, P(u), P(v), P(valid), P(exx), P(eyy), P(exy))
    assert rej == 0
    ok = exx > -999
    assert ok.sum() > 0.5 * gw * gh
    assert np.abs(exx[ok] - 0.5 * (2 * ux + ux * ux + vx * vx)).max() < 1e-5
    assert np.abs(eyy[ok] - 0.5 * (2 * vy + uy * uy + vy * vy)).max() < 1e-5
    assert np.abs(exy[ok] - 0.5 * (uy + vx + ux * uy + vx * vy)).max() < 1e-5


def test_median_test_rejects_a_blunder():
    """strain_calculator.cpp:174-224: one wrong match in a smooth field is
    rejected, its neighbours are not."""
    gw = gh = 9
    u = np.full(gw * gh, 1.0, np.float32) + np.linspace(0, 0.05, gw * gh, dtype=np.float32)
    v = np.zeros(gw * gh, np.float32)
    u[4 * gw + 4] += 3.0
    valid = np.ones(gw * gh, np.uint8)
    e = [np.zeros(gw * gh, np.float32) for _ in range(3)]
    rej = oracle().orc_outliers_and_vsg(gw, gh, 5, 21, P(u), P(v), P(valid), P(e[0]), P(e[1]), P(e[2]))
    assert rej == 1 and valid[4 * gw + 4] == 0 and valid.sum() == gw * gh - 1


def test_vsg_window_node_count():
    """strain_calculator.cpp:12-25 closed form on a few (step, window) pairs."""
    f = oracle().orc_vsg_window_node_count
    assert f(12, 48) == 13
    assert f(10, 21) == 5
    assert f(8, 31) == 9
    assert f(16, 16) == 1


def test_fp64_oracle_translation_truth():
    """The FP64 IC-GN recovers an analytic sub-pixel translation to the
    interpolation bias of the model (<= 0.02 px, test_optimization_engine.cpp:43-48),
    and agrees with the FP32 engine to its stop rule."""
    W = H = 160
    ref = ref_speckle(1234, W, H, 400)
    dfm = ref_speckle(1234, W, H, 400, [2.25, -1.5, 0, 0, 0, 0, 80, 80])
    nodes = [(80, 80), (70, 90), (90, 70)]
    io, fo = oracle_solve_nodes(ref, dfm, 31, nodes, [[2.0, -2.0, 0, 0, 0, 0]] * 3, [2] * 3)
    assert (io[:, 1] == 0).all()
    assert np.abs(fo[:, 0] - 2.25).max() <= 0.02 and np.abs(fo[:, 1] + 1.5).max() <= 0.02
    si, po, _ = fp64_polish(ref, dfm, 31, nodes, fo[:, :6].astype(np.float64))
    assert (si[:, 0] == 0).all()
    assert np.abs(po[:, 0] - 2.25).max() <= 0.02
    # the FP32 solve stops at ||dp|| < 1e-3; it sits within ~1e-3 of the FP64 optimum
    assert np.abs(po[:, 0] - fo[:, 0]).max() <= 2e-3 and np.abs(po[:, 1] - fo[:, 1]).max() <= 2e-3


def test_fp64_oracle_fixed_point_is_znssd_minimum():
    """The FP64 optimum is a stationary point: a second polish from it does
    not move, and a perturbed start returns to it."""
    ref, dfm = corpus_images(0)
    nodes = corpus_nodes()[40:60]
    p0 = np.zeros((len(nodes), 6))
    p0[:, 0], p0[:, 1] = 3.2, -1.7
    si, po, _ = fp64_polish(ref, dfm, 27, nodes, p0)
    ok = si[:, 0] == 0
    assert ok.sum() >= 8   # the corpus speckle is sparse: flat subsets do not converge
    si2, po2, _ = fp64_polish(ref, dfm, 27, nodes, po + np.array([0.2, -0.2, 0.002, 0, 0, -0.002]))
    ok2 = ok & (si2[:, 0] == 0)
    assert np.abs(po2[ok2] - po[ok2]).max() < 1e-6
