"""The reference's seeding contracts (tests/integration/test_seeding_contracts.cpp)
restated for the engine's global-motion front end, which replaces AKAZE and
its homography: the rotation gate (:128-177) at 5, 15, 120 and 180 degrees
and a 10 degree x 1.05 similarity shifted by (20, -12) px, the guess at the
grid edge under an 85 degree turn (:249-271), a half turn solved end to end
(:273-296) and a 45% stretch seeded without blunders (:179-247). Images are the
reference suite's speckle field (640 x 640, 1600 blobs, noise-free), the ROI
its big-motion one (160, 160, 320, 320), step 5, subset 41, strain window 27;
motions about the image centre. The bounds are the reference's."""
import ctypes
import math

import numpy as np
import pytest

from helpers import P, aff_truth, engine_module, gray8, ref_speckle, stretch_rotate_aff

pytestmark = pytest.mark.gpu
sa = engine_module()
W = H = 640
PRM = (160, 160, 320, 320, 5, 41, 27, 0)


@pytest.fixture(scope="module")
def dev():
    lib = sa.load_dev()
    assert lib.sdk_device_count() >= 1, "no GPU visible"
    return lib


def similarity(deg, s=1.0, u=0.0, v=0.0):
    t = math.radians(deg)
    return [u, v, s * math.cos(t) - 1.0, -s * math.sin(t), s * math.sin(t), s * math.cos(t) - 1.0, W / 2, H / 2]


def pair(aff, seed=11, w=W, h=H, blobs=1600):
    ref = gray8(ref_speckle(seed, w, h, blobs), 0.0, 101)
    dfm = gray8(ref_speckle(seed, w, h, blobs, aff), 0.0, 202)
    return ref, dfm


def run(dev, ref, dfm, prm):
    h, w = ref.shape
    e = dev.sdk_engine_create()
    assert dev.sdk_engine_set_flag(e, 12, 1) == 0
    assert dev.sdk_engine_set_reference(e, P(np.ascontiguousarray(ref)), 0, w, h, None) == 0
    pr = np.array(prm, np.int32)
    gw, gh = prm[2] // prm[4], prm[3] // prm[4]
    cap = (gw + 1) * (gh + 1) * 8
    out = np.zeros(cap, np.float32)
    met = np.zeros(24, np.float32)
    n = dev.sdk_engine_run(e, P(np.ascontiguousarray(dfm)), 0, P(pr), P(out), cap, P(met), 24, ctypes.c_float(1e-3))
    gm = np.zeros(12, np.float64)
    assert dev.sdk_engine_global_model(e, P(gm)) == 0
    rep = {}
    if n >= 0:
        pl = np.zeros(7 * gw * gh, np.float32)
        tag = np.zeros(gw * gh, np.uint8)
        path = np.zeros(gw * gh, np.uint8)
        assert dev.sdk_engine_node_planes(e, 0, P(pl), pl.size) == pl.size
        assert dev.sdk_engine_node_report(e, P(tag), P(path), tag.size) == tag.size
        rep = dict(planes=pl.reshape(7, gh, gw), tag=tag.reshape(gh, gw), path=path.reshape(gh, gw))
    dev.sdk_engine_destroy(e)
    return int(n), out[: max(n, 0) * 8].reshape(-1, 8), met, gm, rep


def anchor_errors(prm, rep, aff):
    ys, xs = np.nonzero(rep["path"] == 2)
    px = prm[0] + xs * prm[4]
    py = prm[1] + ys * prm[4]
    tu, tv, *_ = aff_truth(aff, px, py)
    return np.hypot(rep["planes"][0, ys, xs] - tu, rep["planes"][1, ys, xs] - tv)


def field_rms(pts, aff):
    tu, tv, *_ = aff_truth(aff, pts[:, 0], pts[:, 1])
    return float(np.sqrt(np.mean((pts[:, 2] - tu) ** 2 + (pts[:, 3] - tv) ** 2)))


GATES = {"R_rot_5deg": similarity(5.0), "R_rot_15deg": similarity(15.0),
         "RST_10deg_1.05_20px": similarity(10.0, 1.05, 20.0, -12.0), "R_rot_120deg": similarity(120.0),
         "R_rot_180deg": similarity(180.0)}


@pytest.mark.parametrize("name", list(GATES))
def test_rotation_gate(dev, name):
    """AnchorSeedingRotationGate (:128-177): locked, seeded with the full
    model, FULL quality, hull coverage >= 0.90, >= 95% of the grid inside the
    mesh or solved as anchors, median anchor error <= 0.05 px."""
    aff = GATES[name]
    ref, dfm = pair(aff)
    n, pts, met, gm, rep = run(dev, ref, dfm, PRM)
    err = anchor_errors(PRM, rep, aff)
    tag = rep["tag"]
    mesh_frac = float(((tag == 1) | (tag == 3)).mean())
    rot = math.degrees(math.atan2(gm[6], 1.0 + gm[3]))
    print(f"{name}: pass {int(gm[11])} locked {int(gm[0])} full {int(gm[1])} model rot {rot:+.3f} deg, "
          f"quality {met[16]:.0f} coverage {met[22]:.3f} mesh {mesh_frac:.3f} anchors {len(err)} vertex median "
          f"{np.median(err):.4f} px; field {n} points rms {field_rms(pts, aff) if n > 0 else -1:.4f} px",
          flush=True)
    assert met[21] == 1.0 and gm[0] == 1.0
    assert gm[1] == 1.0                       # the full model seeds the lattice
    assert met[16] == 2.0 and met[22] >= 0.90 and mesh_frac >= 0.95
    assert len(err) > 0 and np.median(err) <= 0.05


@pytest.mark.parametrize("deg", [85.0, 180.0])
def test_big_turn_solves_the_field(dev, deg):
    """GuessSmoothingKeepsEdgesInReach (85 deg) and HalfTurnSolvesTheField
    (180 deg): points returned, at most 5 Simplex rescues, field within
    0.05 px rms of the truth."""
    aff = similarity(deg)
    ref, dfm = pair(aff)
    n, pts, met, gm, _ = run(dev, ref, dfm, PRM)
    rms = field_rms(pts, aff) if n > 0 else -1.0
    print(f"{deg:.0f} deg: {n} points, {met[5]:.0f} Simplex rescues, {met[8]:.2f} mean iterations, rms {rms:.4f} px",
          flush=True)
    assert n > 0 and met[5] <= 5 and rms <= 0.05


def test_stretch_seeds_without_blunders(dev):
    """StretchSeedsWithoutBlunders (:179-247): a 45% stretch with Poisson 0.3
    about (192, 128) on a 384 x 256 frame, 110 px ROI, step 5, subset 21,
    window 15. FULL quality, >= 200 anchors kept, the 95th percentile of the
    anchor error <= 0.25 px, the field within 0.05 px rms. (The reference
    renders this case with its smooth speckle field; the blob field of the
    other contracts stands in for it here.)"""
    w, h = 384, 256
    aff = stretch_rotate_aff(0.45, 0.3, 0.0, 0.0, 0.0, 192.0, 128.0)
    ref, dfm = pair(aff, seed=101, w=w, h=h, blobs=1200)
    prm = ((w - 110) // 2, (h - 110) // 2, 110, 110, 5, 21, 15, 0)
    n, pts, met, gm, rep = run(dev, ref, dfm, prm)
    err = anchor_errors(prm, rep, aff)
    rms = field_rms(pts, aff) if n > 0 else -1.0
    print(f"stretch 45%: pass {int(gm[11])} locked {int(gm[0])} quality {met[16]:.0f} anchors kept {len(err)} "
          f"p95 {np.percentile(err, 95) if len(err) else -1:.4f} px; field {n} points rms {rms:.4f} px", flush=True)
    assert met[16] == 2.0 and len(err) >= 200
    assert np.percentile(err, 95) <= 0.25
    assert n > 0 and rms <= 0.05
