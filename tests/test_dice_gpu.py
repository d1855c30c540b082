"""The MI355X engine on real images: DICe's open-hole tension experiment
(oht_cfrp) as encoded TIFF bytes through the C ABI and the kernel-level node
solver, against DICe's solved fields with the reference's agreement bounds
(tests/dice/test_field_agreement.cpp, tests/dice/test_incremental_oht.cpp) and
against the oracle on the same images. Also the boundary's handling of
encoded masks of another size (full_field_solver.cpp:190-194)."""
import ctypes
import io

import numpy as np
import pytest

from helpers import (DICE_LADDER, P, OracleEngine, OracleIncremental, dice_agreement, dice_bytes, dice_gold,
                     dice_grid_params, dice_image, engine_module, oracle_solve_nodes, ref_speckle, gray8)

pytestmark = pytest.mark.gpu

RMS_TOL, MAX_TOL, MIN_FRACTION = 0.005, 0.02, 0.95          # test_field_agreement.cpp:54-56
LADDER_RMS_TOL, LADDER_MIN_FRACTION = 0.05, 0.90           # test_field_agreement.cpp:150-151
INC_RMS_TOL, INC_MIN_FRACTION = 0.1, 0.90                  # test_incremental_oht.cpp:58-59


@pytest.fixture(scope="module")
def sa():
    m = engine_module()
    if m.load_dev().sdk_device_count() < 1:
        pytest.skip("no GPU")
    return m


def solve_nodes(dev, ref, dfm, nodes, mode, use6=True):
    n = len(nodes)
    cxy = np.ascontiguousarray(nodes, np.int32).reshape(-1)
    g = np.zeros(6 * n, np.float32)
    md = np.full(n, mode, np.int32)
    iout = np.zeros((n, 5), np.int32)
    fout = np.zeros((n, 7), np.float32)
    r = np.ascontiguousarray(ref, np.float32)
    d = np.ascontiguousarray(dfm, np.uint8)
    assert dev.sdk_solve_nodes(P(r), P(d), 0, ref.shape[1], ref.shape[0], 27, int(use6), 0, ctypes.c_float(1e-3), n, P(cxy),
                               P(g), P(md), P(iout), P(fout)) == 0
    return iout, fout


@pytest.mark.parametrize("step", [0, 1, 2, 3])
def test_subsets_agree_with_dice_and_oracle(sa, step):
    """calculate_deformation at DICe's 230 points, subset 27, Keys 6x6, no LM,
    zero guess: step 01 with INIT_NO_SIMPLEX (OhtCfrp_AgreesWithDiceSolution),
    steps 03 / 06 / 11 with INIT_AUTO_SEARCH (OhtCfrp_LoadStepLadder). Images
    decoded by the engine's codec from the TIFF bytes. Per point against the
    oracle: the same status, u, v within the reference's cross-toolchain bound
    (5e-3 px)."""
    dev = sa.load_dev()
    img, sol = DICE_LADDER[step]
    ref = sa.decode(dice_bytes("oht_cfrp_00.tiff"), lib=dev)
    dfm = sa.decode(dice_bytes(img), lib=dev)
    assert np.array_equal(ref, dice_image("oht_cfrp_00.tiff")) and np.array_equal(dfm, dice_image(img))
    gold = dice_gold(sol)
    nodes = gold[:, :2].astype(np.int32)
    mode = 2 if step == 0 else 0
    io_g, fo_g = solve_nodes(dev, ref, dfm, nodes, mode)
    io_o, fo_o = oracle_solve_nodes(ref.astype(np.float32), dfm.astype(np.float32), 27, nodes,
                                    [[0] * 6] * len(nodes), [mode] * len(nodes), use6=True, lm=False)
    assert np.array_equal(io_g[:, :2], io_o[:, :2])
    ok = (io_g[:, 0] == 1) & (io_g[:, 1] == 0)
    d_or = np.abs(fo_g[ok, :2] - fo_o[ok, :2]).max()
    du, dv = fo_g[ok, 0] - gold[ok, 2], fo_g[ok, 1] - gold[ok, 3]
    rms_u, rms_v = np.sqrt((du ** 2).mean()), np.sqrt((dv ** 2).mean())
    print(f"{img}: {ok.mean():.3f} compared, rms {rms_u:.5f} / {rms_v:.5f}, max {np.abs(du).max():.4f} / "
          f"{np.abs(dv).max():.4f}; vs oracle {d_or:.2e}", flush=True)
    assert d_or <= 5e-3
    if step == 0:
        assert ok.mean() >= MIN_FRACTION and rms_u < RMS_TOL and rms_v < RMS_TOL
        assert np.abs(du).max() < MAX_TOL and np.abs(dv).max() < MAX_TOL
    else:
        assert ok.mean() >= LADDER_MIN_FRACTION and rms_u < LADDER_RMS_TOL and rms_v < LADDER_RMS_TOL


def test_translation_real_image(sa):
    """DiceTranslationReal.CustomApp_0p4px_RealSpeckle
    (tests/dice/test_translation_real_image.cpp) on the TIFF bytes: four
    subsets, subset 27, bicubic, INIT_NO_SIMPLEX, each converged with
    |u - 0.4| <= 0.1 px; per subset within 5e-3 px of the oracle."""
    dev = sa.load_dev()
    ref, dfm = sa.decode(dice_bytes("ref.tif"), lib=dev), sa.decode(dice_bytes("def.tif"), lib=dev)
    nodes = np.array([(100, 100), (200, 200), (300, 300), (400, 400)], np.int32)
    io_g, fo_g = solve_nodes(dev, ref, dfm, nodes, 2, use6=False)
    io_o, fo_o = oracle_solve_nodes(ref.astype(np.float32), dfm.astype(np.float32), 27, nodes, [[0] * 6] * 4, [2] * 4)
    print("u", fo_g[:, 0], "v", fo_g[:, 1], flush=True)
    assert (io_g[:, 0] == 1).all() and (io_g[:, 1] == 0).all() and np.array_equal(io_g[:, :2], io_o[:, :2])
    assert (np.abs(fo_g[:, 0] - 0.4) <= 0.1).all()
    assert np.abs(fo_g[:, :2] - fo_o[:, :2]).max() <= 5e-3


def full_field_common(pts_g, pts_o, n_o):
    go = {(a, b): r for a, b, *r in pts_o.tolist()}
    gg = {(a, b): r for a, b, *r in pts_g.tolist()}
    common = sorted(set(go) & set(gg))
    assert len(common) >= n_o - max(2, n_o // 100)
    return np.array([np.abs(np.array(go[k]) - np.array(gg[k])) for k in common])


def test_full_field_from_tiff_bytes(sa):
    """semper_set_reference / semper_run on the TIFF files' bytes (expected
    size 0: the encoded path) on DICe's step-7 grid: the same output, bit for
    bit, as on the decoded planes; against the oracle's run_full_field point
    counts within 1%, u, v within 5e-3 px, strain within 1e-4; against DICe
    within 0.0035 px rms at every one of its points on every step (the
    reference: 0.003 px, test_incremental_oht.cpp:48-57)."""
    prm = dice_grid_params()
    p = sa.params(prm[:4], prm[4], prm[5], prm[6], True)
    eng, eng_raw = sa.Engine(), sa.Engine()
    assert eng.set_reference(dice_bytes("oht_cfrp_00.tiff")) == 0
    ref = dice_image("oht_cfrp_00.tiff")
    assert eng_raw.set_reference(ref) == 0
    oe = OracleEngine()
    assert oe.set_reference(ref) == 0
    for img, sol in DICE_LADDER:
        n, pts, met = eng.run(dice_bytes(img), p, capacity=(prm[2] // 7 + 1) * (prm[3] // 7 + 1) * 8)
        dfm = dice_image(img)
        n_r, pts_r, _ = eng_raw.run(dfm, p)
        n_o, pts_o, _ = oe.run(dfm, prm)
        a = dice_agreement(pts, dice_gold(sol))
        print(f"{img}: gpu {n} pts, oracle {n_o}; DICe {a}", flush=True)
        assert n == n_r > 0 and np.array_equal(pts, pts_r)
        assert abs(n - n_o) <= max(2, n_o // 100)
        d = full_field_common(pts, pts_o, n_o)
        assert d[:, :2].max() <= 5e-3 and d[:, 2:5].max() <= 1e-4
        assert a[0] >= INC_MIN_FRACTION and max(a[1], a[2]) < 0.0035


def test_incremental_ladder_from_tiff_bytes(sa):
    """OhtCfrp_LadderTotalsAgreeWithDice through semper_incremental_* on the
    TIFF bytes, update_interval 1: totals within 0.1 px rms of DICe at >= 90%
    of its points, and frame by frame against the oracle's session: the same
    reference schedule, counts within 1%, u, v within 5e-3 px."""
    prm = dice_grid_params()
    p = sa.params(prm[:4], prm[4], prm[5], prm[6], True)
    s = sa.Incremental()
    assert s.init(dice_bytes("oht_cfrp_00.tiff"), p, 1, expected_w=0, expected_h=0) == 0
    s.cap = (prm[2] // 7 + 1) * (prm[3] // 7 + 1) * 8
    o = OracleIncremental()
    assert o.init(dice_image("oht_cfrp_00.tiff"), prm, 1) == 0
    for img, sol in DICE_LADDER:
        n, pts, _, info = s.process(dice_bytes(img))
        n_o, pts_o, _, info_o = o.process(dice_image(img))
        a = dice_agreement(pts, dice_gold(sol))
        ig = [info.reference_frame, info.reference_updated, info.tracked_points, info.frame0_points,
              info.taylor_fallbacks, info.lost_this_frame]
        print(f"{img}: gpu {n} pts {ig}, oracle {n_o} pts {info_o.tolist()}; DICe {a}", flush=True)
        assert ig[:2] == info_o.tolist()[:2] and ig[3] == info_o[3]
        assert abs(n - n_o) <= max(2, n_o // 100)
        d = full_field_common(pts, pts_o, n_o)
        assert d[:, :2].max() <= 5e-3
        assert a[0] >= INC_MIN_FRACTION and a[1] < INC_RMS_TOL and a[2] < INC_RMS_TOL


def test_encoded_mask_of_another_size(sa):
    """A mask given as an encoded PNG at half the image's size is brought to
    the image's size by nearest neighbour: the same field, bit for bit, as the
    raw mask resized on the host by the same rule (source pixel
    min(floor(x * sw / dw), sw - 1))."""
    from PIL import Image

    w = h = 256
    ref = gray8(ref_speckle(5, w, h, 900), 2.0, 1)
    dfm = gray8(ref_speckle(5, w, h, 900, (0.7, -0.4, 0.01, 0.0, 0.0, -0.005, 128.0, 128.0)), 2.0, 2)
    small = np.zeros((h // 2, w // 2), np.uint8)
    yy, xx = np.mgrid[: h // 2, : w // 2]
    small[(xx - 64) ** 2 + (yy - 64) ** 2 < 50 ** 2] = 255
    small[(xx - 64) ** 2 + (yy - 64) ** 2 < 8 ** 2] = 0
    bio = io.BytesIO()
    Image.fromarray(small).save(bio, "PNG")
    big = small[np.minimum(np.arange(h) // 2, h // 2 - 1)][:, np.minimum(np.arange(w) // 2, w // 2 - 1)]
    p = sa.params((16, 16, 224, 224), 8, 21, 17)
    e1, e2 = sa.Engine(), sa.Engine()
    assert e1.set_reference(ref, mask=bio.getvalue()) == 0
    assert e2.set_reference(ref, mask=big) == 0
    n1, p1, _ = e1.run(dfm, p, mask=bio.getvalue())
    n2, p2, _ = e2.run(dfm, p, mask=big)
    print(f"masked points {n1} / {n2}", flush=True)
    assert n1 == n2 > 0 and np.array_equal(p1, p2)
    # every point sits on material of the resized mask
    assert (big[p1[:, 1].astype(int), p1[:, 0].astype(int)] == 255).all()
