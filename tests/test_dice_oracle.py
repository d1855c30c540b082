"""The oracle on real images: DICe's open-hole tension experiment (oht_cfrp,
DICe's dic_challenge_12 case) against DICe's own solved fields, with the
reference's agreement bounds (tests/dice/test_field_agreement.cpp,
tests/dice/test_incremental_oht.cpp). The images are decoded by PIL here, so
these tests do not depend on the engine's codec. Beyond the bounds, each test
checks the oracle against the spread the reference measured for itself and
wrote next to its bounds (test_field_agreement.cpp:47-56,
test_incremental_oht.cpp:48-57)."""
import numpy as np

from helpers import (DICE_LADDER, OracleEngine, OracleIncremental, dice_agreement, dice_gold, dice_grid_params,
                     dice_image, oracle_solve_nodes)

RMS_TOL, MAX_TOL, MIN_FRACTION = 0.005, 0.02, 0.95          # test_field_agreement.cpp:54-56
LADDER_RMS_TOL, LADDER_MIN_FRACTION = 0.05, 0.90           # test_field_agreement.cpp:150-151
INC_RMS_TOL, INC_MIN_FRACTION = 0.1, 0.90                  # test_incremental_oht.cpp:58-59


def subset_agreement(ref, dfm, gold, mode):
    """calculate_deformation at DICe's points: subset 27, Keys 6x6, a bare
    engine (no LM), zero guess (test_field_agreement.cpp:101-113)."""
    nodes = gold[:, :2].astype(np.int32)
    io, fo = oracle_solve_nodes(ref.astype(np.float32), dfm.astype(np.float32), 27, nodes,
                                [[0] * 6] * len(nodes), [mode] * len(nodes), use6=True, lm=False)
    ok = (io[:, 0] == 1) & (io[:, 1] == 0)
    du, dv = fo[ok, 0] - gold[ok, 2], fo[ok, 1] - gold[ok, 3]
    return ok.mean(), np.sqrt((du ** 2).mean()), np.sqrt((dv ** 2).mean()), np.abs(du).max(), np.abs(dv).max()


def test_step01_agrees_with_dice():
    """OhtCfrp_AgreesWithDiceSolution: INIT_NO_SIMPLEX from a zero guess. The
    reference measured rms 0.0006 px, max 0.0033 px, 230/230 points."""
    ref = dice_image("oht_cfrp_00.tiff")
    gold = dice_gold("DICe_solution_01.txt")
    assert len(gold) == 230
    frac, rms_u, rms_v, max_u, max_v = subset_agreement(ref, dice_image("oht_cfrp_01.tiff"), gold, 2)
    print(f"step 01: {frac:.3f} compared, rms {rms_u:.5f} / {rms_v:.5f}, max {max_u:.5f} / {max_v:.5f}")
    assert frac >= MIN_FRACTION and rms_u < RMS_TOL and rms_v < RMS_TOL and max_u < MAX_TOL and max_v < MAX_TOL
    assert frac == 1.0 and max(rms_u, rms_v) < 0.0008 and max(max_u, max_v) < 0.0040


def test_load_step_ladder_by_coarse_search():
    """OhtCfrp_LoadStepLadder: steps 03 / 06 / 11 (up to 12 px) with
    INIT_AUTO_SEARCH, so the coarse search has to find the shift."""
    ref = dice_image("oht_cfrp_00.tiff")
    for img, sol in DICE_LADDER[1:]:
        gold = dice_gold(sol)
        frac, rms_u, rms_v, max_u, max_v = subset_agreement(ref, dice_image(img), gold, 0)
        print(f"{img}: {frac:.3f} compared, rms {rms_u:.5f} / {rms_v:.5f}, max {max_u:.4f} / {max_v:.4f}")
        assert frac >= LADDER_MIN_FRACTION and rms_u < LADDER_RMS_TOL and rms_v < LADDER_RMS_TOL


def test_incremental_ladder_totals_agree_with_dice():
    """OhtCfrp_LadderTotalsAgreeWithDice: 00 -> 01 -> 03 -> 06 -> 11 through the
    session with the reference moved every frame, totals at DICe's points. The
    reference measured rms du 0.058 / dv 0.026 px at step 11, mostly a uniform
    u bias (mean du -0.054 px), and the fixed reference on the same grid within
    0.003 px rms."""
    ref = dice_image("oht_cfrp_00.tiff")
    prm = dice_grid_params()
    inc = OracleIncremental()
    assert inc.init(ref, prm, 1) == 0
    fixed = OracleEngine()
    assert fixed.set_reference(ref) == 0
    for img, sol in DICE_LADDER:
        gold = dice_gold(sol)
        dfm = dice_image(img)
        n, pts, _, info = inc.process(dfm)
        n_f, pts_f, _ = fixed.run(dfm, prm)
        a, af = dice_agreement(pts, gold), dice_agreement(pts_f, gold)
        print(f"{img}: incremental {n} pts {info.tolist()} {a}; fixed {n_f} pts {af}")
        assert n > 0 and n_f > 0
        assert a[0] >= INC_MIN_FRACTION and a[1] < INC_RMS_TOL and a[2] < INC_RMS_TOL
        assert af[0] >= INC_MIN_FRACTION and max(af[1], af[2]) < 0.0035
    # step 11: the spread the reference measured
    assert 0.045 < a[1] < 0.07 and 0.018 < a[2] < 0.034 and -0.065 < a[3] < -0.045


def test_translation_real_image():
    """DiceTranslationReal.CustomApp_0p4px_RealSpeckle
    (tests/dice/test_translation_real_image.cpp): DICe's ref.tif / def.tif, the
    four subsets of its custom_app, subset 27, bicubic, INIT_NO_SIMPLEX: every
    subset converges with |u - 0.4| <= 0.1 px (DICe's published contract)."""
    ref, dfm = dice_image("ref.tif"), dice_image("def.tif")
    assert ref.shape == dfm.shape == (512, 512)
    nodes = [(100, 100), (200, 200), (300, 300), (400, 400)]
    io, fo = oracle_solve_nodes(ref.astype(np.float32), dfm.astype(np.float32), 27, nodes, [[0] * 6] * 4, [2] * 4)
    print("u", fo[:, 0], "v", fo[:, 1])
    assert (io[:, 0] == 1).all() and (io[:, 1] == 0).all()
    assert (np.abs(fo[:, 0] - 0.4) <= 0.1).all()
