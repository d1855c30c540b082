"""The stop-split analysis that the GPU parity tests use, checked on the CPU
oracle against itself: the same frame solved with the reference's stop rule
(||dp|| < 1e-3) and with a rule 3% tighter differs exactly at the nodes whose
last step lands between the two thresholds, and each of those differs by one
IC-GN step of the oracle (optimization_engine.cpp:272-465: a fixed-Hessian
step depends on the current warp alone)."""
import numpy as np

from helpers import OracleEngine, gray8, oracle_vsg, ref_speckle, split_analysis


def frame_pair():
    W = H = 384
    ref = gray8(ref_speckle(4321, W, H, 3600), 2.0, 11)
    dfm = gray8(ref_speckle(4321, W, H, 3600, [1.3, -0.7, 0.004, -0.002, 0.003, 0.002, W / 2, H / 2]), 2.0, 12)
    return ref, dfm


def test_planted_stop_splits_are_one_step():
    ref, dfm = frame_pair()
    prm = (0, 0, 384, 384, 6, 21, 25, 0)
    oe = OracleEngine()
    assert oe.set_reference(ref) == 0
    n_a, pts_a, _, pl_a = oe.run_planes(dfm, prm, tol=1e-3)
    n_b, pts_b, _, pl_b = oe.run_planes(dfm, prm, tol=0.97e-3)
    assert n_a > 2000 and n_a == n_b
    res = split_analysis(ref, dfm, prm, pl_a, pl_b, 21, 1e-3)
    assert (res["valid_g"] == res["valid_o"]).all()
    # the tighter rule stops later at some nodes, never earlier
    assert len(res["nodes"]) >= 5, len(res["nodes"])
    assert all(res["split"]), [(n, f) for n, f in zip(res["nodes"], res["how"])]
    assert set(res["first"]) == {"gpu"}          # side a (the 1e-3 rule) stopped first
    # with the splits resolved the strains agree to rounding
    pl = pl_a.copy()
    for x, y in res["nodes"]:
        pl[:, y, x] = pl_b[:, y, x]
    ok, exx, eyy, exy = oracle_vsg(pl, 6, 25)
    okb, exxb, eyyb, exyb = oracle_vsg(pl_b, 6, 25)
    assert (ok == okb).all()
    d = max(np.abs(exx - exxb)[ok].max(), np.abs(eyy - eyyb)[ok].max(), np.abs(exy - exyb)[ok].max())
    assert d <= 1e-6, d


def test_unrelated_difference_is_not_a_split():
    """A planted difference that is not one IC-GN step (0.01 px added to u) is
    reported as a difference, neither a stop split nor a solution the
    oracle's stop rule accepts."""
    ref, dfm = frame_pair()
    prm = (0, 0, 384, 384, 6, 21, 25, 0)
    oe = OracleEngine()
    assert oe.set_reference(ref) == 0
    _, _, _, pl = oe.run_planes(dfm, prm, tol=1e-3)
    bad = pl.copy()
    ys, xs = np.nonzero(pl[6] >= 0)
    k = len(xs) // 2
    bad[0, ys[k], xs[k]] += 0.01
    res = split_analysis(ref, dfm, prm, bad, pl, 21, 1e-3)
    assert res["nodes"] == [(int(xs[k]), int(ys[k]))]
    assert res["split"] == [False] and res["how"] == ["-"]
