"""The oracle's IncrementalSession against the reference's own session
contracts (tests/integration/test_incremental.cpp) on sequences with exact
ground truth, on the CPU: interval 0 is run_full_field bit for bit, the first
frame of any interval is too, the reference moves on the documented schedule,
lost nodes never return, and a stretching, turning ROI is tracked to the
truth."""
import numpy as np
import pytest

from helpers import OracleEngine, OracleIncremental, aff_truth, centred_params, speckle_sequence

W = H = 256


@pytest.fixture(scope="module")
def seq():
    return speckle_sequence(8, W, H, e_max=0.16, theta_max=8.0)


def run_session(ref, frames, prm, interval, mask=None):
    s = OracleIncremental()
    assert s.init(ref, prm, interval, mask) == 0
    return [s.process(f) for f in frames]


def test_interval_zero_is_run_full_field(seq):
    """Incremental.IntervalZero_BitIdenticalToRunFullField (test_incremental.cpp:202-217)."""
    ref, frames, _ = seq
    prm = centred_params(W, H, 128)
    inc = run_session(ref, frames[:3], prm, 0)
    oe = OracleEngine()
    assert oe.set_reference(ref) == 0
    for (n, pts, _, info), f in zip(inc, frames[:3]):
        n_f, pts_f, _ = oe.run(f, prm)
        assert n == n_f > 0 and np.array_equal(pts, pts_f)
        assert info[0] == 0 and info[1] == 0


def test_first_frame_is_run_full_field(seq):
    """Incremental.FirstFrame_BitIdenticalToRunFullField (:219-228)."""
    ref, frames, _ = seq
    prm = centred_params(W, H, 128)
    oe = OracleEngine()
    assert oe.set_reference(ref) == 0
    n_f, pts_f, _ = oe.run(frames[0], prm)
    for interval in (1, 3):
        n, pts, _, _ = run_session(ref, frames[:1], prm, interval)[0]
        assert n == n_f and np.array_equal(pts, pts_f)


def test_update_schedule(seq):
    """Incremental.UpdateSchedule_FollowsInterval (:230-246)."""
    ref, frames, _ = seq
    prm = centred_params(W, H, 128)
    n1 = run_session(ref, frames[:6], prm, 1)
    n3 = run_session(ref, frames[:6], prm, 3)
    for t in range(6):
        assert n1[t][3][0] == t and n1[t][3][1] == 1 and n1[t][0] > 0
        assert n3[t][3][0] == (t // 3) * 3 and n3[t][3][1] == int((t + 1) % 3 == 0) and n3[t][0] > 0


def accuracy(pts, aff):
    u, v, exx, eyy, exy = aff_truth(aff, pts[:, 0], pts[:, 1])
    err = np.hypot(pts[:, 2] - u, pts[:, 3] - v)
    es = np.sqrt(((pts[:, 4] - exx) ** 2 + (pts[:, 5] - eyy) ** 2 + (pts[:, 6] - exy) ** 2) / 3)
    return np.sqrt(np.mean(err ** 2)), err.max(), np.sqrt(np.mean(es ** 2))


@pytest.mark.parametrize("interval", [1, 3])
def test_tracks_stretch_and_rotation(seq, interval):
    """Incremental.StretchRotation_TracksTheStretchingRoi (:248-303), scaled down:
    16% stretch and 8 degrees over 8 frames; every frame reports points near
    the truth and the tracked set keeps >= 98% of the frame-0 nodes."""
    ref, frames, affs = seq
    prm = centred_params(W, H, 128)
    inc = run_session(ref, frames, prm, interval)
    full = inc[0][0]
    assert full > 0
    for (n, pts, _, info), aff in zip(inc, affs):
        rms, mx, rms_e = accuracy(pts, aff)
        assert n >= 0.95 * full, (n, full)
        assert rms <= 0.1 and rms_e <= 5e-3, (rms, mx, rms_e)
        assert info[2] >= 0.98 * info[3]


def test_lost_nodes_never_return():
    """Incremental.LostNodesNeverReturn (:377-391): a drift that carries part of
    the ROI past the solver's border margin loses nodes for good under N = 1."""
    ref, frames, _ = speckle_sequence(8, W, H, e_max=0.02, theta_max=1.0, tx_max=100.0)
    inc = run_session(ref, frames, centred_params(W, H, 160), 1)
    for t in range(1, len(inc)):
        assert inc[t][3][2] <= inc[t - 1][3][2] and inc[t][3][5] >= 0
    assert inc[-1][3][2] < inc[0][3][3]
