"""The subset solver's Keys 4x4 sampling forms against the oracle's
interpolate_bicubic (image_processor.cpp:188-242): the tap weights over the
FP32 and the u8 plane, and the row-cubic coefficient plane (four fp16 per
pixel, exact for u8 taps, evaluated by Horner)."""
import numpy as np
import pytest

from helpers import P, engine_module, oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    lib = engine_module().load_dev()
    assert lib.sdk_device_count() >= 1, "no GPU visible"
    return lib


def probe(dev, img, form, xs, ys):
    h, w = img.shape
    out = np.zeros(len(xs), np.float32)
    xs = np.ascontiguousarray(xs, np.float32)
    ys = np.ascontiguousarray(ys, np.float32)
    assert dev.sdk_interp_probe(P(np.ascontiguousarray(img, np.uint8)), w, h, form, len(xs), P(xs), P(ys), P(out)) == 0
    return out


def oracle_keys(img, xs, ys):
    h, w = img.shape
    out = np.zeros(len(xs), np.float32)
    f = np.ascontiguousarray(img, np.float32)
    oracle().orc_interp(P(f), w, h, 0, len(xs), P(np.ascontiguousarray(xs, np.float32)),
                        P(np.ascontiguousarray(ys, np.float32)), P(out))
    return out


def images():
    rng = np.random.default_rng(17)
    h, w = 200, 300
    yield "random", rng.integers(0, 256, size=(h, w), dtype=np.uint8)
    # the largest coefficients: 0 / 255 taps alternating along rows
    yield "alternating", np.tile(np.array([255, 0, 255, 0, 0, 255, 0, 0], np.uint8), (h, w // 8 + 1))[:, :w]
    yield "ramp", np.tile(np.arange(w, dtype=np.int64) % 256, (h, 1)).astype(np.uint8)


@pytest.mark.parametrize("name,img", list(images()), ids=[n for n, _ in images()])
def test_sampling_forms_against_oracle(dev, name, img):
    """Every form matches the oracle within a few float ulps of 255 (5e-4);
    the tap weights over the u8 and the FP32 plane agree bit for bit; at
    integer points every form returns the pixel exactly; on a linear ramp
    (x mod 256, away from its wrap) every form reproduces x (Keys is exact on
    linear data) within 5e-4."""
    h, w = img.shape
    rng = np.random.default_rng(3)
    n = 4000
    # (below w - 6 and h - 6, so that the cell-end fractions below stay
    # inside the guard after rounding to float)
    xs = rng.uniform(4.0, w - 6.0, n).astype(np.float32)
    ys = rng.uniform(4.0, h - 6.0, n).astype(np.float32)
    # fractions at the ends of the cell
    xs[:200] = np.floor(xs[:200]) + np.float32(0.99999)
    ys[200:400] = np.floor(ys[200:400]) + np.float32(1e-6)
    ref = oracle_keys(img, xs, ys)
    f0, f1, f2 = (probe(dev, img, k, xs, ys) for k in range(3))
    d0, d1 = np.abs(f0 - ref).max(), np.abs(f1 - ref).max()
    print(f"{name}: FP32 taps vs oracle {d0:.2e}, coefficient plane vs oracle {d1:.2e}, "
          f"coefficient plane vs FP32 taps {np.abs(f1 - f0).max():.2e}", flush=True)
    assert np.array_equal(f0, f2)
    assert d0 <= 5e-4 and d1 <= 5e-4
    ix = rng.integers(4, w - 5, 500).astype(np.float32)
    iy = rng.integers(4, h - 5, 500).astype(np.float32)
    px = img[iy.astype(int), ix.astype(int)].astype(np.float32)
    for k in range(3):
        assert np.array_equal(probe(dev, img, k, ix, iy), px), k
    if name == "ramp":
        ok = (xs < 250.0) & (xs >= 5.0)
        for f in (f0, f1):
            assert np.abs(f[ok] - xs[ok]).max() <= 5e-4


def test_probe_rejects_points_outside_the_guard(dev):
    img = np.zeros((64, 64), np.uint8)
    for x, y in [(3.9, 10.0), (10.0, 59.0), (59.5, 10.0), (10.0, 3.0)]:
        out = np.zeros(1, np.float32)
        xs, ys = np.array([x], np.float32), np.array([y], np.float32)
        assert dev.sdk_interp_probe(P(img), 64, 64, 1, 1, P(xs), P(ys), P(out)) == -3
