"""The incremental session's FP64 composition on the device (no reference
counterpart): the kernel that samples the increment field at every followed
node and adds the sample to the node's total in double. Checked bit for bit
against a float32 restatement of the reference's sample_increment
(incremental_math.cpp:61-123; the session's host form, incremental.cpp) on a
field with invalid nodes, positions on exact node hits, inside cells with
four valid corners, next to invalid corners (the first-order expansion within
1.5 cells) and beyond every node (not sampled)."""
import numpy as np
import pytest

from helpers import P, engine_module

pytestmark = pytest.mark.gpu
f32 = np.float32


@pytest.fixture(scope="module")
def dev():
    lib = engine_module().load_dev()
    assert lib.sdk_device_count() >= 1, "no GPU visible"
    return lib


def sample_increment(pl, med, fgw, fgh, frx, fry, fstep, x, y):
    """(ok, taylor, u, v, corr) in float32 arithmetic, the reference's order."""
    u, v, ux, uy, vx, vy, c = (pl[k] for k in range(7))
    gx = (x - f32(frx)) / f32(fstep)
    gy = (y - f32(fry)) / f32(fstep)
    i0, j0 = int(np.floor(gx)), int(np.floor(gy))
    fx, fy = gx - f32(i0), gy - f32(j0)
    inn = lambda i, j: 0 <= i < fgw and 0 <= j < fgh
    ok = lambda i, j: inn(i, j) and med[j, i] != 0
    if fx == 0 and fy == 0 and ok(i0, j0):
        return True, False, u[j0, i0], v[j0, i0], c[j0, i0]
    if inn(i0, j0) and inn(i0 + 1, j0 + 1) and ok(i0, j0) and ok(i0 + 1, j0) and ok(i0, j0 + 1) and ok(i0 + 1, j0 + 1):
        one = f32(1)
        wa, wb, wc, wd = (one - fx) * (one - fy), fx * (one - fy), (one - fx) * fy, fx * fy
        su = wa * u[j0, i0] + wb * u[j0, i0 + 1] + wc * u[j0 + 1, i0] + wd * u[j0 + 1, i0 + 1]
        sv = wa * v[j0, i0] + wb * v[j0, i0 + 1] + wc * v[j0 + 1, i0] + wd * v[j0 + 1, i0 + 1]
        m = lambda a, b: b if a < b else a
        return True, False, su, sv, m(m(c[j0, i0], c[j0, i0 + 1]), m(c[j0 + 1, i0], c[j0 + 1, i0 + 1]))
    max_d = f32(1.5) * f32(fstep)
    best, bi, bj = f32(np.inf), -1, -1
    for j in range(j0 - 1, j0 + 3):
        for i in range(i0 - 1, i0 + 3):
            if not ok(i, j):
                continue
            dx, dy = x - f32(frx + i * fstep), y - f32(fry + j * fstep)
            d2 = dx * dx + dy * dy
            if d2 < best:
                best, bi, bj = d2, i, j
    if bi < 0 or best > max_d * max_d:
        return False, False, f32(0), f32(0), f32(0)
    dx, dy = x - f32(frx + bi * fstep), y - f32(fry + bj * fstep)
    return (True, True, u[bj, bi] + ux[bj, bi] * dx + uy[bj, bi] * dy, v[bj, bi] + vx[bj, bi] * dx + vy[bj, bi] * dy,
            c[bj, bi])


def test_device_composition_is_the_reference_sampling_in_double(dev):
    rng = np.random.default_rng(11)
    fgw, fgh, frx, fry, fstep = 24, 18, 13, 9, 8
    pl = np.empty((7, fgh, fgw), np.float32)
    pl[0:2] = rng.normal(0.0, 3.0, (2, fgh, fgw))
    pl[2:6] = rng.normal(0.0, 0.02, (4, fgh, fgw))
    pl[6] = rng.uniform(0.0, 0.15, (fgh, fgw))
    med = (rng.random((fgh, fgw)) < 0.8).astype(np.uint8)
    gw, gh, rx, ry, step = 31, 23, 4, 2, 7
    n0 = gw * gh
    uf = rng.uniform(-6.0, 9.0, n0).astype(np.float32)
    vf = rng.uniform(-6.0, 9.0, n0).astype(np.float32)
    # exact node hits: totals that land a node on an increment node
    gx, gy = np.arange(n0) % gw, np.arange(n0) // gw
    hit = rng.random(n0) < 0.15
    ti = rng.integers(0, fgw, n0)
    tj = rng.integers(0, fgh, n0)
    uf[hit] = (frx + ti[hit] * fstep - (rx + gx[hit] * step)).astype(np.float32)
    vf[hit] = (fry + tj[hit] * fstep - (ry + gy[hit] * step)).astype(np.float32)
    # a few far outside the field
    far = rng.random(n0) < 0.05
    uf[far] = np.float32(400.0)
    # totals carried in double: off their float roundings, except the hits
    u64 = uf.astype(np.float64) + np.where(hit, 0.0, rng.normal(0.0, 1e-6, n0))
    v64 = vf.astype(np.float64) + np.where(hit, 0.0, rng.normal(0.0, 1e-6, n0))
    uf = u64.astype(np.float32)
    vf = v64.astype(np.float32)
    alive = (rng.random(n0) < 0.9).astype(np.uint8)
    flags = np.zeros(n0, np.uint8)
    ou64, ov64 = np.zeros(n0), np.zeros(n0)
    ouf, ovf, oc = (np.zeros(n0, np.float32) for _ in range(3))
    plc = np.ascontiguousarray(pl)
    assert dev.sdk_compose64_probe(P(plc), P(med), fgw, fgh, frx, fry, fstep, P(uf), P(vf), P(u64), P(v64), P(alive),
                                   gw, gh, rx, ry, step, P(flags), P(ou64), P(ov64), P(ouf), P(ovf), P(oc)) == 0
    kinds = {"exact": 0, "bilinear": 0, "taylor": 0, "none": 0}
    for k in range(n0):
        if not alive[k]:
            assert flags[k] == 0
            continue
        x = f32(rx + gx[k] * step) + uf[k]
        y = f32(ry + gy[k] * step) + vf[k]
        ok, taylor, su, sv, sc = sample_increment(pl, med, fgw, fgh, frx, fry, fstep, x, y)
        assert flags[k] == ((1 if ok else 0) | (2 if taylor else 0)), k
        if not ok:
            kinds["none"] += 1
            continue
        kinds["taylor" if taylor else ("exact" if hit[k] and med[tj[k], ti[k]] else "bilinear")] += 1
        U = u64[k] + np.float64(su)
        V = v64[k] + np.float64(sv)
        assert ou64[k] == U and ov64[k] == V, k
        assert ouf[k] == np.float32(U) and ovf[k] == np.float32(V) and oc[k] == sc, k
    print(kinds, flush=True)
    assert min(kinds.values()) >= 10


def test_probe_rejects_bad_grids(dev):
    z = np.zeros(64, np.float32)
    b = np.zeros(64, np.uint8)
    d = np.zeros(64)
    assert dev.sdk_compose64_probe(P(z), P(b), 0, 4, 0, 0, 8, P(z), P(z), P(d), P(d), P(b), 2, 2, 0, 0, 8, P(b), P(d),
                                   P(d), P(z), P(z), P(z)) == -3
