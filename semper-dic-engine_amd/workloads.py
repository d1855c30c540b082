"""The benchmark workloads: synthetic speckle pairs and solve parameters.

Speckle (all workloads): seeded Gaussian blobs, sigma U[2, 4] px, 0.5 blobs per
px^2 over the image padded by 16 px, signs +-1, mapped linearly with +-3
standard deviations of the blob field onto [0, 200] (peak 200), 2% (4 grey
levels) Gaussian noise, rounded to u8. The deformed image samples the continuous field at the analytic inverse
of the displacement field (fixed-point iteration in FP64), never by
resampling the reference, so the ground truth is exact -- except the 512^2
translation, whose deformed frame is the Keys bicubic resampling of the u8
reference its configuration names. The pairs are rendered on the GPU by
sdk_synth_pair (csrc/synth.hip).
"""
import ctypes
import math

import numpy as np

IDENTITY, AFFINE, SIN_GAUSS, BUMPS, UNIAXIAL = 0, 1, 2, 3, 4
NOISE_SIGMA = 4.0


def affine_params(rot_deg, exx, eyy, exy, tu, tv):
    """Displacement t + (F - I)(p - c) with F = R(rot) [[1+exx, exy], [exy, 1+eyy]]."""
    c, s = math.cos(math.radians(rot_deg)), math.sin(math.radians(rot_deg))
    a11, a12, a21, a22 = 1 + exx, exy, exy, 1 + eyy
    f00, f01 = c * a11 - s * a21, c * a12 - s * a22
    f10, f11 = s * a11 + c * a21, s * a12 + c * a22
    return [tu, tv, f00 - 1.0, f01, f10, f11 - 1.0]


def bumps_params(seed, w, h, n=8):
    """Sum of n Gaussian bumps (amplitude U[-30, 30] px per component, width
    U[200, 800] px) plus a global rotation U[-5, 5] deg about the centre."""
    rng = np.random.default_rng(seed)
    p = [math.radians(rng.uniform(-5.0, 5.0)), float(n)]
    for _ in range(n):
        p += [rng.uniform(0, w), rng.uniform(0, h), rng.uniform(200, 800), rng.uniform(-30, 30),
              rng.uniform(-30, 30)]
    return p


CONFIGS = {
    # 512^2, u = 0.4 px rigid translation applied by Keys bicubic resampling of
    # the u8 reference, step 10, subset 21, window 21
    "c1_512_translate": dict(w=512, h=512, seed=1001, kind=AFFINE, p=[0.4, 0.0, 0, 0, 0, 0], step=10, subset=21,
                             window=21, tol=1e-3, resample=True),
    # 2048^2 homogeneous affine: rot 2 deg, exx 1%, eyy -0.3%, exy 0.5%, t (12.3, -7.8)
    "c2_2048_affine": dict(w=2048, h=2048, seed=2002, kind=AFFINE, p=affine_params(2.0, 0.01, -0.003, 0.005, 12.3, -7.8),
                           step=5, subset=31, window=31, tol=1e-4),
    # 4096^2 heterogeneous: u = 3 sin(2 pi x / 1024) cos(2 pi y / 1536), v = 2 exp(-r^2 / (2 600^2))
    "c4_4096_heterogeneous": dict(w=4096, h=4096, seed=4004, kind=SIN_GAUSS, p=[3.0, 1024.0, 1536.0, 2.0, 600.0],
                                  step=8, subset=31, window=31, tol=1e-3),
    # 1024^2 heterogeneous field at a low signal-to-noise ratio (noise sigma 9,
    # subset 21): a few percent of the subsets miss the ZNSSD acceptance, so the
    # Path B flood has real work
    "t_1024_noisy": dict(w=1024, h=1024, seed=777, kind=SIN_GAUSS, p=[3.0, 512.0, 384.0, 2.0, 300.0], step=8,
                         subset=21, window=31, tol=1e-3, noise=9.0),
    # 1024^2 frame of the batched / incremental sequence at 10% stretch
    # (exx 0.1, Poisson 0.3), step 8, window 21, subset sizes 15 / 21 / 31 /
    # 41 assigned to a 4 x 4 region map drawn from region_seed
    "c3_1024_mixed": dict(w=1024, h=1024, seed=3003, kind=UNIAXIAL, p=[0.1, 0.3], step=8, subset=21, window=21,
                          tol=1e-3, sizes=(15, 21, 31, 41), region_seed=30, frames=64, ramp=0.2),
    # 4096^2 bumps + rotation (the learned-seeding workload)
    "c5_4096_bumps": dict(w=4096, h=4096, seed=5005, kind=BUMPS, p=None, step=8, subset=31, window=31, tol=1e-3),
}


def config(name):
    analytic = name.endswith("@analytic")
    c = dict(CONFIGS[name[: -len("@analytic")] if analytic else name])
    if analytic:
        c.pop("resample", None)
    if c["p"] is None:
        c["p"] = bumps_params(c["seed"], c["w"], c["h"])
    return c


def keys_weights(s):
    """Keys (a = -0.5) cubic weights of the four taps at -1, 0, 1, 2 for a
    fraction s in [0, 1)."""
    s2, s3 = s * s, s * s * s
    return np.array([-0.5 * s3 + s2 - 0.5 * s, 1.5 * s3 - 2.5 * s2 + 1.0, -1.5 * s3 + 2.0 * s2 + 0.5 * s,
                     0.5 * s3 - 0.5 * s2])


def resample_shift_x(ref, u):
    """The deformed frame of a rigid shift by u px along x made by Keys
    bicubic resampling of the u8 reference: D(x, y) = R(x - u, y) with the
    four taps of each row (edge pixels repeated), computed in double, rounded
    and clamped to u8. A pixel of the deformed frame at x then shows the
    reference point x - u, which is the displacement +u."""
    h, w = ref.shape
    r = ref.astype(np.float64)
    x = np.arange(w) - u
    x0 = np.floor(x).astype(int)
    wt = keys_weights(x - x0)
    out = np.zeros((h, w))
    for k in range(4):
        out += r[:, np.clip(x0 - 1 + k, 0, w - 1)] * wt[k][None, :]
    return np.clip(np.rint(out), 0, 255).astype(np.uint8)


def pair(dev, name, on_device=False):
    """(ref, def) of a workload: host uint8 arrays, or device pointers
    (allocated with sdk_malloc, to be freed by the caller) when on_device."""
    c = config(name)
    if c.get("resample"):
        ra, _ = pair(dev, name + "@analytic")
        da = resample_shift_x(ra, c["p"][0])
        if not on_device:
            return ra, da
        r, d = dev.sdk_malloc(ra.size), dev.sdk_malloc(da.size)
        dev.sdk_memcpy(r, ra.ctypes.data_as(ctypes.c_void_p), ra.size, 0)
        dev.sdk_memcpy(d, da.ctypes.data_as(ctypes.c_void_p), da.size, 0)
        return r, d
    w, h = c["w"], c["h"]
    prm = np.array(c["p"], np.float64)
    if on_device:
        r = dev.sdk_malloc(w * h)
        d = dev.sdk_malloc(w * h)
    else:
        ra = np.zeros((h, w), np.uint8)
        da = np.zeros((h, w), np.uint8)
        r = ra.ctypes.data_as(ctypes.c_void_p)
        d = da.ctypes.data_as(ctypes.c_void_p)
    rc = dev.sdk_synth_pair(c["seed"], w, h, c["kind"], prm.ctypes.data_as(ctypes.c_void_p), len(prm), w / 2.0,
                            h / 2.0, ctypes.c_float(c.get("noise", NOISE_SIGMA)), r, d, 1 if on_device else 0)
    if rc != 0:
        raise RuntimeError(f"sdk_synth_pair failed ({rc})")
    return (r, d) if on_device else (ra, da)


def region_map(name):
    """The 4 x 4 subset-size map of a mixed-size workload over the frame
    (row-major int32), drawn (seeded) from the workload's sizes."""
    c = config(name)
    rng = np.random.default_rng(c["region_seed"])
    return np.ascontiguousarray(rng.choice(np.array(c["sizes"], np.int32), size=(4, 4)), np.int32)


def sequence_strain(name, k):
    """exx of frame k of a sequence workload: a linear ramp from 0 at frame 0
    to `ramp` at the last frame."""
    c = config(name)
    return c["ramp"] * k / (c["frames"] - 1)


def sequence_frame(dev, name, k):
    """Frame k of a sequence workload (host uint8, h x w): the workload's
    speckle under uniaxial stretch exx_k with its Poisson ratio, about the
    frame centre, with noise of its own (frames never share a noise pattern,
    which would pull each increment toward zero); frame 0 is unstretched."""
    c = config(name)
    w, h = c["w"], c["h"]
    prm = np.array([sequence_strain(name, k), c["p"][1]], np.float64)
    out = np.zeros((h, w), np.uint8)
    rc = dev.sdk_synth_frame(c["seed"], c["seed"] * 1000 + 7 * k + 1, w, h, c["kind"],
                             prm.ctypes.data_as(ctypes.c_void_p), len(prm), w / 2.0, h / 2.0,
                             ctypes.c_float(c.get("noise", NOISE_SIGMA)), out.ctypes.data_as(ctypes.c_void_p), 0)
    if rc != 0:
        raise RuntimeError(f"sdk_synth_frame failed ({rc})")
    return out


def subset_map(name, gw, gh):
    """Per-node subset sizes of a mixed-size workload: a 4 x 4 map of regions
    over the grid, each region's size drawn (seeded) from the workload's sizes;
    row-major gw x gh int32."""
    regions = region_map(name)
    gy, gx = np.mgrid[0:gh, 0:gw]
    return np.ascontiguousarray(regions[gy * 4 // gh, gx * 4 // gw], np.int32)


def solve_params(name):
    c = config(name)
    return (0, 0, c["w"], c["h"], c["step"], c["subset"], c["window"], 0)


def displacement(name, x, y):
    """Ground-truth displacement of a reference point (same fields as synth.hip)."""
    c = config(name)
    p, cx, cy = c["p"], c["w"] / 2.0, c["h"] / 2.0
    dx, dy = x - cx, y - cy
    if c["kind"] == AFFINE:
        return p[0] + p[2] * dx + p[3] * dy, p[1] + p[4] * dx + p[5] * dy
    if c["kind"] == SIN_GAUSS:
        u = p[0] * np.sin(2 * np.pi * x / p[1]) * np.cos(2 * np.pi * y / p[2])
        v = p[3] * np.exp(-(dx * dx + dy * dy) / (2 * p[4] ** 2))
        return u, v
    if c["kind"] == BUMPS:
        th = p[0]
        u = (math.cos(th) - 1) * dx - math.sin(th) * dy
        v = math.sin(th) * dx + (math.cos(th) - 1) * dy
        for k in range(int(p[1])):
            bx, by, sg, au, av = p[2 + 5 * k: 7 + 5 * k]
            g = np.exp(-((x - bx) ** 2 + (y - by) ** 2) / (2 * sg * sg))
            u = u + au * g
            v = v + av * g
        return u, v
    if c["kind"] == UNIAXIAL:
        return p[0] * dx, -p[1] * p[0] * dy
    raise ValueError(name)
