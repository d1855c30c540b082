"""The boundary's image decode (codec.cpp, through libsemper_dev.so's host-only
sdk_decode / sdk_to_gray8 / sdk_to_mask8 / sdk_image_dimensions; no GPU).

Pinned by the reference's own codec suite (tests/unit/test_image_codec.cpp:
every case restated below, with PIL in place of cv::imencode), by PIL as an
independent decoder of the same files, by small encoders written here whose
output PIL reads back (Adam7 PNG, tiled / planar / big-endian / bit-reversed
TIFF), and by closed forms: the percentile stretch of an affine image of an
8-bit ramp gives the ramp back exactly, and the luma of a colour lies within
0.5 + 0.012 grey levels of 0.299 R + 0.587 G + 0.114 B (the 14-bit fixed point's
rounding plus its coefficient error)."""
import io
import struct
import zlib

import numpy as np
import pytest
from PIL import Image

from helpers import DICE_LADDER, dice_bytes, dice_image, engine_module

sa = engine_module()


@pytest.fixture(scope="module")
def lib():
    return sa.load_dev()


def dec(lib, data, ew=0, eh=0, mask=False):
    return sa.decode(data, ew, eh, mask=mask, lib=lib)


def enc(img, fmt, **kw):
    b = io.BytesIO()
    img.save(b, fmt, **kw)
    return b.getvalue()


def pil_gray(data):
    """PIL's decode of a file, with colour collapsed by the 14-bit fixed
    point luma (checked separately against the real-valued formula)."""
    im = Image.open(io.BytesIO(data))
    if im.mode in ("L", "I;16", "I", "F"):
        return np.array(im)
    a = np.array(im.convert("RGB")).astype(np.uint32)
    return luma(a)


def luma(a):
    a = a.astype(np.uint32)
    return ((a[..., 0] * 4899 + a[..., 1] * 9617 + a[..., 2] * 1868 + 8192) >> 14).astype(np.uint8)


def make_gray(w, h):
    y, x = np.mgrid[:h, :w]
    return ((x * 3 + y * 11) & 0xFF).astype(np.uint8)


def make_rgba(w, h):
    y, x = np.mgrid[:h, :w]
    return np.stack([(x * 7) & 0xFF, (y * 5) & 0xFF, (x + y) & 0xFF, np.full_like(x, 255)], -1).astype(np.uint8)


def ramp_image(k=100):
    """Each grey level 0..255 exactly k times (160 x 160 for k = 100), shuffled:
    the 0.1th and 99.9th percentiles are the levels 0 and 255 themselves."""
    v = np.repeat(np.arange(256), k)
    np.random.default_rng(3).shuffle(v)
    return v.reshape(160, -1)


# ---- the luma constants ------------------------------------------------------------

def test_luma_is_itu601(lib):
    """Raw RGBA (RawRgba_MatchesOpenCvGray): within rounding of
    0.299 R + 0.587 G + 0.114 B on 200k random colours and on the extremes."""
    rng = np.random.default_rng(0)
    px = rng.integers(0, 256, (400, 500, 4), dtype=np.uint8)
    px[0, :8, :3] = [[0, 0, 0], [255, 0, 0], [0, 255, 0], [0, 0, 255], [255, 255, 255], [255, 255, 0],
                     [0, 255, 255], [255, 0, 255]]
    g = dec(lib, px.tobytes(), 500, 400)
    y = 0.299 * px[..., 0] + 0.587 * px[..., 1] + 0.114 * px[..., 2]
    assert np.abs(g - y).max() <= 0.5 + 0.012
    assert g[0, :8].tolist() == [0, 76, 150, 29, 255, 226, 179, 105]
    # the alpha channel is ignored
    px2 = px.copy()
    px2[..., 3] = 0
    assert np.array_equal(dec(lib, px2.tobytes(), 500, 400), g)


# ---- tests/unit/test_image_codec.cpp ---------------------------------------------------

def test_raw_rgba_and_raw_gray(lib):
    rgba = make_rgba(8, 6)
    got = dec(lib, rgba.tobytes(), 8, 6)
    assert got.shape == (6, 8) and np.array_equal(got, luma(rgba))
    gray = make_gray(9, 7)
    got = dec(lib, gray.tobytes(), 9, 7)
    assert got.shape == (7, 9) and np.array_equal(got, gray)


def test_encoded_png_round_trips_and_colour_collapses(lib):
    gray = make_gray(12, 10)
    assert np.array_equal(dec(lib, enc(Image.fromarray(gray), "PNG")), gray)
    y, x = np.mgrid[:8, :10]
    bgr = np.stack([x * 9, y * 13, (x + y) * 3], -1).astype(np.uint8)
    got = dec(lib, enc(Image.fromarray(bgr[..., ::-1].copy()), "PNG"))
    assert got.shape == (8, 10) and np.array_equal(got, luma(bgr[..., ::-1]))


def test_encoded_16bit_to_8bit(lib):
    """EncodedNon8U_ConvertsTo8U and Encoded16Bit_PreservesContrast."""
    y, x = np.mgrid[:5, :6]
    got = dec(lib, enc(Image.fromarray(((x + y) * 1000).astype(np.uint16)), "PNG"))
    assert got.dtype == np.uint8 and got.shape == (5, 6)
    y, x = np.mgrid[:8, :64]
    got = dec(lib, enc(Image.fromarray((x * 64 + y).astype(np.uint16)), "PNG"))
    assert got.min() == 0 and got.max() == 255
    assert (np.diff(got[0].astype(int)) > 0).all()


def test_float_and_signed_scaled_to_gray8(lib):
    """FloatAndSigned_ScaledToGray8."""
    f = np.tile((np.arange(16) / 15.0).astype(np.float32), (4, 1))
    g = sa.to_gray8(f, lib=lib)
    assert g[0, 0] == 0 and g[0, 15] == 255 and g[2, 5] == 85
    assert (sa.to_gray8(np.full((3, 3), 0.25), lib=lib) == 128).all()
    with_inf = f.copy()
    with_inf[1, 3] = np.inf
    assert (sa.to_gray8(with_inf, lib=lib) == 128).all()
    g8 = make_gray(5, 4)
    assert np.array_equal(sa.to_gray8(g8, lib=lib), g8)
    assert sa.to_gray8(np.full((4, 5, 3), (30, 20, 10), np.uint8), lib=lib).shape == (4, 5)


def test_encoded_with_raw_length_still_decoded(lib):
    """EncodedWithRawLength_StillDecoded: a PNG whose length is w * h for the
    expected size decodes as a PNG; raw bytes that merely start like a JPEG
    stay raw."""
    gray = make_gray(11, 9)
    png = enc(Image.fromarray(gray), "PNG")
    got = dec(lib, png, len(png), 1)
    assert got.shape == (9, 11) and np.array_equal(got, gray)
    raw = np.full(11 * 9, 7, np.uint8)
    raw[:3] = [0xFF, 0xD8, 0xFF]
    r = dec(lib, raw.tobytes(), 11, 9)
    assert r.shape == (9, 11) and r[0, 0] == 0xFF and r[0, 3] == 7


def test_mismatched_length_falls_through_to_encoded(lib):
    gray = make_gray(11, 9)
    png = enc(Image.fromarray(gray), "PNG")
    assert len(png) not in (99, 396)
    assert np.array_equal(dec(lib, png, 11, 9), gray)


def test_null_empty_and_garbage(lib):
    assert lib.sdk_decode(None, 16, 4, 4, 0, None, 0, None, None) == -3
    assert dec(lib, b"", 4, 4) is None
    junk = bytes((i * 7 + 1) & 0xFF for i in range(37))
    assert dec(lib, junk) is None and dec(lib, junk, 100, 100) is None


def test_image_dimensions(lib):
    """ImageDimensions_EncodedBuffer / ZerosOnFailureAndNull /
    RawBufferYieldsZeros_KnownTrap."""
    assert sa.image_dimensions(enc(Image.fromarray(make_gray(13, 5)), "PNG"), lib=lib) == (13, 5)
    assert sa.image_dimensions(bytes([0xAB] * 29), lib=lib) == (0, 0)
    assert sa.image_dimensions(None, lib=lib) == (0, 0)
    assert sa.image_dimensions(make_rgba(8, 8).tobytes(), lib=lib) == (0, 0)
    # PNM has no header reader: sized by a decode
    assert sa.image_dimensions(enc(Image.fromarray(make_gray(7, 3)), "PPM"), lib=lib) == (7, 3)


def test_masks_are_boolean(lib):
    """Mask16BitAndFloat_AreBoolean."""
    png = enc(Image.fromarray(np.zeros((6, 8), np.uint16)), "PNG")
    m = dec(lib, png, mask=True)
    assert m.shape == (6, 8) and (m == 0).all()
    assert (dec(lib, png) == 128).sum() == 48
    labels = np.zeros((4, 4), np.uint16)
    labels[1, 1], labels[2, 2] = 1, 2
    lm = sa.to_mask8(labels, lib=lib)
    assert lm[1, 1] == 255 and lm[2, 2] == 255 and lm[0, 0] == 0
    assert (sa.to_mask8(np.full((3, 3), 0.2, np.float32), lib=lib) == 255).all()
    g8 = make_gray(5, 4)
    assert np.array_equal(sa.to_mask8(g8, lib=lib), g8)


def test_16bit_hot_pixel_does_not_set_the_scale(lib):
    """Encoded16Bit_HotPixelDoesNotSetTheScale."""
    u16 = np.tile((1000 + 20 * np.arange(100)).astype(np.uint16), (50, 1))
    u16[7, 7] = 65535
    for g in (sa.to_gray8(u16, lib=lib), dec(lib, enc(Image.fromarray(u16), "PNG"))):
        assert g[7, 7] == 255 and g[0, 0] <= 3 and g[0, 99] >= 250
        assert (np.diff(g[0].astype(int)) != 0).sum() >= 95


def test_oversized_header_refused_before_decode(lib):
    """OversizedHeader_RefusedBeforeDecode: 40000 x 40000 in the IHDR is
    refused (image and mask) yet reported by image_dimensions; TIFF, BMP and
    JPEG headers are read too."""
    png = enc(Image.fromarray(make_gray(16, 9)), "PNG")
    assert sa.image_dimensions(png, lib=lib) == (16, 9)
    big = bytearray(png)
    big[16:24] = struct.pack(">II", 40000, 40000)
    assert dec(lib, bytes(big)) is None and dec(lib, bytes(big), mask=True) is None
    assert sa.image_dimensions(bytes(big), lib=lib) == (40000, 40000)
    for fmt in ("JPEG", "TIFF", "BMP"):
        assert sa.image_dimensions(enc(Image.fromarray(make_gray(16, 9)), fmt), lib=lib) == (16, 9)


# ---- the percentile stretch: closed forms -----------------------------------------

@pytest.mark.parametrize("dtype,a,b", [(np.uint16, 1000, 4), (np.int16, -500, 3), (np.int32, -70000, 16),
                                       (np.float32, 0.5, 1 / 512), (np.float64, -2.0, 1 / 1024)])
def test_stretch_inverts_an_affine_ramp(lib, dtype, a, b):
    v = ramp_image()
    img = (a + b * v).astype(dtype)
    assert np.array_equal(sa.to_gray8(img, lib=lib), v)
    # and as three equal channels (the luma of equal channels is the channel)
    assert np.array_equal(sa.to_gray8(np.repeat(img[..., None], 3, -1), lib=lib), v)
    # outliers beyond the percentiles saturate and do not move the scale
    img2 = img.copy()
    img2[0, 0], img2[0, 1] = img.max() + 40 * b, img.min() - 40 * b
    out = sa.to_gray8(img2, lib=lib)
    assert out[0, 0] == 255 and out[0, 1] == 0 and np.array_equal(out.ravel()[2:], v.ravel()[2:])


def test_small_image_range_is_min_max(lib):
    """Up to 1000 pixels the percentiles are the extremes: a 10 x 10 ramp
    0..99 maps 0 -> 0 and 99 -> 255, linearly."""
    img = np.arange(100, dtype=np.uint16).reshape(10, 10) * 7 + 3
    out = sa.to_gray8(img, lib=lib).ravel().astype(int)
    assert out[0] == 0 and out[-1] == 255
    assert np.abs(out - np.arange(100) * 255 / 99).max() <= 0.5 + 1e-6


# ---- formats against PIL ------------------------------------------------------------

def test_dice_fixtures_decode_as_pil(lib):
    for name in ["oht_cfrp_00.tiff"] + [i for i, _ in DICE_LADDER] + ["ref.tif", "def.tif"]:
        got = dec(lib, dice_bytes(name))
        assert np.array_equal(got, dice_image(name)), name


RNG = np.random.default_rng(11)
G8 = RNG.integers(0, 256, (37, 53), dtype=np.uint8)
RGB = RNG.integers(0, 256, (37, 53, 3), dtype=np.uint8)
RGBA = RNG.integers(0, 256, (37, 53, 4), dtype=np.uint8)


@pytest.mark.parametrize("fmt,kw", [("PNG", {}), ("PNG", {"optimize": True}), ("TIFF", {}),
                                    ("TIFF", {"compression": "tiff_lzw"}),
                                    ("TIFF", {"compression": "tiff_lzw", "tiffinfo": {317: 2}}),
                                    ("TIFF", {"compression": "tiff_adobe_deflate"}),
                                    ("TIFF", {"compression": "tiff_adobe_deflate", "tiffinfo": {317: 2}}),
                                    ("TIFF", {"compression": "packbits"}), ("BMP", {}), ("PPM", {})])
@pytest.mark.parametrize("mode", ["L", "RGB", "P", "1"])
def test_formats_match_pil(lib, fmt, kw, mode):
    if mode == "P":
        if fmt == "PPM":
            pytest.skip("PNM has no palette")
        img = Image.fromarray(RGB).convert("P", palette=Image.ADAPTIVE, colors=200)
    elif mode == "1":
        if 317 in kw.get("tiffinfo", {}):
            pytest.skip("libtiff has no predictor for 1-bit samples")
        img = Image.fromarray(G8).convert("1")
    else:
        img = Image.fromarray(G8 if mode == "L" else RGB)
    data = enc(img, fmt, **kw)
    assert np.array_equal(dec(lib, data), pil_gray(data))


@pytest.mark.parametrize("fmt,kw", [("PNG", {}), ("TIFF", {}), ("TIFF", {"compression": "tiff_lzw", "tiffinfo": {317: 2}}),
                                    ("TIFF", {"compression": "tiff_adobe_deflate"}), ("PPM", {})])
def test_16bit_gray_files(lib, fmt, kw):
    """16-bit gray PNG, TIFF (plain, LZW + horizontal predictor, Deflate) and
    PGM: the affine ramp comes back exactly through the stretch."""
    v = ramp_image()
    assert np.array_equal(dec(lib, enc(Image.fromarray((1000 + 4 * v).astype(np.uint16)), fmt, **kw)), v)


@pytest.mark.parametrize("kw", [{}, {"compression": "tiff_lzw", "tiffinfo": {317: 3}},
                                {"compression": "tiff_adobe_deflate", "tiffinfo": {317: 3}}])
def test_float_tiff(lib, kw):
    v = ramp_image()
    assert np.array_equal(dec(lib, enc(Image.fromarray((0.5 + v / 512).astype(np.float32)), "TIFF", **kw)), v)


def test_alpha_rules(lib):
    """RGBA PNG: alpha dropped. RGBA TIFF with unassociated alpha (what PIL
    writes): premultiplied as (v * a + 127) / 255 before the luma. LA PNG: the
    gray channel."""
    a = RGBA.astype(np.uint32)
    assert np.array_equal(dec(lib, enc(Image.fromarray(RGBA), "PNG")), luma(RGBA[..., :3]))
    prem = (a[..., :3] * a[..., 3:4] + 127) // 255
    tif = enc(Image.fromarray(RGBA), "TIFF")
    assert Image.open(io.BytesIO(tif)).tag_v2.get(338) == (2,)
    assert np.array_equal(dec(lib, tif), luma(prem))
    la = np.stack([G8, RNG.integers(0, 256, G8.shape, dtype=np.uint8)], -1)
    assert np.array_equal(dec(lib, enc(Image.fromarray(la, "LA"), "PNG")), G8)


# ---- PNG layouts PIL does not write: a small encoder, read back by PIL -------------

ADAM7 = [(0, 0, 8, 8), (4, 0, 8, 8), (0, 4, 4, 8), (2, 0, 4, 4), (0, 2, 2, 4), (1, 0, 2, 2), (0, 1, 1, 2)]


def _paeth(a, b, c):
    p = a + b - c
    pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
    return a if pa <= pb and pa <= pc else (b if pb <= pc else c)


def png_bytes(samples, ct, bd, interlace=False, palette=None):
    """samples: (h, w, channels) integer sample values. Rows filtered with
    types 0..4 in turn."""
    h, w, ch = samples.shape
    bpp = max(1, ch * bd // 8)

    def pack(row):
        flat = [int(v) for v in row.reshape(-1)]
        if bd == 16:
            return b"".join(v.to_bytes(2, "big") for v in flat)
        if bd == 8:
            return bytes(flat)
        bits = "".join(format(v, f"0{bd}b") for v in flat)
        bits += "0" * (-len(bits) % 8)
        return int(bits, 2).to_bytes(len(bits) // 8, "big")

    raw, k = bytearray(), 0
    for x0, y0, dx, dy in (ADAM7 if interlace else [(0, 0, 1, 1)]):
        sub = samples[y0::dy, x0::dx]
        if sub.size == 0:
            continue
        prev = None
        for r in sub:
            cur = pack(r)
            f = k % 5
            k += 1
            out = bytearray(len(cur))
            for i in range(len(cur)):
                a = cur[i - bpp] if i >= bpp else 0
                b = prev[i] if prev else 0
                c = prev[i - bpp] if prev and i >= bpp else 0
                out[i] = (cur[i] - [0, a, b, (a + b) // 2, _paeth(a, b, c)][f]) & 255
            raw += bytes([f]) + out
            prev = cur
    chunks = [(b"IHDR", struct.pack(">IIBBBBB", w, h, bd, ct, 0, 0, int(interlace)))]
    if palette is not None:
        chunks.append((b"PLTE", bytes(np.asarray(palette, np.uint8).reshape(-1))))
    chunks += [(b"IDAT", zlib.compress(bytes(raw))), (b"IEND", b"")]
    out = b"\x89PNG\r\n\x1a\n"
    for t, d in chunks:
        out += struct.pack(">I", len(d)) + t + d + struct.pack(">I", zlib.crc32(t + d))
    return out


@pytest.mark.parametrize("interlace", [False, True])
@pytest.mark.parametrize("ct,bd", [(0, 1), (0, 2), (0, 4), (0, 8), (0, 16), (2, 8), (2, 16), (3, 1), (3, 2), (3, 4),
                                   (3, 8), (4, 8), (6, 8), (6, 16)])
def test_png_layouts(lib, ct, bd, interlace):
    h, w = 13, 19
    ch = {0: 1, 2: 3, 3: 1, 4: 2, 6: 4}[ct]
    rng = np.random.default_rng(ct * 100 + bd)
    s = rng.integers(0, 1 << bd, (h, w, ch))
    pal = rng.integers(0, 256, (1 << bd, 3)) if ct == 3 else None
    data = png_bytes(s, ct, bd, interlace, pal)
    # the encoder is read back by PIL
    im = Image.open(io.BytesIO(data))
    if ct == 3:
        want = luma(pal[s[..., 0]])
        assert np.array_equal(np.array(im), s[..., 0])
    elif ct in (0, 4):
        g = s[..., 0]
        if bd == 16:
            assert np.array_equal(np.array(im)[..., 0] if ct == 4 else np.array(im), g)
            want = None
        else:
            want = (g * 255 // ((1 << bd) - 1)).astype(np.uint8)
            assert np.array_equal(np.array(im.convert("L")), want)
    else:
        rgb = s[..., :3]
        want = None if bd == 16 else luma(rgb)
        if bd == 8:
            assert np.array_equal(np.array(im.convert("RGB")), rgb)
    got = dec(lib, data)
    assert got.shape == (h, w)
    if want is not None:
        assert np.array_equal(got, want)
    else:
        # 16-bit: the stretch of the gray plane (or of the fixed-point luma
        # of the RGB samples), as to_gray8 gives it for the raster
        ref16 = s[..., 0] if ct in (0, 4) else (s[..., 0] * 4899 + s[..., 1] * 9617 + s[..., 2] * 1868 + 8192) >> 14
        assert np.array_equal(got, sa.to_gray8(ref16.astype(np.uint16), lib=lib))


def test_png_damage_is_refused(lib):
    data = png_bytes(np.random.default_rng(1).integers(0, 256, (9, 9, 1)), 0, 8)
    assert dec(lib, data) is not None
    assert dec(lib, data[:-12]) is None                      # no IEND
    bad = bytearray(data)
    bad[41] ^= 1                                             # inside IDAT: its CRC fails
    assert dec(lib, bytes(bad)) is None
    short = png_bytes(np.zeros((9, 9, 1), int), 0, 8)
    i = short.index(b"IDAT")
    n = struct.unpack(">I", short[i - 4:i])[0]
    d = zlib.compress(zlib.decompress(short[i + 4:i + 4 + n])[:-5])   # five bytes short
    cut = short[:i - 4] + struct.pack(">I", len(d)) + b"IDAT" + d + struct.pack(">I", zlib.crc32(b"IDAT" + d)) + \
        short[i + 8 + n:]
    assert dec(lib, cut) is None


# ---- TIFF layouts PIL does not write: a small writer, read back by PIL -------------

def tiff_bytes(samples, bps, photometric, fmt=1, big_endian=False, tile=None, planar=1, fill_order=1,
               compression=1, rows_per_strip=None):
    """samples: (h, w, spp). Strips (or tiles) of uncompressed / Deflate data."""
    h, w, spp = samples.shape
    e = ">" if big_endian else "<"
    code = {8: "B", 16: "H", 32: {1: "I", 2: "i", 3: "f"}[fmt], 64: "d"}[bps]
    dt = np.dtype(e + code if bps > 8 else code)
    planes = [samples[..., k:k + 1] for k in range(spp)] if planar == 2 else [samples]
    chunks = []
    if tile:
        tw, tl = tile
        for pl in planes:
            for ty in range(0, h, tl):
                for tx in range(0, w, tw):
                    block = np.zeros((tl, tw, pl.shape[2]), pl.dtype)
                    part = pl[ty:ty + tl, tx:tx + tw]
                    block[:part.shape[0], :part.shape[1]] = part
                    chunks.append(block.astype(dt).tobytes())
    else:
        rps = rows_per_strip or h
        for pl in planes:
            for y in range(0, h, rps):
                chunks.append(pl[y:y + rps].astype(dt).tobytes())
    if fill_order == 2:
        rev = bytes(int(f"{b:08b}"[::-1], 2) for b in range(256))
        chunks = [c.translate(rev) for c in chunks]
    if compression == 8:
        chunks = [zlib.compress(c) for c in chunks]
    tags = {256: (4, [w]), 257: (4, [h]), 258: (3, [bps] * spp), 259: (3, [compression]), 262: (3, [photometric]),
            266: (3, [fill_order]), 277: (3, [spp]), 284: (3, [planar]), 339: (3, [fmt] * spp)}
    if tile:
        tags[322], tags[323] = (4, [tile[0]]), (4, [tile[1]])
    else:
        tags[278] = (4, [rows_per_strip or h])
    off_tag, cnt_tag = (324, 325) if tile else (273, 279)
    tags[cnt_tag] = (4, [len(c) for c in chunks])
    tags[off_tag] = (4, [0] * len(chunks))
    n = len(tags)
    ifd_size = 2 + 12 * n + 4
    extra, extra_off = bytearray(), 8 + ifd_size
    # the chunk data after the IFD and the out-of-line values
    arrays = {}
    for t, (ty, vals) in sorted(tags.items()):
        size = (2 if ty == 3 else 4) * len(vals)
        if size > 4:
            arrays[t] = extra_off + len(extra)
            extra += b"\0" * size
    data_off = extra_off + len(extra)
    offs, pos = [], data_off
    for c in chunks:
        offs.append(pos)
        pos += len(c)
    tags[off_tag] = (4, offs)
    ent, extra = b"", bytearray()
    for t, (ty, vals) in sorted(tags.items()):
        f = "H" if ty == 3 else "I"
        packed = struct.pack(e + f * len(vals), *vals)
        if len(packed) <= 4:
            ent += struct.pack(e + "HHI", t, ty, len(vals)) + packed.ljust(4, b"\0")
        else:
            ent += struct.pack(e + "HHII", t, ty, len(vals), extra_off + len(extra))
            extra += packed
    head = (b"MM\0*" if big_endian else b"II*\0") + struct.pack(e + "I", 8)
    return head + struct.pack(e + "H", n) + ent + struct.pack(e + "I", 0) + bytes(extra) + b"".join(chunks)


@pytest.mark.parametrize("layout", [dict(), dict(big_endian=True), dict(tile=(16, 16)), dict(planar=2),
                                    dict(planar=2, tile=(16, 32), big_endian=True), dict(rows_per_strip=7),
                                    dict(compression=8, tile=(32, 16)), dict(fill_order=2)])
@pytest.mark.parametrize("kind", ["u8", "rgb8", "u16", "rgb16", "s16", "s32", "f32", "f64"])
def test_tiff_layouts(lib, layout, kind):
    v = ramp_image()                                          # 160 x 160, levels 0..255
    rng = np.random.default_rng(7)
    if kind == "u8":
        s, bps, ph, fmt = v[..., None], 8, 1, 1
        want = v
    elif kind == "rgb8":
        s = rng.integers(0, 256, v.shape + (3,))
        bps, ph, fmt = 8, 2, 1
        want = luma(s)
    else:
        bps = {"u16": 16, "rgb16": 16, "s16": 16, "s32": 32, "f32": 32, "f64": 64}[kind]
        fmt = {"s16": 2, "s32": 2, "f32": 3, "f64": 3}.get(kind, 1)
        a, b = {"u16": (1000, 4), "rgb16": (1000, 4), "s16": (-500, 3), "s32": (-70000, 16), "f32": (0.5, 1 / 512),
                "f64": (-2.0, 1 / 1024)}[kind]
        g = a + b * v
        s = np.repeat(g[..., None], 3, -1) if kind == "rgb16" else g[..., None]
        ph = 2 if kind == "rgb16" else 1
        want = v
    if layout.get("planar") == 2 and s.shape[2] == 1:
        pytest.skip("one sample per pixel")
    data = tiff_bytes(s, bps, ph, fmt, **layout)
    if kind in ("u8", "rgb8") and layout.get("fill_order", 1) == 1:
        assert np.array_equal(pil_gray(data), want)           # the writer, read back by PIL
    assert np.array_equal(dec(lib, data), want)


def test_tiff_min_is_white_and_bilevel(lib):
    v = RNG.integers(0, 256, (21, 30))
    assert np.array_equal(dec(lib, tiff_bytes(v[..., None], 8, 0)), 255 - v)
    data = enc(Image.fromarray(G8).convert("1"), "TIFF", compression="group4")
    assert dec(lib, data) is None                            # CCITT: not decoded


def test_truncated_tiff_is_refused(lib):
    data = tiff_bytes(ramp_image()[..., None], 8, 1)
    assert dec(lib, data) is not None and dec(lib, data[:-100]) is None


def test_tiff_tiles_beyond_the_data_are_refused_before_allocation(lib):
    """A few hundred bytes that declare a 1 x 1 image in 65536 x 65536 Deflate
    tiles: the tile would need 4 GiB, while its coded bytes can inflate to at
    most 1032 times their size. It is refused without touching that memory
    (as fast as any other refusal)."""
    import time
    data = bytearray(tiff_bytes(np.array([[[7]]], np.uint8), 8, 1, tile=(16, 16), compression=8))
    n = struct.unpack("<H", data[8:10])[0]
    for k in range(n):
        e = 10 + 12 * k
        if struct.unpack("<H", data[e:e + 2])[0] in (322, 323):
            data[e + 8:e + 12] = struct.pack("<I", 65536)
    assert len(data) < 400
    t0 = time.perf_counter()
    assert dec(lib, bytes(data)) is None
    assert time.perf_counter() - t0 < 0.05


def test_unsupported_formats_fail(lib):
    """WebP and CMYK JPEG are not decoded here: the call fails, as a file
    that does not decode fails in the reference."""
    assert dec(lib, enc(Image.fromarray(RGB), "WEBP")) is None
    assert dec(lib, enc(Image.fromarray(RGB).convert("CMYK"), "JPEG")) is None


def speckle(h, w, seed, rgb):
    """A smooth random texture (JPEG's natural case) with sharp edges mixed in."""
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[:h, :w]
    chans = []
    for c in range(3 if rgb else 1):
        f = sum(np.sin(x * rng.uniform(0.05, 0.9) + y * rng.uniform(0.05, 0.9) + rng.uniform(0, 6)) for _ in range(5))
        f = 128 + 25 * f + rng.normal(0, 12, (h, w))
        f[(x // 7 + y // 5 + c) % 4 == 0] = 255 * (c % 2)
        chans.append(np.clip(f, 0, 255))
    a = np.stack(chans, -1).astype(np.uint8)
    return a if rgb else a[..., 0]


JPEG_CASES = [
    ("grey", dict(quality=75)), ("grey", dict(quality=100)), ("grey", dict(quality=20, optimize=True)),
    ("grey", dict(quality=85, progressive=True)),
    ("444", dict(quality=90, subsampling=0)), ("422", dict(quality=80, subsampling=1)),
    ("420", dict(quality=75, subsampling=2)), ("420", dict(quality=95, subsampling=2, optimize=True)),
    ("420", dict(quality=70, subsampling=2, progressive=True)), ("444", dict(quality=60, subsampling=0, progressive=True)),
    ("422", dict(quality=90, subsampling=1, restart_marker_blocks=3)),
    ("420", dict(quality=85, subsampling=2, restart_marker_rows=1, progressive=True)),
]


@pytest.mark.parametrize("size", [(23, 37), (48, 64), (9, 17), (1, 1), (8, 8), (2, 3), (61, 3)])
@pytest.mark.parametrize("kind,kw", JPEG_CASES, ids=[f"{k}-{i}" for i, (k, _) in enumerate(JPEG_CASES)])
def test_jpeg_matches_libjpeg_turbo(lib, kind, kw, size):
    """JPEG as libjpeg-turbo decodes it with its defaults (ISLOW IDCT, fancy
    upsampling, JFIF YCbCr -> RGB), which is what cv::imdecode hands the
    reference: PIL's decode of the same file, through the same luma, bit for
    bit -- sequential and progressive, optimised Huffman tables, restart
    intervals, 4:4:4 / 4:2:2 / 4:2:0, odd sizes down to 1 x 1."""
    h, w = size
    a = speckle(h, w, h * 131 + w, kind != "grey")
    data = enc(Image.fromarray(a), "JPEG", **kw)
    got = dec(lib, data)
    assert got is not None
    want = pil_gray(data)
    assert got.shape == want.shape and np.array_equal(got, want), np.abs(got.astype(int) - want).max()


@pytest.mark.parametrize("sub", [0, 1, 2])
def test_jpeg_mask_is_the_y_plane(lib, sub):
    """Masks are read with IMREAD_GRAYSCALE: libjpeg then delivers the Y
    component itself, not a luma of the decoded RGB (PIL's YCbCr draft gives
    the decoder's Y plane)."""
    a = speckle(40, 56, 7 + sub, True)
    data = enc(Image.fromarray(a), "JPEG", quality=80, subsampling=sub)
    got = dec(lib, data, mask=True)
    im = Image.open(io.BytesIO(data))
    im.draft("YCbCr", im.size)
    y = np.array(im)[..., 0]
    assert im.mode == "YCbCr" and np.array_equal(got, y)
    assert not np.array_equal(got, pil_gray(data))   # the luma of RGB differs


@pytest.mark.parametrize("orientation", range(1, 9))
def test_jpeg_exif_orientation(lib, orientation):
    """OpenCV applies the EXIF orientation after decode (these flags do not
    ask it to ignore it): the eight orientations against PIL's transpose."""
    from PIL import ImageOps
    a = speckle(30, 44, orientation, True)
    exif = Image.Exif()
    exif[0x0112] = orientation
    data = enc(Image.fromarray(a), "JPEG", quality=90, exif=exif.tobytes())
    got = dec(lib, data)
    im = ImageOps.exif_transpose(Image.open(io.BytesIO(data)))
    want = luma(np.array(im.convert("RGB")))
    assert got.shape == want.shape and np.array_equal(got, want)


def test_jpeg_truncated_and_garbage(lib):
    """A JPEG cut short decodes with the missing data read as zero bits, as
    libjpeg does (it warns and goes on); bytes that only start like a JPEG
    fail."""
    data = enc(Image.fromarray(speckle(64, 64, 3, True)), "JPEG", quality=90)
    cut = data[: len(data) * 2 // 3]
    got = dec(lib, cut)
    full = dec(lib, data)
    assert got is not None and got.shape == full.shape
    assert np.array_equal(got[:8], full[:8])
    assert dec(lib, b"\xff\xd8\xff\xe0" + b"\x00" * 40) is None


def test_pnm_variants(lib):
    g = RNG.integers(0, 256, (5, 7))
    ascii_pgm = ("P2\n# a comment\n7 5\n255\n" + " ".join(str(int(x)) for x in g.ravel()) + "\n").encode()
    assert np.array_equal(dec(lib, ascii_pgm), g)
    g15 = RNG.integers(0, 16, (5, 7))
    bin15 = b"P5 7 5 15\n" + bytes(int(x) for x in g15.ravel())
    assert np.array_equal(dec(lib, bin15), (g15 * 255 // 15).astype(np.uint8))
    bits = RNG.integers(0, 2, (4, 9))
    p1 = ("P1\n9 4\n" + "\n".join("".join(str(int(b)) for b in row) for row in bits) + "\n").encode()
    assert np.array_equal(dec(lib, p1), np.where(bits == 1, 0, 255))
    rgb = RNG.integers(0, 256, (3, 4, 3))
    p3 = ("P3 4 3 255 " + " ".join(str(int(x)) for x in rgb.ravel())).encode()
    assert np.array_equal(dec(lib, p3), luma(rgb))
    assert dec(lib, b"P5\n7 5\n255\n" + bytes(10)) is None   # short data


def test_bmp_16bit_and_top_down(lib):
    """5-5-5 BMP (channels shifted left three bits) and a top-down (negative
    height) 24-bit BMP, written here."""
    rng = np.random.default_rng(5)
    w, h = 6, 4
    px = rng.integers(0, 1 << 15, (h, w))
    row_bytes = (w * 2 + 3) & ~3
    body = b"".join(struct.pack("<" + "H" * w, *map(int, px[y])).ljust(row_bytes, b"\0") for y in range(h - 1, -1, -1))
    info = struct.pack("<IiiHHIIiiII", 40, w, h, 1, 16, 0, len(body), 0, 0, 0, 0)
    bmp = b"BM" + struct.pack("<IHHI", 14 + 40 + len(body), 0, 0, 54) + info + body
    r, g, b = (px >> 10 & 31) << 3, (px >> 5 & 31) << 3, (px & 31) << 3
    assert np.array_equal(dec(lib, bmp), luma(np.stack([r, g, b], -1)))
    rgb = rng.integers(0, 256, (h, w, 3))
    row_bytes = (w * 3 + 3) & ~3
    body = b"".join(bytes(rgb[y, :, ::-1].astype(np.uint8).ravel()).ljust(row_bytes, b"\0") for y in range(h))
    info = struct.pack("<IiiHHIIiiII", 40, w, -h, 1, 24, 0, len(body), 0, 0, 0, 0)
    bmp = b"BM" + struct.pack("<IHHI", 14 + 40 + len(body), 0, 0, 54) + info + body
    assert np.array_equal(dec(lib, bmp), luma(rgb))


def test_jpeg_damage_never_breaks_the_decoder(lib):
    """Byte flips, truncations and spliced markers in valid JPEG files
    (sequential, progressive, restart intervals): each either decodes to an
    image of a plausible size or fails, never faults or hangs."""
    rng = np.random.default_rng(5)
    bases = [enc(Image.fromarray(speckle(37, 53, 1, True)), "JPEG", quality=80, subsampling=2),
             enc(Image.fromarray(speckle(40, 33, 2, True)), "JPEG", quality=70, progressive=True),
             enc(Image.fromarray(speckle(24, 64, 3, False)), "JPEG", quality=90, restart_marker_blocks=2)]
    for base in bases:
        for trial in range(400):
            b = bytearray(base)
            for _ in range(int(rng.integers(1, 6))):
                i = int(rng.integers(2, len(b)))
                b[i] = int(rng.integers(0, 256)) if rng.random() < 0.8 else 0xFF
            if trial % 5 == 0:
                b = b[: int(rng.integers(4, len(b)))]
            got = dec(lib, bytes(b))
            assert got is None or (got.ndim == 2 and got.size <= 1 << 28)
