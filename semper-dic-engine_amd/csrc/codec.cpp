// Host-side image decode for the C ABI. The conversion rules are the
// reference's (src/io/image_codec.cpp); the per-format sample handling
// restates what its decoder stack does with each layout:
//   PNG   libpng through OpenCV's PNG decoder with IMREAD_ANYDEPTH |
//         IMREAD_ANYCOLOR: gray 1/2/4-bit expanded to 8 (v * 255 / max),
//         16-bit kept, palette expanded to RGB (indices past the palette
//         black), alpha dropped; CRCs of critical chunks checked, IEND
//         required, at most 1,000,000 columns or rows (libpng's default
//         user limit).
//   TIFF  libtiff through OpenCV's TIFF decoder. Samples of 1-8 bits take
//         libtiff's RGBA route: min-is-black / min-is-white through the
//         map v * 255 / max (inverted for min-is-white), palettes through a
//         colour map read as 16-bit when any entry exceeds 255 (then >> 8),
//         RGB with unassociated alpha premultiplied as (v * a + 127) / 255.
//         16-bit min-is-black and RGB stay 16-bit; 32-bit int and IEEE
//         float and 64-bit float keep their type. Compression none, LZW,
//         Deflate and PackBits; horizontal and floating-point predictors;
//         strips or tiles, contiguous or planar; fill order 2.
//   BMP   OpenCV's BMP decoder: 1/4/8-bit palettes, 15/16-bit (5-5-5 or,
//         with bit fields, 5-6-5; channels shifted left, not replicated),
//         24 and 32-bit (alpha dropped); bottom-up or top-down rows.
//   PNM   OpenCV's PxM decoder: P1-P6, 8-bit samples rescaled v * 255 /
//         maxval when maxval < 255, 16-bit samples (maxval > 255) kept,
//         PBM ones black.
//   JPEG  libjpeg-turbo's default decompression (jpeg.cpp): sequential and
//         progressive Huffman, 8-bit, grey or YCbCr / RGB, ISLOW IDCT, fancy
//         upsampling, EXIF orientation; masks take the Y plane.
// Not decoded (the call fails with the reference's decode-failure result):
// CMYK / arithmetic / lossless / 12-bit JPEG, WebP, BigTIFF, BMP run-length
// encoding, TIFF JPEG / CCITT / YCbCr / CMYK / 12-bit samples, 16-bit palette
// or min-is-white TIFF.
#include "codec.hpp"

#include <zlib.h>

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <limits>

namespace sdic {
namespace codec {

namespace {

inline uint32_t be16(const uint8_t* p) { return (uint32_t)p[0] << 8 | p[1]; }
inline uint32_t be32(const uint8_t* p) {
    return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3];
}
inline uint32_t le16(const uint8_t* p) { return (uint32_t)p[1] << 8 | p[0]; }
inline uint32_t le32(const uint8_t* p) {
    return (uint32_t)p[3] << 24 | (uint32_t)p[2] << 16 | (uint32_t)p[1] << 8 | p[0];
}

size_t depth_bytes(Depth d) {
    switch (d) {
        case Depth::U8: return 1;
        case Depth::U16:
        case Depth::S16: return 2;
        case Depth::S32:
        case Depth::F32: return 4;
        case Depth::F64: return 8;
    }
    return 1;
}

bool dims_ok(uint64_t w, uint64_t h) {
    return w > 0 && h > 0 && w <= kMaxPixels && h <= kMaxPixels && w * h <= kMaxPixels;
}

Raster make_raster(int w, int h, int cn, Depth d) {
    Raster r;
    r.w = w;
    r.h = h;
    r.cn = cn;
    r.depth = d;
    r.px.assign((size_t)w * h * cn * depth_bytes(d), 0);
    return r;
}

// ITU-R 601 luma in 14-bit fixed point, rounded: cv::cvtColor *2GRAY on 8U
// and 16U data.
inline uint32_t luma_fixed(uint32_t r, uint32_t g, uint32_t b) {
    return (r * 4899u + g * 9617u + b * 1868u + 8192u) >> 14;
}

// Inflates a zlib stream into exactly `want` bytes; trailing input is
// ignored, a short stream fails.
bool inflate_to(const uint8_t* src, size_t n, uint8_t* dst, size_t want) {
    if (want == 0) return true;
    z_stream zs;
    std::memset(&zs, 0, sizeof zs);
    if (inflateInit(&zs) != Z_OK) return false;
    size_t in_off = 0, out_off = 0;
    bool ok = false;
    for (;;) {
        if (zs.avail_in == 0 && in_off < n) {
            const size_t c = std::min<size_t>(n - in_off, size_t(1) << 30);
            zs.next_in = const_cast<Bytef*>(src + in_off);
            zs.avail_in = (uInt)c;
            in_off += c;
        }
        const size_t c = std::min<size_t>(want - out_off, size_t(1) << 30);
        zs.next_out = dst + out_off;
        zs.avail_out = (uInt)c;
        const int rc = inflate(&zs, Z_NO_FLUSH);
        out_off += c - zs.avail_out;
        if (out_off == want) {
            ok = true;
            break;
        }
        if (rc == Z_STREAM_END) break;
        if (rc == Z_OK) continue;
        if (rc == Z_BUF_ERROR && zs.avail_in == 0 && in_off < n) continue;
        break;
    }
    inflateEnd(&zs);
    return ok;
}

// ---------------------------------------------------------------------------
// PNG
// ---------------------------------------------------------------------------

inline int paeth(int a, int b, int c) {
    const int p = a + b - c, pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
    if (pa <= pb && pa <= pc) return a;
    return pb <= pc ? b : c;
}

// Reverses the five PNG row filters in place over `rows` rows of `rb` bytes,
// each preceded by its filter byte; `bpp` = bytes per complete pixel (>= 1).
bool png_unfilter(uint8_t* p, size_t rows, size_t rb, size_t bpp) {
    const uint8_t* prior = nullptr;
    for (size_t y = 0; y < rows; ++y) {
        const uint8_t f = p[0];
        uint8_t* row = p + 1;
        switch (f) {
            case 0: break;
            case 1:
                for (size_t i = bpp; i < rb; ++i) row[i] = (uint8_t)(row[i] + row[i - bpp]);
                break;
            case 2:
                if (prior)
                    for (size_t i = 0; i < rb; ++i) row[i] = (uint8_t)(row[i] + prior[i]);
                break;
            case 3:
                for (size_t i = 0; i < rb; ++i) {
                    const int a = i >= bpp ? row[i - bpp] : 0, b = prior ? prior[i] : 0;
                    row[i] = (uint8_t)(row[i] + ((a + b) >> 1));
                }
                break;
            case 4:
                for (size_t i = 0; i < rb; ++i) {
                    const int a = i >= bpp ? row[i - bpp] : 0, b = prior ? prior[i] : 0,
                              c = (prior && i >= bpp) ? prior[i - bpp] : 0;
                    row[i] = (uint8_t)(row[i] + paeth(a, b, c));
                }
                break;
            default: return false;
        }
        prior = row;
        p += rb + 1;
    }
    return true;
}

// Sample `i` of a packed row at `bits` per sample (MSB first below 8 bits,
// big-endian at 16).
inline uint32_t packed_sample(const uint8_t* row, size_t i, int bits) {
    switch (bits) {
        case 8: return row[i];
        case 16: return be16(row + 2 * i);
        default: {
            const size_t bit = i * (size_t)bits;
            const int shift = 8 - bits - (int)(bit & 7);
            return (uint32_t)(row[bit >> 3] >> shift) & ((1u << bits) - 1u);
        }
    }
}

Raster decode_png(const uint8_t* d, size_t len) {
    if (len < 8 || std::memcmp(d, "\x89PNG\r\n\x1a\n", 8) != 0) return {};
    size_t pos = 8;
    uint32_t w = 0, h = 0;
    int bd = 0, ct = 0, il = 0;
    bool have_ihdr = false, have_iend = false;
    std::vector<uint8_t> idat;
    uint8_t plte[256 * 3];
    std::memset(plte, 0, sizeof plte);
    int plte_n = 0;
    while (pos + 12 <= len) {
        const uint32_t clen = be32(d + pos);
        const uint8_t* type = d + pos + 4;
        if (clen > 0x7fffffffu || (uint64_t)pos + 12 + clen > len) return {};
        const uint8_t* data = d + pos + 8;
        const bool critical = (type[0] & 0x20) == 0;
        if (critical) {
            uLong crc = crc32(0L, Z_NULL, 0);
            crc = crc32(crc, type, (uInt)(4 + clen));
            if ((uint32_t)crc != be32(data + clen)) return {};
        }
        const bool is_ihdr = std::memcmp(type, "IHDR", 4) == 0;
        if (!have_ihdr && !is_ihdr) return {};
        if (is_ihdr) {
            if (have_ihdr || clen != 13) return {};
            w = be32(data);
            h = be32(data + 4);
            bd = data[8];
            ct = data[9];
            il = data[12];
            if (data[10] != 0 || data[11] != 0 || il > 1) return {};
            if (w == 0 || h == 0 || w > 0x7fffffffu || h > 0x7fffffffu) return {};
            if (w > 1000000u || h > 1000000u) return {};
            if (!dims_ok(w, h)) return {};
            bool valid = false;
            switch (ct) {
                case 0: valid = bd == 1 || bd == 2 || bd == 4 || bd == 8 || bd == 16; break;
                case 3: valid = bd == 1 || bd == 2 || bd == 4 || bd == 8; break;
                case 2:
                case 4:
                case 6: valid = bd == 8 || bd == 16; break;
                default: break;
            }
            if (!valid) return {};
            have_ihdr = true;
        } else if (std::memcmp(type, "PLTE", 4) == 0) {
            if (clen % 3 != 0 || clen / 3 > 256 || clen == 0) return {};
            plte_n = (int)(clen / 3);
            std::memcpy(plte, data, clen);
        } else if (std::memcmp(type, "IDAT", 4) == 0) {
            idat.insert(idat.end(), data, data + clen);
        } else if (std::memcmp(type, "IEND", 4) == 0) {
            have_iend = true;
            break;
        } else if (critical) {
            return {};   // an unknown critical chunk
        }
        pos += 12 + (size_t)clen;
    }
    if (!have_ihdr || !have_iend || idat.empty()) return {};
    if (ct == 3 && plte_n == 0) return {};

    const int chans = ct == 0 ? 1 : ct == 2 ? 3 : ct == 3 ? 1 : ct == 4 ? 2 : 4;
    const int bits_px = chans * bd;
    const size_t bpp = (size_t)std::max(1, bits_px / 8);
    static const int x0[7] = {0, 4, 0, 2, 0, 1, 0}, y0[7] = {0, 0, 4, 0, 2, 0, 1};
    static const int dx[7] = {8, 8, 4, 4, 2, 2, 1}, dy[7] = {8, 8, 8, 4, 4, 2, 2};
    const int npass = il ? 7 : 1;
    size_t pw[7], ph[7], prb[7], total = 0;
    for (int p = 0; p < npass; ++p) {
        if (il) {
            pw[p] = w > (uint32_t)x0[p] ? (w - x0[p] + dx[p] - 1) / dx[p] : 0;
            ph[p] = h > (uint32_t)y0[p] ? (h - y0[p] + dy[p] - 1) / dy[p] : 0;
        } else {
            pw[p] = w;
            ph[p] = h;
        }
        prb[p] = (pw[p] * (size_t)bits_px + 7) / 8;
        if (pw[p] && ph[p]) total += ph[p] * (1 + prb[p]);
    }
    std::vector<uint8_t> raw(total);
    if (!inflate_to(idat.data(), idat.size(), raw.data(), total)) return {};
    idat.clear();
    idat.shrink_to_fit();

    // gray and gray+alpha come out as one channel (an RGB collapse of equal
    // channels is the gray value itself), RGB / RGBA / palette as three.
    const bool color = ct == 2 || ct == 3 || ct == 6;
    const Depth depth = bd == 16 ? Depth::U16 : Depth::U8;
    Raster r = make_raster((int)w, (int)h, color ? 3 : 1, depth);
    const uint32_t gmax = (1u << (bd < 8 ? bd : 8)) - 1u;
    size_t off = 0;
    for (int p = 0; p < npass; ++p) {
        if (!pw[p] || !ph[p]) continue;
        uint8_t* pass = raw.data() + off;
        if (!png_unfilter(pass, ph[p], prb[p], bpp)) return {};
        for (size_t py = 0; py < ph[p]; ++py) {
            const uint8_t* row = pass + py * (1 + prb[p]) + 1;
            const size_t y = il ? (size_t)y0[p] + py * dy[p] : py;
            for (size_t px = 0; px < pw[p]; ++px) {
                const size_t x = il ? (size_t)x0[p] + px * dx[p] : px;
                const size_t o = (y * w + x) * (size_t)r.cn;
                uint32_t v[3];
                if (ct == 3) {
                    const uint32_t idx = packed_sample(row, px, bd);
                    const uint8_t* e = plte + 3 * idx;   // zero past the palette
                    v[0] = e[0];
                    v[1] = e[1];
                    v[2] = e[2];
                } else if (ct == 0 || ct == 4) {
                    uint32_t g = packed_sample(row, px * (size_t)chans, bd);
                    if (bd < 8) g = g * 255u / gmax;
                    v[0] = g;
                } else {
                    for (int c = 0; c < 3; ++c) v[c] = packed_sample(row, px * (size_t)chans + c, bd);
                }
                if (depth == Depth::U8) {
                    for (int c = 0; c < r.cn; ++c) r.px[o + c] = (uint8_t)v[c];
                } else {
                    uint16_t* q = reinterpret_cast<uint16_t*>(r.px.data()) + o;
                    for (int c = 0; c < r.cn; ++c) q[c] = (uint16_t)v[c];
                }
            }
        }
        off += ph[p] * (1 + prb[p]);
    }
    return r;
}

// ---------------------------------------------------------------------------
// TIFF
// ---------------------------------------------------------------------------

struct TiffReader {
    const uint8_t* d;
    size_t len;
    bool le;
    uint32_t u16(size_t o) const { return le ? le16(d + o) : be16(d + o); }
    uint32_t u32(size_t o) const { return le ? le32(d + o) : be32(d + o); }
};

size_t tiff_type_size(uint32_t t) {
    switch (t) {
        case 1: case 2: case 6: case 7: return 1;
        case 3: case 8: return 2;
        case 4: case 9: case 11: case 13: return 4;
        case 5: case 10: case 12: return 8;
        default: return 0;
    }
}

// The unsigned integer values of an IFD entry (BYTE, SHORT, LONG).
bool tiff_uints(const TiffReader& t, size_t entry, std::vector<uint64_t>& out) {
    const uint32_t type = t.u16(entry + 2), count = t.u32(entry + 4);
    const size_t sz = tiff_type_size(type);
    if (sz == 0 || (type != 1 && type != 3 && type != 4)) return false;
    if ((uint64_t)count * sz > t.len) return false;
    const size_t bytes = (size_t)count * sz;
    const size_t at = bytes <= 4 ? entry + 8 : (size_t)t.u32(entry + 8);
    if ((uint64_t)at + bytes > t.len) return false;
    out.resize(count);
    for (uint32_t i = 0; i < count; ++i) {
        const size_t o = at + (size_t)i * sz;
        out[i] = type == 1 ? t.d[o] : type == 3 ? t.u16(o) : t.u32(o);
    }
    return true;
}

uint8_t reverse_bits(uint8_t b) {
    b = (uint8_t)((b & 0xF0) >> 4 | (b & 0x0F) << 4);
    b = (uint8_t)((b & 0xCC) >> 2 | (b & 0x33) << 2);
    return (uint8_t)((b & 0xAA) >> 1 | (b & 0x55) << 1);
}

// TIFF 6.0 LZW (MSB-first codes of 9-12 bits, the width growing one code
// early, 256 = clear, 257 = end of information) into exactly `want` bytes.
bool lzw_decode(const uint8_t* src, size_t n, uint8_t* dst, size_t want) {
    if (n >= 2 && src[0] == 0 && (src[1] & 1)) return false;   // the pre-6.0 bit order
    struct Entry {
        uint16_t prefix;
        uint16_t len;
        uint8_t suffix, first;
    };
    std::vector<Entry> tab(4096);
    for (int i = 0; i < 256; ++i) tab[i] = Entry{0xFFFF, 1, (uint8_t)i, (uint8_t)i};
    size_t out = 0, bitpos = 0;
    const size_t nbits_total = n * 8;
    int width = 9, next = 258, prev = -1;
    auto emit = [&](int code) {
        const size_t l = tab[code].len;
        size_t end = out + l;
        int c = code;
        for (size_t k = end; k-- > out;) {
            if (k < want) dst[k] = tab[c].suffix;
            c = tab[c].prefix;
        }
        out = end;
    };
    while (out < want) {
        if (bitpos + (size_t)width > nbits_total) break;
        uint32_t code = 0;
        for (int b = 0; b < width; ++b, ++bitpos)
            code = code << 1 | ((src[bitpos >> 3] >> (7 - (bitpos & 7))) & 1u);
        if (code == 257) break;
        if (code == 256) {
            width = 9;
            next = 258;
            prev = -1;
            continue;
        }
        if (prev < 0) {
            if (code > 255) return false;
            emit((int)code);
            prev = (int)code;
            continue;
        }
        uint8_t first;
        if ((int)code < next) {
            first = tab[code].first;
            emit((int)code);
        } else if ((int)code == next && next < 4096) {
            // the code being defined: prev's string, then its first byte
            first = tab[prev].first;
            emit(prev);
            if (out < want) dst[out] = first;
            out += 1;
        } else {
            return false;
        }
        if (next < 4096) {
            tab[next] = Entry{(uint16_t)prev, (uint16_t)(tab[prev].len + 1), first, tab[prev].first};
            ++next;
            if (next >= (1 << width) - 1 && width < 12) ++width;
        }
        prev = (int)code;
    }
    return out >= want;
}

// PackBits (TIFF compression 32773) into exactly `want` bytes.
bool packbits_decode(const uint8_t* src, size_t n, uint8_t* dst, size_t want) {
    size_t in = 0, out = 0;
    while (out < want && in < n) {
        const int c = (int8_t)src[in++];
        if (c >= 0) {
            size_t cnt = (size_t)c + 1;
            cnt = std::min(cnt, n - in);
            const size_t take = std::min(cnt, want - out);
            std::memcpy(dst + out, src + in, take);
            in += cnt;
            out += take;
        } else if (c != -128) {
            if (in >= n) break;
            const size_t cnt = std::min<size_t>((size_t)(1 - c), want - out);
            std::memset(dst + out, src[in++], cnt);
            out += cnt;
        }
    }
    return out == want;
}

Raster decode_tiff(const uint8_t* d, size_t len) {
    if (len < 8) return {};
    const bool le = d[0] == 'I';
    if (!((le && std::memcmp(d, "II*\0", 4) == 0) || (!le && std::memcmp(d, "MM\0*", 4) == 0))) return {};
    TiffReader t{d, len, le};
    const size_t ifd = t.u32(4);
    if ((uint64_t)ifd + 2 > len) return {};
    const size_t nent = t.u16(ifd);
    if ((uint64_t)ifd + 2 + 12 * (uint64_t)nent > len) return {};

    uint64_t w = 0, h = 0, rps = 0, tw = 0, tl = 0;
    uint32_t comp = 1, photo = 0xFFFF, planar = 1, pred = 1, fill = 1;
    bool have_w = false, have_h = false, tiled = false, have_counts = false;
    std::vector<uint64_t> bps_v{1}, spp_v{1}, fmt_v{1}, extra_v, offs, counts, cmap;
    std::vector<uint64_t> tmp;
    for (size_t k = 0; k < nent; ++k) {
        const size_t e = ifd + 2 + 12 * k;
        const uint32_t tag = t.u16(e);
        auto one = [&](uint64_t& v) {
            if (!tiff_uints(t, e, tmp) || tmp.empty()) return false;
            v = tmp[0];
            return true;
        };
        uint64_t v = 0;
        switch (tag) {
            case 256: if (!one(w)) return {}; have_w = true; break;
            case 257: if (!one(h)) return {}; have_h = true; break;
            case 258: if (!tiff_uints(t, e, bps_v) || bps_v.empty()) return {}; break;
            case 259: if (!one(v)) return {}; comp = (uint32_t)v; break;
            case 262: if (!one(v)) return {}; photo = (uint32_t)v; break;
            case 266: if (!one(v)) return {}; fill = (uint32_t)v; break;
            case 273: if (!tiff_uints(t, e, offs)) return {}; break;
            case 277: if (!tiff_uints(t, e, spp_v) || spp_v.empty()) return {}; break;
            case 278: if (!one(rps)) return {}; break;
            case 279: if (!tiff_uints(t, e, counts)) return {}; have_counts = true; break;
            case 284: if (!one(v)) return {}; planar = (uint32_t)v; break;
            case 317: if (!one(v)) return {}; pred = (uint32_t)v; break;
            case 320: if (!tiff_uints(t, e, cmap)) return {}; break;
            case 322: if (!one(tw)) return {}; tiled = true; break;
            case 323: if (!one(tl)) return {}; tiled = true; break;
            case 324: if (!tiff_uints(t, e, offs)) return {}; break;
            case 325: if (!tiff_uints(t, e, counts)) return {}; have_counts = true; break;
            case 338: if (!tiff_uints(t, e, extra_v)) return {}; break;
            case 339: if (!tiff_uints(t, e, fmt_v) || fmt_v.empty()) return {}; break;
            default: break;
        }
    }
    if (!have_w || !have_h || !dims_ok(w, h)) return {};
    const uint32_t spp = (uint32_t)spp_v[0];
    if (spp < 1 || spp > 16) return {};
    for (uint64_t b : bps_v)
        if (b != bps_v[0]) return {};
    for (uint64_t f : fmt_v)
        if (f != fmt_v[0]) return {};
    const uint32_t bps = (uint32_t)bps_v[0], fmt = (uint32_t)fmt_v[0];
    if (photo == 0xFFFF) photo = spp >= 3 ? 2 : 1;
    if (planar != 1 && planar != 2) return {};
    if (fill != 1 && fill != 2) return {};
    if (comp != 1 && comp != 5 && comp != 8 && comp != 32946 && comp != 32773) return {};
    if (pred != 1 && pred != 2 && pred != 3) return {};
    if (!(bps == 1 || bps == 2 || bps == 4 || bps == 8 || bps == 16 || bps == 32 || bps == 64)) return {};

    // What the samples become (the reference decoder stack's type choice).
    enum class Route { Rgba8, Gray16, Color16, Gray32, Color32 } route;
    Depth depth = Depth::U8;
    if (bps <= 8) {
        if (fmt != 1) return {};
        if (photo == 0 || photo == 1) {
            if (planar == 1 && spp != 1 && bps < 8) return {};
        } else if (photo == 3) {
            if (spp != 1 || cmap.size() != 3 * (size_t(1) << bps)) return {};
        } else if (photo == 2) {
            if (bps != 8 || spp < 3) return {};
        } else {
            return {};
        }
        route = Route::Rgba8;
    } else if (bps == 16) {
        if (fmt != 1 && fmt != 2) return {};
        depth = fmt == 2 ? Depth::S16 : Depth::U16;
        if (photo == 1 && spp == 1) route = Route::Gray16;
        else if (photo == 2 && (spp == 3 || spp == 4)) route = Route::Color16;
        else return {};
    } else {
        if (bps == 32 && fmt == 3) depth = Depth::F32;
        else if (bps == 32 && fmt == 2) depth = Depth::S32;
        else if (bps == 64 && fmt == 3) depth = Depth::F64;
        else return {};
        if (photo == 1 && spp == 1) route = Route::Gray32;
        else if (photo == 2 && (spp == 3 || spp == 4)) route = Route::Color32;
        else return {};
    }
    if (pred == 2 && bps < 8) return {};
    if (pred == 3 && fmt != 3) return {};

    // Chunk geometry.
    const uint32_t nplanes = planar == 2 ? spp : 1, cspp = planar == 2 ? 1 : spp;
    uint64_t cw, ch, across, down;
    if (tiled) {
        if (tw == 0 || tl == 0 || tw > (1u << 24) || tl > (1u << 24)) return {};
        cw = tw;
        ch = tl;
        across = (w + tw - 1) / tw;
        down = (h + tl - 1) / tl;
    } else {
        if (rps == 0 || rps > h) rps = h;
        cw = w;
        ch = rps;
        across = 1;
        down = (h + rps - 1) / rps;
    }
    const uint64_t nchunks = across * down * nplanes;
    if (offs.size() < nchunks) return {};
    const size_t rowbytes = (size_t)((cw * cspp * bps + 7) / 8);
    if (!have_counts) {
        if (comp != 1 || nchunks != 1) return {};
        counts.assign(1, (uint64_t)rowbytes * ch);
    }
    if (counts.size() < nchunks) return {};

    // Unpacked samples: one byte each below 16 bits, else bps / 8 bytes in
    // host order.
    const size_t sbytes = bps <= 8 ? 1 : bps / 8;
    std::vector<uint8_t> full((size_t)w * h * spp * sbytes);
    std::vector<uint8_t> chunk, rawc;
    for (uint64_t ci = 0; ci < nchunks; ++ci) {
        const uint64_t plane = ci / (across * down), rest = ci % (across * down);
        const uint64_t cy = rest / across, cx = rest % across;
        const uint64_t rows = tiled ? ch : std::min<uint64_t>(ch, h - cy * ch);
        const size_t want = rowbytes * (size_t)rows;
        const uint64_t off = offs[ci], cnt = counts[ci];
        if (off > len || cnt > len - off) return {};
        const uint8_t* src = d + off;
        if (fill == 2) {
            rawc.assign(src, src + cnt);
            for (auto& b : rawc) b = reverse_bits(b);
            src = rawc.data();
        }
        // Every codec here must deliver the whole chunk, and none can expand
        // its input beyond a fixed ratio: PackBits 64 (a 2-byte run of 128),
        // LZW about 3400 (9-bit codes whose strings grow by one byte per
        // code, up to a 4094-entry table), Deflate 1032. A chunk that
        // declares more than that is refused before it is allocated.
        const uint64_t most = comp == 1          ? cnt
                              : comp == 32773    ? cnt * 64 + 128
                              : comp == 5        ? cnt * 3641 + 4096
                              : (comp == 8 || comp == 32946) ? cnt * 1032 + 1024
                                                             : 0;
        if ((uint64_t)want > most) return {};
        chunk.assign(want, 0);
        bool ok = false;
        switch (comp) {
            case 1:
                ok = cnt >= want;
                if (ok) std::memcpy(chunk.data(), src, want);
                break;
            case 5: ok = lzw_decode(src, (size_t)cnt, chunk.data(), want); break;
            case 8:
            case 32946: ok = inflate_to(src, (size_t)cnt, chunk.data(), want); break;
            case 32773: ok = packbits_decode(src, (size_t)cnt, chunk.data(), want); break;
            default: break;
        }
        if (!ok) return {};
        // Byte order and predictor, row by row.
        const size_t ns = (size_t)cw * cspp;   // samples per chunk row
        for (uint64_t r = 0; r < rows; ++r) {
            uint8_t* row = chunk.data() + r * rowbytes;
            if (pred == 3) {
                const size_t bb = bps / 8, stride = cspp;
                for (size_t i = stride; i < rowbytes; ++i) row[i] = (uint8_t)(row[i] + row[i - stride]);
                std::vector<uint8_t> tmpb(row, row + rowbytes);
                for (size_t i = 0; i < ns; ++i)
                    for (size_t b = 0; b < bb; ++b) row[bb * i + b] = tmpb[(bb - b - 1) * ns + i];   // host is little-endian
                continue;
            }
            if (bps >= 16 && !le) {
                const size_t bb = bps / 8;
                for (size_t i = 0; i < ns; ++i) std::reverse(row + bb * i, row + bb * (i + 1));
            }
            if (pred == 2) {
                switch (bps) {
                    case 8:
                        for (size_t i = cspp; i < ns; ++i) row[i] = (uint8_t)(row[i] + row[i - cspp]);
                        break;
                    case 16: {
                        uint16_t* q = reinterpret_cast<uint16_t*>(row);
                        for (size_t i = cspp; i < ns; ++i) q[i] = (uint16_t)(q[i] + q[i - cspp]);
                        break;
                    }
                    case 32: {
                        uint32_t* q = reinterpret_cast<uint32_t*>(row);
                        for (size_t i = cspp; i < ns; ++i) q[i] = q[i] + q[i - cspp];
                        break;
                    }
                    case 64: {
                        uint64_t* q = reinterpret_cast<uint64_t*>(row);
                        for (size_t i = cspp; i < ns; ++i) q[i] = q[i] + q[i - cspp];
                        break;
                    }
                    default: return {};
                }
            }
        }
        // Scatter into the full sample image.
        const uint64_t ox = cx * cw, oy = cy * ch;
        const uint64_t vw = std::min<uint64_t>(cw, w - ox), vh = std::min<uint64_t>(rows, h - oy);
        for (uint64_t r = 0; r < vh; ++r) {
            const uint8_t* row = chunk.data() + r * rowbytes;
            uint8_t* dst = full.data() + (((oy + r) * w + ox) * spp) * sbytes;
            if (bps < 8) {
                for (uint64_t x = 0; x < vw; ++x)
                    for (uint32_t s = 0; s < cspp; ++s)
                        dst[(x * spp + plane + s)] = (uint8_t)packed_sample(row, (size_t)(x * cspp + s), (int)bps);
            } else if (planar == 1) {
                std::memcpy(dst, row, (size_t)vw * spp * sbytes);
            } else {
                for (uint64_t x = 0; x < vw; ++x)
                    std::memcpy(dst + (x * spp + plane) * sbytes, row + x * sbytes, sbytes);
            }
        }
    }

    const int W = (int)w, H = (int)h;
    const size_t npx = (size_t)w * h;
    switch (route) {
        case Route::Rgba8: {
            const bool gray = photo == 0 || photo == 1;
            Raster r = make_raster(W, H, gray ? 1 : 3, Depth::U8);
            if (gray) {
                const uint32_t range = (1u << bps) - 1u;
                uint8_t map[256];
                for (uint32_t v = 0; v <= range; ++v)
                    map[v] = (uint8_t)(photo == 0 ? (range - v) * 255u / range : v * 255u / range);
                for (size_t i = 0; i < npx; ++i) r.px[i] = map[full[i * spp]];
            } else if (photo == 3) {
                const size_t nc = size_t(1) << bps;
                bool wide = false;
                for (uint64_t c : cmap) wide = wide || c >= 256;
                uint8_t pal[256 * 3];
                for (size_t i = 0; i < nc; ++i)
                    for (int c = 0; c < 3; ++c) {
                        const uint64_t e = cmap[c * nc + i];
                        pal[3 * i + c] = (uint8_t)(wide ? e >> 8 : e);
                    }
                for (size_t i = 0; i < npx; ++i) std::memcpy(&r.px[3 * i], pal + 3 * full[i], 3);
            } else {
                // RGB: a fourth sample with no ExtraSamples is associated
                // alpha; unassociated alpha is premultiplied.
                const bool unassoc = spp >= 4 && !extra_v.empty() && extra_v[0] == 2;
                for (size_t i = 0; i < npx; ++i) {
                    const uint8_t* s = &full[i * spp];
                    for (int c = 0; c < 3; ++c)
                        r.px[3 * i + c] = unassoc ? (uint8_t)((s[c] * (uint32_t)s[3] + 127u) / 255u) : s[c];
                }
            }
            return r;
        }
        case Route::Gray16:
        case Route::Gray32: {
            Raster r = make_raster(W, H, 1, depth);
            std::memcpy(r.px.data(), full.data(), r.px.size());
            return r;
        }
        case Route::Color16:
        case Route::Color32: {
            Raster r = make_raster(W, H, 3, depth);
            const size_t sb = depth_bytes(depth);
            for (size_t i = 0; i < npx; ++i) std::memcpy(&r.px[3 * i * sb], &full[i * spp * sb], 3 * sb);
            return r;
        }
    }
    return {};
}

// ---------------------------------------------------------------------------
// BMP
// ---------------------------------------------------------------------------

Raster decode_bmp(const uint8_t* d, size_t len) {
    if (len < 26 || d[0] != 'B' || d[1] != 'M') return {};
    const size_t data_off = le32(d + 10);
    const uint32_t hsize = le32(d + 14);
    int64_t w = 0, h = 0;
    uint32_t bpp = 0, comp = 0, clrused = 0;
    size_t pal_off = 0, pal_entry = 4, pal_n = 0;
    if (hsize >= 36) {
        if (len < 14 + 36) return {};
        w = (int32_t)le32(d + 18);
        h = (int32_t)le32(d + 22);
        bpp = le32(d + 26) >> 16;
        comp = le32(d + 30);
        clrused = le32(d + 46);
        pal_off = 14 + (size_t)hsize;
    } else if (hsize == 12) {
        w = le16(d + 18);
        h = (int16_t)le16(d + 20);
        bpp = le32(d + 22) >> 16;
        comp = 0;
        pal_off = 26;
        pal_entry = 3;
    } else {
        return {};
    }
    if (w <= 0 || h == 0) return {};
    const bool top_down = h < 0;
    h = std::llabs(h);
    if (!dims_ok((uint64_t)w, (uint64_t)h)) return {};
    const bool paletted = bpp == 1 || bpp == 4 || bpp == 8;
    if (comp == 0) {
        if (!(paletted || bpp == 16 || bpp == 24 || bpp == 32)) return {};
    } else if (comp == 3) {
        if (!(bpp == 16 || bpp == 32)) return {};
    } else {
        return {};   // run-length encoded
    }
    int b16 = 0;   // 15: 5-5-5, 16: 5-6-5
    if (bpp == 16) {
        if (comp == 0) {
            b16 = 15;
        } else {
            const size_t mo = hsize >= 52 ? 14 + 40 : pal_off;
            if (mo + 12 > len) return {};
            const uint32_t rm = le32(d + mo), gm = le32(d + mo + 4), bm = le32(d + mo + 8);
            if (bm == 0x1f && gm == 0x3e0 && rm == 0x7c00) b16 = 15;
            else if (bm == 0x1f && gm == 0x7e0 && rm == 0xf800) b16 = 16;
            else return {};
        }
    }
    uint8_t pal[256 * 3];   // R, G, B; zero past the entries read
    std::memset(pal, 0, sizeof pal);
    if (paletted) {
        if (clrused > 256) return {};
        pal_n = (hsize == 12 || clrused == 0) ? (size_t(1) << bpp) : clrused;
        if (pal_off + pal_n * pal_entry > len) return {};
        for (size_t i = 0; i < pal_n; ++i) {
            const uint8_t* e = d + pal_off + i * pal_entry;
            pal[3 * i] = e[2];
            pal[3 * i + 1] = e[1];
            pal[3 * i + 2] = e[0];
        }
    }
    const size_t pitch = (((size_t)w * (bpp == 15 ? 16 : bpp) + 7) / 8 + 3) & ~size_t(3);
    if (data_off > len || (uint64_t)pitch * (uint64_t)h > len - data_off) return {};
    Raster r = make_raster((int)w, (int)h, 3, Depth::U8);
    for (int64_t y = 0; y < h; ++y) {
        const uint8_t* row = d + data_off + (size_t)y * pitch;
        const int64_t oy = top_down ? y : h - 1 - y;
        uint8_t* dst = r.px.data() + (size_t)oy * w * 3;
        for (int64_t x = 0; x < w; ++x) {
            uint8_t* q = dst + 3 * x;
            if (paletted) {
                const uint32_t idx = packed_sample(row, (size_t)x, (int)bpp);
                std::memcpy(q, pal + 3 * idx, 3);
            } else if (bpp == 16) {
                const uint32_t v = le16(row + 2 * x);
                const uint8_t b = (uint8_t)(v << 3);
                const uint8_t g = b16 == 15 ? (uint8_t)((v >> 2) & ~7u) : (uint8_t)((v >> 3) & ~3u);
                const uint8_t rr = b16 == 15 ? (uint8_t)((v >> 7) & ~7u) : (uint8_t)((v >> 8) & ~7u);
                q[0] = rr;
                q[1] = g;
                q[2] = b;
            } else {
                const uint8_t* s = row + x * (bpp / 8);
                q[0] = s[2];
                q[1] = s[1];
                q[2] = s[0];
            }
        }
    }
    return r;
}

// ---------------------------------------------------------------------------
// PNM (P1-P6)
// ---------------------------------------------------------------------------

struct PnmReader {
    const uint8_t* d;
    size_t len, pos = 0;
    bool get(int& c) {
        if (pos >= len) return false;
        c = d[pos++];
        return true;
    }
    // A decimal number after whitespace and '#' comments; consumes the one
    // byte that ends it. maxdigits > 0 stops after that many digits.
    bool number(int64_t& v, int maxdigits = 0) {
        int c;
        if (!get(c)) return false;
        while (!(c >= '0' && c <= '9')) {
            if (c == '#') {
                do {
                    if (!get(c)) return false;
                } while (c != '\n' && c != '\r');
                if (!get(c)) return false;
            } else if (std::isspace(c)) {
                while (std::isspace(c))
                    if (!get(c)) return false;
            } else {
                return false;
            }
        }
        v = 0;
        int digits = 0;
        for (;;) {
            v = v * 10 + (c - '0');
            if (v > std::numeric_limits<int32_t>::max()) return false;
            ++digits;
            if (maxdigits != 0 && digits >= maxdigits) break;
            if (!get(c)) {
                c = -1;
                break;
            }
            if (!(c >= '0' && c <= '9')) break;
        }
        return true;
    }
};

Raster decode_pnm(const uint8_t* d, size_t len) {
    if (len < 3 || d[0] != 'P' || d[1] < '1' || d[1] > '6' || !std::isspace(d[2])) return {};
    const int kind = d[1] - '0';
    const int bits = (kind == 1 || kind == 4) ? 1 : (kind == 2 || kind == 5) ? 8 : 24;
    const bool binary = kind >= 4;
    PnmReader rd{d, len, 2};
    int64_t w = 0, h = 0, maxval = 1;
    if (!rd.number(w) || !rd.number(h)) return {};
    if (bits > 1 && !rd.number(maxval)) return {};
    if (maxval > 65535 || maxval <= 0 || w <= 0 || h <= 0) return {};
    if (!dims_ok((uint64_t)w, (uint64_t)h)) return {};
    const bool wide = maxval > 255;
    const int cn = bits == 24 ? 3 : 1;
    Raster r = make_raster((int)w, (int)h, cn, wide ? Depth::U16 : Depth::U8);
    uint8_t lut[256];
    std::memset(lut, 0, sizeof lut);
    if (!wide)
        for (int64_t i = 0; i <= maxval; ++i) lut[i] = (uint8_t)((i * 255 / maxval) ^ (bits == 1 ? 255 : 0));
    const size_t nsamp = (size_t)w * h * cn;
    if (!binary) {
        for (size_t i = 0; i < nsamp; ++i) {
            int64_t v = 0;
            if (!rd.number(v, bits == 1 ? 1 : 0)) return {};
            if (bits == 1) v = v != 0;
            if (v > maxval) v = maxval;
            if (wide) reinterpret_cast<uint16_t*>(r.px.data())[i] = (uint16_t)v;
            else r.px[i] = lut[v];
        }
        return r;
    }
    const uint8_t* p = d + rd.pos;
    const size_t avail = len - rd.pos;
    if (bits == 1) {
        const size_t pitch = ((size_t)w + 7) / 8;
        if (pitch * (size_t)h > avail) return {};
        for (int64_t y = 0; y < h; ++y)
            for (int64_t x = 0; x < w; ++x)
                r.px[(size_t)y * w + x] = lut[packed_sample(p + (size_t)y * pitch, (size_t)x, 1)];
        return r;
    }
    const size_t sb = wide ? 2 : 1;
    if (nsamp * sb > avail) return {};
    if (wide) {
        uint16_t* q = reinterpret_cast<uint16_t*>(r.px.data());
        for (size_t i = 0; i < nsamp; ++i) q[i] = (uint16_t)be16(p + 2 * i);
    } else {
        for (size_t i = 0; i < nsamp; ++i) r.px[i] = maxval < 255 ? lut[p[i]] : p[i];
    }
    return r;
}

// ---------------------------------------------------------------------------
// Depth conversion
// ---------------------------------------------------------------------------

template <typename T>
const T* as(const Raster& r) {
    return reinterpret_cast<const T*>(r.px.data());
}

// One channel of any depth, as doubles for the non-8/16-bit types
// (to_one_channel, image_codec.cpp:164-175): 8U and 16U collapse in fixed
// point; float collapses in float with 0.299 / 0.587 / 0.114; the other
// types are widened to float first.
template <typename T>
float fetch_f(const Raster& r, size_t i) {
    return (float)as<T>(r)[i];
}

float luma_float(float rr, float g, float b) { return b * 0.114f + g * 0.587f + rr * 0.299f; }

// 0..255 from a value the way cv::saturate_cast<uchar> of a rounded float
// does: round half to even; NaN and values beyond int become 0.
inline uint8_t sat_round(double v) {
    if (!(v > -2147483648.0 && v < 2147483648.0)) return 0;
    const double r = std::nearbyint(v);
    return (uint8_t)(r < 0 ? 0 : r > 255 ? 255 : r);
}

}  // namespace

namespace {

// The containers the reference's imdecode route recognises, told apart by
// their leading magic bytes (PNG spec 5.2; JPEG T.81 B.1.1.2 SOI followed by
// a marker; TIFF 6.0 section 2 byte-order mark and 42; BMP file header "BM";
// RIFF/WEBP; Netpbm "P1".."P6").
enum class Container { None, Png, Jpeg, Tiff, Bmp, Webp, Pnm };

constexpr uint8_t kPngMagic[8] = {0x89, 'P', 'N', 'G', 0x0D, 0x0A, 0x1A, 0x0A};

bool prefix_is(const uint8_t* d, size_t len, size_t at, const uint8_t* bytes, size_t n) {
    if (len < at + n) return false;
    for (size_t k = 0; k < n; ++k)
        if (d[at + k] != bytes[k]) return false;
    return true;
}

Container sniff(const uint8_t* d, size_t len) {
    if (d == nullptr) return Container::None;
    static const uint8_t soi[3] = {0xFF, 0xD8, 0xFF};
    static const uint8_t tiff_le[4] = {'I', 'I', 42, 0}, tiff_be[4] = {'M', 'M', 0, 42};
    static const uint8_t bm[2] = {'B', 'M'}, riff[4] = {'R', 'I', 'F', 'F'}, webp[4] = {'W', 'E', 'B', 'P'};
    if (prefix_is(d, len, 0, kPngMagic, 8)) return Container::Png;
    if (prefix_is(d, len, 0, soi, 3)) return Container::Jpeg;
    if (prefix_is(d, len, 0, tiff_le, 4) || prefix_is(d, len, 0, tiff_be, 4)) return Container::Tiff;
    if (prefix_is(d, len, 0, bm, 2)) return Container::Bmp;
    if (prefix_is(d, len, 0, riff, 4) && prefix_is(d, len, 8, webp, 4)) return Container::Webp;
    if (len >= 2 && d[0] == 'P' && d[1] >= '1' && d[1] <= '6') return Container::Pnm;
    return Container::None;
}

// PNG: the first chunk is IHDR, whose data opens with the width and height as
// big-endian 32-bit integers (PNG spec 11.2.2).
bool png_size(const uint8_t* d, size_t len, uint64_t& w, uint64_t& h) {
    static const uint8_t ihdr[4] = {'I', 'H', 'D', 'R'};
    if (len < 24 || !prefix_is(d, len, 12, ihdr, 4)) return false;
    w = be32(d + 16);
    h = be32(d + 20);
    return true;
}

// BMP: after the 14-byte file header, the info header's signed 32-bit width
// and height (a negative height is a top-down bitmap).
bool bmp_size(const uint8_t* d, size_t len, uint64_t& w, uint64_t& h) {
    if (len < 26) return false;
    const int64_t bw = (int32_t)le32(d + 18), bh = (int32_t)le32(d + 22);
    w = (uint64_t)(bw < 0 ? -bw : bw);
    h = (uint64_t)(bh < 0 ? -bh : bh);
    return true;
}

// JPEG (T.81 annex B): marker segments after SOI up to the first frame
// header. Fill bytes 0xFF may precede a marker; SOI, TEM and RST0..7 stand
// alone; every other marker carries a 16-bit length that counts itself. A
// frame header (SOF0..15 other than DHT C4, JPG C8 and DAC CC) holds the
// sample precision, then Y (lines) and X (samples per line). A scan or the
// end of the image before any frame header leaves the size unknown.
bool jpeg_size(const uint8_t* d, size_t len, uint64_t& w, uint64_t& h) {
    auto standalone = [](uint8_t m) { return m == 0xD8 || m == 0x01 || (m >= 0xD0 && m <= 0xD7); };
    auto frame = [](uint8_t m) { return m >= 0xC0 && m <= 0xCF && m != 0xC4 && m != 0xC8 && m != 0xCC; };
    size_t at = 2;   // past SOI
    for (;;) {
        if (at + 4 > len || d[at] != 0xFF) return false;
        const uint8_t code = d[at + 1];
        if (code == 0xFF) {   // a fill byte before the marker
            at += 1;
            continue;
        }
        if (standalone(code)) {
            at += 2;
            continue;
        }
        if (code == 0xDA || code == 0xD9) return false;
        if (frame(code)) {
            if (at + 9 > len) return false;
            h = be16(d + at + 5);
            w = be16(d + at + 7);
            return true;
        }
        const size_t seglen = be16(d + at + 2);
        if (seglen < 2) return false;
        at += 2 + seglen;
    }
}

// TIFF (6.0, section 2): the first IFD's ImageWidth (256) and ImageLength
// (257) entries, of type SHORT (3) or LONG (4); the value sits left-justified
// in the entry's last four bytes in either byte order.
bool tiff_size(const uint8_t* d, size_t len, uint64_t& w, uint64_t& h) {
    if (len < 8) return false;
    const TiffReader t{d, len, d[0] == 'I'};
    const uint64_t ifd = t.u32(4);
    if (ifd + 2 > len) return false;
    const uint64_t entries = t.u16((size_t)ifd);
    int found = 0;   // bit 0: width, bit 1: height
    for (uint64_t k = 0; k < entries; ++k) {
        const uint64_t at = ifd + 2 + 12 * k;
        if (at + 12 > len) return false;
        const uint32_t tag = t.u16((size_t)at);
        if (tag != 256 && tag != 257) continue;
        const uint32_t type = t.u16((size_t)at + 2);
        const bool integral = type == 3 || type == 4;
        const uint64_t v = type == 3 ? t.u16((size_t)at + 8) : type == 4 ? t.u32((size_t)at + 8) : 0;
        const int bit = tag == 256 ? 1 : 2;
        if (tag == 256)
            w = v;
        else
            h = v;
        found = integral ? (found | bit) : (found & ~bit);
    }
    return found == 3;
}

}  // namespace

bool has_image_signature(const uint8_t* d, size_t len) { return sniff(d, len) != Container::None; }

bool header_dimensions(const uint8_t* d, size_t len, uint64_t& w, uint64_t& h) {
    if (d == nullptr) return false;
    switch (sniff(d, len)) {
        case Container::Png: return png_size(d, len, w, h);
        case Container::Bmp: return bmp_size(d, len, w, h);
        case Container::Tiff: return tiff_size(d, len, w, h);
        default: break;
    }
    // a JPEG's size needs only SOI and a marker (sniff wants three bytes)
    if (len >= 4 && d[0] == 0xFF && d[1] == 0xD8) return jpeg_size(d, len, w, h);
    return false;
}

Raster imdecode(const uint8_t* d, size_t len, bool grey) {
    if (d == nullptr || len == 0 || len > (size_t)INT32_MAX) return {};
    uint64_t hw = 0, hh = 0;
    if (header_dimensions(d, len, hw, hh) && !dims_ok(hw, hh)) return {};
    Raster r;
    if (len >= 8 && std::memcmp(d, "\x89PNG\r\n\x1a\n", 8) == 0) r = decode_png(d, len);
    else if (len >= 3 && d[0] == 0xFF && d[1] == 0xD8 && d[2] == 0xFF) r = decode_jpeg(d, len, grey);
    else if (len >= 4 && (std::memcmp(d, "II*\0", 4) == 0 || std::memcmp(d, "MM\0*", 4) == 0)) r = decode_tiff(d, len);
    else if (len >= 2 && d[0] == 'B' && d[1] == 'M') r = decode_bmp(d, len);
    else if (len >= 3 && d[0] == 'P') r = decode_pnm(d, len);
    if (r.empty() || !dims_ok((uint64_t)r.w, (uint64_t)r.h)) return {};
    return r;
}

void image_dimensions(const uint8_t* d, size_t len, int& w, int& h) {
    w = h = 0;
    if (d == nullptr || len == 0) return;
    uint64_t hw = 0, hh = 0;
    if (header_dimensions(d, len, hw, hh)) {
        if (hw > 0 && hh > 0 && hw <= (uint64_t)INT32_MAX && hh <= (uint64_t)INT32_MAX) {
            w = (int)hw;
            h = (int)hh;
        }
        return;
    }
    const Raster r = imdecode(d, len);
    if (!r.empty()) {
        w = r.w;
        h = r.h;
    }
}

bool to_gray8(const Raster& r, std::vector<uint8_t>& out) {
    if (r.empty() || (r.cn != 1 && r.cn != 3)) return false;
    const size_t n = (size_t)r.w * r.h;
    out.resize(n);
    if (r.depth == Depth::U8) {
        if (r.cn == 1) {
            std::memcpy(out.data(), r.px.data(), n);
        } else {
            for (size_t i = 0; i < n; ++i) {
                const uint8_t* q = &r.px[3 * i];
                out[i] = (uint8_t)luma_fixed(q[0], q[1], q[2]);
            }
        }
        return true;
    }
    if (r.depth == Depth::U16) {
        std::vector<uint16_t> g(n);
        const uint16_t* s = as<uint16_t>(r);
        for (size_t i = 0; i < n; ++i)
            g[i] = r.cn == 1 ? s[i] : (uint16_t)luma_fixed(s[3 * i], s[3 * i + 1], s[3 * i + 2]);
        uint16_t mn = 65535, mx = 0;
        for (uint16_t v : g) {
            mn = std::min(mn, v);
            mx = std::max(mx, v);
        }
        if (!(mx > mn)) {
            std::fill(out.begin(), out.end(), (uint8_t)128);
            return true;
        }
        // robust_range (image_codec.cpp:183-199): the 0.1th and 99.9th
        // percentiles from a histogram.
        const size_t k_lo = (size_t)std::floor(0.001 * (double)(n - 1));
        const size_t k_hi = (size_t)std::ceil(0.999 * (double)(n - 1));
        std::vector<size_t> hist(65536, 0);
        for (uint16_t v : g) hist[v]++;
        double lo = mn, hi = mx;
        size_t seen = 0;
        bool have_lo = false;
        for (size_t v = 0; v < hist.size(); ++v) {
            seen += hist[v];
            if (!have_lo && seen > k_lo) {
                lo = (double)v;
                have_lo = true;
            }
            if (seen > k_hi) {
                hi = (double)v;
                break;
            }
        }
        if (!(hi > lo)) {
            lo = mn;
            hi = mx;
        }
        const double scale = 255.0 / (hi - lo);
        const float a = (float)scale, b = (float)(-lo * scale);
        for (size_t i = 0; i < n; ++i) out[i] = sat_round((double)std::fmaf((float)g[i], a, b));
        return true;
    }
    // Floating point and signed types: one float (or double) channel.
    const bool f64 = r.depth == Depth::F64 && r.cn == 1;
    std::vector<double> g(n);
    for (size_t i = 0; i < n; ++i) {
        if (r.cn == 1) {
            switch (r.depth) {
                case Depth::S16: g[i] = as<int16_t>(r)[i]; break;
                case Depth::S32: g[i] = as<int32_t>(r)[i]; break;
                case Depth::F32: g[i] = as<float>(r)[i]; break;
                case Depth::F64: g[i] = as<double>(r)[i]; break;
                default: return false;
            }
        } else {
            float c[3];
            for (int k = 0; k < 3; ++k) {
                const size_t j = 3 * i + k;
                switch (r.depth) {
                    case Depth::S16: c[k] = fetch_f<int16_t>(r, j); break;
                    case Depth::S32: c[k] = fetch_f<int32_t>(r, j); break;
                    case Depth::F32: c[k] = fetch_f<float>(r, j); break;
                    case Depth::F64: c[k] = fetch_f<double>(r, j); break;
                    default: return false;
                }
            }
            g[i] = luma_float(c[0], c[1], c[2]);
        }
    }
    double mn = std::numeric_limits<double>::infinity(), mx = -mn;
    bool any = false;
    for (double v : g) {
        if (std::isnan(v)) continue;
        any = true;
        mn = std::min(mn, v);
        mx = std::max(mx, v);
    }
    if (!any || !std::isfinite(mn) || !std::isfinite(mx) || !(mx > mn)) {
        std::fill(out.begin(), out.end(), (uint8_t)128);
        return true;
    }
    const size_t k_lo = (size_t)std::floor(0.001 * (double)(n - 1));
    const size_t k_hi = (size_t)std::ceil(0.999 * (double)(n - 1));
    double lo = mn, hi = mx;
    {
        std::vector<double> v;
        v.reserve(n);
        for (double x : g)
            if (!std::isnan(x)) v.push_back(x);
        if (k_hi < v.size()) {
            std::nth_element(v.begin(), v.begin() + (std::ptrdiff_t)k_lo, v.end());
            lo = v[k_lo];
            std::nth_element(v.begin(), v.begin() + (std::ptrdiff_t)k_hi, v.end());
            hi = v[k_hi];
        }
        if (!(hi > lo)) {
            lo = mn;
            hi = mx;
        }
    }
    const double scale = 255.0 / (hi - lo);
    if (f64) {
        const double b = -lo * scale;
        for (size_t i = 0; i < n; ++i) out[i] = sat_round(std::fma(g[i], scale, b));
    } else {
        const float a = (float)scale, b = (float)(-lo * scale);
        for (size_t i = 0; i < n; ++i) out[i] = sat_round((double)std::fmaf((float)g[i], a, b));
    }
    return true;
}

bool to_mask8(const Raster& r, std::vector<uint8_t>& out) {
    if (r.empty() || (r.cn != 1 && r.cn != 3)) return false;
    const size_t n = (size_t)r.w * r.h;
    out.resize(n);
    if (r.depth == Depth::U8) return to_gray8(r, out);
    for (size_t i = 0; i < n; ++i) {
        bool nz = false;
        if (r.depth == Depth::U16) {
            const uint16_t* s = as<uint16_t>(r);
            nz = r.cn == 1 ? s[i] != 0 : luma_fixed(s[3 * i], s[3 * i + 1], s[3 * i + 2]) != 0;
        } else if (r.cn == 1) {
            switch (r.depth) {
                case Depth::S16: nz = as<int16_t>(r)[i] != 0; break;
                case Depth::S32: nz = as<int32_t>(r)[i] != 0; break;
                case Depth::F32: nz = as<float>(r)[i] != 0.0f; break;
                case Depth::F64: nz = as<double>(r)[i] != 0.0; break;
                default: return false;
            }
        } else {
            float c[3];
            for (int k = 0; k < 3; ++k) {
                const size_t j = 3 * i + k;
                switch (r.depth) {
                    case Depth::S16: c[k] = fetch_f<int16_t>(r, j); break;
                    case Depth::S32: c[k] = fetch_f<int32_t>(r, j); break;
                    case Depth::F32: c[k] = fetch_f<float>(r, j); break;
                    case Depth::F64: c[k] = fetch_f<double>(r, j); break;
                    default: return false;
                }
            }
            nz = luma_float(c[0], c[1], c[2]) != 0.0f;
        }
        out[i] = nz ? 255 : 0;
    }
    return true;
}

namespace {

template <typename Convert>
bool decode_any(const uint8_t* d, size_t len, int ew, int eh, std::vector<uint8_t>& out, int& w, int& h,
                bool grey, Convert convert) {
    w = h = 0;
    if (d == nullptr || len == 0) return false;
    auto encoded = [&] {
        const Raster r = imdecode(d, len, grey);
        if (r.empty() || !convert(r, out)) return false;
        w = r.w;
        h = r.h;
        return true;
    };
    if (ew > 0 && eh > 0) {
        const uint64_t px = (uint64_t)ew * (uint64_t)eh;
        const bool fits = px <= kMaxPixels;
        const bool raw_rgba = fits && (uint64_t)len == px * 4, raw_gray = fits && (uint64_t)len == px;
        if (raw_rgba || raw_gray) {
            if (has_image_signature(d, len) && encoded()) return true;
            out.resize(px);
            if (raw_gray) {
                std::memcpy(out.data(), d, px);
            } else {
                for (uint64_t i = 0; i < px; ++i) {
                    const uint8_t* q = d + 4 * i;
                    out[i] = (uint8_t)luma_fixed(q[0], q[1], q[2]);
                }
            }
            w = ew;
            h = eh;
            return true;
        }
    }
    return encoded();
}

}  // namespace

bool decode_gray(const uint8_t* d, size_t len, int ew, int eh, std::vector<uint8_t>& out, int& w, int& h) {
    return decode_any(d, len, ew, eh, out, w, h, false,
                      [](const Raster& r, std::vector<uint8_t>& o) { return to_gray8(r, o); });
}

bool decode_mask(const uint8_t* d, size_t len, int ew, int eh, std::vector<uint8_t>& out, int& w, int& h) {
    // the reference reads masks with IMREAD_GRAYSCALE: a colour JPEG gives its Y plane
    return decode_any(d, len, ew, eh, out, w, h, true,
                      [](const Raster& r, std::vector<uint8_t>& o) { return to_mask8(r, o); });
}

void resize_nearest(const std::vector<uint8_t>& in, int sw, int sh, int dw, int dh, std::vector<uint8_t>& out) {
    out.resize((size_t)dw * dh);
    const double ifx = 1.0 / ((double)dw / sw), ify = 1.0 / ((double)dh / sh);
    std::vector<int> xo(dw);
    for (int x = 0; x < dw; ++x) xo[x] = std::min((int)std::floor(x * ifx), sw - 1);
    for (int y = 0; y < dh; ++y) {
        const int sy = std::min((int)std::floor(y * ify), sh - 1);
        const uint8_t* s = in.data() + (size_t)sy * sw;
        uint8_t* o = out.data() + (size_t)y * dw;
        for (int x = 0; x < dw; ++x) o[x] = s[xo[x]];
    }
}

}  // namespace codec
}  // namespace sdic
