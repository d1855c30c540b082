// JPEG decode for the C ABI's image boundary. The reference hands encoded
// bytes to cv::imdecode (src/io/image_codec.cpp:117-133,268-286), whose JPEG
// decoder is libjpeg-turbo with its default decompression parameters; this
// restates that path from ITU-T T.81 and the published libjpeg algorithms:
//   - Huffman-coded sequential (SOF0, SOF1) and progressive (SOF2) frames,
//     8-bit samples, one or three components, restart intervals; data past a
//     marker or the end of the buffer reads as zero bits (libjpeg warns and
//     goes on);
//   - the "ISLOW" inverse DCT: the Loeffler-Ligtenberg-Moschytz factorisation
//     in 13-bit fixed point with 2 extra bits between the passes, its
//     zero-AC shortcuts, and the 1024-entry wrap-around range limit;
//   - "fancy" upsampling of subsampled components: triangular filters for
//     2:1 horizontal, 2:1 vertical and 2:1 both ways (the first and last
//     columns and the rows beyond the image edges replicated), plain
//     replication for other integral factors and for components at most two
//     samples wide;
//   - the JFIF YCbCr -> RGB conversion in 16-bit fixed point (R = Y + 1.402
//     Cr', B = Y + 1.772 Cb', G = Y - 0.34414 Cb' - 0.71414 Cr'), or no
//     conversion for frames coded as RGB (an Adobe marker with transform 0, or
//     component ids 'R', 'G', 'B' with no JFIF marker);
//   - a grey request (the reference's mask decode, IMREAD_GRAYSCALE) takes
//     the Y component alone, or libjpeg's RGB -> grey weights for RGB frames;
//   - the EXIF orientation (APP1), applied as OpenCV applies it after decode.
// Four-component (CMYK / YCCK) frames, arithmetic coding, lossless and
// 12-bit frames are refused.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "codec.hpp"

namespace sdic {
namespace codec {

namespace {

// zig-zag position -> natural (row-major) index, with 16 guard entries of 63
// for corrupt runs that step past the end of a block
constexpr int kNatural[64 + 16] = {
    0,  1,  8,  16, 9,  2,  3,  10, 17, 24, 32, 25, 18, 11, 4,  5,  12, 19, 26, 33, 40, 48, 41, 34, 27, 20, 13,
    6,  7,  14, 21, 28, 35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23, 30, 37, 44, 51, 58, 59, 52, 45, 38, 31,
    39, 46, 53, 60, 61, 54, 47, 55, 62, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63};

struct Huffman {
    bool present = false;
    uint8_t vals[256] = {};
    int maxcode[18] = {};   // largest code of each length, -1 when none (18: sentinel)
    int valptr[17] = {};
    int mincode[17] = {};
    // 9-bit lookahead: (length << 8) | value, 0 when the code is longer
    uint16_t look[512] = {};
    bool build(const uint8_t* counts, const uint8_t* v, int n) {
        int code = 0, k = 0;
        std::memset(look, 0, sizeof(look));
        for (int len = 1; len <= 16; ++len) {
            valptr[len] = k;
            mincode[len] = code;
            for (int i = 0; i < counts[len - 1]; ++i) {
                if (code >= (1 << len)) return false;   // over-subscribed
                if (len <= 9) {
                    const int shift = 9 - len;
                    for (int f = 0; f < (1 << shift); ++f) look[(code << shift) | f] = (uint16_t)(len << 8 | v[k]);
                }
                ++code;
                ++k;
            }
            maxcode[len] = counts[len - 1] ? code - 1 : -1;
            code <<= 1;
        }
        maxcode[17] = 0x7fffffff;
        std::memcpy(vals, v, (size_t)n);
        present = true;
        return true;
    }
};

struct Component {
    int id = 0, h = 1, v = 1, tq = 0;
    int td = 0, ta = 0;         // Huffman tables of the current scan
    int bw = 0, bh = 0;         // blocks across / down, padded to whole MCUs
    int dw = 0, dh = 0;         // samples across / down (the downsampled size)
    int pred = 0;
    std::vector<int16_t> coef;  // progressive frames: bw x bh blocks of 64
    std::vector<uint8_t> plane; // bw * 8 x bh * 8 samples
};

class Reader {
public:
    Reader(const uint8_t* d, size_t n) : d_(d), n_(n) {}
    void seek(size_t p) {
        pos_ = p;
        bits_ = 0;
        nbits_ = 0;
        hit_marker_ = false;
    }
    size_t pos() const { return pos_; }
    void reset_bits() {
        bits_ = 0;
        nbits_ = 0;
    }
    // next byte of entropy-coded data; zeros once a marker or the end is met
    void fill() {
        while (nbits_ <= 24) {
            uint32_t b = 0;
            if (!hit_marker_ && pos_ < n_) {
                b = d_[pos_];
                if (b == 0xFF) {
                    const uint32_t nx = pos_ + 1 < n_ ? d_[pos_ + 1] : 0xD9;
                    if (nx == 0x00) {
                        pos_ += 2;
                    } else {
                        hit_marker_ = true;
                        b = 0;
                    }
                } else {
                    ++pos_;
                }
            }
            bits_ |= b << (24 - nbits_);
            nbits_ += 8;
        }
    }
    int bit() {
        if (nbits_ < 1) fill();
        const int r = (int)(bits_ >> 31);
        bits_ <<= 1;
        --nbits_;
        return r;
    }
    // n <= 16 (callers clamp sizes read from a table)
    int get(int n) {
        if (n <= 0) return 0;
        if (nbits_ < n) fill();
        const int r = (int)(bits_ >> (32 - n));
        bits_ <<= n;
        nbits_ -= n;
        return r;
    }
    int decode(const Huffman& t) {
        if (nbits_ < 16) fill();
        const uint16_t e = t.look[bits_ >> 23];
        if (e) {
            const int len = e >> 8;
            bits_ <<= len;
            nbits_ -= len;
            return e & 0xFF;
        }
        int code = 0, len = 0;
        do {
            code = (code << 1) | bit();
            ++len;
        } while (len <= 16 && code > t.maxcode[len]);
        if (len > 16) return 0;   // corrupt: libjpeg substitutes zero
        return t.vals[t.valptr[len] + code - t.mincode[len]];
    }
    // past a restart interval: the marker the data now stands at
    bool restart() {
        reset_bits();
        hit_marker_ = false;
        while (pos_ + 1 < n_ && !(d_[pos_] == 0xFF && d_[pos_ + 1] >= 0xD0 && d_[pos_ + 1] <= 0xD7)) ++pos_;
        if (pos_ + 1 >= n_) return false;
        pos_ += 2;
        return true;
    }

private:
    const uint8_t* d_;
    size_t n_, pos_ = 0;
    uint32_t bits_ = 0;
    int nbits_ = 0;
    bool hit_marker_ = false;
};

inline int extend(int v, int s) { return s == 0 ? 0 : (v < (1 << (s - 1)) ? v - (1 << s) + 1 : v); }

// a DC size category from a (possibly corrupt) table: at most 16 bits follow
inline int dc_size(int t) { return t > 16 ? 16 : t; }

// ---- inverse DCT (ISLOW) ------------------------------------------------------
constexpr int kConstBits = 13, kPass1Bits = 2;
// round(x * 2^13) of the LL&M factors (c_k = cos(k pi / 16))
constexpr int32_t F0_298631336 = 2446;    // sqrt(2) (-c1 + c3 + c5 - c7)
constexpr int32_t F0_390180644 = 3196;    // sqrt(2) (c5 - c3)
constexpr int32_t F0_541196100 = 4433;    // sqrt(2) c6
constexpr int32_t F0_765366865 = 6270;    // sqrt(2) (c2 - c6)
constexpr int32_t F0_899976223 = 7373;    // sqrt(2) (c7 - c3)
constexpr int32_t F1_175875602 = 9633;    // sqrt(2) c3
constexpr int32_t F1_501321110 = 12299;   // sqrt(2) (c1 + c3 - c5 - c7)
constexpr int32_t F1_847759065 = 15137;   // sqrt(2) (c2 + c6)
constexpr int32_t F1_961570560 = 16069;   // sqrt(2) (c3 + c5)
constexpr int32_t F2_053119869 = 16819;   // sqrt(2) (c1 + c3 - c5 + c7)
constexpr int32_t F2_562915447 = 20995;   // sqrt(2) (c1 + c3)
constexpr int32_t F3_072711026 = 25172;   // sqrt(2) (c1 + c3 + c5 - c7)

inline int32_t descale(int64_t x, int n) { return (int32_t)((x + ((int64_t)1 << (n - 1))) >> n); }

// libjpeg's post-IDCT range limit: the descaled value masked to 10 bits,
// [-128, 127] -> +128, [128, 511] -> 255, [-512, -129] -> 0 (and wrapping past)
inline uint8_t range_limit(int32_t v) {
    const int i = v & 1023;
    return (uint8_t)(i < 128 ? i + 128 : i < 512 ? 255 : i < 896 ? 0 : i - 896);
}

// coef: 64 coefficients in natural order; q: quantisation table in natural order
void idct_islow(const int16_t* coef, const uint16_t* q, uint8_t* out, int stride) {
    int32_t ws[64];
    for (int c = 0; c < 8; ++c) {
        const int16_t* in = coef + c;
        const uint16_t* qq = q + c;
        if (in[8] == 0 && in[16] == 0 && in[24] == 0 && in[32] == 0 && in[40] == 0 && in[48] == 0 && in[56] == 0) {
            const int32_t dc = (int32_t)in[0] * qq[0] * (1 << kPass1Bits);
            for (int k = 0; k < 8; ++k) ws[8 * k + c] = dc;
            continue;
        }
        int64_t z2 = (int64_t)in[16] * qq[16], z3 = (int64_t)in[48] * qq[48];
        int64_t z1 = (z2 + z3) * F0_541196100;
        int64_t tmp2 = z1 + z3 * -F1_847759065;
        int64_t tmp3 = z1 + z2 * F0_765366865;
        z2 = (int64_t)in[0] * qq[0];
        z3 = (int64_t)in[32] * qq[32];
        int64_t tmp0 = (z2 + z3) * (1 << kConstBits);
        int64_t tmp1 = (z2 - z3) * (1 << kConstBits);
        const int64_t tmp10 = tmp0 + tmp3, tmp13 = tmp0 - tmp3, tmp11 = tmp1 + tmp2, tmp12 = tmp1 - tmp2;
        tmp0 = (int64_t)in[56] * qq[56];
        tmp1 = (int64_t)in[40] * qq[40];
        tmp2 = (int64_t)in[24] * qq[24];
        tmp3 = (int64_t)in[8] * qq[8];
        z1 = tmp0 + tmp3;
        z2 = tmp1 + tmp2;
        z3 = tmp0 + tmp2;
        int64_t z4 = tmp1 + tmp3;
        const int64_t z5 = (z3 + z4) * F1_175875602;
        tmp0 *= F0_298631336;
        tmp1 *= F2_053119869;
        tmp2 *= F3_072711026;
        tmp3 *= F1_501321110;
        z1 *= -F0_899976223;
        z2 *= -F2_562915447;
        z3 *= -F1_961570560;
        z4 *= -F0_390180644;
        z3 += z5;
        z4 += z5;
        tmp0 += z1 + z3;
        tmp1 += z2 + z4;
        tmp2 += z2 + z3;
        tmp3 += z1 + z4;
        constexpr int s = kConstBits - kPass1Bits;
        ws[c] = descale(tmp10 + tmp3, s);
        ws[56 + c] = descale(tmp10 - tmp3, s);
        ws[8 + c] = descale(tmp11 + tmp2, s);
        ws[48 + c] = descale(tmp11 - tmp2, s);
        ws[16 + c] = descale(tmp12 + tmp1, s);
        ws[40 + c] = descale(tmp12 - tmp1, s);
        ws[24 + c] = descale(tmp13 + tmp0, s);
        ws[32 + c] = descale(tmp13 - tmp0, s);
    }
    for (int r = 0; r < 8; ++r) {
        const int32_t* w = ws + 8 * r;
        uint8_t* o = out + (size_t)r * stride;
        if (w[1] == 0 && w[2] == 0 && w[3] == 0 && w[4] == 0 && w[5] == 0 && w[6] == 0 && w[7] == 0) {
            const uint8_t v = range_limit(descale(w[0], kPass1Bits + 3));
            for (int k = 0; k < 8; ++k) o[k] = v;
            continue;
        }
        int64_t z2 = w[2], z3 = w[6];
        int64_t z1 = (z2 + z3) * F0_541196100;
        int64_t tmp2 = z1 + z3 * -F1_847759065;
        int64_t tmp3 = z1 + z2 * F0_765366865;
        int64_t tmp0 = ((int64_t)w[0] + w[4]) * (1 << kConstBits);
        int64_t tmp1 = ((int64_t)w[0] - w[4]) * (1 << kConstBits);
        const int64_t tmp10 = tmp0 + tmp3, tmp13 = tmp0 - tmp3, tmp11 = tmp1 + tmp2, tmp12 = tmp1 - tmp2;
        tmp0 = w[7];
        tmp1 = w[5];
        tmp2 = w[3];
        tmp3 = w[1];
        z1 = tmp0 + tmp3;
        z2 = tmp1 + tmp2;
        z3 = tmp0 + tmp2;
        int64_t z4 = tmp1 + tmp3;
        const int64_t z5 = (z3 + z4) * F1_175875602;
        tmp0 *= F0_298631336;
        tmp1 *= F2_053119869;
        tmp2 *= F3_072711026;
        tmp3 *= F1_501321110;
        z1 *= -F0_899976223;
        z2 *= -F2_562915447;
        z3 *= -F1_961570560;
        z4 *= -F0_390180644;
        z3 += z5;
        z4 += z5;
        tmp0 += z1 + z3;
        tmp1 += z2 + z4;
        tmp2 += z2 + z3;
        tmp3 += z1 + z4;
        constexpr int s = kConstBits + kPass1Bits + 3;
        o[0] = range_limit(descale(tmp10 + tmp3, s));
        o[7] = range_limit(descale(tmp10 - tmp3, s));
        o[1] = range_limit(descale(tmp11 + tmp2, s));
        o[6] = range_limit(descale(tmp11 - tmp2, s));
        o[2] = range_limit(descale(tmp12 + tmp1, s));
        o[5] = range_limit(descale(tmp12 - tmp1, s));
        o[3] = range_limit(descale(tmp13 + tmp0, s));
        o[4] = range_limit(descale(tmp13 - tmp0, s));
    }
}

// ---- upsampling ----------------------------------------------------------------
// The component's dw x dh samples (stride st) to (dw * fx) x (dh * fy).
std::vector<uint8_t> upsample(const uint8_t* in, int st, int dw, int dh, int fx, int fy, int& ow, int& oh) {
    ow = dw * fx;
    oh = dh * fy;
    std::vector<uint8_t> out((size_t)ow * oh);
    auto row = [&](int y) { return in + (size_t)std::min(std::max(y, 0), dh - 1) * st; };
    if (fx == 1 && fy == 1) {
        for (int y = 0; y < dh; ++y) std::memcpy(&out[(size_t)y * ow], row(y), (size_t)dw);
    } else if (fx == 2 && fy == 1 && dw > 2) {
        for (int y = 0; y < dh; ++y) {
            const uint8_t* p = row(y);
            uint8_t* o = &out[(size_t)y * ow];
            o[0] = p[0];
            o[1] = (uint8_t)((p[0] * 3 + p[1] + 2) >> 2);
            for (int x = 1; x < dw - 1; ++x) {
                const int t = p[x] * 3;
                o[2 * x] = (uint8_t)((t + p[x - 1] + 1) >> 2);
                o[2 * x + 1] = (uint8_t)((t + p[x + 1] + 2) >> 2);
            }
            o[2 * dw - 2] = (uint8_t)((p[dw - 1] * 3 + p[dw - 2] + 1) >> 2);
            o[2 * dw - 1] = p[dw - 1];
        }
    } else if (fx == 1 && fy == 2) {
        for (int y = 0; y < dh; ++y)
            for (int v = 0; v < 2; ++v) {
                const uint8_t* p0 = row(y);
                const uint8_t* p1 = row(v == 0 ? y - 1 : y + 1);
                const int bias = v == 0 ? 1 : 2;
                uint8_t* o = &out[(size_t)(2 * y + v) * ow];
                for (int x = 0; x < dw; ++x) o[x] = (uint8_t)((p0[x] * 3 + p1[x] + bias) >> 2);
            }
    } else if (fx == 2 && fy == 2 && dw > 2) {
        for (int y = 0; y < dh; ++y)
            for (int v = 0; v < 2; ++v) {
                const uint8_t* p0 = row(y);
                const uint8_t* p1 = row(v == 0 ? y - 1 : y + 1);
                uint8_t* o = &out[(size_t)(2 * y + v) * ow];
                int thiscs = p0[0] * 3 + p1[0], nextcs = p0[1] * 3 + p1[1];
                o[0] = (uint8_t)((thiscs * 4 + 8) >> 4);
                o[1] = (uint8_t)((thiscs * 3 + nextcs + 7) >> 4);
                int lastcs = thiscs;
                thiscs = nextcs;
                for (int x = 2; x < dw; ++x) {
                    nextcs = p0[x] * 3 + p1[x];
                    o[2 * x - 2] = (uint8_t)((thiscs * 3 + lastcs + 8) >> 4);
                    o[2 * x - 1] = (uint8_t)((thiscs * 3 + nextcs + 7) >> 4);
                    lastcs = thiscs;
                    thiscs = nextcs;
                }
                o[2 * dw - 2] = (uint8_t)((thiscs * 3 + lastcs + 8) >> 4);
                o[2 * dw - 1] = (uint8_t)((thiscs * 4 + 7) >> 4);
            }
    } else {
        for (int y = 0; y < oh; ++y) {
            const uint8_t* p = row(y / fy);
            uint8_t* o = &out[(size_t)y * ow];
            for (int x = 0; x < ow; ++x) o[x] = p[x / fx];
        }
    }
    return out;
}

// ---- EXIF orientation -------------------------------------------------------------
int exif_orientation(const uint8_t* p, size_t n) {
    if (n < 14 || std::memcmp(p, "Exif\0\0", 6) != 0) return 1;
    const uint8_t* t = p + 6;
    const size_t tn = n - 6;
    const bool le = t[0] == 'I' && t[1] == 'I';
    if (!le && !(t[0] == 'M' && t[1] == 'M')) return 1;
    auto u16 = [&](size_t o) -> uint32_t {
        if (o + 2 > tn) return 0;
        return le ? (uint32_t)t[o] | (uint32_t)t[o + 1] << 8 : (uint32_t)t[o] << 8 | t[o + 1];
    };
    auto u32 = [&](size_t o) -> uint32_t {
        if (o + 4 > tn) return 0;
        return le ? (uint32_t)t[o] | (uint32_t)t[o + 1] << 8 | (uint32_t)t[o + 2] << 16 | (uint32_t)t[o + 3] << 24
                  : (uint32_t)t[o] << 24 | (uint32_t)t[o + 1] << 16 | (uint32_t)t[o + 2] << 8 | t[o + 3];
    };
    const size_t ifd = u32(4);
    if (ifd + 2 > tn) return 1;
    const uint32_t cnt = u16(ifd);
    for (uint32_t i = 0; i < cnt; ++i) {
        const size_t e = ifd + 2 + 12 * (size_t)i;
        if (e + 12 > tn) break;
        if (u16(e) == 0x0112 && u16(e + 2) == 3) {
            const uint32_t v = u16(e + 8);
            return v >= 1 && v <= 8 ? (int)v : 1;
        }
    }
    return 1;
}

// OpenCV's ApplyExifOrientation: 2 flip x, 3 flip both, 4 flip y, 5 transpose,
// 6 transpose + flip x, 7 transpose + flip both, 8 transpose + flip y
void orient(Raster& r, int o) {
    if (o <= 1 || o > 8) return;
    const int w = r.w, h = r.h, cn = r.cn;
    const bool tr = o >= 5;
    const int nw = tr ? h : w, nh = tr ? w : h;
    const bool fx = o == 2 || o == 3 || o == 6 || o == 7, fy = o == 3 || o == 4 || o == 7 || o == 8;
    std::vector<uint8_t> out((size_t)nw * nh * cn);
    for (int y = 0; y < nh; ++y)
        for (int x = 0; x < nw; ++x) {
            const int tx = fx ? nw - 1 - x : x, ty = fy ? nh - 1 - y : y;   // in the transposed (or plain) image
            const int sx = tr ? ty : tx, sy = tr ? tx : ty;
            std::memcpy(&out[((size_t)y * nw + x) * cn], &r.px[((size_t)sy * w + sx) * cn], (size_t)cn);
        }
    r.w = nw;
    r.h = nh;
    r.px.swap(out);
}

}  // namespace

Raster decode_jpeg(const uint8_t* d, size_t len, bool grey) {
    if (len < 4 || d[0] != 0xFF || d[1] != 0xD8) return {};
    uint16_t qt[4][64] = {};
    bool qt_ok[4] = {};
    Huffman dc[4], ac[4];
    std::vector<Component> comps;
    int W = 0, H = 0, hmax = 1, vmax = 1, mcux = 0, mcuy = 0, restart = 0, orientation = 1;
    bool progressive = false, have_frame = false, jfif = false, adobe = false;
    int adobe_transform = -1;
    Reader rd(d, len);
    size_t p = 2;
    auto seg = [&](size_t at) -> size_t { return at + 2 <= len ? (size_t)(d[at] << 8 | d[at + 1]) : 0; };

    for (;;) {
        // next marker
        while (p < len && d[p] != 0xFF) ++p;
        while (p < len && d[p] == 0xFF) ++p;
        if (p >= len) break;
        const int m = d[p++];
        if (m == 0xD9) break;                        // EOI
        if (m == 0x01 || (m >= 0xD0 && m <= 0xD7)) continue;
        const size_t L = seg(p);
        if (L < 2 || p + L > len) return {};
        const uint8_t* s = d + p + 2;
        const size_t sn = L - 2;
        if (m == 0xDB) {                             // DQT
            size_t o = 0;
            while (o < sn) {
                const int pq = s[o] >> 4, tq = s[o] & 15;
                if (tq > 3 || pq > 1) return {};
                const size_t need = 1 + 64 * (pq ? 2 : 1);
                if (o + need > sn) return {};
                for (int k = 0; k < 64; ++k)
                    qt[tq][kNatural[k]] = pq ? (uint16_t)(s[o + 1 + 2 * k] << 8 | s[o + 2 + 2 * k]) : s[o + 1 + k];
                qt_ok[tq] = true;
                o += need;
            }
        } else if (m == 0xC4) {                      // DHT
            size_t o = 0;
            while (o < sn) {
                if (o + 17 > sn) return {};
                const int tc = s[o] >> 4, th = s[o] & 15;
                if (tc > 1 || th > 3) return {};
                int n = 0;
                for (int k = 0; k < 16; ++k) n += s[o + 1 + k];
                if (n > 256 || o + 17 + n > sn) return {};
                Huffman& t = tc ? ac[th] : dc[th];
                if (!t.build(s + o + 1, s + o + 17, n)) return {};
                o += 17 + n;
            }
        } else if (m == 0xDD) {                      // DRI
            if (sn < 2) return {};
            restart = s[0] << 8 | s[1];
        } else if (m == 0xE0) {                      // APP0
            if (sn >= 5 && std::memcmp(s, "JFIF\0", 5) == 0) jfif = true;
        } else if (m == 0xE1) {                      // APP1
            const int o = exif_orientation(s, sn);
            if (o != 1 || (sn >= 6 && std::memcmp(s, "Exif\0\0", 6) == 0)) orientation = o;
        } else if (m == 0xEE) {                      // APP14
            if (sn >= 12 && std::memcmp(s, "Adobe", 5) == 0) {
                adobe = true;
                adobe_transform = s[11];
            }
        } else if (m == 0xC0 || m == 0xC1 || m == 0xC2) {   // SOF: sequential / progressive Huffman
            if (have_frame || sn < 6 || s[0] != 8) return {};
            progressive = m == 0xC2;
            H = s[1] << 8 | s[2];
            W = s[3] << 8 | s[4];
            const int nc = s[5];
            if (W <= 0 || H <= 0 || (nc != 1 && nc != 3) || sn < (size_t)(6 + 3 * nc)) return {};
            comps.resize(nc);
            for (int i = 0; i < nc; ++i) {
                Component& c = comps[i];
                c.id = s[6 + 3 * i];
                c.h = s[7 + 3 * i] >> 4;
                c.v = s[7 + 3 * i] & 15;
                c.tq = s[8 + 3 * i];
                if (c.h < 1 || c.h > 4 || c.v < 1 || c.v > 4 || c.tq > 3) return {};
                hmax = std::max(hmax, c.h);
                vmax = std::max(vmax, c.v);
            }
            if ((uint64_t)W * H > ((uint64_t)1 << 28)) return {};
            mcux = (W + 8 * hmax - 1) / (8 * hmax);
            mcuy = (H + 8 * vmax - 1) / (8 * vmax);
            for (Component& c : comps) {
                if (hmax % c.h || vmax % c.v) return {};   // fractional sampling
                c.bw = mcux * c.h;
                c.bh = mcuy * c.v;
                c.dw = (int)(((int64_t)W * c.h + hmax - 1) / hmax);
                c.dh = (int)(((int64_t)H * c.v + vmax - 1) / vmax);
                c.plane.assign((size_t)c.bw * 8 * c.bh * 8, 0);
                if (progressive) c.coef.assign((size_t)c.bw * c.bh * 64, 0);
            }
            have_frame = true;
        } else if ((m >= 0xC3 && m <= 0xCF) && m != 0xC4 && m != 0xC8 && m != 0xCC) {
            return {};                               // lossless, arithmetic, hierarchical
        } else if (m == 0xDA) {                      // SOS
            if (!have_frame || sn < 1) return {};
            const int ns = s[0];
            if (ns < 1 || ns > 4 || sn < (size_t)(4 + 2 * ns)) return {};
            std::vector<Component*> sc;
            for (int i = 0; i < ns; ++i) {
                Component* c = nullptr;
                for (Component& k : comps)
                    if (k.id == s[1 + 2 * i]) c = &k;
                if (!c) return {};
                c->td = s[2 + 2 * i] >> 4;
                c->ta = s[2 + 2 * i] & 15;
                if (c->td > 3 || c->ta > 3) return {};
                sc.push_back(c);
            }
            const int Ss = s[1 + 2 * ns], Se = s[2 + 2 * ns], Ah = s[3 + 2 * ns] >> 4, Al = s[3 + 2 * ns] & 15;
            if (!progressive && (Ss != 0 || Se != 63 || Ah || Al)) return {};
            if (progressive && (Se > 63 || Ss > Se || (Ss == 0 && Se != 0) || (Ss > 0 && ns != 1))) return {};
            for (Component* c : sc) {
                if (!qt_ok[c->tq]) return {};
                const bool need_dc = !progressive || (Ss == 0 && Ah == 0);
                const bool need_ac = !progressive ? true : Ss > 0;
                if (need_dc && !dc[c->td].present) return {};
                if (need_ac && !ac[c->ta].present) return {};
                c->pred = 0;
            }
            rd.seek(p + L);
            int eobrun = 0;
            // one block of a scan, decoded into blk (natural order)
            auto block = [&](Component& c, int16_t* blk) {
                if (!progressive) {
                    int16_t z[64] = {};
                    const int t = dc_size(rd.decode(dc[c.td]));
                    c.pred += extend(rd.get(t), t);
                    z[0] = (int16_t)c.pred;
                    for (int k = 1; k < 64;) {
                        const int rs = rd.decode(ac[c.ta]), r = rs >> 4, sz = rs & 15;
                        if (sz) {
                            k += r;
                            z[kNatural[std::min(k, 79)]] = (int16_t)extend(rd.get(sz), sz);
                            ++k;
                        } else {
                            if (r != 15) break;
                            k += 16;
                        }
                    }
                    std::memcpy(blk, z, sizeof(z));
                    return;
                }
                if (Ss == 0) {                       // DC scans
                    if (Ah == 0) {
                        const int t = dc_size(rd.decode(dc[c.td]));
                        c.pred += extend(rd.get(t), t);
                        blk[0] = (int16_t)(c.pred * (1 << Al));
                    } else if (rd.bit()) {
                        blk[0] = (int16_t)(blk[0] | (1 << Al));
                    }
                    return;
                }
                if (Ah == 0) {                       // AC first pass
                    if (eobrun > 0) {
                        --eobrun;
                        return;
                    }
                    for (int k = Ss; k <= Se;) {
                        const int rs = rd.decode(ac[c.ta]), r = rs >> 4, sz = rs & 15;
                        if (sz) {
                            k += r;
                            blk[kNatural[std::min(k, 79)]] = (int16_t)(extend(rd.get(sz), sz) * (1 << Al));
                            ++k;
                        } else if (r == 15) {
                            k += 16;
                        } else {
                            eobrun = (1 << r) - 1;
                            if (r) eobrun += rd.get(r);
                            break;
                        }
                    }
                    return;
                }
                // AC refinement (T.81 G.1.2.3)
                const int p1 = 1 << Al, m1 = -1 * (1 << Al);
                int k = Ss;
                if (eobrun <= 0) {
                    for (; k <= Se;) {
                        const int rs = rd.decode(ac[c.ta]);
                        int r = rs >> 4;
                        const int sz = rs & 15;
                        int val = 0;
                        if (sz) {
                            val = rd.bit() ? p1 : m1;   // sz is 1 in a valid stream
                        } else if (r != 15) {
                            eobrun = 1 << r;
                            if (r) eobrun += rd.get(r);
                            break;
                        }
                        while (k <= Se) {
                            int16_t& z = blk[kNatural[k]];
                            if (z != 0) {
                                if (rd.bit() && (z & p1) == 0) z = (int16_t)(z >= 0 ? z + p1 : z + m1);
                            } else {
                                if (r == 0) {
                                    if (val) z = (int16_t)val;
                                    ++k;
                                    break;
                                }
                                --r;
                            }
                            ++k;
                        }
                    }
                }
                if (eobrun > 0) {
                    for (; k <= Se; ++k) {
                        int16_t& z = blk[kNatural[k]];
                        if (z != 0 && rd.bit() && (z & p1) == 0) z = (int16_t)(z >= 0 ? z + p1 : z + m1);
                    }
                    --eobrun;
                }
            };
            auto emit = [&](Component& c, int bx, int by, const int16_t* blk) {
                idct_islow(blk, qt[c.tq], &c.plane[((size_t)by * 8 * c.bw + (size_t)bx) * 8], c.bw * 8);
            };
            auto do_block = [&](Component& c, int bx, int by) {
                if (progressive) {
                    block(c, &c.coef[((size_t)by * c.bw + bx) * 64]);
                } else {
                    int16_t z[64];
                    block(c, z);
                    emit(c, bx, by, z);
                }
            };
            // a restart comes before the unit that opens a new interval (never
            // after the scan's last unit)
            int units = 0;
            auto unit_start = [&]() -> bool {
                if (restart && units > 0 && units % restart == 0) {
                    if (!rd.restart()) return false;
                    for (Component* c : sc) c->pred = 0;
                    eobrun = 0;
                }
                ++units;
                return true;
            };
            if (ns == 1) {
                Component& c = *sc[0];
                const int cbw = (c.dw + 7) / 8, cbh = (c.dh + 7) / 8;
                for (int by = 0; by < cbh; ++by)
                    for (int bx = 0; bx < cbw; ++bx) {
                        if (!unit_start()) goto scan_end;
                        do_block(c, bx, by);
                    }
            } else {
                for (int my = 0; my < mcuy; ++my)
                    for (int mx = 0; mx < mcux; ++mx) {
                        if (!unit_start()) goto scan_end;
                        for (Component* c : sc)
                            for (int v = 0; v < c->v; ++v)
                                for (int h = 0; h < c->h; ++h) do_block(*c, mx * c->h + h, my * c->v + v);
                    }
            }
        scan_end:
            // continue after the entropy-coded data: the next marker that is not a restart
            p = rd.pos();
            while (p + 1 < len && !(d[p] == 0xFF && d[p + 1] != 0x00 && !(d[p + 1] >= 0xD0 && d[p + 1] <= 0xD7))) ++p;
            continue;
        }
        p += L;
    }
    if (!have_frame) return {};
    if (progressive)
        for (Component& c : comps)
            for (int by = 0; by < c.bh; ++by)
                for (int bx = 0; bx < c.bw; ++bx)
                    idct_islow(&c.coef[((size_t)by * c.bw + bx) * 64], qt[c.tq],
                               &c.plane[((size_t)by * 8 * c.bw + (size_t)bx) * 8], c.bw * 8);

    // colour space (libjpeg's default_decompress_parms)
    bool rgb = false;
    if (comps.size() == 3) {
        if (jfif) rgb = false;
        else if (adobe) rgb = adobe_transform == 0;
        else rgb = comps[0].id == 'R' && comps[1].id == 'G' && comps[2].id == 'B';
    }
    // each component at full resolution, cropped to W x H
    std::vector<std::vector<uint8_t>> full;
    const size_t nc_out = grey && !rgb ? 1 : comps.size();
    for (size_t i = 0; i < nc_out; ++i) {
        const Component& c = comps[i];
        int ow = 0, oh = 0;
        std::vector<uint8_t> u = upsample(c.plane.data(), c.bw * 8, c.dw, c.dh, hmax / c.h, vmax / c.v, ow, oh);
        std::vector<uint8_t> cr((size_t)W * H);
        for (int y = 0; y < H; ++y) std::memcpy(&cr[(size_t)y * W], &u[(size_t)y * ow], (size_t)W);
        full.push_back(std::move(cr));
    }
    Raster r;
    r.w = W;
    r.h = H;
    r.depth = Depth::U8;
    const size_t n = (size_t)W * H;
    if (comps.size() == 1 || (grey && !rgb)) {
        r.cn = 1;
        r.px = std::move(full[0]);
    } else if (grey) {
        // libjpeg's RGB -> grey: (0.299 R + 0.587 G + 0.114 B) in 16-bit fixed point
        r.cn = 1;
        r.px.resize(n);
        const int64_t fr = 19595, fg = 38470, fb = 7471;   // round(x * 65536)
        for (size_t i = 0; i < n; ++i)
            r.px[i] = (uint8_t)((fr * full[0][i] + fg * full[1][i] + fb * full[2][i] + 32768) >> 16);
    } else {
        r.cn = 3;
        r.px.resize(3 * n);
        if (rgb) {
            for (size_t i = 0; i < n; ++i)
                for (int k = 0; k < 3; ++k) r.px[3 * i + k] = full[k][i];
        } else {
            // jdcolor.c's tables, written out: x = sample - 128
            auto fix = [](double v) { return (int64_t)(v * 65536.0 + 0.5); };
            const int64_t cr_r = fix(1.40200), cb_b = fix(1.77200), cr_g = fix(0.71414), cb_g = fix(0.34414);
            const int64_t half = (int64_t)1 << 15;
            int rr[256], bb[256];
            int64_t gr[256], gb[256];
            for (int i = 0; i < 256; ++i) {
                const int64_t x = i - 128;
                rr[i] = (int)((cr_r * x + half) >> 16);
                bb[i] = (int)((cb_b * x + half) >> 16);
                gr[i] = -cr_g * x;
                gb[i] = -cb_g * x + half;
            }
            auto clamp = [](int v) { return (uint8_t)(v < 0 ? 0 : v > 255 ? 255 : v); };
            for (size_t i = 0; i < n; ++i) {
                const int y = full[0][i], cb = full[1][i], cr = full[2][i];
                r.px[3 * i] = clamp(y + rr[cr]);
                r.px[3 * i + 1] = clamp(y + (int)((gb[cb] + gr[cr]) >> 16));
                r.px[3 * i + 2] = clamp(y + bb[cb]);
            }
        }
    }
    orient(r, orientation);
    return r;
}

}  // namespace codec
}  // namespace sdic
