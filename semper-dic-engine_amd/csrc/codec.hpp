// Image bytes at the boundary: raw gray / RGBA buffers and encoded PNG, TIFF,
// BMP and PNM files, to the 8-bit gray plane (or 0/255-style mask) the engine
// solves on. The rules follow the reference's io layer
// (src/io/image_codec.cpp): the raw-length test, the file signatures that
// veto it, the pixel cap checked against the header before anything is
// decoded, the one-channel collapse and the percentile stretch of images that
// are not 8-bit. The file formats are decoded here, on the host, with zlib for
// the deflate streams; what the reference's decoder library does with each
// sample layout is restated per format in codec.cpp.
#pragma once

#include <cstddef>
#include <cstdint>
#include <vector>

namespace sdic {
namespace codec {

// The most pixels one image may have (src/io/image_codec.cpp:18-25).
constexpr uint64_t kMaxPixels = uint64_t(1) << 28;

enum class Depth : uint8_t { U8, U16, S16, S32, F32, F64 };

// A decoded image of any depth: 1 or 3 interleaved channels (R, G, B order),
// rows packed, samples in host byte order.
struct Raster {
    int w = 0, h = 0, cn = 0;
    Depth depth = Depth::U8;
    std::vector<uint8_t> px;
    bool empty() const { return w <= 0 || h <= 0 || px.empty(); }
};

// Whether the bytes start like an image file (image_codec.cpp:29-40).
bool has_image_signature(const uint8_t* d, size_t len);
// Width and height from a PNG, BMP, JPEG or classic TIFF header without
// decoding (image_codec.cpp:47-105).
bool header_dimensions(const uint8_t* d, size_t len, uint64_t& w, uint64_t& h);
// image_dimensions (image_codec.cpp:265-287): the header, else a decode; zeros
// when neither gives a size.
void image_dimensions(const uint8_t* d, size_t len, int& w, int& h);

// An encoded file to a raster (empty when the bytes are not a file this
// decoder reads, or declare more than kMaxPixels).
// grey: the caller wants one channel (the reference's IMREAD_GRAYSCALE): a
// colour JPEG then yields its Y plane; other formats ignore it.
Raster imdecode(const uint8_t* d, size_t len, bool grey = false);
// JPEG (jpeg.cpp): an empty raster when the file does not decode.
Raster decode_jpeg(const uint8_t* d, size_t len, bool grey);

// to_gray8 / to_mask8 (image_codec.cpp:221-252). False on an empty raster.
bool to_gray8(const Raster& r, std::vector<uint8_t>& out);
bool to_mask8(const Raster& r, std::vector<uint8_t>& out);

// decode_gray / decode_mask (image_codec.cpp:129-161,254-262): with
// expected_w, expected_h > 0 a buffer of w*h bytes is raw gray and one of
// 4*w*h raw RGBA, unless it starts like a file that decodes; anything else is
// decoded as a file. Writes the plane and its size; false when nothing decodes.
bool decode_gray(const uint8_t* d, size_t len, int expected_w, int expected_h, std::vector<uint8_t>& out, int& w,
                 int& h);
bool decode_mask(const uint8_t* d, size_t len, int expected_w, int expected_h, std::vector<uint8_t>& out, int& w,
                 int& h);

// Nearest-neighbour resize of an 8-bit plane, source pixel
// min(floor(x * sw / dw), sw - 1) per axis (the rule masks of another size
// are brought to the image's size with, full_field_solver.cpp:190-194).
void resize_nearest(const std::vector<uint8_t>& in, int sw, int sh, int dw, int dh, std::vector<uint8_t>& out);

}  // namespace codec
}  // namespace sdic
