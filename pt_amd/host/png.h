// PNG decoding for the texture loader: 8 bits per channel, with the channel conventions the
// reference's loader gets from stb_image (src/scene/texturedata.cpp:30-46): one, two or four
// channels (three are widened to four with alpha 255).
#pragma once

#include <cstddef>
#include <cstdint>
#include <vector>

namespace pth {

struct Image {
    uint32_t width = 0, height = 0;
    uint32_t channels = 0;     // of px: 1 (R8), 2 (R8G8) or 4 (RGBA8)
    std::vector<uint8_t> px;   // rows top to bottom
};

// Channel count the file holds: 1 gray, 2 gray+alpha (or gray with a tRNS key), 3 RGB or
// palette, 4 RGBA (or RGB / palette with tRNS). False if the bytes are not a PNG.
bool png_info(const uint8_t* data, size_t size, uint32_t& width, uint32_t& height, uint32_t& channels);

// Decodes to `want` channels (1, 2, 3 or 4). Throws std::runtime_error on malformed input.
Image png_decode(const uint8_t* data, size_t size, uint32_t want);

// The texture loader's rule: the file's channel count, three widened to four.
Image load_texture_image(const uint8_t* data, size_t size);

// zlib stream (RFC 1950/1951) to bytes; exposed for tests.
std::vector<uint8_t> zlib_inflate(const uint8_t* data, size_t size);

// Bytes to a zlib stream (one DEFLATE block, fixed Huffman codes, LZ77 matches).
std::vector<uint8_t> zlib_deflate(const uint8_t* data, size_t size);

// An 8-bit RGBA PNG file of width x height pixels whose rows are row_pitch bytes apart.
std::vector<uint8_t> png_encode_rgba8(const uint8_t* px, uint32_t width, uint32_t height, size_t row_pitch);

}  // namespace pth
