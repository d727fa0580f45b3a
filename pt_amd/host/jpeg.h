// JPEG decoding for the texture loader (see jpeg.cpp): the pixels stb_image 2.30 gives the
// reference's loader (src/scene/texturedata.cpp:30-46) for the same file.
#pragma once

#include <cstddef>
#include <cstdint>

#include "png.h"  // Image

namespace pth {

// Size and channel count as stb_image reports them (1 for gray, 3 for three or four
// components). False if the bytes are not a readable JPEG header.
bool jpeg_info(const uint8_t* data, size_t size, uint32_t& width, uint32_t& height, uint32_t& channels);

// Decodes to RGBA8 (three or four components) or one gray channel. Throws
// std::runtime_error on malformed input.
Image jpeg_decode(const uint8_t* data, size_t size);

}  // namespace pth
