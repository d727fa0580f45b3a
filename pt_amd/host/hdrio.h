#pragma once

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace pth {

std::vector<uint8_t> encode_hdr(const float* rgba, uint32_t w, uint32_t h, size_t row_pitch);
void write_hdr_file(const std::string& path, const float* rgba, uint32_t w, uint32_t h, size_t row_pitch);
std::vector<float> decode_hdr(const std::vector<uint8_t>& bytes, uint32_t& w, uint32_t& h);
std::vector<float> read_hdr_file(const std::string& path, uint32_t& w, uint32_t& h);
double meansq(const float* img, const float* ref, uint32_t w, uint32_t h);

}  // namespace pth
