// Radiance RGBE (.hdr) writer and reader, and the meansq metric.
//
// The writer emits the byte stream of stb_image_write 1.16's stbi_write_hdr, which the
// reference calls from write_hdr (src/io/imagewrite.cpp:19-37): the same header text, the
// same RGBE quantisation (mantissa scaled by 256/max and truncated) and the same run-length
// scanline encoding for widths 8..32767. Byte compatibility is the point, so the encoding
// steps follow the file format as that writer emits it.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "hdrio.h"

namespace pth {

static void rgbe_encode(const float* rgb, uint8_t* out)
{
    const float m = std::max(rgb[0], std::max(rgb[1], rgb[2]));
    if (m < 1e-32f) {
        out[0] = out[1] = out[2] = out[3] = 0;
        return;
    }
    int e;
    const float scale = float(std::frexp(m, &e)) * 256.0f / m;
    out[0] = (uint8_t)(rgb[0] * scale);
    out[1] = (uint8_t)(rgb[1] * scale);
    out[2] = (uint8_t)(rgb[2] * scale);
    out[3] = (uint8_t)(e + 128);
}

static void encode_channel_rle(const uint8_t* c, int w, std::vector<uint8_t>& out)
{
    int x = 0;
    while (x < w) {
        // Start of the next run of at least three equal bytes (or the end).
        int r = x;
        while (r + 2 < w && !(c[r] == c[r + 1] && c[r] == c[r + 2]))
            r++;
        if (r + 2 >= w)
            r = w;
        while (x < r) {  // literal block(s) up to the run
            int n = std::min(r - x, 128);
            out.push_back(uint8_t(n));
            out.insert(out.end(), c + x, c + x + n);
            x += n;
        }
        if (r + 2 < w) {  // the run itself
            while (r < w && c[r] == c[x])
                r++;
            while (x < r) {
                int n = std::min(r - x, 127);
                out.push_back(uint8_t(n + 128));
                out.push_back(c[x]);
                x += n;
            }
        }
    }
}

std::vector<uint8_t> encode_hdr(const float* rgba, uint32_t w, uint32_t h, size_t row_pitch)
{
    std::vector<uint8_t> out;
    const char* header = "#?RADIANCE\n# Written by stb_image_write.h\nFORMAT=32-bit_rle_rgbe\n";
    out.insert(out.end(), header, header + strlen(header));
    char buf[128];
    int len = snprintf(buf, sizeof(buf), "EXPOSURE=          1.0000000000000\n\n-Y %u +X %u\n", h, w);
    out.insert(out.end(), buf, buf + len);
    std::vector<uint8_t> planes(size_t(w) * 4);
    for (uint32_t y = 0; y < h; y++) {
        const float* row = (const float*)((const uint8_t*)rgba + size_t(y) * row_pitch);
        if (w < 8 || w >= 32768) {
            for (uint32_t x = 0; x < w; x++) {
                uint8_t e[4];
                rgbe_encode(row + size_t(x) * 4, e);
                out.insert(out.end(), e, e + 4);
            }
            continue;
        }
        for (uint32_t x = 0; x < w; x++) {
            uint8_t e[4];
            rgbe_encode(row + size_t(x) * 4, e);
            for (int c = 0; c < 4; c++)
                planes[size_t(c) * w + x] = e[c];
        }
        const uint8_t head[4] = {2, 2, uint8_t((w & 0xff00) >> 8), uint8_t(w & 0xff)};
        out.insert(out.end(), head, head + 4);
        for (int c = 0; c < 4; c++)
            encode_channel_rle(&planes[size_t(c) * w], int(w), out);
    }
    return out;
}

void write_hdr_file(const std::string& path, const float* rgba, uint32_t w, uint32_t h, size_t row_pitch)
{
    if (w == 0 || h == 0 || !rgba)
        throw std::runtime_error("Could not write HDR file.");
    std::vector<uint8_t> bytes = encode_hdr(rgba, w, h, row_pitch);
    std::ofstream f(path, std::ios::binary);
    if (!f)
        throw std::runtime_error("Could not write HDR file.");
    f.write((const char*)bytes.data(), std::streamsize(bytes.size()));
    if (!f)
        throw std::runtime_error("Could not write HDR file.");
}

static void rgbe_decode(const uint8_t* e, float* out)
{
    if (e[3] != 0) {
        const float f = float(std::ldexp(1.0f, int(e[3]) - (128 + 8)));
        out[0] = e[0] * f;
        out[1] = e[1] * f;
        out[2] = e[2] * f;
    } else {
        out[0] = out[1] = out[2] = 0.0f;
    }
    out[3] = 1.0f;
}

std::vector<float> decode_hdr(const std::vector<uint8_t>& bytes, uint32_t& w, uint32_t& h)
{
    size_t p = 0;
    auto line = [&]() {
        std::string s;
        while (p < bytes.size() && bytes[p] != '\n')
            s += char(bytes[p++]);
        if (p < bytes.size())
            p++;
        return s;
    };
    std::string first = line();
    if (first != "#?RADIANCE" && first != "#?RGBE")
        throw std::runtime_error("not a Radiance HDR file");
    bool fmt_ok = false;
    while (true) {
        if (p >= bytes.size())
            throw std::runtime_error("HDR header without resolution");
        std::string s = line();
        if (s.empty())
            break;
        if (s == "FORMAT=32-bit_rle_rgbe")
            fmt_ok = true;
    }
    if (!fmt_ok)
        throw std::runtime_error("unsupported HDR format");
    std::string res = line();
    int hh = 0, ww = 0;
    if (sscanf(res.c_str(), "-Y %d +X %d", &hh, &ww) != 2 || hh <= 0 || ww <= 0)
        throw std::runtime_error("unsupported HDR resolution line");
    w = uint32_t(ww);
    h = uint32_t(hh);
    std::vector<float> out(size_t(w) * h * 4);
    std::vector<uint8_t> scan(size_t(w) * 4);
    for (uint32_t y = 0; y < h; y++) {
        const bool rle = w >= 8 && w < 32768 && p + 4 <= bytes.size() && bytes[p] == 2 && bytes[p + 1] == 2 &&
                         !(bytes[p + 2] & 0x80);
        if (!rle) {
            for (uint32_t x = 0; x < w; x++) {
                if (p + 4 > bytes.size())
                    throw std::runtime_error("truncated HDR data");
                rgbe_decode(&bytes[p], &out[(size_t(y) * w + x) * 4]);
                p += 4;
            }
            continue;
        }
        const uint32_t sw = (uint32_t(bytes[p + 2]) << 8) | bytes[p + 3];
        if (sw != w)
            throw std::runtime_error("HDR scanline width mismatch");
        p += 4;
        for (int c = 0; c < 4; c++) {
            uint32_t x = 0;
            while (x < w) {
                if (p >= bytes.size())
                    throw std::runtime_error("truncated HDR data");
                uint32_t n = bytes[p++];
                if (n > 128) {
                    n -= 128;
                    if (n == 0 || x + n > w || p >= bytes.size())
                        throw std::runtime_error("bad HDR run");
                    uint8_t v = bytes[p++];
                    for (uint32_t k = 0; k < n; k++)
                        scan[size_t(x++) * 4 + c] = v;
                } else {
                    if (n == 0 || x + n > w || p + n > bytes.size())
                        throw std::runtime_error("bad HDR dump");
                    for (uint32_t k = 0; k < n; k++)
                        scan[size_t(x++) * 4 + c] = bytes[p++];
                }
            }
        }
        for (uint32_t x = 0; x < w; x++)
            rgbe_decode(&scan[size_t(x) * 4], &out[(size_t(y) * w + x) * 4]);
    }
    return out;
}

std::vector<float> read_hdr_file(const std::string& path, uint32_t& w, uint32_t& h)
{
    std::ifstream f(path, std::ios::binary);
    if (!f)
        throw std::runtime_error("Could not open file " + path);
    std::vector<uint8_t> bytes((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    return decode_hdr(bytes, w, h);
}

double meansq(const float* img, const float* ref, uint32_t w, uint32_t h)
{
    double sse = 0.0;
    const size_t n = size_t(w) * h * 4;
    for (size_t i = 0; i < n; i++) {
        if ((i % 4) == 3)
            continue;
        const double e = double(img[i]) - double(ref[i]);
        sse += e * e;
    }
    return sse / (double(w) * h * 3 - 1);
}

}  // namespace pth
