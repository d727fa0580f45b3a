// PNG (ISO/IEC 15948) and zlib/DEFLATE (RFC 1950/1951) decoding, written from the
// specifications. Output conventions follow what the reference's loader receives from
// stb_image 2.30 (include/stb/stb_image.h, used by src/scene/texturedata.cpp:30-46):
// samples of 1, 2 and 4 bits are scaled to 8 bits (gray) or used as palette indices, 16-bit
// samples keep their high byte, a tRNS key gives alpha 0 to matching pixels, palettes expand
// to RGB(A), and channel-count conversion to the requested count uses the luma weights
// (77, 150, 29) / 256.
#include "png.h"

#include "jpeg.h"

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>

namespace pth {

namespace {

[[noreturn]] void bad(const char* what) { throw std::runtime_error(std::string("PNG: ") + what); }

// --- DEFLATE -------------------------------------------------------------------------------

struct Bits {
    const uint8_t* p;
    size_t n, pos = 0;
    uint64_t buf = 0;
    int cnt = 0;
    uint32_t get(int k)
    {
        while (cnt < k) {
            if (pos >= n)
                bad("truncated compressed data");
            buf |= uint64_t(p[pos++]) << cnt;
            cnt += 8;
        }
        const uint32_t v = uint32_t(buf & ((1ull << k) - 1));
        buf >>= k;
        cnt -= k;
        return v;
    }
    void align()
    {
        buf >>= (cnt & 7);
        cnt -= cnt & 7;
    }
};

// Canonical Huffman code: number of codes per length and the symbols in code order.
struct Huff {
    uint16_t count[16];
    uint16_t sym[320];
    void build(const uint8_t* len, int n)
    {
        memset(count, 0, sizeof(count));
        for (int i = 0; i < n; i++)
            count[len[i]]++;
        count[0] = 0;
        uint16_t offs[16];
        offs[1] = 0;
        for (int l = 1; l < 15; l++)
            offs[l + 1] = uint16_t(offs[l] + count[l]);
        for (int i = 0; i < n; i++)
            if (len[i])
                sym[offs[len[i]]++] = uint16_t(i);
    }
    // Reads one symbol bit by bit: codes of length l are the consecutive integers after
    // those of length l - 1 (shifted left), in symbol order.
    int decode(Bits& b) const
    {
        int code = 0, first = 0, index = 0;
        for (int l = 1; l < 16; l++) {
            code |= int(b.get(1));
            const int c = count[l];
            if (code - first < c)
                return sym[index + code - first];
            index += c;
            first = (first + c) << 1;
            code <<= 1;
        }
        bad("invalid Huffman code");
    }
};

const uint16_t LEN_BASE[29] = {3,  4,  5,  6,  7,  8,  9,  10, 11,  13,  15,  17,  19,  23, 27,
                               31, 35, 43, 51, 59, 67, 83, 99, 115, 131, 163, 195, 227, 258};
const uint8_t LEN_EXTRA[29] = {0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 2, 2, 2, 2, 3, 3, 3, 3, 4, 4, 4, 4, 5, 5, 5, 5, 0};
const uint16_t DIST_BASE[30] = {1,   2,   3,   4,   5,   7,    9,    13,   17,   25,   33,   49,   65,    97,    129,
                                193, 257, 385, 513, 769, 1025, 1537, 2049, 3073, 4097, 6145, 8193, 12289, 16385, 24577};
const uint8_t DIST_EXTRA[30] = {0, 0, 0, 0, 1, 1, 2, 2, 3, 3, 4, 4, 5, 5, 6, 6, 7, 7, 8, 8, 9, 9, 10, 10, 11, 11, 12, 12, 13, 13};

void inflate_block(Bits& b, const Huff& lit, const Huff& dist, std::vector<uint8_t>& out)
{
    while (true) {
        const int s = lit.decode(b);
        if (s < 256) {
            out.push_back(uint8_t(s));
        } else if (s == 256) {
            return;
        } else {
            const int li = s - 257;
            if (li >= 29)
                bad("invalid length symbol");
            const size_t len = LEN_BASE[li] + b.get(LEN_EXTRA[li]);
            const int di = dist.decode(b);
            if (di >= 30)
                bad("invalid distance symbol");
            const size_t d = DIST_BASE[di] + b.get(DIST_EXTRA[di]);
            if (d > out.size())
                bad("distance too far back");
            const size_t from = out.size() - d;
            for (size_t k = 0; k < len; k++)
                out.push_back(out[from + k]);
        }
    }
}

}  // namespace

std::vector<uint8_t> zlib_inflate(const uint8_t* data, size_t size)
{
    if (size < 2)
        bad("truncated zlib header");
    const uint8_t cmf = data[0], flg = data[1];
    if ((cmf & 15) != 8 || ((cmf << 8) | flg) % 31 != 0 || (flg & 32))
        bad("bad zlib header");
    Bits b{data + 2, size - 2};
    std::vector<uint8_t> out;
    bool last = false;
    while (!last) {
        last = b.get(1) != 0;
        const uint32_t type = b.get(2);
        if (type == 0) {  // stored
            b.align();
            const uint32_t len = b.get(16), nlen = b.get(16);
            if ((len ^ 0xFFFFu) != nlen)
                bad("bad stored block length");
            for (uint32_t k = 0; k < len; k++)
                out.push_back(uint8_t(b.get(8)));
        } else if (type == 1) {  // fixed Huffman codes
            uint8_t l[288];
            for (int i = 0; i < 144; i++)
                l[i] = 8;
            for (int i = 144; i < 256; i++)
                l[i] = 9;
            for (int i = 256; i < 280; i++)
                l[i] = 7;
            for (int i = 280; i < 288; i++)
                l[i] = 8;
            Huff lit, dist;
            lit.build(l, 288);
            uint8_t dl[30];
            memset(dl, 5, sizeof(dl));
            dist.build(dl, 30);
            inflate_block(b, lit, dist, out);
        } else if (type == 2) {  // dynamic Huffman codes
            const int hlit = int(b.get(5)) + 257, hdist = int(b.get(5)) + 1, hclen = int(b.get(4)) + 4;
            static const uint8_t order[19] = {16, 17, 18, 0, 8, 7, 9, 6, 10, 5, 11, 4, 12, 3, 13, 2, 14, 1, 15};
            uint8_t cl[19] = {0};
            for (int i = 0; i < hclen; i++)
                cl[order[i]] = uint8_t(b.get(3));
            Huff clh;
            clh.build(cl, 19);
            uint8_t lens[320] = {0};
            int n = 0;
            while (n < hlit + hdist) {
                const int s = clh.decode(b);
                if (s < 16) {
                    lens[n++] = uint8_t(s);
                } else {
                    int rep = 0;
                    uint8_t v = 0;
                    if (s == 16) {
                        if (n == 0)
                            bad("repeat with no previous length");
                        v = lens[n - 1];
                        rep = 3 + int(b.get(2));
                    } else if (s == 17) {
                        rep = 3 + int(b.get(3));
                    } else {
                        rep = 11 + int(b.get(7));
                    }
                    if (n + rep > hlit + hdist)
                        bad("code lengths overflow");
                    while (rep--)
                        lens[n++] = v;
                }
            }
            Huff lit, dist;
            lit.build(lens, hlit);
            dist.build(lens + hlit, hdist);
            inflate_block(b, lit, dist, out);
        } else {
            bad("invalid block type");
        }
    }
    return out;
}

namespace {

uint32_t be32(const uint8_t* p) { return (uint32_t(p[0]) << 24) | (uint32_t(p[1]) << 16) | (uint32_t(p[2]) << 8) | p[3]; }

struct Png {
    uint32_t w = 0, h = 0;
    int depth = 0, color = 0, interlace = 0;
    int samples = 0;  // per pixel in the file: 1, 2, 3 or 4
    std::vector<uint8_t> plte;  // RGB triples
    std::vector<uint8_t> trns;
    std::vector<uint8_t> idat;
};

bool parse(const uint8_t* d, size_t n, Png& p, bool header_only)
{
    static const uint8_t sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
    if (n < 8 || memcmp(d, sig, 8) != 0)
        return false;
    size_t pos = 8;
    bool have_ihdr = false;
    while (pos + 8 <= n) {
        const uint32_t len = be32(d + pos);
        const uint8_t* type = d + pos + 4;
        if (len > n - pos - 8 || n - pos - 8 - len < 4)
            bad("truncated chunk");
        const uint8_t* c = d + pos + 8;
        if (!memcmp(type, "IHDR", 4)) {
            if (len != 13)
                bad("bad IHDR");
            p.w = be32(c);
            p.h = be32(c + 4);
            p.depth = c[8];
            p.color = c[9];
            p.interlace = c[12];
            if (c[10] != 0 || c[11] != 0 || p.interlace > 1)
                bad("unsupported compression, filter or interlace method");
            if (p.w == 0 || p.h == 0 || p.w > (1u << 24) || p.h > (1u << 24))
                bad("bad image size");
            switch (p.color) {
            case 0: p.samples = 1; break;
            case 2: p.samples = 3; break;
            case 3: p.samples = 1; break;
            case 4: p.samples = 2; break;
            case 6: p.samples = 4; break;
            default: bad("bad colour type");
            }
            const bool ok_depth = p.color == 0 ? (p.depth == 1 || p.depth == 2 || p.depth == 4 || p.depth == 8 ||
                                                   p.depth == 16)
                                  : p.color == 3 ? (p.depth == 1 || p.depth == 2 || p.depth == 4 || p.depth == 8)
                                                 : (p.depth == 8 || p.depth == 16);
            if (!ok_depth)
                bad("bad bit depth");
            have_ihdr = true;
        } else if (!memcmp(type, "PLTE", 4)) {
            if (len % 3 || len > 768)
                bad("bad palette");
            p.plte.assign(c, c + len);
        } else if (!memcmp(type, "tRNS", 4)) {
            p.trns.assign(c, c + len);
        } else if (!memcmp(type, "IDAT", 4)) {
            if (header_only)
                return have_ihdr;
            p.idat.insert(p.idat.end(), c, c + len);
        } else if (!memcmp(type, "IEND", 4)) {
            break;
        } else if (!(type[0] & 32)) {
            bad("unknown critical chunk");
        }
        pos += 12 + size_t(len);
    }
    if (!have_ihdr)
        bad("no IHDR");
    return true;
}

int file_channels(const Png& p)
{
    if (p.color == 3)
        return p.trns.empty() ? 3 : 4;
    return p.samples + ((p.color == 0 || p.color == 2) && !p.trns.empty() ? 1 : 0);
}

uint8_t paeth(int a, int b, int c)
{
    const int pp = a + b - c;
    const int pa = pp > a ? pp - a : a - pp, pb = pp > b ? pp - b : b - pp, pc = pp > c ? pp - c : c - pp;
    if (pa <= pb && pa <= pc)
        return uint8_t(a);
    return uint8_t(pb <= pc ? b : c);
}

// Undoes the row filters of one (sub-)image of w x h pixels starting at raw[pos].
std::vector<uint8_t> unfilter(const std::vector<uint8_t>& raw, size_t& pos, uint32_t w, uint32_t h, int bits_pp)
{
    const size_t stride = (size_t(w) * bits_pp + 7) / 8;
    const size_t bpp = std::max(1, bits_pp / 8);
    std::vector<uint8_t> out(stride * h);
    for (uint32_t y = 0; y < h; y++) {
        if (pos + 1 + stride > raw.size())
            bad("not enough image data");
        const int f = raw[pos++];
        uint8_t* row = out.data() + y * stride;
        const uint8_t* prev = y ? row - stride : nullptr;
        for (size_t i = 0; i < stride; i++) {
            const int x = raw[pos + i];
            const int a = i >= bpp ? row[i - bpp] : 0;
            const int b = prev ? prev[i] : 0;
            const int c = (prev && i >= bpp) ? prev[i - bpp] : 0;
            int v;
            switch (f) {
            case 0: v = x; break;
            case 1: v = x + a; break;
            case 2: v = x + b; break;
            case 3: v = x + ((a + b) >> 1); break;
            case 4: v = x + paeth(a, b, c); break;
            default: bad("bad filter type");
            }
            row[i] = uint8_t(v);
        }
        pos += stride;
    }
    return out;
}

}  // namespace

bool png_info(const uint8_t* data, size_t size, uint32_t& width, uint32_t& height, uint32_t& channels)
{
    Png p;
    try {
        if (!parse(data, size, p, true))
            return false;
    } catch (const std::exception&) {
        return false;
    }
    width = p.w;
    height = p.h;
    channels = uint32_t(file_channels(p));
    return true;
}

Image png_decode(const uint8_t* data, size_t size, uint32_t want)
{
    Png p;
    if (!parse(data, size, p, false))
        bad("not a PNG file");
    if (want < 1 || want > 4)
        bad("bad channel count");
    if (p.color == 3 && p.plte.empty())
        bad("palette image without PLTE");
    const std::vector<uint8_t> raw = zlib_inflate(p.idat.data(), p.idat.size());
    const int bits_pp = p.samples * p.depth;
    const uint32_t w = p.w, h = p.h;
    // samples of every pixel at their native depth (up to 16 bits), in raster order
    std::vector<uint16_t> smp(size_t(w) * h * p.samples);
    auto put_rows = [&](const std::vector<uint8_t>& img, uint32_t pw, uint32_t ph, uint32_t x0, uint32_t y0,
                        uint32_t dx, uint32_t dy) {
        const size_t stride = (size_t(pw) * bits_pp + 7) / 8;
        for (uint32_t y = 0; y < ph; y++) {
            const uint8_t* row = img.data() + y * stride;
            for (uint32_t x = 0; x < pw; x++)
                for (int k = 0; k < p.samples; k++) {
                    const size_t si = size_t(x) * p.samples + k;
                    uint16_t v;
                    if (p.depth == 16) {
                        v = uint16_t((row[2 * si] << 8) | row[2 * si + 1]);
                    } else if (p.depth == 8) {
                        v = row[si];
                    } else {
                        const size_t bit = si * p.depth;
                        v = uint16_t((row[bit / 8] >> (8 - p.depth - int(bit % 8))) & ((1 << p.depth) - 1));
                    }
                    smp[(size_t(y0 + y * dy) * w + x0 + x * dx) * p.samples + k] = v;
                }
        }
    };
    size_t pos = 0;
    if (p.interlace == 0) {
        put_rows(unfilter(raw, pos, w, h, bits_pp), w, h, 0, 0, 1, 1);
    } else {  // Adam7
        static const uint32_t X0[7] = {0, 4, 0, 2, 0, 1, 0}, Y0[7] = {0, 0, 4, 0, 2, 0, 1};
        static const uint32_t DX[7] = {8, 8, 4, 4, 2, 2, 1}, DY[7] = {8, 8, 8, 4, 4, 2, 2};
        for (int k = 0; k < 7; k++) {
            const uint32_t pw = w > X0[k] ? (w - X0[k] + DX[k] - 1) / DX[k] : 0;
            const uint32_t ph = h > Y0[k] ? (h - Y0[k] + DY[k] - 1) / DY[k] : 0;
            if (pw == 0 || ph == 0)
                continue;
            put_rows(unfilter(raw, pos, pw, ph, bits_pp), pw, ph, X0[k], Y0[k], DX[k], DY[k]);
        }
    }

    // to 8-bit channels: out_n of them
    const size_t npx = size_t(w) * h;
    int out_n;
    std::vector<uint8_t> px;
    if (p.color == 3) {
        const size_t entries = p.plte.size() / 3;
        out_n = p.trns.empty() ? 3 : 4;
        if (want >= 3)
            out_n = int(want);
        px.resize(npx * out_n);
        for (size_t i = 0; i < npx; i++) {
            const uint16_t idx = smp[i];
            uint8_t rgba[4] = {0, 0, 0, 255};
            if (idx < entries) {
                rgba[0] = p.plte[idx * 3];
                rgba[1] = p.plte[idx * 3 + 1];
                rgba[2] = p.plte[idx * 3 + 2];
                if (idx < p.trns.size())
                    rgba[3] = p.trns[idx];
            }
            for (int k = 0; k < out_n; k++)
                px[i * out_n + k] = rgba[k];
        }
    } else {
        const bool key = !p.trns.empty() && (p.color == 0 || p.color == 2);
        if (key && p.trns.size() < size_t(2 * p.samples))
            bad("bad tRNS");
        out_n = p.samples + (key ? 1 : 0);
        px.resize(npx * out_n);
        static const int scale[9] = {0, 0xFF, 0x55, 0, 0x11, 0, 0, 0, 0x01};
        uint16_t tc[3] = {0, 0, 0};
        if (key)
            for (int k = 0; k < p.samples; k++) {
                const uint16_t v = uint16_t((p.trns[2 * k] << 8) | p.trns[2 * k + 1]);
                tc[k] = p.depth == 16 ? v : uint16_t((v & 255) * scale[p.depth]);
            }
        for (size_t i = 0; i < npx; i++) {
            bool match = key;
            for (int k = 0; k < p.samples; k++) {
                const uint16_t v = smp[i * p.samples + k];
                const uint16_t s8 = p.depth == 16 ? v : uint16_t(v * scale[p.depth]);
                if (key)
                    match = match && (p.depth == 16 ? v : s8) == tc[k];
                px[i * out_n + k] = p.depth == 16 ? uint8_t(v >> 8) : uint8_t(s8);
            }
            if (key)
                px[i * out_n + p.samples] = match ? 0 : 255;
        }
    }

    Image img;
    img.width = w;
    img.height = h;
    img.channels = want;
    if (out_n == int(want)) {
        img.px = std::move(px);
        return img;
    }
    img.px.resize(npx * want);
    for (size_t i = 0; i < npx; i++) {
        const uint8_t* s = &px[i * out_n];
        uint8_t* o = &img.px[i * want];
        auto luma = [](const uint8_t* c) { return uint8_t((c[0] * 77 + c[1] * 150 + c[2] * 29) >> 8); };
        uint8_t g, a;
        switch (out_n) {
        case 1: g = s[0], a = 255; break;
        case 2: g = s[0], a = s[1]; break;
        case 3: g = luma(s), a = 255; break;
        default: g = luma(s), a = s[3]; break;
        }
        if (want == 1) {
            o[0] = g;
        } else if (want == 2) {
            o[0] = g, o[1] = a;
        } else {
            const bool color = out_n >= 3;
            o[0] = color ? s[0] : g;
            o[1] = color ? s[1] : g;
            o[2] = color ? s[2] : g;
            if (want == 4)
                o[3] = a;
        }
    }
    return img;
}

Image load_texture_image(const uint8_t* data, size_t size)
{
    uint32_t w, h, c;
    if (png_info(data, size, w, h, c))
        return png_decode(data, size, c == 3 ? 4 : c);
    if (jpeg_info(data, size, w, h, c))
        return jpeg_decode(data, size);
    throw std::runtime_error("texture: unsupported image (PNG and JPEG are decoded)");
}

// --- encoding (the PNG path of write_png, src/io/imagewrite.cpp:44-67) -----------------------
//
// The reference writes with stb_image_write; the file here is another valid PNG of the same
// pixels (8-bit RGBA, per-row filter chosen by the smallest sum of absolute filtered bytes,
// one DEFLATE block with the fixed Huffman codes and LZ77 matches from hash chains).

namespace {

struct BitWriter {
    std::vector<uint8_t>& out;
    uint64_t buf = 0;
    int cnt = 0;
    void put(uint32_t v, int k)  // k bits of v, least significant first
    {
        buf |= uint64_t(v) << cnt;
        cnt += k;
        while (cnt >= 8) {
            out.push_back(uint8_t(buf));
            buf >>= 8;
            cnt -= 8;
        }
    }
    void put_code(uint32_t code, int len)  // a Huffman code, most significant bit first
    {
        uint32_t r = 0;
        for (int i = 0; i < len; i++)
            r |= ((code >> i) & 1u) << (len - 1 - i);
        put(r, len);
    }
    void flush()
    {
        if (cnt > 0)
            out.push_back(uint8_t(buf));
        buf = 0;
        cnt = 0;
    }
};

// Fixed literal/length codes (RFC 1951 3.2.6).
void put_symbol(BitWriter& w, int s)
{
    if (s < 144)
        w.put_code(uint32_t(0x30 + s), 8);
    else if (s < 256)
        w.put_code(uint32_t(0x190 + (s - 144)), 9);
    else if (s < 280)
        w.put_code(uint32_t(s - 256), 7);
    else
        w.put_code(uint32_t(0xC0 + (s - 280)), 8);
}

void put_match(BitWriter& w, int len, int dist)
{
    int li = 28;
    while (LEN_BASE[li] > len)
        li--;
    put_symbol(w, 257 + li);
    if (LEN_EXTRA[li])
        w.put(uint32_t(len - LEN_BASE[li]), LEN_EXTRA[li]);
    int di = 29;
    while (DIST_BASE[di] > dist)
        di--;
    w.put_code(uint32_t(di), 5);
    if (DIST_EXTRA[di])
        w.put(uint32_t(dist - DIST_BASE[di]), DIST_EXTRA[di]);
}

uint32_t crc32_of(const uint8_t* p, size_t n, uint32_t crc = 0)
{
    static uint32_t table[256];
    static bool ready = false;
    if (!ready) {
        for (uint32_t i = 0; i < 256; i++) {
            uint32_t c = i;
            for (int k = 0; k < 8; k++)
                c = (c & 1u) ? 0xEDB88320u ^ (c >> 1) : c >> 1;
            table[i] = c;
        }
        ready = true;
    }
    crc = ~crc;
    for (size_t i = 0; i < n; i++)
        crc = table[(crc ^ p[i]) & 0xFFu] ^ (crc >> 8);
    return ~crc;
}

void put_be32(std::vector<uint8_t>& v, uint32_t x)
{
    v.push_back(uint8_t(x >> 24));
    v.push_back(uint8_t(x >> 16));
    v.push_back(uint8_t(x >> 8));
    v.push_back(uint8_t(x));
}

void put_chunk(std::vector<uint8_t>& png, const char* type, const std::vector<uint8_t>& data)
{
    put_be32(png, uint32_t(data.size()));
    const size_t start = png.size();
    png.insert(png.end(), type, type + 4);
    png.insert(png.end(), data.begin(), data.end());
    put_be32(png, crc32_of(png.data() + start, png.size() - start));
}

uint8_t paeth_pred(int a, int b, int c)
{
    const int p = a + b - c, pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
    if (pa <= pb && pa <= pc)
        return uint8_t(a);
    return uint8_t(pb <= pc ? b : c);
}

}  // namespace

std::vector<uint8_t> zlib_deflate(const uint8_t* data, size_t n)
{
    std::vector<uint8_t> out = {0x78, 0x01};
    BitWriter w{out};
    w.put(1, 1);  // BFINAL
    w.put(1, 2);  // BTYPE 01: fixed Huffman codes
    constexpr int HBITS = 15;
    constexpr size_t WINDOW = 32768, WMASK = WINDOW - 1;
    constexpr int CHAIN = 32;
    std::vector<int64_t> head(size_t(1) << HBITS, -1), prev(WINDOW, -1);
    auto hash3 = [&](size_t i) {
        const uint32_t v = (uint32_t(data[i]) << 16) | (uint32_t(data[i + 1]) << 8) | data[i + 2];
        return (v * 2654435761u) >> (32 - HBITS);
    };
    auto insert = [&](size_t i) {
        if (i + 3 > n)
            return;
        const uint32_t h = hash3(i);
        prev[i & WMASK] = head[h];
        head[h] = int64_t(i);
    };
    size_t i = 0;
    while (i < n) {
        size_t best_len = 0, best_dist = 0;
        if (i + 3 <= n) {
            const size_t max_len = std::min<size_t>(258, n - i);
            int64_t cand = head[hash3(i)];
            for (int chain = 0; cand >= 0 && i - size_t(cand) <= WINDOW && chain < CHAIN; chain++) {
                size_t l = 0;
                while (l < max_len && data[size_t(cand) + l] == data[i + l])
                    l++;
                if (l > best_len) {
                    best_len = l;
                    best_dist = i - size_t(cand);
                    if (l == max_len)
                        break;
                }
                const int64_t next = prev[size_t(cand) & WMASK];
                if (next >= cand)  // the ring slot was reused by a newer position
                    break;
                cand = next;
            }
        }
        if (best_len >= 3) {
            put_match(w, int(best_len), int(best_dist));
            for (size_t k = 0; k < best_len; k++)
                insert(i + k);
            i += best_len;
        } else {
            put_symbol(w, data[i]);
            insert(i);
            i++;
        }
    }
    put_symbol(w, 256);
    w.flush();
    uint32_t a = 1, b = 0;  // Adler-32
    for (size_t k = 0; k < n; k++) {
        a = (a + data[k]) % 65521u;
        b = (b + a) % 65521u;
    }
    put_be32(out, (b << 16) | a);
    return out;
}

std::vector<uint8_t> png_encode_rgba8(const uint8_t* px, uint32_t width, uint32_t height, size_t row_pitch)
{
    const size_t rb = size_t(width) * 4;
    std::vector<uint8_t> raw;
    raw.reserve((rb + 1) * height);
    std::vector<uint8_t> cand(rb), best(rb);
    for (uint32_t y = 0; y < height; y++) {
        const uint8_t* row = px + size_t(y) * row_pitch;
        const uint8_t* up = y ? px + size_t(y - 1) * row_pitch : nullptr;
        uint64_t best_sum = ~0ull;
        int best_f = 0;
        for (int f = 0; f < 5; f++) {
            uint64_t sum = 0;
            for (size_t i = 0; i < rb; i++) {
                const int a = i >= 4 ? row[i - 4] : 0, b = up ? up[i] : 0, c = (up && i >= 4) ? up[i - 4] : 0;
                const int pred = f == 0 ? 0 : f == 1 ? a : f == 2 ? b : f == 3 ? (a + b) >> 1 : paeth_pred(a, b, c);
                cand[i] = uint8_t(row[i] - pred);
                sum += uint64_t(std::abs(int(int8_t(cand[i]))));
            }
            if (sum < best_sum) {
                best_sum = sum;
                best_f = f;
                best.swap(cand);
            }
        }
        raw.push_back(uint8_t(best_f));
        raw.insert(raw.end(), best.begin(), best.end());
    }
    std::vector<uint8_t> png = {137, 80, 78, 71, 13, 10, 26, 10};
    std::vector<uint8_t> ihdr;
    put_be32(ihdr, width);
    put_be32(ihdr, height);
    ihdr.insert(ihdr.end(), {8, 6, 0, 0, 0});  // 8 bits, RGBA, deflate, filter method 0, no interlace
    put_chunk(png, "IHDR", ihdr);
    put_chunk(png, "IDAT", zlib_deflate(raw.data(), raw.size()));
    put_chunk(png, "IEND", {});
    return png;
}

}  // namespace pth
