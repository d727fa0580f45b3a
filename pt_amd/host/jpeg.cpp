// JPEG (ITU-T T.81) decoding for the texture loader: baseline and extended sequential and
// progressive Huffman-coded 8-bit images with one, three or four components, any integer
// sampling ratios up to 4, restart intervals. Written from the standard; the reconstruction
// choices that decide the output bytes are those of stb_image 2.30, the reference's texture
// decoder (include/stb/stb_image.h, used by src/scene/texturedata.cpp:30-46):
//   - coefficients dequantised into 16-bit integers (baseline: as each block is decoded;
//     progressive: after the last scan);
//   - the "islow" integer IDCT (Loeffler-Ligtenberg-Moschytz, as in the IJG library) with
//     12-bit constants round(c * 4096) truncated toward zero, two extra bits kept between the
//     column and row passes, rows rounded by 2^16 and offset by 128;
//   - chroma upsampling: 2x horizontal, vertical or both by the 3:1 triangle filter centred
//     between samples (edges replicate), other ratios by replication;
//   - YCbCr to RGB in 20-bit fixed point with the coefficients rounded to 12 bits, and the
//     Cb contribution to green truncated to its upper 16 bits;
//   - four components: CMYK (Adobe transform 0) or YCCK (transform 2) multiplied by K with
//     the rounded 8-bit product;
//   - colour images come out as RGBA with alpha 255, gray images as one channel.
#include "jpeg.h"

#include <algorithm>
#include <cstring>
#include <stdexcept>
#include <string>

namespace pth {

namespace {

[[noreturn]] void bad(const char* what) { throw std::runtime_error(std::string("JPEG: ") + what); }

constexpr int NO_MARKER = 0xFF;  // "no marker pending" (0xFF never follows a marker prefix)

// zig-zag position -> row-major index in the 8x8 block; positions past 63 (corrupt runs)
// land on the last coefficient
constexpr uint8_t ZIGZAG[80] = {
    0,  1,  8,  16, 9,  2,  3,  10, 17, 24, 32, 25, 18, 11, 4,  5,  12, 19, 26, 33,
    40, 48, 41, 34, 27, 20, 13, 6,  7,  14, 21, 28, 35, 42, 49, 56, 57, 50, 43, 36,
    29, 22, 15, 23, 30, 37, 44, 51, 58, 59, 52, 45, 38, 31, 39, 46, 53, 60, 61, 54,
    47, 55, 62, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63};

struct Huffman {
    // canonical code of each symbol index, from the code-length counts (T.81 Annex C)
    uint8_t symbols[256];
    uint8_t len_of[257];
    uint32_t max_code[18];  // largest code of each length + 1, left-aligned to 16 bits
    int offset[17];         // symbol index = code + offset[len]
    uint8_t fast[512];      // 9-bit prefix -> symbol index (255: longer code)
    bool defined = false;

    void build(const uint8_t counts[16], const uint8_t* syms, int n)
    {
        int k = 0;
        for (int l = 1; l <= 16; l++)
            for (int i = 0; i < counts[l - 1]; i++)
                len_of[k++] = uint8_t(l);
        len_of[k] = 0;
        memcpy(symbols, syms, size_t(n));
        uint32_t code = 0;
        uint16_t codes[256];
        k = 0;
        for (int l = 1; l <= 16; l++) {
            offset[l] = k - int(code);
            while (len_of[k] == l)
                codes[k++] = uint16_t(code++);
            if (code - 1 >= (1u << l) && counts[l - 1])
                bad("bad code lengths");
            max_code[l] = code << (16 - l);
            code <<= 1;
        }
        max_code[17] = 0xFFFFFFFFu;
        memset(fast, 255, sizeof(fast));
        for (int i = 0; i < k; i++) {
            const int l = len_of[i];
            if (l <= 9) {
                const int first = codes[i] << (9 - l);
                for (int j = 0; j < (1 << (9 - l)); j++)
                    fast[first + j] = uint8_t(i);
            }
        }
        defined = true;
    }
};

struct Component {
    int id = 0, h = 1, v = 1, tq = 0;
    int hd = 0, ha = 0;
    int dc_pred = 0;
    int x = 0, y = 0;    // effective samples
    int w2 = 0, h2 = 0;  // plane size: whole MCUs
    std::vector<uint8_t> plane;
    std::vector<int16_t> coeff;  // progressive: 64 per block, row-major blocks of w2 / 8
    int coeff_w = 0;
};

class Decoder {
public:
    Decoder(const uint8_t* data, size_t size) : p_(data), end_(data + size) {}

    void read_header(bool full)
    {
        if (marker() != 0xD8)
            bad("no SOI");
        int m = marker();
        while (!(m == 0xC0 || m == 0xC1 || m == 0xC2)) {
            process_marker(m);
            m = marker();
            while (m == NO_MARKER) {  // padding between segments
                if (p_ >= end_)
                    bad("no SOF");
                m = marker();
            }
        }
        progressive_ = m == 0xC2;
        frame_header(full);
    }

    Image decode()
    {
        read_header(true);
        int m = marker();
        while (m != 0xD9) {
            if (m == 0xDA) {
                scan_header();
                entropy_coded_data();
                if (pending_ == NO_MARKER)
                    pending_ = skip_junk();
                m = marker();
                if (m >= 0xD0 && m <= 0xD7)
                    m = marker();
            } else if (m == 0xDC) {  // DNL
                const int ld = get16(), nl = get16();
                if (ld != 4)
                    bad("bad DNL length");
                if (nl != height_)
                    bad("bad DNL height");
                m = marker();
            } else {
                // a segment that cannot be read after the frame (unknown, malformed, or the end
                // of the file without EOI) ends the image as it stands
                bool ok = false;
                try {
                    ok = m != NO_MARKER && try_marker(m);
                } catch (const std::runtime_error&) {
                    ok = false;
                }
                if (!ok)
                    break;
                m = marker();
            }
        }
        if (progressive_)
            finish_progressive();
        return output();
    }

    int width() const { return width_; }
    int height() const { return height_; }
    int components() const { return ncomp_; }

private:
    const uint8_t* p_;
    const uint8_t* end_;
    int width_ = 0, height_ = 0, ncomp_ = 0;
    Component comp_[4];
    uint16_t dequant_[4][64] = {};
    Huffman huff_dc_[4], huff_ac_[4];
    int hmax_ = 1, vmax_ = 1, mcu_x_ = 0, mcu_y_ = 0;
    bool progressive_ = false;
    int restart_interval_ = 0, todo_ = 0;
    int scan_n_ = 0, order_[4] = {};
    int spec_start_ = 0, spec_end_ = 63, succ_high_ = 0, succ_low_ = 0, eob_run_ = 0;
    bool jfif_ = false;
    int adobe_transform_ = -1;
    int rgb_ids_ = 0;
    // entropy-coded bits: the next bits left-aligned in buf_
    uint32_t buf_ = 0;
    int bits_ = 0;
    int pending_ = NO_MARKER;  // marker met inside entropy-coded data
    bool stopped_ = false;     // past a marker: zero bits from here on

    int get8() { return p_ < end_ ? *p_++ : 0; }
    int get16()
    {
        const int a = get8();
        return (a << 8) | get8();
    }
    void skip(int n)
    {
        if (n < 0 || n > end_ - p_)
            p_ = end_;
        else
            p_ += n;
    }

    int marker()
    {
        if (pending_ != NO_MARKER) {
            const int m = pending_;
            pending_ = NO_MARKER;
            return m;
        }
        int x = get8();
        if (x != 0xFF)
            return NO_MARKER;
        while (x == 0xFF)  // fill bytes
            x = get8();
        return x;
    }

    // Bytes after the entropy-coded data that are not a marker are skipped until one is.
    int skip_junk()
    {
        while (p_ < end_) {
            int x = get8();
            while (x == 0xFF) {
                if (p_ >= end_)
                    return NO_MARKER;
                x = get8();
                if (x != 0x00 && x != 0xFF)
                    return x;
            }
        }
        return NO_MARKER;
    }

    void process_marker(int m)
    {
        if (!try_marker(m))
            bad("unknown marker");
    }

    bool try_marker(int m)
    {
        switch (m) {
        case NO_MARKER:
            bad("expected marker");
        case 0xDD:  // DRI
            if (get16() != 4)
                bad("bad DRI length");
            restart_interval_ = get16();
            return true;
        case 0xDB: {  // DQT
            int len = get16() - 2;
            while (len > 0) {
                const int pq = get8();
                const int prec = pq >> 4, t = pq & 15;
                if (prec > 1 || t > 3)
                    bad("bad DQT");
                for (int i = 0; i < 64; i++)
                    dequant_[t][ZIGZAG[i]] = uint16_t(prec ? get16() : get8());
                len -= prec ? 129 : 65;
            }
            if (len != 0)
                bad("bad DQT length");
            return true;
        }
        case 0xC4: {  // DHT
            int len = get16() - 2;
            while (len > 0) {
                const int tcth = get8();
                const int tc = tcth >> 4, th = tcth & 15;
                if (tc > 1 || th > 3)
                    bad("bad DHT");
                uint8_t counts[16];
                int n = 0;
                for (int i = 0; i < 16; i++) {
                    counts[i] = uint8_t(get8());
                    n += counts[i];
                }
                if (n > 256)
                    bad("bad DHT");
                uint8_t syms[256];
                for (int i = 0; i < n; i++)
                    syms[i] = uint8_t(get8());
                (tc ? huff_ac_[th] : huff_dc_[th]).build(counts, syms, n);
                len -= 17 + n;
            }
            if (len != 0)
                bad("bad DHT length");
            return true;
        }
        default:
            break;
        }
        if ((m >= 0xE0 && m <= 0xEF) || m == 0xFE) {  // APPn, COM
            int len = get16();
            if (len < 2)
                bad("bad segment length");
            len -= 2;
            if (m == 0xE0 && len >= 5) {
                static const uint8_t tag[5] = {'J', 'F', 'I', 'F', 0};
                bool ok = true;
                for (int i = 0; i < 5; i++)
                    ok &= get8() == tag[i];
                len -= 5;
                jfif_ |= ok;
            } else if (m == 0xEE && len >= 12) {
                static const uint8_t tag[6] = {'A', 'd', 'o', 'b', 'e', 0};
                bool ok = true;
                for (int i = 0; i < 6; i++)
                    ok &= get8() == tag[i];
                len -= 6;
                if (ok) {
                    get8();   // version
                    get16();  // flags
                    get16();
                    adobe_transform_ = get8();
                    len -= 6;
                }
            }
            skip(len);
            return true;
        }
        return false;
    }

    void frame_header(bool full)
    {
        const int lf = get16();
        if (lf < 11)
            bad("bad SOF length");
        if (get8() != 8)
            bad("only 8-bit samples");
        height_ = get16();
        width_ = get16();
        if (height_ == 0 || width_ == 0)
            bad("zero image size");
        if (height_ > (1 << 24) || width_ > (1 << 24))
            bad("image too large");
        ncomp_ = get8();
        if (ncomp_ != 1 && ncomp_ != 3 && ncomp_ != 4)
            bad("bad component count");
        if (lf != 8 + 3 * ncomp_)
            bad("bad SOF length");
        static const int rgb[3] = {'R', 'G', 'B'};
        for (int i = 0; i < ncomp_; i++) {
            Component& c = comp_[i];
            c.id = get8();
            if (ncomp_ == 3 && c.id == rgb[i])
                rgb_ids_++;
            const int hv = get8();
            c.h = hv >> 4;
            c.v = hv & 15;
            if (c.h < 1 || c.h > 4 || c.v < 1 || c.v > 4)
                bad("bad sampling factors");
            c.tq = get8();
            if (c.tq > 3)
                bad("bad quantisation table index");
        }
        if (!full)
            return;
        if (uint64_t(width_) * uint64_t(height_) * uint64_t(ncomp_) > (uint64_t(1) << 31))
            bad("image too large");
        for (int i = 0; i < ncomp_; i++) {
            hmax_ = std::max(hmax_, comp_[i].h);
            vmax_ = std::max(vmax_, comp_[i].v);
        }
        for (int i = 0; i < ncomp_; i++)
            if (hmax_ % comp_[i].h || vmax_ % comp_[i].v)
                bad("fractional sampling ratio");
        mcu_x_ = (width_ + hmax_ * 8 - 1) / (hmax_ * 8);
        mcu_y_ = (height_ + vmax_ * 8 - 1) / (vmax_ * 8);
        for (int i = 0; i < ncomp_; i++) {
            Component& c = comp_[i];
            c.x = (width_ * c.h + hmax_ - 1) / hmax_;
            c.y = (height_ * c.v + vmax_ - 1) / vmax_;
            c.w2 = mcu_x_ * c.h * 8;
            c.h2 = mcu_y_ * c.v * 8;
            c.plane.assign(size_t(c.w2) * c.h2, 0);
            if (progressive_) {
                c.coeff_w = c.w2 / 8;
                c.coeff.assign(size_t(c.w2) * c.h2, 0);
            }
        }
    }

    void scan_header()
    {
        const int ls = get16();
        scan_n_ = get8();
        if (scan_n_ < 1 || scan_n_ > 4 || scan_n_ > ncomp_)
            bad("bad SOS component count");
        if (ls != 6 + 2 * scan_n_)
            bad("bad SOS length");
        for (int i = 0; i < scan_n_; i++) {
            const int id = get8(), tables = get8();
            int which = 0;
            while (which < ncomp_ && comp_[which].id != id)
                which++;
            if (which == ncomp_)
                bad("scan names an unknown component");
            comp_[which].hd = tables >> 4;
            comp_[which].ha = tables & 15;
            if (comp_[which].hd > 3 || comp_[which].ha > 3)
                bad("bad Huffman table index");
            order_[i] = which;
        }
        spec_start_ = get8();
        spec_end_ = get8();
        const int a = get8();
        succ_high_ = a >> 4;
        succ_low_ = a & 15;
        if (progressive_) {
            if (spec_start_ > 63 || spec_end_ > 63 || spec_start_ > spec_end_ || succ_high_ > 13 || succ_low_ > 13)
                bad("bad progressive scan");
        } else {
            if (spec_start_ != 0 || succ_high_ != 0 || succ_low_ != 0)
                bad("bad sequential scan");
            spec_end_ = 63;
        }
    }

    // --- entropy-coded data ---------------------------------------------------------------

    void restart()
    {
        buf_ = 0;
        bits_ = 0;
        stopped_ = false;
        for (Component& c : comp_)
            c.dc_pred = 0;
        pending_ = NO_MARKER;
        todo_ = restart_interval_ ? restart_interval_ : 0x7FFFFFFF;
        eob_run_ = 0;
    }

    // Tops the bit buffer up to more than 24 bits; a marker stops the data (zeros follow).
    void fill()
    {
        do {
            int b = stopped_ ? 0 : get8();
            if (b == 0xFF) {
                int c = get8();
                while (c == 0xFF)
                    c = get8();
                if (c != 0) {
                    pending_ = c;
                    stopped_ = true;
                    return;
                }
            }
            buf_ |= uint32_t(b) << (24 - bits_);
            bits_ += 8;
        } while (bits_ <= 24);
    }

    int decode_symbol(const Huffman& h)
    {
        if (!h.defined)
            bad("undefined Huffman table");
        if (bits_ < 16)
            fill();
        const int k = h.fast[buf_ >> 23];
        if (k < 255) {
            const int l = h.len_of[k];
            if (l > bits_)
                bad("bad Huffman code");
            buf_ <<= l;
            bits_ -= l;
            return h.symbols[k];
        }
        const uint32_t top = buf_ >> 16;
        int l = 10;
        while (top >= h.max_code[l])
            l++;
        if (l == 17 || l > bits_)
            bad("bad Huffman code");
        const int idx = int(buf_ >> (32 - l)) + h.offset[l];
        if (idx < 0 || idx >= 256)
            bad("bad Huffman code");
        buf_ <<= l;
        bits_ -= l;
        return h.symbols[idx];
    }

    // n raw bits as the signed magnitude category value (T.81 F.2.2.1 EXTEND)
    int receive_extend(int n)
    {
        if (bits_ < n)
            fill();
        if (bits_ < n)
            return 0;
        const uint32_t v = buf_ >> (32 - n);
        buf_ <<= n;
        bits_ -= n;
        return v >= (1u << (n - 1)) ? int(v) : int(v) - (1 << n) + 1;
    }

    int get_bits(int n)
    {
        if (bits_ < n)
            fill();
        if (bits_ < n)
            return 0;
        const uint32_t v = buf_ >> (32 - n);
        buf_ <<= n;
        bits_ -= n;
        return int(v);
    }

    bool get_bit() { return get_bits(1) != 0; }

    // Coefficient arithmetic is modulo 2^32 and stored as its low 16 bits (the C conversion to
    // short): a corrupt stream's out-of-range DC sums and products wrap instead of overflowing.
    static int16_t wrap16(uint32_t v) { return int16_t(uint16_t(v)); }

    void decode_block(int16_t data[64], Component& c)
    {
        const uint16_t* q = dequant_[c.tq];
        const int t = decode_symbol(huff_dc_[c.hd]);
        if (t > 15)
            bad("bad DC category");
        memset(data, 0, 64 * sizeof(int16_t));
        const uint32_t dc = uint32_t(c.dc_pred) + uint32_t(t ? receive_extend(t) : 0);
        c.dc_pred = int32_t(dc);
        data[0] = wrap16(dc * q[0]);
        int k = 1;
        do {
            const int rs = decode_symbol(huff_ac_[c.ha]);
            const int r = rs >> 4, s = rs & 15;
            if (s == 0) {
                if (rs != 0xF0)
                    break;  // end of block
                k += 16;
            } else {
                k += r;
                const int z = ZIGZAG[k++];
                data[z] = wrap16(uint32_t(receive_extend(s)) * q[z]);
            }
        } while (k < 64);
    }

    void decode_prog_dc(int16_t data[64], Component& c)
    {
        if (spec_end_ != 0)
            bad("DC scan with AC coefficients");
        if (succ_high_ == 0) {
            memset(data, 0, 64 * sizeof(int16_t));
            const int t = decode_symbol(huff_dc_[c.hd]);
            if (t > 15)
                bad("bad DC category");
            const uint32_t dc = uint32_t(c.dc_pred) + uint32_t(t ? receive_extend(t) : 0);
            c.dc_pred = int32_t(dc);
            data[0] = wrap16(dc << succ_low_);
        } else if (get_bit()) {
            data[0] = wrap16(uint32_t(data[0]) + (1u << succ_low_));
        }
    }

    // refinement of one coefficient that is already nonzero: a correction bit
    void refine(int16_t* p, int16_t bit)
    {
        if (get_bit() && (*p & bit) == 0)
            *p = int16_t(*p > 0 ? *p + bit : *p - bit);
    }

    void decode_prog_ac(int16_t data[64], Component& c)
    {
        if (spec_start_ == 0)
            bad("AC scan with the DC coefficient");
        const Huffman& h = huff_ac_[c.ha];
        if (succ_high_ == 0) {
            if (eob_run_) {
                eob_run_--;
                return;
            }
            int k = spec_start_;
            do {
                const int rs = decode_symbol(h);
                const int r = rs >> 4, s = rs & 15;
                if (s == 0) {
                    if (r < 15) {
                        eob_run_ = (1 << r) + (r ? get_bits(r) : 0) - 1;
                        break;
                    }
                    k += 16;
                } else {
                    k += r;
                    data[ZIGZAG[k++]] = wrap16(uint32_t(receive_extend(s)) << succ_low_);
                }
            } while (k <= spec_end_);
            return;
        }
        const int16_t bit = int16_t(1 << succ_low_);
        if (eob_run_) {
            eob_run_--;
            for (int k = spec_start_; k <= spec_end_; k++) {
                int16_t* p = &data[ZIGZAG[k]];
                if (*p != 0)
                    refine(p, bit);
            }
            return;
        }
        int k = spec_start_;
        do {
            const int rs = decode_symbol(h);
            int r = rs >> 4, s = rs & 15;
            if (s == 0) {
                if (r < 15) {
                    eob_run_ = (1 << r) - 1 + (r ? get_bits(r) : 0);
                    r = 64;  // the rest of the band only refines
                }
                // r == 15: sixteen zeros (fifteen skipped below, the sixteenth "written" as 0)
            } else {
                if (s != 1)
                    bad("bad refinement magnitude");
                s = get_bit() ? bit : -bit;
            }
            while (k <= spec_end_) {
                int16_t* p = &data[ZIGZAG[k++]];
                if (*p != 0) {
                    refine(p, bit);
                } else {
                    if (r == 0) {
                        *p = int16_t(s);
                        break;
                    }
                    r--;
                }
            }
        } while (k <= spec_end_);
    }

    // After a restart interval's MCUs: false when the data does not go on with an RSTn.
    bool next_interval()
    {
        if (--todo_ > 0)
            return true;
        if (bits_ < 24)
            fill();
        if (!(pending_ >= 0xD0 && pending_ <= 0xD7))
            return false;
        restart();
        return true;
    }

    void entropy_coded_data()
    {
        restart();
        alignas(16) int16_t data[64];
        if (scan_n_ == 1) {
            // one component: its blocks in raster order, each block an MCU
            Component& c = comp_[order_[0]];
            const int bw = (c.x + 7) >> 3, bh = (c.y + 7) >> 3;
            for (int j = 0; j < bh; j++)
                for (int i = 0; i < bw; i++) {
                    if (!progressive_) {
                        decode_block(data, c);
                        idct(&c.plane[size_t(c.w2) * j * 8 + i * 8], c.w2, data);
                    } else {
                        int16_t* d = &c.coeff[64 * (size_t(i) + size_t(j) * c.coeff_w)];
                        if (spec_start_ == 0)
                            decode_prog_dc(d, c);
                        else
                            decode_prog_ac(d, c);
                    }
                    if (!next_interval())
                        return;
                }
            return;
        }
        for (int j = 0; j < mcu_y_; j++)
            for (int i = 0; i < mcu_x_; i++) {
                for (int k = 0; k < scan_n_; k++) {
                    Component& c = comp_[order_[k]];
                    for (int y = 0; y < c.v; y++)
                        for (int x = 0; x < c.h; x++) {
                            const int bx = i * c.h + x, by = j * c.v + y;
                            if (!progressive_) {
                                decode_block(data, c);
                                idct(&c.plane[size_t(c.w2) * by * 8 + bx * 8], c.w2, data);
                            } else {
                                decode_prog_dc(&c.coeff[64 * (size_t(bx) + size_t(by) * c.coeff_w)], c);
                            }
                        }
                }
                if (!next_interval())
                    return;
            }
    }

    void finish_progressive()
    {
        for (int n = 0; n < ncomp_; n++) {
            Component& c = comp_[n];
            const int bw = (c.x + 7) >> 3, bh = (c.y + 7) >> 3;
            for (int j = 0; j < bh; j++)
                for (int i = 0; i < bw; i++) {
                    int16_t* d = &c.coeff[64 * (size_t(i) + size_t(j) * c.coeff_w)];
                    for (int k = 0; k < 64; k++)
                        d[k] = int16_t(d[k] * dequant_[c.tq][k]);
                    idct(&c.plane[size_t(c.w2) * j * 8 + i * 8], c.w2, d);
                }
        }
    }

    // --- reconstruction -------------------------------------------------------------------

    static constexpr int fix12(float c) { return int(c * 4096 + 0.5); }

    // 1-D islow IDCT of s[0..7] into the even part x0..x3 and the odd part t0..t3 (outputs:
    // x0+t3, x1+t2, x2+t1, x3+t0, x3-t0, x2-t1, x1-t2, x0-t3), constants scaled by 2^12. The
    // arithmetic is modulo 2^32: the bits of 32-bit int arithmetic, whose overflow on the
    // out-of-range coefficients of a corrupt stream is thereby defined.
    static constexpr uint32_t k12(float c) { return uint32_t(fix12(c)); }
    struct Idct1 {
        uint32_t x0, x1, x2, x3, t0, t1, t2, t3;
        Idct1(int i0, int i1, int i2, int i3, int i4, int i5, int i6, int i7)
        {
            const uint32_t s0 = uint32_t(i0), s1 = uint32_t(i1), s2 = uint32_t(i2), s3 = uint32_t(i3),
                           s4 = uint32_t(i4), s5 = uint32_t(i5), s6 = uint32_t(i6), s7 = uint32_t(i7);
            const uint32_t e1 = (s2 + s6) * k12(0.5411961f);
            const uint32_t e2 = e1 + s6 * k12(-1.847759065f);
            const uint32_t e3 = e1 + s2 * k12(0.765366865f);
            const uint32_t e0 = (s0 + s4) * 4096u, e4 = (s0 - s4) * 4096u;
            x0 = e0 + e3;
            x3 = e0 - e3;
            x1 = e4 + e2;
            x2 = e4 - e2;
            uint32_t a0 = s7, a1 = s5, a2 = s3, a3 = s1;
            const uint32_t z3 = a0 + a2, z4 = a1 + a3, z1 = a0 + a3, z2 = a1 + a2;
            const uint32_t z5 = (z3 + z4) * k12(1.175875602f);
            a0 *= k12(0.298631336f);
            a1 *= k12(2.053119869f);
            a2 *= k12(3.072711026f);
            a3 *= k12(1.501321110f);
            const uint32_t y1 = z5 + z1 * k12(-0.899976223f);
            const uint32_t y2 = z5 + z2 * k12(-2.562915447f);
            const uint32_t y3 = z3 * k12(-1.961570560f);
            const uint32_t y4 = z4 * k12(-0.390180644f);
            t3 = a3 + y1 + y4;
            t2 = a2 + y2 + y3;
            t1 = a1 + y2 + y4;
            t0 = a0 + y1 + y3;
        }
    };
    // arithmetic right shift of a modulo-2^32 value read as a signed 32-bit int
    static int sar(uint32_t v, int n) { return int32_t(v) >> n; }

    static uint8_t clamp255(int x) { return uint8_t(x < 0 ? 0 : (x > 255 ? 255 : x)); }

    static void idct(uint8_t* out, int stride, const int16_t d[64])
    {
        int v[64];
        for (int i = 0; i < 8; i++) {  // columns
            const int16_t* c = d + i;
            if (!c[8] && !c[16] && !c[24] && !c[32] && !c[40] && !c[48] && !c[56]) {
                for (int r = 0; r < 8; r++)
                    v[i + 8 * r] = c[0] * 4;
                continue;
            }
            Idct1 f(c[0], c[8], c[16], c[24], c[32], c[40], c[48], c[56]);
            f.x0 += 512u, f.x1 += 512u, f.x2 += 512u, f.x3 += 512u;
            v[i + 0] = sar(f.x0 + f.t3, 10);
            v[i + 56] = sar(f.x0 - f.t3, 10);
            v[i + 8] = sar(f.x1 + f.t2, 10);
            v[i + 48] = sar(f.x1 - f.t2, 10);
            v[i + 16] = sar(f.x2 + f.t1, 10);
            v[i + 40] = sar(f.x2 - f.t1, 10);
            v[i + 24] = sar(f.x3 + f.t0, 10);
            v[i + 32] = sar(f.x3 - f.t0, 10);
        }
        for (int r = 0; r < 8; r++, out += stride) {  // rows
            const int* s = v + 8 * r;
            Idct1 f(s[0], s[1], s[2], s[3], s[4], s[5], s[6], s[7]);
            const uint32_t bias = 65536u + (128u << 17);
            f.x0 += bias, f.x1 += bias, f.x2 += bias, f.x3 += bias;
            out[0] = clamp255(sar(f.x0 + f.t3, 17));
            out[7] = clamp255(sar(f.x0 - f.t3, 17));
            out[1] = clamp255(sar(f.x1 + f.t2, 17));
            out[6] = clamp255(sar(f.x1 - f.t2, 17));
            out[2] = clamp255(sar(f.x2 + f.t1, 17));
            out[5] = clamp255(sar(f.x2 - f.t1, 17));
            out[3] = clamp255(sar(f.x3 + f.t0, 17));
            out[4] = clamp255(sar(f.x3 - f.t0, 17));
        }
    }

    // One output row of a component upsampled by (hs, vs) from the rows `near` and `far`.
    static const uint8_t* upsample(uint8_t* out, const uint8_t* near, const uint8_t* far, int w, int hs, int vs)
    {
        if (hs == 1 && vs == 1)
            return near;
        if (hs == 1 && vs == 2) {
            for (int i = 0; i < w; i++)
                out[i] = uint8_t((3 * near[i] + far[i] + 2) >> 2);
            return out;
        }
        if (hs == 2 && vs == 1) {
            if (w == 1) {
                out[0] = out[1] = near[0];
                return out;
            }
            out[0] = near[0];
            out[1] = uint8_t((near[0] * 3 + near[1] + 2) >> 2);
            int i = 1;
            for (; i < w - 1; i++) {
                const int n = 3 * near[i] + 2;
                out[2 * i] = uint8_t((n + near[i - 1]) >> 2);
                out[2 * i + 1] = uint8_t((n + near[i + 1]) >> 2);
            }
            out[2 * i] = uint8_t((near[w - 2] * 3 + near[w - 1] + 2) >> 2);
            out[2 * i + 1] = near[w - 1];
            return out;
        }
        if (hs == 2 && vs == 2) {
            if (w == 1) {
                out[0] = out[1] = uint8_t((3 * near[0] + far[0] + 2) >> 2);
                return out;
            }
            int cur = 3 * near[0] + far[0];
            out[0] = uint8_t((cur + 2) >> 2);
            for (int i = 1; i < w; i++) {
                const int prev = cur;
                cur = 3 * near[i] + far[i];
                out[2 * i - 1] = uint8_t((3 * prev + cur + 8) >> 4);
                out[2 * i] = uint8_t((3 * cur + prev + 8) >> 4);
            }
            out[2 * w - 1] = uint8_t((cur + 2) >> 2);
            return out;
        }
        for (int i = 0; i < w; i++)  // other ratios: replication
            for (int j = 0; j < hs; j++)
                out[i * hs + j] = near[i];
        return out;
    }

    static constexpr int fix20(float c) { return int(c * 4096.0f + 0.5f) << 8; }

    static void ycbcr_row(uint8_t* out, const uint8_t* y, const uint8_t* cb, const uint8_t* cr, int n, int step)
    {
        for (int i = 0; i < n; i++, out += step) {
            const int yf = (y[i] << 20) + (1 << 19);
            const int r1 = cr[i] - 128, b1 = cb[i] - 128;
            int r = yf + r1 * fix20(1.40200f);
            int g = yf + r1 * -fix20(0.71414f) + int(uint32_t(b1 * -fix20(0.34414f)) & 0xFFFF0000u);
            int b = yf + b1 * fix20(1.77200f);
            out[0] = clamp255(r >> 20);
            out[1] = clamp255(g >> 20);
            out[2] = clamp255(b >> 20);
            out[3] = 255;
        }
    }

    // rounded x * y / 255 of two bytes
    static uint8_t mul8(uint8_t x, uint8_t y)
    {
        const unsigned t = unsigned(x) * y + 128;
        return uint8_t((t + (t >> 8)) >> 8);
    }

    Image output()
    {
        const int n = ncomp_ >= 3 ? 4 : 1;  // the texture loader's channel count
        const bool is_rgb = ncomp_ == 3 && (rgb_ids_ == 3 || (adobe_transform_ == 0 && !jfif_));
        struct Up {
            int hs, vs, ystep, w_lores, ypos;
            const uint8_t* line0;
            const uint8_t* line1;
            std::vector<uint8_t> buf;
        } up[4];
        for (int k = 0; k < ncomp_; k++) {
            Up& u = up[k];
            u.hs = hmax_ / comp_[k].h;
            u.vs = vmax_ / comp_[k].v;
            u.ystep = u.vs >> 1;
            u.w_lores = (width_ + u.hs - 1) / u.hs;
            u.ypos = 0;
            u.line0 = u.line1 = comp_[k].plane.data();
            u.buf.assign(size_t(width_) + 3, 0);
        }
        Image img;
        img.width = uint32_t(width_);
        img.height = uint32_t(height_);
        img.channels = uint32_t(n);
        img.px.resize(size_t(n) * width_ * height_);
        const uint8_t* row[4] = {nullptr, nullptr, nullptr, nullptr};
        for (int j = 0; j < height_; j++) {
            uint8_t* out = img.px.data() + size_t(n) * width_ * j;
            for (int k = 0; k < ncomp_; k++) {
                Up& u = up[k];
                // the nearer source row first: rows are centred between the coarse samples
                const bool bottom = u.ystep >= (u.vs >> 1);
                row[k] = upsample(u.buf.data(), bottom ? u.line1 : u.line0, bottom ? u.line0 : u.line1, u.w_lores,
                                  u.hs, u.vs);
                if (++u.ystep >= u.vs) {
                    u.ystep = 0;
                    u.line0 = u.line1;
                    if (++u.ypos < comp_[k].y)
                        u.line1 += comp_[k].w2;
                }
            }
            if (n == 1) {
                memcpy(out, row[0], size_t(width_));
                continue;
            }
            if (ncomp_ == 3) {
                if (is_rgb) {
                    for (int i = 0; i < width_; i++, out += 4) {
                        out[0] = row[0][i];
                        out[1] = row[1][i];
                        out[2] = row[2][i];
                        out[3] = 255;
                    }
                } else {
                    ycbcr_row(out, row[0], row[1], row[2], width_, 4);
                }
            } else if (adobe_transform_ == 0) {  // CMYK
                for (int i = 0; i < width_; i++, out += 4) {
                    const uint8_t m = row[3][i];
                    out[0] = mul8(row[0][i], m);
                    out[1] = mul8(row[1][i], m);
                    out[2] = mul8(row[2][i], m);
                    out[3] = 255;
                }
            } else if (adobe_transform_ == 2) {  // YCCK
                ycbcr_row(out, row[0], row[1], row[2], width_, 4);
                for (int i = 0; i < width_; i++, out += 4) {
                    const uint8_t m = row[3][i];
                    out[0] = mul8(uint8_t(255 - out[0]), m);
                    out[1] = mul8(uint8_t(255 - out[1]), m);
                    out[2] = mul8(uint8_t(255 - out[2]), m);
                }
            } else {  // YCbCr with a fourth component that is not used
                ycbcr_row(out, row[0], row[1], row[2], width_, 4);
            }
        }
        return img;
    }
};

}  // namespace

bool jpeg_info(const uint8_t* data, size_t size, uint32_t& width, uint32_t& height, uint32_t& channels)
{
    if (size < 4 || data[0] != 0xFF || data[1] != 0xD8)
        return false;
    try {
        Decoder d(data, size);
        d.read_header(false);
        width = uint32_t(d.width());
        height = uint32_t(d.height());
        channels = d.components() >= 3 ? 3 : 1;
        return true;
    } catch (const std::runtime_error&) {
        return false;
    }
}

Image jpeg_decode(const uint8_t* data, size_t size)
{
    Decoder d(data, size);
    return d.decode();
}

}  // namespace pth
