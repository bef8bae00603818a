// Environment image files without OpenImageIO.
//
// EnvironmentMap::Load (reference src/render/EnvironmentMap.cpp:49-92) asks
// OIIO for the file's pixels as float, keeps the first three channels
// (r, g = nch > 1 ? c1 : r, b = nch > 2 ? c2 : g) with row 0 the top of the
// image, and builds the importance table.  This file decodes the formats HDR
// environments ship in, from their published specifications, into the same
// RGB float rows:
//  - PFM (Portable Float Map): "PF" colour / "Pf" grey, the sign of the
//    scale token gives the byte order, rows stored bottom to top.
//  - Radiance RGBE (.hdr/.pic): header lines up to an empty line, a
//    "[+-]Y H [+-]X W" resolution line, flat or adaptive run-length
//    scanlines; a pixel (r, g, b, e) is c * 2^(e - 136), 0 when e = 0 (the
//    conversion of Walter's rgbe.c, which OIIO's reader uses).
//  - OpenEXR scanline and tiled files, single- or multi-part (the first
//    part, as OIIO's subimage 0; level 0 of mipmapped or ripmapped tiles):
//    NONE, RLE, ZIPS, ZIP, PIZ, PXR24, B44 and B44A compression; HALF,
//    FLOAT and UINT channels; colour from R/G/B, grey from Y.  Deep files,
//    the DWAA/DWAB codecs, subsampled channels and B44 channels with
//    pLinear set are rejected with an error.
// Every read is bounds-checked; malformed input yields an error, never an
// out-of-range access.
#include <zlib.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

#include "../../include/shiro_hip.h"

namespace {

struct Fail {
    std::string msg;
};

[[noreturn]] void fail(const std::string& m) { throw Fail{m}; }

constexpr uint64_t kMaxPixels = uint64_t(1) << 28;  // 268M pixels: 3.2 GB of RGB floats

struct Image {
    uint32_t w = 0, h = 0, channels = 0, format = 0;
    std::vector<float> rgb;  // row 0 = top
    void resize(uint64_t W, uint64_t H) {
        if (W == 0 || H == 0) fail("empty image");
        if (W * H > kMaxPixels) fail("image too large");
        w = static_cast<uint32_t>(W);
        h = static_cast<uint32_t>(H);
        rgb.assign(static_cast<size_t>(W * H) * 3u, 0.0f);
    }
    float* px(uint32_t x, uint32_t y) { return rgb.data() + (static_cast<size_t>(y) * w + x) * 3u; }
};

// Load's channel expansion: g falls back to r, b to g.
void put(float* o, const float* c, uint32_t n) {
    o[0] = c[0];
    o[1] = n > 1 ? c[1] : o[0];
    o[2] = n > 2 ? c[2] : o[1];
}

struct Cursor {
    const uint8_t* p;
    size_t n, at = 0;
    size_t left() const { return n - at; }
    void need(size_t k) const {
        if (k > n - at) fail("truncated file");
    }
    uint8_t u8() {
        need(1);
        return p[at++];
    }
    uint32_t u32le() {
        need(4);
        uint32_t v;
        std::memcpy(&v, p + at, 4);
        at += 4;
        return v;
    }
    int32_t i32le() { return static_cast<int32_t>(u32le()); }
    uint64_t u64le() {
        need(8);
        uint64_t v;
        std::memcpy(&v, p + at, 8);
        at += 8;
        return v;
    }
    const uint8_t* take(size_t k) {
        need(k);
        const uint8_t* r = p + at;
        at += k;
        return r;
    }
};

// ---------------------------------------------------------------------------
// PFM

bool is_space(uint8_t c) { return c == ' ' || c == '\t' || c == '\n' || c == '\r'; }

std::string token(Cursor& c) {
    while (c.left() && is_space(c.p[c.at])) ++c.at;
    std::string t;
    while (c.left() && !is_space(c.p[c.at])) {
        t.push_back(static_cast<char>(c.p[c.at++]));
        if (t.size() > 64) fail("malformed PFM header");
    }
    if (t.empty()) fail("truncated PFM header");
    return t;
}

uint64_t parse_uint(const std::string& t, const char* what) {
    if (t.empty() || t.size() > 9 || t.find_first_not_of("0123456789") != std::string::npos)
        fail(std::string("malformed ") + what);
    return std::strtoull(t.c_str(), nullptr, 10);
}

void read_pfm(Cursor& c, Image& img) {
    const std::string magic = token(c);
    uint32_t ch;
    if (magic == "PF")
        ch = 3;
    else if (magic == "Pf")
        ch = 1;
    else
        fail("not a PFM file");
    const uint64_t W = parse_uint(token(c), "PFM width");
    const uint64_t H = parse_uint(token(c), "PFM height");
    const std::string st = token(c);
    char* end = nullptr;
    const double scale = std::strtod(st.c_str(), &end);
    if (end != st.c_str() + st.size() || !(scale != 0.0) || !std::isfinite(scale)) fail("malformed PFM scale");
    if (!c.left() || !is_space(c.p[c.at])) fail("truncated PFM header");
    ++c.at;  // the single whitespace byte before the raster
    img.resize(W, H);
    img.channels = ch;
    const bool big = scale > 0.0;
    const uint8_t* d = c.take(static_cast<size_t>(W * H * ch * 4u));
    for (uint32_t r = 0; r < img.h; ++r) {
        const uint32_t y = img.h - 1u - r;  // first stored row is the bottom
        for (uint32_t x = 0; x < img.w; ++x) {
            float v[3];
            for (uint32_t k = 0; k < ch; ++k) {
                const uint8_t* b = d + ((static_cast<size_t>(r) * img.w + x) * ch + k) * 4u;
                const uint32_t u = big ? (uint32_t(b[0]) << 24 | uint32_t(b[1]) << 16 | uint32_t(b[2]) << 8 | b[3])
                                       : (uint32_t(b[3]) << 24 | uint32_t(b[2]) << 16 | uint32_t(b[1]) << 8 | b[0]);
                std::memcpy(&v[k], &u, 4);
            }
            put(img.px(x, y), v, ch);
        }
    }
}

// ---------------------------------------------------------------------------
// Radiance RGBE

float rgbe_channel(uint8_t v, uint8_t e) { return e ? std::ldexp(static_cast<float>(v), int(e) - 136) : 0.0f; }

std::string line(Cursor& c) {
    std::string s;
    while (true) {
        const uint8_t b = c.u8();
        if (b == '\n') break;
        s.push_back(static_cast<char>(b));
        if (s.size() > 4096) fail("RGBE header line too long");
    }
    if (!s.empty() && s.back() == '\r') s.pop_back();
    return s;
}

void read_rgbe(Cursor& c, Image& img) {
    std::string l = line(c);
    if (l.rfind("#?", 0) != 0) fail("not a Radiance RGBE file");
    while (true) {
        l = line(c);
        if (l.empty()) break;
        if (l.rfind("FORMAT=", 0) == 0 && l != "FORMAT=32-bit_rle_rgbe")
            fail("unsupported RGBE pixel format: " + l.substr(7));
    }
    // Resolution string: "<s>Y <H> <s>X <W>" (no transposed forms).
    l = line(c);
    char sy = 0, sx = 0;
    unsigned long H = 0, W = 0;
    char tail = 0;
    if (std::sscanf(l.c_str(), "%cY %lu %cX %lu%c", &sy, &H, &sx, &W, &tail) != 4 || (sy != '-' && sy != '+') ||
        (sx != '-' && sx != '+'))
        fail("unsupported RGBE resolution line: " + l);
    img.resize(W, H);
    img.channels = 3;
    std::vector<uint8_t> scan(static_cast<size_t>(W) * 4u);
    bool flat = W < 8 || W > 0x7fff;
    for (uint32_t r = 0; r < img.h; ++r) {
        if (!flat) {
            const uint8_t* hd = c.take(4);
            if (hd[0] != 2 || hd[1] != 2 || (hd[2] & 0x80)) {
                flat = true;  // not run-length encoded: the rest of the file is flat
                c.at -= 4;
            } else if ((uint32_t(hd[2]) << 8 | hd[3]) != W) {
                fail("RGBE scanline width mismatch");
            }
        }
        if (flat) {
            std::memcpy(scan.data(), c.take(scan.size()), scan.size());
        } else {
            // Four component planes, each a sequence of runs (count > 128:
            // count - 128 copies of one byte) and literals (count bytes).
            std::vector<uint8_t> plane(W);
            for (int k = 0; k < 4; ++k) {
                uint32_t i = 0;
                while (i < W) {
                    uint32_t n = c.u8();
                    if (n > 128) {
                        n -= 128;
                        if (n > W - i) fail("bad RGBE run length");
                        std::fill(plane.begin() + i, plane.begin() + i + n, c.u8());
                    } else {
                        if (n == 0 || n > W - i) fail("bad RGBE run length");
                        std::memcpy(plane.data() + i, c.take(n), n);
                    }
                    i += n;
                }
                for (uint32_t x = 0; x < W; ++x) scan[4u * x + k] = plane[x];
            }
        }
        const uint32_t y = sy == '-' ? r : img.h - 1u - r;
        for (uint32_t i = 0; i < img.w; ++i) {
            const uint32_t x = sx == '+' ? i : img.w - 1u - i;
            const uint8_t* q = scan.data() + 4u * i;
            const float v[3] = {rgbe_channel(q[0], q[3]), rgbe_channel(q[1], q[3]), rgbe_channel(q[2], q[3])};
            put(img.px(x, y), v, 3);
        }
    }
}

// ---------------------------------------------------------------------------
// OpenEXR

float half_to_float(uint16_t h) {
    const uint32_t s = uint32_t(h >> 15) << 31;
    const uint32_t e = (h >> 10) & 31u;
    const uint32_t m = h & 1023u;
    uint32_t bits;
    if (e == 0) {
        const float v = std::ldexp(static_cast<float>(m), -24);  // zero or subnormal, exact
        std::memcpy(&bits, &v, 4);
        bits |= s;
    } else if (e == 31) {
        bits = s | 0x7f800000u | (m << 13);  // infinity or NaN (payload kept)
    } else {
        bits = s | ((e + 112u) << 23) | (m << 13);
    }
    float f;
    std::memcpy(&f, &bits, 4);
    return f;
}

struct Channel {
    std::string name;
    int32_t type;  // 0 UINT, 1 HALF, 2 FLOAT
    bool pLinear;
    uint32_t bytes() const { return type == 1 ? 2u : 4u; }
};

enum : int { kExrNone, kExrRle, kExrZips, kExrZip, kExrPiz, kExrPxr24, kExrB44, kExrB44a, kExrDwaa, kExrDwab };
const char* const kExrCodec[] = {"NONE", "RLE", "ZIPS", "ZIP", "PIZ", "PXR24", "B44", "B44A", "DWAA", "DWAB"};

// Scanlines per chunk of a scanline file (a tile is always one chunk).
uint32_t exr_lines_per_chunk(int comp) {
    switch (comp) {
        case kExrZip:
        case kExrPxr24: return 16u;
        case kExrPiz:
        case kExrB44:
        case kExrB44a:
        case kExrDwaa: return 32u;
        case kExrDwab: return 256u;
        default: return 1u;
    }
}

// Undo the byte predictor and the even/odd split that the ZIP and RLE codecs
// apply before compressing (OpenEXR file layout: compressed data).
void unpredict(std::vector<uint8_t>& t, std::vector<uint8_t>& out) {
    for (size_t i = 1; i < t.size(); ++i) t[i] = static_cast<uint8_t>(int(t[i - 1]) + int(t[i]) - 128);
    out.resize(t.size());
    const size_t half = (t.size() + 1) / 2;
    for (size_t i = 0; i < t.size(); ++i) out[i] = (i & 1) ? t[half + i / 2] : t[i / 2];
}

void exr_rle(const uint8_t* in, size_t n, std::vector<uint8_t>& out, size_t expect) {
    out.clear();
    size_t i = 0;
    while (i < n) {
        const int8_t c = static_cast<int8_t>(in[i++]);
        if (c < 0) {
            const size_t k = static_cast<size_t>(-int(c));
            if (k > n - i || out.size() + k > expect) fail("corrupt EXR RLE data");
            out.insert(out.end(), in + i, in + i + k);
            i += k;
        } else {
            const size_t k = static_cast<size_t>(c) + 1u;
            if (i >= n || out.size() + k > expect) fail("corrupt EXR RLE data");
            out.insert(out.end(), k, in[i++]);
        }
    }
    if (out.size() != expect) fail("corrupt EXR RLE data");
}

// ---- PIZ: the 16-bit words of a block (a float channel's samples as two
// words each) mapped through a table of the values present, a two-level
// Haar-like wavelet per channel and component, then Huffman coding with a
// run-length symbol.

constexpr uint32_t kHufSymbols = (1u << 16) + 1u;  // 16-bit values and the run-length symbol
constexpr uint32_t kHufShortZeroRun = 59, kHufLongZeroRun = 63;
constexpr uint32_t kHufShortestLongRun = 2 + kHufLongZeroRun - kHufShortZeroRun;
constexpr int kHufMaxLen = 58;
constexpr int kHufFastBits = 12;

// Most-significant-bit-first reader over n bytes; reads past the end give
// zeros (callers bound what they consume by the stream's bit count).
struct MsbBits {
    const uint8_t* p;
    size_t n;
    uint64_t pos = 0;
    // The next k <= 57 bits, not consumed.
    uint64_t peek(int k) const {
        const size_t b = static_cast<size_t>(pos >> 3);
        uint64_t v = 0;
        for (size_t i = 0; i < 8; ++i) v = (v << 8) | (b + i < n ? p[b + i] : 0u);
        return (v << (pos & 7u)) >> (64 - k);
    }
    uint64_t get(int k) {
        const uint64_t v = peek(k);
        pos += static_cast<uint64_t>(k);
        return v;
    }
};

// Decodes nOut 16-bit words from OpenEXR's Huffman stream: a 20-byte header
// (little-endian: first and last symbol of the code table, table bytes,
// stream bit count, 0), the code lengths of symbols im..iM in 6-bit fields
// (59-62: runs of 2-5 zero lengths; 63 + 8 bits: runs of 6-261), then the
// codes, canonical with the longest codes first from 0.  Symbol iM is the
// run-length symbol: an 8-bit count follows, and the previous word is
// repeated that many times.
void huf_decode(const uint8_t* in, size_t nIn, uint16_t* out, size_t nOut) {
    if (nIn == 0) {
        if (nOut) fail("corrupt OpenEXR PIZ data");
        return;
    }
    if (nIn < 20) fail("corrupt OpenEXR PIZ data");
    auto u32 = [&](size_t o) {
        uint32_t v;
        std::memcpy(&v, in + o, 4);
        return v;
    };
    const uint32_t im = u32(0), iM = u32(4), nBits = u32(12);
    if (im >= kHufSymbols || iM >= kHufSymbols || im > iM) fail("corrupt OpenEXR PIZ code table");
    std::vector<uint8_t> len(kHufSymbols, 0);
    MsbBits tb{in + 20, nIn - 20};
    const uint64_t tableBits = 8ull * (nIn - 20);
    for (uint32_t i = im; i <= iM; ++i) {
        if (tb.pos + 6 > tableBits) fail("corrupt OpenEXR PIZ code table");
        const uint32_t l = static_cast<uint32_t>(tb.get(6));
        uint32_t run = 0;
        if (l == kHufLongZeroRun) {
            if (tb.pos + 8 > tableBits) fail("corrupt OpenEXR PIZ code table");
            run = static_cast<uint32_t>(tb.get(8)) + kHufShortestLongRun;
        } else if (l >= kHufShortZeroRun) {
            run = l - kHufShortZeroRun + 2u;
        } else {
            len[i] = static_cast<uint8_t>(l);
            continue;
        }
        if (run > iM + 1u - i) fail("corrupt OpenEXR PIZ code table");
        i += run - 1u;  // those lengths stay 0
    }
    const size_t tableBytes = static_cast<size_t>((tb.pos + 7) / 8);
    // Canonical codes: per length the first code and the symbols in order.
    uint64_t count[kHufMaxLen + 1] = {}, first[kHufMaxLen + 1] = {};
    for (uint32_t i = im; i <= iM; ++i) ++count[len[i]];
    uint64_t c = 0;
    for (int l = kHufMaxLen; l > 0; --l) {
        first[l] = c;
        c = (c + count[l]) >> 1;
    }
    uint64_t base[kHufMaxLen + 2] = {};
    for (int l = 1; l <= kHufMaxLen; ++l) base[l + 1] = base[l] + count[l];
    std::vector<uint32_t> sym(static_cast<size_t>(base[kHufMaxLen + 1]));
    {
        uint64_t fill[kHufMaxLen + 1];
        for (int l = 1; l <= kHufMaxLen; ++l) fill[l] = base[l];
        for (uint32_t i = im; i <= iM; ++i)
            if (len[i]) sym[static_cast<size_t>(fill[len[i]]++)] = i;
    }
    int maxLen = 0;
    for (int l = 1; l <= kHufMaxLen; ++l) {
        if (!count[l]) continue;
        maxLen = l;
        // A complete prefix code has its codes of each length below 2^l.
        if (first[l] + count[l] > (1ull << l)) fail("corrupt OpenEXR PIZ code table");
    }
    if (maxLen > 57) fail("OpenEXR PIZ code longer than 57 bits");
    // Codes of up to kHufFastBits bits through a table (length << 20 | index).
    std::vector<uint32_t> fast(1u << kHufFastBits, 0u);
    for (int l = 1; l <= kHufFastBits && l <= maxLen; ++l)
        for (uint64_t k = 0; k < count[l]; ++k) {
            const uint64_t code = first[l] + k;
            const uint32_t lo = static_cast<uint32_t>(code << (kHufFastBits - l));
            const uint32_t hi = static_cast<uint32_t>((code + 1) << (kHufFastBits - l));
            for (uint32_t e = lo; e < hi; ++e) fast[e] = (static_cast<uint32_t>(l) << 20) | static_cast<uint32_t>(base[l] + k);
        }
    const uint8_t* data = in + 20 + tableBytes;
    const size_t nData = nIn - 20 - tableBytes;
    if (tableBytes > nIn - 20 || nBits > 8ull * nData) fail("corrupt OpenEXR PIZ data");
    MsbBits db{data, nData};
    size_t o = 0;
    while (db.pos < nBits) {
        uint32_t s;
        int l;
        const uint32_t f = fast[static_cast<size_t>(db.peek(kHufFastBits))];
        if (f) {
            l = static_cast<int>(f >> 20);
            s = sym[f & 0xfffffu];
        } else {
            l = 0;
            for (int k = kHufFastBits + 1; k <= maxLen; ++k) {
                const uint64_t v = db.peek(k);
                if (count[k] && v >= first[k]) {
                    if (v - first[k] >= count[k]) fail("corrupt OpenEXR PIZ code");
                    l = k;
                    s = sym[static_cast<size_t>(base[k] + (v - first[k]))];
                    break;
                }
            }
            if (!l) fail("corrupt OpenEXR PIZ code");
        }
        if (db.pos + static_cast<uint64_t>(l) > nBits) fail("corrupt OpenEXR PIZ code");
        db.pos += static_cast<uint64_t>(l);
        if (s == iM) {
            if (db.pos + 8 > nBits || o == 0) fail("corrupt OpenEXR PIZ run");
            const size_t r = static_cast<size_t>(db.get(8));
            if (r > nOut - o) fail("corrupt OpenEXR PIZ run");
            std::fill(out + o, out + o + r, out[o - 1]);
            o += r;
        } else {
            if (o >= nOut) fail("corrupt OpenEXR PIZ data");
            out[o++] = static_cast<uint16_t>(s);
        }
    }
    if (o != nOut) fail("corrupt OpenEXR PIZ data");
}

// Inverse of the wavelet's two-sample step: 14-bit form (values below 2^14:
// signed average and difference) and 16-bit form (modular).
inline void wdec14(uint16_t l, uint16_t h, uint16_t& a, uint16_t& b) {
    const int hi = static_cast<int16_t>(h);
    const int ai = static_cast<int16_t>(l) + (hi & 1) + (hi >> 1);
    a = static_cast<uint16_t>(ai);
    b = static_cast<uint16_t>(ai - hi);
}

inline void wdec16(uint16_t l, uint16_t h, uint16_t& a, uint16_t& b) {
    const int m = l, d = h;
    const int bb = (m - (d >> 1)) & 0xffff;
    const int aa = (d + bb - 0x8000) & 0xffff;
    b = static_cast<uint16_t>(bb);
    a = static_cast<uint16_t>(aa);
}

// Inverse 2D wavelet of an nx x ny array of words at stride ox (along x) and
// oy (along y), coarsest level first; mx is the largest value the table
// maps to (below 2^14 the 14-bit step was used).
void wav2_decode(uint16_t* in, int nx, int ox, int ny, int oy, uint16_t mx) {
    const bool w14 = mx < (1 << 14);
    auto dec = [w14](uint16_t l, uint16_t h, uint16_t& a, uint16_t& b) {
        if (w14) {
            wdec14(l, h, a, b);
        } else {
            wdec16(l, h, a, b);
        }
    };
    const int n = nx > ny ? ny : nx;
    int p = 1;
    while (p <= n) p <<= 1;
    p >>= 1;
    int p2 = p;
    p >>= 1;
    while (p >= 1) {
        const ptrdiff_t oy1 = ptrdiff_t(oy) * p, oy2 = ptrdiff_t(oy) * p2;
        const ptrdiff_t ox1 = ptrdiff_t(ox) * p, ox2 = ptrdiff_t(ox) * p2;
        const ptrdiff_t ey = ptrdiff_t(oy) * (ny - p2);
        uint16_t i00, i01, i10, i11;
        ptrdiff_t py = 0;
        for (; py <= ey; py += oy2) {
            ptrdiff_t px = py;
            const ptrdiff_t ex = py + ptrdiff_t(ox) * (nx - p2);
            for (; px <= ex; px += ox2) {
                const ptrdiff_t p01 = px + ox1, p10 = px + oy1, p11 = p10 + ox1;
                dec(in[px], in[p10], i00, i10);
                dec(in[p01], in[p11], i01, i11);
                dec(i00, i01, in[px], in[p01]);
                dec(i10, i11, in[p10], in[p11]);
            }
            if (nx & p) {  // odd column
                const ptrdiff_t p10 = px + oy1;
                dec(in[px], in[p10], i00, in[p10]);
                in[px] = i00;
            }
        }
        if (ny & p) {  // odd line
            ptrdiff_t px = py;
            const ptrdiff_t ex = py + ptrdiff_t(ox) * (nx - p2);
            for (; px <= ex; px += ox2) {
                const ptrdiff_t p01 = px + ox1;
                dec(in[px], in[p01], i00, in[p01]);
                in[px] = i00;
            }
        }
        p2 = p;
        p >>= 1;
    }
}

// One PIZ block of w x h pixels into `out` (scanline layout: per line, per
// channel, w samples little-endian).
void piz_decode(const uint8_t* in, size_t n, const std::vector<Channel>& chans, uint32_t w, uint32_t h,
                std::vector<uint8_t>& out) {
    size_t words = 0;
    for (const Channel& ch : chans) words += static_cast<size_t>(w) * h * (ch.bytes() / 2u);
    Cursor c{in, n};
    auto u16 = [&]() {
        const uint8_t* b = c.take(2);
        return static_cast<uint16_t>(b[0] | (b[1] << 8));
    };
    const uint16_t minNz = u16(), maxNz = u16();
    if (maxNz >= 8192) fail("corrupt OpenEXR PIZ bitmap");
    std::vector<uint8_t> bitmap(8192, 0);
    if (minNz <= maxNz) std::memcpy(bitmap.data() + minNz, c.take(size_t(maxNz) - minNz + 1u), size_t(maxNz) - minNz + 1u);
    std::vector<uint16_t> lut(65536, 0);
    uint32_t k = 0;
    for (uint32_t i = 0; i < 65536; ++i)
        if (i == 0 || (bitmap[i >> 3] & (1u << (i & 7u)))) lut[k++] = static_cast<uint16_t>(i);
    const uint16_t maxValue = static_cast<uint16_t>(k - 1u);
    const int32_t len = c.i32le();
    if (len < 0) fail("corrupt OpenEXR PIZ data");
    const uint8_t* huf = c.take(static_cast<size_t>(len));
    std::vector<uint16_t> buf(words);
    huf_decode(huf, static_cast<size_t>(len), buf.data(), words);
    size_t off = 0;
    std::vector<size_t> start(chans.size());
    for (size_t i = 0; i < chans.size(); ++i) {
        const int size = static_cast<int>(chans[i].bytes() / 2u);
        start[i] = off;
        for (int j = 0; j < size; ++j)
            wav2_decode(buf.data() + off + j, static_cast<int>(w), size, static_cast<int>(h),
                        static_cast<int>(w) * size, maxValue);
        off += static_cast<size_t>(w) * h * static_cast<size_t>(size);
    }
    for (uint16_t& v : buf) v = lut[v];
    out.resize(words * 2u);
    uint8_t* o = out.data();
    for (uint32_t y = 0; y < h; ++y)
        for (size_t i = 0; i < chans.size(); ++i) {
            const size_t nw = static_cast<size_t>(w) * (chans[i].bytes() / 2u);
            const uint16_t* src = buf.data() + start[i] + nw * y;
            for (size_t x = 0; x < nw; ++x) {
                *o++ = static_cast<uint8_t>(src[x] & 0xffu);
                *o++ = static_cast<uint8_t>(src[x] >> 8);
            }
        }
}

// ---- PXR24: zlib over, per line and channel, the byte planes of the
// differences between neighbouring samples (HALF 2 planes, UINT 4, FLOAT
// 3: the float's top 24 bits).
void pxr24_decode(const uint8_t* in, size_t n, const std::vector<Channel>& chans, uint32_t w, uint32_t h,
                  std::vector<uint8_t>& out) {
    size_t planeBytes = 0, lineBytes = 0;
    for (const Channel& ch : chans) {
        planeBytes += static_cast<size_t>(w) * (ch.type == 1 ? 2u : (ch.type == 2 ? 3u : 4u));
        lineBytes += static_cast<size_t>(w) * ch.bytes();
    }
    std::vector<uint8_t> tmp(planeBytes * h);
    uLongf got = static_cast<uLongf>(tmp.size());
    if (uncompress(tmp.data(), &got, in, static_cast<uLong>(n)) != Z_OK || got != tmp.size())
        fail("corrupt OpenEXR PXR24 data");
    out.resize(lineBytes * h);
    const uint8_t* t = tmp.data();
    uint8_t* o = out.data();
    for (uint32_t y = 0; y < h; ++y)
        for (const Channel& ch : chans) {
            uint32_t pixel = 0;
            if (ch.type == 1) {
                const uint8_t *p0 = t, *p1 = t + w;
                t += 2u * w;
                for (uint32_t x = 0; x < w; ++x) {
                    pixel += (uint32_t(p0[x]) << 8) | p1[x];
                    const uint16_t v = static_cast<uint16_t>(pixel);
                    std::memcpy(o, &v, 2);
                    o += 2;
                }
            } else if (ch.type == 2) {
                const uint8_t *p0 = t, *p1 = t + w, *p2 = t + 2u * w;
                t += 3u * w;
                for (uint32_t x = 0; x < w; ++x) {
                    pixel += (uint32_t(p0[x]) << 24) | (uint32_t(p1[x]) << 16) | (uint32_t(p2[x]) << 8);
                    std::memcpy(o, &pixel, 4);
                    o += 4;
                }
            } else {
                const uint8_t *p0 = t, *p1 = t + w, *p2 = t + 2u * w, *p3 = t + 3u * w;
                t += 4u * w;
                for (uint32_t x = 0; x < w; ++x) {
                    pixel += (uint32_t(p0[x]) << 24) | (uint32_t(p1[x]) << 16) | (uint32_t(p2[x]) << 8) | p3[x];
                    std::memcpy(o, &pixel, 4);
                    o += 4;
                }
            }
        }
}

// ---- B44 / B44A: HALF channels in 4x4 blocks of 14 bytes (a 16-bit start
// value, a 6-bit shift and 15 6-bit differences along the columns and the
// first row, of the samples mapped to an ordered integer form) or, for
// flat blocks, 3 bytes; UINT and FLOAT channels stored as they are.
void b44_unpack14(const uint8_t* b, uint16_t s[16]) {
    s[0] = static_cast<uint16_t>((b[0] << 8) | b[1]);
    const uint32_t shift = b[2] >> 2;
    const uint32_t bias = 0x20u << shift;
    auto at = [&](int i, int from, uint32_t r) {
        s[i] = static_cast<uint16_t>(s[from] + (r << shift) - bias);
    };
    at(4, 0, ((uint32_t(b[2]) << 4) | (b[3] >> 4)) & 0x3fu);
    at(8, 4, ((uint32_t(b[3]) << 2) | (b[4] >> 6)) & 0x3fu);
    at(12, 8, b[4] & 0x3fu);
    at(1, 0, b[5] >> 2);
    at(5, 4, ((uint32_t(b[5]) << 4) | (b[6] >> 4)) & 0x3fu);
    at(9, 8, ((uint32_t(b[6]) << 2) | (b[7] >> 6)) & 0x3fu);
    at(13, 12, b[7] & 0x3fu);
    at(2, 1, b[8] >> 2);
    at(6, 5, ((uint32_t(b[8]) << 4) | (b[9] >> 4)) & 0x3fu);
    at(10, 9, ((uint32_t(b[9]) << 2) | (b[10] >> 6)) & 0x3fu);
    at(14, 13, b[10] & 0x3fu);
    at(3, 2, b[11] >> 2);
    at(7, 6, ((uint32_t(b[11]) << 4) | (b[12] >> 4)) & 0x3fu);
    at(11, 10, ((uint32_t(b[12]) << 2) | (b[13] >> 6)) & 0x3fu);
    at(15, 14, b[13] & 0x3fu);
    for (int i = 0; i < 16; ++i) s[i] = (s[i] & 0x8000u) ? static_cast<uint16_t>(s[i] & 0x7fffu) : static_cast<uint16_t>(~s[i]);
}

void b44_unpack3(const uint8_t* b, uint16_t s[16]) {
    uint16_t v = static_cast<uint16_t>((b[0] << 8) | b[1]);
    v = (v & 0x8000u) ? static_cast<uint16_t>(v & 0x7fffu) : static_cast<uint16_t>(~v);
    for (int i = 0; i < 16; ++i) s[i] = v;
}

void b44_decode(const uint8_t* in, size_t n, const std::vector<Channel>& chans, uint32_t w, uint32_t h,
                std::vector<uint8_t>& out) {
    Cursor c{in, n};
    std::vector<std::vector<uint8_t>> planes(chans.size());
    for (size_t i = 0; i < chans.size(); ++i) {
        const Channel& ch = chans[i];
        std::vector<uint8_t>& pl = planes[i];
        pl.resize(static_cast<size_t>(w) * h * ch.bytes());
        if (ch.type != 1) {
            std::memcpy(pl.data(), c.take(pl.size()), pl.size());
            continue;
        }
        if (ch.pLinear) fail("OpenEXR B44 channels with pLinear set are not supported");
        for (uint32_t y = 0; y < h; y += 4)
            for (uint32_t x = 0; x < w; x += 4) {
                uint16_t s[16];
                c.need(3);
                if (c.p[c.at + 2] >= (13u << 2)) {
                    b44_unpack3(c.take(3), s);
                } else {
                    b44_unpack14(c.take(14), s);
                }
                for (uint32_t dy = 0; dy < 4 && y + dy < h; ++dy)
                    for (uint32_t dx = 0; dx < 4 && x + dx < w; ++dx)
                        std::memcpy(pl.data() + 2u * ((static_cast<size_t>(y) + dy) * w + x + dx), &s[dy * 4 + dx], 2);
            }
    }
    size_t lineBytes = 0;
    for (const Channel& ch : chans) lineBytes += static_cast<size_t>(w) * ch.bytes();
    out.resize(lineBytes * h);
    uint8_t* o = out.data();
    for (uint32_t y = 0; y < h; ++y)
        for (size_t i = 0; i < chans.size(); ++i) {
            const size_t rb = static_cast<size_t>(w) * chans[i].bytes();
            std::memcpy(o, planes[i].data() + rb * y, rb);
            o += rb;
        }
}

// One chunk's pixels (a scanline block or a tile of w x h) in the scanline
// layout; chunks the codec did not shrink are stored uncompressed.
const uint8_t* exr_block(int comp, const std::vector<Channel>& chans, const uint8_t* data, size_t size, uint32_t w,
                         uint32_t h, std::vector<uint8_t>& raw, std::vector<uint8_t>& tmp) {
    size_t lineBytes = 0;
    for (const Channel& ch : chans) lineBytes += static_cast<size_t>(w) * ch.bytes();
    const size_t expect = lineBytes * h;
    if (size == expect) return data;
    switch (comp) {
        case kExrRle:
            exr_rle(data, size, tmp, expect);
            unpredict(tmp, raw);
            break;
        case kExrZips:
        case kExrZip: {
            tmp.resize(expect);
            uLongf len = static_cast<uLongf>(expect);
            if (uncompress(tmp.data(), &len, data, static_cast<uLong>(size)) != Z_OK || len != expect)
                fail("corrupt OpenEXR zip data");
            unpredict(tmp, raw);
            break;
        }
        case kExrPiz: piz_decode(data, size, chans, w, h, raw); break;
        case kExrPxr24: pxr24_decode(data, size, chans, w, h, raw); break;
        case kExrB44:
        case kExrB44a: b44_decode(data, size, chans, w, h, raw); break;
        default: fail("bad OpenEXR chunk size");
    }
    if (raw.size() != expect) fail("corrupt OpenEXR chunk");
    return raw.data();
}

// Header of one part: the attributes this reader uses.
struct ExrPart {
    std::vector<Channel> chans;
    int comp = -1;
    bool haveWindow = false;
    int32_t xmin = 0, ymin = 0, xmax = -1, ymax = -1;
    bool haveTiles = false;
    uint32_t tileW = 0, tileH = 0, levelMode = 0;
    std::string type;
    int64_t chunkCount = -1;
};

std::string exr_cstr(Cursor& c) {
    std::string s;
    while (true) {
        const uint8_t b = c.u8();
        if (!b) break;
        s.push_back(static_cast<char>(b));
        if (s.size() > 255) fail("malformed OpenEXR header");
    }
    return s;
}

// Reads attributes up to the empty name that ends a header; false for an
// empty header (the end of a multi-part file's header list).
bool exr_header(Cursor& c, ExrPart& part) {
    bool any = false;
    while (true) {
        const std::string name = exr_cstr(c);
        if (name.empty()) return any;
        any = true;
        const std::string type = exr_cstr(c);
        const int32_t size = c.i32le();
        if (size < 0) fail("malformed OpenEXR attribute");
        Cursor v{c.take(static_cast<size_t>(size)), static_cast<size_t>(size)};
        if (name == "channels" && type == "chlist") {
            while (true) {
                const std::string cn = exr_cstr(v);
                if (cn.empty()) break;
                Channel ch{cn, v.i32le(), false};
                ch.pLinear = v.u8() != 0;
                v.take(3);  // reserved
                const int32_t xs = v.i32le(), ys = v.i32le();
                if (ch.type < 0 || ch.type > 2) fail("unknown OpenEXR pixel type");
                if (xs != 1 || ys != 1) fail("subsampled OpenEXR channels are not supported");
                part.chans.push_back(ch);
            }
        } else if (name == "compression" && type == "compression") {
            part.comp = v.u8();
        } else if (name == "dataWindow" && type == "box2i") {
            part.xmin = v.i32le();
            part.ymin = v.i32le();
            part.xmax = v.i32le();
            part.ymax = v.i32le();
            part.haveWindow = true;
        } else if (name == "tiles" && type == "tiledesc") {
            part.tileW = v.u32le();
            part.tileH = v.u32le();
            part.levelMode = v.u8() & 0x0fu;
            part.haveTiles = true;
        } else if (name == "type" && type == "string") {
            part.type.assign(reinterpret_cast<const char*>(v.p), v.n);
        } else if (name == "chunkCount" && type == "int") {
            part.chunkCount = v.i32le();
        }
    }
}

// Decodes the first part (OIIO's subimage 0): scanline or tiled (level 0 of
// one-level, mipmap or ripmap files), single- or multi-part.
void read_exr(Cursor& c, Image& img) {
    if (c.u32le() != 20000630u) fail("not an OpenEXR file");
    const uint32_t ver = c.u32le();
    if ((ver & 0xffu) != 2u) fail("unsupported OpenEXR version");
    if (ver & 0x800u) fail("deep OpenEXR files are not supported");
    const bool multi = (ver & 0x1000u) != 0;
    std::vector<ExrPart> parts;
    if (multi) {
        while (true) {
            ExrPart p;
            if (!exr_header(c, p)) break;
            parts.push_back(std::move(p));
            if (parts.size() > 4096) fail("malformed OpenEXR multi-part header");
        }
        if (parts.empty()) fail("OpenEXR multi-part file without parts");
    } else {
        parts.emplace_back();
        if (!exr_header(c, parts[0])) fail("empty OpenEXR header");
    }
    ExrPart& P = parts[0];
    bool tiled;
    if (multi) {
        if (P.type == "deepscanline" || P.type == "deeptile") fail("deep OpenEXR files are not supported");
        if (P.type != "scanlineimage" && P.type != "tiledimage") fail("unknown OpenEXR part type '" + P.type + "'");
        tiled = P.type == "tiledimage";
        if (P.chunkCount < 0) fail("OpenEXR part without chunkCount");
    } else {
        tiled = (ver & 0x200u) != 0;
    }
    if (P.chans.empty() || P.comp < 0 || !P.haveWindow)
        fail("OpenEXR header lacks channels, compression or dataWindow");
    if (P.comp > kExrDwab) fail("unknown OpenEXR compression");
    if (P.comp == kExrDwaa || P.comp == kExrDwab)
        fail(std::string("unsupported OpenEXR compression ") + kExrCodec[P.comp]);
    if (tiled && (!P.haveTiles || P.tileW == 0 || P.tileH == 0 || P.levelMode > 2))
        fail("malformed OpenEXR tile description");
    const int64_t W = int64_t(P.xmax) - P.xmin + 1, H = int64_t(P.ymax) - P.ymin + 1;
    if (W <= 0 || H <= 0) fail("empty OpenEXR data window");
    img.resize(static_cast<uint64_t>(W), static_cast<uint64_t>(H));
    img.channels = static_cast<uint32_t>(P.chans.size());
    // Which stored channels give r, g, b (Load's expansion for missing ones).
    int src[3] = {-1, -1, -1};
    uint32_t nsrc = 0;
    auto find = [&](const char* n) {
        for (size_t i = 0; i < P.chans.size(); ++i)
            if (P.chans[i].name == n) return static_cast<int>(i);
        return -1;
    };
    if (find("R") >= 0) {
        src[0] = find("R");
        nsrc = 1;
        if (find("G") >= 0) {
            src[1] = find("G");
            nsrc = 2;
            if (find("B") >= 0) {
                src[2] = find("B");
                nsrc = 3;
            }
        }
    } else if (find("Y") >= 0) {
        src[0] = find("Y");
        nsrc = 1;
    } else {
        fail("OpenEXR file has no R or Y channel");
    }
    // Chunks of level 0: scanline blocks, or tiles in row-major order (the
    // first entries of a tiled part's offset table).
    uint64_t numX = 1, chunks;
    uint32_t lpb = 1;
    if (tiled) {
        numX = (static_cast<uint64_t>(W) + P.tileW - 1u) / P.tileW;
        const uint64_t numY = (static_cast<uint64_t>(H) + P.tileH - 1u) / P.tileH;
        chunks = numX * numY;
    } else {
        lpb = exr_lines_per_chunk(P.comp);
        chunks = (static_cast<uint64_t>(H) + lpb - 1u) / lpb;
    }
    if (multi && (tiled ? static_cast<uint64_t>(P.chunkCount) < chunks : static_cast<uint64_t>(P.chunkCount) != chunks))
        fail("OpenEXR chunkCount does not match the data window");
    if (chunks > c.left() / 8u) fail("truncated OpenEXR offset table");
    std::vector<uint64_t> offsets(chunks);
    for (auto& o : offsets) o = c.u64le();
    std::vector<uint8_t> raw, tmp;
    std::vector<uint8_t> seen(chunks, 0);
    for (uint64_t k = 0; k < chunks; ++k) {
        if (offsets[k] > c.n) fail("OpenEXR chunk offset out of range");
        Cursor ck{c.p, c.n, static_cast<size_t>(offsets[k])};
        if (multi && ck.i32le() != 0) fail("OpenEXR chunk of another part");
        uint64_t chunk;
        uint32_t x0 = 0, y0, bw, bh;
        if (tiled) {
            const int32_t tx = ck.i32le(), ty = ck.i32le(), lx = ck.i32le(), ly = ck.i32le();
            if (lx != 0 || ly != 0) fail("OpenEXR tile of another level among level 0's");
            if (tx < 0 || ty < 0 || static_cast<uint64_t>(tx) >= numX || static_cast<uint64_t>(ty) >= chunks / numX)
                fail("bad OpenEXR tile coordinates");
            chunk = static_cast<uint64_t>(ty) * numX + static_cast<uint64_t>(tx);
            x0 = static_cast<uint32_t>(tx) * P.tileW;
            y0 = static_cast<uint32_t>(ty) * P.tileH;
            bw = static_cast<uint32_t>(std::min<int64_t>(P.tileW, W - x0));
            bh = static_cast<uint32_t>(std::min<int64_t>(P.tileH, H - y0));
        } else {
            const int32_t yc = ck.i32le();
            if (yc < P.ymin || yc > P.ymax || (int64_t(yc) - P.ymin) % lpb != 0) fail("bad OpenEXR chunk line");
            y0 = static_cast<uint32_t>(int64_t(yc) - P.ymin);
            chunk = y0 / lpb;
            bw = static_cast<uint32_t>(W);
            bh = static_cast<uint32_t>(std::min<int64_t>(lpb, H - y0));
        }
        if (seen[chunk]) fail("duplicate OpenEXR chunk");
        seen[chunk] = 1;
        const int32_t size = ck.i32le();
        if (size < 0) fail("bad OpenEXR chunk size");
        const uint8_t* data = ck.take(static_cast<size_t>(size));
        if (P.comp == kExrNone) {
            size_t expect = 0;
            for (const Channel& ch : P.chans) expect += static_cast<size_t>(bw) * ch.bytes();
            if (static_cast<size_t>(size) != expect * bh) fail("bad OpenEXR chunk size");
        }
        const uint8_t* px = exr_block(P.comp, P.chans, data, static_cast<size_t>(size), bw, bh, raw, tmp);
        std::vector<size_t> chanOff(P.chans.size());
        size_t lineBytes = 0;
        for (size_t i = 0; i < P.chans.size(); ++i) {
            chanOff[i] = lineBytes;
            lineBytes += static_cast<size_t>(bw) * P.chans[i].bytes();
        }
        for (uint32_t l = 0; l < bh; ++l) {
            const uint8_t* row = px + lineBytes * l;
            for (uint32_t x = 0; x < bw; ++x) {
                float v[3];
                for (uint32_t j = 0; j < nsrc; ++j) {
                    const Channel& ch = P.chans[static_cast<size_t>(src[j])];
                    const uint8_t* b = row + chanOff[static_cast<size_t>(src[j])] + static_cast<size_t>(x) * ch.bytes();
                    if (ch.type == 1) {
                        uint16_t hv;
                        std::memcpy(&hv, b, 2);
                        v[j] = half_to_float(hv);
                    } else if (ch.type == 2) {
                        std::memcpy(&v[j], b, 4);
                    } else {
                        uint32_t u;
                        std::memcpy(&u, b, 4);
                        v[j] = static_cast<float>(u);
                    }
                }
                put(img.px(x0 + x, y0 + l), v, nsrc);
            }
        }
    }
}

int load(const uint8_t* data, size_t size, shiro_hip_image* out, char* err, size_t errSize) {
    auto report = [&](const std::string& m) {
        if (err && errSize) {
            std::snprintf(err, errSize, "%s", m.c_str());
        }
        return static_cast<int>(SHIRO_HIP_INVALID);
    };
    if (!out) return report("null output");
    std::memset(out, 0, sizeof(*out));
    if (!data || size < 4) return report("not an image file");
    try {
        Image img;
        Cursor c{data, size};
        if (data[0] == 'P' && (data[1] == 'F' || data[1] == 'f')) {
            img.format = SHIRO_HIP_IMAGE_PFM;
            read_pfm(c, img);
        } else if (data[0] == '#' && data[1] == '?') {
            img.format = SHIRO_HIP_IMAGE_RGBE;
            read_rgbe(c, img);
        } else if (data[0] == 0x76 && data[1] == 0x2f && data[2] == 0x31 && data[3] == 0x01) {
            img.format = SHIRO_HIP_IMAGE_EXR;
            read_exr(c, img);
        } else {
            return report("unrecognised image format (PFM, Radiance RGBE and OpenEXR are read)");
        }
        float* pixels = static_cast<float*>(std::malloc(img.rgb.size() * sizeof(float)));
        if (!pixels) return static_cast<int>(SHIRO_HIP_OOM);
        std::memcpy(pixels, img.rgb.data(), img.rgb.size() * sizeof(float));
        out->pixels = pixels;
        out->width = img.w;
        out->height = img.h;
        out->channels = img.channels;
        out->format = img.format;
        return SHIRO_HIP_OK;
    } catch (const Fail& f) {
        return report(f.msg);
    } catch (const std::bad_alloc&) {
        return report("out of host memory");
    }
}

}  // namespace

extern "C" int shiro_hip_load_image_memory(const void* data, size_t size, shiro_hip_image* out, char* err,
                                           size_t errSize) {
    return load(static_cast<const uint8_t*>(data), size, out, err, errSize);
}

extern "C" int shiro_hip_load_image(const char* path, shiro_hip_image* out, char* err, size_t errSize) {
    if (out) std::memset(out, 0, sizeof(*out));
    if (!path) {
        if (err && errSize) std::snprintf(err, errSize, "null path");
        return SHIRO_HIP_INVALID;
    }
    FILE* f = std::fopen(path, "rb");
    if (!f) {
        if (err && errSize) std::snprintf(err, errSize, "cannot open %s", path);
        return SHIRO_HIP_INVALID;
    }
    std::vector<uint8_t> bytes;
    uint8_t buf[1 << 16];
    size_t n;
    while ((n = std::fread(buf, 1, sizeof(buf), f)) > 0) bytes.insert(bytes.end(), buf, buf + n);
    const bool bad = std::ferror(f) != 0;
    std::fclose(f);
    if (bad) {
        if (err && errSize) std::snprintf(err, errSize, "cannot read %s", path);
        return SHIRO_HIP_INVALID;
    }
    return load(bytes.data(), bytes.size(), out, err, errSize);
}

extern "C" void shiro_hip_free_image(shiro_hip_image* img) {
    if (!img) return;
    std::free(img->pixels);
    std::memset(img, 0, sizeof(*img));
}
