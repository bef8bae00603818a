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
//  - OpenEXR single-part scanline files: NONE, RLE, ZIPS and ZIP
//    compression; HALF, FLOAT and UINT channels; colour from R/G/B, grey
//    from Y.  Tiled, deep, multi-part files and the PIZ/PXR24/B44/DWA codecs
//    are rejected with an error.
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
    uint32_t bytes() const { return type == 1 ? 2u : 4u; }
};

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

void read_exr(Cursor& c, Image& img) {
    if (c.u32le() != 20000630u) fail("not an OpenEXR file");
    const uint32_t ver = c.u32le();
    if ((ver & 0xffu) != 2u) fail("unsupported OpenEXR version");
    if (ver & 0x200u) fail("tiled OpenEXR files are not supported");
    if (ver & 0x800u) fail("deep OpenEXR files are not supported");
    if (ver & 0x1000u) fail("multi-part OpenEXR files are not supported");
    std::vector<Channel> chans;
    int comp = -1;
    bool haveWindow = false;
    int32_t xmin = 0, ymin = 0, xmax = -1, ymax = -1;
    auto cstr = [&](size_t maxLen) {
        std::string s;
        while (true) {
            const uint8_t b = c.u8();
            if (!b) break;
            s.push_back(static_cast<char>(b));
            if (s.size() > maxLen) fail("malformed OpenEXR header");
        }
        return s;
    };
    while (true) {
        const std::string name = cstr(255);
        if (name.empty()) break;
        const std::string type = cstr(255);
        const int32_t size = c.i32le();
        if (size < 0) fail("malformed OpenEXR attribute");
        Cursor v{c.take(static_cast<size_t>(size)), static_cast<size_t>(size)};
        if (name == "channels" && type == "chlist") {
            while (true) {
                std::string cn;
                while (true) {
                    const uint8_t b = v.u8();
                    if (!b) break;
                    cn.push_back(static_cast<char>(b));
                    if (cn.size() > 255) fail("malformed OpenEXR channel list");
                }
                if (cn.empty()) break;
                Channel ch{cn, v.i32le()};
                v.take(4);  // pLinear + reserved
                const int32_t xs = v.i32le(), ys = v.i32le();
                if (ch.type < 0 || ch.type > 2) fail("unknown OpenEXR pixel type");
                if (xs != 1 || ys != 1) fail("subsampled OpenEXR channels are not supported");
                chans.push_back(ch);
            }
        } else if (name == "compression" && type == "compression") {
            comp = v.u8();
        } else if (name == "dataWindow" && type == "box2i") {
            xmin = v.i32le();
            ymin = v.i32le();
            xmax = v.i32le();
            ymax = v.i32le();
            haveWindow = true;
        } else if (name == "tiles") {
            fail("tiled OpenEXR files are not supported");
        }
    }
    if (chans.empty() || comp < 0 || !haveWindow) fail("OpenEXR header lacks channels, compression or dataWindow");
    static const char* kCodec[] = {"NONE", "RLE", "ZIPS", "ZIP", "PIZ", "PXR24", "B44", "B44A", "DWAA", "DWAB"};
    if (comp > 3) fail(std::string("unsupported OpenEXR compression ") + (comp <= 9 ? kCodec[comp] : "?"));
    const uint32_t lpb = comp == 3 ? 16u : 1u;
    const int64_t W = int64_t(xmax) - xmin + 1, H = int64_t(ymax) - ymin + 1;
    if (W <= 0 || H <= 0) fail("empty OpenEXR data window");
    img.resize(static_cast<uint64_t>(W), static_cast<uint64_t>(H));
    img.channels = static_cast<uint32_t>(chans.size());
    // Which stored channels give r, g, b (Load's expansion for missing ones).
    int src[3] = {-1, -1, -1};
    uint32_t nsrc = 0;
    auto find = [&](const char* n) {
        for (size_t i = 0; i < chans.size(); ++i)
            if (chans[i].name == n) return static_cast<int>(i);
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
    std::vector<size_t> chanOff(chans.size());
    size_t lineBytes = 0;
    for (size_t i = 0; i < chans.size(); ++i) {
        chanOff[i] = lineBytes;
        lineBytes += static_cast<size_t>(W) * chans[i].bytes();
    }
    const uint64_t chunks = (static_cast<uint64_t>(H) + lpb - 1u) / lpb;
    std::vector<uint64_t> offsets(chunks);
    for (auto& o : offsets) o = c.u64le();
    std::vector<uint8_t> raw, tmp;
    std::vector<uint8_t> seen(chunks, 0);
    for (uint64_t k = 0; k < chunks; ++k) {
        if (offsets[k] > c.n) fail("OpenEXR chunk offset out of range");
        Cursor ck{c.p, c.n, static_cast<size_t>(offsets[k])};
        const int32_t y0 = ck.i32le();
        const int32_t size = ck.i32le();
        if (y0 < ymin || y0 > ymax || (int64_t(y0) - ymin) % lpb != 0) fail("bad OpenEXR chunk line");
        const uint64_t chunk = static_cast<uint64_t>(int64_t(y0) - ymin) / lpb;
        if (seen[chunk]) fail("duplicate OpenEXR chunk");
        seen[chunk] = 1;
        if (size < 0) fail("bad OpenEXR chunk size");
        const uint32_t lines = static_cast<uint32_t>(std::min<int64_t>(lpb, int64_t(ymax) - y0 + 1));
        const size_t expect = lineBytes * lines;
        const uint8_t* data = ck.take(static_cast<size_t>(size));
        const uint8_t* px;
        if (static_cast<size_t>(size) == expect || comp == 0) {
            if (static_cast<size_t>(size) != expect) fail("bad OpenEXR chunk size");
            px = data;  // stored uncompressed
        } else if (comp == 1) {
            exr_rle(data, static_cast<size_t>(size), tmp, expect);
            unpredict(tmp, raw);
            px = raw.data();
        } else {
            tmp.resize(expect);
            uLongf len = static_cast<uLongf>(expect);
            if (uncompress(tmp.data(), &len, data, static_cast<uLong>(size)) != Z_OK || len != expect)
                fail("corrupt OpenEXR zip data");
            unpredict(tmp, raw);
            px = raw.data();
        }
        for (uint32_t l = 0; l < lines; ++l) {
            const uint8_t* row = px + lineBytes * l;
            const uint32_t y = static_cast<uint32_t>(int64_t(y0) - ymin + l);
            for (uint32_t x = 0; x < img.w; ++x) {
                float v[3];
                for (uint32_t j = 0; j < nsrc; ++j) {
                    const Channel& ch = chans[static_cast<size_t>(src[j])];
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
                put(img.px(x, y), v, nsrc);
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
