// Image files for rendered frames.
//
// 8-bit formats follow ImageWriter::save (image_write.cpp:10-100): each
// channel goes through to_byte (NaN -> 0, pow(v, 1/2.2), clamp to
// [0, 0.999], times 256, truncated), rows optionally flipped (the frame
// loop stores row 0 at the bottom, src/main.cpp:151 saves flipped).  PNG is
// written with stored (uncompressed) deflate blocks, so the file bytes differ
// from stb_image_write's compressed stream while the decoded pixels are the
// same; PPM is the reference's plain-text P3.  The reference flips only the
// PNG / JPEG bytes: PPM and HDR keep the rows in memory order
// (image_write.cpp:35-50,62-70,81).  HDR is Radiance RGBE (Ward's
// shared-exponent encoding, flat scanlines).  Lossless float formats: PFM
// (little endian, bottom row first) and raw little-endian float32 with a
// JSON sidecar.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "host.hpp"

namespace lrt {
namespace {

uint32_t crc_table[256];
bool crc_ready = false;

uint32_t crc32(const uint8_t* p, size_t n, uint32_t c = 0xffffffffu) {
    if (!crc_ready) {
        for (uint32_t i = 0; i < 256; ++i) {
            uint32_t v = i;
            for (int k = 0; k < 8; ++k) v = (v & 1u) ? 0xedb88320u ^ (v >> 1) : v >> 1;
            crc_table[i] = v;
        }
        crc_ready = true;
    }
    for (size_t i = 0; i < n; ++i) c = crc_table[(c ^ p[i]) & 0xffu] ^ (c >> 8);
    return c;
}

void put_be32(std::vector<uint8_t>& b, uint32_t v) {
    b.push_back((uint8_t)(v >> 24));
    b.push_back((uint8_t)(v >> 16));
    b.push_back((uint8_t)(v >> 8));
    b.push_back((uint8_t)v);
}

void png_chunk(std::vector<uint8_t>& out, const char* type, const std::vector<uint8_t>& data) {
    put_be32(out, (uint32_t)data.size());
    const size_t start = out.size();
    out.insert(out.end(), type, type + 4);
    out.insert(out.end(), data.begin(), data.end());
    put_be32(out, crc32(out.data() + start, out.size() - start) ^ 0xffffffffu);
}

bool write_file(const char* path, const void* data, size_t n) {
    FILE* f = std::fopen(path, "wb");
    if (!f) return false;
    const bool ok = std::fwrite(data, 1, n, f) == n;
    return std::fclose(f) == 0 && ok;
}

// (unsigned char)v as the reference's x86-64 build computes it: truncation
// to a 32-bit integer (NaN and out-of-range values give INT_MIN), low byte
uint8_t trunc_u8(float v) {
    if (!(v > -2147483648.0f && v < 2147483648.0f)) return 0;
    return (uint8_t)(uint32_t)(int32_t)v;
}

// OpenEXR 2 single-part scanline file: uncompressed, one scanline per block,
// channels B, G, R (the format lists channels alphabetically) as 32-bit float,
// rows top first (lineOrder INCREASING_Y).  Frame rows are stored bottom first,
// so file row y is frame row h - 1 - y (row y when the input is flipped).
void put_le32(std::vector<uint8_t>& o, uint32_t v) {
    for (int i = 0; i < 4; ++i) o.push_back((uint8_t)(v >> (8 * i)));
}
void put_le64(std::vector<uint8_t>& o, uint64_t v) {
    for (int i = 0; i < 8; ++i) o.push_back((uint8_t)(v >> (8 * i)));
}
void put_f32(std::vector<uint8_t>& o, float v) {
    uint32_t b;
    std::memcpy(&b, &v, 4);
    put_le32(o, b);
}
void exr_attr(std::vector<uint8_t>& o, const char* name, const char* type, const std::vector<uint8_t>& value) {
    o.insert(o.end(), name, name + std::strlen(name) + 1);
    o.insert(o.end(), type, type + std::strlen(type) + 1);
    put_le32(o, (uint32_t)value.size());
    o.insert(o.end(), value.begin(), value.end());
}

bool write_exr(const char* path, const float* rgb, uint32_t w, uint32_t h, bool flip, std::string& err) {
    std::vector<uint8_t> o;
    put_le32(o, 20000630u);  // magic
    put_le32(o, 2u);         // version 2, single-part scanline
    std::vector<uint8_t> v;
    for (const char* ch : {"B", "G", "R"}) {
        v.push_back((uint8_t)ch[0]);
        v.push_back(0);
        put_le32(v, 2u);  // FLOAT
        for (int i = 0; i < 4; ++i) v.push_back(0);  // pLinear, reserved
        put_le32(v, 1u);  // x sampling
        put_le32(v, 1u);  // y sampling
    }
    v.push_back(0);
    exr_attr(o, "channels", "chlist", v);
    exr_attr(o, "compression", "compression", {0});
    v.clear();
    for (uint32_t x : {0u, 0u, w - 1, h - 1}) put_le32(v, x);
    exr_attr(o, "dataWindow", "box2i", v);
    exr_attr(o, "displayWindow", "box2i", v);
    exr_attr(o, "lineOrder", "lineOrder", {0});
    v.clear();
    put_f32(v, 1.0f);
    exr_attr(o, "pixelAspectRatio", "float", v);
    v.clear();
    put_f32(v, 0.0f);
    put_f32(v, 0.0f);
    exr_attr(o, "screenWindowCenter", "v2f", v);
    v.clear();
    put_f32(v, 1.0f);
    exr_attr(o, "screenWindowWidth", "float", v);
    o.push_back(0);  // end of header
    const uint32_t line_bytes = 3 * 4 * w;
    const uint64_t table = o.size(), first = table + 8ull * h;
    for (uint32_t y = 0; y < h; ++y) put_le64(o, first + (uint64_t)y * (8 + line_bytes));
    for (uint32_t y = 0; y < h; ++y) {
        put_le32(o, y);
        put_le32(o, line_bytes);
        const uint32_t src = flip ? y : h - 1 - y;
        const float* row = rgb + 3 * (size_t)w * src;
        for (int c : {2, 1, 0})
            for (uint32_t x = 0; x < w; ++x) put_f32(o, row[3 * (size_t)x + c]);
    }
    if (!write_file(path, o.data(), o.size())) { err = std::string("cannot write ") + path; return false; }
    return true;
}

}  // namespace

uint8_t image_to_byte(float value) {
    if (std::isnan(value)) value = 0.0f;
    value = std::pow(value, 1.0f / 2.2f);
    value = value < 0.0f ? 0.0f : (0.999f < value ? 0.999f : value);  // std::clamp(value, 0, 0.999)
    return trunc_u8(value * 256.0f);
}

void image_to_bytes(const float* rgb, uint32_t w, uint32_t h, bool flip, uint8_t* out) {
    for (uint32_t y = 0; y < h; ++y) {
        const uint32_t src = flip ? h - 1 - y : y;
        for (uint32_t i = 0; i < 3 * w; ++i) out[3 * (size_t)w * y + i] = image_to_byte(rgb[3 * (size_t)w * src + i]);
    }
}

bool write_image(const char* path, const float* rgb, uint32_t w, uint32_t h, int format, bool flip, std::string& err) {
    const size_t n = (size_t)w * h * 3;
    if (format == LRT_IMAGE_PNG || format == LRT_IMAGE_PPM) {
        std::vector<uint8_t> px(n);
        // the reference writes PPM before its row flip (image_write.cpp:35-50)
        image_to_bytes(rgb, w, h, flip && format == LRT_IMAGE_PNG, px.data());
        if (format == LRT_IMAGE_PPM) {
            std::string s = "P3\n" + std::to_string(w) + " " + std::to_string(h) + "\n255\n";
            for (size_t i = 0; i < n; i += 3)
                s += std::to_string(px[i]) + " " + std::to_string(px[i + 1]) + " " + std::to_string(px[i + 2]) + "\n";
            if (!write_file(path, s.data(), s.size())) { err = std::string("cannot write ") + path; return false; }
            return true;
        }
        // zlib stream of stored deflate blocks over the filtered scanlines
        std::vector<uint8_t> raw;
        raw.reserve((size_t)h * (3 * w + 1));
        for (uint32_t y = 0; y < h; ++y) {
            raw.push_back(0);  // filter: none
            raw.insert(raw.end(), px.begin() + 3 * (size_t)w * y, px.begin() + 3 * (size_t)w * (y + 1));
        }
        std::vector<uint8_t> z = {0x78, 0x01};
        for (size_t at = 0; at < raw.size() || raw.empty();) {
            const size_t len = std::min<size_t>(65535, raw.size() - at);
            const bool last = at + len >= raw.size();
            z.push_back(last ? 1 : 0);
            z.push_back((uint8_t)len);
            z.push_back((uint8_t)(len >> 8));
            z.push_back((uint8_t)~len);
            z.push_back((uint8_t)(~len >> 8));
            z.insert(z.end(), raw.begin() + at, raw.begin() + at + len);
            at += len;
            if (last) break;
        }
        uint32_t a = 1, b = 0;  // Adler-32
        for (uint8_t c : raw) {
            a = (a + c) % 65521u;
            b = (b + a) % 65521u;
        }
        put_be32(z, (b << 16) | a);
        std::vector<uint8_t> out = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};
        std::vector<uint8_t> ihdr;
        put_be32(ihdr, w);
        put_be32(ihdr, h);
        for (uint8_t v : {8, 2, 0, 0, 0}) ihdr.push_back(v);  // 8-bit RGB, deflate, no filter set, no interlace
        png_chunk(out, "IHDR", ihdr);
        png_chunk(out, "IDAT", z);
        png_chunk(out, "IEND", {});
        if (!write_file(path, out.data(), out.size())) { err = std::string("cannot write ") + path; return false; }
        return true;
    }
    if (format == LRT_IMAGE_HDR) {
        std::string hdr = "#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y " + std::to_string(h) + " +X " + std::to_string(w) + "\n";
        std::vector<uint8_t> out(hdr.begin(), hdr.end());
        out.reserve(out.size() + (size_t)w * h * 4);
        for (uint32_t y = 0; y < h; ++y) {
            const uint32_t src = y;
            for (uint32_t x = 0; x < w; ++x) {
                // stb_image_write's stbiw__linear_to_rgbe (stb_image_write.h:639-654),
                // whose max is (a > b ? a : b)
                const float* p = rgb + 3 * ((size_t)w * src + x);
                const float gb = p[1] > p[2] ? p[1] : p[2];
                const float m = p[0] > gb ? p[0] : gb;
                if (m < 1e-32f) {
                    for (int k = 0; k < 4; ++k) out.push_back(0);
                    continue;
                }
                int e = 0;
                const float scale = (float)std::frexp((double)m, &e) * 256.0f / m;
                for (int k = 0; k < 3; ++k) out.push_back(trunc_u8(p[k] * scale));
                out.push_back((uint8_t)(e + 128));
            }
        }
        if (!write_file(path, out.data(), out.size())) { err = std::string("cannot write ") + path; return false; }
        return true;
    }
    if (format == LRT_IMAGE_PFM || format == LRT_IMAGE_RAW) {
        // float rows bottom first: the frame's own order, or flipped input
        std::vector<float> rows(n);
        for (uint32_t y = 0; y < h; ++y) {
            const uint32_t src = flip ? h - 1 - y : y;
            std::memcpy(rows.data() + 3 * (size_t)w * y, rgb + 3 * (size_t)w * src, 3 * (size_t)w * sizeof(float));
        }
        std::vector<uint8_t> out;
        if (format == LRT_IMAGE_PFM) {
            const std::string head = "PF\n" + std::to_string(w) + " " + std::to_string(h) + "\n-1.0\n";
            out.assign(head.begin(), head.end());
        }
        const uint8_t* b = (const uint8_t*)rows.data();
        out.insert(out.end(), b, b + n * sizeof(float));
        if (!write_file(path, out.data(), out.size())) { err = std::string("cannot write ") + path; return false; }
        if (format == LRT_IMAGE_RAW) {
            const std::string side = "{\"width\": " + std::to_string(w) + ", \"height\": " + std::to_string(h) +
                                     ", \"channels\": 3, \"dtype\": \"float32\", \"byte_order\": \"little\", "
                                     "\"row_order\": \"bottom_up\"}\n";
            const std::string jp = std::string(path) + ".json";
            if (!write_file(jp.c_str(), side.data(), side.size())) { err = "cannot write " + jp; return false; }
        }
        return true;
    }
    if (format == LRT_IMAGE_EXR) return write_exr(path, rgb, w, h, flip, err);
    err = "unsupported image format";
    return false;
}

// ---- accumulators of a progressive render ----------------------------------
// path: little-endian float64, the per-pixel sums (3 * w * h, the frame's
// rows bottom first) followed by the sums of squares when kept; path + ".json":
// the shape and what a resumed render needs (samples done, next sample index,
// seed, depth limits).
bool write_accum(const char* path, const double* sum, const double* sumsq, const lrt_accum_info& info,
                 std::string& err) {
    const size_t n = 3 * (size_t)info.width * info.height;
    FILE* f = std::fopen(path, "wb");
    if (!f) { err = std::string("cannot write ") + path; return false; }
    bool ok = std::fwrite(sum, sizeof(double), n, f) == n;
    if (ok && sumsq) ok = std::fwrite(sumsq, sizeof(double), n, f) == n;
    ok = std::fclose(f) == 0 && ok;
    if (!ok) { err = std::string("cannot write ") + path; return false; }
    char side[768];
    std::snprintf(side, sizeof(side),
                  "{\"width\": %u, \"height\": %u, \"channels\": 3, \"dtype\": \"float64\", \"byte_order\": \"little\", "
                  "\"row_order\": \"bottom_up\", \"arrays\": [%s], \"spp_done\": %llu, \"next_sample\": %llu, "
                  "\"seed\": %llu, \"max_depth\": %u, \"rr_depth\": %u}\n",
                  info.width, info.height, sumsq ? "\"sum\", \"sumsq\"" : "\"sum\"", (unsigned long long)info.spp_done,
                  (unsigned long long)info.next_sample, (unsigned long long)info.seed, info.max_depth, info.rr_depth);
    const std::string jp = std::string(path) + ".json";
    if (!write_file(jp.c_str(), side, std::strlen(side))) { err = "cannot write " + jp; return false; }
    return true;
}

namespace {
// the unsigned integer after "key": in a flat JSON object; false if absent
bool json_uint(const std::string& doc, const char* key, unsigned long long& out) {
    const std::string k = std::string("\"") + key + "\"";
    size_t at = doc.find(k);
    if (at == std::string::npos) return false;
    at = doc.find(':', at + k.size());
    if (at == std::string::npos) return false;
    const char* p = doc.c_str() + at + 1;
    while (*p == ' ') ++p;
    if (*p < '0' || *p > '9') return false;
    char* end = nullptr;
    out = std::strtoull(p, &end, 10);
    return end != p;
}
}  // namespace

bool read_accum(const char* path, double* sum, double* sumsq, lrt_accum_info& info, std::string& err) {
    const std::string jp = std::string(path) + ".json";
    FILE* f = std::fopen(jp.c_str(), "rb");
    if (!f) { err = "cannot read " + jp; return false; }
    std::string doc;
    char buf[512];
    size_t got;
    while ((got = std::fread(buf, 1, sizeof(buf), f)) > 0) doc.append(buf, got);
    std::fclose(f);
    unsigned long long w, h, done, next, seed, md, rr;
    if (doc.find("\"float64\"") == std::string::npos || !json_uint(doc, "width", w) || !json_uint(doc, "height", h) ||
        !json_uint(doc, "spp_done", done) || !json_uint(doc, "next_sample", next) || !json_uint(doc, "seed", seed) ||
        !json_uint(doc, "max_depth", md) || !json_uint(doc, "rr_depth", rr) || w == 0 || h == 0 || w > 65536 || h > 65536) {
        err = "not a float64 accumulator sidecar: " + jp;
        return false;
    }
    info.width = (uint32_t)w;
    info.height = (uint32_t)h;
    info.spp_done = done;
    info.next_sample = next;
    info.seed = seed;
    info.max_depth = (uint32_t)md;
    info.rr_depth = (uint32_t)rr;
    info.has_sumsq = doc.find("\"sumsq\"") != std::string::npos ? 1 : 0;
    if (!sum) return true;
    const size_t n = 3 * (size_t)w * h;
    f = std::fopen(path, "rb");
    if (!f) { err = std::string("cannot read ") + path; return false; }
    bool ok = std::fread(sum, sizeof(double), n, f) == n;
    if (ok && sumsq) {
        if (info.has_sumsq) ok = std::fread(sumsq, sizeof(double), n, f) == n;
        else std::memset(sumsq, 0, n * sizeof(double));
    }
    std::fclose(f);
    if (!ok) { err = std::string("truncated accumulator file ") + path; return false; }
    return true;
}

}  // namespace lrt
