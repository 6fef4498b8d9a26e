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
    err = "unsupported image format";
    return false;
}

}  // namespace lrt
