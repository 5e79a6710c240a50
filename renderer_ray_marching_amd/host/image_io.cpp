// Image writers of the per-view outputs: colour PNG, depth PNG and depth .npy
// (src/utils.cu:151-192). The PNG encoder is a plain one (filter 0 on every row, zlib
// deflate); decoders see the same pixels as in the reference's stb / OpenCV files.
#include <zlib.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>

#include "rm_io.h"

namespace rmh {

namespace {

void put_be32(std::vector<unsigned char>& v, uint32_t x) {
    v.push_back((unsigned char)(x >> 24));
    v.push_back((unsigned char)(x >> 16));
    v.push_back((unsigned char)(x >> 8));
    v.push_back((unsigned char)x);
}

void chunk(std::vector<unsigned char>& out, const char type[4], const unsigned char* data,
           size_t len) {
    put_be32(out, (uint32_t)len);
    const size_t start = out.size();
    out.insert(out.end(), type, type + 4);
    if (len) out.insert(out.end(), data, data + len);
    const uLong crc = crc32(0L, out.data() + start, (uInt)(len + 4));
    put_be32(out, (uint32_t)crc);
}

void write_file(const std::string& path, const void* data, size_t len) {
    std::FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) throw std::runtime_error("could not open '" + path + "' for writing");
    const size_t w = len ? std::fwrite(data, 1, len, f) : 0;
    if (std::fclose(f) != 0 || w != len) throw std::runtime_error("write failed for '" + path + "'");
}

// Rows of the texture are bottom-up (row 0 = bottom); outputs are top-down.
inline int64_t flipped(int y, int H) { return (int64_t)(H - 1 - y); }

}  // namespace

void write_png8(const std::string& path, const uint8_t* pix, int W, int H, int channels) {
    if (W <= 0 || H <= 0 || (channels != 1 && channels != 4))
        throw std::runtime_error("write_png8: bad image shape");
    const size_t row = (size_t)W * channels;
    std::vector<unsigned char> raw((row + 1) * (size_t)H);
    for (int y = 0; y < H; ++y) {
        raw[(row + 1) * y] = 0;   // filter type 0 (none)
        std::memcpy(&raw[(row + 1) * y + 1], pix + row * y, row);
    }
    uLongf zlen = compressBound((uLong)raw.size());
    std::vector<unsigned char> z(zlen);
    if (compress2(z.data(), &zlen, raw.data(), (uLong)raw.size(), 6) != Z_OK)
        throw std::runtime_error("write_png8: deflate failed");
    std::vector<unsigned char> out = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1A, '\n'};
    std::vector<unsigned char> ihdr;
    put_be32(ihdr, (uint32_t)W);
    put_be32(ihdr, (uint32_t)H);
    ihdr.push_back(8);                                // bit depth
    ihdr.push_back(channels == 4 ? 6 : 0);            // colour type: RGBA or grey
    ihdr.push_back(0);                                // deflate
    ihdr.push_back(0);                                // adaptive filtering
    ihdr.push_back(0);                                // no interlace
    chunk(out, "IHDR", ihdr.data(), ihdr.size());
    chunk(out, "IDAT", z.data(), zlen);
    chunk(out, "IEND", nullptr, 0);
    write_file(path, out.data(), out.size());
}

// Byte layout of the .npy files the reference writes through npy.hpp (src/npy.hpp:398-440):
// magic, version 1.0, little-endian u16 header length, the dict
// "{'descr': '<f4', 'fortran_order': False, 'shape': (H, W), }", then 16 - (L mod 16) spaces
// (L = 10 + dict length + 1, so 16 spaces when L is already a multiple of 16) and a newline.
void write_npy_f32(const std::string& path, const float* data, int H, int W) {
    if (W < 0 || H < 0) throw std::runtime_error("write_npy_f32: bad shape");
    const std::string dict = "{'descr': '<f4', 'fortran_order': False, 'shape': (" +
                             std::to_string(H) + ", " + std::to_string(W) + "), }";
    const size_t len = 6 + 2 + 2 + dict.size() + 1;
    const std::string pad(16 - len % 16, ' ');
    const uint16_t hlen = (uint16_t)(dict.size() + pad.size() + 1);
    std::string head = "\x93NUMPY";
    head += (char)1;
    head += (char)0;
    head += (char)(hlen & 0xff);
    head += (char)(hlen >> 8);
    head += dict + pad + "\n";
    std::vector<unsigned char> out(head.begin(), head.end());
    const size_t bytes = sizeof(float) * (size_t)W * (size_t)H;
    out.resize(head.size() + bytes);
    if (bytes) std::memcpy(out.data() + head.size(), data, bytes);
    write_file(path, out.data(), out.size());
}

void save_png(const std::string& path, const float* rgba, int W, int H) {
    std::vector<uint8_t> png((size_t)W * H * 4);
    for (int y = 0; y < H; ++y) {
        const float* src = rgba + 4 * (size_t)W * flipped(y, H);
        uint8_t* dst = png.data() + 4 * (size_t)W * y;
        for (int x = 0; x < W; ++x) {
            for (int c = 0; c < 3; ++c) dst[4 * x + c] = (unsigned char)(src[4 * x + c] * 255.0f);
            dst[4 * x + 3] = (unsigned char)(1.0f * 255.0f);   // alpha forced to 1.0
        }
    }
    write_png8(path, png.data(), W, H, 4);
}

void save_depth(const std::string& path, const float* rgba, int W, int H) {
    std::vector<float> depth((size_t)W * H);
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x)
            depth[(size_t)W * y + x] = rgba[4 * ((size_t)W * flipped(y, H) + x) + 3];
    std::vector<uint8_t> png(depth.size());
    const float scale = 255.0f / 100.0f;
    for (size_t i = 0; i < depth.size(); ++i) {
        const float d = depth[i];
        // a depth of 1 marks empty space in the original article's textures
        png[i] = std::abs(d - 1.0f) < 0.00001f ? (uint8_t)0
                                               : (uint8_t)std::clamp(scale * d, 0.0f, 255.0f);
    }
    write_npy_f32(path + ".npy", depth.data(), H, W);
    write_png8(path, png.data(), W, H, 1);
}

}  // namespace rmh
