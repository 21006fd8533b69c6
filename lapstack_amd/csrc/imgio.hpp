// imgio.hpp -- PNG and TIFF decoding to interleaved RGB u16, and the writers
// of io.rs:273-346 (host code).
#pragma once
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace lsh {
namespace io {

struct Image {
    uint32_t w = 0, h = 0;
    int bits = 0;               // the file's bits per sample (8 or 16)
    std::vector<uint16_t> rgb;  // h * w * 3, 8-bit samples scaled by 257 (unless ext is set)
    uint16_t *ext = nullptr;    // when set: the samples go here instead, at most ext_cap of them
    size_t ext_cap = 0;
    uint16_t *target(size_t n, const std::string &path) {
        if (!ext) {
            rgb.resize(n);
            return rgb.data();
        }
        if (n > ext_cap)
            throw std::runtime_error("cannot decode " + path + ": " + std::to_string(w) + "x" + std::to_string(h) +
                                     ", larger than the frame buffer");
        return ext;
    }
};

// io.rs:62-97 load_rgb for PNG and TIFF files (throws std::runtime_error)
void decode(const std::string &path, Image &out);
// the size and bits per sample (a PNG from its header; a TIFF decoded)
void probe(const std::string &path, uint32_t &w, uint32_t &h, int &bits);
// io.rs:273-315 save_rgb from planes (3 * w * h f32), 16 or 8 bits, TIFF or PNG by extension
void save_rgb(const std::string &path, const float *img3, uint32_t w, uint32_t h, int bits);
// interleaved RGB u16 as a 16-bit TIFF or PNG
void save_rgb16(const std::string &path, const uint16_t *rgb, uint32_t w, uint32_t h);
// io.rs:325-335 save_gray16: [0, scale] -> 16-bit gray PNG, (clamp(v, 0, scale) * 65535 / scale + 0.5) truncated
void save_gray16(const std::string &path, const float *plane, uint32_t w, uint32_t h, float scale);
// io.rs:337-346 save_gray: min-max normalised 8-bit gray PNG
void save_gray(const std::string &path, const float *depth, uint32_t w, uint32_t h);

}  // namespace io
}  // namespace lsh
