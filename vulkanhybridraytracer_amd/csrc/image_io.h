// image_io.h - host-side decoders shared by the image and asset loaders.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace hr {
namespace img {

struct Image {
    uint32_t w = 0, h = 0;
    std::vector<uint8_t> px;  // RGBA8, row-major, row 0 first
};

// PNG / JPEG / Radiance HDR bytes -> RGBA8 as stbi_load(..., STBI_rgb_alpha)
// returns them. For HDR input, hdr_float (if given) also receives the RGBE
// values as float RGBA (alpha 1). Returns false with a message on error.
bool decode(const uint8_t* data, size_t size, Image& out, std::string& err,
            std::vector<float>* hdr_float = nullptr);
bool read_file(const std::string& path, std::vector<uint8_t>& bytes, std::string& err);

}  // namespace img
}  // namespace hr
