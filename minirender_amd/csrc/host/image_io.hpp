// Output images of the command line (render_cli.cpp:144-154): the linear float RGB image as an
// OpenEXR file (three FLOAT channels B, G, R, ZIP compression in 16-scanline blocks, as the
// reference's writer asks of its codec, exr_io.cpp:36-90) and the display image as an 8-bit sRGB
// PNG (display_image, path_tracer.cpp:407-424; Image::write_png, raster_io.cpp:115-138).
#pragma once

#include <cstdint>
#include <filesystem>
#include <vector>

namespace mrg_host {

// Row-major [height][width][3] linear RGB.
struct RgbImage {
    uint32_t           width = 0, height = 0;
    std::vector<float> rgb;
};

float   linear_to_srgb(float linear);   // image.cpp:26-32
uint8_t encode_unorm8(float value);     // image.cpp:42-47
// Reinhard per channel after 2^exposure (path_tracer.cpp:407-424); throws on nonfinite input.
RgbImage display_image(const RgbImage& linear, double exposure);

void write_exr(const RgbImage& image, const std::filesystem::path& path);
void write_png(const RgbImage& image, const std::filesystem::path& path);

float srgb_to_linear(float encoded);   // image.cpp:34-40

// Texture image decoding (Image::read, raster_io.cpp:40-113) for PNG files: bit depth 8 or 16, gray,
// gray + alpha, RGB, RGBA or 8-bit palette, non-interlaced, decoded per the PNG specification and
// delivered as three channels (gray replicated, alpha ignored); each channel is value / 255 (or
// / 65535 for 16-bit data), then srgb_to_linear when `srgb`. Other formats are refused by name.
RgbImage read_texture(const std::filesystem::path& path, bool srgb);

}  // namespace mrg_host
