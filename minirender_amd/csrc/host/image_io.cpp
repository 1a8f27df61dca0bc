// EXR and PNG writers (see image_io.hpp). Both formats are written from their public
// specifications with zlib for the deflate streams.
#include "image_io.hpp"

#include <zlib.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <fstream>
#include <limits>
#include <stdexcept>
#include <string>

namespace mrg_host {

float linear_to_srgb(float linear) {
    if (!std::isfinite(linear)) return 0.0f;
    linear = std::max(linear, 0.0f);
    return linear <= 0.0031308f ? 12.92f * linear : 1.055f * std::pow(linear, 1.0f / 2.4f) - 0.055f;
}

uint8_t encode_unorm8(float value) {
    if (!std::isfinite(value)) return 0;
    return static_cast<uint8_t>(std::lround(std::clamp(value, 0.0f, 1.0f) * 255.0f));
}

RgbImage display_image(const RgbImage& linear, double exposure) {
    if (!std::isfinite(exposure) || std::abs(exposure) > 64)
        throw std::invalid_argument("Display exposure must be finite and in [-64,64]");
    RgbImage out = linear;
    const double scale = std::exp2(exposure);
    for (size_t i = 0; i < linear.rgb.size(); ++i) {
        const double v = double(linear.rgb[i]);
        if (!std::isfinite(v)) throw std::invalid_argument("Display input contains nonfinite radiance");
        const double positive = std::max(v, 0.0) * scale;
        out.rgb[i]            = float(positive / (1.0 + positive));
    }
    return out;
}

namespace {

void put32(std::string& s, uint32_t v) {
    for (int k = 0; k < 4; ++k) s += char((v >> (8 * k)) & 0xff);
}
void put64(std::string& s, uint64_t v) {
    for (int k = 0; k < 8; ++k) s += char((v >> (8 * k)) & 0xff);
}
void put_be32(std::string& s, uint32_t v) {
    for (int k = 3; k >= 0; --k) s += char((v >> (8 * k)) & 0xff);
}
void attribute(std::string& h, const char* name, const char* type, const std::string& value) {
    h += name;
    h += '\0';
    h += type;
    h += '\0';
    put32(h, uint32_t(value.size()));
    h += value;
}
std::string deflate(const std::string& data) {
    uLongf      size = compressBound(uLong(data.size()));
    std::string out(size, '\0');
    if (compress2(reinterpret_cast<Bytef*>(&out[0]), &size, reinterpret_cast<const Bytef*>(data.data()), uLong(data.size()),
                  Z_DEFAULT_COMPRESSION) != Z_OK)
        throw std::runtime_error("zlib compression failed");
    out.resize(size);
    return out;
}
void write_file(const std::filesystem::path& path, const std::string& bytes) {
    std::ofstream stream(path, std::ios::binary);
    stream.exceptions(std::ios::badbit | std::ios::failbit);
    stream.write(bytes.data(), std::streamsize(bytes.size()));
    stream.close();
}

}  // namespace

void write_exr(const RgbImage& image, const std::filesystem::path& path) {
    const uint32_t w = image.width, h = image.height;
    if (w == 0 || h == 0 || uint64_t(w) * h > uint64_t(std::numeric_limits<int>::max()))
        throw std::invalid_argument("EXR dimensions exceed codec limits");
    for (float v : image.rgb)
        if (!std::isfinite(v)) throw std::invalid_argument("EXR output contains nonfinite radiance");
    // header (single-part scanline file, version 2)
    std::string file;
    put32(file, 20000630u);
    put32(file, 2u);
    std::string chlist;
    for (const char* name : {"B", "G", "R"}) {  // channels in name order
        chlist += name;
        chlist += '\0';
        put32(chlist, 2u);  // FLOAT
        chlist += std::string(4, '\0');  // pLinear + reserved
        put32(chlist, 1u);
        put32(chlist, 1u);
    }
    chlist += '\0';
    attribute(file, "channels", "chlist", chlist);
    attribute(file, "compression", "compression", std::string(1, char(3)));  // ZIP, 16 scanlines per block
    std::string box;
    put32(box, 0u);
    put32(box, 0u);
    put32(box, w - 1);
    put32(box, h - 1);
    attribute(file, "dataWindow", "box2i", box);
    attribute(file, "displayWindow", "box2i", box);
    attribute(file, "lineOrder", "lineOrder", std::string(1, '\0'));
    std::string one;
    const float f1 = 1.0f, f0 = 0.0f;
    one.append(reinterpret_cast<const char*>(&f1), 4);
    attribute(file, "pixelAspectRatio", "float", one);
    std::string center;
    center.append(reinterpret_cast<const char*>(&f0), 4);
    center.append(reinterpret_cast<const char*>(&f0), 4);
    attribute(file, "screenWindowCenter", "v2f", center);
    attribute(file, "screenWindowWidth", "float", one);
    file += '\0';
    // blocks of 16 scanlines: per line, the B, G and R planes of `w` floats
    const uint32_t blocks = (h + 15) / 16;
    const size_t   table  = file.size();
    file += std::string(size_t(blocks) * 8, '\0');
    std::vector<uint64_t> offsets;
    for (uint32_t b = 0; b < blocks; ++b) {
        const uint32_t y0 = b * 16, y1 = std::min(h, y0 + 16);
        std::string    raw;
        raw.reserve(size_t(y1 - y0) * w * 12);
        for (uint32_t y = y0; y < y1; ++y)
            for (int c = 2; c >= 0; --c)
                for (uint32_t x = 0; x < w; ++x) {
                    const float v = image.rgb[(size_t(y) * w + x) * 3 + size_t(c)];
                    raw.append(reinterpret_cast<const char*>(&v), 4);
                }
        // ZIP: bytes split into even / odd halves, then a byte-delta predictor, then zlib
        std::string tmp(raw.size(), '\0');
        size_t      t1 = 0, t2 = (raw.size() + 1) / 2;
        for (size_t i = 0; i < raw.size(); ++i) tmp[i % 2 == 0 ? t1++ : t2++] = raw[i];
        for (size_t i = tmp.size() - 1; i > 0; --i)
            tmp[i] = char(uint8_t(int(uint8_t(tmp[i])) - int(uint8_t(tmp[i - 1])) + 128 + 256));
        std::string z = deflate(tmp);
        if (z.size() >= raw.size()) z = raw;  // stored uncompressed when deflate does not pay
        offsets.push_back(file.size());
        put32(file, y0);
        put32(file, uint32_t(z.size()));
        file += z;
    }
    for (uint32_t b = 0; b < blocks; ++b) {
        std::string o;
        put64(o, offsets[b]);
        std::memcpy(&file[table + size_t(b) * 8], o.data(), 8);
    }
    write_file(path, file);
}

void write_png(const RgbImage& image, const std::filesystem::path& path) {
    const uint32_t w = image.width, h = image.height;
    if (w == 0 || h == 0) throw std::invalid_argument("PNG dimensions must be nonzero");
    std::string raw;
    raw.reserve(size_t(h) * (size_t(w) * 3 + 1));
    for (uint32_t y = 0; y < h; ++y) {
        raw += '\0';  // filter type 0
        for (size_t i = 0; i < size_t(w) * 3; ++i) raw += char(encode_unorm8(linear_to_srgb(image.rgb[size_t(y) * w * 3 + i])));
    }
    auto chunk = [](std::string& out, const char* type, const std::string& data) {
        put_be32(out, uint32_t(data.size()));
        std::string body = std::string(type, 4) + data;
        out += body;
        put_be32(out, uint32_t(crc32(0L, reinterpret_cast<const Bytef*>(body.data()), uInt(body.size()))));
    };
    std::string png("\x89PNG\r\n\x1a\n", 8);
    std::string ihdr;
    put_be32(ihdr, w);
    put_be32(ihdr, h);
    ihdr += char(8);  // bit depth
    ihdr += char(2);  // RGB
    ihdr += std::string(3, '\0');
    chunk(png, "IHDR", ihdr);
    chunk(png, "IDAT", deflate(raw));
    chunk(png, "IEND", std::string());
    write_file(path, png);
}

}  // namespace mrg_host
