// EXR and PNG writers (see image_io.hpp). Both formats are written from their public
// specifications with zlib for the deflate streams.
#include "image_io.hpp"

#include <zlib.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iterator>
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

namespace mrg_host {

float srgb_to_linear(float encoded) {
    if (!std::isfinite(encoded)) return 0.0f;
    encoded = std::max(encoded, 0.0f);
    return encoded <= 0.04045f ? encoded / 12.92f : std::pow((encoded + 0.055f) / 1.055f, 2.4f);
}

namespace {

uint32_t be32(const unsigned char* p) { return (uint32_t(p[0]) << 24) | (uint32_t(p[1]) << 16) | (uint32_t(p[2]) << 8) | p[3]; }

int paeth(int a, int b, int c) {
    const int p = a + b - c, pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
    return pa <= pb && pa <= pc ? a : (pb <= pc ? b : c);
}

RgbImage decode_png(const std::string& bytes, bool srgb) {
    static const unsigned char sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
    if (bytes.size() < 8 || std::memcmp(bytes.data(), sig, 8) != 0) throw std::runtime_error("Unsupported or corrupt texture image");
    const auto* b = reinterpret_cast<const unsigned char*>(bytes.data());
    size_t      p = 8;
    uint32_t    w = 0, h = 0, depth = 0, type = 0, interlace = 0;
    std::string idat, palette;
    bool        header = false;
    while (p + 8 <= bytes.size()) {
        const uint32_t len = be32(b + p);
        const std::string tag(bytes.data() + p + 4, 4);
        if (p + 12 + uint64_t(len) > bytes.size()) throw std::runtime_error("Unsupported or corrupt texture image");
        const unsigned char* d = b + p + 8;
        if (tag == "IHDR") {
            if (len < 13) throw std::runtime_error("Unsupported or corrupt texture image");
            w = be32(d), h = be32(d + 4), depth = d[8], type = d[9], interlace = d[12];
            header = true;
        } else if (tag == "PLTE") {
            palette.assign(reinterpret_cast<const char*>(d), len);
        } else if (tag == "IDAT") {
            idat.append(reinterpret_cast<const char*>(d), len);
        } else if (tag == "IEND") {
            break;
        }
        p += 12 + len;
    }
    if (!header || w == 0 || h == 0) throw std::runtime_error("Unsupported or corrupt texture image");
    if (w > 16384 || h > 16384 || uint64_t(w) * h > 64ull * 1024 * 1024) throw std::invalid_argument("Texture exceeds the 64-megapixel budget");
    if (interlace != 0) throw std::invalid_argument("Interlaced PNG textures are not supported by this loader");
    uint32_t channels = 0;
    switch (type) {
    case 0: channels = 1; break;
    case 2: channels = 3; break;
    case 3: channels = 1; break;
    case 4: channels = 2; break;
    case 6: channels = 4; break;
    default: throw std::runtime_error("Unsupported or corrupt texture image");
    }
    if (!(depth == 8 || (depth == 16 && type != 3)))
        throw std::invalid_argument("PNG textures of bit depth " + std::to_string(depth) + " are not supported by this loader");
    if (type == 3 && palette.size() < 3) throw std::runtime_error("Unsupported or corrupt texture image");
    const uint32_t bpp      = channels * depth / 8;  // bytes per pixel
    const uint64_t row      = uint64_t(w) * bpp;
    uLongf         raw_size = uLongf((row + 1) * h);
    std::string    raw(raw_size, '\0');
    if (uncompress(reinterpret_cast<Bytef*>(&raw[0]), &raw_size, reinterpret_cast<const Bytef*>(idat.data()), uLong(idat.size())) != Z_OK ||
        raw_size != (row + 1) * h)
        throw std::runtime_error("Image decode failed");
    std::vector<unsigned char> cur(row), prev(row, 0);
    RgbImage img;
    img.width  = w;
    img.height = h;
    img.rgb.resize(size_t(w) * h * 3);
    const float scale = depth == 16 ? 65535.0f : 255.0f;
    for (uint32_t y = 0; y < h; ++y) {
        const auto*         line   = reinterpret_cast<const unsigned char*>(raw.data()) + y * (row + 1);
        const unsigned char filter = line[0];
        for (uint64_t i = 0; i < row; ++i) {
            const int x = line[1 + i];
            const int a = i >= bpp ? cur[i - bpp] : 0, up = prev[i], c = i >= bpp ? prev[i - bpp] : 0;
            int       v;
            switch (filter) {
            case 0: v = x; break;
            case 1: v = x + a; break;
            case 2: v = x + up; break;
            case 3: v = x + (a + up) / 2; break;
            case 4: v = x + paeth(a, up, c); break;
            default: throw std::runtime_error("Image decode failed");
            }
            cur[i] = static_cast<unsigned char>(v & 0xff);
        }
        for (uint32_t x = 0; x < w; ++x) {
            float rgb[3];
            auto  sample = [&](uint32_t ch) {
                const unsigned char* q = &cur[size_t(x) * bpp + ch * (depth / 8)];
                return depth == 16 ? float((uint32_t(q[0]) << 8) | q[1]) : float(q[0]);
            };
            if (type == 3) {
                const uint32_t index = cur[x];
                if (3 * index + 2 >= palette.size()) throw std::runtime_error("Image decode failed");
                for (int k = 0; k < 3; ++k) rgb[k] = float(static_cast<unsigned char>(palette[3 * index + k]));
            } else if (channels <= 2) {
                rgb[0] = rgb[1] = rgb[2] = sample(0);
            } else {
                for (uint32_t k = 0; k < 3; ++k) rgb[k] = sample(k);
            }
            for (int k = 0; k < 3; ++k) {
                const float value                       = rgb[k] / scale;
                img.rgb[(size_t(y) * w + x) * 3 + k] = srgb ? srgb_to_linear(value) : value;
            }
        }
        std::swap(cur, prev);
    }
    return img;
}

}  // namespace

RgbImage read_texture(const std::filesystem::path& path, bool srgb) {
    std::string ext = path.extension().string();
    std::transform(ext.begin(), ext.end(), ext.begin(), [](unsigned char c) { return char(std::tolower(c)); });
    if (ext != ".png") throw std::invalid_argument("Texture format " + ext + " is not supported by this loader (PNG only)");
    std::ifstream in(path, std::ios::binary);
    if (!in) throw std::runtime_error("Cannot open texture image");
    const std::string bytes((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
    return decode_png(bytes, srgb);
}

}  // namespace mrg_host
