// Environment-map tables (see mrg_envmap.h). Follows EnvironmentLight::EnvironmentLight
// (src/render/light.cpp:61-114): the sampling grid is the first mip level (Texture's FP64 pixel-area
// averages, src/core/texture.cpp:24-51) of at most 1024 x 512 texels; texel weights are
// luminance(max(texel, 0)) x texel solid angle; the selection power adds strength x the
// level-0 luminance integral to 4 pi luminance(constant).
#include "mrg_envmap.h"

#include <algorithm>
#include <cmath>
#include <stdexcept>

namespace mrg {
namespace {

constexpr double kPi  = 3.14159265358979323846;
constexpr float  kPiF = 3.14159274101257324219f;

struct Level {
    uint32_t            w = 0, h = 0;
    std::vector<float>  rgb;  // [h][w][3]
};

Level downsample(const Level& src) {
    Level t;
    t.w = std::max<uint32_t>(1, src.w / 2);
    t.h = std::max<uint32_t>(1, src.h / 2);
    t.rgb.resize(size_t(t.w) * t.h * 3);
    for (uint32_t y = 0; y < t.h; ++y) {
        for (uint32_t x = 0; x < t.w; ++x) {
            const double left = double(x) * double(src.w) / double(t.w), right = double(x + 1) * double(src.w) / double(t.w);
            const double top = double(y) * double(src.h) / double(t.h), bottom = double(y + 1) * double(src.h) / double(t.h);
            double sum[3] = {0, 0, 0}, weight_sum = 0;
            for (size_t sy = size_t(top); sy < size_t(std::ceil(bottom)); ++sy)
                for (size_t sx = size_t(left); sx < size_t(std::ceil(right)); ++sx) {
                    const double weight = (std::min(right, double(sx + 1)) - std::max(left, double(sx))) *
                                          (std::min(bottom, double(sy + 1)) - std::max(top, double(sy)));
                    for (int c = 0; c < 3; ++c) sum[c] = sum[c] + double(src.rgb[(sy * src.w + sx) * 3 + c]) * weight;
                    weight_sum += weight;
                }
            for (int c = 0; c < 3; ++c) t.rgb[(size_t(y) * t.w + x) * 3 + c] = float(sum[c] / weight_sum);
        }
    }
    return t;
}

double positive_luminance(const float* p) {
    const double r = std::max(double(p[0]), 0.0), g = std::max(double(p[1]), 0.0), b = std::max(double(p[2]), 0.0);
    return (r * 0.2126 + g * 0.7152) + b * 0.0722;
}

double row_angle(uint32_t y, uint32_t w, uint32_t h) {
    const double t0 = kPi * double(y) / double(h), t1 = kPi * double(y + 1) / double(h);
    return (2 * kPi / double(w)) * 2 * std::sin((t0 + t1) * 0.5) * std::sin((t1 - t0) * 0.5);
}

}  // namespace

EnvTables build_env_tables(const float env[3], const mrg_texture* texture, float strength, float rotation_degrees) {
    EnvTables   t;
    const float lum = (env[0] * 0.2126f + env[1] * 0.7152f) + env[2] * 0.0722f;
    t.integral      = 4.0f * kPiF * lum;
    if (!texture) return t;
    t.strength = strength;
    t.rotation = float(std::remainder(double(rotation_degrees), 360.0) / 360.0);
    if (strength == 0) return t;
    Level level;
    level.w = texture->width;
    level.h = texture->height;
    level.rgb.assign(texture->rgb, texture->rgb + size_t(level.w) * level.h * 3);
    double full_integral = 0;
    for (uint32_t y = 0; y < level.h; ++y) {
        const double angle = row_angle(y, level.w, level.h);
        for (uint32_t x = 0; x < level.w; ++x) full_integral += positive_luminance(&level.rgb[(size_t(y) * level.w + x) * 3]) * angle;
    }
    while (level.w > 1024 || level.h > 512) level = downsample(level);
    // DiscreteDistribution (sampling.cpp:109-130) over the grid texels, FP64
    std::vector<double> weights;
    weights.reserve(size_t(level.w) * level.h);
    double map_integral = 0, maximum = 0;
    for (uint32_t y = 0; y < level.h; ++y) {
        const double angle = row_angle(y, level.w, level.h);
        const double t0 = kPi * double(y) / double(level.h), t1 = kPi * double(y + 1) / double(level.h);
        t.rows.push_back(float(angle));
        t.rows.push_back(float(std::cos(t0)));
        t.rows.push_back(float(std::cos(t1)));
        t.rows.push_back(float(2 * std::sin((t0 + t1) * 0.5) * std::sin((t1 - t0) * 0.5)));
        for (uint32_t x = 0; x < level.w; ++x) {
            const double w = positive_luminance(&level.rgb[(size_t(y) * level.w + x) * 3]) * angle;
            weights.push_back(w);
            map_integral += w;
            maximum = std::max(maximum, w);
        }
    }
    std::vector<double> cdf(weights.size() + 1, 0.0);
    for (size_t i = 0; i < weights.size(); ++i) cdf[i + 1] = cdf[i] + (maximum > 0 ? weights[i] / maximum : 1.0);
    const double total = cdf.back();
    for (auto& v : cdf) v /= total;
    t.uniform = map_integral == 0;
    t.gw      = level.w;
    t.gh      = level.h;
    t.cdf.assign(cdf.begin(), cdf.end());
    t.integral = t.integral + t.strength * float(full_integral);
    return t;
}

}  // namespace mrg
