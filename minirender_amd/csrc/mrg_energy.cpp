// Kulla-Conty energy table: the reference's bake_ibl_brdf() at its defaults (64 x 64 nodes, 512
// samples per node, seed 101; src/render/ibl.cpp:240-297) and store_average_energy()
// (src/render/microfacet_energy.cpp:36-54), in FP64 on the host.
//
// Each node integrates the white GGX conductor's VNDF sample weights f cos / pdf (Heitz 2018
// visible-normal sampling; src/render/bsdf.cpp:36-48, :72-85, :148-277) over a Hammersley point
// set shifted by a per-node random offset, split by the Schlick factor into (A, B). The operation
// order follows the reference expression by expression so the float table is reproducible.
#include "mrg_energy.h"

#include <algorithm>
#include <cmath>
#include <cstdint>

namespace mrg {
namespace {

constexpr double kPi = 3.14159265358979323846;

struct D3 {
    double x, y, z;
};
D3     add(D3 a, D3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
D3     sub(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
D3     scale(double s, D3 a) { return {s * a.x, s * a.y, s * a.z}; }
double dot(D3 a, D3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
D3     cross(D3 a, D3 b) { return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y}; }
D3     normalize(D3 v) { return scale(1.0 / std::sqrt(dot(v, v)), v); }

// PCG32 stream of Sampler::for_sample (src/core/sampling.cpp:28-60), FP64 dimensions.
struct Stream {
    uint64_t state = 0, inc = 1;
    static uint64_t mix(uint64_t v) {
        v = (v ^ (v >> 30U)) * 0xbf58476d1ce4e5b9ULL;
        v = (v ^ (v >> 27U)) * 0x94d049bb133111ebULL;
        return v ^ (v >> 31U);
    }
    Stream(uint64_t seed, uint64_t pixel, uint64_t sample) {
        const uint64_t s = mix(seed ^ mix(pixel + 0x9e3779b97f4a7c15ULL));
        inc              = ((mix(sample + 0xd1b54a32d192ed03ULL) ^ mix(pixel)) << 1U) | 1U;
        next();
        state += s;
        next();
    }
    uint32_t next() {
        const uint64_t prev  = state;
        state                = prev * 6364136223846793005ULL + inc;
        const auto     xored = static_cast<uint32_t>(((prev >> 18U) ^ prev) >> 27U);
        const auto     rot   = static_cast<uint32_t>(prev >> 59U);
        return (xored >> rot) | (xored << ((0U - rot) & 31U));
    }
    double uniform() {
        const uint32_t hi = next() >> 5U;
        const uint32_t lo = next() >> 6U;
        return (static_cast<double>(hi) * 67108864.0 + lo) / 9007199254740992.0;
    }
};

double bit_reverse_fraction(uint32_t bits) {
    bits = (bits << 16U) | (bits >> 16U);
    bits = ((bits & 0x55555555U) << 1U) | ((bits & 0xAAAAAAAAU) >> 1U);
    bits = ((bits & 0x33333333U) << 2U) | ((bits & 0xCCCCCCCCU) >> 2U);
    bits = ((bits & 0x0F0F0F0FU) << 4U) | ((bits & 0xF0F0F0F0U) >> 4U);
    bits = ((bits & 0x00FF00FFU) << 8U) | ((bits & 0xFF00FF00U) >> 8U);
    return static_cast<double>(bits) * 0x1p-32;
}

double ggx(D3 m, double alpha) {
    const double a2 = alpha * alpha;
    const double q  = (m.x * m.x + m.y * m.y) + a2 * m.z * m.z;
    return m.z > 0 ? a2 / (kPi * q * q) : 0.0;
}

double smith_over_cos(D3 v, double alpha) {
    if (v.z <= 0) return 0.0;
    return 2 / (v.z + std::sqrt(v.z * v.z + alpha * alpha * (v.x * v.x + v.y * v.y)));
}

// Visible-normal sample of the GGX distribution for the stretched view direction.
D3 visible_normal(D3 wo, double ux, double uy, double alpha) {
    const D3     v     = normalize(D3{alpha * wo.x, alpha * wo.y, wo.z});
    const double lensq = v.x * v.x + v.y * v.y;
    const double len   = std::sqrt(lensq);
    const D3     t1    = lensq > 0 ? D3{-v.y / len, v.x / len, 0.0 / len} : D3{1, 0, 0};
    const D3     t2    = cross(v, t1);
    const double r     = std::sqrt(ux);
    const double px    = r * std::cos(2 * kPi * uy);
    const double blend = 0.5 * (1 + v.z);
    const double py    = (1 - blend) * std::sqrt(std::max(0.0, 1 - px * px)) + blend * r * std::sin(2 * kPi * uy);
    const D3     n     = add(add(scale(px, t1), scale(py, t2)), scale(std::sqrt(std::max(0.0, (1 - px * px) - py * py)), v));
    return normalize(D3{alpha * n.x, alpha * n.y, std::max(0.0, n.z)});
}

// (A, B) of one node: ibl.cpp:240-277 with the white conductor (F0 = 1, so the sample weight is
// D G1 G1 / 4 * cos / pdf without Fresnel).
void node_albedo(double mu, double roughness, uint64_t seed, double& a, double& b) {
    if (roughness * roughness <= 1e-6) {
        const double f = std::pow(1 - mu, 5);
        a              = 1 - f;
        b              = f;
        return;
    }
    const double cosine = std::max(mu, 1e-6);
    const D3     wo     = normalize(D3{std::sqrt(std::max(0.0, 1 - cosine * cosine)), 0, cosine});
    const float  r32    = static_cast<float>(roughness);
    const double alpha  = static_cast<double>(r32) * r32;
    Stream       shift(seed, 0, 0);
    const double sx = shift.uniform();
    const double sy = shift.uniform();
    double       sa = 0, sb = 0;
    for (uint32_t i = 0; i < 512; ++i) {
        const double px = (static_cast<double>(i) + 0.5) / 512u + sx;
        const double py = bit_reverse_fraction(i) + sy;
        const double ux = px - std::floor(px), uy = py - std::floor(py);
        const D3 m  = visible_normal(wo, ux, uy, alpha);
        const D3 wi = normalize(sub(scale(2 * dot(wo, m), m), wo));
        if (wi.z <= 0) continue;
        const D3     ewi     = normalize(wi);
        const D3     h       = normalize(add(wo, ewi));
        const double glossy  = ggx(h, alpha) * smith_over_cos(wo, alpha) / 4;
        // Schlick with F0 = 1 is exactly 1
        const double value   = ((1.0 * ggx(h, alpha)) * smith_over_cos(wo, alpha)) * smith_over_cos(ewi, alpha) / 4.0;
        const double weight  = value * wi.z / glossy;
        const D3     halfway = normalize(add(wo, wi));
        const double fc      = std::pow(1 - std::clamp(dot(wo, halfway), 0.0, 1.0), 5);
        sa += weight * (1 - fc);
        sb += weight * fc;
    }
    a = sa / 512.0;
    b = sb / 512.0;
}

std::vector<float> bake() {
    std::vector<float> t(static_cast<size_t>(kEnergyRes) * kEnergyRes * 4, 0.0f);
    for (int y = 0; y < kEnergyRes; ++y) {
        for (int x = 0; x < kEnergyRes; ++x) {
            double a, b;
            node_albedo(static_cast<double>(x) / (kEnergyRes - 1), static_cast<double>(y) / (kEnergyRes - 1),
                        101 + static_cast<uint64_t>(y * kEnergyRes + x) * 0x9e3779b97f4a7c15ULL, a, b);
            float* n = &t[(static_cast<size_t>(y) * kEnergyRes + x) * 4];
            n[0]     = static_cast<float>(a);
            n[1]     = static_cast<float>(b);
        }
        // 2 * int E(mu) mu dmu over the piecewise-linear E of the row (exact Simpson per interval)
        double       integral = 0;
        const double step     = 1. / static_cast<double>(kEnergyRes - 1);
        for (int x = 1; x < kEnergyRes; ++x) {
            const float* p     = &t[(static_cast<size_t>(y) * kEnergyRes + x - 1) * 4];
            const float* q     = &t[(static_cast<size_t>(y) * kEnergyRes + x) * 4];
            const double left  = std::clamp(static_cast<double>(p[0] + p[1]), 0., 1.);
            const double right = std::clamp(static_cast<double>(q[0] + q[1]), 0., 1.);
            const double mu    = static_cast<double>(x) * step;
            integral += step / 3 * (((mu - step) * left + 2 * (mu - step * .5) * (left + right)) + mu * right);
        }
        for (int x = 0; x < kEnergyRes; ++x) t[(static_cast<size_t>(y) * kEnergyRes + x) * 4 + 2] = static_cast<float>(std::clamp(integral, 0., 1.));
    }
    return t;
}

}  // namespace

const std::vector<float>& conductor_energy_table() {
    static const std::vector<float> table = bake();
    return table;
}

}  // namespace mrg
