// TEST INFRASTRUCTURE ONLY — CPU oracle (FP64 reference restatement + FP32 twin).
// Built by oracle/Makefile with -ffp-contract=off so that every product and sum is rounded
// separately, as written.
//
// Parity pinning (what checks each part; tests/test_oracle_*.py):
//   Sampler / PCG32 / for_sample      KAT Sampler(42,54) (sampling_tests.cpp:59-62); stream
//                                     independence; integer closed form.
//   warps                             analytic moments and bins (sampling_tests.cpp:84-139).
//   DiscreteDistribution              exact masses/indices (sampling_tests.cpp:140-172).
//   RayQuery box + triangle           analytic KATs (intersection_tests.cpp:124-180, :239-246).
//   SAH BVH + traversal               equality with an independent Moeller-Trumbore brute force
//                                     (intersection_tests.cpp:264-311), coincident-centroid cases.
//   BSDFs                             Lambert albedo, GGX null-event mass and 1-ln2 albedo,
//                                     Fresnel/Snell/TIR values (sampling_tests.cpp:174-420).
//   LightSampler                      sqrt(4.5) distance, 4.5^1.5/4 pdf, selection PMFs
//                                     (light_tests.cpp:165-235).
//   trace / sample_pixel / render     constant-environment floor, directional rho*E/pi, mirror and
//                                     closed glass, depth-0 emission, RR mean 2-1/256, thin-lens
//                                     focus, render == mean(sample_pixel) (path_tests.cpp).
//   FP64 pixel bits                   parity unpinned at the bit level (no reference golden image
//                                     exists and the reference cannot be built); pinned only to the
//                                     analytic tolerances above.
#include "mr_oracle.hpp"

#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <cstring>
#include <limits>
#include <numeric>
#include <optional>
#include <stdexcept>
#include <string>
#include <thread>
#include <type_traits>
#include <vector>

namespace orc {

constexpr double kPi = 3.14159265358979323846;
constexpr uint32_t kNoSkip = 0xffffffffu;  // no primitive skipped by a query

template <class R> constexpr R pi_v() { return static_cast<R>(kPi); }

// ----------------------------------------------------------------------------------------------
// Vector algebra with glm's operation order (dot = (x+y)+z, cross as glm, normalize = v*(1/|v|)).
template <class R> struct V3 {
    R x = 0, y = 0, z = 0;
    R&       operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
    const R& operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
template <class R> inline V3<R> operator+(V3<R> a, V3<R> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class R> inline V3<R> operator-(V3<R> a, V3<R> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class R> inline V3<R> operator-(V3<R> a) { return {-a.x, -a.y, -a.z}; }
template <class R> inline V3<R> operator*(V3<R> a, V3<R> b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
template <class R> inline V3<R> operator*(V3<R> a, R s) { return {a.x * s, a.y * s, a.z * s}; }
template <class R> inline V3<R> operator*(R s, V3<R> a) { return {s * a.x, s * a.y, s * a.z}; }
template <class R> inline V3<R> operator/(V3<R> a, R s) { return {a.x / s, a.y / s, a.z / s}; }
template <class R> inline R      dot(V3<R> a, V3<R> b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
template <class R> inline V3<R>  cross(V3<R> a, V3<R> b) {
    return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
template <class R> inline R     length(V3<R> v) { return std::sqrt(dot(v, v)); }
template <class R> inline V3<R> normalize(V3<R> v) { return v * (R(1) / std::sqrt(dot(v, v))); }
template <class R> inline V3<R> vabs(V3<R> v) { return {std::abs(v.x), std::abs(v.y), std::abs(v.z)}; }
template <class R> inline R     vmax(V3<R> v) { return std::max({v.x, v.y, v.z}); }
template <class R> inline bool  finite(V3<R> v) {
    return std::isfinite(v.x) && std::isfinite(v.y) && std::isfinite(v.z);
}
template <class R> inline V3<R> from_float(const float* p) { return {R(p[0]), R(p[1]), R(p[2])}; }
template <class R> inline V3<float> to_float(V3<R> v) { return {float(v.x), float(v.y), float(v.z)}; }
template <class R> inline V3<R> widen(V3<float> v) { return {R(v.x), R(v.y), R(v.z)}; }

// ----------------------------------------------------------------------------------------------
// sin/cos of 2*pi*t for the FP32 twin: quarter-turn reduction (exact in float) followed by the
// Taylor series of sin(2*pi*r) and cos(2*pi*r) for |r| <= 1/8, evaluated in Horner form with
// fmaf. Coefficients are (-1)^n (2*pi)^k / k! rounded to float. The HIP kernels use the same
// method and the same constants, so both sides round identically.
inline void sincos_turns(float t, float& s, float& c) {
    const float q  = std::nearbyint(t * 4.0f);
    const float r  = t - q * 0.25f;
    const float r2 = r * r;
    float ps = std::fma(r2, -0x1.e3075p+3f, 0x1.507834p+5f);
    ps       = std::fma(r2, ps, -0x1.32d2ccp+6f);
    ps       = std::fma(r2, ps, 0x1.466bc6p+6f);
    ps       = std::fma(r2, ps, -0x1.4abbcep+5f);
    ps       = std::fma(r2, ps, 0x1.921fb6p+2f);
    const float s0 = r * ps;
    float pc = std::fma(r2, 0x1.f9d38ap+2f, -0x1.a6d1f2p+4f);
    pc       = std::fma(r2, pc, 0x1.e1f506p+5f);
    pc       = std::fma(r2, pc, -0x1.55d3c8p+6f);
    pc       = std::fma(r2, pc, 0x1.03c1fp+6f);
    pc       = std::fma(r2, pc, -0x1.3bd3ccp+4f);
    const float c0 = std::fma(r2, pc, 1.0f);
    switch (static_cast<int>(q) & 3) {
    case 0: s = s0; c = c0; break;
    case 1: s = c0; c = -s0; break;
    case 2: s = -s0; c = -c0; break;
    default: s = -c0; c = s0; break;
    }
}

// sin/cos(2*pi*u): FP64 follows the reference (std::cos(2*pi*u)); FP32 uses sincos_turns.
template <class R> inline void sincos_2pi(R u, R& s, R& c) {
    if constexpr (std::is_same_v<R, double>) {
        c = std::cos(2 * kPi * u);
        s = std::sin(2 * kPi * u);
    } else {
        sincos_turns(u, s, c);
    }
}

// atan2(y, x) / (2 pi) for the FP32 twin (IEEE sign conventions for zeros): octant reduction to
// t = min/max in [0, 1], t > tan(pi/8) mapped through atan(t) = pi/4 + atan((t - 1) / (t + 1)), then
// the Maclaurin series of arctan (coefficients (-1)^k / (2k + 1), k <= 8; truncation below 3e-9 for
// |r| <= tan(pi/8)) in Horner form with fmaf. The HIP kernels use the same method and constants.
inline float atan2_turns(float y, float x) {
    const float ax = std::fabs(x), ay = std::fabs(y);
    if (ax == 0.0f && ay == 0.0f) return std::signbit(x) ? (std::signbit(y) ? -0.5f : 0.5f) : y * 0.0f;
    const bool  swap = ay > ax;
    const float t    = (swap ? ax : ay) / (swap ? ay : ax);
    const bool  big  = t > 0x1.a8279ap-2f;  // tan(pi/8) = 0.41421356
    const float r    = big ? (t - 1.0f) / (t + 1.0f) : t;
    const float r2   = r * r;
    float p = std::fma(r2, 0x1.e1e1e2p-5f, -0x1.111112p-4f);   // 1/17, -1/15
    p       = std::fma(r2, p, 0x1.3b13b2p-4f);                  // 1/13
    p       = std::fma(r2, p, -0x1.745d18p-4f);                 // -1/11
    p       = std::fma(r2, p, 0x1.c71c72p-4f);                  // 1/9
    p       = std::fma(r2, p, -0x1.24924ap-3f);                 // -1/7
    p       = std::fma(r2, p, 0x1.99999ap-3f);                  // 1/5
    p       = std::fma(r2, p, -0x1.555556p-2f);                 // -1/3
    p       = std::fma(r2, p, 1.0f);
    float a = (r * p) * 0x1.45f306p-3f;                         // / (2 pi)
    if (big) a = 0.125f + a;
    if (swap) a = 0.25f - a;
    if (std::signbit(x)) a = 0.5f - a;
    return std::signbit(y) ? -a : a;
}

// ----------------------------------------------------------------------------------------------
// Sampler (sampling.cpp:28-65).
inline uint64_t mix64(uint64_t v) {
    v = (v ^ (v >> 30U)) * 0xbf58476d1ce4e5b9ULL;
    v = (v ^ (v >> 27U)) * 0x94d049bb133111ebULL;
    return v ^ (v >> 31U);
}

struct Sampler {
    uint64_t state = 0, inc = 1;
    Sampler(uint64_t seed, uint64_t sequence) : inc((sequence << 1U) | 1U) {
        next_u32();
        state += seed;
        next_u32();
    }
    static Sampler for_sample(uint64_t seed, uint64_t pixel, uint64_t sample) {
        return Sampler(mix64(seed ^ mix64(pixel + 0x9e3779b97f4a7c15ULL)),
                       mix64(sample + 0xd1b54a32d192ed03ULL) ^ mix64(pixel));
    }
    uint32_t next_u32() {
        const uint64_t prev = state;
        state               = prev * 6364136223846793005ULL + inc;
        const auto shifted  = static_cast<uint32_t>(((prev >> 18U) ^ prev) >> 27U);
        const auto rot      = static_cast<uint32_t>(prev >> 59U);
        return (shifted >> rot) | (shifted << ((0U - rot) & 31U));
    }
    // 53 random bits per dimension (two u32 draws). The FP32 twin rounds the same 53-bit value to
    // float and clamps it below 1.
    template <class R> R next_1d() {
        const uint32_t high = next_u32() >> 5U;
        const uint32_t low  = next_u32() >> 6U;
        if constexpr (std::is_same_v<R, double>) {
            return (static_cast<double>(high) * 67108864.0 + low) / 9007199254740992.0;
        } else {
            const uint64_t bits  = (static_cast<uint64_t>(high) << 26U) | low;
            const float    value = static_cast<float>(bits) * 0x1p-53f;
            return std::min(value, 0x1.fffffep-1f);
        }
    }
};

template <class R> struct V2 { R x = 0, y = 0; };

template <class R> inline void check_unit(R u) {
    if (!(u >= 0 && u < 1)) throw std::invalid_argument("Sample coordinate must be in [0,1)");
}

// Warps (sampling.cpp:67-107).
template <class R> V2<R> sample_disk(V2<R> u) {
    check_unit(u.x);
    check_unit(u.y);
    const R sx = R(2) * u.x - R(1), sy = R(2) * u.y - R(1);
    if (sx == 0 && sy == 0) return {0, 0};
    const bool horizontal = std::abs(sx) > std::abs(sy);
    const R    radius     = horizontal ? sx : sy;
    R          s, c;
    if constexpr (std::is_same_v<R, double>) {
        const double angle = horizontal ? (kPi / 4) * (sy / sx) : kPi / 2 - (kPi / 4) * (sx / sy);
        c = std::cos(angle);
        s = std::sin(angle);
    } else {
        // the same angle expressed in turns: (pi/4)*q = 2*pi*(q/8)
        const float turns = horizontal ? 0.125f * (sy / sx) : 0.25f - 0.125f * (sx / sy);
        sincos_turns(turns, s, c);
    }
    return {radius * c, radius * s};
}

template <class R> V3<R> sample_sphere(V2<R> u) {
    check_unit(u.x);
    check_unit(u.y);
    const R z      = R(1) - R(2) * u.x;
    const R radius = std::sqrt(std::max(R(0), R(1) - z * z));
    R       s, c;
    sincos_2pi(u.y, s, c);
    return {radius * c, radius * s, z};
}

template <class R> V3<R> sample_hemisphere(V2<R> u) {
    check_unit(u.x);
    check_unit(u.y);
    const R radius = std::sqrt(std::max(R(0), R(1) - u.x * u.x));
    R       s, c;
    sincos_2pi(u.y, s, c);
    return {radius * c, radius * s, u.x};
}

template <class R> V3<R> sample_cosine_hemisphere(V2<R> u) {
    const auto d = sample_disk(u);
    return {d.x, d.y, std::sqrt(std::max(R(0), R(1) - (d.x * d.x + d.y * d.y)))};
}

template <class R> V3<R> sample_triangle(V2<R> u) {
    check_unit(u.x);
    check_unit(u.y);
    const R root = std::sqrt(u.x);
    return {R(1) - root, root * (R(1) - u.y), root * u.y};
}

template <class R> R cosine_hemisphere_pdf(V3<R> d) { return std::max(R(0), d.z) / pi_v<R>(); }

// DiscreteDistribution (sampling.cpp:109-150).
template <class R> struct Discrete {
    std::vector<R> cdf{R(0)};
    Discrete() = default;
    explicit Discrete(const std::vector<R>& w) : cdf(w.size() + 1, R(0)) {
        R maximum = 0;
        for (const R x : w) {
            if (!std::isfinite(x) || x < 0)
                throw std::invalid_argument("Distribution weights must be finite and nonnegative");
            maximum = std::max(maximum, x);
        }
        for (size_t i = 0; i < w.size(); ++i) cdf[i + 1] = cdf[i] + (maximum > 0 ? w[i] / maximum : R(1));
        if (!w.empty()) {
            const R integral = cdf.back();
            for (auto& v : cdf) v /= integral;
        }
    }
    size_t size() const { return cdf.size() - 1; }
    R      probability(size_t i) const {
        if (i >= size()) throw std::out_of_range("Distribution index is out of range");
        return cdf[i + 1] - cdf[i];
    }
    struct Sample { size_t index; R mass; R remapped; };
    std::optional<Sample> sample(R u) const {
        check_unit(u);
        if (size() == 0) return std::nullopt;
        const auto upper = std::upper_bound(cdf.begin(), cdf.end(), u);
        const auto index = static_cast<size_t>(upper - cdf.begin() - 1);
        const R    mass  = probability(index);
        const R    remap = std::min((u - cdf[index]) / mass, std::nextafter(R(1), R(0)));
        return Sample{index, mass, remap};
    }
};

// ----------------------------------------------------------------------------------------------
// Ray query (intersection.cpp). The FP32 twin widens each slab by 8 float ulps relative so the
// box test never culls a triangle the (float) watertight test accepts; the FP64 path keeps the
// reference's one-ulp outward rounding only.
template <class R> struct Ray { V3<R> o, d; };

struct Bounds3 {
    V3<float> lo{std::numeric_limits<float>::infinity(), std::numeric_limits<float>::infinity(),
                 std::numeric_limits<float>::infinity()};
    V3<float> hi{-std::numeric_limits<float>::infinity(), -std::numeric_limits<float>::infinity(),
                 -std::numeric_limits<float>::infinity()};
    bool empty() const { return lo.x > hi.x || lo.y > hi.y || lo.z > hi.z; }
    void expand(V3<float> p) {
        lo = {std::min(lo.x, p.x), std::min(lo.y, p.y), std::min(lo.z, p.z)};
        hi = {std::max(hi.x, p.x), std::max(hi.y, p.y), std::max(hi.z, p.z)};
    }
    void expand(const Bounds3& b) {
        if (!b.empty()) {
            expand(b.lo);
            expand(b.hi);
        }
    }
    double area() const {
        if (empty()) return 0;
        const V3<double> e = widen<double>(hi) - widen<double>(lo);
        return 2 * ((e.x * e.y + e.y * e.z) + e.z * e.x);
    }
};

template <class R> inline R gamma_bound(R n) {
    const R e = n * (std::numeric_limits<R>::epsilon() * R(0.5));
    return e / (R(1) - e);
}

template <class R> inline R difference_of_products(R a, R b, R c, R d) {
    const R cd   = c * d;
    const R corr = std::fma(-c, d, cd);
    return std::fma(a, b, -cd) + corr;
}

template <class R> struct TriangleHit { R t; V3<R> bary; };

template <class R> struct RayQuery {
    V3<R> o, d, inv;
    int   kx = 0, ky = 1, kz = 2;
    R     sx = 0, sy = 0, sz = 1;
    explicit RayQuery(const Ray<R>& ray) : o(ray.o), d(ray.d) {
        for (int a = 0; a < 3; ++a) {
            if (!std::isfinite(o[a]) || !std::isfinite(d[a]))
                throw std::invalid_argument("Ray origin and direction must be finite");
            inv[a] = d[a] == 0 ? R(0) : R(1) / d[a];
        }
        if (std::abs(dot(d, d) - R(1)) > R(1e-4)) throw std::invalid_argument("Ray direction must have unit length");
        const V3<R> ad = vabs(d);
        kz             = ad.x > ad.y ? 0 : 1;
        if (ad.z > ad[kz]) kz = 2;
        kx = (kz + 1) % 3;
        ky = (kx + 1) % 3;
        if (d[kz] < 0) std::swap(kx, ky);
        sx = d[kx] / d[kz];
        sy = d[ky] / d[kz];
        sz = R(1) / d[kz];
    }

    std::optional<R> box(const Bounds3& b, R tmin, R tmax) const {
        if (b.empty()) return std::nullopt;
        R enter = tmin, leave = tmax;
        for (int a = 0; a < 3; ++a) {
            if (d[a] == 0) {
                if (o[a] < R(b.lo[a]) || o[a] > R(b.hi[a])) return std::nullopt;
                continue;
            }
            R n = (R(b.lo[a]) - o[a]) * inv[a];
            R f = (R(b.hi[a]) - o[a]) * inv[a];
            if (n > f) std::swap(n, f);
            n = std::nextafter(n, -std::numeric_limits<R>::infinity());
            f = std::nextafter(f, std::numeric_limits<R>::infinity());
            if constexpr (std::is_same_v<R, float>) {
                n = n - std::abs(n) * 0x1p-20f;
                f = f + std::abs(f) * 0x1p-20f;
            }
            enter = std::max(enter, n);
            leave = std::min(leave, f);
            if (enter > leave) return std::nullopt;
        }
        return enter;
    }

    std::optional<TriangleHit<R>> triangle(const V3<float>* v, R tmin, R tmax) const {
        V3<R> p[3];
        for (int i = 0; i < 3; ++i) {
            const V3<R> t = widen<R>(v[i]) - o;
            p[i]          = {t[kx] - sx * t[kz], t[ky] - sy * t[kz], t[kz] * sz};
        }
        const R a = difference_of_products(p[1].x, p[2].y, p[1].y, p[2].x);
        const R b = difference_of_products(p[2].x, p[0].y, p[2].y, p[0].x);
        const R c = difference_of_products(p[0].x, p[1].y, p[0].y, p[1].x);
        if ((a < 0 || b < 0 || c < 0) && (a > 0 || b > 0 || c > 0)) return std::nullopt;
        const R det = (a + b) + c;
        if (det == 0) return std::nullopt;
        const R t = ((a * p[0].z + b * p[1].z) + c * p[2].z) / det;
        if (t < tmin || t > tmax) return std::nullopt;
        const R g2 = gamma_bound(R(2)), g3 = gamma_bound(R(3)), g5 = gamma_bound(R(5));
        const R x  = std::max({std::abs(p[0].x), std::abs(p[1].x), std::abs(p[2].x)});
        const R y  = std::max({std::abs(p[0].y), std::abs(p[1].y), std::abs(p[2].y)});
        const R z  = std::max({std::abs(p[0].z), std::abs(p[1].z), std::abs(p[2].z)});
        const R ex = g5 * (x + z);
        const R ey = g5 * (y + z);
        const R ee = R(2) * ((g2 * x * y + ey * x) + ex * y);
        const R edge = std::max({std::abs(a), std::abs(b), std::abs(c)});
        const R et   = R(3) * ((g3 * edge * z + ee * z) + g3 * z * edge) / std::abs(det);
        if (t <= et) return std::nullopt;
        return TriangleHit<R>{t, {a / det, b / det, c / det}};
    }
};

// ----------------------------------------------------------------------------------------------
// Scene + SAH BVH (bvh.cpp:31-196).
struct Material {
    uint32_t  kind;
    V3<float> base_color, emission;
    float     roughness, ior;
    bool      double_sided, mirrored;
    bool      compensate_energy = false;  // Kulla-Conty lobe for rough conductors (scene.hpp:125-126)
    int32_t   base_color_texture = -1;    // index into Scene::textures, -1 = none
    uint32_t  texture_filter = 2;         // 0 nearest, 1 bilinear, 2 trilinear (texture.hpp:13)
    int32_t   normal_texture = -1;        // tangent-space normal map (a linear texture), -1 = none
    float     normal_scale = 1;
    bool      normal_flip_y = false;
};

struct Scene;

struct Node {
    Bounds3  bounds;
    uint32_t first = 0, right = 0, count = 0;
};

struct Bvh {
    std::vector<uint32_t> order;
    std::vector<Node>     nodes;
    uint64_t              leaves = 0, max_depth = 0;

    struct BuildPrim { Bounds3 bounds; V3<float> centroid; };

    static uint32_t bin_index(float c, float lo, float hi) {
        const double coord = (static_cast<double>(c) - lo) / (static_cast<double>(hi) - lo);
        return std::min(static_cast<uint32_t>(coord * 12), 11U);
    }

    uint32_t build(uint32_t begin, uint32_t end, uint64_t depth, const std::vector<BuildPrim>& prims) {
        const auto node_index = static_cast<uint32_t>(nodes.size());
        nodes.emplace_back();
        Bounds3 bounds, centroids;
        for (uint32_t i = begin; i < end; ++i) {
            bounds.expand(prims[order[i]].bounds);
            centroids.expand(prims[order[i]].centroid);
        }
        const uint32_t count = end - begin;
        max_depth            = std::max(max_depth, depth);
        auto leaf            = [&] {
            nodes[node_index] = {bounds, begin, 0, count};
            ++leaves;
            return node_index;
        };
        if (count <= 4 || depth >= 48) return leaf();
        struct Bin { Bounds3 b; uint32_t n = 0; };
        double   best_cost = static_cast<double>(count);
        int      best_axis = -1;
        uint32_t best_split = 0;
        const double parent_area = bounds.area();
        for (int axis = 0; axis < 3; ++axis) {
            if (centroids.hi[axis] == centroids.lo[axis]) continue;
            std::array<Bin, 12> bins{};
            for (uint32_t i = begin; i < end; ++i) {
                const auto& p = prims[order[i]];
                auto& bin = bins[bin_index(p.centroid[axis], centroids.lo[axis], centroids.hi[axis])];
                ++bin.n;
                bin.b.expand(p.bounds);
            }
            std::array<Bounds3, 12>  rb;
            std::array<uint32_t, 12> rc{};
            Bounds3                  right;
            uint32_t                 rcount = 0;
            for (uint32_t k = 12; k-- > 0;) {
                right.expand(bins[k].b);
                rcount += bins[k].n;
                rb[k] = right;
                rc[k] = rcount;
            }
            Bounds3  left;
            uint32_t lcount = 0;
            for (uint32_t s = 0; s + 1 < 12; ++s) {
                left.expand(bins[s].b);
                lcount += bins[s].n;
                if (lcount == 0 || rc[s + 1] == 0) continue;
                const double cost = 1 + (left.area() * lcount + rb[s + 1].area() * rc[s + 1]) / parent_area;
                if (cost < best_cost) {
                    best_cost  = cost;
                    best_axis  = axis;
                    best_split = s;
                }
            }
        }
        if (best_axis < 0) return leaf();
        const auto middle = std::partition(order.begin() + begin, order.begin() + end, [&](uint32_t i) {
            return bin_index(prims[i].centroid[best_axis], centroids.lo[best_axis], centroids.hi[best_axis]) <= best_split;
        });
        const auto split = static_cast<uint32_t>(middle - order.begin());
        const auto l     = build(begin, split, depth + 1, prims);
        const auto r     = build(split, end, depth + 1, prims);
        nodes[node_index] = {bounds, l, r, 0};
        return node_index;
    }
};

template <class R> struct PrimHit { uint32_t prim; TriangleHit<R> hit; };

struct Counters {
    uint64_t bounds = 0, triangles = 0;
};

// Light sampler data per precision (light.cpp:197-264).
template <class R> V3<R> vpositive(V3<R> c) { return {std::max(c.x, R(0)), std::max(c.y, R(0)), std::max(c.z, R(0))}; }
template <class R> R luminance(V3<R> c);

// EnvironmentLight (light.cpp:61-191): the constant environment plus an optional latitude-longitude
// map (strength, rotation about +Y) importance-sampled by texel luminance x solid angle on the first
// mip level of at most 1024 x 512, mixed with 5% uniform sphere sampling. The FP64 path follows the
// reference; the FP32 twin evaluates directions with atan2_turns / sincos_turns and float
// arithmetic and takes its per-map tables (selection CDF, per-row solid angles, cos(theta) of the
// row boundaries and their difference 2 sin((t0 + t1) / 2) sin((t1 - t0) / 2)) from the FP64
// construction rounded to float.
struct Texture;
template <class R> struct EnvMap {
    V3<R>              constant{0, 0, 0};
    const Texture*     texture = nullptr;
    R                  strength = 1, rotation = 0;  // FP64: radians; FP32 twin: turns
    R                  integral = 0;
    bool               uniform = true;
    size_t             grid_level = 0;
    uint32_t           gw = 0, gh = 0;
    std::vector<R>     cdf;                          // texel selection CDF over the grid level
    std::vector<R>     solid, cos0, cos1, diff;      // per grid row
};

template <class R> struct Lights {
    struct Area { V3<R> v[3]; V3<R> normal; V3<R> radiance; R area; bool double_sided; uint32_t prim; };
    struct Dir { V3<R> direction; V3<R> radiance; };
    struct Choice { int kind; size_t index; };  // 0 area, 1 directional, 2 environment
    V3<R>               env;
    R                   env_integral = 0;
    std::vector<Area>   areas;
    std::vector<Dir>    dirs;
    std::vector<Choice> choices;
    std::vector<int64_t> area_choice_of_prim;  // replaces the (instance, triangle) map
    int64_t             env_choice = -1;
    Discrete<R>         selection;
    EnvMap<R>           envmap;
};

template <class R> R luminance(V3<R> c) { return dot(c, V3<R>{R(0.2126), R(0.7152), R(0.0722)}); }

// ----------------------------------------------------------------------------------------------
// Texture (texture.cpp): an immutable linear-RGB mip chain, level 0 = the decoded image (row 0 at
// the top), each further level the pixel-area average of the previous one in FP64 (odd sizes keep
// their last row and column). Sampling is float arithmetic with repeat addressing; uv has its
// origin at the bottom-left (OBJ convention), images at the top-left.
struct TexLevel {
    uint32_t               w = 0, h = 0;
    std::vector<V3<float>> px;  // [y * w + x]
    const V3<float>& at(uint32_t x, uint32_t y) const { return px[static_cast<size_t>(y) * w + x]; }
};

inline TexLevel downsample(const TexLevel& src) {
    TexLevel t;
    t.w = std::max<uint32_t>(1, src.w / 2);
    t.h = std::max<uint32_t>(1, src.h / 2);
    t.px.resize(static_cast<size_t>(t.w) * t.h);
    for (uint32_t y = 0; y < t.h; ++y) {
        for (uint32_t x = 0; x < t.w; ++x) {
            const double left   = double(x) * double(src.w) / double(t.w);
            const double right  = double(x + 1) * double(src.w) / double(t.w);
            const double top    = double(y) * double(src.h) / double(t.h);
            const double bottom = double(y + 1) * double(src.h) / double(t.h);
            V3<double>   sum{0, 0, 0};
            double       weight_sum = 0;
            for (size_t sy = static_cast<size_t>(top); sy < static_cast<size_t>(std::ceil(bottom)); ++sy) {
                for (size_t sx = static_cast<size_t>(left); sx < static_cast<size_t>(std::ceil(right)); ++sx) {
                    const double weight = (std::min(right, double(sx + 1)) - std::max(left, double(sx))) *
                                          (std::min(bottom, double(sy + 1)) - std::max(top, double(sy)));
                    sum        = sum + widen<double>(src.at(uint32_t(sx), uint32_t(sy))) * weight;
                    weight_sum += weight;
                }
            }
            t.px[static_cast<size_t>(y) * t.w + x] = to_float(sum / weight_sum);
        }
    }
    return t;
}

struct Texture {
    std::vector<TexLevel> levels;
    Texture(uint32_t w, uint32_t h, const float* rgb) {
        if (w == 0 || h == 0) throw std::invalid_argument("Texture must not be empty");
        if (w > 16384 || h > 16384 || uint64_t(w) * h > 64ull * 1024 * 1024)
            throw std::invalid_argument("Texture exceeds the 64-megapixel budget");
        TexLevel l0;
        l0.w = w;
        l0.h = h;
        l0.px.resize(static_cast<size_t>(w) * h);
        for (size_t i = 0; i < l0.px.size(); ++i) {
            l0.px[i] = {rgb[i * 3], rgb[i * 3 + 1], rgb[i * 3 + 2]};
            if (!finite(l0.px[i])) throw std::invalid_argument("Texture contains nonfinite samples");
        }
        levels.push_back(std::move(l0));
        while (levels.back().w > 1 || levels.back().h > 1) levels.push_back(downsample(levels.back()));
    }
    static int wrap(int c, int extent) { return (c % extent + extent) % extent; }
    static V2<float> image_uv(V2<float> uv) {
        if (!std::isfinite(uv.x) || !std::isfinite(uv.y)) throw std::invalid_argument("Texture coordinates must be finite");
        return {uv.x - std::floor(uv.x), 1.0f - (uv.y - std::floor(uv.y))};
    }
    static V3<float> mix(V3<float> a, V3<float> b, float t) { return a * (1.0f - t) + b * t; }
    V3<float> bilinear(V2<float> uv, size_t level) const {
        const TexLevel& im = levels.at(level);
        const V2<float> c  = image_uv(uv);
        const float     x  = c.x * static_cast<float>(im.w) - 0.5f;
        const float     y  = c.y * static_cast<float>(im.h) - 0.5f;
        const int       ix = static_cast<int>(std::floor(x)), iy = static_cast<int>(std::floor(y));
        const float     tx = x - std::floor(x), ty = y - std::floor(y);
        auto texel = [&](int px, int py) { return im.at(uint32_t(wrap(px, int(im.w))), uint32_t(wrap(py, int(im.h)))); };
        return mix(mix(texel(ix, iy), texel(ix + 1, iy), tx), mix(texel(ix, iy + 1), texel(ix + 1, iy + 1), tx), ty);
    }
    V3<float> sample(V2<float> uv, float lod, uint32_t filter) const {
        if (!std::isfinite(lod)) throw std::invalid_argument("Texture LOD must be finite");
        if (filter == 0) {
            const V2<float> c  = image_uv(uv);
            const TexLevel& im = levels.front();
            const int       x  = static_cast<int>(std::floor(c.x * static_cast<float>(im.w)));
            const int       y  = static_cast<int>(std::floor(c.y * static_cast<float>(im.h)));
            return im.at(uint32_t(wrap(x, int(im.w))), uint32_t(wrap(y, int(im.h))));
        }
        if (filter == 1) return bilinear(uv, 0);
        if (filter != 2) throw std::invalid_argument("Unknown texture filter");
        const float  clamped = std::clamp(lod, 0.0f, static_cast<float>(levels.size() - 1));
        const size_t lower   = static_cast<size_t>(std::floor(clamped));
        const size_t upper   = std::min(lower + 1, levels.size() - 1);
        return mix(bilinear(uv, lower), bilinear(uv, upper), clamped - static_cast<float>(lower));
    }
};

template <class R> V3<R> env_unit(V3<R> d) {
    if (!finite(d) || std::abs(dot(d, d) - R(1)) > R(1e-4)) throw std::invalid_argument("Light direction must be finite and have unit length");
    return normalize(d);
}

template <class R> V2<R> env_coordinates(const EnvMap<R>& E, V3<R> d) {
    d = env_unit(d);
    if constexpr (std::is_same_v<R, double>) {
        double u = (std::atan2(d.z, d.x) + E.rotation) / (2 * kPi);
        u -= std::floor(u);
        return {u, std::atan2(std::hypot(d.x, d.z), d.y) / kPi};
    } else {
        float u = atan2_turns(d.z, d.x) + E.rotation;
        u       = u - std::floor(u);
        return {u, 2.0f * atan2_turns(std::sqrt(d.x * d.x + d.z * d.z), d.y)};
    }
}

template <class R> V3<R> env_lookup(const EnvMap<R>& E, V2<R> uv) {
    const TexLevel& im = E.texture->levels.front();
    const R   x  = uv.x * static_cast<R>(im.w) - R(0.5);
    const R   y  = uv.y * static_cast<R>(im.h) - R(0.5);
    const int ix = static_cast<int>(std::floor(x)), iy = static_cast<int>(std::floor(y));
    const R   tx = x - std::floor(x), ty = y - std::floor(y);
    auto texel = [&](int col, int row) {
        const int w = int(im.w), h = int(im.h);
        return vpositive(widen<R>(im.at(uint32_t((col % w + w) % w), uint32_t(std::clamp(row, 0, h - 1)))));
    };
    auto mix = [](V3<R> a, V3<R> b, R t) { return a * (R(1) - t) + b * t; };
    return mix(mix(texel(ix, iy), texel(ix + 1, iy), tx), mix(texel(ix, iy + 1), texel(ix + 1, iy + 1), tx), ty);
}

template <class R> V3<R> env_evaluate(const EnvMap<R>& E, V3<R> d) {
    if (!E.texture) return E.constant;
    return E.constant + E.strength * env_lookup(E, env_coordinates(E, d));
}

template <class R> R env_pdf(const EnvMap<R>& E, V3<R> d) {
    if (E.uniform) return R(1) / (R(4) * pi_v<R>());
    const V2<R>  uv = env_coordinates(E, d);
    const size_t x  = std::min(static_cast<size_t>(uv.x * static_cast<R>(E.gw)), size_t(E.gw) - 1);
    const size_t y  = std::min(static_cast<size_t>(uv.y * static_cast<R>(E.gh)), size_t(E.gh) - 1);
    const R      probability = E.cdf[y * E.gw + x + 1] - E.cdf[y * E.gw + x];
    return R(0.05) / (R(4) * pi_v<R>()) + ((R(1) - R(0.05)) * probability) / E.solid[y];
}

template <class R> struct EnvSample { V3<R> direction, radiance; R pdf; };

template <class R> EnvSample<R> env_sample(const EnvMap<R>& E, V2<R> u) {
    check_unit(u.x);
    check_unit(u.y);
    V3<R> d;
    if (E.uniform || u.x < R(0.05)) {
        d = sample_sphere(V2<R>{E.uniform ? u.x : u.x / R(0.05), u.y});
    } else {
        const R      t     = (u.x - R(0.05)) / (R(1) - R(0.05));
        check_unit(t);
        const auto   upper = std::upper_bound(E.cdf.begin(), E.cdf.end(), t);
        const size_t index = static_cast<size_t>(upper - E.cdf.begin() - 1);
        const R      mass  = E.cdf[index + 1] - E.cdf[index];
        const R      remap = std::min((t - E.cdf[index]) / mass, std::nextafter(R(1), R(0)));
        const size_t x = index % E.gw, y = index / E.gw;
        const R v = std::clamp(u.y, std::numeric_limits<R>::epsilon() * R(0.5), std::nextafter(R(1), R(0)));
        R cosine, sine;
        if (y < E.gh / 2) {
            const R h = (R(1) - E.cos0[y]) + v * E.diff[y];
            cosine    = R(1) - h;
            sine      = std::sqrt(std::max(R(0), h * (R(2) - h)));
        } else {
            const R h = (R(1) + E.cos1[y]) + (R(1) - v) * E.diff[y];
            cosine    = R(-1) + h;
            sine      = std::sqrt(std::max(R(0), h * (R(2) - h)));
        }
        R sp, cp;
        if constexpr (std::is_same_v<R, double>) {
            const double phi = 2 * kPi * (static_cast<double>(x) + remap) / static_cast<double>(E.gw) - E.rotation;
            cp               = std::cos(phi);
            sp               = std::sin(phi);
        } else {
            sincos_turns((static_cast<float>(x) + remap) / static_cast<float>(E.gw) - E.rotation, sp, cp);
        }
        d = {sine * cp, cosine, sine * sp};
    }
    return {d, env_evaluate(E, d), env_pdf(E, d)};
}

// EnvironmentLight construction (light.cpp:61-114) in FP64; the FP32 twin's tables are its values
// rounded to float.
inline void build_envmap(EnvMap<double>& E, EnvMap<float>& T, V3<float> constant, const Texture* texture, float strength,
                         float rotation_degrees) {
    const double turns = std::remainder(static_cast<double>(rotation_degrees), 360.0);
    E.constant = widen<double>(constant);
    E.texture  = texture;
    E.strength = strength;
    E.rotation = turns * kPi / 180;
    E.integral = 4 * kPi * luminance(E.constant);
    T.constant = constant;
    T.texture  = texture;
    T.strength = strength;
    T.rotation = static_cast<float>(turns / 360.0);
    T.integral = 4.0f * pi_v<float>() * luminance(constant);
    if (!texture || strength == 0) return;
    while (texture->levels[E.grid_level].w > 1024 || texture->levels[E.grid_level].h > 512) ++E.grid_level;
    const TexLevel&     image = texture->levels[E.grid_level];
    std::vector<double> weights;
    weights.reserve(static_cast<size_t>(image.w) * image.h);
    double map_integral = 0;
    for (uint32_t y = 0; y < image.h; ++y) {
        const double t0    = kPi * static_cast<double>(y) / static_cast<double>(image.h);
        const double t1    = kPi * static_cast<double>(y + 1) / static_cast<double>(image.h);
        const double angle = (2 * kPi / static_cast<double>(image.w)) * 2 * std::sin((t0 + t1) * 0.5) * std::sin((t1 - t0) * 0.5);
        E.solid.push_back(angle);
        E.cos0.push_back(std::cos(t0));
        E.cos1.push_back(std::cos(t1));
        E.diff.push_back(2 * std::sin((t0 + t1) * 0.5) * std::sin((t1 - t0) * 0.5));
        for (uint32_t x = 0; x < image.w; ++x) {
            const double weight = luminance(vpositive(widen<double>(image.at(x, y)))) * angle;
            weights.push_back(weight);
            map_integral += weight;
        }
    }
    E.cdf     = Discrete<double>(weights).cdf;
    E.uniform = map_integral == 0;
    E.gw      = image.w;
    E.gh      = image.h;
    double          full_integral = 0;
    const TexLevel& full          = texture->levels.front();
    for (uint32_t y = 0; y < full.h; ++y) {
        const double t0    = kPi * static_cast<double>(y) / static_cast<double>(full.h);
        const double t1    = kPi * static_cast<double>(y + 1) / static_cast<double>(full.h);
        const double angle = (2 * kPi / static_cast<double>(full.w)) * 2 * std::sin((t0 + t1) * 0.5) * std::sin((t1 - t0) * 0.5);
        for (uint32_t x = 0; x < full.w; ++x) full_integral += luminance(vpositive(widen<double>(full.at(x, y)))) * angle;
    }
    E.integral += E.strength * full_integral;
    T.integral = T.integral + T.strength * static_cast<float>(full_integral);
    T.uniform  = E.uniform;
    T.grid_level = E.grid_level;
    T.gw = E.gw;
    T.gh = E.gh;
    for (double v : E.cdf) T.cdf.push_back(static_cast<float>(v));
    for (size_t y = 0; y < E.solid.size(); ++y) {
        T.solid.push_back(static_cast<float>(E.solid[y]));
        T.cos0.push_back(static_cast<float>(E.cos0[y]));
        T.cos1.push_back(static_cast<float>(E.cos1[y]));
        T.diff.push_back(static_cast<float>(E.diff[y]));
    }
}

struct Scene {
    std::vector<std::array<V3<float>, 3>> pos;
    std::vector<std::array<V3<float>, 3>> nrm;
    std::vector<std::array<V2<float>, 3>> uv;   // per-corner texture coordinates (empty: all zero)
    std::vector<Texture>                  textures;
    std::vector<uint32_t>                 inst, tri;
    std::vector<Material>                 materials;
    V3<float>                             environment;
    int32_t                               environment_texture = -1;
    float                                 environment_strength = 1, environment_rotation = 0;
    std::vector<OrcDirectional>           directional;
    Bvh                                   bvh;
    Lights<double>                        lights64;
    Lights<float>                         lights32;

    template <class R> const Lights<R>& lights() const {
        if constexpr (std::is_same_v<R, double>) return lights64; else return lights32;
    }

    template <class R> void build_lights(Lights<R>& L) {
        L.env          = widen<R>(environment);
        L.env_integral = L.envmap.integral;
        L.area_choice_of_prim.assign(pos.size(), -1);
        Bounds3        bounds;
        std::vector<R> weights;
        for (size_t p = 0; p < pos.size(); ++p) {
            const Material& m        = materials[inst[p]];
            const V3<R>     radiance = widen<R>(m.emission);
            V3<R>           v[3];
            for (int k = 0; k < 3; ++k) {
                bounds.expand(pos[p][k]);
                v[k] = widen<R>(pos[p][k]);
            }
            if (luminance(radiance) == 0) continue;
            const V3<R> c    = cross(v[1] - v[0], v[2] - v[0]);
            const R     area = length(c) * R(0.5);
            if (area == 0) continue;
            L.area_choice_of_prim[p] = static_cast<int64_t>(L.choices.size());
            L.choices.push_back({0, L.areas.size()});
            const R orientation = m.mirrored ? R(-1) : R(1);
            typename Lights<R>::Area a;
            for (int k = 0; k < 3; ++k) a.v[k] = v[k];
            a.normal       = (c * orientation) / (R(2) * area);
            a.radiance     = radiance;
            a.area         = area;
            a.double_sided = m.double_sided;
            a.prim         = static_cast<uint32_t>(p);
            L.areas.push_back(a);
            weights.push_back(R(m.double_sided ? 2 : 1) * pi_v<R>() * area * luminance(radiance));
        }
        const V3<R> extent = widen<R>(bounds.hi) - widen<R>(bounds.lo);
        const R projected  = pi_v<R>() * (bounds.empty() ? R(1) : std::max(dot(extent, extent) * R(0.25), R(1e-12)));
        for (const auto& d : directional) {
            const V3<R> radiance = from_float<R>(d.color) * R(d.intensity);
            if (luminance(radiance) > 0) {
                L.choices.push_back({1, L.dirs.size()});
                L.dirs.push_back({-normalize(from_float<R>(d.direction)), radiance});
                weights.push_back(projected * luminance(radiance));
            }
        }
        if (L.env_integral > 0) {
            L.env_choice = static_cast<int64_t>(L.choices.size());
            L.choices.push_back({2, 0});
            weights.push_back(projected * L.env_integral);
        }
        L.selection = Discrete<R>(weights);
    }

    void build() {
        std::vector<Bvh::BuildPrim> prims(pos.size());
        for (size_t p = 0; p < pos.size(); ++p) {
            V3<double> center{0, 0, 0};
            for (int k = 0; k < 3; ++k) {
                const auto w = pos[p][k];
                if (!finite(w)) throw std::invalid_argument("Transformed triangle position is not finite");
                prims[p].bounds.expand(w);
                center = center + widen<double>(w) / 3.0;
            }
            const V3<double> e1 = widen<double>(pos[p][1]) - widen<double>(pos[p][0]);
            const V3<double> e2 = widen<double>(pos[p][2]) - widen<double>(pos[p][0]);
            const V3<double> n  = cross(e1, e2);
            if (dot(n, n) == 0) throw std::invalid_argument("Transform collapses a triangle at float precision");
            prims[p].centroid = to_float(center);
        }
        bvh.order.resize(pos.size());
        std::iota(bvh.order.begin(), bvh.order.end(), 0U);
        if (!pos.empty()) {
            bvh.nodes.reserve(pos.size() * 2);
            bvh.build(0, static_cast<uint32_t>(pos.size()), 0, prims);
        }
        build_envmap(lights64.envmap, lights32.envmap, environment,
                     environment_texture >= 0 ? &textures.at(size_t(environment_texture)) : nullptr, environment_strength,
                     environment_rotation);
        build_lights(lights64);
        build_lights(lights32);
    }

    // Closest (or any) hit; with `second`, also the next-nearest primitive (t2 >= t1).
    // `skip`: a scene-order primitive the query ignores (the emitter a FP32 shadow segment ends on).
    template <class R>
    std::optional<PrimHit<R>> trace(const Ray<R>& ray, R tmin, R tmax, bool any_hit, Counters* st,
                                    std::optional<PrimHit<R>>* second = nullptr, uint32_t skip = kNoSkip) const {
        if (!std::isfinite(tmin) || tmin < 0 || std::isnan(tmax) || tmax < tmin)
            throw std::invalid_argument("Ray interval must satisfy 0 <= minimum <= maximum");
        const RayQuery<R> q(ray);
        if (bvh.nodes.empty()) return std::nullopt;
        auto box = [&](const Bounds3& b, R hi) {
            if (st) ++st->bounds;
            return q.box(b, tmin, hi);
        };
        R          cull = tmax;  // interval.maximum (second-nearest t when tracking two hits)
        const auto root = box(bvh.nodes[0].bounds, cull);
        if (!root) return std::nullopt;
        struct Entry { uint32_t node; R enter; };
        std::array<Entry, 50> stack{};
        size_t                pending = 0;
        stack[pending++]              = {0, *root};
        std::optional<PrimHit<R>> best;
        if (second) second->reset();
        auto less = [](const TriangleHit<R>& h, uint32_t p, const PrimHit<R>& o) {
            return h.t < o.hit.t || (h.t == o.hit.t && p < o.prim);
        };
        while (pending) {
            const Entry e = stack[--pending];
            if (e.enter > cull) continue;
            const Node& node = bvh.nodes[e.node];
            if (node.count) {
                for (uint32_t k = 0; k < node.count; ++k) {
                    const uint32_t p = bvh.order[node.first + k];
                    if (st) ++st->triangles;
                    const auto hit = q.triangle(pos[p].data(), tmin, cull);
                    if (!hit) continue;
                    const Material& m = materials[inst[p]];
                    if (!m.double_sided) {
                        const V3<R> e1 = widen<R>(pos[p][1]) - widen<R>(pos[p][0]);
                        const V3<R> e2 = widen<R>(pos[p][2]) - widen<R>(pos[p][0]);
                        const V3<R> outward = cross(e1, e2) * (m.mirrored ? R(-1) : R(1));
                        if (dot(ray.d, outward) >= 0) continue;
                    }
                    if (p == skip) continue;
                    if (second) {
                        if (!best || less(*hit, p, *best)) {
                            *second = best;
                            best    = PrimHit<R>{p, *hit};
                        } else if (!*second || less(*hit, p, **second)) {
                            *second = PrimHit<R>{p, *hit};
                        }
                        if (*second) cull = (*second)->hit.t;
                        continue;
                    }
                    if (!best || less(*hit, p, *best)) {
                        best = PrimHit<R>{p, *hit};
                        if (any_hit) return best;
                        cull = hit->t;
                    }
                }
            } else {
                const auto l = box(bvh.nodes[node.first].bounds, cull);
                const auto r = box(bvh.nodes[node.right].bounds, cull);
                if (l && r) {
                    const bool ln    = *l <= *r;
                    stack[pending++] = ln ? Entry{node.right, *r} : Entry{node.first, *l};
                    stack[pending++] = ln ? Entry{node.first, *l} : Entry{node.right, *r};
                } else if (l) {
                    stack[pending++] = {node.first, *l};
                } else if (r) {
                    stack[pending++] = {node.right, *r};
                }
            }
        }
        return best;
    }
};

// ----------------------------------------------------------------------------------------------
// Surface hit (bvh.cpp:302-370). Normals are stored as float, as in SurfaceHit.
template <class R> struct SurfaceHit {
    R         t;
    V3<R>     position;
    V3<float> geometric_normal, shading_normal;
    V2<float> uv;
    uint32_t  prim;
    bool      front_face;
    V3<R>     position_error;
};

// uv_tangents (src/geometry/uv_tangents.cpp): unit directions of increasing u and v on a triangle
// from its positions and corner uvs, or nothing for degenerate uvs or geometry.
template <class R> std::optional<std::array<V3<R>, 2>> uv_tangents(const std::array<V3<float>, 3>& p, const std::array<V2<float>, 3>& c) {
    const V3<R> a = widen<R>(p[1]) - widen<R>(p[0]);
    const V3<R> b = widen<R>(p[2]) - widen<R>(p[0]);
    V2<R>       su{R(c[1].x) - R(c[0].x), R(c[1].y) - R(c[0].y)};
    V2<R>       tv{R(c[2].x) - R(c[0].x), R(c[2].y) - R(c[0].y)};
    const R scale = std::max({std::abs(su.x), std::abs(su.y), std::abs(tv.x), std::abs(tv.y)});
    if (!std::isfinite(scale) || scale == 0) return std::nullopt;
    su = {su.x / scale, su.y / scale};
    tv = {tv.x / scale, tv.y / scale};
    const R det = su.x * tv.y - su.y * tv.x;
    if (!std::isfinite(det) || std::abs(det) < R(1e-12)) return std::nullopt;
    const V3<R> u  = (tv.y * a - su.y * b) / det;
    const V3<R> v  = ((-tv.x) * a + su.x * b) / det;
    const R     lu = length(u), lv = length(v);
    if (!std::isfinite(lu) || !std::isfinite(lv) || lu == 0 || lv == 0) return std::nullopt;
    return std::array<V3<R>, 2>{u / lu, v / lv};
}

template <class R> SurfaceHit<R> surface_hit(const Scene& s, const Ray<R>& ray, const PrimHit<R>& h) {
    const auto& P  = s.pos[h.prim];
    const auto& N  = s.nrm[h.prim];
    const V3<R> e1 = widen<R>(P[1]) - widen<R>(P[0]);
    const V3<R> e2 = widen<R>(P[2]) - widen<R>(P[0]);
    V3<R>       geometric = normalize(cross(e1, e2));
    if (s.materials[s.inst[h.prim]].mirrored) geometric = -geometric;
    V3<R> shading{0, 0, 0}, position{0, 0, 0}, absolute{0, 0, 0};
    V2<R> uv{0, 0};
    for (int k = 0; k < 3; ++k) {
        const R w = h.hit.bary[k];
        shading   = shading + w * normalize(widen<R>(N[k]));
        if (!s.uv.empty()) {
            uv.x = uv.x + w * R(s.uv[h.prim][k].x);
            uv.y = uv.y + w * R(s.uv[h.prim][k].y);
        }
        const V3<R> term = w * widen<R>(P[k]);
        position         = position + term;
        absolute         = absolute + vabs(term);
    }
    if (dot(shading, shading) < R(1e-24)) {
        shading = geometric;
    } else {
        shading = normalize(shading);
        if (dot(shading, geometric) < 0) shading = -shading;
    }
    // tangent-space normal map in the triangle's uv tangent frame (bvh.cpp:334-359)
    const Material& mat = s.materials[s.inst[h.prim]];
    if (mat.normal_texture >= 0 && mat.normal_scale > 0) {
        std::array<V2<float>, 3> coords{};
        if (!s.uv.empty()) coords = s.uv[h.prim];
        if (const auto axes = uv_tangents<R>(P, coords)) {
            const V3<R> t0 = (*axes)[0] - shading * dot((*axes)[0], shading);
            if (dot(t0, t0) > R(1e-12)) {
                const V3<R> t = normalize(t0);
                V3<R>       b = cross(shading, t);
                if (dot(b, (*axes)[1]) < 0) b = -b;
                const V3<float> texel = s.textures.at(size_t(mat.normal_texture))
                                            .sample(V2<float>{static_cast<float>(uv.x), static_cast<float>(uv.y)}, 0.0f, mat.texture_filter);
                V3<R> smp = widen<R>(texel) * R(2) - V3<R>{1, 1, 1};
                smp.x     = smp.x * R(mat.normal_scale);
                smp.y     = smp.y * (R(mat.normal_scale) * (mat.normal_flip_y ? R(-1) : R(1)));
                const V3<R> mapped = (t * smp.x + b * smp.y) + shading * smp.z;
                if (dot(mapped, geometric) > R(1e-8)) shading = normalize(mapped);
            }
        }
    }
    SurfaceHit<R> out;
    out.t                = h.hit.t;
    out.position         = position;
    out.geometric_normal = to_float(geometric);
    out.shading_normal   = to_float(shading);
    out.uv               = {static_cast<float>(uv.x), static_cast<float>(uv.y)};
    out.prim             = h.prim;
    out.front_face       = dot(ray.d, geometric) < 0;
    out.position_error   = absolute * (R(7) * std::numeric_limits<R>::epsilon());
    return out;
}

// offset_origin / spawn_ray (path_tracer.cpp:87-109). The numeric floor is FLT_MIN^2 in FP64 and
// FLT_MIN in the FP32 twin (FLT_MIN^2 underflows to zero in float).
template <class R> V3<R> offset_origin(V3<R> position, V3<float> geometric_normal, V3<R> error, V3<R> dir) {
    V3<R> n = normalize(widen<R>(geometric_normal));
    if (dot(n, dir) < 0) n = -n;
    R floor_value;
    if constexpr (std::is_same_v<R, double>)
        floor_value = static_cast<double>(std::numeric_limits<float>::min()) * std::numeric_limits<float>::min();
    else
        floor_value = std::numeric_limits<float>::min();
    const R distance = std::max(dot(vabs(n), error), floor_value);
    V3<R>   origin   = position + distance * n;
    for (int a = 0; a < 3; ++a) {
        if (n[a] != 0)
            origin[a] = std::nextafter(origin[a], std::copysign(std::numeric_limits<R>::infinity(), n[a]));
    }
    return origin;
}

template <class R> Ray<R> spawn_ray(const SurfaceHit<R>& h, V3<R> dir) {
    return {offset_origin(h.position, h.geometric_normal, h.position_error, dir), normalize(dir)};
}

template <class R> R power_heuristic(R a, R b) {
    if (!std::isfinite(a) || !std::isfinite(b) || a < 0 || b < 0)
        throw std::invalid_argument("MIS probabilities must be finite and nonnegative");
    const R scale = std::max(a, b);
    if (scale == 0) return 0;
    const R x = a / scale, y = b / scale;
    return x * x / (x * x + y * y);
}

// ----------------------------------------------------------------------------------------------
// BSDF (bsdf.cpp).
template <class R> struct Frame {
    V3<R> t, b, n;
    explicit Frame(V3<R> normal) {
        const R m = std::max({std::abs(normal.x), std::abs(normal.y), std::abs(normal.z)});
        if (!finite(normal) || m == 0) throw std::invalid_argument("Shading normal must be finite and nonzero");
        n                  = normalize(normal / m);
        const V3<R> helper = std::abs(n.z) < R(0.999) ? V3<R>{0, 0, 1} : V3<R>{0, 1, 0};
        t                  = normalize(cross(helper, n));
        b                  = cross(n, t);
    }
    V3<R> to_local(V3<R> w) const { return {dot(w, t), dot(w, b), dot(w, n)}; }
    V3<R> to_world(V3<R> l) const { return (l.x * t + l.y * b) + l.z * n; }
};

enum Event { diffuse_reflection = 0, glossy_reflection = 1, specular_reflection = 2, specular_transmission = 3 };

template <class R> struct BsdfSample {
    V3<R> direction, weight;
    R     probability;
    bool  discrete;
    int   event;
    R     eta;
};

template <class R> V3<R> unit_direction(V3<R> v) {
    if (!finite(v) || std::abs(dot(v, v) - R(1)) > R(1e-4))
        throw std::invalid_argument("BSDF direction must be finite and have unit length");
    return normalize(v);
}

template <class R> R ggx_d(V3<R> m, R alpha) {
    const R a2    = alpha * alpha;
    const R denom = (m.x * m.x + m.y * m.y) + a2 * m.z * m.z;
    return m.z > 0 ? a2 / (pi_v<R>() * denom * denom) : R(0);
}

template <class R> R g1_over_cos(V3<R> v, R alpha) {
    if (v.z <= 0) return 0;
    return R(2) / (v.z + std::sqrt(v.z * v.z + alpha * alpha * (v.x * v.x + v.y * v.y)));
}

template <class R> R fresnel_from_transmission(R cosine, R sin2, R eta_i, R eta_t) {
    if (eta_i == eta_t) return 0;
    if (sin2 >= 1) return 1;
    const R tc  = std::sqrt(R(1) - sin2);
    const R par = (eta_t * cosine - eta_i * tc) / (eta_t * cosine + eta_i * tc);
    const R per = (eta_i * cosine - eta_t * tc) / (eta_i * cosine + eta_t * tc);
    return R(0.5) * (par * par + per * per);
}

template <class R> V3<R> schlick(V3<R> f0, R cosine) {
    const R t  = R(1) - std::clamp(cosine, R(0), R(1));
    const R t5 = (((t * t) * t) * t) * t;
    return f0 + (V3<R>{1, 1, 1} - f0) * t5;
}

template <class R> V3<R> sample_vndf(V3<R> wo, V2<R> u, R alpha) {
    const V3<R> view  = normalize(V3<R>{alpha * wo.x, alpha * wo.y, wo.z});
    const R     lensq = view.x * view.x + view.y * view.y;
    const V3<R> t1    = lensq > 0 ? V3<R>{-view.y, view.x, 0} / std::sqrt(lensq) : V3<R>{1, 0, 0};
    const V3<R> t2    = cross(view, t1);
    const R     radius = std::sqrt(u.x);
    R           s, c;
    sincos_2pi(u.y, s, c);
    const R x     = radius * c;
    const R blend = R(0.5) * (R(1) + view.z);
    const R y     = (R(1) - blend) * std::sqrt(std::max(R(0), R(1) - x * x)) + blend * radius * s;
    const V3<R> n = (x * t1 + y * t2) + std::sqrt(std::max(R(0), (R(1) - x * x) - y * y)) * view;
    return normalize(V3<R>{alpha * n.x, alpha * n.y, std::max(R(0), n.z)});
}

// ----------------------------------------------------------------------------------------------
// Kulla-Conty energy compensation (microfacet_energy.cpp, ibl.cpp:28-47, :240-297). The table is
// the reference's bake_ibl_brdf() at its defaults: 64 x 64 nodes (x = cos(theta_o) in [0,1], y =
// roughness in [0,1]), 512 Hammersley samples per node with a per-node shift drawn from
// Sampler::for_sample(101 + node * 0x9e3779b97f4a7c15, 0, 0), FP64 sums of the white conductor's
// VNDF sample weights split by Fresnel (A, B), stored as float; the third channel holds the
// cosine-weighted average 2 * int E(mu) mu dmu of the row (Simpson over the piecewise-linear E).
constexpr int kEnergyRes     = 64;
constexpr int kEnergySamples = 512;

struct EnergyTable {
    std::vector<std::array<float, 3>> node;  // [y * kEnergyRes + x]
};

inline double radical_inverse(uint32_t bits) {
    bits = (bits << 16U) | (bits >> 16U);
    bits = ((bits & 0x55555555U) << 1U) | ((bits & 0xAAAAAAAAU) >> 1U);
    bits = ((bits & 0x33333333U) << 2U) | ((bits & 0xCCCCCCCCU) >> 2U);
    bits = ((bits & 0x0F0F0F0FU) << 4U) | ((bits & 0xF0F0F0F0U) >> 4U);
    bits = ((bits & 0x00FF00FFU) << 8U) | ((bits & 0xFF00FF00U) >> 8U);
    return static_cast<double>(bits) * 0x1p-32;
}

template <class R> struct Bsdf;
V2<double> integrate_ibl_brdf(double mu, double roughness, uint32_t samples, uint64_t seed);

inline void store_average_energy(std::vector<std::array<float, 3>>& t, int width, int row) {
    double       integral = 0;
    const double step     = 1. / static_cast<double>(width - 1);
    for (int x = 1; x < width; ++x) {
        const auto&  a = t[static_cast<size_t>(row) * width + x - 1];
        const auto&  b = t[static_cast<size_t>(row) * width + x];
        const double left  = std::clamp(static_cast<double>(a[0] + a[1]), 0., 1.);
        const double right = std::clamp(static_cast<double>(b[0] + b[1]), 0., 1.);
        const double mu    = static_cast<double>(x) * step;
        integral += step / 3 * (((mu - step) * left + 2 * (mu - step * .5) * (left + right)) + mu * right);
    }
    for (int x = 0; x < width; ++x) t[static_cast<size_t>(row) * width + x][2] = static_cast<float>(std::clamp(integral, 0., 1.));
}

inline const EnergyTable& energy_table();

template <class R> V2<R> sample_energy(const EnergyTable& table, R mu, R roughness) {
    const R    x  = std::clamp(mu, R(0), R(1)) * static_cast<R>(kEnergyRes - 1);
    const R    y  = std::clamp(roughness, R(0), R(1)) * static_cast<R>(kEnergyRes - 1);
    const auto ix = static_cast<size_t>(x), iy = static_cast<size_t>(y);
    const auto jx = std::min<size_t>(ix + 1, kEnergyRes - 1), jy = std::min<size_t>(iy + 1, kEnergyRes - 1);
    auto at = [&](size_t cx, size_t cy) {
        const auto& n = table.node[cy * kEnergyRes + cx];
        return V3<R>{R(n[0]), R(n[1]), R(n[2])};
    };
    auto mix = [](V3<R> a, V3<R> b, R t) { return a * (R(1) - t) + b * t; };
    const R     tx = x - static_cast<R>(ix);
    const V3<R> lo = mix(at(ix, iy), at(jx, iy), tx);
    const V3<R> hi = mix(at(ix, jy), at(jx, jy), tx);
    const V3<R> e  = mix(lo, hi, y - static_cast<R>(iy));
    return {std::clamp(e.x + e.y, R(0), R(1)), std::clamp(e.z, R(0), R(1))};
}

template <class R> inline V3<R> operator/(V3<R> a, V3<R> b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }

template <class R> V3<R> multiple_scattering(V3<R> f0, R ei, R eo, R average) {
    if (average >= R(1) - R(1e-6)) return {0, 0, 0};
    const V3<R> one{1, 1, 1};
    const V3<R> f     = f0 + (one - f0) / R(21);
    const V3<R> den   = one - f * (R(1) - average);
    const V3<R> color = ((f * f) * average) / V3<R>{std::max(den.x, R(1e-8)), std::max(den.y, R(1e-8)), std::max(den.z, R(1e-8))};
    return color * (((R(1) - ei) * (R(1) - eo)) / (pi_v<R>() * (R(1) - average)));
}

template <class R> struct Bsdf {
    uint32_t           kind;
    V3<R>              color;
    R                  alpha, eta_i, eta_t, roughness;
    const EnergyTable* energy = nullptr;
    // texture_color: the base-color texture sample (1 without a texture); color = base_color * clamp(texel, 0, 1)
    // (bsdf.cpp:122-125).
    Bsdf(const Material& m, bool front, V3<float> texture_color = {1, 1, 1}) : kind(m.kind) {
        const V3<R> tc = widen<R>(texture_color);
        color = widen<R>(m.base_color) * V3<R>{std::clamp(tc.x, R(0), R(1)), std::clamp(tc.y, R(0), R(1)), std::clamp(tc.z, R(0), R(1))};
        if constexpr (std::is_same_v<R, double>)
            alpha = static_cast<double>(m.roughness) * m.roughness;
        else
            alpha = m.roughness * m.roughness;
        roughness = R(m.roughness);
        eta_i = front ? R(1) : R(m.ior);
        eta_t = front ? R(m.ior) : R(1);
        if (kind > 3) throw std::invalid_argument("Unsupported material kind in the oracle");
        if (kind == 1 && m.compensate_energy && !is_delta()) energy = &energy_table();
    }
    bool is_delta() const { return kind == 2 || kind == 3 || (kind == 1 && alpha <= R(1e-6)); }
    R    single_probability(R cosine) const {
        return energy ? std::clamp(sample_energy(*energy, cosine, roughness).x, R(0.05), R(0.95)) : R(1);
    }

    V3<R> eval_n(V3<R> wo, V3<R> wi) const {
        if (wo.z <= 0 || wi.z <= 0 || is_delta()) return {0, 0, 0};
        if (kind == 0) return color / pi_v<R>();
        const V3<R> m = normalize(wo + wi);
        const V3<R> single = (((schlick(color, dot(wo, m)) * ggx_d(m, alpha)) * g1_over_cos(wo, alpha)) * g1_over_cos(wi, alpha)) / R(4);
        if (!energy) return single;
        const V2<R> outgoing = sample_energy(*energy, wo.z, roughness);
        const V2<R> incoming = sample_energy(*energy, wi.z, roughness);
        return single + multiple_scattering(color, incoming.x, outgoing.x, outgoing.y);
    }
    R pdf_n(V3<R> wo, V3<R> wi) const {
        if (wo.z <= 0 || wi.z <= 0 || is_delta()) return 0;
        if (kind == 0) return cosine_hemisphere_pdf(wi);
        const V3<R> m      = normalize(wo + wi);
        const R     glossy = ggx_d(m, alpha) * g1_over_cos(wo, alpha) / R(4);
        if (!energy) return glossy;
        const R weight = single_probability(wo.z);
        return weight * glossy + (R(1) - weight) * cosine_hemisphere_pdf(wi);
    }
    V3<R> eval(V3<R> wo, V3<R> wi) const { return eval_n(unit_direction(wo), unit_direction(wi)); }
    R     pdf(V3<R> wo, V3<R> wi) const { return pdf_n(unit_direction(wo), unit_direction(wi)); }

    std::optional<BsdfSample<R>> sample(V3<R> wo, R component, V2<R> u) const {
        wo = unit_direction(wo);
        check_unit(component);
        check_unit(u.x);
        check_unit(u.y);
        if (wo.z <= 0) return std::nullopt;
        const V3<R> reflected{-wo.x, -wo.y, wo.z};
        if (kind == 3) {
            const R ratio       = eta_i / eta_t;
            const R sin2        = ratio * ratio * (wo.x * wo.x + wo.y * wo.y);
            const R reflectance = fresnel_from_transmission(std::clamp(wo.z, R(0), R(1)), sin2, eta_i, eta_t);
            if (component < reflectance)
                return BsdfSample<R>{reflected, {1, 1, 1}, reflectance, true, specular_reflection, R(1)};
            const V3<R> transmitted{-ratio * wo.x, -ratio * wo.y, -std::sqrt(std::max(R(0), R(1) - sin2))};
            const R scale = ratio * ratio;
            return BsdfSample<R>{transmitted, color * scale, R(1) - reflectance, true, specular_transmission, eta_t / eta_i};
        }
        if (is_delta()) {
            const V3<R> w = kind == 2 ? color : schlick(color, wo.z);
            return BsdfSample<R>{reflected, w, R(1), true, specular_reflection, R(1)};
        }
        V3<R> wi;
        if (kind == 0 || (energy && component >= single_probability(wo.z))) {
            wi = sample_cosine_hemisphere(u);
        } else {
            const V3<R> m = sample_vndf(wo, u, alpha);
            wi            = normalize((R(2) * dot(wo, m)) * m - wo);
        }
        if (wi.z <= 0) return std::nullopt;
        const V3<R> ewi     = unit_direction(wi);
        const R     density = pdf_n(wo, ewi);
        const V3<R> weight  = eval_n(wo, ewi) * wi.z / density;
        return BsdfSample<R>{wi, weight, density, false, kind == 0 ? diffuse_reflection : glossy_reflection, R(1)};
    }
};

// integrate_ibl_brdf (ibl.cpp:240-277): Fresnel-split albedo of the white conductor at one node.
V2<double> integrate_ibl_brdf(double mu, double roughness, uint32_t samples, uint64_t seed) {
    if (roughness * roughness <= 1e-6) {
        const double fresnel = std::pow(1 - mu, 5);
        return {1 - fresnel, fresnel};
    }
    const double   cosine = std::max(mu, 1e-6);
    const V3<double> wo{std::sqrt(std::max(0.0, 1 - cosine * cosine)), 0, cosine};
    Material       white{1, {1, 1, 1}, {0, 0, 0}, static_cast<float>(roughness), 1.5f, true, false};
    const Bsdf<double> bsdf(white, true);
    Sampler        shift_rng = Sampler::for_sample(seed, 0, 0);
    const double   sx        = shift_rng.next_1d<double>();
    const double   sy        = shift_rng.next_1d<double>();
    V2<double>     sum{0, 0};
    for (uint32_t i = 0; i < samples; ++i) {
        const double px = (static_cast<double>(i) + 0.5) / samples + sx;
        const double py = radical_inverse(i) + sy;
        const V2<double> u{px - std::floor(px), py - std::floor(py)};
        const auto smp = bsdf.sample(wo, 0.5, u);
        if (!smp) continue;
        const V3<double> halfway = normalize(wo + smp->direction);
        const double     fc      = std::pow(1 - std::clamp(dot(wo, halfway), 0.0, 1.0), 5);
        sum.x += smp->weight.x * (1 - fc);
        sum.y += smp->weight.x * fc;
    }
    return {sum.x / static_cast<double>(samples), sum.y / static_cast<double>(samples)};
}

inline const EnergyTable& energy_table() {
    static const EnergyTable table = [] {
        EnergyTable t;
        t.node.resize(static_cast<size_t>(kEnergyRes) * kEnergyRes);
        for (int y = 0; y < kEnergyRes; ++y) {
            for (int x = 0; x < kEnergyRes; ++x) {
                const auto v = integrate_ibl_brdf(static_cast<double>(x) / (kEnergyRes - 1), static_cast<double>(y) / (kEnergyRes - 1),
                                                  kEnergySamples, 101 + static_cast<uint64_t>(y * kEnergyRes + x) * 0x9e3779b97f4a7c15ULL);
                t.node[static_cast<size_t>(y) * kEnergyRes + x] = {static_cast<float>(v.x), static_cast<float>(v.y), 0.0f};
            }
            store_average_energy(t.node, kEnergyRes, y);
        }
        return t;
    }();
    return table;
}

template <class R> R fresnel_dielectric(R cosine, R eta_i, R eta_t) {
    if (!std::isfinite(cosine) || cosine < 0 || cosine > 1 || !(eta_i > 0) || !(eta_t > 0))
        throw std::invalid_argument("Fresnel requires cosine in [0,1] and positive finite IORs");
    const R ratio = eta_i / eta_t;
    const R sin2  = ratio * ratio * std::max(R(0), R(1) - cosine * cosine);
    return fresnel_from_transmission(cosine, sin2, eta_i, eta_t);
}

// ----------------------------------------------------------------------------------------------
// Light sampling (light.cpp:49-59, :266-340).
template <class R> R area_density(V3<R> ref, V3<R> point, V3<R> normal, R area) {
    const V3<R> diff = point - ref;
    const R     d2   = dot(diff, diff);
    if (d2 == 0) return 0;
    const R cosine = std::abs(dot(normal, (-diff) / std::sqrt(d2)));
    return cosine > 0 ? d2 / (area * cosine) : R(0);
}

template <class R> struct LightSample {
    V3<R> direction, radiance;
    R     distance, probability;
    bool  discrete;
    int   kind;
    size_t emitter;
    V3<R> position, normal, error;
    uint32_t prim = kNoSkip;  // scene-order primitive of a sampled area emitter
};

template <class R> std::optional<LightSample<R>> sample_light(const Lights<R>& L, V3<R> ref, R light_u, V2<R> su) {
    if (!finite(ref)) throw std::invalid_argument("Light reference point must be finite");
    check_unit(su.x);
    check_unit(su.y);
    const auto sel = L.selection.sample(light_u);
    if (!sel) return std::nullopt;
    const auto& ch  = L.choices[sel->index];
    const R     inf = std::numeric_limits<R>::infinity();
    LightSample<R> out{};
    if (ch.kind == 1) {
        const auto& d = L.dirs[ch.index];
        out           = {d.direction, d.radiance, inf, sel->mass, true, 1, sel->index, {}, {}, {}};
        return out;
    }
    if (ch.kind == 2) {
        const EnvSample<R> e = env_sample(L.envmap, V2<R>{su.x, su.y});
        out                  = {e.direction, e.radiance, inf, sel->mass * e.pdf, false, 2, sel->index, {}, {}, {}};
        return out;
    }
    const auto& a  = L.areas[ch.index];
    const V3<R> b  = sample_triangle(su);
    const V3<R> pt = (b.x * a.v[0] + b.y * a.v[1]) + b.z * a.v[2];
    if (!a.double_sided && dot(a.normal, ref - pt) <= 0) return std::nullopt;
    const R density = area_density(ref, pt, a.normal, a.area);
    if (density == 0) return std::nullopt;
    const V3<R> dir      = pt - ref;
    const R     distance = length(dir);
    const V3<R> err = ((vabs(b.x * a.v[0]) + vabs(b.y * a.v[1])) + vabs(b.z * a.v[2])) * (R(7) * std::numeric_limits<R>::epsilon());
    out = {dir / distance, a.radiance, distance, sel->mass * density, false, 0, sel->index, pt, a.normal, err, a.prim};
    return out;
}

template <class R> R area_pdf(const Lights<R>& L, V3<R> ref, const SurfaceHit<R>& h) {
    const int64_t c = L.area_choice_of_prim[h.prim];
    if (c < 0) return 0;
    const auto& a = L.areas[L.choices[static_cast<size_t>(c)].index];
    if (!a.double_sided && dot(a.normal, ref - h.position) <= 0) return 0;
    return L.selection.probability(static_cast<size_t>(c)) * area_density(ref, h.position, a.normal, a.area);
}

template <class R> R environment_pdf(const Lights<R>& L, V3<R> direction) {
    return L.env_choice >= 0 ? L.selection.probability(static_cast<size_t>(L.env_choice)) * env_pdf(L.envmap, direction) : R(0);
}

// ----------------------------------------------------------------------------------------------
// Integrator (path_tracer.cpp:126-405).
struct Settings {
    uint32_t mode = 2, max_depth = 8, rr_depth = 3;
    bool     rr = true;
};

struct PathCounters {
    uint64_t paths = 0, shadow = 0, scatter = 0, rr_term = 0, extension = 0;
    Counters geom;
    Counters shadow_geom;
};

template <class R> void check_radiance(V3<R> v) {
    if (!finite(v) || v.x < 0 || v.y < 0 || v.z < 0)
        throw std::runtime_error("Path contribution is nonfinite or negative");
}

// FP32 twin deviation: a shadow segment toward an area emitter ignores that emitter's own triangle.
// The reference ends the segment short of the light by the light point's position error
// (7 eps |p|, light.cpp:310-313) and one ulp (path_tracer.cpp:126-151); that margin scales with the
// light's coordinates, not with the segment's length, so for a light near a coordinate-zero plane
// seen from farther away the rounding of the segment's own t lands past the limit: on the reference's
// emitter case (path_tests.cpp:103-120) 24 of 1000 segments in FP64 and 37 of 1000 in FP32 report the
// light itself as an occluder. The FP64 oracle keeps the reference's rule; the FP32 twin (and the
// kernels, which reproduce it) skip the sampled triangle, which removes the self-occlusion exactly
// and leaves every other occluder in place.
template <class R> uint32_t shadow_skip(const LightSample<R>& light) {
    return std::is_same_v<R, float> ? light.prim : kNoSkip;
}

template <class R>
V3<R> shadow_transmittance(const Scene& s, Ray<R> ray, const LightSample<R>& light, PathCounters* st) {
    if (st) ++st->shadow;
    R limit = std::numeric_limits<R>::infinity();
    if (std::isfinite(light.distance)) {
        const V3<R> target = offset_origin(light.position, to_float(light.normal), light.error, -light.direction);
        const V3<R> diff   = target - ray.o;
        limit              = length(diff);
        if (limit == 0) return {0, 0, 0};
        ray.d = diff / limit;
        limit = std::nextafter(limit, R(0));
    }
    if (s.trace<R>(ray, R(0), limit, true, st ? &st->shadow_geom : nullptr, nullptr, shadow_skip(light))) return {0, 0, 0};
    return {1, 1, 1};
}

template <class R>
V3<R> trace(const Scene& s, const Settings& cfg, Ray<R> ray, Sampler& sampler, R tmin0, R tmax0, PathCounters* st) {
    const Lights<R>& L = s.lights<R>();
    if (st) ++st->paths;
    V3<R> result{0, 0, 0}, beta{1, 1, 1}, prev_ref{0, 0, 0};
    R     prev_pdf = 0, eta_scale = 1;
    bool  prev_delta = true;
    const R inf      = std::numeric_limits<R>::infinity();
    for (uint32_t depth = 0;; ++depth) {
        const R tmin = depth == 0 ? tmin0 : R(0);
        const R tmax = depth == 0 ? tmax0 : inf;
        if (st) ++st->extension;
        const auto ph = s.trace<R>(ray, tmin, tmax, false, st ? &st->geom : nullptr);
        check_radiance(beta);
        if (vmax(beta) == 0) break;
        auto emission_weight = [&](R light_pdf) -> R {
            if (depth == 0 || prev_delta || cfg.mode == 0) return R(1);
            return cfg.mode == 2 ? power_heuristic(prev_pdf, light_pdf) : R(0);
        };
        if (!ph) {
            const R w = emission_weight(environment_pdf(L, ray.d));
            result    = result + (beta * env_evaluate(L.envmap, ray.d)) * w;
            break;
        }
        const SurfaceHit<R> hit = surface_hit(s, ray, *ph);
        const Material&     mat = s.materials[s.inst[hit.prim]];
        {
            const R w = emission_weight(area_pdf(L, prev_ref, hit));
            result    = result + (beta * widen<R>(mat.emission)) * w;
        }
        if (depth == cfg.max_depth) break;
        V3<R> normal = widen<R>(mat.kind == 3 ? hit.geometric_normal : hit.shading_normal);
        const V3<R> wo_world = -ray.d;
        if (dot(normal, wo_world) < 0) normal = -normal;
        const Frame<R> frame(normal);
        const V3<R>    wo = frame.to_local(wo_world);
        const V3<float> texture_color =
            mat.base_color_texture >= 0 ? s.textures.at(size_t(mat.base_color_texture)).sample(hit.uv, 0.0f, mat.texture_filter)
                                        : V3<float>{1, 1, 1};
        const Bsdf<R>  bsdf(mat, hit.front_face, texture_color);
        const R        light_u   = sampler.next_1d<R>();
        const R        su_x      = sampler.next_1d<R>();
        const R        su_y      = sampler.next_1d<R>();
        const R        component = sampler.next_1d<R>();
        const R        du_x      = sampler.next_1d<R>();
        const R        du_y      = sampler.next_1d<R>();
        const V3<R>    ng        = widen<R>(hit.geometric_normal);
        if (!bsdf.is_delta()) {
            const auto light = sample_light(L, hit.position, light_u, V2<R>{su_x, su_y});
            if (light && (cfg.mode != 0 || light->discrete)) {
                const V3<R> wi   = frame.to_local(light->direction);
                const bool  side = dot(wo_world, ng) * dot(light->direction, ng) > 0;
                if (wi.z > 0 && side) {
                    const V3<R> tr = shadow_transmittance(s, spawn_ray(hit, light->direction), *light, st);
                    if (vmax(tr) > 0) {
                        R weight = 1;
                        if (cfg.mode == 2 && !light->discrete) weight = power_heuristic(light->probability, bsdf.pdf(wo, wi));
                        result = result + ((((beta * tr) * bsdf.eval(wo, wi)) * wi.z) * light->radiance) * (weight / light->probability);
                    }
                }
            }
        }
        const auto smp = bsdf.sample(wo, component, V2<R>{du_x, du_y});
        if (!smp || vmax(smp->weight) == 0) break;
        const V3<R> wi_world     = normalize(frame.to_world(smp->direction));
        const R     sides        = dot(wo_world, ng) * dot(wi_world, ng);
        const bool  transmission = smp->event == specular_transmission;
        if ((transmission && sides >= 0) || (!transmission && sides <= 0)) break;
        beta = beta * smp->weight;
        check_radiance(beta);
        prev_ref   = hit.position;
        prev_pdf   = smp->probability;
        prev_delta = smp->discrete;
        eta_scale  = eta_scale * (smp->eta * smp->eta);
        if (st) ++st->scatter;
        ray                = spawn_ray(hit, wi_world);
        const R roulette_u = sampler.next_1d<R>();
        if (cfg.rr && depth + 1 >= cfg.rr_depth) {
            const R survival = std::clamp(vmax(beta * eta_scale), R(0.05), R(0.95));
            if (roulette_u >= survival) {
                if (st) ++st->rr_term;
                break;
            }
            beta = beta / survival;
        }
    }
    check_radiance(result);
    return result;
}

// Camera (scene.cpp:171-199). The FP32 twin computes tan(fov/2) in double and rounds it.
template <class R> struct CameraRays {
    V3<R> position, forward, right, up;
    R     half_height, aperture, focus, near_plane, far_plane;
    // Camera::validate (scene.cpp:141-155), in float as the reference's glm::vec3 members.
    static void validate(const OrcCamera& c) {
        const V3<float> p{c.position[0], c.position[1], c.position[2]}, t{c.target[0], c.target[1], c.target[2]},
            u{c.up[0], c.up[1], c.up[2]};
        if (!(finite(p) && finite(t) && finite(u))) throw std::invalid_argument("Camera vectors must be finite");
        if (!(length(t - p) > 1e-6f)) throw std::invalid_argument("Camera position and target must differ");
        if (!(length(cross(t - p, u)) > 1e-6f)) throw std::invalid_argument("Camera up must not be parallel to view direction");
        if (!(std::isfinite(c.vertical_fov_degrees) && c.vertical_fov_degrees > 0 && c.vertical_fov_degrees < 179))
            throw std::invalid_argument("Camera FOV must be between 0 and 179 degrees");
        if (!(std::isfinite(c.near_plane) && std::isfinite(c.far_plane) && c.near_plane > 0 && c.far_plane > c.near_plane))
            throw std::invalid_argument("Camera clipping planes must satisfy 0 < near < far");
        if (!(std::isfinite(c.aperture_radius) && c.aperture_radius >= 0 && std::isfinite(c.focus_distance) && c.focus_distance > 0))
            throw std::invalid_argument("Camera lens requires nonnegative aperture and positive focus distance");
    }
    explicit CameraRays(const OrcCamera& c) {
        validate(c);
        position     = from_float<R>(c.position);
        forward      = normalize(from_float<R>(c.target) - from_float<R>(c.position));
        right        = normalize(cross(forward, from_float<R>(c.up)));
        up           = cross(right, forward);
        half_height  = static_cast<R>(std::tan(static_cast<double>(c.vertical_fov_degrees) * (kPi / 180.0) * 0.5));
        aperture     = R(c.aperture_radius);
        focus        = R(c.focus_distance);
        near_plane   = R(c.near_plane);
        far_plane    = R(c.far_plane);
    }
    Ray<R> pinhole(V2<R> uv, R aspect) const {
        const V3<R> d = (forward + (((R(2) * uv.x - R(1)) * aspect) * half_height) * right) +
                        ((R(1) - R(2) * uv.y) * half_height) * up;
        return {position, normalize(d)};
    }
    Ray<R> lens(V2<R> uv, R aspect, V2<R> lu) const {
        const Ray<R> p = pinhole(uv, aspect);
        if (aperture == 0) return p;
        const V2<R> dk0 = sample_disk(lu);
        const V2<R> dk{dk0.x * aperture, dk0.y * aperture};
        const V3<R> focus_point = p.o + p.d * (focus / dot(p.d, forward));
        const V3<R> origin      = (position + dk.x * right) + dk.y * up;
        return {origin, normalize(focus_point - origin)};
    }
};

template <class R>
V3<R> sample_pixel(const Scene& s, const Settings& cfg, const CameraRays<R>& cam, uint32_t w, uint32_t h,
                   uint64_t seed, uint32_t x, uint32_t y, uint64_t sample, PathCounters* st) {
    Sampler     smp = Sampler::for_sample(seed, static_cast<uint64_t>(y) * w + x, sample);
    const R     fx  = smp.next_1d<R>();
    const R     fy  = smp.next_1d<R>();
    const R     lx  = smp.next_1d<R>();
    const R     ly  = smp.next_1d<R>();
    const V2<R> uv{(R(x) + fx) / R(w), (R(y) + fy) / R(h)};
    const Ray<R> ray = cam.lens(uv, R(w) / R(h), V2<R>{lx, ly});
    const R cosine   = dot(ray.d, cam.forward);
    return trace(s, cfg, ray, smp, cam.near_plane / cosine, cam.far_plane / cosine, st);
}

}  // namespace orc

// ==============================================================================================
// C ABI for the Python test harness.
using namespace orc;

static thread_local std::string g_error;

extern "C" const char* orc_last_error(void) { return g_error.c_str(); }

template <class F> static int guarded(F&& f) {
    try {
        f();
        return 0;
    } catch (const std::exception& e) {
        g_error = e.what();
        return 1;
    }
}

static Material to_material(const OrcMaterial& m) {
    return Material{m.kind, {m.base_color[0], m.base_color[1], m.base_color[2]}, {m.emission[0], m.emission[1], m.emission[2]},
                    m.roughness, m.ior, m.double_sided != 0, m.mirrored != 0, m.compensate_energy != 0,
                    static_cast<int32_t>(m.base_color_texture) - 1, m.texture_filter, static_cast<int32_t>(m.normal_texture) - 1,
                    m.normal_scale, m.normal_flip_y != 0};
}

extern "C" int orc_texture_sample(uint32_t w, uint32_t h, const float* rgb, const float* uv, uint64_t n, float lod, uint32_t filter,
                                  float* out) {
    return guarded([&] {
        const Texture t(w, h, rgb);
        for (uint64_t i = 0; i < n; ++i) {
            const V3<float> v = t.sample(V2<float>{uv[i * 2], uv[i * 2 + 1]}, lod, filter);
            out[i * 3] = v.x; out[i * 3 + 1] = v.y; out[i * 3 + 2] = v.z;
        }
    });
}

extern "C" int orc_texture_level(uint32_t w, uint32_t h, const float* rgb, uint32_t level, uint32_t* out_w, uint32_t* out_h,
                                 float* out) {
    return guarded([&] {
        const Texture   t(w, h, rgb);
        const TexLevel& l = t.levels.at(level);
        *out_w = l.w;
        *out_h = l.h;
        if (out)
            for (size_t i = 0; i < l.px.size(); ++i) { out[i * 3] = l.px[i].x; out[i * 3 + 1] = l.px[i].y; out[i * 3 + 2] = l.px[i].z; }
    });
}

template <class R> static int surface_query(const Scene& s, const double* ray, double* out) {
    const Ray<R> r{{R(ray[0]), R(ray[1]), R(ray[2])}, {R(ray[3]), R(ray[4]), R(ray[5])}};
    const auto   ph = s.trace<R>(r, R(0), std::numeric_limits<R>::infinity(), false, nullptr);
    if (!ph) return 0;
    const SurfaceHit<R> h = surface_hit(s, r, *ph);
    const double v[14] = {double(h.position.x), double(h.position.y), double(h.position.z), h.geometric_normal.x,
                          h.geometric_normal.y, h.geometric_normal.z, h.shading_normal.x, h.shading_normal.y,
                          h.shading_normal.z, h.uv.x, h.uv.y, double(h.t), h.front_face ? 1.0 : 0.0, double(h.prim)};
    std::copy(v, v + 14, out);
    return 1;
}

// The shadow segment PathTracer::trace hands to shadow_transmittance (path_tracer.cpp:126-151) for a light
// sampled from `origin` (LightSampler::sample, light.cpp:266-322): origin, unit direction toward the offset
// light endpoint, and the nextafter-shortened limit (+inf for infinite lights). out[0..5] ray, out[6] limit,
// out[7] the sampled light distance, out[8] 1 when the segment is occluded in this precision.
template <class R> static int segment_query(const Scene& s, const double* origin, double lu, const double* u, double* out) {
    const V3<R> o{R(origin[0]), R(origin[1]), R(origin[2])};
    const auto  l = sample_light(s.lights<R>(), o, R(lu), V2<R>{R(u[0]), R(u[1])});
    if (!l) return 0;
    Ray<R> ray{o, l->direction};
    R      limit = std::numeric_limits<R>::infinity();
    if (std::isfinite(l->distance)) {
        const V3<R> target = offset_origin(l->position, to_float(l->normal), l->error, -l->direction);
        const V3<R> diff   = target - ray.o;
        limit              = length(diff);
        if (limit == 0) return 0;
        ray.d = diff / limit;
        limit = std::nextafter(limit, R(0));
    }
    const bool blocked = s.trace<R>(ray, R(0), limit, true, nullptr, nullptr, shadow_skip(*l)).has_value();
    const double v[9]  = {double(ray.o.x), double(ray.o.y), double(ray.o.z), double(ray.d.x), double(ray.d.y), double(ray.d.z),
                          double(limit), double(l->distance), blocked ? 1.0 : 0.0};
    std::copy(v, v + 9, out);
    return 1;
}

extern "C" int orc_shadow_segment(void* sp, int precision, const double* origin, double lu, const double* u, double* out) {
    int found = 0;
    const int rc = guarded([&] {
        const Scene& s = *static_cast<Scene*>(sp);
        found = precision == 32 ? segment_query<float>(s, origin, lu, u, out) : segment_query<double>(s, origin, lu, u, out);
    });
    return rc ? -1 : found;
}

extern "C" int orc_surface_hit(void* sp, int precision, const double* ray, double* out) {
    int hit = 0;
    const int rc = guarded([&] {
        hit = precision == 32 ? surface_query<float>(*static_cast<Scene*>(sp), ray, out)
                              : surface_query<double>(*static_cast<Scene*>(sp), ray, out);
    });
    return rc ? -1 : hit;
}

extern "C" void orc_atan2_turns(const float* y, const float* x, uint64_t n, float* out) {
    for (uint64_t i = 0; i < n; ++i) out[i] = atan2_turns(y[i], x[i]);
}

template <class R> static void env_query(const Scene& s, const double* dir, uint64_t n, double* radiance, double* pdf) {
    const Lights<R>& L = s.lights<R>();
    for (uint64_t i = 0; i < n; ++i) {
        const V3<R> d{R(dir[i * 3]), R(dir[i * 3 + 1]), R(dir[i * 3 + 2])};
        const V3<R> v = env_evaluate(L.envmap, d);
        radiance[i * 3] = v.x; radiance[i * 3 + 1] = v.y; radiance[i * 3 + 2] = v.z;
        pdf[i] = env_pdf(L.envmap, d);
    }
}

template <class R> static void env_draw(const Scene& s, const double* u, uint64_t n, double* out) {
    const Lights<R>& L = s.lights<R>();
    for (uint64_t i = 0; i < n; ++i) {
        const EnvSample<R> e = env_sample(L.envmap, V2<R>{R(u[i * 2]), R(u[i * 2 + 1])});
        double* o = out + i * 7;
        o[0] = e.direction.x; o[1] = e.direction.y; o[2] = e.direction.z;
        o[3] = e.radiance.x; o[4] = e.radiance.y; o[5] = e.radiance.z; o[6] = e.pdf;
    }
}

extern "C" int orc_env_eval(void* sp, int precision, const double* dir, uint64_t n, double* radiance, double* pdf) {
    return guarded([&] {
        if (precision == 32) env_query<float>(*static_cast<Scene*>(sp), dir, n, radiance, pdf);
        else env_query<double>(*static_cast<Scene*>(sp), dir, n, radiance, pdf);
    });
}

extern "C" int orc_env_sample(void* sp, int precision, const double* u, uint64_t n, double* out) {
    return guarded([&] {
        if (precision == 32) env_draw<float>(*static_cast<Scene*>(sp), u, n, out);
        else env_draw<double>(*static_cast<Scene*>(sp), u, n, out);
    });
}

extern "C" double orc_env_integral(void* sp, int precision) {
    const Scene* s = static_cast<Scene*>(sp);
    return precision == 32 ? double(s->lights32.envmap.integral) : s->lights64.envmap.integral;
}

extern "C" void orc_energy_table(float* out) {
    const auto& t = energy_table();
    for (size_t i = 0; i < t.node.size(); ++i)
        for (int c = 0; c < 3; ++c) out[i * 3 + c] = t.node[i][c];
}

extern "C" void orc_integrate_ibl_brdf(double mu, double roughness, uint32_t samples, uint64_t seed, double* out) {
    const auto v = integrate_ibl_brdf(mu, roughness, samples, seed);
    out[0]       = v.x;
    out[1]       = v.y;
}

extern "C" double orc_average_energy_of_ramp(const float* e, uint32_t width) {
    std::vector<std::array<float, 3>> row(width);
    for (uint32_t x = 0; x < width; ++x) row[x] = {e[x], 0.0f, 0.0f};
    store_average_energy(row, static_cast<int>(width), 0);
    return row[0][2];
}

extern "C" void* orc_scene_new(const OrcSceneDesc* d) {
    Scene* s = nullptr;
    const int rc = guarded([&] {
        auto sc = std::make_unique<Scene>();
        sc->pos.resize(d->triangle_count);
        sc->nrm.resize(d->triangle_count);
        sc->inst.assign(d->instance, d->instance + d->triangle_count);
        sc->tri.assign(d->triangle, d->triangle + d->triangle_count);
        for (uint64_t p = 0; p < d->triangle_count; ++p) {
            for (int k = 0; k < 3; ++k) {
                sc->pos[p][k] = {d->positions[p * 9 + k * 3], d->positions[p * 9 + k * 3 + 1], d->positions[p * 9 + k * 3 + 2]};
                sc->nrm[p][k] = {d->normals[p * 9 + k * 3], d->normals[p * 9 + k * 3 + 1], d->normals[p * 9 + k * 3 + 2]};
            }
            if (d->instance[p] >= d->instance_count) throw std::invalid_argument("instance index out of range");
        }
        for (uint32_t i = 0; i < d->instance_count; ++i) {
            const auto& m = d->materials[i];
            sc->materials.push_back(to_material(m));
        }
        for (uint32_t i = 0; i < d->texture_count; ++i) sc->textures.emplace_back(d->textures[i].width, d->textures[i].height, d->textures[i].rgb);
        for (const auto& m : sc->materials) {
            if (m.base_color_texture >= int64_t(sc->textures.size()) || m.normal_texture >= int64_t(sc->textures.size()))
                throw std::invalid_argument("Material references missing texture");
            if (!(std::isfinite(m.normal_scale) && m.normal_scale >= 0)) throw std::invalid_argument("Normal scale must be finite and nonnegative");
        }
        if (d->uvs) {
            sc->uv.resize(d->triangle_count);
            for (uint64_t p = 0; p < d->triangle_count; ++p)
                for (int k = 0; k < 3; ++k) sc->uv[p][k] = {d->uvs[p * 6 + k * 2], d->uvs[p * 6 + k * 2 + 1]};
        }
        if (d->environment_texture > sc->textures.size()) throw std::invalid_argument("Environment references missing texture");
        if (!(std::isfinite(d->environment_strength) && d->environment_strength >= 0 && std::isfinite(d->environment_rotation_degrees)))
            throw std::invalid_argument("Environment strength and rotation must be finite; strength must be nonnegative");
        sc->environment_texture  = static_cast<int32_t>(d->environment_texture) - 1;
        sc->environment_strength = d->environment_texture ? d->environment_strength : 1.0f;
        sc->environment_rotation = d->environment_texture ? d->environment_rotation_degrees : 0.0f;
        sc->environment = {d->environment[0], d->environment[1], d->environment[2]};
        sc->directional.assign(d->directional, d->directional + d->directional_count);
        sc->build();
        s = sc.release();
    });
    return rc == 0 ? s : nullptr;
}

extern "C" void orc_scene_free(void* s) { delete static_cast<Scene*>(s); }

extern "C" void orc_scene_build_stats(void* sp, uint64_t out[4]) {
    const Scene* s = static_cast<Scene*>(sp);
    out[0]         = s->pos.size();
    out[1]         = s->bvh.nodes.size();
    out[2]         = s->bvh.leaves;
    out[3]         = s->bvh.max_depth;
}

static void add_stats(OrcStats* out, const PathCounters& c) {
    if (!out) return;
    out->paths += c.paths;
    out->shadow_rays += c.shadow;
    out->scattering_events += c.scatter;
    out->roulette_terminations += c.rr_term;
    out->bounds_tests += c.geom.bounds + c.shadow_geom.bounds;
    out->triangle_tests += c.geom.triangles + c.shadow_geom.triangles;
    out->extension_rays += c.extension;
    out->shadow_bounds_tests += c.shadow_geom.bounds;
    out->shadow_triangle_tests += c.shadow_geom.triangles;
}

template <class F> static void parallel_for(uint64_t n, int threads, F&& body) {
    threads = std::max(1, threads);
    std::atomic<uint64_t> next{0};
    std::vector<std::thread> pool;
    std::vector<std::string> errors(threads);
    for (int w = 0; w < threads; ++w) {
        pool.emplace_back([&, w] {
            try {
                for (;;) {
                    const uint64_t begin = next.fetch_add(1024);
                    if (begin >= n) break;
                    body(w, begin, std::min<uint64_t>(n, begin + 1024));
                }
            } catch (const std::exception& e) {
                errors[w] = e.what();
                next.store(n);
            }
        });
    }
    for (auto& t : pool) t.join();
    for (auto& e : errors)
        if (!e.empty()) throw std::runtime_error(e);
}

// The FP32 twin's reported distance of a winning hit: the FP64 ray/plane solve of that triangle
// (n = (v1 - v0) x (v2 - v0), t = n.(v0 - o) / n.d), rounded to float; the FP32 t if n.d == 0.
// The hit decision itself stays with the FP32 watertight test.
static double refine_distance(const std::array<V3<float>, 3>& v, const Ray<float>& ray, float t32) {
    const double e1x = double(v[1].x) - v[0].x, e1y = double(v[1].y) - v[0].y, e1z = double(v[1].z) - v[0].z;
    const double e2x = double(v[2].x) - v[0].x, e2y = double(v[2].y) - v[0].y, e2z = double(v[2].z) - v[0].z;
    const double nx = e1y * e2z - e2y * e1z, ny = e1z * e2x - e2z * e1x, nz = e1x * e2y - e2x * e1y;
    const double den = (nx * ray.d.x + ny * ray.d.y) + nz * ray.d.z;
    if (den == 0.0) return t32;
    const double num = (nx * (double(v[0].x) - ray.o.x) + ny * (double(v[0].y) - ray.o.y)) + nz * (double(v[0].z) - ray.o.z);
    return double(static_cast<float>(num / den));
}
static double refine_distance(const std::array<V3<float>, 3>&, const Ray<double>&, double t) { return t; }

template <class R>
static void intersect_impl(const Scene& s, const double* rays, uint64_t n, double tmin, double tmax, int threads,
                           int32_t* prim, double* t, double* t2, int32_t* p2, double* bary, OrcStats* stats) {
    std::vector<Counters> local(std::max(1, threads));
    parallel_for(n, threads, [&](int w, uint64_t b, uint64_t e) {
        for (uint64_t i = b; i < e; ++i) {
            const double* r = rays + i * 6;
            const Ray<R>  ray{{R(r[0]), R(r[1]), R(r[2])}, {R(r[3]), R(r[4]), R(r[5])}};
            std::optional<PrimHit<R>> second;
            const auto hit = s.trace<R>(ray, R(tmin), R(tmax), false, &local[w], (t2 || p2) ? &second : nullptr);
            prim[i] = hit ? static_cast<int32_t>(hit->prim) : -1;
            if (t) {
                double reported = hit ? double(hit->hit.t) : std::numeric_limits<double>::infinity();
                if (hit && std::is_same_v<R, float>) reported = refine_distance(s.pos[hit->prim], ray, hit->hit.t);
                t[i] = reported;
            }
            if (t2) t2[i] = second ? double(second->hit.t) : std::numeric_limits<double>::infinity();
            if (p2) p2[i] = second ? static_cast<int32_t>(second->prim) : -1;
            if (bary) {
                for (int k = 0; k < 3; ++k) bary[i * 3 + k] = hit ? double(hit->hit.bary[k]) : 0.0;
            }
        }
    });
    if (stats)
        for (auto& c : local) {
            stats->bounds_tests += c.bounds;
            stats->triangle_tests += c.triangles;
        }
}

extern "C" int orc_intersect(void* sp, int precision, const double* rays, uint64_t n, double tmin, double tmax,
                             int threads, int32_t* prim, double* t, double* t2, int32_t* p2, double* bary,
                             OrcStats* stats) {
    return guarded([&] {
        const Scene& s = *static_cast<Scene*>(sp);
        if (precision == 32)
            intersect_impl<float>(s, rays, n, tmin, tmax, threads, prim, t, t2, p2, bary, stats);
        else
            intersect_impl<double>(s, rays, n, tmin, tmax, threads, prim, t, t2, p2, bary, stats);
    });
}

extern "C" int orc_occluded(void* sp, int precision, const double* rays, uint64_t n, double tmin, double tmax,
                            int threads, uint8_t* out, OrcStats* stats) {
    return guarded([&] {
        const Scene& s = *static_cast<Scene*>(sp);
        std::vector<Counters> local(std::max(1, threads));
        parallel_for(n, threads, [&](int w, uint64_t b, uint64_t e) {
            for (uint64_t i = b; i < e; ++i) {
                const double* r = rays + i * 6;
                if (precision == 32) {
                    const Ray<float> ray{{float(r[0]), float(r[1]), float(r[2])}, {float(r[3]), float(r[4]), float(r[5])}};
                    out[i] = s.trace<float>(ray, float(tmin), float(tmax), true, &local[w]).has_value();
                } else {
                    const Ray<double> ray{{r[0], r[1], r[2]}, {r[3], r[4], r[5]}};
                    out[i] = s.trace<double>(ray, tmin, tmax, true, &local[w]).has_value();
                }
            }
        });
        if (stats)
            for (auto& c : local) {
                stats->bounds_tests += c.bounds;
                stats->triangle_tests += c.triangles;
            }
    });
}

static Settings to_settings(const OrcRender* r) {
    if (r->mode > 2) throw std::invalid_argument("Sampling mode must be bsdf, light or mis");
    if (r->maximum_depth > 128 || r->roulette_depth > 128) throw std::invalid_argument("Path and roulette depths must not exceed 128");
    Settings s;
    s.mode      = r->mode;
    s.max_depth = r->maximum_depth;
    s.rr        = r->russian_roulette != 0;
    s.rr_depth  = r->roulette_depth;
    return s;
}

template <class R>
static void render_impl(const Scene& s, const OrcCamera& camera, const OrcRender& rs, int threads, uint32_t tile,
                        uint32_t batch, float* image, double* sums_out, OrcStats* stats) {
    if (rs.width == 0 || rs.height == 0 || rs.width > 16384 || rs.height > 16384 || rs.spp == 0)
        throw std::invalid_argument("Render dimensions or sample count exceed the image budget");
    const Settings       cfg = to_settings(&rs);
    const CameraRays<R>  cam(camera);
    const uint32_t x0 = rs.x0, y0 = rs.y0;
    const uint32_t x1 = rs.x1 ? rs.x1 : rs.width, y1 = rs.y1 ? rs.y1 : rs.height;
    if (x0 >= x1 || y0 >= y1 || x1 > rs.width || y1 > rs.height) throw std::invalid_argument("bad crop");
    const uint32_t cw = x1 - x0, chh = y1 - y0;
    const uint32_t cols = (cw + tile - 1) / tile, rows = (chh + tile - 1) / tile;
    const uint64_t tiles = uint64_t(cols) * rows;
    threads              = std::max(1, threads);
    std::atomic<uint64_t>     next{0};
    std::vector<PathCounters> local(threads);
    std::vector<std::string>  errors(threads);
    std::vector<std::thread>  pool;
    for (int w = 0; w < threads; ++w) {
        pool.emplace_back([&, w] {
            try {
                for (;;) {
                    const uint64_t t = next.fetch_add(1);
                    if (t >= tiles) break;
                    const uint32_t tx = x0 + uint32_t(t % cols) * tile, ty = y0 + uint32_t(t / cols) * tile;
                    const uint32_t ww = std::min(tile, x1 - tx), hh = std::min(tile, y1 - ty);
                    std::vector<V3<double>> sums(size_t(ww) * hh, V3<double>{0, 0, 0});
                    for (uint32_t base = 0; base < rs.spp; base += batch) {
                        const uint32_t end = std::min(rs.spp, base + batch);
                        for (size_t p = 0; p < sums.size(); ++p) {
                            for (uint32_t smp = base; smp < end; ++smp) {
                                const V3<R> v = sample_pixel<R>(s, cfg, cam, rs.width, rs.height, rs.seed,
                                                                tx + uint32_t(p % ww), ty + uint32_t(p / ww), smp, &local[w]);
                                sums[p] = sums[p] + V3<double>{double(v.x), double(v.y), double(v.z)};
                            }
                        }
                    }
                    for (size_t p = 0; p < sums.size(); ++p) {
                        const uint32_t px = tx + uint32_t(p % ww) - x0, py = ty + uint32_t(p / ww) - y0;
                        const size_t   o  = (size_t(py) * cw + px) * 3;
                        const V3<double> mean = sums[p] / static_cast<double>(rs.spp);
                        if (vmax(mean) > std::numeric_limits<float>::max())
                            throw std::overflow_error("Rendered radiance exceeds the float image range");
                        image[o] = float(mean.x);
                        image[o + 1] = float(mean.y);
                        image[o + 2] = float(mean.z);
                        if (sums_out) {
                            sums_out[o] = sums[p].x;
                            sums_out[o + 1] = sums[p].y;
                            sums_out[o + 2] = sums[p].z;
                        }
                    }
                }
            } catch (const std::exception& e) {
                errors[w] = e.what();
                next.store(tiles);
            }
        });
    }
    for (auto& t : pool) t.join();
    for (auto& e : errors)
        if (!e.empty()) throw std::runtime_error(e);
    for (auto& c : local) add_stats(stats, c);
}

extern "C" int orc_render(void* sp, int precision, const OrcCamera* camera, const OrcRender* settings, int threads,
                          uint32_t tile, uint32_t batch, float* image, double* sums, OrcStats* stats) {
    return guarded([&] {
        if (tile == 0 || tile > 128 || batch == 0 || batch > 1024) throw std::invalid_argument("bad tile/batch");
        const Scene& s = *static_cast<Scene*>(sp);
        if (precision == 32)
            render_impl<float>(s, *camera, *settings, threads, tile, batch, image, sums, stats);
        else
            render_impl<double>(s, *camera, *settings, threads, tile, batch, image, sums, stats);
    });
}

extern "C" int orc_sample_pixel(void* sp, int precision, const OrcCamera* camera, const OrcRender* rs, uint32_t x,
                                uint32_t y, uint64_t sample, double* out, OrcStats* stats) {
    return guarded([&] {
        const Scene&   s   = *static_cast<Scene*>(sp);
        const Settings cfg = to_settings(rs);
        if (x >= rs->width || y >= rs->height) throw std::out_of_range("Pixel coordinates are outside the render image");
        PathCounters   c;
        if (precision == 32) {
            const auto v = sample_pixel<float>(s, cfg, CameraRays<float>(*camera), rs->width, rs->height, rs->seed, x, y, sample, &c);
            out[0] = v.x; out[1] = v.y; out[2] = v.z;
        } else {
            const auto v = sample_pixel<double>(s, cfg, CameraRays<double>(*camera), rs->width, rs->height, rs->seed, x, y, sample, &c);
            out[0] = v.x; out[1] = v.y; out[2] = v.z;
        }
        add_stats(stats, c);
    });
}

extern "C" int orc_trace(void* sp, int precision, const OrcRender* rs, const double* r, double tmin, double tmax,
                         int sampler_kind, uint64_t a, uint64_t b, uint64_t c, double* out, OrcStats* stats) {
    return guarded([&] {
        const Scene&   s   = *static_cast<Scene*>(sp);
        const Settings cfg = to_settings(rs);
        Sampler        smp = sampler_kind == 1 ? Sampler::for_sample(a, b, c) : Sampler(a, b);
        PathCounters   pc;
        if (precision == 32) {
            const Ray<float> ray{{float(r[0]), float(r[1]), float(r[2])}, {float(r[3]), float(r[4]), float(r[5])}};
            const auto v = trace<float>(s, cfg, ray, smp, float(tmin), float(tmax), &pc);
            out[0] = v.x; out[1] = v.y; out[2] = v.z;
        } else {
            const Ray<double> ray{{r[0], r[1], r[2]}, {r[3], r[4], r[5]}};
            const auto v = trace<double>(s, cfg, ray, smp, tmin, tmax, &pc);
            out[0] = v.x; out[1] = v.y; out[2] = v.z;
        }
        add_stats(stats, pc);
    });
}

extern "C" void orc_pcg(uint64_t seed, uint64_t sequence, uint64_t n, uint32_t* out) {
    Sampler s(seed, sequence);
    for (uint64_t i = 0; i < n; ++i) out[i] = s.next_u32();
}

extern "C" void orc_for_sample(uint64_t seed, uint64_t pixel, uint64_t sample, uint64_t n, uint32_t* out) {
    Sampler s = Sampler::for_sample(seed, pixel, sample);
    for (uint64_t i = 0; i < n; ++i) out[i] = s.next_u32();
}

extern "C" void orc_next_1d(uint64_t seed, uint64_t pixel, uint64_t sample, uint64_t n, double* out64, float* out32) {
    Sampler a = Sampler::for_sample(seed, pixel, sample), b = a;
    for (uint64_t i = 0; i < n; ++i) {
        if (out64) out64[i] = a.next_1d<double>();
        if (out32) out32[i] = b.next_1d<float>();
    }
}

template <class R> static void warp_impl(int which, const double* u, uint64_t n, double* out) {
    for (uint64_t i = 0; i < n; ++i) {
        const V2<R> uu{R(u[i * 2]), R(u[i * 2 + 1])};
        V3<R>       v{0, 0, 0};
        switch (which) {
        case 0: { const auto d = sample_disk(uu); v = {d.x, d.y, 0}; break; }
        case 1: v = sample_sphere(uu); break;
        case 2: v = sample_hemisphere(uu); break;
        case 3: v = sample_cosine_hemisphere(uu); break;
        case 4: v = sample_triangle(uu); break;
        default: throw std::invalid_argument("unknown warp");
        }
        out[i * 3] = v.x;
        out[i * 3 + 1] = v.y;
        out[i * 3 + 2] = v.z;
    }
}

extern "C" int orc_warp(int which, int precision, const double* u, uint64_t n, double* out) {
    return guarded([&] {
        if (precision == 32) warp_impl<float>(which, u, n, out); else warp_impl<double>(which, u, n, out);
    });
}

extern "C" void orc_sincos_turns(const float* t, uint64_t n, float* s, float* c) {
    for (uint64_t i = 0; i < n; ++i) sincos_turns(t[i], s[i], c[i]);
}

extern "C" int orc_discrete(const double* weights, uint64_t count, double u, double* probabilities, uint64_t* index,
                            double* mass, double* remapped) {
    return guarded([&] {
        const Discrete<double> d(std::vector<double>(weights, weights + count));
        for (uint64_t i = 0; i < count; ++i) probabilities[i] = d.probability(i);
        const auto s = d.sample(u);
        if (!s) throw std::invalid_argument("empty distribution");
        *index    = s->index;
        *mass     = s->mass;
        *remapped = s->remapped;
    });
}

extern "C" double orc_power_heuristic(double a, double b) {
    double v = std::nan("");
    guarded([&] { v = power_heuristic(a, b); });
    return v;
}

extern "C" double orc_fresnel(double cosine, double eta_i, double eta_t) {
    double v = std::nan("");
    guarded([&] { v = fresnel_dielectric(cosine, eta_i, eta_t); });
    return v;
}

static Material to_material(const OrcMaterial* m) { return to_material(*m); }

template <class R> static int bsdf_sample_impl(const OrcMaterial* m, int front, const double* wo, double comp, const double* u, double* out) {
    const Bsdf<R> b(to_material(m), front != 0);
    const auto    s = b.sample(V3<R>{R(wo[0]), R(wo[1]), R(wo[2])}, R(comp), V2<R>{R(u[0]), R(u[1])});
    if (!s) return 0;
    out[0] = s->direction.x; out[1] = s->direction.y; out[2] = s->direction.z;
    out[3] = s->weight.x; out[4] = s->weight.y; out[5] = s->weight.z;
    out[6] = s->probability; out[7] = s->discrete ? 1 : 0; out[8] = s->event; out[9] = s->eta;
    return 1;
}

extern "C" int orc_bsdf_sample(const OrcMaterial* m, int front, int precision, const double* wo, double comp,
                               const double* u, double* out) {
    int found = 0;
    const int rc = guarded([&] {
        found = precision == 32 ? bsdf_sample_impl<float>(m, front, wo, comp, u, out)
                                : bsdf_sample_impl<double>(m, front, wo, comp, u, out);
    });
    return rc ? -1 : found;
}

extern "C" int orc_bsdf_eval(const OrcMaterial* m, int front, int precision, const double* wo, const double* wi,
                             double* value, double* pdf) {
    return guarded([&] {
        if (precision == 32) {
            const Bsdf<float> b(to_material(m), front != 0);
            const V3<float> o{float(wo[0]), float(wo[1]), float(wo[2])}, i{float(wi[0]), float(wi[1]), float(wi[2])};
            const auto v = b.eval(o, i);
            value[0] = v.x; value[1] = v.y; value[2] = v.z;
            *pdf = b.pdf(o, i);
        } else {
            const Bsdf<double> b(to_material(m), front != 0);
            const V3<double> o{wo[0], wo[1], wo[2]}, i{wi[0], wi[1], wi[2]};
            const auto v = b.eval(o, i);
            value[0] = v.x; value[1] = v.y; value[2] = v.z;
            *pdf = b.pdf(o, i);
        }
    });
}

template <class R> static int light_sample_impl(const Scene& s, const double* ref, double lu, const double* u, double* out) {
    const auto l = sample_light(s.lights<R>(), V3<R>{R(ref[0]), R(ref[1]), R(ref[2])}, R(lu), V2<R>{R(u[0]), R(u[1])});
    if (!l) return 0;
    out[0] = l->direction.x; out[1] = l->direction.y; out[2] = l->direction.z;
    out[3] = l->radiance.x; out[4] = l->radiance.y; out[5] = l->radiance.z;
    out[6] = l->distance; out[7] = l->probability; out[8] = l->discrete ? 1 : 0; out[9] = l->kind;
    out[10] = double(l->emitter);
    return 1;
}

extern "C" int orc_light_sample(void* sp, int precision, const double* ref, double lu, const double* u, double* out) {
    int found = 0;
    const int rc = guarded([&] {
        const Scene& s = *static_cast<Scene*>(sp);
        found = precision == 32 ? light_sample_impl<float>(s, ref, lu, u, out) : light_sample_impl<double>(s, ref, lu, u, out);
    });
    return rc ? -1 : found;
}

extern "C" double orc_light_selection(void* sp, uint64_t index) {
    double v = std::nan("");
    guarded([&] { v = static_cast<Scene*>(sp)->lights64.selection.probability(index); });
    return v;
}

extern "C" uint64_t orc_light_count(void* sp) { return static_cast<Scene*>(sp)->lights64.choices.size(); }

extern "C" int orc_camera_ray(const OrcCamera* c, int precision, const double* uv, double aspect, const double* lens, double* out) {
    return guarded([&] {
        if (precision == 32) {
            const CameraRays<float> cam(*c);
            const auto r = cam.lens(V2<float>{float(uv[0]), float(uv[1])}, float(aspect), V2<float>{float(lens[0]), float(lens[1])});
            out[0] = r.o.x; out[1] = r.o.y; out[2] = r.o.z; out[3] = r.d.x; out[4] = r.d.y; out[5] = r.d.z;
        } else {
            const CameraRays<double> cam(*c);
            const auto r = cam.lens(V2<double>{uv[0], uv[1]}, aspect, V2<double>{lens[0], lens[1]});
            out[0] = r.o.x; out[1] = r.o.y; out[2] = r.o.z; out[3] = r.d.x; out[4] = r.d.y; out[5] = r.d.z;
        }
    });
}
