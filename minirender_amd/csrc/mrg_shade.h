// Surface interaction, BSDFs and light sampling in FP32, operation for operation as the oracle's FP32
// twin. Reference: src/render/bvh.cpp:302-370 (surface_hit), src/render/bsdf.cpp (ShadingFrame,
// diffuse / GGX conductor / mirror / dielectric), src/render/light.cpp:49-59, :266-340 (light
// sampling, area pdf), src/render/path_tracer.cpp:73-109 (power heuristic, offset_origin).
#pragma once

#include "mrg_layout.h"
#include "mrg_math.h"

namespace mrg {

__device__ __forceinline__ float power_heuristic(float a, float b, bool& bad) {
    if (!__builtin_isfinite(a) || !__builtin_isfinite(b) || a < 0.0f || b < 0.0f) bad = true;
    const float scale = smax(a, b);
    if (scale == 0.0f) return 0.0f;
    const float x = a / scale, y = b / scale;
    return x * x / (x * x + y * y);
}

// offset_origin (path_tracer.cpp:87-105) with the FP32 numeric floor FLT_MIN.
__device__ __forceinline__ F3 offset_origin(F3 position, F3 geometric_normal, F3 error, F3 dir) {
    F3 n = normalize(geometric_normal);
    if (dot(n, dir) < 0.0f) n = -n;
    const float distance = smax(dot(vabs(n), error), kFltMin);
    F3          o        = position + distance * n;
    if (n.x != 0.0f) o.x = next_after(o.x, __builtin_copysignf(__builtin_inff(), n.x));
    if (n.y != 0.0f) o.y = next_after(o.y, __builtin_copysignf(__builtin_inff(), n.y));
    if (n.z != 0.0f) o.z = next_after(o.z, __builtin_copysignf(__builtin_inff(), n.z));
    return o;
}

struct Frame {
    F3 t, b, n;
};

__device__ __forceinline__ Frame make_frame(F3 normal) {
    Frame     f;
    const float m = max3(__builtin_fabsf(normal.x), __builtin_fabsf(normal.y), __builtin_fabsf(normal.z));
    f.n           = normalize(normal / m);
    const F3 helper = __builtin_fabsf(f.n.z) < 0.999f ? mk(0.0f, 0.0f, 1.0f) : mk(0.0f, 1.0f, 0.0f);
    f.t             = normalize(cross(helper, f.n));
    f.b             = cross(f.n, f.t);
    return f;
}

__device__ __forceinline__ F3 to_local(const Frame& f, F3 w) { return mk(dot(w, f.t), dot(w, f.b), dot(w, f.n)); }
__device__ __forceinline__ F3 to_world(const Frame& f, F3 l) { return (l.x * f.t + l.y * f.b) + l.z * f.n; }

// --- warps (sampling.cpp:67-107), FP32 twin form
__device__ __forceinline__ void sample_disk(float ux, float uy, float& dx, float& dy) {
    const float sx = 2.0f * ux - 1.0f, sy = 2.0f * uy - 1.0f;
    if (sx == 0.0f && sy == 0.0f) {
        dx = 0.0f;
        dy = 0.0f;
        return;
    }
    const bool  horizontal = __builtin_fabsf(sx) > __builtin_fabsf(sy);
    const float radius     = horizontal ? sx : sy;
    const float turns      = horizontal ? 0.125f * (sy / sx) : 0.25f - 0.125f * (sx / sy);
    float       s, c;
    sincos_turns(turns, s, c);
    dx = radius * c;
    dy = radius * s;
}

__device__ __forceinline__ F3 sample_sphere(float ux, float uy) {
    const float z = 1.0f - 2.0f * ux;
    const float r = __builtin_sqrtf(smax(0.0f, 1.0f - z * z));
    float       s, c;
    sincos_turns(uy, s, c);
    return mk(r * c, r * s, z);
}

__device__ __forceinline__ F3 sample_cosine_hemisphere(float ux, float uy) {
    float dx, dy;
    sample_disk(ux, uy, dx, dy);
    return mk(dx, dy, __builtin_sqrtf(smax(0.0f, 1.0f - (dx * dx + dy * dy))));
}

__device__ __forceinline__ F3 sample_triangle(float ux, float uy) {
    const float root = __builtin_sqrtf(ux);
    return mk(1.0f - root, root * (1.0f - uy), root * uy);
}

// --- BSDF pieces (bsdf.cpp:36-85)
__device__ __forceinline__ float ggx_d(F3 m, float alpha) {
    const float a2    = alpha * alpha;
    const float denom = (m.x * m.x + m.y * m.y) + a2 * m.z * m.z;
    return m.z > 0.0f ? a2 / (kPiF * denom * denom) : 0.0f;
}

__device__ __forceinline__ float g1_over_cos(F3 v, float alpha) {
    if (v.z <= 0.0f) return 0.0f;
    return 2.0f / (v.z + __builtin_sqrtf(v.z * v.z + alpha * alpha * (v.x * v.x + v.y * v.y)));
}

__device__ __forceinline__ float fresnel_ft(float cosine, float sin2, float eta_i, float eta_t) {
    if (eta_i == eta_t) return 0.0f;
    if (sin2 >= 1.0f) return 1.0f;
    const float tc  = __builtin_sqrtf(1.0f - sin2);
    const float par = (eta_t * cosine - eta_i * tc) / (eta_t * cosine + eta_i * tc);
    const float per = (eta_i * cosine - eta_t * tc) / (eta_i * cosine + eta_t * tc);
    return 0.5f * (par * par + per * per);
}

__device__ __forceinline__ F3 schlick(F3 f0, float cosine) {
    const float t  = 1.0f - sclamp(cosine, 0.0f, 1.0f);
    const float t5 = (((t * t) * t) * t) * t;
    return f0 + (mk(1.0f, 1.0f, 1.0f) - f0) * t5;
}

__device__ __forceinline__ F3 sample_vndf(F3 wo, float ux, float uy, float alpha) {
    const F3    view  = normalize(mk(alpha * wo.x, alpha * wo.y, wo.z));
    const float lensq = view.x * view.x + view.y * view.y;
    const F3    t1    = lensq > 0.0f ? mk(-view.y, view.x, 0.0f) / __builtin_sqrtf(lensq) : mk(1.0f, 0.0f, 0.0f);
    const F3    t2    = cross(view, t1);
    const float radius = __builtin_sqrtf(ux);
    float       s, c;
    sincos_turns(uy, s, c);
    const float x     = radius * c;
    const float blend = 0.5f * (1.0f + view.z);
    const float y     = (1.0f - blend) * __builtin_sqrtf(smax(0.0f, 1.0f - x * x)) + blend * radius * s;
    const F3    n     = (x * t1 + y * t2) + __builtin_sqrtf(smax(0.0f, (1.0f - x * x) - y * y)) * view;
    return normalize(mk(alpha * n.x, alpha * n.y, smax(0.0f, n.z)));
}

// --- base-color texture (texture.cpp:74-137), float arithmetic with repeat addressing; uv has its
// origin at the bottom-left, the image at the top-left. Trilinear filtering at lod 0 (the only lod
// the path tracer asks for, path_tracer.cpp:279-282) is the level-0 bilinear value.
__device__ __forceinline__ int wrap_texel(int c, int extent) { return (c % extent + extent) % extent; }

__device__ __forceinline__ F3 texel(const GpuTexture& t, const float* texels, int x, int y) {
    const float4 v = gld(reinterpret_cast<const float4*>(texels) + t.offset +
                         static_cast<uint64_t>(wrap_texel(y, int(t.height))) * t.width + uint32_t(wrap_texel(x, int(t.width))));
    return mk(v.x, v.y, v.z);
}

__device__ __forceinline__ F3 texture_sample(const GpuScene& sc, int32_t index, uint32_t filter, float u, float v) {
    const GpuTexture t  = sc.textures[index];
    const float      cu = u - __builtin_floorf(u);
    const float      cv = 1.0f - (v - __builtin_floorf(v));
    if (filter == 0u) {
        const int x = static_cast<int>(__builtin_floorf(cu * static_cast<float>(t.width)));
        const int y = static_cast<int>(__builtin_floorf(cv * static_cast<float>(t.height)));
        return texel(t, sc.texels, x, y);
    }
    const float x  = cu * static_cast<float>(t.width) - 0.5f;
    const float y  = cv * static_cast<float>(t.height) - 0.5f;
    const float fx = __builtin_floorf(x), fy = __builtin_floorf(y);
    const int   ix = static_cast<int>(fx), iy = static_cast<int>(fy);
    const float tx = x - fx, ty = y - fy;
    const float sx = 1.0f - tx, sy = 1.0f - ty;
    const F3    top    = texel(t, sc.texels, ix, iy) * sx + texel(t, sc.texels, ix + 1, iy) * tx;
    const F3    bottom = texel(t, sc.texels, ix, iy + 1) * sx + texel(t, sc.texels, ix + 1, iy + 1) * tx;
    return top * sy + bottom * ty;
}

// --- environment map (EnvironmentLight, light.cpp:61-191) in the FP32 twin's form: directions to
// latitude-longitude coordinates with atan2_turns, bilinear lookup of max(texel, 0) on level 0
// (longitude wraps, latitude clamps), importance sampling of the grid texels with the host-built
// float CDF, 5% uniform sphere sampling.

// atan2(y, x) / (2 pi): octant reduction to t in [0, 1], t > tan(pi/8) mapped through
// atan(t) = pi/4 + atan((t - 1) / (t + 1)), then the Maclaurin series of arctan (coefficients
// (-1)^k / (2k + 1), k <= 8) in Horner form with fmaf. IEEE sign conventions for zeros.
__device__ __forceinline__ float atan2_turns(float y, float x) {
    const float ax = __builtin_fabsf(x), ay = __builtin_fabsf(y);
    if (ax == 0.0f && ay == 0.0f) return __builtin_signbitf(x) ? (__builtin_signbitf(y) ? -0.5f : 0.5f) : y * 0.0f;
    const bool  swap = ay > ax;
    const float t    = (swap ? ax : ay) / (swap ? ay : ax);
    const bool  big  = t > 0x1.a8279ap-2f;
    const float r    = big ? (t - 1.0f) / (t + 1.0f) : t;
    const float r2   = r * r;
    float p = __builtin_fmaf(r2, 0x1.e1e1e2p-5f, -0x1.111112p-4f);
    p       = __builtin_fmaf(r2, p, 0x1.3b13b2p-4f);
    p       = __builtin_fmaf(r2, p, -0x1.745d18p-4f);
    p       = __builtin_fmaf(r2, p, 0x1.c71c72p-4f);
    p       = __builtin_fmaf(r2, p, -0x1.24924ap-3f);
    p       = __builtin_fmaf(r2, p, 0x1.99999ap-3f);
    p       = __builtin_fmaf(r2, p, -0x1.555556p-2f);
    p       = __builtin_fmaf(r2, p, 1.0f);
    float a = (r * p) * 0x1.45f306p-3f;
    if (big) a = 0.125f + a;
    if (swap) a = 0.25f - a;
    if (__builtin_signbitf(x)) a = 0.5f - a;
    return __builtin_signbitf(y) ? -a : a;
}

__device__ __forceinline__ void env_coordinates(const GpuScene& sc, F3 d, float& u, float& v) {
    d = normalize(d);
    u = atan2_turns(d.z, d.x) + sc.env_rotation;
    u = u - __builtin_floorf(u);
    v = 2.0f * atan2_turns(__builtin_sqrtf(d.x * d.x + d.z * d.z), d.y);
}

__device__ __forceinline__ F3 env_texel(const GpuTexture& t, const float* texels, int col, int row) {
    const int    w = int(t.width), h = int(t.height);
    const int    r = row < 0 ? 0 : (row > h - 1 ? h - 1 : row);
    const float4 c = gld(reinterpret_cast<const float4*>(texels) + t.offset + static_cast<uint64_t>(r) * t.width +
                         uint32_t((col % w + w) % w));
    return mk(smax(c.x, 0.0f), smax(c.y, 0.0f), smax(c.z, 0.0f));
}

// Radiance arriving from direction d (EnvironmentLight::evaluate).
__device__ __forceinline__ F3 env_evaluate(const GpuScene& sc, F3 d) {
    if (sc.env_map < 0) return ld3(sc.env);
    float u, v;
    env_coordinates(sc, d, u, v);
    const GpuTexture t  = sc.textures[sc.env_map];
    const float      x  = u * static_cast<float>(t.width) - 0.5f;
    const float      y  = v * static_cast<float>(t.height) - 0.5f;
    const float      fx = __builtin_floorf(x), fy = __builtin_floorf(y);
    const int        ix = static_cast<int>(fx), iy = static_cast<int>(fy);
    const float      tx = x - fx, ty = y - fy, sx = 1.0f - tx, sy = 1.0f - ty;
    const F3 top    = env_texel(t, sc.texels, ix, iy) * sx + env_texel(t, sc.texels, ix + 1, iy) * tx;
    const F3 bottom = env_texel(t, sc.texels, ix, iy + 1) * sx + env_texel(t, sc.texels, ix + 1, iy + 1) * tx;
    return ld3(sc.env) + sc.env_strength * (top * sy + bottom * ty);
}

// Solid-angle density of sampling d (EnvironmentLight::pdf) for a map with power.
__device__ __forceinline__ float env_map_pdf(const GpuScene& sc, F3 d) {
    float u, v;
    env_coordinates(sc, d, u, v);
    const uint32_t x  = smin_u(static_cast<uint32_t>(u * static_cast<float>(sc.env_gw)), sc.env_gw - 1u);
    const uint32_t y  = smin_u(static_cast<uint32_t>(v * static_cast<float>(sc.env_gh)), sc.env_gh - 1u);
    const uint32_t i  = y * sc.env_gw + x;
    const float    pr = gld(sc.env_cdf + i + 1) - gld(sc.env_cdf + i);
    return 0.05f / (4.0f * kPiF) + ((1.0f - 0.05f) * pr) / gld(sc.env_rows + y * 4u);
}

// Selection probability times the solid-angle density of the environment for direction d
// (LightSampler::environment_pdf).
__device__ __forceinline__ float env_light_pdf(const GpuScene& sc, F3 d) {
    if (sc.env_map < 0 || sc.env_uniform) return sc.env_pdf;
    return sc.env_sel * env_map_pdf(sc, d);
}

// EnvironmentLight::sample: direction, radiance and density.
__device__ __forceinline__ F3 env_sample(const GpuScene& sc, float ux, float uy, F3& radiance, float& pdf) {
    F3 d;
    if (sc.env_map < 0 || sc.env_uniform || ux < 0.05f) {
        d = sample_sphere(sc.env_map < 0 || sc.env_uniform ? ux : ux / 0.05f, uy);
    } else {
        const float t  = (ux - 0.05f) / (1.0f - 0.05f);
        uint32_t    lo = 0, hi = sc.env_gw * sc.env_gh + 1u;  // first index with t < cdf[index]
        while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (t < gld(sc.env_cdf + mid)) hi = mid; else lo = mid + 1;
        }
        const uint32_t index = lo - 1u;
        const float    c0    = gld(sc.env_cdf + index);
        const float    remap = smin((t - c0) / (gld(sc.env_cdf + index + 1) - c0), 0x1.fffffep-1f);
        const uint32_t x = index % sc.env_gw, y = index / sc.env_gw;
        const float4   row = gld(reinterpret_cast<const float4*>(sc.env_rows) + y);
        const float    v   = sclamp(uy, 0x1p-24f, 0x1.fffffep-1f);
        float cosine, sine;
        if (y < sc.env_gh / 2u) {
            const float h = (1.0f - row.y) + v * row.w;
            cosine        = 1.0f - h;
            sine          = __builtin_sqrtf(smax(0.0f, h * (2.0f - h)));
        } else {
            const float h = (1.0f + row.z) + (1.0f - v) * row.w;
            cosine        = -1.0f + h;
            sine          = __builtin_sqrtf(smax(0.0f, h * (2.0f - h)));
        }
        float sp, cp;
        sincos_turns((static_cast<float>(x) + remap) / static_cast<float>(sc.env_gw) - sc.env_rotation, sp, cp);
        d = mk(sine * cp, cosine, sine * sp);
    }
    radiance = env_evaluate(sc, d);
    pdf      = sc.env_map < 0 || sc.env_uniform ? 1.0f / (4.0f * kPiF) : env_map_pdf(sc, d);
    return d;
}

// --- tangent-space normal map (bvh.cpp:334-359) on the uv tangent frame of the triangle
// (uv_tangents, src/geometry/uv_tangents.cpp); `shading` is replaced when the mapped normal stays
// above the geometric surface.
__device__ __forceinline__ void apply_normal_map(const GpuScene& sc, const GpuMaterial& mat, F3 v0, F3 v1, F3 v2, const float2* uvc,
                                                 float u, float v, F3 geometric, F3& shading) {
    const F3    a  = v1 - v0, b = v2 - v0;
    float       sx = uvc[1].x - uvc[0].x, sy = uvc[1].y - uvc[0].y;
    float       tx = uvc[2].x - uvc[0].x, ty = uvc[2].y - uvc[0].y;
    const float scale = smax(smax(smax(__builtin_fabsf(sx), __builtin_fabsf(sy)), __builtin_fabsf(tx)), __builtin_fabsf(ty));
    if (!__builtin_isfinite(scale) || scale == 0.0f) return;
    sx = sx / scale, sy = sy / scale, tx = tx / scale, ty = ty / scale;
    const float det = sx * ty - sy * tx;
    if (!__builtin_isfinite(det) || __builtin_fabsf(det) < 1e-12f) return;
    const F3    tu = (ty * a - sy * b) / det;
    const F3    tv = ((-tx) * a + sx * b) / det;
    const float lu = length(tu), lv = length(tv);
    if (!__builtin_isfinite(lu) || !__builtin_isfinite(lv) || lu == 0.0f || lv == 0.0f) return;
    const F3 au = tu / lu, av = tv / lv;
    const F3 t0 = au - shading * dot(au, shading);
    if (!(dot(t0, t0) > 1e-12f)) return;
    const F3 t = normalize(t0);
    F3       bt = cross(shading, t);
    if (dot(bt, av) < 0.0f) bt = -bt;
    const F3 texel = texture_sample(sc, mat.normal_map, mat.filter, u, v);
    F3       m     = texel * 2.0f - mk(1.0f, 1.0f, 1.0f);
    m.x            = m.x * mat.normal_scale;
    m.y            = m.y * (mat.normal_scale * (mat.normal_flip ? -1.0f : 1.0f));
    const F3 mapped = (t * m.x + bt * m.y) + shading * m.z;
    if (dot(mapped, geometric) > 1e-8f) shading = normalize(mapped);
}

// --- Kulla-Conty energy compensation (microfacet_energy.cpp:69-89, bsdf.cpp:128-189)
// Bilinear lookup of the energy table ([64][64] float4 nodes (A, B, average, 0); x = cos(theta),
// y = roughness): returns E = clamp(A + B, 0, 1) and the row average clamped to [0, 1].
__device__ __forceinline__ void sample_energy(const float* table, float mu, float roughness, float& e, float& average) {
    const float    x  = sclamp(mu, 0.0f, 1.0f) * 63.0f;
    const float    y  = sclamp(roughness, 0.0f, 1.0f) * 63.0f;
    const uint32_t ix = static_cast<uint32_t>(x), iy = static_cast<uint32_t>(y);
    const uint32_t jx = ix + 1u < 63u ? ix + 1u : 63u, jy = iy + 1u < 63u ? iy + 1u : 63u;
    const float4*  t  = reinterpret_cast<const float4*>(table);
    const float4   a = gld(t + iy * 64u + ix), b = gld(t + iy * 64u + jx);
    const float4   c = gld(t + jy * 64u + ix), d = gld(t + jy * 64u + jx);
    const float    tx = x - static_cast<float>(ix), ty = y - static_cast<float>(iy);
    const float    sx = 1.0f - tx, sy = 1.0f - ty;
    const F3       lo = mk(a.x, a.y, a.z) * sx + mk(b.x, b.y, b.z) * tx;
    const F3       hi = mk(c.x, c.y, c.z) * sx + mk(d.x, d.y, d.z) * tx;
    const F3       v  = lo * sy + hi * ty;
    e                 = sclamp(v.x + v.y, 0.0f, 1.0f);
    average           = sclamp(v.z, 0.0f, 1.0f);
}

__device__ __forceinline__ F3 multiple_scattering(F3 f0, float ei, float eo, float average) {
    if (average >= 1.0f - 1e-6f) return mk(0.0f, 0.0f, 0.0f);
    const F3 one   = mk(1.0f, 1.0f, 1.0f);
    const F3 f     = f0 + (one - f0) / 21.0f;
    const F3 den   = one - f * (1.0f - average);
    const F3 color = ((f * f) * average) / mk(smax(den.x, 1e-8f), smax(den.y, 1e-8f), smax(den.z, 1e-8f));
    return color * (((1.0f - ei) * (1.0f - eo)) / (kPiF * (1.0f - average)));
}

// BSDF of one shading point. With energy compensation the outgoing direction's table values (E(wo),
// the row average and the single-lobe probability clamp(E(wo), 0.05, 0.95), bsdf.cpp:170-172) are
// looked up once per shading point by bsdf_prepare: every eval / pdf / sample of the point uses the
// same normalised wo.
struct BsdfParams {
    uint32_t     kind;
    F3           color;
    float        alpha, eta_i, eta_t, roughness;
    const float* table;     // the device's energy table (kernel-uniform)
    bool         energy;    // compensated rough conductor
    float        e_out;     // E(wo)
    float        average;   // row average
    __device__ __forceinline__ bool  is_delta() const { return kind == 2u || kind == 3u || (kind == 1u && alpha <= 1e-6f); }
    // probability of the single-scattering lobe
    __device__ __forceinline__ float p_single() const { return sclamp(e_out, 0.05f, 0.95f); }
};

__device__ __forceinline__ void bsdf_prepare(BsdfParams& b, F3 wo) {
    b.e_out   = 0.0f;
    b.average = 0.0f;
    if (b.energy) sample_energy(b.table, wo.z, b.roughness, b.e_out, b.average);
}

__device__ __forceinline__ F3 bsdf_eval_n(const BsdfParams& b, F3 wo, F3 wi) {
    if (wo.z <= 0.0f || wi.z <= 0.0f || b.is_delta()) return mk(0.0f, 0.0f, 0.0f);
    if (b.kind == 0u) return b.color / kPiF;
    const F3 m      = normalize(wo + wi);
    const F3 single = (((schlick(b.color, dot(wo, m)) * ggx_d(m, b.alpha)) * g1_over_cos(wo, b.alpha)) * g1_over_cos(wi, b.alpha)) / 4.0f;
    if (!b.energy) return single;
    float ei, unused;
    sample_energy(b.table, wi.z, b.roughness, ei, unused);
    return single + multiple_scattering(b.color, ei, b.e_out, b.average);
}

__device__ __forceinline__ float bsdf_pdf_n(const BsdfParams& b, F3 wo, F3 wi) {
    if (wo.z <= 0.0f || wi.z <= 0.0f || b.is_delta()) return 0.0f;
    if (b.kind == 0u) return smax(0.0f, wi.z) / kPiF;
    const F3    m      = normalize(wo + wi);
    const float glossy = ggx_d(m, b.alpha) * g1_over_cos(wo, b.alpha) / 4.0f;
    if (!b.energy) return glossy;
    const float p = b.p_single();
    return p * glossy + (1.0f - p) * (smax(0.0f, wi.z) / kPiF);
}

struct BsdfSample {
    F3    dir, weight;
    float probability, eta;
    bool  discrete, transmission;
};

// Bsdf::sample (bsdf.cpp:198-277) with wo already normalised.
__device__ __forceinline__ bool bsdf_sample(const BsdfParams& b, F3 wo, float component, float ux, float uy, BsdfSample& s) {
    if (wo.z <= 0.0f) return false;
    const F3 reflected = mk(-wo.x, -wo.y, wo.z);
    s.transmission     = false;
    if (b.kind == 3u) {
        const float ratio = b.eta_i / b.eta_t;
        const float sin2  = ratio * ratio * (wo.x * wo.x + wo.y * wo.y);
        const float refl  = fresnel_ft(sclamp(wo.z, 0.0f, 1.0f), sin2, b.eta_i, b.eta_t);
        if (component < refl) {
            s.dir         = reflected;
            s.weight      = mk(1.0f, 1.0f, 1.0f);
            s.probability = refl;
            s.discrete    = true;
            s.eta         = 1.0f;
            return true;
        }
        s.dir          = mk(-ratio * wo.x, -ratio * wo.y, -__builtin_sqrtf(smax(0.0f, 1.0f - sin2)));
        s.weight       = b.color * (ratio * ratio);
        s.probability  = 1.0f - refl;
        s.discrete     = true;
        s.transmission = true;
        s.eta          = b.eta_t / b.eta_i;
        return true;
    }
    if (b.is_delta()) {
        s.dir         = reflected;
        s.weight      = b.kind == 2u ? b.color : schlick(b.color, wo.z);
        s.probability = 1.0f;
        s.discrete    = true;
        s.eta         = 1.0f;
        return true;
    }
    F3 wi;
    if (b.kind == 0u || (b.energy && component >= b.p_single())) {
        wi = sample_cosine_hemisphere(ux, uy);
    } else {
        const F3 m = sample_vndf(wo, ux, uy, b.alpha);
        wi         = normalize((2.0f * dot(wo, m)) * m - wo);
    }
    if (wi.z <= 0.0f) return false;
    const F3    ewi     = normalize(wi);
    const float density = bsdf_pdf_n(b, wo, ewi);
    s.weight            = bsdf_eval_n(b, wo, ewi) * wi.z / density;
    s.dir               = wi;
    s.probability       = density;
    s.discrete          = false;
    s.eta               = 1.0f;
    return true;
}

// --- lights (light.cpp)
__device__ __forceinline__ float area_density(F3 ref, F3 point, F3 normal, float area) {
    const F3    diff = point - ref;
    const float d2   = dot(diff, diff);
    if (d2 == 0.0f) return 0.0f;
    const float cosine = __builtin_fabsf(dot(normal, (-diff) / __builtin_sqrtf(d2)));
    return cosine > 0.0f ? d2 / (area * cosine) : 0.0f;
}

struct LightSample {
    F3       dir, radiance;
    float    distance, probability;
    bool     discrete;
    F3       position, normal, error;
    uint32_t prim;  // area emitter's scene-order primitive (its shadow ray skips it), else 0xffffffff
};

// DiscreteDistribution::sample index (upper_bound over the CDF) and mass.
__device__ __forceinline__ uint32_t select_light(const GpuScene& sc, float u, float& mass) {
    uint32_t lo = 0, hi = sc.choice_count + 1;  // first index with u < cdf[index]
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (u < sc.cdf[mid]) hi = mid; else lo = mid + 1;
    }
    const uint32_t index = lo == 0 ? 0u : (lo > sc.choice_count ? sc.choice_count - 1u : lo - 1u);
    mass                 = sc.cdf[index + 1] - sc.cdf[index];
    return index;
}

template <bool EXT>
__device__ __forceinline__ bool sample_light(const GpuScene& sc, F3 ref, float light_u, float sux, float suy, LightSample& out) {
    if (sc.choice_count == 0) return false;
    float          mass;
    const uint32_t idx = select_light(sc, light_u, mass);
    const GpuLightChoice ch = sc.choices[idx];
    if (ch.kind == 1u) {
        const GpuDirectional d = sc.directional[ch.index];
        out.dir         = ld3(d.direction);
        out.radiance    = ld3(d.radiance);
        out.distance    = __builtin_inff();
        out.probability = mass;
        out.discrete    = true;
        out.prim        = 0xffffffffu;
        return true;
    }
    if (ch.kind == 2u) {
        float pdf = 1.0f / (4.0f * kPiF);
        if (EXT) {
            out.dir = env_sample(sc, sux, suy, out.radiance, pdf);
        } else {
            out.dir      = sample_sphere(sux, suy);
            out.radiance = ld3(sc.env);
        }
        out.distance    = __builtin_inff();
        out.probability = mass * pdf;
        out.discrete    = false;
        out.prim        = 0xffffffffu;
        return true;
    }
    const GpuAreaLight& a  = sc.areas[ch.index];
    const F3            v0 = ld3(a.v0), v1 = ld3(a.v1), v2 = ld3(a.v2), nrm = ld3(a.normal);
    const F3            b  = sample_triangle(sux, suy);
    const F3            pt = (b.x * v0 + b.y * v1) + b.z * v2;
    if (!a.double_sided && dot(nrm, ref - pt) <= 0.0f) return false;
    const float density = area_density(ref, pt, nrm, a.area);
    if (density == 0.0f) return false;
    const F3    dir      = pt - ref;
    const float distance = length(dir);
    out.error       = ((vabs(b.x * v0) + vabs(b.y * v1)) + vabs(b.z * v2)) * (7.0f * kEpsF);
    out.dir         = dir / distance;
    out.radiance    = ld3(a.radiance);
    out.distance    = distance;
    out.probability = mass * density;
    out.discrete    = false;
    out.position    = pt;
    out.normal      = nrm;
    out.prim        = a.prim;
    return true;
}

__device__ __forceinline__ float area_pdf(const GpuScene& sc, int32_t choice, F3 ref, F3 position) {
    if (choice < 0) return 0.0f;
    const GpuAreaLight& a   = sc.areas[sc.choices[choice].index];
    const F3            nrm = ld3(a.normal);
    if (!a.double_sided && dot(nrm, ref - position) <= 0.0f) return 0.0f;
    return (sc.cdf[choice + 1] - sc.cdf[choice]) * area_density(ref, position, nrm, a.area);
}

}  // namespace mrg
