// Device arithmetic with the FP32 twin's exact operation order.
//
// Every expression below is written in the order the reference's glm/std code evaluates it
// (dot = (x*x' + y*y') + z*z', cross as glm::cross, normalize = v * (1/sqrt(v.v)),
// std::max(a,b) = a < b ? b : a, std::max({a,b,c}) = first largest). The library is compiled with
// -ffp-contract=off, so no product is fused unless an fmaf is written.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mrg {

struct F3 {
    float x, y, z;
};

__device__ __forceinline__ F3 mk(float x, float y, float z) { return F3{x, y, z}; }
__device__ __forceinline__ F3 operator+(F3 a, F3 b) { return F3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ F3 operator-(F3 a, F3 b) { return F3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ F3 operator-(F3 a) { return F3{-a.x, -a.y, -a.z}; }
__device__ __forceinline__ F3 operator*(F3 a, F3 b) { return F3{a.x * b.x, a.y * b.y, a.z * b.z}; }
__device__ __forceinline__ F3 operator*(F3 a, float s) { return F3{a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ F3 operator*(float s, F3 a) { return F3{s * a.x, s * a.y, s * a.z}; }
__device__ __forceinline__ F3 operator/(F3 a, float s) { return F3{a.x / s, a.y / s, a.z / s}; }
__device__ __forceinline__ F3 operator/(F3 a, F3 b) { return F3{a.x / b.x, a.y / b.y, a.z / b.z}; }
__device__ __forceinline__ float dot(F3 a, F3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
__device__ __forceinline__ F3 cross(F3 a, F3 b) {
    return F3{a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
__device__ __forceinline__ float length(F3 v) { return __builtin_sqrtf(dot(v, v)); }
__device__ __forceinline__ F3 normalize(F3 v) { return v * (1.0f / __builtin_sqrtf(dot(v, v))); }
__device__ __forceinline__ F3 vabs(F3 v) { return F3{__builtin_fabsf(v.x), __builtin_fabsf(v.y), __builtin_fabsf(v.z)}; }
__device__ __forceinline__ float smax(float a, float b) { return a < b ? b : a; }
__device__ __forceinline__ float smin(float a, float b) { return b < a ? b : a; }
__device__ __forceinline__ uint32_t smin_u(uint32_t a, uint32_t b) { return b < a ? b : a; }
__device__ __forceinline__ float max3(float a, float b, float c) {
    float m = a;
    if (m < b) m = b;
    if (m < c) m = c;
    return m;
}
__device__ __forceinline__ float vmax(F3 v) { return max3(v.x, v.y, v.z); }
__device__ __forceinline__ float sclamp(float v, float lo, float hi) { return v < lo ? lo : (hi < v ? hi : v); }
__device__ __forceinline__ bool finite3(F3 v) {
    return __builtin_isfinite(v.x) && __builtin_isfinite(v.y) && __builtin_isfinite(v.z);
}
__device__ __forceinline__ F3 ld3(const float* p) { return F3{p[0], p[1], p[2]}; }

// Load through the global address space. Scene pointers reach the kernels inside structures held in
// device memory, so the compiler cannot infer their address space and would emit flat loads, which
// also count against the LDS wait counter and serialise with the LDS traversal stack.
// (The vector forms load a clang ext-vector: a HIP vector class would be copied through a generic
// reference, which gives the flat load back.)
typedef float    gf4 __attribute__((ext_vector_type(4)));
typedef uint32_t gu4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ float4 gld(const float4* p) {
    const gf4 v = *(const __attribute__((address_space(1))) gf4*)p;
    return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ uint4 gld(const uint4* p) {
    const gu4 v = *(const __attribute__((address_space(1))) gu4*)p;
    return make_uint4(v.x, v.y, v.z, v.w);
}
typedef float    gf2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ float2 gld(const float2* p) {
    const gf2 v = *(const __attribute__((address_space(1))) gf2*)p;
    return make_float2(v.x, v.y);
}
__device__ __forceinline__ float gld(const float* p) { return *(const __attribute__((address_space(1))) float*)p; }
__device__ __forceinline__ uint32_t gld(const uint32_t* p) { return *(const __attribute__((address_space(1))) uint32_t*)p; }
__device__ __forceinline__ void gst(float* p, float v) { *(__attribute__((address_space(1))) float*)p = v; }
__device__ __forceinline__ void gst(uint32_t* p, uint32_t v) { *(__attribute__((address_space(1))) uint32_t*)p = v; }
__device__ __forceinline__ float comp(F3 v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : v.z); }

constexpr float kPiF     = 3.14159274101257324219f;  // float(pi)
constexpr float kEpsF    = 1.1920928955078125e-07f;  // FLT_EPSILON
constexpr float kFltMin  = 1.17549435082228750797e-38f;

// std::nextafter(x, y) for float.
__device__ __forceinline__ float next_after(float x, float y) {
    if (__builtin_isnan(x) || __builtin_isnan(y)) return x + y;
    if (x == y) return y;
    if (x == 0.0f) return y > 0.0f ? __int_as_float(1) : __int_as_float(int(0x80000001u));
    int bits = __float_as_int(x);
    if ((x < y) == (x > 0.0f)) ++bits; else --bits;
    return __int_as_float(bits);
}

// sin/cos(2*pi*t): quarter-turn reduction (exact) + Taylor series of sin/cos(2*pi*r), |r| <= 1/8,
// coefficients (-1)^n (2*pi)^k / k! rounded to float, Horner form with fmaf.
__device__ __forceinline__ void sincos_turns(float t, float& s, float& c) {
    const float q  = __builtin_rintf(t * 4.0f);
    const float r  = t - q * 0.25f;
    const float r2 = r * r;
    float ps = __builtin_fmaf(r2, -0x1.e3075p+3f, 0x1.507834p+5f);
    ps       = __builtin_fmaf(r2, ps, -0x1.32d2ccp+6f);
    ps       = __builtin_fmaf(r2, ps, 0x1.466bc6p+6f);
    ps       = __builtin_fmaf(r2, ps, -0x1.4abbcep+5f);
    ps       = __builtin_fmaf(r2, ps, 0x1.921fb6p+2f);
    const float s0 = r * ps;
    float pc = __builtin_fmaf(r2, 0x1.f9d38ap+2f, -0x1.a6d1f2p+4f);
    pc       = __builtin_fmaf(r2, pc, 0x1.e1f506p+5f);
    pc       = __builtin_fmaf(r2, pc, -0x1.55d3c8p+6f);
    pc       = __builtin_fmaf(r2, pc, 0x1.03c1fp+6f);
    pc       = __builtin_fmaf(r2, pc, -0x1.3bd3ccp+4f);
    const float c0 = __builtin_fmaf(r2, pc, 1.0f);
    switch (static_cast<int>(q) & 3) {
    case 0: s = s0; c = c0; break;
    case 1: s = c0; c = -s0; break;
    case 2: s = -s0; c = -c0; break;
    default: s = -c0; c = s0; break;
    }
}

// ---------------------------------------------------------------------------------------------
// Sampler streams (src/core/sampling.cpp:28-65).
struct Pcg {
    uint64_t state, inc;
};

__device__ __forceinline__ uint64_t mix64(uint64_t v) {
    v = (v ^ (v >> 30)) * 0xbf58476d1ce4e5b9ull;
    v = (v ^ (v >> 27)) * 0x94d049bb133111ebull;
    return v ^ (v >> 31);
}

__device__ __forceinline__ uint32_t next_u32(Pcg& p) {
    const uint64_t prev = p.state;
    p.state             = prev * 6364136223846793005ull + p.inc;
    const uint32_t sh   = static_cast<uint32_t>(((prev >> 18) ^ prev) >> 27);
    const uint32_t rot  = static_cast<uint32_t>(prev >> 59);
    return (sh >> rot) | (sh << ((0u - rot) & 31u));
}

__device__ __forceinline__ Pcg pcg_init(uint64_t seed, uint64_t sequence) {
    Pcg p{0, (sequence << 1) | 1u};
    next_u32(p);
    p.state += seed;
    next_u32(p);
    return p;
}

// The pixel's share of a sample stream's set-up (three of its four SplitMix finalisers): a thread
// that draws several samples of one pixel forms it once.
struct PixelStream {
    uint64_t seed;      // PCG seed of every sample of the pixel
    uint64_t sequence;  // the pixel's term of the PCG sequence
};

__device__ __forceinline__ PixelStream pixel_stream(uint64_t seed, uint64_t pixel) {
    return PixelStream{mix64(seed ^ mix64(pixel + 0x9e3779b97f4a7c15ull)), mix64(pixel)};
}

__device__ __forceinline__ Pcg for_sample(const PixelStream& px, uint64_t sample) {
    return pcg_init(px.seed, mix64(sample + 0xd1b54a32d192ed03ull) ^ px.sequence);
}

__device__ __forceinline__ Pcg for_sample(uint64_t seed, uint64_t pixel, uint64_t sample) {
    return for_sample(pixel_stream(seed, pixel), sample);
}

// 53 random bits rounded to float and clamped below one.
__device__ __forceinline__ float next_1d(Pcg& p) {
    const uint32_t high = next_u32(p) >> 5;
    const uint32_t low  = next_u32(p) >> 6;
    const uint64_t bits = (static_cast<uint64_t>(high) << 26) | low;
    const float    v    = static_cast<float>(bits) * 0x1p-53f;
    return smin(v, 0x1.fffffep-1f);
}

}  // namespace mrg
