// Device data layout shared by the host builder/scheduler (C++) and the HIP kernels.
//
// HBM layout of one resident scene:
//   nodes      [node_count]    GpuNodeQ, 64 B each (half a cache line): the four child boxes of an
//                              inner node, quantized to 8 bits per plane against an exact per-node
//                              grid, plus the four child references, so one node fetch gives the
//                              whole four-box test. Depth-first order, root = 0.
//   tris       [n]             GpuTri, 48 B: three float4 vertices in leaf order; v0.w carries the
//                              scene-order primitive index (the tie-break key of bvh.cpp:273-276),
//                              v1.w the sidedness flags and v2.w the instance (material) index.
//   normals    [n]             GpuTriNormals, 48 B: the three corner normals normalized, with the
//                              triangle's unit geometric normal in the w components, in leaf order.
//   slot_emitter [n]           area-light choice of each triangle or -1, in leaf order.
//   materials  [instances]     GpuMaterial.
//   uvs        [n]             per-corner texture coordinates in leaf order (3 x float2), if any.
//   textures                   level-0 texels of the base-color textures (float4 each).
//   lights                     light choices, area emitters, directional lights and the selection
//                              CDF, all float, built on the host exactly as the FP32 twin builds them.
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define MRG_HD __host__ __device__
#else
#define MRG_HD
#endif

namespace mrg {

constexpr uint32_t kLeafBit    = 0x80000000u;
constexpr uint32_t kEmptyRef   = 0xffffffffu;
constexpr uint32_t kMaxLeaf    = 16;   // triangles per leaf (count-1 packed in 4 bits)
constexpr uint32_t kMaxDepth   = 30;   // binary-tree depth bound of the builder
constexpr uint32_t kStackLds   = 11;   // stack entries per lane kept in LDS by the path kernels: 22 KB per 256-lane block, seven blocks per CU
constexpr uint32_t kStackLdsQuery = 13;  // the same for ray queries (incoherent rays stack deeper): 26 KB per block, six blocks per CU
constexpr uint32_t kStackTotal = 64;   // LDS + global spill entries per lane (builder-checked bound)

MRG_HD inline uint32_t leaf_ref(uint32_t first, uint32_t count) { return kLeafBit | ((count - 1u) << 27) | first; }
MRG_HD inline bool     is_leaf(uint32_t ref) { return (ref & kLeafBit) != 0u; }
MRG_HD inline uint32_t leaf_first(uint32_t ref) { return ref & 0x07ffffffu; }
MRG_HD inline uint32_t leaf_count(uint32_t ref) { return ((ref >> 27) & 15u) + 1u; }

// Four-wide node as the builder produces it (host only): exact child boxes in SoA order (lo.x[4],
// lo.y[4], lo.z[4], hi.x[4], hi.y[4], hi.z[4]) then the four child references (kEmptyRef for unused
// slots). Uploaded as GpuNodeQ.
struct GpuNode4 {
    float    lox[4], loy[4], loz[4], hix[4], hiy[4], hiz[4];
    uint32_t ref[4];
};

// Four-wide node with quantized child boxes, 64 B. Per axis a, every child plane is
// origin[a] + q * scale[a] with an 8-bit q (byte k of a q word belongs to child k): scale[a] is a
// power of two and origin[a] a multiple of it with |origin[a] / scale[a]| + 255 <= 2^24, so each
// such plane is an exactly representable float and the device decodes it without rounding. The
// builder rounds q_lo down and q_hi up, so every decoded box contains its child's exact box and the
// conservative box test never culls a box the exact one would keep.
struct alignas(64) GpuNodeQ {
    float    origin[3];
    float    scale_x;
    float    scale_y, scale_z;
    uint32_t qlo_x, qlo_y;
    uint32_t qlo_z, qhi_x, qhi_y, qhi_z;
    uint32_t ref[4];
};
static_assert(sizeof(GpuNodeQ) == 64, "GpuNodeQ must be 64 bytes");

struct alignas(16) GpuTri {
    float v0[3];
    uint32_t prim;
    float v1[3];
    uint32_t flags;     // bit 0 single-sided, bit 1 mirrored instance
    float v2[3];
    uint32_t instance;  // material index
};
static_assert(sizeof(GpuTri) == 48, "GpuTri must be 48 bytes");

// Shading normals of one triangle, computed once at upload with the shading kernel's float operations:
// n_i = normalize(corner normal i) and g = normalize(cross(v1 - v0, v2 - v0)) (before the flip of a
// mirrored instance), so surface_hit reads what it would otherwise recompute at every hit.
struct alignas(16) GpuTriNormals {
    float n0[3], gx;
    float n1[3], gy;
    float n2[3], gz;
};
static_assert(sizeof(GpuTriNormals) == 48, "GpuTriNormals must be 48 bytes");

struct GpuMaterial {
    uint32_t kind;
    float    color[3];
    float    emission[3];
    float    roughness;
    float    alpha;       // roughness * roughness (float)
    float    ior;
    uint32_t double_sided;
    uint32_t mirrored;
    uint32_t compensate;  // Kulla-Conty lobe (rough conductors only; see BsdfParams)
    int32_t  texture;     // base-color texture index into GpuScene::textures, -1 = none
    uint32_t filter;      // 0 nearest, 1 bilinear, 2 trilinear (lod 0: the bilinear value)
    int32_t  normal_map;  // tangent-space normal map texture index, -1 = none
    float    normal_scale;
    uint32_t normal_flip; // negate the map's y
    uint32_t pad[2];
};

// Level 0 of a base-color texture: texels[offset + y * width + x] (float4, rgb + 0), row 0 at the
// top of the image.
struct GpuTexture {
    uint64_t offset;
    uint32_t width, height;
};

struct GpuAreaLight {
    float    v0[3], v1[3], v2[3];
    float    normal[3];
    float    radiance[3];
    float    area;
    uint32_t double_sided;
    uint32_t prim;  // scene-order primitive index of the emitter triangle (skipped by its shadow rays)
};

struct GpuLightChoice {
    uint32_t kind;   // 0 area, 1 directional, 2 environment
    uint32_t index;
};

struct GpuDirectional {
    float direction[3];  // toward the light (-normalize(propagation))
    float radiance[3];
};

// Per-scene pointers and constants handed to the kernels by value.
struct GpuScene {
    const GpuNodeQ*       nodes;
    const GpuTri*         tris;
    const GpuTriNormals*  normals;
    const int32_t*        slot_emitter;
    const GpuMaterial*    materials;
    const GpuLightChoice* choices;
    const float*          cdf;          // [choice_count + 1]
    const GpuAreaLight*   areas;
    const GpuDirectional* directional;
    const float*          uvs;          // [tri_count][3][2] per-corner texture coordinates in leaf order, or null
    const GpuTexture*     textures;
    const float*          texels;       // float4 per texel
    uint32_t              node_count;
    uint32_t              tri_count;
    uint32_t              choice_count;
    int32_t               env_choice;
    float                 env[3];
    float                 env_pdf;      // selection(env) * 1/(4 pi), 0 without environment
    // environment map (EnvironmentLight, light.cpp:61-191); env_map < 0: constant environment only
    int32_t               env_map;      // texture index
    uint32_t              env_uniform;  // map without power: uniform sphere sampling
    float                 env_strength;
    float                 env_rotation; // turns
    float                 env_sel;      // selection probability of the environment choice
    uint32_t              env_gw, env_gh;
    const float*          env_cdf;      // [gw * gh + 1]
    const float*          env_rows;     // [gh][4]: solid angle, cos(theta0), cos(theta1), cos difference
};

// Camera constants computed on the host with the FP32 twin's operations (scene.cpp:171-199).
struct GpuCamera {
    float position[3];
    float forward[3];
    float right[3];
    float up[3];
    float half_height;
    float aperture;
    float focus;
    float near_plane;
    float far_plane;
};

struct GpuJob {
    GpuCamera   camera;
    GpuScene    scene;
    uint32_t    width, height;
    uint32_t    mode, max_depth;
    uint32_t    rr, rr_depth;
    uint64_t    seed;
    double*     film_sums;    // [w*h*3]
    uint32_t*   film_counts;  // [w*h]
    uint32_t*   error_flag;   // [0] error bits, [1] cancel requested (host), [2] this batch's commit gate
    uint32_t    pad;
};

// A contiguous range of samples for a contiguous range of pixels of one job. The kernels map a path
// (pixel, sample) of the chunk to its path slot with chunk_slot / chunk_xy: a chunk of whole
// image rows (tile_rows != 0) holds its samples in groups of 4, each group pixel-major with the 4
// samples of a pixel side by side and the pixels in 4x4 tiles (tile rows top to bottom, tiles left to
// right, the remainder rows and columns in the last tile row and column); samples past the last whole
// group follow one at a time in the same pixel order. Any other chunk is sample-major in scanline
// order: slot i -> sample = sample_begin + i / pixel_count, pixel = pixel_begin + i % pixel_count.
struct GpuChunk {
    uint32_t job;
    uint32_t pixel_begin;
    uint32_t pixel_count;
    uint32_t sample_begin;
    uint32_t sample_count;
    uint32_t slot_begin;    // first path slot of this chunk in the pool
    uint32_t pixel_prefix;  // pixels of all earlier chunks of the batch
    uint32_t tile_rows;     // image rows of a chunk of whole rows (4x4 pixel tiles x 4 samples), else 0 (scanline order)
    uint32_t unit_prefix;   // camera-ray work units of all earlier chunks of the batch (k_generate: 4 per pixel of a
                            // chunk of whole rows, 1 per pixel otherwise)
};

}  // namespace mrg
