// TEST INFRASTRUCTURE ONLY — the CPU oracle for the MI355X path tracer.
//
// A from-scratch restatement of MiniRender's offline CPU path tracer (read, not copied; the
// reference cannot be built here because its dependencies are fetched by URL). It is templated on
// the real type:
//   R = double  follows the reference semantics (FP64 end to end, float mesh storage);
//   R = float   is the "FP32 twin": the same source with float arithmetic, which the HIP kernels
//               must reproduce bit for bit (same operation order, no FP contraction, the same
//               project-written sincos, correctly rounded sqrt/div).
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline may load this code; the product
// never links it.
//
// Reference files followed (paths relative to the reference root):
//   src/core/sampling.cpp            Sampler, warps, DiscreteDistribution
//   src/render/intersection.cpp      RayQuery slab + watertight triangle test
//   src/render/bvh.cpp               SAH build, closest/any-hit traversal, surface_hit
//   src/render/bsdf.cpp              ShadingFrame, diffuse/conductor/mirror/dielectric
//   src/render/light.cpp             LightSampler (area, directional, constant environment)
//   src/render/path_tracer.cpp       power_heuristic, offset_origin, spawn_ray, trace, sample_pixel
//   src/scene/scene.cpp:171-199      Camera::generate_ray / generate_lens_ray
//   src/render/render_session.cpp    tile/batch order of the Film
//
// Parity status of each part is stated in the header of mr_oracle.cpp.
#pragma once

#include <cstdint>

extern "C" {

// Flattened world-space scene (identity instance transforms: positions and normals are already in
// world space, in scene order = instance-major, triangle-minor; bvh.cpp:53-89).
struct OrcMaterial {
    uint32_t kind;           // 0 diffuse, 1 conductor, 2 mirror, 3 dielectric (scene.hpp:105)
    float    base_color[3];
    float    emission[3];
    float    roughness;
    float    ior;
    uint32_t double_sided;
    uint32_t mirrored;       // determinant(object_to_world) < 0
    uint32_t compensate_energy;   // Kulla-Conty multiple-scattering lobe (scene.hpp:125-126)
    uint32_t base_color_texture;  // 1 + texture index, 0 = none
    uint32_t texture_filter;      // 0 nearest, 1 bilinear, 2 trilinear
    uint32_t normal_texture;      // tangent-space normal map: 1 + texture index, 0 = none
    float    normal_scale;
    uint32_t normal_flip_y;
};

struct OrcDirectional {
    float direction[3];
    float color[3];
    float intensity;
};

// Linear RGB texture, level 0 (row 0 = top row of the image).
struct OrcTexture {
    uint32_t     width, height;
    const float* rgb;  // [height][width][3]
};

struct OrcSceneDesc {
    uint64_t               triangle_count;
    const float*           positions;       // [n][3][3]
    const float*           normals;         // [n][3][3] per-corner world normals
    const uint32_t*        instance;        // [n] instance index of each triangle
    const uint32_t*        triangle;        // [n] triangle index inside its instance
    uint32_t               instance_count;
    const OrcMaterial*     materials;       // [instance_count]
    float                  environment[3];
    uint32_t               directional_count;
    const OrcDirectional*  directional;
    const float*           uvs;             // [n][3][2] per-corner texture coordinates or null
    uint32_t               texture_count;
    const OrcTexture*      textures;
    uint32_t               environment_texture;  // 1 + texture index, 0 = none
    float                  environment_strength;
    float                  environment_rotation_degrees;
};

struct OrcCamera {
    float position[3];
    float target[3];
    float up[3];
    float vertical_fov_degrees;
    float near_plane;
    float far_plane;
    float aperture_radius;
    float focus_distance;
};

// Render settings + integrator settings + crop (path_tracer.hpp:21-50).
struct OrcRender {
    uint32_t width, height, spp;
    uint64_t seed;
    uint32_t mode;            // 0 bsdf, 1 light, 2 mis
    uint32_t maximum_depth;
    uint32_t russian_roulette;
    uint32_t roulette_depth;
    uint32_t x0, y0, x1, y1;  // pixel crop [x0,x1) x [y0,y1); x1 = y1 = 0 means full image
};

// PathStatistics (path_tracer.hpp:52-68) followed by TraceStatistics (bvh.hpp:37-40).
struct OrcStats {
    uint64_t paths;
    uint64_t shadow_rays;
    uint64_t scattering_events;
    uint64_t roulette_terminations;
    uint64_t bounds_tests;
    uint64_t triangle_tests;
    uint64_t extension_rays;
    uint64_t shadow_bounds_tests;    // the part of bounds_tests spent on shadow (any-hit) rays
    uint64_t shadow_triangle_tests;  // the part of triangle_tests spent on shadow rays
};

const char* orc_last_error(void);

void* orc_scene_new(const OrcSceneDesc* desc);
void  orc_scene_free(void* scene);
// triangles, nodes, leaves, maximum depth
void orc_scene_build_stats(void* scene, uint64_t out[4]);

// precision: 64 = reference FP64, 32 = FP32 twin. rays: [n][6] (origin, direction) in double; the
// FP32 twin rounds them to float first. Returns 0 on success.
int orc_intersect(void* scene, int precision, const double* rays, uint64_t n, double tmin,
                  double tmax, int threads, int32_t* prim, double* t, double* t_second,
                  int32_t* prim_second, double* barycentric, OrcStats* stats);
int orc_occluded(void* scene, int precision, const double* rays, uint64_t n, double tmin,
                 double tmax, int threads, uint8_t* out, OrcStats* stats);

// Session-order render of a crop (16^2 tiles, 4-sample batches, render_session.cpp:386-459).
// image: [crop_h][crop_w][3] float mean; sums: [crop_h][crop_w][3] double or null.
int orc_render(void* scene, int precision, const OrcCamera* camera, const OrcRender* settings,
               int threads, uint32_t tile, uint32_t batch, float* image, double* sums,
               OrcStats* stats);
int orc_sample_pixel(void* scene, int precision, const OrcCamera* camera,
                     const OrcRender* settings, uint32_t x, uint32_t y, uint64_t sample,
                     double* out, OrcStats* stats);
// trace() with a caller ray. sampler_kind 0: Sampler(a, b); 1: Sampler::for_sample(a, b, c).
int orc_trace(void* scene, int precision, const OrcRender* settings, const double* ray,
              double tmin, double tmax, int sampler_kind, uint64_t a, uint64_t b, uint64_t c,
              double* out, OrcStats* stats);

// Kulla-Conty energy table [64][64][3] (A, B, row average; microfacet_energy.cpp, ibl.cpp:279-297),
// one node's Fresnel-split albedo (ibl.cpp:240-277) and store_average_energy of a one-row table
// whose A channel is e[] (B = 0); returns the row average.
void   orc_energy_table(float* out);
// Texture (texture.cpp) of a [h][w][3] linear image: sample n uvs ([n][2]) at lod with filter
// (0 nearest, 1 bilinear, 2 trilinear) into out [n][3]; or read mip level `level` (out may be null
// to query its size). Return 0 on success.
int    orc_texture_sample(uint32_t w, uint32_t h, const float* rgb, const float* uv, uint64_t n, float lod,
                          uint32_t filter, float* out);
int    orc_texture_level(uint32_t w, uint32_t h, const float* rgb, uint32_t level, uint32_t* out_w,
                         uint32_t* out_h, float* out);
void   orc_integrate_ibl_brdf(double mu, double roughness, uint32_t samples, uint64_t seed, double* out);
double orc_average_energy_of_ramp(const float* e, uint32_t width);

// Environment light of a scene (light.cpp:61-191): radiance and pdf of n directions ([n][3]),
// n samples of u ([n][2]) -> out [n][7] (direction, radiance, pdf), and the luminance integral.
// atan2_turns: the FP32 twin's atan2(y, x) / (2 pi).
int    orc_env_eval(void* scene, int precision, const double* dir, uint64_t n, double* radiance, double* pdf);
int    orc_env_sample(void* scene, int precision, const double* u, uint64_t n, double* out);
double orc_env_integral(void* scene, int precision);
void   orc_atan2_turns(const float* y, const float* x, uint64_t n, float* out);
// SceneBvh::intersect + surface_hit of one ray ([6]): out = position[3], geometric normal[3],
// shading normal[3], uv[2]. Returns 1 on a hit, 0 on a miss, -1 on error.
int    orc_surface_hit(void* scene, int precision, const double* ray, double* out);

// Primitive pieces for known-answer tests.
void   orc_pcg(uint64_t seed, uint64_t sequence, uint64_t n, uint32_t* out);
void   orc_for_sample(uint64_t seed, uint64_t pixel, uint64_t sample, uint64_t n, uint32_t* out);
void   orc_next_1d(uint64_t seed, uint64_t pixel, uint64_t sample, uint64_t n, double* out64,
                   float* out32);
// which: 0 disk, 1 sphere, 2 hemisphere, 3 cosine hemisphere, 4 triangle. out [n][3].
int    orc_warp(int which, int precision, const double* u, uint64_t n, double* out);
void   orc_sincos_turns(const float* t, uint64_t n, float* s, float* c);
int    orc_discrete(const double* weights, uint64_t count, double u, double* probabilities,
                    uint64_t* index, double* mass, double* remapped);
double orc_power_heuristic(double a, double b);
double orc_fresnel(double cosine, double eta_i, double eta_t);
// BSDF sample in the local frame. out: direction[3], weight[3], probability, measure, event, eta.
int    orc_bsdf_sample(const OrcMaterial* material, int front_face, int precision,
                       const double* wo, double component, const double* u, double* out);
int    orc_bsdf_eval(const OrcMaterial* material, int front_face, int precision, const double* wo,
                     const double* wi, double* value, double* pdf);
// Light sample from a reference point. out: direction[3], radiance[3], distance, probability,
// measure, kind, emitter.
int    orc_light_sample(void* scene, int precision, const double* reference, double light_u,
                        const double* u, double* out);
double orc_light_selection(void* scene, uint64_t index);
uint64_t orc_light_count(void* scene);
// Camera ray for film uv, aspect and lens sample. out: origin[3], direction[3].
int    orc_camera_ray(const OrcCamera* camera, int precision, const double* uv, double aspect,
                      const double* lens, double* out);
}
