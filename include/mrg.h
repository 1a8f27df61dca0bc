/*
 * mrg.h — C ABI of the MI355X-native offline path tracer (libmrg.so).
 *
 * Drop-in boundary for MiniRender's offline path-tracing hot path. Host C++ in the reference
 * (RenderSession / PathTracer / SceneBvh) would call these entry points for a new integrator kind
 * (e.g. IntegratorKind::gpu_path_tracing, dispatched in RenderSession::Impl::run,
 * src/render/render_session.cpp:377-381). Plain pointers and sizes only; no C++ or torch types.
 *
 * Every function returns an int status: MRG_OK (0), MRG_INVALID_ARGUMENT (1), MRG_CANCELLED (2),
 * MRG_NOT_ADMITTED (3, over the HBM budget), MRG_NONFINITE (4, nonfinite/negative radiance or a
 * film value above FLT_MAX), MRG_DEVICE_ERROR (5), MRG_TIMEOUT (6). mrg_last_error() returns the
 * calling thread's last message. Handles are opaque and owned by the caller until released.
 */
#ifndef MRG_H
#define MRG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    MRG_OK               = 0,
    MRG_INVALID_ARGUMENT = 1,
    MRG_CANCELLED        = 2,
    MRG_NOT_ADMITTED     = 3,
    MRG_NONFINITE        = 4,
    MRG_DEVICE_ERROR     = 5,
    MRG_TIMEOUT          = 6
};

/* RenderState (include/minirender/render/render_session.hpp:12-23), same order. */
enum {
    MRG_STATE_IDLE       = 0,
    MRG_STATE_PREPARING  = 1,
    MRG_STATE_RUNNING    = 2,
    MRG_STATE_PAUSING    = 3,
    MRG_STATE_PAUSED     = 4,
    MRG_STATE_FINALIZING = 5,
    MRG_STATE_CANCELLING = 6,
    MRG_STATE_CANCELLED  = 7,
    MRG_STATE_COMPLETED  = 8,
    MRG_STATE_FAILED     = 9
};

/* MaterialKind subset on the offline path (scene.hpp:105). */
enum { MRG_DIFFUSE = 0, MRG_CONDUCTOR = 1, MRG_MIRROR = 2, MRG_DIELECTRIC = 3 };

/* SamplingMode (path_tracer.hpp:13). */
enum { MRG_MODE_BSDF = 0, MRG_MODE_LIGHT = 1, MRG_MODE_MIS = 2 };

/*
 * Scheduling policy of a device's job loop.
 *   PLAIN       one job at a time in submission order, one progressive pass (pass_samples spp, or as
 *               many pixels of it as fit the path pool) per batch: the plain drop-in path.
 *   BATCHED     up to max_jobs_in_flight jobs share every batch (fair share of the path pool, jobs
 *               still waiting for their first pass first, leftover capacity handed back out).
 *   PLAIN_FILL  one job at a time, each batch filled with as many samples of that job as fit.
 */
enum { MRG_SCHEDULE_PLAIN = 0, MRG_SCHEDULE_BATCHED = 1, MRG_SCHEDULE_PLAIN_FILL = 2 };
/* Traversal kernels of a bounce: split = one closest-hit and one any-hit launch per bounce; fused =
 * one launch tracing the previous bounce's shadow rays together with this bounce's extension rays;
 * refill = the fused launch where a lane that finishes its ray takes the next queue entry at once. */
enum { MRG_TRAVERSAL_AUTO = 0, MRG_TRAVERSAL_SPLIT = 1, MRG_TRAVERSAL_FUSED = 2, MRG_TRAVERSAL_REFILL = 3 };

typedef struct mrg_device mrg_device;
typedef struct mrg_scene  mrg_scene;
typedef struct mrg_job    mrg_job;

/* TextureFilter (texture.hpp:13). */
enum { MRG_FILTER_NEAREST = 0, MRG_FILTER_BILINEAR = 1, MRG_FILTER_TRILINEAR = 2 };

/* Material (scene.hpp:108-129) of one instance; 68 bytes. */
typedef struct mrg_material {
    uint32_t kind;
    float    base_color[3];
    float    emission[3];
    float    roughness;
    float    ior;
    uint32_t double_sided;
    uint32_t mirrored;           /* determinant(object_to_world) < 0 (bvh.cpp:63) */
    uint32_t compensate_energy;  /* Kulla-Conty lobe of rough conductors (scene.hpp:125-126, bsdf.cpp:128) */
    uint32_t base_color_texture; /* 1 + index into mrg_scene_desc.textures; 0 = none (RenderInstance::base_color_texture) */
    uint32_t texture_filter;     /* MRG_FILTER_* */
    uint32_t normal_texture;     /* tangent-space normal map: 1 + texture index (a linear texture), 0 = none */
    float    normal_scale;       /* scales the map's x and y (scene.hpp:131-133) */
    uint32_t normal_flip_y;      /* negate the map's y (glTF convention) */
} mrg_material;

/* Decoded linear-RGB texture, level 0 of the reference's mip chain (Texture, texture.hpp:22-40):
 * row 0 is the top row of the image, as the reference's Image. */
typedef struct mrg_texture {
    uint32_t     width;
    uint32_t     height;
    const float* rgb; /* [height][width][3] */
} mrg_texture;

/* DirectionalLight (scene.hpp:180-186): direction is the propagation direction. */
typedef struct mrg_directional_light {
    float direction[3];
    float color[3];
    float intensity;
} mrg_directional_light;

/* Camera (scene.hpp:85-103). */
typedef struct mrg_camera {
    float position[3];
    float target[3];
    float up[3];
    float vertical_fov_degrees;
    float near_plane;
    float far_plane;
    float aperture_radius;
    float focus_distance;
} mrg_camera;

/*
 * Flattened SceneSnapshot (scene.hpp:220-230 as flattened by SceneBvh, bvh.cpp:53-89): world-space
 * FP32 triangles in scene order (instance-major, triangle-minor) with per-corner world normals.
 */
typedef struct mrg_scene_desc {
    uint64_t                     triangle_count;
    const float*                 positions;  /* [n][3][3] */
    const float*                 normals;    /* [n][3][3] */
    const uint32_t*              instance;   /* [n] */
    const uint32_t*              triangle;   /* [n] */
    uint32_t                     instance_count;
    const mrg_material*          materials;  /* [instance_count] */
    float                        environment[3];
    uint32_t                     directional_count;
    const mrg_directional_light* directional;
    const float*                 uvs;           /* [n][3][2] per-corner texture coordinates, or NULL (all zero) */
    uint32_t                     texture_count;
    const mrg_texture*           textures;      /* [texture_count] */
    uint32_t                     environment_texture; /* EnvironmentMap (scene.hpp:188-193): 1 + texture index, 0 = none */
    float                        environment_strength;
    float                        environment_rotation_degrees;
} mrg_scene_desc;

typedef struct mrg_scene_info {
    uint64_t triangles;
    uint64_t nodes;
    uint64_t leaves;
    uint64_t maximum_depth;
    uint64_t device_bytes;
    uint64_t reference_count;
    double   build_seconds;
    char     content_hash[33];
} mrg_scene_info;

/* One job: RenderSettings + IntegratorSettings (path_tracer.hpp:21-50) + camera + pass size. */
typedef struct mrg_job_desc {
    mrg_camera camera;
    uint32_t   width;
    uint32_t   height;
    uint32_t   samples_per_pixel;
    uint32_t   mode;
    uint64_t   seed;
    uint32_t   maximum_depth;
    uint32_t   russian_roulette;
    uint32_t   roulette_depth;
    uint32_t   pass_samples; /* samples per pixel per progressive pass; 0 = automatic */
} mrg_job_desc;

/* RenderProgress (render_session.hpp:41-65) + PathStatistics (path_tracer.hpp:52-68) + latency. */
typedef struct mrg_progress {
    uint32_t state;
    int32_t  status;
    uint64_t job_id;
    uint64_t recorded_samples;
    uint64_t total_samples;
    uint64_t completed_pixels;
    double   elapsed_seconds;
    double   queued_seconds;      /* submit -> admission */
    double   first_pass_seconds;  /* submit -> first complete pass published (0 if none yet) */
    double   completion_seconds;  /* submit -> terminal state (0 while running) */
    uint64_t paths;
    uint64_t extension_rays;
    uint64_t shadow_rays;
    uint64_t scattering_events;
    uint64_t roulette_terminations;
    char     error[256];
} mrg_progress;

typedef struct mrg_device_options {
    uint64_t hbm_budget_bytes;   /* admission budget for scenes + films + path pools; 0 = 64 GiB for
                                    scenes and films on top of the path pools */
    uint64_t path_capacity;      /* paths in flight per batch; 0 = automatic (2^29 with >= 256 GiB HBM and
                                    one batch in flight, 2^28 with >= 192 GiB, 2^26 with >= 64 GiB,
                                    else 2^23; halved while the pools would take more than half of the
                                    memory free at open); 176 B of HBM per path */
    uint32_t schedule;           /* MRG_SCHEDULE_PLAIN, _BATCHED or _PLAIN_FILL */
    uint32_t max_jobs_in_flight; /* batched: jobs co-scheduled per batch; 0 = 64 */
    uint32_t use_graph;          /* capture the bounce loop in a hipGraph */
    uint32_t block_size;         /* kernel block size (64/128/256); 0 = 256 */
    uint32_t persistent_blocks_per_cu; /* 0 = automatic */
    uint32_t traversal;          /* MRG_TRAVERSAL_*; AUTO = refill for a batch of >= 2^23 paths with a
                                    scene of >= 2^14 triangles, fused otherwise */
    uint32_t batches_in_flight;  /* 2: a second path pool renders the next batch while the previous one
                                    drains its tail of late bounces (films still commit in batch order);
                                    0 or 1: one batch at a time */
    uint32_t query_precision;    /* hit decisions of the ray-query entry points: 64 (or 0, the default) =
                                    FP64, the reference intersector's arithmetic (RayQuery in double,
                                    intersection.cpp:108-152; bvh.cpp:259-276), so ids, distances and
                                    barycentrics are its own rounded to float; 32 = the render kernels'
                                    FP32 watertight test (the FP32 twin's) */
} mrg_device_options;

/* Hit of the intersection entry points: closest t and scene-order primitive index (-1 = miss). */
typedef struct mrg_hit {
    float   t;
    int32_t primitive;
} mrg_hit;

/* Ray-query counters (TraceStatistics, bvh.hpp:37-40) collected when requested. */
typedef struct mrg_trace_stats {
    uint64_t bounds_tests;
    uint64_t triangle_tests;
} mrg_trace_stats;

const char* mrg_last_error(void);
const char* mrg_version(void);

/* Device lifetime (one process per GPU). */
int mrg_device_open(int ordinal, const mrg_device_options* options, mrg_device** out);
int mrg_device_close(mrg_device* device);
/* Properties: name (<=256 bytes), CU count, HBM bytes, L2 bytes. */
int mrg_device_query(mrg_device* device, char* name, size_t name_capacity, uint32_t* compute_units,
                     uint64_t* hbm_bytes, uint64_t* l2_bytes);

/*
 * Scene upload: builds the SAH BVH on the host (replaces SceneBvh's constructor, bvh.cpp:31-99),
 * flattens it into the device layout and uploads it. Scenes are deduplicated by content hash on a
 * device: uploading identical content returns the resident scene with its reference count raised.
 */
int mrg_scene_upload(mrg_device* device, const mrg_scene_desc* desc, mrg_scene** out);
int mrg_scene_release(mrg_scene* scene);
int mrg_scene_info_get(mrg_scene* scene, mrg_scene_info* out);

/*
 * Closest-hit / any-hit queries (SceneBvh::intersect / occluded, bvh.hpp:53-59; bvh.cpp:206-300).
 * rays: [n][6] FP32 (origin, unit direction); interval [tmin, tmax] closed, tmax may be +inf.
 * Host-buffer variants copy in and out. *_device variants take device pointers already resident
 * in HBM and run on the library's stream; they return after the kernel completes.
 * barycentric (optional): [n][3]. stats (optional): counters (slower instrumented kernel).
 */
int mrg_intersect(mrg_scene* scene, const float* rays, uint64_t n, float tmin, float tmax, mrg_hit* hits,
                  float* barycentric, mrg_trace_stats* stats);
int mrg_occluded(mrg_scene* scene, const float* rays, uint64_t n, float tmin, float tmax, uint8_t* occluded,
                 mrg_trace_stats* stats);
int mrg_intersect_device(mrg_scene* scene, const void* device_rays, uint64_t n, float tmin, float tmax,
                         void* device_hits, mrg_trace_stats* stats, float* kernel_ms);

/*
 * mrg_intersect without any allocation: no counterpart in the reference. For callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). Same arguments and results as mrg_intersect (host rays in, host hits / barycentrics /
 * counters out), but the rays, hits, barycentrics and the query's counters are staged through one
 * fixed workspace per GPU (a 44 MiB static device array inside libmrg, held from load until the process
 * ends) instead of buffers allocated per call. At most 1,048,576 rays per call (MRG_INVALID_ARGUMENT
 * beyond). Calls through one device handle are serialised; calls through two handles of the same GPU
 * share the workspace and race.
 */
int mrg_intersect_noalloc(mrg_scene* scene, const float* rays, uint64_t n, float tmin, float tmax, mrg_hit* hits,
                          float* barycentric, mrg_trace_stats* stats);

/*
 * Jobs (RenderSession::start/pause/resume/cancel/wait/progress/preview/result,
 * render_session.hpp:82-115). Submitting admits the job when the HBM budget allows, otherwise
 * queues it; the device loop renders it progressively in passes of pass_samples spp.
 */
int mrg_job_submit(mrg_device* device, mrg_scene* scene, const mrg_job_desc* desc, mrg_job** out);
int mrg_job_pause(mrg_job* job, int* accepted);
int mrg_job_resume(mrg_job* job, int* accepted);
int mrg_job_cancel(mrg_job* job);
/* Waits until the job is terminal; timeout_ms < 0 waits forever. MRG_TIMEOUT if still running. */
int mrg_job_wait(mrg_job* job, int timeout_ms);
int mrg_job_progress(mrg_job* job, mrg_progress* out);
/* Film read-back: FP64 per-pixel sums [h][w][3] and sample counts [h][w] (either may be null). */
int mrg_job_read_film(mrg_job* job, double* sums, uint32_t* counts);
/* Mean image [h][w][3] float (sum / count, 0 where count = 0). */
int mrg_job_read_image(mrg_job* job, float* rgb);
/* Nearest-neighbour preview with the longest side <= max_dimension (0 = full); writes w, h. */
int mrg_job_preview(mrg_job* job, uint32_t max_dimension, float* rgb, uint32_t* width, uint32_t* height,
                    uint64_t* recorded_samples);
int mrg_job_release(mrg_job* job);

/* Synchronous render of one job (plain path): submit + wait + read image. */
int mrg_render(mrg_device* device, mrg_scene* scene, const mrg_job_desc* desc, float* rgb, mrg_progress* progress);

/* Device-wide kernel milliseconds by kernel class (accumulated since the last reset; measured only
 * when the bounce loop is not graph-captured). With fused traversal, extend_ms and extend_launches
 * cover the fused launches (shadow rays included) and shadow_ms / shadow_launches stay 0;
 * refill_launches counts those of the fused launches that ran the refill driver (k_trace_refill). */
typedef struct mrg_kernel_times {
    double generate_ms;
    double extend_ms;
    double shade_ms;
    double shadow_ms;
    double accumulate_ms;
    uint64_t batches;
    uint64_t extension_rays;
    uint64_t shadow_rays;
    uint64_t extend_launches;
    uint64_t shadow_launches;
    uint64_t refill_launches;
} mrg_kernel_times;
int mrg_device_kernel_times(mrg_device* device, mrg_kernel_times* out, int reset);

/* HBM accounting of the admission policy: budget, bytes held by resident scenes + admitted films,
 * and the fixed path-pool / traversal-stack bytes. */
int mrg_device_memory(mrg_device* device, uint64_t* budget_bytes, uint64_t* resident_bytes, uint64_t* pool_bytes);

/*
 * Host-only BVH build of [n][3][3] FP32 triangles (no device needed): out[0] four-wide nodes,
 * out[1] leaves, out[2] maximum binary depth, out[3] largest leaf, out[4] primitives referenced by
 * leaves, out[5] structural errors (a primitive not referenced exactly once, a reference out of range,
 * or a worst-case stack above the traversal bound), out[6] worst-case traversal stack entries.
 * Diagnostic of the builder behind mrg_scene_upload (the BvhBuildStatistics of bvh.hpp:29-34).
 */
int mrg_bvh_build_host(const float* positions, uint64_t n, uint64_t* out);

/*
 * The Kulla-Conty energy table the compensated conductors use (conductor_energy_table(),
 * microfacet_energy.hpp:14-15; baked as bake_ibl_brdf() at its defaults, ibl.hpp:39-45, 72):
 * out [64][64][3] = (A, B, average) per (roughness row, cos(theta) column). Host only.
 */
int mrg_conductor_energy_table(float* out);

#ifdef __cplusplus
}
#endif

#endif /* MRG_H */
