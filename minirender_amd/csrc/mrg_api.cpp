// Host side of libmrg: device context, scene upload (BVH build, light table, dedup cache), ray
// queries, and the job scheduler that renders RenderSession-style jobs in batched wavefront passes.
//
// Reference interfaces replaced (see include/mrg.h): SceneBvh (src/render/bvh.cpp:31-99, 372-384),
// LightSampler construction (src/render/light.cpp:197-264), PathTracer::render / sample_pixel
// (src/render/path_tracer.cpp:368-405) and RenderSession (src/render/render_session.cpp:200-656).
#include "mrg.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "mrg_bvh.h"
#include "mrg_energy.h"
#include "mrg_envmap.h"
#include "mrg_kernels.h"
#include "mrg_layout.h"

using namespace mrg;

namespace {

thread_local std::string g_error;

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void check_hip(hipError_t e, const char* what) {
    if (e != hipSuccess) throw Error(MRG_DEVICE_ERROR, std::string(what) + ": " + hipGetErrorString(e));
}

template <class F> int guarded(F&& f) {
    try {
        f();
        return MRG_OK;
    } catch (const Error& e) {
        g_error = e.what();
        return e.code;
    } catch (const std::invalid_argument& e) {
        g_error = e.what();
        return MRG_INVALID_ARGUMENT;
    } catch (const std::exception& e) {
        g_error = e.what();
        return MRG_DEVICE_ERROR;
    }
}

using Clock = std::chrono::steady_clock;
double seconds_since(Clock::time_point t) { return std::chrono::duration<double>(Clock::now() - t).count(); }

// Device allocation owned by RAII.
struct DevBuf {
    void*  ptr   = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    explicit DevBuf(size_t n) { alloc(n); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void alloc(size_t n) {
        release();
        if (n == 0) return;
        check_hip(hipMalloc(&ptr, n), "hipMalloc");
        bytes = n;
    }
    void release() {
        if (ptr) (void)hipFree(ptr);
        ptr   = nullptr;
        bytes = 0;
    }
    template <class T> T* as() const { return static_cast<T*>(ptr); }
};

// Pinned host allocation owned by RAII: staging of a batch's uploads and read-backs, so that copies
// queued behind a running batch do not block the host thread.
struct PinnedBuf {
    void*  ptr   = nullptr;
    size_t bytes = 0;
    PinnedBuf() = default;
    PinnedBuf(const PinnedBuf&) = delete;
    PinnedBuf& operator=(const PinnedBuf&) = delete;
    ~PinnedBuf() { release(); }
    void alloc(size_t n) {
        release();
        if (n == 0) return;
        check_hip(hipHostMalloc(&ptr, n, hipHostMallocDefault), "hipHostMalloc");
        bytes = n;
    }
    void release() {
        if (ptr) (void)hipHostFree(ptr);
        ptr   = nullptr;
        bytes = 0;
    }
    template <class T> T* as() const { return static_cast<T*>(ptr); }
};

// ---------------------------------------------------------------------------------------------
// Float vector arithmetic in the FP32 twin's order (host side: camera constants, light table).
struct V3f {
    float x, y, z;
};
V3f   sub(V3f a, V3f b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
V3f   mul(V3f a, float s) { return {a.x * s, a.y * s, a.z * s}; }
V3f   divs(V3f a, float s) { return {a.x / s, a.y / s, a.z / s}; }
V3f   neg(V3f a) { return {-a.x, -a.y, -a.z}; }
float dot(V3f a, V3f b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
V3f   cross(V3f a, V3f b) { return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y}; }
V3f   normalize(V3f v) { return mul(v, 1.0f / std::sqrt(dot(v, v))); }
float length(V3f v) { return std::sqrt(dot(v, v)); }
V3f   ld(const float* p) { return {p[0], p[1], p[2]}; }
float luminance(V3f c) { return dot(c, V3f{0.2126f, 0.7152f, 0.0722f}); }
void  st(float* d, V3f v) { d[0] = v.x; d[1] = v.y; d[2] = v.z; }

constexpr float  kPiF = 3.14159274101257324219f;
constexpr double kPi  = 3.14159265358979323846;

// 128-bit content hash (two independent 64-bit multiply-mix lanes over the bytes).
struct Hasher {
    uint64_t a = 0x243f6a8885a308d3ull, b = 0x13198a2e03707344ull;
    static uint64_t mix(uint64_t v) {
        v = (v ^ (v >> 30)) * 0xbf58476d1ce4e5b9ull;
        v = (v ^ (v >> 27)) * 0x94d049bb133111ebull;
        return v ^ (v >> 31);
    }
    void bytes(const void* p, size_t n) {
        const auto* c = static_cast<const unsigned char*>(p);
        size_t      i = 0;
        for (; i + 8 <= n; i += 8) {
            uint64_t w;
            std::memcpy(&w, c + i, 8);
            a = mix(a ^ w) + 0x9e3779b97f4a7c15ull;
            b = mix(b + w * 0xff51afd7ed558ccdull) ^ (b >> 17);
        }
        uint64_t w = n;
        for (; i < n; ++i) w = (w << 8) | c[i];
        a = mix(a ^ w);
        b = mix(b + w);
    }
    std::string hex() const {
        char s[33];
        std::snprintf(s, sizeof s, "%016llx%016llx", static_cast<unsigned long long>(a), static_cast<unsigned long long>(b));
        return s;
    }
};

}  // namespace

// ---------------------------------------------------------------------------------------------
struct mrg_scene {
    mrg_device*    device = nullptr;
    std::string    hash;
    int            refs = 1;
    GpuScene       gpu{};
    mrg_scene_info info{};
    DevBuf         nodes, tris, normals, slot_emitter, materials, choices, cdf, areas, directional;
    DevBuf         uvs, textures, texels, env_cdf, env_rows;
    bool           extended = false;  // uses textures, an environment map or energy compensation
};

struct mrg_job {
    mrg_device*   device = nullptr;
    mrg_scene*    scene  = nullptr;
    mrg_job_desc  desc{};
    GpuCamera     camera{};
    uint64_t      id = 0;
    // guarded by device->mutex
    uint32_t      state = MRG_STATE_PREPARING;
    bool          admitted = false, cancel_requested = false, pause_requested = false;
    int           in_flight = 0;   // batches holding samples of this job that have not landed
    int           readers = 0;     // film reads waiting for the job's in-flight batches to land
    bool          fail_pending = false;  // a landed batch failed; the job fails when the others land
    int32_t       status = MRG_OK;
    std::string   error;
    uint32_t      next_sample = 0;   // first sample of the next chunk
    uint32_t      next_pixel  = 0;   // pixel cursor inside the current sample range
    uint32_t      range_samples = 0; // sample count of the current range (when pixels are split)
    uint64_t      recorded = 0, completed_pixels = 0;
    uint64_t      paths = 0, ext_rays = 0, shadow_rays = 0, scatter = 0, rr_term = 0;
    Clock::time_point submitted, admitted_at, first_pass_at, finished_at;
    bool          first_pass_done = false, finished = false;
    uint64_t      film_bytes = 0;
    bool          film_counted = false;  // film_bytes are part of device->resident_bytes
    int           refs = 1;  // caller + scheduler
    DevBuf        film_sums, film_counts, error_flag;
    // A finished job's film, copied out of HBM (the device buffers are freed at that point).
    std::vector<double>   host_sums;
    std::vector<uint32_t> host_counts;
    bool                  film_on_host = false;
    std::mutex    film_mutex;  // film reads against the film's move to host memory
};

// The chunks of one batch and the jobs they sample.
struct BatchPlan {
    std::vector<mrg_job*>  jobs;
    std::vector<GpuJob>    gjobs;
    std::vector<GpuChunk>  chunks;
    uint32_t               slots = 0, pixels = 0, max_depth = 0;
    bool                   extended = false;  // some job's scene needs the extended shading kernel
    bool                   refill   = false;  // bounce traversal with the refill driver (k_trace_refill)
    bool                   one_scene = true;  // every job of the batch renders the same scene
};

// The resources of one batch in flight: a path pool with its queues, counters and stack-spill area,
// the job and chunk tables, pinned staging for uploads and read-backs, a stream, timing events and
// the captured bounce-loop graphs. A device runs its batches on one or two tracks; with two, the next
// batch renders while the previous one drains its tail of late bounces, and each batch's film commit
// waits for the previous batch's (so every pixel still sums its samples in index order).
struct Track {
    hipStream_t                        stream = nullptr;
    Pool                               pool{};
    DevBuf                             pool_mem, spill_ref, spill_t, counters, stats, header, jobs_buf, chunks_buf;
    PinnedBuf                          up_jobs, up_chunks, up_header, down_counters, down_stats, down_flags;
    std::vector<hipEvent_t>            events;
    hipEvent_t                         committed = nullptr;  // after the batch's validate + accumulate
    std::map<uint32_t, hipGraphExec_t> graphs;  // by bounce count, extended shading, refill traversal, one scene
    // the batch in flight
    bool                               busy = false;
    BatchPlan                          plan;
    uint32_t                           iters = 0;
    bool                               timed = false;
    std::vector<std::pair<int, int>>   marks;
};

struct mrg_device {
    int                ordinal = 0;
    mrg_device_options opt{};
    hipStream_t        stream = nullptr;       // admission work (film clears)
    hipStream_t        copy_stream = nullptr;
    hipStream_t        query_stream = nullptr;  // ray queries: own stream and stack-spill area
    Pool               query_pool{};
    hipDeviceProp_t    props{};
    LaunchShape        shape{};
    // batch tracks (path pools)
    Track              tracks[2];
    uint32_t           track_count = 1;
    DevBuf             query_work, query_spill_ref, query_spill_t;
    DevBuf             energy;  // conductor energy table (mrg_energy.h), float4 per node
    void*              noalloc_scratch = nullptr;  // this GPU's address of the allocation-free query's workspace
    std::mutex         query_mutex;
    uint32_t           max_chunks = 0, max_jobs = 0;
    uint64_t           pool_bytes = 0;
    // scene cache + scheduler
    std::mutex                         mutex;
    std::condition_variable            cv;
    std::map<std::string, mrg_scene*>  scenes;
    std::deque<mrg_job*>               jobs;  // admission order
    // Finished jobs whose film is still in HBM (and in the budget), oldest first: admission moves
    // them to host memory only when a waiting job needs the room, so finishing a job costs nothing.
    std::deque<mrg_job*>               resident_films;
    uint64_t                           next_job_id = 1;
    bool                               shutdown = false;
    // Set once mrg_device_close has stopped the scheduler and freed the device's own buffers and
    // streams. The object itself lives on while scenes or jobs still reference it (`live` counts
    // them), so their entry points keep working on host state (progress, films) or fail cleanly.
    std::atomic<bool>                  closed{false};
    bool                               released = false;  // mrg_device_close has returned (guarded by mutex)
    uint64_t                           live = 0;
    std::thread                        worker;
    uint64_t                           resident_bytes = 0;  // scenes + admitted films
    mrg_kernel_times                   times{};
    hipEvent_t                         query_events[2] = {nullptr, nullptr};
};

namespace {

// An explicit budget covers path pools, scenes and films; the default gives scenes and films 64 GiB
// on top of the path pools (the pools are allocated once, at device open).
uint64_t budget(const mrg_device* d) {
    return d->opt.hbm_budget_bytes ? d->opt.hbm_budget_bytes : (uint64_t(64) << 30) + d->pool_bytes;
}

// Pool bytes per path slot: ray_o, ray_d, hit, beta, rad, prev, rng, sh_o, sh_d, sh_c (16 B each) + 4 u32.
constexpr uint64_t kPathSlotBytes = 10 * 16 + 4 * 4;

void alloc_track(mrg_device* d, Track& t) {
    const uint64_t P   = d->opt.path_capacity;
    const uint64_t per = kPathSlotBytes;
    t.pool_mem.alloc(P * per);
    char* base = t.pool_mem.as<char>();
    auto take  = [&](uint64_t bytes) {
        char* p = base;
        base += bytes;
        return p;
    };
    Pool& pl    = t.pool;
    pl.ray_o    = reinterpret_cast<float4*>(take(P * 16));
    pl.ray_d    = reinterpret_cast<float4*>(take(P * 16));
    pl.hit      = reinterpret_cast<float4*>(take(P * 16));
    pl.beta     = reinterpret_cast<float4*>(take(P * 16));
    pl.rad      = reinterpret_cast<float4*>(take(P * 16));
    pl.prev     = reinterpret_cast<float4*>(take(P * 16));
    pl.rng      = reinterpret_cast<uint4*>(take(P * 16));
    pl.sh_o     = reinterpret_cast<float4*>(take(P * 16));
    pl.sh_d     = reinterpret_cast<float4*>(take(P * 16));
    pl.sh_c     = reinterpret_cast<float4*>(take(P * 16));
    pl.hit_slot = reinterpret_cast<uint32_t*>(take(P * 4));
    pl.q_ext0   = reinterpret_cast<uint32_t*>(take(P * 4));
    pl.q_ext1   = reinterpret_cast<uint32_t*>(take(P * 4));
    pl.q_sh     = reinterpret_cast<uint32_t*>(take(P * 4));
    // spill columns for the larger of the two persistent grids (a lane's column is its global id)
    const uint64_t lanes = uint64_t(std::max({d->shape.traversal_blocks, d->shape.refill_blocks, d->shape.query_blocks, 1u})) * 256;
    t.spill_ref.alloc(lanes * (kStackTotal - kStackLds) * 4);
    t.spill_t.alloc(lanes * (kStackTotal - kStackLds) * 4);
    t.counters.alloc(kCounterWords * 4);
    t.stats.alloc(kStatWords * 8);
    t.header.alloc(kHdrWords * 4);
    pl.spill_ref   = t.spill_ref.as<uint32_t>();
    pl.spill_t     = t.spill_t.as<float>();
    pl.spill_lanes = uint32_t(lanes);
    pl.counters    = t.counters.as<uint32_t>();
    pl.stats       = t.stats.as<uint64_t>();
    pl.header      = t.header.as<uint32_t>();
    pl.capacity    = uint32_t(P);
    t.jobs_buf.alloc(sizeof(GpuJob) * d->max_jobs);
    t.chunks_buf.alloc(sizeof(GpuChunk) * d->max_chunks);
    t.up_jobs.alloc(sizeof(GpuJob) * d->max_jobs);
    t.up_chunks.alloc(sizeof(GpuChunk) * d->max_chunks);
    t.up_header.alloc(kHdrWords * 4);
    t.down_counters.alloc(kCounterWords * 4);
    t.down_stats.alloc(kStatWords * 8);
    t.down_flags.alloc(size_t(d->max_jobs) * 16);
    check_hip(hipStreamCreateWithFlags(&t.stream, hipStreamNonBlocking), "hipStreamCreate");
    t.events.resize(2 * (3 * kMaxBounces + 8));
    for (auto& e : t.events) check_hip(hipEventCreate(&e), "hipEventCreate");
    check_hip(hipEventCreateWithFlags(&t.committed, hipEventDisableTiming), "hipEventCreate");
}

uint64_t track_bytes(const Track& t) { return t.pool_mem.bytes + t.spill_ref.bytes + t.spill_t.bytes; }

void free_track(Track& t) {
    if (t.stream) (void)hipStreamSynchronize(t.stream);
    for (auto& g : t.graphs) (void)hipGraphExecDestroy(g.second);
    t.graphs.clear();
    for (auto& e : t.events) (void)hipEventDestroy(e);
    t.events.clear();
    if (t.committed) (void)hipEventDestroy(t.committed);
    t.committed = nullptr;
    if (t.stream) (void)hipStreamDestroy(t.stream);
    t.stream = nullptr;
    for (DevBuf* b : {&t.pool_mem, &t.spill_ref, &t.spill_t, &t.counters, &t.stats, &t.header, &t.jobs_buf, &t.chunks_buf}) b->release();
    for (PinnedBuf* b : {&t.up_jobs, &t.up_chunks, &t.up_header, &t.down_counters, &t.down_stats, &t.down_flags}) b->release();
}

// The device's batch tracks and the ray queries' own spill area. Two tracks (two path pools) when the
// options ask for two batches in flight. One is the default: on the served stream (64 uneven jobs,
// 8 scenes) two overlapped 2^26-path pools gave 3,005 Msamples/s at a first-pass p50 of 0.136 s, one
// 2^27-path pool 3,095 at 0.154 s and one 2^26-path pool 2,900 at 0.098 s: a queued second batch adds
// a batch of latency for about what a larger single pool gains.
void alloc_pool(mrg_device* d) {
    d->max_jobs   = std::max(1u, d->opt.max_jobs_in_flight);
    d->max_chunks = d->max_jobs * 4;
    alloc_track(d, d->tracks[0]);
    d->track_count = d->opt.batches_in_flight == 2 ? 2u : 1u;
    if (d->track_count == 2) alloc_track(d, d->tracks[1]);
    const uint64_t lanes = uint64_t(std::max({d->shape.traversal_blocks, d->shape.refill_blocks, d->shape.query_blocks, 1u})) * 256;
    d->query_work.alloc(64);
    d->query_spill_ref.alloc(lanes * (kStackTotal - kStackLds) * 4);
    d->query_spill_t.alloc(lanes * (kStackTotal - kStackLds) * 4);
    d->query_pool           = d->tracks[0].pool;
    d->query_pool.spill_ref = d->query_spill_ref.as<uint32_t>();
    d->query_pool.spill_t   = d->query_spill_t.as<float>();
    d->pool_bytes = d->query_spill_ref.bytes + d->query_spill_t.bytes;
    for (uint32_t k = 0; k < d->track_count; ++k) d->pool_bytes += track_bytes(d->tracks[k]);
}

// Camera constants: CameraRays<float> of the oracle (scene.cpp:171-199 in FP32; tan in FP64).
GpuCamera make_camera(const mrg_camera& c) {
    const V3f p = ld(c.position), t = ld(c.target), u = ld(c.up);
    if (!(std::isfinite(p.x) && std::isfinite(p.y) && std::isfinite(p.z) && std::isfinite(t.x) && std::isfinite(t.y) &&
          std::isfinite(t.z) && std::isfinite(u.x) && std::isfinite(u.y) && std::isfinite(u.z)))
        throw std::invalid_argument("Camera vectors must be finite");
    if (!(length(sub(t, p)) > 1e-6f)) throw std::invalid_argument("Camera position and target must differ");
    if (!(length(cross(sub(t, p), u)) > 1e-6f)) throw std::invalid_argument("Camera up must not be parallel to view direction");
    if (!(c.vertical_fov_degrees > 0 && c.vertical_fov_degrees < 179)) throw std::invalid_argument("Camera FOV must be between 0 and 179 degrees");
    if (!(std::isfinite(c.near_plane) && std::isfinite(c.far_plane) && c.near_plane > 0 && c.far_plane > c.near_plane))
        throw std::invalid_argument("Camera clipping planes must satisfy 0 < near < far");
    if (!(std::isfinite(c.aperture_radius) && c.aperture_radius >= 0 && std::isfinite(c.focus_distance) && c.focus_distance > 0))
        throw std::invalid_argument("Camera lens requires nonnegative aperture and positive focus distance");
    GpuCamera g{};
    const V3f fwd   = normalize(sub(t, p));
    const V3f right = normalize(cross(fwd, u));
    const V3f up    = cross(right, fwd);
    st(g.position, p);
    st(g.forward, fwd);
    st(g.right, right);
    st(g.up, up);
    g.half_height = static_cast<float>(std::tan(static_cast<double>(c.vertical_fov_degrees) * (kPi / 180.0) * 0.5));
    g.aperture    = c.aperture_radius;
    g.focus       = c.focus_distance;
    g.near_plane  = c.near_plane;
    g.far_plane   = c.far_plane;
    return g;
}

template <class T> void upload(DevBuf& b, const std::vector<T>& v) {
    b.alloc(std::max<size_t>(v.size() * sizeof(T), 16));
    if (!v.empty()) check_hip(hipMemcpy(b.ptr, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy");
}

mrg_scene* build_scene(mrg_device* d, const mrg_scene_desc& s, const std::string& hash) {
    const auto t0 = Clock::now();
    const uint64_t n = s.triangle_count;
    if (n && (!s.positions || !s.normals || !s.instance || !s.triangle)) throw std::invalid_argument("scene arrays are missing");
    if (s.instance_count && !s.materials) throw std::invalid_argument("materials are missing");
    if (s.directional_count && !s.directional) throw std::invalid_argument("directional lights are missing");
    for (uint64_t p = 0; p < n; ++p) {
        if (s.instance[p] >= s.instance_count) throw std::invalid_argument("instance index out of range");
        const float* v = s.positions + p * 9;
        const double e1[3] = {double(v[3]) - v[0], double(v[4]) - v[1], double(v[5]) - v[2]};
        const double e2[3] = {double(v[6]) - v[0], double(v[7]) - v[1], double(v[8]) - v[2]};
        const double c[3]  = {e1[1] * e2[2] - e2[1] * e1[2], e1[2] * e2[0] - e2[2] * e1[0], e1[0] * e2[1] - e2[0] * e1[1]};
        for (int k = 0; k < 9; ++k)
            if (!std::isfinite(v[k])) throw std::invalid_argument("Transformed triangle position is not finite");
        if ((c[0] * c[0] + c[1] * c[1]) + c[2] * c[2] == 0) throw std::invalid_argument("Transform collapses a triangle at float precision");
    }
    for (uint32_t i = 0; i < s.instance_count; ++i) {
        const mrg_material& m = s.materials[i];
        if (m.kind > 3) throw std::invalid_argument("Unsupported material kind");
        for (int k = 0; k < 3; ++k) {
            if (!(m.base_color[k] >= 0 && m.base_color[k] <= 1)) throw std::invalid_argument("Material base color must be in [0,1]");
            if (!(m.emission[k] >= 0 && std::isfinite(m.emission[k]))) throw std::invalid_argument("Material emission must be nonnegative");
        }
        if (!(m.roughness >= 0 && m.roughness <= 1)) throw std::invalid_argument("Roughness must be in [0,1]");
        if (!(m.ior > 0 && std::isfinite(m.ior))) throw std::invalid_argument("IOR must be positive");
        if (m.compensate_energy > 1) throw std::invalid_argument("compensate_energy must be 0 or 1");
        if (m.texture_filter > MRG_FILTER_TRILINEAR) throw std::invalid_argument("Unknown texture filter");
        if (m.base_color_texture > s.texture_count || m.normal_texture > s.texture_count)
            throw std::invalid_argument("Material references missing texture");
        if (m.normal_texture && !(std::isfinite(m.normal_scale) && m.normal_scale >= 0))
            throw std::invalid_argument("Normal scale must be finite and nonnegative");
    }
    if (s.texture_count && !s.textures) throw std::invalid_argument("textures are missing");
    for (uint32_t i = 0; i < s.texture_count; ++i) {
        // Texture::Texture (texture.cpp:53-72)
        const mrg_texture& t = s.textures[i];
        if (t.width == 0 || t.height == 0 || !t.rgb) throw std::invalid_argument("Texture must not be empty");
        if (t.width > 16384 || t.height > 16384 || uint64_t(t.width) * t.height > 64ull * 1024 * 1024)
            throw std::invalid_argument("Texture exceeds the 64-megapixel budget");
        for (uint64_t k = 0; k < uint64_t(t.width) * t.height * 3; ++k)
            if (!std::isfinite(t.rgb[k])) throw std::invalid_argument("Texture contains nonfinite samples");
    }
    if (n && s.uvs)
        for (uint64_t k = 0; k < n * 6; ++k)
            if (!std::isfinite(s.uvs[k])) throw std::invalid_argument("Texture coordinates must be finite");
    if (s.environment_texture > s.texture_count) throw std::invalid_argument("Environment references missing texture");
    if (s.environment_texture != 0 && !(std::isfinite(s.environment_strength) && s.environment_strength >= 0 &&
                                        std::isfinite(s.environment_rotation_degrees)))
        throw std::invalid_argument("Environment strength and rotation must be finite; strength must be nonnegative");
    auto sc    = std::make_unique<mrg_scene>();
    sc->device = d;
    sc->hash   = hash;
    const BvhBuildResult bvh = build_bvh(s.positions, n, 0);
    if (bvh.max_depth > kMaxDepth || bvh.max_stack > kStackTotal)
        throw Error(MRG_DEVICE_ERROR, "BVH exceeds the traversal stack bound");
    // leaf-order triangle records
    std::vector<GpuTri>        tris(n);
    std::vector<GpuTriNormals> nrms(n);
    std::vector<int32_t>  slot_emitter(n, -1);
    // light table (LightSampler<float> of the oracle, light.cpp:197-264)
    std::vector<GpuLightChoice> choices;
    std::vector<GpuAreaLight>   areas;
    std::vector<GpuDirectional> dirs;
    std::vector<float>          weights;
    std::vector<int32_t>        area_choice(n, -1);
    const V3f env      = ld(s.environment);
    const mrg_texture* env_texture = s.environment_texture ? &s.textures[s.environment_texture - 1] : nullptr;
    const EnvTables    env_tables  = build_env_tables(s.environment, env_texture, env_texture ? s.environment_strength : 1.0f,
                                                      env_texture ? s.environment_rotation_degrees : 0.0f);
    const float env_integral = env_tables.integral;
    float lo[3] = {std::numeric_limits<float>::infinity(), std::numeric_limits<float>::infinity(), std::numeric_limits<float>::infinity()};
    float hi[3] = {-lo[0], -lo[1], -lo[2]};
    for (uint64_t p = 0; p < n; ++p) {
        const mrg_material& m = s.materials[s.instance[p]];
        const V3f radiance    = ld(m.emission);
        V3f v[3];
        for (int k = 0; k < 3; ++k) {
            v[k] = ld(s.positions + p * 9 + k * 3);
            lo[0] = std::min(lo[0], v[k].x); lo[1] = std::min(lo[1], v[k].y); lo[2] = std::min(lo[2], v[k].z);
            hi[0] = std::max(hi[0], v[k].x); hi[1] = std::max(hi[1], v[k].y); hi[2] = std::max(hi[2], v[k].z);
        }
        if (luminance(radiance) == 0) continue;
        const V3f   c    = cross(sub(v[1], v[0]), sub(v[2], v[0]));
        const float area = length(c) * 0.5f;
        if (area == 0) continue;
        area_choice[p] = int32_t(choices.size());
        choices.push_back({0u, uint32_t(areas.size())});
        GpuAreaLight a{};
        st(a.v0, v[0]);
        st(a.v1, v[1]);
        st(a.v2, v[2]);
        st(a.normal, divs(mul(c, m.mirrored ? -1.0f : 1.0f), 2.0f * area));
        st(a.radiance, radiance);
        a.area         = area;
        a.double_sided = m.double_sided ? 1u : 0u;
        a.prim         = uint32_t(p);
        areas.push_back(a);
        weights.push_back(float(m.double_sided ? 2 : 1) * kPiF * area * luminance(radiance));
    }
    const bool empty_bounds = n == 0;
    const V3f  extent       = empty_bounds ? V3f{0, 0, 0} : V3f{hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]};
    float      q            = dot(extent, extent) * 0.25f;
    q                       = q < 1e-12f ? 1e-12f : q;
    const float projected   = kPiF * (empty_bounds ? 1.0f : q);
    for (uint32_t i = 0; i < s.directional_count; ++i) {
        const mrg_directional_light& l = s.directional[i];
        const V3f radiance = mul(ld(l.color), l.intensity);
        if (luminance(radiance) > 0) {
            choices.push_back({1u, uint32_t(dirs.size())});
            GpuDirectional g{};
            st(g.direction, neg(normalize(ld(l.direction))));
            st(g.radiance, radiance);
            dirs.push_back(g);
            weights.push_back(projected * luminance(radiance));
        }
    }
    int32_t env_choice = -1;
    if (env_integral > 0) {
        env_choice = int32_t(choices.size());
        choices.push_back({2u, 0u});
        weights.push_back(projected * env_integral);
    }
    // DiscreteDistribution (sampling.cpp:109-130) in float
    std::vector<float> cdf(weights.size() + 1, 0.0f);
    float maximum = 0;
    for (float w : weights) {
        if (!std::isfinite(w) || w < 0) throw std::invalid_argument("Distribution weights must be finite and nonnegative");
        maximum = maximum < w ? w : maximum;
    }
    for (size_t i = 0; i < weights.size(); ++i) cdf[i + 1] = cdf[i] + (maximum > 0 ? weights[i] / maximum : 1.0f);
    if (!weights.empty()) {
        const float integral = cdf.back();
        for (auto& v : cdf) v /= integral;
    }
    for (uint64_t slot = 0; slot < n; ++slot) {
        const uint32_t p = bvh.order[slot];
        const mrg_material& m = s.materials[s.instance[p]];
        const float* v = s.positions + uint64_t(p) * 9;
        const float* nm = s.normals + uint64_t(p) * 9;
        GpuTri& t = tris[slot];
        t.v0[0] = v[0]; t.v0[1] = v[1]; t.v0[2] = v[2];
        t.v1[0] = v[3]; t.v1[1] = v[4]; t.v1[2] = v[5];
        t.v2[0] = v[6]; t.v2[1] = v[7]; t.v2[2] = v[8];
        t.prim = p;
        t.flags = (m.double_sided ? 0u : 1u) | (m.mirrored ? 2u : 0u);
        t.instance = s.instance[p];
        // surface_hit's per-triangle constants in its float operations (normalize = v * (1 / sqrt(v.v))):
        // the normalized corner normals and the geometric normal before a mirrored instance's flip
        GpuTriNormals& r  = nrms[slot];
        const V3f      gn = normalize(cross(sub(ld(v + 3), ld(v)), sub(ld(v + 6), ld(v))));
        st(r.n0, normalize(ld(nm)));
        st(r.n1, normalize(ld(nm + 3)));
        st(r.n2, normalize(ld(nm + 6)));
        r.gx = gn.x;
        r.gy = gn.y;
        r.gz = gn.z;
        slot_emitter[slot]  = area_choice[p];
    }
    // per-corner uvs in leaf order; level-0 texels of every texture
    std::vector<float> uvs(s.uvs ? n * 6 : 0);
    if (s.uvs)
        for (uint64_t slot = 0; slot < n; ++slot) std::memcpy(&uvs[slot * 6], s.uvs + uint64_t(bvh.order[slot]) * 6, 24);
    std::vector<GpuTexture> texs(s.texture_count);
    uint64_t                texel_count = 0;
    for (uint32_t i = 0; i < s.texture_count; ++i) {
        texs[i] = GpuTexture{texel_count, s.textures[i].width, s.textures[i].height};
        texel_count += uint64_t(s.textures[i].width) * s.textures[i].height;
    }
    std::vector<float> texels(texel_count * 4, 0.0f);
    for (uint32_t i = 0; i < s.texture_count; ++i)
        for (uint64_t k = 0; k < uint64_t(texs[i].width) * texs[i].height; ++k)
            std::memcpy(&texels[(texs[i].offset + k) * 4], s.textures[i].rgb + k * 3, 12);
    std::vector<GpuMaterial> mats(s.instance_count);
    for (uint32_t i = 0; i < s.instance_count; ++i) {
        const mrg_material& m = s.materials[i];
        GpuMaterial&        g = mats[i];
        g.kind = m.kind;
        std::memcpy(g.color, m.base_color, 12);
        std::memcpy(g.emission, m.emission, 12);
        g.roughness    = m.roughness;
        g.alpha        = m.roughness * m.roughness;
        g.ior          = m.ior;
        g.double_sided = m.double_sided;
        g.mirrored     = m.mirrored;
        g.compensate   = m.compensate_energy;
        g.texture      = int32_t(m.base_color_texture) - 1;
        g.filter       = m.texture_filter;
        g.normal_map   = int32_t(m.normal_texture) - 1;
        g.normal_scale = m.normal_texture ? m.normal_scale : 0.0f;
        g.normal_flip  = m.normal_flip_y ? 1u : 0u;
        g.pad[0] = g.pad[1] = 0;
    }
    check_hip(hipSetDevice(d->ordinal), "hipSetDevice");
    upload(sc->nodes, bvh.qnodes);
    upload(sc->tris, tris);
    upload(sc->normals, nrms);
    upload(sc->slot_emitter, slot_emitter);
    upload(sc->materials, mats);
    upload(sc->choices, choices);
    upload(sc->cdf, cdf);
    upload(sc->areas, areas);
    upload(sc->directional, dirs);
    upload(sc->uvs, uvs);
    upload(sc->textures, texs);
    upload(sc->texels, texels);
    upload(sc->env_cdf, env_tables.cdf);
    upload(sc->env_rows, env_tables.rows);
    GpuScene& g     = sc->gpu;
    g.nodes         = sc->nodes.as<GpuNodeQ>();
    g.tris          = sc->tris.as<GpuTri>();
    g.normals       = sc->normals.as<GpuTriNormals>();
    g.slot_emitter  = sc->slot_emitter.as<int32_t>();
    g.materials     = sc->materials.as<GpuMaterial>();
    g.choices       = sc->choices.as<GpuLightChoice>();
    g.cdf           = sc->cdf.as<float>();
    g.areas         = sc->areas.as<GpuAreaLight>();
    g.directional   = sc->directional.as<GpuDirectional>();
    g.uvs           = s.uvs ? sc->uvs.as<float>() : nullptr;
    g.textures      = sc->textures.as<GpuTexture>();
    g.texels        = sc->texels.as<float>();
    g.node_count    = uint32_t(bvh.nodes.size());
    g.tri_count     = uint32_t(n);
    g.choice_count  = uint32_t(choices.size());
    g.env_choice    = env_choice;
    st(g.env, env);
    g.env_pdf = env_choice >= 0 ? (cdf[env_choice + 1] - cdf[env_choice]) * (1.0f / (4.0f * kPiF)) : 0.0f;
    g.env_sel      = env_choice >= 0 ? cdf[env_choice + 1] - cdf[env_choice] : 0.0f;
    g.env_map      = env_texture ? int32_t(s.environment_texture) - 1 : -1;
    g.env_uniform  = env_tables.uniform ? 1u : 0u;
    g.env_strength = env_tables.strength;
    g.env_rotation = env_tables.rotation;
    g.env_gw       = env_tables.gw;
    g.env_gh       = env_tables.gh;
    g.env_cdf      = sc->env_cdf.as<float>();
    g.env_rows     = sc->env_rows.as<float>();
    sc->extended   = env_texture != nullptr || s.texture_count != 0;
    for (uint32_t i = 0; i < s.instance_count; ++i)
        sc->extended = sc->extended || s.materials[i].compensate_energy != 0 || s.materials[i].base_color_texture != 0 ||
                       s.materials[i].normal_texture != 0;
    mrg_scene_info& info = sc->info;
    info.triangles       = n;
    info.nodes           = bvh.nodes.size();
    info.leaves          = bvh.leaves;
    info.maximum_depth   = bvh.max_depth;
    info.device_bytes    = sc->nodes.bytes + sc->tris.bytes + sc->normals.bytes + sc->slot_emitter.bytes +
                        sc->materials.bytes + sc->choices.bytes + sc->cdf.bytes + sc->areas.bytes + sc->directional.bytes +
                        sc->uvs.bytes + sc->textures.bytes + sc->texels.bytes + sc->env_cdf.bytes + sc->env_rows.bytes;
    info.build_seconds = seconds_since(t0);
    std::snprintf(info.content_hash, sizeof info.content_hash, "%s", hash.c_str());
    return sc.release();
}

// ---------------------------------------------------------------------------------------------
// Scheduler.

bool terminal(uint32_t s) {
    return s == MRG_STATE_IDLE || s == MRG_STATE_CANCELLED || s == MRG_STATE_COMPLETED || s == MRG_STATE_FAILED;
}

uint64_t film_bytes_of(const mrg_job_desc& d) { return uint64_t(d.width) * d.height * (24 + 4); }

bool release_job_locked(mrg_job* j);

void uncount_film_locked(mrg_device* d, mrg_job* j) {
    if (!j->film_counted) return;
    d->resident_bytes -= std::min(d->resident_bytes, j->film_bytes);
    j->film_counted = false;
}

// Copies a finished job's film into host memory and frees its HBM; the film stays readable. The
// caller holds the job's film_mutex, and no batch that writes the film is in flight. On a copy
// failure the film stays on the device (and in the budget) until the job is released.
void retire_film(mrg_device* d, mrg_job* j) {
    if (!j->film_sums.ptr || j->film_on_host) return;
    const uint64_t px = uint64_t(j->desc.width) * j->desc.height;
    j->host_sums.resize(px * 3);
    j->host_counts.resize(px);
    if (hipMemcpyAsync(j->host_sums.data(), j->film_sums.ptr, px * 24, hipMemcpyDeviceToHost, d->copy_stream) != hipSuccess ||
        hipMemcpyAsync(j->host_counts.data(), j->film_counts.ptr, px * 4, hipMemcpyDeviceToHost, d->copy_stream) != hipSuccess ||
        hipStreamSynchronize(d->copy_stream) != hipSuccess) {
        j->host_sums.clear();
        j->host_counts.clear();
        return;
    }
    j->film_on_host = true;
    j->film_sums.release();
    j->film_counts.release();
    j->error_flag.release();
    uncount_film_locked(d, j);
}

// Terminal state. An admitted job's film stays in HBM, counted in the budget, on the device's list of
// resident films, until admission needs its room, the job is released or the device closes.
void finish_job_locked(mrg_device* d, mrg_job* j, uint32_t state) {
    j->state       = state;
    j->finished    = true;
    j->finished_at = Clock::now();
    if (j->admitted) {
        if (j->film_counted) d->resident_films.push_back(j);
        j->admitted = false;
    }
    d->cv.notify_all();
}

// Raises an admitted, unfinished job's cancel word for the kernels of its batches in flight: their
// shading ends its paths at the next bounce and their film commit skips it. Device mutex held.
void signal_cancel_locked(mrg_device* d, mrg_job* j) {
    if (!j->admitted || j->finished || d->closed) return;
    std::lock_guard<std::mutex> fl(j->film_mutex);
    if (j->film_on_host || !j->error_flag.ptr) return;
    (void)hipSetDevice(d->ordinal);
    if (launch_signal(j->error_flag.as<uint32_t>() + 1, 1u, d->copy_stream) == hipSuccess) (void)hipStreamSynchronize(d->copy_stream);
}

// Moves the oldest finished films to host memory until `needed` more bytes fit the budget (or
// none are left); returns whether they fit. Scheduler thread, between batches.
bool make_room_locked(mrg_device* d, uint64_t needed) {
    while (d->resident_bytes + needed + d->pool_bytes > budget(d) && !d->resident_films.empty()) {
        mrg_job* j = d->resident_films.front();
        d->resident_films.pop_front();
        std::lock_guard<std::mutex> fl(j->film_mutex);
        retire_film(d, j);
    }
    return d->resident_bytes + needed + d->pool_bytes <= budget(d);
}

// Admission in submission order while the budget allows (films; scenes are already resident).
void admit_locked(mrg_device* d) {
    for (mrg_job* j : d->jobs) {
        if (j->admitted || j->finished) continue;
        if (j->cancel_requested) {
            finish_job_locked(d, j, MRG_STATE_CANCELLED);
            continue;
        }
        if (!make_room_locked(d, j->film_bytes)) break;  // FIFO: later jobs wait too
        j->film_sums.alloc(uint64_t(j->desc.width) * j->desc.height * 3 * 8);
        j->film_counts.alloc(uint64_t(j->desc.width) * j->desc.height * 4);
        j->error_flag.alloc(16);
        check_hip(hipMemsetAsync(j->film_sums.ptr, 0, j->film_sums.bytes, d->stream), "hipMemsetAsync");
        check_hip(hipMemsetAsync(j->film_counts.ptr, 0, j->film_counts.bytes, d->stream), "hipMemsetAsync");
        check_hip(hipMemsetAsync(j->error_flag.ptr, 0, 16, d->stream), "hipMemsetAsync");
        // film reads run on the copy stream: the zeroed film is complete before the job is visible as
        // admitted (the scheduler stream is idle between batches, so this waits for the fills only)
        check_hip(hipStreamSynchronize(d->stream), "hipStreamSynchronize");
        j->admitted     = true;
        j->admitted_at  = Clock::now();
        j->film_counted = true;
        d->resident_bytes += j->film_bytes;
        if (j->state == MRG_STATE_PREPARING) j->state = j->pause_requested ? MRG_STATE_PAUSED : MRG_STATE_RUNNING;
    }
}

// Traversal kernel of a batch under MRG_TRAVERSAL_AUTO: the refill driver (parked leaves, seven waves
// per SIMD) when the batch holds a scene of at least 2^14 triangles and at least 2^23 paths, the
// fused launch otherwise. Measured refill/fused throughput: by scene (768 x 768 x 64 spp batches,
// and the 1080p terrain job) 0.85-0.90 at 36-5k triangles, 1.03-1.10 at 21k-82k, 1.09-1.21 at
// 141k-408k, 1.31 at 2.2M; by batch (16-spp passes, 21k-408k triangles) 0.46-0.67 at 128 x 128,
// 0.60-0.75 at 256 x 256, 0.79-0.91 at 512 x 512 (4.2M paths), 1.03-1.13 at 1024 x 1024 (16.8M).
// A refill pays for every ray it starts and per step of its finer loop; it gains where ray lengths
// vary most within a wave, in larger scenes, and where a launch has rays to refill with.
constexpr uint32_t kRefillMinTriangles = 1u << 14;
constexpr uint32_t kRefillMinPaths     = 1u << 23;

bool refill_scene(const mrg_device* d, const mrg_scene* s) {
    if (d->opt.traversal == MRG_TRAVERSAL_REFILL) return true;
    return d->opt.traversal == MRG_TRAVERSAL_AUTO && s->gpu.tri_count >= kRefillMinTriangles;
}

bool runnable(const mrg_job* j) {
    return j->admitted && !j->finished && !j->cancel_requested && !j->pause_requested && !j->fail_pending && j->readers == 0 &&
           j->state == MRG_STATE_RUNNING && j->next_sample < j->desc.samples_per_pixel;
}

uint32_t pass_of(const mrg_job* j) {
    return j->desc.pass_samples ? j->desc.pass_samples : 16u;
}

// Adds the next chunk of job `j` within `room` path slots; returns false when nothing fits.
bool plan_chunk(mrg_job* j, uint32_t job_slot, uint32_t room, bool fill, BatchPlan& b) {
    const uint32_t pixels    = j->desc.width * j->desc.height;
    const uint32_t remaining = j->desc.samples_per_pixel - j->next_sample;
    GpuChunk c{};
    c.job = job_slot;
    if (j->next_pixel == 0 && pixels <= room) {
        uint32_t k = std::min(remaining, pass_of(j));
        if (fill) k = std::min(remaining, std::max(k, room / pixels));
        k = std::min(k, room / pixels);
        if (k == 0) return false;
        c.pixel_begin  = 0;
        c.pixel_count  = pixels;
        c.sample_begin = j->next_sample;
        c.sample_count = k;
        j->next_sample += k;
    } else {
        // pixel-split range: fixed sample count for the whole range
        if (j->next_pixel == 0) j->range_samples = std::min(remaining, pass_of(j));
        const uint32_t k     = j->range_samples;
        const uint32_t count = std::min(pixels - j->next_pixel, room / k);
        if (count == 0) return false;
        c.pixel_begin  = j->next_pixel;
        c.pixel_count  = count;
        c.sample_begin = j->next_sample;
        c.sample_count = k;
        j->next_pixel += count;
        if (j->next_pixel == pixels) {
            j->next_pixel = 0;
            j->next_sample += k;
        }
    }
    c.slot_begin   = b.slots;
    c.pixel_prefix = b.pixels;
    b.slots += c.pixel_count * c.sample_count;
    b.pixels += c.pixel_count;
    b.chunks.push_back(c);
    return true;
}

GpuJob gpu_job(mrg_job* j) {
    GpuJob g{};
    g.camera      = j->camera;
    g.scene       = j->scene->gpu;
    g.width       = j->desc.width;
    g.height      = j->desc.height;
    g.mode        = j->desc.mode;
    g.max_depth   = j->desc.maximum_depth;
    g.rr          = j->desc.russian_roulette;
    g.rr_depth    = j->desc.roulette_depth;
    g.seed        = j->desc.seed;
    g.film_sums   = j->film_sums.as<double>();
    g.film_counts = j->film_counts.as<uint32_t>();
    g.error_flag  = j->error_flag.as<uint32_t>();
    return g;
}

// Hands the capacity left after the fair shares back out: whole-image chunks take more samples (up to a
// pass each round), pixel-split chunks take more pixels of their current sample range.
void grow_chunks(BatchPlan& b, uint32_t cap) {
    bool grew = true;
    while (grew && b.slots < cap) {
        grew = false;
        for (size_t i = 0; i < b.chunks.size() && b.slots < cap; ++i) {
            GpuChunk& c = b.chunks[i];
            mrg_job*  j = b.jobs[c.job];
            const uint32_t pixels = j->desc.width * j->desc.height;
            if (c.pixel_count == pixels) {
                if (j->next_sample >= j->desc.samples_per_pixel) continue;
                const uint32_t add = std::min({pass_of(j), j->desc.samples_per_pixel - j->next_sample, (cap - b.slots) / pixels});
                if (add == 0) continue;
                c.sample_count += add;
                j->next_sample += add;
                b.slots += add * pixels;
                grew = true;
            } else if (j->next_pixel == c.pixel_begin + c.pixel_count && j->next_pixel != 0) {
                const uint32_t add = std::min(pixels - j->next_pixel, (cap - b.slots) / c.sample_count);
                if (add == 0) continue;
                c.pixel_count += add;
                j->next_pixel += add;
                b.slots += add * c.sample_count;
                b.pixels += add;
                if (j->next_pixel == pixels) {
                    j->next_pixel = 0;
                    j->next_sample += c.sample_count;
                }
                grew = true;
            }
        }
    }
}

// Plans one chunk per job (a pixel never appears twice in one accumulate launch) for each job of `group`
// within a fair share of the capacity still free.
void plan_group(mrg_device* d, const std::vector<mrg_job*>& group, uint32_t cap, bool fill, BatchPlan& b) {
    if (group.empty() || b.slots >= cap) return;
    const uint32_t share = std::max<uint32_t>((cap - b.slots) / uint32_t(group.size()), 1u);
    for (mrg_job* j : group) {
        if (b.slots >= cap) break;
        const uint32_t slot = uint32_t(b.jobs.size());
        if (!plan_chunk(j, slot, std::min(cap - b.slots, share), fill, b)) continue;
        b.jobs.push_back(j);
        b.gjobs.push_back(gpu_job(j));
        b.extended  = b.extended || j->scene->extended;
        b.refill    = b.refill || refill_scene(d, j->scene);
        b.one_scene = b.one_scene && j->scene == b.jobs.front()->scene;
        b.max_depth = std::max(b.max_depth, j->desc.maximum_depth);
    }
}

BatchPlan plan_batch_locked(mrg_device* d, uint32_t cap) {
    BatchPlan b;
    const bool     batched  = d->opt.schedule == MRG_SCHEDULE_BATCHED;
    const bool     fill     = d->opt.schedule == MRG_SCHEDULE_PLAIN_FILL;
    const uint32_t max_jobs = batched ? d->max_jobs : 1u;
    // Plain schedules: the oldest runnable job only (FIFO, one job at a time). Batched: candidates in
    // arrival order, jobs still waiting for their first pass ahead of the others.
    std::vector<mrg_job*> cand;
    for (int pass = 0; pass < (batched ? 2 : 1); ++pass)
        for (mrg_job* j : d->jobs)
            if (runnable(j) && (!batched || j->first_pass_done == (pass == 1)) && cand.size() < max_jobs) cand.push_back(j);
    if (!batched || cand.empty()) {
        plan_group(d, cand, cap, fill, b);
    } else {
        // Scene-affine batches: the batch renders the scene of its most urgent job, shared by every
        // runnable job of that scene, and so runs the one-scene traversal and shading kernels (scene
        // pointers in scalar registers, no per-ray job lookup). Jobs of other scenes join only when the
        // lead scene's jobs leave more than half of the pool empty; they lead the following batches.
        mrg_scene*            lead = cand.front()->scene;
        std::vector<mrg_job*> same, other;
        for (mrg_job* j : cand) (j->scene == lead ? same : other).push_back(j);
        plan_group(d, same, cap, fill, b);
        grow_chunks(b, cap);
        if (b.slots < cap / 2) {
            plan_group(d, other, cap, fill, b);
            grow_chunks(b, cap);
        }
        // recompute slot and pixel offsets
        uint32_t s = 0, px = 0;
        for (auto& c : b.chunks) {
            c.slot_begin   = s;
            c.pixel_prefix = px;
            s += c.pixel_count * c.sample_count;
            px += c.pixel_count;
        }
    }
    if (d->opt.traversal == MRG_TRAVERSAL_AUTO && b.slots < kRefillMinPaths) b.refill = false;
    for (mrg_job* j : b.jobs) {
        j->in_flight++;
        j->refs++;
    }
    return b;
}

bool fused_traversal(const mrg_device* d) { return d->opt.traversal != MRG_TRAVERSAL_SPLIT; }

// The batch's bounce loop on track t: reset, camera rays, (trace, shade) per bounce, the last shadow rays.
void launch_bounce_loop(mrg_device* d, Track& t, bool extended, bool refill, bool one_scene, bool timed) {
    auto trace = [&](uint32_t it) {
        const GpuJob* jobs = t.jobs_buf.as<GpuJob>();
        return refill ? launch_trace_refill(t.pool, jobs, it, one_scene, d->shape, t.stream)
                      : launch_trace(t.pool, jobs, it, one_scene, d->shape, t.stream);
    };
    const GpuJob*   jobs   = t.jobs_buf.as<GpuJob>();
    const GpuChunk* chunks = t.chunks_buf.as<GpuChunk>();
    auto            rec    = [&](int cls) {
        if (!timed) return;
        const int ev = int(t.marks.size());
        check_hip(hipEventRecord(t.events[ev], t.stream), "hipEventRecord");
        t.marks.push_back({ev, cls});
    };
    check_hip(launch_reset(t.pool, t.stream), "k_reset");
    rec(-1);
    check_hip(launch_generate(t.pool, jobs, chunks, d->shape, t.stream), "k_generate");
    rec(0);
    if (fused_traversal(d)) {
        for (uint32_t it = 0; it < t.iters; ++it) {
            check_hip(trace(it), "k_trace");
            rec(1);
            check_hip(launch_shade(t.pool, jobs, it, extended, one_scene, d->shape, t.stream), "k_shade");
            rec(2);
        }
        // the last shading pass's shadow rays (its extension queue is empty: no path continues
        // beyond the batch's deepest bounce)
        check_hip(trace(t.iters), "k_trace");
        rec(1);
    } else {
        for (uint32_t it = 0; it < t.iters; ++it) {
            check_hip(launch_extend(t.pool, jobs, it, d->shape, t.stream), "k_extend");
            rec(1);
            check_hip(launch_shade(t.pool, jobs, it, extended, one_scene, d->shape, t.stream), "k_shade");
            rec(2);
            check_hip(launch_shadow(t.pool, jobs, it, d->shape, t.stream), "k_shadow");
            rec(3);
        }
    }
}

// Queues the batch planned into t.plan on track t without waiting for it. `before`: the commit event
// of the batch queued just before (on the other track), or null; this batch's film commit waits for it.
void enqueue_batch(mrg_device* d, Track& t, hipEvent_t before) {
    BatchPlan& b = t.plan;
    t.iters      = b.max_depth + 1;
    t.timed      = d->opt.use_graph == 0;
    t.marks.clear();
    uint32_t units = 0;
    for (GpuChunk& c : b.chunks) {
        const uint32_t w = b.gjobs[c.job].width;
        c.tile_rows      = c.pixel_begin % w == 0 && c.pixel_count % w == 0 ? c.pixel_count / w : 0u;
        c.unit_prefix    = units;
        units += c.tile_rows ? c.pixel_count * 4u : c.pixel_count;  // k_generate's units (4 per pixel: kSampleGroup)
    }
    const uint32_t hdr[kHdrWords] = {uint32_t(b.chunks.size()), b.slots, b.pixels, t.iters, b.gjobs.size() == 1 ? 1u : 0u, units, 0u, 0u};
    std::memcpy(t.up_jobs.ptr, b.gjobs.data(), b.gjobs.size() * sizeof(GpuJob));
    std::memcpy(t.up_chunks.ptr, b.chunks.data(), b.chunks.size() * sizeof(GpuChunk));
    std::memcpy(t.up_header.ptr, hdr, sizeof hdr);
    check_hip(hipMemcpyAsync(t.jobs_buf.ptr, t.up_jobs.ptr, b.gjobs.size() * sizeof(GpuJob), hipMemcpyHostToDevice, t.stream), "upload jobs");
    check_hip(hipMemcpyAsync(t.chunks_buf.ptr, t.up_chunks.ptr, b.chunks.size() * sizeof(GpuChunk), hipMemcpyHostToDevice, t.stream),
              "upload chunks");
    check_hip(hipMemcpyAsync(t.header.ptr, t.up_header.ptr, sizeof hdr, hipMemcpyHostToDevice, t.stream), "upload header");
    if (d->opt.use_graph) {
        const uint32_t key = ((t.iters * 2u + (b.extended ? 1u : 0u)) * 2u + (b.refill ? 1u : 0u)) * 2u + (b.one_scene ? 1u : 0u);
        auto           it  = t.graphs.find(key);
        if (it == t.graphs.end()) {
            hipGraph_t graph;
            check_hip(hipStreamBeginCapture(t.stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
            launch_bounce_loop(d, t, b.extended, b.refill, b.one_scene, false);
            check_hip(hipStreamEndCapture(t.stream, &graph), "hipStreamEndCapture");
            hipGraphExec_t exec;
            check_hip(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0), "hipGraphInstantiate");
            (void)hipGraphDestroy(graph);
            it = t.graphs.emplace(key, exec).first;
        }
        check_hip(hipGraphLaunch(it->second, t.stream), "hipGraphLaunch");
    } else {
        launch_bounce_loop(d, t, b.extended, b.refill, b.one_scene, t.timed);
    }
    // film commit, in batch order: pixels shared with the previous batch get its samples first
    if (before) check_hip(hipStreamWaitEvent(t.stream, before, 0), "hipStreamWaitEvent");
    if (t.timed) {  // the commit's interval starts after the wait
        const int ev = int(t.marks.size());
        check_hip(hipEventRecord(t.events[ev], t.stream), "hipEventRecord");
        t.marks.push_back({ev, -1});
    }
    const GpuJob*   jobs   = t.jobs_buf.as<GpuJob>();
    const GpuChunk* chunks = t.chunks_buf.as<GpuChunk>();
    check_hip(launch_commit_gate(jobs, uint32_t(b.jobs.size()), t.stream), "k_commit_gate");
    check_hip(launch_validate(t.pool, jobs, chunks, d->shape, t.stream), "k_validate");
    check_hip(launch_accumulate(t.pool, jobs, chunks, d->shape, t.stream), "k_accumulate");
    if (t.timed) {
        const int ev = int(t.marks.size());
        check_hip(hipEventRecord(t.events[ev], t.stream), "hipEventRecord");
        t.marks.push_back({ev, 4});
    }
    check_hip(hipEventRecord(t.committed, t.stream), "hipEventRecord");
    check_hip(hipMemcpyAsync(t.down_counters.ptr, t.counters.ptr, kCounterWords * 4, hipMemcpyDeviceToHost, t.stream), "read counters");
    check_hip(hipMemcpyAsync(t.down_stats.ptr, t.stats.ptr, kStatWords * 8, hipMemcpyDeviceToHost, t.stream), "read stats");
    for (size_t i = 0; i < b.jobs.size(); ++i)
        check_hip(hipMemcpyAsync(t.down_flags.as<uint32_t>() + i * 4, b.jobs[i]->error_flag.ptr, 16, hipMemcpyDeviceToHost, t.stream),
                  "read flags");
    t.busy = true;
}

// A landed batch's jobs: samples recorded, statistics, first pass, and the terminal or paused state once
// the job has no other batch in flight. Device mutex held.
void settle_batch_locked(mrg_device* d, Track& t, const double* acc, bool failed, const std::string& why) {
    BatchPlan&      b     = t.plan;
    const uint32_t* ctr   = t.down_counters.as<uint32_t>();
    const uint64_t* st    = t.down_stats.as<uint64_t>();
    const uint32_t* flags = t.down_flags.as<uint32_t>();
    uint64_t        ext = 0, sh = 0;
    if (!failed)
        for (uint32_t i = 0; i < t.iters; ++i) {
            ext += ctr[kCtrExt + i];
            sh += ctr[kCtrShadow + i];
        }
    d->times.generate_ms += acc[0];
    d->times.extend_ms += acc[1];
    d->times.shade_ms += acc[2];
    d->times.shadow_ms += acc[3];
    d->times.accumulate_ms += acc[4];
    d->times.batches++;
    d->times.extend_launches += fused_traversal(d) ? t.iters + 1 : t.iters;
    d->times.shadow_launches += fused_traversal(d) ? 0 : t.iters;
    d->times.refill_launches += fused_traversal(d) && b.refill ? t.iters + 1 : 0;
    d->times.extension_rays += ext;
    d->times.shadow_rays += sh;
    const bool shared = b.jobs.size() > 1;
    for (size_t i = 0; i < b.jobs.size(); ++i) {
        mrg_job* j = b.jobs[i];
        j->in_flight--;
        const uint32_t flag  = failed ? 0u : flags[i * 4];
        const bool     gated = !failed && flags[i * 4 + 2] != 0u;  // cancelled during the batch: nothing committed
        if (failed) {
            if (!j->fail_pending) {
                j->status = MRG_DEVICE_ERROR;
                j->error  = why;
            }
            j->fail_pending = true;
        } else if (flag != 0 && !gated) {
            if (!j->fail_pending) {
                j->status = MRG_NONFINITE;
                j->error  = (flag & 1u) ? "Path contribution is nonfinite or negative" : "Render sample exceeds the float image range";
            }
            j->fail_pending = true;
        } else if (!gated && !j->fail_pending && !j->finished) {
            uint64_t samples = 0;
            for (const auto& c : b.chunks)
                if (c.job == i) samples += uint64_t(c.pixel_count) * c.sample_count;
            j->recorded += samples;
            j->paths += samples;
            if (!shared) {
                j->ext_rays += ext;
                j->shadow_rays += st[kStatShadowRays];
                j->scatter += st[kStatScatter];
                j->rr_term += st[kStatRoulette];
            }
        }
        const uint64_t pixels = uint64_t(j->desc.width) * j->desc.height;
        if (!j->first_pass_done && j->recorded >= uint64_t(std::min(pass_of(j), j->desc.samples_per_pixel)) * pixels) {
            j->first_pass_done = true;
            j->first_pass_at   = Clock::now();
        }
        const bool all_planned = j->next_sample >= j->desc.samples_per_pixel && j->next_pixel == 0;
        if (j->in_flight == 0 && !j->finished) {
            if (j->fail_pending) {
                finish_job_locked(d, j, MRG_STATE_FAILED);
            } else if (j->cancel_requested) {
                j->status = MRG_CANCELLED;
                finish_job_locked(d, j, MRG_STATE_CANCELLED);
            } else if (all_planned) {
                j->completed_pixels = pixels;
                j->state            = MRG_STATE_FINALIZING;
                finish_job_locked(d, j, MRG_STATE_COMPLETED);
            } else if (j->pause_requested) {
                j->state = MRG_STATE_PAUSED;
            }
        }
        release_job_locked(j);
    }
    b = BatchPlan{};
    t.busy = false;
    d->cv.notify_all();
}

// Waits for track t's batch, reads its kernel times and settles its jobs.
void land_batch(mrg_device* d, Track& t) {
    double      acc[5] = {0, 0, 0, 0, 0};
    bool        failed = false;
    std::string why;
    try {
        check_hip(hipStreamSynchronize(t.stream), "hipStreamSynchronize");
        for (size_t i = 1; t.timed && i < t.marks.size(); ++i) {
            float ms = 0;
            check_hip(hipEventElapsedTime(&ms, t.events[t.marks[i - 1].first], t.events[t.marks[i].first]), "hipEventElapsedTime");
            if (t.marks[i].second >= 0) acc[t.marks[i].second] += ms;
        }
    } catch (const std::exception& e) {
        failed = true;
        why    = e.what();
    }
    std::lock_guard<std::mutex> lock(d->mutex);
    settle_batch_locked(d, t, acc, failed, why);
}

// Enqueue failure: the planned batch never ran; its jobs fail once nothing else of theirs is in flight.
void abandon_batch(mrg_device* d, Track& t, const std::string& why) {
    (void)hipStreamSynchronize(t.stream);
    std::lock_guard<std::mutex> lock(d->mutex);
    const double acc[5] = {0, 0, 0, 0, 0};
    settle_batch_locked(d, t, acc, true, why);
}

void scheduler(mrg_device* d) {
    (void)hipSetDevice(d->ordinal);
    std::deque<Track*> inflight;  // in queue order
    hipEvent_t         last_commit = nullptr;  // commit event of the most recently queued batch
    for (;;) {
        Track* free_track = nullptr;
        for (uint32_t k = 0; k < d->track_count && !free_track; ++k)
            if (!d->tracks[k].busy) free_track = &d->tracks[k];
        bool planned = false;
        {
            std::unique_lock<std::mutex> lock(d->mutex);
            for (;;) {
                if (d->shutdown && inflight.empty()) {
                    // every job still queued ends cancelled (its film, if admitted, retired to host memory
                    // so it stays readable), waiters wake, and the queue drops its references
                    for (mrg_job* j : d->jobs) {
                        if (!j->finished) {
                            j->status = MRG_CANCELLED;
                            finish_job_locked(d, j, MRG_STATE_CANCELLED);
                        }
                    }
                    // the device buffers go away with the device: every film still in HBM moves to host memory
                    for (mrg_job* j : d->resident_films) {
                        std::lock_guard<std::mutex> fl(j->film_mutex);
                        retire_film(d, j);
                    }
                    d->resident_films.clear();
                    for (mrg_job* j : d->jobs) release_job_locked(j);
                    d->jobs.clear();
                    d->cv.notify_all();
                    return;
                }
                if (!d->shutdown) {
                    try {
                        admit_locked(d);
                    } catch (const std::exception& e) {
                        for (mrg_job* j : d->jobs)
                            if (!j->admitted && !j->finished) {
                                j->status = MRG_DEVICE_ERROR;
                                j->error  = e.what();
                                finish_job_locked(d, j, MRG_STATE_FAILED);
                                break;
                            }
                    }
                    // settle pause/cancel of jobs with no batch in flight
                    for (mrg_job* j : d->jobs) {
                        if (j->finished || j->in_flight > 0) continue;
                        if (j->cancel_requested && j->admitted) {
                            j->status = MRG_CANCELLED;
                            finish_job_locked(d, j, MRG_STATE_CANCELLED);
                        } else if (j->pause_requested && j->state == MRG_STATE_PAUSING) {
                            j->state = MRG_STATE_PAUSED;
                            d->cv.notify_all();
                        }
                    }
                    // drop finished jobs from the queue (the caller's handle stays valid)
                    for (auto it = d->jobs.begin(); it != d->jobs.end();) {
                        if ((*it)->finished) {
                            mrg_job* j = *it;
                            it         = d->jobs.erase(it);
                            release_job_locked(j);
                        } else {
                            ++it;
                        }
                    }
                    if (free_track) {
                        free_track->plan = plan_batch_locked(d, free_track->pool.capacity);
                        if (!free_track->plan.jobs.empty()) {
                            planned = true;
                            break;
                        }
                    }
                }
                if (!inflight.empty()) break;  // land the oldest batch
                d->cv.wait(lock);
            }
        }
        if (planned) {
            try {
                enqueue_batch(d, *free_track, last_commit);
                last_commit = free_track->committed;
                inflight.push_back(free_track);
            } catch (const std::exception& e) {
                abandon_batch(d, *free_track, e.what());
            }
            // keep a second batch rendering while the first drains its tail
            bool more = false;
            for (uint32_t k = 0; k < d->track_count; ++k) more = more || !d->tracks[k].busy;
            if (more) continue;
        }
        if (!inflight.empty()) {
            Track* t = inflight.front();
            inflight.pop_front();
            land_batch(d, *t);
        }
    }
}

// Drops one reference of a job; deletes the job (and its scene with the scene's last reference).
// Returns true when the device is closed and nothing references it any more: the caller deletes it
// once it has let go of the device mutex.
bool release_job_locked(mrg_job* j) {
    mrg_device* d = j->device;
    if (--j->refs == 0) {
        mrg_scene* s = j->scene;
        uncount_film_locked(d, j);
        d->resident_films.erase(std::remove(d->resident_films.begin(), d->resident_films.end(), j), d->resident_films.end());
        delete j;
        d->live--;
        if (s && --s->refs == 0) {
            d->scenes.erase(s->hash);
            d->resident_bytes -= std::min(d->resident_bytes, s->info.device_bytes);
            delete s;
            d->live--;
        }
    }
    return d->released && d->live == 0;
}

void validate_job(const mrg_job_desc& d) {
    if (d.width == 0 || d.height == 0 || d.width > 16384 || d.height > 16384 || uint64_t(d.width) * d.height > 64ull * 1024 * 1024 ||
        d.samples_per_pixel == 0)
        throw std::invalid_argument("Render dimensions or sample count exceed the image budget");
    if (d.mode > 2) throw std::invalid_argument("Sampling mode must be bsdf, light or mis");
    if (d.maximum_depth > 128 || d.roulette_depth > 128) throw std::invalid_argument("Path and roulette depths must not exceed 128");
}

}  // namespace

// =============================================================================================
extern "C" {

const char* mrg_last_error(void) { return g_error.c_str(); }
const char* mrg_version(void) { return "mrg 0.1 (gfx950 wavefront path tracer)"; }

int mrg_device_open(int ordinal, const mrg_device_options* options, mrg_device** out) {
    return guarded([&] {
        if (!out) throw std::invalid_argument("out is null");
        int count = 0;
        check_hip(hipGetDeviceCount(&count), "hipGetDeviceCount");
        if (ordinal < 0 || ordinal >= count) throw std::invalid_argument("device ordinal out of range");
        auto d     = std::make_unique<mrg_device>();
        d->ordinal = ordinal;
        if (options) d->opt = *options;
        check_hip(hipSetDevice(ordinal), "hipSetDevice");
        check_hip(hipGetDeviceProperties(&d->props, ordinal), "hipGetDeviceProperties");
        check_hip(hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking), "hipStreamCreate");
        check_hip(hipStreamCreateWithFlags(&d->copy_stream, hipStreamNonBlocking), "hipStreamCreate");
        check_hip(hipStreamCreateWithFlags(&d->query_stream, hipStreamNonBlocking), "hipStreamCreate");
        const uint32_t cus = uint32_t(std::max(1, d->props.multiProcessorCount));
        // Resident 256-thread blocks per CU of the persistent grids, unless the options fix one count
        // for both: 4 for the fused / split traversal kernels (at most 5 waves per SIMD by their
        // registers; 4 measured best), 7 for the refill-driver kernels (72 VGPRs and 22 KB of LDS:
        // seven blocks fill the CU; 6 vs 4 measured +8%, 7 vs 6 +3% on the 1080p terrain job).
        const uint32_t per    = d->opt.persistent_blocks_per_cu ? d->opt.persistent_blocks_per_cu : 4u;
        const uint32_t per_rf = d->opt.persistent_blocks_per_cu ? d->opt.persistent_blocks_per_cu : 7u;
        const uint32_t per_q  = d->opt.persistent_blocks_per_cu ? d->opt.persistent_blocks_per_cu : 6u;  // k_query: six
        d->shape              = LaunchShape{cus, cus * per, cus * 8, cus * per_rf, cus * per_q};
        // Automatic pool: 2^28 paths (47 GB of path state, within the default 64 GiB budget) on a
        // device of at least 192 GiB, 2^29 (94.5 GB) from 256 GiB with one batch in flight, 2^26
        // (12 GB) from 64 GiB, else 2^23. Fewer, fuller batches pay each batch's latency-bound tail
        // of late bounces (a few ms, whatever the batch size) and each launch's drain less often.
        // Measured on the 1080p terrain job (530.8M samples): 2^23 -> 2^26 paths +30% (per-ray
        // traversal); with the refill traversal 2^26: 2,303, 2^27: 2,486, 2^28: 2,602 Msamples/s
        // (earlier kernels); 2^27: 2,735, 2^28: 2,917, 2^29 (the job in one batch): 2,982.
        // Latency-oriented callers (progressive serving) pass a smaller pool.
        // The automatic pools (one per batch in flight) also keep within half of the memory free now,
        // so that a second device handle in the process, or another process, still fits.
        if (d->opt.path_capacity == 0) {
            const uint64_t mem = d->props.totalGlobalMem;
            uint64_t       cap = mem >= (uint64_t(256) << 30) && d->opt.batches_in_flight != 2 ? uint64_t(1) << 29
                                 : mem >= (uint64_t(192) << 30)                                ? uint64_t(1) << 28
                                 : mem >= (uint64_t(64) << 30)                                 ? uint64_t(1) << 26
                                                                                               : uint64_t(1) << 23;
            size_t free_bytes = 0, total_bytes = 0;
            check_hip(hipMemGetInfo(&free_bytes, &total_bytes), "hipMemGetInfo");
            const uint64_t tracks = d->opt.batches_in_flight == 2 ? 2u : 1u;
            while (cap > (uint64_t(1) << 20) && cap * kPathSlotBytes * tracks > free_bytes / 2) cap >>= 1;
            d->opt.path_capacity = cap;
        }
        if (d->opt.path_capacity > (uint64_t(1) << 30)) throw std::invalid_argument("path capacity too large");
        if (d->opt.traversal > MRG_TRAVERSAL_REFILL) throw std::invalid_argument("unknown traversal kernel mode");
        if (d->opt.max_jobs_in_flight == 0) d->opt.max_jobs_in_flight = 64;
        if (d->opt.max_jobs_in_flight > 4096) throw std::invalid_argument("max_jobs_in_flight too large");
        if (d->opt.batches_in_flight > 2) throw std::invalid_argument("batches_in_flight must be 0 (automatic), 1 or 2");
        if (d->opt.query_precision == 0) d->opt.query_precision = 64;
        if (d->opt.query_precision != 32 && d->opt.query_precision != 64) throw std::invalid_argument("query_precision must be 0, 32 or 64");
        alloc_pool(d.get());
        // the allocation-free query's workspace: one static array per GPU, looked up with the device current
        check_hip(noalloc_scratch_address(&d->noalloc_scratch), "hipGetSymbolAddress(scratch)");
        upload(d->energy, conductor_energy_table());
        for (Track& t : d->tracks) t.pool.energy = d->energy.as<float>();
        d->query_pool.energy = d->energy.as<float>();
        for (auto& e : d->query_events) check_hip(hipEventCreate(&e), "hipEventCreate");
        d->resident_bytes = 0;
        mrg_device* raw   = d.release();
        raw->worker       = std::thread(scheduler, raw);
        *out              = raw;
    });
}

int mrg_device_close(mrg_device* d) {
    return guarded([&] {
        if (!d) return;
        {
            std::lock_guard<std::mutex> lock(d->mutex);
            if (d->shutdown) throw std::invalid_argument("device is already closed");
            for (mrg_job* j : d->jobs) {
                j->cancel_requested = true;
                signal_cancel_locked(d, j);
            }
            d->shutdown = true;
            d->cv.notify_all();
        }
        // the scheduler cancels every unfinished job (films retired to host memory) before it returns
        if (d->worker.joinable()) d->worker.join();
        (void)hipSetDevice(d->ordinal);
        {
            std::lock_guard<std::mutex> q(d->query_mutex);  // no ray query is using the streams
            (void)hipStreamSynchronize(d->stream);
            (void)hipStreamSynchronize(d->query_stream);
            for (Track& t : d->tracks) free_track(t);
            for (auto& e : d->query_events) (void)hipEventDestroy(e);
            (void)hipStreamDestroy(d->stream);
            (void)hipStreamDestroy(d->copy_stream);
            (void)hipStreamDestroy(d->query_stream);
            d->stream = d->copy_stream = d->query_stream = nullptr;
            for (DevBuf* b : {&d->query_work, &d->query_spill_ref, &d->query_spill_t, &d->energy}) b->release();
            d->closed = true;
        }
        bool last;
        {
            std::lock_guard<std::mutex> lock(d->mutex);
            d->released = true;
            last        = d->live == 0;
        }
        // with scenes or jobs still held by the caller, the last of their releases deletes the device
        if (last) delete d;
    });
}

int mrg_device_query(mrg_device* d, char* name, size_t cap, uint32_t* cus, uint64_t* hbm, uint64_t* l2) {
    return guarded([&] {
        if (!d) throw std::invalid_argument("device is null");
        if (name && cap) std::snprintf(name, cap, "%s (%s)", d->props.name, d->props.gcnArchName);
        if (cus) *cus = uint32_t(d->props.multiProcessorCount);
        if (hbm) *hbm = d->props.totalGlobalMem;
        if (l2) *l2 = uint64_t(d->props.l2CacheSize);
    });
}

int mrg_conductor_energy_table(float* out) {
    return guarded([&] {
        if (!out) throw std::invalid_argument("null argument");
        const std::vector<float>& t = conductor_energy_table();
        for (size_t i = 0; i < t.size() / 4; ++i)
            for (int c = 0; c < 3; ++c) out[i * 3 + c] = t[i * 4 + c];
    });
}

int mrg_scene_upload(mrg_device* d, const mrg_scene_desc* s, mrg_scene** out) {
    return guarded([&] {
        if (!d || !s || !out) throw std::invalid_argument("null argument");
        if (d->closed) throw std::invalid_argument("device is closed");
        Hasher h;
        const uint64_t n = s->triangle_count;
        h.bytes(&n, 8);
        if (n) {
            h.bytes(s->positions, n * 36);
            h.bytes(s->normals, n * 36);
            h.bytes(s->instance, n * 4);
            h.bytes(s->triangle, n * 4);
        }
        h.bytes(&s->instance_count, 4);
        if (s->instance_count) h.bytes(s->materials, s->instance_count * sizeof(mrg_material));
        h.bytes(s->environment, 12);
        h.bytes(&s->directional_count, 4);
        if (s->directional_count) h.bytes(s->directional, s->directional_count * sizeof(mrg_directional_light));
        const uint32_t has_uv = s->uvs ? 1u : 0u;
        h.bytes(&has_uv, 4);
        if (n && s->uvs) h.bytes(s->uvs, n * 24);
        h.bytes(&s->texture_count, 4);
        for (uint32_t i = 0; s->textures && i < s->texture_count; ++i) {
            const mrg_texture& t = s->textures[i];
            h.bytes(&t.width, 4);
            h.bytes(&t.height, 4);
            if (t.rgb) h.bytes(t.rgb, uint64_t(t.width) * t.height * 12);
        }
        h.bytes(&s->environment_texture, 4);
        h.bytes(&s->environment_strength, 4);
        h.bytes(&s->environment_rotation_degrees, 4);
        const std::string key = h.hex();
        {
            std::lock_guard<std::mutex> lock(d->mutex);
            auto it = d->scenes.find(key);
            if (it != d->scenes.end()) {
                it->second->refs++;
                *out = it->second;
                return;
            }
        }
        mrg_scene* sc = build_scene(d, *s, key);
        std::lock_guard<std::mutex> lock(d->mutex);
        auto it = d->scenes.find(key);
        if (it != d->scenes.end()) {  // raced with an identical upload
            delete sc;
            it->second->refs++;
            *out = it->second;
            return;
        }
        d->scenes[key] = sc;
        d->resident_bytes += sc->info.device_bytes;
        d->live++;
        *out = sc;
    });
}

int mrg_scene_release(mrg_scene* s) {
    return guarded([&] {
        if (!s) return;
        mrg_device* d    = s->device;
        bool        last = false;
        {
            std::lock_guard<std::mutex> lock(d->mutex);
            if (--s->refs == 0) {
                d->scenes.erase(s->hash);
                d->resident_bytes -= std::min(d->resident_bytes, s->info.device_bytes);
                delete s;
                d->live--;
                last = d->released && d->live == 0;
            }
        }
        if (last) delete d;
    });
}

int mrg_scene_info_get(mrg_scene* s, mrg_scene_info* out) {
    return guarded([&] {
        if (!s || !out) throw std::invalid_argument("null argument");
        std::lock_guard<std::mutex> lock(s->device->mutex);
        *out                 = s->info;
        out->reference_count = uint64_t(s->refs);
    });
}

static void check_interval(uint64_t n, float tmin, float tmax) {
    if (!(tmin >= 0) || !std::isfinite(tmin) || std::isnan(tmax) || tmax < tmin)
        throw std::invalid_argument("Ray interval must satisfy 0 <= minimum <= maximum");
    if (n >= (uint64_t(1) << 32) - 64) throw std::invalid_argument("too many rays for one call");
}

// One query launch on the device's query stream, the caller holding query_mutex. `work` (64 B) and
// `stats_dev` (16 B) are the launch's counters; they are zeroed here.
static void query_locked(mrg_scene* s, const void* d_rays, uint64_t n, float tmin, float tmax, void* d_hits, float* d_bary,
                         uint8_t* d_occ, bool any, uint32_t* work, uint64_t* stats_dev, mrg_trace_stats* stats, float* kernel_ms) {
    mrg_device* d = s->device;
    check_hip(hipMemsetAsync(stats_dev, 0, 16, d->query_stream), "hipMemsetAsync");
    check_hip(hipMemsetAsync(work, 0, 64, d->query_stream), "hipMemsetAsync");
    hipEvent_t e0 = d->query_events[0], e1 = d->query_events[1];
    check_hip(hipEventRecord(e0, d->query_stream), "hipEventRecord");
    check_hip(launch_query(s->gpu, static_cast<const float*>(d_rays), uint32_t(n), tmin, tmax, static_cast<mrg_hit_raw*>(d_hits), d_bary,
                           d_occ, work, d->query_pool, stats_dev, any, stats != nullptr, d->opt.query_precision == 64, d->shape,
                           d->query_stream),
              "k_query");
    check_hip(hipEventRecord(e1, d->query_stream), "hipEventRecord");
    uint64_t h[2] = {0, 0};
    if (stats) check_hip(hipMemcpyAsync(h, stats_dev, 16, hipMemcpyDeviceToHost, d->query_stream), "read stats");
    check_hip(hipStreamSynchronize(d->query_stream), "hipStreamSynchronize");
    if (kernel_ms) check_hip(hipEventElapsedTime(kernel_ms, e0, e1), "hipEventElapsedTime");
    if (stats) {
        stats->bounds_tests   = h[0];
        stats->triangle_tests = h[1];
    }
}

static void run_query(mrg_scene* s, const void* d_rays, uint64_t n, float tmin, float tmax, void* d_hits, float* d_bary,
                      uint8_t* d_occ, bool any, mrg_trace_stats* stats, float* kernel_ms) {
    check_interval(n, tmin, tmax);
    mrg_device* d = s->device;
    std::lock_guard<std::mutex> lock(d->query_mutex);  // one query at a time per device (shared spill area)
    if (d->closed) throw std::invalid_argument("device is closed");
    check_hip(hipSetDevice(d->ordinal), "hipSetDevice");
    DevBuf st(16);
    query_locked(s, d_rays, n, tmin, tmax, d_hits, d_bary, d_occ, any, d->query_work.as<uint32_t>(), st.as<uint64_t>(), stats,
                 kernel_ms);
}

int mrg_intersect(mrg_scene* s, const float* rays, uint64_t n, float tmin, float tmax, mrg_hit* hits, float* bary, mrg_trace_stats* stats) {
    return guarded([&] {
        if (!s || (n && (!rays || !hits))) throw std::invalid_argument("null argument");
        if (n == 0) return;
        if (s->device->closed) throw std::invalid_argument("device is closed");
        check_hip(hipSetDevice(s->device->ordinal), "hipSetDevice");
        DevBuf dr(n * 24), dh(n * 8), db(bary ? n * 12 : 0);
        check_hip(hipMemcpy(dr.ptr, rays, n * 24, hipMemcpyHostToDevice), "hipMemcpy");
        run_query(s, dr.ptr, n, tmin, tmax, dh.ptr, bary ? db.as<float>() : nullptr, nullptr, false, stats, nullptr);
        check_hip(hipMemcpy(hits, dh.ptr, n * 8, hipMemcpyDeviceToHost), "hipMemcpy");
        if (bary) check_hip(hipMemcpy(bary, db.ptr, n * 12, hipMemcpyDeviceToHost), "hipMemcpy");
    });
}

int mrg_occluded(mrg_scene* s, const float* rays, uint64_t n, float tmin, float tmax, uint8_t* occ, mrg_trace_stats* stats) {
    return guarded([&] {
        if (!s || (n && (!rays || !occ))) throw std::invalid_argument("null argument");
        if (n == 0) return;
        if (s->device->closed) throw std::invalid_argument("device is closed");
        check_hip(hipSetDevice(s->device->ordinal), "hipSetDevice");
        DevBuf dr(n * 24), dout(n);
        check_hip(hipMemcpy(dr.ptr, rays, n * 24, hipMemcpyHostToDevice), "hipMemcpy");
        run_query(s, dr.ptr, n, tmin, tmax, nullptr, nullptr, dout.as<uint8_t>(), true, stats, nullptr);
        check_hip(hipMemcpy(occ, dout.ptr, n, hipMemcpyDeviceToHost), "hipMemcpy");
    });
}

int mrg_intersect_device(mrg_scene* s, const void* rays, uint64_t n, float tmin, float tmax, void* hits, mrg_trace_stats* stats,
                         float* kernel_ms) {
    return guarded([&] {
        if (!s || (n && (!rays || !hits))) throw std::invalid_argument("null argument");
        if (n == 0) return;
        run_query(s, rays, n, tmin, tmax, hits, nullptr, nullptr, false, stats, kernel_ms);
    });
}

int mrg_intersect_noalloc(mrg_scene* s, const float* rays, uint64_t n, float tmin, float tmax, mrg_hit* hits, float* bary,
                          mrg_trace_stats* stats) {
    return guarded([&] {
        if (!s || (n && (!rays || !hits))) throw std::invalid_argument("null argument");
        if (n > kNoAllocMaxRays)
            throw std::invalid_argument("mrg_intersect_noalloc takes at most 1048576 rays per call (its workspace is fixed)");
        check_interval(n, tmin, tmax);
        if (n == 0) return;
        mrg_device* d = s->device;
        // the workspace is this GPU's static array: query_mutex serialises the calls on one device handle
        std::lock_guard<std::mutex> lock(d->query_mutex);
        if (d->closed) throw std::invalid_argument("device is closed");
        if (!d->noalloc_scratch) throw Error(MRG_DEVICE_ERROR, "the allocation-free query's workspace is not available");
        check_hip(hipSetDevice(d->ordinal), "hipSetDevice");
        char* base = static_cast<char*>(d->noalloc_scratch);
        check_hip(hipMemcpyAsync(base + kNoAllocRaysOffset, rays, n * 24, hipMemcpyHostToDevice, d->query_stream), "stage rays");
        query_locked(s, base + kNoAllocRaysOffset, n, tmin, tmax, base + kNoAllocHitsOffset,
                     bary ? reinterpret_cast<float*>(base + kNoAllocBaryOffset) : nullptr, nullptr, false,
                     reinterpret_cast<uint32_t*>(base + kNoAllocWorkOffset), reinterpret_cast<uint64_t*>(base + kNoAllocStatsOffset),
                     stats, nullptr);
        check_hip(hipMemcpyAsync(hits, base + kNoAllocHitsOffset, n * 8, hipMemcpyDeviceToHost, d->query_stream), "read hits");
        if (bary)
            check_hip(hipMemcpyAsync(bary, base + kNoAllocBaryOffset, n * 12, hipMemcpyDeviceToHost, d->query_stream), "read barycentrics");
        check_hip(hipStreamSynchronize(d->query_stream), "hipStreamSynchronize");
    });
}

int mrg_job_submit(mrg_device* d, mrg_scene* s, const mrg_job_desc* desc, mrg_job** out) {
    return guarded([&] {
        if (!d || !s || !desc || !out) throw std::invalid_argument("null argument");
        if (s->device != d) throw std::invalid_argument("scene belongs to another device");
        validate_job(*desc);
        auto j        = std::make_unique<mrg_job>();
        j->device     = d;
        j->scene      = s;
        j->desc       = *desc;
        j->camera     = make_camera(desc->camera);
        j->film_bytes = film_bytes_of(*desc);
        j->submitted  = Clock::now();
        j->refs       = 2;  // caller + scheduler queue
        std::lock_guard<std::mutex> lock(d->mutex);
        if (d->shutdown) throw std::invalid_argument("device is closed");
        if (j->film_bytes + d->pool_bytes + s->info.device_bytes > budget(d))
            throw Error(MRG_NOT_ADMITTED, "job exceeds the device HBM budget");
        s->refs++;
        d->live++;
        j->id = d->next_job_id++;
        d->jobs.push_back(j.get());
        *out = j.release();
        d->cv.notify_all();
    });
}

int mrg_job_pause(mrg_job* j, int* accepted) {
    return guarded([&] {
        if (!j) throw std::invalid_argument("job is null");
        std::lock_guard<std::mutex> lock(j->device->mutex);
        const bool ok = j->state == MRG_STATE_RUNNING;
        if (ok) {
            j->pause_requested = true;
            j->state           = j->in_flight ? MRG_STATE_PAUSING : MRG_STATE_PAUSED;
            j->device->cv.notify_all();
        }
        if (accepted) *accepted = ok;
    });
}

int mrg_job_resume(mrg_job* j, int* accepted) {
    return guarded([&] {
        if (!j) throw std::invalid_argument("job is null");
        std::lock_guard<std::mutex> lock(j->device->mutex);
        const bool ok = j->state == MRG_STATE_PAUSED || j->state == MRG_STATE_PAUSING;
        if (ok) {
            j->pause_requested = false;
            j->state           = MRG_STATE_RUNNING;
            j->device->cv.notify_all();
        }
        if (accepted) *accepted = ok;
    });
}

int mrg_job_cancel(mrg_job* j) {
    return guarded([&] {
        if (!j) throw std::invalid_argument("job is null");
        std::lock_guard<std::mutex> lock(j->device->mutex);
        if (terminal(j->state)) return;
        j->cancel_requested = true;
        j->pause_requested  = false;
        j->state            = MRG_STATE_CANCELLING;
        signal_cancel_locked(j->device, j);
        j->device->cv.notify_all();
    });
}

int mrg_job_wait(mrg_job* j, int timeout_ms) {
    int rc = MRG_OK;
    const int g = guarded([&] {
        if (!j) throw std::invalid_argument("job is null");
        std::unique_lock<std::mutex> lock(j->device->mutex);
        auto done = [&] { return terminal(j->state); };
        if (timeout_ms < 0) {
            j->device->cv.wait(lock, done);
        } else if (!j->device->cv.wait_for(lock, std::chrono::milliseconds(timeout_ms), done)) {
            rc = MRG_TIMEOUT;
        }
    });
    return g != MRG_OK ? g : rc;
}

int mrg_job_progress(mrg_job* j, mrg_progress* p) {
    return guarded([&] {
        if (!j || !p) throw std::invalid_argument("null argument");
        std::lock_guard<std::mutex> lock(j->device->mutex);
        std::memset(p, 0, sizeof *p);
        p->state            = j->state;
        p->status           = j->status;
        p->job_id           = j->id;
        p->recorded_samples = j->recorded;
        p->total_samples    = uint64_t(j->desc.width) * j->desc.height * j->desc.samples_per_pixel;
        p->completed_pixels = j->completed_pixels;
        const auto end      = j->finished ? j->finished_at : Clock::now();
        p->elapsed_seconds  = std::chrono::duration<double>(end - j->submitted).count();
        p->queued_seconds   = j->admitted_at.time_since_epoch().count() ? std::chrono::duration<double>(j->admitted_at - j->submitted).count() : 0;
        p->first_pass_seconds = j->first_pass_done ? std::chrono::duration<double>(j->first_pass_at - j->submitted).count() : 0;
        p->completion_seconds = j->finished ? std::chrono::duration<double>(j->finished_at - j->submitted).count() : 0;
        p->paths                 = j->paths;
        p->extension_rays        = j->ext_rays;
        p->shadow_rays           = j->shadow_rays;
        p->scattering_events     = j->scatter;
        p->roulette_terminations = j->rr_term;
        std::snprintf(p->error, sizeof p->error, "%s", j->error.c_str());
    });
}

// A film read waits until the job's batches in flight have landed (no new batch of the job starts
// meanwhile), so it never sees a partly committed batch.
static void read_film_raw(mrg_job* j, std::vector<double>& sums, std::vector<uint32_t>& counts) {
    const uint64_t px = uint64_t(j->desc.width) * j->desc.height;
    sums.assign(px * 3, 0.0);
    counts.assign(px, 0);
    mrg_device* d = j->device;
    {
        std::unique_lock<std::mutex> lock(d->mutex);
        j->readers++;
        d->cv.wait(lock, [&] { return j->in_flight == 0; });
    }
    struct Done {
        mrg_device* d;
        mrg_job*    j;
        ~Done() {
            {
                std::lock_guard<std::mutex> lock(d->mutex);
                j->readers--;
            }
            d->cv.notify_all();
        }
    } done{d, j};
    std::lock_guard<std::mutex> fl(j->film_mutex);
    if (j->film_on_host) {  // finished: the film was copied out of HBM
        sums   = j->host_sums;
        counts = j->host_counts;
        return;
    }
    if (!j->film_sums.ptr) return;  // not admitted yet
    check_hip(hipSetDevice(j->device->ordinal), "hipSetDevice");
    check_hip(hipMemcpyAsync(sums.data(), j->film_sums.ptr, px * 24, hipMemcpyDeviceToHost, j->device->copy_stream), "read film");
    check_hip(hipMemcpyAsync(counts.data(), j->film_counts.ptr, px * 4, hipMemcpyDeviceToHost, j->device->copy_stream), "read film");
    check_hip(hipStreamSynchronize(j->device->copy_stream), "hipStreamSynchronize");
}

int mrg_job_read_film(mrg_job* j, double* sums, uint32_t* counts) {
    return guarded([&] {
        if (!j) throw std::invalid_argument("job is null");
        std::vector<double>   s;
        std::vector<uint32_t> c;
        read_film_raw(j, s, c);
        if (sums) std::memcpy(sums, s.data(), s.size() * 8);
        if (counts) std::memcpy(counts, c.data(), c.size() * 4);
    });
}

int mrg_job_read_image(mrg_job* j, float* rgb) {
    return guarded([&] {
        if (!j || !rgb) throw std::invalid_argument("null argument");
        std::vector<double>   s;
        std::vector<uint32_t> c;
        read_film_raw(j, s, c);
        for (size_t p = 0; p < c.size(); ++p)
            for (int k = 0; k < 3; ++k) rgb[p * 3 + k] = c[p] ? float(s[p * 3 + k] / double(c[p])) : 0.0f;
    });
}

int mrg_job_preview(mrg_job* j, uint32_t max_dim, float* rgb, uint32_t* width, uint32_t* height, uint64_t* recorded) {
    return guarded([&] {
        if (!j || !width || !height) throw std::invalid_argument("null argument");
        const uint32_t W = j->desc.width, H = j->desc.height;
        const double scale = max_dim ? std::max(1.0, double(std::max(W, H)) / double(max_dim)) : 1.0;
        const uint32_t w = std::max<uint32_t>(1, uint32_t(double(W) / scale));
        const uint32_t h = std::max<uint32_t>(1, uint32_t(double(H) / scale));
        *width  = w;
        *height = h;
        if (!rgb) return;
        uint64_t rec;
        {
            std::lock_guard<std::mutex> lock(j->device->mutex);
            rec = j->recorded;
        }
        std::vector<double>   s;
        std::vector<uint32_t> c;
        read_film_raw(j, s, c);
        for (uint32_t y = 0; y < h; ++y) {
            const uint32_t sy = std::min<uint32_t>(H - 1, uint32_t((uint64_t(y) * H + H / 2) / h));
            for (uint32_t x = 0; x < w; ++x) {
                const uint32_t sx = std::min<uint32_t>(W - 1, uint32_t((uint64_t(x) * W + W / 2) / w));
                const uint64_t p  = uint64_t(sy) * W + sx;
                for (int k = 0; k < 3; ++k) rgb[(uint64_t(y) * w + x) * 3 + k] = c[p] ? float(s[p * 3 + k] / double(c[p])) : 0.0f;
            }
        }
        if (recorded) *recorded = rec;
    });
}

int mrg_job_release(mrg_job* j) {
    return guarded([&] {
        if (!j) return;
        mrg_device* d    = j->device;
        bool        last = false;
        {
            std::lock_guard<std::mutex> lock(d->mutex);
            if (!terminal(j->state)) {
                j->cancel_requested = true;
                j->state            = MRG_STATE_CANCELLING;
                signal_cancel_locked(d, j);
                d->cv.notify_all();
            }
            last = release_job_locked(j);
        }
        if (last) delete d;
    });
}

int mrg_render(mrg_device* d, mrg_scene* s, const mrg_job_desc* desc, float* rgb, mrg_progress* progress) {
    mrg_job* j = nullptr;
    int rc = mrg_job_submit(d, s, desc, &j);
    if (rc) return rc;
    rc = mrg_job_wait(j, -1);
    mrg_progress p{};
    if (!rc) rc = mrg_job_progress(j, &p);
    if (!rc && p.state != MRG_STATE_COMPLETED) {
        g_error = p.error[0] ? p.error : "render did not complete";
        rc      = p.status ? p.status : MRG_DEVICE_ERROR;
    }
    if (!rc && rgb) rc = mrg_job_read_image(j, rgb);
    if (progress) *progress = p;
    const std::string keep = g_error;
    mrg_job_release(j);
    g_error = keep;
    return rc;
}

int mrg_bvh_build_host(const float* positions, uint64_t n, uint64_t* out) {
    return guarded([&] {
        if (!out || (n && !positions)) throw std::invalid_argument("null argument");
        const BvhBuildResult b = build_bvh(positions, n, 0);
        // structural check: every primitive exactly once, leaves <= kMaxLeaf, and every quantized child
        // box (as the device decodes it) contains the exact child box
        std::vector<uint32_t> seen(n, 0);
        uint64_t max_leaf = 0, leaf_prims = 0, bad = 0;
        if (b.qnodes.size() != b.nodes.size()) ++bad;
        for (size_t i = 0; i < b.nodes.size() && i < b.qnodes.size(); ++i) {
            const GpuNode4& g = b.nodes[i];
            const GpuNodeQ& q = b.qnodes[i];
            for (int k = 0; k < 4; ++k) {
                const uint32_t r = g.ref[k];
                if (q.ref[k] != r) ++bad;
                if (r == kEmptyRef) continue;
                float lo[3], hi[3];
                decode_child(q, k, lo, hi);
                if (!(lo[0] <= g.lox[k] && lo[1] <= g.loy[k] && lo[2] <= g.loz[k] && hi[0] >= g.hix[k] && hi[1] >= g.hiy[k] &&
                      hi[2] >= g.hiz[k]))
                    ++bad;
                if (is_leaf(r)) {
                    const uint32_t f = leaf_first(r), c = leaf_count(r);
                    max_leaf = std::max<uint64_t>(max_leaf, c);
                    for (uint32_t m = 0; m < c; ++m) {
                        if (f + m >= n) { ++bad; continue; }
                        seen[b.order[f + m]]++;
                        ++leaf_prims;
                    }
                } else if (r >= b.nodes.size()) {
                    ++bad;
                }
            }
        }
        for (uint64_t i = 0; i < n; ++i) bad += seen[i] != 1;
        out[0] = b.nodes.size();
        out[1] = b.leaves;
        out[2] = b.max_depth;
        out[3] = max_leaf;
        out[4] = leaf_prims;
        out[5] = bad + (b.max_stack > kStackTotal ? 1 : 0);
        out[6] = b.max_stack;
    });
}

int mrg_device_memory(mrg_device* d, uint64_t* budget_bytes, uint64_t* resident_bytes, uint64_t* pool_bytes) {
    return guarded([&] {
        if (!d) throw std::invalid_argument("device is null");
        std::lock_guard<std::mutex> lock(d->mutex);
        if (budget_bytes) *budget_bytes = budget(d);
        if (resident_bytes) *resident_bytes = d->resident_bytes;
        if (pool_bytes) *pool_bytes = d->pool_bytes;
    });
}

int mrg_device_kernel_times(mrg_device* d, mrg_kernel_times* out, int reset) {
    return guarded([&] {
        if (!d || !out) throw std::invalid_argument("null argument");
        std::lock_guard<std::mutex> lock(d->mutex);
        *out = d->times;
        if (reset) d->times = mrg_kernel_times{};
    });
}

}  // extern "C"
