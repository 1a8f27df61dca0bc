// C++ face of libmrg for the reference's host code: the RenderSession contract of
// include/minirender/render/render_session.hpp:82-115 and the SceneBvh query interface of
// include/minirender/render/bvh.hpp:45-108, with the reference's state names and exception types,
// on top of the C ABI in mrg.h. Header-only; link libmrg.so.
//
//   minirender::gpu::Device  device(0);
//   auto scene = std::make_shared<minirender::gpu::SceneBvh>(device, desc);   // mrg_scene_desc
//   minirender::gpu::RenderSession session(device);
//   session.start(scene, camera, {1920, 1080, 256, 0}, {MRG_MODE_MIS, 16, true, 3});
//   session.wait();
//   std::shared_ptr<const minirender::gpu::Image> image = session.result();
#pragma once

#include <chrono>
#include <cmath>
#include <cstdint>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "mrg.h"

namespace minirender::gpu {

// RenderState (render_session.hpp:12-23), same order as the MRG_STATE_* codes.
enum class RenderState { idle, preparing, running, pausing, paused, finalizing, cancelling, cancelled, completed, failed };

inline std::string render_state_name(RenderState state) {
    static const char* names[] = {"idle",       "preparing",  "running",   "pausing",   "paused",
                                  "finalizing", "cancelling", "cancelled", "completed", "failed"};
    return names[static_cast<int>(state)];
}

// render_session.cpp:161-164
inline bool render_terminal(RenderState state) noexcept {
    return state == RenderState::idle || state == RenderState::cancelled || state == RenderState::completed ||
           state == RenderState::failed;
}

// OperationCancelled (cancellation.hpp:13-18)
class OperationCancelled : public std::runtime_error {
  public:
    OperationCancelled() : std::runtime_error("Operation cancelled") {}
};

// Status codes back to the reference's exceptions.
inline void check(int status) {
    switch (status) {
    case MRG_OK:
        return;
    case MRG_INVALID_ARGUMENT:
        throw std::invalid_argument(mrg_last_error());
    case MRG_CANCELLED:
        throw OperationCancelled();
    case MRG_NONFINITE: {
        const std::string message = mrg_last_error();
        if (message.find("float image range") != std::string::npos) throw std::overflow_error(message);
        throw std::runtime_error(message);
    }
    default:
        throw std::runtime_error(mrg_last_error());
    }
}

// One GPU with its own scheduler thread inside libmrg.
class Device {
  public:
    explicit Device(int ordinal = 0, const mrg_device_options* options = nullptr) {
        check(mrg_device_open(ordinal, options, &device_));
    }
    ~Device() {
        if (device_) mrg_device_close(device_);
    }
    Device(const Device&)            = delete;
    Device& operator=(const Device&) = delete;
    mrg_device* get() const { return device_; }

  private:
    mrg_device* device_ = nullptr;
};

// A scene resident on the device (deduplicated by content), with the SceneBvh queries.
class SceneBvh {
  public:
    SceneBvh(Device& device, const mrg_scene_desc& desc) { check(mrg_scene_upload(device.get(), &desc, &scene_)); }
    ~SceneBvh() {
        if (scene_) mrg_scene_release(scene_);
    }
    SceneBvh(const SceneBvh&)            = delete;
    SceneBvh& operator=(const SceneBvh&) = delete;
    mrg_scene* get() const { return scene_; }

    // Closest hits of rays [n][6] (origin, unit direction) in [tmin, tmax] (bvh.hpp:53-56).
    std::vector<mrg_hit> intersect(const std::vector<float>& rays, float tmin = 0.0f, float tmax = INFINITY) const {
        std::vector<mrg_hit> hits(rays.size() / 6);
        check(mrg_intersect(scene_, rays.data(), hits.size(), tmin, tmax, hits.data(), nullptr, nullptr));
        return hits;
    }
    // Any-hit flags (bvh.hpp:57-59).
    std::vector<uint8_t> occluded(const std::vector<float>& rays, float tmin = 0.0f, float tmax = INFINITY) const {
        std::vector<uint8_t> flags(rays.size() / 6);
        check(mrg_occluded(scene_, rays.data(), flags.size(), tmin, tmax, flags.data(), nullptr));
        return flags;
    }

  private:
    mrg_scene* scene_ = nullptr;
};

// RenderSettings / IntegratorSettings (path_tracer.hpp:21-50); mode is MRG_MODE_*.
struct RenderSettings {
    uint32_t width             = 256;
    uint32_t height            = 256;
    uint32_t samples_per_pixel = 64;
    uint64_t seed              = 1;
};

struct IntegratorSettings {
    uint32_t mode             = MRG_MODE_MIS;
    uint32_t maximum_depth    = 8;
    bool     russian_roulette = true;
    uint32_t roulette_depth   = 3;
};

// SessionSettings (render_session.hpp:30-38): on the GPU the batch is a progressive pass of
// `batch_samples` samples per pixel (0 = 16); threads and tiles have no GPU meaning.
struct SessionSettings {
    uint32_t batch_samples = 0;
};

struct PathStatistics {
    uint64_t paths = 0, extension_rays = 0, shadow_rays = 0, scattering_events = 0, roulette_terminations = 0;
};

// RenderProgress (render_session.hpp:41-65) plus the GPU's latency figures.
struct RenderProgress {
    RenderState    state            = RenderState::idle;
    uint64_t       job_id           = 0;
    uint64_t       recorded_samples = 0;
    uint64_t       total_samples    = 0;
    uint64_t       completed_pixels = 0;
    double         elapsed_seconds  = 0;
    std::string    error;
    PathStatistics statistics;
    double         queued_seconds = 0, first_pass_seconds = 0, completion_seconds = 0;
};

// Linear float RGB image, row-major [height][width][3].
struct Image {
    uint32_t           width = 0, height = 0;
    std::vector<float> rgb;
};

// RenderPreview (render_session.hpp:68-77).
struct RenderPreview {
    std::shared_ptr<const Image> image;
    uint64_t                     job_id           = 0;
    uint64_t                     revision         = 0;
    uint64_t                     recorded_samples = 0;
};

// The asynchronous session. Lifecycle calls come from one control thread; progress() and preview()
// may be called concurrently (render_session.hpp:79-81).
class RenderSession {
  public:
    explicit RenderSession(Device& device) : device_(device) {}
    // Cancels and waits for the job (render_session.cpp:511-514); never throws.
    ~RenderSession() {
        try {
            cancel();
            wait();
        } catch (...) {
        }
        release_job();
    }
    RenderSession(const RenderSession&)            = delete;
    RenderSession& operator=(const RenderSession&) = delete;

    // render_session.cpp:516-565: validates, refuses while a job is active, then runs the job.
    void start(std::shared_ptr<const SceneBvh> scene, const mrg_camera& camera, RenderSettings render,
               IntegratorSettings integrator = {}, SessionSettings session = {}) {
        if (!scene) throw std::invalid_argument("Render session requires a scene snapshot");
        if (!render_terminal(progress().state)) throw std::logic_error("Render session is still active");
        mrg_job_desc d{};
        d.camera            = camera;
        d.width             = render.width;
        d.height            = render.height;
        d.samples_per_pixel = render.samples_per_pixel;
        d.seed              = render.seed;
        d.mode              = integrator.mode;
        d.maximum_depth     = integrator.maximum_depth;
        d.russian_roulette  = integrator.russian_roulette ? 1u : 0u;
        d.roulette_depth    = integrator.roulette_depth;
        d.pass_samples      = session.batch_samples;
        mrg_job* job        = nullptr;
        check(mrg_job_submit(device_.get(), scene->get(), &d, &job));
        std::lock_guard<std::mutex> lock(mutex_);
        release_job_locked();
        job_    = job;
        scene_  = std::move(scene);
        width_  = render.width;
        height_ = render.height;
        result_.reset();
    }

    bool pause() { return toggle(&mrg_job_pause); }
    bool resume() { return toggle(&mrg_job_resume); }

    void cancel() {
        std::lock_guard<std::mutex> lock(mutex_);
        if (job_) check(mrg_job_cancel(job_));
    }

    void wait() {
        mrg_job* job = current();
        if (job) check(mrg_job_wait(job, -1));
    }

    bool wait_for(std::chrono::milliseconds timeout) const {
        mrg_job* job = current();
        if (!job) return true;
        const int rc = mrg_job_wait(job, static_cast<int>(timeout.count()));
        if (rc == MRG_TIMEOUT) return false;
        check(rc);
        return true;
    }

    RenderProgress progress() const {
        mrg_job* job = current();
        RenderProgress out;
        if (!job) return out;
        mrg_progress p{};
        check(mrg_job_progress(job, &p));
        out.state                            = static_cast<RenderState>(p.state);
        out.job_id                           = p.job_id;
        out.recorded_samples                 = p.recorded_samples;
        out.total_samples                    = p.total_samples;
        out.completed_pixels                 = p.completed_pixels;
        out.elapsed_seconds                  = p.elapsed_seconds;
        out.error                            = p.error;
        out.statistics.paths                 = p.paths;
        out.statistics.extension_rays        = p.extension_rays;
        out.statistics.shadow_rays           = p.shadow_rays;
        out.statistics.scattering_events     = p.scattering_events;
        out.statistics.roulette_terminations = p.roulette_terminations;
        out.queued_seconds                   = p.queued_seconds;
        out.first_pass_seconds               = p.first_pass_seconds;
        out.completion_seconds               = p.completion_seconds;
        return out;
    }

    // Nearest-neighbour preview, longest side <= maximum_dimension (0 = full resolution).
    RenderPreview preview(size_t maximum_dimension = 1024) const {
        RenderPreview out;
        mrg_job* job = current();
        if (!job) return out;
        uint32_t w = 0, h = 0;
        uint64_t recorded = 0;
        check(mrg_job_preview(job, uint32_t(maximum_dimension), nullptr, &w, &h, nullptr));
        auto image    = std::make_shared<Image>();
        image->width  = w;
        image->height = h;
        image->rgb.resize(size_t(w) * h * 3);
        check(mrg_job_preview(job, uint32_t(maximum_dimension), image->rgb.data(), &w, &h, &recorded));
        out.image            = std::move(image);
        out.job_id           = progress().job_id;
        out.revision         = recorded;
        out.recorded_samples = recorded;
        return out;
    }

    // The full image, only once the job has completed (render_session.cpp:653-656).
    std::shared_ptr<const Image> result() const {
        if (progress().state != RenderState::completed) return nullptr;
        std::lock_guard<std::mutex> lock(mutex_);
        if (!result_) {
            auto image    = std::make_shared<Image>();
            image->width  = width_;
            image->height = height_;
            image->rgb.resize(size_t(width_) * height_ * 3);
            check(mrg_job_read_image(job_, image->rgb.data()));
            result_ = std::move(image);
        }
        return result_;
    }

  private:
    mrg_job* current() const {
        std::lock_guard<std::mutex> lock(mutex_);
        return job_;
    }
    bool toggle(int (*fn)(mrg_job*, int*)) {
        std::lock_guard<std::mutex> lock(mutex_);
        if (!job_) return false;
        int accepted = 0;
        check(fn(job_, &accepted));
        return accepted != 0;
    }
    void release_job_locked() {
        if (job_) mrg_job_release(job_);
        job_ = nullptr;
    }
    void release_job() {
        std::lock_guard<std::mutex> lock(mutex_);
        release_job_locked();
    }

    Device&                              device_;
    mutable std::mutex                   mutex_;
    mrg_job*                             job_ = nullptr;
    std::shared_ptr<const SceneBvh>      scene_;
    uint32_t                             width_ = 0, height_ = 0;
    mutable std::shared_ptr<const Image> result_;
};

}  // namespace minirender::gpu
