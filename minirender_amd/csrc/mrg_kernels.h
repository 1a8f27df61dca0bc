// Kernel-facing declarations shared by mrg_kernels.hip and the host library.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mrg_layout.h"

namespace mrg {

struct mrg_hit_raw {
    float   t;
    int32_t primitive;
};

// Workspace of the allocation-free ray query (mrg_intersect_noalloc): one file-scope device array
// per GPU, laid out as the query's work counters, its statistics counters, then the staged rays, hits
// and barycentrics of at most kNoAllocMaxRays rays.
constexpr uint64_t kNoAllocMaxRays       = uint64_t(1) << 20;
constexpr uint64_t kNoAllocWorkOffset    = 0;    // 64 B of work-claim counters (zeroed per call)
constexpr uint64_t kNoAllocStatsOffset   = 128;  // 2 x u64 traversal counters (zeroed per call)
constexpr uint64_t kNoAllocRaysOffset    = 256;
constexpr uint64_t kNoAllocHitsOffset    = kNoAllocRaysOffset + kNoAllocMaxRays * 24;
constexpr uint64_t kNoAllocBaryOffset    = kNoAllocHitsOffset + kNoAllocMaxRays * 8;
constexpr uint64_t kNoAllocScratchBytes  = kNoAllocBaryOffset + kNoAllocMaxRays * 12;  // 44 MiB + 256 B
// Device address of the workspace on the current device (hipGetSymbolAddress of the array).
hipError_t noalloc_scratch_address(void** out);

// Path-state flags kept in prev.w: bit 0 = previous event was discrete (delta); bits 16.. = job slot.
constexpr uint32_t kPrevDelta = 1u;
constexpr uint32_t kMaxBounces = 130;  // depth <= 128 (IntegratorSettings::validate) plus one

// Counter words (uint32): extension-queue counts per bounce, shadow counts per bounce, and the
// persistent-kernel work counters.
constexpr uint32_t kCtrExt        = 0;
constexpr uint32_t kCtrShadow     = kCtrExt + kMaxBounces + 1;
constexpr uint32_t kCtrWorkExt    = kCtrShadow + kMaxBounces + 1;
constexpr uint32_t kCtrWorkShadow = kCtrWorkExt + kMaxBounces + 1;
constexpr uint32_t kCounterWords  = kCtrWorkShadow + kMaxBounces + 1;

// Batch header words (device memory, written by the host before each batch) so a captured graph
// needs no per-batch kernel arguments.
constexpr uint32_t kHdrChunks = 0;
constexpr uint32_t kHdrSlots  = 1;
constexpr uint32_t kHdrPixels = 2;
constexpr uint32_t kHdrOneJob = 4;  // 1 when every chunk of the batch belongs to one job: bounce 0 has no path records of
                                   // `prev` (the job is 0, the flags are a camera path's), k_generate writes none
constexpr uint32_t kHdrUnits  = 5;  // camera-ray work units of the batch (GpuChunk::unit_prefix)
constexpr uint32_t kHdrWords  = 8;

// Batch statistics (uint64).
constexpr uint32_t kStatShadowRays = 0;
constexpr uint32_t kStatScatter    = 1;
constexpr uint32_t kStatRoulette   = 2;
constexpr uint32_t kStatWords      = 8;

// Path pool: structure of arrays over `capacity` path slots.
struct Pool {
    float4*   ray_o;     // origin, tmin
    float4*   ray_d;     // direction, tmax
    float4*   hit;       // closest hit: edge functions e0, e1, e2 and their sum (barycentric i = e_i / sum)
    uint32_t* hit_slot;  // leaf-order slot or 0xffffffff
    float4*   beta;      // throughput, previous pdf
    float4*   rad;       // radiance, eta scale
    float4*   prev;      // previous reference point, flags
    uint4*    rng;       // PCG state, increment
    uint32_t* q_ext0;
    uint32_t* q_ext1;
    uint32_t* q_sh;
    float4*   sh_o;      // shadow origin, limit
    float4*   sh_d;
    float4*   sh_c;      // contribution
    uint32_t* counters;
    uint64_t* stats;
    const uint32_t* header;  // kHdr* words of the current batch
    uint32_t* spill_ref;
    float*    spill_t;
    const float* energy;  // conductor energy table [64][64][4] of the device (mrg_energy.h)
    uint32_t  capacity;
    uint32_t  spill_lanes;
};

struct LaunchShape {
    uint32_t compute_units;
    uint32_t traversal_blocks;  // persistent grid of k_trace, k_extend and k_shadow (256-thread blocks)
    uint32_t shade_blocks;
    uint32_t refill_blocks;     // persistent grid of the refill-driver bounce kernel (k_trace_refill)
    uint32_t query_blocks;      // persistent grid of the ray-query kernel (k_query)
};

// exact: hit decisions in FP64 (the reference intersector's arithmetic); otherwise the render kernels' FP32 test.
hipError_t launch_query(const GpuScene& sc, const float* rays, uint32_t n, float tmin, float tmax, mrg_hit_raw* hits, float* bary,
                        uint8_t* occluded, uint32_t* work, const Pool& P, uint64_t* stats, bool any, bool with_stats, bool exact,
                        const LaunchShape& shape, hipStream_t stream);
hipError_t launch_reset(const Pool& P, hipStream_t stream);
hipError_t launch_generate(const Pool& P, const GpuJob* jobs, const GpuChunk* chunks, const LaunchShape& shape, hipStream_t stream);
hipError_t launch_extend(const Pool& P, const GpuJob* jobs, uint32_t iter, const LaunchShape& shape, hipStream_t stream);
// extended: the shading kernel with base-color textures, environment maps and energy-compensated
// conductors (the plain one covers scenes that use none of them).
// one_scene: every job of the batch renders jobs[0]'s scene.
hipError_t launch_shade(const Pool& P, const GpuJob* jobs, uint32_t iter, bool extended, bool one_scene, const LaunchShape& shape,
                        hipStream_t stream);
hipError_t launch_shadow(const Pool& P, const GpuJob* jobs, uint32_t iter, const LaunchShape& shape, hipStream_t stream);
// Fused traversal: shadow rays of bounce iter - 1 and extension rays of bounce iter in one launch.
hipError_t launch_trace(const Pool& P, const GpuJob* jobs, uint32_t iter, bool one_scene, const LaunchShape& shape,
                        hipStream_t stream);
// The same with the refill driver (finished lanes take the next queue entry at once).
// one_scene: every job of the batch renders jobs[0]'s scene (the kernel reads it once per wave).
hipError_t launch_trace_refill(const Pool& P, const GpuJob* jobs, uint32_t iter, bool one_scene, const LaunchShape& shape,
                               hipStream_t stream);
hipError_t launch_validate(const Pool& P, const GpuJob* jobs, const GpuChunk* chunks, const LaunchShape& shape, hipStream_t stream);
// Cancellation of a job while its batch renders: the host raises word 1 of the job's flags (k_signal);
// shading ends the job's paths at their next bounce, and before the film commit the batch freezes the
// word into word 2 (k_commit_gate) so that validation and accumulation either commit all of the job's
// samples of the batch or none.
hipError_t launch_signal(uint32_t* word, uint32_t value, hipStream_t stream);
hipError_t launch_commit_gate(const GpuJob* jobs, uint32_t job_count, hipStream_t stream);
hipError_t launch_accumulate(const Pool& P, const GpuJob* jobs, const GpuChunk* chunks, const LaunchShape& shape, hipStream_t stream);

}  // namespace mrg
