// HIP kernels for gfx950: ray queries and the wavefront path integrator.
//
// One batch renders a list of chunks (job, pixel range, sample range) for one or many jobs:
//   reset -> generate -> { trace -> shade } x (max depth + 1) -> trace -> validate -> accumulate
// where `trace` of bounce i traces the shadow rays of bounce i - 1 and the extension rays of bounce
// i in one persistent launch (k_trace: a per-ray loop; k_trace_refill: the step driver with lane
// refill and parked leaves), or, in split mode, { extend -> shade -> shadow } per bounce.
// Queues: the extension queue (path ids, ping-pong by bounce) and the shadow queue (path id, shadow
// ray, contribution) are compacted in `shade` block-wide in thread order, with one atomic per queue
// per block. All bounce kernels are persistent (fixed grid, counts read from device
// memory), so the bounce loop can be captured once in a hipGraph.
//
// Reference: src/render/path_tracer.cpp:153-386 (trace, sample_pixel), scene.cpp:171-199 (camera),
// render_session.cpp:39-74 (film commit order: per pixel, samples in index order, FP64 sums).
#include <hip/hip_runtime.h>

#include "mrg_kernels.h"
#include "mrg_shade.h"
#include "mrg_traverse.h"

namespace mrg {

// ---------------------------------------------------------------------------------------------
// Wave helpers (wave64).
__device__ __forceinline__ uint32_t lane_id() { return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u)); }

__device__ __forceinline__ uint32_t prefix_in_mask(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(mask >> 32), __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mask), 0u));
}

// Reserves popcount(mask) consecutive slots of `counter` for the lanes in `mask`; returns each
// lane's slot (all lanes of the wave must call).
__device__ __forceinline__ uint32_t wave_reserve(uint32_t* counter, uint64_t mask) {
    uint32_t base = 0;
    if (mask) {
        const uint32_t leader = static_cast<uint32_t>(__builtin_ctzll(mask));
        if (lane_id() == leader) base = atomicAdd(counter, static_cast<uint32_t>(__builtin_popcountll(mask)));
        base = __shfl(base, static_cast<int>(leader), 64);
    }
    return base + prefix_in_mask(mask);
}

__device__ __forceinline__ void wave_add64(uint64_t* counter, uint32_t value) {
    // wave reduction then one atomic per wave
    uint32_t v = value;
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if (lane_id() == 0 && v) atomicAdd(reinterpret_cast<unsigned long long*>(counter), static_cast<unsigned long long>(v));
}

// Work distribution of the persistent traversal kernels: wave w first takes items [64 w, 64 w + 64)
// without any atomic, then claims CLAIM-item ranges from `work` (counted from the end of the static
// range). A queue that the static range covers (the late bounces) needs no atomic at all. The claim
// sizes were measured on the 2M-triangle terrain scene (64/128/256 items; claiming the next range
// ahead of time was slower): larger claims pay fewer contended atomics, smaller ones balance the
// tail of the more divergent shadow rays better.
constexpr uint32_t kExtendClaim = 128;  // minimum items per claim, extension rays (a multiple of 64)
constexpr uint32_t kShadowClaim = 64;   // minimum items per claim, shadow rays
constexpr uint32_t kTraceClaim  = 128;  // minimum items per claim, fused traversal
// Guided claims: a claim takes 1/kGuidedDivisor of this wave's share of the work still unclaimed
// (as of its last claim), at least CLAIM items, at most kGuidedMax: large launches pay fewer
// atomics early, and the tail is still balanced in CLAIM-sized pieces. 0 disables it (fixed CLAIM).
// Measured on the 1080p terrain scene: divisor 16 +2%, 8 +1%; 4 and 2 are slower than fixed
// claims (long private ranges spread the concurrently traced rays over the image and lose cache
// locality between neighbouring waves).
constexpr uint32_t kGuidedDivisor = 16;
constexpr uint32_t kGuidedMax     = 4096;

template <uint32_t CLAIM>
struct WaveWork {
    uint32_t base, static_end, claim_end, waves, seen;

    __device__ __forceinline__ void begin() {
        const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
        waves               = (gridDim.x * blockDim.x) >> 6;
        base                = wave * 64u;
        static_end          = waves * 64u;
        claim_end           = base + 64u;
        seen                = 0;
    }

    __device__ __forceinline__ void next(uint32_t* work, uint32_t n) {
        if (base + 64u < claim_end) {  // the rest of the current range
            base += 64u;
            return;
        }
        if (static_end >= n) {
            base = n;
            return;
        }
        uint32_t size = CLAIM;
        if (kGuidedDivisor) {
            const uint32_t claimed = static_end + seen;
            const uint32_t rest    = claimed < n ? n - claimed : 0u;
            const uint32_t guided  = (rest / (waves * kGuidedDivisor)) & ~63u;
            size = guided > size ? (guided < kGuidedMax ? guided : kGuidedMax) : size;
        }
        uint32_t t = 0;
        if (lane_id() == 0) t = atomicAdd(work, size);
        seen      = __shfl(t, 0, 64);
        base      = static_end + seen;
        claim_end = base + size;
    }
};

// Threads per shading block: the compaction works over the whole block (measured: 512 beats 256 by
// 2.5% and 1024 by 0.3% on the 1080p terrain scene).
constexpr uint32_t kShadeThreads = 512;

// Block-wide compaction into two queues at once, order-preserving: the block's entries of each queue
// keep thread order, so a queue holds continuing paths (and shadow rays) in the slot order the block
// read them in, and the next bounce reads their path records nearly in address order. (Grouping the
// entries by direction octant instead made the next traversal no faster and the next shading 12%
// slower: 45.9 -> 40.2 ms per 1080p job without it.) One atomic per queue per block. WAVES =
// blockDim.x / 64. Every thread of the block must call (it synchronises the block).
struct BlockSlots {
    uint32_t ext, shadow;
};

template <uint32_t WAVES>
__device__ __forceinline__ BlockSlots block_reserve(uint32_t* ext_counter, bool ext, uint32_t* sh_counter, bool sh,
                                                    uint32_t* s_off /*[2 * WAVES]*/, uint32_t* s_base /*[2]*/) {
    static_assert(2 * WAVES <= 64, "one lane of wave 0 per count");
    const uint32_t wave = threadIdx.x >> 6;
    const uint64_t me   = __ballot(ext);
    const uint64_t ms   = __ballot(sh);
    if (lane_id() == 0) {
        s_off[wave]         = static_cast<uint32_t>(__builtin_popcountll(me));
        s_off[WAVES + wave] = static_cast<uint32_t>(__builtin_popcountll(ms));
    }
    __syncthreads();
    if (wave == 0) {
        // exclusive scan of the counts, per queue (lanes 0..WAVES-1 queue 0, WAVES..2 WAVES-1 queue 1)
        const uint32_t l = lane_id();
        const uint32_t x = l < 2 * WAVES ? s_off[l] : 0u;
        uint32_t       v = x;
        for (int d = 1; d < static_cast<int>(2 * WAVES); d <<= 1) {
            const uint32_t y = __shfl_up(v, d, 64);
            if (l >= static_cast<uint32_t>(d)) v += y;
        }
        const uint32_t total0 = __shfl(v, WAVES - 1, 64);
        const uint32_t total1 = __shfl(v, 2 * WAVES - 1, 64) - total0;
        if (l < 2 * WAVES) s_off[l] = v - x - (l >= WAVES ? total0 : 0u);
        if (l == 0) {
            s_base[0] = total0 ? atomicAdd(ext_counter, total0) : 0u;
            s_base[1] = total1 ? atomicAdd(sh_counter, total1) : 0u;
        }
    }
    __syncthreads();
    BlockSlots g{0u, 0u};
    if (ext) g.ext = s_base[0] + s_off[wave] + prefix_in_mask(me);
    if (sh) g.shadow = s_base[1] + s_off[WAVES + wave] + prefix_in_mask(ms);
    __syncthreads();  // s_off / s_base are reused by the next call
    return g;
}

__device__ __forceinline__ StackView make_stack(uint32_t* lds_ref, float* lds_t, const Pool& P, uint32_t depth = kStackLds) {
    StackView s;
    s.depth     = depth;
    s.lds_ref   = lds_ref;
    s.lds_t     = lds_t;
    s.spill_ref = P.spill_ref;
    s.spill_t   = P.spill_t;
    s.lane      = threadIdx.x;
    s.glane     = blockIdx.x * blockDim.x + threadIdx.x;
    s.lanes     = gridDim.x * blockDim.x;
    return s;
}

// Persistent traversal with dynamic refill: every lane keeps its own ray; whenever at least
// REFILL_LANES lanes of the wave are idle, they take the next entries of the wave's claimed range
// and start new rays, so a wave does not wait for its slowest ray. A claim is one atomic for
// kRefillClaim entries, or for 64 when the launch has fewer than two items per lane of the grid:
// a small launch then spreads one item per lane over as many waves as it can fill, instead of two
// per lane over half as many. `start(i, st, any)` initialises item i, says whether it is an any-hit
// query (lanes of one wave may trace both kinds) and returns its scene; `finish(i, st)` consumes its
// result at the lane's next refill (or when the wave drains), so the lanes that finished since the
// last refill run it together. UNIFORM: every item traces `uscene` (kept in scalar registers, so node and triangle
// addresses are a scalar base plus a lane offset), and start's scene is ignored.
// Thresholds, measured on the 16.7M-ray icosphere microbench and the 1080p terrain job: ray queries
// start a ray from one packed record and refill at 16 idle lanes; bounce traversal gathers each ray
// through the path queue and refills at 40 (32: -2.5%, 48: -1.3%, 56: -16%). LEAF_BATCH is the number
// of lanes with a parked leaf that starts a triangle step (trace_step_dyn): 24 for both (16: -3%
// queries, -1% bounces; 32: -1.5% queries, bounces equal).
constexpr int      kQueryRefillLanes  = 16;
constexpr int      kBounceRefillLanes = 40;
constexpr int      kQueryLeafBatch    = 24;
constexpr int      kBounceLeafBatch   = 24;
constexpr uint32_t kRefillClaim = 128;  // entries per atomic claim (>= 64; 128 measured best, +61% over an atomic per refill)

template <int REFILL_LANES, int LEAF_BATCH, bool STATS, bool UNIFORM, bool EXACT, class Start, class Finish, class Keep>
__device__ __forceinline__ void persistent_trace_mixed(const StackView& stk, uint32_t n, uint32_t* work, const TraceScene& uscene,
                                                       Start start, Finish finish, const Keep& keep, uint64_t& nb, uint64_t& nt) {
    TraceState st;
    TraceScene ts{nullptr, nullptr, 0};
    bool       active  = false;
    bool       pending = false;  // the lane's finished ray still awaits finish()
    bool       any     = false;
    bool       drained = false;  // wave-uniform: the queue has no entries left
    uint32_t   item    = 0;
    uint32_t   cur = 0, end = 0;  // wave-uniform: the unused part of the wave's claimed range
    const uint32_t claim_size = n >= 2u * gridDim.x * blockDim.x ? kRefillClaim : 64u;
    for (;;) {
        const uint64_t idle = __ballot(!active);
        if (!drained && __builtin_popcountll(idle) >= REFILL_LANES) {
            // finished rays are consumed here, together, rather than one lane at a time in the
            // iteration each one ends
            if (pending) {
                finish(item, st);
                pending = false;
            }
            // idle lanes take the next entries of the wave's range; a new range of claim_size
            // entries is claimed (one atomic) when the current one runs short
            const uint32_t need  = static_cast<uint32_t>(__builtin_popcountll(idle));
            const uint32_t avail = end - cur;
            uint32_t       claim = 0;
            if (avail < need) {
                if (lane_id() == 0) claim = atomicAdd(work, claim_size);
                claim = __shfl(claim, 0, 64);
            }
            const uint32_t j    = prefix_in_mask(idle);
            const uint32_t slot = j < avail ? cur + j : claim + (j - avail);
            if (avail < need) {
                cur = claim + (need - avail);
                end = claim + claim_size;
            } else {
                cur += need;
            }
            drained = __ballot(!active && slot >= n) != 0;
            if (!active && slot < n) {
                item                = slot;
                const TraceScene sc = start(item, st, any);
                if (!UNIFORM) ts = sc;
                active = true;
            }
        }
        if (__ballot(active) == 0) break;
        if (active && trace_step_dyn<LEAF_BATCH, STATS, EXACT>(UNIFORM ? uscene : ts, st, stk, nb, nt, any, keep)) {
            active  = false;
            pending = true;
        }
    }
    if (pending) finish(item, st);
}

// Every item of one kind against one scene: closest hit (ANY = false) or any hit (ANY = true).
template <bool ANY, bool STATS, bool EXACT, class Start, class Finish>
__device__ __forceinline__ void persistent_trace(const StackView& stk, uint32_t n, uint32_t* work, const TraceScene& scene,
                                                 Start start, Finish finish, uint64_t& nb, uint64_t& nt) {
    persistent_trace_mixed<kQueryRefillLanes, kQueryLeafBatch, STATS, true, EXACT>(
        stk, n, work, scene,
        [&](uint32_t i, TraceState& st, bool& any) {
            any = ANY;
            start(i, st);
            return scene;
        },
        finish, KeepInResult{}, nb, nt);
}

// Reported distance of the winning hit: the ray/plane solve of that one triangle in FP64
// (t = n.(v0 - o) / n.d with n = (v1 - v0) x (v2 - v0)), rounded to float. The hit decision stays
// with the FP32 watertight test; this only removes the FP32 cancellation error of t for hits close
// to the ray origin. Falls back to the FP32 t for a ray parallel to the plane.
__device__ __forceinline__ float refine_distance(const GpuScene& sc, uint32_t slot, const RayPre& r, float t32) {
    const float4* tp = reinterpret_cast<const float4*>(sc.tris + slot);
    const float4  a = gld(tp + 0), b = gld(tp + 1), c = gld(tp + 2);
    const double  e1x = double(b.x) - a.x, e1y = double(b.y) - a.y, e1z = double(b.z) - a.z;
    const double  e2x = double(c.x) - a.x, e2y = double(c.y) - a.y, e2z = double(c.z) - a.z;
    const double  nx = e1y * e2z - e2y * e1z, ny = e1z * e2x - e2z * e1x, nz = e1x * e2y - e2x * e1y;
    const double  den = (nx * r.d.x + ny * r.d.y) + nz * r.d.z;
    if (den == 0.0) return t32;
    const double num = (nx * (double(a.x) - r.o.x) + ny * (double(a.y) - r.o.y)) + nz * (double(a.z) - r.o.z);
    return static_cast<float>(num / den);
}

// ---------------------------------------------------------------------------------------------
// Intersection microbenchmark / SceneBvh::intersect and ::occluded over a packed ray array.
// Six waves per SIMD (at most 80 VGPRs, 26 KB of LDS stack per block: the incoherent rays of queries
// stack deeper than path rays, and measured 2.5% slower at seven waves with 22 KB stacks).
// EXACT: hit decisions in FP64 as the reference's intersector (tri_test64): ids, distances and
// barycentrics are the FP64 ones rounded to float. Otherwise the FP32 watertight test of the render
// kernels, with the winner's distance refined in FP64 (refine_distance).
template <bool ANY, bool STATS, bool EXACT>
__device__ __forceinline__ void query_body(const GpuScene& sc, const float* __restrict__ rays, uint32_t n, float tmin, float tmax,
                                           mrg_hit_raw* __restrict__ hits, float* __restrict__ bary, uint8_t* __restrict__ occluded,
                                           uint32_t* __restrict__ work, const Pool& P, uint64_t* __restrict__ stats) {
    __shared__ uint32_t s_ref[kStackLdsQuery * kTraversalBlock];
    __shared__ float    s_t[kStackLdsQuery * kTraversalBlock];
    const StackView stk = make_stack(s_ref, s_t, P, kStackLdsQuery);
    uint64_t        nb = 0, nt = 0;
    const TraceScene ts = trace_scene(sc);
    persistent_trace<ANY, STATS, EXACT>(
        stk, n, work, ts,
        [&](uint32_t i, TraceState& st) {
            const float* r = rays + static_cast<size_t>(i) * 6;
            trace_init(st, mk(r[0], r[1], r[2]), mk(r[3], r[4], r[5]), tmin, tmax);
        },
        [&](uint32_t i, const TraceState& st) {
            TraceResult res = st.res;
            if (ANY) {
                occluded[i] = res.slot != 0xffffffffu ? 1 : 0;
            } else if (EXACT) {
                mrg_hit_raw h;
                h.t         = res.slot != 0xffffffffu ? static_cast<float>(st.best64) : __builtin_inff();
                h.primitive = res.slot != 0xffffffffu ? static_cast<int32_t>(res.prim) : -1;
                hits[i]     = h;
                if (bary) {
                    float b0 = 0.0f, b1 = 0.0f, b2 = 0.0f;
                    if (res.slot != 0xffffffffu) {
                        // the winner's FP64 hit again (same inputs, so it passes with t = best64)
                        const float4* tp = reinterpret_cast<const float4*>(sc.tris + res.slot);
                        const float4  a = gld(tp + 0), b = gld(tp + 1), c = gld(tp + 2);
                        TriHit64      w;
                        if (tri_test64(st.r, mk(a.x, a.y, a.z), mk(b.x, b.y, b.z), mk(c.x, c.y, c.z), st.tmin, st.best64, w)) {
                            b0 = static_cast<float>(w.a / w.det);
                            b1 = static_cast<float>(w.b / w.det);
                            b2 = static_cast<float>(w.c / w.det);
                        }
                    }
                    bary[static_cast<size_t>(i) * 3 + 0] = b0;
                    bary[static_cast<size_t>(i) * 3 + 1] = b1;
                    bary[static_cast<size_t>(i) * 3 + 2] = b2;
                }
            } else {
                mrg_hit_raw h;
                h.t         = res.slot != 0xffffffffu ? refine_distance(sc, res.slot, st.r, res.t) : __builtin_inff();
                h.primitive = res.slot != 0xffffffffu ? static_cast<int32_t>(res.prim) : -1;
                hits[i]     = h;
                if (bary) {
                    finish_barycentrics(res);
                    bary[static_cast<size_t>(i) * 3 + 0] = res.b0;
                    bary[static_cast<size_t>(i) * 3 + 1] = res.b1;
                    bary[static_cast<size_t>(i) * 3 + 2] = res.b2;
                }
            }
        },
        nb, nt);
    if (STATS) {
        atomicAdd(reinterpret_cast<unsigned long long*>(stats), static_cast<unsigned long long>(nb));
        atomicAdd(reinterpret_cast<unsigned long long*>(stats + 1), static_cast<unsigned long long>(nt));
    }
}

template <bool ANY, bool STATS>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6))) k_query(GpuScene sc, const float* __restrict__ rays, uint32_t n, float tmin, float tmax,
                                               mrg_hit_raw* __restrict__ hits, float* __restrict__ bary,
                                               uint8_t* __restrict__ occluded, uint32_t* __restrict__ work, Pool P,
                                               uint64_t* __restrict__ stats) {
    query_body<ANY, STATS, false>(sc, rays, n, tmin, tmax, hits, bary, occluded, work, P, stats);
}

// FP64 hit decisions at five waves per SIMD (96 VGPRs: the out-of-line FP64 test's call needs the
// registers; measured 2.71 ms per 16.7M-ray launch against 2.96 ms at six waves).
template <bool ANY, bool STATS>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5))) k_query_exact(GpuScene sc, const float* __restrict__ rays, uint32_t n, float tmin, float tmax,
                                               mrg_hit_raw* __restrict__ hits, float* __restrict__ bary,
                                               uint8_t* __restrict__ occluded, uint32_t* __restrict__ work, Pool P,
                                               uint64_t* __restrict__ stats) {
    query_body<ANY, STATS, true>(sc, rays, n, tmin, tmax, hits, bary, occluded, work, P, stats);
}

// ---------------------------------------------------------------------------------------------
// Batch start: clear counters (a fill kernel, so a captured graph holds no memset node).
__global__ void k_reset(Pool P) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < kCounterWords; i += gridDim.x * blockDim.x)
        P.counters[i] = i == kCtrExt ? P.header[kHdrSlots] : 0u;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < kStatWords; i += gridDim.x * blockDim.x) P.stats[i] = 0;
}

// Pixel tile of the slot order (a power of two wide; kTileW * kTileH = one wave).
constexpr uint32_t kTileW = 4, kTileH = 4;
// Samples of one pixel held side by side in a tiled chunk (a power of two): a wave then holds
// kSampleGroup samples of each pixel of one kTileW x kTileH tile.
constexpr uint32_t kSampleGroup = 4;
static_assert((kTileW & (kTileW - 1u)) == 0 && (kSampleGroup & (kSampleGroup - 1u)) == 0 && kTileW * kTileH * kSampleGroup == 64,
              "tile: power-of-two width, one wave");

// Slot layout of a chunk: samples [0, grouped) in groups of kSampleGroup, each group pixel-major
// (slot = group base + q * kSampleGroup + sample within the group), then the remaining samples one
// after the other, each over all pixel positions q. Chunks in scanline order group no samples.
__device__ __forceinline__ uint32_t grouped_samples(const GpuChunk& ch) {
    return ch.tile_rows ? ch.sample_count & ~(kSampleGroup - 1u) : 0u;
}

// Slot of sample s at pixel position q.
__device__ __forceinline__ uint32_t chunk_slot(const GpuChunk& ch, uint32_t s, uint32_t q) {
    const uint32_t n       = ch.pixel_count;
    const uint32_t grouped = grouped_samples(ch);
    if (s < grouped) return ch.slot_begin + (s / kSampleGroup) * (n * kSampleGroup) + q * kSampleGroup + (s & (kSampleGroup - 1u));
    return ch.slot_begin + grouped * n + (s - grouped) * n + q;
}

// Pixel (x, y) of pixel position q of a chunk. A chunk of whole image rows (tile_rows != 0) lays its
// pixels out in kTileW x kTileH tiles, tile rows top to bottom and tiles left to right (the last tile
// row and column hold the remainder rows and columns); with kSampleGroup samples of each pixel side
// by side (chunk_slot), the 64 paths of a wave start from one compact patch of the image instead
// of a 64-pixel strip. Other chunks keep scanline order. k_generate, k_validate and k_accumulate
// use the same maps, so every sample lands on its own pixel.
__device__ __forceinline__ void chunk_xy(const GpuChunk& ch, uint32_t width, uint32_t q, uint32_t& x, uint32_t& y) {
    if (ch.tile_rows == 0u) {
        const uint32_t pixel = ch.pixel_begin + q;
        y = pixel / width;
        x = pixel - y * width;
        return;
    }
    constexpr uint32_t TW = kTileW, TH = kTileH;
    const uint32_t band = q / (width * TH);
    const uint32_t y0   = band * TH;
    const uint32_t h    = min(TH, ch.tile_rows - y0);
    const uint32_t in   = q - band * (width * TH);
    const uint32_t full = width & ~(TW - 1u);
    if (in < full * h) {
        const uint32_t t = h == TH ? in / (TW * TH) : in / (TW * h);
        const uint32_t w = in - t * (TW * h);
        y = w / TW;
        x = t * TW + (w & (TW - 1u));
    } else {
        const uint32_t rest = width - full;
        const uint32_t w    = in - full * h;
        y = w / rest;
        x = full + (w - y * rest);
    }
    y += y0 + ch.pixel_begin / width;
}

__device__ __forceinline__ uint32_t find_chunk_by_pixel(const GpuChunk* chunks, uint32_t count, uint32_t pix) {
    uint32_t lo = 0, hi = count;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (chunks[mid].pixel_prefix <= pix) lo = mid; else hi = mid;
    }
    return lo;
}

// The camera ray of one path slot, drawn from the path's sample stream, and the stream's state after
// the four camera draws (PathTracer::sample_pixel, path_tracer.cpp:368-386; Camera::generate_lens_ray).
__device__ __forceinline__ void start_camera_path(const Pool& P, const GpuJob& job, uint32_t job_index, uint32_t slot, uint32_t x,
                                                  uint32_t y, Pcg rng, bool one_job) {
    const float      fx     = next_1d(rng);
    const float      fy     = next_1d(rng);
    const float      lx     = next_1d(rng);
    const float      ly     = next_1d(rng);
    const GpuCamera& cam    = job.camera;
    const float      uvx    = (static_cast<float>(x) + fx) / static_cast<float>(job.width);
    const float      uvy    = (static_cast<float>(y) + fy) / static_cast<float>(job.height);
    const float      aspect = static_cast<float>(job.width) / static_cast<float>(job.height);
    const F3         fwd    = ld3(cam.forward), right = ld3(cam.right), up = ld3(cam.up), pos = ld3(cam.position);
    const F3         pd = normalize((fwd + (((2.0f * uvx - 1.0f) * aspect) * cam.half_height) * right) +
                                    ((1.0f - 2.0f * uvy) * cam.half_height) * up);
    F3 o = pos, d = pd;
    if (cam.aperture != 0.0f) {
        float dx, dy;
        sample_disk(lx, ly, dx, dy);
        const float kx = dx * cam.aperture, ky = dy * cam.aperture;
        const F3    focus_point = pos + pd * (cam.focus / dot(pd, fwd));
        o                       = (pos + kx * right) + ky * up;
        d                       = normalize(focus_point - o);
    }
    const float cosine = dot(d, fwd);
    P.ray_o[slot] = make_float4(o.x, o.y, o.z, cam.near_plane / cosine);
    P.ray_d[slot] = make_float4(d.x, d.y, d.z, cam.far_plane / cosine);
    // beta = 1, pdf 0, radiance 0 and eta scale 1 are implied for bounce 0 (k_shade), and so is the
    // previous-event record of a one-job batch (job 0, a camera path's flags)
    if (!one_job) P.prev[slot] = make_float4(0.0f, 0.0f, 0.0f, __uint_as_float(kPrevDelta | (job_index << 16)));
    P.rng[slot] = make_uint4(static_cast<uint32_t>(rng.state), static_cast<uint32_t>(rng.state >> 32), static_cast<uint32_t>(rng.inc),
                             static_cast<uint32_t>(rng.inc >> 32));
}

__device__ __forceinline__ uint32_t find_chunk_by_unit(const GpuChunk* chunks, uint32_t count, uint32_t unit) {
    uint32_t lo = 0, hi = count;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (chunks[mid].unit_prefix <= unit) lo = mid; else hi = mid;
    }
    return lo;
}

// Camera rays of every path slot of the batch, pixel-stationary: a thread takes a work unit (one
// pixel of a chunk in scanline order, or one pixel and one position within the sample groups of a
// chunk of whole rows) and walks that unit's samples, so the pixel's position, its share of the
// stream set-up (pixel_stream) and the job's camera are formed once per unit instead of once per
// path. Consecutive units are consecutive slots of each sample (group), so the 64 lanes of a wave
// write 64 consecutive slots per step.
__global__ void __launch_bounds__(256) k_generate(Pool P, const GpuJob* __restrict__ jobs, const GpuChunk* __restrict__ chunks) {
    const uint32_t chunk_count = P.header[kHdrChunks];
    const uint32_t units       = P.header[kHdrUnits];
    const bool     one_job     = P.header[kHdrOneJob] != 0u;
    for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < units; u += gridDim.x * blockDim.x) {
        const GpuChunk ch    = chunks[chunk_count == 1u ? 0u : find_chunk_by_unit(chunks, chunk_count, u)];
        const GpuJob&  job   = jobs[ch.job];
        const uint32_t n     = ch.pixel_count;
        const uint32_t local = u - ch.unit_prefix;
        if (ch.tile_rows != 0u) {
            const uint32_t groups = ch.sample_count / kSampleGroup;  // grouped_samples(ch) / kSampleGroup
            const uint32_t rest   = ch.sample_count - groups * kSampleGroup;
            const uint32_t pitch  = n * kSampleGroup;                // slots of one sample group (chunk_slot)
            const uint32_t q = local / kSampleGroup, j = local & (kSampleGroup - 1u);
            uint32_t       x, y;
            chunk_xy(ch, job.width, q, x, y);
            const PixelStream px   = pixel_stream(job.seed, static_cast<uint64_t>(y) * job.width + x);
            uint32_t          slot = ch.slot_begin + local;
            for (uint32_t g = 0; g < groups; ++g, slot += pitch)
                start_camera_path(P, job, ch.job, slot, x, y, for_sample(px, ch.sample_begin + g * kSampleGroup + j), one_job);
            if (j < rest)  // the samples after the last whole group, one over all pixel positions each
                start_camera_path(P, job, ch.job, ch.slot_begin + groups * pitch + j * n + q, x, y,
                                  for_sample(px, ch.sample_begin + groups * kSampleGroup + j), one_job);
        } else {  // scanline order, sample-major: slot = first slot + s * n + q (chunk_slot)
            uint32_t x, y;
            chunk_xy(ch, job.width, local, x, y);
            const PixelStream px = pixel_stream(job.seed, static_cast<uint64_t>(y) * job.width + x);
            for (uint32_t s = 0; s < ch.sample_count; ++s)
                start_camera_path(P, job, ch.job, ch.slot_begin + s * n + local, x, y, for_sample(px, ch.sample_begin + s), one_job);
        }
    }
}

// Extension queue of bounce `iter` (ping-pong between two arrays); null at bounce 0, where every
// path slot of the batch continues in slot order (k_generate writes no queue).
__device__ __forceinline__ const uint32_t* ext_queue(const Pool& P, uint32_t iter) {
    return iter == 0 ? nullptr : (iter & 1u) ? P.q_ext1 : P.q_ext0;
}

// One extension ray: closest hit of path queue[i] (path i at bounce 0, whose queue is the identity:
// a null queue). ONE_SCENE: every job of the batch renders
// jobs[0]'s scene, read through a uniform index (scalar loads) without the path's job lookup.
// job0: the path's job is 0 without a record to read (bounce 0 of a one-job batch).
template <bool ONE_SCENE>
__device__ __forceinline__ void extend_one(const Pool& P, const GpuJob* __restrict__ jobs, const uint32_t* queue, uint32_t i,
                                           const StackView& stk, bool job0) {
    uint64_t       dummy = 0;
    const uint32_t path  = queue ? queue[i] : i;
    const float4   ro    = P.ray_o[path];
    const float4   rd    = P.ray_d[path];
    const uint32_t job   = ONE_SCENE || job0 ? 0u : __float_as_uint(P.prev[path].w) >> 16;
    const RayPre   rp    = ray_setup(mk(ro.x, ro.y, ro.z), mk(rd.x, rd.y, rd.z));
    const TraceResult res = trace_ray<false, false>(jobs[job].scene, rp, ro.w, rd.w, stk, &dummy, &dummy);
    P.hit[path]      = make_float4(res.b0, res.b1, res.b2, 1.0f);  // already divided: their sum stands in as 1
    P.hit_slot[path] = res.slot;
}

// One shadow ray (entry i of the shadow queue): any hit; unoccluded rays add their contribution.
template <bool ONE_SCENE>
__device__ __forceinline__ void shadow_one(const Pool& P, const GpuJob* __restrict__ jobs, uint32_t i, const StackView& stk) {
    uint64_t       dummy = 0;
    const uint32_t path  = P.q_sh[i];
    const float4   o     = P.sh_o[i];
    const float4   d     = P.sh_d[i];
    const uint32_t job   = ONE_SCENE ? 0u : __float_as_uint(P.prev[path].w) >> 16;
    const RayPre   rp    = ray_setup(mk(o.x, o.y, o.z), mk(d.x, d.y, d.z));
    const TraceResult res = trace_ray<true, false>(jobs[job].scene, rp, 0.0f, o.w, stk, &dummy, &dummy, __float_as_uint(d.w));
    if (res.slot == 0xffffffffu) {
        const float4 c = P.sh_c[i];
        float4       l = P.rad[path];
        l.x = l.x + c.x;
        l.y = l.y + c.y;
        l.z = l.z + c.z;
        P.rad[path] = l;
    }
}

// Closest hit for every path in the extension queue of this bounce (split traversal kernels).
__global__ void __launch_bounds__(256) k_extend(Pool P, const GpuJob* __restrict__ jobs, uint32_t iter) {
    __shared__ uint32_t s_ref[kStackLds * kTraversalBlock];
    __shared__ float    s_t[kStackLds * kTraversalBlock];
    const StackView stk   = make_stack(s_ref, s_t, P);
    const uint32_t  n     = P.counters[kCtrExt + iter];
    const uint32_t* queue = ext_queue(P, iter);
    uint32_t*       work  = P.counters + kCtrWorkExt + iter;
    const bool      job0  = iter == 0 && P.header[kHdrOneJob] != 0u;
    WaveWork<kExtendClaim> w;
    for (w.begin(); w.base < n; w.next(work, n)) {
        const uint32_t i = w.base + lane_id();
        if (i < n) extend_one<false>(P, jobs, queue, i, stk, job0);
    }
}

// Fused traversal of one bounce: the shadow rays of bounce iter - 1 and the extension rays of
// bounce iter (both produced by the previous shading pass) share one persistent launch. The two
// ranges are padded to whole 64-item chunks, so every wave traces one kind of ray at a time.
// Radiance accumulation keeps its order: a path's shadow contribution of bounce iter - 1 lands
// before the shading of bounce iter reads it. ONE_SCENE: the batch renders one scene (extend_one).
template <bool ONE_SCENE>
__global__ void __launch_bounds__(256) k_trace(Pool P, const GpuJob* __restrict__ jobs, uint32_t iter) {
    __shared__ uint32_t s_ref[kStackLds * kTraversalBlock];
    __shared__ float    s_t[kStackLds * kTraversalBlock];
    const StackView stk    = make_stack(s_ref, s_t, P);
    const uint32_t  n_ext  = P.counters[kCtrExt + iter];
    const uint32_t  n_sh   = iter ? P.counters[kCtrShadow + iter - 1] : 0u;
    const uint32_t  sh_end = (n_sh + 63u) & ~63u;  // shadow rays first: the longer, more divergent work
    const uint32_t  n      = sh_end + n_ext;
    const uint32_t* queue  = ext_queue(P, iter);
    uint32_t*       work   = P.counters + kCtrWorkExt + iter;
    const bool      job0   = iter == 0 && P.header[kHdrOneJob] != 0u;
    WaveWork<kTraceClaim> w;
    for (w.begin(); w.base < n; w.next(work, n)) {
        if (w.base < sh_end) {
            const uint32_t i = w.base + lane_id();
            if (i < n_sh) shadow_one<ONE_SCENE>(P, jobs, i, stk);
        } else {
            const uint32_t i = w.base - sh_end + lane_id();
            if (i < n_ext) extend_one<ONE_SCENE>(P, jobs, queue, i, stk, job0);
        }
    }
}

// Fused traversal of one bounce with the refill driver: the same work as k_trace (shadow rays of
// bounce iter - 1 first, then the extension rays of bounce iter), but lanes that finish their ray
// take the next queue entry at once, so a wave can trace shadow and extension rays side by side.
// ONE_SCENE: every job of the batch renders the same scene (jobs[0]'s), read once per wave.
// Seven waves per SIMD (at most 72 VGPRs, 22 KB of LDS stack per block): seven resident blocks per CU;
// the seventh wave hides more of the dependent node-fetch latency: 137.3 -> 131.3 ms of traversal per
// 1080p job.
template <bool ONE_SCENE>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(7))) k_trace_refill(Pool P, const GpuJob* __restrict__ jobs, uint32_t iter) {
    __shared__ uint32_t s_ref[kStackLds * kTraversalBlock];
    __shared__ float    s_t[kStackLds * kTraversalBlock];
    const StackView stk   = make_stack(s_ref, s_t, P);
    const uint32_t  n_ext = P.counters[kCtrExt + iter];
    const uint32_t  n_sh  = iter ? P.counters[kCtrShadow + iter - 1] : 0u;
    const uint32_t* queue = ext_queue(P, iter);
    uint32_t*       work  = P.counters + kCtrWorkExt + iter;
    uint64_t        nb = 0, nt = 0;
    uint32_t        path = 0;
    const bool      job0 = iter == 0 && P.header[kHdrOneJob] != 0u;
    const TraceScene one = ONE_SCENE ? trace_scene(jobs[0].scene) : TraceScene{nullptr, nullptr, 0};
    persistent_trace_mixed<kBounceRefillLanes, kBounceLeafBatch, false, ONE_SCENE, false>(
        stk, n_sh + n_ext, work, one,
        [&](uint32_t i, TraceState& st, bool& any) {
            any = i < n_sh;
            if (any) {
                path           = P.q_sh[i];
                const float4 o = P.sh_o[i];
                const float4 d = P.sh_d[i];
                trace_init(st, mk(o.x, o.y, o.z), mk(d.x, d.y, d.z), 0.0f, o.w, __float_as_uint(d.w));
            } else {
                path            = queue ? queue[i - n_sh] : i - n_sh;
                const float4 ro = P.ray_o[path];
                const float4 rd = P.ray_d[path];
                trace_init(st, mk(ro.x, ro.y, ro.z), mk(rd.x, rd.y, rd.z), ro.w, rd.w);
            }
            if (ONE_SCENE) return one;
            const uint32_t job = job0 ? 0u : __float_as_uint(P.prev[path].w) >> 16;
            return trace_scene(jobs[job].scene);
        },
        [&](uint32_t i, const TraceState& st) {
            TraceResult res = st.res;
            if (i < n_sh) {
                if (res.slot == 0xffffffffu) {
                    const float4 c = P.sh_c[i];
                    float4       l = P.rad[path];
                    l.x = l.x + c.x;
                    l.y = l.y + c.y;
                    l.z = l.z + c.z;
                    P.rad[path] = l;
                }
            } else {
                P.hit_slot[path] = res.slot;  // the hit record holds the winner's edge functions already
            }
        },
        [&](TraceResult&, const TriHit& h) { P.hit[path] = make_float4(h.a, h.b, h.c, h.det); },
        nb, nt);
}

// The path state and hit record of one shading thread, loaded before its job is known (they
// depend on the path slot alone), and the hit triangle, loaded as soon as the hit slot arrives.
struct ShadeIn {
    float4   pv, rd, hv;
    uint4    rg;
    uint32_t slot;
    float4   ta, tb, tc;  // hit triangle (vertices; v1.w = flags, v2.w = instance)
};

// One path of the shading pass (the loop body of k_shade): takes the path state and its hit, adds
// emission, samples a light and the BSDF, and leaves the continuation and shadow-ray outputs in the
// references. `job` is the path's job and `sc` its scene; a caller whose wave holds a single job
// passes references through wave-uniform indices, so their fields load through the scalar cache.
template <bool EXT>
__device__ __forceinline__ void shade_path(const Pool& P, const GpuJob& job, const GpuScene& sc, uint32_t path, const ShadeIn& in,
                                           uint32_t iter, bool& cont, bool& emit,
                                           float4* s_shadow, uint32_t& n_scatter, uint32_t& n_rr, uint32_t& n_shadow) {
        // the cancel word, read first and used last so that its latency hides behind the shading; the host
        // raises it while kernels run, so it is read past the per-XCD L2s
        const bool     cancelled = __hip_atomic_load(job.error_flag + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u;
        const float4   pv    = in.pv;
        const uint32_t flags = __float_as_uint(pv.w);
        const uint32_t depth = iter;
        const float4  rd = in.rd;
        const F3      ray_d = mk(rd.x, rd.y, rd.z);
        // camera paths (bounce 0) start with beta = 1, pdf 0, L = 0, eta scale 1: k_generate
        // writes none of them and they are not read here
        const float4  bt = iter == 0 ? make_float4(1.0f, 1.0f, 1.0f, 0.0f) : P.beta[path];
        const float4  lr = iter == 0 ? make_float4(0.0f, 0.0f, 0.0f, 1.0f) : P.rad[path];
        F3            beta = mk(bt.x, bt.y, bt.z);
        float         prev_pdf  = bt.w;
        F3            L         = mk(lr.x, lr.y, lr.z);
        float         eta_scale = lr.w;
        F3            prev_ref  = mk(pv.x, pv.y, pv.z);
        bool          prev_delta = (flags & kPrevDelta) != 0u;
        const uint4   rg = in.rg;
        Pcg           rng{static_cast<uint64_t>(rg.x) | (static_cast<uint64_t>(rg.y) << 32),
                          static_cast<uint64_t>(rg.z) | (static_cast<uint64_t>(rg.w) << 32)};
        const float4  hv   = in.hv;
        const uint32_t slot = in.slot;
        bool bad = false;
        do {
            if (!finite3(beta) || beta.x < 0.0f || beta.y < 0.0f || beta.z < 0.0f) { bad = true; break; }
            if (vmax(beta) == 0.0f) break;
            const uint32_t mode = job.mode;
            const bool unit_weight = depth == 0 || prev_delta || mode == 0u;
            if (slot == 0xffffffffu) {
                float w = 1.0f;
                if (!unit_weight) w = mode == 2u ? power_heuristic(prev_pdf, EXT ? env_light_pdf(sc, ray_d) : sc.env_pdf, bad) : 0.0f;
                L = L + (beta * (EXT ? env_evaluate(sc, ray_d) : ld3(sc.env))) * w;
                break;
            }
            // surface_hit (bvh.cpp:302-370)
            const float4* tp = reinterpret_cast<const float4*>(sc.tris + slot);
            const float4  ta = EXT ? gld(tp + 0) : in.ta, tb = EXT ? gld(tp + 1) : in.tb, tc = EXT ? gld(tp + 2) : in.tc;
            const F3      v0 = mk(ta.x, ta.y, ta.z), v1 = mk(tb.x, tb.y, tb.z), v2 = mk(tc.x, tc.y, tc.z);
            const uint32_t tflags = __float_as_uint(tb.w);
            const float4* np = reinterpret_cast<const float4*>(sc.normals + slot);
            const float4  na = gld(np + 0), nb = gld(np + 1), nc = gld(np + 2);
            const uint32_t inst = __float_as_uint(tc.w);
            const GpuMaterial& mat = sc.materials[inst];
            // corner normals arrive normalized and the geometric normal in their w components
            // (computed at upload with these same float operations)
            F3 geometric = mk(na.w, nb.w, nc.w);
            if (tflags & kTriMirrored) geometric = -geometric;
            // the hit record holds the edge functions and their sum: barycentric i = e_i / sum, the
        // division the watertight test does when it accepts the hit
        const float b0 = hv.x / hv.w, b1 = hv.y / hv.w, b2 = hv.z / hv.w;
            F3 shading  = mk(0.0f, 0.0f, 0.0f);
            F3 position = mk(0.0f, 0.0f, 0.0f);
            F3 absolute = mk(0.0f, 0.0f, 0.0f);
            shading = shading + b0 * mk(na.x, na.y, na.z);
            { const F3 term = b0 * v0; position = position + term; absolute = absolute + vabs(term); }
            shading = shading + b1 * mk(nb.x, nb.y, nb.z);
            { const F3 term = b1 * v1; position = position + term; absolute = absolute + vabs(term); }
            shading = shading + b2 * mk(nc.x, nc.y, nc.z);
            { const F3 term = b2 * v2; position = position + term; absolute = absolute + vabs(term); }
            if (dot(shading, shading) < 1e-24f) {
                shading = geometric;
            } else {
                shading = normalize(shading);
                if (dot(shading, geometric) < 0.0f) shading = -shading;
            }
            // base color: the material's, times the clamped texel at the hit uv = sum of b_i uv_i
            // (bvh.cpp:318-321, path_tracer.cpp:279-282, bsdf.cpp:122-125)
            // and the tangent-space normal map of surface_hit (bvh.cpp:334-359)
            F3 base_color = ld3(mat.color);
            if (EXT && (mat.texture >= 0 || mat.normal_map >= 0)) {
                float  u = 0.0f, v = 0.0f;
                float2 q[3] = {make_float2(0.0f, 0.0f), make_float2(0.0f, 0.0f), make_float2(0.0f, 0.0f)};
                if (sc.uvs) {
                    const float2* qp = reinterpret_cast<const float2*>(sc.uvs) + static_cast<uint64_t>(slot) * 3u;
                    q[0] = gld(qp + 0), q[1] = gld(qp + 1), q[2] = gld(qp + 2);
                    u = ((u + b0 * q[0].x) + b1 * q[1].x) + b2 * q[2].x;
                    v = ((v + b0 * q[0].y) + b1 * q[1].y) + b2 * q[2].y;
                }
                if (mat.normal_map >= 0 && mat.normal_scale > 0.0f) apply_normal_map(sc, mat, v0, v1, v2, q, u, v, geometric, shading);
                if (mat.texture >= 0) {
                    const F3 tc = texture_sample(sc, mat.texture, mat.filter, u, v);
                    base_color  = base_color * mk(sclamp(tc.x, 0.0f, 1.0f), sclamp(tc.y, 0.0f, 1.0f), sclamp(tc.z, 0.0f, 1.0f));
                }
            }
            const bool front_face = dot(ray_d, geometric) < 0.0f;
            const F3   perr = absolute * (7.0f * kEpsF);
            const F3   ng   = geometric;
            // emission with its MIS weight
            const F3 emission = ld3(mat.emission);
            if (emission.x != 0.0f || emission.y != 0.0f || emission.z != 0.0f) {
                float w = 1.0f;
                if (!unit_weight) {
                    w = mode == 2u ? power_heuristic(prev_pdf, area_pdf(sc, sc.slot_emitter[slot], prev_ref, position), bad) : 0.0f;
                }
                L = L + (beta * emission) * w;
            }
            if (depth == job.max_depth) break;
            // shading frame and BSDF
            F3 normal = mat.kind == 3u ? ng : shading;
            const F3 wo_world = -ray_d;
            if (dot(normal, wo_world) < 0.0f) normal = -normal;
            const Frame f  = make_frame(normal);
            const F3    wo = to_local(f, wo_world);
            BsdfParams  bp;
            bp.kind  = mat.kind;
            bp.color = base_color;
            bp.alpha = mat.alpha;
            bp.eta_i = front_face ? 1.0f : mat.ior;
            bp.eta_t = front_face ? mat.ior : 1.0f;
            bp.roughness = mat.roughness;
            bp.table     = P.energy;
            bp.energy    = EXT && mat.kind == 1u && mat.compensate != 0u && !bp.is_delta();
            const float light_u   = next_1d(rng);
            const float su_x      = next_1d(rng);
            const float su_y      = next_1d(rng);
            const float component = next_1d(rng);
            const float du_x      = next_1d(rng);
            const float du_y      = next_1d(rng);
            const F3    won       = normalize(wo);
            bsdf_prepare(bp, won);
            if (!bp.is_delta()) {
                LightSample ls;
                if (sample_light<EXT>(sc, position, light_u, su_x, su_y, ls) && (mode != 0u || ls.discrete)) {
                    const F3   wi   = to_local(f, ls.dir);
                    const bool side = dot(wo_world, ng) * dot(ls.dir, ng) > 0.0f;
                    if (wi.z > 0.0f && side) {
                        ++n_shadow;
                        F3    origin = offset_origin(position, ng, perr, ls.dir);
                        F3    sdir   = normalize(ls.dir);
                        float limit  = __builtin_inff();
                        bool  zero   = false;
                        if (__builtin_isfinite(ls.distance)) {
                            const F3 target = offset_origin(ls.position, ls.normal, ls.error, -ls.dir);
                            const F3 diff   = target - origin;
                            limit           = length(diff);
                            if (limit == 0.0f) {
                                zero = true;
                            } else {
                                sdir  = diff / limit;
                                limit = next_after(limit, 0.0f);
                            }
                        }
                        if (!zero) {
                            const F3 win = normalize(wi);
                            float weight = 1.0f;
                            if (mode == 2u && !ls.discrete) weight = power_heuristic(ls.probability, bsdf_pdf_n(bp, won, win), bad);
                            const F3 c = (((beta * bsdf_eval_n(bp, won, win)) * wi.z) * ls.radiance) * (weight / ls.probability);
                            emit   = true;
                            // the shadow record waits in LDS (this thread's row) until the block has
                            // reserved its queue slots, so it holds no registers through BSDF sampling
                            s_shadow[threadIdx.x] = make_float4(origin.x, origin.y, origin.z, limit);
                            // w: the emitter triangle the segment ends on, which its any-hit skips
                            s_shadow[kShadeThreads + threadIdx.x]     = make_float4(sdir.x, sdir.y, sdir.z, __uint_as_float(ls.prim));
                            s_shadow[2 * kShadeThreads + threadIdx.x] = make_float4(c.x, c.y, c.z, 0.0f);
                        }
                    }
                }
            }
            BsdfSample bs;
            if (!bsdf_sample(bp, won, component, du_x, du_y, bs)) break;
            if (vmax(bs.weight) == 0.0f) break;
            const F3    wi_world = normalize(to_world(f, bs.dir));
            const float sides    = dot(wo_world, ng) * dot(wi_world, ng);
            if ((bs.transmission && sides >= 0.0f) || (!bs.transmission && sides <= 0.0f)) break;
            beta = beta * bs.weight;
            if (!finite3(beta) || beta.x < 0.0f || beta.y < 0.0f || beta.z < 0.0f) { bad = true; break; }
            prev_ref   = position;
            prev_pdf   = bs.probability;
            prev_delta = bs.discrete;
            eta_scale  = eta_scale * (bs.eta * bs.eta);
            ++n_scatter;
            const F3 no = offset_origin(position, ng, perr, wi_world);
            const F3 nd = normalize(wi_world);
            const float roulette_u = next_1d(rng);
            if (job.rr && depth + 1 >= job.rr_depth) {
                const float survival = sclamp(vmax(beta * eta_scale), 0.05f, 0.95f);
                if (roulette_u >= survival) {
                    ++n_rr;
                    break;
                }
                beta = beta / survival;
            }
            cont    = true;
            P.ray_o[path] = make_float4(no.x, no.y, no.z, 0.0f);
            P.ray_d[path] = make_float4(nd.x, nd.y, nd.z, __builtin_inff());
            P.beta[path]  = make_float4(beta.x, beta.y, beta.z, prev_pdf);
            P.prev[path]  = make_float4(prev_ref.x, prev_ref.y, prev_ref.z,
                                        __uint_as_float((flags & ~kPrevDelta) | (prev_delta ? kPrevDelta : 0u)));
            P.rng[path]   = make_uint4(static_cast<uint32_t>(rng.state), static_cast<uint32_t>(rng.state >> 32),
                                       static_cast<uint32_t>(rng.inc), static_cast<uint32_t>(rng.inc >> 32));
        } while (false);
        if (cancelled) {  // the job was cancelled: its paths stop here and its batch commits none of them
            cont = false;
            emit = false;
        }
        P.rad[path] = make_float4(L.x, L.y, L.z, eta_scale);
        if (bad) atomicOr(job.error_flag, 1u);
}

// Shading of one bounce (PathTracer::trace loop body, path_tracer.cpp:164-363).
// At most 128 VGPRs (4 waves per SIMD): the kernel is bound by the latency of its path-state
// and scene loads, so occupancy pays more than the registers the compiler would otherwise use.
// EXT = false compiles out base-color textures, environment maps and energy compensation: scenes
// without them run a kernel with fewer live registers.
// ONE_SCENE: every job of the batch renders jobs[0]'s scene, whose pointers are read once (scalar),
// so the hit triangle's loads wait for the hit slot alone.
template <bool EXT, bool ONE_SCENE>
__global__ void __launch_bounds__(kShadeThreads) __attribute__((amdgpu_waves_per_eu(4))) k_shade(Pool P, const GpuJob* __restrict__ jobs, uint32_t iter) {
    const uint32_t  n       = P.counters[kCtrExt + iter];
    const uint32_t* queue   = ext_queue(P, iter);
    uint32_t*       next_q  = (iter & 1u) ? P.q_ext0 : P.q_ext1;
    uint32_t*       next_n  = P.counters + kCtrExt + iter + 1;
    uint32_t*       sh_n    = P.counters + kCtrShadow + iter;
    const uint32_t  stride  = gridDim.x * blockDim.x;
    const bool      job0    = iter == 0 && P.header[kHdrOneJob] != 0u;  // bounce 0 of a one-job batch: no prev record
    uint32_t        n_scatter = 0, n_rr = 0, n_shadow = 0;
    __shared__ uint32_t s_off[2 * (kShadeThreads / 64)];
    __shared__ uint32_t s_base[2];
    __shared__ float4   s_shadow[3 * kShadeThreads];  // origin + limit, direction + emitter, contribution
    // block-uniform loop: the grouped compaction below synchronises the block
    for (uint32_t bb = blockIdx.x * blockDim.x; bb < n; bb += stride) {
        const uint32_t i        = bb + threadIdx.x;
        bool           cont     = false;
        bool           emit     = false;
        uint32_t       path     = 0;
        if (i < n) {
            path = queue ? queue[i] : i;
            // everything that depends on the path slot alone is in flight before the job is known
            ShadeIn in;
            in.pv   = job0 ? make_float4(0.0f, 0.0f, 0.0f, __uint_as_float(kPrevDelta)) : P.prev[path];
            in.rd   = P.ray_d[path];
            in.hv   = P.hit[path];
            in.slot = P.hit_slot[path];
            in.rg   = P.rng[path];
            const uint32_t jid = __float_as_uint(in.pv.w) >> 16;
            const uint32_t j0  = __builtin_amdgcn_readfirstlane(jid);
            const bool     uniform = __ballot(jid != j0) == 0;
            const GpuScene& sc = ONE_SCENE ? jobs[0].scene : (uniform ? jobs[j0].scene : jobs[jid].scene);
            // the hit triangle, fetched before the shading decides it needs it (a
            // missed ray reads the first triangle instead, unused; a scene without triangles has no hits)
            if (!EXT && sc.tri_count != 0u) {  // (the extended kernel has no registers to spare for it)
                const uint32_t t  = in.slot != 0xffffffffu ? in.slot : 0u;
                const float4*  tp = reinterpret_cast<const float4*>(sc.tris + t);
                in.ta = gld(tp + 0), in.tb = gld(tp + 1), in.tc = gld(tp + 2);
            }
            if (uniform)
                shade_path<EXT>(P, jobs[j0], sc, path, in, iter, cont, emit, s_shadow, n_scatter, n_rr, n_shadow);
            else
                shade_path<EXT>(P, jobs[jid], sc, path, in, iter, cont, emit, s_shadow, n_scatter, n_rr, n_shadow);
        }
        const BlockSlots g = block_reserve<kShadeThreads / 64>(next_n, cont, sh_n, emit, s_off, s_base);
        const uint32_t cpos = g.ext, spos = g.shadow;
        if (cont) next_q[cpos] = path;
        if (emit) {
            P.q_sh[spos] = path;
            P.sh_o[spos] = s_shadow[threadIdx.x];
            P.sh_d[spos] = s_shadow[kShadeThreads + threadIdx.x];
            P.sh_c[spos] = s_shadow[2 * kShadeThreads + threadIdx.x];
        }
    }
    wave_add64(P.stats + kStatShadowRays, n_shadow);
    wave_add64(P.stats + kStatScatter, n_scatter);
    wave_add64(P.stats + kStatRoulette, n_rr);
}

// Any-hit for every shadow ray of this bounce (split traversal kernels).
__global__ void __launch_bounds__(256) k_shadow(Pool P, const GpuJob* __restrict__ jobs, uint32_t iter) {
    __shared__ uint32_t s_ref[kStackLds * kTraversalBlock];
    __shared__ float    s_t[kStackLds * kTraversalBlock];
    const StackView stk  = make_stack(s_ref, s_t, P);
    const uint32_t  n    = P.counters[kCtrShadow + iter];
    uint32_t*       work = P.counters + kCtrWorkShadow + iter;
    WaveWork<kShadowClaim> w;
    for (w.begin(); w.base < n; w.next(work, n)) {
        const uint32_t i = w.base + lane_id();
        if (i < n) shadow_one<false>(P, jobs, i, stk);
    }
}

// Film validation of a batch (Film::commit checks the whole batch before publishing any of it,
// render_session.cpp:39-51): nonfinite or negative sample radiance, or a pixel mean above FLT_MAX,
// flags the job; k_accumulate then leaves that job's film untouched. The validation already forms
// each pixel's new FP64 sums (film + samples in index order), so it stages them for k_accumulate in
// the batch's ray arrays, which no kernel reads again before the next batch's k_generate rewrites
// them: batch pixel g's sums go to ray_o[g] (x, y) and ray_d[g] (z) as doubles (a batch has at most
// one pixel per path slot, so g < capacity).
__global__ void __launch_bounds__(256) k_validate(Pool P, const GpuJob* __restrict__ jobs, const GpuChunk* __restrict__ chunks) {
    const uint32_t chunk_count  = P.header[kHdrChunks];
    const uint32_t total_pixels = P.header[kHdrPixels];
    for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < total_pixels; g += gridDim.x * blockDim.x) {
        const uint32_t ci = find_chunk_by_pixel(chunks, chunk_count, g);
        const GpuChunk ch = chunks[ci];
        const uint32_t p  = g - ch.pixel_prefix;
        const GpuJob&  job = jobs[ch.job];
        uint32_t       px, py;
        chunk_xy(ch, job.width, p, px, py);
        if (job.error_flag[2] != 0u) continue;  // cancelled during the batch: nothing of it is committed
        const uint32_t pixel = py * job.width + px;
        double sx = job.film_sums[static_cast<size_t>(pixel) * 3 + 0];
        double sy = job.film_sums[static_cast<size_t>(pixel) * 3 + 1];
        double sz = job.film_sums[static_cast<size_t>(pixel) * 3 + 2];
        uint32_t bad = 0;
        for (uint32_t s = 0; s < ch.sample_count; ++s) {
            const float4 v = P.rad[chunk_slot(ch, s, p)];
            if (!(__builtin_isfinite(v.x) && __builtin_isfinite(v.y) && __builtin_isfinite(v.z)) || v.x < 0.0f || v.y < 0.0f || v.z < 0.0f)
                bad |= 1u;
            sx = sx + static_cast<double>(v.x);
            sy = sy + static_cast<double>(v.y);
            sz = sz + static_cast<double>(v.z);
        }
        const double count = static_cast<double>(job.film_counts[pixel] + ch.sample_count);
        const double fm    = 3.4028234663852886e38;
        if (!(sx / count <= fm && sy / count <= fm && sz / count <= fm)) bad |= 2u;
        if (bad) atomicOr(job.error_flag, bad);
        reinterpret_cast<double2*>(P.ray_o)[g] = make_double2(sx, sy);
        reinterpret_cast<double2*>(P.ray_d)[g] = make_double2(sz, 0.0);
    }
}

// One word raised for kernels that are running (device-scope store).
__global__ void k_signal(uint32_t* word, uint32_t value) {
    if (threadIdx.x == 0) __hip_atomic_store(word, value, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Before a batch's film commit: each job's cancel word, frozen for validation and accumulation.
__global__ void k_commit_gate(const GpuJob* __restrict__ jobs, uint32_t job_count) {
    for (uint32_t i = threadIdx.x; i < job_count; i += blockDim.x) {
        uint32_t* f = jobs[i].error_flag;
        f[2] = __hip_atomic_load(f + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// Film commit: per pixel, samples in index order, FP64 sums (render_session.cpp:39-74), as k_validate
// formed and staged them.
__global__ void __launch_bounds__(256) k_accumulate(Pool P, const GpuJob* __restrict__ jobs, const GpuChunk* __restrict__ chunks) {
    const uint32_t chunk_count  = P.header[kHdrChunks];
    const uint32_t total_pixels = P.header[kHdrPixels];
    for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < total_pixels; g += gridDim.x * blockDim.x) {
        const uint32_t ci = find_chunk_by_pixel(chunks, chunk_count, g);
        const GpuChunk ch = chunks[ci];
        const uint32_t p  = g - ch.pixel_prefix;
        const GpuJob&  job = jobs[ch.job];
        if (job.error_flag[0] != 0u || job.error_flag[2] != 0u) continue;  // a failed or cancelled batch publishes nothing
        uint32_t       px, py;
        chunk_xy(ch, job.width, p, px, py);
        const uint32_t pixel = py * job.width + px;
        const double2  sxy   = reinterpret_cast<const double2*>(P.ray_o)[g];
        const double2  sz    = reinterpret_cast<const double2*>(P.ray_d)[g];
        job.film_sums[static_cast<size_t>(pixel) * 3 + 0] = sxy.x;
        job.film_sums[static_cast<size_t>(pixel) * 3 + 1] = sxy.y;
        job.film_sums[static_cast<size_t>(pixel) * 3 + 2] = sz.x;
        job.film_counts[pixel] = job.film_counts[pixel] + ch.sample_count;
    }
}

// ---------------------------------------------------------------------------------------------
// Host-side launchers.

static uint32_t grid_for(const LaunchShape& s, uint32_t per_cu) { return s.compute_units * per_cu; }

// The allocation-free query's workspace. Host code never passes `scratch` itself to a kernel or a copy
// (in host code the name is a stand-in, not the device address): it looks the address up once per
// device with noalloc_scratch_address and uses that pointer.
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

hipError_t noalloc_scratch_address(void** out) { return hipGetSymbolAddress(out, HIP_SYMBOL(scratch)); }

hipError_t launch_query(const GpuScene& sc, const float* rays, uint32_t n, float tmin, float tmax, mrg_hit_raw* hits, float* bary,
                        uint8_t* occluded, uint32_t* work, const Pool& P, uint64_t* stats, bool any, bool with_stats, bool exact,
                        const LaunchShape& shape, hipStream_t stream) {
    const dim3 grid(shape.query_blocks), block(256);
#define MRG_QUERY(A, S, E) hipLaunchKernelGGL((E ? k_query_exact<A, S> : k_query<A, S>), grid, block, 0, stream, sc, rays, n, tmin, tmax, hits, bary, occluded, work, P, stats)
    if (exact) {
        if (any) {
            if (with_stats) MRG_QUERY(true, true, true);
            else MRG_QUERY(true, false, true);
        } else {
            if (with_stats) MRG_QUERY(false, true, true);
            else MRG_QUERY(false, false, true);
        }
    } else {
        if (any) {
            if (with_stats) MRG_QUERY(true, true, false);
            else MRG_QUERY(true, false, false);
        } else {
            if (with_stats) MRG_QUERY(false, true, false);
            else MRG_QUERY(false, false, false);
        }
    }
#undef MRG_QUERY
    return hipGetLastError();
}

hipError_t launch_reset(const Pool& P, hipStream_t stream) {
    hipLaunchKernelGGL(k_reset, dim3(4), dim3(256), 0, stream, P);
    return hipGetLastError();
}

hipError_t launch_generate(const Pool& P, const GpuJob* jobs, const GpuChunk* chunks, const LaunchShape& shape, hipStream_t stream) {
    hipLaunchKernelGGL(k_generate, dim3(grid_for(shape, 8)), dim3(256), 0, stream, P, jobs, chunks);
    return hipGetLastError();
}

hipError_t launch_extend(const Pool& P, const GpuJob* jobs, uint32_t iter, const LaunchShape& shape, hipStream_t stream) {
    hipLaunchKernelGGL(k_extend, dim3(shape.traversal_blocks), dim3(256), 0, stream, P, jobs, iter);
    return hipGetLastError();
}

hipError_t launch_shade(const Pool& P, const GpuJob* jobs, uint32_t iter, bool extended, bool one_scene, const LaunchShape& shape,
                        hipStream_t stream) {
    const dim3 grid(shape.shade_blocks * 256 / kShadeThreads), block(kShadeThreads);
    if (extended) {
        if (one_scene) hipLaunchKernelGGL((k_shade<true, true>), grid, block, 0, stream, P, jobs, iter);
        else hipLaunchKernelGGL((k_shade<true, false>), grid, block, 0, stream, P, jobs, iter);
    } else {
        if (one_scene) hipLaunchKernelGGL((k_shade<false, true>), grid, block, 0, stream, P, jobs, iter);
        else hipLaunchKernelGGL((k_shade<false, false>), grid, block, 0, stream, P, jobs, iter);
    }
    return hipGetLastError();
}

hipError_t launch_trace(const Pool& P, const GpuJob* jobs, uint32_t iter, bool one_scene, const LaunchShape& shape,
                        hipStream_t stream) {
    if (one_scene) hipLaunchKernelGGL(k_trace<true>, dim3(shape.traversal_blocks), dim3(256), 0, stream, P, jobs, iter);
    else hipLaunchKernelGGL(k_trace<false>, dim3(shape.traversal_blocks), dim3(256), 0, stream, P, jobs, iter);
    return hipGetLastError();
}

hipError_t launch_trace_refill(const Pool& P, const GpuJob* jobs, uint32_t iter, bool one_scene, const LaunchShape& shape,
                               hipStream_t stream) {
    if (one_scene) hipLaunchKernelGGL(k_trace_refill<true>, dim3(shape.refill_blocks), dim3(256), 0, stream, P, jobs, iter);
    else hipLaunchKernelGGL(k_trace_refill<false>, dim3(shape.refill_blocks), dim3(256), 0, stream, P, jobs, iter);
    return hipGetLastError();
}

hipError_t launch_shadow(const Pool& P, const GpuJob* jobs, uint32_t iter, const LaunchShape& shape, hipStream_t stream) {
    hipLaunchKernelGGL(k_shadow, dim3(shape.traversal_blocks), dim3(256), 0, stream, P, jobs, iter);
    return hipGetLastError();
}

hipError_t launch_validate(const Pool& P, const GpuJob* jobs, const GpuChunk* chunks, const LaunchShape& shape, hipStream_t stream) {
    hipLaunchKernelGGL(k_validate, dim3(grid_for(shape, 4)), dim3(256), 0, stream, P, jobs, chunks);
    return hipGetLastError();
}

hipError_t launch_accumulate(const Pool& P, const GpuJob* jobs, const GpuChunk* chunks, const LaunchShape& shape, hipStream_t stream) {
    hipLaunchKernelGGL(k_accumulate, dim3(grid_for(shape, 4)), dim3(256), 0, stream, P, jobs, chunks);
    return hipGetLastError();
}

hipError_t launch_signal(uint32_t* word, uint32_t value, hipStream_t stream) {
    hipLaunchKernelGGL(k_signal, dim3(1), dim3(64), 0, stream, word, value);
    return hipGetLastError();
}

hipError_t launch_commit_gate(const GpuJob* jobs, uint32_t job_count, hipStream_t stream) {
    hipLaunchKernelGGL(k_commit_gate, dim3(1), dim3(64), 0, stream, jobs, job_count);
    return hipGetLastError();
}

}  // namespace mrg
