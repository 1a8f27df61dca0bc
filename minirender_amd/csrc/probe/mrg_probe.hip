// Cache-hierarchy read-bandwidth probe for gfx950 (measurement only; no part of the render path).
//
// k_read streams a resident buffer: each wave sweeps the whole buffer in consecutive 1 KB rows (its
// 64 lanes read one float4 each per row, four rows in flight), starting at a wave-specific row and
// wrapping around, and writes one checksum per thread so the loads cannot be removed. A wave's sweep
// does not fit the 32 KB vector L1, so with a buffer that fits each XCD's L2 (4 MB) every read
// after the first sweep is an L2 hit; a buffer that fits the Infinity Cache but not the L2
// measures that level; one far larger than both measures HBM. The host entry point times the
// kernel with HIP events and reports bytes read per second.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>

namespace {

__global__ void __launch_bounds__(256) k_read(const float4* __restrict__ buf, uint32_t rows, uint32_t iters,
                                              float* __restrict__ out) {
    const uint32_t tid  = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t wave = tid >> 6;
    uint32_t       row  = (wave * 4099u) % rows;  // rows of 64 float4; n4 = 64 * rows
    float4         acc  = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    for (uint32_t it = 0; it < iters; it += 4) {
        float4 v[4];
        for (int k = 0; k < 4; ++k) {
            uint32_t r = row + static_cast<uint32_t>(k);
            if (r >= rows) r -= rows;
            v[k] = buf[static_cast<uint64_t>(r) * 64u + lane];
        }
        for (int k = 0; k < 4; ++k) {
            acc.x += v[k].x;
            acc.y += v[k].y;
            acc.z += v[k].z;
            acc.w += v[k].w;
        }
        row += 4u;
        if (row >= rows) row -= rows;
    }
    out[tid] = (acc.x + acc.y) + (acc.z + acc.w);
}

}  // namespace

// Reads a buffer of `bytes` (rounded down to 1 KB, at least 4 KB) on device `ordinal` with `blocks_per_cu`
// 256-thread blocks per compute unit, `iters` float4 loads per thread (a multiple of 4), `repeats`
// timed launches after one warm-up launch. Writes the read bandwidth in GB/s (1e9 bytes per
// second), the mean launch time and the bytes read per launch. Returns 0, or the HIP error code.
extern "C" int mrg_probe_read_bandwidth(int ordinal, uint64_t bytes, uint32_t blocks_per_cu, uint32_t iters, uint32_t repeats,
                                        double* gb_per_s, double* ms_per_launch, uint64_t* bytes_per_launch) {
    const uint64_t rows = bytes / 1024;  // 1 KB rows (64 float4)
    if (rows < 4 || rows > 0xffffffffull || blocks_per_cu == 0 || blocks_per_cu > 64 || iters == 0 || iters % 4 != 0 ||
        repeats == 0 || !gb_per_s || !ms_per_launch || !bytes_per_launch)
        return static_cast<int>(hipErrorInvalidValue);
    float4*     buf = nullptr;
    float*      out = nullptr;
    hipEvent_t  e0 = nullptr, e1 = nullptr;
    int         cus = 0;
    hipError_t  rc  = hipSetDevice(ordinal);
    if (rc == hipSuccess) rc = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ordinal);
    if (rc == hipSuccess && cus <= 0) rc = hipErrorInvalidDevice;
    if (rc == hipSuccess) rc = hipMalloc(&buf, rows * 1024);
    const uint32_t blocks  = static_cast<uint32_t>(cus > 0 ? cus : 1) * blocks_per_cu;
    const uint32_t threads = blocks * 256u;
    if (rc == hipSuccess) rc = hipMalloc(&out, static_cast<size_t>(threads) * 4);
    if (rc == hipSuccess) rc = hipMemset(buf, 0, rows * 1024);
    if (rc == hipSuccess) rc = hipEventCreate(&e0);
    if (rc == hipSuccess) rc = hipEventCreate(&e1);
    if (rc == hipSuccess) {
        hipLaunchKernelGGL(k_read, dim3(blocks), dim3(256), 0, nullptr, buf, static_cast<uint32_t>(rows), iters, out);
        rc = hipGetLastError();
    }
    if (rc == hipSuccess) rc = hipEventRecord(e0, nullptr);
    for (uint32_t r = 0; rc == hipSuccess && r < repeats; ++r) {
        hipLaunchKernelGGL(k_read, dim3(blocks), dim3(256), 0, nullptr, buf, static_cast<uint32_t>(rows), iters, out);
        rc = hipGetLastError();
    }
    if (rc == hipSuccess) rc = hipEventRecord(e1, nullptr);
    if (rc == hipSuccess) rc = hipEventSynchronize(e1);
    float ms = 0.0f;
    if (rc == hipSuccess) rc = hipEventElapsedTime(&ms, e0, e1);
    if (rc == hipSuccess) {
        const double per = static_cast<double>(ms) / repeats;
        *ms_per_launch    = per;
        *bytes_per_launch = static_cast<uint64_t>(threads) * iters * 16u;
        *gb_per_s         = static_cast<double>(*bytes_per_launch) / (per * 1e-3) / 1e9;
    }
    if (rc != hipSuccess) {
        int count = -1;
        (void)hipGetDeviceCount(&count);
        std::fprintf(stderr, "mrg_probe_read_bandwidth: %s (%d devices visible)\n", hipGetErrorString(rc), count);
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (out) (void)hipFree(out);
    if (buf) (void)hipFree(buf);
    return static_cast<int>(rc);
}
