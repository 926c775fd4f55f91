// Achievable 32-bit integer add rate of the GPU, the denominator of the
// channel sum's roofline (its algorithmic unit is one integer add per DM,
// sample and channel). Not part of libdedisp.so: bench.py loads it from
// libdd_peaks.so to report the measured peak beside the nominal one.
//
// Every thread keeps NACC accumulators and updates each one per step with a
// plain v_add_u32 from another accumulator (a_i += a_{i+1 mod NACC}): a linear
// recurrence the compiler cannot fold into fewer adds, with NACC independent
// adds per step. The grid holds 8 waves on every SIMD of 256 CUs.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

namespace {

constexpr int NACC = 8;
constexpr int TPB = 256;

__global__ __launch_bounds__(TPB) void int_add_k(uint32_t iters, uint32_t seed, uint32_t* __restrict__ out)
{
    uint32_t a[NACC];
#pragma unroll
    for (int i = 0; i < NACC; ++i) a[i] = seed + threadIdx.x * (uint32_t)(2 * i + 1);
#pragma unroll 4
    for (uint32_t n = 0; n < iters; ++n) {
#pragma unroll
        for (int i = 0; i < NACC; ++i) a[i] += a[(i + 1) % NACC];
    }
    uint32_t s = 0;
#pragma unroll
    for (int i = 0; i < NACC; ++i) s ^= a[i];
    out[(size_t)blockIdx.x * TPB + threadIdx.x] = s;
}

}  // namespace

// Runs int_add_k reps times over nblocks workgroups of 256 threads, iters
// steps each (iters a multiple of 4), and returns the fastest run as
// lane-adds/s and wave-instructions/s (one v_add_u32 per wave and add).
// Returns 0, or the HIP error code.
extern "C" int dd_peak_int_add(uint32_t iters, uint32_t nblocks, int reps, double* lane_adds_per_s,
                               double* wave_insts_per_s)
{
    if (iters == 0 || iters % 4 != 0 || nblocks == 0 || reps < 1) return (int)hipErrorInvalidValue;
    uint32_t* out = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t err = hipMalloc(&out, (size_t)nblocks * TPB * sizeof(uint32_t));
    if (err == hipSuccess) err = hipEventCreate(&e0);
    if (err == hipSuccess) err = hipEventCreate(&e1);
    float best = 0.f;
    for (int r = 0; err == hipSuccess && r <= reps; ++r) {  // run 0 warms up
        hipLaunchKernelGGL(int_add_k, dim3(nblocks), dim3(TPB), 0, nullptr, iters, (uint32_t)r, out);
        err = hipGetLastError();
        if (err == hipSuccess) err = hipEventRecord(e0, nullptr);
        if (err != hipSuccess) break;
        hipLaunchKernelGGL(int_add_k, dim3(nblocks), dim3(TPB), 0, nullptr, iters, (uint32_t)r + 1u, out);
        err = hipGetLastError();
        if (err == hipSuccess) err = hipEventRecord(e1, nullptr);
        if (err == hipSuccess) err = hipEventSynchronize(e1);
        float ms = 0.f;
        if (err == hipSuccess) err = hipEventElapsedTime(&ms, e0, e1);
        if (err == hipSuccess && r > 0) best = (best == 0.f) ? ms : std::min(best, ms);
    }
    if (err == hipSuccess) {
        const double adds = (double)nblocks * TPB * iters * NACC;
        *lane_adds_per_s = adds / (best * 1e-3);
        *wave_insts_per_s = adds / 64.0 / (best * 1e-3);
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (out) (void)hipFree(out);
    return (int)err;
}
