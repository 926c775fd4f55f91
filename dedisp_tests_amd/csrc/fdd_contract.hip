// gfx950 phasor contraction of the FDD path: for every Fourier bin k of a
// chunked transform, Y_k[dm][chunk] = sum_c P_k[dm][c] X_k[c][chunk] with
// P_k[dm][c] = e^{+2 pi i k d(dm,c) / nfft} (the phase rotation of
// src/fdd/dedisperse/fdd_kernel.cuh:226-235, d the truncated integer delay).
// The phase is reduced exactly in integers, (k*d) mod nfft, and looked up in a
// table built in double precision, so phasors carry only float32 rounding.
#include <hip/hip_runtime.h>

#include "fdd_kernels.h"

namespace {

// Lane = bin, wave = DPW DMs, CG chunks per pass over the channels.
template <int DPW, int CG>
__global__ __launch_bounds__(256) void fdd_contract_fp32_k(
    const float2* __restrict__ X, size_t Xp, const int32_t* __restrict__ delays, uint32_t ndm,
    uint32_t nchan, uint32_t nchunk, const float2* __restrict__ ph, uint32_t nfft,
    float2* __restrict__ Y, size_t Yp)
{
    const uint32_t M = nfft / 2;
    const uint32_t k = blockIdx.x * 64 + (threadIdx.x & 63);
    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint32_t dm0 = blockIdx.y * (4 * DPW) + wave * DPW;
    if (k > M) return;
    uint32_t drow[DPW];
#pragma unroll
    for (int i = 0; i < DPW; ++i) drow[i] = min(dm0 + i, ndm - 1) * nchan;

    for (uint32_t g0 = 0; g0 < nchunk; g0 += CG) {
        float2 acc[DPW][CG];
#pragma unroll
        for (int i = 0; i < DPW; ++i)
#pragma unroll
            for (int j = 0; j < CG; ++j) acc[i][j] = make_float2(0.f, 0.f);
        for (uint32_t c = 0; c < nchan; ++c) {
            float2 x[CG];
#pragma unroll
            for (int j = 0; j < CG; ++j) {
                const uint32_t ch = min(g0 + j, nchunk - 1);
                x[j] = X[((size_t)c * nchunk + ch) * Xp + k];
            }
#pragma unroll
            for (int i = 0; i < DPW; ++i) {
                const int32_t d = delays[drow[i] + c];
                const float2 p = ph[((uint32_t)k * (uint32_t)d) & (nfft - 1)];
#pragma unroll
                for (int j = 0; j < CG; ++j) {
                    acc[i][j].x = fmaf(p.x, x[j].x, fmaf(-p.y, x[j].y, acc[i][j].x));
                    acc[i][j].y = fmaf(p.x, x[j].y, fmaf(p.y, x[j].x, acc[i][j].y));
                }
            }
        }
#pragma unroll
        for (int i = 0; i < DPW; ++i) {
            if (dm0 + i >= ndm) continue;
#pragma unroll
            for (int j = 0; j < CG; ++j)
                if (g0 + j < nchunk) Y[((size_t)(dm0 + i) * nchunk + g0 + j) * Yp + k] = acc[i][j];
        }
    }
}

}  // namespace

extern "C" hipError_t dd_fdd_contract_fp32(const float2* X, size_t Xp, const int32_t* delays, size_t ndm,
                                           size_t nchan, size_t nchunk, const float2* ph, size_t nfft,
                                           float2* Y, size_t Yp, hipStream_t stream)
{
    constexpr int DPW = 4, CG = 8;
    const size_t M = nfft / 2;
    dim3 grid((unsigned)((M + 1 + 63) / 64), (unsigned)((ndm + 4 * DPW - 1) / (4 * DPW)));
    hipLaunchKernelGGL((fdd_contract_fp32_k<DPW, CG>), grid, dim3(256), 0, stream, X, Xp, delays,
                       (uint32_t)ndm, (uint32_t)nchan, (uint32_t)nchunk, ph, (uint32_t)nfft, Y, Yp);
    return hipGetLastError();
}
