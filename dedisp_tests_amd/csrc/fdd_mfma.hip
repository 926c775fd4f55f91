// gfx950 matrix-core phasor contraction of the FDD path.
//
// For every Fourier bin k the dedispersed spectra are a complex GEMM
//   Y_k[dm][chunk] = sum_c P_k[dm][c] X_k[c][chunk],  P_k[dm][c] = e^{+2 pi i k d(dm,c)/nfft}
// (the phase rotation of src/fdd/dedisperse/fdd_kernel.cuh:226-235 with the
// truncated integer delay d). A wave owns a 32 (DM) x 32 (chunk) tile of NB
// consecutive bins and walks the channels 16 at a time:
//   A operand: phasors generated in registers — exact phase m = (k*d) mod nfft,
//              e^{2 pi i m/nfft} = T_hi[m>>8] * T_lo[m & 255] from two 256-entry
//              LDS tables (built in double precision on the host);
//   B operand: X_k[c][chunk] staged in LDS for the workgroup's bins, formed on
//              load from the packed forward transform (real-input pairing);
//   complex MAC as four real MFMAs, Yr += Pr Xr - Pi Xi, Yi += Pr Xi + Pi Xr.
// Forms: float32 (v_mfma_f32_32x32x2_f32, 8 MFMAs per real product and
// K-step), bfloat16 (v_mfma_f32_32x32x16_bf16, phasors and spectra split into
// hi + lo, three products Ph Xh + Ph Xl + Pl Xh), float16
// (v_mfma_f32_32x32x16_f16, one product). Float32 accumulation throughout.
// The k-positions of a K-step are assigned the same way in A and B (lane group
// g = lane/32, element s -> channel 8g+s), so the sum runs over channels
// whatever the hardware's k order.
#include <hip/hip_runtime.h>

#include "fdd_kernels.h"

namespace {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(8 * sizeof(_Float16)))) _Float16 f16x8;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float f32x16;

constexpr int WAVES = 8;   // waves per workgroup, one 32-DM tile each
constexpr int NB = 4;      // bins per workgroup
constexpr int KS = 16;     // channels per K-step
constexpr int NJ = 32;     // chunks per tile
constexpr int NX = NB * NJ * KS;  // staged X values per K-step

__device__ __forceinline__ float2 cmul(float2 a, float2 b)
{
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// X[k] of the real series from the packed transform (see fft_kernels.hip).
__device__ __forceinline__ float2 unpack_bin(const float2* __restrict__ zrow, uint32_t k, uint32_t M,
                                             float2 wk /* e^{-2 pi i k/nfft} */)
{
    const float2 a = zrow[k & (M - 1)];
    const float2 bz = zrow[(M - k) & (M - 1)];
    const float2 e = make_float2(0.5f * (a.x + bz.x), 0.5f * (a.y - bz.y));
    const float2 d = make_float2(a.x - bz.x, a.y + bz.y);
    const float2 o = make_float2(0.5f * d.y, -0.5f * d.x);
    const float2 r = cmul(o, wk);
    return make_float2(e.x + r.x, e.y + r.y);
}

__device__ __forceinline__ uint16_t bf16_trunc_bits(float x)
{
    return (uint16_t)(__builtin_bit_cast(uint32_t, x) >> 16);
}
__device__ __forceinline__ float bf16_bits_to_float(uint16_t h)
{
    return __builtin_bit_cast(float, (uint32_t)h << 16);
}
__device__ __forceinline__ uint16_t bf16_rne_bits(float x)
{
    return __builtin_bit_cast(uint16_t, (__bf16)x);
}
__device__ __forceinline__ uint16_t f16_bits(float x)
{
    return __builtin_bit_cast(uint16_t, (_Float16)x);
}

template <int PREC>
__global__ __launch_bounds__(WAVES * 64) void fdd_mfma_k(
    const float2* __restrict__ Z, size_t Zp, const int32_t* __restrict__ delays, uint32_t ndm,
    uint32_t nchan, uint32_t nchunk, const float2* __restrict__ ph, uint32_t nfft,
    float2* __restrict__ Y, size_t Yp, uint32_t nbb, uint32_t ndb)
{
    // X staging planes: fp32 -> 2 planes of float; bf16 -> 4 planes of
    // 16-bit (re hi, re lo, im hi, im lo); fp16 -> 2 planes of 16-bit.
    constexpr int PLANES = (PREC == 1) ? 4 : 2;
    constexpr int ELEM = (PREC == 0) ? 4 : 2;
    __shared__ __attribute__((aligned(16))) unsigned char xs_raw[PLANES * NX * ELEM];
    __shared__ float2 tw_hi[256], tw_lo[256];

    const uint32_t M = nfft / 2;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const uint32_t wave = __builtin_amdgcn_readfirstlane(tid >> 6);

    // XCD-aware order: consecutive workgroups of an XCD share a bin block
    // (DM block fastest), so the staged spectra come from that XCD's L2.
    const uint32_t nb = gridDim.x, bid = blockIdx.x;
    const uint32_t xcd = bid & 7u, per = nb >> 3, rem = nb & 7u;
    const uint32_t L = xcd * per + min(xcd, rem) + (bid >> 3);
    const uint32_t dblk = L % ndb;
    const uint32_t rest = L / ndb;
    const uint32_t bblk = rest % nbb;
    const uint32_t cgrp = rest / nbb;
    const uint32_t k0 = bblk * NB;
    const uint32_t ch0 = cgrp * NJ;

    for (int t = tid; t < 256; t += WAVES * 64) {
        tw_lo[t] = ph[t & (nfft - 1)];
        tw_hi[t] = ph[(t * 256) & (nfft - 1)];
    }

    const int i = lane & 31, g = lane >> 5;
    const uint32_t dm = dblk * (WAVES * 32) + wave * 32 + i;
    const int32_t* drow = delays + (size_t)min(dm, ndm - 1) * nchan;

    f32x16 acc_r[NB], acc_i[NB];
#pragma unroll
    for (int b = 0; b < NB; ++b)
#pragma unroll
        for (int r = 0; r < 16; ++r) acc_r[b][r] = acc_i[b][r] = 0.f;

    // e^{-2 pi i k/nfft} of the bins this thread stages (bin = tid % NB)
    const uint32_t kst = k0 + (tid % NB);
    const float2 pkst = ph[kst & (nfft - 1)];
    const float2 wst = make_float2(pkst.x, -pkst.y);

    const uint32_t nks = (nchan + KS - 1) / KS;
    for (uint32_t ks = 0; ks < nks; ++ks) {
        const uint32_t c0 = ks * KS;
        __syncthreads();
        // ---- stage X for NB bins x NJ chunks x KS channels ----
#pragma unroll
        for (int r = 0; r < NX / (WAVES * 64); ++r) {
            const int e = tid + r * WAVES * 64;
            const int b = e % NB, j = (e / NB) % NJ, cl = e / (NB * NJ);
            const uint32_t c = c0 + cl, ch = ch0 + j;
            float2 x = make_float2(0.f, 0.f);
            if (c < nchan && ch < nchunk && kst <= M)
                x = unpack_bin(Z + ((size_t)c * nchunk + ch) * Zp, kst, M, wst);
            const int idx = (b * NJ + j) * KS + cl;
            if constexpr (PREC == 0) {
                float* xs = reinterpret_cast<float*>(xs_raw);
                xs[idx] = x.x;
                xs[NX + idx] = x.y;
            } else if constexpr (PREC == 1) {
                uint16_t* xs = reinterpret_cast<uint16_t*>(xs_raw);
                const uint16_t rh = bf16_trunc_bits(x.x), ih = bf16_trunc_bits(x.y);
                xs[idx] = rh;
                xs[NX + idx] = bf16_rne_bits(x.x - bf16_bits_to_float(rh));
                xs[2 * NX + idx] = ih;
                xs[3 * NX + idx] = bf16_rne_bits(x.y - bf16_bits_to_float(ih));
            } else {
                uint16_t* xs = reinterpret_cast<uint16_t*>(xs_raw);
                xs[idx] = f16_bits(x.x);
                xs[NX + idx] = f16_bits(x.y);
            }
        }
        __syncthreads();

        // ---- delays of this lane's (DM, 8 channels) ----
        int32_t d[8];
#pragma unroll
        for (int s = 0; s < 8; ++s) {
            const uint32_t c = c0 + 8 * g + s;
            d[s] = (c < nchan) ? drow[c] : 0;
        }

#pragma unroll
        for (int b = 0; b < NB; ++b) {
            const uint32_t k = k0 + b;
            float2 p[8];
#pragma unroll
            for (int s = 0; s < 8; ++s) {
                const uint32_t m = (k * (uint32_t)d[s]) & (nfft - 1);
                p[s] = cmul(tw_hi[m >> 8], tw_lo[m & 255]);
            }
            const int base = (b * NJ + i) * KS + 8 * g;  // B operand: chunk j = lane%32
            if constexpr (PREC == 0) {
                const float* xs = reinterpret_cast<const float*>(xs_raw);
#pragma unroll
                for (int s = 0; s < 8; ++s) {
                    const float xr = xs[base + s], xi = xs[NX + base + s];
                    acc_r[b] = __builtin_amdgcn_mfma_f32_32x32x2f32(p[s].x, xr, acc_r[b], 0, 0, 0);
                    acc_r[b] = __builtin_amdgcn_mfma_f32_32x32x2f32(-p[s].y, xi, acc_r[b], 0, 0, 0);
                    acc_i[b] = __builtin_amdgcn_mfma_f32_32x32x2f32(p[s].x, xi, acc_i[b], 0, 0, 0);
                    acc_i[b] = __builtin_amdgcn_mfma_f32_32x32x2f32(p[s].y, xr, acc_i[b], 0, 0, 0);
                }
            } else if constexpr (PREC == 1) {
                const uint16_t* xs = reinterpret_cast<const uint16_t*>(xs_raw);
                const bf16x8 xrh = *reinterpret_cast<const bf16x8*>(xs + base);
                const bf16x8 xrl = *reinterpret_cast<const bf16x8*>(xs + NX + base);
                const bf16x8 xih = *reinterpret_cast<const bf16x8*>(xs + 2 * NX + base);
                const bf16x8 xil = *reinterpret_cast<const bf16x8*>(xs + 3 * NX + base);
                bf16x8 prh, prl, pih, pil, nih, nil;
#pragma unroll
                for (int s = 0; s < 8; ++s) {
                    const uint16_t rh = bf16_trunc_bits(p[s].x), ih = bf16_trunc_bits(p[s].y);
                    const uint16_t rl = bf16_rne_bits(p[s].x - bf16_bits_to_float(rh));
                    const uint16_t il = bf16_rne_bits(p[s].y - bf16_bits_to_float(ih));
                    prh[s] = __builtin_bit_cast(__bf16, rh);
                    prl[s] = __builtin_bit_cast(__bf16, rl);
                    pih[s] = __builtin_bit_cast(__bf16, ih);
                    pil[s] = __builtin_bit_cast(__bf16, il);
                    nih[s] = __builtin_bit_cast(__bf16, (uint16_t)(ih ^ 0x8000u));
                    nil[s] = __builtin_bit_cast(__bf16, (uint16_t)(il ^ 0x8000u));
                }
                f32x16 ar = acc_r[b], ai = acc_i[b];
                ar = __builtin_amdgcn_mfma_f32_32x32x16_bf16(prh, xrh, ar, 0, 0, 0);
                ar = __builtin_amdgcn_mfma_f32_32x32x16_bf16(prh, xrl, ar, 0, 0, 0);
                ar = __builtin_amdgcn_mfma_f32_32x32x16_bf16(prl, xrh, ar, 0, 0, 0);
                ar = __builtin_amdgcn_mfma_f32_32x32x16_bf16(nih, xih, ar, 0, 0, 0);
                ar = __builtin_amdgcn_mfma_f32_32x32x16_bf16(nih, xil, ar, 0, 0, 0);
                ar = __builtin_amdgcn_mfma_f32_32x32x16_bf16(nil, xih, ar, 0, 0, 0);
                ai = __builtin_amdgcn_mfma_f32_32x32x16_bf16(prh, xih, ai, 0, 0, 0);
                ai = __builtin_amdgcn_mfma_f32_32x32x16_bf16(prh, xil, ai, 0, 0, 0);
                ai = __builtin_amdgcn_mfma_f32_32x32x16_bf16(prl, xih, ai, 0, 0, 0);
                ai = __builtin_amdgcn_mfma_f32_32x32x16_bf16(pih, xrh, ai, 0, 0, 0);
                ai = __builtin_amdgcn_mfma_f32_32x32x16_bf16(pih, xrl, ai, 0, 0, 0);
                ai = __builtin_amdgcn_mfma_f32_32x32x16_bf16(pil, xrh, ai, 0, 0, 0);
                acc_r[b] = ar;
                acc_i[b] = ai;
            } else {
                const uint16_t* xs = reinterpret_cast<const uint16_t*>(xs_raw);
                const f16x8 xr = *reinterpret_cast<const f16x8*>(xs + base);
                const f16x8 xi = *reinterpret_cast<const f16x8*>(xs + NX + base);
                f16x8 pr, pi, ni;
#pragma unroll
                for (int s = 0; s < 8; ++s) {
                    pr[s] = (_Float16)p[s].x;
                    pi[s] = (_Float16)p[s].y;
                    ni[s] = -pi[s];
                }
                acc_r[b] = __builtin_amdgcn_mfma_f32_32x32x16_f16(pr, xr, acc_r[b], 0, 0, 0);
                acc_r[b] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ni, xi, acc_r[b], 0, 0, 0);
                acc_i[b] = __builtin_amdgcn_mfma_f32_32x32x16_f16(pr, xi, acc_i[b], 0, 0, 0);
                acc_i[b] = __builtin_amdgcn_mfma_f32_32x32x16_f16(pi, xr, acc_i[b], 0, 0, 0);
            }
        }
    }

    // ---- epilogue: 32x32 accumulator layout, column j = lane%32,
    // row = 8*(r/4) + 4*(lane/32) + r%4 for element r ----
    const uint32_t ch = ch0 + i;
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        const uint32_t k = k0 + b;
        if (k > M || ch >= nchunk) continue;
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const uint32_t row = 8 * (r / 4) + 4 * g + (r % 4);
            const uint32_t dmo = dblk * (WAVES * 32) + wave * 32 + row;
            if (dmo < ndm) Y[((size_t)dmo * nchunk + ch) * Yp + k] = make_float2(acc_r[b][r], acc_i[b][r]);
        }
    }
}

}  // namespace

extern "C" hipError_t dd_fdd_contract_mfma(int prec, const float2* Z, size_t Zp, const int32_t* delays,
                                           size_t ndm, size_t nchan, size_t nchunk, const float2* ph,
                                           size_t nfft, float2* Y, size_t Yp, hipStream_t stream)
{
    if (ndm == 0 || nchunk == 0) return hipSuccess;
    const size_t M = nfft / 2;
    const uint32_t nbb = (uint32_t)((M + 1 + NB - 1) / NB);
    const uint32_t ndb = (uint32_t)((ndm + WAVES * 32 - 1) / (WAVES * 32));
    const uint32_t ncg = (uint32_t)((nchunk + NJ - 1) / NJ);
    const size_t nblocks = (size_t)nbb * ndb * ncg;
    if (nblocks >= (size_t(1) << 31)) return hipErrorInvalidValue;
#define DD_MFMA(P)                                                                                \
    hipLaunchKernelGGL(fdd_mfma_k<P>, dim3((unsigned)nblocks), dim3(WAVES * 64), 0, stream, Z, Zp, \
                       delays, (uint32_t)ndm, (uint32_t)nchan, (uint32_t)nchunk, ph, (uint32_t)nfft, \
                       Y, Yp, nbb, ndb)
    switch (prec) {
    case 0: DD_MFMA(0); break;
    case 1: DD_MFMA(1); break;
    case 2: DD_MFMA(2); break;
    default: return hipErrorInvalidValue;
    }
#undef DD_MFMA
    return hipGetLastError();
}
