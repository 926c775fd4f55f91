// gfx950 matrix-core phasor contraction of the FDD path.
//
// For every Fourier bin k the dedispersed spectra are a complex GEMM
//   Y_k[dm][chunk] = sum_c P_k[dm][c] X_k[c][chunk],  P_k[dm][c] = e^{+2 pi i k d(dm,c)/nfft}
// (the phase rotation of src/fdd/dedisperse/fdd_kernel.cuh:226-235 with the
// truncated integer delay d). A workgroup owns NB consecutive bins (4; 2 for
// the float32 form), 32 chunks and WAVES*32 DMs; a wave owns a 32 (DM) x 32 (chunk) tile of every bin and
// walks the channels 16 at a time (one K-step).
//
// Phasors (A operand), generated in registers:
//   the phase of bin k is the exact integer (k*q) mod nfft with
//   q = d mod nfft per (DM, channel) (uint32), in units of 1/nfft; per (DM,
//   channel) and workgroup P at the first bin and the step W = e^{2 pi i q/nfft}
//   come from the hardware cosine and sine (v_cos_f32 / v_sin_f32 take the
//   angle in revolutions, and m/nfft is exact in float for nfft <= 2^24); the
//   following bins are P <- P*W (NB-1 complex products). An earlier form read
//   P and W as T_hi[m>>8] * T_lo[m&255] from two 256-entry LDS tables built in
//   double: the same accuracy at the oracle (FP32 rel-L2 6.5e-7 vs 6.7e-7) but
//   slower (bf16 59.9 vs 56.9 ms, fp16 36.6 vs 34.4 ms), the random-address
//   table reads costing LDS time and bank conflicts.
// Spectra (B operand): X_k[c][chunk] for the K-step, staged in LDS in the
//   planes of the form (float32: re, im - re, -(re + im); bfloat16: re hi, re lo, im hi,
//   im lo and the negated im planes; float16: re, im, -im), formed on load from
//   the real spectra of the forward transform. The next K-step's
//   spectra and delays are loaded into registers while the current one is
//   multiplied; one barrier per K-step (two LDS buffers).
// Complex MAC as real MFMAs: Yr += Pr Xr + Pi (-Xi), Yi += Pr Xi + Pi Xr;
// the float32 form uses three products (Gauss): K = (Pr + Pi) Xr,
// A = Pi (-(Xr + Xi)), B = Pr (Xi - Xr), Yr = K + A, Yi = K + B, each summed in
// its own accumulators (three sets, hence two bins per workgroup; with
// DD_MFMA_GAUSS=0 it runs the four-product form on four bins).
// Forms: float32 (v_mfma_f32_32x32x2_f32), bfloat16 (v_mfma_f32_32x32x16_bf16,
// phasors and spectra split into truncated hi + rounded lo, products
// Ph Xh + Ph Xl + Pl Xh), float16 (v_mfma_f32_32x32x16_f16, one product,
// spectra scaled by a power of two into the float16 range and scaled back in
// the epilogue). Float32 accumulation throughout.
// Operand layout: the k-positions of a K-step are assigned the same way in A
// and B (lane group g = lane/32, element s -> channel 8g+s), so the sum runs
// over channels whatever the hardware's k order. Accumulator element r of lane
// l is row 8*(r/4) + 4*(l/32) + r%4, column l%32.
#include <hip/hip_runtime.h>
#include <type_traits>

#include "fdd_kernels.h"

namespace {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(2 * sizeof(__bf16)))) __bf16 bf16x2;
typedef __attribute__((__vector_size__(8 * sizeof(_Float16)))) _Float16 f16x8;
typedef __attribute__((__vector_size__(2 * sizeof(_Float16)))) _Float16 f16x2;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float f32x16;
typedef __attribute__((__vector_size__(4 * sizeof(uint32_t)))) uint32_t u32x4;

#ifndef DD_MFMA_NEGA
#define DD_MFMA_NEGA 1
#endif
#ifndef DD_MFMA_BPREF
#define DD_MFMA_BPREF 1
#endif
#ifndef DD_MFMA_STAGGER
#define DD_MFMA_STAGGER 1
#endif

#ifndef DD_MFMA_WAVES
#define DD_MFMA_WAVES 8
#endif
// float32 form: complex products with three real products (Gauss), two bins
// per workgroup (three accumulator sets)
#ifndef DD_MFMA_GAUSS
#define DD_MFMA_GAUSS 1
#endif

constexpr int WAVES = DD_MFMA_WAVES;   // waves per workgroup, one 32-DM tile each
#ifndef DD_MFMA_NB
#define DD_MFMA_NB 4
#endif
constexpr int NB = DD_MFMA_NB;         // bins per workgroup
constexpr int KS = 16;                 // channels per K-step
constexpr int NJ = 32;                 // chunks per tile
constexpr int SLAB = NB * 2 * NJ * 8;  // elements of one plane of a K-step: [b][g][j][s]
constexpr int NT = WAVES * 64;

// Bins per workgroup, plane size and channels a thread stages, per form.
template <int PREC>
struct Blk {
    static constexpr bool GAUSS = PREC == 0 && DD_MFMA_GAUSS;
    static constexpr int NB = GAUSS ? 2 : ::NB;
    static constexpr int SLAB = NB * 2 * NJ * 8;
    static constexpr int CPT = NB * NJ * KS / NT;
    static_assert(CPT * NT == NB * NJ * KS && (CPT == 4 || (GAUSS && CPT == 2)), "staging layout");
};

template <int PREC>
struct Form;
template <>
struct Form<0> {  // float32: re, im - re, -(re + im) (Gauss) or re, im, -im
    static constexpr int PLANES = 3, ELEM = 4;
};
// The bfloat16 and float16 forms stage re and im only when DD_MFMA_NEGA is
// set (the products with -im then take the negated A operand, two xor per
// register, instead of a staged -im plane).
template <>
struct Form<1> {  // bfloat16: re hi, re lo, im hi, im lo [, -im hi, -im lo]
    static constexpr int PLANES = DD_MFMA_NEGA ? 4 : 6, ELEM = 2;
};
template <>
struct Form<2> {  // float16: re, im [, -im]
    static constexpr int PLANES = DD_MFMA_NEGA ? 2 : 3, ELEM = 2;
};

__device__ __forceinline__ float2 cmul(float2 a, float2 b)
{
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

__device__ __forceinline__ uint32_t fbits(float x) { return __builtin_bit_cast(uint32_t, x); }

// hi: truncated bfloat16 of a (low half) and b (high half); lo: the rounded
// bfloat16 of the remainders a - hi(a), b - hi(b), which are exact in float32.
// The remainders are formed with v_and_b32 + v_sub_f32, which issue alongside
// MFMAs; a dot product against (-1, 0) would take one instruction per value
// but is serialised with the matrix pipe on gfx950 (measured: every
// v_dot2c_f32_bf16 adds ~7.5 cycles to a stream of MFMAs, while and/sub
// under an MFMA are nearly free).
__device__ __forceinline__ void split_bf16x2(float a, float b, uint32_t& hi, uint32_t& lo)
{
    hi = __builtin_amdgcn_perm(fbits(b), fbits(a), 0x07060302u);
    const float ra = a - __builtin_bit_cast(float, fbits(a) & 0xffff0000u);
    const float rb = b - __builtin_bit_cast(float, fbits(b) & 0xffff0000u);
    const bf16x2 l = {(__bf16)ra, (__bf16)rb};
    lo = __builtin_bit_cast(uint32_t, l);
}

// Round-to-nearest float16 pair (a in the low half). The empty asm pins the
// inputs in registers so the conversion stays one v_cvt_pk_f16_f32 instead of
// being fused into the producing FMA (which then has to be computed twice);
// the conversion itself is left to the compiler, whose hazard tracking covers
// the MFMA that reads it.
__device__ __forceinline__ uint32_t pack_f16x2(float a, float b)
{
    asm volatile("" : "+v"(a), "+v"(b));
    const f16x2 v = {(_Float16)a, (_Float16)b};
    return __builtin_bit_cast(uint32_t, v);
}

template <int PREC>
__global__ __launch_bounds__(NT)
#if DD_MFMA_OCC2
__attribute__((amdgpu_waves_per_eu(2, 2)))
#elif DD_MFMA_OCC1
__attribute__((amdgpu_waves_per_eu(1, 1)))
#endif
void fdd_mfma_k(
    const float2* __restrict__ X, size_t Xp, const uint32_t* __restrict__ q, size_t qpitch,
    uint32_t ndm, uint32_t nchan, uint32_t nchunk,
    uint32_t nfft, float xscale, float yscale,
    float2* __restrict__ Y, size_t Yp, uint32_t nbb, uint32_t ndb)
{
    using F = Form<PREC>;
    // per form: these hide the namespace-scope defaults
    constexpr bool GAUSS = Blk<PREC>::GAUSS;
    constexpr int NB = Blk<PREC>::NB;
    constexpr int SLAB = Blk<PREC>::SLAB;
    constexpr int BUF = F::PLANES * SLAB * F::ELEM;  // bytes of one staged K-step
    __shared__ __attribute__((aligned(16))) char xs[2 * BUF];

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


    // ---- staging role: bin b, chunk j, channel group g, half h (4 channels) ----
    // Every load is unconditional and in bounds (X has its channel count
    // padded to a multiple of 16 with zero rows; out-of-range bins and chunks
    // read a valid element and are multiplied by 0), so the compiler has no
    // reason to wait for the loads before they are consumed at the end of the
    // K-step.
    // A thread stages bin sb of chunk sj for CPT consecutive channels sc0 ..
    // (consecutive threads take consecutive bins: contiguous in X).
    constexpr int CPT = Blk<PREC>::CPT;
    static_assert(NT % (NB * NJ) == 0 && (NT / (NB * NJ)) * CPT == KS, "staging covers a K-step");
    const int sb = tid % NB, sj = (tid / NB) % NJ, sc0 = (tid / (NB * NJ)) * CPT;
    const uint32_t sk = k0 + sb, sch = ch0 + sj;
    const float xs_scale = (sk <= M && sch < nchunk) ? xscale : 0.f;
    // element offset in a plane of channels sc0 + 4u .. +3: [b][g][j][s], channel 8g + s
    auto s_off = [&](int u) {
        const int c = sc0 + 4 * u;
        return ((sb * 2 + (c >> 3)) * NJ + sj) * 8 + (c & 7);
    };
    // element (channel sc0 + q, chunk, bin) of the K-step: xsrc + q*rs, a
    // K-step further 16*rs
    const size_t rs = (size_t)nchunk * Xp;
    const float2* xsrc = X + ((size_t)sc0 * nchunk + min(sch, nchunk - 1)) * Xp + min(sk, M);

    float2 xv[CPT];
    auto load_x = [&]() {
#pragma unroll
        for (int q = 0; q < CPT; ++q) xv[q] = xsrc[q * rs];
    };
    auto store_x = [&](char* buf) {
      if constexpr (GAUSS) {
        // planes Xr, Xi - Xr, -(Xr + Xi) of two channels
        float xr[2], xi[2];
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            xr[q] = xv[q].x * xs_scale;
            xi[q] = xv[q].y * xs_scale;
        }
        float2* p = reinterpret_cast<float2*>(buf) + s_off(0) / 2;
        p[0] = make_float2(xr[0], xr[1]);
        p[SLAB / 2] = make_float2(xi[0] - xr[0], xi[1] - xr[1]);
        p[2 * SLAB / 2] = make_float2(-(xr[0] + xi[0]), -(xr[1] + xi[1]));
        return;
      }
#pragma unroll
      for (int u = 0; u < CPT / 4; ++u) {
        float xr[4], xi[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            xr[q] = xv[4 * u + q].x * xs_scale;
            xi[q] = xv[4 * u + q].y * xs_scale;
        }
        const int off = s_off(u);
        if constexpr (PREC == 0) {
            float4* p = reinterpret_cast<float4*>(buf) + off / 4;
            p[0] = make_float4(xr[0], xr[1], xr[2], xr[3]);
            p[SLAB / 4] = make_float4(xi[0], xi[1], xi[2], xi[3]);
            p[2 * SLAB / 4] = make_float4(-xi[0], -xi[1], -xi[2], -xi[3]);
        } else if constexpr (PREC == 1) {
            uint32_t rh0, rl0, rh1, rl1, ih0, il0, ih1, il1;
            split_bf16x2(xr[0], xr[1], rh0, rl0);
            split_bf16x2(xr[2], xr[3], rh1, rl1);
            split_bf16x2(xi[0], xi[1], ih0, il0);
            split_bf16x2(xi[2], xi[3], ih1, il1);
            uint2* p = reinterpret_cast<uint2*>(buf) + off / 4;
            p[0] = make_uint2(rh0, rh1);
            p[SLAB / 4] = make_uint2(rl0, rl1);
            p[2 * SLAB / 4] = make_uint2(ih0, ih1);
            p[3 * SLAB / 4] = make_uint2(il0, il1);
            if constexpr (!DD_MFMA_NEGA) {
                p[4 * SLAB / 4] = make_uint2(ih0 ^ 0x80008000u, ih1 ^ 0x80008000u);
                p[5 * SLAB / 4] = make_uint2(il0 ^ 0x80008000u, il1 ^ 0x80008000u);
            }
        } else {
            const uint32_t r0 = pack_f16x2(xr[0], xr[1]), r1 = pack_f16x2(xr[2], xr[3]);
            const uint32_t i0 = pack_f16x2(xi[0], xi[1]), i1 = pack_f16x2(xi[2], xi[3]);
            uint2* p = reinterpret_cast<uint2*>(buf) + off / 4;
            p[0] = make_uint2(r0, r1);
            p[SLAB / 4] = make_uint2(i0, i1);
            if constexpr (!DD_MFMA_NEGA) p[2 * SLAB / 4] = make_uint2(i0 ^ 0x80008000u, i1 ^ 0x80008000u);
        }
      }
    };

    // ---- compute role: DM row i, channel group g ----
    const int i = lane & 31, g = lane >> 5;
    const uint32_t dm = dblk * (WAVES * 32) + wave * 32 + i;
    const uint32_t* qrow = q + (size_t)min(dm, ndm - 1) * qpitch + 8 * g;
    const uint32_t qmask = nfft - 1;
    const float qinv = 1.0f / (float)nfft;
    const int b_off = (g * NJ + i) * 8;  // B operand: chunk j = lane%32 (same index as i)

    // Gauss: acc_k = sum (Pr + Pi) Xr, acc_r = -sum Pi (Xr + Xi),
    // acc_i = sum Pr (Xi - Xr); Y = (acc_k + acc_r, acc_k + acc_i)
    f32x16 acc_r[NB], acc_i[NB], acc_k[GAUSS ? NB : 1];
#pragma unroll
    for (int b = 0; b < NB; ++b)
#pragma unroll
        for (int r = 0; r < 16; ++r) acc_r[b][r] = acc_i[b][r] = 0.f;
    if constexpr (GAUSS) {
#pragma unroll
        for (int b = 0; b < NB; ++b)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc_k[b][r] = 0.f;
    }

    // Phasors of bin k0 and the per-bin steps for this lane's 8 channels. The
    // second half of the waves computes them for the next K-step at the end
    // of the current one (the registers are free by then), so after each
    // barrier one wave of a SIMD issues MFMAs while the other generates
    // phasors.
    float2 p[8], w[8];
    // v_cos_f32 / v_sin_f32 take the angle in revolutions: (k0*q mod nfft)/nfft
    // and q/nfft are exact in float
    auto lookups = [&](const u32x4 (&qq)[2]) {
#pragma unroll
        for (int s = 0; s < 8; ++s) {
            const uint32_t qs = qq[s >> 2][s & 3];
            const float am = (float)(__umul24(k0, qs) & qmask) * qinv;
            const float aw = (float)qs * qinv;
            p[s] = make_float2(__builtin_amdgcn_cosf(am), __builtin_amdgcn_sinf(am));
            w[s] = make_float2(__builtin_amdgcn_cosf(aw), __builtin_amdgcn_sinf(aw));
        }
    };
    const bool late = DD_MFMA_STAGGER && WAVES == 8 && wave >= WAVES / 2;

    const uint32_t nks = (nchan + KS - 1) / KS;
    const size_t kstep = KS * rs;
    load_x();
    u32x4 qv[2] = {*reinterpret_cast<const u32x4*>(qrow), *reinterpret_cast<const u32x4*>(qrow + 4)};
    store_x(xs);
    if (late) lookups(qv);
    if (nks > 1) xsrc += kstep;
    __syncthreads();

    for (uint32_t ks = 0; ks < nks; ++ks) {
        const char* cur = xs + (ks & 1) * BUF;
        const bool more = ks + 1 < nks;
        // prefetch K-step min(ks + 1, nks - 1): the last iteration re-reads
        // its own K-step (in bounds, unused)
        load_x();
        const uint32_t* qnext = qrow + min(ks + 1, nks - 1) * KS;
        const u32x4 qn[2] = {*reinterpret_cast<const u32x4*>(qnext), *reinterpret_cast<const u32x4*>(qnext + 4)};
        xsrc += ks + 2 < nks ? kstep : 0;

        if (!late) lookups(qv);

        // spectra planes of a bin (bfloat16: re hi, re lo, im hi, im lo;
        // float16: re, im); with DD_MFMA_BPREF the next bin's are read before
        // the current bin's MFMAs, so their LDS latency is not waited on
        constexpr bool bpref = DD_MFMA_BPREF && DD_MFMA_NEGA && PREC != 0;
        using Pl = typename std::conditional<PREC == 1, bf16x8, f16x8>::type;
        constexpr int NPL = (PREC == 1) ? 4 : 2;
        auto planes = [&](int b, Pl (&x)[NPL]) {
            const Pl* X = reinterpret_cast<const Pl*>(cur + (size_t)(b * 2 * NJ * 8 + b_off) * F::ELEM);
#pragma unroll
            for (int t = 0; t < NPL; ++t) x[t] = X[t * SLAB / 8];
        };
        Pl xc[NPL];
        if constexpr (bpref) planes(0, xc);

#pragma unroll
        for (int b = 0; b < NB; ++b) {
            const char* bb = cur + (size_t)(b * 2 * NJ * 8 + b_off) * F::ELEM;
            Pl xn[NPL];
            if constexpr (bpref) {
                if (b + 1 < NB) planes(b + 1, xn);
            }
            // phasors of the next bin first: the operand formation below may
            // then overwrite p in place
            float2 pn[8];
            if (b + 1 < NB) {
#pragma unroll
                for (int s = 0; s < 8; ++s) pn[s] = cmul(p[s], w[s]);
            }
            if constexpr (PREC == 0) {
                const float4* X = reinterpret_cast<const float4*>(bb);
                float xr[8], xi[8], xn[8];
                *reinterpret_cast<float4*>(xr) = X[0];
                *reinterpret_cast<float4*>(xr + 4) = X[1];
                *reinterpret_cast<float4*>(xi) = X[SLAB / 4];
                *reinterpret_cast<float4*>(xi + 4) = X[SLAB / 4 + 1];
                *reinterpret_cast<float4*>(xn) = X[2 * SLAB / 4];
                *reinterpret_cast<float4*>(xn + 4) = X[2 * SLAB / 4 + 1];
                if constexpr (GAUSS) {
                    // planes: xr = Xr, xi = Xi - Xr, xn = -(Xr + Xi)
#pragma unroll
                    for (int s = 0; s < 8; ++s) {
                        acc_k[b] = __builtin_amdgcn_mfma_f32_32x32x2f32(p[s].x + p[s].y, xr[s], acc_k[b], 0, 0, 0);
                        acc_r[b] = __builtin_amdgcn_mfma_f32_32x32x2f32(p[s].y, xn[s], acc_r[b], 0, 0, 0);
                        acc_i[b] = __builtin_amdgcn_mfma_f32_32x32x2f32(p[s].x, xi[s], acc_i[b], 0, 0, 0);
                    }
                } else {
#pragma unroll
                    for (int s = 0; s < 8; ++s) {
                        acc_r[b] = __builtin_amdgcn_mfma_f32_32x32x2f32(p[s].x, xr[s], acc_r[b], 0, 0, 0);
                        acc_r[b] = __builtin_amdgcn_mfma_f32_32x32x2f32(p[s].y, xn[s], acc_r[b], 0, 0, 0);
                        acc_i[b] = __builtin_amdgcn_mfma_f32_32x32x2f32(p[s].x, xi[s], acc_i[b], 0, 0, 0);
                        acc_i[b] = __builtin_amdgcn_mfma_f32_32x32x2f32(p[s].y, xr[s], acc_i[b], 0, 0, 0);
                    }
                }
            } else if constexpr (PREC == 1) {
                const bf16x8* X = reinterpret_cast<const bf16x8*>(bb);
                uint32_t prh[4], prl[4], pih[4], pil[4];
#pragma unroll
                for (int s = 0; s < 4; ++s) {
                    split_bf16x2(p[2 * s].x, p[2 * s + 1].x, prh[s], prl[s]);
                    split_bf16x2(p[2 * s].y, p[2 * s + 1].y, pih[s], pil[s]);
                }
                const bf16x8 Arh = __builtin_bit_cast(bf16x8, prh), Arl = __builtin_bit_cast(bf16x8, prl);
                const bf16x8 Aih = __builtin_bit_cast(bf16x8, pih), Ail = __builtin_bit_cast(bf16x8, pil);
                f32x16 ar = acc_r[b], ai = acc_i[b];
                if constexpr (bpref) {
                    uint32_t nih[4], nil[4];
#pragma unroll
                    for (int s = 0; s < 4; ++s) nih[s] = pih[s] ^ 0x80008000u, nil[s] = pil[s] ^ 0x80008000u;
                    const bf16x8 nAih = __builtin_bit_cast(bf16x8, nih), nAil = __builtin_bit_cast(bf16x8, nil);
                    ar = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Arh, xc[0], ar, 0, 0, 0);
                    ai = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Aih, xc[0], ai, 0, 0, 0);
                    ar = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Arl, xc[0], ar, 0, 0, 0);
                    ai = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Ail, xc[0], ai, 0, 0, 0);
                    ar = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Arh, xc[1], ar, 0, 0, 0);
                    ai = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Aih, xc[1], ai, 0, 0, 0);
                    ai = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Arh, xc[2], ai, 0, 0, 0);
                    ar = __builtin_amdgcn_mfma_f32_32x32x16_bf16(nAih, xc[2], ar, 0, 0, 0);
                    ai = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Arl, xc[2], ai, 0, 0, 0);
                    ar = __builtin_amdgcn_mfma_f32_32x32x16_bf16(nAil, xc[2], ar, 0, 0, 0);
                    ai = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Arh, xc[3], ai, 0, 0, 0);
                    ar = __builtin_amdgcn_mfma_f32_32x32x16_bf16(nAih, xc[3], ar, 0, 0, 0);
                    acc_r[b] = ar;
                    acc_i[b] = ai;
                } else {
                // grouped by spectra plane (each read once, used, then dead)
                bf16x8 x = X[0];  // re hi
                ar = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Arh, x, ar, 0, 0, 0);
                ai = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Aih, x, ai, 0, 0, 0);
                ar = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Arl, x, ar, 0, 0, 0);
                ai = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Ail, x, ai, 0, 0, 0);
                x = X[SLAB / 8];  // re lo
                ar = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Arh, x, ar, 0, 0, 0);
                ai = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Aih, x, ai, 0, 0, 0);
                if constexpr (DD_MFMA_NEGA) {
                    // (-Aim) * Xim: negated A operands (sign bits of both halves)
                    uint32_t nih[4], nil[4];
#pragma unroll
                    for (int s = 0; s < 4; ++s) nih[s] = pih[s] ^ 0x80008000u, nil[s] = pil[s] ^ 0x80008000u;
                    const bf16x8 nAih = __builtin_bit_cast(bf16x8, nih), nAil = __builtin_bit_cast(bf16x8, nil);
                    x = X[2 * SLAB / 8];  // im hi
                    ai = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Arh, x, ai, 0, 0, 0);
                    ar = __builtin_amdgcn_mfma_f32_32x32x16_bf16(nAih, x, ar, 0, 0, 0);
                    ai = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Arl, x, ai, 0, 0, 0);
                    ar = __builtin_amdgcn_mfma_f32_32x32x16_bf16(nAil, x, ar, 0, 0, 0);
                    x = X[3 * SLAB / 8];  // im lo
                    ai = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Arh, x, ai, 0, 0, 0);
                    ar = __builtin_amdgcn_mfma_f32_32x32x16_bf16(nAih, x, ar, 0, 0, 0);
                } else {
                    x = X[2 * SLAB / 8];  // im hi
                    ai = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Arh, x, ai, 0, 0, 0);
                    ai = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Arl, x, ai, 0, 0, 0);
                    x = X[4 * SLAB / 8];  // -im hi
                    ar = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Aih, x, ar, 0, 0, 0);
                    ar = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Ail, x, ar, 0, 0, 0);
                    x = X[3 * SLAB / 8];  // im lo
                    ai = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Arh, x, ai, 0, 0, 0);
                    x = X[5 * SLAB / 8];  // -im lo
                    ar = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Aih, x, ar, 0, 0, 0);
                }
                acc_r[b] = ar;
                acc_i[b] = ai;
                }
            } else {
                const f16x8* X = reinterpret_cast<const f16x8*>(bb);
                const f16x8 xr = bpref ? xc[0] : X[0], xi = bpref ? xc[1] : X[SLAB / 8];
                uint32_t pr[4], pi[4];
#pragma unroll
                for (int s = 0; s < 4; ++s) {
                    pr[s] = pack_f16x2(p[2 * s].x, p[2 * s + 1].x);
                    pi[s] = pack_f16x2(p[2 * s].y, p[2 * s + 1].y);
                }
                const f16x8 Ar = __builtin_bit_cast(f16x8, pr), Ai = __builtin_bit_cast(f16x8, pi);
                acc_r[b] = __builtin_amdgcn_mfma_f32_32x32x16_f16(Ar, xr, acc_r[b], 0, 0, 0);
                acc_i[b] = __builtin_amdgcn_mfma_f32_32x32x16_f16(Ar, xi, acc_i[b], 0, 0, 0);
                if constexpr (DD_MFMA_NEGA) {
                    uint32_t ni[4];
#pragma unroll
                    for (int s = 0; s < 4; ++s) ni[s] = pi[s] ^ 0x80008000u;
                    acc_r[b] = __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(f16x8, ni), xi, acc_r[b], 0, 0, 0);
                } else {
                    acc_r[b] = __builtin_amdgcn_mfma_f32_32x32x16_f16(Ai, X[2 * SLAB / 8], acc_r[b], 0, 0, 0);
                }
                acc_i[b] = __builtin_amdgcn_mfma_f32_32x32x16_f16(Ai, xr, acc_i[b], 0, 0, 0);
            }
            if constexpr (bpref) {
                if (b + 1 < NB) {
#pragma unroll
                    for (int t = 0; t < NPL; ++t) xc[t] = xn[t];
                }
            }
            if (b + 1 < NB) {
#pragma unroll
                for (int s = 0; s < 8; ++s) p[s] = pn[s];
            }
        }

        if (late && more) lookups(qn);
        if (more) store_x(xs + ((ks + 1) & 1) * BUF);
        qv[0] = qn[0];
        qv[1] = qn[1];
        __syncthreads();
    }

    // ---- epilogue: rows 8*(r/4) + 4*g + r%4 of the tile, column = chunk ----
    const uint32_t ch = ch0 + i;
    if (ch >= nchunk) return;
    const bool all_bins = k0 + NB - 1 <= M;
    if constexpr (GAUSS) {
#pragma unroll
        for (int b = 0; b < NB; ++b)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                acc_r[b][r] += acc_k[b][r];
                acc_i[b][r] += acc_k[b][r];
            }
    }
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const uint32_t row = 8 * (r / 4) + 4 * g + (r % 4);
        const uint32_t dmo = dblk * (WAVES * 32) + wave * 32 + row;
        if (dmo >= ndm) continue;
        float2* dst = Y + ((size_t)dmo * nchunk + ch) * Yp + k0;
        if (all_bins) {
            float4* d4 = reinterpret_cast<float4*>(dst);
#pragma unroll
            for (int b = 0; b < NB; b += 2)
                d4[b / 2] = make_float4(acc_r[b][r] * yscale, acc_i[b][r] * yscale,
                                        acc_r[b + 1][r] * yscale, acc_i[b + 1][r] * yscale);
        } else {
#pragma unroll
            for (int b = 0; b < NB; ++b)
                if (k0 + b <= M) dst[b] = make_float2(acc_r[b][r] * yscale, acc_i[b][r] * yscale);
        }
    }
}


}  // namespace

extern "C" hipError_t dd_fdd_contract_mfma(int prec, const float2* X, size_t Xp, const uint32_t* q,
                                           size_t qpitch, size_t ndm, size_t nchan, size_t nchunk,
                                           size_t nfft,
                                           float xscale, float2* Y, size_t Yp, hipStream_t stream)
{
    if (ndm == 0 || nchunk == 0) return hipSuccess;
    // layout preconditions of the vector loads and stores
    if (qpitch % 16 || qpitch < (nchan + 15) / 16 * 16 || Yp % 2 || nfft < 2 * NB || nfft > (size_t(1) << 18) ||
        (nfft & (nfft - 1)) || Xp < nfft / 2 + 1)
        return hipErrorInvalidValue;
    const size_t M = nfft / 2;
    const uint32_t ndb = (uint32_t)((ndm + WAVES * 32 - 1) / (WAVES * 32));
    const uint32_t ncg = (uint32_t)((nchunk + NJ - 1) / NJ);
    const float yscale = 1.0f / xscale;
    // bin blocks of the form's workgroup
#define DD_MFMA(P)                                                                                  \
    {                                                                                               \
        const uint32_t nbb = (uint32_t)((M + Blk<P>::NB) / Blk<P>::NB);                             \
        const size_t nblocks = (size_t)nbb * ndb * ncg;                                             \
        if (nblocks >= (size_t(1) << 31)) return hipErrorInvalidValue;                              \
        hipLaunchKernelGGL(fdd_mfma_k<P>, dim3((unsigned)nblocks), dim3(NT), 0, stream, X, Xp, q,   \
                           qpitch, (uint32_t)ndm, (uint32_t)nchan, (uint32_t)nchunk,                \
                           (uint32_t)nfft, xscale, yscale, Y, Yp, nbb, ndb);                        \
    }
    switch (prec) {
    case 0: DD_MFMA(0); break;
    case 1: DD_MFMA(1); break;
    case 2: DD_MFMA(2); break;
    default: return hipErrorInvalidValue;
    }
#undef DD_MFMA
    return hipGetLastError();
}
