// gfx950 phasor contraction of the FDD path on 16 x 16 matrix-core tiles, re
// and im side by side along N (float32, bfloat16 three-product and float16
// single-product forms).
//
// For every Fourier bin k the dedispersed spectra are a complex GEMM
//   Y_k[dm][chunk] = sum_c P_k[dm][c] X_k[c][chunk],  P_k[dm][c] = e^{+2 pi i k d(dm,c)/nfft}
// (the phase rotation of src/fdd/dedisperse/fdd_kernel.cuh:226-235 with the
// truncated integer delay d). As a real GEMM over K = channels the complex
// product is two products with the phasor planes as A operands,
//   [Yr | Yi] += Pr [Xr | Xi] + Pi [-Xi | Xr],
// so a chunk's real and imaginary outputs are two columns of one accumulator
// tile: an N-tile of 16 columns holds 8 chunks (columns 0-7 the real parts,
// 8-15 the imaginary parts). 23 chunks fill 46 of 48 columns; a 32-column
// tile of one part per column filled 23 of 32.
//
// Workgroup: 8 waves, each 32 DMs (two 16-row tiles sharing every B read),
// CG = 8*NTL chunks, 4 bin slots; the K loop walks the channels 32 at a time.
// With MIRROR the slots are bins k0, k0+1 and their mirrors M-k0, M-k0-1
// (M = nfft/2): P_{M-k} = (-1)^d conj(P_k) exactly, so a mirror phasor is two
// sign flips of the direct one instead of a cosine, a sine and the hi/lo
// split. Without MIRROR the slots are four consecutive bins.
//
// Phasors (A operand), generated in registers: the phase of bin k is the
// exact integer (k*q) mod nfft with q = d mod nfft (q[dm][c], uint32), and
// v_cos_f32 / v_sin_f32 take it in revolutions, (k*q mod nfft)/nfft, exact in
// float for nfft <= 2^24.
// float32: v_mfma_f32_16x16x4_f32, one phasor component per lane;
// bfloat16 (v_mfma_f32_16x16x32_bf16): phasors and spectra split into
// truncated hi + rounded lo, products Ph Xh + Ph Xl + Pl Xh (the remainder
// x - hi is exact in float32); float16 (v_mfma_f32_16x16x32_f16):
// round-to-nearest, spectra scaled by a power of two into the float16 range.
// Float32 accumulation.
//
// Operand layout: lane l of A holds DM row l%16 and channels 8*(l/16) .. +7 of
// the K-step; lane l of B holds column l%16 and the same channels, so the sum
// runs over channels whatever the hardware's k order. Accumulator element r of
// lane l is row 4*(l/16) + r, column l%16.
#include <hip/hip_runtime.h>

#include "fdd_kernels.h"

namespace {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(2 * sizeof(__bf16)))) __bf16 bf16x2;
typedef __attribute__((__vector_size__(8 * sizeof(_Float16)))) _Float16 f16x8;
typedef __attribute__((__vector_size__(2 * sizeof(_Float16)))) _Float16 f16x2;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;
typedef __attribute__((__vector_size__(4 * sizeof(uint32_t)))) uint32_t u32x4;

constexpr int KS = 32;          // channels per K-step

__device__ __forceinline__ uint32_t fbits(float x) { return __builtin_bit_cast(uint32_t, x); }

// hi: truncated bfloat16 of a (low half) and b (high half); lo: the rounded
// bfloat16 of the remainders a - hi(a), b - hi(b), which are exact in float32.
__device__ __forceinline__ void split_bf16x2(float a, float b, uint32_t& hi, uint32_t& lo)
{
    hi = __builtin_amdgcn_perm(fbits(b), fbits(a), 0x07060302u);
    const float ra = a - __builtin_bit_cast(float, fbits(a) & 0xffff0000u);
    const float rb = b - __builtin_bit_cast(float, fbits(b) & 0xffff0000u);
    const bf16x2 l = {(__bf16)ra, (__bf16)rb};
    lo = __builtin_bit_cast(uint32_t, l);
}

// hi: the round-to-nearest bfloat16 of a (low half) and b (high half); lo:
// the rounded bfloat16 of the remainders, a - hi(a) from one bfloat16 dot
// product (-1 * hi(a) + 0 * hi(b) + a, exact: the difference is representable).
// Four instructions per pair instead of six. The selector pairs (-1, 0) and
// (0, -1) come in registers (DotSel): as an inline constant the compiler's
// encoding of -1.0 for a bfloat16 operand computes a different product on
// gfx950 (measured), a register operand the exact one.
struct DotSel {
    bf16x2 m1, m2;
    __device__ __forceinline__ DotSel()
    {
        uint32_t r1 = 0x0000bf80u, r2 = 0xbf800000u;  // (-1, 0), (0, -1)
        asm volatile("" : "+v"(r1), "+v"(r2));
        m1 = __builtin_bit_cast(bf16x2, r1);
        m2 = __builtin_bit_cast(bf16x2, r2);
    }
};
__device__ __forceinline__ void split_bf16x2_rn(float a, float b, uint32_t& hi, uint32_t& lo, const DotSel& ds)
{
    const bf16x2 h = {(__bf16)a, (__bf16)b};
    const float ra = __builtin_amdgcn_fdot2_f32_bf16(h, ds.m1, a, false);
    const float rb = __builtin_amdgcn_fdot2_f32_bf16(h, ds.m2, b, false);
    const bf16x2 l = {(__bf16)ra, (__bf16)rb};
    hi = __builtin_bit_cast(uint32_t, h);
    lo = __builtin_bit_cast(uint32_t, l);
}

// Round-to-nearest float16 pair (a in the low half); the empty asm keeps the
// conversion one v_cvt_pk_f16_f32 instead of being fused into its producer.
__device__ __forceinline__ uint32_t pack_f16x2(float a, float b)
{
    asm volatile("" : "+v"(a), "+v"(b));
    const f16x2 v = {(_Float16)a, (_Float16)b};
    return __builtin_bit_cast(uint32_t, v);
}

// Staged spectra planes per form and bytes per element; a staged column
// holds the 32 channels of a K-step plus 16 bytes, so that the 16-byte reads
// of 8 consecutive columns cover all banks (column strides 80 and 144 bytes
// are odd multiples of 16).
template <int PREC>
struct Planes;
template <>
struct Planes<0> {  // float32: [Xr|Xi], [-Xi|Xr]
    static constexpr int N = 2, ELEM = 4;
    using V = float;
};
template <>
struct Planes<1> {  // bfloat16: [Xr|Xi] hi, [Xr|Xi] lo, [-Xi|Xr] hi, [-Xi|Xr] lo
    static constexpr int N = 4, ELEM = 2;
    using V = bf16x8;
};
template <>
struct Planes<2> {  // float16: [Xr|Xi], [-Xi|Xr]
    static constexpr int N = 2, ELEM = 2;
    using V = f16x8;
};

// Staging layout. A staging item is (slot s, channel pair cp, chunk jj),
// s fastest, then cp: a wave's 64 lanes write 16 channel pairs of NS slots
// (of one or two chunks). The slot stride is an odd multiple of 128/NS bytes
// beyond a multiple of 128 (NCOL*COLB is one), so the 32 lanes of a half-wave
// write 32 distinct banks.
template <int PREC, int NTL, int NS, int NT>
struct Geo {
    static constexpr int COLB = KS * Planes<PREC>::ELEM + 16;  // bytes per staged column
    static constexpr int CG = 8 * NTL;                    // chunks per workgroup
    static constexpr int NCOL = 16 * NTL;                 // columns per bin slot
    static constexpr int SSTR = NCOL * COLB + 128 / NS;   // bytes per slot
    static constexpr int PSTR = NS * SSTR;                // bytes per plane
    static constexpr int ITEMS = NS * CG * (KS / 2);      // staging items per K-step (channel pairs)
    static constexpr int IPT = ITEMS / NT;                // items per thread
    static_assert(IPT * NT == ITEMS, "staging items divide over the threads");
    static_assert((NCOL * COLB) % 128 == 0 && (NS == 2 || NS == 4), "bank spread of the staging writes");
};

// WAVES waves per workgroup, each MT 16-row DM tiles (sharing every B read);
// NS bin slots (with MIRROR NS/2 direct bins and their mirrors).
//
// PIPE (bfloat16 and float16 with the delays computed in the kernel, 8 waves
// of 2 DM tiles): software pipelined. A K-step is NDIR units, one computed bin
// each (its direct slot and, with MIRROR, its mirror slot); between the
// product groups of a unit the wave generates pieces of the next unit's
// phasors (the next K-step's first bin in the last unit) and stages half of
// the next K-step's spectra after the first and the last unit, so that the
// phasor arithmetic issues beside the matrix products instead of before them.
// Measured on 1024 ch x 2048 DMs x 23 chunks of 65536 (bfloat16 with mirror
// bins): 45.6 ms against 49.6 ms for cfg 0; 4 waves of 4 tiles (one wave per
// SIMD, 512 registers) ran 56.5 ms, its waits on the LDS exposed.
//
// BLK (PIPE with MIRROR and the packed output): the spectra and the packed
// output in the bin-block layout (bins k0, k0+1, M-k0, M-k0-1 of a workgroup
// contiguous for every chunk), so that the staging loads read and the
// epilogue writes whole cache lines instead of 8- and 16-byte pieces.
template <int PREC, int NTL, bool MIRROR, int WAVES, int MT, int NS, bool PIPE = false, bool BLK = false>
__global__ __launch_bounds__(WAVES * 64, 512 / (WAVES * 64)) void fdd_mfma16_k(
    const float2* __restrict__ X, size_t Xp, const uint32_t* __restrict__ q, size_t qpitch,
    uint32_t ndm, uint32_t nchan, uint32_t nchunk, uint32_t nfft, float xscale, float yscale,
    float2* __restrict__ Y, size_t Yp, uint32_t nbb, uint32_t ndb, const float2* __restrict__ zph,
    const float* __restrict__ dmlist, const float* __restrict__ dtab)
{
    static_assert(!PIPE || (PREC != 0 && NS == 4 && WAVES == 8), "pipelined form: bfloat16 / float16");
    static_assert(!BLK || (PIPE && MIRROR), "bin-block layout: pipelined form with mirror bins");
    constexpr int NT = WAVES * 64;        // threads
    constexpr int DPB = WAVES * MT * 16;  // DMs per workgroup
    constexpr int NDIR = MIRROR ? NS / 2 : NS;  // bins whose phasors are computed
    using G = Geo<PREC, NTL, NS, NT>;
    using PL = Planes<PREC>;
    using V = typename PL::V;
    constexpr int COLB = G::COLB;
    constexpr int BUF = PL::N * G::PSTR;
    __shared__ __attribute__((aligned(16))) char xs[2 * BUF];

    const uint32_t M = nfft / 2;
    const uint32_t mask = nfft - 1;
    const float inv = 1.0f / (float)nfft;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const uint32_t wave = __builtin_amdgcn_readfirstlane(tid >> 6);

    // XCD-aware order: consecutive workgroups of an XCD share a bin block (DM
    // block fastest), so the staged spectra come from that XCD's L2.
    const uint32_t nb = gridDim.x, bid = blockIdx.x;
    const uint32_t xcd = bid & 7u, per = nb >> 3, rem = nb & 7u;
    const uint32_t L = xcd * per + min(xcd, rem) + (bid >> 3);
    const uint32_t dblk = L % ndb;
    const uint32_t rest = L / ndb;
    const uint32_t bblk = rest % nbb;
    const uint32_t cgrp = rest / nbb;
    const uint32_t ch0 = cgrp * G::CG;
    // bin of slot s and whether it is computed and written here
    const uint32_t k0 = NDIR * bblk;
    auto slot_bin = [&](int s) -> uint32_t {
        if (MIRROR) return s < NDIR ? k0 + s : M - (k0 + s - NDIR);
        return k0 + s;
    };
    auto slot_valid = [&](int s) -> bool {
        if (MIRROR) return s < NDIR ? (k0 + s <= M / 2) : (k0 + s - NDIR < M / 2);
        return k0 + s <= M;
    };

    // ---- staging role: items (slot s, channel pair cp, chunk jj) ----
    // Every load is unconditional and in bounds (X has its channel count
    // padded to a multiple of 32 with zero rows; out-of-range bins and chunks
    // read a valid element and are multiplied by 0).
    const size_t rs = (size_t)nchunk * Xp;  // channel stride of X
    const DotSel dsel;  // used by the pipelined bfloat16 form
    const float2* src[G::IPT];
    float sscale[G::IPT];
    int soff[G::IPT];
#pragma unroll
    for (int u = 0; u < G::IPT; ++u) {
        const int it = tid + u * NT;
        const int s = it % NS, cp = (it / NS) % (KS / 2), jj = it / (NS * (KS / 2));
        const uint32_t chunk = ch0 + jj;
        const uint32_t bin = min(slot_bin(s), M);
        src[u] = X + ((size_t)(2 * cp) * nchunk + min(chunk, nchunk - 1)) * Xp + bin;
        sscale[u] = (slot_valid(s) && chunk < nchunk) ? xscale : 0.f;
        const int col = 16 * (jj / 8) + (jj % 8);
        soff[u] = s * G::SSTR + col * COLB + 2 * PL::ELEM * cp;  // the imaginary column is 8 further
    }
    float2 xv[G::IPT][2];
    auto load_x = [&](size_t kofs) {
#pragma unroll
        for (int u = 0; u < G::IPT; ++u) {
            xv[u][0] = src[u][kofs];
            xv[u][1] = src[u][kofs + rs];
        }
    };
    // staged form of item u (channels 2cp, 2cp+1 of one chunk and slot)
    auto store_item = [&](char* buf, int u, float2 x0, float2 x1) __attribute__((always_inline)) {
        const float xr0 = x0.x * sscale[u], xi0 = x0.y * sscale[u];
        const float xr1 = x1.x * sscale[u], xi1 = x1.y * sscale[u];
        char* b = buf + soff[u];
        if constexpr (PREC == 0) {
            *reinterpret_cast<float2*>(b) = make_float2(xr0, xr1);
            *reinterpret_cast<float2*>(b + 8 * COLB) = make_float2(xi0, xi1);
            *reinterpret_cast<float2*>(b + G::PSTR) = make_float2(-xi0, -xi1);
            *reinterpret_cast<float2*>(b + G::PSTR + 8 * COLB) = make_float2(xr0, xr1);
        } else if constexpr (PREC == 1) {
            uint32_t rh, rl, ih, il;
            if constexpr (PIPE) {
                split_bf16x2_rn(xr0, xr1, rh, rl, dsel);
                split_bf16x2_rn(xi0, xi1, ih, il, dsel);
            } else {
                split_bf16x2(xr0, xr1, rh, rl);
                split_bf16x2(xi0, xi1, ih, il);
            }
            *reinterpret_cast<uint32_t*>(b) = rh;
            *reinterpret_cast<uint32_t*>(b + 8 * COLB) = ih;
            *reinterpret_cast<uint32_t*>(b + G::PSTR) = rl;
            *reinterpret_cast<uint32_t*>(b + G::PSTR + 8 * COLB) = il;
            *reinterpret_cast<uint32_t*>(b + 2 * G::PSTR) = ih ^ 0x80008000u;
            *reinterpret_cast<uint32_t*>(b + 2 * G::PSTR + 8 * COLB) = rh;
            *reinterpret_cast<uint32_t*>(b + 3 * G::PSTR) = il ^ 0x80008000u;
            *reinterpret_cast<uint32_t*>(b + 3 * G::PSTR + 8 * COLB) = rl;
        } else {
            const uint32_t r = pack_f16x2(xr0, xr1), i = pack_f16x2(xi0, xi1);
            *reinterpret_cast<uint32_t*>(b) = r;
            *reinterpret_cast<uint32_t*>(b + 8 * COLB) = i;
            *reinterpret_cast<uint32_t*>(b + G::PSTR) = i ^ 0x80008000u;
            *reinterpret_cast<uint32_t*>(b + G::PSTR + 8 * COLB) = r;
        }
    };
    auto store_x = [&](char* buf) {
#pragma unroll
        for (int u = 0; u < G::IPT; ++u) store_item(buf, u, xv[u][0], xv[u][1]);
    };

    // ---- compute role: DM rows lane%16 of tiles mt, channels 8g .. 8g+7 ----
    const int i16 = lane & 15, g = lane >> 4;
    const uint32_t dmw = dblk * DPB + wave * (MT * 16);
    const uint32_t* qrow[MT];
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
        qrow[mt] = q + (size_t)min(dmw + mt * 16 + i16, ndm - 1) * qpitch + 8 * g;
    const int b_off = i16 * COLB + 8 * PL::ELEM * g;  // B operand: column 16t + lane%16

    f32x4 acc[NS][MT][NTL];
#pragma unroll
    for (int s = 0; s < NS; ++s)
#pragma unroll
        for (int mt = 0; mt < MT; ++mt)
#pragma unroll
            for (int t = 0; t < NTL; ++t) acc[s][mt][t] = f32x4{0.f, 0.f, 0.f, 0.f};

    const uint32_t nks = (nchan + KS - 1) / KS;
    const size_t kstep = KS * rs;
    // Integer delays d(dm, c) of the K-step's channels: from the phase-step
    // table q, or (dmlist given) computed as (int)(dm * delay[c]) in float, the
    // reference's truncated delay (src/fdd/dedisperse/fdd_kernel.cuh:226-229),
    // from the DM of the lane's rows and 8 channel delays shared by every row
    // (d mod nfft is all the phase needs, and d < nfft).
    const bool from_dm = dmlist != nullptr;
    float dmv[MT];
#pragma unroll
    for (int mt = 0; mt < MT; ++mt) dmv[mt] = from_dm ? dmlist[min(dmw + mt * 16 + i16, ndm - 1)] : 0.f;
    u32x4 qv[MT][2];
    float4 dlv[2];
    auto load_q = [&](uint32_t ks) {
        if (from_dm) {
            const float* dp = dtab + ks * KS + 8 * g;
            dlv[0] = *reinterpret_cast<const float4*>(dp);
            dlv[1] = *reinterpret_cast<const float4*>(dp + 4);
            return;
        }
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) {
            qv[mt][0] = *reinterpret_cast<const u32x4*>(qrow[mt] + ks * KS);
            qv[mt][1] = *reinterpret_cast<const u32x4*>(qrow[mt] + ks * KS + 4);
        }
    };
    auto delays_now = [&](u32x4 (&qc)[MT][2]) {
        if (from_dm) {
            const float dl[8] = {dlv[0].x, dlv[0].y, dlv[0].z, dlv[0].w, dlv[1].x, dlv[1].y, dlv[1].z, dlv[1].w};
#pragma unroll
            for (int mt = 0; mt < MT; ++mt)
#pragma unroll
                for (int e = 0; e < 8; ++e)
                    qc[mt][e >> 2][e & 3] = (uint32_t)__float2int_rz(__fmul_rn(dmv[mt], dl[e]));
        } else {
#pragma unroll
            for (int mt = 0; mt < MT; ++mt) {
                qc[mt][0] = qv[mt][0];
                qc[mt][1] = qv[mt][1];
            }
        }
    };
    if constexpr (!PIPE) {
        load_x(0);
        load_q(0);
        store_x(xs);
        __syncthreads();

        for (uint32_t ks = 0; ks < nks; ++ks) {
            const char* cur = xs + (ks & 1) * BUF;
            const bool more = ks + 1 < nks;
            u32x4 qc[MT][2];
            delays_now(qc);
            // prefetch K-step min(ks + 1, nks - 1): the last iteration re-reads
            // its own K-step (in bounds, unused)
            const uint32_t kn = more ? ks + 1 : ks;
            load_x((size_t)kn * kstep);
            load_q(kn);

            // sign masks of the mirror phasors: (-1)^d per channel, d odd <=> q odd
            uint32_t smask[MT][4];
            if constexpr (MIRROR) {
    #pragma unroll
                for (int mt = 0; mt < MT; ++mt)
    #pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const uint32_t qa = qc[mt][j >> 1][2 * (j & 1)], qb = qc[mt][j >> 1][2 * (j & 1) + 1];
                        smask[mt][j] = ((qa & 1u) << 15) | ((qb & 1u) << 31);
                    }
            }

            if constexpr (PREC == 0) {
                // float32: A is one phasor component per lane and 16 x 16 x 4
                // MFMA; MFMA e of a K-step sums channels 8g + e of the lane groups
    #pragma unroll
                for (int b = 0; b < NDIR; ++b) {
                    const uint32_t kb = k0 + b;
                    float pr[MT][8], pim[MT][8];
    #pragma unroll
                    for (int mt = 0; mt < MT; ++mt)
    #pragma unroll
                        for (int e = 0; e < 8; ++e) {
                            const uint32_t qq = qc[mt][e >> 2][e & 3];
                            const float a = (float)(__umul24(kb, qq) & mask) * inv;
                            pr[mt][e] = __builtin_amdgcn_cosf(a);
                            pim[mt][e] = __builtin_amdgcn_sinf(a);
                        }
    #pragma unroll
                    for (int pass = 0; pass < (MIRROR ? 2 : 1); ++pass) {
                        const int s = pass ? NDIR + b : b;
                        if (pass) {
                            // re' = (-1)^d re, im' = -(-1)^d im
    #pragma unroll
                            for (int mt = 0; mt < MT; ++mt)
    #pragma unroll
                                for (int e = 0; e < 8; ++e) {
                                    const uint32_t sg = (qc[mt][e >> 2][e & 3] & 1u) << 31;
                                    pr[mt][e] = __builtin_bit_cast(float, fbits(pr[mt][e]) ^ sg);
                                    pim[mt][e] = __builtin_bit_cast(float, fbits(pim[mt][e]) ^ sg ^ 0x80000000u);
                                }
                        }
                        const char* sb = cur + s * G::SSTR + b_off;
    #pragma unroll
                        for (int t = 0; t < NTL; ++t) {
                            float xr[8], xi[8];
                            *reinterpret_cast<float4*>(xr) = *reinterpret_cast<const float4*>(sb + t * 16 * COLB);
                            *reinterpret_cast<float4*>(xr + 4) =
                                *reinterpret_cast<const float4*>(sb + t * 16 * COLB + 16);
                            *reinterpret_cast<float4*>(xi) =
                                *reinterpret_cast<const float4*>(sb + G::PSTR + t * 16 * COLB);
                            *reinterpret_cast<float4*>(xi + 4) =
                                *reinterpret_cast<const float4*>(sb + G::PSTR + t * 16 * COLB + 16);
    #pragma unroll
                            for (int e = 0; e < 8; ++e)
    #pragma unroll
                                for (int mt = 0; mt < MT; ++mt) {
                                    acc[s][mt][t] =
                                        __builtin_amdgcn_mfma_f32_16x16x4f32(pr[mt][e], xr[e], acc[s][mt][t], 0, 0, 0);
                                    acc[s][mt][t] =
                                        __builtin_amdgcn_mfma_f32_16x16x4f32(pim[mt][e], xi[e], acc[s][mt][t], 0, 0, 0);
                                }
                        }
                    }
                }
            } else {
        #pragma unroll
                for (int b = 0; b < NDIR; ++b) {
                    const uint32_t kb = k0 + b;
                    // phasor planes of bin kb for the two DM tiles
                    uint32_t A[MT][PL::N][4];
        #pragma unroll
                    for (int mt = 0; mt < MT; ++mt)
        #pragma unroll
                        for (int j = 0; j < 4; ++j) {
                            float cr[2], ci[2];
        #pragma unroll
                            for (int h = 0; h < 2; ++h) {
                                const uint32_t qq = qc[mt][j >> 1][2 * (j & 1) + h];
                                const float a = (float)(__umul24(kb, qq) & mask) * inv;
                                cr[h] = __builtin_amdgcn_cosf(a);
                                ci[h] = __builtin_amdgcn_sinf(a);
                            }
                            if constexpr (PREC == 1) {
                                split_bf16x2(cr[0], cr[1], A[mt][0][j], A[mt][1][j]);
                                split_bf16x2(ci[0], ci[1], A[mt][2][j], A[mt][3][j]);
                            } else {
                                A[mt][0][j] = pack_f16x2(cr[0], cr[1]);
                                A[mt][1][j] = pack_f16x2(ci[0], ci[1]);
                            }
                        }
                    // the direct slot, then (MIRROR) its mirror with the signs flipped
        #pragma unroll
                    for (int pass = 0; pass < (MIRROR ? 2 : 1); ++pass) {
                        const int s = pass ? NDIR + b : b;
                        if (pass) {
        #pragma unroll
                            for (int mt = 0; mt < MT; ++mt)
        #pragma unroll
                                for (int j = 0; j < 4; ++j) {
                                    // re' = (-1)^d re, im' = -(-1)^d im
        #pragma unroll
                                    for (int pl = 0; pl < PL::N; ++pl)
                                        A[mt][pl][j] ^= smask[mt][j] ^ ((pl >= PL::N / 2) ? 0x80008000u : 0u);
                                }
                        }
                        V a[MT][PL::N];
        #pragma unroll
                        for (int mt = 0; mt < MT; ++mt)
        #pragma unroll
                            for (int pl = 0; pl < PL::N; ++pl) a[mt][pl] = __builtin_bit_cast(V, A[mt][pl]);
                        const char* sb = cur + s * G::SSTR + b_off;
        #pragma unroll
                        for (int t = 0; t < NTL; ++t) {
                            V x[PL::N];
        #pragma unroll
                            for (int pl = 0; pl < PL::N; ++pl)
                                x[pl] = *reinterpret_cast<const V*>(sb + pl * G::PSTR + t * 16 * COLB);
                            if constexpr (PREC == 1) {
                                // planes: 0 Ar hi, 1 Ar lo, 2 Ai hi, 3 Ai lo; x: 0 BR hi, 1 BR lo, 2 BI hi, 3 BI lo
        #pragma unroll
                                for (int mt = 0; mt < MT; ++mt)
                                    acc[s][mt][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[mt][0], x[0], acc[s][mt][t], 0, 0, 0);
        #pragma unroll
                                for (int mt = 0; mt < MT; ++mt)
                                    acc[s][mt][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[mt][2], x[2], acc[s][mt][t], 0, 0, 0);
        #pragma unroll
                                for (int mt = 0; mt < MT; ++mt)
                                    acc[s][mt][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[mt][0], x[1], acc[s][mt][t], 0, 0, 0);
        #pragma unroll
                                for (int mt = 0; mt < MT; ++mt)
                                    acc[s][mt][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[mt][2], x[3], acc[s][mt][t], 0, 0, 0);
        #pragma unroll
                                for (int mt = 0; mt < MT; ++mt)
                                    acc[s][mt][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[mt][1], x[0], acc[s][mt][t], 0, 0, 0);
        #pragma unroll
                                for (int mt = 0; mt < MT; ++mt)
                                    acc[s][mt][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[mt][3], x[2], acc[s][mt][t], 0, 0, 0);
                            } else {
        #pragma unroll
                                for (int mt = 0; mt < MT; ++mt)
                                    acc[s][mt][t] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a[mt][0], x[0], acc[s][mt][t], 0, 0, 0);
        #pragma unroll
                                for (int mt = 0; mt < MT; ++mt)
                                    acc[s][mt][t] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a[mt][1], x[1], acc[s][mt][t], 0, 0, 0);
                            }
                        }
                    }
                }
            }

            if (more) store_x(xs + ((ks + 1) & 1) * BUF);
            __syncthreads();
        }
    } else {
        // Delays of the lane's 8 channels for its MT tiles: (int)(dm * delay[c])
        // in float (the conversion truncates), the reference's integer delay.
        constexpr int NU = NDIR;                // units (computed bins) per K-step
        constexpr bool BPREF = false;           // B fragments read one product group ahead
        uint32_t d[MT][8];
        float4 dlv[2];
        auto load_dl = [&](uint32_t ks) __attribute__((always_inline)) {
            const float* dp = dtab + ks * KS + 8 * g;
            dlv[0] = *reinterpret_cast<const float4*>(dp);
            dlv[1] = *reinterpret_cast<const float4*>(dp + 4);
        };
        auto make_d = [&]() __attribute__((always_inline)) {
            const float dl[8] = {dlv[0].x, dlv[0].y, dlv[0].z, dlv[0].w, dlv[1].x, dlv[1].y, dlv[1].z, dlv[1].w};
#pragma unroll
            for (int mt = 0; mt < MT; ++mt)
#pragma unroll
                for (int e = 0; e < 8; ++e) d[mt][e] = (uint32_t)(int)(dmv[mt] * dl[e]);
        };
        // Phase of bin kb as a float in [1, 2): (kb << (23 - lgn)) * d keeps
        // (kb*d mod nfft) / nfft in mantissa bits 22 .. 23-lgn (exact), and
        // v_cos_f32 / v_sin_f32 take revolutions, so 1 + phase is as good as
        // the phase itself: one multiply and one and-or per phasor.
        const uint32_t sh = 23u - (31u - (uint32_t)__builtin_clz(nfft));
        using APl = uint32_t[MT][PL::N][4];
        // piece pc = 4*mt + j of the phasor planes of bin kb: two channels of tile mt
        auto gen_piece = [&](uint32_t kb, APl& A, int pc) __attribute__((always_inline)) {
            const int mt = pc >> 2, j = pc & 3;
            const uint32_t kbs = kb << sh;
            float cr[2], ci[2];
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const uint32_t n = __umul24(kbs, d[mt][2 * j + h]);
                const float x = __builtin_bit_cast(float, (n & 0x7fffffu) | 0x3f800000u);
                cr[h] = __builtin_amdgcn_cosf(x);
                ci[h] = __builtin_amdgcn_sinf(x);
            }
            if constexpr (PREC == 1) {
                split_bf16x2_rn(cr[0], cr[1], A[mt][0][j], A[mt][1][j], dsel);
                split_bf16x2_rn(ci[0], ci[1], A[mt][2][j], A[mt][3][j], dsel);
            } else {
                A[mt][0][j] = pack_f16x2(cr[0], cr[1]);
                A[mt][1][j] = pack_f16x2(ci[0], ci[1]);
            }
        };
        // the B fragments (staged planes) of slot s, N-tile t in buffer cur
        using BFr = V[PL::N];
        auto read_b = [&](const char* cur, int s, int t, BFr& x) __attribute__((always_inline)) {
            const char* sb = cur + s * G::SSTR + b_off + t * 16 * COLB;
#pragma unroll
            for (int pl = 0; pl < PL::N; ++pl) x[pl] = *reinterpret_cast<const V*>(sb + pl * G::PSTR);
        };
        // the products of slot s, N-tile t with phasor planes A and fragments x
        auto group_products = [&](int s, int t, const APl& A, const BFr& x) __attribute__((always_inline)) {
            V a[MT][PL::N];
#pragma unroll
            for (int mt = 0; mt < MT; ++mt)
#pragma unroll
                for (int pl = 0; pl < PL::N; ++pl) a[mt][pl] = __builtin_bit_cast(V, A[mt][pl]);
            if constexpr (PREC == 1) {
                // planes: 0 Ar hi, 1 Ar lo, 2 Ai hi, 3 Ai lo; x: 0 BR hi, 1 BR lo, 2 BI hi, 3 BI lo
#pragma unroll
                for (int mt = 0; mt < MT; ++mt)
                    acc[s][mt][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[mt][0], x[0], acc[s][mt][t], 0, 0, 0);
#pragma unroll
                for (int mt = 0; mt < MT; ++mt)
                    acc[s][mt][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[mt][2], x[2], acc[s][mt][t], 0, 0, 0);
#pragma unroll
                for (int mt = 0; mt < MT; ++mt)
                    acc[s][mt][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[mt][0], x[1], acc[s][mt][t], 0, 0, 0);
#pragma unroll
                for (int mt = 0; mt < MT; ++mt)
                    acc[s][mt][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[mt][2], x[3], acc[s][mt][t], 0, 0, 0);
#pragma unroll
                for (int mt = 0; mt < MT; ++mt)
                    acc[s][mt][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[mt][1], x[0], acc[s][mt][t], 0, 0, 0);
#pragma unroll
                for (int mt = 0; mt < MT; ++mt)
                    acc[s][mt][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[mt][3], x[2], acc[s][mt][t], 0, 0, 0);
            } else {
#pragma unroll
                for (int mt = 0; mt < MT; ++mt)
                    acc[s][mt][t] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a[mt][0], x[0], acc[s][mt][t], 0, 0, 0);
#pragma unroll
                for (int mt = 0; mt < MT; ++mt)
                    acc[s][mt][t] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a[mt][1], x[1], acc[s][mt][t], 0, 0, 0);
            }
        };
        // The spectra are staged in two halves: items [0, IPT/2) loaded at the
        // top of the K-step and staged after unit 0, the rest loaded then and
        // staged after the last unit (half the staging registers live at a
        // time). Element offsets from X + kofs fit 32 bits (the launcher checks).
        constexpr int IH = (G::IPT + 1) / 2;  // items of the first half
        uint32_t xoff[G::IPT];
        // bin blocks of the bin-block layout (M/4 + 1, the grid's nbb)
        const uint32_t nbk = M / 4 + 1;
        if constexpr (BLK) {
            // item (slot s, chunk jj, channel pair cp), s fastest, then jj: a
            // wave's loads cover 16 chunks x 4 slots, 512 contiguous bytes
#pragma unroll
            for (int u = 0; u < G::IPT; ++u) {
                const int it = tid + u * NT;
                const int s = it & 3, jj = (it >> 2) % G::CG, cp = (it >> 2) / G::CG;
                const uint32_t chunk = ch0 + jj;
                xoff[u] = ((2 * cp * nbk + bblk) * nchunk + min(chunk, nchunk - 1)) * 4 + s;
                sscale[u] = (slot_valid(s) && chunk < nchunk) ? xscale : 0.f;
                soff[u] = s * G::SSTR + (16 * (jj / 8) + (jj % 8)) * COLB + 2 * PL::ELEM * cp;
            }
        } else {
#pragma unroll
            for (int u = 0; u < G::IPT; ++u) xoff[u] = (uint32_t)(src[u] - X);
        }
        // channel stride and K-step stride of the spectra, elements
        const uint32_t rs32 = BLK ? nbk * nchunk * 4 : (uint32_t)rs;
        const size_t kstep_p = BLK ? (size_t)KS * rs32 : kstep;
        float2 xh[IH][2];
        auto load_half = [&](size_t kofs, int p) __attribute__((always_inline)) {
            const float2* xb = X + kofs;
#pragma unroll
            for (int v = 0; v < IH; ++v)
                if (p * IH + v < G::IPT) {
                    xh[v][0] = xb[xoff[p * IH + v]];
                    xh[v][1] = xb[xoff[p * IH + v] + rs32];
                }
        };
        auto store_half = [&](char* buf, int p) __attribute__((always_inline)) {
#pragma unroll
            for (int v = 0; v < IH; ++v)
                if (p * IH + v < G::IPT) store_item(buf, p * IH + v, xh[v][0], xh[v][1]);
        };
        // unit u of K-step ks: the products of bin k0 + u (and its mirror) with
        // phasors Acur while the phasors of the next unit go into Anext
        auto unit = [&](uint32_t ks, int u, APl& Acur, APl& Anext) __attribute__((always_inline)) {
            const char* cur = xs + (ks & 1) * BUF;
            const bool last = u == NU - 1;
            // sign masks of the mirror phasors from this K-step's delays:
            // (-1)^d per channel, d odd <=> q odd
            uint32_t sm[MT][4];
            if constexpr (MIRROR) {
#pragma unroll
                for (int mt = 0; mt < MT; ++mt)
#pragma unroll
                    for (int j = 0; j < 4; ++j) sm[mt][j] = (d[mt][2 * j] & 1u) << 15 | (d[mt][2 * j + 1] << 31);
            }
            if (last) {
                // the next K-step's delays (loaded one K-step ago), then the
                // load of the one after (the last K-steps re-read the last)
                make_d();
                load_dl(min(ks + 2, nks - 1));
            }
            // the product groups (slot, N-tile) of the unit in order, the direct
            // slot's first; the pieces of the next unit's phasors spread
            // between them, so that they issue beside the matrix products
            constexpr int SPU = MIRROR ? 2 : 1;
            constexpr int NG = SPU * NTL, NPC = MT * 4;
            const uint32_t kbn = last ? k0 : k0 + u + 1;
            // B fragments one group ahead (BPREF), else read before use
            BFr xb[BPREF ? 2 : 1];
            auto gslot = [&](int gi) { return gi < NTL ? u : NDIR + u; };
            if (BPREF) read_b(cur, gslot(0), 0, xb[0]);
#pragma unroll
            for (int gi = 0; gi < NG; ++gi) {
                if (BPREF && gi + 1 < NG) read_b(cur, gslot(gi + 1), (gi + 1) % NTL, xb[(gi + 1) & 1]);
#pragma unroll
                for (int pc = gi * NPC / NG; pc < (gi + 1) * NPC / NG; ++pc) gen_piece(kbn, Anext, pc);
                if (MIRROR && gi == NTL) {
                    // the mirror slot: re' = (-1)^d re, im' = -(-1)^d im
#pragma unroll
                    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
                        for (int j = 0; j < 4; ++j)
#pragma unroll
                            for (int pl = 0; pl < PL::N; ++pl)
                                Acur[mt][pl][j] ^= sm[mt][j] ^ ((pl >= PL::N / 2) ? 0x80008000u : 0u);
                }
                if (!BPREF) read_b(cur, gslot(gi), gi % NTL, xb[0]);
                group_products(gslot(gi), gi % NTL, Acur, xb[BPREF ? (gi & 1) : 0]);
            }
            // the next K-step's spectra into the other buffer (the last K-step
            // stores its own again, unread)
            char* nbuf = xs + ((ks + 1) & 1) * BUF;
            if (u == 0) {
                store_half(nbuf, 0);
                load_half((size_t)((ks + 1 < nks) ? ks + 1 : ks) * kstep_p, 1);
            }
            if (last) store_half(nbuf, 1);
        };

        load_dl(0);
        make_d();
        load_dl(min(1u, nks - 1));
        load_half(0, 0);
        store_half(xs, 0);
        load_half(0, 1);
        store_half(xs, 1);
        APl A0, A1;
#pragma unroll
        for (int pc = 0; pc < MT * 4; ++pc) gen_piece(k0, A0, pc);
        __syncthreads();
        for (uint32_t ks = 0; ks < nks; ++ks) {
            const uint32_t kn = (ks + 1 < nks) ? ks + 1 : ks;
            load_half((size_t)kn * kstep_p, 0);
            unit(ks, 0, A0, A1);
            unit(ks, 1, A1, A0);
            if constexpr (NU == 4) {
                unit(ks, 2, A0, A1);
                unit(ks, 3, A1, A0);
            }
            __syncthreads();
        }
    }

    // ---- epilogue ----
    // lane l holds rows 4g + r (r < 4) and column c = l%16 of each tile: chunk
    // 8t + c%8, real part for c < 8, imaginary for c >= 8. Lanes l and l^8
    // swap values (DPP row rotate by 8), so both hold (re, im); the real-part
    // lane writes rows r = 0, 1 and its partner rows 2, 3.
    const bool re_lane = (lane & 8) == 0;
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
        for (int t = 0; t < NTL; ++t) {
            const uint32_t chunk = ch0 + 8 * t + (lane & 7);
            float vr[NS][2], vi[NS][2];
#pragma unroll
            for (int s = 0; s < NS; ++s)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const float own = acc[s][mt][t][r];
                    const float other = __builtin_bit_cast(
                        float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, own), 0x128, 0xf, 0xf, false));
                    // rows 0, 1 on the real lane, 2, 3 on the imaginary lane
                    if ((r < 2) == re_lane) {
                        vr[s][r & 1] = re_lane ? own : other;
                        vi[s][r & 1] = re_lane ? other : own;
                    }
                }
            if (chunk >= nchunk) continue;
#pragma unroll
            for (int rr = 0; rr < 2; ++rr) {
                const uint32_t row = 4 * g + (re_lane ? rr : 2 + rr);
                const uint32_t dmo = dmw + mt * 16 + row;
                if (dmo >= ndm) continue;
                float2* dst = Y + ((size_t)dmo * nchunk + chunk) * Yp;
                if constexpr (BLK) {
                    // the four packed values of the bin block, 32 contiguous bytes
                    float2 y[NS], z[NS];
#pragma unroll
                    for (int s = 0; s < NS; ++s) y[s] = make_float2(vr[s][rr] * yscale, vi[s][rr] * yscale);
                    auto zu = [&](uint32_t kk, float2 a, float2 m) -> float2 {
                        float2 b = make_float2(m.x, -m.y);
                        if (kk == 0) a.y = 0.f, b.y = 0.f;
                        const float2 e = make_float2(a.x + b.x, a.y + b.y);
                        const float2 d = make_float2(a.x - b.x, a.y - b.y);
                        const float2 w = zph[kk];
                        const float2 o = make_float2(d.x * w.x - d.y * w.y, d.x * w.y + d.y * w.x);
                        return make_float2(e.x - o.y, e.y + o.x);
                    };
#pragma unroll
                    for (int s = 0; s < NDIR; ++s) {
                        const uint32_t k = k0 + s;
                        z[s] = slot_valid(s) ? zu(k, y[s], k == M / 2 ? y[s] : y[s + NDIR]) : make_float2(0.f, 0.f);
                    }
#pragma unroll
                    for (int s = NDIR; s < NS; ++s) {
                        const uint32_t kk = M - (k0 + s - NDIR);
                        z[s] = (slot_valid(s) && kk < M) ? zu(kk, y[s], y[s - NDIR]) : make_float2(0.f, 0.f);
                    }
                    float4* o = reinterpret_cast<float4*>(Y + (((size_t)dmo * (M / 4 + 1) + bblk) * nchunk + chunk) * 4);
                    o[0] = make_float4(z[0].x, z[0].y, z[1].x, z[1].y);
                    o[1] = make_float4(z[2].x, z[2].y, z[3].x, z[3].y);
                    continue;
                }
                if constexpr (MIRROR) {
                    if (zph) {
                        // packed spectrum of the inverse transform:
                        // Zu[k] = (Y[k] + conj Y[M-k]) + i e^{2 pi i k/nfft} (Y[k] - conj Y[M-k]),
                        // k < M (k = 0: Re Y[0] and Re Y[M] only)
                        float2 y[NS];
#pragma unroll
                        for (int s = 0; s < NS; ++s) y[s] = make_float2(vr[s][rr] * yscale, vi[s][rr] * yscale);
                        auto zu = [&](uint32_t kk, float2 a, float2 m) -> float2 {
                            float2 b = make_float2(m.x, -m.y);
                            if (kk == 0) a.y = 0.f, b.y = 0.f;
                            const float2 e = make_float2(a.x + b.x, a.y + b.y);
                            const float2 d = make_float2(a.x - b.x, a.y - b.y);
                            const float2 w = zph[kk];
                            const float2 o = make_float2(d.x * w.x - d.y * w.y, d.x * w.y + d.y * w.x);
                            return make_float2(e.x - o.y, e.y + o.x);
                        };
#pragma unroll
                        for (int s = 0; s < NDIR; ++s) {
                            const uint32_t k = k0 + s;
                            if (slot_valid(s)) dst[k] = zu(k, y[s], k == M / 2 ? y[s] : y[s + NDIR]);
                        }
#pragma unroll
                        for (int s = NDIR; s < NS; ++s) {
                            const uint32_t kk = M - (k0 + s - NDIR);
                            if (slot_valid(s) && kk < M) dst[kk] = zu(kk, y[s], y[s - NDIR]);
                        }
                    } else {
#pragma unroll
                        for (int s = 0; s < NS; ++s)
                            if (slot_valid(s))
                                dst[slot_bin(s)] = make_float2(vr[s][rr] * yscale, vi[s][rr] * yscale);
                    }
                } else {
                    // consecutive bins k0 .. k0+NS-1: 16-byte stores of bin pairs
                    if (k0 + NS - 1 <= M) {
#pragma unroll
                        for (int s = 0; s < NS; s += 2)
                            *reinterpret_cast<float4*>(dst + k0 + s) =
                                make_float4(vr[s][rr] * yscale, vi[s][rr] * yscale, vr[s + 1][rr] * yscale,
                                            vi[s + 1][rr] * yscale);
                    } else {
#pragma unroll
                        for (int s = 0; s < NS; ++s)
                            if (slot_valid(s)) dst[k0 + s] = make_float2(vr[s][rr] * yscale, vi[s][rr] * yscale);
                    }
                }
            }
        }
}

}  // namespace

extern "C" int dd_fdd_mfma16_tiles(size_t nchunk, int prec, int* ntl_out, size_t* ngroups_out)
{
    // N-tiles per chunk group: fewest tiles plus the phasor cost of each group
    // (every group regenerates the same phasors), about 1.5 tiles' worth
    const int max_ntl = 3;
    (void)prec;
    double best = 1e300;
    int best_ntl = 1;
    for (int ntl = 1; ntl <= max_ntl; ++ntl) {
        const size_t groups = (nchunk + 8 * ntl - 1) / (8 * ntl);
        const double cost = (double)(groups * ntl) + 1.5 * (double)groups;
        if (cost < best || (cost == best && ntl > best_ntl)) best = cost, best_ntl = ntl;
    }
    if (ntl_out) *ntl_out = best_ntl;
    if (ngroups_out) *ngroups_out = (nchunk + 8 * best_ntl - 1) / (8 * best_ntl);
    return best_ntl;
}

extern "C" hipError_t dd_fdd_contract_mfma16(int prec, int mirror, const float2* X, size_t Xp, const uint32_t* q,
                                             size_t qpitch, size_t ndm, size_t nchan, size_t nchunk, size_t nfft,
                                             float xscale, float2* Y, size_t Yp, const float2* ph_packed,
                                             const float* dmlist, const float* delay32, int cfg, int blocked,
                                             hipStream_t stream)
{
    if (ph_packed && !mirror) return hipErrorInvalidValue;
    if ((dmlist == nullptr) != (delay32 == nullptr)) return hipErrorInvalidValue;
    if (!dmlist && !q) return hipErrorInvalidValue;
    if (ndm == 0 || nchunk == 0) return hipSuccess;
    // layout preconditions of the vector loads and stores
    if ((!dmlist && (qpitch % 32 || qpitch < (nchan + 31) / 32 * 32)) || (!blocked && Yp % 2) || nfft < 64 ||
        nfft > (size_t(1) << 18) || (nfft & (nfft - 1)) || (!blocked && Xp < nfft / 2 + 1))
        return hipErrorInvalidValue;
    if (prec != 0 && prec != 1 && prec != 2) return hipErrorInvalidValue;
    if (cfg < 0 || cfg > 2 || (cfg == 2 && (prec == 0 || !dmlist))) return hipErrorInvalidValue;
    if (blocked && (cfg != 2 || !mirror || !ph_packed || nfft > 65536)) return hipErrorInvalidValue;
    // the pipelined form addresses the spectra of a K-step by 32-bit element offsets
    const size_t rowp = blocked ? nfft / 2 + 4 : Xp;
    if (cfg == 2 && (size_t)KS * nchunk * rowp + rowp >= (size_t(1) << 32)) return hipErrorInvalidValue;
    const size_t M = nfft / 2;
    // cfg 0: 8 waves x 2 DM tiles, 4 bin slots; cfg 1: 4 waves x 4 DM tiles, 2
    // bin slots; cfg 2: pipelined, 8 waves x 2 DM tiles, 4 bin slots
    const size_t waves = (cfg == 1) ? 4 : 8, mt = (cfg == 1) ? 4 : 2, ns = (cfg == 1) ? 2 : 4;
    const size_t dpb = waves * mt * 16, ndir = mirror ? ns / 2 : ns;
    const uint32_t ndb = (uint32_t)((ndm + dpb - 1) / dpb);
    int ntl = 1;
    size_t ngroups = 0;
    dd_fdd_mfma16_tiles(nchunk, prec, &ntl, &ngroups);
    // bin blocks: ndir direct bins k <= M/2 (mirror) or ns consecutive bins
    const uint32_t nbb = (uint32_t)(mirror ? (M / 2 + ndir) / ndir : (M + ns) / ns);
    const size_t nblocks = (size_t)nbb * ndb * ngroups;
    if (nblocks >= (size_t(1) << 31)) return hipErrorInvalidValue;
    const float yscale = 1.0f / xscale;
#define DD_M16(P, T, MI, W, MT_, NS_, PI, BL)                                                                    \
    hipLaunchKernelGGL((fdd_mfma16_k<P, T, MI, W, MT_, NS_, PI, BL>), dim3((unsigned)nblocks), dim3(W * 64), 0,      \
                       stream, X, Xp, q, qpitch, (uint32_t)ndm, (uint32_t)nchan, (uint32_t)nchunk, (uint32_t)nfft,   \
                       xscale, yscale, Y, Yp, nbb, ndb, ph_packed, dmlist, delay32)
#define DD_M16_C(P, T, MI)                                                               \
    if (cfg == 1) DD_M16(P, T, MI, 4, 4, 2, false, false);                               \
    else if (cfg == 2 && blocked) DD_M16(P, T, MI, 8, 2, 4, P != 0, P != 0 && MI);       \
    else if (cfg == 2) DD_M16(P, T, MI, 8, 2, 4, P != 0, false);                         \
    else DD_M16(P, T, MI, 8, 2, 4, false, false);
#define DD_M16_T(P, MI)                    \
    switch (ntl) {                         \
    case 1: DD_M16_C(P, 1, MI) break;      \
    case 2: DD_M16_C(P, 2, MI) break;      \
    default: DD_M16_C(P, 3, MI) break;     \
    }
    if (prec == 0) {
        if (mirror) { DD_M16_T(0, true) } else { DD_M16_T(0, false) }
    } else if (prec == 1) {
        if (mirror) { DD_M16_T(1, true) } else { DD_M16_T(1, false) }
    } else {
        if (mirror) { DD_M16_T(2, true) } else { DD_M16_T(2, false) }
    }
#undef DD_M16_T
#undef DD_M16_C
#undef DD_M16
    return hipGetLastError();
}
