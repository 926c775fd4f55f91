// gfx950 kernels of the time-domain dedispersion (TDD) path.
//
//  dd_tdd_transpose      time-major packed input -> channel-major rows packed
//                        along time (one pass, LDS tile, zero fill of the tail)
//  tdd_generic_k         one thread per (DM, output sample); any nbits and
//                        integer killmask weights (exact uint32/uint64 sums)
//  tdd_lds_k             the fast path: a wave owns one DM at a time, so every
//                        per-(DM, channel) delay is wave-uniform (scalar
//                        registers); delay-shifted channel windows are staged in
//                        LDS; two aligned LDS words and one v_alignbit (or
//                        v_perm for 8-bit) give a lane 32/nbits consecutive
//                        samples; sums accumulate in packed 8-bit lanes, then
//                        16-bit lanes, then 32-bit (SWAR), bit-exact to the
//                        uint32 sums of the reference.
//
// Arithmetic restated from the reference: the integer delay
// __float2uint_rn(dm*delay[c]) (src/tdd/dedisperse/dedisperse_kernel.cuh:163-165)
// is precomputed on the host with the same float product and rounding; sums
// and the output epilogue follow dedisperse_kernel.cuh:45-87, 151-193.
#include <hip/hip_runtime.h>

#include "kernels.h"

#ifndef TDD_EARLY_READS
#define TDD_EARLY_READS 1
#endif

namespace {

// a word of global memory: loads through such pointers are global (vector
// memory only) rather than flat
typedef const __attribute__((address_space(1))) uint32_t gword;
// a word of LDS (the destination operand of global -> LDS loads)
typedef __attribute__((address_space(3))) uint32_t lword;

// ---------------------------------------------------------------------------
// Epilogue (src/tdd/dedisperse/dedisperse_kernel.cuh:45-87)
// ---------------------------------------------------------------------------

__device__ __forceinline__ float tdd_scale(float fsum, const dd_tdd_epilogue& ep)
{
    // (float)sum * out_range / (in_range * nchans) * factor, IEEE division
    float num = fsum * ep.out_range;
    float q = __fdiv_rn(num, ep.den);
    return q * ep.factor;
}

__device__ __forceinline__ void tdd_store(uint8_t* row, size_t t, float scaled,
                                          const dd_tdd_epilogue& ep)
{
    if (ep.out_nbits == 32) {
        reinterpret_cast<float*>(row)[t] = scaled;
    } else {
        float v = fminf(fmaxf(scaled, 0.f), ep.out_range);  // clip, then truncate
        if (ep.out_nbits == 16) reinterpret_cast<uint16_t*>(row)[t] = (uint16_t)__float2uint_rz(v);
        else row[t] = (uint8_t)__float2uint_rz(v);
    }
}

// ---------------------------------------------------------------------------
// Transpose: time-major packed rows -> channel-major packed-along-time rows
// ---------------------------------------------------------------------------

constexpr int TR_TS = 256;  // time samples per tile
constexpr int TR_CB = 128;  // channel bytes per tile: whole 128-byte lines of a sample row
constexpr int TR_PITCH = TR_CB + 4;

template <int NBITS>
__global__ __launch_bounds__(256) void tdd_transpose_k(
    const uint8_t* __restrict__ in, size_t nsamps, size_t in_stride, size_t row_bytes,
    size_t nchans, uint32_t* __restrict__ out, size_t row_words)
{
    constexpr int SPW = 32 / NBITS;                // samples per output word
    constexpr int CH_PER_TILE = TR_CB * 8 / NBITS; // channels per tile
    constexpr int WPC = TR_TS / SPW;               // output words per channel per tile
    __shared__ __attribute__((aligned(16))) uint8_t tile[TR_TS][TR_PITCH];

    const size_t t0 = (size_t)blockIdx.x * TR_TS;
    const size_t b0 = (size_t)blockIdx.y * TR_CB;
    const int tid = threadIdx.x;

    // Load TR_TS rows x TR_CB bytes (zero outside the input): 32-bit words
    // when the rows are word-aligned (32 lanes per 128-byte line), else bytes.
    const bool words = ((in_stride | reinterpret_cast<uintptr_t>(in)) & 3u) == 0;
    if (words) {
        for (int i = tid; i < TR_TS * (TR_CB / 4); i += 256) {
            const int tl = i / (TR_CB / 4), bl = 4 * (i % (TR_CB / 4));
            const size_t t = t0 + tl, b = b0 + bl;
            uint32_t v = 0;
            if (t < nsamps) {
                if (b + 3 < row_bytes) {
                    v = *reinterpret_cast<const uint32_t*>(in + t * in_stride + b);
                } else {
#pragma unroll
                    for (int q = 0; q < 4; ++q)
                        if (b + q < row_bytes) v |= (uint32_t)in[t * in_stride + b + q] << (8 * q);
                }
            }
            *reinterpret_cast<uint32_t*>(&tile[tl][bl]) = v;
        }
    } else {
        for (int i = tid; i < TR_TS * TR_CB; i += 256) {
            int tl = i / TR_CB, bl = i % TR_CB;
            size_t t = t0 + tl, b = b0 + bl;
            uint8_t v = 0;
            if (t < nsamps && b < row_bytes) v = in[t * in_stride + b];
            tile[tl][bl] = v;
        }
    }
    __syncthreads();

    const size_t c0 = (size_t)blockIdx.y * CH_PER_TILE;
    const size_t w0 = t0 / SPW;
    for (int i = tid; i < CH_PER_TILE * WPC; i += 256) {
        int cl = i / WPC, wl = i % WPC;
        size_t c = c0 + cl;
        if (c >= nchans || w0 + wl >= row_words) continue;
        uint32_t word = 0;
#pragma unroll
        for (int j = 0; j < SPW; ++j) {
            const uint8_t* r = tile[wl * SPW + j];
            uint32_t s;
            if (NBITS == 32) {
                s = (uint32_t)r[4 * cl] | ((uint32_t)r[4 * cl + 1] << 8) |
                    ((uint32_t)r[4 * cl + 2] << 16) | ((uint32_t)r[4 * cl + 3] << 24);
            } else if (NBITS == 16) {
                s = (uint32_t)r[2 * cl] | ((uint32_t)r[2 * cl + 1] << 8);
            } else if (NBITS == 8) {
                s = r[cl];
            } else {
                int bit = cl * NBITS;
                s = (r[bit >> 3] >> (bit & 7)) & ((1u << NBITS) - 1u);
            }
            word |= (NBITS == 32) ? s : (s << (j * NBITS));
        }
        out[c * row_words + w0 + wl] = word;
    }
}

// Transpose for 2-, 4- and 8-bit samples, where one 32-bit word of a sample
// row holds SPW = 32/NBITS channels and one output word SPW samples: a thread
// reads one word from each of SPW consecutive sample rows (an SPW x SPW matrix
// of NBITS-bit elements, row = sample, column = channel) and transposes it in
// registers, log2(SPW) stages that swap the off-diagonal blocks of each
// pair of rows (shift + bitfield insert per word), giving SPW output words.
// The tile is stored 32 words per sample row with word d of row t at
// d ^ ((t / SPW) & 31): the 32 lanes that read word d of rows SPW*wl + r for
// wl = 0..31 hit 32 different banks, and a row's stores stay a permutation.

constexpr uint32_t low_blocks_mask(int sh)
{
    // the low sh bits of every 2*sh-bit group
    uint32_t m = 0;
    for (int b = 0; b < 32; ++b)
        if ((b / sh) % 2 == 0) m |= 1u << b;
    return m;
}

template <int NBITS>
__global__ __launch_bounds__(256) void tdd_transpose_reg_k(
    const uint8_t* __restrict__ in, size_t nsamps, size_t in_stride, size_t row_bytes,
    size_t nchans, uint32_t* __restrict__ out, size_t row_words)
{
    static_assert(NBITS == 2 || NBITS == 4 || NBITS == 8, "register transpose: 2-, 4- or 8-bit");
    constexpr int SPW = 32 / NBITS;          // samples per output word = channels per row word
    constexpr int RW = TR_CB / 4;            // words per tile row (32)
    constexpr int WPC = TR_TS / SPW;         // output words per channel per tile
    __shared__ __attribute__((aligned(16))) uint32_t tile[TR_TS * RW];

    const size_t t0 = (size_t)blockIdx.x * TR_TS;
    const size_t b0 = (size_t)blockIdx.y * TR_CB;
    const int tid = threadIdx.x;
    auto slot = [](int t, int d) { return t * RW + (d ^ ((t / SPW) & (RW - 1))); };

    const bool words = ((in_stride | reinterpret_cast<uintptr_t>(in)) & 3u) == 0;
    if (words && t0 + TR_TS <= nsamps && b0 + TR_CB <= row_bytes) {
        // interior tile: every load in bounds, all issued before the first is
        // waited on (thread tid takes word tid % 32 of rows tid / 32 + 8k)
        constexpr int PER = TR_TS * RW / 256;
        const int d = tid % RW, tl0 = tid / RW;
        const uint8_t* src = in + (t0 + tl0) * in_stride + b0 + 4 * d;
        uint32_t v[PER];
#pragma unroll
        for (int k = 0; k < PER; ++k)
            v[k] = *reinterpret_cast<const uint32_t*>(src + (size_t)(8 * k) * in_stride);
#pragma unroll
        for (int k = 0; k < PER; ++k) tile[slot(tl0 + 8 * k, d)] = v[k];
    } else {
        for (int i = tid; i < TR_TS * RW; i += 256) {
            const int tl = i / RW, d = i % RW;
            const size_t t = t0 + tl, b = b0 + 4 * (size_t)d;
            uint32_t v = 0;
            if (t < nsamps) {
                if (words && b + 3 < row_bytes) {
                    v = *reinterpret_cast<const uint32_t*>(in + t * in_stride + b);
                } else {
#pragma unroll
                    for (int q = 0; q < 4; ++q)
                        if (b + q < row_bytes) v |= (uint32_t)in[t * in_stride + b + q] << (8 * q);
                }
            }
            tile[slot(tl, d)] = v;
        }
    }
    __syncthreads();

    const size_t c0 = (size_t)blockIdx.y * (TR_CB * 8 / NBITS);
    const size_t w0 = t0 / SPW;
    for (int i = tid; i < RW * WPC; i += 256) {
        const int wl = i % WPC, d = i / WPC;
        const size_t cb = c0 + (size_t)d * SPW;
        if (cb >= nchans || w0 + wl >= row_words) continue;
        uint32_t R[SPW];
#pragma unroll
        for (int r = 0; r < SPW; ++r) R[r] = tile[slot(wl * SPW + r, d)];
        // element q of R[r] (bits q*NBITS) is channel cb+q at sample SPW*wl+r;
        // after the stages it is channel cb+r at sample SPW*wl+q
#pragma unroll
        for (int B = SPW / 2; B >= 1; B /= 2) {
            const int sh = B * NBITS;
            const uint32_t m = low_blocks_mask(sh);
#pragma unroll
            for (int r = 0; r < SPW; ++r) {
                if (r & B) continue;
                const uint32_t a = R[r], b = R[r + B];
                R[r] = (a & m) | ((b << sh) & ~m);
                R[r + B] = ((a >> sh) & m) | (b & ~m);
            }
        }
        uint32_t* o = out + cb * row_words + w0 + wl;
        if (cb + SPW <= nchans) {
#pragma unroll
            for (int q = 0; q < SPW; ++q) o[q * row_words] = R[q];
        } else {
#pragma unroll
            for (int q = 0; q < SPW; ++q)
                if (cb + q < nchans) o[q * row_words] = R[q];
        }
    }
}

// ---------------------------------------------------------------------------
// Generic channel sum
// ---------------------------------------------------------------------------

template <int NBITS>
__global__ __launch_bounds__(256) void tdd_generic_k(
    const uint32_t* __restrict__ rows, size_t row_words,
    const uint32_t* __restrict__ delays, const int32_t* __restrict__ chan,
    const int32_t* __restrict__ weight, uint32_t nact, size_t nsamps_out,
    uint8_t* __restrict__ out, size_t out_stride, dd_tdd_epilogue ep)
{
    constexpr int SPW = 32 / NBITS;
    constexpr uint32_t MASK = (NBITS == 32) ? 0xffffffffu : ((1u << NBITS) - 1u);
    const size_t t = (size_t)blockIdx.x * 256 + threadIdx.x;
    const size_t i = blockIdx.y;
    if (t >= nsamps_out) return;
    const uint32_t* drow = delays + i * nact;
    uint32_t sum32 = 0;
    unsigned long long sum64 = 0;
    for (uint32_t j = 0; j < nact; ++j) {
        size_t ts = t + drow[j];
        uint32_t w = rows[(size_t)chan[j] * row_words + ts / SPW];
        uint32_t v = (NBITS == 32) ? w : ((w >> ((ts % SPW) * NBITS)) & MASK);
        uint32_t term = (uint32_t)weight[j] * v;  // int * unsigned wraps mod 2^32
        if (NBITS == 32) sum64 += term;
        else sum32 += term;
    }
    float fsum = (NBITS == 32) ? (float)sum64 : (float)sum32;
    tdd_store(out + i * out_stride, t, tdd_scale(fsum, ep), ep);
}

// ---------------------------------------------------------------------------
// Packed-lane LDS channel sum
// ---------------------------------------------------------------------------
//
// Tile: a workgroup owns 64*G*SPW consecutive output samples of WAVES*DPW DMs.
// Lane l of a wave owns the G input words l, l+64, .., l+64(G-1) of the tile
// (SPW samples each), so a wave's LDS reads are unit-stride.
// code[group][j][i] (group = DPW consecutive DMs of one wave, DM i of the
// group, channel j) = (LDS byte offset of the delayed window: staged channel
// slot j % chunk times the window plus the word offset inside it, times 4)
// << 5 | bit shift (sub-byte; v_alignbit reads only these 5 bits) or byte
// offset (8-bit). The 2*DPW codes of a channel pair are consecutive, so a
// wave fetches them with one scalar load from a pointer that only advances.
//
// Lane layout of the packed accumulators for one input word k:
//   P = 8/NBITS byte-lane registers m (P = 1 for 8-bit); byte b of register m
//   holds sample b*P + m of the word. acc16[k][m][h] holds the sums of bytes
//   h (low half) and h+2 (high half), i.e. samples P*h+m and P*(h+2)+m.

template <int NBITS, int G, int DPW, int WAVES, bool NEED32>
__global__ __launch_bounds__(WAVES * 64) void tdd_lds_k(
    const uint32_t* __restrict__ rows, size_t row_words,
    const uint32_t* __restrict__ code, const int32_t* __restrict__ chan,
    const uint32_t* __restrict__ wstart, uint32_t nact, uint32_t ndm,
    uint32_t nsamps_out, uint32_t ndmblk, uint32_t cc_max, uint32_t W,
    uint8_t* __restrict__ out, size_t out_stride, dd_tdd_epilogue ep,
    const dd_tdd_beam* __restrict__ beams, const uint32_t* __restrict__ tile_beam)
{
    constexpr int SPW = 32 / NBITS;
    constexpr int P = (NBITS < 8) ? 8 / NBITS : 1;
    constexpr int TILE_WORDS = 64 * G;
    constexpr uint32_t MAXV = (1u << NBITS) - 1u;
    // byte lanes hold at most 255: flush every FL8 channels (even count)
    constexpr uint32_t FL8 = (NBITS < 8) ? ((255u / MAXV) & ~1u) : 0xffffffffu;
    constexpr uint32_t M8 = (NBITS == 4) ? 0x0F0F0F0Fu : (NBITS == 2) ? 0x03030303u : 0x01010101u;

    extern __shared__ uint32_t lds[];

    // XCD-aware mapping: workgroups are dispatched round-robin over the 8 XCDs;
    // give each XCD a contiguous range of (time tile, DM block) pairs with the
    // DM block fastest, so the blocks sharing a time tile's input run on the
    // same L2.
    const uint32_t nb = gridDim.x, bid = blockIdx.x;
    const uint32_t xcd = bid & 7u, per = nb >> 3, rem = nb & 7u;
    const uint32_t L = xcd * per + min(xcd, rem) + (bid >> 3);
    const uint32_t tile_g = L / ndmblk, dblk = L % ndmblk;
    // batched launch: the time tile's beam supplies rows, output and length
    uint32_t tile = tile_g;
    if (beams) {
        const dd_tdd_beam bm = beams[tile_beam[tile_g]];
        tile = tile_g - bm.tile0;
        rows = bm.rows;
        row_words = bm.row_words;
        nsamps_out = bm.nsamps_out;
        out = bm.out;
        out_stride = bm.out_stride;
    }

    const int lane = threadIdx.x & 63;
    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint32_t tw0 = tile * TILE_WORDS;  // first input word of the tile
    const uint32_t dmb = dblk * (WAVES * DPW) + wave * DPW;

    uint32_t acc8[DPW][G][P];
    uint32_t acc16[DPW][G][P][2];
    uint32_t acc32[NEED32 ? DPW : 1][NEED32 ? G : 1][P][4];
#pragma unroll
    for (int i = 0; i < DPW; ++i)
#pragma unroll
        for (int k = 0; k < G; ++k)
#pragma unroll
            for (int m = 0; m < P; ++m) {
                acc8[i][k][m] = 0;
                acc16[i][k][m][0] = acc16[i][k][m][1] = 0;
            }
    if constexpr (NEED32) {
#pragma unroll
        for (int i = 0; i < DPW; ++i)
#pragma unroll
            for (int k = 0; k < G; ++k)
#pragma unroll
                for (int m = 0; m < P; ++m)
#pragma unroll
                    for (int q = 0; q < 4; ++q) acc32[i][k][m][q] = 0;
    }

    // byte lanes -> 16-bit lanes: bytes (0,2) and (1,3) moved by one v_perm each
    auto flush8 = [&]() {
#pragma unroll
        for (int i = 0; i < DPW; ++i)
#pragma unroll
            for (int k = 0; k < G; ++k)
#pragma unroll
                for (int m = 0; m < P; ++m) {
                    uint32_t a = acc8[i][k][m];
                    acc16[i][k][m][0] += __builtin_amdgcn_perm(0u, a, 0x0C020C00u);
                    acc16[i][k][m][1] += __builtin_amdgcn_perm(0u, a, 0x0C030C01u);
                    acc8[i][k][m] = 0;
                }
    };
    auto flush16 = [&]() {
        if constexpr (NEED32) {
#pragma unroll
            for (int i = 0; i < DPW; ++i)
#pragma unroll
                for (int k = 0; k < G; ++k)
#pragma unroll
                    for (int m = 0; m < P; ++m)
#pragma unroll
                        for (int h = 0; h < 2; ++h) {
                            uint32_t a = acc16[i][k][m][h];
                            acc32[i][k][m][h] += a & 0xFFFFu;
                            acc32[i][k][m][h + 2] += a >> 16;
                            acc16[i][k][m][h] = 0;
                        }
        }
    };

    // Adds the samples of one or two channels for DM i: each lane reads two
    // aligned LDS words per input word k and funnel-shifts them to the
    // delayed position (v_alignbit; v_perm for bytes). Two channels are summed
    // before touching the accumulator, so the adds fuse into v_add3_u32.
    auto extract = [&](const uint32_t* wbase, uint32_t shift, int k, int m) -> uint32_t {
        const uint32_t* p = wbase + lane + 64 * k;
        uint32_t w0 = p[0], w1 = p[1];
        if constexpr (NBITS == 8) {
            // byte offset (low 5 bits of the code) -> selectors for bytes (r, r+2) and (r+1, r+3)
            uint32_t sel = 0x0C000C00u + (shift & 31u) * 0x00010001u + 0x00020000u + (uint32_t)m * 0x00010001u;
            return __builtin_amdgcn_perm(w1, w0, sel);
        } else {
            uint32_t x = __builtin_amdgcn_alignbit(w1, w0, shift);
            return (x >> (m * NBITS)) & M8;
        }
    };
    auto accumulate2 = [&](int i, const uint32_t* b0, uint32_t s0, const uint32_t* b1, uint32_t s1) {
#pragma unroll
        for (int k = 0; k < G; ++k) {
            if constexpr (NBITS == 8) {
#pragma unroll
                for (int h = 0; h < 2; ++h)
                    acc16[i][k][0][h] += extract(b0, s0, k, h) + extract(b1, s1, k, h);
            } else {
                const uint32_t* p0 = b0 + lane + 64 * k;
                const uint32_t* p1 = b1 + lane + 64 * k;
                // the funnel shift takes its amount modulo 32: the code's low 5 bits
                uint32_t x0 = __builtin_amdgcn_alignbit(p0[1], p0[0], s0);
                uint32_t x1 = __builtin_amdgcn_alignbit(p1[1], p1[0], s1);
#pragma unroll
                for (int m = 0; m < P; ++m)
                    acc8[i][k][m] += ((x0 >> (m * NBITS)) & M8) + ((x1 >> (m * NBITS)) & M8);
            }
        }
    };
    auto accumulate1 = [&](int i, const uint32_t* b0, uint32_t s0) {
#pragma unroll
        for (int k = 0; k < G; ++k) {
            if constexpr (NBITS == 8) {
#pragma unroll
                for (int h = 0; h < 2; ++h) acc16[i][k][0][h] += extract(b0, s0, k, h);
            } else {
#pragma unroll
                for (int m = 0; m < P; ++m) acc8[i][k][m] += extract(b0, s0, k, m);
            }
        }
    };

    // This wave's rows of the code table (wave-uniform: scalar loads).
    const uint32_t coff = __builtin_amdgcn_readfirstlane((dblk * WAVES + wave) * nact * DPW);
    const uint32_t wrow = __builtin_amdgcn_readfirstlane(dblk * nact);
    auto window = [&](uint32_t c) { return reinterpret_cast<const uint32_t*>(reinterpret_cast<const char*>(lds) + (c >> 5)); };

    // channels per byte-lane group (sub-byte): the byte lanes are widened
    // after every group; 8-bit input goes straight to 16-bit lanes
    constexpr uint32_t GROUP = (NBITS < 8) ? FL8 : 0x7fffffffu;
    for (uint32_t j0 = 0; j0 < nact; j0 += cc_max) {
        const uint32_t cc = min(cc_max, nact - j0);
        __syncthreads();
        // Stage the delayed windows of channels j0 .. j0+cc-1, a wave taking
        // channels wave, wave+WAVES, ...: global -> LDS loads (LDS address =
        // the wave-uniform row base + 4*lane), which need no registers and no
        // wait before the barrier, so every staging load of the wave is in
        // flight at once.
        for (uint32_t jj = wave; jj < cc; jj += WAVES) {
            const uint32_t j = j0 + jj;
            // wave-uniform row start; lane is the only per-lane part of the
            // address, and only the last, partial 64 words need a bound
            const gword* src = (const gword*)rows + (size_t)chan[j] * row_words + tw0 + wstart[wrow + j];
            lword* dst = (lword*)(lds + jj * W);
            uint32_t w0 = 0;
            for (; w0 + 64 <= W; w0 += 64) __builtin_amdgcn_global_load_lds(src + w0 + lane, dst + w0, 4, 0, 0);
            if (lane < W - w0) __builtin_amdgcn_global_load_lds(src + w0 + lane, dst + w0, 4, 0, 0);
        }
        __syncthreads();

        // Adds the two channels of a pair whose codes are c (2*DPW of them).
        auto pair = [&](const uint32_t (&c)[2 * DPW]) {
            if constexpr (NBITS < 8 && TDD_EARLY_READS) {
                // all LDS words of the pair first, so their latencies overlap
                uint32_t wv[2 * DPW][G][2];
#pragma unroll
                for (int t = 0; t < 2 * DPW; ++t)
#pragma unroll
                    for (int k = 0; k < G; ++k) {
                        const uint32_t* q = window(c[t]) + lane + 64 * k;
                        wv[t][k][0] = q[0];
                        wv[t][k][1] = q[1];
                    }
#pragma unroll
                for (int i = 0; i < DPW; ++i)
#pragma unroll
                    for (int k = 0; k < G; ++k) {
                        const uint32_t x0 = __builtin_amdgcn_alignbit(wv[i][k][1], wv[i][k][0], c[i]);
                        const uint32_t x1 = __builtin_amdgcn_alignbit(wv[DPW + i][k][1], wv[DPW + i][k][0], c[DPW + i]);
#pragma unroll
                        for (int m = 0; m < P; ++m)
                            acc8[i][k][m] += ((x0 >> (m * NBITS)) & M8) + ((x1 >> (m * NBITS)) & M8);
                    }
            } else {
#pragma unroll
                for (int i = 0; i < DPW; ++i)
                    accumulate2(i, window(c[i]), c[i], window(c[DPW + i]), c[DPW + i]);
            }
        };
        // codes of pair jj (scalar loads; LDS and scalar loads share one
        // counter, so the first wait of a pair covers the prefetch too)
        auto load_codes = [&](uint32_t (&c)[2 * DPW], uint32_t jj) {
            const uint32_t* cp = code + (coff + (j0 + jj) * DPW);
#pragma unroll
            for (int t = 0; t < 2 * DPW; ++t) c[t] = cp[t];
        };

        for (uint32_t g0 = 0; g0 < cc; g0 += GROUP) {
            const uint32_t g1 = min(cc, g0 + GROUP);
            uint32_t jj = g0;
            // two pairs per iteration with two code sets, each fetched one pair
            // ahead; the last prefetch may read one pair past the group (the
            // code table is padded, so it stays in bounds and is not used)
            uint32_t ca[2 * DPW], cb[2 * DPW];
            if (jj + 1 < g1) load_codes(ca, jj);
            // the first codes are in registers before the loop (lgkmcnt(0)), so
            // no path into the loop leaves a scalar load to wait for at its top
            __builtin_amdgcn_s_waitcnt(0xc07f);
            for (; jj + 3 < g1; jj += 4) {
                load_codes(cb, jj + 2);
                __builtin_amdgcn_sched_barrier(0);
                pair(ca);
                load_codes(ca, jj + 4);
                __builtin_amdgcn_sched_barrier(0);
                pair(cb);
            }
            if (jj + 1 < g1) {
                pair(ca);
                jj += 2;
            }
            if (jj < g1) {
                const uint32_t* cp = code + (coff + (j0 + jj) * DPW);
#pragma unroll
                for (int i = 0; i < DPW; ++i) {
                    const uint32_t c0 = cp[i];
                    accumulate1(i, window(c0), c0);
                }
            }
            if constexpr (NBITS < 8) flush8();
        }
        flush16();
    }

    // Epilogue: lane l owns samples (l + 64k)*SPW + s of the tile.
    const uint32_t t_tile = tile * TILE_WORDS * SPW;
#pragma unroll
    for (int i = 0; i < DPW; ++i) {
        const uint32_t dm = dmb + i;
        if (dm >= ndm) continue;
        uint8_t* orow = out + (size_t)dm * out_stride;
#pragma unroll
        for (int k = 0; k < G; ++k) {
            const uint32_t tbase = t_tile + (uint32_t)(lane + 64 * k) * SPW;
#pragma unroll
            for (int s = 0; s < SPW; ++s) {
                // sample s of word k: s = P*(h + 2*half) + m
                const int m = s % P, q = s / P;   // q = h + 2*half
                const int h = q & 1, half = q >> 1;
                uint32_t sum;
                if constexpr (NEED32) sum = acc32[i][k][m][h + 2 * half];
                else sum = half ? (acc16[i][k][m][h] >> 16) : (acc16[i][k][m][h] & 0xFFFFu);
                const uint32_t t = tbase + s;
                if (t < nsamps_out) tdd_store(orow, t, tdd_scale((float)sum, ep), ep);
            }
        }
    }
}

template <int NBITS, int G, int DPW, int WAVES, bool NEED32>
hipError_t launch_lds(const uint32_t* rows, size_t row_words, const uint32_t* code,
                      const int32_t* chan, const uint32_t* wstart, size_t nact, size_t ndm,
                      size_t nsamps_out, uint8_t* out, size_t out_stride,
                      dd_tdd_epilogue ep, const dd_tdd_tiling& tl, hipStream_t stream,
                      const dd_tdd_beam* beams = nullptr, const uint32_t* tile_beam = nullptr,
                      size_t batch_tiles = 0)
{
    constexpr int SPW = 32 / NBITS;
    const size_t tile_samps = (size_t)64 * G * SPW;
    const size_t ntiles = beams ? batch_tiles : (nsamps_out + tile_samps - 1) / tile_samps;
    const size_t ndmblk = (ndm + WAVES * DPW - 1) / (WAVES * DPW);
    const size_t nblocks = ntiles * ndmblk;
    const size_t lds_bytes = (size_t)tl.chunk_channels * tl.window_words * 4;
    auto kern = tdd_lds_k<NBITS, G, DPW, WAVES, NEED32>;
    if (lds_bytes > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute((const void*)kern,
                                           hipFuncAttributeMaxDynamicSharedMemorySize,
                                           (int)lds_bytes);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(kern, dim3((unsigned)nblocks), dim3(WAVES * 64), lds_bytes, stream,
                       rows, row_words, code, chan, wstart, (uint32_t)nact, (uint32_t)ndm,
                       (uint32_t)nsamps_out, (uint32_t)ndmblk, tl.chunk_channels,
                       tl.window_words, out, out_stride, ep, beams, tile_beam);
    return hipGetLastError();
}

}  // namespace

extern "C" hipError_t dd_tdd_transpose(const uint8_t* in, size_t nsamps, size_t in_stride,
                                       unsigned nbits, size_t nchans, uint32_t* out,
                                       size_t row_words, hipStream_t stream)
{
    const size_t row_bytes = (nchans * nbits + 7) / 8;
    const size_t spw = 32 / nbits;
    const size_t total_samps = row_words * spw;
    dim3 grid((unsigned)((total_samps + TR_TS - 1) / TR_TS),
              (unsigned)((row_bytes + TR_CB - 1) / TR_CB));
    const size_t rb = (in_stride < row_bytes) ? in_stride : row_bytes;
    switch (nbits) {
    case 1:  hipLaunchKernelGGL(tdd_transpose_k<1>, grid, dim3(256), 0, stream, in, nsamps, in_stride, rb, nchans, out, row_words); break;
    case 2:  hipLaunchKernelGGL(tdd_transpose_reg_k<2>, grid, dim3(256), 0, stream, in, nsamps, in_stride, rb, nchans, out, row_words); break;
    case 4:  hipLaunchKernelGGL(tdd_transpose_reg_k<4>, grid, dim3(256), 0, stream, in, nsamps, in_stride, rb, nchans, out, row_words); break;
    case 8:  hipLaunchKernelGGL(tdd_transpose_reg_k<8>, grid, dim3(256), 0, stream, in, nsamps, in_stride, rb, nchans, out, row_words); break;
    case 16: hipLaunchKernelGGL(tdd_transpose_k<16>, grid, dim3(256), 0, stream, in, nsamps, in_stride, rb, nchans, out, row_words); break;
    case 32: hipLaunchKernelGGL(tdd_transpose_k<32>, grid, dim3(256), 0, stream, in, nsamps, in_stride, rb, nchans, out, row_words); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

extern "C" hipError_t dd_tdd_dedisperse_generic(const uint32_t* rows, size_t row_words, unsigned nbits,
                                                const uint32_t* delays, const int32_t* chan,
                                                const int32_t* weight, size_t nact, size_t ndm,
                                                size_t nsamps_out, uint8_t* out, size_t out_stride,
                                                dd_tdd_epilogue ep, hipStream_t stream)
{
    if (nsamps_out == 0 || ndm == 0) return hipSuccess;
    dim3 grid((unsigned)((nsamps_out + 255) / 256), (unsigned)ndm);
#define DD_GENERIC(NB)                                                                         \
    hipLaunchKernelGGL(tdd_generic_k<NB>, grid, dim3(256), 0, stream, rows, row_words, delays, \
                       chan, weight, (uint32_t)nact, nsamps_out, out, out_stride, ep)
    switch (nbits) {
    case 1:  DD_GENERIC(1); break;
    case 2:  DD_GENERIC(2); break;
    case 4:  DD_GENERIC(4); break;
    case 8:  DD_GENERIC(8); break;
    case 16: DD_GENERIC(16); break;
    case 32: DD_GENERIC(32); break;
    default: return hipErrorInvalidValue;
    }
#undef DD_GENERIC
    return hipGetLastError();
}

// Instantiated tilings: (nbits, words per lane G, DMs per wave, waves).
#define DD_LDS_TILINGS(X) \
    X(1, 1, 2, 4) X(1, 1, 4, 4) X(1, 1, 2, 8) \
    X(2, 1, 2, 4) X(2, 2, 2, 4) X(2, 1, 4, 4) X(2, 1, 2, 8) \
    X(4, 2, 2, 4) X(4, 2, 4, 4) X(4, 1, 4, 4) X(4, 2, 2, 8) X(4, 4, 2, 4) X(4, 2, 4, 8) \
    X(4, 2, 1, 4) X(4, 1, 2, 8) \
    X(8, 2, 2, 4) X(8, 4, 2, 4) X(8, 2, 4, 4) X(8, 2, 2, 8)

extern "C" int dd_tdd_lds_tilings(unsigned nbits, dd_tdd_tiling* out, int max)
{
    int n = 0;
    if (nbits == 4) {
        // the runs-layout kernel (tdd_runs.hip)
        if (n < max && out) out[n] = dd_tdd_tiling{DD_TDD_RUNS_SPL, 4, 4, 16, 0};
        ++n;
        if (n < max && out) out[n] = dd_tdd_tiling{DD_TDD_RUNS_SPL, 2, 8, 16, 0};
        ++n;
    }
#define DD_LIST(NB, GG, DPW, WAVES)                                        \
    if (nbits == NB) {                                                     \
        if (n < max && out) {                                              \
            out[n] = dd_tdd_tiling{(unsigned)(GG * (32 / NB)), DPW, WAVES, 16, 0}; \
        }                                                                  \
        ++n;                                                               \
    }
    DD_LDS_TILINGS(DD_LIST)
#undef DD_LIST
    return n;
}

extern "C" int dd_tdd_lds_supported(unsigned nbits, dd_tdd_tiling tl)
{
    if (dd_tdd_tiling_is_runs(nbits, tl))
        return (tl.dms_per_wave == 4 && tl.waves == 4) || (tl.dms_per_wave == 2 && tl.waves == 8);
    const unsigned spw = 32 / nbits;
    if (tl.samples_per_lane % spw) return 0;
    const unsigned G = tl.samples_per_lane / spw;
#define DD_MATCH(NB, GG, DPW, WAVES) \
    if (nbits == NB && G == GG && tl.dms_per_wave == DPW && tl.waves == WAVES) return 1;
    DD_LDS_TILINGS(DD_MATCH)
#undef DD_MATCH
    return 0;
}

extern "C" hipError_t dd_tdd_dedisperse_lds(const uint32_t* rows, size_t row_words, unsigned nbits,
                                            const uint32_t* code, const int32_t* chan,
                                            const uint32_t* wstart, const uint32_t* chunks, unsigned maxc,
                                            size_t nact, size_t ndm,
                                            size_t nsamps_out, uint8_t* out, size_t out_stride,
                                            dd_tdd_epilogue ep, dd_tdd_tiling tl,
                                            hipStream_t stream)
{
    if (nsamps_out == 0 || ndm == 0) return hipSuccess;
    if (dd_tdd_tiling_is_runs(nbits, tl))
        return dd_tdd_dedisperse_runs(rows, row_words, code, chan, wstart, chunks, maxc, nact, ndm, nsamps_out,
                                      out, out_stride, ep, tl.dms_per_wave, tl.waves, nullptr, nullptr, 0,
                                      stream);
    const unsigned spw = 32 / nbits;
    const unsigned G = tl.samples_per_lane / spw;
    const uint32_t maxv = (1u << nbits) - 1u;
    // 16-bit lanes overflow beyond 65535: then sums are widened every staged chunk
    const bool need32 = (uint64_t)nact * maxv > 65535u;
#define DD_CALL(NB, GG, DPW, WAVES)                                                             \
    if (nbits == NB && G == GG && tl.dms_per_wave == DPW && tl.waves == WAVES) {                \
        if (need32)                                                                             \
            return launch_lds<NB, GG, DPW, WAVES, true>(rows, row_words, code, chan, wstart,    \
                                                        nact, ndm, nsamps_out, out, out_stride, \
                                                        ep, tl, stream);                        \
        return launch_lds<NB, GG, DPW, WAVES, false>(rows, row_words, code, chan, wstart, nact, \
                                                     ndm, nsamps_out, out, out_stride, ep, tl,  \
                                                     stream);                                   \
    }
    DD_LDS_TILINGS(DD_CALL)
#undef DD_CALL
    return hipErrorInvalidValue;
}

extern "C" hipError_t dd_tdd_dedisperse_lds_batch(const dd_tdd_beam* beams, const uint32_t* tile_beam,
                                                  size_t ntiles, unsigned nbits, const uint32_t* code,
                                                  const int32_t* chan, const uint32_t* wstart,
                                                  const uint32_t* chunks, unsigned maxc, size_t nact,
                                                  size_t ndm, dd_tdd_epilogue ep, dd_tdd_tiling tl,
                                                  hipStream_t stream)
{
    if (ntiles == 0 || ndm == 0) return hipSuccess;
    if (!beams || !tile_beam) return hipErrorInvalidValue;
    if (dd_tdd_tiling_is_runs(nbits, tl))
        return dd_tdd_dedisperse_runs(nullptr, 0, code, chan, wstart, chunks, maxc, nact, ndm, 0, nullptr, 0, ep,
                                      tl.dms_per_wave, tl.waves, beams, tile_beam, ntiles, stream);
    const unsigned spw = 32 / nbits;
    const unsigned G = tl.samples_per_lane / spw;
    const uint32_t maxv = (1u << nbits) - 1u;
    const bool need32 = (uint64_t)nact * maxv > 65535u;
#define DD_CALL(NB, GG, DPW, WAVES)                                                              \
    if (nbits == NB && G == GG && tl.dms_per_wave == DPW && tl.waves == WAVES) {                 \
        if (need32)                                                                              \
            return launch_lds<NB, GG, DPW, WAVES, true>(nullptr, 0, code, chan, wstart, nact,    \
                                                        ndm, 0, nullptr, 0, ep, tl, stream,      \
                                                        beams, tile_beam, ntiles);               \
        return launch_lds<NB, GG, DPW, WAVES, false>(nullptr, 0, code, chan, wstart, nact, ndm,  \
                                                     0, nullptr, 0, ep, tl, stream, beams,       \
                                                     tile_beam, ntiles);                         \
    }
    DD_LDS_TILINGS(DD_CALL)
#undef DD_CALL
    return hipErrorInvalidValue;
}
