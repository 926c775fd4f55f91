// Channel-sum kernel of the TDD path for 4-bit input in the "runs" layout.
//
// The time tile of a workgroup is TILE = 7 * T output samples (T = 256), split
// into 7 runs of T consecutive samples. In LDS one 32-bit word holds, at one
// position p (0 <= p < T + span), the sample p of each run: nibble r = sample
// t0 + a + p + r*T of the channel (a: the window start of the channel in the
// DM block), nibble 7 = 0. A DM whose delay is d then reads, for its output
// position m, the word at position d - a + m: every delay is a whole-word
// offset, so the inner loop has no funnel shift and one LDS word gives 7
// samples. The packed-lane kernel (tdd_lds_k) reads two words and shifts them
// for every 8 samples instead.
//
// Sums (bit-exact to the uint32 sums of src/tdd/dedisperse/dedisperse_kernel.cuh:151-193):
// for a pair of channels with words x0, x1
//     lo  += (x0 & 0x0F0F0F0F) + (x1 & 0x0F0F0F0F)    runs 0,2,4,6 in bytes 0..3
//     all += x0 + x1                                  every nibble, mod 2^32
// After at most 16 channels all - lo = sum of (x & 0xF0F0F0F0): runs 1,3,5 at
// bits 4..11, 12..19, 20..27 (each sum <= 240 < 256; nibble 7 being 0 keeps
// the sum below 2^28, so the subtraction is exact), and both are widened into
// 16-bit lanes. Two channels cost one v_and each and two v_add3_u32.
//
// Staging per chunk of channels (one LDS buffer pair per workgroup):
//  1. global -> LDS loads of each channel's time-ordered window (6T + P
//     samples, P = positions of the chunk) into lds1, 64 words per load with 8
//     words of skew after each (so the run rows of a transposing read spread
//     over the LDS banks: at most 5 of its 128 words share a bank);
//  2. ds_read_b64_tr_b4: in each group of 16 lanes, lane n supplies the
//     address of row n (16 samples) and lane j receives nibble j of every row,
//     row n at nibble n. Rows 0..6 are runs 0..6 at one position block, rows
//     8..14 the same runs 64 positions later, rows 7 and 15 an all-zero row:
//     one read makes the run words of 128 positions, written to lds2;
//  3. the next chunk's global -> LDS loads are issued, then the sums of this
//     chunk run from lds2 while they are in flight.
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace {

typedef const __attribute__((address_space(1))) uint32_t gword;
typedef __attribute__((address_space(3))) uint32_t lword;
typedef int v2i __attribute__((ext_vector_type(2)));
typedef __attribute__((address_space(3))) v2i lv2i;

constexpr int RUNS = 7;
constexpr int G = 4;                   // positions per lane
constexpr int T = 64 * G;              // positions per tile = run length (samples)
constexpr int TILE = RUNS * T;         // output samples per tile
constexpr int PIECE = 64;              // words per global -> LDS load
constexpr int PSTRIDE = 72;            // LDS words per staged piece (8 words of skew)
constexpr int L1W = DD_TDD_RUNS_L1_BYTES / 4;
constexpr int L2W = DD_TDD_RUNS_L2_BYTES / 4;
constexpr uint32_t FLUSH = 16;         // channels per byte-lane group
constexpr int CODES = DD_TDD_RUNS_MAXCC * 4 + 8;  // LDS words of codes per 4 DMs (a wave's or two waves')
constexpr int L1Z = L1W;                          // the zero row (2 words) in lds1
constexpr int L1C = L1W + 4;                      // first word of the staged codes in lds1
constexpr uint32_t M4 = 0x0F0F0F0Fu;

static_assert(T == 256 && PIECE == 64, "the transposing-read constants assume runs of 32 words, pieces of 64");
static_assert(PSTRIDE % 2 == 0, "transposing reads need 8-byte aligned rows");

__device__ __forceinline__ float tdd_scale(float fsum, const dd_tdd_epilogue& ep)
{
    float num = fsum * ep.out_range;
    float q = __fdiv_rn(num, ep.den);
    return q * ep.factor;
}

__device__ __forceinline__ void tdd_store(uint8_t* row, size_t t, float scaled, const dd_tdd_epilogue& ep)
{
    if (ep.out_nbits == 32) {
        reinterpret_cast<float*>(row)[t] = scaled;
    } else {
        float v = fminf(fmaxf(scaled, 0.f), ep.out_range);  // clip, then truncate
        if (ep.out_nbits == 16) reinterpret_cast<uint16_t*>(row)[t] = (uint16_t)__float2uint_rz(v);
        else row[t] = (uint8_t)__float2uint_rz(v);
    }
}

// chunk entry: first channel (bits 0..15), channels (16..23), positions / 128 (24..31)
__device__ __forceinline__ uint32_t chunk_j0(uint32_t e) { return e & 0xffffu; }
__device__ __forceinline__ uint32_t chunk_cc(uint32_t e) { return (e >> 16) & 0xffu; }
__device__ __forceinline__ uint32_t chunk_p(uint32_t e) { return (e >> 24) * 128u; }

// DPW DMs per wave, WAVES waves: a block of DPW * WAVES DMs. MINW: waves per
// SIMD the register allocation must allow.
template <int DPW, int WAVES, int MINW>
__global__ __launch_bounds__(WAVES * 64, MINW) void tdd_runs_k(
    const uint32_t* __restrict__ rows, size_t row_words, const uint32_t* __restrict__ code,
    const int32_t* __restrict__ chan, const uint32_t* __restrict__ wstart,
    const uint32_t* __restrict__ chunks, uint32_t maxc, uint32_t nact, uint32_t ndm,
    uint32_t nsamps_out, uint32_t ndmblk, uint8_t* __restrict__ out, size_t out_stride,
    dd_tdd_epilogue ep, const dd_tdd_beam* __restrict__ beams, const uint32_t* __restrict__ tile_beam)
{
    constexpr int DPB = DPW * WAVES;                 // DMs per block
    constexpr int CW = DD_TDD_RUNS_MAXCC * DPW + 8;  // LDS words of codes per wave
    // lds1: time-ordered windows (+ one zero row of 8 bytes at the end);
    // lds2: run words, P per channel of the chunk
    __shared__ __attribute__((aligned(16))) uint32_t lds1[L1W + 4 + WAVES * CW];
    __shared__ __attribute__((aligned(16))) uint32_t lds2[L2W];
    // codes of the chunk, per wave: staged into lds1 beside the windows, then
    // copied here by their wave (so no load into LDS that is in flight during
    // the sums can alias what the sums read)
    __shared__ __attribute__((aligned(16))) uint32_t ldsc[WAVES * CW];

    // XCD-aware mapping: workgroups go round-robin over the 8 XCDs; each XCD
    // gets a contiguous range of (time tile, DM block) pairs, DM block fastest,
    // so the blocks sharing a tile's input run on the same L2.
    const uint32_t nb = gridDim.x, bid = blockIdx.x;
    const uint32_t xcd = bid & 7u, per = nb >> 3, rem = nb & 7u;
    const uint32_t L = xcd * per + min(xcd, rem) + (bid >> 3);
    const uint32_t tile_g = L / ndmblk, dblk = L % ndmblk;
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
    const uint32_t tw0 = tile * (TILE / 8);  // first input word of the tile
    if (threadIdx.x < 2) lds1[L1Z + threadIdx.x] = 0u;
    // the sums read the codes one pair ahead, past a chunk's last pair too:
    // such words must be in-range offsets (0) before any chunk wrote them
    for (uint32_t q = threadIdx.x; q < WAVES * CW; q += WAVES * 64) ldsc[q] = 0u;

    // Per-lane constants of the transposing reads. Lane n of a group of 16
    // supplies row n: run n & 7 of position block 16g (g = lane >> 4), 64
    // positions later for n >= 8; n = 7, 15: the zero row. A unit u = 4a + b
    // (128 positions) reads time-ordered word w = 16u + 2g + 8(n >> 3) +
    // 32(n & 7), at LDS word w + 8 floor(w / 64) of the channel's slot, i.e.
    // 72a + rk[b].
    const uint32_t rrow = lane & 15, rgrp = lane >> 4;
    const bool rdata = (rrow & 7) != 7;
    const uint32_t rmask = rdata ? 0xffffffffu : 0u;
    uint32_t rk[4];
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        const uint32_t v = 16 * b + 2 * rgrp + 8 * (rrow >> 3) + 32 * (rrow & 7);
        rk[b] = rdata ? (v + 8 * (v / PIECE)) * 4u : (uint32_t)L1Z * 4u;
    }

    uint32_t lo8[DPW][G], all8[DPW][G], s16[DPW][G][4];
#pragma unroll
    for (int i = 0; i < DPW; ++i)
#pragma unroll
        for (int k = 0; k < G; ++k) {
            lo8[i][k] = all8[i][k] = 0;
#pragma unroll
            for (int q = 0; q < 4; ++q) s16[i][k][q] = 0;
        }

    // byte lanes -> 16-bit lanes: s16[0] = runs (0, 4), [1] = (2, 6), [2] =
    // (1, 5), [3] = (3, -)
    auto flush = [&]() {
#pragma unroll
        for (int i = 0; i < DPW; ++i)
#pragma unroll
            for (int k = 0; k < G; ++k) {
                const uint32_t lo = lo8[i][k];
                const uint32_t hi = (all8[i][k] - lo) >> 4;
                s16[i][k][0] += __builtin_amdgcn_perm(0u, lo, 0x0C020C00u);
                s16[i][k][1] += __builtin_amdgcn_perm(0u, lo, 0x0C030C01u);
                s16[i][k][2] += __builtin_amdgcn_perm(0u, hi, 0x0C020C00u);
                s16[i][k][3] += __builtin_amdgcn_perm(0u, hi, 0x0C0C0C01u);
                lo8[i][k] = all8[i][k] = 0;
            }
    };

    const uint32_t* chk = chunks + dblk * maxc;
    const uint32_t coff = __builtin_amdgcn_readfirstlane((dblk * WAVES + wave) * nact * DPW);
    const uint32_t wrow = __builtin_amdgcn_readfirstlane(dblk * nact);

    // The channels of a chunk are split into WAVES contiguous ranges; wave w
    // stages and transposes range w.
    auto wave_range = [&](uint32_t cc, uint32_t& jb, uint32_t& je) {
        const uint32_t q = (cc + WAVES - 1) / WAVES;
        jb = min(cc, wave * q);
        je = min(cc, jb + q);
    };

    // 1. global -> LDS: the wave's channels, 64 words per load (the last one
    //    partial), and the wave's DMs' codes of the chunk
    auto stage = [&](uint32_t e) {
        const uint32_t j0 = chunk_j0(e), cc = chunk_cc(e), P = chunk_p(e);
        const uint32_t n1 = 6 * T / 8 + P / 8;  // window words
        const uint32_t np = (n1 + PIECE - 1) / PIECE;
        uint32_t jb, je;
        wave_range(cc, jb, je);
        // channel numbers and window starts of up to 8 channels in one go
        // (both tables have 8 spare entries)
        uint32_t chs[8], wss[8];
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            chs[t] = (uint32_t)chan[j0 + jb + t];
            wss[t] = wstart[wrow + j0 + jb + t];
        }
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            if (jb + t >= je) break;
            const gword* src = (const gword*)rows + (size_t)chs[t] * row_words + tw0 + wss[t];
            lword* dst = (lword*)(lds1 + (jb + t) * np * PSTRIDE);
            for (uint32_t m = 0; m + 1 < np; ++m)
                __builtin_amdgcn_global_load_lds(src + m * PIECE + lane, dst + m * PSTRIDE, 4, 0, 0);
            if ((uint32_t)lane < n1 - (np - 1) * PIECE)
                __builtin_amdgcn_global_load_lds(src + (np - 1) * PIECE + lane, dst + (np - 1) * PSTRIDE, 4, 0, 0);
        }
        const gword* csrc = (const gword*)code + coff + j0 * DPW;
        lword* cdst = (lword*)(lds1 + L1C + wave * CW);
        for (uint32_t q = 0; q < cc * DPW; q += 64)
            if ((uint32_t)lane < cc * DPW - q) __builtin_amdgcn_global_load_lds(csrc + q + lane, cdst + q, 4, 0, 0);
    };

    // 2. time-ordered windows -> run words: the wave's channels, four
    //    transposing reads in flight
    auto transpose = [&](uint32_t e) {
        const uint32_t cc = chunk_cc(e), P = chunk_p(e);
        const uint32_t nu = P / 128;
        const uint32_t s1b = ((6 * T / 8 + P / 8 + PIECE - 1) / PIECE) * PSTRIDE * 4;
        for (uint32_t q = lane; q < cc * DPW; q += 64) ldsc[wave * CW + q] = lds1[L1C + wave * CW + q];
        uint32_t jb, je;
        wave_range(cc, jb, je);
        for (uint32_t jj = jb; jj < je; ++jj) {
            for (uint32_t a = 0; 4 * a < nu; ++a) {
                const uint32_t base = jj * s1b + a * (PSTRIDE * 4);
                v2i r[4];
#pragma unroll
                for (int b = 0; b < 4; ++b)
                    r[b] = __builtin_amdgcn_ds_read_tr4_b64_v2i32(
                        (lv2i*)((__attribute__((address_space(3))) char*)((lword*)lds1) + (base & rmask) + rk[b]));
#pragma unroll
                for (int b = 0; b < 4; ++b)
                    if (4 * a + b < nu) {
                        uint32_t* d = lds2 + jj * P + 128 * (4 * a + b) + lane;
                        d[0] = (uint32_t)r[b].x;
                        d[64] = (uint32_t)r[b].y;
                    }
            }
        }
    };

    uint32_t cnt = 0;  // channels in the byte lanes since the last flush (wave-uniform)

    // 3. sums of one chunk from lds2. A channel pair is summed in two halves
    //    of DPW / 2 DMs, software-pipelined: the words of one half (and the
    //    codes of the next pair) are read while the other half is summed.
    constexpr int HD = DPW / 2;  // DMs per half
    auto read_half = [&](const uint32_t (&c)[2 * DPW], int h, uint32_t (&x)[2 * HD][G]) {
#pragma unroll
        for (int t = 0; t < 2 * HD; ++t) {
            // t < HD: channel jj, DM h*HD + t; else channel jj+1, DM h*HD + t - HD
            const uint32_t ct = (t < HD) ? c[h * HD + t] : c[DPW + h * HD + t - HD];
            const uint32_t* p = (const uint32_t*)((const char*)lds2 + ct) + lane;
#pragma unroll
            for (int k = 0; k < G; ++k) x[t][k] = p[64 * k];
        }
    };
    auto sum_half = [&](int h, const uint32_t (&x)[2 * HD][G]) {
#pragma unroll
        for (int t = 0; t < HD; ++t)
#pragma unroll
            for (int k = 0; k < G; ++k) {
                const uint32_t a = x[t][k], b = x[HD + t][k];
                lo8[h * HD + t][k] += (a & M4) + (b & M4);
                all8[h * HD + t][k] += a + b;
            }
    };
    auto compute = [&](uint32_t e) {
        const uint32_t cc = chunk_cc(e);
        const uint32_t cw0 = wave * CW;
        uint32_t c[2 * DPW];
#pragma unroll
        for (int t = 0; t < 2 * DPW; ++t) c[t] = ldsc[cw0 + t];
        uint32_t jj = 0;
        if (cc >= 2) {
            uint32_t xa[2 * HD][G], xb[2 * HD][G];
            read_half(c, 0, xa);
            for (; jj + 1 < cc; jj += 2) {
                if (cnt + 2 > FLUSH) {
                    flush();
                    cnt = 0;
                }
                cnt += 2;
                read_half(c, 1, xb);
                // codes of the next pair (past the chunk's last pair they are
                // stale words of the code buffer, read but not used). The index
                // is hidden from loop strength reduction: a running pointer
                // would lose which LDS array it reads, and the compiler would
                // then wait for the next chunk's loads into lds1 here.
                uint32_t ci4 = cw0 + (jj + 2) * DPW;
                asm volatile("" : "+s"(ci4));
#pragma unroll
                for (int t = 0; t < 2 * DPW; ++t) c[t] = ldsc[ci4 + t];
                sum_half(0, xa);
                read_half(c, 0, xa);
                sum_half(1, xb);
            }
        }
        if (jj < cc) {
            if (cnt + 1 > FLUSH) {
                flush();
                cnt = 0;
            }
            cnt += 1;
#pragma unroll
            for (int i = 0; i < DPW; ++i) {
                const uint32_t* p = (const uint32_t*)((const char*)lds2 + c[i]) + lane;
#pragma unroll
                for (int k = 0; k < G; ++k) {
                    const uint32_t a = p[64 * k];
                    lo8[i][k] += a & M4;
                    all8[i][k] += a;
                }
            }
        }
    };

    uint32_t e = chk[0];
    stage(e);
    for (uint32_t ci = 0;; ++ci) {
        __syncthreads();  // this chunk's windows are in lds1; lds2 is free
        transpose(e);
        __syncthreads();  // run words in lds2; lds1 is free
        const uint32_t en = chk[ci + 1];
        if (chunk_cc(en)) stage(en);
        compute(e);
        if (!chunk_cc(en)) break;
        e = en;
    }
    flush();

    // Epilogue: lane l owns positions l + 64k; run r of position m is sample
    // t0 + r*T + m.
    const uint32_t dmb = dblk * DPB + wave * DPW;
    const uint32_t t0 = tile * TILE;
#pragma unroll
    for (int i = 0; i < DPW; ++i) {
        const uint32_t dm = dmb + i;
        if (dm >= ndm) continue;
        uint8_t* orow = out + (size_t)dm * out_stride;
#pragma unroll
        for (int k = 0; k < G; ++k) {
            const uint32_t m = (uint32_t)lane + 64u * k;
#pragma unroll
            for (int r = 0; r < RUNS; ++r) {
                // r = 0,4: s16[0]; 2,6: s16[1]; 1,5: s16[2]; 3: s16[3]
                const int q = (r & 1) ? ((r == 3) ? 3 : 2) : ((r & 2) ? 1 : 0);
                const bool high = (r == 4 || r == 6 || r == 5);
                const uint32_t v = s16[i][k][q];
                const uint32_t sum = high ? (v >> 16) : (v & 0xFFFFu);
                const uint32_t t = t0 + (uint32_t)r * T + m;
                if (t < nsamps_out) tdd_store(orow, t, tdd_scale((float)sum, ep), ep);
            }
        }
    }
}

}  // namespace

extern "C" hipError_t dd_tdd_dedisperse_runs(const uint32_t* rows, size_t row_words, const uint32_t* code,
                                             const int32_t* chan, const uint32_t* wstart, const uint32_t* chunks,
                                             unsigned maxc, size_t nact, size_t ndm, size_t nsamps_out,
                                             uint8_t* out, size_t out_stride, dd_tdd_epilogue ep,
                                             unsigned dms_per_wave, unsigned waves, const dd_tdd_beam* beams,
                                             const uint32_t* tile_beam, size_t batch_tiles, hipStream_t stream)
{
    if (ndm == 0) return hipSuccess;
    if (nact * 15 > 65535 || nact > 65535) return hipErrorInvalidValue;
    const size_t ntiles = beams ? batch_tiles : (nsamps_out + TILE - 1) / TILE;
    if (ntiles == 0) return hipSuccess;
    const size_t dpb = (size_t)dms_per_wave * waves;
    if (dpb == 0) return hipErrorInvalidValue;
    const size_t ndmblk = (ndm + dpb - 1) / dpb;
    const size_t nblocks = ntiles * ndmblk;
    if (nblocks >= (size_t(1) << 31)) return hipErrorInvalidValue;
#define DD_RUNS(DPW, WAVES, MINW)                                                                              \
    if (dms_per_wave == DPW && waves == WAVES) {                                                               \
        hipLaunchKernelGGL((tdd_runs_k<DPW, WAVES, MINW>), dim3((unsigned)nblocks), dim3(WAVES * 64), 0, stream, \
                           rows, row_words, code, chan, wstart, chunks, (uint32_t)maxc, (uint32_t)nact,          \
                           (uint32_t)ndm, (uint32_t)nsamps_out, (uint32_t)ndmblk, out, out_stride, ep, beams,    \
                           tile_beam);                                                                         \
        return hipGetLastError();                                                                              \
    }
    DD_RUNS(4, 4, 3)
    DD_RUNS(2, 8, 6)
#undef DD_RUNS
    return hipErrorInvalidValue;
}
