// gfx950 FFT kernels of the Fourier-domain dedispersion (FDD) path.
//
// One workgroup computes one complex transform of M = A*B*C points entirely
// on chip: three passes of small in-register DFTs (sizes C, B, A), with the
// inter-pass exchanges staged through LDS (in P rounds when the data exceeds
// the LDS budget). Input and output are in natural order and every global
// access is unit-stride across the workgroup, so a transform costs one HBM
// read and one HBM write.
//
//   n = n0 + A*n1 + A*B*n2,   k = k2 + C*k1 + C*B*k0
//   pass 1: DFT_C over n2 for each (n0,n1), twiddle W_M^{(n0+A n1) k2}
//   pass 2: DFT_B over n1 for each (n0,k2), twiddle W_{AB}^{n0 k1}
//   pass 3: DFT_A over n0 for each (k1,k2)
//
// Real transforms of length nfft = 2M use the half-length packing
// z[m] = x[2m] + i x[2m+1]; the forward kernel pairs its packed result Z into
// X[k] = (Z[k] + conj Z[M-k])/2 - i/2 W^k (Z[k] - conj Z[M-k]),
// W = e^{-2 pi i/nfft}, and writes bins 0..M. The inverse kernel forms its
// packed input from the M+1 bins of each DM row at load time.
//
// Twiddles come from one table ph[m] = e^{+2 pi i m / nfft}, m < nfft, built in
// double precision on the host: W_M^j = conj(ph[2j]).
#include <hip/hip_runtime.h>

#include <cstdlib>

#include <algorithm>

#include "fdd_kernels.h"

namespace {

// Fourth word of a raw (unswizzled, stride 0) buffer resource on gfx9-family
// GPUs: 32-bit data format. Accesses at byte offsets at or past the record
// count are dropped (stores) or read as zero (loads).
constexpr int kRawBufferWord3 = 0x00020000;

__device__ __forceinline__ float2 cmul(float2 a, float2 b)
{
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 conjf2(float2 a) { return make_float2(a.x, -a.y); }

// In-register DFT of size R (power of two), natural order in and out,
// Stockham radix-2; DIR = -1 forward, +1 inverse. The twiddles are compile
// time constants after unrolling.
template <int R, int DIR>
__device__ __forceinline__ void dft_reg(float2 (&v)[R])
{
    float2 t[R];
#pragma unroll
    for (int ns = 1; ns < R; ns *= 2) {
#pragma unroll
        for (int j = 0; j < R / 2; ++j) {
            const int jm = j % ns;
            const double ang = DIR * M_PI * (double)jm / (double)ns;
            const float2 w = make_float2((float)__builtin_cos(ang), (float)__builtin_sin(ang));
            float2 a = v[j];
            float2 b = (jm == 0) ? v[j + R / 2] : cmul(v[j + R / 2], w);
            const int d = (j / ns) * ns * 2 + jm;
            t[d] = cadd(a, b);
            t[d + ns] = csub(a, b);
        }
#pragma unroll
        for (int j = 0; j < R; ++j) v[j] = t[j];
    }
}

template <int A, int B, int C>
struct FftShape {
    static constexpr int M = A * B * C;
    static constexpr int T1 = A * B;  // pass-1 threads (each C values)
    static constexpr int T2 = A * C;  // pass-2 threads (each B values)
    static constexpr int T3 = B * C;  // pass-3 threads (each A values)
    static constexpr int T = (T1 > T2 ? (T1 > T3 ? T1 : T3) : (T2 > T3 ? T2 : T3));
    // LDS floats of one exchange round (one component of every value).
    // Exchange 1 -> 2: rows of B values padded by one. Exchange 2 -> 3:
    // groups of B*A values (one k2) padded by G2, chosen (by a bank
    // simulation over 32 and 64 banks) so that both the pass-2 stores (lanes
    // along n0) and the pass-3 loads (lanes along k2) are at most 2-way
    // conflicted; unpadded rows of A+1 put every pass-3 lane of a half-wave
    // in one bank.
    static constexpr int X1 = A * C * (B + 1);
    static constexpr int G2 = (A == 32) ? 1 : 9;
    static constexpr int X2 = C * (B * A + G2);
    static constexpr int LDS_FLOATS = X1 > X2 ? X1 : X2;
};

// Bin-block layout of the real spectra (the float16 / bfloat16 matrix-core
// contraction with mirror bins reads and writes it): bins k and k+1 (k even,
// k <= M/2) and their mirrors M-k, M-k-1 form block k/2, the four values of a
// block for every chunk contiguous, so a workgroup of the contraction reads
// and writes whole cache lines. Element (row, k, chunk) of a launch of nchunk
// chunks sits at ((row*NBB + block)*nchunk + chunk)*4 + slot, NBB = M/4 + 1:
// slot k&1 of block k/2 for k <= M/2, else slot 2 + (j&1) of block j/2 with
// j = M - k.
__device__ __forceinline__ uint32_t blk_offset(uint32_t k, uint32_t M, uint32_t bstride)
{
    if (k <= M / 2) return (k >> 1) * bstride + (k & 1);
    const uint32_t j = M - k;
    return (j >> 1) * bstride + 2 + (j & 1);
}

// Workgroup order across the 8 XCDs: consecutive ids (bid) go to different
// XCDs; L gives consecutive values to the workgroups of one XCD, so those
// that share cache lines (neighbouring chunks in the bin-block layout) run
// together in that XCD's L2.
__device__ __forceinline__ uint32_t xcd_linear(uint32_t bid, uint32_t nb)
{
    const uint32_t xcd = bid & 7u, per = nb >> 3, rem = nb & 7u;
    return xcd * per + min(xcd, rem) + (bid >> 3);
}

// Twiddle w^j for j = 0..R-1 given w = base: reseeded from the table every 8
// steps (tw_at(j) returns the exact table value for j multiple of 8), so each
// value is at most 7 complex products away from a correctly rounded one.
template <int DIR>
__device__ __forceinline__ float2 table_tw(const float2* __restrict__ ph, uint32_t idx, uint32_t mask)
{
    float2 tw = ph[idx & mask];
    if (DIR < 0) tw = conjf2(tw);
    return tw;
}

// The transform engine. load(n) gives input point n (pass-1 threads call it
// for n = t + A*B*n2); prep(v, active) then sees the C loaded points of every
// thread before pass 1 (all threads call it); on return q[k0] of thread t < T3 holds output point
// k = t + T3*k0 (T3 = B*C, the block size for every supported shape). ph is the
// twiddle table of e^{+2 pi i m / (2M)}. Each exchange moves the real parts
// of all values through LDS, then the imaginary parts, so at most half of a
// thread's values are in flight (and the LDS holds one component of the
// whole transform).
struct NoPrep {
    template <class V>
    __device__ void operator()(V&, bool) const {}
};

template <int A, int B, int C, int DIR, class Load, class Prep = NoPrep>
__device__ __forceinline__ void fft3_core(const float2* __restrict__ ph, float* lds, Load load, float2 (&q)[A],
                                          Prep prep = Prep(), int t1 = -1, int t3 = -1)
{
    using S = FftShape<A, B, C>;
    constexpr int M = S::M;
    constexpr uint32_t MASK = 2 * M - 1;
    const int t = threadIdx.x;
    // pass-1 index of this thread (a permutation of 0..T1-1 over the threads
    // t < T1 may be given; pass 2 uses t itself) and pass-3 index (likewise a
    // permutation of 0..T3-1 over t < T3: thread t then holds the output
    // points k = t3 + T3*k0)
    const int tp = (t1 < 0) ? t : t1;
    const int tq = (t3 < 0) ? t : t3;

    // pass 1: DFT over n2, then twiddle W_M^{tp k2} (recurrence reseeded every 8)
    float2 v[C];
    const bool act1 = t < S::T1;
    if (act1) {
#pragma unroll
        for (int n2 = 0; n2 < C; ++n2) v[n2] = load(tp + S::T1 * n2);
    }
    prep(v, act1);  // every thread (it may use LDS and barriers)
    if (act1) {
        dft_reg<C, DIR>(v);
        const float2 w1 = table_tw<DIR>(ph, 2u * (uint32_t)tp, MASK);
        float2 tw = make_float2(1.f, 0.f);
#pragma unroll
        for (int k2 = 1; k2 < C; ++k2) {
            if (k2 % 8 == 0) tw = table_tw<DIR>(ph, 2u * (uint32_t)tp * (uint32_t)k2, MASK);
            else tw = cmul(tw, w1);
            v[k2] = cmul(v[k2], tw);
        }
    }
    const int n0_1 = tp % A, n1_1 = tp / A;

    // exchange 1 -> 2: element (n0, n1, k2) at (k2*A + n0)*(B+1) + n1
    float2 u[B];
    const int n0_2 = t % A, k2_2 = t / A;  // pass-2 thread (n0, k2)
    const bool act2 = t < S::T2;
#pragma unroll
    for (int comp = 0; comp < 2; ++comp) {
        __syncthreads();
        if (act1) {
#pragma unroll
            for (int k2 = 0; k2 < C; ++k2)
                lds[(k2 * A + n0_1) * (B + 1) + n1_1] = comp ? v[k2].y : v[k2].x;
        }
        __syncthreads();
        if (act2) {
#pragma unroll
            for (int n1 = 0; n1 < B; ++n1) {
                const float val = lds[(k2_2 * A + n0_2) * (B + 1) + n1];
                if (comp) u[n1].y = val;
                else u[n1].x = val;
            }
        }
    }
    // pass 2: DFT over n1, then twiddle W_M^{C n0 k1}
    if (act2) {
        dft_reg<B, DIR>(u);
        const float2 w2 = table_tw<DIR>(ph, 2u * (uint32_t)(C * n0_2), MASK);
        float2 tw = make_float2(1.f, 0.f);
#pragma unroll
        for (int k1 = 1; k1 < B; ++k1) {
            if (k1 % 8 == 0) tw = table_tw<DIR>(ph, 2u * (uint32_t)(C * n0_2) * (uint32_t)k1, MASK);
            else tw = cmul(tw, w2);
            u[k1] = cmul(u[k1], tw);
        }
    }

    // exchange 2 -> 3: element (n0, k1, k2) at k2*(B*A + G2) + k1*A + n0
    const int k2_3 = tq % C, k1_3 = tq / C;
    const bool act3 = t < S::T3;
#pragma unroll
    for (int comp = 0; comp < 2; ++comp) {
        __syncthreads();
        if (act2) {
#pragma unroll
            for (int k1 = 0; k1 < B; ++k1)
                lds[k2_2 * (B * A + S::G2) + k1 * A + n0_2] = comp ? u[k1].y : u[k1].x;
        }
        __syncthreads();
        if (act3) {
#pragma unroll
            for (int n0 = 0; n0 < A; ++n0) {
                const float val = lds[k2_3 * (B * A + S::G2) + k1_3 * A + n0];
                if (comp) q[n0].y = val;
                else q[n0].x = val;
            }
        }
    }
    // pass 3: DFT over n0; output k = k2 + C k1 + C B k0 = tq + T3 k0
    if (act3) dft_reg<A, DIR>(q);
}

// The engine with per-point output: store(k, value) receives output point k.
template <int A, int B, int C, int DIR, class Load, class Store>
__device__ __forceinline__ void fft3(const float2* __restrict__ ph, float* lds, Load load, Store store)
{
    float2 q[A];
    fft3_core<A, B, C, DIR>(ph, lds, load, q);
    const int t = threadIdx.x;
    if (t < FftShape<A, B, C>::T3) {
#pragma unroll
        for (int k0 = 0; k0 < A; ++k0) store(t + FftShape<A, B, C>::T3 * k0, q[k0]);
    }
}

// Forward: channel c, chunk i of the channel-major 8-bit rows; writes the
// real spectrum X[0..M] of the chunk (M+1 bins, row pitch Xp). The packed
// transform Z of z[m] = x[2m] + i x[2m+1] is paired in the epilogue,
//   X[k] = (Z[k] + conj Z[M-k])/2 - i/2 W^k (Z[k] - conj Z[M-k]),
// with Z[M-k] taken from the lane that holds it: pass 3 runs index tq on lane
// i of a wave and T3 - tq on lane i + 32 (the inverse's mapping), so point
// k = tq + T3*e has its mirror M-k at element A-1-e of the partner lane and
// one v_permlane32_swap of that register serves both lanes (tq = T3/2: its
// own element A-1-e; tq = 0: its own element (A-e) mod A). Measured as fast as
// the earlier exchange of the mirrors through LDS (two rounds per component,
// 3.73 vs 3.80 ms for 1024 ch x 23 chunks of 65536) with 92 instead of 127
// registers.
// TAB: the chunk's rows, first sample and valid length come from tab[chunk]
// (the chunks of several beams in one launch) instead of rows/chunk_step/nvalid.
// BLK: X in the bin-block layout (one-dimensional grid in XCD order, chunk
// fastest) instead of rows of Xp.
template <int A, int B, int C, bool TAB = false, bool BLK = false, int NT = FftShape<A, B, C>::T>
__global__ __launch_bounds__(NT) void fdd_forward_k(
    const uint8_t* __restrict__ rows, size_t row_bytes, size_t chunk_step, size_t nvalid,
    size_t period, const float* __restrict__ chan_scale, const float2* __restrict__ ph,
    float2* __restrict__ X, size_t nchunk, size_t Xp, uint32_t nbits,
    const dd_fdd_chunk* __restrict__ tab)
{
    using S = FftShape<A, B, C>;
    static_assert(S::T3 == NT, "pass 3 uses every thread");
    constexpr int M = S::M, T3 = S::T3;
    __shared__ float lds[S::LDS_FLOATS];
    size_t chunk = blockIdx.x, c = blockIdx.y;
    if constexpr (BLK) {
        const uint32_t l = xcd_linear(blockIdx.x, gridDim.x);
        chunk = l % nchunk;
        c = l / nchunk;
    }
    size_t t0 = chunk * chunk_step;
    if constexpr (TAB) {
        const dd_fdd_chunk e = tab[chunk];
        rows = e.rows;
        row_bytes = e.row_bytes;
        nvalid = e.nvalid;
        t0 = e.t0;
    }
    const uint8_t* src = rows + c * row_bytes;
    const float s = chan_scale[c];
    float2* dst = X + (c * nchunk + chunk) * Xp;
    // bin k of this transform
    constexpr uint32_t NBB = M / 4 + 1;
    float2* const dstb = X + (c * NBB * nchunk + chunk) * 4;
    const uint32_t bstride = (uint32_t)nchunk * 4u;
    auto bin = [&](uint32_t k) -> float2* { return BLK ? dstb + blk_offset(k, M, bstride) : dst + k; };
    // x[t] = (u8 - 127.5) * scale (src/fdd/unpack/unpack_kernel.cuh:60) for
    // t < nvalid, zero beyond (the zero padding of src/fdd/FDDGPUPlan.cpp:476-483);
    // with period > 0 the series is read periodically (t mod period). Sub-byte
    // samples (an extension: the reference FDD reads 8 bits only) come packed
    // LSB-first in 32-bit words and are centred on (2^nbits - 1)/2 likewise.
    const float mid = 0.5f * (float)((1u << nbits) - 1u);
    const uint32_t lspw = 5 - __builtin_ctz(nbits);  // log2 samples per word
    const uint32_t mask = (1u << nbits) - 1u;
    const uint32_t* wsrc = reinterpret_cast<const uint32_t*>(src);
    auto value = [&](size_t t) -> uint32_t {
        if (nbits == 8) return src[t];
        return (wsrc[t >> lspw] >> ((uint32_t)(t & ((1u << lspw) - 1u)) * nbits)) & mask;
    };
    auto sample = [&](size_t t) -> float {
        if (period) {
            // t < 2 period for every chunk of a series at least one segment
            // long: a subtraction, the 64-bit remainder only beyond that
            if (t >= period) t -= period;
            if (t >= period) t %= period;
        }
        return (t < nvalid) ? ((float)value(t) - mid) * s : 0.f;
    };
    auto load = [&](int n) -> float2 {
        const size_t t = t0 + 2 * (size_t)n;
        if (!period && t + 1 < nvalid) {
            if (nbits == 8) {
                const uint16_t pair = *reinterpret_cast<const uint16_t*>(src + t);
                return make_float2(((float)(pair & 0xff) - mid) * s, ((float)(pair >> 8) - mid) * s);
            }
            // t even: both samples sit in one word
            const uint32_t w = wsrc[t >> lspw] >> ((uint32_t)(t & ((1u << lspw) - 1u)) * nbits);
            return make_float2(((float)(w & mask) - mid) * s, ((float)((w >> nbits) & mask) - mid) * s);
        }
        return make_float2(sample(t), sample(t + 1));
    };
    static_assert(T3 % 64 == 0, "mirror indices pair lanes i and i + 32 of whole waves");
    // pass-3 index: index tq on lane i of a wave, its mirror T3 - tq on lane
    // i + 32 (T3/2 is its own mirror, 0 pairs within its own elements)
    const int t = threadIdx.x;
    const int hi = (t >> 5) & 1;
    const int base = (t >> 6) * 32 + (t & 31);
    const int tq = hi ? (base ? T3 - base : T3 / 2) : base;
    const bool self0 = (tq == 0), selfh = (tq == T3 / 2);
    float2 q[A];
    if (nbits == 8 && t0 + 2 * (size_t)M <= nvalid && (!period || t0 + 2 * (size_t)M <= period)) {
        // interior 8-bit chunk: every pair in bounds (and, for a periodic
        // series, before the wrap, where t mod period = t), no branch between
        // the loads, so all of a thread's loads are in flight together
        auto interior = [&](int n) -> float2 {
            const uint16_t pair = *reinterpret_cast<const uint16_t*>(src + t0 + 2 * (size_t)n);
            return make_float2(((float)(pair & 0xff) - mid) * s, ((float)(pair >> 8) - mid) * s);
        };
        fft3_core<A, B, C, -1>(ph, lds, interior, q, NoPrep(), -1, tq);
    } else {
        fft3_core<A, B, C, -1>(ph, lds, load, q, NoPrep(), -1, tq);
    }
    // the value of the lane i ^ 32
    auto other_half = [&](float x) -> float {
        const auto r = __builtin_amdgcn_permlane32_swap(__builtin_bit_cast(uint32_t, x),
                                                        __builtin_bit_cast(uint32_t, x), false, false);
        return __builtin_bit_cast(float, hi ? r[0] : r[1]);
    };
    // point k = tq + T3*e has its mirror M - k at element A-1-e of the partner
    // lane (one swap of that register serves both lanes)
    // BLK: the upper lane computes its element A-1-e in step e instead of e,
    // so that lanes i, i+1, i+32, i+33 store the four values of a bin block
    // (k, k+1 and their mirrors) in one instruction; each lane then sends the
    // partner its own operand.
#pragma unroll
    for (int e = 0; e < A; ++e) {
        const bool rev = BLK && hi;
        const int ee = rev ? A - 1 - e : e;
        float2 a, mz;
        if constexpr (BLK) {
            a = make_float2(rev ? q[A - 1 - e].x : q[e].x, rev ? q[A - 1 - e].y : q[e].y);
            mz = make_float2(other_half(a.x), other_half(a.y));
            if (selfh) mz = q[e];
        } else {
            const float2 own = q[A - 1 - e];
            mz = make_float2(other_half(own.x), other_half(own.y));
            if (selfh) mz = own;
            a = q[e];
        }
        if (self0) mz = q[(A - e) % A];
        const int k = tq + T3 * ee;
        const float2 pk = ph[k];  // W^k = conj(pk)
        const float ux = a.x + mz.x, uy = a.y - mz.y;  // Z[k] + conj Z[M-k]
        const float vx = a.x - mz.x, vy = a.y + mz.y;  // Z[k] - conj Z[M-k]
        // -i/2 W^k v = 1/2 (Im(W^k v), -Re(W^k v))
        const float wr = vx * pk.x + vy * pk.y, wi = vy * pk.x - vx * pk.y;
        *bin(k) = make_float2(0.5f * (ux + wi), 0.5f * (uy - wr));
    }
    if (self0) *bin(M) = make_float2(q[0].x - q[0].y, 0.f);  // X[M] = Re Z[0] - Im Z[0]
}

// Inverse: DM row j, chunk i of Y (M+1 bins); writes the first `good`
// samples of the chunk's real series, times 1/nfft, into out. Pass 1 takes
// the packed spectrum of the even/odd samples,
//   Zu[k] = (Y[k] + conj Y[M-k]) + i e^{+2 pi i k/nfft} (Y[k] - conj Y[M-k]).
// Every bin is read from HBM once: pass-1 index tp (points k = tp + T1 n2)
// runs on lane i of a wave and T1 - tp on lane i + 32 of the same wave, so
// the mirror Y[M-k] of element n2, element C-1-n2 of index T1 - tp, comes
// from the other half of the wave (v_permlane32_swap). Index T1/2 is its own
// mirror partner (element C-1-n2 of itself); index 0 pairs element n2 with
// its own element C-n2 (k = 0 with Y[M]), read back from a small LDS area.
// TAB: the chunk's output row (row row0 + blockIdx.y of tab[chunk].out), first
// output sample and output length come from tab[chunk].
// PACKED: Y already holds the packed spectrum Zu[0..M-1] (the contraction's
// mirror form writes it), so pass 1 loads it as it is.
// BLK (with PACKED): Zu in the bin-block layout (one-dimensional grid in XCD
// order, chunk fastest).
template <int A, int B, int C, bool TAB = false, bool PACKED = false, bool BLK = false,
          int NT = FftShape<A, B, C>::T>
__global__ __launch_bounds__(NT) void fdd_inverse_k(
    const float2* __restrict__ Y, size_t nchunk, size_t Yp, const float2* __restrict__ ph,
    float* __restrict__ out, size_t out_stride, size_t good, size_t nsamps_out,
    const dd_fdd_chunk* __restrict__ tab, size_t row0)
{
    using S = FftShape<A, B, C>;
    constexpr int M = S::M, T1 = S::T1;
    static_assert(T1 % 64 == 0 && C % 2 == 0, "pass-1 threads fill whole waves");
    static_assert(S::LDS_FLOATS >= 2 * C, "index 0 stages its points");
    static_assert(!BLK || PACKED, "the bin-block layout holds the packed spectrum");
    __shared__ float lds[S::LDS_FLOATS];
    size_t chunk = blockIdx.x, row = blockIdx.y;
    if constexpr (BLK) {
        const uint32_t l = xcd_linear(blockIdx.x, gridDim.x);
        chunk = l % nchunk;
        row = l / nchunk;
    }
    const float2* src = Y + (row * nchunk + chunk) * Yp;
    size_t t0 = chunk * good;
    if constexpr (TAB) {
        const dd_fdd_chunk e = tab[chunk];
        out = e.out;
        out_stride = e.out_stride;
        nsamps_out = e.nsamps_out;
        t0 = e.t0;
    }
    float* dst = out + (row0 + row) * out_stride;
    const float inv_n = 1.0f / (float)(2 * M);
    const int t = threadIdx.x;
    float2 q[A];
    if constexpr (PACKED && BLK) {
        constexpr uint32_t NBB = M / 4 + 1;
        const float2* srcb = Y + (row * NBB * nchunk + chunk) * 4;
        const uint32_t bstride = (uint32_t)nchunk * 4u;
        // pass-1 index tp on lane i, its mirror T1 - tp on lane i + 32; the
        // upper lane loads its element C-1-n2 in step n2 (the mirrors of the
        // lower lane's points), so that one load instruction reads the four
        // values of a bin block, and reverses its elements before pass 1
        const int hi = (t >> 5) & 1;
        const int base = (t >> 6) * 32 + (t & 31);
        const int tp = (t < T1) ? (hi ? (base ? T1 - base : T1 / 2) : base) : t;
        auto zload = [&](int n) -> float2 {
            const int m = hi ? 2 * tp + T1 * (C - 1) - n : n;
            return srcb[blk_offset((uint32_t)m, M, bstride)];
        };
        auto reverse = [&](float2 (&v)[C], bool act) {
            if (!act) return;
#pragma unroll
            for (int n2 = 0; n2 < C / 2; ++n2) {
                const float2 a = v[n2], b = v[C - 1 - n2];
                v[n2] = make_float2(hi ? b.x : a.x, hi ? b.y : a.y);
                v[C - 1 - n2] = make_float2(hi ? a.x : b.x, hi ? a.y : b.y);
            }
        };
        fft3_core<A, B, C, +1>(ph, lds, zload, q, reverse, tp);
    } else if constexpr (PACKED) {
        // an opaque 32-bit byte offset from the uniform row base (one offset
        // register per load; k < M, so 8k fits)
        const char* const srcb = reinterpret_cast<const char*>(src);
        auto zload = [&](int k) -> float2 {
            uint32_t off = (uint32_t)k * 8u;
            asm volatile("" : "+v"(off));
            return *reinterpret_cast<const float2*>(srcb + off);
        };
        fft3_core<A, B, C, +1>(ph, lds, zload, q);
    } else {
    const float ymx = src[M].x;  // Y[M]: its imaginary part is not used
    const int hi = (t >> 5) & 1;
    const int base = (t >> 6) * 32 + (t & 31);
    const int tp = (t < T1) ? (hi ? (base ? T1 - base : T1 / 2) : base) : t;
    const bool self0 = (tp == 0), selfh = (tp == T1 / 2);
    // the value of the lane i ^ 32
    auto other_half = [&](float x) -> float {
        const auto r = __builtin_amdgcn_permlane32_swap(__builtin_bit_cast(uint32_t, x),
                                                        __builtin_bit_cast(uint32_t, x), false, false);
        return __builtin_bit_cast(float, hi ? r[0] : r[1]);
    };
    // e^{+2 pi i k/nfft} for k = tp + T1 n2 is e^{+2 pi i tp/nfft} e^{i pi n2/C}:
    // one table read per thread, the second factor a constant per element
    const float2 ptp = ph[tp];
    auto wconst = [](int n2) -> float2 {
        const double ang = M_PI * (double)n2 / (double)C;
        return make_float2((float)__builtin_cos(ang), (float)__builtin_sin(ang));
    };
    // Zu from a = Y[k], ym = Y[M-k] and the twiddle e^{+2 pi i k/nfft}
    auto zu = [&](float2 a, float2 ym, float2 tw) -> float2 {
        const float2 b = conjf2(ym);
        const float2 e = cadd(a, b), d = csub(a, b);
        const float2 o = cmul(d, tw);
        return make_float2(e.x - o.y, e.y + o.x);
    };
    // a 32-bit byte offset from the workgroup-uniform row base (the
    // scalar-base load form: one offset register per load instead of a
    // 64-bit address; k <= M, so 8k fits). The offset is opaque to the
    // compiler, which would otherwise fold the per-thread part into a 64-bit
    // address and add each element's constant to it (two registers a load)
    const char* const srcb = reinterpret_cast<const char*>(src);
    auto load = [&](int k) -> float2 {
        uint32_t off = (uint32_t)k * 8u;
        asm volatile("" : "+v"(off));
        return *reinterpret_cast<const float2*>(srcb + off);
    };
    auto pair = [&](float2 (&v)[C], bool act) {
        if (!act) return;
        // Mirrors come from the partner lane i ^ 32 (element m2 for n2,
        // element n2 for m2) by v_permlane32_swap, except in wave 0, which
        // holds the two self-paired indices: 0 (lane 0: element n2 pairs with
        // its own element C - n2, n2 = 0 with Y[M]) and T1/2 (lane 32: its own
        // element C-1-n2). Wave 0 stages its values in LDS ([element][lane];
        // free before pass 1) and takes every mirror from there. The loop has
        // no branch: all waves read both sources and select with the
        // wave-uniform w0 (a branch per source kept both paths' values live)
        const bool w0 = __builtin_amdgcn_readfirstlane(t >> 6) == 0;
        const int lane = t & 63;
        static_assert(S::LDS_FLOATS >= 2 * 64 * C, "wave 0 stages its values");
        if (w0) {
#pragma unroll
            for (int e = 0; e < C; ++e) lds[e * 64 + lane] = v[e].x, lds[(C + e) * 64 + lane] = v[e].y;
        }
        const int sl = (self0 || selfh) ? lane : (lane ^ 32);
        auto staged = [&](int n2) -> float2 {  // wave 0: the mirror of element n2
            const int se = self0 ? (C - n2) % C : C - 1 - n2;
            return make_float2(lds[se * 64 + sl], lds[(C + se) * 64 + sl]);
        };
#pragma unroll
        for (int n2 = 0; n2 < C / 2; ++n2) {
            const int m2 = C - 1 - n2;
            float2 a1 = v[n2];
            const float2 a2 = v[m2];
            float2 b1 = make_float2(other_half(a2.x), other_half(a2.y));
            float2 b2 = make_float2(other_half(a1.x), other_half(a1.y));
            const float2 s1 = staged(n2), s2 = staged(m2);
            b1 = w0 ? s1 : b1;
            b2 = w0 ? s2 : b2;
            if (n2 == 0 && self0) a1.y = 0.f, b1 = make_float2(ymx, 0.f);
            v[n2] = zu(a1, b1, n2 ? cmul(ptp, wconst(n2)) : ptp);
            v[m2] = zu(a2, b2, cmul(ptp, wconst(m2)));
            __builtin_amdgcn_sched_barrier(0);  // one pair's temporaries at a time
        }
    };
    fft3_core<A, B, C, +1>(ph, lds, load, q, pair, tp);
    }
    if (t < S::T3) {
        // one workgroup-uniform base and limit (t0 < nsamps_out for every
        // chunk processed). The limit is the record count of a raw buffer
        // resource: the hardware drops stores at byte offsets past it, so the
        // stores need no branch (a branch per sample split the pass-3 sums
        // into small blocks and kept their partial values live across them)
        float* const base = dst + t0;
        const uint32_t lim = (uint32_t)min(good, nsamps_out - t0);
        const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(base, 0, (int)(lim * 4u), kRawBufferWord3);
#pragma unroll
        for (int k0 = 0; k0 < A; ++k0) {
            const uint32_t ta = 2u * ((uint32_t)t + (uint32_t)S::T3 * k0);
            __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, q[k0].x * inv_n), rs, (int)(4u * ta), 0, 0);
            __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, q[k0].y * inv_n), rs, (int)(4u * ta + 4u), 0, 0);
        }
    }
}

// Grid of a launch over nchunk x nrows transforms: two-dimensional, or (the
// bin-block layout) one-dimensional in XCD order.
inline bool blk_grid(bool blk, size_t nchunk, size_t nrows, dim3& grid)
{
    if (!blk) {
        grid = dim3((unsigned)nchunk, (unsigned)nrows);
        return nchunk < (size_t(1) << 31) && nrows < 65536;
    }
    grid = dim3((unsigned)(nchunk * nrows));
    return nchunk * nrows < (size_t(1) << 31);
}

template <int A, int B, int C>
hipError_t launch_forward(const uint8_t* rows, size_t row_bytes, size_t chunk_step, size_t nvalid,
                          size_t period, const float* chan_scale, const float2* ph, float2* X,
                          size_t nchan, size_t nchunk, size_t Xp, unsigned nbits, hipStream_t stream,
                          bool blk = false)
{
    if (!blk && Xp < (size_t)FftShape<A, B, C>::M + 1) return hipErrorInvalidValue;
    if (nbits != 1 && nbits != 2 && nbits != 4 && nbits != 8) return hipErrorInvalidValue;
    if (nbits != 8 && (row_bytes % 4)) return hipErrorInvalidValue;
    if (!nchunk || !nchan) return hipSuccess;
    dim3 grid;
    if (!blk_grid(blk, nchunk, nchan, grid)) return hipErrorInvalidValue;
    if (blk)
        hipLaunchKernelGGL((fdd_forward_k<A, B, C, false, true>), grid, dim3(FftShape<A, B, C>::T), 0, stream,
                           rows, row_bytes, chunk_step, nvalid, period, chan_scale, ph, X, nchunk, Xp,
                           (uint32_t)nbits, (const dd_fdd_chunk*)nullptr);
    else
        hipLaunchKernelGGL((fdd_forward_k<A, B, C>), grid, dim3(FftShape<A, B, C>::T), 0, stream, rows,
                           row_bytes, chunk_step, nvalid, period, chan_scale, ph, X, nchunk, Xp, (uint32_t)nbits,
                           (const dd_fdd_chunk*)nullptr);
    return hipGetLastError();
}

template <int A, int B, int C>
hipError_t launch_forward_chunks(const dd_fdd_chunk* chunks, size_t nchunk, const float* chan_scale,
                                 const float2* ph, float2* X, size_t nchan, size_t Xp, unsigned nbits,
                                 hipStream_t stream, bool blk = false)
{
    if ((!blk && Xp < (size_t)FftShape<A, B, C>::M + 1) || !chunks) return hipErrorInvalidValue;
    if (nbits != 1 && nbits != 2 && nbits != 4 && nbits != 8) return hipErrorInvalidValue;
    if (!nchunk || !nchan) return hipSuccess;
    dim3 grid;
    if (!blk_grid(blk, nchunk, nchan, grid)) return hipErrorInvalidValue;
    if (blk)
        hipLaunchKernelGGL((fdd_forward_k<A, B, C, true, true>), grid, dim3(FftShape<A, B, C>::T), 0, stream,
                           (const uint8_t*)nullptr, (size_t)0, (size_t)0, (size_t)0, (size_t)0, chan_scale, ph,
                           X, nchunk, Xp, (uint32_t)nbits, chunks);
    else
        hipLaunchKernelGGL((fdd_forward_k<A, B, C, true>), grid, dim3(FftShape<A, B, C>::T), 0, stream,
                           (const uint8_t*)nullptr, (size_t)0, (size_t)0, (size_t)0, (size_t)0, chan_scale, ph,
                           X, nchunk, Xp, (uint32_t)nbits, chunks);
    return hipGetLastError();
}

// packed: 0 bins 0..M in rows of Yp, 1 the packed spectrum in rows of Yp,
// 2 the packed spectrum in the bin-block layout.
template <int A, int B, int C>
hipError_t launch_inverse(const float2* Y, size_t nrows, size_t nchunk, size_t Yp, const float2* ph,
                          float* out, size_t out_stride, size_t good, size_t nsamps_out,
                          hipStream_t stream, int packed = 0)
{
    if (packed < 0 || packed > 2) return hipErrorInvalidValue;
    if (!nchunk || !nrows) return hipSuccess;
    dim3 grid;
    if (!blk_grid(packed == 2, nchunk, nrows, grid)) return hipErrorInvalidValue;
    const dim3 block(FftShape<A, B, C>::T);
    if (packed == 2)
        hipLaunchKernelGGL((fdd_inverse_k<A, B, C, false, true, true>), grid, block, 0, stream, Y, nchunk, Yp, ph,
                           out, out_stride, good, nsamps_out, (const dd_fdd_chunk*)nullptr, (size_t)0);
    else if (packed)
        hipLaunchKernelGGL((fdd_inverse_k<A, B, C, false, true>), grid, block, 0, stream, Y, nchunk, Yp, ph, out,
                           out_stride, good, nsamps_out, (const dd_fdd_chunk*)nullptr, (size_t)0);
    else
        hipLaunchKernelGGL((fdd_inverse_k<A, B, C>), grid, block, 0, stream, Y, nchunk, Yp, ph, out, out_stride,
                           good, nsamps_out, (const dd_fdd_chunk*)nullptr, (size_t)0);
    return hipGetLastError();
}

template <int A, int B, int C>
hipError_t launch_inverse_chunks(const float2* Y, size_t nrows, size_t row0, const dd_fdd_chunk* chunks,
                                 size_t nchunk, size_t Yp, const float2* ph, size_t good,
                                 hipStream_t stream, int packed = 0)
{
    if (!chunks || packed < 0 || packed > 2) return hipErrorInvalidValue;
    if (!nchunk || !nrows) return hipSuccess;
    dim3 grid;
    if (!blk_grid(packed == 2, nchunk, nrows, grid)) return hipErrorInvalidValue;
    const dim3 block(FftShape<A, B, C>::T);
    if (packed == 2)
        hipLaunchKernelGGL((fdd_inverse_k<A, B, C, true, true, true>), grid, block, 0, stream, Y, nchunk, Yp, ph,
                           (float*)nullptr, (size_t)0, good, (size_t)0, chunks, row0);
    else if (packed)
        hipLaunchKernelGGL((fdd_inverse_k<A, B, C, true, true>), grid, block, 0, stream, Y, nchunk, Yp, ph,
                           (float*)nullptr, (size_t)0, good, (size_t)0, chunks, row0);
    else
        hipLaunchKernelGGL((fdd_inverse_k<A, B, C, true>), grid, block, 0, stream, Y, nchunk, Yp, ph,
                           (float*)nullptr, (size_t)0, good, (size_t)0, chunks, row0);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Segments of nfft = 2^17 and 2^18 samples (M = nfft/2 = R*L packed points,
// L = 32768, R = 2 or 4): one component of M points no longer fits the LDS, so
// each transform is split in time (decimation by R) into R on-chip L-point
// transforms plus a radix-R combine, with one round trip through a work
// buffer of M complex values per transform.
//
// Forward, for n = r + R*j and k' = k + L*s (k < L, s < R):
//   Z_r[k]   = sum_j z[r + R j] W_L^{jk}                         (split_fwd_k)
//   Z[k+L s] = sum_r W_R^{rs} W_M^{rk} Z_r[k]                     (combine_fwd_k)
// and the combine kernel also pairs Z into the real spectrum X[0..M]; the
// mirror M - (k + L s) = (L - k) + L (R-1-s), so a thread taking k and L - k
// holds every value its outputs need.
// Inverse: Zu[k'] is formed from Y[k'] and Y[M - k'] as in fdd_inverse_k, then
//   Zr'[k]   = W_M^{-rk} sum_s W_R^{-rs} Zu[k + L s]                (split_inv_k)
//   z[r+R j] = sum_k Zr'[k] W_L^{-jk}                             (combine_inv_k)
// Twiddles: ph[m] = e^{+2 pi i m/nfft}; the on-chip engine takes the table
// of its own size, ph64k[m] = e^{+2 pi i m/65536}.
// ---------------------------------------------------------------------------

constexpr int LBIG = 32768;  // on-chip points per transform (32 x 32 x 32)

// radix-R DFT of v (R = 2 or 4), DIR = -1 forward, +1 inverse
template <int R, int DIR>
__device__ __forceinline__ void dft_small(float2 (&v)[R])
{
    if constexpr (R == 2) {
        const float2 a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    } else {
        const float2 a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
        const float2 b0 = cadd(v[1], v[3]), b1 = csub(v[1], v[3]);
        // DIR * i * b1
        const float2 jb1 = (DIR < 0) ? make_float2(b1.y, -b1.x) : make_float2(-b1.y, b1.x);
        v[0] = cadd(a0, b0);
        v[2] = csub(a0, b0);
        v[1] = cadd(a1, jb1);
        v[3] = csub(a1, jb1);
    }
}

// split_fwd_k: residue r = blockIdx.x % R of chunk blockIdx.x / R, channel
// blockIdx.y; writes Z_r (L points) to work[((c*nchunk + chunk)*R + r)*L].
// TAB: the chunk's rows, first sample and valid length come from tab[chunk].
template <int R, bool TAB = false>
__global__ __launch_bounds__(1024) void split_fwd_k(
    const uint8_t* __restrict__ rows, size_t row_bytes, size_t chunk_step, size_t nvalid, size_t period,
    const float* __restrict__ chan_scale, const float2* __restrict__ ph64k, float2* __restrict__ work,
    size_t nchunk, uint32_t nbits, const dd_fdd_chunk* __restrict__ tab)
{
    using S = FftShape<32, 32, 32>;
    __shared__ float lds[S::LDS_FLOATS];
    const size_t chunk = blockIdx.x / R, c = blockIdx.y;
    const uint32_t r = blockIdx.x % R;
    size_t t0 = chunk * chunk_step;
    if constexpr (TAB) {
        const dd_fdd_chunk e = tab[chunk];
        rows = e.rows;
        row_bytes = e.row_bytes;
        nvalid = e.nvalid;
        t0 = e.t0;
    }
    const uint8_t* src = rows + c * row_bytes;
    const float s = chan_scale[c];
    const float mid = 0.5f * (float)((1u << nbits) - 1u);
    const uint32_t lspw = 5 - __builtin_ctz(nbits);
    const uint32_t mask = (1u << nbits) - 1u;
    const uint32_t* wsrc = reinterpret_cast<const uint32_t*>(src);
    auto value = [&](size_t t) -> uint32_t {
        if (nbits == 8) return src[t];
        return (wsrc[t >> lspw] >> ((uint32_t)(t & ((1u << lspw) - 1u)) * nbits)) & mask;
    };
    auto sample = [&](size_t t) -> float {
        if (period) {
            // t < 2 period for every chunk of a series at least one segment
            // long: a subtraction, the 64-bit remainder only beyond that
            if (t >= period) t -= period;
            if (t >= period) t %= period;
        }
        return (t < nvalid) ? ((float)value(t) - mid) * s : 0.f;
    };
    // packed point r + R*n: samples t0 + 2(r + R n) and the next
    auto load = [&](int n) -> float2 {
        const size_t t = t0 + 2 * ((size_t)r + (size_t)R * (size_t)n);
        if (!period && nbits == 8 && t + 1 < nvalid) {
            const uint16_t pair = *reinterpret_cast<const uint16_t*>(src + t);
            return make_float2(((float)(pair & 0xff) - mid) * s, ((float)(pair >> 8) - mid) * s);
        }
        return make_float2(sample(t), sample(t + 1));
    };
    float2 q[32];
    fft3_core<32, 32, 32, -1>(ph64k, lds, load, q);
    float2* dst = work + ((c * nchunk + chunk) * R + r) * (size_t)LBIG;
    const int t = threadIdx.x;
#pragma unroll
    for (int k0 = 0; k0 < 32; ++k0) dst[t + S::T3 * k0] = q[k0];
}

// combine_fwd_k: transform (c, chunk) = blockIdx.x; threads take k = 0 ..
// L/2 (and L - k); writes X[0..M] of the transform at X + blockIdx.x * Xp.
template <int R>
__global__ __launch_bounds__(256) void combine_fwd_k(const float2* __restrict__ work, const float2* __restrict__ ph,
                                                     uint32_t nfft, float2* __restrict__ X, size_t Xp)
{
    constexpr uint32_t L = LBIG, M = R * LBIG;
    const float2* zr = work + (size_t)blockIdx.x * M;
    float2* dst = X + (size_t)blockIdx.x * Xp;
    const uint32_t mask = nfft - 1;
    // W_M^{rk} = conj(ph[2 r k])
    auto zk = [&](uint32_t k, float2 (&z)[R]) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const float2 v = zr[(size_t)r * L + (k & (L - 1))];
            z[r] = r ? cmul(v, conjf2(ph[(2u * r * k) & mask])) : v;
        }
        dft_small<R, -1>(z);  // z[s] = Z[k + L s]
    };
    // X[kk] from Z[kk] = a and Z[M - kk] = m
    auto pair_out = [&](uint32_t kk, float2 a, float2 m) {
        const float2 pk = ph[kk];
        const float ux = a.x + m.x, uy = a.y - m.y;
        const float vx = a.x - m.x, vy = a.y + m.y;
        const float wr = vx * pk.x + vy * pk.y, wi = vy * pk.x - vx * pk.y;
        dst[kk] = make_float2(0.5f * (ux + wi), 0.5f * (uy - wr));
    };
    for (uint32_t k = threadIdx.x; k <= L / 2; k += 256) {
        float2 a[R];
        zk(k, a);
        if (k == 0) {
            // Z[L s] pairs with Z[M - L s] = Z[L (R - s)] (Z[M] = Z[0])
#pragma unroll
            for (int s = 0; s < R; ++s) pair_out((uint32_t)s * L, a[s], a[(R - s) % R]);
            dst[M] = make_float2(a[0].x - a[0].y, 0.f);
        } else if (k == L / 2) {
#pragma unroll
            for (int s = 0; s < R; ++s) pair_out(k + (uint32_t)s * L, a[s], a[R - 1 - s]);
        } else {
            float2 b[R];
            zk(L - k, b);
#pragma unroll
            for (int s = 0; s < R; ++s) {
                pair_out(k + (uint32_t)s * L, a[s], b[R - 1 - s]);
                pair_out(L - k + (uint32_t)s * L, b[s], a[R - 1 - s]);
            }
        }
    }
}

// split_inv_k: row / chunk transform blockIdx.x of Y (bins 0..M at Yp); writes
// Zr' (R x L points) to work[blockIdx.x*M].
template <int R, bool PACKED = false>
__global__ __launch_bounds__(256) void split_inv_k(const float2* __restrict__ Y, size_t Yp,
                                                   const float2* __restrict__ ph, uint32_t nfft,
                                                   float2* __restrict__ work)
{
    constexpr uint32_t L = LBIG, M = R * LBIG;
    const float2* src = Y + (size_t)blockIdx.x * Yp;
    float2* dst = work + (size_t)blockIdx.x * M;
    const uint32_t mask = nfft - 1;
    if constexpr (PACKED) {
        // Y holds Zu[0..M-1]: the radix-R split only
        for (uint32_t k = threadIdx.x; k < L; k += 256) {
            float2 u[R];
#pragma unroll
            for (int s = 0; s < R; ++s) u[s] = src[k + (uint32_t)s * L];
            dft_small<R, +1>(u);
#pragma unroll
            for (int r = 0; r < R; ++r) dst[(size_t)r * L + k] = r ? cmul(u[r], ph[(2u * r * k) & mask]) : u[r];
        }
        return;
    }
    const float ymx = src[M].x;
    // Zu[k'] = (Y[k'] + conj Y[M-k']) + i e^{+2 pi i k'/nfft} (Y[k'] - conj Y[M-k'])
    auto zu = [&](uint32_t kk, float2 a, float2 m) -> float2 {
        float2 b = conjf2(m);
        if (kk == 0) a.y = 0.f, b = make_float2(ymx, 0.f);
        const float2 e = cadd(a, b), d = csub(a, b);
        const float2 o = cmul(d, ph[kk]);
        return make_float2(e.x - o.y, e.y + o.x);
    };
    // Zr'[k] = W_M^{-rk} sum_s W_R^{-rs} Zu[k + L s]
    auto out = [&](uint32_t k, float2 (&u)[R]) {
        dft_small<R, +1>(u);
#pragma unroll
        for (int r = 0; r < R; ++r) dst[(size_t)r * L + k] = r ? cmul(u[r], ph[(2u * r * k) & mask]) : u[r];
    };
    for (uint32_t k = threadIdx.x; k <= L / 2; k += 256) {
        float2 ya[R], u[R];
#pragma unroll
        for (int s = 0; s < R; ++s) ya[s] = src[k + (uint32_t)s * L];
        if (k == 0) {
            // Y[M - L s] = Y[L (R - s)] for s > 0; Y[M] for s = 0 (taken by zu)
#pragma unroll
            for (int s = 0; s < R; ++s) u[s] = zu((uint32_t)s * L, ya[s], ya[(R - s) % R]);
            out(0, u);
        } else if (k == L / 2) {
#pragma unroll
            for (int s = 0; s < R; ++s) u[s] = zu(k + (uint32_t)s * L, ya[s], ya[R - 1 - s]);
            out(k, u);
        } else {
            float2 yb[R], w[R];
#pragma unroll
            for (int s = 0; s < R; ++s) yb[s] = src[L - k + (uint32_t)s * L];
#pragma unroll
            for (int s = 0; s < R; ++s) {
                u[s] = zu(k + (uint32_t)s * L, ya[s], yb[R - 1 - s]);
                w[s] = zu(L - k + (uint32_t)s * L, yb[s], ya[R - 1 - s]);
            }
            out(k, u);
            out(L - k, w);
        }
    }
}

// combine_inv_k: residue r = blockIdx.x % R of chunk blockIdx.x / R, row
// blockIdx.y: inverse L-point transform of Zr', samples 2(r + R j) and the
// next of the chunk, times 1/nfft, the first `good` of them written. TAB: the
// chunk's output row (row row0 + blockIdx.y of tab[chunk].out), first output
// sample and output length come from tab[chunk].
template <int R, bool TAB = false>
__global__ __launch_bounds__(1024) void combine_inv_k(const float2* __restrict__ work, size_t nchunk,
                                                      const float2* __restrict__ ph64k, float* __restrict__ out,
                                                      size_t out_stride, size_t good, size_t nsamps_out,
                                                      const dd_fdd_chunk* __restrict__ tab, size_t row0)
{
    using S = FftShape<32, 32, 32>;
    __shared__ float lds[S::LDS_FLOATS];
    const size_t chunk = blockIdx.x / R, row = blockIdx.y;
    const uint32_t r = blockIdx.x % R;
    const float2* src = work + ((row * nchunk + chunk) * R + r) * (size_t)LBIG;
    size_t t0 = chunk * good;
    if constexpr (TAB) {
        const dd_fdd_chunk e = tab[chunk];
        out = e.out;
        out_stride = e.out_stride;
        nsamps_out = e.nsamps_out;
        t0 = e.t0;
    }
    auto load = [&](int k) -> float2 { return src[(uint32_t)k]; };
    float2 q[32];
    fft3_core<32, 32, 32, +1>(ph64k, lds, load, q);
    float* const base = out + (row0 + row) * out_stride + t0;
    const uint32_t lim = (uint32_t)min(good, nsamps_out - t0);
    const float inv_n = 1.0f / (float)(2 * R * LBIG);
    const int t = threadIdx.x;
    // stores bounded by the record count of a raw buffer resource (dropped
    // past lim by the hardware), with no branch per sample
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(base, 0, (int)(lim * 4u), kRawBufferWord3);
#pragma unroll
    for (int k0 = 0; k0 < 32; ++k0) {
        const uint32_t ta = 2u * (r + (uint32_t)R * ((uint32_t)t + (uint32_t)S::T3 * k0));
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, q[k0].x * inv_n), rs, (int)(4u * ta), 0, 0);
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, q[k0].y * inv_n), rs, (int)(4u * ta + 4u), 0, 0);
    }
}

template <int R>
hipError_t launch_forward_split(const uint8_t* rows, size_t row_bytes, size_t chunk_step, size_t nvalid,
                                size_t period, const float* chan_scale, const float2* ph, const float2* ph64k,
                                float2* X, size_t nchan, size_t nchunk, size_t Xp, unsigned nbits, float2* work,
                                hipStream_t stream, const dd_fdd_chunk* tab = nullptr)
{
    constexpr size_t M = (size_t)R * LBIG;
    if (Xp < M + 1 || !work || !ph64k) return hipErrorInvalidValue;
    if (nbits != 1 && nbits != 2 && nbits != 4 && nbits != 8) return hipErrorInvalidValue;
    if (nbits != 8 && (row_bytes % 4)) return hipErrorInvalidValue;
    if (!nchunk || !nchan) return hipSuccess;
    if (nchunk * R >= (size_t(1) << 31) || nchan * nchunk >= (size_t(1) << 31)) return hipErrorInvalidValue;
    if (tab)
        hipLaunchKernelGGL((split_fwd_k<R, true>), dim3((unsigned)(nchunk * R), (unsigned)nchan), dim3(1024), 0,
                           stream, (const uint8_t*)nullptr, (size_t)0, (size_t)0, (size_t)0, (size_t)0, chan_scale,
                           ph64k, work, nchunk, (uint32_t)nbits, tab);
    else
        hipLaunchKernelGGL((split_fwd_k<R>), dim3((unsigned)(nchunk * R), (unsigned)nchan), dim3(1024), 0, stream,
                           rows, row_bytes, chunk_step, nvalid, period, chan_scale, ph64k, work, nchunk,
                           (uint32_t)nbits, (const dd_fdd_chunk*)nullptr);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    // X rows of transform (c, chunk) at (c*nchunk + chunk)*Xp: one row per blockIdx.y
    hipLaunchKernelGGL((combine_fwd_k<R>), dim3((unsigned)(nchan * nchunk)), dim3(256), 0, stream, work, ph,
                       (uint32_t)(2 * M), X, Xp);
    return hipGetLastError();
}

template <int R>
hipError_t launch_inverse_split(const float2* Y, size_t nrows, size_t nchunk, size_t Yp, const float2* ph,
                                const float2* ph64k, float* out, size_t out_stride, size_t good, size_t nsamps_out,
                                float2* work, hipStream_t stream, const dd_fdd_chunk* tab = nullptr,
                                size_t row0 = 0, bool packed = false)
{
    constexpr size_t M = (size_t)R * LBIG;
    if (Yp < M + 1 || !work || !ph64k) return hipErrorInvalidValue;
    if (!nchunk || !nrows) return hipSuccess;
    if (nrows * nchunk >= (size_t(1) << 31) || nchunk * R >= (size_t(1) << 31)) return hipErrorInvalidValue;
    if (packed)
        hipLaunchKernelGGL((split_inv_k<R, true>), dim3((unsigned)(nrows * nchunk)), dim3(256), 0, stream, Y, Yp,
                           ph, (uint32_t)(2 * M), work);
    else
        hipLaunchKernelGGL((split_inv_k<R>), dim3((unsigned)(nrows * nchunk)), dim3(256), 0, stream, Y, Yp, ph,
                           (uint32_t)(2 * M), work);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (tab)
        hipLaunchKernelGGL((combine_inv_k<R, true>), dim3((unsigned)(nchunk * R), (unsigned)nrows), dim3(1024), 0,
                           stream, work, nchunk, ph64k, (float*)nullptr, (size_t)0, good, (size_t)0, tab, row0);
    else
        hipLaunchKernelGGL((combine_inv_k<R>), dim3((unsigned)(nchunk * R), (unsigned)nrows), dim3(1024), 0, stream,
                           work, nchunk, ph64k, out, out_stride, good, nsamps_out, (const dd_fdd_chunk*)nullptr,
                           (size_t)0);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Real transform of length N = 16384*m per row (whole-series Fourier output).
// Packed z[n] = s[2n] + i s[2n+1], L = N/2 = 8192*m points, four-step with
// n = n1 + m*n2, k = k2 + 8192*k1:
//   columns: z viewed as [n2][n1] transposed to [n1][n2] (tiles through LDS),
//            so each column n1 is contiguous
//   step A: 8192-point DFT over n2 for each (row, n1), times W_L^{n1 k2}
//   step B: m-point DFT over n1 for each (row, k2): a direct sum, O(m^2) per
//           column (one radix-2 split for even m halves it); with the
//           twiddles in LDS for m <= 4096 (N <= 2^26), read from twL above.
//           At the reference workloads m is small (66 for 2^20 samples), where
//           the step is bound by re-reading its column, not by the sums; long
//           series grow quadratically in this step (a mixed-radix m-point
//           transform would make it O(m log m))
//   step C: real-input pairing X[k] = (Z[k] + conj Z[L-k])/2
//                                     - i/2 W_N^k (Z[k] - conj Z[L-k])
// twL[j] = e^{-2 pi i j/L}, twN[k] = e^{-2 pi i k/N} (j, k < L), built in double.
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(256) void fourier_cols_k(const float2* __restrict__ Z, size_t z_stride, size_t m,
                                                      float2* __restrict__ T, size_t t_stride)
{
    __shared__ float2 tile[32][33];
    const size_t row = blockIdx.z;
    const size_t n1_0 = (size_t)blockIdx.x * 32, n2_0 = (size_t)blockIdx.y * 32;
    const float2* src = Z + row * z_stride;
    float2* dst = T + row * t_stride;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    for (int r = ty; r < 32; r += 8) {
        const size_t n1 = n1_0 + tx;
        if (n1 < m) tile[r][tx] = src[n1 + m * (n2_0 + r)];
    }
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {
        const size_t n1 = n1_0 + r;
        if (n1 < m) dst[n1 * 8192 + n2_0 + tx] = tile[tx][r];
    }
}

__global__ __launch_bounds__(512) void fourier_step_a_k(
    const float2* __restrict__ T, size_t t_stride, size_t m, const float2* __restrict__ ph8192,
    const float2* __restrict__ twL, float2* __restrict__ A, size_t a_stride)
{
    using Sh = FftShape<16, 16, 32>;
    __shared__ float lds[Sh::LDS_FLOATS];
    // XCD-aware order: workgroups are dispatched round-robin over the 8 XCDs;
    // each XCD takes a contiguous range of (row, n1), so the neighbouring
    // columns n1, n1+1, ... that share the strided input's cache lines run on
    // one L2
    const uint32_t nb = gridDim.x * gridDim.y, bid = blockIdx.x + gridDim.x * blockIdx.y;
    const uint32_t xcd = bid & 7u, per = nb >> 3, rem = nb & 7u;
    const uint32_t lin = xcd * per + min(xcd, rem) + (bid >> 3);
    const size_t n1 = lin % m, row = lin / m;
    const float2* src = T + row * t_stride + n1 * 8192;  // column n1, contiguous
    float2* dst = A + row * a_stride + n1 * 8192;
    // W_L^{n1 k2} for k2 = 64 h + l as W_L^{64 n1 h} W_L^{n1 l}: two exact
    // table values per workgroup in LDS (both indices < L) instead of a
    // gather from the L-entry table per output; written before the engine's
    // first barrier, read after its last
    __shared__ float2 tw_hi[128], tw_lo[64];
    const int t = threadIdx.x;
    if (t < 128) tw_hi[t] = twL[64 * n1 * (size_t)t];
    else if (t < 192) tw_lo[t - 128] = twL[n1 * (size_t)(t - 128)];
    // loads and stores take a 32-bit byte offset from the uniform column base
    // (opaque, so that it is not folded into a 64-bit address per access)
    const char* const srcb = reinterpret_cast<const char*>(src);
    auto load = [&](int n2) -> float2 {
        uint32_t off = (uint32_t)n2 * 8u;
        asm volatile("" : "+v"(off));
        return *reinterpret_cast<const float2*>(srcb + off);
    };
    char* const dstb = reinterpret_cast<char*>(dst);
    auto store = [&](int k2, float2 v) {
        const float2 w = cmul(tw_hi[k2 >> 6], tw_lo[k2 & 63]);
        uint32_t off = (uint32_t)k2 * 8u;
        asm volatile("" : "+v"(off));
        *reinterpret_cast<float2*>(dstb + off) = cmul(v, w);
    };
    fft3<16, 16, 32, -1>(ph8192, lds, load, store);
}

__global__ __launch_bounds__(256) void fourier_step_b_k(const float2* __restrict__ A, size_t a_stride,
                                                        size_t m, const float2* __restrict__ twL,
                                                        float2* __restrict__ Z, size_t z_stride)
{
    // the m distinct twiddles W_m^e = twL[8192 e] (e < m), staged in LDS when
    // they fit: each is then a broadcast LDS read rather than a scalar load
    // waited on before its use
    __shared__ float2 wt[4096];
    const bool tw_lds = m <= 4096;
    if (tw_lds) {
        for (size_t i = threadIdx.x; i < m; i += 256) wt[i] = twL[i * 8192];
        __syncthreads();
    }
    const size_t k2 = (size_t)blockIdx.x * 256 + threadIdx.x;
    const size_t row = blockIdx.y;
    if (k2 >= 8192) return;
    const float2* a = A + row * a_stride + k2;
    float2* z = Z + row * z_stride + k2;
    if (tw_lds && m % 2 == 0) {
        // even m: one radix-2 split, E[k] / O[k] = the h-point DFTs (h = m/2)
        // of the even / odd n1, X[k] = E[k] + W_m^k O[k], X[k+h] = E[k] - W_m^k O[k];
        // W_h^e = W_m^{2e}. KH outputs k per pass over the inputs: the pass
        // count sets how often the column is re-read from HBM (the kernel is
        // bound by those reads: 17 instead of 8 outputs per pass takes the
        // 66-point step of the 2048-DM workload from 5 passes to 2)
        constexpr int KH = 17;
        const uint32_t h = (uint32_t)(m / 2);
        for (uint32_t kb = 0; kb < h; kb += KH) {
            float2 E[KH], O[KH];
            uint32_t e[KH];  // (p * (kb + j)) mod h
#pragma unroll
            for (int j = 0; j < KH; ++j) E[j] = O[j] = make_float2(0.f, 0.f), e[j] = 0;
            for (uint32_t p = 0; p < h; ++p) {
                const float2 ve = a[(size_t)(2 * p) * 8192], vo = a[(size_t)(2 * p + 1) * 8192];
#pragma unroll
                for (int j = 0; j < KH; ++j) {
                    if (kb + j >= h) break;
                    const float2 w = wt[2 * e[j]];
                    E[j].x = fmaf(ve.x, w.x, fmaf(-ve.y, w.y, E[j].x));
                    E[j].y = fmaf(ve.x, w.y, fmaf(ve.y, w.x, E[j].y));
                    O[j].x = fmaf(vo.x, w.x, fmaf(-vo.y, w.y, O[j].x));
                    O[j].y = fmaf(vo.x, w.y, fmaf(vo.y, w.x, O[j].y));
                    e[j] += kb + j;
                    if (e[j] >= h) e[j] -= h;
                }
            }
#pragma unroll
            for (int j = 0; j < KH; ++j) {
                const uint32_t k = kb + j;
                if (k >= h) break;
                const float2 t = cmul(O[j], wt[k]);
                z[(size_t)k * 8192] = cadd(E[j], t);
                z[(size_t)(k + h) * 8192] = csub(E[j], t);
            }
        }
        return;
    }
    // KB outputs k1 per pass over the inputs, so each input is loaded once per
    // pass instead of once per output; every output still sums n1 = 0, 1, ...
    // in order with the same operations
    constexpr int KB = 16;
    for (size_t k1b = 0; k1b < m; k1b += KB) {
        float2 acc[KB];
        uint32_t e[KB];  // (n1 * (k1b + j)) mod m
#pragma unroll
        for (int j = 0; j < KB; ++j) acc[j] = make_float2(0.f, 0.f), e[j] = 0;
        for (size_t n1 = 0; n1 < m; ++n1) {
            const float2 v = a[n1 * 8192];
#pragma unroll
            for (int j = 0; j < KB; ++j) {
                if (k1b + j >= m) break;
                const float2 w = tw_lds ? wt[e[j]] : twL[(size_t)e[j] * 8192];
                acc[j].x = fmaf(v.x, w.x, fmaf(-v.y, w.y, acc[j].x));
                acc[j].y = fmaf(v.x, w.y, fmaf(v.y, w.x, acc[j].y));
                e[j] += (uint32_t)(k1b + j);
                if (e[j] >= m) e[j] -= (uint32_t)m;
            }
        }
#pragma unroll
        for (int j = 0; j < KB; ++j)
            if (k1b + j < m) z[(k1b + j) * 8192] = acc[j];
    }
}

__global__ __launch_bounds__(256) void fourier_step_c_k(const float2* __restrict__ Z, size_t z_stride,
                                                        size_t L, const float2* __restrict__ twN,
                                                        float2* __restrict__ out, size_t out_stride)
{
    const size_t k = (size_t)blockIdx.x * 256 + threadIdx.x;
    const size_t row = blockIdx.y;
    if (k >= L) return;
    const float2* z = Z + row * z_stride;
    const float2 a = z[k];
    const float2 bz = z[k ? L - k : 0];
    const float2 e = make_float2(0.5f * (a.x + bz.x), 0.5f * (a.y - bz.y));
    const float2 d = make_float2(a.x - bz.x, a.y + bz.y);
    const float2 o = make_float2(0.5f * d.y, -0.5f * d.x);
    const float2 r = cmul(o, twN[k]);
    out[row * out_stride + k] = make_float2(e.x + r.x, e.y + r.y);
}

// Steps B and C in one pass for even m <= 144. A tile is columns A = 64w + l
// and their mirrors B = 8192 - A (l < 64, all m rows) of one row: the
// workgroup stages them in LDS, and each of its BC_WAVES waves sums BC_KH of the
// radix-2 outputs k (E[k] / O[k] over the even / odd rows, as step B) for
// column A and, at h-1-k, for column B, then pairs them. The real-input
// pairing of bin kk = k2 + 8192 k1 takes Z[L - kk], which for k2 = A >= 1 is
// row m-1-k1 of column B: rows k and k+h of A pair with rows 2h-1-k and h-1-k
// of B, and the other way round. Column 4096 is its own mirror (B = A; its B
// outputs are not written again); column 0 is paired within itself by
// fourier_col0_k. Each column is read from HBM once, and the intermediate Z
// never goes to HBM. The grid is persistent (tiles t = blockIdx.x + i
// gridDim.x, w = t % 65, row = t / 65): the next tile's loads are in flight
// in registers while the current one is summed.
// Two shapes: BC_MAXM = 72 with 5 outputs per wave (LDS of two workgroups
// per CU) and 144 with 9 (one workgroup per CU).
constexpr int BC_WAVES = 8;
constexpr int BC_NT = BC_WAVES * 64;
constexpr int BC_TILES_PER_ROW = 65;

template <int BC_MAXM, int BC_KH>
__global__ __launch_bounds__(BC_NT) void fourier_step_bc_k(const float2* __restrict__ A, size_t a_stride,
                                                         size_t m, uint32_t nrows,
                                                         const float2* __restrict__ twL,
                                                         const float2* __restrict__ twN, float2* __restrict__ out,
                                                         size_t out_stride)
{
    constexpr int BC_RPT = BC_MAXM * 128 / BC_NT;              // staged values per thread
    constexpr int BC_LDS = 2 * BC_MAXM + BC_MAXM * 128;        // float2: W_m^e, W_N^{8192 k1}, [row][128] values
    static_assert(BC_MAXM / 2 <= BC_WAVES * BC_KH && BC_NT % 128 == 0 && BC_MAXM % (BC_NT / 128) == 0,
                  "the waves cover h = m/2 outputs, the threads the staged rows");
    __shared__ float2 bc_lds[BC_LDS];
    float2* const wt = bc_lds;                   // W_m^e, e < m
    float2* const wk = bc_lds + BC_MAXM;         // W_N^{8192 k1}, k1 < m
    float2* const col = bc_lds + 2 * BC_MAXM;    // [row][128]: lane l's column A at l, B at 64 + l
    const uint32_t h = (uint32_t)(m / 2);
    const int t = threadIdx.x, lane = t & 63;
    const uint32_t wave = __builtin_amdgcn_readfirstlane(t >> 6);
    const uint32_t kb = wave * BC_KH;
    const uint32_t ntile = BC_TILES_PER_ROW * nrows;
    for (uint32_t i = t; i < m; i += BC_NT) wt[i] = twL[(size_t)i * 8192], wk[i] = twN[(size_t)i * 8192];
    // staging: thread t holds tile column c = t & 127 at rows RS i + (t >> 7)
    // (RS = BC_NT / 128 rows a step), i < BC_RPT, the rows past m - 1 clamped
    // to m - 1 (loaded and stored, never summed); columns clamped into
    // [1, 4096] (the lanes past them compute but do not write). Byte offsets
    // from the row base fit 32 bits (< BC_MAXM x 8192 x 8 <= 144 x 8192 x 8)
    constexpr uint32_t RS = BC_NT / 128;
    const uint32_t sc = t & 127, sr0 = t >> 7;
    float2 v[BC_RPT];
    auto load = [&](uint32_t tile) {
        const uint32_t w = tile % BC_TILES_PER_ROW, row = tile / BC_TILES_PER_ROW;
        const uint32_t ca = min(max(64 * w + (sc & 63), 1u), 4096u);
        const char* base = reinterpret_cast<const char*>(A + row * a_stride);
        const uint32_t off0 = ((sc < 64) ? ca : 8192 - ca) * 8;
        // opaque copy of the last row: keeps the clamped offsets from being
        // hoisted out of the tile loop into as many scalar registers
        uint32_t rlast = 2 * h - 1;
        asm volatile("" : "+s"(rlast));
#pragma unroll
        for (int i = 0; i < BC_RPT; ++i) {
            const uint32_t r = min((uint32_t)(RS * i) + sr0, rlast);
            v[i] = *reinterpret_cast<const float2*>(base + off0 + r * (8192 * 8));
        }
    };
    // pairing of bin kk = k2 + 8192 k1 with Z[L - kk] (the step-C formula);
    // W_N^kk = W_N^{k2} W_N^{8192 k1}: the lane's column factor times an LDS
    // entry (two exact table values) rather than a gather from the L-entry
    // table per output
    auto pair_out = [&](float2* o_row, uint32_t k2, uint32_t k1, float2 wc, float2 za, float2 zb) {
        const float2 ev = make_float2(0.5f * (za.x + zb.x), 0.5f * (za.y - zb.y));
        const float2 d = make_float2(za.x - zb.x, za.y + zb.y);
        const float2 o = make_float2(0.5f * d.y, -0.5f * d.x);
        const float2 r = cmul(o, cmul(wc, wk[k1]));
        o_row[k2 + 8192 * (size_t)k1] = make_float2(ev.x + r.x, ev.y + r.y);
    };
    uint32_t tile = blockIdx.x;
    if (tile < ntile) load(tile);
    for (; tile < ntile; tile += gridDim.x) {
        __syncthreads();  // the previous tile's sums have read col (and wt is written)
#pragma unroll
        for (int i = 0; i < BC_RPT; ++i) col[(RS * i + sr0) * 128 + sc] = v[i];
        __syncthreads();
        if (tile + gridDim.x < ntile) load(tile + gridDim.x);
        if (kb < h) {
            // all BC_KH outputs are summed (the ones past h - 1 are not written):
            // no branch in the loop, and e stays in scalar registers
            float2 EA[BC_KH], OA[BC_KH], EB[BC_KH], OB[BC_KH];
            uint32_t e[BC_KH + 1];  // (p * (kb + j)) mod h, j <= BC_KH
#pragma unroll
            for (int j = 0; j < BC_KH; ++j) EA[j] = OA[j] = EB[j] = OB[j] = make_float2(0.f, 0.f);
#pragma unroll
            for (int j = 0; j <= BC_KH; ++j) e[j] = 0;
            for (uint32_t p = 0; p < h; ++p) {
                const float2 ae = col[(2 * p) * 128 + lane], ao = col[(2 * p + 1) * 128 + lane];
                const float2 be = col[(2 * p) * 128 + 64 + lane], bo = col[(2 * p + 1) * 128 + 64 + lane];
#pragma unroll
                for (int j = 0; j < BC_KH; ++j) {
                    // A at k = kb + j: W_h^{p k}; B at h-1-k: W_h^{p (h-1-k)} = conj W_h^{p (k+1)}
                    const float2 wa = wt[2 * e[j]], wb0 = wt[2 * e[j + 1]];
                    const float2 wb = make_float2(wb0.x, -wb0.y);
                    EA[j].x = fmaf(ae.x, wa.x, fmaf(-ae.y, wa.y, EA[j].x));
                    EA[j].y = fmaf(ae.x, wa.y, fmaf(ae.y, wa.x, EA[j].y));
                    OA[j].x = fmaf(ao.x, wa.x, fmaf(-ao.y, wa.y, OA[j].x));
                    OA[j].y = fmaf(ao.x, wa.y, fmaf(ao.y, wa.x, OA[j].y));
                    EB[j].x = fmaf(be.x, wb.x, fmaf(-be.y, wb.y, EB[j].x));
                    EB[j].y = fmaf(be.x, wb.y, fmaf(be.y, wb.x, EB[j].y));
                    OB[j].x = fmaf(bo.x, wb.x, fmaf(-bo.y, wb.y, OB[j].x));
                    OB[j].y = fmaf(bo.x, wb.y, fmaf(bo.y, wb.x, OB[j].y));
                }
#pragma unroll
                for (int j = 0; j <= BC_KH; ++j) {
                    // x mod h for x < 2h as a minimum (x - h wraps when x < h):
                    // selects in scalar registers, where a compare-and-subtract
                    // was lowered to vector instructions
                    const uint32_t x = e[j] + kb + j;
                    e[j] = min(x, x - h);
                }
            }
            const uint32_t w = tile % BC_TILES_PER_ROW, row = tile / BC_TILES_PER_ROW;
            const uint32_t cA = 64 * w + lane, cB = 8192 - cA;
            if (cA >= 1 && cA <= 4096) {
                float2* o_row = out + row * out_stride;
                const float2 wA = twN[cA], wB = twN[cB];
#pragma unroll
                for (int j = 0; j < BC_KH; ++j) {
                    const uint32_t k = kb + j;
                    if (k < h) {
                        const uint32_t kB = h - 1 - k;
                        const float2 ta = cmul(OA[j], wt[k]), tb = cmul(OB[j], wt[kB]);
                        const float2 zA0 = cadd(EA[j], ta), zA1 = csub(EA[j], ta);  // rows k, k+h of A
                        const float2 zB0 = cadd(EB[j], tb), zB1 = csub(EB[j], tb);  // rows h-1-k, 2h-1-k of B
                        pair_out(o_row, cA, k, wA, zA0, zB1);
                        pair_out(o_row, cA, k + h, wA, zA1, zB0);
                        if (cA < 4096) {
                            pair_out(o_row, cB, kB, wB, zB0, zA1);
                            pair_out(o_row, cB, kB + h, wB, zB1, zA0);
                        }
                    }
                }
            }
        }
    }
}

// Column 0 of steps B and C: the m-point DFT of the column as direct sums,
// then row k1 paired with row (m - k1) mod m (m <= 256, one thread per row).
__global__ __launch_bounds__(256) void fourier_col0_k(const float2* __restrict__ A, size_t a_stride, size_t m,
                                                      const float2* __restrict__ twL, const float2* __restrict__ twN,
                                                      float2* __restrict__ out, size_t out_stride)
{
    __shared__ float2 zs[256], wt[256], av[256];
    const size_t row = blockIdx.x;
    const uint32_t k1 = threadIdx.x;
    if (k1 < m) {
        wt[k1] = twL[(size_t)k1 * 8192];
        av[k1] = A[row * a_stride + (size_t)k1 * 8192];
    }
    __syncthreads();
    if (k1 < m) {
        float2 acc = make_float2(0.f, 0.f);
        uint32_t e = 0;
        for (uint32_t n1 = 0; n1 < m; ++n1) {
            const float2 v = av[n1], tw = wt[e];
            acc.x = fmaf(v.x, tw.x, fmaf(-v.y, tw.y, acc.x));
            acc.y = fmaf(v.x, tw.y, fmaf(v.y, tw.x, acc.y));
            e += k1;
            if (e >= m) e -= (uint32_t)m;
        }
        zs[k1] = acc;
    }
    __syncthreads();
    if (k1 < m) {
        const float2 za = zs[k1], zb = zs[k1 ? m - k1 : 0];
        const size_t kk = 8192 * (size_t)k1;
        const float2 ev = make_float2(0.5f * (za.x + zb.x), 0.5f * (za.y - zb.y));
        const float2 d = make_float2(za.x - zb.x, za.y + zb.y);
        const float2 o = make_float2(0.5f * d.y, -0.5f * d.x);
        const float2 r = cmul(o, twN[kk]);
        out[row * out_stride + kk] = make_float2(ev.x + r.x, ev.y + r.y);
    }
}

}  // namespace

// Whether steps B and C run fused: even m <= 144, whose h = m/2 outputs the
// waves of fourier_step_bc_k cover (DEDISP_FDD_BC=0 keeps the two passes).
inline bool bc_fused(size_t m)
{
    static const bool off = [] {
        const char* v = std::getenv("DEDISP_FDD_BC");
        return v && v[0] == '0';
    }();
    return !off && m % 2 == 0 && m <= 144;
}

extern "C" hipError_t dd_fdd_rfft_rows(const float* S, size_t s_stride, size_t nrows, size_t N,
                                       const float2* ph16384, const float2* twL, const float2* twN,
                                       float2* work, float2* out, size_t out_stride,
                                       hipStream_t stream)
{
    if (N % 16384 || s_stride % 2 || s_stride < N) return hipErrorInvalidValue;
    const size_t m = N / 16384, L = N / 2;
    // S rows hold L complex each (pitch s_stride/2); work holds nrows x L complex
    float2* Sc = reinterpret_cast<float2*>(const_cast<float*>(S));
    const size_t sp = s_stride / 2;
    // rows in slices that keep every grid dimension within its limit
    constexpr size_t kRows = 32768;
    for (size_t r0 = 0; r0 < nrows; r0 += kRows) {
        const unsigned nr = (unsigned)std::min(kRows, nrows - r0);
        float2* Sr = Sc + r0 * sp;
        float2* Wr = work + r0 * L;
        // columns: S -> work; step A: work -> S; step B: S -> work; step C: work -> out
        hipLaunchKernelGGL(fourier_cols_k, dim3((unsigned)((m + 31) / 32), 8192 / 32, nr), dim3(256), 0, stream,
                           Sr, sp, m, Wr, L);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        // the 8192-point engine uses the 16384-point phasor table (ph[2j] = W_8192^-j)
        hipLaunchKernelGGL(fourier_step_a_k, dim3((unsigned)m, nr), dim3(512), 0, stream, Wr, L, m, ph16384,
                           twL, Sr, sp);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
        if (bc_fused(m)) {
            // steps B and C in one pass (column 0 on its own)
            // persistent: two workgroups per CU for m <= 72 (the LDS tile
            // allows two), one for m <= 144
            int dev = 0, ncu = 0;
            e = hipGetDevice(&dev);
            if (e == hipSuccess) e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
            if (e != hipSuccess) return e;
            const size_t ntile = (size_t)BC_TILES_PER_ROW * nr;
            const bool small = m <= 72;
            const unsigned grid = (unsigned)std::min<size_t>(ntile, (small ? 2 : 1) * (size_t)std::max(ncu, 1));
            if (small)
                hipLaunchKernelGGL((fourier_step_bc_k<72, 5>), dim3(grid), dim3(BC_NT), 0, stream, Sr, sp, m,
                                   (uint32_t)nr, twL, twN, out + r0 * out_stride, out_stride);
            else
                hipLaunchKernelGGL((fourier_step_bc_k<144, 9>), dim3(grid), dim3(BC_NT), 0, stream, Sr, sp, m,
                                   (uint32_t)nr, twL, twN, out + r0 * out_stride, out_stride);
            e = hipGetLastError();
            if (e != hipSuccess) return e;
            hipLaunchKernelGGL(fourier_col0_k, dim3(nr), dim3(256), 0, stream, Sr, sp, m, twL, twN,
                               out + r0 * out_stride, out_stride);
            e = hipGetLastError();
            if (e != hipSuccess) return e;
            continue;
        }
        hipLaunchKernelGGL(fourier_step_b_k, dim3(32, nr), dim3(256), 0, stream, Sr, sp, m, twL, Wr, L);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(fourier_step_c_k, dim3((unsigned)((L + 255) / 256), nr), dim3(256), 0, stream, Wr, L,
                           L, twN, out + r0 * out_stride, out_stride);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// Zero fill by a kernel (16-byte stores), usable inside captured graphs.
__global__ __launch_bounds__(256) void zero_fill_k(uint4* __restrict__ p, size_t n16)
{
    const size_t stride = (size_t)gridDim.x * 256;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n16; i += stride) p[i] = make_uint4(0, 0, 0, 0);
}

extern "C" hipError_t dd_zero_fill(void* p, size_t bytes, hipStream_t stream)
{
    if (((uintptr_t)p | bytes) & 15) return hipErrorInvalidValue;
    const size_t n16 = bytes / 16;
    if (!n16) return hipSuccess;
    const unsigned grid = (unsigned)std::min<size_t>((n16 + 255) / 256, 2048);
    hipLaunchKernelGGL(zero_fill_k, dim3(grid), dim3(256), 0, stream, static_cast<uint4*>(p), n16);
    return hipGetLastError();
}

extern "C" int dd_fdd_fft_supported(size_t nfft)
{
    return nfft == 8192 || nfft == 16384 || nfft == 32768 || nfft == 65536 || nfft == 131072 || nfft == 262144;
}

extern "C" size_t dd_fdd_split_work(size_t nfft, size_t ntransforms)
{
    return nfft > 65536 ? ntransforms * (nfft / 2) * sizeof(float2) : 0;
}

extern "C" hipError_t dd_fdd_forward_split(const uint8_t* rows, size_t row_bytes, size_t chunk_step, size_t nvalid,
                                           size_t period, const float* chan_scale, const float2* ph,
                                           const float2* ph64k, size_t nfft, float2* X, size_t nchan,
                                           size_t nchunk, size_t Xp, unsigned nbits, float2* work,
                                           int blocked, hipStream_t stream)
{
    if (blocked) {
#define DD_FWD(A, B, C) launch_forward<A, B, C>(rows, row_bytes, chunk_step, nvalid, period, chan_scale, ph, X, nchan, nchunk, Xp, nbits, stream, true)
        switch (nfft) {
        case 8192:  return DD_FWD(16, 16, 16);
        case 16384: return DD_FWD(16, 16, 32);
        case 32768: return DD_FWD(16, 32, 32);
        case 65536: return DD_FWD(32, 32, 32);
#undef DD_FWD
        default: return hipErrorInvalidValue;
        }
    }
    switch (nfft) {
    case 131072:
        return launch_forward_split<2>(rows, row_bytes, chunk_step, nvalid, period, chan_scale, ph, ph64k, X, nchan,
                                       nchunk, Xp, nbits, work, stream);
    case 262144:
        return launch_forward_split<4>(rows, row_bytes, chunk_step, nvalid, period, chan_scale, ph, ph64k, X, nchan,
                                       nchunk, Xp, nbits, work, stream);
    default:
        return dd_fdd_forward(rows, row_bytes, chunk_step, nvalid, period, chan_scale, ph, nfft, X, nchan, nchunk,
                              Xp, nbits, stream);
    }
}

extern "C" hipError_t dd_fdd_forward_chunks_split(const dd_fdd_chunk* chunks, size_t nchunk, const float* chan_scale,
                                                  const float2* ph, const float2* ph64k, size_t nfft, float2* X,
                                                  size_t nchan, size_t Xp, unsigned nbits, float2* work,
                                                  int blocked, hipStream_t stream)
{
    if (!chunks) return hipErrorInvalidValue;
    if (blocked) {
#define DD_FWD(A, B, C) launch_forward_chunks<A, B, C>(chunks, nchunk, chan_scale, ph, X, nchan, Xp, nbits, stream, true)
        switch (nfft) {
        case 8192:  return DD_FWD(16, 16, 16);
        case 16384: return DD_FWD(16, 16, 32);
        case 32768: return DD_FWD(16, 32, 32);
        case 65536: return DD_FWD(32, 32, 32);
#undef DD_FWD
        default: return hipErrorInvalidValue;
        }
    }
    switch (nfft) {
    case 131072:
        return launch_forward_split<2>(nullptr, 0, 0, 0, 0, chan_scale, ph, ph64k, X, nchan, nchunk, Xp, nbits, work,
                                       stream, chunks);
    case 262144:
        return launch_forward_split<4>(nullptr, 0, 0, 0, 0, chan_scale, ph, ph64k, X, nchan, nchunk, Xp, nbits, work,
                                       stream, chunks);
    default:
        return dd_fdd_forward_chunks(chunks, nchunk, chan_scale, ph, nfft, X, nchan, Xp, nbits, stream);
    }
}

extern "C" hipError_t dd_fdd_inverse_chunks_split(const float2* Y, size_t nrows, size_t row0,
                                                  const dd_fdd_chunk* chunks, size_t nchunk, size_t Yp,
                                                  const float2* ph, const float2* ph64k, size_t nfft, size_t good,
                                                  float2* work, int packed, hipStream_t stream)
{
    if (!chunks || packed < 0 || packed > 2 || (packed == 2 && nfft > 65536)) return hipErrorInvalidValue;
    switch (nfft) {
    case 8192:  return launch_inverse_chunks<16, 16, 16>(Y, nrows, row0, chunks, nchunk, Yp, ph, good, stream, packed);
    case 16384: return launch_inverse_chunks<16, 16, 32>(Y, nrows, row0, chunks, nchunk, Yp, ph, good, stream, packed);
    case 32768: return launch_inverse_chunks<16, 32, 32>(Y, nrows, row0, chunks, nchunk, Yp, ph, good, stream, packed);
    case 65536: return launch_inverse_chunks<32, 32, 32>(Y, nrows, row0, chunks, nchunk, Yp, ph, good, stream, packed);
    case 131072:
        return launch_inverse_split<2>(Y, nrows, nchunk, Yp, ph, ph64k, nullptr, 0, good, 0, work, stream, chunks,
                                       row0, packed);
    case 262144:
        return launch_inverse_split<4>(Y, nrows, nchunk, Yp, ph, ph64k, nullptr, 0, good, 0, work, stream, chunks,
                                       row0, packed);
    default:
        return hipErrorInvalidValue;
    }
}

extern "C" hipError_t dd_fdd_inverse_split(const float2* Y, size_t nrows, size_t nchunk, size_t Yp, const float2* ph,
                                           const float2* ph64k, size_t nfft, float* out, size_t out_stride,
                                           size_t good, size_t nsamps_out, float2* work, int packed,
                                           hipStream_t stream)
{
    if (packed < 0 || packed > 2 || (packed == 2 && nfft > 65536)) return hipErrorInvalidValue;
    switch (nfft) {
    case 8192:
        return launch_inverse<16, 16, 16>(Y, nrows, nchunk, Yp, ph, out, out_stride, good, nsamps_out, stream, packed);
    case 16384:
        return launch_inverse<16, 16, 32>(Y, nrows, nchunk, Yp, ph, out, out_stride, good, nsamps_out, stream, packed);
    case 32768:
        return launch_inverse<16, 32, 32>(Y, nrows, nchunk, Yp, ph, out, out_stride, good, nsamps_out, stream, packed);
    case 65536:
        return launch_inverse<32, 32, 32>(Y, nrows, nchunk, Yp, ph, out, out_stride, good, nsamps_out, stream, packed);
    case 131072:
        return launch_inverse_split<2>(Y, nrows, nchunk, Yp, ph, ph64k, out, out_stride, good, nsamps_out, work,
                                       stream, nullptr, 0, packed);
    case 262144:
        return launch_inverse_split<4>(Y, nrows, nchunk, Yp, ph, ph64k, out, out_stride, good, nsamps_out, work,
                                       stream, nullptr, 0, packed);
    default:
        return hipErrorInvalidValue;
    }
}

extern "C" hipError_t dd_fdd_forward(const uint8_t* rows, size_t row_bytes, size_t chunk_step,
                                     size_t nvalid, size_t period, const float* chan_scale,
                                     const float2* ph, size_t nfft, float2* X, size_t nchan,
                                     size_t nchunk, size_t Xp, unsigned nbits, hipStream_t stream)
{
#define DD_FWD(A, B, C) launch_forward<A, B, C>(rows, row_bytes, chunk_step, nvalid, period, chan_scale, ph, X, nchan, nchunk, Xp, nbits, stream)
    switch (nfft) {
    case 8192:  return DD_FWD(16, 16, 16);
    case 16384: return DD_FWD(16, 16, 32);
    case 32768: return DD_FWD(16, 32, 32);
    case 65536: return DD_FWD(32, 32, 32);
#undef DD_FWD
    default: return hipErrorInvalidValue;
    }
}

extern "C" hipError_t dd_fdd_forward_chunks(const dd_fdd_chunk* chunks, size_t nchunk,
                                            const float* chan_scale, const float2* ph, size_t nfft,
                                            float2* X, size_t nchan, size_t Xp, unsigned nbits,
                                            hipStream_t stream)
{
#define DD_FWD(A, B, C) launch_forward_chunks<A, B, C>(chunks, nchunk, chan_scale, ph, X, nchan, Xp, nbits, stream)
    switch (nfft) {
    case 8192:  return DD_FWD(16, 16, 16);
    case 16384: return DD_FWD(16, 16, 32);
    case 32768: return DD_FWD(16, 32, 32);
    case 65536: return DD_FWD(32, 32, 32);
#undef DD_FWD
    default: return hipErrorInvalidValue;
    }
}

extern "C" hipError_t dd_fdd_inverse_chunks(const float2* Y, size_t nrows, size_t row0,
                                            const dd_fdd_chunk* chunks, size_t nchunk, size_t Yp,
                                            const float2* ph, size_t nfft, size_t good,
                                            hipStream_t stream)
{
#define DD_INV(A, B, C) launch_inverse_chunks<A, B, C>(Y, nrows, row0, chunks, nchunk, Yp, ph, good, stream)
    switch (nfft) {
    case 8192:  return DD_INV(16, 16, 16);
    case 16384: return DD_INV(16, 16, 32);
    case 32768: return DD_INV(16, 32, 32);
    case 65536: return DD_INV(32, 32, 32);
#undef DD_INV
    default: return hipErrorInvalidValue;
    }
}

extern "C" hipError_t dd_fdd_inverse(const float2* Y, size_t nrows, size_t nchunk, size_t Yp,
                                     const float2* ph, size_t nfft, float* out, size_t out_stride,
                                     size_t good, size_t nsamps_out, hipStream_t stream)
{
    switch (nfft) {
    case 8192:  return launch_inverse<16, 16, 16>(Y, nrows, nchunk, Yp, ph, out, out_stride, good, nsamps_out, stream);
    case 16384: return launch_inverse<16, 16, 32>(Y, nrows, nchunk, Yp, ph, out, out_stride, good, nsamps_out, stream);
    case 32768: return launch_inverse<16, 32, 32>(Y, nrows, nchunk, Yp, ph, out, out_stride, good, nsamps_out, stream);
    case 65536: return launch_inverse<32, 32, 32>(Y, nrows, nchunk, Yp, ph, out, out_stride, good, nsamps_out, stream);
    default: return hipErrorInvalidValue;
    }
}
