// FDDGPUPlan: host side of the Fourier-domain dedispersion path.
//
// Reference: src/fdd/FDDGPUPlan.cpp:39-706 (whole-series transform) and
// :725-1271 (time-segmented, overlap-save). The design here:
//  - the 8-bit input is made resident in HBM once and transposed to
//    channel-major rows (the reference re-uploads and re-transforms it for
//    every outer DM batch);
//  - time series are produced by overlap-save over chunks of nfft samples
//    (power of two, nfft > max delay): chunk i starts at sample i*good,
//    good = nfft - max delay rounded down to even, and keeps its first good
//    outputs. Every chunk is processed (the reference's segmented kernel stops
//    at bin nsamp/2+1 of the flattened chunk layout, leaving trailing chunks
//    empty);
//  - per (channel, chunk) forward transform, per-bin phasor contraction
//    Y_k[dm][chunk] = sum_c P_k[dm][c] X_k[c][chunk] with exact integer phase
//    reduction (k*d) mod nfft, per (DM, chunk) inverse transform;
//  - the killmask zeroes channels (the reference FDD ignores it).
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <vector>

#include "fdd_kernels.h"
#include "internal.hpp"
#include "kernels.h"

namespace dedisp {

struct FDDState {
    dedisp_amd_fdd_output output = DEDISP_AMD_FDD_FOURIER;
    dedisp_amd_fdd_precision precision = DEDISP_AMD_FDD_FP32;
    int segmented = 0;
    dedisp_size nfft_req = 0;

    DeviceBuffer d_in;      // staged 8-bit input (host-pointer mode)
    DeviceBuffer d_rows;    // channel-major 8-bit rows
    DeviceBuffer d_Z;       // [nchan padded to 16][nchunk][Xp] spectra, bins 0..nfft/2
    DeviceBuffer d_Y;       // [dm batch][nchunk][Yp] dedispersed spectra
    DeviceBuffer d_out;     // staged output (host-pointer mode)
    DeviceBuffer d_chunks;  // [chunks of a batch] dd_fdd_chunk (host-pointer batches)

    // Captured device-pointer batches, the least recently used replaced first;
    // each keeps its own chunk table.
    struct Graph {
        std::vector<uint64_t> key;
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        DeviceBuffer d_chunks;
        uint64_t used = 0;
        void drop()
        {
            if (exec) (void)hipGraphExecDestroy(exec);
            if (graph) (void)hipGraphDestroy(graph);
            exec = nullptr;
            graph = nullptr;
            key.clear();
        }
        ~Graph() { drop(); }
    };
    static constexpr size_t kGraphs = 4;
    std::vector<std::unique_ptr<Graph>> graphs;
    uint64_t graph_clock = 0;

    // end of each DM batch (host-pointer executes)
    std::vector<hipEvent_t> blk_events;
    ~FDDState()
    {
        for (hipEvent_t e : blk_events) (void)hipEventDestroy(e);
    }
    DeviceBuffer d_delays;  // [ndm][nchan] int32 truncated delays
    DeviceBuffer d_q32;     // [ndm][qpitch32] uint32 phase steps d mod nfft (matrix-core forms)
    DeviceBuffer d_dm32;    // [ndm] DM list (16 x 16 tile kernel)
    DeviceBuffer d_delay32; // [nchan rounded up to 32] delay table, zero padded
    DeviceBuffer d_ph64k;   // [65536] e^{+2 pi i m/65536} (split transforms, nfft > 65536)
    DeviceBuffer d_fwork;   // forward work of the split transforms
    DeviceBuffer d_iwork;   // inverse work of the split transforms
    DeviceBuffer d_scale;   // [nchan] per-channel input scale
    DeviceBuffer d_ph;      // [nfft] e^{+2 pi i m / nfft}

    // whole-series Fourier output: row transforms of length N
    DeviceBuffer d_S;       // [DM batch][N] circular shift-sums, then packed spectra
    DeviceBuffer d_work;    // [DM batch][N/2] complex
    DeviceBuffer d_ph16k;   // [16384] e^{+2 pi i j/16384}
    DeviceBuffer d_twL;     // [N/2] e^{-2 pi i j/(N/2)}
    DeviceBuffer d_twN;     // [N/2] e^{-2 pi i k/N}
    size_t tw_N = 0;

    // contraction choices of the autotuner for precision tuned_prec (-1:
    // none): form 0/1/2 = the 16 x 16 tile kernel's workgroup shape, 3 = the
    // 32 x 32 tile kernel; mirror bins; bin-block layout (0 = rows)
    int tuned_prec = -1, tuned_form = -1, tuned_mirror = -1, tuned_blocked = -1;

    unsigned long version = ~0ul;
    int32_t dmin = 0, dmax = 0;
    size_t ph_nfft = 0;
    unsigned long q32_version = ~0ul;
    size_t q32_nfft = 0, qpitch32 = 0;
    std::vector<int32_t> h_delays;  // host copy of d_delays
    float max_scale = 0.f;
};

namespace {

// Memory the dedispersed spectra of one DM batch may take.
constexpr size_t kYBudget = size_t(24) << 30;
// Memory the channel spectra of one group of beams may take (batched beams).
constexpr size_t kZBudget = size_t(32) << 30;

bool env_flag(const char* name)
{
    const char* s = std::getenv(name);
    return s && std::atoi(s) != 0;
}

// The contraction kernel of a precision: the bfloat16 and float16 forms run
// on 16 x 16 x 32 tiles with re/im along N (DEDISP_FDD_MFMA32=1 selects the
// 32 x 32 tile kernel instead), float32 on 32 x 32 x 2 tiles
// (DEDISP_FDD_FP32_TILE16=1: 16 x 16 x 4 tiles with re/im along N), and
// DEDISP_FDD_VALU=1 the float32 vector-ALU form.
// The autotuner's choices (FDDGPUPlan::autotune) replace these defaults for
// the precision they were timed with; the environment variables override both.
enum class Contraction { VALU, MFMA32, MFMA16 };

// Whether the autotuner's choices apply (timed for the plan's precision).
bool tuned(const FDDState& st) { return st.tuned_prec == (int)st.precision; }

Contraction contraction_kind(const FDDState& st)
{
    const bool t = tuned(st) && st.tuned_form >= 0;
    if (st.precision == DEDISP_AMD_FDD_FP32) {
        if (env_flag("DEDISP_FDD_VALU")) return Contraction::VALU;
        if (env_flag("DEDISP_FDD_FP32_TILE16")) return Contraction::MFMA16;
        return (t && st.tuned_form != 3) ? Contraction::MFMA16 : Contraction::MFMA32;
    }
    if (env_flag("DEDISP_FDD_MFMA32")) return Contraction::MFMA32;
    return (t && st.tuned_form == 3) ? Contraction::MFMA32 : Contraction::MFMA16;
}

// Channel rows of the spectra: a multiple of 32 (the K-step of the 16 x 16
// tile kernel, a multiple of the 32 x 32 kernel's 16), extra rows zero.
size_t padded_channels(size_t nchan) { return round_up(nchan, 32); }

// DMs per batch of the contraction and inverse when at most `most` fit the
// memory budget: as many batches as needed, of equal size rounded up to the
// contraction's 256-DM blocks (2048 DMs with room for 1730: 2 x 1024, not
// 1730 + 318, whose second batch leaves most of its DM blocks idle).
size_t balanced_batch(size_t ndm, size_t most)
{
    most = std::max<size_t>(1, std::min(ndm, most));
    const size_t nbat = div_up(ndm, most);
    const size_t even = div_up(ndm, nbat);
    const size_t blk = round_up(even, 256);
    return (blk <= most) ? blk : even;
}

}  // namespace

FDDGPUPlan::FDDGPUPlan(size_type nchans, float_type dt, float_type f0, float_type df, int device_idx)
    : GPUPlan(nchans, dt, f0, df, device_idx), m_state(new FDDState)
{
    if (nchans > kFddMaxChans) DD_THROW(DEDISP_NCHANS_EXCEEDS_LIMIT);
}

FDDGPUPlan::~FDDGPUPlan() {}

// The CPU FDD algorithm is not part of this GPU library (src/fdd/FDDCPUPlan.cpp
// is restated only as a test oracle): constructing the plan fails loudly.
FDDCPUPlan::FDDCPUPlan(size_type nchans, float_type dt, float_type f0, float_type df, int)
    : Plan(nchans, dt, f0, df)
{
    DD_THROW(DEDISP_UNKNOWN_ERROR);
}

FDDCPUPlan::~FDDCPUPlan() {}

void FDDCPUPlan::execute(size_type, const byte_type*, size_type, byte_type*, size_type, unsigned)
{
    DD_THROW(DEDISP_UNKNOWN_ERROR);
}

void FDDGPUPlan::set_output_mode(dedisp_amd_fdd_output mode)
{
    if (mode != DEDISP_AMD_FDD_FOURIER && mode != DEDISP_AMD_FDD_TIMESERIES) DD_THROW(DEDISP_UNKNOWN_ERROR);
    m_state->output = mode;
}
dedisp_amd_fdd_output FDDGPUPlan::get_output_mode() const { return m_state->output; }

void FDDGPUPlan::set_precision(dedisp_amd_fdd_precision p)
{
    if (p != DEDISP_AMD_FDD_FP32 && p != DEDISP_AMD_FDD_BF16 && p != DEDISP_AMD_FDD_FP16)
        DD_THROW(DEDISP_UNKNOWN_ERROR);
    m_state->precision = p;
}
dedisp_amd_fdd_precision FDDGPUPlan::get_precision() const { return m_state->precision; }

void FDDGPUPlan::set_segmented(int enable, dedisp_size nfft)
{
    if (nfft && !dd_fdd_fft_supported(nfft)) DD_THROW(DEDISP_UNKNOWN_ERROR);
    m_state->segmented = enable;
    m_state->nfft_req = nfft;
}

double FDDGPUPlan::autotune(size_type nsamps, size_type in_nbits, size_type out_nbits,
                            dedisp_amd_tuning* chosen)
{
    if (m_dm_count == 0) DD_THROW(DEDISP_NO_DM_LIST_SET);
    if (nsamps <= m_max_delay) DD_THROW(DEDISP_TOO_FEW_NSAMPS);
    if (in_nbits != 1 && in_nbits != 2 && in_nbits != 4 && in_nbits != 8) DD_THROW(DEDISP_UNSUPPORTED_IN_NBITS);
    if (out_nbits != 32) DD_THROW(DEDISP_UNSUPPORTED_OUT_NBITS);
    if (m_state->output != DEDISP_AMD_FDD_TIMESERIES) DD_THROW(DEDISP_INVALID_PLAN);
    GPUContext& ctx = *m_ctx;
    ctx.make_current();
    const size_t row_bytes = m_nchans * in_nbits / 8, nsc = nsamps - m_max_delay;
    DeviceBuffer d_in, d_out;
    d_in.reserve(nsamps * row_bytes);
    d_out.reserve(m_dm_count * nsc * 4);
    DD_HIP(hipMemsetAsync(d_in.ptr, 0, nsamps * row_bytes, ctx.exec));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    DD_HIP(hipEventCreate(&e0));
    DD_HIP(hipEventCreate(&e1));
    const int saved_seg = m_state->segmented;
    const dedisp_size saved_nfft = m_state->nfft_req;
    double best = -1;
    dedisp_size best_n = 0;
    unsigned timed = 0;
    try {
        for (dedisp_size n = 8192; n <= 262144; n *= 2) {
            if (n <= m_max_delay || !dd_fdd_fft_supported(n)) continue;
            set_segmented(1, n);
            execute(nsamps, d_in.as<uint8_t>(), in_nbits, d_out.as<uint8_t>(), 32, DEDISP_DEVICE_POINTERS);
            DD_HIP(hipEventRecord(e0, ctx.exec));
            execute(nsamps, d_in.as<uint8_t>(), in_nbits, d_out.as<uint8_t>(), 32, DEDISP_DEVICE_POINTERS);
            DD_HIP(hipEventRecord(e1, ctx.exec));
            DD_HIP(hipEventSynchronize(e1));
            float ms = 0.f;
            DD_HIP(hipEventElapsedTime(&ms, e0, e1));
            ++timed;
            if (best < 0 || ms < best) best = ms, best_n = n;
        }
    } catch (...) {
        m_state->segmented = saved_seg;
        m_state->nfft_req = saved_nfft;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        throw;
    }
    if (best >= 0) {
        set_segmented(1, best_n);
    } else {
        m_state->segmented = saved_seg;
        m_state->nfft_req = saved_nfft;
    }
    // the contraction at that segment length: workgroup form (0, 1, 2 = the
    // 16 x 16 tile kernel's 8-wave, 4-wave and pipelined shapes, 3 = the
    // 32 x 32 tile kernel), mirror bins, bin-block layout (-1: not used)
    FDDState& st = *m_state;
    struct Cand {
        int form, mirror, blocked;
    };
    std::vector<Cand> cands;
    if (st.precision == DEDISP_AMD_FDD_FP32)
        cands = {{3, -1, -1}, {0, 0, -1}, {0, 1, -1}};
    else
        cands = {{2, 1, 1}, {2, 1, 0}, {2, 0, -1}, {0, 1, -1}, {0, 0, -1}, {3, -1, -1}};
    const int saved[4] = {st.tuned_prec, st.tuned_form, st.tuned_mirror, st.tuned_blocked};
    Cand pick = {-1, -1, -1};
    double pick_ms = -1;
    try {
        if (best >= 0) {
            for (const Cand& c : cands) {
                st.tuned_prec = (int)st.precision;
                st.tuned_form = c.form;
                st.tuned_mirror = c.mirror;
                st.tuned_blocked = c.blocked;
                execute(nsamps, d_in.as<uint8_t>(), in_nbits, d_out.as<uint8_t>(), 32, DEDISP_DEVICE_POINTERS);
                DD_HIP(hipEventRecord(e0, ctx.exec));
                execute(nsamps, d_in.as<uint8_t>(), in_nbits, d_out.as<uint8_t>(), 32, DEDISP_DEVICE_POINTERS);
                DD_HIP(hipEventRecord(e1, ctx.exec));
                DD_HIP(hipEventSynchronize(e1));
                float ms = 0.f;
                DD_HIP(hipEventElapsedTime(&ms, e0, e1));
                ++timed;
                if (pick_ms < 0 || ms < pick_ms) pick_ms = ms, pick = c;
            }
        }
    } catch (...) {
        st.tuned_prec = saved[0], st.tuned_form = saved[1], st.tuned_mirror = saved[2], st.tuned_blocked = saved[3];
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (pick_ms >= 0) {
        st.tuned_prec = (int)st.precision;
        st.tuned_form = pick.form;
        st.tuned_mirror = pick.mirror;
        st.tuned_blocked = pick.blocked;
        best = pick_ms;
    } else {
        st.tuned_prec = saved[0], st.tuned_form = saved[1], st.tuned_mirror = saved[2], st.tuned_blocked = saved[3];
    }
    if (chosen) {
        *chosen = dedisp_amd_tuning{};
        chosen->nfft = best_n;
        chosen->best_ms = best;
        chosen->candidates = timed;
        chosen->fdd_form = pick.form;
        chosen->fdd_mirror = pick.mirror;
        chosen->fdd_layout = pick.blocked;
    }
    return best;
}

dedisp_size FDDGPUPlan::output_length(dedisp_size nsamps) const
{
    const bool segmented = m_state->segmented || env_flag("USE_SEGMENTED");
    if (m_state->output == DEDISP_AMD_FDD_FOURIER && !segmented)
        return round_up(nsamps + m_max_delay, 16384);  // src/fdd/FDDGPUPlan.cpp:91
    return nsamps - m_max_delay;
}

// Per-plan tables: truncated delays, channel scales (rebuilt when the DM list,
// killmask or channels change).
void FDDGPUPlan::prepare_tables()
{
    FDDState& st = *m_state;
    hipStream_t stream = m_ctx->exec;
    const size_t nchan = m_nchans, ndm = m_dm_count;
    if (st.version != m_tables_version) {
        std::vector<int32_t> d(ndm * nchan);
        int32_t dmin = 0, dmax = 0;
        for (size_t i = 0; i < ndm; ++i)
            for (size_t c = 0; c < nchan; ++c) {
                int32_t v = fdd_delay(h_dm_list[i], h_delay_table[c]);
                d[i * nchan + c] = v;
                dmin = std::min(dmin, v);
                dmax = std::max(dmax, v);
            }
        // x = (u8 - 127.5) * (float)(1.0/nchan) (src/fdd/FDDGPUPlan.cpp:465-473);
        // killed channels contribute nothing
        std::vector<float> scale(nchan);
        for (size_t c = 0; c < nchan; ++c) scale[c] = h_killmask[c] ? (float)(1.0 / nchan) : 0.f;
        std::vector<float> dl32(round_up(nchan, 32), 0.f);
        std::copy(h_delay_table.begin(), h_delay_table.end(), dl32.begin());
        st.d_dm32.reserve(std::max<size_t>(ndm, 1) * 4);
        st.d_delay32.reserve(dl32.size() * 4);
        DD_HIP(hipMemcpyAsync(st.d_dm32.ptr, h_dm_list.data(), ndm * 4, hipMemcpyHostToDevice, stream));
        DD_HIP(hipMemcpyAsync(st.d_delay32.ptr, dl32.data(), dl32.size() * 4, hipMemcpyHostToDevice, stream));
        st.d_delays.reserve(d.size() * 4);
        st.d_scale.reserve(nchan * 4);
        DD_HIP(hipMemcpyAsync(st.d_delays.ptr, d.data(), d.size() * 4, hipMemcpyHostToDevice, stream));
        DD_HIP(hipMemcpyAsync(st.d_scale.ptr, scale.data(), nchan * 4, hipMemcpyHostToDevice, stream));
        DD_HIP(hipStreamSynchronize(stream));
        st.dmin = dmin;
        st.dmax = dmax;
        st.h_delays.swap(d);
        st.max_scale = nchan ? *std::max_element(scale.begin(), scale.end()) : 0.f;
        st.version = m_tables_version;
    }
    if (st.dmin < 0) DD_THROW(DEDISP_UNKNOWN_ERROR);  // negative DMs: not supported
}

// Overlap-save segment length for series with nsc[0..n) outputs per DM: the
// requested one, else the supported length (2^13 .. 2^18) with the smallest
// contraction volume (chunk columns of the contraction kernel's tiles, padding
// included, times nfft/2 + 1). The reference instead
// doubles nfft from 16384 until a fifth of it covers the maximum delay
// (src/fdd/FDDGPUPlan.cpp:756-762); the smallest volume needs fewer complex
// multiply-accumulates for the same outputs.
size_t FDDGPUPlan::segment_length(const size_t* nsc, size_t n) const
{
    const FDDState& st = *m_state;
    const size_t dmax = (size_t)st.dmax;
    size_t nfft = st.nfft_req;
    if (!nfft) {
        // smallest contraction volume (chunk columns the kernel computes,
        // padding included) x (nfft/2 + 1) among the supported lengths
        const Contraction kind = contraction_kind(st);
        size_t best = ~size_t(0);
        for (size_t cand = 8192; cand <= 262144; cand *= 2) {
            if (cand < dmax + 2 + 2) continue;
            const size_t good = (cand - dmax) & ~size_t(1);
            size_t chunks = 0;
            for (size_t i = 0; i < n; ++i) chunks += div_up(nsc[i], good);
            size_t cols = chunks;
            if (kind == Contraction::MFMA32) {
                cols = round_up(chunks, 32);
            } else if (kind == Contraction::MFMA16) {
                int ntl = 1;
                size_t groups = 0;
                dd_fdd_mfma16_tiles(chunks, (int)st.precision, &ntl, &groups);
                cols = groups * (size_t)ntl * 8;
            }
            const size_t cost = cols * (cand / 2 + 1);
            if (cost < best) { best = cost; nfft = cand; }
        }
        if (!nfft) DD_THROW(DEDISP_UNKNOWN_ERROR);  // max delay beyond the largest segment
    }
    if (nfft < dmax + 4) DD_THROW(DEDISP_UNKNOWN_ERROR);
    return nfft;
}

// Phase tables of segment length nfft: e^{2 pi i m/nfft} for the transforms
// (and the 65536-entry table of the on-chip engine when segments are split)
// and, for the matrix-core contraction (valu false), the phase steps
// d mod nfft. Returns the float16 spectrum scale.
float FDDGPUPlan::prepare_phasors(size_t nfft, bool valu)
{
    FDDState& st = *m_state;
    hipStream_t stream = m_ctx->exec;
    const size_t nchan = m_nchans, ndm = m_dm_count;
    if (st.ph_nfft != nfft) {
        std::vector<float> ph(2 * nfft);
        for (size_t m = 0; m < nfft; ++m) {
            const double a = 2.0 * M_PI * (double)m / (double)nfft;
            ph[2 * m] = (float)std::cos(a);
            ph[2 * m + 1] = (float)std::sin(a);
        }
        st.d_ph.reserve(nfft * 8);
        DD_HIP(hipMemcpyAsync(st.d_ph.ptr, ph.data(), nfft * 8, hipMemcpyHostToDevice, stream));
        DD_HIP(hipStreamSynchronize(stream));
        st.ph_nfft = nfft;
    }

    // the on-chip transform engine's table when segments are split (nfft > 65536)
    if (nfft > 65536 && !st.d_ph64k.ptr) {
        std::vector<float> ph(2 * 65536);
        for (size_t m = 0; m < 65536; ++m) {
            const double a = 2.0 * M_PI * (double)m / 65536.0;
            ph[2 * m] = (float)std::cos(a);
            ph[2 * m + 1] = (float)std::sin(a);
        }
        st.d_ph64k.reserve(65536 * 8);
        DD_HIP(hipMemcpyAsync(st.d_ph64k.ptr, ph.data(), 65536 * 8, hipMemcpyHostToDevice, stream));
        DD_HIP(hipStreamSynchronize(stream));
    }
    // phase steps of the matrix-core contraction: q = d mod nfft (phases in
    // units of 1/nfft)
    if (!valu && (st.q32_version != m_tables_version || st.q32_nfft != nfft)) {
        const size_t qpitch = round_up(nchan, 32);
        std::vector<uint32_t> q(ndm * qpitch, 0);
        for (size_t i = 0; i < ndm; ++i)
            for (size_t c = 0; c < nchan; ++c)
                q[i * qpitch + c] = (uint32_t)st.h_delays[i * nchan + c] & (uint32_t)(nfft - 1);
        st.d_q32.reserve(q.size() * 4);
        DD_HIP(hipMemcpyAsync(st.d_q32.ptr, q.data(), q.size() * 4, hipMemcpyHostToDevice, stream));
        DD_HIP(hipStreamSynchronize(stream));
        st.qpitch32 = qpitch;
        st.q32_nfft = nfft;
        st.q32_version = m_tables_version;
    }
    // float16 spectra: |X| <= nfft * 127.5 * max scale; a power of two brings
    // that bound to at most 2^15 (float16 max 65504)
    float xscale = 1.f;
    if (st.precision == DEDISP_AMD_FDD_FP16 && st.max_scale > 0.f) {
        const double bound = (double)nfft * 127.5 * (double)st.max_scale;
        xscale = (float)std::ldexp(1.0, (int)std::floor(std::log2(32768.0 / bound)));
    }
    return xscale;
}

namespace {

// Whether the contraction writes the inverse transform's packed spectrum
// (16 x 16 tile kernel with mirror bins; DEDISP_FDD_PACKED=0 keeps bins).
bool env_off(const char* name)
{
    const char* s = std::getenv(name);
    return s && std::atoi(s) == 0;
}

// DEDISP_FDD_M16CFG as set (0, 1 or 2), -1 when unset.
int m16cfg_env()
{
    const char* s = std::getenv("DEDISP_FDD_M16CFG");
    if (!s) return -1;
    const int v = std::atoi(s);
    return (v >= 0 && v <= 2) ? v : -1;
}

// Workgroup shape of the 16 x 16 tile kernel: 2, the software-pipelined form
// (8 waves of 2 DM tiles, 4 bin slots), for bfloat16 and float16 with the
// delays computed in the kernel while a K-step's spectra stay within 32-bit
// element offsets; else 0 (8 waves of 2 DM tiles, 4 slots). DEDISP_FDD_M16CFG
// forces 0, 1 (4 waves of 4 DM tiles, 2 slots) or 2 where it applies.
int mfma16_cfg(const FDDState& st, size_t nchunk, size_t Xp)
{
    const bool pipe = st.precision != DEDISP_AMD_FDD_FP32 && !env_flag("DEDISP_FDD_QTABLE") &&
                      32 * nchunk * Xp + Xp < (size_t(1) << 32);
    const int e = m16cfg_env();
    if (e == 0 || e == 1 || (e == 2 && pipe)) return e;
    if (tuned(st) && (st.tuned_form == 0 || st.tuned_form == 1 || (st.tuned_form == 2 && pipe)))
        return st.tuned_form;
    return pipe ? 2 : 0;
}

// Whether the 16 x 16 tile kernel pairs each bin with its mirror
// (DEDISP_FDD_MIRROR=0/1 forces it): by default for bfloat16 and float16,
// whose mirror phasors save a cosine, a sine (and the hi/lo split) for half
// the bins and which then run in the bin-block layout (1024 ch x 2048 DMs x
// 23 chunks of 65536: float16 21.2 ms against 28.7 ms on four consecutive
// bins); not for float32 (117.1 ms on consecutive bins against 119.5 ms).
bool mfma16_mirror(const FDDState& st)
{
    const char* s = std::getenv("DEDISP_FDD_MIRROR");
    if (s) return std::atoi(s) != 0;
    if (tuned(st) && st.tuned_mirror >= 0) return st.tuned_mirror != 0;
    return st.precision != DEDISP_AMD_FDD_FP32;
}

bool contraction_packs(const FDDState& st)
{
    return contraction_kind(st) == Contraction::MFMA16 && mfma16_mirror(st) &&
           !env_off("DEDISP_FDD_PACKED");
}

// Whether the spectra and the packed output use the bin-block layout (the
// pipelined float16 / bfloat16 contraction with mirror bins and the packed
// output, nfft <= 65536): a workgroup of the contraction then reads and
// writes whole cache lines (DEDISP_FDD_BLOCKED=0 keeps rows of Xp / Yp).
bool blocked_layout(const FDDState& st, size_t nfft, size_t nchunk, size_t Xp)
{
    return contraction_packs(st) && mfma16_cfg(st, nchunk, Xp) == 2 && nfft <= 65536 &&
           !env_off("DEDISP_FDD_BLOCKED") && !(tuned(st) && st.tuned_blocked == 0);
}

// Elements of one channel's spectra: rows of Xp per chunk, or the bin blocks.
size_t chan_elems(bool blocked, size_t nfft, size_t nchunk, size_t Xp)
{
    return blocked ? (nfft / 8 + 1) * 4 * nchunk : nchunk * Xp;
}

// The inverse transforms' input form: 0 bins, 1 packed rows, 2 packed bin blocks.
int inverse_form(const FDDState& st, bool blocked) { return blocked ? 2 : (contraction_packs(st) ? 1 : 0); }

// Contraction of DMs b0 .. b0+nb of the plan's tables over the spectra in
// st.d_Z ([padded channels][nchunk][Xp]) into st.d_Y ([nb][nchunk][Yp]):
// dedispersed bins 0..M, or the packed spectrum when contraction_packs.
void contract(FDDState& st, size_t nchan, size_t b0, size_t nb, size_t nchunk, size_t nfft, size_t Xp, size_t Yp,
              float xscale, bool blocked, hipStream_t stream)
{
    switch (contraction_kind(st)) {
    case Contraction::VALU:
        DD_HIP(dd_fdd_contract_fp32(st.d_Z.as<float2>(), Xp, st.d_delays.as<int32_t>() + b0 * nchan, nb, nchan,
                                    nchunk, st.d_ph.as<float2>(), nfft, st.d_Y.as<float2>(), Yp, stream));
        break;
    case Contraction::MFMA32:
        DD_HIP(dd_fdd_contract_mfma((int)st.precision, st.d_Z.as<float2>(), Xp,
                                    st.d_q32.as<uint32_t>() + b0 * st.qpitch32, st.qpitch32, nb, nchan, nchunk,
                                    nfft, xscale, st.d_Y.as<float2>(), Yp, stream));
        break;
    case Contraction::MFMA16:
        // DEDISP_FDD_MIRROR=0: four consecutive bins per workgroup instead of
        // two bins and their mirrors
        // delays from the DM list and delay table in the kernel
        // (DEDISP_FDD_QTABLE=1: read from the phase-step table instead)
        DD_HIP(dd_fdd_contract_mfma16((int)st.precision, mfma16_mirror(st), st.d_Z.as<float2>(), Xp,
                                      st.d_q32.as<uint32_t>() + b0 * st.qpitch32, st.qpitch32, nb, nchan, nchunk,
                                      nfft, xscale, st.d_Y.as<float2>(), Yp,
                                      contraction_packs(st) ? st.d_ph.as<float2>() : nullptr,
                                      env_flag("DEDISP_FDD_QTABLE") ? nullptr : st.d_dm32.as<float>() + b0,
                                      env_flag("DEDISP_FDD_QTABLE") ? nullptr : st.d_delay32.as<float>(),
                                      mfma16_cfg(st, nchunk, Xp), blocked ? 1 : 0, stream));
        break;
    }
}

}  // namespace

void FDDGPUPlan::execute(size_type nsamps, const byte_type* in, size_type in_nbits,
                         byte_type* out, size_type out_nbits, unsigned flags)
{
    if (!in || !out) DD_THROW(DEDISP_INVALID_POINTER);
    if (m_dm_count == 0) DD_THROW(DEDISP_NO_DM_LIST_SET);
    if (nsamps < m_max_delay) DD_THROW(DEDISP_TOO_FEW_NSAMPS);
    if ((flags & DEDISP_ASYNC) && (flags & DEDISP_WAIT)) DD_THROW(DEDISP_INVALID_FLAG_COMBINATION);
    if ((flags & DEDISP_HOST_POINTERS) && (flags & DEDISP_DEVICE_POINTERS))
        DD_THROW(DEDISP_INVALID_FLAG_COMBINATION);
    // 8-bit input, float output (src/fdd/FDDGPUPlan.cpp:66-67); 1, 2 and 4-bit
    // input are an extension (same centring, (u - (2^nbits-1)/2) / nchan)
    if (in_nbits != 1 && in_nbits != 2 && in_nbits != 4 && in_nbits != 8) DD_THROW(DEDISP_UNSUPPORTED_IN_NBITS);
    if ((m_nchans * in_nbits) % 8) DD_THROW(DEDISP_UNSUPPORTED_IN_NBITS);
    if (out_nbits != 32) DD_THROW(DEDISP_UNSUPPORTED_OUT_NBITS);
    if (nsamps >= (size_t(1) << 31)) DD_THROW(DEDISP_UNKNOWN_ERROR);

    const bool segmented = m_state->segmented || env_flag("USE_SEGMENTED");
    const bool fourier = (m_state->output == DEDISP_AMD_FDD_FOURIER) && !segmented;
    execute_fdd(nsamps, in, in_nbits, out, flags, fourier);
}

// Time-series output: overlap-save over the input, outputs t < nsamps - max_delay.
// Fourier output (src/fdd/FDDGPUPlan.cpp:85-93, 394-407, 618-650): with
// N = round_up(nsamps + max_delay, 16384), sum_c X_c[k] e^{2 pi i k d_c/N} is the
// N-point transform of the circular shift-sum s[t] = sum_c x_c[(t + d_c) mod N]
// (x zero beyond nsamps). s is produced by the same overlap-save path over the
// N-periodic extension of the input (t = 0 .. N-1), then transformed per DM;
// bins 0..N/2-1 are written as N floats (re, im).
void FDDGPUPlan::execute_fdd(size_type nsamps, const byte_type* in, size_type in_nbits, byte_type* out,
                             unsigned flags, bool fourier)
{
    GPUContext& ctx = *m_ctx;
    FDDState& st = *m_state;
    ctx.make_current();
    hipStream_t stream = ctx.exec;
    ctx.timer.reset(ctx.profiling);
    ctx.profile = dedisp_amd_profile{};
    const bool device_ptrs = (flags & DEDISP_DEVICE_POINTERS) != 0;
    const size_t nchan = m_nchans, ndm = m_dm_count;
    const size_t in_row = nchan * in_nbits / 8;  // bytes per input sample
    const size_t Nf = round_up(nsamps + m_max_delay, 16384);   // whole-series length
    const size_t nsc = fourier ? Nf : nsamps - m_max_delay;    // outputs per DM (t)
    const size_t out_floats = fourier ? Nf : nsc;              // floats per output row
    if (nsc == 0) return;
    if (fourier && Nf >= (size_t(1) << 31)) DD_THROW(DEDISP_UNKNOWN_ERROR);

    prepare_tables();
    const size_t dmax = (size_t)st.dmax;
    const size_t nfft = segment_length(&nsc, 1);
    const size_t good = (nfft - dmax) & ~size_t(1);
    const size_t nchunk = div_up(nsc, good);
    const size_t M = nfft / 2;
    const size_t Xp = M + 64, Yp = M + 64;  // bins 0..M, rows 512-byte aligned

    const bool valu = env_flag("DEDISP_FDD_VALU") && st.precision == DEDISP_AMD_FDD_FP32;
    const float xscale = prepare_phasors(nfft, valu);
    const float2* ph = st.d_ph.as<float2>();

    // ---- input ----
    const uint8_t* d_in;
    if (device_ptrs) {
        d_in = in;
    } else {
        st.d_in.reserve(nsamps * in_row + 64);
        ctx.timer.begin(StageTimer::H2D, stream);
        ctx.io.upload(stream, st.d_in.as<uint8_t>(), in_row, in, in_row, in_row, nsamps);
        ctx.timer.end(StageTimer::H2D, stream);
        d_in = st.d_in.as<uint8_t>();
    }

    if (fourier && st.tw_N != Nf) {
        const size_t L = Nf / 2;
        std::vector<float> a(2 * L), b(2 * L), c(2 * 16384);
        for (size_t j = 0; j < L; ++j) {
            const double t1 = -2.0 * M_PI * (double)j / (double)L;
            const double t2 = -2.0 * M_PI * (double)j / (double)Nf;
            a[2 * j] = (float)std::cos(t1);
            a[2 * j + 1] = (float)std::sin(t1);
            b[2 * j] = (float)std::cos(t2);
            b[2 * j + 1] = (float)std::sin(t2);
        }
        for (size_t j = 0; j < 16384; ++j) {
            const double t = 2.0 * M_PI * (double)j / 16384.0;
            c[2 * j] = (float)std::cos(t);
            c[2 * j + 1] = (float)std::sin(t);
        }
        st.d_twL.reserve(L * 8);
        st.d_twN.reserve(L * 8);
        st.d_ph16k.reserve(16384 * 8);
        DD_HIP(hipMemcpyAsync(st.d_twL.ptr, a.data(), L * 8, hipMemcpyHostToDevice, stream));
        DD_HIP(hipMemcpyAsync(st.d_twN.ptr, b.data(), L * 8, hipMemcpyHostToDevice, stream));
        DD_HIP(hipMemcpyAsync(st.d_ph16k.ptr, c.data(), 16384 * 8, hipMemcpyHostToDevice, stream));
        DD_HIP(hipStreamSynchronize(stream));
        st.tw_N = Nf;
    }

    // ---- transpose to channel-major rows, forward transforms ----
    const size_t in_span = fourier ? nsamps : std::max(nsamps, (nchunk - 1) * good + nfft);
    const size_t row_words = div_up(in_span, (32 / in_nbits) * 64) * 64;
    st.d_rows.reserve(nchan * row_words * 4);
    // channel rows padded to a multiple of 32 with zeros (matrix-core staging)
    const size_t nchan_pad = padded_channels(nchan);
    const bool blocked = blocked_layout(st, nfft, nchunk, Xp);
    const size_t zrow = chan_elems(blocked, nfft, nchunk, Xp);
    st.d_Z.reserve(nchan_pad * nchunk * Xp * 8);
    if (nchan_pad > nchan)
        DD_HIP(hipMemsetAsync(st.d_Z.as<float2>() + nchan * zrow, 0, (nchan_pad - nchan) * zrow * 8, stream));
    ctx.timer.begin(StageTimer::UNPACK, stream);
    DD_HIP(dd_tdd_transpose(d_in, nsamps, in_row, (unsigned)in_nbits, nchan, st.d_rows.as<uint32_t>(),
                            row_words, stream));
    ctx.timer.end(StageTimer::UNPACK, stream);
    st.d_fwork.reserve(std::max<size_t>(dd_fdd_split_work(nfft, nchan * nchunk), 16));
    ctx.timer.begin(StageTimer::FFT, stream);
    DD_HIP(dd_fdd_forward_split(st.d_rows.as<uint8_t>(), row_words * 4, good, nsamps, fourier ? Nf : 0,
                                st.d_scale.as<float>(), ph, st.d_ph64k.as<float2>(), nfft, st.d_Z.as<float2>(),
                                nchan, nchunk, Xp, (unsigned)in_nbits, st.d_fwork.as<float2>(), blocked ? 1 : 0,
                                stream));
    ctx.timer.end(StageTimer::FFT, stream);

    // ---- per DM batch: contraction, inverse transforms ----
    const size_t per_dm = nchunk * Yp * 8 + (fourier ? 2 * Nf * 4 : 0) + dd_fdd_split_work(nfft, nchunk);
    size_t dm_batch = balanced_batch(ndm, kYBudget / per_dm);
    if (!device_ptrs) {
        // host outputs: DM batches of about 1 GB of output (whole 256-DM tiles
        // of the contraction), so the download overlaps the later batches
        const size_t want = round_up(std::max<size_t>(1, (size_t(1) << 30) / (out_floats * 4)), 256);
        dm_batch = std::min(dm_batch, want);
        const size_t nbat = div_up(ndm, dm_batch);
        while (st.blk_events.size() < nbat) {
            hipEvent_t e = nullptr;
            DD_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
            st.blk_events.push_back(e);
        }
    }
    st.d_Y.reserve(dm_batch * nchunk * Yp * 8);
    st.d_iwork.reserve(std::max<size_t>(dd_fdd_split_work(nfft, dm_batch * nchunk), 16));
    if (fourier) {
        st.d_S.reserve(dm_batch * Nf * 4);
        st.d_work.reserve(dm_batch * Nf * 4);
    }
    float* d_out;
    if (device_ptrs) {
        d_out = reinterpret_cast<float*>(out);
    } else {
        st.d_out.reserve(ndm * out_floats * 4);
        d_out = st.d_out.as<float>();
    }
    // DEDISP_FDD_VALU=1 selects the float32 vector-ALU contraction instead of
    // the matrix cores (same arithmetic as the float32 MFMA form)
    for (size_t b0 = 0; b0 < ndm; b0 += dm_batch) {
        const size_t nb = std::min(dm_batch, ndm - b0);
        ctx.timer.begin(StageTimer::DEDISPERSE, stream);
        contract(st, nchan, b0, nb, nchunk, nfft, Xp, Yp, xscale, blocked, stream);
        ctx.timer.end(StageTimer::DEDISPERSE, stream);
        ctx.timer.begin(StageTimer::IFFT, stream);
        if (!fourier) {
            DD_HIP(dd_fdd_inverse_split(st.d_Y.as<float2>(), nb, nchunk, Yp, ph, st.d_ph64k.as<float2>(), nfft,
                                        d_out + b0 * nsc, nsc, good, nsc, st.d_iwork.as<float2>(),
                                        inverse_form(st, blocked), stream));
        } else {
            DD_HIP(dd_fdd_inverse_split(st.d_Y.as<float2>(), nb, nchunk, Yp, ph, st.d_ph64k.as<float2>(), nfft,
                                        st.d_S.as<float>(), Nf, good, Nf, st.d_iwork.as<float2>(),
                                        inverse_form(st, blocked), stream));
            DD_HIP(dd_fdd_rfft_rows(st.d_S.as<float>(), Nf, nb, Nf, st.d_ph16k.as<float2>(),
                                    st.d_twL.as<float2>(), st.d_twN.as<float2>(),
                                    st.d_work.as<float2>(),
                                    reinterpret_cast<float2*>(d_out + b0 * out_floats), Nf / 2, stream));
        }
        ctx.timer.end(StageTimer::IFFT, stream);
        if (!device_ptrs) DD_HIP(hipEventRecord(st.blk_events[b0 / dm_batch], stream));
    }

    if (!device_ptrs) {
        hipStream_t d2h = ctx.d2h;
        size_t waited = 0;  // DM batches the copy stream already waits for
        ctx.timer.begin(StageTimer::D2H, d2h);
        ctx.io.download(d2h, out, out_floats * 4, reinterpret_cast<const uint8_t*>(d_out), out_floats * 4,
                        out_floats * 4, ndm, [&](size_t rows) {
                            const size_t need = div_up(rows, dm_batch);
                            for (; waited < need; ++waited)
                                DD_HIP(hipStreamWaitEvent(d2h, st.blk_events[waited], 0));
                        });
        ctx.timer.end(StageTimer::D2H, d2h);
    }
    if (!(device_ptrs && (flags & DEDISP_ASYNC))) {
        DD_HIP(hipStreamSynchronize(stream));
        ctx.timer.collect(ctx.profile);
        ctx.profile.work_units = (double)ndm * (double)nchan * (double)nchunk * (double)(M + 1);
    }
}

// Batched beams, time-series output: every beam's overlap-save chunks (one
// segment length for the batch) are concatenated along the chunk axis, so the
// forward transforms, the contraction over the DM list and the inverse
// transforms each run once per group of beams (a group's channel spectra
// within kZBudget), with per-chunk tables giving each chunk its beam's rows
// and output. A (DM, chunk) output sums the same products in the same order
// as in a single-beam execute, so each beam's output equals execute() on it
// alone with the same segment length.
void FDDGPUPlan::execute_batch(size_type nbeams, const size_type* nsamps, const byte_type* const* in,
                               size_type in_nbits, byte_type* const* out, size_type out_nbits,
                               unsigned flags)
{
    if (nbeams == 0) return;
    if (!nsamps || !in || !out) DD_THROW(DEDISP_INVALID_POINTER);
    if (m_dm_count == 0) DD_THROW(DEDISP_NO_DM_LIST_SET);
    if ((flags & DEDISP_ASYNC) && (flags & DEDISP_WAIT)) DD_THROW(DEDISP_INVALID_FLAG_COMBINATION);
    if ((flags & DEDISP_HOST_POINTERS) && (flags & DEDISP_DEVICE_POINTERS))
        DD_THROW(DEDISP_INVALID_FLAG_COMBINATION);
    if (in_nbits != 1 && in_nbits != 2 && in_nbits != 4 && in_nbits != 8) DD_THROW(DEDISP_UNSUPPORTED_IN_NBITS);
    if ((m_nchans * in_nbits) % 8) DD_THROW(DEDISP_UNSUPPORTED_IN_NBITS);
    if (out_nbits != 32) DD_THROW(DEDISP_UNSUPPORTED_OUT_NBITS);
    for (size_t b = 0; b < nbeams; ++b) {
        if (!in[b] || !out[b]) DD_THROW(DEDISP_INVALID_POINTER);
        if (nsamps[b] < m_max_delay) DD_THROW(DEDISP_TOO_FEW_NSAMPS);
        if (nsamps[b] >= (size_t(1) << 31)) DD_THROW(DEDISP_UNKNOWN_ERROR);
    }
    const bool segmented = m_state->segmented || env_flag("USE_SEGMENTED");
    if (m_state->output == DEDISP_AMD_FDD_FOURIER && !segmented) {
        // whole-series spectra: one beam after another
        for (size_t b = 0; b < nbeams; ++b) execute_fdd(nsamps[b], in[b], in_nbits, out[b], flags, true);
        return;
    }

    GPUContext& ctx = *m_ctx;
    FDDState& st = *m_state;
    ctx.make_current();
    hipStream_t stream = ctx.exec;
    ctx.timer.reset(ctx.profiling);
    ctx.profile = dedisp_amd_profile{};
    const bool device_ptrs = (flags & DEDISP_DEVICE_POINTERS) != 0;
    const size_t nchan = m_nchans, ndm = m_dm_count;
    const size_t in_row = nchan * in_nbits / 8;

    prepare_tables();
    const size_t dmax = (size_t)st.dmax;
    std::vector<size_t> nsc(nbeams);
    for (size_t b = 0; b < nbeams; ++b) nsc[b] = nsamps[b] - m_max_delay;
    const size_t nfft = segment_length(nsc.data(), nbeams);
    const size_t good = (nfft - dmax) & ~size_t(1);
    const size_t M = nfft / 2;
    const size_t Xp = M + 64, Yp = M + 64;
    const bool valu = env_flag("DEDISP_FDD_VALU") && st.precision == DEDISP_AMD_FDD_FP32;
    const float xscale = prepare_phasors(nfft, valu);
    const float2* ph = st.d_ph.as<float2>();
    const size_t nchan_pad = padded_channels(nchan);

    // ---- layout: groups of beams, each beam's rows / staged input / output ----
    // DEDISP_FDD_GROUP_CHUNKS caps the chunks of a group below the memory bound
    size_t max_chunks =
        std::max<size_t>(1, kZBudget / (nchan_pad * Xp * 8 + dd_fdd_split_work(nfft, nchan)));
    if (const char* e = std::getenv("DEDISP_FDD_GROUP_CHUNKS"))
        if (std::atol(e) > 0) max_chunks = std::min(max_chunks, (size_t)std::atol(e));
    struct Group {
        size_t b0 = 0, b1 = 0, chunk0 = 0, nchunk = 0, rows = 0, in_bytes = 0, out_floats = 0;
    };
    std::vector<Group> groups;
    std::vector<size_t> nch(nbeams), rw(nbeams), roff(nbeams), ioff(nbeams), ooff(nbeams);
    const size_t spw = 32 / in_nbits;  // samples per 32-bit row word
    Group cur;
    size_t total_chunks = 0;
    double work = 0;
    for (size_t b = 0; b < nbeams; ++b) {
        nch[b] = div_up(nsc[b], good);
        if (cur.b1 > cur.b0 && cur.nchunk + nch[b] > max_chunks) {
            groups.push_back(cur);
            cur = Group{};
            cur.b0 = cur.b1 = b;
            cur.chunk0 = total_chunks;
        }
        const size_t span = nch[b] ? std::max(nsamps[b], (nch[b] - 1) * good + nfft) : 0;
        rw[b] = div_up(span, spw * 64) * 64;
        roff[b] = cur.rows;
        cur.rows += nchan * rw[b];
        ioff[b] = cur.in_bytes;
        cur.in_bytes += round_up(nsamps[b] * in_row, 256);
        ooff[b] = cur.out_floats;
        cur.out_floats += round_up(ndm * nsc[b], 64);
        cur.nchunk += nch[b];
        cur.b1 = b + 1;
        total_chunks += nch[b];
        work += (double)ndm * (double)nchan * (double)nch[b] * (double)(M + 1);
    }
    groups.push_back(cur);
    if (total_chunks == 0) return;
    size_t g_rows = 0, g_in = 0, g_out = 0, g_chunks = 0;
    for (const Group& g : groups) {
        g_rows = std::max(g_rows, g.rows);
        g_in = std::max(g_in, g.in_bytes);
        g_out = std::max(g_out, g.out_floats);
        g_chunks = std::max(g_chunks, g.nchunk);
    }
    const size_t per_dm = g_chunks * Yp * 8 + dd_fdd_split_work(nfft, g_chunks);
    const size_t dm_batch = balanced_batch(ndm, kYBudget / per_dm);
    st.d_rows.reserve(g_rows * 4);
    st.d_Z.reserve(nchan_pad * g_chunks * Xp * 8);
    st.d_Y.reserve(dm_batch * g_chunks * Yp * 8);
    st.d_fwork.reserve(std::max<size_t>(dd_fdd_split_work(nfft, nchan * g_chunks), 16));
    st.d_iwork.reserve(std::max<size_t>(dd_fdd_split_work(nfft, dm_batch * g_chunks), 16));
    if (!device_ptrs) {
        st.d_in.reserve(g_in + 64);
        st.d_out.reserve(g_out * 4);
    }

    // device pointers: a captured graph per batch (lengths, pointers, segment
    // length, precision, tables and buffers), replayed while they stay the same
    std::vector<uint64_t> key = {nbeams, in_nbits, nfft, (uint64_t)st.precision, valu ? 1u : 0u,
                                 m_tables_version, max_chunks, dm_batch, __builtin_bit_cast(uint32_t, xscale),
                                 (uint64_t)st.d_rows.ptr, (uint64_t)st.d_Z.ptr, (uint64_t)st.d_Y.ptr,
                                 (uint64_t)st.d_q32.ptr, (uint64_t)st.d_ph.ptr, (uint64_t)st.d_ph64k.ptr,
                                 (uint64_t)st.d_fwork.ptr, (uint64_t)st.d_iwork.ptr,
                                 (uint64_t)contraction_kind(st), (uint64_t)contraction_packs(st),
                                 (uint64_t)mfma16_mirror(st), (uint64_t)(m16cfg_env() + 1),
                                 (uint64_t)(st.tuned_prec + 1), (uint64_t)(st.tuned_form + 1),
                                 (uint64_t)(st.tuned_mirror + 1), (uint64_t)(st.tuned_blocked + 1),
                                 (uint64_t)env_flag("DEDISP_FDD_QTABLE"), (uint64_t)env_off("DEDISP_FDD_BLOCKED"),
                                 (uint64_t)st.d_scale.ptr, (uint64_t)st.d_delays.ptr};
    for (size_t b = 0; b < nbeams; ++b) {
        key.push_back(nsamps[b]);
        key.push_back((uint64_t)in[b]);
        key.push_back((uint64_t)out[b]);
    }
    FDDState::Graph* ge = nullptr;
    bool replay = false;
    if (device_ptrs) {
        for (auto& g : st.graphs)
            if (g->exec && g->key == key) {
                ge = g.get();
                replay = true;
                break;
            }
        if (!ge) {
            if (st.graphs.size() < FDDState::kGraphs) st.graphs.emplace_back(new FDDState::Graph);
            ge = std::min_element(st.graphs.begin(), st.graphs.end(),
                                  [](const auto& x, const auto& y) { return x->used < y->used; })
                     ->get();
            ge->drop();
        }
        ge->used = ++st.graph_clock;
    }
    DeviceBuffer& d_chunks = ge ? ge->d_chunks : st.d_chunks;

    // ---- chunk tables (one upload for the batch) ----
    if (!replay) {
        d_chunks.reserve(total_chunks * sizeof(dd_fdd_chunk));
        std::vector<dd_fdd_chunk> tab(total_chunks);
        for (const Group& g : groups) {
            size_t i = g.chunk0;
            for (size_t b = g.b0; b < g.b1; ++b)
                for (size_t j = 0; j < nch[b]; ++j, ++i) {
                    dd_fdd_chunk& e = tab[i];
                    e.rows = st.d_rows.as<uint8_t>() + roff[b] * 4;
                    e.out = device_ptrs ? reinterpret_cast<float*>(out[b]) : st.d_out.as<float>() + ooff[b];
                    e.row_bytes = rw[b] * 4;
                    e.out_stride = nsc[b];
                    e.t0 = (uint32_t)(j * good);
                    e.nvalid = (uint32_t)nsamps[b];
                    e.nsamps_out = (uint32_t)nsc[b];
                    e.reserved = 0;
                }
        }
        DD_HIP(hipMemcpyAsync(d_chunks.ptr, tab.data(), tab.size() * sizeof(dd_fdd_chunk),
                              hipMemcpyHostToDevice, stream));
        DD_HIP(hipStreamSynchronize(stream));
    }
    const dd_fdd_chunk* d_tab = d_chunks.as<dd_fdd_chunk>();

    // the kernels of one group (no copies: they may be captured)
    auto group_kernels = [&](const Group& g, bool timed) {
        const size_t nc = g.nchunk;
        const bool blocked = blocked_layout(st, nfft, nc, Xp);
        const size_t zrow = chan_elems(blocked, nfft, nc, Xp);
        if (nchan_pad > nchan)
            DD_HIP(dd_zero_fill(st.d_Z.as<float2>() + nchan * zrow, (nchan_pad - nchan) * zrow * 8, stream));
        if (timed) ctx.timer.begin(StageTimer::UNPACK, stream);
        for (size_t b = g.b0; b < g.b1; ++b) {
            if (!nch[b]) continue;
            const uint8_t* src = device_ptrs ? in[b] : st.d_in.as<uint8_t>() + ioff[b];
            DD_HIP(dd_tdd_transpose(src, nsamps[b], in_row, (unsigned)in_nbits, nchan,
                                    st.d_rows.as<uint32_t>() + roff[b], rw[b], stream));
        }
        if (timed) ctx.timer.end(StageTimer::UNPACK, stream);
        if (timed) ctx.timer.begin(StageTimer::FFT, stream);
        DD_HIP(dd_fdd_forward_chunks_split(d_tab + g.chunk0, nc, st.d_scale.as<float>(), ph,
                                           st.d_ph64k.as<float2>(), nfft, st.d_Z.as<float2>(), nchan, Xp,
                                           (unsigned)in_nbits, st.d_fwork.as<float2>(), blocked ? 1 : 0, stream));
        if (timed) ctx.timer.end(StageTimer::FFT, stream);
        for (size_t b0 = 0; b0 < ndm; b0 += dm_batch) {
            const size_t nb = std::min(dm_batch, ndm - b0);
            if (timed) ctx.timer.begin(StageTimer::DEDISPERSE, stream);
            contract(st, nchan, b0, nb, nc, nfft, Xp, Yp, xscale, blocked, stream);
            if (timed) ctx.timer.end(StageTimer::DEDISPERSE, stream);
            if (timed) ctx.timer.begin(StageTimer::IFFT, stream);
            DD_HIP(dd_fdd_inverse_chunks_split(st.d_Y.as<float2>(), nb, b0, d_tab + g.chunk0, nc, Yp, ph,
                                               st.d_ph64k.as<float2>(), nfft, good, st.d_iwork.as<float2>(),
                                               inverse_form(st, blocked), stream));
            if (timed) ctx.timer.end(StageTimer::IFFT, stream);
        }
    };

    if (device_ptrs) {
        if (!replay) {
            DD_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
            try {
                for (const Group& g : groups)
                    if (g.nchunk) group_kernels(g, false);
            } catch (...) {
                hipGraph_t gr = nullptr;
                (void)hipStreamEndCapture(stream, &gr);
                if (gr) (void)hipGraphDestroy(gr);
                throw;
            }
            DD_HIP(hipStreamEndCapture(stream, &ge->graph));
            DD_HIP(hipGraphInstantiate(&ge->exec, ge->graph, nullptr, nullptr, 0));
            ge->key = key;
        }
        ctx.timer.begin(StageTimer::DEDISPERSE, stream);
        DD_HIP(hipGraphLaunch(ge->exec, stream));
        ctx.timer.end(StageTimer::DEDISPERSE, stream);
    } else {
        for (const Group& g : groups) {
            if (g.nchunk == 0) continue;
            ctx.timer.begin(StageTimer::H2D, stream);
            for (size_t b = g.b0; b < g.b1; ++b)
                if (nch[b])
                    DD_HIP(hipMemcpyAsync(st.d_in.as<uint8_t>() + ioff[b], in[b], nsamps[b] * in_row,
                                          hipMemcpyHostToDevice, stream));
            ctx.timer.end(StageTimer::H2D, stream);
            group_kernels(g, true);
            ctx.timer.begin(StageTimer::D2H, stream);
            for (size_t b = g.b0; b < g.b1; ++b)
                if (nch[b])
                    DD_HIP(hipMemcpyAsync(out[b], st.d_out.as<float>() + ooff[b], ndm * nsc[b] * 4,
                                          hipMemcpyDeviceToHost, stream));
            ctx.timer.end(StageTimer::D2H, stream);
        }
    }
    if (!(device_ptrs && (flags & DEDISP_ASYNC))) {
        DD_HIP(hipStreamSynchronize(stream));
        ctx.timer.collect(ctx.profile);
        ctx.profile.work_units = work;
    }
}

}  // namespace dedisp
