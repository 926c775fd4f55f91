// TDDPlan: host side of the time-domain dedispersion path.
//
// Reference: src/tdd/TDDPlan.cpp:43-487. The reference streams gulps of
// m_gulp_size (65536) output samples through pinned double buffers because its
// device copy of the input is gulp-sized; here the whole series is made
// resident in HBM (288 GB per GPU), transposed once, and every output sample
// of every DM is produced by one launch of the channel-sum kernel. Gulps are
// used when the caller sets a gulp size (set_gulp_size), when the resident
// buffers of a series would exceed 80 % of the free device memory, or when a
// series has 2^31 samples or more: the series is then processed in gulps of
// output samples, each reading its max_delay samples of overlap again
// (src/tdd/TDDPlan.cpp:225-263, 354-420).
#include <algorithm>
#include <atomic>
#include <cstdlib>
#include <cstring>
#include <memory>

#include "internal.hpp"
#include "kernels.h"

namespace dedisp {

struct TDDState {
    // device buffers, grown on demand and kept between executes
    DeviceBuffer d_in;      // staged input (host-pointer mode)
    DeviceBuffer d_rows;    // channel-major packed rows
    DeviceBuffer d_out;     // staged output (host-pointer mode)
    DeviceBuffer d_code;    // fast path [DM group][nact][DMs per wave] LDS offset | shift; generic [ndm][nact] delay
    DeviceBuffer d_chan;    // [nact] channel index
    DeviceBuffer d_weight;  // [nact] killmask weight (generic path)
    DeviceBuffer d_wstart;  // [ndmblk][nact] staged window start (fast path)
    DeviceBuffer d_chunks;  // [ndmblk][maxc] staging chunks (runs-layout kernel)

    // host copies of the tables (what the device buffers above hold once
    // uploaded): fast path delay codes or generic [ndm][nact] delays, channel
    // indices, weights, window starts
    std::vector<uint32_t> h_tab, h_wstart, h_chunks;
    std::vector<int32_t> h_chan, h_weight;
    bool uploaded = false;   // d_code / d_chan / d_weight / d_wstart hold the host tables
    uint64_t table_uid = 0;  // process-unique id of the current host tables

    // cache key of the tables
    unsigned long version = ~0ul;
    unsigned nbits = 0;
    size_t first_dm = ~size_t(0), dm_count = 0;

    // derived from the tables
    bool fast = false;
    dd_tdd_tiling tiling{};
    size_t nact = 0;
    size_t dmax = 0;        // largest delay (samples) over the DMs in use
    size_t max_wstart = 0;  // fast path
    unsigned maxc = 0;      // runs-layout kernel: chunk entries per DM block (terminator included)
    size_t row_words = 0;

    // the caller chose the gulp size (set_gulp_size): execute in gulps of it
    bool gulp_explicit = false;

    // tilings chosen by the autotuner, per input width (waves == 0: none)
    dd_tdd_tiling tuned[33] = {};
    dd_tdd_tiling tiling_key{};

    // batched execution (own buffers, so a cached graph's pointers stay valid
    // whatever single executes do in between)
    DeviceBuffer d_brows, d_bin, d_bout, d_beams, d_tile_beam;
    std::vector<dd_tdd_beam> h_beams;
    std::vector<uint32_t> h_tile_beam;

    // Captured device-pointer batches, the least recently used replaced first.
    // Each keeps its own beam tables, so callers alternating between a few
    // buffer sets (a stream's double-buffered histories) replay every one.
    struct Graph {
        std::vector<uint64_t> key;
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        DeviceBuffer d_beams, d_tile_beam;
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

    // end of each DM block's channel sum (host-pointer executes)
    std::vector<hipEvent_t> blk_events;
    ~TDDState()
    {
        for (hipEvent_t e : blk_events) (void)hipEventDestroy(e);
    }
};

namespace {

// Default tiling per input width (words per lane x SPW samples, DMs per wave,
// waves); the chunk size is fitted to the LDS budget below. Overridden by the
// environment variable DEDISP_TDD_TILING="samples_per_lane,dms_per_wave,waves,chunk".
dd_tdd_tiling default_tiling(unsigned nbits)
{
    dd_tdd_tiling t{};
    // Measured on MI355X (channel-sum kernel, 2^20 samples, 16-channel
    // staging chunks): 2 DMs per wave beat 4 at every width once a channel
    // pair's codes come in one scalar load. 4-bit, 1024 ch x 1167 DMs:
    // 29.5 ms with 8 waves (4 DMs/wave, 4 waves: 32.4 ms); 8-bit, 1024 ch x
    // 1000 DMs, 16 samples/lane: 44.4 ms (8 samples/lane, 4 DMs: 50.8 ms);
    // 2-bit, 512 ch x 1024 DMs: 12.8 ms (13.3 ms); 1-bit, 512 ch x 1024 DMs:
    // 11.9 ms (13.7 ms). 4-bit with 32-channel chunks: 28.8 ms (16: 29.9,
    // 64: 42.5). With the windows staged by global-to-LDS loads the 4-bit
    // optimum moved: 4 DMs per wave, 4 waves, 32-channel chunks 25.7 ms
    // (2 DMs x 8 waves: 26.1; 16-channel chunks 26.4, 64: 42.5).
    t.dms_per_wave = 2;
    t.waves = 4;
    t.chunk_channels = 16;
    switch (nbits) {
    case 1: t.samples_per_lane = 32; break;
    case 2: t.samples_per_lane = 16; break;
    // 4-bit: the runs-layout kernel (7 runs x 4 positions per lane, 4 DMs x
    // 4 waves): 22.6 ms on 1024 ch x 2^20 x 1167 DMs, against 25.8 ms for
    // the packed-lane kernel's best tiling (16 samples per lane, 4 DMs x 4
    // waves, 32-channel chunks).
    case 4: t.samples_per_lane = DD_TDD_RUNS_SPL; t.dms_per_wave = 4; t.chunk_channels = 32; break;
    default: t.samples_per_lane = 16; break;
    }
    return t;
}

constexpr size_t kLdsBudget = 64 * 1024;  // bytes of LDS per workgroup

// Output samples per time tile of the channel-sum kernel: 7 runs of 256 for
// the runs-layout kernel, 64 lanes x samples_per_lane for the packed-lane one.
size_t tile_samples(unsigned nbits, const dd_tdd_tiling& t)
{
    return dd_tdd_tiling_is_runs(nbits, t) ? size_t(7 * 256) : size_t(64) * t.samples_per_lane;
}

// Tables of the runs-layout channel sum (tdd_runs.hip) from the integer delays
// d[dm][j] (samples): per DM block (the tiling's waves x DMs per wave) and channel, the staged window
// start (the smallest delay rounded down to 16 samples) and the positions
// the block's DMs read (T = 256 plus the delay span, in whole 128s); channels
// are grouped, in order, into chunks whose staged windows and run words fit
// the kernel's two LDS buffers (at most cap channels each); codes are the
// LDS byte offsets of every DM's window in its chunk's run words. False when
// one channel's window does not fit (then the packed-lane kernel is used).
bool build_runs_tables(const std::vector<uint32_t>& d, size_t dm_count, size_t nact, const dd_tdd_tiling& tl,
                       std::vector<uint32_t>& code, std::vector<uint32_t>& wstart, std::vector<uint32_t>& chunks,
                       unsigned& maxc, size_t& max_wstart, size_t& window_words)
{
    constexpr size_t T = 256;
    const size_t DPW = tl.dms_per_wave, WAVES = tl.waves, DPB = DPW * WAVES;
    unsigned cap = tl.chunk_channels;
    constexpr size_t L1 = DD_TDD_RUNS_L1_BYTES, L2 = DD_TDD_RUNS_L2_BYTES;
    if (nact == 0 || nact > 4369 || dm_count == 0) return false;
    cap = std::max(1u, std::min(cap, (unsigned)DD_TDD_RUNS_MAXCC));
    auto slot1 = [](size_t P) { return div_up(6 * T / 8 + P / 8, size_t(64)) * 72 * 4; };
    const size_t ndmblk = div_up(dm_count, DPB);
    wstart.assign(ndmblk * nact, 0);
    std::vector<std::vector<uint32_t>> blk(ndmblk);
    std::vector<uint32_t> chunk_of(ndmblk * nact);  // chunk index within the block
    max_wstart = 0;
    window_words = 0;
    size_t most = 0;
    for (size_t b = 0; b < ndmblk; ++b) {
        const size_t i0 = b * DPB, i1 = std::min(dm_count, i0 + DPB);
        size_t j0 = 0, cc = 0, P = 0;
        auto close = [&]() {
            blk[b].push_back((uint32_t)j0 | (uint32_t)cc << 16 | (uint32_t)(P / 128) << 24);
            window_words = std::max(window_words, 6 * T / 8 + P / 8);
        };
        for (size_t j = 0; j < nact; ++j) {
            uint32_t lo = ~0u, hi = 0;
            for (size_t i = i0; i < i1; ++i) {
                lo = std::min(lo, d[i * nact + j]);
                hi = std::max(hi, d[i * nact + j]);
            }
            const uint32_t a = lo & ~15u;
            wstart[b * nact + j] = a / 8;
            max_wstart = std::max<size_t>(max_wstart, a / 8);
            const size_t need = round_up(T + (hi - a), size_t(128));
            size_t P2 = std::max(P, need);
            if (cc > 0 && ((cc + 1) * slot1(P2) > L1 || (cc + 1) * P2 * 4 > L2 || cc + 1 > cap)) {
                close();
                j0 = j;
                cc = 0;
                P2 = need;
            }
            if (slot1(P2) > L1 || P2 * 4 > L2 || P2 / 128 > 255) return false;
            P = P2;
            chunk_of[b * nact + j] = (uint32_t)blk[b].size();
            ++cc;
        }
        close();
        most = std::max(most, blk[b].size());
    }
    maxc = (unsigned)(most + 1);
    chunks.assign(ndmblk * maxc, 0);
    for (size_t b = 0; b < ndmblk; ++b)
        std::copy(blk[b].begin(), blk[b].end(), chunks.begin() + b * maxc);
    // code[group][j][i]: DMs past the end repeat the last one
    const size_t ngroups = ndmblk * WAVES;
    code.assign(ngroups * nact * DPW, 0);
    for (size_t g = 0; g < ngroups; ++g) {
        const size_t b = g / WAVES;
        for (size_t j = 0; j < nact; ++j) {
            const uint32_t e = blk[b][chunk_of[b * nact + j]];
            const size_t cj0 = e & 0xffffu, P = (size_t)(e >> 24) * 128;
            const uint32_t a = wstart[b * nact + j] * 8;
            for (size_t i = 0; i < DPW; ++i) {
                const size_t dm = std::min(g * DPW + i, dm_count - 1);
                code[(g * nact + j) * DPW + i] = (uint32_t)(((j - cj0) * P + (d[dm * nact + j] - a)) * 4);
            }
        }
    }
    return true;
}

}  // namespace

TDDPlan::TDDPlan(size_type nchans, float_type dt, float_type f0, float_type df, int device_idx)
    : GPUPlan(nchans, dt, f0, df, device_idx), m_state(new TDDState)
{
    if (nchans > kTddMaxChans) DD_THROW(DEDISP_NCHANS_EXCEEDS_LIMIT);
    m_gulp_size = 65536;  // src/tdd/TDDPlan.cpp:71-74
}

TDDPlan::~TDDPlan() {}

void TDDPlan::set_gulp_size(size_type gulp_size)
{
    if (gulp_size == 0) DD_THROW(DEDISP_UNKNOWN_ERROR);
    m_gulp_size = gulp_size;
    m_state->gulp_explicit = true;
}

namespace {
std::atomic<uint64_t> g_table_uid{0};
}

// Tables for input width nbits and DMs first_dm_idx .. +dm_count: active
// channels, integer delays, and for the packed-lane kernel the window starts
// and delay codes. Built on the host and cached per DM list / killmask / width
// / DM range / tiling (independent of the series length); with upload they are
// also copied to the plan's device buffers.
void TDDPlan::prepare_tables(unsigned nbits, dedisp_size first_dm_idx, dedisp_size dm_count, bool upload)
{
    GPUContext& ctx = *m_ctx;
    TDDState& st = *m_state;
    hipStream_t stream = ctx.exec;
    const size_t spw = 32 / nbits;
    dd_tdd_tiling want = default_tiling(nbits);
    if (st.tuned[nbits].waves) want = st.tuned[nbits];
    if (const char* env = std::getenv("DEDISP_TDD_TILING")) {
        unsigned a, b, c, e;
        if (std::sscanf(env, "%u,%u,%u,%u", &a, &b, &c, &e) == 4) {
            want.samples_per_lane = a; want.dms_per_wave = b; want.waves = c; want.chunk_channels = e;
        }
    }
    if (st.version != m_tables_version || st.nbits != nbits || st.first_dm != first_dm_idx ||
        st.dm_count != dm_count || std::memcmp(&st.tiling_key, &want, sizeof want) != 0) {
        st.tiling_key = want;
        std::vector<int32_t> chan, weight;
        bool binary = true;
        for (size_t c = 0; c < m_nchans; ++c) {
            int32_t k = h_killmask[c];
            if (k == 0) continue;
            if (k != 1) binary = false;
            chan.push_back((int32_t)c);
            weight.push_back(k);
        }
        const size_t nact = chan.size();
        std::vector<uint32_t> d((size_t)dm_count * nact);
        size_t dmax = 0;
        for (size_t i = 0; i < dm_count; ++i) {
            float dm = h_dm_list[first_dm_idx + i];
            for (size_t j = 0; j < nact; ++j) {
                uint32_t v = tdd_delay(dm, h_delay_table[chan[j]]);
                d[i * nact + j] = v;
                dmax = std::max<size_t>(dmax, v);
            }
        }

        dd_tdd_tiling tl = want;
        bool fast = binary && nact > 0 && nbits <= 8 && dd_tdd_lds_supported(nbits, tl) &&
                    (uint64_t)dm_count * nact < (uint64_t(1) << 31);

        std::vector<uint32_t> code, wstart, chunks;
        size_t max_wstart = 0, W = 0;
        unsigned maxc = 0;
        if (fast && dd_tdd_tiling_is_runs(nbits, tl)) {
            size_t ww = 0;
            if (build_runs_tables(d, dm_count, nact, tl, code, wstart, chunks, maxc, max_wstart, ww)) {
                tl.window_words = (unsigned)ww;
            } else {
                // a window too long for the runs kernel's LDS buffers
                tl = dd_tdd_tiling{16, 4, 4, 32, 0};
                code.clear();
                wstart.clear();
                chunks.clear();
                maxc = 0;
                max_wstart = 0;
            }
        }
        if (fast && !dd_tdd_tiling_is_runs(nbits, tl)) {
            const size_t G = tl.samples_per_lane / spw;
            const size_t dpb = (size_t)tl.waves * tl.dms_per_wave;
            const size_t ndmblk = div_up(dm_count, dpb);
            wstart.assign(ndmblk * nact, 0);
            size_t span = 0;
            for (size_t b = 0; b < ndmblk; ++b) {
                const size_t i0 = b * dpb, i1 = std::min(dm_count, i0 + dpb);
                for (size_t j = 0; j < nact; ++j) {
                    uint32_t lo = ~0u, hi = 0;
                    for (size_t i = i0; i < i1; ++i) {
                        uint32_t q = d[i * nact + j] / (uint32_t)spw;
                        lo = std::min(lo, q);
                        hi = std::max(hi, q);
                    }
                    wstart[b * nact + j] = lo;
                    max_wstart = std::max<size_t>(max_wstart, lo);
                    span = std::max<size_t>(span, hi - lo);
                }
            }
            W = span + 64 * G + 1;
            // channels per staged chunk: LDS budget, and for 8-bit input the
            // 16-bit lanes hold at most 257 channels of 255
            size_t cc = std::min<size_t>(tl.chunk_channels, nact);
            if (nbits == 8) cc = std::min<size_t>(cc, 256);
            while (cc > 1 && cc * W * 4 > kLdsBudget) cc /= 2;
            if (cc * W * 4 > kLdsBudget || span >= (1u << 26)) {
                fast = false;
            } else {
                tl.chunk_channels = (unsigned)cc;
                tl.window_words = (unsigned)W;
                // code[group][j][i]: DM i of the wave's group of dms_per_wave
                // (DMs past the end repeat the last one, whose windows lie in
                // the same last block), LDS byte offset of its delayed window
                // in staged slot j % cc, shifted left 5, | bit or byte shift
                const size_t dpw = tl.dms_per_wave, ngroups = ndmblk * tl.waves;
                code.assign(ngroups * nact * dpw, 0);
                for (size_t g = 0; g < ngroups; ++g) {
                    const size_t b = g / tl.waves;
                    for (size_t j = 0; j < nact; ++j)
                        for (size_t i = 0; i < dpw; ++i) {
                            const size_t dm = std::min(g * dpw + i, dm_count - 1);
                            const uint32_t v = d[dm * nact + j];
                            const uint32_t rel = v / (uint32_t)spw - wstart[b * nact + j];
                            const uint32_t sh = (nbits == 8) ? (v % 4u) : (v % (uint32_t)spw) * nbits;
                            const uint32_t byte = (uint32_t)(((j % cc) * W + rel) * 4);
                            code[(g * nact + j) * dpw + i] = (byte << 5) | sh;
                        }
                }
            }
        }

        st.h_chan = std::move(chan);
        st.h_weight = std::move(weight);
        st.h_tab = fast ? std::move(code) : std::move(d);
        st.h_wstart = fast ? std::move(wstart) : std::vector<uint32_t>();
        st.h_chunks = fast ? std::move(chunks) : std::vector<uint32_t>();
        st.maxc = fast ? maxc : 0;
        st.uploaded = false;
        st.table_uid = ++g_table_uid;

        st.version = m_tables_version;
        st.nbits = nbits;
        st.first_dm = first_dm_idx;
        st.dm_count = dm_count;
        st.fast = fast;
        st.tiling = tl;
        st.nact = nact;
        st.dmax = dmax;
        st.max_wstart = max_wstart;
    }
    if (upload && !st.uploaded) {
        const size_t nact = st.nact;
        // 8 spare entries: the runs-layout kernel reads channel numbers and
        // window starts in groups of 8
        st.d_chan.reserve((nact + 8) * 4);
        st.d_weight.reserve(std::max<size_t>(nact, 1) * 4);
        if (nact) {
            DD_HIP(hipMemcpyAsync(st.d_chan.ptr, st.h_chan.data(), nact * 4, hipMemcpyHostToDevice, stream));
            DD_HIP(hipMemcpyAsync(st.d_weight.ptr, st.h_weight.data(), nact * 4, hipMemcpyHostToDevice, stream));
        }
        // 64 spare entries: the channel-sum kernel may prefetch one pair of
        // codes past its last channel
        st.d_code.reserve((st.h_tab.size() + 64) * 4);
        if (!st.h_tab.empty())
            DD_HIP(hipMemcpyAsync(st.d_code.ptr, st.h_tab.data(), st.h_tab.size() * 4, hipMemcpyHostToDevice,
                                  stream));
        if (st.fast) {
            st.d_wstart.reserve((st.h_wstart.size() + 8) * 4);
            DD_HIP(hipMemcpyAsync(st.d_wstart.ptr, st.h_wstart.data(), st.h_wstart.size() * 4,
                                  hipMemcpyHostToDevice, stream));
            st.d_chunks.reserve(std::max<size_t>(st.h_chunks.size(), 1) * 4);
            if (!st.h_chunks.empty())
                DD_HIP(hipMemcpyAsync(st.d_chunks.ptr, st.h_chunks.data(), st.h_chunks.size() * 4,
                                      hipMemcpyHostToDevice, stream));
        }
        DD_HIP(hipStreamSynchronize(stream));
        st.uploaded = true;
    }
}

// Words per channel row of the transposed input for nsamps samples with nsc
// output samples: every (DM, output) window of every time tile lies inside.
size_t TDDPlan::row_words_for(unsigned nbits, size_t nsamps, size_t nsc) const
{
    const TDDState& st = *m_state;
    const size_t spw = 32 / nbits;
    size_t row_words;
    if (st.fast) {
        const size_t tile_words = tile_samples(nbits, st.tiling) / spw;
        const size_t ntiles = div_up(nsc, tile_words * spw);
        row_words = ntiles * tile_words + st.max_wstart + st.tiling.window_words;
    } else {
        row_words = (nsc + st.dmax) / spw + 2;
    }
    row_words = std::max(row_words, div_up(nsamps, spw));
    return round_up(row_words, 64);
}

void TDDPlan::execute(size_type nsamps, const byte_type* in, size_type in_nbits,
                      byte_type* out, size_type out_nbits, unsigned flags)
{
    // src/tdd/TDDPlan.cpp:102-127: unpadded strides
    dedisp_size out_bytes_per_sample = out_nbits / 8;
    dedisp_size in_stride = m_nchans * in_nbits / 8;
    dedisp_size out_stride = (nsamps - m_max_delay) * out_bytes_per_sample;
    execute_adv(nsamps, in, in_nbits, in_stride, out, out_nbits, out_stride, flags);
}

void TDDPlan::execute_adv(size_type nsamps, const byte_type* in, size_type in_nbits,
                          size_type in_stride, byte_type* out, size_type out_nbits,
                          size_type out_stride, unsigned flags)
{
    execute_guru(nsamps, in, in_nbits, in_stride, out, out_nbits, out_stride, 0, m_dm_count, flags);
}

void TDDPlan::execute_guru(size_type nsamps, const byte_type* in, size_type in_nbits,
                           size_type in_stride, byte_type* out, size_type out_nbits,
                           size_type out_stride, dedisp_size first_dm_idx,
                           dedisp_size dm_count, unsigned flags)
{
    // Validation in the order of src/tdd/TDDPlan.cpp:167-195 (the unsigned
    // wrap of nsamps - max_delay in the stride test is kept, so nsamps <
    // max_delay reports DEDISP_INVALID_STRIDE as the reference does).
    dedisp_size out_bytes_per_sample = out_nbits / 8;
    if (!in || !out) DD_THROW(DEDISP_INVALID_POINTER);
    if (in_stride < m_nchans * in_nbits / 8 ||
        out_stride < (nsamps - m_max_delay) * out_bytes_per_sample)
        DD_THROW(DEDISP_INVALID_STRIDE);
    if (m_dm_count == 0) DD_THROW(DEDISP_NO_DM_LIST_SET);
    if (nsamps < m_max_delay) DD_THROW(DEDISP_TOO_FEW_NSAMPS);
    // src/dedisp/DedispPlan.cpp:137-139
    if ((flags & DEDISP_ASYNC) && (flags & DEDISP_WAIT)) DD_THROW(DEDISP_INVALID_FLAG_COMBINATION);
    if ((flags & DEDISP_HOST_POINTERS) && (flags & DEDISP_DEVICE_POINTERS))
        DD_THROW(DEDISP_INVALID_FLAG_COMBINATION);
    if (in_nbits != 1 && in_nbits != 2 && in_nbits != 4 && in_nbits != 8 && in_nbits != 16 &&
        in_nbits != 32)
        DD_THROW(DEDISP_UNSUPPORTED_IN_NBITS);
    if (out_nbits != 8 && out_nbits != 16 && out_nbits != 32) DD_THROW(DEDISP_UNSUPPORTED_OUT_NBITS);
    if (first_dm_idx > m_dm_count || dm_count > m_dm_count - first_dm_idx)
        DD_THROW(DEDISP_UNKNOWN_ERROR);

    const size_t nsc = nsamps - m_max_delay;
    if (nsc == 0 || dm_count == 0) return;

    // Output samples per gulp: the caller's gulp size; else the whole series
    // unless its resident buffers (rows, and for host pointers the staged input
    // and output) exceed 80 % of the free device memory or it reaches 2^31
    // samples (the kernels index samples in 32 bits).
    const unsigned nbits = (unsigned)in_nbits;
    const bool device_ptrs = (flags & DEDISP_DEVICE_POINTERS) != 0;
    const double row_bytes_per_sample = (double)m_nchans * nbits / 8.0;
    size_t gulp = nsc;
    if (m_state->gulp_explicit) gulp = std::min(nsc, (size_t)m_gulp_size);
    const size_t kMaxSpan = (size_t(1) << 31) - (size_t(1) << 20);  // samples per span, with tile rounding room
    if (gulp + m_max_delay > kMaxSpan) {
        if (m_max_delay + 65536 > kMaxSpan) DD_THROW(DEDISP_UNKNOWN_ERROR);
        gulp = kMaxSpan - m_max_delay;
    }
    {
        m_ctx->make_current();
        size_t free_b = 0, total_b = 0;
        DD_HIP(hipMemGetInfo(&free_b, &total_b));
        auto need = [&](size_t g) {
            const double span = (double)(g + m_max_delay);
            double b = span * row_bytes_per_sample * 1.25;  // rows (tile rounding, 1-bit rows padded)
            if (!device_ptrs) b += span * row_bytes_per_sample + (double)dm_count * g * (out_nbits / 8);
            return b;
        };
        const double budget = 0.8 * (double)free_b;
        while (gulp > 65536 && need(gulp) > budget) gulp = (gulp + 1) / 2;
    }
    m_ctx->timer.reset(m_ctx->profiling);
    m_ctx->profile = dedisp_amd_profile{};
    const size_t out_bytes = out_nbits / 8;
    for (size_t o0 = 0; o0 < nsc; o0 += gulp) {
        const size_t g = std::min(gulp, nsc - o0);
        execute_span(g + m_max_delay, in + o0 * in_stride, in_nbits, in_stride, out + o0 * out_bytes, out_nbits,
                     out_stride, first_dm_idx, dm_count, flags);
    }
}

// One resident pass: nsamps input samples at in (nsamps - max_delay outputs
// per DM at out), validated by execute_guru; stage times and work add to the
// plan's profile.
void TDDPlan::execute_span(size_type nsamps, const byte_type* in, size_type in_nbits, size_type in_stride,
                           byte_type* out, size_type out_nbits, size_type out_stride, dedisp_size first_dm_idx,
                           dedisp_size dm_count, unsigned flags)
{
    const dedisp_size out_bytes_per_sample = out_nbits / 8;
    const bool device_ptrs = (flags & DEDISP_DEVICE_POINTERS) != 0;
    const size_t nsc = nsamps - m_max_delay;

    GPUContext& ctx = *m_ctx;
    TDDState& st = *m_state;
    ctx.make_current();
    hipStream_t stream = ctx.exec;

    const unsigned nbits = (unsigned)in_nbits;
    prepare_tables(nbits, first_dm_idx, dm_count);
    st.row_words = row_words_for(nbits, nsamps, nsc);

    // ---- input ----
    const size_t row_bytes = (m_nchans * nbits + 7) / 8;
    const uint8_t* d_in;
    size_t d_in_stride = in_stride;
    if (device_ptrs) {
        d_in = in;
    } else {
        // rows packed on the device; the host rows go through page-locked slots
        st.d_in.reserve(nsamps * row_bytes + 64);
        ctx.timer.begin(StageTimer::H2D, stream);
        ctx.io.upload(stream, st.d_in.as<uint8_t>(), row_bytes, in, in_stride, row_bytes, nsamps);
        ctx.timer.end(StageTimer::H2D, stream);
        d_in = st.d_in.as<uint8_t>();
        d_in_stride = row_bytes;
    }

    // ---- transpose ----
    st.d_rows.reserve(m_nchans * st.row_words * 4);
    ctx.timer.begin(StageTimer::UNPACK, stream);
    DD_HIP(dd_tdd_transpose(d_in, nsamps, d_in_stride, nbits, m_nchans, st.d_rows.as<uint32_t>(),
                            st.row_words, stream));
    ctx.timer.end(StageTimer::UNPACK, stream);

    // ---- channel sum ----
    uint8_t* d_out;
    size_t d_out_stride;
    if (device_ptrs) {
        d_out = out;
        d_out_stride = out_stride;
    } else {
        d_out_stride = nsc * out_bytes_per_sample;
        st.d_out.reserve(dm_count * d_out_stride);
        d_out = st.d_out.as<uint8_t>();
    }
    dd_tdd_epilogue ep;
    {
        // scale_output constants (src/tdd/dedisperse/dedisperse_kernel.cuh:50-58)
        uint32_t in_max = (nbits == 32) ? 0xffffffffu : ((1u << nbits) - 1u);
        float in_range = (float)in_max;
        ep.out_range = (out_nbits == 32) ? 1.f : (float)((out_nbits == 16) ? 65535u : 255u);
        ep.den = in_range * (float)m_nchans;
        ep.factor = (3.f * 1024.f) / 255.f / 16.f;
        ep.out_nbits = (unsigned)out_nbits;
    }
    // Host outputs: the channel sum runs in DM blocks of about 256 MB of output
    // (whole kernel DM blocks), each followed by an event, so the download of
    // the first blocks overlaps the sums of the later ones.
    size_t blk = dm_count;
    if (!device_ptrs && st.fast) {
        const size_t grp = (size_t)st.tiling.waves * st.tiling.dms_per_wave;
        const size_t want = std::max<size_t>(1, (size_t(256) << 20) / (nsc * out_bytes_per_sample));
        blk = std::min(dm_count, round_up(want, grp));
    }
    const size_t nblk = div_up(dm_count, blk);
    if (!device_ptrs) {
        while (st.blk_events.size() < nblk) {
            hipEvent_t e = nullptr;
            DD_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
            st.blk_events.push_back(e);
        }
    }
    ctx.timer.begin(StageTimer::DEDISPERSE, stream);
    for (size_t k = 0; k < nblk; ++k) {
        const size_t b0 = k * blk, nb = std::min(blk, dm_count - b0);
        if (st.fast) {
            // the tables of DM block b0 / (waves * dms_per_wave) start at DM b0
            const size_t b0blk = b0 / ((size_t)st.tiling.waves * st.tiling.dms_per_wave);
            DD_HIP(dd_tdd_dedisperse_lds(st.d_rows.as<uint32_t>(), st.row_words, nbits,
                                         st.d_code.as<uint32_t>() + b0 * st.nact, st.d_chan.as<int32_t>(),
                                         st.d_wstart.as<uint32_t>() + b0blk * st.nact,
                                         st.d_chunks.as<uint32_t>() + b0blk * st.maxc, st.maxc, st.nact, nb,
                                         nsc, d_out + b0 * d_out_stride, d_out_stride, ep, st.tiling, stream));
        } else {
            DD_HIP(dd_tdd_dedisperse_generic(st.d_rows.as<uint32_t>(), st.row_words, nbits,
                                             st.d_code.as<uint32_t>(), st.d_chan.as<int32_t>(),
                                             st.d_weight.as<int32_t>(), st.nact, dm_count, nsc,
                                             d_out, d_out_stride, ep, stream));
        }
        if (!device_ptrs) DD_HIP(hipEventRecord(st.blk_events[k], stream));
    }
    ctx.timer.end(StageTimer::DEDISPERSE, stream);

    // ---- output ----
    if (!device_ptrs) {
        hipStream_t d2h = ctx.d2h;
        size_t waited = 0;  // blocks the copy stream already waits for
        ctx.timer.begin(StageTimer::D2H, d2h);
        ctx.io.download(d2h, out, out_stride, d_out, d_out_stride, nsc * out_bytes_per_sample, dm_count,
                        [&](size_t rows) {
                            const size_t need = div_up(rows, blk);
                            for (; waited < need; ++waited)
                                DD_HIP(hipStreamWaitEvent(d2h, st.blk_events[waited], 0));
                        });
        ctx.timer.end(StageTimer::D2H, d2h);
    }
    if (!(device_ptrs && (flags & DEDISP_ASYNC))) {
        DD_HIP(hipStreamSynchronize(stream));
        ctx.timer.collect(ctx.profile);
        ctx.profile.work_units += (double)dm_count * (double)nsc * (double)m_nchans;
    }
}


namespace {

dd_tdd_epilogue make_epilogue(unsigned nbits, size_t out_nbits, size_t nchans)
{
    // scale_output constants (src/tdd/dedisperse/dedisperse_kernel.cuh:50-58)
    dd_tdd_epilogue ep;
    uint32_t in_max = (nbits == 32) ? 0xffffffffu : ((1u << nbits) - 1u);
    float in_range = (float)in_max;
    ep.out_range = (out_nbits == 32) ? 1.f : (float)((out_nbits == 16) ? 65535u : 255u);
    ep.den = in_range * (float)nchans;
    ep.factor = (3.f * 1024.f) / 255.f / 16.f;
    ep.out_nbits = (unsigned)out_nbits;
    return ep;
}

}  // namespace

void TDDPlan::execute_noalloc(size_type nsamps, const byte_type* in, size_type in_nbits, byte_type* out,
                              size_type out_nbits, unsigned flags, void* scratch, size_type scratch_bytes,
                              unsigned long long* scratch_tag)
{
    // execute()'s checks (src/tdd/TDDPlan.cpp:167-195) for device pointers and
    // unpadded rows
    if (!in || !out || !scratch || !scratch_tag) DD_THROW(DEDISP_INVALID_POINTER);
    if (m_dm_count == 0) DD_THROW(DEDISP_NO_DM_LIST_SET);
    if (nsamps < m_max_delay) DD_THROW(DEDISP_TOO_FEW_NSAMPS);
    if ((flags & DEDISP_ASYNC) && (flags & DEDISP_WAIT)) DD_THROW(DEDISP_INVALID_FLAG_COMBINATION);
    if (flags & DEDISP_HOST_POINTERS) DD_THROW(DEDISP_INVALID_FLAG_COMBINATION);
    if (in_nbits != 1 && in_nbits != 2 && in_nbits != 4 && in_nbits != 8 && in_nbits != 16 &&
        in_nbits != 32)
        DD_THROW(DEDISP_UNSUPPORTED_IN_NBITS);
    if (out_nbits != 8 && out_nbits != 16 && out_nbits != 32) DD_THROW(DEDISP_UNSUPPORTED_OUT_NBITS);
    if (nsamps >= (size_t(1) << 31)) DD_THROW(DEDISP_UNKNOWN_ERROR);
    const size_t nsc = nsamps - m_max_delay;
    if (nsc == 0) return;

    GPUContext& ctx = *m_ctx;
    TDDState& st = *m_state;
    ctx.make_current();
    hipStream_t stream = ctx.exec;
    const unsigned nbits = (unsigned)in_nbits;
    prepare_tables(nbits, 0, m_dm_count, /*upload=*/false);
    const size_t row_words = row_words_for(nbits, nsamps, nsc);

    // scratch layout: delay codes (64 spare entries), channels, weights,
    // window starts, then the transposed rows; 256-byte aligned pieces
    size_t used = 0;
    auto carve = [&](size_t bytes) {
        const size_t at = used;
        used += round_up(std::max<size_t>(bytes, 4), 256);
        return at;
    };
    const size_t o_code = carve((st.h_tab.size() + 64) * 4);
    const size_t o_chan = carve((st.nact + 8) * 4);
    const size_t o_weight = carve(st.nact * 4);
    const size_t o_wstart = carve((st.h_wstart.size() + 8) * 4);
    const size_t o_chunks = carve(st.h_chunks.size() * 4);
    const size_t o_rows = carve(m_nchans * row_words * 4);
    if (used > scratch_bytes) DD_THROW(DEDISP_MEM_ALLOC_FAILED);
    uint8_t* const base = static_cast<uint8_t*>(scratch);
    auto* d_code = reinterpret_cast<uint32_t*>(base + o_code);
    auto* d_chan = reinterpret_cast<int32_t*>(base + o_chan);
    auto* d_weight = reinterpret_cast<int32_t*>(base + o_weight);
    auto* d_wstart = reinterpret_cast<uint32_t*>(base + o_wstart);
    auto* d_chunks = reinterpret_cast<uint32_t*>(base + o_chunks);
    auto* d_rows = reinterpret_cast<uint32_t*>(base + o_rows);

    // tables: uploaded only when the scratch holds another plan's (or an
    // older) set; the rows that follow them are rewritten by every call
    if (*scratch_tag != st.table_uid) {
        *scratch_tag = 0;
        if (st.nact) {
            DD_HIP(hipMemcpyAsync(d_chan, st.h_chan.data(), st.nact * 4, hipMemcpyHostToDevice, stream));
            DD_HIP(hipMemcpyAsync(d_weight, st.h_weight.data(), st.nact * 4, hipMemcpyHostToDevice, stream));
        }
        if (!st.h_tab.empty())
            DD_HIP(hipMemcpyAsync(d_code, st.h_tab.data(), st.h_tab.size() * 4, hipMemcpyHostToDevice, stream));
        if (!st.h_wstart.empty())
            DD_HIP(hipMemcpyAsync(d_wstart, st.h_wstart.data(), st.h_wstart.size() * 4, hipMemcpyHostToDevice,
                                  stream));
        if (!st.h_chunks.empty())
            DD_HIP(hipMemcpyAsync(d_chunks, st.h_chunks.data(), st.h_chunks.size() * 4, hipMemcpyHostToDevice,
                                  stream));
        DD_HIP(hipStreamSynchronize(stream));
        *scratch_tag = st.table_uid;
    }

    // the transpose writes every word of every row (zero beyond the data), so
    // nothing in the scratch needs clearing
    const size_t row_bytes = (m_nchans * nbits + 7) / 8;
    DD_HIP(dd_tdd_transpose(in, nsamps, row_bytes, nbits, m_nchans, d_rows, row_words, stream));
    const size_t out_stride = nsc * (out_nbits / 8);
    const dd_tdd_epilogue ep = make_epilogue(nbits, out_nbits, m_nchans);
    if (st.fast)
        DD_HIP(dd_tdd_dedisperse_lds(d_rows, row_words, nbits, d_code, d_chan, d_wstart, d_chunks, st.maxc,
                                     st.nact, m_dm_count, nsc, out, out_stride, ep, st.tiling, stream));
    else
        DD_HIP(dd_tdd_dedisperse_generic(d_rows, row_words, nbits, d_code, d_chan, d_weight, st.nact, m_dm_count,
                                         nsc, out, out_stride, ep, stream));
    if (!(flags & DEDISP_ASYNC)) DD_HIP(hipStreamSynchronize(stream));
}

void TDDPlan::execute_batch(size_type nbeams, const size_type* nsamps, const byte_type* const* in,
                            size_type in_nbits, byte_type* const* out, size_type out_nbits,
                            unsigned flags)
{
    if (nbeams == 0) return;
    if (!nsamps || !in || !out) DD_THROW(DEDISP_INVALID_POINTER);
    if (m_dm_count == 0) DD_THROW(DEDISP_NO_DM_LIST_SET);
    if ((flags & DEDISP_ASYNC) && (flags & DEDISP_WAIT)) DD_THROW(DEDISP_INVALID_FLAG_COMBINATION);
    if ((flags & DEDISP_HOST_POINTERS) && (flags & DEDISP_DEVICE_POINTERS))
        DD_THROW(DEDISP_INVALID_FLAG_COMBINATION);
    if (in_nbits != 1 && in_nbits != 2 && in_nbits != 4 && in_nbits != 8 && in_nbits != 16 &&
        in_nbits != 32)
        DD_THROW(DEDISP_UNSUPPORTED_IN_NBITS);
    if (out_nbits != 8 && out_nbits != 16 && out_nbits != 32) DD_THROW(DEDISP_UNSUPPORTED_OUT_NBITS);
    for (size_t b = 0; b < nbeams; ++b) {
        if (!in[b] || !out[b]) DD_THROW(DEDISP_INVALID_POINTER);
        if (nsamps[b] < m_max_delay) DD_THROW(DEDISP_TOO_FEW_NSAMPS);
        if (nsamps[b] >= (size_t(1) << 31)) DD_THROW(DEDISP_UNKNOWN_ERROR);
    }
    const unsigned nbits = (unsigned)in_nbits;
    const size_t row_bytes = (m_nchans * nbits + 7) / 8;
    const size_t out_bytes = out_nbits / 8;

    GPUContext& ctx = *m_ctx;
    TDDState& st = *m_state;
    ctx.make_current();
    hipStream_t stream = ctx.exec;
    prepare_tables(nbits, 0, m_dm_count);
    if (!st.fast || st.nact == 0) {
        // weighted killmask or 16/32-bit samples: one beam after another
        for (size_t b = 0; b < nbeams; ++b)
            execute_guru(nsamps[b], in[b], in_nbits, row_bytes, out[b], out_nbits,
                         (nsamps[b] - m_max_delay) * out_bytes, 0, m_dm_count, flags);
        return;
    }
    ctx.timer.reset(ctx.profiling);
    ctx.profile = dedisp_amd_profile{};
    const bool device_ptrs = (flags & DEDISP_DEVICE_POINTERS) != 0;

    // ---- layout: rows, staged input/output and time tiles of every beam ----
    const size_t tile_samps = tile_samples(nbits, st.tiling);
    std::vector<size_t> nsc(nbeams), rw(nbeams), roff(nbeams), ioff(nbeams), ooff(nbeams), nt(nbeams);
    size_t rows_total = 0, in_total = 0, out_total = 0, ntiles = 0;
    double work = 0;
    for (size_t b = 0; b < nbeams; ++b) {
        nsc[b] = nsamps[b] - m_max_delay;
        rw[b] = row_words_for(nbits, nsamps[b], nsc[b]);
        roff[b] = rows_total;
        rows_total += m_nchans * rw[b];
        ioff[b] = in_total;
        in_total += round_up(nsamps[b] * row_bytes, 256);
        ooff[b] = out_total;
        out_total += round_up(m_dm_count * nsc[b] * out_bytes, 256);
        nt[b] = div_up(nsc[b], tile_samps);
        ntiles += nt[b];
        work += (double)m_dm_count * (double)nsc[b] * (double)m_nchans;
    }
    if (ntiles == 0) return;
    st.d_brows.reserve(rows_total * 4);
    if (!device_ptrs) {
        st.d_bin.reserve(in_total);
        st.d_bout.reserve(out_total);
    }
    uint8_t* const bout = device_ptrs ? nullptr : st.d_bout.as<uint8_t>();
    const uint8_t* const bin = device_ptrs ? nullptr : st.d_bin.as<uint8_t>();

    std::vector<uint64_t> key = {nbeams, nbits, out_nbits, m_tables_version,
                                 st.tiling.samples_per_lane, st.tiling.dms_per_wave, st.tiling.waves,
                                 st.tiling.chunk_channels, st.tiling.window_words,
                                 (uint64_t)st.d_code.ptr, (uint64_t)st.d_chan.ptr, (uint64_t)st.d_wstart.ptr,
                                 (uint64_t)st.d_chunks.ptr, st.maxc,
                                 (uint64_t)st.d_brows.ptr};
    for (size_t b = 0; b < nbeams; ++b) {
        key.push_back(nsamps[b]);
        key.push_back((uint64_t)in[b]);
        key.push_back((uint64_t)out[b]);
    }
    const dd_tdd_epilogue ep = make_epilogue(nbits, out_nbits, m_nchans);
    // device pointers: the cached capture of this batch, else the least
    // recently used slot for a new one
    TDDState::Graph* ge = nullptr;
    bool replay = false;
    if (device_ptrs) {
        for (auto& g : st.graphs)
            if (g->exec && g->key == key) {
                ge = g.get();
                replay = true;
                break;
            }
        if (!ge) {
            if (st.graphs.size() < TDDState::kGraphs) st.graphs.emplace_back(new TDDState::Graph);
            ge = std::min_element(st.graphs.begin(), st.graphs.end(),
                                  [](const auto& a, const auto& b) { return a->used < b->used; })
                     ->get();
            ge->drop();
        }
        ge->used = ++st.graph_clock;
    }
    DeviceBuffer& d_beams = ge ? ge->d_beams : st.d_beams;
    DeviceBuffer& d_tile_beam = ge ? ge->d_tile_beam : st.d_tile_beam;

    if (!replay) {
        d_beams.reserve(nbeams * sizeof(dd_tdd_beam));
        d_tile_beam.reserve(ntiles * 4);
        st.h_beams.resize(nbeams);
        st.h_tile_beam.clear();
        st.h_tile_beam.reserve(ntiles);
        size_t t0 = 0;
        for (size_t b = 0; b < nbeams; ++b) {
            dd_tdd_beam& e = st.h_beams[b];
            e.rows = st.d_brows.as<uint32_t>() + roff[b];
            e.out = device_ptrs ? out[b] : bout + ooff[b];
            e.row_words = rw[b];
            e.out_stride = nsc[b] * out_bytes;
            e.nsamps_out = (uint32_t)nsc[b];
            e.tile0 = (uint32_t)t0;
            st.h_tile_beam.insert(st.h_tile_beam.end(), nt[b], (uint32_t)b);
            t0 += nt[b];
        }
        DD_HIP(hipMemcpyAsync(d_beams.ptr, st.h_beams.data(), nbeams * sizeof(dd_tdd_beam),
                              hipMemcpyHostToDevice, stream));
        DD_HIP(hipMemcpyAsync(d_tile_beam.ptr, st.h_tile_beam.data(), ntiles * 4,
                              hipMemcpyHostToDevice, stream));
        DD_HIP(hipStreamSynchronize(stream));
    }

    // transposes of every beam, then one channel-sum launch over all tiles
    auto launches = [&](bool timed) {
        if (timed) ctx.timer.begin(StageTimer::UNPACK, stream);
        for (size_t b = 0; b < nbeams; ++b) {
            const uint8_t* src = device_ptrs ? in[b] : bin + ioff[b];
            DD_HIP(dd_tdd_transpose(src, nsamps[b], row_bytes, nbits, m_nchans,
                                    st.d_brows.as<uint32_t>() + roff[b], rw[b], stream));
        }
        if (timed) ctx.timer.end(StageTimer::UNPACK, stream);
        if (timed) ctx.timer.begin(StageTimer::DEDISPERSE, stream);
        DD_HIP(dd_tdd_dedisperse_lds_batch(d_beams.as<dd_tdd_beam>(), d_tile_beam.as<uint32_t>(),
                                           ntiles, nbits, st.d_code.as<uint32_t>(), st.d_chan.as<int32_t>(),
                                           st.d_wstart.as<uint32_t>(), st.d_chunks.as<uint32_t>(), st.maxc,
                                           st.nact, m_dm_count, ep,
                                           st.tiling, stream));
        if (timed) ctx.timer.end(StageTimer::DEDISPERSE, stream);
    };

    if (device_ptrs) {
        if (!replay) {
            // capture the kernels once (no copies or fills inside the graph)
            DD_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
            try {
                launches(false);
            } catch (...) {
                hipGraph_t g = nullptr;
                (void)hipStreamEndCapture(stream, &g);
                if (g) (void)hipGraphDestroy(g);
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
        ctx.timer.begin(StageTimer::H2D, stream);
        for (size_t b = 0; b < nbeams; ++b)
            DD_HIP(hipMemcpyAsync(st.d_bin.as<uint8_t>() + ioff[b], in[b], nsamps[b] * row_bytes,
                                  hipMemcpyHostToDevice, stream));
        ctx.timer.end(StageTimer::H2D, stream);
        launches(true);
        ctx.timer.begin(StageTimer::D2H, stream);
        for (size_t b = 0; b < nbeams; ++b)
            DD_HIP(hipMemcpyAsync(out[b], bout + ooff[b], m_dm_count * nsc[b] * out_bytes,
                                  hipMemcpyDeviceToHost, stream));
        ctx.timer.end(StageTimer::D2H, stream);
    }
    if (!(device_ptrs && (flags & DEDISP_ASYNC))) {
        DD_HIP(hipStreamSynchronize(stream));
        ctx.timer.collect(ctx.profile);
        ctx.profile.work_units = work;
    }
}

double TDDPlan::autotune(size_type nsamps, size_type in_nbits, size_type out_nbits,
                         dedisp_amd_tuning* chosen)
{
    if (m_dm_count == 0) DD_THROW(DEDISP_NO_DM_LIST_SET);
    if (nsamps <= m_max_delay) DD_THROW(DEDISP_TOO_FEW_NSAMPS);
    if (out_nbits != 8 && out_nbits != 16 && out_nbits != 32) DD_THROW(DEDISP_UNSUPPORTED_OUT_NBITS);
    const unsigned nbits = (unsigned)in_nbits;
    if (nbits != 1 && nbits != 2 && nbits != 4 && nbits != 8) DD_THROW(DEDISP_UNSUPPORTED_IN_NBITS);
    GPUContext& ctx = *m_ctx;
    TDDState& st = *m_state;
    ctx.make_current();
    // every instantiated shape with 16- and 32-channel staging chunks
    dd_tdd_tiling cand[128];
    const int nshape = std::min(dd_tdd_lds_tilings(nbits, cand, 64), 64);
    for (int i = 0; i < nshape; ++i) {
        cand[nshape + i] = cand[i];
        cand[nshape + i].chunk_channels = 32;
    }
    const int ncand = 2 * nshape;
    const size_t row_bytes = (m_nchans * nbits + 7) / 8;
    const size_t nsc = nsamps - m_max_delay, out_bytes = out_nbits / 8;
    DeviceBuffer d_in, d_out;
    d_in.reserve(nsamps * row_bytes);
    d_out.reserve(m_dm_count * nsc * out_bytes);
    DD_HIP(hipMemsetAsync(d_in.ptr, 0, nsamps * row_bytes, ctx.exec));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    DD_HIP(hipEventCreate(&e0));
    DD_HIP(hipEventCreate(&e1));
    const dd_tdd_tiling saved = st.tuned[nbits];
    double best = -1;
    dd_tdd_tiling best_t{};
    int timed = 0;
    try {
        for (int i = 0; i < ncand; ++i) {
            st.tuned[nbits] = cand[i];
            auto run = [&] {
                execute_guru(nsamps, d_in.as<uint8_t>(), in_nbits, row_bytes, d_out.as<uint8_t>(), out_nbits,
                             nsc * out_bytes, 0, m_dm_count, DEDISP_DEVICE_POINTERS);
            };
            run();  // builds the tables for this tiling
            if (!st.fast) break;  // weighted killmask: nothing to choose
            DD_HIP(hipEventRecord(e0, ctx.exec));
            run();
            DD_HIP(hipEventRecord(e1, ctx.exec));
            DD_HIP(hipEventSynchronize(e1));
            float ms = 0.f;
            DD_HIP(hipEventElapsedTime(&ms, e0, e1));
            ++timed;
            if (best < 0 || ms < best) best = ms, best_t = cand[i];
        }
    } catch (...) {
        st.tuned[nbits] = saved;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    st.tuned[nbits] = (best >= 0) ? best_t : saved;
    if (chosen) {
        const dd_tdd_tiling t = (best >= 0) ? best_t : default_tiling(nbits);
        *chosen = dedisp_amd_tuning{};
        chosen->samples_per_lane = t.samples_per_lane;
        chosen->dms_per_wave = t.dms_per_wave;
        chosen->waves = t.waves;
        chosen->chunk_channels = t.chunk_channels;
        chosen->best_ms = best;
        chosen->candidates = (unsigned)timed;
    }
    return best;
}

}  // namespace dedisp
