// Streaming front end: fixed-size chunks of many beams through one plan.
//
// No reference counterpart: this is the loop a real-time pipeline runs around
// dedisp_execute (src/clib/dedisp.h:345-350 for the output convention). Each
// beam owns two device buffers of max_delay + chunk samples (time-major, the
// plan's input layout). A push copies the chunk behind the beam's history,
// runs the plan on the whole buffer with device pointers (chunk output
// samples per DM), and copies the last max_delay samples to the head of the
// other buffer, which becomes the beam's buffer for the next push. Nothing is
// recomputed: the history is only re-read by the channel sum. A batched push
// takes one chunk of each of several beams through the plan's batched
// execution (one channel-sum launch, or one FDD transform / contraction
// sequence over the chunks of all of them).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <memory>
#include <new>
#include <vector>

#include "dedisp_amd.h"
#include "internal.hpp"

struct dedisp_amd_stream_struct {
    dedisp::GPUPlan* plan = nullptr;
    int device = 0;
    size_t nbeams = 0, chunk = 0, in_nbits = 0, out_nbits = 0;
    size_t row_bytes = 0;  // bytes per input sample (all channels)
    size_t hist = 0;       // max_delay at creation
    size_t ndm = 0;
    std::vector<dedisp::DeviceBuffer> bufs;  // 2 per beam
    std::vector<int> cur;                    // which of the two is live, per beam
    dedisp::DeviceBuffer d_out;              // output staging for host outputs
    std::vector<double> lat_ms;
    double busy_s = 0;
    double samples = 0;  // input samples pushed (all beams)
    int fill = 0;        // byte value of the samples before a beam's first push
};

namespace {

size_t stream_bytes(const dedisp::Plan& p, size_t nbeams, size_t chunk, size_t in_nbits,
                    size_t out_nbits)
{
    const size_t row = p.get_channel_count() * in_nbits / 8;
    const size_t hist = p.get_max_delay();
    return nbeams * 2 * (hist + chunk) * row + p.get_dm_count() * chunk * (out_nbits / 8);
}

template <typename F>
dedisp_error guarded(F&& f)
{
    try {
        f();
    } catch (const dedisp::Error& e) {
        return e.code();
    } catch (const std::bad_alloc&) {
        return DEDISP_MEM_ALLOC_FAILED;
    } catch (...) {
        return DEDISP_UNKNOWN_ERROR;
    }
    return DEDISP_NO_ERROR;
}

// The byte the history starts as: 0 for TDD (its outputs at t >= 0 never
// read the history); for FDD the sample value nearest the zero level
// (2^nbits - 1)/2 of every packed sample, since the history enters each
// segment's transform and a constant offset there would raise the rounding
// error of the outputs that follow.
int history_fill(const dedisp::GPUPlan& p, size_t in_nbits)
{
    if (!dynamic_cast<const dedisp::FDDGPUPlan*>(&p)) return 0;
    switch (in_nbits) {
    case 1: return 0xff;  // value 1 (zero level 0.5)
    case 2: return 0xaa;  // value 2 (zero level 1.5)
    case 4: return 0x88;  // value 8 (zero level 7.5)
    default: return 0x80; // value 128 (zero level 127.5)
    }
}

// A push writes chunk output samples per DM, which holds while the plan's DM
// list is the one the stream was created with and an FDD plan still writes
// time series (Fourier output is N floats per DM and would overrun the
// stream's output rows).
void check_plan_unchanged(const dedisp_amd_stream_struct& s)
{
    if (s.plan->get_max_delay() != s.hist || s.plan->get_dm_count() != s.ndm)
        DD_THROW(DEDISP_INVALID_PLAN);
    if (auto* f = dynamic_cast<const dedisp::FDDGPUPlan*>(s.plan))
        if (f->get_output_mode() != DEDISP_AMD_FDD_TIMESERIES) DD_THROW(DEDISP_INVALID_PLAN);
}

}  // namespace

extern "C" {

dedisp_size dedisp_amd_stream_bytes(const dedisp_plan plan, dedisp_size nbeams, dedisp_size chunk_nsamps,
                                    dedisp_size in_nbits, dedisp_size out_nbits)
{
    if (!plan || !plan->ptr) return 0;
    return stream_bytes(*plan->ptr, nbeams, chunk_nsamps, in_nbits, out_nbits);
}

dedisp_error dedisp_amd_stream_create(dedisp_amd_stream* stream, dedisp_plan plan, dedisp_size nbeams,
                                      dedisp_size chunk_nsamps, dedisp_size in_nbits, dedisp_size out_nbits)
{
    if (!stream) return DEDISP_INVALID_POINTER;
    *stream = nullptr;
    if (!plan || !plan->ptr) return DEDISP_INVALID_PLAN;
    auto* gp = dynamic_cast<dedisp::GPUPlan*>(plan->ptr.get());
    if (!gp) return DEDISP_INVALID_PLAN;
    if (gp->get_dm_count() == 0) return DEDISP_NO_DM_LIST_SET;
    if (nbeams == 0 || chunk_nsamps == 0) return DEDISP_INVALID_POINTER;
    if (in_nbits != 1 && in_nbits != 2 && in_nbits != 4 && in_nbits != 8 && in_nbits != 16 && in_nbits != 32)
        return DEDISP_UNSUPPORTED_IN_NBITS;
    if (out_nbits != 8 && out_nbits != 16 && out_nbits != 32) return DEDISP_UNSUPPORTED_OUT_NBITS;
    if (gp->get_channel_count() * in_nbits % 8) return DEDISP_UNSUPPORTED_IN_NBITS;

    auto s = std::unique_ptr<dedisp_amd_stream_struct>(new (std::nothrow) dedisp_amd_stream_struct());
    if (!s) return DEDISP_MEM_ALLOC_FAILED;
    dedisp_error err = guarded([&] {
        if (auto* fdd = dynamic_cast<dedisp::FDDGPUPlan*>(gp)) fdd->set_output_mode(DEDISP_AMD_FDD_TIMESERIES);
        dedisp::GPUContext& ctx = gp->context();
        ctx.make_current();
        // admission: the stream's buffers against free HBM less a 10 % reserve
        size_t free_b = 0, total_b = 0;
        DD_HIP(hipMemGetInfo(&free_b, &total_b));
        const size_t need = stream_bytes(*gp, nbeams, chunk_nsamps, in_nbits, out_nbits);
        if (need > free_b - free_b / 10) DD_THROW(DEDISP_MEM_ALLOC_FAILED);

        s->plan = gp;
        s->device = gp->get_device();
        s->nbeams = nbeams;
        s->chunk = chunk_nsamps;
        s->in_nbits = in_nbits;
        s->out_nbits = out_nbits;
        s->row_bytes = gp->get_channel_count() * in_nbits / 8;
        s->hist = gp->get_max_delay();
        s->ndm = gp->get_dm_count();
        s->bufs = std::vector<dedisp::DeviceBuffer>(2 * nbeams);
        s->cur.assign(nbeams, 0);
        s->fill = history_fill(*gp, in_nbits);
        const size_t bytes = (s->hist + s->chunk) * s->row_bytes;
        for (auto& b : s->bufs) {
            b.reserve(bytes);
            DD_HIP(hipMemsetAsync(b.ptr, s->fill, bytes, ctx.exec));
        }
        DD_HIP(hipStreamSynchronize(ctx.exec));
    });
    if (err != DEDISP_NO_ERROR) return err;
    *stream = s.release();
    return DEDISP_NO_ERROR;
}

dedisp_error dedisp_amd_stream_push(dedisp_amd_stream s, dedisp_size beam, const dedisp_byte* chunk,
                                    dedisp_byte* out, unsigned flags)
{
    if (!s) return DEDISP_INVALID_PLAN;
    if (!chunk || !out) return DEDISP_INVALID_POINTER;
    if (beam >= s->nbeams) return DEDISP_INVALID_POINTER;
    return guarded([&] {
        check_plan_unchanged(*s);
        const auto t0 = std::chrono::steady_clock::now();
        dedisp::GPUContext& ctx = s->plan->context();
        ctx.make_current();
        const bool dev = (flags & DEDISP_DEVICE_POINTERS) != 0;
        dedisp::DeviceBuffer& a = s->bufs[2 * beam + s->cur[beam]];
        dedisp::DeviceBuffer& b = s->bufs[2 * beam + 1 - s->cur[beam]];
        auto* buf = a.as<dedisp_byte>();
        const size_t hist_bytes = s->hist * s->row_bytes, chunk_bytes = s->chunk * s->row_bytes;
        DD_HIP(hipMemcpyAsync(buf + hist_bytes, chunk, chunk_bytes,
                              dev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, ctx.exec));
        // next history: the last max_delay samples of [history | chunk]
        if (s->hist)
            DD_HIP(hipMemcpyAsync(b.ptr, buf + chunk_bytes, hist_bytes, hipMemcpyDeviceToDevice, ctx.exec));
        dedisp_byte* dst = out;
        const size_t out_bytes = s->ndm * s->chunk * (s->out_nbits / 8);
        if (!dev) {
            s->d_out.reserve(out_bytes);
            dst = s->d_out.as<dedisp_byte>();
        }
        s->plan->execute(s->hist + s->chunk, buf, s->in_nbits, dst, s->out_nbits, DEDISP_DEVICE_POINTERS);
        if (!dev) DD_HIP(hipMemcpyAsync(out, dst, out_bytes, hipMemcpyDeviceToHost, ctx.exec));
        DD_HIP(hipStreamSynchronize(ctx.exec));
        s->cur[beam] ^= 1;
        const double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        s->lat_ms.push_back(dt * 1e3);
        s->busy_s += dt;
        s->samples += (double)s->chunk;
    });
}

dedisp_error dedisp_amd_stream_push_batch(dedisp_amd_stream s, dedisp_size n, const dedisp_size* beams,
                                          const dedisp_byte* const* chunks, dedisp_byte* const* outs,
                                          unsigned flags)
{
    if (!s) return DEDISP_INVALID_PLAN;
    if (n == 0) return DEDISP_NO_ERROR;
    if (!beams || !chunks || !outs) return DEDISP_INVALID_POINTER;
    std::vector<char> seen(s->nbeams, 0);
    for (size_t i = 0; i < n; ++i) {
        if (!chunks[i] || !outs[i] || beams[i] >= s->nbeams || seen[beams[i]]) return DEDISP_INVALID_POINTER;
        seen[beams[i]] = 1;
    }
    return guarded([&] {
        check_plan_unchanged(*s);
        const auto t0 = std::chrono::steady_clock::now();
        dedisp::GPUContext& ctx = s->plan->context();
        ctx.make_current();
        const bool dev = (flags & DEDISP_DEVICE_POINTERS) != 0;
        const size_t hist_bytes = s->hist * s->row_bytes, chunk_bytes = s->chunk * s->row_bytes;
        const size_t out_bytes = s->ndm * s->chunk * (s->out_nbits / 8);
        std::vector<dedisp_size> nsamps(n, s->hist + s->chunk);
        std::vector<const dedisp_byte*> in(n);
        std::vector<dedisp_byte*> dst(n);
        if (!dev) s->d_out.reserve(n * out_bytes);
        for (size_t i = 0; i < n; ++i) {
            const size_t beam = beams[i];
            dedisp::DeviceBuffer& a = s->bufs[2 * beam + s->cur[beam]];
            dedisp::DeviceBuffer& b = s->bufs[2 * beam + 1 - s->cur[beam]];
            auto* buf = a.as<dedisp_byte>();
            DD_HIP(hipMemcpyAsync(buf + hist_bytes, chunks[i], chunk_bytes,
                                  dev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, ctx.exec));
            if (s->hist)
                DD_HIP(hipMemcpyAsync(b.ptr, buf + chunk_bytes, hist_bytes, hipMemcpyDeviceToDevice, ctx.exec));
            in[i] = buf;
            dst[i] = dev ? outs[i] : s->d_out.as<dedisp_byte>() + i * out_bytes;
        }
        // one batched execution: a channel-sum launch over every beam's tiles
        // (TDD) or one chunk-concatenated transform/contraction sequence (FDD)
        if (auto* t = dynamic_cast<dedisp::TDDPlan*>(s->plan))
            t->execute_batch(n, nsamps.data(), in.data(), s->in_nbits, dst.data(), s->out_nbits,
                             DEDISP_DEVICE_POINTERS);
        else if (auto* f = dynamic_cast<dedisp::FDDGPUPlan*>(s->plan))
            f->execute_batch(n, nsamps.data(), in.data(), s->in_nbits, dst.data(), s->out_nbits,
                             DEDISP_DEVICE_POINTERS);
        else
            for (size_t i = 0; i < n; ++i)
                s->plan->execute(nsamps[i], in[i], s->in_nbits, dst[i], s->out_nbits, DEDISP_DEVICE_POINTERS);
        if (!dev)
            for (size_t i = 0; i < n; ++i)
                DD_HIP(hipMemcpyAsync(outs[i], dst[i], out_bytes, hipMemcpyDeviceToHost, ctx.exec));
        DD_HIP(hipStreamSynchronize(ctx.exec));
        for (size_t i = 0; i < n; ++i) s->cur[beams[i]] ^= 1;
        const double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        s->lat_ms.push_back(dt * 1e3);
        s->busy_s += dt;
        s->samples += (double)n * (double)s->chunk;
    });
}

dedisp_error dedisp_amd_stream_reset(dedisp_amd_stream s)
{
    if (!s) return DEDISP_INVALID_PLAN;
    return guarded([&] {
        dedisp::GPUContext& ctx = s->plan->context();
        ctx.make_current();
        for (auto& b : s->bufs) DD_HIP(hipMemsetAsync(b.ptr, s->fill, b.bytes, ctx.exec));
        DD_HIP(hipStreamSynchronize(ctx.exec));
        std::fill(s->cur.begin(), s->cur.end(), 0);
        s->lat_ms.clear();
        s->busy_s = 0;
        s->samples = 0;
    });
}

dedisp_error dedisp_amd_stream_get_stats(const dedisp_amd_stream s, dedisp_amd_stream_stats* st)
{
    if (!s) return DEDISP_INVALID_PLAN;
    if (!st) return DEDISP_INVALID_POINTER;
    *st = dedisp_amd_stream_stats{};
    st->pushes = s->lat_ms.size();
    if (s->lat_ms.empty()) return DEDISP_NO_ERROR;
    std::vector<double> v = s->lat_ms;
    std::sort(v.begin(), v.end());
    auto pct = [&](double q) {  // nearest-rank percentile
        size_t r = (size_t)std::ceil(q * (double)v.size());
        return v[std::min(v.size() - 1, r ? r - 1 : 0)];
    };
    double sum = 0;
    for (double x : v) sum += x;
    st->mean_ms = sum / (double)v.size();
    st->p50_ms = pct(0.50);
    st->p99_ms = pct(0.99);
    st->max_ms = v.back();
    st->samples_per_s = s->busy_s > 0 ? s->samples / s->busy_s : 0;
    return DEDISP_NO_ERROR;
}

dedisp_size dedisp_amd_stream_history(const dedisp_amd_stream s) { return s ? s->hist : 0; }

void dedisp_amd_stream_destroy(dedisp_amd_stream s)
{
    if (!s) return;
    // the plan is not touched: it may already be gone
    if (hipSetDevice(s->device) == hipSuccess) (void)hipDeviceSynchronize();
    delete s;
}

}  // extern "C"
