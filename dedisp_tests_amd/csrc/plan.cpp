// Host plan core: delay table, DM list, max delay, killmask (Plan), and the
// device context shared by the GPU plans (GPUPlan).
//
// Compiled with -ffp-contract=off: the float/double operation order of the
// reference host code (src/Plan.cpp) is reproduced exactly, so delay tables,
// DM lists and max_delay are bit-identical to the reference's.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <sstream>

#include "internal.hpp"

namespace dedisp {

// ---------------------------------------------------------------------------
// Errors
// ---------------------------------------------------------------------------

// Message format of _throw_error (src/common/dedisp_error.hpp:60-76).
void throw_code(dedisp_error code, const char* file, int line)
{
    std::stringstream message;
    message << "An error occurred within dedisp on line " << line << " of " << file
            << ": " << dedisp_get_error_string(code);
    throw Error(code, message.str());
}

void check_hip(hipError_t err, const char* what, const char* file, int line)
{
    if (err == hipSuccess) return;
    dedisp_error code;
    switch (err) {
    case hipErrorOutOfMemory:
        code = DEDISP_MEM_ALLOC_FAILED; break;
    case hipErrorInvalidDevice:
    case hipErrorNoDevice:
        code = DEDISP_INVALID_DEVICE_INDEX; break;
    default:
        code = (std::strstr(what, "Memcpy") != nullptr) ? DEDISP_MEM_COPY_FAILED
                                                        : DEDISP_INTERNAL_GPU_ERROR;
    }
    (void)hipGetLastError();  // the failure is reported here, not to the next call
    std::stringstream message;
    message << "An error occurred within dedisp on line " << line << " of " << file
            << ": " << dedisp_get_error_string(code) << " (" << what << ": "
            << hipGetErrorString(err) << ")";
    throw Error(code, message.str());
}

// ---------------------------------------------------------------------------
// Buffers
// ---------------------------------------------------------------------------

DeviceBuffer::~DeviceBuffer() { release(); }

void DeviceBuffer::release()
{
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
}

void DeviceBuffer::reserve(size_t n)
{
    if (n <= bytes && ptr) return;
    release();
    if (n == 0) n = 256;
    DD_HIP(hipMalloc(&ptr, n));
    bytes = n;
}

PinnedBuffer::~PinnedBuffer() { release(); }

void PinnedBuffer::release()
{
    if (ptr) (void)hipHostFree(ptr);
    ptr = nullptr;
    bytes = 0;
}

void PinnedBuffer::reserve(size_t n)
{
    if (n <= bytes && ptr) return;
    release();
    if (n == 0) n = 256;
    DD_HIP(hipHostMalloc(&ptr, n, hipHostMallocDefault));
    bytes = n;
}

// ---------------------------------------------------------------------------
// Stage timer
// ---------------------------------------------------------------------------

StageTimer::~StageTimer()
{
    for (hipEvent_t e : m_pool) (void)hipEventDestroy(e);
}

hipEvent_t StageTimer::get_event()
{
    if (m_pool_used == m_pool.size()) {
        hipEvent_t e;
        DD_HIP(hipEventCreate(&e));
        m_pool.push_back(e);
    }
    return m_pool[m_pool_used++];
}

void StageTimer::reset(bool enabled)
{
    m_enabled = enabled;
    m_marks.clear();
    m_pool_used = 0;
    for (int i = 0; i < NSTAGES; ++i) m_open[i] = false;
}

void StageTimer::begin(Stage s, hipStream_t stream)
{
    if (!m_enabled) return;
    hipEvent_t e = get_event();
    DD_HIP(hipEventRecord(e, stream));
    m_open[s] = true;
    m_open_start[s] = e;
}

void StageTimer::end(Stage s, hipStream_t stream)
{
    if (!m_enabled || !m_open[s]) return;
    hipEvent_t e = get_event();
    DD_HIP(hipEventRecord(e, stream));
    m_marks.push_back({s, m_open_start[s], e});
    m_open[s] = false;
}

void StageTimer::collect(dedisp_amd_profile& out)
{
    if (!m_enabled) return;
    double* ms[NSTAGES] = {&out.h2d_ms, &out.unpack_ms, &out.fft_ms,
                           &out.dedisperse_ms, &out.ifft_ms, &out.d2h_ms};
    unsigned long* cnt[NSTAGES] = {nullptr, &out.unpack_launches, &out.fft_launches,
                                   &out.dedisperse_launches, &out.ifft_launches, nullptr};
    hipEvent_t first = nullptr, last = nullptr;
    for (const Mark& m : m_marks) {
        DD_HIP(hipEventSynchronize(m.stop));
        float t = 0.f;
        DD_HIP(hipEventElapsedTime(&t, m.start, m.stop));
        *ms[m.stage] += t;
        if (cnt[m.stage]) *cnt[m.stage] += 1;
        if (!first) first = m.start;
        last = m.stop;
    }
    if (first && last) {
        float t = 0.f;
        if (hipEventElapsedTime(&t, first, last) == hipSuccess) out.total_ms = t;
        else (void)hipGetLastError();
    }
    m_marks.clear();
    m_pool_used = 0;
}

// ---------------------------------------------------------------------------
// Host math
// ---------------------------------------------------------------------------

uint32_t tdd_delay(float dm, float delay)
{
    float p = dm * delay;
    if (!(p > 0.f)) return 0u;
    float r = std::nearbyint(p);  // round to nearest, ties to even
    if (r >= 4294967296.0f) return 0xffffffffu;
    return (uint32_t)r;
}

int32_t fdd_delay(float dm, float delay)
{
    float p = dm * delay;
    return (int32_t)p;  // truncation toward zero
}

// Plan::generate_delay_table (src/Plan.cpp:109-119) with df already negative.
void make_delay_table(dedisp_size nchans, dedisp_float dt, dedisp_float f0,
                      dedisp_float df, dedisp_float* out)
{
    for (dedisp_size c = 0; c < nchans; ++c) {
        dedisp_float a = 1.f / (f0 + c * df);
        dedisp_float b = 1.f / f0;
        out[c] = 4.15e3 / dt * (a * a - b * b);
    }
}

// ---------------------------------------------------------------------------
// Plan (src/Plan.cpp:16-154)
// ---------------------------------------------------------------------------

Plan::Plan(size_type nchans, float_type dt, float_type f0, float_type df)
{
    m_dm_count = 0;
    m_nchans = nchans;
    m_max_delay = 0;
    m_dt = dt;
    m_f0 = f0;
    m_df = df;
    // freq[c] = f0 + c*df with df forced negative (src/Plan.cpp:29-31)
    df = -std::abs(df);
    h_delay_table.resize(nchans);
    make_delay_table(nchans, dt, f0, df, h_delay_table.data());
    h_killmask.assign(nchans, (dedisp_bool)1);
}

Plan::~Plan() {}

void Plan::generate_dm_list(std::vector<dedisp_float>& dm_table,
                            dedisp_float dm_start, dedisp_float dm_end,
                            double dt, double ti, double f0, double df,
                            dedisp_size nchans, double tol)
{
    // Lina Levin's recurrence in double precision (src/Plan.cpp:45-70).
    dt *= 1e6;
    double f    = (f0 + ((nchans / 2) - 0.5) * df) * 1e-3;
    double tol2 = tol * tol;
    double a    = 8.3 * df / (f * f * f);
    double a2   = a * a;
    double b2   = a2 * (double)(nchans * nchans / 16.0);
    double c    = (dt * dt + ti * ti) * (tol2 - 1.0);

    dm_table.push_back(dm_start);
    while (dm_table.back() < dm_end) {
        double prev  = dm_table.back();
        double prev2 = prev * prev;
        double k     = c + tol2 * a2 * prev2;
        double dm    = ((b2 * prev + std::sqrt(-a2 * b2 * prev2 + (a2 + b2) * k)) / (a2 + b2));
        // A step that does not advance the float value (or a NaN) would make
        // the reference loop forever; stop with the list as it stands.
        if (!((dedisp_float)dm > dm_table.back())) break;
        dm_table.push_back((dedisp_float)dm);
    }
}

void Plan::generate_dm_list(float_type dm_start, float_type dm_end,
                            float_type ti, float_type tol)
{
    h_dm_list.clear();
    generate_dm_list(h_dm_list, dm_start, dm_end, m_dt, ti, m_f0, m_df, m_nchans, tol);
    m_dm_count = h_dm_list.size();
    m_max_delay = dedisp_size(h_dm_list[m_dm_count - 1] * h_delay_table[m_nchans - 1] + 0.5);
    ++m_tables_version;
}

void Plan::set_dm_list(const float_type* dm_list, size_type count)
{
    if (!dm_list) DD_THROW(DEDISP_INVALID_POINTER);
    if (count == 0) DD_THROW(DEDISP_NO_DM_LIST_SET);
    m_dm_count = count;
    h_dm_list.assign(dm_list, dm_list + count);
    m_max_delay = dedisp_size(h_dm_list[m_dm_count - 1] * h_delay_table[m_nchans - 1] + 0.5);
    ++m_tables_version;
}

void Plan::set_killmask(const bool_type* killmask)
{
    if (killmask) h_killmask.assign(killmask, killmask + m_nchans);
    else std::fill(h_killmask.begin(), h_killmask.end(), (dedisp_bool)1);
    ++m_tables_version;
}

void Plan::sync()
{
    DD_HIP(hipDeviceSynchronize());
}

// ---------------------------------------------------------------------------
// GPUPlan (src/GPUPlan.cpp:12-82)
// ---------------------------------------------------------------------------

void GPUContext::make_current() const
{
    DD_HIP(hipSetDevice(device));
}

GPUPlan::GPUPlan(size_type nchans, float_type dt, float_type f0, float_type df,
                 int device_idx)
    : Plan(nchans, dt, f0, df), m_ctx(new GPUContext)
{
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
        (void)hipGetLastError();
        DD_THROW(DEDISP_INVALID_DEVICE_INDEX);
    }
    if (device_idx < 0 || device_idx >= count) DD_THROW(DEDISP_INVALID_DEVICE_INDEX);
    m_ctx->device = device_idx;
    m_ctx->make_current();
    // Default-priority, non-blocking streams.
    DD_HIP(hipStreamCreateWithFlags(&m_ctx->h2d, hipStreamNonBlocking));
    DD_HIP(hipStreamCreateWithFlags(&m_ctx->exec, hipStreamNonBlocking));
    DD_HIP(hipStreamCreateWithFlags(&m_ctx->d2h, hipStreamNonBlocking));
    DD_HIP(hipEventCreateWithFlags(&m_ctx->ev_a, hipEventDisableTiming));
    DD_HIP(hipEventCreateWithFlags(&m_ctx->ev_b, hipEventDisableTiming));

    m_ctx->d_delay_table.reserve(nchans * sizeof(dedisp_float));
    DD_HIP(hipMemcpyAsync(m_ctx->d_delay_table.ptr, h_delay_table.data(),
                          nchans * sizeof(dedisp_float), hipMemcpyHostToDevice, m_ctx->h2d));
    m_ctx->d_killmask.reserve(nchans * sizeof(dedisp_bool));
    DD_HIP(hipMemcpyAsync(m_ctx->d_killmask.ptr, h_killmask.data(),
                          nchans * sizeof(dedisp_bool), hipMemcpyHostToDevice, m_ctx->h2d));
    DD_HIP(hipStreamSynchronize(m_ctx->h2d));
}

GPUPlan::~GPUPlan()
{
    if (!m_ctx) return;
    (void)hipSetDevice(m_ctx->device);
    if (m_ctx->h2d) (void)hipStreamSynchronize(m_ctx->h2d);
    if (m_ctx->exec) (void)hipStreamSynchronize(m_ctx->exec);
    if (m_ctx->d2h) (void)hipStreamSynchronize(m_ctx->d2h);
    m_ctx->d_dm_list.release();
    m_ctx->d_delay_table.release();
    m_ctx->d_killmask.release();
    if (m_ctx->ev_a) (void)hipEventDestroy(m_ctx->ev_a);
    if (m_ctx->ev_b) (void)hipEventDestroy(m_ctx->ev_b);
    if (m_ctx->h2d) (void)hipStreamDestroy(m_ctx->h2d);
    if (m_ctx->exec) (void)hipStreamDestroy(m_ctx->exec);
    if (m_ctx->d2h) (void)hipStreamDestroy(m_ctx->d2h);
}

int GPUPlan::get_device() const { return m_ctx->device; }

void GPUPlan::generate_dm_list(float_type dm_start, float_type dm_end,
                               float_type ti, float_type tol)
{
    Plan::generate_dm_list(dm_start, dm_end, ti, tol);
    m_ctx->make_current();
    m_ctx->d_dm_list.reserve(m_dm_count * sizeof(dedisp_float));
    DD_HIP(hipMemcpyAsync(m_ctx->d_dm_list.ptr, h_dm_list.data(),
                          m_dm_count * sizeof(dedisp_float), hipMemcpyHostToDevice, m_ctx->h2d));
    DD_HIP(hipStreamSynchronize(m_ctx->h2d));
}

void GPUPlan::set_dm_list(const float_type* dm_list, size_type count)
{
    Plan::set_dm_list(dm_list, count);
    m_ctx->make_current();
    m_ctx->d_dm_list.reserve(m_dm_count * sizeof(dedisp_float));
    DD_HIP(hipMemcpyAsync(m_ctx->d_dm_list.ptr, h_dm_list.data(),
                          m_dm_count * sizeof(dedisp_float), hipMemcpyHostToDevice, m_ctx->h2d));
    DD_HIP(hipStreamSynchronize(m_ctx->h2d));
}

void GPUPlan::set_killmask(const bool_type* killmask)
{
    Plan::set_killmask(killmask);
    m_ctx->make_current();
    DD_HIP(hipMemcpyAsync(m_ctx->d_killmask.ptr, h_killmask.data(),
                          m_nchans * sizeof(dedisp_bool), hipMemcpyHostToDevice, m_ctx->h2d));
    DD_HIP(hipStreamSynchronize(m_ctx->h2d));
}

void GPUPlan::set_profiling(bool enable) { m_ctx->profiling = enable; }

dedisp_amd_profile GPUPlan::get_profile() const { return m_ctx->profile; }

}  // namespace dedisp
