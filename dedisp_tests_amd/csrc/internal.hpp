// Internal declarations shared by the host sources of libdedisp.so.
// Host code only: kernels live in *.hip files and are reached through the
// launcher functions declared in kernels.h.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstddef>
#include <cstdint>
#include <functional>
#include <string>
#include <vector>

#include "dedisp.hpp"

namespace dedisp {

[[noreturn]] void throw_code(dedisp_error code, const char* file, int line);
#define DD_THROW(code) ::dedisp::throw_code((code), __FILE__, __LINE__)

// Maps a failed HIP call to a dedisp error and throws.
void check_hip(hipError_t err, const char* what, const char* file, int line);
#define DD_HIP(call) ::dedisp::check_hip((call), #call, __FILE__, __LINE__)

// Device allocation that grows on demand and is kept between executes.
struct DeviceBuffer {
    void*  ptr = nullptr;
    size_t bytes = 0;
    DeviceBuffer() = default;
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    ~DeviceBuffer();
    void reserve(size_t n);   // ensures at least n bytes (contents not kept)
    void release();
    template <typename T> T* as() const { return static_cast<T*>(ptr); }
};

// Page-locked host allocation (hipHostMalloc), grows on demand.
struct PinnedBuffer {
    void*  ptr = nullptr;
    size_t bytes = 0;
    PinnedBuffer() = default;
    PinnedBuffer(const PinnedBuffer&) = delete;
    PinnedBuffer& operator=(const PinnedBuffer&) = delete;
    ~PinnedBuffer();
    void reserve(size_t n);
    void release();
};

// Transfers between the caller's pageable host memory and the device through
// two page-locked slots (host copies on a worker pool overlap the DMA).
class HostIO {
public:
    HostIO() = default;
    HostIO(const HostIO&) = delete;
    HostIO& operator=(const HostIO&) = delete;
    ~HostIO();
    // rows x row_bytes from host src (src_stride apart) to device dst
    // (dst_stride apart), queued on stream; returns when src may be reused
    void upload(hipStream_t stream, uint8_t* dst, size_t dst_stride, const uint8_t* src, size_t src_stride,
                size_t row_bytes, size_t rows);
    // rows x row_bytes from device src to host dst; before the copy of rows
    // below r is queued on stream, ready(r) is called (to make stream wait for
    // the work that writes them). Returns when dst holds every row.
    void download(hipStream_t stream, uint8_t* dst, size_t dst_stride, const uint8_t* src, size_t src_stride,
                  size_t row_bytes, size_t rows, const std::function<void(size_t)>& ready);

private:
    void prepare();
    PinnedBuffer m_slot[2];
    hipEvent_t m_ev[2] = {nullptr, nullptr};
};

// Records event pairs around launches while profiling is on and sums them
// into a dedisp_amd_profile after the stream has drained.
class StageTimer {
public:
    enum Stage { H2D = 0, UNPACK, FFT, DEDISPERSE, IFFT, D2H, NSTAGES };
    ~StageTimer();
    void reset(bool enabled);
    void begin(Stage s, hipStream_t stream);
    void end(Stage s, hipStream_t stream);
    // Waits for every recorded event and accumulates the stage times.
    void collect(dedisp_amd_profile& out);
    bool enabled() const { return m_enabled; }
private:
    struct Mark { Stage stage; hipEvent_t start; hipEvent_t stop; };
    hipEvent_t get_event();
    bool m_enabled = false;
    std::vector<Mark> m_marks;
    std::vector<hipEvent_t> m_pool;
    size_t m_pool_used = 0;
    bool m_open[NSTAGES] = {};
    hipEvent_t m_open_start[NSTAGES] = {};
};

struct GPUContext {
    int device = 0;
    hipStream_t h2d = nullptr;   // host -> device copies
    hipStream_t exec = nullptr;  // kernels
    hipStream_t d2h = nullptr;   // device -> host copies
    hipEvent_t  ev_a = nullptr, ev_b = nullptr;  // cross-stream ordering
    DeviceBuffer d_dm_list;      // float[ndm]
    DeviceBuffer d_delay_table;  // float[nchans]
    DeviceBuffer d_killmask;     // int[nchans]
    HostIO io;                   // host-pointer transfers
    bool profiling = false;
    StageTimer timer;
    dedisp_amd_profile profile{};
    void make_current() const;
};

// Largest channel counts the reference accepts (TDD: constant memory / 8,
// src/tdd/TDDPlan.cpp:76-82 on a 64 KiB device; FDD: the 16384-entry constant
// table of src/fdd/dedisperse/fdd_kernel.cuh:7).
constexpr dedisp_size kTddMaxChans = 8192;
constexpr dedisp_size kFddMaxChans = 16384;

// Rounding helpers.
inline size_t div_up(size_t a, size_t b) { return (a + b - 1) / b; }
inline size_t round_up(size_t a, size_t b) { return div_up(a, b) * b; }

// Host copy of the TDD integer delay: __float2uint_rn(dm * delay)
// (src/tdd/dedisperse/dedisperse_kernel.cuh:163-165).
uint32_t tdd_delay(float dm, float delay);
// Host copy of the FDD integer delay: (int)(dm * delay)
// (src/fdd/dedisperse/fdd_kernel.cuh:226-229).
int32_t fdd_delay(float dm, float delay);

// Delay table of Plan.cpp:109-119 (host float/double order).
void make_delay_table(dedisp_size nchans, dedisp_float dt, dedisp_float f0,
                      dedisp_float df, dedisp_float* out);

}  // namespace dedisp

// The object behind a C-API dedisp_plan handle.
struct dedisp_plan_struct {
    std::unique_ptr<dedisp::Plan> ptr;
    dedisp_implementation impl = DEDISP_TDD;
};

