// Host <-> device transfers of host-pointer executes.
//
// The reference stages its input gulps and output blocks through page-locked
// double buffers with a host output thread (src/tdd/TDDPlan.cpp:237-447). The
// caller's memory is ordinary pageable memory that is never registered, so
// every byte goes through two page-locked slots here too: the host copies
// rows into (or out of) one slot, with a small worker pool, while the DMA
// engine moves the other. Downloads wait per piece for the kernels that wrote
// it, so the D2H of the first DM rows overlaps the channel sums of the later
// ones.
#include <algorithm>
#include <condition_variable>
#include <cstring>
#include <mutex>
#include <thread>

#include "internal.hpp"

namespace dedisp {

namespace {

// Bytes of one page-locked slot.
constexpr size_t kSlotBytes = size_t(64) << 20;

// A fixed pool of copy workers; run() splits a job over the workers and the
// calling thread and returns when every part is done.
class CopyPool {
public:
    explicit CopyPool(unsigned n)
    {
        for (unsigned i = 0; i < n; ++i) m_threads.emplace_back([this] { worker(); });
    }
    ~CopyPool()
    {
        {
            std::lock_guard<std::mutex> lk(m_mutex);
            m_stop = true;
        }
        m_wake.notify_all();
        for (auto& t : m_threads) t.join();
    }
    void run(size_t nparts, const std::function<void(size_t)>& f)
    {
        if (m_threads.empty() || nparts <= 1) {
            for (size_t i = 0; i < nparts; ++i) f(i);
            return;
        }
        std::lock_guard<std::mutex> serial(m_run);  // one job at a time
        {
            std::lock_guard<std::mutex> lk(m_mutex);
            m_job = &f;
            m_nparts = nparts;
            m_next = 0;
            m_active = m_threads.size();
            ++m_gen;
        }
        m_wake.notify_all();
        work();
        std::unique_lock<std::mutex> lk(m_mutex);
        m_done.wait(lk, [this] { return m_active == 0; });
        m_job = nullptr;
    }
    size_t width() const { return m_threads.size() + 1; }

private:
    void work()
    {
        for (;;) {
            size_t i;
            {
                std::lock_guard<std::mutex> lk(m_mutex);
                if (m_next >= m_nparts) return;
                i = m_next++;
            }
            (*m_job)(i);
        }
    }
    void worker()
    {
        uint64_t seen = 0;
        for (;;) {
            {
                std::unique_lock<std::mutex> lk(m_mutex);
                m_wake.wait(lk, [&] { return m_stop || m_gen != seen; });
                if (m_stop) return;
                seen = m_gen;
            }
            work();
            std::lock_guard<std::mutex> lk(m_mutex);
            if (--m_active == 0) m_done.notify_all();
        }
    }
    std::vector<std::thread> m_threads;
    std::mutex m_mutex, m_run;
    std::condition_variable m_wake, m_done;
    const std::function<void(size_t)>* m_job = nullptr;
    size_t m_nparts = 0, m_next = 0, m_active = 0;
    uint64_t m_gen = 0;
    bool m_stop = false;
};

CopyPool& pool()
{
    static CopyPool p([] {
        const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
        return std::min(7u, hw > 1 ? hw - 1 : 0u);
    }());
    return p;
}

// Copies nrows rows of nbytes between strided host buffers, split over the
// pool (by rows, or by bytes when there are fewer rows than parts).
void copy_rows(uint8_t* dst, size_t dst_stride, const uint8_t* src, size_t src_stride, size_t nrows,
               size_t nbytes)
{
    const size_t total = nrows * nbytes;
    if (total == 0) return;
    const size_t parts = std::min<size_t>(pool().width() * 2, std::max<size_t>(1, total >> 20));
    if (nrows >= parts) {
        pool().run(parts, [&](size_t p) {
            const size_t a = nrows * p / parts, b = nrows * (p + 1) / parts;
            if (dst_stride == nbytes && src_stride == nbytes) {
                std::memcpy(dst + a * nbytes, src + a * nbytes, (b - a) * nbytes);
                return;
            }
            for (size_t r = a; r < b; ++r) std::memcpy(dst + r * dst_stride, src + r * src_stride, nbytes);
        });
    } else {
        pool().run(parts, [&](size_t p) {
            const size_t a = total * p / parts, b = total * (p + 1) / parts;
            for (size_t r = a / nbytes; r < nrows && r * nbytes < b; ++r) {
                const size_t lo = std::max(a, r * nbytes) - r * nbytes;
                const size_t hi = std::min(b, (r + 1) * nbytes) - r * nbytes;
                std::memcpy(dst + r * dst_stride + lo, src + r * src_stride + lo, hi - lo);
            }
        });
    }
}

// Pieces of a rows x row_bytes transfer that fit one slot: whole rows when a
// row fits, else slot-sized segments of one row.
struct Piece {
    size_t row, nrows, off, bytes;  // rows [row, row+nrows), bytes [off, off+bytes) of each
};

template <typename F>
void for_each_piece(size_t rows, size_t row_bytes, F&& f)
{
    if (row_bytes == 0) return;
    if (row_bytes <= kSlotBytes) {
        const size_t per = kSlotBytes / row_bytes;
        for (size_t r = 0; r < rows; r += per) f(Piece{r, std::min(per, rows - r), 0, row_bytes});
    } else {
        for (size_t r = 0; r < rows; ++r)
            for (size_t o = 0; o < row_bytes; o += kSlotBytes)
                f(Piece{r, 1, o, std::min(kSlotBytes, row_bytes - o)});
    }
}

}  // namespace

HostIO::~HostIO()
{
    for (auto& e : m_ev)
        if (e) (void)hipEventDestroy(e);
}

void HostIO::prepare()
{
    for (int i = 0; i < 2; ++i) {
        // no transfer of an earlier call still uses the slot (an event never
        // recorded is complete)
        if (!m_ev[i]) DD_HIP(hipEventCreateWithFlags(&m_ev[i], hipEventDisableTiming));
        DD_HIP(hipEventSynchronize(m_ev[i]));
        m_slot[i].reserve(kSlotBytes);
    }
}

void HostIO::upload(hipStream_t stream, uint8_t* dst, size_t dst_stride, const uint8_t* src,
                    size_t src_stride, size_t row_bytes, size_t rows)
{
    prepare();
    size_t n = 0;
    for_each_piece(rows, row_bytes, [&](const Piece& p) {
        const int s = (int)(n++ & 1);
        // the slot's previous DMA has finished before it is refilled
        DD_HIP(hipEventSynchronize(m_ev[s]));
        auto* slot = static_cast<uint8_t*>(m_slot[s].ptr);
        copy_rows(slot, p.bytes, src + p.row * src_stride + p.off, src_stride, p.nrows, p.bytes);
        uint8_t* d = dst + p.row * dst_stride + p.off;
        if (dst_stride == p.bytes)  // contiguous rows: one linear copy
            DD_HIP(hipMemcpyAsync(d, slot, p.nrows * p.bytes, hipMemcpyHostToDevice, stream));
        else
            DD_HIP(hipMemcpy2DAsync(d, dst_stride, slot, p.bytes, p.bytes, p.nrows, hipMemcpyHostToDevice, stream));
        DD_HIP(hipEventRecord(m_ev[s], stream));
    });
}

void HostIO::download(hipStream_t stream, uint8_t* dst, size_t dst_stride, const uint8_t* src,
                      size_t src_stride, size_t row_bytes, size_t rows,
                      const std::function<void(size_t)>& ready)
{
    prepare();
    // issue piece k's DMA, then drain piece k-1 from the other slot while it runs
    Piece prev{0, 0, 0, 0};
    size_t n = 0;
    auto drain = [&](const Piece& p, int s) {
        DD_HIP(hipEventSynchronize(m_ev[s]));
        copy_rows(dst + p.row * dst_stride + p.off, dst_stride, static_cast<const uint8_t*>(m_slot[s].ptr),
                  p.bytes, p.nrows, p.bytes);
    };
    for_each_piece(rows, row_bytes, [&](const Piece& p) {
        const int s = (int)(n & 1);
        if (ready) ready(p.row + p.nrows);
        const uint8_t* d = src + p.row * src_stride + p.off;
        if (src_stride == p.bytes)  // contiguous rows: one linear copy
            DD_HIP(hipMemcpyAsync(m_slot[s].ptr, d, p.nrows * p.bytes, hipMemcpyDeviceToHost, stream));
        else
            DD_HIP(hipMemcpy2DAsync(m_slot[s].ptr, p.bytes, d, src_stride, p.bytes, p.nrows, hipMemcpyDeviceToHost,
                                    stream));
        DD_HIP(hipEventRecord(m_ev[s], stream));
        if (n > 0) drain(prev, 1 - s);
        prev = p;
        ++n;
    });
    if (n > 0) drain(prev, (int)((n - 1) & 1));
}

}  // namespace dedisp
