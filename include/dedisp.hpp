/*
 * dedisp.hpp — C++ plan classes of the MI355X-native dedispersion library.
 *
 * Same class names, constructors and member functions as the reference
 * (src/Plan.hpp:18-114, src/GPUPlan.hpp:16-75, src/tdd/TDDPlan.hpp:13-71,
 * src/fdd/FDDGPUPlan.hpp:14-71), so code written against the reference's C++
 * API compiles against this header. HIP state is kept behind opaque pointers:
 * the header needs no HIP include.
 *
 * Errors are thrown as dedisp::Error (a std::runtime_error, as the reference
 * throws at src/common/dedisp_error.hpp:60-76) carrying the dedisp_error code.
 */
#ifndef DEDISP_HPP_INCLUDE_GUARD
#define DEDISP_HPP_INCLUDE_GUARD

#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "dedisp.h"
#include "dedisp_amd.h"

namespace dedisp
{

class Error : public std::runtime_error {
public:
    Error(dedisp_error code, const std::string& what)
        : std::runtime_error(what), m_code(code) {}
    dedisp_error code() const { return m_code; }
private:
    dedisp_error m_code;
};

/* Host part of a plan: delay table, DM list, max delay and killmask.
 * reference src/Plan.hpp:18-114, src/Plan.cpp:16-154. */
class Plan {
public:
    typedef dedisp_size  size_type;
    typedef dedisp_byte  byte_type;
    typedef dedisp_float float_type;
    typedef dedisp_bool  bool_type;

    Plan(size_type nchans, float_type dt, float_type f0, float_type df);
    Plan(const Plan& other) = delete;
    Plan& operator=(const Plan& other) = delete;
    virtual ~Plan();

    virtual void generate_dm_list(float_type dm_start, float_type dm_end,
                                  float_type ti, float_type tol);
    virtual void set_dm_list(const float_type* dm_list, size_type count);
    virtual void set_killmask(const bool_type* killmask);

    /* get_max_delay returns float_type as the reference does (Plan.hpp:68). */
    float_type get_max_delay()     const { return (float_type)m_max_delay; }
    size_type  get_channel_count() const { return m_nchans; }
    size_type  get_dm_count()      const { return m_dm_count; }
    float_type get_dt()            const { return m_dt; }
    float_type get_df()            const { return m_df; }
    float_type get_f0()            const { return m_f0; }

    const float_type* get_dm_list()     const { return h_dm_list.data(); }
    const float_type* get_delay_table() const { return h_delay_table.data(); }
    const bool_type*  get_killmask()    const { return h_killmask.data(); }

    void sync();

    virtual void execute(size_type nsamps, const byte_type* in, size_type in_nbits,
                         byte_type* out, size_type out_nbits, unsigned flags = 0) = 0;

    /* The DM-list generator of Plan.cpp:45-70, usable without a plan. */
    static void generate_dm_list(std::vector<dedisp_float>& dm_table,
                                 dedisp_float dm_start, dedisp_float dm_end,
                                 double dt, double ti, double f0, double df,
                                 dedisp_size nchans, double tol);

protected:
    dedisp_size  m_dm_count;
    dedisp_size  m_nchans;
    dedisp_size  m_max_delay;
    dedisp_float m_dt;
    dedisp_float m_f0;
    dedisp_float m_df;

    std::vector<dedisp_float> h_dm_list;
    std::vector<dedisp_float> h_delay_table;
    std::vector<dedisp_bool>  h_killmask;

    /* Bumped whenever the DM list or killmask changes; plans use it to
     * refresh their device-side tables lazily. */
    unsigned long m_tables_version = 0;
};

struct GPUContext; /* device, streams, events and device tables (opaque) */

/* Device part shared by the GPU plans: one HIP device, its streams and the
 * device copies of the tables. reference src/GPUPlan.hpp:16-75. */
class GPUPlan : public Plan {
protected:
    GPUPlan(size_type nchans, float_type dt, float_type f0, float_type df,
            int device_idx = 0);
    virtual ~GPUPlan();

    std::unique_ptr<GPUContext> m_ctx;

public:
    virtual void generate_dm_list(float_type dm_start, float_type dm_end,
                                  float_type ti, float_type tol) override;
    virtual void set_dm_list(const float_type* dm_list, size_type count) override;
    virtual void set_killmask(const bool_type* killmask) override;

    int get_device() const;
    /* Device context (streams, device buffers) of the plan; opaque to
     * callers, used by the streaming front end. */
    GPUContext& context() const { return *m_ctx; }

    /* Stage timings of the last execute (extension, dedisp_amd.h). */
    void set_profiling(bool enable);
    dedisp_amd_profile get_profile() const;
};

struct TDDState;

/* Time-domain dedispersion. reference src/tdd/TDDPlan.hpp:13-71,
 * src/tdd/TDDPlan.cpp:43-487. execute_adv / execute_guru take an optional
 * trailing flags argument (the reference versions have none). */
class TDDPlan : public GPUPlan {
public:
    TDDPlan(size_type nchans, float_type dt, float_type f0, float_type df,
            int device_idx = 0);
    ~TDDPlan();

    /* Output samples per gulp (reference default 65536). This library keeps
     * the whole series resident and gulps only when a gulp size has been set
     * here, when the series' device buffers would exceed 80 % of the free
     * device memory, or from 2^31 samples on; 0 is refused. */
    void set_gulp_size(size_type gulp_size);
    dedisp_size get_gulp_size() const { return m_gulp_size; }

    virtual void execute(size_type nsamps, const byte_type* in, size_type in_nbits,
                         byte_type* out, size_type out_nbits,
                         unsigned flags = 0) override;

    void execute_adv(size_type nsamps, const byte_type* in, size_type in_nbits,
                     size_type in_stride, byte_type* out, size_type out_nbits,
                     size_type out_stride, unsigned flags = 0);

    void execute_guru(size_type nsamps, const byte_type* in, size_type in_nbits,
                      size_type in_stride, byte_type* out, size_type out_nbits,
                      size_type out_stride, dedisp_size first_dm_idx,
                      dedisp_size dm_count, unsigned flags = 0);

    /* Batched execution (extension, dedisp_amd_execute_batch): nbeams series
     * of uneven length through the same DM list; beam b has nsamps[b]
     * time-major samples at in[b] and gets (nsamps[b] - max_delay) outputs per
     * DM at out[b] (DM-major, unpadded). With the packed-lane kernel one
     * channel-sum launch covers every beam; with DEDISP_DEVICE_POINTERS the
     * launches are captured once in a HIP graph and replayed while the
     * batch (lengths, pointers, widths, tables) stays the same. */
    void execute_batch(size_type nbeams, const size_type* nsamps, const byte_type* const* in,
                       size_type in_nbits, byte_type* const* out, size_type out_nbits,
                       unsigned flags = 0);

    /* Times every instantiated tiling of the channel-sum kernel for this
     * plan's DM list on nsamps samples of in_nbits input (device scratch) and
     * keeps the fastest for later executes at that width. Returns the time of
     * the chosen tiling in ms and the tiling in *chosen. */
    double autotune(size_type nsamps, size_type in_nbits, size_type out_nbits,
                    dedisp_amd_tuning* chosen);

    /* execute() on device pointers with every device buffer it needs (the
     * transposed input rows and the delay tables) carved out of the caller's
     * scratch of scratch_bytes, allocating no device memory
     * (dedisp_amd_execute_noalloc). *scratch_tag names the tables the scratch
     * holds: they are uploaded only when it differs from this plan's current
     * tables, and it is updated. Throws DEDISP_MEM_ALLOC_FAILED when rows and
     * tables do not fit. */
    void execute_noalloc(size_type nsamps, const byte_type* in, size_type in_nbits,
                         byte_type* out, size_type out_nbits, unsigned flags,
                         void* scratch, size_type scratch_bytes, unsigned long long* scratch_tag);

private:
    void execute_span(size_type nsamps, const byte_type* in, size_type in_nbits, size_type in_stride,
                      byte_type* out, size_type out_nbits, size_type out_stride, dedisp_size first_dm_idx,
                      dedisp_size dm_count, unsigned flags);
    void prepare_tables(unsigned nbits, dedisp_size first_dm_idx, dedisp_size dm_count,
                        bool upload = true);
    size_t row_words_for(unsigned nbits, size_t nsamps, size_t nsc) const;

    dedisp_size m_gulp_size;
    std::unique_ptr<TDDState> m_state;
};

struct FDDState;

/* The legacy plan (DEDISP_DEDISP, the C API's default selector). reference
 * src/dedisp/DedispPlan.hpp:9-74, DedispPlan.cpp:24-50: same interface as
 * TDDPlan (gulp size 65536, at most 8192 channels, execute / execute_adv /
 * execute_guru) and the same per-sample integer arithmetic, so it runs the
 * TDD kernels. */
class DedispPlan : public TDDPlan {
public:
    DedispPlan(size_type nchans, float_type dt, float_type f0, float_type df, int device_idx = 0)
        : TDDPlan(nchans, dt, f0, df, device_idx) {}
};

/* Fourier-domain dedispersion on the GPU. reference src/fdd/FDDGPUPlan.hpp,
 * src/fdd/FDDGPUPlan.cpp:39-706 (whole-series transform) and :725-1271
 * (time-segmented, overlap-save). Output formats and precision are chosen
 * with dedisp_amd_fdd_* (dedisp_amd.h); the default is the reference fork's
 * Fourier-coefficient output. */
class FDDGPUPlan : public GPUPlan {
public:
    FDDGPUPlan(size_type nchans, float_type dt, float_type f0, float_type df,
               int device_idx = 0);
    ~FDDGPUPlan();

    virtual void execute(size_type nsamps, const byte_type* in, size_type in_nbits,
                         byte_type* out, size_type out_nbits,
                         unsigned flags = 0) override;

    void set_output_mode(dedisp_amd_fdd_output mode);
    dedisp_amd_fdd_output get_output_mode() const;
    void set_precision(dedisp_amd_fdd_precision precision);
    dedisp_amd_fdd_precision get_precision() const;
    /* nfft = 0 selects the supported segment length (2^13 .. 2^18) with the
     * smallest contraction volume (dedisp_amd_fdd_set_segmented); segmented
     * mode writes time series. */
    void set_segmented(int enable, dedisp_size nfft);

    /* Output length per DM, in floats, for nsamps input samples. */
    dedisp_size output_length(dedisp_size nsamps) const;

    /* Times the overlap-save segment lengths 2^13 .. 2^18 longer than
     * max_delay on nsamps samples of in_nbits input (device scratch,
     * time-series output) and keeps the fastest (set_segmented(1, nfft)).
     * Returns its time in ms and the choice in *chosen. */
    double autotune(size_type nsamps, size_type in_nbits, size_type out_nbits,
                    dedisp_amd_tuning* chosen);

    /* Batched execution (extension, dedisp_amd_execute_batch): as
     * TDDPlan::execute_batch. In time-series mode the overlap-save chunks of
     * all beams go through one forward transform, one contraction and one
     * inverse transform per group of beams (one segment length for the
     * batch), so short beams share the contraction's 32-chunk tiles; each
     * beam's output equals execute() on that beam alone with the same segment
     * length. Whole-series Fourier output runs beam by beam. */
    void execute_batch(size_type nbeams, const size_type* nsamps, const byte_type* const* in,
                       size_type in_nbits, byte_type* const* out, size_type out_nbits,
                       unsigned flags = 0);

private:
    void execute_fdd(size_type nsamps, const byte_type* in, size_type in_nbits, byte_type* out,
                     unsigned flags, bool fourier);
    void prepare_tables();
    size_t segment_length(const size_t* nsc, size_t n) const;
    float prepare_phasors(size_t nfft, bool valu);

    std::unique_ptr<FDDState> m_state;
};

/* The reference's CPU FDD plan (src/fdd/FDDCPUPlan.hpp:14-40), declared so
 * code written against the reference's C++ API compiles and links. This
 * library computes every plan on the GPU and does not provide the CPU
 * algorithm: constructing one throws dedisp::Error with
 * DEDISP_UNKNOWN_ERROR, as dedisp_create_plan refuses USE_CPU=1 for FDD. */
class FDDCPUPlan : public Plan {
public:
    FDDCPUPlan(size_type nchans, float_type dt, float_type f0, float_type df, int device_index = 0);
    ~FDDCPUPlan();
    virtual void execute(size_type nsamps, const byte_type* in, size_type in_nbits,
                         byte_type* out, size_type out_nbits,
                         unsigned flags = 0) override;
};

} // namespace dedisp

#endif // DEDISP_HPP_INCLUDE_GUARD
