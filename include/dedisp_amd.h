/*
 * dedisp_amd.h — extensions of the MI355X-native dedispersion library.
 *
 * None of these entry points has a counterpart in the reference C API
 * (src/clib/dedisp.h); they expose choices the reference makes with compile
 * time macros or environment variables, plus per-stage device timings.
 * Plain C types only.
 */
#ifndef DEDISP_AMD_H_INCLUDE_GUARD
#define DEDISP_AMD_H_INCLUDE_GUARD

#include "dedisp.h"

#ifdef __cplusplus
extern "C" {
#endif

/* FDD output format.
 *  FOURIER:    the reference fork's default (src/fdd/FDDGPUPlan.cpp:85-93,
 *              394-407): N = round_up(nsamps + max_delay, 16384) floats per DM,
 *              interleaved re/im of Fourier bins 0 .. N/2-1 of the dedispersed
 *              series (Nyquist bin dropped).
 *  TIMESERIES: the reference's EXPORT_DEDISP_TIME_SERIES build
 *              (FDDGPUPlan.hpp:6, FDDGPUPlan.cpp:618-636): nsamps - max_delay
 *              floats per DM. The segmented mode always writes time series. */
typedef enum {
    DEDISP_AMD_FDD_FOURIER    = 0,
    DEDISP_AMD_FDD_TIMESERIES = 1
} dedisp_amd_fdd_output;

/* Arithmetic of the FDD phasor contraction.
 *  FP32: plain float32 complex multiply-accumulate (the drop-in form);
 *  BF16: bfloat16 MFMA with float32 accumulation, split hi+lo phasors and
 *        spectra (three-product compensation);
 *  FP16: float16 MFMA with float32 accumulation, per-channel spectrum
 *        scaling, single product. */
typedef enum {
    DEDISP_AMD_FDD_FP32 = 0,
    DEDISP_AMD_FDD_BF16 = 1,
    DEDISP_AMD_FDD_FP16 = 2
} dedisp_amd_fdd_precision;

/* Device time per stage of the last execute, in milliseconds, measured with
 * HIP events on the stream each kernel runs on (only while profiling is on).
 * "dedisperse" is the TDD channel-sum kernel or the FDD contraction. */
typedef struct {
    double        h2d_ms;
    double        unpack_ms;
    unsigned long unpack_launches;
    double        fft_ms;
    unsigned long fft_launches;
    double        dedisperse_ms;
    unsigned long dedisperse_launches;
    double        ifft_ms;
    unsigned long ifft_launches;
    double        d2h_ms;
    double        total_ms;
    /* TDD: channel-adds; FDD: complex multiply-accumulates of the contraction
     * (algorithmic counts, no padding). */
    double        work_units;
} dedisp_amd_profile;

dedisp_error dedisp_amd_set_profiling(dedisp_plan plan, int enable);
dedisp_error dedisp_amd_get_profile(const dedisp_plan plan, dedisp_amd_profile* out);

dedisp_error dedisp_amd_fdd_set_output(dedisp_plan plan, dedisp_amd_fdd_output mode);
dedisp_error dedisp_amd_fdd_set_precision(dedisp_plan plan, dedisp_amd_fdd_precision precision);
/* enable=1 selects the time-segmented (overlap-save) path; nfft=0 picks the
 * reference's automatic segment length (src/fdd/FDDGPUPlan.cpp:742-768),
 * otherwise nfft must be a power of two larger than max_delay. The reference
 * selects this path with the USE_SEGMENTED environment variable, which is
 * honoured too. */
dedisp_error dedisp_amd_fdd_set_segmented(dedisp_plan plan, int enable, dedisp_size nfft);
/* Floats written per DM by dedisp_execute on an FDD plan for nsamps samples. */
dedisp_size  dedisp_amd_fdd_output_length(const dedisp_plan plan, dedisp_size nsamps);

/* The implementation a plan was created with. */
dedisp_error dedisp_amd_get_implementation(const dedisp_plan plan, dedisp_implementation* impl);

/* Host-only helper (no device needed): the plan's delay table for the given
 * geometry, identical to the table a plan builds (src/Plan.cpp:109-119). */
dedisp_error dedisp_amd_generate_delay_table(dedisp_size nchans, dedisp_float dt,
                                             dedisp_float f0, dedisp_float df,
                                             dedisp_float* out);

/* Library version string. */
const char* dedisp_amd_version(void);

#ifdef __cplusplus
}
#endif

#endif /* DEDISP_AMD_H_INCLUDE_GUARD */
