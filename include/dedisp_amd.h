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
/* enable=1 selects the time-segmented (overlap-save) path. nfft is the
 * segment length: a power of two from 2^13 to 2^18 larger than max_delay + 3,
 * or 0 for the supported length with the smallest contraction volume (chunk
 * columns of the contraction's tiles x (nfft/2 + 1)); the reference instead
 * doubles nfft from 16384 until
 * a fifth of it covers max_delay (src/fdd/FDDGPUPlan.cpp:756-762). The same
 * choice picks the segments of the time-series and Fourier outputs when no
 * length is set, so any max_delay below 2^18 - 3 samples is supported. The
 * reference selects this path with the USE_SEGMENTED environment variable,
 * which is honoured too. */
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

/* ---------------------------------------------------------------------------
 * Streaming front end (no reference counterpart: the caller-side loop a
 * real-time pipeline runs around dedisp_execute).
 *
 * A stream serves nbeams independent beams that deliver fixed chunks of
 * chunk_nsamps samples. Each beam keeps the last max_delay samples of its
 * input on the device, so every push of one chunk yields chunk_nsamps output
 * samples per DM: after the n-th push (n = 0, 1, ...) of a beam the outputs are
 * the dedispersed samples t = n*chunk_nsamps - max_delay + i, i < chunk_nsamps,
 * of the beam's concatenated input (samples before the first push read as 0
 * for TDD plans and as the zero level, 2^(in_nbits-1), for FDD plans).
 * Outputs for t >= 0 equal those of one dedisp_execute over the whole series.
 * The plan (TDD, legacy or FDD) must have its DM list set; an FDD plan is
 * switched to time-series output. max_delay is fixed at creation.
 * --------------------------------------------------------------------------- */
typedef struct dedisp_amd_stream_struct* dedisp_amd_stream;

typedef struct {
    unsigned long pushes;       /* push calls since creation or the last reset */
    double        mean_ms;      /* wall time per push call, host-synchronous */
    double        p50_ms;
    double        p99_ms;
    double        max_ms;
    double        samples_per_s;  /* input samples (all beams) per second of push time */
} dedisp_amd_stream_stats;

/* Device bytes a stream with these parameters allocates (ring buffers and
 * output staging), for admission against free HBM. */
dedisp_size  dedisp_amd_stream_bytes(const dedisp_plan plan, dedisp_size nbeams,
                                     dedisp_size chunk_nsamps, dedisp_size in_nbits,
                                     dedisp_size out_nbits);
/* Fails with DEDISP_MEM_ALLOC_FAILED when dedisp_amd_stream_bytes exceeds the
 * device's free memory less a 10 % reserve (admission). */
dedisp_error dedisp_amd_stream_create(dedisp_amd_stream* stream, dedisp_plan plan,
                                      dedisp_size nbeams, dedisp_size chunk_nsamps,
                                      dedisp_size in_nbits, dedisp_size out_nbits);
/* chunk: chunk_nsamps time-major samples of the beam (nchans*in_nbits/8 bytes
 * each); out: ndm rows of chunk_nsamps outputs (out_nbits/8 bytes each).
 * flags: DEDISP_DEVICE_POINTERS when both are device pointers (else host). */
dedisp_error dedisp_amd_stream_push(dedisp_amd_stream stream, dedisp_size beam,
                                    const dedisp_byte* chunk, dedisp_byte* out,
                                    unsigned flags);
/* One chunk for each of n distinct beams (beams[i] < nbeams) in one batched
 * execution (dedisp_amd_execute_batch over the beams' [history | chunk]
 * buffers): the outputs equal n single pushes. The statistics record one
 * latency per call (the time until every output of the call is written).
 * A repeated beam index is rejected with DEDISP_INVALID_POINTER. */
dedisp_error dedisp_amd_stream_push_batch(dedisp_amd_stream stream, dedisp_size n,
                                          const dedisp_size* beams, const dedisp_byte* const* chunks,
                                          dedisp_byte* const* outs, unsigned flags);
/* Clears every beam's history (to its initial value) and the statistics. */
dedisp_error dedisp_amd_stream_reset(dedisp_amd_stream stream);
dedisp_error dedisp_amd_stream_get_stats(const dedisp_amd_stream stream,
                                         dedisp_amd_stream_stats* stats);
dedisp_size  dedisp_amd_stream_history(const dedisp_amd_stream stream);
void         dedisp_amd_stream_destroy(dedisp_amd_stream stream);

/* ---------------------------------------------------------------------------
 * Filterbank front end: SIGPROC filterbank input, PRESTO .dat/.inf output
 * (the reference's bin/fil/fil.hpp as library calls; host only, no device).
 * File errors return DEDISP_UNKNOWN_ERROR.
 * --------------------------------------------------------------------------- */
typedef struct {
    long   header_bytes;   /* bytes before the samples */
    long   data_bytes;     /* bytes of samples after the header */
    int    nchans, nsamps, nbits, nifs;  /* nsamps from the file size */
    int    machine_id, telescope_id, nbeams, ibeam;
    double tstart, tsamp, fch1, foff;
    double src_raj, src_dej, az_start, za_start;
    char   source_name[80];
} dedisp_amd_fil_header;

/* reference fil.hpp:288-350 (read_header) */
dedisp_error dedisp_amd_fil_read_header(const char* path, dedisp_amd_fil_header* header);
/* The first nbytes sample bytes (time-major, channels fastest). */
dedisp_error dedisp_amd_fil_read_data(const char* path, const dedisp_amd_fil_header* header,
                                      dedisp_byte* data, dedisp_size nbytes);
/* reference fil.hpp:31-52: 8-bit samples averaged over groups of factor in
 * place; returns the new sample count nsamps / factor. */
dedisp_size  dedisp_amd_fil_decimate(dedisp_byte* data, dedisp_size nchans, dedisp_size nsamps,
                                     dedisp_size factor);
/* <prefix>_DM<dm %.3f>.dat: numout float samples (fil.hpp:262-281). */
dedisp_error dedisp_amd_write_dat(const char* prefix, dedisp_float dm, const dedisp_float* series,
                                  dedisp_size numout);
/* <prefix>_DM<dm %.3f>.inf: PRESTO info file (fil.hpp:468-539). */
dedisp_error dedisp_amd_write_inf(const char* prefix, const dedisp_amd_fil_header* header,
                                  dedisp_float dm, dedisp_size numout);

/* ---------------------------------------------------------------------------
 * Batched beams and autotuning (no reference counterpart: the reference runs
 * one series per dedisp_execute and fixes its tile sizes at compile time).
 * --------------------------------------------------------------------------- */

/* nbeams series of uneven length through the plan's DM list: beam b has
 * nsamps[b] time-major samples of in_nbits at in[b] (unpadded rows of
 * nchans*in_nbits/8 bytes) and receives (nsamps[b] - max_delay) output
 * samples per DM at out[b] (DM-major, unpadded rows), exactly as dedisp_execute
 * on that beam alone would write them (FDD plans: their configured output).
 * flags as for dedisp_execute; with DEDISP_DEVICE_POINTERS every in[b] and
 * out[b] is device memory. TDD and legacy plans with a 0/1 killmask and
 * in_nbits <= 8 transpose each beam and run one channel-sum launch over all
 * beams; with DEDISP_DEVICE_POINTERS that sequence is captured in a HIP graph
 * and replayed while lengths, pointers, widths and DM list stay the same.
 * FDD plans writing time series use one segment length for the batch (the
 * set one, else the one with the smallest summed contraction volume) and run
 * the overlap-save chunks of all beams through one forward transform, one
 * contraction and one inverse transform per group of beams; each beam's
 * output equals dedisp_execute on it alone with that segment length. Other
 * cases (weighted killmasks, 16/32-bit TDD input, whole-series FDD spectra)
 * run the beams one after another. */
dedisp_error dedisp_amd_execute_batch(dedisp_plan plan, dedisp_size nbeams,
                                      const dedisp_size* nsamps, const dedisp_byte* const* in,
                                      dedisp_size in_nbits, dedisp_byte* const* out,
                                      dedisp_size out_nbits, unsigned flags);

typedef struct {
    /* TDD / legacy: tiling of the channel-sum kernel */
    unsigned    samples_per_lane, dms_per_wave, waves, chunk_channels;
    /* FDD: overlap-save segment length (time-series output) */
    dedisp_size nfft;
    double      best_ms;      /* execute time of the chosen configuration */
    unsigned    candidates;   /* configurations timed */
    /* FDD contraction at that segment length: fdd_form 0, 1, 2 = the 16 x 16
     * tile kernel with 8 waves, 4 waves, software-pipelined 8 waves; 3 = the
     * 32 x 32 tile kernel. fdd_mirror: bins paired with their mirrors (1) or
     * consecutive (0); fdd_layout: spectra in bin blocks (1) or rows (0);
     * -1 where the form has no such choice. 0 for TDD plans. */
    int         fdd_form, fdd_mirror, fdd_layout;
} dedisp_amd_tuning;

/* Times the plan's configurations on nsamps samples of in_nbits input in
 * device scratch (contents irrelevant to the timing) and keeps the fastest for
 * later executes: TDD / legacy plans every instantiated channel-sum tiling for
 * that width, FDD plans the overlap-save segment lengths 2^13 .. 2^18 longer
 * than max_delay (time-series output) and then, at the chosen length, the
 * contraction forms of the plan's precision (workgroup shape, mirror bins,
 * spectra layout; kept for that precision, the environment variables of the
 * forms still override). Requires the DM list; returns the choice in *result. */
dedisp_error dedisp_amd_autotune(dedisp_plan plan, dedisp_size nsamps, dedisp_size in_nbits,
                                 dedisp_size out_nbits, dedisp_amd_tuning* result);

/* ---------------------------------------------------------------------------
 * Allocation-free execution. No counterpart in the reference: for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).
 *
 * dedisp_execute on a TDD or legacy plan with device pointers (in: nsamps
 * unpadded time-major rows of nchans*in_nbits/8 bytes; out: DM-major unpadded
 * rows of nsamps - max_delay outputs), giving the same output, but every
 * device buffer it needs (the transposed input rows and the delay tables)
 * lives in one fixed device array of DEDISP_AMD_NOALLOC_SCRATCH_BYTES per
 * device, held from the library's load until the process ends; no device
 * memory is allocated or freed. It needs
 *     nchans * 4 * row_words + the tables (about ndm * nchans * 4 bytes)
 * with row_words ~ (nsamps + max_delay) * in_nbits / 32 plus a tile; calls
 * whose rows and tables do not fit return DEDISP_MEM_ALLOC_FAILED before any
 * device work. The tables are uploaded from the host only when the array holds
 * another plan's (or an older DM list's) tables. flags: DEDISP_DEVICE_POINTERS
 * (implied), DEDISP_ASYNC (return without waiting for the plan's stream);
 * DEDISP_HOST_POINTERS is refused. Every call on one device shares that
 * device's array: calls that overlap on one device (several host threads, or
 * DEDISP_ASYNC calls of different plans) race. */
#define DEDISP_AMD_NOALLOC_SCRATCH_BYTES (128ul << 20)
dedisp_error dedisp_amd_execute_noalloc(dedisp_plan plan, dedisp_size nsamps, const dedisp_byte* in,
                                        dedisp_size in_nbits, dedisp_byte* out, dedisp_size out_nbits,
                                        unsigned flags);

#ifdef __cplusplus
}
#endif

#endif /* DEDISP_AMD_H_INCLUDE_GUARD */
