/*
 * dedisp.h — C interface of the MI355X-native dedispersion library.
 *
 * Drop-in for the C API of dedisp (reference: src/clib/dedisp.h). Types, enum
 * values and prototypes are kept identical so that a caller compiled against
 * the reference header links against this library unchanged. Every function
 * lists the reference declaration it replaces.
 *
 * Conventions (reference src/clib/dedisp.h:330-374):
 *   - input is time-major: channels are the fastest dimension, no padding;
 *     sub-byte samples are packed LSB-first (channel c sits at bit
 *     (c % (32/nbits)) * nbits of little-endian 32-bit word c / (32/nbits));
 *   - output is DM-major: time is the fastest dimension, (nsamps - max_delay)
 *     samples per DM for the time-domain plans.
 *   - functions returning dedisp_error return the specific error code of the
 *     failure (the reference collapses most failures to DEDISP_UNKNOWN_ERROR;
 *     the numeric values are unchanged).
 */
#ifndef DEDISP_H_INCLUDE_GUARD
#define DEDISP_H_INCLUDE_GUARD

#ifdef __cplusplus
extern "C" {
#endif

/* Types: reference src/common/dedisp_types.h:6-9 and src/clib/dedisp.h:45. */
typedef float         dedisp_float;
typedef unsigned char dedisp_byte;
typedef unsigned long dedisp_size;
typedef int           dedisp_bool;
typedef struct dedisp_plan_struct* dedisp_plan;

/* Internal 32-bit packing word and interleaved complex (reference
 * src/common/dedisp_types.h:22-26). */
typedef unsigned int dedisp_word;
typedef struct { float real; float imag; } dedisp_float2;

/* Flags: reference src/common/dedisp_types.h:28-36. */
typedef enum {
    DEDISP_USE_DEFAULT     = 0,
    DEDISP_HOST_POINTERS   = 1 << 1,
    DEDISP_DEVICE_POINTERS = 1 << 2,
    DEDISP_WAIT            = 1 << 3,
    DEDISP_ASYNC           = 1 << 4
} dedisp_flag;

/* Error codes: reference src/clib/dedisp.h:76-94 (same order, same values). */
typedef enum {
    DEDISP_NO_ERROR,
    DEDISP_MEM_ALLOC_FAILED,
    DEDISP_MEM_COPY_FAILED,
    DEDISP_NCHANS_EXCEEDS_LIMIT,
    DEDISP_INVALID_PLAN,
    DEDISP_INVALID_POINTER,
    DEDISP_INVALID_STRIDE,
    DEDISP_NO_DM_LIST_SET,
    DEDISP_TOO_FEW_NSAMPS,
    DEDISP_INVALID_FLAG_COMBINATION,
    DEDISP_UNSUPPORTED_IN_NBITS,
    DEDISP_UNSUPPORTED_OUT_NBITS,
    DEDISP_INVALID_DEVICE_INDEX,
    DEDISP_DEVICE_ALREADY_SET,
    DEDISP_PRIOR_GPU_ERROR,
    DEDISP_INTERNAL_GPU_ERROR,
    DEDISP_UNKNOWN_ERROR
} dedisp_error;

/* Implementation selector: reference src/clib/dedisp.h:116-120.
 * DEDISP_DEDISP (the legacy default) runs the TDD kernels: the legacy kernel
 * computes the same per-sample arithmetic. */
typedef enum {
    DEDISP_DEDISP,
    DEDISP_TDD,
    DEDISP_FDD
} dedisp_implementation;

/* ---- Plan management ---------------------------------------------------- */

/* reference src/clib/dedisp.h:139; src/clib/dedisp.cpp:291-296 */
dedisp_error dedisp_select_implementation(dedisp_implementation imp);

/* reference src/clib/dedisp.h:155-159; src/clib/dedisp.cpp:104-160 */
dedisp_error dedisp_create_plan(dedisp_plan* plan,
                                dedisp_size  nchans,
                                dedisp_float dt,
                                dedisp_float f0,
                                dedisp_float df);

/* reference src/clib/dedisp.h:165; src/clib/dedisp.cpp:544-549 */
void dedisp_destroy_plan(dedisp_plan plan);

/* reference src/clib/dedisp.h:181; src/clib/dedisp.cpp:284-289 */
dedisp_error dedisp_set_device(int device_idx);

/* reference src/clib/dedisp.h:192-199; src/clib/dedisp.cpp:162-222 */
dedisp_error dedisp_set_gulp_size(dedisp_plan plan, dedisp_size gulp_size);
dedisp_size  dedisp_get_gulp_size(dedisp_plan plan);

/* reference src/clib/dedisp.h:214-216; src/clib/dedisp.cpp:298-317 */
dedisp_error dedisp_set_killmask(dedisp_plan plan, const dedisp_bool* killmask);

/* reference src/clib/dedisp.h:234-236; src/clib/dedisp.cpp:224-244 */
dedisp_error dedisp_set_dm_list(dedisp_plan plan,
                                const dedisp_float* dm_list,
                                dedisp_size count);

/* reference src/clib/dedisp.h:254-257; src/clib/dedisp.cpp:246-267 */
dedisp_error dedisp_generate_dm_list(dedisp_plan  plan,
                                     dedisp_float dm_start, dedisp_float dm_end,
                                     dedisp_float ti, dedisp_float tol);

/* reference src/clib/dedisp.h:261-264 (declared and exported there, body
 * commented out at src/clib/dedisp.cpp:269-282). Here it is implemented: the
 * returned array holds *dm_count values and is released with
 * dedisp_free_dm_list_guru. */
dedisp_float* dedisp_generate_dm_list_guru(dedisp_float dm_start, dedisp_float dm_end,
                                           double dt, double ti, double f0, double df,
                                           dedisp_size nchans, double tol,
                                           dedisp_size* dm_count);
/* Releases an array returned by dedisp_generate_dm_list_guru (no reference
 * counterpart: the reference never returned an owned array). */
void dedisp_free_dm_list_guru(dedisp_float* dm_list);

/* ---- Getters: reference src/clib/dedisp.h:276-329; dedisp.cpp:321-370 ---- */
dedisp_size         dedisp_get_max_delay(const dedisp_plan plan);
dedisp_size         dedisp_get_dm_delay(const dedisp_plan plan, int dm_trial);
dedisp_size         dedisp_get_channel_count(const dedisp_plan plan);
dedisp_size         dedisp_get_dm_count(const dedisp_plan plan);
const dedisp_float* dedisp_get_dm_list(const dedisp_plan plan);
const dedisp_bool*  dedisp_get_killmask(const dedisp_plan plan);
dedisp_float        dedisp_get_dt(const dedisp_plan plan);
dedisp_float        dedisp_get_f0(const dedisp_plan plan);
dedisp_float        dedisp_get_df(const dedisp_plan plan);
const char*         dedisp_get_error_string(dedisp_error error);

/* ---- Execution ------------------------------------------------------------
 * reference src/clib/dedisp.h:375-381; src/clib/dedisp.cpp:501-526.
 * flags: DEDISP_DEVICE_POINTERS makes in/out device pointers (the reference
 * documents this flag but ignores it, src/tdd/TDDPlan.cpp:102-127);
 * DEDISP_ASYNC returns once the work is queued on the plan's stream
 * (dedisp_sync waits for it). */
dedisp_error dedisp_execute(const dedisp_plan  plan,
                            dedisp_size        nsamps,
                            const dedisp_byte* in,
                            dedisp_size        in_nbits,
                            dedisp_byte*       out,
                            dedisp_size        out_nbits,
                            unsigned           flags);

/* reference src/clib/dedisp.h:425-433; src/clib/dedisp.cpp:440-498
 * (TDD / DEDISP plans only; FDD plans return DEDISP_INVALID_PLAN). */
dedisp_error dedisp_execute_adv(const dedisp_plan  plan,
                                dedisp_size        nsamps,
                                const dedisp_byte* in,
                                dedisp_size        in_nbits,
                                dedisp_size        in_stride,
                                dedisp_byte*       out,
                                dedisp_size        out_nbits,
                                dedisp_size        out_stride,
                                unsigned           flags);

/* reference src/clib/dedisp.h:480-490; src/clib/dedisp.cpp:374-438
 * (TDD / DEDISP plans only). DMs first_dm_idx .. first_dm_idx+dm_count-1 are
 * computed; row 0 of out holds DM first_dm_idx. */
dedisp_error dedisp_execute_guru(const dedisp_plan  plan,
                                 dedisp_size        nsamps,
                                 const dedisp_byte* in,
                                 dedisp_size        in_nbits,
                                 dedisp_size        in_stride,
                                 dedisp_byte*       out,
                                 dedisp_size        out_nbits,
                                 dedisp_size        out_stride,
                                 dedisp_size        first_dm_idx,
                                 dedisp_size        dm_count,
                                 unsigned           flags);

/* reference src/clib/dedisp.h:501; src/clib/dedisp.cpp:530-542 */
dedisp_error dedisp_sync(void);

/* Adaptive time resolution: reference src/clib/dedisp.h:516-545, declared and
 * exported but commented out in src/clib/dedisp.cpp. Kept as exported stubs:
 * enable returns DEDISP_UNKNOWN_ERROR (not supported), disable succeeds,
 * using returns 0, get_dt_factors returns NULL. */
dedisp_error       dedisp_enable_adaptive_dt(dedisp_plan plan, dedisp_float pulse_width,
                                             dedisp_float tol);
dedisp_error       dedisp_disable_adaptive_dt(dedisp_plan plan);
dedisp_bool        dedisp_using_adaptive_dt(const dedisp_plan plan);
const dedisp_size* dedisp_get_dt_factors(const dedisp_plan plan);

#ifdef __cplusplus
}
#endif

#endif /* DEDISP_H_INCLUDE_GUARD */
