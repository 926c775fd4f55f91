/*
 * dedisp_oracle.c — CPU restatement of the reference dedispersion algorithms.
 *
 * TEST INFRASTRUCTURE ONLY. This file is the checker: it is built into
 * oracle/libdedisp_oracle.so and may be loaded only by tests/, by
 * __graft_entry__.smoke() and by the cpu_baseline leg of bench.py. The product
 * library (dedisp_tests_amd/libdedisp.so) never links, loads or calls it.
 *
 * Every function names the reference file:line it restates. Host floating
 * point follows the reference's operation order; the file is compiled with
 * -ffp-contract=off and without -ffast-math so float/double rounding matches
 * a plain x86-64 SSE build of the reference host code.
 *
 * Pinning (what checks each function independently of this text) is listed in
 * tests/test_oracle.py; functions without such a check are marked
 * "parity unpinned" below.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* Plan core                                                                 */
/* ------------------------------------------------------------------------ */

/* Delay table in samples per unit DM.
 * Restates Plan::Plan (src/Plan.cpp:16-40: df forced negative) and
 * Plan::generate_delay_table (src/Plan.cpp:109-119):
 *   a = 1.f/(f0 + c*df)   (c converted to float, float arithmetic)
 *   b = 1.f/f0
 *   delay[c] = (float)((4.15e3/(double)dt) * (double)(a*a - b*b))           */
void oracle_delay_table(unsigned long nchans, float dt, float f0, float df,
                        float* out)
{
    float dfn = -fabsf(df);
    for (unsigned long c = 0; c < nchans; ++c) {
        float a = 1.f / (f0 + (float)c * dfn);
        float b = 1.f / f0;
        float diff = a * a - b * b;
        double k = 4.15e3 / (double)dt;
        out[c] = (float)(k * (double)diff);
    }
}

/* DM list generator (Lina Levin recurrence).
 * Restates Plan::generate_dm_list (src/Plan.cpp:45-70) as called from
 * src/Plan.cpp:136-154 (dt, f0, df are the plan's float values promoted to
 * double; df keeps the caller's sign). Values are pushed as float and the
 * loop condition compares floats. Returns the number of DMs; writes at most
 * cap values into out (pass out=NULL to count only).                       */
unsigned long oracle_generate_dm_list(float dm_start, float dm_end,
                                      double dt, double ti, double f0, double df,
                                      unsigned long nchans, double tol,
                                      float* out, unsigned long cap)
{
    dt *= 1e6;
    double f    = (f0 + ((double)(nchans / 2) - 0.5) * df) * 1e-3;
    double tol2 = tol * tol;
    double a    = 8.3 * df / (f * f * f);
    double a2   = a * a;
    double b2   = a2 * (double)((double)(nchans * nchans) / 16.0);
    double c    = (dt * dt + ti * ti) * (tol2 - 1.0);

    unsigned long n = 0;
    float back = dm_start;
    if (out && n < cap) out[n] = back;
    n++;
    while (back < dm_end) {
        double prev  = (double)back;
        double prev2 = prev * prev;
        double k     = c + tol2 * a2 * prev2;
        double dm    = ((b2 * prev + sqrt(-a2 * b2 * prev2 + (a2 + b2) * k)) / (a2 + b2));
        /* a step that does not advance the float value never ends in the
         * reference; stop there */
        if (!((float)dm > back)) break;
        back = (float)dm;
        if (out && n < cap) out[n] = back;
        n++;
    }
    return n;
}

/* Maximum delay of a DM list: src/Plan.cpp:85-86 and :152-153,
 *   (dedisp_size)(dm_list[last] * delay_table[nchans-1] + 0.5)
 * float product, promoted to double for the +0.5, truncated.               */
unsigned long oracle_max_delay(float dm_last, float delay_last)
{
    float p = dm_last * delay_last;
    return (unsigned long)((double)p + 0.5);
}

/* ------------------------------------------------------------------------ */
/* TDD (time-domain dedispersion)                                            */
/* ------------------------------------------------------------------------ */

/* Integer delay used by the TDD kernel:
 * __float2uint_rn(dm * frac_delay) at src/tdd/dedisperse/dedisperse_kernel.cuh:163-165
 * float product, round-half-even to unsigned (negative and NaN saturate to 0,
 * values beyond UINT_MAX saturate).                                          */
unsigned oracle_tdd_delay(float dm, float delay)
{
    float p = dm * delay;
    if (!(p > 0.f)) return 0u;
    float r = nearbyintf(p); /* default rounding mode: to nearest even */
    if (r >= 4294967296.0f) return 0xffffffffu;
    return (unsigned)r;
}

/* Sample of channel c at time t from a time-major packed input row.
 * Layout: src/clib/dedisp.h:338-344 and the subword extraction at
 * src/tdd/dedisperse/dedisperse_kernel.cuh:38-43 (channel c at bit
 * (c % (32/nbits))*nbits of little-endian word c/(32/nbits)).               */
static inline uint32_t sample_at(const unsigned char* in, unsigned long in_stride,
                                 unsigned nbits, unsigned long t, unsigned long c)
{
    const unsigned char* row = in + t * in_stride;
    switch (nbits) {
    case 32: {
        uint32_t v; memcpy(&v, row + 4 * c, 4); return v; }
    case 16: {
        uint16_t v; memcpy(&v, row + 2 * c, 2); return v; }
    default: {
        unsigned long bit = c * nbits;
        return (row[bit >> 3] >> (bit & 7)) & ((1u << nbits) - 1u); }
    }
}

/* Output epilogue scale_output<IN_NBITS,T> + set_out_val
 * (src/tdd/dedisperse/dedisperse_kernel.cuh:29-31, 45-87):
 *   in_range  = (float)(2^nbits - 1)
 *   out_range = 1.f (32-bit float out) or (float)(2^out_nbits - 1)
 *   factor    = (3.f*1024.f)/255.f/16.f
 *   scaled    = (float)sum * out_range / (in_range * (float)nchans) * factor
 *   8/16-bit: clamp to [0, out_range] in double, then truncate.              */
static void store_out(unsigned char* out_row, unsigned out_nbits, unsigned long t,
                      uint64_t sum, unsigned in_nbits, unsigned long nchans)
{
    uint32_t in_max = (in_nbits == 32) ? 0xffffffffu : ((1u << in_nbits) - 1u);
    float in_range = (float)in_max;
    float out_range = (out_nbits == 32) ? 1.f
                    : (float)((out_nbits == 16) ? 65535u : 255u);
    float factor = (3.f * 1024.f) / 255.f / 16.f;
    float fsum = (in_nbits == 32) ? (float)sum : (float)(uint32_t)sum;
    float num = fsum * out_range;
    float den = in_range * (float)nchans;
    float scaled = num / den * factor;
    if (out_nbits == 32) {
        memcpy(out_row + 4 * t, &scaled, 4);
        return;
    }
    double v = fmax((double)scaled, 0.);
    v = fmin(v, (double)out_range);
    scaled = (float)v;
    if (out_nbits == 16) {
        uint16_t o = (uint16_t)scaled;
        memcpy(out_row + 2 * t, &o, 2);
    } else {
        out_row[t] = (unsigned char)scaled;
    }
}

/* Brute-force TDD: restates TDDPlan::execute_guru (src/tdd/TDDPlan.cpp:148-487)
 * and dedisperse_kernel (src/tdd/dedisperse/dedisperse_kernel.cuh:98-198):
 *   out[i][t] = E( sum_c  (uint32)(killmask[c] * sample(c, t + d(dm_i, c))) )
 * with d = __float2uint_rn(dm * delay[c]), sums in uint32 (uint64 for 32-bit
 * input, SumType at :34-36) and E the epilogue above. Computes DMs
 * first_dm .. first_dm+dm_count-1 into rows 0..dm_count-1 (out_stride bytes
 * apart), t = 0 .. nsamps-max_delay-1. Samples at t >= nsamps read as zero.
 * OpenMP over DMs.                                                          */
int oracle_tdd(unsigned long nchans, const float* delay_table,
               const int* killmask, const float* dm_list,
               unsigned long first_dm, unsigned long dm_count,
               unsigned long nsamps, unsigned long max_delay,
               const unsigned char* in, unsigned in_nbits, unsigned long in_stride,
               unsigned char* out, unsigned out_nbits, unsigned long out_stride)
{
    if (nsamps < max_delay) return -1;
    unsigned long nsc = nsamps - max_delay;
    long i;
#pragma omp parallel for schedule(dynamic, 1)
    for (i = 0; i < (long)dm_count; ++i) {
        float dm = dm_list[first_dm + (unsigned long)i];
        unsigned* d = (unsigned*)malloc(nchans * sizeof(unsigned));
        for (unsigned long c = 0; c < nchans; ++c) d[c] = oracle_tdd_delay(dm, delay_table[c]);
        unsigned char* orow = out + (unsigned long)i * out_stride;
        for (unsigned long t = 0; t < nsc; ++t) {
            uint64_t sum64 = 0;
            uint32_t sum32 = 0;
            for (unsigned long c = 0; c < nchans; ++c) {
                unsigned long ts = t + d[c];
                uint32_t v = (ts < nsamps) ? sample_at(in, in_stride, in_nbits, ts, c) : 0u;
                uint32_t term = (uint32_t)killmask[c] * v;
                if (in_nbits == 32) sum64 += term; else sum32 += term;
            }
            store_out(orow, out_nbits, t, (in_nbits == 32) ? sum64 : sum32, in_nbits, nchans);
        }
        free(d);
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* FDD (Fourier-domain dedispersion)                                         */
/* ------------------------------------------------------------------------ */

/* Integer delay used by the FDD GPU kernel: (int)(dm * delay[c]) with a float
 * product, src/fdd/dedisperse/fdd_kernel.cuh:226-229 (truncation toward zero). */
int oracle_fdd_delay(float dm, float delay)
{
    float p = dm * delay;
    return (int)p;
}

/* FP64 FDD time series (the ideal result of the FDD GPU path in time-series
 * mode, src/fdd/FDDGPUPlan.cpp:618-636 with EXPORT_DEDISP_TIME_SERIES, and of
 * the segmented path, :1185-1226):
 *   y[i][t] = sum_c killmask-free  x_c[t + d_c],  x_c[t] = (u8[t][c] - 127.5)/nchan
 * with d_c = (int)(dm_i * delay[c]) (fdd_kernel.cuh:226-229), t < nsamps-max_delay.
 * Zero-padded transforms of length N >= nsamps make the Fourier shift a plain
 * shift for these t (circular shift-sum identity). Output float64 rows of
 * nsamps-max_delay values. in holds one sample value per byte; offset is 127.5
 * for the reference's 8-bit input (src/fdd/unpack/unpack_kernel.cuh:60) and
 * (2^nbits - 1)/2 for the sub-byte extension. */
int oracle_fdd_timeseries(unsigned long nchans, const float* delay_table,
                          const float* dm_list, unsigned long ndm,
                          unsigned long nsamps, unsigned long max_delay,
                          const unsigned char* in, double offset, double* out)
{
    if (nsamps < max_delay) return -1;
    unsigned long nsc = nsamps - max_delay;
    double inv = 1.0 / (double)nchans;
    long i;
#pragma omp parallel for schedule(dynamic, 1)
    for (i = 0; i < (long)ndm; ++i) {
        double* orow = out + (unsigned long)i * nsc;
        for (unsigned long t = 0; t < nsc; ++t) orow[t] = 0.0;
        for (unsigned long c = 0; c < nchans; ++c) {
            long d = oracle_fdd_delay(dm_list[i], delay_table[c]);
            for (unsigned long t = 0; t < nsc; ++t) {
                long ts = (long)t + d;
                if (ts >= 0 && (unsigned long)ts < nsamps)
                    orow[t] += ((double)in[(unsigned long)ts * nchans + c] - offset) * inv;
            }
        }
    }
    return 0;
}

/* FP64 circular shift-sum over a transform length N (the time-domain image of
 * the reference's Fourier-domain output, src/fdd/FDDGPUPlan.cpp:85-93, 394-407):
 *   s[i][t] = sum_c x_c[(t + d_c) mod N],  t = 0..N-1, x_c zero for t >= nsamps.
 * Its FP64 DFT is Y[i][k] = sum_c X_c[k] e^{+2 pi i k d_c / N}, the quantity
 * the fork's FDD GPU path writes (bins 0..N/2-1). Output float64 [ndm][N].   */
int oracle_fdd_circular(unsigned long nchans, const float* delay_table,
                        const float* dm_list, unsigned long ndm,
                        unsigned long nsamps, unsigned long N,
                        const unsigned char* in, double offset, double* out)
{
    if (nsamps > N) return -1;
    double inv = 1.0 / (double)nchans;
    long i;
#pragma omp parallel for schedule(dynamic, 1)
    for (i = 0; i < (long)ndm; ++i) {
        double* orow = out + (unsigned long)i * N;
        for (unsigned long t = 0; t < N; ++t) orow[t] = 0.0;
        for (unsigned long c = 0; c < nchans; ++c) {
            long d = oracle_fdd_delay(dm_list[i], delay_table[c]);
            unsigned long dm = (unsigned long)(((d % (long)N) + (long)N) % (long)N);
            for (unsigned long ts = 0; ts < nsamps; ++ts) {
                unsigned long t = (ts + N - dm) % N;
                orow[t] += ((double)in[ts * nchans + c] - offset) * inv;
            }
        }
    }
    return 0;
}
