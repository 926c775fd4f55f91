/*
 * fdd_cpu.c — restatement of the reference CPU FDD plan, as written.
 *
 * TEST INFRASTRUCTURE ONLY (see dedisp_oracle.c). This is the CPU baseline
 * that bench.py times beside the GPU path, and it is checked in
 * tests/test_oracle.py against an FP64 evaluation of its own formula.
 *
 * Restates FDDCPUPlan::execute_cpu (src/fdd/FDDCPUPlan.cpp:439-624) with
 * dedisperse_optimized<float,float,32,false> (:277-403), transpose_data and
 * copy_data (src/fdd/helper.h:18-60), the spin-frequency table (:934-946) and
 * the FFTW wrappers (:949-1015). Its semantics are kept as the reference has
 * them, including where they differ from the GPU path: phasors use
 * f_k = k/(nsamp*dt) and fractional delays
 * dm*delay*dt while the FFT length is round_up(nsamp+1, 16384); only bins
 * 0..nsamp/2 are filled. The result is therefore not the dedispersed series
 * of the GPU path; it is a timing baseline.
 *
 * Differences from the reference text, none of which changes a computed value:
 *   - the transposed spectra live on the heap (the reference declares two
 *     nfreq x nchan float VLAs on the stack, :289-290, which overflow the
 *     default 8 MiB stack beyond toy sizes);
 *   - FFTW (src/fdd/FDDCPUPlan.cpp:949-1015: one FFTW_ESTIMATE plan per
 *     call, rows executed in parallel) is loaded at run time with
 *     dlopen("libfftw3f.so.3") when the host has it; otherwise
 *     oracle_rfft_f / oracle_irfft_f (fft_cpu.c) stand in, whose float
 *     rounding differs from FFTW's. oracle_fdd_cpu_fft_backend() names the
 *     one used.
 * A DM subset [dm_first, dm_first+dm_count) may be computed (the DM loop is
 * independent per DM, :305-402); the timing sample uses this.
 */
#include <dlfcn.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

void oracle_rfft_f(unsigned long n, const float* in, float* out);
void oracle_irfft_f(unsigned long n, const float* in, float* out);

static unsigned long round_up_ul(unsigned long a, unsigned long b) { return ((a + b - 1) / b) * b; }

/* FFTW single-precision entry points (the public fftw3 API), if present. */
typedef void* fftwf_plan_t;
static struct {
    int tried, ok;
    fftwf_plan_t (*plan_r2c)(int n, float* in, float* out, unsigned flags);
    fftwf_plan_t (*plan_c2r)(int n, float* in, float* out, unsigned flags);
    void (*exec_r2c)(const fftwf_plan_t p, float* in, float* out);
    void (*exec_c2r)(const fftwf_plan_t p, float* in, float* out);
    void (*destroy)(fftwf_plan_t p);
} g_fftw;
#define ORACLE_FFTW_ESTIMATE (1U << 6)

static void load_fftw(void)
{
    if (g_fftw.tried) return;
    g_fftw.tried = 1;
    void* h = dlopen("libfftw3f.so.3", RTLD_NOW | RTLD_LOCAL);
    if (!h) return;
    *(void**)&g_fftw.plan_r2c = dlsym(h, "fftwf_plan_dft_r2c_1d");
    *(void**)&g_fftw.plan_c2r = dlsym(h, "fftwf_plan_dft_c2r_1d");
    *(void**)&g_fftw.exec_r2c = dlsym(h, "fftwf_execute_dft_r2c");
    *(void**)&g_fftw.exec_c2r = dlsym(h, "fftwf_execute_dft_c2r");
    *(void**)&g_fftw.destroy = dlsym(h, "fftwf_destroy_plan");
    g_fftw.ok = g_fftw.plan_r2c && g_fftw.plan_c2r && g_fftw.exec_r2c && g_fftw.exec_c2r && g_fftw.destroy;
}

/* "fftw3f" when the transforms run through the host's FFTW, else "own". */
const char* oracle_fdd_cpu_fft_backend(void)
{
    load_fftw();
    return g_fftw.ok ? "fftw3f" : "own";
}

static double now_s(void)
{
#ifdef _OPENMP
    return omp_get_wtime();
#else
    return 0.0;
#endif
}

/* times (optional, 3 doubles, seconds): [0] input transpose + forward
 * transforms (independent of the DM count), [1] phasor contraction, [2]
 * inverse transforms + output copy. */
int oracle_fdd_cpu_asis(unsigned long nchan, float dt, const float* delays,
                        const float* dms, unsigned long dm_first, unsigned long dm_count,
                        unsigned long nsamp, unsigned long max_delay,
                        const unsigned char* in, float* out, double* times)
{
    double t_start = now_s();
    if (nsamp < max_delay || nchan % 32 != 0) return -1;
    const unsigned nchan_batch = 32;
    unsigned long nfreq = nsamp / 2 + 1;
    unsigned long nsamp_computed = nsamp - max_delay;
    unsigned long nsamp_fft = round_up_ul(nsamp + 1, 16384);
    unsigned long nsamp_padded = round_up_ul(nsamp_fft + 1, 1024);

    /* generate_spin_frequency_table(nfreq, nsamp, dt): (:934-946) */
    float* f = (float*)malloc(nfreq * sizeof(float));
    float nsdt = (float)nsamp * dt;
    for (unsigned long k = 0; k < nfreq; ++k) f[k] = (float)((double)k * (1.0 / (double)nsdt));

    /* transpose_data<u8,float>(nchan, nsamp, nchan, nsamp_padded, 127.5, nchan) */
    float* data_nu = (float*)calloc(nchan * nsamp_padded, sizeof(float));
    const float offset = 127.5f, scale = (float)nchan;
    long c;
#pragma omp parallel for
    for (c = 0; c < (long)nchan; ++c)
        for (unsigned long t = 0; t < nsamp; ++t)
            data_nu[(unsigned long)c * nsamp_padded + t] =
                ((float)in[t * nchan + (unsigned long)c] - offset) / scale;

    /* fft_r2c_inplace(nsamp_fft, nchan, nsamp_padded) */
    load_fftw();
    if (g_fftw.ok) {
        fftwf_plan_t plan = g_fftw.plan_r2c((int)nsamp_fft, data_nu, data_nu, ORACLE_FFTW_ESTIMATE);
#pragma omp parallel for
        for (c = 0; c < (long)nchan; ++c) {
            float* row = data_nu + (unsigned long)c * nsamp_padded;
            g_fftw.exec_r2c(plan, row, row);
        }
        g_fftw.destroy(plan);
    } else {
#pragma omp parallel for
        for (c = 0; c < (long)nchan; ++c) {
            float* row = data_nu + (unsigned long)c * nsamp_padded;
            float* tmp = (float*)malloc((nsamp_fft + 2) * sizeof(float));
            oracle_rfft_f(nsamp_fft, row, tmp);
            memcpy(row, tmp, (nsamp_fft + 2) * sizeof(float));
            free(tmp);
        }
    }

    double t_fft = now_s();
    /* dedisperse_optimized: transpose the spectra to [nfreq][nchan] */
    float* in_real = (float*)malloc(nfreq * nchan * sizeof(float));
    float* in_imag = (float*)malloc(nfreq * nchan * sizeof(float));
    long kk;
#pragma omp parallel for
    for (kk = 0; kk < (long)nfreq; ++kk)
        for (unsigned long ch = 0; ch < nchan; ++ch) {
            const float* s = data_nu + ch * nsamp_padded;
            in_real[(unsigned long)kk * nchan + ch] = s[2 * kk];
            in_imag[(unsigned long)kk * nchan + ch] = s[2 * kk + 1];
        }

    float* data_dm = (float*)calloc(dm_count * nsamp_padded, sizeof(float));
    long i;
#pragma omp parallel for schedule(dynamic, 1)
    for (i = 0; i < (long)dm_count; ++i) {
        float* tdms = (float*)malloc(nchan * sizeof(float));
        float dm = dms[dm_first + (unsigned long)i];
        for (unsigned long ch = 0; ch < nchan; ++ch) tdms[ch] = dm * delays[ch] * dt;
        float* orow = data_dm + (unsigned long)i * nsamp_padded;
        for (unsigned long k = 0; k < nfreq; ++k) {
            float sums_real[32], sums_imag[32];
            memset(sums_real, 0, sizeof(sums_real));
            memset(sums_imag, 0, sizeof(sums_imag));
            for (unsigned long co = 0; co < nchan; co += nchan_batch) {
                float phases[32], pr[32], pi[32];
                for (unsigned j = 0; j < nchan_batch; ++j)
                    phases[j] = (float)(2.0 * M_PI * (double)f[k] * (double)tdms[co + j]);
                for (unsigned j = 0; j < nchan_batch; ++j) {
                    pr[j] = cosf(phases[j]);
                    pi[j] = sinf(phases[j]);
                }
                const float* sr = in_real + k * nchan + co;
                const float* si = in_imag + k * nchan + co;
                for (unsigned j = 0; j < nchan_batch; ++j) {
                    sums_real[j] += sr[j] * pr[j];
                    sums_real[j] -= si[j] * pi[j];
                    sums_imag[j] += sr[j] * pi[j];
                    sums_imag[j] += si[j] * pr[j];
                }
            }
            float s_re = 0, s_im = 0;
            for (unsigned j = 0; j < nchan_batch; ++j) { s_re += sums_real[j]; s_im += sums_imag[j]; }
            orow[2 * k] = s_re;
            orow[2 * k + 1] = s_im;
        }
        free(tdms);
    }
    free(in_real); free(in_imag); free(data_nu); free(f);
    double t_dd = now_s();

    /* fft_c2r_inplace(nsamp_fft, ndm, nsamp_padded) with the 1/n scale, then
     * copy_data of the first nsamp_computed samples per DM */
    fftwf_plan_t iplan = g_fftw.ok ? g_fftw.plan_c2r((int)nsamp_fft, data_dm, data_dm, ORACLE_FFTW_ESTIMATE) : 0;
#pragma omp parallel for
    for (i = 0; i < (long)dm_count; ++i) {
        float* row = data_dm + (unsigned long)i * nsamp_padded;
        const double scale_n = 1.0 / (double)nsamp_fft;
        if (iplan) {
            g_fftw.exec_c2r(iplan, row, row);
            for (unsigned long t = 0; t < nsamp_computed; ++t)
                out[(unsigned long)i * nsamp_computed + t] = (float)((double)row[t] * scale_n);
        } else {
            float* tmp = (float*)malloc(nsamp_fft * sizeof(float));
            oracle_irfft_f(nsamp_fft, row, tmp);
            for (unsigned long t = 0; t < nsamp_computed; ++t)
                out[(unsigned long)i * nsamp_computed + t] = (float)((double)tmp[t] * scale_n);
            free(tmp);
        }
    }
    if (iplan) g_fftw.destroy(iplan);
    free(data_dm);
    if (times) {
        times[0] = t_fft - t_start;
        times[1] = t_dd - t_fft;
        times[2] = now_s() - t_dd;
    }
    return 0;
}
