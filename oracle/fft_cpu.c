/*
 * fft_cpu.c — mixed-radix Stockham FFT for the oracle (float and double).
 *
 * TEST INFRASTRUCTURE ONLY (see dedisp_oracle.c). The reference uses FFTW3f
 * (src/fdd/FDDCPUPlan.cpp:949-1015), which is absent from this image; this is
 * the oracle's own FFT, pinned in tests/test_oracle.py against numpy.fft and
 * against textbook DFT identities. Conventions match FFTW: the forward
 * transform uses e^{-2 pi i k t / n}, r2c returns bins 0..n/2, c2r is
 * unnormalized (returns n times the inverse).
 *
 * Algorithm: Stockham autosort (Govindaraju et al., "High performance discrete
 * Fourier transforms on graphics processors", SC'08), generalised to a mixed
 * radix sequence; any radix p is handled by a direct length-p DFT.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

static int factorize(unsigned long n, unsigned* radix, int max)
{
    int k = 0;
    while (n % 4 == 0 && k < max) { radix[k++] = 4; n /= 4; }
    while (n % 2 == 0 && k < max) { radix[k++] = 2; n /= 2; }
    for (unsigned p = 3; n > 1 && k < max; p += 2) {
        while (n % p == 0 && k < max) { radix[k++] = p; n /= p; }
        if ((unsigned long)p * p > n && n > 1) { radix[k++] = (unsigned)n; n = 1; }
    }
    return k;
}

#define DEFINE_FFT(REAL, SUF, COS, SIN)                                          \
/* In-place complex FFT of length n on interleaved (re,im) data.            */   \
static void cfft_##SUF(unsigned long n, REAL* data, int sign)                    \
{                                                                               \
    if (n <= 1) return;                                                         \
    unsigned radix[64];                                                         \
    int nr = factorize(n, radix, 64);                                           \
    REAL* buf = (REAL*)malloc(2 * n * sizeof(REAL));                            \
    REAL* a = data; REAL* b = buf;                                              \
    unsigned pmax = 0;                                                          \
    for (int s = 0; s < nr; ++s) if (radix[s] > pmax) pmax = radix[s];          \
    double* vr = (double*)malloc(2 * pmax * sizeof(double));                    \
    double* vi = vr + pmax;                                                     \
    double* wr = (double*)malloc(2 * pmax * sizeof(double));                    \
    double* wi = wr + pmax;                                                     \
    unsigned long Ns = 1;                                                       \
    for (int s = 0; s < nr; ++s) {                                              \
        unsigned p = radix[s];                                                  \
        unsigned long stride = n / p;                                           \
        for (unsigned q = 0; q < p; ++q) {                                      \
            double ang = sign * 2.0 * M_PI * q / p;                             \
            wr[q] = cos(ang); wi[q] = sin(ang);                                 \
        }                                                                       \
        for (unsigned long j = 0; j < stride; ++j) {                            \
            unsigned long jm = j % Ns;                                          \
            double ang = sign * 2.0 * M_PI * (double)jm / (double)(Ns * p);     \
            for (unsigned r = 0; r < p; ++r) {                                  \
                double xr = a[2 * (j + r * stride)];                            \
                double xi = a[2 * (j + r * stride) + 1];                        \
                double c = cos(r * ang), sn = sin(r * ang);                     \
                vr[r] = xr * c - xi * sn;                                       \
                vi[r] = xr * sn + xi * c;                                       \
            }                                                                   \
            unsigned long idxD = (j / Ns) * Ns * p + jm;                        \
            for (unsigned k = 0; k < p; ++k) {                                  \
                double sr = 0, si = 0;                                          \
                for (unsigned r = 0; r < p; ++r) {                              \
                    unsigned e = (unsigned)(((unsigned long)r * k) % p);        \
                    sr += vr[r] * wr[e] - vi[r] * wi[e];                        \
                    si += vr[r] * wi[e] + vi[r] * wr[e];                        \
                }                                                               \
                b[2 * (idxD + k * Ns)] = (REAL)sr;                              \
                b[2 * (idxD + k * Ns) + 1] = (REAL)si;                          \
            }                                                                   \
        }                                                                       \
        REAL* t = a; a = b; b = t;                                              \
        Ns *= p;                                                                \
    }                                                                           \
    if (a != data) memcpy(data, a, 2 * n * sizeof(REAL));                       \
    free(buf); free(vr); free(wr);                                              \
}                                                                               \
                                                                                \
/* Real-to-complex transform of even length n: in[n] -> out[n/2+1] complex. */  \
void oracle_rfft_##SUF(unsigned long n, const REAL* in, REAL* out)             \
{                                                                               \
    unsigned long M = n / 2;                                                    \
    REAL* z = (REAL*)malloc(2 * M * sizeof(REAL));                              \
    memcpy(z, in, n * sizeof(REAL));                                            \
    cfft_##SUF(M, z, -1);                                                       \
    for (unsigned long k = 0; k <= M; ++k) {                                    \
        unsigned long k1 = k % M, k2 = (M - k) % M;                             \
        double zr = z[2 * k1], zi = z[2 * k1 + 1];                              \
        double cr = z[2 * k2], ci = -z[2 * k2 + 1];                             \
        double er = 0.5 * (zr + cr), ei = 0.5 * (zi + ci);                      \
        double dr = zr - cr, di = zi - ci;                                      \
        double orr = 0.5 * di, oi = -0.5 * dr; /* (d)/(2i) */                   \
        double ang = -2.0 * M_PI * (double)k / (double)n;                       \
        double c = cos(ang), s = sin(ang);                                      \
        out[2 * k] = (REAL)(er + orr * c - oi * s);                             \
        out[2 * k + 1] = (REAL)(ei + orr * s + oi * c);                         \
    }                                                                           \
    free(z);                                                                    \
}                                                                               \
                                                                                \
/* Complex-to-real transform of even length n: in[n/2+1] -> out[n], returns  */ \
/* n times the inverse (FFTW c2r convention); Im of bins 0 and n/2 ignored.  */ \
void oracle_irfft_##SUF(unsigned long n, const REAL* in, REAL* out)            \
{                                                                               \
    unsigned long M = n / 2;                                                    \
    REAL* z = (REAL*)malloc(2 * M * sizeof(REAL));                              \
    for (unsigned long k = 0; k < M; ++k) {                                     \
        double xr = in[2 * k], xi = (k == 0) ? 0.0 : in[2 * k + 1];             \
        double yr = in[2 * (M - k)], yi = (k == 0) ? 0.0 : -in[2 * (M - k) + 1]; \
        double er = xr + yr, ei = xi + yi;                                      \
        double dr = xr - yr, di = xi - yi;                                      \
        double ang = 2.0 * M_PI * (double)k / (double)n;                        \
        double c = cos(ang), s = sin(ang);                                      \
        double orr = dr * c - di * s, oi = dr * s + di * c;                     \
        z[2 * k] = (REAL)(er - oi);                                             \
        z[2 * k + 1] = (REAL)(ei + orr);                                        \
    }                                                                           \
    cfft_##SUF(M, z, +1);                                                       \
    memcpy(out, z, n * sizeof(REAL));                                           \
    free(z);                                                                    \
}                                                                               \
                                                                                \
void oracle_cfft_##SUF(unsigned long n, REAL* data, int sign)                  \
{                                                                               \
    cfft_##SUF(n, data, sign);                                                  \
}

DEFINE_FFT(double, d, cos, sin)
DEFINE_FFT(float, f, cosf, sinf)
