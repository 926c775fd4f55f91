// Internal C-ABI of the FDD kernels (see kernels.h for conventions).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

// Segment lengths the transforms handle: 2^13 .. 2^16 on chip, 2^17 and
// 2^18 split in time (dd_fdd_forward_split / dd_fdd_inverse_split).
int dd_fdd_fft_supported(size_t nfft);
// Bytes of work buffer the split transforms need for ntransforms transforms
// of length nfft (0 for nfft <= 65536).
size_t dd_fdd_split_work(size_t nfft, size_t ntransforms);

// Zeroes bytes at p (both multiples of 16) with a kernel: the form that may
// be recorded in a captured graph.
hipError_t dd_zero_fill(void* p, size_t bytes, hipStream_t stream);

// One overlap-save chunk of one beam, for transforms over the chunks of
// several beams at once (all beams share nfft, the channels and the DM list;
// the chunk's good length is the launch's). Forward: the beam's channel rows
// (row_bytes per channel), the chunk's first sample t0, the beam's sample
// count nvalid (zero beyond). Inverse: the beam's output (DM-major, out_stride
// floats per DM row), the chunk's first output sample t0, the beam's outputs
// per DM nsamps_out.
typedef struct {
    const uint8_t* rows;
    float* out;
    uint64_t row_bytes;
    uint64_t out_stride;
    uint32_t t0;
    uint32_t nvalid;
    uint32_t nsamps_out;
    uint32_t reserved;
} dd_fdd_chunk;

// dd_fdd_forward over a chunk table: chunks [nchunk] on the device; X:
// [nchan][nchunk][Xp], chunk i of X from chunks[i]. No periodic reads.
hipError_t dd_fdd_forward_chunks(const dd_fdd_chunk* chunks, size_t nchunk, const float* chan_scale,
                                 const float2* ph, size_t nfft, float2* X, size_t nchan, size_t Xp,
                                 unsigned nbits, hipStream_t stream);

// dd_fdd_inverse over a chunk table: Y row r (r < nrows), chunk i goes to DM
// row row0 + r of chunks[i].out, samples chunks[i].t0 + s for s < good and
// below chunks[i].nsamps_out.
hipError_t dd_fdd_inverse_chunks(const float2* Y, size_t nrows, size_t row0, const dd_fdd_chunk* chunks,
                                 size_t nchunk, size_t Yp, const float2* ph, size_t nfft, size_t good,
                                 hipStream_t stream);

// Forward transforms of every (channel, chunk): rows are channel-major
// nbits-bit samples (1, 2, 4 or 8; packed LSB-first along time in 32-bit
// words, row_bytes per channel), chunk i starts at sample i*chunk_step and
// spans nfft samples; x = (u - (2^nbits-1)/2) * chan_scale[c] for sample
// t < nvalid and 0 beyond. period > 0 reads the series periodically (sample
// t mod period; the chunks then span at most two periods). X:
// [nchan][nchunk][Xp] complex, the real spectrum bins 0..nfft/2
// (Xp >= nfft/2 + 1). ph: [nfft] e^{+2 pi i m/nfft}.
hipError_t dd_fdd_forward(const uint8_t* rows, size_t row_bytes, size_t chunk_step,
                          size_t nvalid, size_t period, const float* chan_scale,
                          const float2* ph, size_t nfft, float2* X, size_t nchan,
                          size_t nchunk, size_t Xp, unsigned nbits, hipStream_t stream);

// Inverse transforms of every (row, chunk) of Y ([nrows][nchunk][Yp], bins
// 0..nfft/2): writes samples t < good of each chunk, scaled by 1/nfft, to
// out[row*out_stride + chunk*good + t] for chunk*good + t < nsamps_out.
hipError_t dd_fdd_inverse(const float2* Y, size_t nrows, size_t nchunk, size_t Yp,
                          const float2* ph, size_t nfft, float* out, size_t out_stride,
                          size_t good, size_t nsamps_out, hipStream_t stream);

// dd_fdd_forward / dd_fdd_inverse for every supported nfft: up to 65536 the
// on-chip kernels; 2^17 and 2^18 as R = nfft/65536 transforms of 32768
// points per segment (decimation in time) and a radix-R combine, through
// work (dd_fdd_split_work(nfft, nchan*nchunk) or (nrows*nchunk) bytes).
// ph64k: e^{+2 pi i m/65536} (m < 65536), the on-chip engine's table.
// blocked != 0 (nfft <= 65536): X in the bin-block layout instead of rows of
// Xp: bins k, k+1 (k even, k <= M/2) and their mirrors M-k, M-k-1 form block
// k/2 of NBB = M/4 + 1 blocks, element (c, k, chunk) at
// ((c*NBB + block)*nchunk + chunk)*4 + slot, slot k&1 for k <= M/2 and
// 2 + ((M-k)&1) above (the layout dd_fdd_contract_mfma16 reads with blocked).
hipError_t dd_fdd_forward_split(const uint8_t* rows, size_t row_bytes, size_t chunk_step, size_t nvalid,
                                size_t period, const float* chan_scale, const float2* ph, const float2* ph64k,
                                size_t nfft, float2* X, size_t nchan, size_t nchunk, size_t Xp, unsigned nbits,
                                float2* work, int blocked, hipStream_t stream);
// packed 1: Y holds the packed half-length spectrum Zu[0..M-1] of each
// transform (dd_fdd_contract_mfma16 with ph_packed) instead of bins 0..M;
// packed 2 (nfft <= 65536): the packed spectrum in the bin-block layout.
hipError_t dd_fdd_inverse_split(const float2* Y, size_t nrows, size_t nchunk, size_t Yp, const float2* ph,
                                const float2* ph64k, size_t nfft, float* out, size_t out_stride, size_t good,
                                size_t nsamps_out, float2* work, int packed, hipStream_t stream);
// The chunk-table forms (dd_fdd_forward_chunks / dd_fdd_inverse_chunks) for
// every supported nfft (the layouts of a launch's nchunk chunks as above).
hipError_t dd_fdd_forward_chunks_split(const dd_fdd_chunk* chunks, size_t nchunk, const float* chan_scale,
                                       const float2* ph, const float2* ph64k, size_t nfft, float2* X, size_t nchan,
                                       size_t Xp, unsigned nbits, float2* work, int blocked, hipStream_t stream);
hipError_t dd_fdd_inverse_chunks_split(const float2* Y, size_t nrows, size_t row0, const dd_fdd_chunk* chunks,
                                       size_t nchunk, size_t Yp, const float2* ph, const float2* ph64k, size_t nfft,
                                       size_t good, float2* work, int packed, hipStream_t stream);

// Real transforms of length N = 16384*m of nrows rows of S (float, row
// stride s_stride >= N, even), bins 0..N/2-1 written as complex to out (row
// stride out_stride complex). ph16384: e^{+2 pi i j/16384}; twL: e^{-2 pi i j/L},
// twN: e^{-2 pi i k/N} for j, k < L = N/2. work: nrows*L complex. S is
// overwritten.
hipError_t dd_fdd_rfft_rows(const float* S, size_t s_stride, size_t nrows, size_t N,
                            const float2* ph16384, const float2* twL, const float2* twN,
                            float2* work, float2* out, size_t out_stride, hipStream_t stream);

// Phasor contraction, float32 on the vector ALUs:
//   Y[i][ch][k] = sum_c e^{+2 pi i ((k*d[i][c]) mod nfft)/nfft} X_c[ch][k],  k = 0..nfft/2
// with X the spectra of dd_fdd_forward. delays: [ndm][nchan] int32.
hipError_t dd_fdd_contract_fp32(const float2* X, size_t Xp, const int32_t* delays, size_t ndm,
                                size_t nchan, size_t nchunk, const float2* ph, size_t nfft,
                                float2* Y, size_t Yp, hipStream_t stream);

// Phasor contraction on the matrix cores. prec: 0 = float32 MFMA
// (v_mfma_f32_32x32x2_f32), 1 = bfloat16 MFMA with hi+lo split phasors and
// spectra (three products), 2 = float16 MFMA (one product); float32
// accumulation in all three. Same output as dd_fdd_contract_fp32; the delays
// enter as phase steps q[i][c] = d[i][c] mod nfft (uint32, row pitch qpitch, a
// multiple of 16 and >= nchan rounded up to 16, zero padded). The spectra are
// multiplied by xscale before the MFMA inputs are formed and Y by 1/xscale
// after (a power of two keeping float16 in range). X must hold nchan rounded
// up to 16 channel rows, the extra rows zero. nfft a power of two up to 2^18.
hipError_t dd_fdd_contract_mfma(int prec, const float2* X, size_t Xp, const uint32_t* q,
                                size_t qpitch, size_t ndm, size_t nchan, size_t nchunk, size_t nfft,
                                float xscale, float2* Y, size_t Yp, hipStream_t stream);

// Phasor contraction on 16 x 16 matrix-core tiles with the real and
// imaginary outputs of a chunk side by side along N. prec: 1 = bfloat16 with
// hi+lo split phasors and spectra (three products), 2 = float16 (one
// product); float32 accumulation. q: [ndm][qpitch] uint32 phase steps
// d mod nfft (qpitch a multiple of 32, >= nchan rounded up to 32, zero
// padded); X must hold nchan rounded up to 32 channel rows, the extra rows
// zero. mirror != 0: each workgroup takes two bins k, k+1 <= nfft/4 and their
// mirrors nfft/2 - k, whose phasors are (-1)^d conj(P_k). nfft a power of two
// up to 2^18. Same output as dd_fdd_contract_mfma; with ph_packed
// (e^{+2 pi i m/nfft}, mirror only) each row instead holds the packed
// half-length spectrum the inverse transform starts from,
// Zu[k] = (Y[k] + conj Y[M-k]) + i e^{2 pi i k/nfft} (Y[k] - conj Y[M-k]), k < M.
// prec 0 = float32 (v_mfma_f32_16x16x4_f32). With dmlist (ndm floats) and
// delay32 (the delay table, nchan rounded up to 32 entries, zero padded) the
// integer delays are computed in the kernel as (int)(dm * delay[c]) and q is
// not read.
// cfg selects the workgroup shape: 0 = 8 waves of 2 16-row DM tiles and 4
// bin slots; 1 = 4 waves of 4 DM tiles and 2 bin slots (half the B-operand
// LDS reads per product, two workgroups per CU); 2 = the software-pipelined
// form of 0 (bfloat16 / float16 with dmlist; a K-step's spectra within 32-bit
// element offsets). blocked != 0 (cfg 2 with mirror and ph_packed, nfft <=
// 65536): X and the packed output in the bin-block layout of
// dd_fdd_forward_split (Xp and Yp unused).
hipError_t dd_fdd_contract_mfma16(int prec, int mirror, const float2* X, size_t Xp, const uint32_t* q,
                                  size_t qpitch, size_t ndm, size_t nchan, size_t nchunk, size_t nfft,
                                  float xscale, float2* Y, size_t Yp, const float2* ph_packed,
                                  const float* dmlist, const float* delay32, int cfg, int blocked,
                                  hipStream_t stream);
// N-tiles (8 chunks each) per workgroup dd_fdd_contract_mfma16 uses for
// nchunk chunks, and the number of chunk groups.
int dd_fdd_mfma16_tiles(size_t nchunk, int prec, int* ntl, size_t* ngroups);

#ifdef __cplusplus
}
#endif
