// Internal C-ABI between the host plans and the gfx950 kernels: plain
// pointers, sizes and a stream. Every launcher returns the hipError_t of its
// launch (hipSuccess on success).
#pragma once

#include <hip/hip_runtime_api.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

// ---- TDD -----------------------------------------------------------------

// Time-major packed input -> channel-major rows packed along time.
// in: nsamps rows of in_stride bytes (channel c of row t at bit (c*nbits) of
// the row, LSB first). out: nchans rows of row_words 32-bit words; sample t of
// channel c sits at bits (t % (32/nbits))*nbits of word t/(32/nbits). Samples
// t in [nsamps, row_words*32/nbits) are written as zero.
hipError_t dd_tdd_transpose(const uint8_t* in, size_t nsamps, size_t in_stride,
                            unsigned nbits, size_t nchans, uint32_t* out,
                            size_t row_words, hipStream_t stream);

// Epilogue constants of scale_output (src/tdd/dedisperse/dedisperse_kernel.cuh:45-64).
typedef struct {
    float out_range;  // 1 (float out) or 2^out_nbits - 1
    float den;        // in_range * (float)nchans
    float factor;     // (3.f*1024.f)/255.f/16.f
    unsigned out_nbits;
} dd_tdd_epilogue;

// Generic channel sum: any nbits (1..32), integer killmask weights.
// delays: [ndm][nact] integer sample delays; chan: [nact] channel indices;
// weight: [nact] killmask values. out row i at out + i*out_stride bytes.
hipError_t dd_tdd_dedisperse_generic(const uint32_t* rows, size_t row_words, unsigned nbits,
                                     const uint32_t* delays, const int32_t* chan,
                                     const int32_t* weight, size_t nact, size_t ndm,
                                     size_t nsamps_out, uint8_t* out, size_t out_stride,
                                     dd_tdd_epilogue ep, hipStream_t stream);

// Tiling of the LDS (packed-lane) channel-sum kernel.
typedef struct {
    unsigned samples_per_lane;  // S: output samples one lane owns (multiple of 32/nbits)
    unsigned dms_per_wave;      // DMs each wave accumulates
    unsigned waves;             // waves per workgroup (DM block = waves*dms_per_wave)
    unsigned chunk_channels;    // channels staged in LDS per pass
    unsigned window_words;      // LDS words per staged channel
} dd_tdd_tiling;

// Packed-lane channel sum for nbits in {1,2,4,8} and a 0/1 killmask (only the
// nact active channels are summed). wstart: [ndmblk][nact] first staged word
// of each channel relative to the time tile's first word.
hipError_t dd_tdd_dedisperse_lds(const uint32_t* rows, size_t row_words, unsigned nbits,
                                 const uint32_t* delays, const int32_t* chan,
                                 const uint32_t* wstart, const uint32_t* chunks, unsigned maxc,
                                 size_t nact, size_t ndm,
                                 size_t nsamps_out, uint8_t* out, size_t out_stride,
                                 dd_tdd_epilogue ep, dd_tdd_tiling tiling,
                                 hipStream_t stream);

// Whether dd_tdd_dedisperse_lds has an instantiation for this tiling.
int dd_tdd_lds_supported(unsigned nbits, dd_tdd_tiling tiling);

// Runs-layout channel sum (tdd_runs.hip): 4-bit input, the tiling with
// samples_per_lane 28 (7 time runs x 4 positions per lane), 4 DMs per wave
// and 4 waves or 2 DMs per wave and 8 waves. Each DM block (16 DMs) stages its
// channels in chunks listed in
// chunks[block * maxc ...] (entry: first channel | channels << 16 |
// (positions / 128) << 24; an entry with no channels ends the list); wstart:
// [ndmblk][nact] first staged word of each channel relative to the tile's
// first word (a multiple of 2); code: [DM group][nact][DMs per wave] LDS byte offsets of
// each DM's window in the run words of its chunk. A chunk's staged windows fit
// in DD_TDD_RUNS_L1_BYTES (time-ordered words, 72 LDS words per 64) and its
// run words in DD_TDD_RUNS_L2_BYTES, with at most DD_TDD_RUNS_MAXCC channels.
// At most 4369 active channels (16-bit sums); chan and wstart are read up to 8
// entries past a chunk's last channel. beams / tile_beam / batch_tiles:
// as dd_tdd_dedisperse_lds_batch (or null / 0 for one series).
#define DD_TDD_RUNS_L1_BYTES 18432
#define DD_TDD_RUNS_L2_BYTES 30720
#define DD_TDD_RUNS_SPL 28
#define DD_TDD_RUNS_MAXCC 32
static inline int dd_tdd_tiling_is_runs(unsigned nbits, dd_tdd_tiling t)
{
    return nbits == 4 && t.samples_per_lane == DD_TDD_RUNS_SPL;
}

// The instantiated tilings for nbits (16-channel chunks): writes up to max
// into out and returns how many there are.
int dd_tdd_lds_tilings(unsigned nbits, dd_tdd_tiling* out, int max);

// One beam of a batched channel sum: its transposed rows, output rows and
// the index of its first time tile among all beams' tiles.
typedef struct {
    const uint32_t* rows;   // [nchans][row_words]
    uint8_t* out;           // DM row i at out + i*out_stride
    uint64_t row_words;
    uint64_t out_stride;    // bytes
    uint32_t nsamps_out;    // > 0
    uint32_t tile0;
} dd_tdd_beam;

// Packed-lane channel sum over several beams in one launch (same DM list,
// channels, tables and tiling). beams: [nbeams] on the device; tile_beam:
// [ntiles] beam index of every time tile (device), ntiles = sum over beams of
// ceil(nsamps_out / (64 * samples_per_lane)).
hipError_t dd_tdd_dedisperse_lds_batch(const dd_tdd_beam* beams, const uint32_t* tile_beam,
                                       size_t ntiles, unsigned nbits, const uint32_t* code,
                                       const int32_t* chan, const uint32_t* wstart,
                                       const uint32_t* chunks, unsigned maxc, size_t nact,
                                       size_t ndm, dd_tdd_epilogue ep, dd_tdd_tiling tiling,
                                       hipStream_t stream);

// The runs-layout channel sum (see dd_tdd_tiling_is_runs).
hipError_t dd_tdd_dedisperse_runs(const uint32_t* rows, size_t row_words, const uint32_t* code,
                                  const int32_t* chan, const uint32_t* wstart, const uint32_t* chunks,
                                  unsigned maxc, size_t nact, size_t ndm, size_t nsamps_out,
                                  uint8_t* out, size_t out_stride, dd_tdd_epilogue ep,
                                  unsigned dms_per_wave, unsigned waves, const dd_tdd_beam* beams,
                                  const uint32_t* tile_beam, size_t batch_tiles, hipStream_t stream);

#ifdef __cplusplus
}
#endif
