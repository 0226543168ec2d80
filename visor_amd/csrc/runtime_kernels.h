// runtime_kernels.h - ahead-of-time device helpers (runtime_kernels.hip).
#pragma once

#include <hip/hip_runtime_api.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Exclusive scan of n values (out may alias in); temp holds va_scan_temp_words(n) words. */
size_t va_scan_temp_words(uint32_t n);
hipError_t va_scan_u32(uint32_t *temp, const uint32_t *in, uint32_t *out, uint32_t n, hipStream_t stream);
/* Stable LSD radix sort of the first min(*n_dev, cap) (key, value) pairs over the low `bits` key
 * bits, ping-ponging between (keys, vals) and (keys_alt, vals_alt); *result_in_alt says which
 * pair of arrays holds the result. temp holds va_sort_temp_words(cap) words. */
size_t va_sort_temp_words(uint32_t cap);
hipError_t va_sort_pairs(uint32_t *temp, uint32_t *keys, uint32_t *vals, uint32_t *keys_alt, uint32_t *vals_alt,
                         const uint32_t *n_dev, uint32_t cap, int bits, int *result_in_alt, hipStream_t stream);
/* va_sort_pairs without a workspace: the histogram words and digit totals come from a device array of
 * the library sized for VA_SORT_NOALLOC_MAX_PAIRS pairs (8 MiB), so the call allocates nothing.
 * cap above the maximum returns hipErrorInvalidValue. Calls on one device share the array:
 * calls that overlap on one device race. */
#define VA_SORT_NOALLOC_MAX_PAIRS (1u << 24)
hipError_t va_sort_pairs_noalloc(uint32_t *keys, uint32_t *vals, uint32_t *keys_alt, uint32_t *vals_alt,
                                 const uint32_t *n_dev, uint32_t cap, int bits, int *result_in_alt, hipStream_t stream);
hipError_t va_fill_u32_async(void *p, uint64_t count, uint32_t value, hipStream_t stream);
hipError_t va_fill_u64_async(void *p, uint64_t count, uint32_t lo, uint32_t hi, hipStream_t stream);

#ifdef __cplusplus
}
#endif

/* Texel transcode: writes, for every texel of every level and layer, the 4 bytes the reference's
 * texture fetch returns (texel_decode.h) into an RGBA8 image of the same extent, levels and
 * layers (levels packed per layer, layers back to back). */
#define VA_TRANSCODE_MAX_LEVELS 16
typedef struct VaTranscode {
  uint64_t src_bytes;   /* whole source image */
  uint64_t src_slice;   /* source bytes per layer */
  uint64_t dst_slice;   /* shadow bytes per layer */
  uint32_t width, height, mips, layers, bpp, fmt;
  uint64_t src_level_off[VA_TRANSCODE_MAX_LEVELS];
  uint64_t dst_level_off[VA_TRANSCODE_MAX_LEVELS];
} VaTranscode;

#ifdef __cplusplus
extern "C" {
#endif
hipError_t va_transcode_async(const uint8_t *src, uint8_t *dst, const VaTranscode *t, hipStream_t stream);
hipError_t va_stamp_async(uint64_t *out, hipStream_t stream);
/* Counter sums, pair total and stage stamps of a pass into host-visible `out` (15 words): rows
 * holds nrows (at most 64) counter rows of 4 words for setup, then as many for the raster; the
 * launch sums them and leaves them zero. Last, out[14] = ++*seq (the number of launches so far,
 * *seq zero before the first), written after every other word of out is visible to the host. */
hipError_t va_finish_async(uint64_t *rows, uint32_t nrows, const uint32_t *pair_total, const uint64_t *stamps,
                           uint64_t *out, uint32_t *seq, hipStream_t stream);
#ifdef __cplusplus
}
#endif
