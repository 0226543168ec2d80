/*
 * visor_bins.h - the tile-binning primitives exported by libvisor_mi355x.so.
 *
 * No counterpart in the reference: its rasterizer hands 32x32-pixel blocks to worker threads
 * through one mutex-guarded FIFO (rasterizer.cpp:36-98, 454-515). Here every (tile, triangle)
 * pair of a pass is binned on the device instead: an exclusive scan of per-triangle tile counts
 * gives each triangle its first pair, and a stable radix sort by tile key orders the pairs into
 * per-tile bins that keep API primitive order (stability is what keeps blending in order). The
 * batched pass calls these internally; they are exported so that they can be tested alone.
 *
 * All pointers are device pointers; `stream` is a HIP stream (NULL = the default stream). The
 * functions return a HIP error code (0 = success) and only enqueue work on the stream.
 */
#ifndef VISOR_BINS_H
#define VISOR_BINS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ihipStream_t *hipStream_t;

/* Words of device workspace va_scan_u32 needs for n items. */
size_t va_scan_temp_words(uint32_t n);
/* Exclusive scan of n uint32 values (out may alias in). */
int va_scan_u32(uint32_t *temp, const uint32_t *in, uint32_t *out, uint32_t n, hipStream_t stream);

/* Words of device workspace va_sort_pairs needs for a capacity of cap pairs. */
size_t va_sort_temp_words(uint32_t cap);
/* Stable LSD radix sort of the first min(*n_dev, cap) (key, value) pairs by the low `bits` key
 * bits (8 bits per pass), ping-ponging between (keys, vals) and (keys_alt, vals_alt);
 * *result_in_alt (host memory) says which pair of arrays holds the result. The pair count is
 * read on the device, so the call needs no host synchronisation. */
int va_sort_pairs(uint32_t *temp, uint32_t *keys, uint32_t *vals, uint32_t *keys_alt, uint32_t *vals_alt,
                  const uint32_t *n_dev, uint32_t cap, int bits, int *result_in_alt, hipStream_t stream);

/* No counterpart in the reference. For callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): va_sort_pairs without
 * a workspace argument. The histogram words and digit totals live in a device array of the library sized
 * for VA_SORT_NOALLOC_MAX_PAIRS pairs (8 MiB per GPU, held until the process ends); the call
 * allocates nothing. cap above the maximum returns a nonzero error before any device work. Every
 * call on one GPU shares that GPU's array, so calls that overlap on one GPU (two streams) race. */
#define VA_SORT_NOALLOC_MAX_PAIRS (1u << 24)
int va_sort_pairs_noalloc(uint32_t *keys, uint32_t *vals, uint32_t *keys_alt, uint32_t *vals_alt,
                          const uint32_t *n_dev, uint32_t cap, int bits, int *result_in_alt, hipStream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* VISOR_BINS_H */
