// runtime_kernels.h - ahead-of-time device helpers (runtime_kernels.hip).
#pragma once

#include <hip/hip_runtime_api.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

hipError_t va_scan_u32(void *temp, size_t *temp_bytes, const uint32_t *in, uint32_t *out, uint32_t n,
                       hipStream_t stream);
hipError_t va_sort_pairs(void *temp, size_t *temp_bytes, const uint32_t *keys_in, uint32_t *keys_out,
                         const uint32_t *vals_in, uint32_t *vals_out, uint32_t n, int end_bit,
                         hipStream_t stream);
hipError_t va_fill_u32_async(void *p, uint64_t count, uint32_t value, hipStream_t stream);
hipError_t va_fill_u64_async(void *p, uint64_t count, uint32_t lo, uint32_t hi, hipStream_t stream);

#ifdef __cplusplus
}
#endif

#ifdef __cplusplus
extern "C" {
#endif
hipError_t va_reduce_cols_async(const uint64_t *part, uint32_t rows, uint32_t cols, uint64_t *out, hipStream_t stream);
#ifdef __cplusplus
}
#endif
