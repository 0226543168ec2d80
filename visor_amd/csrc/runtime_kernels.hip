// runtime_kernels.hip - pipeline-independent device work, compiled ahead of time for gfx950:
// the exclusive scan of per-primitive tile counts, the stable radix sort of (tile, primitive)
// pairs that turns them into API-ordered bins, and render-target clears.
#include <cstring>

#include <hip/hip_runtime.h>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "runtime_kernels.h"

extern "C" hipError_t va_scan_u32(void *temp, size_t *temp_bytes, const uint32_t *in, uint32_t *out,
                                  uint32_t n, hipStream_t stream)
{
  return rocprim::exclusive_scan(temp, *temp_bytes, in, out, 0u, (size_t)n, rocprim::plus<uint32_t>(), stream);
}

// LSD radix sort is stable: pairs emitted in primitive order stay in primitive order per tile.
extern "C" hipError_t va_sort_pairs(void *temp, size_t *temp_bytes, const uint32_t *keys_in,
                                    uint32_t *keys_out, const uint32_t *vals_in, uint32_t *vals_out,
                                    uint32_t n, int end_bit, hipStream_t stream)
{
  return rocprim::radix_sort_pairs(temp, *temp_bytes, keys_in, keys_out, vals_in, vals_out, (size_t)n, 0,
                                   end_bit, stream);
}

__global__ __launch_bounds__(256) void va_fill32_kernel(uint32_t *p, uint64_t n, uint32_t v)
{
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for(; i < n; i += stride)
    p[i] = v;
}

__global__ __launch_bounds__(256) void va_fill64_kernel(uint2 *p, uint64_t n, uint2 v)
{
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for(; i < n; i += stride)
    p[i] = v;
}

static unsigned fill_blocks(uint64_t n)
{
  uint64_t b = (n + 255) / 256;
  return (unsigned)(b > 8192 ? 8192 : (b ? b : 1));
}

// A fill kernel rather than hipMemsetAsync so the sequence stays graph-capture safe.
extern "C" hipError_t va_fill_u32_async(void *p, uint64_t count, uint32_t value, hipStream_t stream)
{
  if(count == 0)
    return hipSuccess;
  hipLaunchKernelGGL(va_fill32_kernel, dim3(fill_blocks(count)), dim3(256), 0, stream, (uint32_t *)p, count, value);
  return hipGetLastError();
}

extern "C" hipError_t va_fill_u64_async(void *p, uint64_t count, uint32_t lo, uint32_t hi, hipStream_t stream)
{
  if(count == 0)
    return hipSuccess;
  hipLaunchKernelGGL(va_fill64_kernel, dim3(fill_blocks(count)), dim3(256), 0, stream, (uint2 *)p, count,
                     make_uint2(lo, hi));
  return hipGetLastError();
}

// out[c] += sum over rows of part[row * 4 + c] (four columns). Each workgroup reduces a slice
// of rows and adds its four totals atomically: a few dozen atomics instead of one per wave.
__global__ __launch_bounds__(256) void va_reduce4_kernel(const uint64_t *part, uint32_t rows, uint64_t *out)
{
  __shared__ uint64_t s[4][256];
  uint64_t t[4] = {0, 0, 0, 0};
  for(uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < rows; r += gridDim.x * blockDim.x)
  {
    const ulonglong2 *p = (const ulonglong2 *)(part + (uint64_t)r * 4u);
    ulonglong2 a = p[0], b = p[1];
    t[0] += a.x;
    t[1] += a.y;
    t[2] += b.x;
    t[3] += b.y;
  }
  for(int c = 0; c < 4; c++)
    s[c][threadIdx.x] = t[c];
  __syncthreads();
  for(uint32_t o = blockDim.x / 2; o > 0; o >>= 1)
  {
    if(threadIdx.x < o)
      for(int c = 0; c < 4; c++)
        s[c][threadIdx.x] += s[c][threadIdx.x + o];
    __syncthreads();
  }
  if(threadIdx.x < 4)
    atomicAdd((unsigned long long *)&out[threadIdx.x], (unsigned long long)s[threadIdx.x][0]);
}

// Adds the column sums of a rows x 4 partials array into out[0..3] (out must start at zero).
extern "C" hipError_t va_reduce_cols_async(const uint64_t *part, uint32_t rows, uint32_t cols, uint64_t *out,
                                           hipStream_t stream)
{
  if(cols != 4)
    return hipErrorInvalidValue;
  uint32_t blocks = (rows + 1023) / 1024;
  if(blocks < 1) blocks = 1;
  if(blocks > 64) blocks = 64;
  hipLaunchKernelGGL(va_reduce4_kernel, dim3(blocks), dim3(256), 0, stream, part, rows, out);
  return hipGetLastError();
}
