// runtime_kernels.hip - pipeline-independent device work, compiled ahead of time for gfx950:
// the exclusive scan of per-primitive tile counts, the stable radix sort of (tile, primitive)
// pairs that turns them into API-ordered bins, fills and counter reductions. Every launch here
// is a plain kernel (no library calls, no memset nodes), so a frame's sequence can be captured
// into a HIP graph.
#include <hip/hip_runtime.h>

#include <mutex>

#include "runtime_kernels.h"
#include "texel_decode.h"

typedef uint32_t u32;
typedef uint64_t u64;

// ---------------------------------------------------------------------------------------------
// Wave and block scans (64-lane wavefronts)

// Inclusive prefix sum over the wave's 64 lanes (all active) with DPP moves only, no LDS-routed
// shuffles: a scan inside each row of 16 lanes (row_shr by 1, 2, 4, 8; a lane whose source lies
// before its row's start adds 0), then row_bcast:15 adds lane 15 of rows 0 and 2 to rows 1 and 3,
// and row_bcast:31 adds lane 31 to rows 2 and 3.
__device__ __forceinline__ u32 rk_wave_inclusive(u32 v, u32 lane)
{
  (void)lane;
  u32 x = v;
  x += (u32)__builtin_amdgcn_update_dpp(0, (int)x, 0x111, 0xf, 0xf, false);  // row_shr:1
  x += (u32)__builtin_amdgcn_update_dpp(0, (int)x, 0x112, 0xf, 0xf, false);  // row_shr:2
  x += (u32)__builtin_amdgcn_update_dpp(0, (int)x, 0x114, 0xf, 0xf, false);  // row_shr:4
  x += (u32)__builtin_amdgcn_update_dpp(0, (int)x, 0x118, 0xf, 0xf, false);  // row_shr:8
  x += (u32)__builtin_amdgcn_update_dpp(0, (int)x, 0x142, 0xa, 0xf, false);  // row_bcast:15
  x += (u32)__builtin_amdgcn_update_dpp(0, (int)x, 0x143, 0xc, 0xf, false);  // row_bcast:31
  return x;
}

// Exclusive scan over the block (blockDim.x = 64 * NW threads); s_w holds NW words.
template <u32 NW>
__device__ __forceinline__ u32 rk_block_exclusive(u32 v, u32 *s_w, u32 &total)
{
  const u32 tid = threadIdx.x, lane = tid & 63u, w = tid >> 6;
  const u32 incl = rk_wave_inclusive(v, lane);
  if(lane == 63u)
    s_w[w] = incl;
  __syncthreads();
  u32 off = 0, tot = 0;
#pragma unroll
  for(u32 k = 0; k < NW; k++)
  {
    const u32 x = s_w[k];
    off += k < w ? x : 0u;
    tot += x;
  }
  total = tot;
  __syncthreads();
  return off + incl - v;
}

__device__ __forceinline__ void rk_wave_sync()
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ---------------------------------------------------------------------------------------------
// Exclusive scan of n u32 (reduce, scan of block sums, down-sweep). 2048 items per block.

#define RK_SCAN_T 256u
#define RK_SCAN_I 8u
#define RK_SCAN_B (RK_SCAN_T * RK_SCAN_I)

__global__ __launch_bounds__(256) void va_scan_reduce_kernel(const u32 *__restrict__ in, u32 n, u32 *__restrict__ bsum)
{
  __shared__ u32 s_w[4];
  const u32 base = blockIdx.x * RK_SCAN_B;
  u32 t = 0;
#pragma unroll
  for(u32 i = 0; i < RK_SCAN_I; i++)
  {
    const u32 idx = base + i * RK_SCAN_T + threadIdx.x;
    t += idx < n ? in[idx] : 0u;
  }
  u32 total;
  rk_block_exclusive<4>(t, s_w, total);
  if(threadIdx.x == 0)
    bsum[blockIdx.x] = total;
}

// One workgroup of 1024 threads scans the block sums in place, 1024 at a time.
__global__ __launch_bounds__(1024) void va_scan_top_kernel(u32 *__restrict__ bsum, u32 nb)
{
  __shared__ u32 s_w[16];
  u32 carry = 0;
  for(u32 c = 0; c < nb; c += 1024u)
  {
    const u32 idx = c + threadIdx.x;
    const u32 v = idx < nb ? bsum[idx] : 0u;
    u32 total;
    const u32 ex = rk_block_exclusive<16>(v, s_w, total);
    if(idx < nb)
      bsum[idx] = carry + ex;
    carry += total;
  }
}

// out may alias in: every block reads its whole range into LDS before writing it. With
// prefixed != 0, bsum already holds exclusive block offsets; otherwise each block sums the
// block totals before it itself (cheap for up to a few thousand blocks; saves a launch).
__global__ __launch_bounds__(256) void va_scan_down_kernel(const u32 *in, u32 n, const u32 *__restrict__ bsum, u32 *out,
                                                           u32 prefixed)
{
  __shared__ u32 s_v[RK_SCAN_B];
  __shared__ u32 s_w[4];
  const u32 base = blockIdx.x * RK_SCAN_B;
  u32 boff;
  if(prefixed)
    boff = bsum[blockIdx.x];
  else
  {
    u32 pre = 0;
    for(u32 i = threadIdx.x; i < blockIdx.x; i += RK_SCAN_T)
      pre += bsum[i];
    rk_block_exclusive<4>(pre, s_w, boff);
  }
#pragma unroll
  for(u32 i = 0; i < RK_SCAN_I; i++)
  {
    const u32 idx = base + i * RK_SCAN_T + threadIdx.x;
    s_v[i * RK_SCAN_T + threadIdx.x] = idx < n ? in[idx] : 0u;
  }
  __syncthreads();
  u32 v[RK_SCAN_I], t = 0;
#pragma unroll
  for(u32 i = 0; i < RK_SCAN_I; i++)
  {
    v[i] = s_v[threadIdx.x * RK_SCAN_I + i];
    t += v[i];
  }
  u32 total;
  u32 run = boff + rk_block_exclusive<4>(t, s_w, total);
#pragma unroll
  for(u32 i = 0; i < RK_SCAN_I; i++)
  {
    s_v[threadIdx.x * RK_SCAN_I + i] = run;
    run += v[i];
  }
  __syncthreads();
#pragma unroll
  for(u32 i = 0; i < RK_SCAN_I; i++)
  {
    const u32 idx = base + i * RK_SCAN_T + threadIdx.x;
    if(idx < n)
      out[idx] = s_v[i * RK_SCAN_T + threadIdx.x];
  }
}

static u32 scan_blocks(u32 n) { return (n + RK_SCAN_B - 1) / RK_SCAN_B; }

extern "C" size_t va_scan_temp_words(uint32_t n) { return (size_t)scan_blocks(n) + 64; }

extern "C" hipError_t va_scan_u32(uint32_t *temp, const uint32_t *in, uint32_t *out, uint32_t n, hipStream_t stream)
{
  if(n == 0)
    return hipSuccess;
  const u32 nb = scan_blocks(n);
  const u32 prefixed = nb > 2048u ? 1u : 0u;  // many blocks: scan the block totals separately
  (void)hipGetLastError();  // report this launch's error only
  hipLaunchKernelGGL(va_scan_reduce_kernel, dim3(nb), dim3(RK_SCAN_T), 0, stream, in, n, temp);
  if(prefixed)
    hipLaunchKernelGGL(va_scan_top_kernel, dim3(1), dim3(1024), 0, stream, temp, nb);
  hipLaunchKernelGGL(va_scan_down_kernel, dim3(nb), dim3(RK_SCAN_T), 0, stream, in, n, (const u32 *)temp, out,
                     prefixed);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------------------------
// Stable LSD radix sort of (key, value) pairs, 8-bit digits, 2048 pairs per workgroup (twice the
// workgroups of 4096-pair ones: binning 0.105 vs 0.110 ms on 2.1M pairs; 1024: 0.113):
//   histogram per (digit, workgroup) -> exclusive scan of each digit's row (one workgroup per
//   digit; the scatter adds the scan of the digit totals) -> stable scatter.
// The pair count is read on the device (min(*n_dev, cap)), so the launches are sized by the
// capacity and need no host readback.

#define RK_SORT_T 256u
#define RK_SORT_I 8u
#define RK_SORT_B (RK_SORT_T * RK_SORT_I)

// Workgroup i of a sort kernel takes block rk_sort_block(i, n): workgroups are dispatched
// round-robin over the 8 XCDs, each with its own L2, and this deals runs of 16 consecutive blocks
// to the XCDs in turn (the tail past the last whole round keeps i). The 16 words that a run's
// blocks write next to each other in a digit's row of the histogram (and the neighbouring runs
// of a digit in the scatter's output) then meet in one L2 instead of as partial cache lines
// from eight.
__device__ __forceinline__ u32 rk_sort_block(u32 i, u32 n)
{
#ifdef RK_SORT_NO_XCD_RUNS
  return i;
#else
  const u32 full = n & ~127u;
  if(i >= full)
    return i;
  const u32 x = i & 7u, s = i >> 3;  // s-th workgroup of XCD x
  return (((s >> 4) << 3) + x) * 16u + (s & 15u);
#endif
}

__global__ __launch_bounds__(256) void va_sort_hist_kernel(const u32 *__restrict__ keys, const u32 *__restrict__ n_dev,
                                                           u32 cap, u32 shift, u32 *__restrict__ hist, u32 nbs)
{
  __shared__ u32 s_h[256];
  s_h[threadIdx.x] = 0u;
  __syncthreads();
  const u32 n = min(*n_dev, cap);
  const u32 blk = rk_sort_block(blockIdx.x, nbs);
  const u32 base = blk * RK_SORT_B;
  // the thread's eight keys are loaded before the first LDS atomic, so their loads are in flight
  // together
  u32 k[RK_SORT_I];
#pragma unroll
  for(u32 i = 0; i < RK_SORT_I; i++)
  {
    const u32 idx = base + i * RK_SORT_T + threadIdx.x;
    k[i] = idx < n ? keys[idx] : 0xffffffffu;
  }
#pragma unroll
  for(u32 i = 0; i < RK_SORT_I; i++)
    if(base + i * RK_SORT_T + threadIdx.x < n)
      atomicAdd(&s_h[(k[i] >> shift) & 255u], 1u);
  __syncthreads();
  hist[threadIdx.x * nbs + blk] = s_h[threadIdx.x];
}

// Wave w of the workgroup ranks pairs [base + 512 w, base + 512 (w + 1)) in order, 64 at a time:
// lanes with equal digits find each other with eight ballots, and per-wave digit counters in LDS
// carry the rank across steps. Waves are then ordered by a per-digit prefix over the waves. The
// ranked pairs are staged in LDS in sorted order (keys, then values through the same buffer)
// and stored from there, consecutive threads to consecutive addresses within each digit's run.
__global__ __launch_bounds__(256) void va_sort_scatter_kernel(const u32 *__restrict__ kin, const u32 *__restrict__ vin,
                                                              u32 *__restrict__ kout, u32 *__restrict__ vout,
                                                              const u32 *__restrict__ n_dev, u32 cap, u32 shift,
                                                              const u32 *__restrict__ hist, const u32 *__restrict__ tot,
                                                              u32 nbs)
{
  __shared__ u32 s_cnt[4][256];
  __shared__ u32 s_goff[256];
  __shared__ u32 s_dstart[256];
  __shared__ u32 s_stage[RK_SORT_B];
  __shared__ u32 s_w[4];
  const u32 tid = threadIdx.x, lane = tid & 63u, w = tid >> 6;
  const u32 n = min(*n_dev, cap);
  const u32 blk = rk_sort_block(blockIdx.x, nbs);
  const u32 base = blk * RK_SORT_B;
  if(base >= n)
    return;
  // The wave's 512 pairs are loaded first, all sixteen loads per lane in flight together: the
  // wave syncs of the ranking loop below order memory, so loads inside it would be issued one
  // step at a time, each step waiting out a full memory latency.
  u32 key[RK_SORT_I], val[RK_SORT_I], rank[RK_SORT_I];
#pragma unroll
  for(u32 j = 0; j < RK_SORT_I; j++)
  {
    const u32 idx = base + w * (RK_SORT_I * 64u) + j * 64u + lane;
    key[j] = idx < n ? kin[idx] : 0u;
    val[j] = idx < n ? vin[idx] : 0u;
  }
  // this block's first slot for digit tid: the pairs of lower digits (an exclusive scan of the
  // digit totals) + those of digit tid in earlier blocks (the row scan)
  {
    u32 all;
    s_goff[tid] = rk_block_exclusive<4>(tot[tid], s_w, all) + hist[tid * nbs + blk];
  }
#pragma unroll
  for(u32 k = 0; k < 4u; k++)
    s_cnt[k][tid] = 0u;
  __syncthreads();
  const u64 below = lane == 0u ? 0ull : (~0ull >> (64u - lane));
#pragma unroll
  for(u32 j = 0; j < RK_SORT_I; j++)
  {
    const u32 idx = base + w * (RK_SORT_I * 64u) + j * 64u + lane;
    const bool valid = idx < n;
    const u32 d = (key[j] >> shift) & 255u;
    u64 peers = __ballot(valid);
#pragma unroll
    for(u32 b = 0; b < 8u; b++)
    {
      const u64 m = __ballot((d >> b) & 1u);
      peers &= ((d >> b) & 1u) ? m : ~m;
    }
    const u32 r = (u32)__popcll(peers & below);
    const u32 cur = valid ? s_cnt[w][d] : 0u;
    rk_wave_sync();
    if(valid && r == 0u)
      s_cnt[w][d] = cur + (u32)__popcll(peers);
    rk_wave_sync();
    rank[j] = valid ? cur + r : 0xffffffffu;
  }
  __syncthreads();
  u32 dcount = 0;  // this workgroup's pairs with digit tid
#pragma unroll
  for(u32 k = 0; k < 4u; k++)
  {
    const u32 c = s_cnt[k][tid];
    s_cnt[k][tid] = dcount;
    dcount += c;
  }
  // The workgroup's pairs in sorted order are staged in LDS, so that the global stores go out
  // in runs of consecutive addresses: local index i with digit d goes to i + (goff[d] - start[d]).
  u32 nvalid;
  const u32 dstart = rk_block_exclusive<4>(dcount, s_w, nvalid);
  s_dstart[tid] = dstart;
  s_goff[tid] -= dstart;
  __syncthreads();
#pragma unroll
  for(u32 j = 0; j < RK_SORT_I; j++)
    if(rank[j] != 0xffffffffu)
    {
      const u32 d = (key[j] >> shift) & 255u;
      rank[j] += s_dstart[d] + s_cnt[w][d];  // now the local sorted index
      s_stage[rank[j]] = key[j];
    }
  __syncthreads();
  u32 gpos[RK_SORT_I];
#pragma unroll
  for(u32 j = 0; j < RK_SORT_I; j++)
  {
    const u32 i = j * RK_SORT_T + tid;
    if(i < nvalid)
    {
      const u32 k = s_stage[i];
      gpos[j] = i + s_goff[(k >> shift) & 255u];
      kout[gpos[j]] = k;
    }
  }
  __syncthreads();
#pragma unroll
  for(u32 j = 0; j < RK_SORT_I; j++)
    if(rank[j] != 0xffffffffu)
      s_stage[rank[j]] = val[j];
  __syncthreads();
#pragma unroll
  for(u32 j = 0; j < RK_SORT_I; j++)
  {
    const u32 i = j * RK_SORT_T + tid;
    if(i < nvalid)
      vout[gpos[j]] = s_stage[i];
  }
}

// Exclusive scan of each digit's row of the histogram in place (row d = the nbs blocks' counts of
// digit d, contiguous in the digit-major layout), and tot[d] = the row's total; the scatter adds
// the scan of the 256 totals itself. One workgroup per digit, eight counts per thread per step.
#define RK_ROW_I 8u
__global__ __launch_bounds__(256) void va_sort_rowscan_kernel(u32 *__restrict__ hist, u32 nbs, u32 *__restrict__ tot)
{
  __shared__ u32 s_w[4];
  u32 *row = hist + (u64)blockIdx.x * nbs;
  u32 carry = 0;
  for(u32 c = 0; c < nbs; c += 256u * RK_ROW_I)
  {
    const u32 i0 = c + threadIdx.x * RK_ROW_I;
    u32 v[RK_ROW_I], t = 0;
#pragma unroll
    for(u32 k = 0; k < RK_ROW_I; k++)
    {
      v[k] = i0 + k < nbs ? row[i0 + k] : 0u;
      t += v[k];
    }
    u32 total;
    u32 run = carry + rk_block_exclusive<4>(t, s_w, total);
#pragma unroll
    for(u32 k = 0; k < RK_ROW_I; k++)
    {
      if(i0 + k < nbs)
        row[i0 + k] = run;
      run += v[k];
    }
    carry += total;
  }
  if(threadIdx.x == 0u)
    tot[blockIdx.x] = carry;
}

static u32 sort_blocks(u32 cap) { return (cap + RK_SORT_B - 1) / RK_SORT_B; }

extern "C" size_t va_sort_temp_words(uint32_t cap) { return 256u * (size_t)sort_blocks(cap) + 256u; }

extern "C" hipError_t va_sort_pairs(uint32_t *temp, uint32_t *keys, uint32_t *vals, uint32_t *keys_alt,
                                    uint32_t *vals_alt, const uint32_t *n_dev, uint32_t cap, int bits,
                                    int *result_in_alt, hipStream_t stream)
{
  *result_in_alt = 0;
  if(cap == 0)
    return hipSuccess;
  const u32 nbs = sort_blocks(cap), nh = 256u * nbs;
  u32 *hist = temp, *tot = temp + nh;
  u32 *ki = keys, *vi = vals, *ko = keys_alt, *vo = vals_alt;
  for(int shift = 0; shift < bits; shift += 8)
  {
    (void)hipGetLastError();  // report this launch's error only
    hipLaunchKernelGGL(va_sort_hist_kernel, dim3(nbs), dim3(RK_SORT_T), 0, stream, (const u32 *)ki, n_dev, cap,
                       (u32)shift, hist, nbs);
    hipLaunchKernelGGL(va_sort_rowscan_kernel, dim3(256), dim3(256), 0, stream, hist, nbs, tot);
    hipLaunchKernelGGL(va_sort_scatter_kernel, dim3(nbs), dim3(RK_SORT_T), 0, stream, (const u32 *)ki,
                       (const u32 *)vi, ko, vo, n_dev, cap, (u32)shift, (const u32 *)hist, (const u32 *)tot, nbs);
    u32 *t = ki; ki = ko; ko = t;
    t = vi; vi = vo; vo = t;
    *result_in_alt ^= 1;
  }
  return hipGetLastError();
}

// ---------------------------------------------------------------------------------------------
// The same sort without a workspace argument and without allocating: the histogram words and
// digit totals live in `scratch`, a file-scope device array sized for VA_SORT_NOALLOC_MAX_PAIRS pairs.
// Host code reaches the array through the address hipGetSymbolAddress returns for the current
// device (the array's name is only a host-side stand-in), looked up once per device on the first
// call there. Every call on one device shares that device's array; nothing in it needs clearing
// (the histogram kernel writes every word the scan and scatter read), and calls that overlap on
// one device (two streams) race on it.

static constexpr size_t kNoallocSortWords = 256u * ((VA_SORT_NOALLOC_MAX_PAIRS + RK_SORT_B - 1) / RK_SORT_B) + 256u;
static __device__ __align__(256) uint32_t scratch[kNoallocSortWords];

#define RK_MAX_DEVICES 64
static std::mutex g_scratch_mu;
static uint32_t *g_scratch_ptr[RK_MAX_DEVICES];

static hipError_t rk_scratch_for_current_device(uint32_t **out)
{
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if(e != hipSuccess)
    return e;
  if(dev < 0 || dev >= RK_MAX_DEVICES)
    return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if(!g_scratch_ptr[dev])
  {
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if(e != hipSuccess)
      return e;
    g_scratch_ptr[dev] = (uint32_t *)p;
  }
  *out = g_scratch_ptr[dev];
  return hipSuccess;
}

extern "C" hipError_t va_sort_pairs_noalloc(uint32_t *keys, uint32_t *vals, uint32_t *keys_alt, uint32_t *vals_alt,
                                            const uint32_t *n_dev, uint32_t cap, int bits, int *result_in_alt,
                                            hipStream_t stream)
{
  *result_in_alt = 0;
  if(cap > VA_SORT_NOALLOC_MAX_PAIRS)
    return hipErrorInvalidValue;
  if(va_sort_temp_words(cap) > kNoallocSortWords)  // holds for every cap within the maximum
    return hipErrorInvalidValue;
  uint32_t *temp = nullptr;
  const hipError_t e = rk_scratch_for_current_device(&temp);
  if(e != hipSuccess)
    return e;
  return va_sort_pairs(temp, keys, vals, keys_alt, vals_alt, n_dev, cap, bits, result_in_alt, stream);
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

// Device wall clock (constant rate, hipDeviceAttributeWallClockRate kHz) at this point of the
// stream: stage boundaries inside a captured graph, where event records give no timings.
__global__ __launch_bounds__(64) void va_stamp_kernel(uint64_t *out)
{
  const uint64_t t = (uint64_t)wall_clock64();
  if(threadIdx.x == 0)
    out[0] = t;
}

extern "C" hipError_t va_stamp_async(uint64_t *out, hipStream_t stream)
{
  (void)hipGetLastError();  // report this launch's error only
  hipLaunchKernelGGL(va_stamp_kernel, dim3(1), dim3(64), 0, stream, out);
  return hipGetLastError();
}

// One thread per shadow texel, grid-stride over the level; blockIdx.y = layer * mips + level.
__global__ __launch_bounds__(256) void va_transcode_kernel(const uint8_t *__restrict__ src, uint8_t *__restrict__ dst,
                                                           VaTranscode t)
{
  const uint32_t level = blockIdx.y % t.mips, layer = blockIdx.y / t.mips;
  const uint32_t w = t.width >> level ? t.width >> level : 1u;
  const uint32_t h = t.height >> level ? t.height >> level : 1u;
  const uint64_t n = (uint64_t)w * h;
  const uint64_t sub = t.src_slice * layer + t.src_level_off[level];
  uint32_t *out = (uint32_t *)(dst + t.dst_slice * layer + t.dst_level_off[level]);
  for(uint64_t i = (uint64_t)blockIdx.x * 256u + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256u)
  {
    const uint32_t y = (uint32_t)(i / w), x = (uint32_t)(i - (uint64_t)y * w);
    out[i] = va_texel_bytes(src, t.src_bytes, sub, w, t.bpp, t.fmt, x, y);
  }
}

extern "C" hipError_t va_transcode_async(const uint8_t *src, uint8_t *dst, const VaTranscode *t, hipStream_t stream)
{
  if(t->mips == 0u || t->mips > VA_TRANSCODE_MAX_LEVELS || t->layers == 0u || t->width == 0u || t->height == 0u ||
     (uint64_t)t->mips * t->layers > 65535u)
    return hipErrorInvalidValue;
  const uint64_t texels = (uint64_t)t->width * t->height;
  uint64_t bx = (texels + 255u) / 256u;
  if(bx > 4096u)
    bx = 4096u;
  (void)hipGetLastError();
  hipLaunchKernelGGL(va_transcode_kernel, dim3((unsigned)bx, t->mips * t->layers), dim3(256), 0, stream, src, dst, *t);
  return hipGetLastError();
}

// End of a pass: sums the counter rows of setup and raster (nrows rows of 4 u64 per stage, setup's
// first; in VaCounters order from tris_in and from pixels_written), clears them for the next
// pass, and writes the sums, the pair total and the stage stamps straight into host-visible
// memory, then the pass's sequence number: *seq counts the passes that reached this kernel, and
// out[14] receives the new count once every other word of out is visible to the host, so the
// host learns that the pass is complete by reading host memory (no event or stream wait, which
// would make every pass end with a system-scope release).
// out: [0] pair total, [2..9] counter sums, [10..12] stamps 0-2, [13] this launch's start,
// [14] sequence number.
constexpr uint32_t kFinishMaxRows = 64;

__global__ __launch_bounds__(256) void va_finish_kernel(uint64_t *__restrict__ rows, uint32_t nrows,
                                                        const uint32_t *__restrict__ pair_total,
                                                        const uint64_t *__restrict__ stamps, uint64_t *out,
                                                        uint32_t *__restrict__ seq)
{
  __shared__ uint64_t s_v[2 * kFinishMaxRows * 4];
  const uint64_t t_start = (uint64_t)wall_clock64();
  // the words that do not depend on the sums are loaded together with the rows
  uint64_t w0 = 0, w10 = 0, w11 = 0, w12 = 0;
  uint32_t n = 0;
  if(threadIdx.x == 0u)
  {
    w0 = *pair_total;
    w10 = stamps[0];
    w11 = stamps[1];
    w12 = stamps[2];
    n = *seq + 1u;  // one workgroup per launch, launches in stream order
  }
  const uint32_t nw = 2u * nrows * 4u;
  for(uint32_t i = threadIdx.x; i < nw; i += 256u)
  {
    s_v[i] = rows[i];
    rows[i] = 0ull;
  }
  __syncthreads();
  if(threadIdx.x < 8u)
  {
    const uint32_t stage = threadIdx.x >> 2, c = threadIdx.x & 3u;
    uint64_t t = 0;
    for(uint32_t r = 0; r < nrows; r++)
      t += s_v[(stage * nrows + r) * 4u + c];
    out[2 + threadIdx.x] = t;
  }
  if(threadIdx.x == 0u)
  {
    out[0] = w0;
    out[10] = w10;
    out[11] = w11;
    out[12] = w12;
    out[13] = t_start;
    *seq = n;
  }
  // Every word of out is written by the first wave: only its lanes fence their words before the
  // sequence number (a system-scope fence also writes back the L2, once per wave that runs it).
  if(threadIdx.x < 64u)
    __threadfence_system();
  __syncthreads();
  if(threadIdx.x == 0u)
  {
    __threadfence_system();
    *(volatile uint64_t *)&out[14] = n;
  }
}

extern "C" hipError_t va_finish_async(uint64_t *rows, uint32_t nrows, const uint32_t *pair_total,
                                      const uint64_t *stamps, uint64_t *out, uint32_t *seq, hipStream_t stream)
{
  if(nrows == 0u || nrows > kFinishMaxRows)
    return hipErrorInvalidValue;
  (void)hipGetLastError();
  hipLaunchKernelGGL(va_finish_kernel, dim3(1), dim3(256), 0, stream, rows, nrows, pair_total, stamps, out, seq);
  return hipGetLastError();
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
  (void)hipGetLastError();  // report this launch's error only
  hipLaunchKernelGGL(va_fill32_kernel, dim3(fill_blocks(count)), dim3(256), 0, stream, (uint32_t *)p, count, value);
  return hipGetLastError();
}

extern "C" hipError_t va_fill_u64_async(void *p, uint64_t count, uint32_t lo, uint32_t hi, hipStream_t stream)
{
  if(count == 0)
    return hipSuccess;
  (void)hipGetLastError();  // report this launch's error only
  hipLaunchKernelGGL(va_fill64_kernel, dim3(fill_blocks(count)), dim3(256), 0, stream, (uint2 *)p, count,
                     make_uint2(lo, hi));
  return hipGetLastError();
}

