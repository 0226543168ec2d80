// pipeline_kernels.hip - the draw-execution kernels of one pipeline module.
//
// Included after the translated shaders, which define:
//   VA_NSLOTS, VA_MAXNV, VA_REC_STRIDE, VA_FS_DERIV
//   va_vs(slot, draw, vertexIndex, pos[4], vary[VA_MAXNV])   vertex program + varying gather
//   va_fs(slot, draw, vary, b[3][4], out[4]) -> keep          fragment program
//
// Stages (reference rasterizer.cpp:363-696 in GPU form):
//   va_setup   one thread per primitive: index fetch, 3 VS invocations, ToWindow, area / facing
//              / cull, clamped bbox, 64..128-byte triangle record, per-primitive tile count
//   va_emit    one thread per (tile, primitive) pair, keys = global tile id, values = primitive
//   (stable radix sort by tile, in the host runtime)  -> bins keep API primitive order
//   va_ranges  per-tile [start, end) into the sorted values
//   va_raster  one workgroup per tile; each wave owns an 8x8 pixel block with colour and depth in
//              registers, walks the tile's bin in order: integer edge functions + ballot,
//              early depth, fragment program, blend, and writes the block back once.

#ifndef VA_TILE_MAX_SHIFT
#define VA_TILE_MAX_SHIFT 5
#endif

VA_INLINE u32 va_find_draw(const VaDraw *__restrict__ draws, u32 ndraws, u32 g)
{
  u32 lo = 0, hi = ndraws;
  while(hi - lo > 1u)
  {
    u32 mid = (lo + hi) >> 1;
    if(draws[mid].tri_base <= g) lo = mid; else hi = mid;
  }
  return lo;
}

VA_INLINE u32 va_find_frame(const VaFrame *__restrict__ frames, u32 nframes, u32 tile)
{
  u32 lo = 0, hi = nframes;
  while(hi - lo > 1u)
  {
    u32 mid = (lo + hi) >> 1;
    if(frames[mid].tile_base <= tile) lo = mid; else hi = mid;
  }
  return lo;
}

VA_INLINE u64 va_wave_sum(u64 v)
{
#pragma unroll
  for(int o = 32; o > 0; o >>= 1)
    v += __shfl_xor(v, o);
  return v;
}

// reference rasterizer.cpp:100-119 (GetIndex, vertexOffset not applied)
VA_INLINE u32 va_index(const VaDraw &d, u32 i)
{
  if(!(d.flags & VA_DRAW_INDEXED)) return i;
  // an index past the end of the bound index buffer reads as 0 instead of faulting
  if(d.flags & VA_DRAW_INDEX32) return 4ull * i + 4ull <= d.ib_bytes ? ((const u32 *)d.ib)[i] : 0u;
  return 2ull * i + 2ull <= d.ib_bytes ? ((const u16 *)d.ib)[i] : 0u;
}

extern "C" __global__ __launch_bounds__(256) void va_setup(
    const VaDraw *__restrict__ draws, u32 ndraws, const VaFrame *__restrict__ frames, u32 total_tris,
    u32 tile_shift, u32 *__restrict__ rec, u32 *__restrict__ tile_count, VaCounters *__restrict__ ctr)
{
  const u32 g = blockIdx.x * blockDim.x + threadIdx.x;
  u64 n_in = 0, n_out = 0, n_tested = 0, n_blocks = 0;
  if(g < total_tris)
  {
    const VaDraw &d = draws[va_find_draw(draws, ndraws, g)];
    const VaFrame &fr = frames[d.frame];
    const u32 prim = g - d.tri_base;
    u32 p0, p1, p2;
    if(d.flags & VA_DRAW_STRIP)
    {
      // reference rasterizer.cpp:150-232: (N,N+1,N+2), (N+2,N+1,N+3), ...
      if(prim & 1u) { p0 = prim + 1u; p1 = prim; } else { p0 = prim; p1 = prim + 1u; }
      p2 = prim + 2u;
    }
    else
    {
      p0 = 3u * prim; p1 = p0 + 1u; p2 = p0 + 2u;
    }
    float pos[3][4];
    float vary[3][VA_MAXNV > 0 ? VA_MAXNV : 1];
    va_vs(d.slot, d, va_index(d, d.first + p0), pos[0], vary[0]);
    va_vs(d.slot, d, va_index(d, d.first + p1), pos[1], vary[1]);
    va_vs(d.slot, d, va_index(d, d.first + p2), pos[2], vary[2]);

    // ToWindow (rasterizer.cpp:239-253)
    const float Wf = (float)fr.width, Hf = (float)fr.height;
    i32 X[3], Y[3];
#pragma unroll
    for(int k = 0; k < 3; k++)
    {
      X[k] = va_f2i((pos[k][0] / pos[k][3] + 1.0f) * 0.5f * Wf);
      Y[k] = va_f2i((pos[k][1] * -1.0f / pos[k][3] + 1.0f) * 0.5f * Hf);
    }
    n_in = 1;
    u32 cnt = 0;
    // double_triarea (rasterizer.cpp:272-275), wrapping int32
    const i32 area2 = (i32)((u32)(X[1] - X[0]) * (u32)(Y[2] - Y[0]) - (u32)(Y[1] - Y[0]) * (u32)(X[2] - X[0]));
    if(area2 != 0)
    {
      // facing and cull (rasterizer.cpp:401-424)
      i32 flipped = area2, barymul = 1;
      bool culled = false;
      if(d.front_face == 1u) { barymul = -1; flipped = -flipped; }
      if(flipped > 0 && (d.cull_mode & 1u)) culled = true;
      if(!culled && flipped < 0)
      {
        if(d.cull_mode & 2u) culled = true;
        else { barymul = -barymul; flipped = -flipped; }
      }
      if(!culled)
      {
        n_out = 1;
        i32 minx = min(X[0], min(X[1], X[2])), maxx = max(X[0], max(X[1], X[2]));
        i32 miny = min(Y[0], min(Y[1], Y[2])), maxy = max(Y[0], max(Y[1], Y[2]));
        minx = max(0, minx);
        miny = max(0, miny);
        maxx = min((i32)fr.width - 1, maxx);
        maxy = min((i32)fr.height - 1, maxy);
        // 32x32 job count as the reference computes it (C truncating division), for the counter
        i32 xb = 1 + (maxx - minx) / 32, yb = 1 + (maxy - miny) / 32;
        n_blocks = (u64)(xb > 0 ? xb : 0) * (u64)(yb > 0 ? yb : 0);
        if(maxx > minx && maxy > miny)
        {
          n_tested = (u64)(maxx - minx) * (u64)(maxy - miny);
          u32 *r = rec + (u64)g * VA_REC_STRIDE;
          const u32 bm = (u32)barymul;
          r[VA_R_X0] = (u32)X[0];
          r[VA_R_Y0] = (u32)Y[0];
          // barymul folded into the edge coefficients: exact in wrapping arithmetic
          r[VA_R_ABX] = (u32)(X[1] - X[0]) * bm;
          r[VA_R_ABY] = (u32)(Y[1] - Y[0]) * bm;
          r[VA_R_ACX] = (u32)(X[2] - X[0]) * bm;
          r[VA_R_ACY] = (u32)(Y[2] - Y[0]) * bm;
          r[VA_R_AREA] = (u32)flipped;
          r[VA_R_INVAREA] = va_f2u(1.0f / (float)flipped);
          const float iw0 = 1.0f / pos[0][3], iw1 = 1.0f / pos[1][3], iw2 = 1.0f / pos[2][3];
          r[VA_R_Z0 + 0] = va_f2u(pos[0][2] * iw0);
          r[VA_R_Z0 + 1] = va_f2u(pos[1][2] * iw1);
          r[VA_R_Z0 + 2] = va_f2u(pos[2][2] * iw2);
          r[VA_R_IW0 + 0] = va_f2u(iw0);
          r[VA_R_IW0 + 1] = va_f2u(iw1);
          r[VA_R_IW0 + 2] = va_f2u(iw2);
          r[VA_R_BBMIN] = (u32)minx | ((u32)miny << 16);
          r[VA_R_BBMAX] = (u32)maxx | ((u32)maxy << 16);
          r[VA_R_DRAW] = va_find_draw(draws, ndraws, g);
#pragma unroll
          for(int j = 0; j < VA_MAXNV; j++)
          {
            r[VA_REC_HDR + 3 * j + 0] = va_f2u(vary[0][j]);
            r[VA_REC_HDR + 3 * j + 1] = va_f2u(vary[1][j]);
            r[VA_REC_HDR + 3 * j + 2] = va_f2u(vary[2][j]);
          }
          const u32 tx0 = (u32)minx >> tile_shift, tx1 = (u32)(maxx - 1) >> tile_shift;
          const u32 ty0 = (u32)miny >> tile_shift, ty1 = (u32)(maxy - 1) >> tile_shift;
          cnt = (tx1 - tx0 + 1u) * (ty1 - ty0 + 1u);
        }
      }
    }
    tile_count[g] = cnt;
  }
  n_in = va_wave_sum(n_in);
  n_out = va_wave_sum(n_out);
  n_tested = va_wave_sum(n_tested);
  n_blocks = va_wave_sum(n_blocks);
  if((threadIdx.x & 63u) == 0u && n_in)
  {
    atomicAdd(&ctr->tris_in, n_in);
    atomicAdd(&ctr->tris_out, n_out);
    atomicAdd(&ctr->pixels_tested, n_tested);
    atomicAdd(&ctr->blocks, n_blocks);
  }
}

// One thread per (tile, primitive) pair. offsets = exclusive scan of tile counts (ntris + 1).
extern "C" __global__ __launch_bounds__(256) void va_emit(
    const u32 *__restrict__ rec, const u32 *__restrict__ offsets, u32 total_tris, u32 total_pairs,
    const VaDraw *__restrict__ draws, const VaFrame *__restrict__ frames, u32 tile_shift,
    u32 *__restrict__ keys, u32 *__restrict__ vals)
{
  const u32 p = blockIdx.x * blockDim.x + threadIdx.x;
  if(p >= total_pairs) return;
  u32 lo = 0, hi = total_tris;
  while(hi - lo > 1u)
  {
    u32 mid = (lo + hi) >> 1;
    if(offsets[mid] <= p) lo = mid; else hi = mid;
  }
  const u32 t = lo;
  const u32 k = p - offsets[t];
  const u32 *r = rec + (u64)t * VA_REC_STRIDE;
  const u32 bmin = r[VA_R_BBMIN], bmax = r[VA_R_BBMAX];
  const u32 tx0 = (bmin & 0xffffu) >> tile_shift, tx1 = ((bmax & 0xffffu) - 1u) >> tile_shift;
  const u32 ty0 = (bmin >> 16) >> tile_shift;
  const u32 nx = tx1 - tx0 + 1u;
  const VaFrame &fr = frames[draws[r[VA_R_DRAW]].frame];
  const u32 tx = tx0 + k % nx, ty = ty0 + k / nx;
  keys[p] = fr.tile_base + ty * fr.tiles_x + tx;
  vals[p] = t;
}

extern "C" __global__ __launch_bounds__(256) void va_ranges(const u32 *__restrict__ keys, u32 n,
                                                            u32 *__restrict__ start,
                                                            u32 *__restrict__ end)
{
  const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
  if(i >= n) return;
  const u32 k = keys[i];
  if(i == 0u || keys[i - 1u] != k) start[k] = i;
  if(i == n - 1u || keys[i + 1u] != k) end[k] = i + 1u;
}

extern "C" __global__ __launch_bounds__(256) void va_fill_u32(u32 *__restrict__ p, u32 n, u32 v)
{
  const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
  if(i < n) p[i] = v;
}

VA_INLINE float va_unorm_in(u32 byte) { return (float)byte / 255.0f; }

// reference rasterizer.cpp:277-280,674-676 (clamp01, truncate)
VA_INLINE u32 va_unorm_out(float v)
{
  v = v > 1.0f ? 1.0f : (v < 0.0f ? 0.0f : v);
  return (u32)(v * 255.0f);
}

VA_INLINE float va_blend_factor(u32 f, float a)
{
  return f == 0u ? 0.0f : (f == 6u ? a : (f == 7u ? 1.0f - a : 1.0f));
}

VA_INLINE u32 va_half_bits(float f) { _Float16 h = (_Float16)f; return (u32)__builtin_bit_cast(u16, h); }
VA_INLINE float va_half_val(u32 b) { return (float)__builtin_bit_cast(_Float16, (u16)b); }

// Perspective-correct barycentrics at integer corner (x, y) from the record
// (rasterizer.cpp:282-310, 551-588). Returns coverage; writes the pre-perspective depth.
VA_INLINE bool va_bary(const u32 *__restrict__ r, i32 x, i32 y, float b[4], float *depth)
{
  const i32 PAx = (i32)r[VA_R_X0] - x;
  const i32 PAy = (i32)r[VA_R_Y0] - y;
  const u32 ux = r[VA_R_ACX] * (u32)PAy - r[VA_R_ACY] * (u32)PAx;
  const u32 uy = (u32)PAx * r[VA_R_ABY] - (u32)PAy * r[VA_R_ABX];
  const i32 b0 = (i32)(r[VA_R_AREA] - (ux + uy)), b1 = (i32)ux, b2 = (i32)uy;
  const bool covered = b0 >= 0 && b1 >= 0 && b2 >= 0;
  const float invarea = va_u2f(r[VA_R_INVAREA]);
  float n0 = (float)b0, n1 = (float)b1, n2 = (float)b2;
  n0 *= invarea;
  n1 *= invarea;
  n2 *= invarea;
  if(depth)
    *depth = n0 * va_u2f(r[VA_R_Z0]) + n1 * va_u2f(r[VA_R_Z0 + 1]) + n2 * va_u2f(r[VA_R_Z0 + 2]);
  n0 *= va_u2f(r[VA_R_IW0]);
  n1 *= va_u2f(r[VA_R_IW0 + 1]);
  n2 *= va_u2f(r[VA_R_IW0 + 2]);
  const float invlen = 1.0f / (n0 + n1 + n2);
  b[0] = n0 * invlen;
  b[1] = n1 * invlen;
  b[2] = n2 * invlen;
  b[3] = 0.0f;
  return covered;
}

template <int T>
VA_INLINE void va_raster_tile(const VaDraw *__restrict__ draws, const VaFrame *__restrict__ frames,
                              u32 nframes, const u32 *__restrict__ rec, const u32 *__restrict__ bins,
                              const u32 *__restrict__ bin_start, const u32 *__restrict__ bin_end,
                              u32 use_fused_clear, VaCounters *__restrict__ ctr, u32 tile)
{
  constexpr int SB = T / 8;
  const VaFrame &fr = frames[va_find_frame(frames, nframes, tile)];
  const u32 lt = tile - fr.tile_base;
  const u32 tx = lt % fr.tiles_x, ty = lt / fr.tiles_x;
  const u32 wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
  const i32 bx = (i32)(tx * T + (wave % SB) * 8), by = (i32)(ty * T + (wave / SB) * 8);
  const i32 px = bx + (i32)(lane & 7u), py = by + (i32)(lane >> 3);
  const u32 W = fr.width, H = fr.height;
  const bool inside = (u32)px < W && (u32)py < H;
  const u32 beg = bin_start[tile], en = bin_end[tile];
  const bool fused = use_fused_clear != 0u && fr.clear_flags != 0u;
  if(beg == en && !fused) return;

  const u64 pix = (u64)(u32)py * W + (u32)px;
  const bool has_depth = fr.depth != 0ull;
  const bool hdr = fr.color_bpp == 8u;
  u32 c0 = 0, c1 = 0;
  float dep = 0.0f;
  if(inside)
  {
    if(fused && (fr.clear_flags & VA_CLEAR_COLOR))
    {
      if(hdr)
      {
        c0 = va_half_bits(fr.clear_color[0]) | (va_half_bits(fr.clear_color[1]) << 16);
        c1 = va_half_bits(fr.clear_color[2]) | (va_half_bits(fr.clear_color[3]) << 16);
      }
      else
      {
        // reference rasterizer.cpp:341-345: [2]=R [1]=G [0]=B [3]=A, truncated, no clamp
        c0 = ((u32)(u8)va_f2i(fr.clear_color[2] * 255.0f)) | ((u32)(u8)va_f2i(fr.clear_color[1] * 255.0f) << 8) |
             ((u32)(u8)va_f2i(fr.clear_color[0] * 255.0f) << 16) | ((u32)(u8)va_f2i(fr.clear_color[3] * 255.0f) << 24);
      }
    }
    else if(hdr)
    {
      const u32 *cp = (const u32 *)fr.color + pix * 2u;
      c0 = cp[0];
      c1 = cp[1];
    }
    else
      c0 = ((const u32 *)fr.color)[pix];
    if(has_depth)
      dep = (fused && (fr.clear_flags & VA_CLEAR_DEPTH)) ? fr.clear_depth : ((const float *)fr.depth)[pix];
  }

  u32 n_cov = 0, n_pass = 0, n_blend = 0, n_dw = 0;
  for(u32 i = beg; i < en; ++i)
  {
    const u32 t = __builtin_amdgcn_readfirstlane(bins[i]);
    const u32 *__restrict__ r = rec + (u64)t * VA_REC_STRIDE;
    const u32 bmin = r[VA_R_BBMIN], bmax = r[VA_R_BBMAX];
    const i32 minx = (i32)(bmin & 0xffffu), miny = (i32)(bmin >> 16);
    const i32 maxx = (i32)(bmax & 0xffffu), maxy = (i32)(bmax >> 16);
    if(maxx <= bx || minx >= bx + 8 || maxy <= by || miny >= by + 8) continue;
    float b[3][4];
    float pixdepth;
    bool cov = va_bary(r, px, py, b[0], &pixdepth);
    cov = cov && inside && px >= minx && px < maxx && py >= miny && py < maxy;
    if(__ballot(cov) == 0ull) continue;
    if(!cov) continue;
    const VaDraw &d = draws[r[VA_R_DRAW]];
    n_cov++;
    bool passed = true;
    const u32 op = d.depth_op;
    if(op != 7u && has_depth)
    {
      switch(op)
      {
        case 0u: passed = false; break;
        case 1u: passed = pixdepth < dep; break;
        case 2u: passed = pixdepth == dep; break;
        case 3u: passed = pixdepth <= dep; break;
        case 4u: passed = pixdepth > dep; break;
        case 5u: passed = pixdepth != dep; break;
        case 6u: passed = pixdepth >= dep; break;
        default: break;
      }
    }
    if(!passed) continue;
#if VA_FS_DERIV
    va_bary(r, px + 1, py, b[1], nullptr);
    va_bary(r, px, py + 1, b[2], nullptr);
#endif
    float o[4];
    const bool keep = va_fs(d.slot, d, (const float *)(r + VA_REC_HDR), b, o);
    if(!keep) continue;
    n_pass++;
    n_blend += d.blend_enable ? 1u : 0u;
    if(hdr)
    {
      if(d.blend_enable)
      {
        const float sf = va_blend_factor(d.src_factor, o[3]);
        const float df_ = va_blend_factor(d.dst_factor, o[3]);
        const float e0 = va_half_val(c0 & 0xffffu), e1 = va_half_val(c0 >> 16);
        const float e2 = va_half_val(c1 & 0xffffu), e3 = va_half_val(c1 >> 16);
        o[0] = sf * o[0] + df_ * e0;
        o[1] = sf * o[1] + df_ * e1;
        o[2] = sf * o[2] + df_ * e2;
        o[3] = sf * o[3] + df_ * e3;
      }
      c0 = va_half_bits(o[0]) | (va_half_bits(o[1]) << 16);
      c1 = va_half_bits(o[2]) | (va_half_bits(o[3]) << 16);
    }
    else
    {
      if(d.blend_enable)
      {
        // rasterizer.cpp:593-672: dst = (byte[2], byte[1], byte[0]) / 255, dst alpha 1
        const float sf = va_blend_factor(d.src_factor, o[3]);
        const float df_ = va_blend_factor(d.dst_factor, o[3]);
        const float e0 = va_unorm_in((c0 >> 16) & 0xffu);
        const float e1 = va_unorm_in((c0 >> 8) & 0xffu);
        const float e2 = va_unorm_in(c0 & 0xffu);
        const float r0 = sf * o[0] + df_ * e0;
        const float r1 = sf * o[1] + df_ * e1;
        const float r2 = sf * o[2] + df_ * e2;
        o[0] = r0; o[1] = r1; o[2] = r2;
      }
      c0 = (c0 & 0xff000000u) | (va_unorm_out(o[0]) << 16) | (va_unorm_out(o[1]) << 8) | va_unorm_out(o[2]);
    }
    if(d.depth_write && has_depth)
    {
      dep = pixdepth;
      n_dw++;
    }
  }

  if(inside)
  {
    if(hdr)
    {
      u32 *cp = (u32 *)fr.color + pix * 2u;
      cp[0] = c0;
      cp[1] = c1;
    }
    else
      ((u32 *)fr.color)[pix] = c0;
    if(has_depth)
      ((float *)fr.depth)[pix] = dep;
  }
  u64 s_cov = va_wave_sum((u64)n_cov), s_pass = va_wave_sum((u64)n_pass);
  u64 s_blend = va_wave_sum((u64)n_blend), s_dw = va_wave_sum((u64)n_dw);
  if(lane == 0u && (s_cov | s_pass))
  {
    atomicAdd(&ctr->pixels_written, s_cov);
    atomicAdd(&ctr->depth_passed, s_pass);
    atomicAdd(&ctr->shaded_blend, s_blend);
    atomicAdd(&ctr->shaded_dwrite, s_dw);
  }
}

#define VA_RASTER_KERNEL(T)                                                                                 \
  extern "C" __global__ __launch_bounds__((T / 8) * (T / 8) * 64) void va_raster_##T(                      \
      const VaDraw *__restrict__ draws, const VaFrame *__restrict__ frames, u32 nframes,                  \
      const u32 *__restrict__ rec, const u32 *__restrict__ bins, const u32 *__restrict__ bin_start,        \
      const u32 *__restrict__ bin_end, u32 ntiles, u32 use_fused_clear, VaCounters *__restrict__ ctr)      \
  {                                                                                                         \
    const u32 tile = blockIdx.x;                                                                            \
    if(tile < ntiles)                                                                                       \
      va_raster_tile<T>(draws, frames, nframes, rec, bins, bin_start, bin_end, use_fused_clear, ctr, tile); \
  }

VA_RASTER_KERNEL(8)
VA_RASTER_KERNEL(16)
VA_RASTER_KERNEL(32)
