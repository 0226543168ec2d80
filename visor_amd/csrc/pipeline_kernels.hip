// pipeline_kernels.hip - the draw-execution kernels of one pipeline module.
//
// Included after the translated shaders, which define:
//   VA_NSLOTS, VA_MAXNV, VA_REC_STRIDE, VA_FS_DERIV
//   va_vs3(slot, draw, vertexIndex[3], pos[3][4], vary[3][VA_MAXNV])
//                                                   vertex program + varying gather, 3 vertices
//   va_fs(slot, draw, vary, b[3][4], out[4]) -> keep          fragment program
//
// Stages (reference rasterizer.cpp:363-696 in GPU form):
//   va_setup   one thread per primitive: index fetch, 3 VS invocations, ToWindow, area / facing
//              / cull, clamped bbox, 64..128-byte triangle record, per-primitive tile count
//   va_emit    (tile, primitive) pairs, keys = global tile id, values = primitive, expanded per
//              group of 1024 primitives with an LDS search
//   (stable radix sort by tile, runtime_kernels.hip)  -> bins keep API primitive order
//   va_ranges  per-tile [start, end) into the sorted values
//   va_raster  one workgroup per tile; each wave owns an 8x8 pixel block with colour and depth in
//              registers, walks the tile's bin in order: integer edge functions + ballot,
//              early depth, fragment program, blend, and writes the block back once.

#ifndef VA_TILE_MAX_SHIFT
#define VA_TILE_MAX_SHIFT 5
#endif

// int32 difference modulo 2^32 (window coordinates of far-off vertices reach the wrapping range;
// the reference's x64 arithmetic wraps there)
VA_INLINE u32 va_wsub(i32 a, i32 b) { return (u32)a - (u32)b; }

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

extern "C" __device__ __attribute__((const)) u32 __ockl_wfred_add_u32(u32);

// The wave's sum of a 32-bit value per lane, exact: the 16-bit halves are summed separately
// (a sum of 64 halves stays below 2^22) by the DPP wave reduction.
VA_INLINE u64 va_wave_sum32(u32 v)
{
  return (u64)__ockl_wfred_add_u32(v & 0xffffu) + ((u64)__ockl_wfred_add_u32(v >> 16) << 16);
}

// Exclusive prefix sum over the wave's 64 lanes with DPP moves only (no LDS-routed shuffles):
// an inclusive scan inside each row of 16 lanes (row_shr by 1, 2, 4, 8; a lane whose source lies
// before its row's start adds 0), then row_bcast:15 adds lane 15 of rows 0 and 2 to rows 1 and 3,
// and row_bcast:31 adds lane 31 to rows 2 and 3.
VA_INLINE u32 va_wave_exclusive_sum(u32 v)
{
  u32 x = v;
  x += (u32)__builtin_amdgcn_update_dpp(0, (int)x, 0x111, 0xf, 0xf, false);  // row_shr:1
  x += (u32)__builtin_amdgcn_update_dpp(0, (int)x, 0x112, 0xf, 0xf, false);  // row_shr:2
  x += (u32)__builtin_amdgcn_update_dpp(0, (int)x, 0x114, 0xf, 0xf, false);  // row_shr:4
  x += (u32)__builtin_amdgcn_update_dpp(0, (int)x, 0x118, 0xf, 0xf, false);  // row_shr:8
  x += (u32)__builtin_amdgcn_update_dpp(0, (int)x, 0x142, 0xa, 0xf, false);  // row_bcast:15
  x += (u32)__builtin_amdgcn_update_dpp(0, (int)x, 0x143, 0xc, 0xf, false);  // row_bcast:31
  return x - v;
}

// Adds the row's four values into a counter row (device-scope atomics; the rows are zero when
// the pass begins and the finish kernel sums and clears them).
VA_INLINE void va_row_add(u64 *row, u32 i, u64 t)
{
  if(t != 0ull)
    atomicAdd((unsigned long long *)&row[i], (unsigned long long)t);
}

// Sums four counters over the workgroup and adds them into a counter row.
VA_INLINE void va_block_partials(u64 *row, u32 na, u32 nb, u32 nc, u32 ne)
{
  __shared__ u64 s_part[16][4];
  const u64 a = va_wave_sum32(na), b = va_wave_sum32(nb), c = va_wave_sum32(nc), e = va_wave_sum32(ne);
  const u32 wave = threadIdx.x >> 6, nw = (blockDim.x + 63u) >> 6;
  if((threadIdx.x & 63u) == 0u)
  {
    s_part[wave][0] = a;
    s_part[wave][1] = b;
    s_part[wave][2] = c;
    s_part[wave][3] = e;
  }
  __syncthreads();
  if(threadIdx.x < 4u)
  {
    u64 t = 0;
    for(u32 w = 0; w < nw; w++)
      t += s_part[w][threadIdx.x];
    va_row_add(row, threadIdx.x, t);
  }
}

// va_setup's counters over the workgroup, added into a counter row (tris_in, tris_out, blocks,
// pixels_tested): the two flags are counted by ballot, blocks and pixels_tested summed as 32-bit
// values across the wave (pixels_tested in 16-bit halves), then across the waves in LDS. Exact while a
// primitive's blocks and tested pixels stay below 2^26 and 2^42, i.e. for targets of at most
// 65536 pixels per side (the device allows 16384).
VA_INLINE void va_setup_partials(u64 *row, bool fin, bool fout, u32 blocks, u64 tested)
{
  __shared__ u64 s_part[VA_SETUP_PRIMS / 64u][4];
  const u64 a = (u64)__popcll(__ballot(fin));
  const u64 b = (u64)__popcll(__ballot(fout));
  const u64 c = (u64)__ockl_wfred_add_u32(blocks);
  const u64 e = (u64)__ockl_wfred_add_u32((u32)tested & 0xffffu) + ((u64)__ockl_wfred_add_u32((u32)(tested >> 16)) << 16);
  const u32 wave = threadIdx.x >> 6;
  if((threadIdx.x & 63u) == 0u)
  {
    s_part[wave][0] = a;
    s_part[wave][1] = b;
    s_part[wave][2] = c;
    s_part[wave][3] = e;
  }
  __syncthreads();
  if(threadIdx.x < 4u)
  {
    u64 t = 0;
#pragma unroll
    for(u32 w = 0; w < VA_SETUP_PRIMS / 64u; w++)
      t += s_part[w][threadIdx.x];
    va_row_add(row, threadIdx.x, t);
  }
}

// Fixed-function state of a record's draw in 9 bits: depth op (0-2), depth write (3), blend
// enable (4), source and destination colour factor codes (5-6, 7-8: 0 ZERO, 1 ONE and every
// other factor, 2 SRC_ALPHA, 3 ONE_MINUS_SRC_ALPHA; reference rasterizer.cpp:600-640).
#define VA_F_DWRITE 8u
#define VA_F_BLEND 16u
// record word VA_R_DRAW: draw index in the low VA_DRAW_BITS bits, the 9 flag bits above
#define VA_DRAW_BITS 23u
#define VA_DRAW_MASK ((1u << VA_DRAW_BITS) - 1u)
VA_INLINE u32 va_factor_code(u32 f) { return f == 0u ? 0u : (f == 6u ? 2u : (f == 7u ? 3u : 1u)); }
VA_INLINE u16 va_rec_flags(const VaDraw &d)
{
  return (u16)((d.depth_op & 7u) | (d.depth_write ? VA_F_DWRITE : 0u) | (d.blend_enable ? VA_F_BLEND : 0u) |
               (va_factor_code(d.src_factor) << 5) | (va_factor_code(d.dst_factor) << 7));
}
// reference rasterizer.cpp:100-119 (GetIndex, vertexOffset not applied), for the three vertices
// of a primitive at once: an index past the end of the bound index buffer reads as 0 instead of
// faulting. Such an index is loaded from va_zero_words, so every load is unconditional and needs
// no select afterwards, and the three are issued together.
VA_INLINE void va_index3(const VaDraw &d, const u32 p[3], u32 v[3])
{
  if(!(d.flags & VA_DRAW_INDEXED))
  {
#pragma unroll
    for(int k = 0; k < 3; k++)
      v[k] = p[k];
    return;
  }
  const VA_GLOBAL u8 *zw = (const VA_GLOBAL u8 *)va_zero_words;
  const VA_GLOBAL u8 *ib = (const VA_GLOBAL u8 *)d.ib;
  const u32 sh = (d.flags & VA_DRAW_INDEX32) ? 2u : 1u;
  const VA_GLOBAL u8 *a[3];
#pragma unroll
  for(int k = 0; k < 3; k++)
    a[k] = ((u64)p[k] << sh) + (1ull << sh) <= d.ib_bytes ? ib + ((u64)p[k] << sh) : zw;
  if(sh == 2u)
  {
#pragma unroll
    for(int k = 0; k < 3; k++)
      v[k] = *(const VA_GLOBAL u32 *)a[k];
  }
  else
  {
#pragma unroll
    for(int k = 0; k < 3; k++)
      v[k] = *(const VA_GLOBAL u16 *)a[k];
  }
}

// True when the raster's 32-bit edge arithmetic cannot wrap for this triangle: window
// coordinates within +-8191 and a target of at most 8192 pixels per side bound every product and
// sum of the edge functions below 2^31, so exact integer tests equal the raster's.
VA_INLINE bool va_small_coords(const i32 X[3], const i32 Y[3], const VaFrame &fr)
{
  bool ok = fr.width <= 8192u && fr.height <= 8192u;
#pragma unroll
  for(int k = 0; k < 3; k++)
    ok = ok && X[k] >= -8191 && X[k] <= 8191 && Y[k] >= -8191 && Y[k] <= 8191;
  return ok;
}

// Tiles of the bbox's tile rectangle that may hold a covered pixel: a tile is dropped when one of
// the three edge functions of the coverage test (va_covered) is negative at every pixel of the
// tile's part of the bbox, i.e. when its maximum over that pixel rectangle (at a corner chosen by
// the signs of its x and y coefficients) is negative. Exact for triangles passing
// va_small_coords, so a dropped tile never holds a covered pixel.
VA_INLINE i32 va_edge_max(i32 a, i32 b, i32 c, i32 xa, i32 xb, i32 ya, i32 yb)
{
  return c + (a >= 0 ? a * xb : a * xa) + (b >= 0 ? b * yb : b * ya);
}

// 32-bit arithmetic: within va_small_coords the coefficients are below 2^15 in magnitude and
// every product, constant term and partial sum stays below 8.1e8 (< 2^31).
VA_INLINE u64 va_tile_mask(const u32 *r, i32 minx, i32 miny, i32 maxx, i32 maxy, u32 tx0, u32 ty0, u32 nx, u32 ny,
                           u32 tile_shift)
{
  const i32 X0 = (i32)r[VA_R_X0], Y0 = (i32)r[VA_R_Y0];
  const i32 ABX = (i32)r[VA_R_ABX], ABY = (i32)r[VA_R_ABY], ACX = (i32)r[VA_R_ACX], ACY = (i32)r[VA_R_ACY];
  const i32 AR = (i32)r[VA_R_AREA];
  // ux = ACX (Y0 - y) - ACY (X0 - x), uy = ABY (X0 - x) - ABX (Y0 - y), b0 = AR - ux - uy
  const i32 au = ACY, bu = -ACX, cu = ACX * Y0 - ACY * X0;
  const i32 av = -ABY, bv = ABX, cv = ABY * X0 - ABX * Y0;
  const i32 a0 = -(au + av), b0 = -(bu + bv), c0 = AR - cu - cv;
  const i32 T = 1 << tile_shift;
  u64 mask = 0ull;
  for(u32 j = 0; j < ny; j++)
  {
    const i32 ya = max(miny, (i32)(ty0 + j) * T), yb = min(maxy, (i32)(ty0 + j + 1u) * T) - 1;
    for(u32 i = 0; i < nx; i++)
    {
      const i32 xa = max(minx, (i32)(tx0 + i) * T), xb = min(maxx, (i32)(tx0 + i + 1u) * T) - 1;
      const bool keep = va_edge_max(au, bu, cu, xa, xb, ya, yb) >= 0 && va_edge_max(av, bv, cv, xa, xb, ya, yb) >= 0 &&
                        va_edge_max(a0, b0, c0, xa, xb, ya, yb) >= 0;
      if(keep)
        mask |= 1ull << (j * nx + i);
    }
  }
  return mask;
}

// One workgroup covers VA_SETUP_PRIMS primitives of a single draw (block_draw maps workgroup ->
// draw), so the draw's state (push constants, vertex layout, descriptors) is read with scalar
// loads.
// Records are assembled in registers and leave through LDS: each wave stages the first half of
// its records (then the second) and stores them as whole record halves from consecutive lanes,
// 16 segments of contiguous bytes per store instead of 64 scattered 16-byte pieces (setup 0.082
// vs 0.095 ms on 1.85M triangles; staging whole records at once needs twice the LDS, which costs
// resident workgroups: 0.086 ms).
// emit_info[g] = (first tile key, tiles per bbox row | frame tiles per row << 16, tile mask) for
// va_emit: bit i of the mask = tile i of the bbox's tile rectangle (row-major) may hold a covered
// pixel; all ones = every tile of the rectangle.
// Workgroup 0 also zeroes the count-scan's last element and stamps the start of the pass.
extern "C" __global__ __launch_bounds__(VA_SETUP_PRIMS) void va_setup(
    const VaDraw *__restrict__ draws, const u32 *__restrict__ block_draw, const VaFrame *__restrict__ frames,
    u32 tile_shift, u32 *__restrict__ rec, u32 *__restrict__ tile_count, u64 *__restrict__ part,
    uint4 *__restrict__ emit_info, u32 *__restrict__ block_pairs, u64 *__restrict__ stamps)
{
  constexpr u32 Q = VA_REC_STRIDE / 4u;
  if(blockIdx.x == 0u && threadIdx.x == 0u)
    stamps[0] = (u64)wall_clock64();
  const u32 di = __builtin_amdgcn_readfirstlane(block_draw[blockIdx.x]);
  const VaDraw &d = draws[di];
  const u32 prim = (blockIdx.x - d.block_base) * VA_SETUP_PRIMS + threadIdx.x;
  const u32 g = d.tri_base + prim;
  constexpr u32 H0 = (Q + 1u) / 2u, H1 = Q - H0;
  constexpr u32 SQ = H0 | 1u;  // odd record stride in LDS chunks: no bank conflicts
  __shared__ uint4 s_stage[VA_SETUP_PRIMS * SQ];
  bool stored = false;
  uint4 hi[H1 > 0 ? H1 : 1];
  u32 n_in = 0, n_out = 0, n_blocks = 0;
  u64 n_tested = 0;
  u32 cnt = 0;  // tiles of the primitive's pair set
  if(prim < d.ntris)
  {
    const VaFrame &fr = frames[d.frame];
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
    const u32 pi[3] = {d.first + p0, d.first + p1, d.first + p2};
    u32 vi[3];
    va_index3(d, pi, vi);
    va_vs3(d.slot, d, vi, pos, vary);

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
    // double_triarea (rasterizer.cpp:272-275), wrapping int32
    const i32 area2 = (i32)(va_wsub(X[1], X[0]) * va_wsub(Y[2], Y[0]) - va_wsub(Y[1], Y[0]) * va_wsub(X[2], X[0]));
    if(area2 != 0)
    {
      // facing and cull (rasterizer.cpp:401-424)
      i32 flipped = area2, barymul = 1;
      bool culled = false;
      if(d.front_face == 1u) { barymul = -1; flipped = (i32)(0u - (u32)flipped); }
      if(flipped > 0 && (d.cull_mode & 1u)) culled = true;
      if(!culled && flipped < 0)
      {
        if(d.cull_mode & 2u) culled = true;
        else { barymul = -barymul; flipped = (i32)(0u - (u32)flipped); }
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
        n_blocks = (u32)(xb > 0 ? xb : 0) * (u32)(yb > 0 ? yb : 0);
        if(maxx > minx && maxy > miny)
        {
          n_tested = (u64)(maxx - minx) * (u64)(maxy - miny);
          u32 r[VA_REC_STRIDE];
#pragma unroll
          for(u32 w = 0; w < VA_REC_STRIDE; w++)
            r[w] = 0u;
          const u32 bm = (u32)barymul;
          r[VA_R_X0] = (u32)X[0];
          r[VA_R_Y0] = (u32)Y[0];
          // barymul folded into the edge coefficients: exact in wrapping arithmetic
          r[VA_R_ABX] = va_wsub(X[1], X[0]) * bm;
          r[VA_R_ABY] = va_wsub(Y[1], Y[0]) * bm;
          r[VA_R_ACX] = va_wsub(X[2], X[0]) * bm;
          r[VA_R_ACY] = va_wsub(Y[2], Y[0]) * bm;
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
          r[VA_R_DRAW] = di | ((u32)va_rec_flags(d) << VA_DRAW_BITS);
#pragma unroll
          for(int j = 0; j < VA_MAXNV; j++)
          {
            r[VA_REC_HDR + 3 * j + 0] = va_f2u(vary[0][j]);
            r[VA_REC_HDR + 3 * j + 1] = va_f2u(vary[1][j]);
            r[VA_REC_HDR + 3 * j + 2] = va_f2u(vary[2][j]);
          }
#pragma unroll
          for(u32 q = 0; q < H0; q++)
            s_stage[threadIdx.x * SQ + q] = make_uint4(r[4 * q], r[4 * q + 1], r[4 * q + 2], r[4 * q + 3]);
#pragma unroll
          for(u32 q = H0; q < Q; q++)
            hi[q - H0] = make_uint4(r[4 * q], r[4 * q + 1], r[4 * q + 2], r[4 * q + 3]);
          stored = true;
          const u32 tx0 = (u32)minx >> tile_shift, tx1 = (u32)(maxx - 1) >> tile_shift;
          const u32 ty0 = (u32)miny >> tile_shift, ty1 = (u32)(maxy - 1) >> tile_shift;
          const u32 nx = tx1 - tx0 + 1u, ny = ty1 - ty0 + 1u;
          cnt = nx * ny;
          u64 mask = ~0ull;
          if(cnt <= 64u && va_small_coords(X, Y, fr))
          {
            mask = va_tile_mask(r, minx, miny, maxx, maxy, tx0, ty0, nx, ny, tile_shift);
            cnt = (u32)__popcll(mask);
          }
          emit_info[g] = make_uint4(fr.tile_base + ty0 * fr.tiles_x + tx0, nx | (fr.tiles_x << 16), (u32)mask,
                                    (u32)(mask >> 32));
        }
      }
    }
    tile_count[g] = cnt;
  }
  {
    // the wave's records leave LDS in two rounds (first chunks, then the rest) as stores of whole
    // record halves (lane l of store i writes chunk i * 64 + l of the round's chunks), records of
    // lanes that wrote none skipped
    const u64 sb = __ballot(stored);
    const u32 lane = threadIdx.x & 63u, wbase = threadIdx.x & ~63u;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    uint4 *dst = (uint4 *)(rec + (u64)(d.tri_base + (blockIdx.x - d.block_base) * VA_SETUP_PRIMS + wbase) * VA_REC_STRIDE);
#pragma unroll
    for(u32 i = 0; i < H0; i++)
    {
      const u32 k = i * 64u + lane;
      const u32 rr = k / H0, c = k - rr * H0;
      if((sb >> rr) & 1ull)
        dst[rr * Q + c] = s_stage[(wbase + rr) * SQ + c];
    }
    if(H1 > 0u)
    {
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      if(stored)
      {
#pragma unroll
        for(u32 q = 0; q < H1; q++)
          s_stage[threadIdx.x * SQ + q] = hi[q];
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
      for(u32 i = 0; i < H1; i++)
      {
        const u32 k = i * 64u + lane;
        const u32 rr = k / H1, c = k - rr * H1;
        if((sb >> rr) & 1ull)
          dst[rr * Q + H0 + c] = s_stage[(wbase + rr) * SQ + c];
      }
    }
  }
  // the block's pair count, from which va_emit finds where its pairs start
  __shared__ u32 s_pairs[VA_SETUP_PRIMS / 64u];
  const u32 wave_pairs = __ockl_wfred_add_u32(cnt);
  if((threadIdx.x & 63u) == 0u)
    s_pairs[threadIdx.x >> 6] = wave_pairs;
  // counters: one row of partial sums per workgroup (no contended atomics), columns in
  // VaCounters order from tris_in; its barrier also completes s_pairs
  va_setup_partials(part + (u64)(blockIdx.x % VA_COUNTER_ROWS) * 4u, n_in != 0u, n_out != 0u, n_blocks, n_tested);
  if(threadIdx.x == 0u)
  {
    u32 t = 0;
#pragma unroll
    for(u32 k = 0; k < VA_SETUP_PRIMS / 64u; k++)
      t += s_pairs[k];
    block_pairs[blockIdx.x] = t;
  }
}

// Position of the k-th set bit of m (k < popcount(m)).
VA_INLINE u32 va_nth_bit(u64 m, u32 k)
{
  u32 pos = 0;
#pragma unroll
  for(u32 w = 32u; w > 0u; w >>= 1)
  {
    const u32 c = (u32)__popcll((m >> pos) & ((1ull << w) - 1ull));
    if(k >= c)
    {
      k -= c;
      pos += w;
    }
  }
  return pos;
}

// Pair generation, load-balanced: workgroup w expands the primitives of setup blocks [4 w,
// 4 w + 4), a contiguous range of at most 1024 primitives, one thread per primitive (the work per
// primitive is small, and 1024-wide groups need a quarter of the workgroups of 256-wide ones: 31
// vs 46 us on 1.85M triangles). The group's first pair slot is the sum of the pair counts of the
// blocks before it (setup's block_pairs; workgroup 0 also sums them all into *pair_total); the
// primitives' offsets inside the group are an exclusive scan in LDS, next to their emit
// descriptors, and each thread takes pairs of the group by stride, finding its primitive by a
// binary search over the G + 1 offsets. A group whose primitives have many pairs each (few large
// triangles) is expanded by `split` workgroups: each repeats the group's offsets and takes every
// split-th G-pair step of its pairs, so that a pass of a few thousand primitives still spreads its
// pairs over the GPU (workgroup b: group b / split, part b % split). Pairs at or beyond cap are not
// written (the host sees the total and grows the buffers). Workgroup 0 stamps the start of binning.
// *poison stays set from a pass whose pairs overflowed until the host has seen it: a pass queued
// behind that one before then reports all ones as its total, so its raster writes nothing
// either and the host runs both again, in order.
#define VA_EMIT_G 1024u
extern "C" __global__ __launch_bounds__(VA_EMIT_G) void va_emit(
    const u32 *__restrict__ block_first, const u32 *__restrict__ block_pairs, u32 nblocks,
    const uint4 *__restrict__ emit_info, const u32 *__restrict__ tile_count, u32 cap, u32 *__restrict__ keys,
    u32 *__restrict__ vals, u32 *__restrict__ pair_total, u64 *__restrict__ stamps, u32 *__restrict__ poison,
    u32 split)
{
  static_assert(VA_EMIT_G == 4u * VA_SETUP_PRIMS, "four setup blocks per group");
  __shared__ u32 s_off[VA_EMIT_G + 1u];
  __shared__ uint4 s_info[VA_EMIT_G];
  __shared__ u32 s_w[2][VA_EMIT_G / 64u];
  const u32 tid = threadIdx.x, lane = tid & 63u, w = tid >> 6;
  if(blockIdx.x == 0u && tid == 0u)
    stamps[1] = (u64)wall_clock64();
  const u32 group = blockIdx.x / split, part = blockIdx.x - group * split;
  // block_first[b] = first primitive of setup block b, block_first[nblocks] = the total
  const u32 b0 = group * 4u;
  const u32 p0 = block_first[b0], p1 = block_first[min(b0 + 4u, nblocks)];
  // pairs of the blocks before this group (workgroup 0: of all blocks, the total)
  const u32 lim = blockIdx.x == 0u ? nblocks : b0;
  // this thread's primitive, loaded before the prefix so that both latencies overlap (emit_info
  // holds 16 bytes for every primitive, so the load is in bounds also for a culled one)
  const u32 g = p0 + tid;
  const bool in = g < p1;
  const u32 c = in ? tile_count[g] : 0u;
  uint4 info = make_uint4(0u, 0u, 0u, 0u);
  if(in)
    info = emit_info[g];
  u32 pre = 0;
  for(u32 bb = 0; bb < lim; bb += 8u * VA_EMIT_G)
  {
    // eight independent loads in flight per thread
    u32 x[8];
#pragma unroll
    for(u32 k = 0; k < 8u; k++)
    {
      const u32 b = bb + k * VA_EMIT_G + tid;
      x[k] = b < lim ? block_pairs[b] : 0u;
    }
#pragma unroll
    for(u32 k = 0; k < 8u; k++)
      pre += x[k];
  }
  if(c != 0u)
    s_info[tid] = info;
  const u32 ex = va_wave_exclusive_sum(c);
  const u32 wpre = __ockl_wfred_add_u32(pre);
  if(lane == 63u)
  {
    s_w[0][w] = ex + c;
    s_w[1][w] = wpre;
  }
  __syncthreads();
  u32 off = 0, L = 0, base = 0;
#pragma unroll
  for(u32 k = 0; k < VA_EMIT_G / 64u; k++)
  {
    const u32 x = s_w[0][k];
    off += k < w ? x : 0u;
    L += x;
    base += s_w[1][k];
  }
  if(blockIdx.x == 0u)
  {
    if(tid == 0u)
    {
      const bool over = base > cap;
      if(over)
        *poison = 1u;
      *pair_total = over || *poison == 0u ? base : 0xffffffffu;
    }
    base = 0u;
  }
  s_off[tid] = off + ex;
  if(tid == 0u)
    s_off[VA_EMIT_G] = L;
  __syncthreads();
  const u32 qend = base < cap ? min(L, cap - base) : 0u;
  for(u32 q = part * VA_EMIT_G + tid; q < qend; q += VA_EMIT_G * split)
  {
    // last i with s_off[i] <= q (the steps sum to G - 1, so i + step stays below G)
    u32 i = 0;
#pragma unroll
    for(u32 step = VA_EMIT_G / 2u; step > 0u; step >>= 1)
      if(s_off[i + step] <= q)
        i += step;
    const uint4 e = s_info[i];
    const u64 mask = (u64)e.z | ((u64)e.w << 32);
    // the k-th pair of the primitive is the k-th tile of its mask (the k-th tile of the
    // rectangle when the mask is all ones)
    const u32 k = mask == ~0ull ? q - s_off[i] : va_nth_bit(mask, q - s_off[i]);
    const u32 nx = e.y & 0xffffu, tiles_x = e.y >> 16;
    const u32 row = k / nx;
    keys[base + q] = e.x + row * tiles_x + (k - row * nx);
    vals[base + q] = p0 + i;
  }
}

// Bin bounds from the sorted keys: start[t] = first index whose key is >= t, for every tile
// t in [0, ntiles] (so tile t's bin is [start[t], start[t + 1])), n = min(*n_dev, cap). One
// coalesced pass over the keys: index i is the start of every tile t with
// key[i - 1] < t <= key[i] (key[-1] = -1, key[n] = ntiles), so each entry is written once. A lane
// writes a gap of up to 16 tiles itself; a longer gap (empty tiles, e.g. all tiles outside a
// single draw's few) is written by its whole wave, 64 entries per step.
extern "C" __global__ __launch_bounds__(256) void va_ranges(const u32 *__restrict__ keys, const u32 *__restrict__ n_dev,
                                                            u32 cap, u32 ntiles, u32 *__restrict__ start)
{
  const u32 n = min(*n_dev, cap);
  const u32 lane = threadIdx.x & 63u;
  for(u32 base = blockIdx.x * 256u + (threadIdx.x & ~63u); base <= n; base += gridDim.x * 256u)
  {
    const u32 i = base + lane;
    u32 lo = 1u, hi = 0u;  // no tile
    if(i <= n)
    {
      // keys are below ntiles; the clamps keep every write inside start[0..ntiles] also when a
      // pass's raster is to write nothing (pairs overflowed, or a pass ahead of it overflowed:
      // then the count may reach past the pairs this pass wrote)
      lo = i == 0u ? 0u : min(keys[i - 1u], ntiles) + 1u;
      hi = i == n ? ntiles : min(keys[i], ntiles);
    }
    const bool big = hi >= lo && hi - lo >= 16u;
    if(!big)
      for(u32 t = lo; t <= hi; t++)
        start[t] = i;
    u64 bigs = __ballot(big);
    while(bigs)
    {
      const u32 l = (u32)__builtin_ctzll(bigs);
      bigs &= bigs - 1ull;
      const u32 bl = __builtin_amdgcn_readlane(lo, l), bh = __builtin_amdgcn_readlane(hi, l);
      const u32 bi = __builtin_amdgcn_readlane(i, l);
      for(u32 t = bl + lane; t <= bh; t += 64u)
        start[t] = bi;
    }
  }
}

extern "C" __global__ __launch_bounds__(256) void va_fill_u32(u32 *__restrict__ p, u32 n, u32 v)
{
  const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
  if(i < n) p[i] = v;
}

VA_INLINE float va_unorm_in(u32 byte) { return va_unorm8(byte); }

// reference rasterizer.cpp:277-280,674-676 (clamp01, truncate)
VA_INLINE u32 va_unorm_out(float v)
{
  v = v > 1.0f ? 1.0f : (v < 0.0f ? 0.0f : v);
  return (u32)(v * 255.0f);
}

VA_INLINE u32 va_half_bits(float f) { _Float16 h = (_Float16)f; return (u32)__builtin_bit_cast(u16, h); }
VA_INLINE float va_half_val(u32 b) { return (float)__builtin_bit_cast(_Float16, (u16)b); }

// Perspective-correct barycentrics at integer corner (x, y) from the record
// (rasterizer.cpp:282-310, 551-588). Returns coverage; writes the pre-perspective depth.
// The barycentric weights from a pixel's two integer edge functions ux, uy (b1 and b2; b0 =
// area - ux - uy), as va_bary computes them.
VA_INLINE bool va_bary_uv(const u32 *__restrict__ r, u32 ux, u32 uy, float b[4], float *depth)
{
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

VA_INLINE void va_edges(const u32 *__restrict__ r, i32 x, i32 y, u32 &ux, u32 &uy)
{
  const u32 PAx = r[VA_R_X0] - (u32)x, PAy = r[VA_R_Y0] - (u32)y;
  ux = r[VA_R_ACX] * PAy - r[VA_R_ACY] * PAx;
  uy = PAx * r[VA_R_ABY] - PAy * r[VA_R_ABX];
}

VA_INLINE bool va_bary(const u32 *__restrict__ r, i32 x, i32 y, float b[4], float *depth)
{
  u32 ux, uy;
  va_edges(r, x, y, ux, uy);
  return va_bary_uv(r, ux, uy, b, depth);
}

// The weights at the pixel and at its right and lower neighbours. The neighbours' edge functions
// are the pixel's stepped by one (x + 1 adds ACY to ux and -ABY to uy, y + 1 adds -ACX and ABX):
// the same values modulo 2^32 as evaluating them there, without the four 32-bit products.
VA_INLINE bool va_bary3(const u32 *__restrict__ r, i32 x, i32 y, float b[3][4], float *depth)
{
  u32 ux, uy;
  va_edges(r, x, y, ux, uy);
  const bool cov = va_bary_uv(r, ux, uy, b[0], depth);
  va_bary_uv(r, ux + r[VA_R_ACY], uy - r[VA_R_ABY], b[1], nullptr);
  va_bary_uv(r, ux - r[VA_R_ACX], uy + r[VA_R_ABX], b[2], nullptr);
  return cov;
}

VA_INLINE float va_factor_value(u32 code, float a)
{
  return code == 0u ? 0.0f : (code == 2u ? a : (code == 3u ? 1.0f - a : 1.0f));
}

// Blend (rasterizer.cpp:593-672) and store (:674-676) one shaded fragment into the pixel's
// registers. 8-bit targets: dst = (byte[2], byte[1], byte[0]) / 255 with dst alpha 1, factors
// from the colour factors, bytes B,G,R truncated, alpha byte untouched. RGBA16F targets (beyond
// the reference) blend all four channels and store halves.
// The blended values are computed for every fragment and selected (lanes of one step blend
// fragments of different draws, so a branch would diverge); without blending the fragment's own
// values are stored, as the reference does (not 1 * src + 0 * dst, which differs for inf / NaN).
VA_INLINE void va_blend_store_flags(u32 f, bool hdr, float o[4], u32 &c0, u32 &c1)
{
  const bool blend = (f & VA_F_BLEND) != 0u;
  const float sf = va_factor_value((f >> 5) & 3u, o[3]);
  const float df_ = va_factor_value((f >> 7) & 3u, o[3]);
  if(hdr)
  {
    const float e0 = va_half_val(c0 & 0xffffu), e1 = va_half_val(c0 >> 16);
    const float e2 = va_half_val(c1 & 0xffffu), e3 = va_half_val(c1 >> 16);
    const float b0 = sf * o[0] + df_ * e0, b1 = sf * o[1] + df_ * e1;
    const float b2 = sf * o[2] + df_ * e2, b3 = sf * o[3] + df_ * e3;
    o[0] = blend ? b0 : o[0];
    o[1] = blend ? b1 : o[1];
    o[2] = blend ? b2 : o[2];
    o[3] = blend ? b3 : o[3];
    c0 = va_half_bits(o[0]) | (va_half_bits(o[1]) << 16);
    c1 = va_half_bits(o[2]) | (va_half_bits(o[3]) << 16);
  }
  else
  {
    const float e0 = va_unorm_in((c0 >> 16) & 0xffu);
    const float e1 = va_unorm_in((c0 >> 8) & 0xffu);
    const float e2 = va_unorm_in(c0 & 0xffu);
    const float b0 = sf * o[0] + df_ * e0, b1 = sf * o[1] + df_ * e1, b2 = sf * o[2] + df_ * e2;
    o[0] = blend ? b0 : o[0];
    o[1] = blend ? b1 : o[1];
    o[2] = blend ? b2 : o[2];
    c0 = (c0 & 0xff000000u) | (va_unorm_out(o[0]) << 16) | (va_unorm_out(o[1]) << 8) | va_unorm_out(o[2]);
  }
}

VA_INLINE void va_blend_store(const VaDraw &d, bool hdr, float o[4], u32 &c0, u32 &c1)
{
  va_blend_store_flags(va_rec_flags(d), hdr, o, c0, c1);
}

// depth compare (rasterizer.cpp:562-576)
// Without branches: VkCompareOp's bits 0, 1 and 2 select less, equal and greater (NEVER 0 ...
// ALWAYS 7), and an unordered pair (a NaN) passes NOT_EQUAL and ALWAYS only. The same results as
// the reference's switch (NEVER false, LESS <, EQUAL ==, LEQUAL <=, GREATER >, NOT_EQUAL !=,
// GEQUAL >=, ALWAYS true); a switch over a per-fragment op compiles to a divergent branch tree.
VA_INLINE bool va_depth_pass(u32 op, float z, float cur)
{
  const u32 rel = (z < cur ? 1u : 0u) | (z == cur ? 2u : 0u) | (z > cur ? 4u : 0u);
  return (op & rel) != 0u || (rel == 0u && (op == 5u || op == 7u));
}

// Orders this wave's LDS writes before its later LDS reads by other lanes.
VA_INLINE void va_wave_sync()
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

#ifndef VA_CHUNK
#define VA_CHUNK 64u       // bin entries staged in LDS per round (at most 64: one bit each)
#endif
#ifndef VA_SMALL
#define VA_SMALL 64        // bbox-in-block overlaps up to this many pixels are tested per triangle
#endif
static_assert(VA_CHUNK <= 64u, "record sets are 64-bit masks");

// A set of records of one chunk (bit k = record k): 32 bits when a chunk holds at most 32.
#if VA_CHUNK <= 32
typedef u32 va_cmask;
VA_INLINE u32 va_cmask_ctz(u32 m) { return (u32)__builtin_ctz(m); }
VA_INLINE u32 va_cmask_pop(u32 m) { return (u32)__builtin_popcount(m); }
VA_INLINE u32 va_cmask_top(u32 m) { return 31u - (u32)__builtin_clz(m); }
#else
typedef u64 va_cmask;
VA_INLINE u32 va_cmask_ctz(u64 m) { return (u32)__builtin_ctzll(m); }
VA_INLINE u32 va_cmask_pop(u64 m) { return (u32)__builtin_popcountll(m); }
VA_INLINE u32 va_cmask_top(u64 m) { return 63u - (u32)__builtin_clzll(m); }
#endif

// Records in LDS at a stride of an odd number of 16-byte groups (VA_REC_STRIDE is a whole number
// of groups; one group is added when that number is even): every record starts 16-byte aligned,
// so the reads of a record's consecutive header words and varyings merge into 64- and 128-bit
// LDS reads and the staging stores are 128-bit, while the same field of 8 consecutive records
// still falls in 8 different groups of banks. (A stride of one word more than VA_REC_STRIDE
// spreads single-word reads over more banks but leaves every read 32-bit: raster 0.274 vs
// 0.270 ms on the 256-draw workload.)
#define VA_LDS_REC_WORDS (VA_REC_STRIDE + ((VA_REC_STRIDE / 4u) & 1u ? 0u : 4u))
static_assert(VA_REC_STRIDE % 4u == 0u, "records are whole 16-byte groups");

// Software-pipelined staging of a tile's bin into LDS, G 16-byte groups of each record: while
// chunk c is processed, the groups of chunk c + 1 (whose bin entries arrived an iteration
// earlier) and the bin entries of chunk c + 2 are in flight in registers. Thread slot j of a
// chunk covers group w = thread + j * NT, i.e. group w % G of the chunk's record w / G.
template <u32 G, u32 NT>
struct VaStager
{
  static constexpr u32 J = (VA_CHUNK * G + NT - 1u) / NT;
  u32 tri[J];
  uint4 grp[J];
  // bin entries of chunk c (nn records)
  VA_INLINE void load_tris(const u32 *__restrict__ bins, u32 beg, u32 nn, u32 c)
  {
    const u32 c0 = beg + c * VA_CHUNK;
#pragma unroll
    for(u32 j = 0; j < J; j++)
    {
      const u32 w = threadIdx.x + j * NT;
      tri[j] = w < nn * G ? bins[c0 + w / G] : 0u;
    }
  }
  // record groups of the chunk whose bin entries were loaded last
  VA_INLINE void load_recs(const u32 *__restrict__ rec, u32 nn)
  {
#pragma unroll
    for(u32 j = 0; j < J; j++)
    {
      const u32 w = threadIdx.x + j * NT;
      if(w < nn * G)
        grp[j] = ((const uint4 *)(rec + (u64)tri[j] * VA_REC_STRIDE))[w % G];
    }
  }
  VA_INLINE void store(u32 *s_recw, u32 n) const
  {
#pragma unroll
    for(u32 j = 0; j < J; j++)
    {
      const u32 w = threadIdx.x + j * NT;
      if(w < n * G)
      {
        const u32 k = w / G, q = w - k * G;
        u32 *dst = s_recw + k * VA_LDS_REC_WORDS + 4u * q;
        dst[0] = grp[j].x;
        dst[1] = grp[j].y;
        dst[2] = grp[j].z;
        dst[3] = grp[j].w;
      }
    }
  }
};

// Transposes the 64x64 bit matrix held one row per lane (bit j of lane i = element (i, j)):
// afterwards bit i of lane j is the former element (i, j). Six block-swap steps, each exchanging
// the off-diagonal s x s blocks of every 2s x 2s block with the partner lane (lane ^ s).
VA_INLINE u64 va_transpose64(u64 x, u32 lane)
{
  const u64 masks[6] = {0x00000000FFFFFFFFull, 0x0000FFFF0000FFFFull, 0x00FF00FF00FF00FFull,
                        0x0F0F0F0F0F0F0F0Full, 0x3333333333333333ull, 0x5555555555555555ull};
#pragma unroll
  for(int i = 0; i < 6; i++)
  {
    const u32 s = 32u >> i;
    const u64 m = masks[i];
    const u32 ylo = (u32)__shfl_xor((int)(u32)x, (int)s);
    const u32 yhi = (u32)__shfl_xor((int)(u32)(x >> 32), (int)s);
    const u64 y = ((u64)yhi << 32) | ylo;
    x = (lane & s) ? ((x & ~m) | ((y >> s) & m)) : ((x & m) | ((y << s) & ~m));
  }
  return x;
}

// The same transpose for at most 32 rows (chunks of at most 32 records), each row held in two
// halves: v of lane i < 32 holds elements 0..31 of row i, v of lane i + 32 elements 32..63 of
// row i. Each half of the wave transposes its 32x32 block in five 32-bit steps; afterwards bit i
// of lane j is the former element (i, j), for i < 32. The partner lane (lane ^ s, s <= 16, inside
// the same 32 lanes) is read with ds_swizzle in bit-mask mode (and 0x1f, xor s), which needs no
// per-lane address.
VA_INLINE u32 va_transpose32(u32 v, u32 lane)
{
#define VA_T32_STEP(S, M)                                                                       \
  {                                                                                             \
    const u32 y = (u32)__builtin_amdgcn_ds_swizzle((int)v, 0x1f | ((S) << 10));                \
    v = (lane & (S)) ? ((v & ~(M)) | ((y >> (S)) & (M))) : ((v & (M)) | ((y << (S)) & ~(M))); \
  }
  VA_T32_STEP(16u, 0x0000FFFFu)
  VA_T32_STEP(8u, 0x00FF00FFu)
  VA_T32_STEP(4u, 0x0F0F0F0Fu)
  VA_T32_STEP(2u, 0x33333333u)
  VA_T32_STEP(1u, 0x55555555u)
#undef VA_T32_STEP
  return v;
}


// Coverage only (the integer part of va_bary).
VA_INLINE bool va_covered(const u32 *__restrict__ r, i32 x, i32 y)
{
  const i32 PAx = (i32)(r[VA_R_X0] - (u32)x);
  const i32 PAy = (i32)(r[VA_R_Y0] - (u32)y);
  const u32 ux = r[VA_R_ACX] * (u32)PAy - r[VA_R_ACY] * (u32)PAx;
  const u32 uy = (u32)PAx * r[VA_R_ABY] - (u32)PAy * r[VA_R_ABX];
  return (i32)(r[VA_R_AREA] - (ux + uy)) >= 0 && (i32)ux >= 0 && (i32)uy >= 0;
}

// One workgroup per tile; wave w owns the 8x8 block w of the tile, one pixel per lane, colour
// and depth in registers from first touch to the final store. The workgroup stages VA_CHUNK bin
// entries' triangle records in LDS (bin order = API primitive order). Each wave then builds the
// coverage matrix of the chunk over its block (triangle-parallel for small overlaps,
// pixel-parallel for large ones), transposes it so each lane holds the ordered set of triangles
// covering its pixel, and runs the early depth test and depth write per pixel in primitive order.
// The fragments that pass are shaded 64 per step regardless of which triangle or pixel they
// belong to, and each pixel blends its own results in primitive order. The arithmetic and order
// of every fragment are those of the reference; only the lane that runs a fragment program
// changes.
// Coverage and interpolated depth only (the first half of va_bary, same operations).
VA_INLINE bool va_cov_depth(const u32 *__restrict__ r, i32 x, i32 y, float *depth)
{
  const i32 PAx = (i32)(r[VA_R_X0] - (u32)x);
  const i32 PAy = (i32)(r[VA_R_Y0] - (u32)y);
  const u32 ux = r[VA_R_ACX] * (u32)PAy - r[VA_R_ACY] * (u32)PAx;
  const u32 uy = (u32)PAx * r[VA_R_ABY] - (u32)PAy * r[VA_R_ABX];
  const i32 b0 = (i32)(r[VA_R_AREA] - (ux + uy)), b1 = (i32)ux, b2 = (i32)uy;
  const float invarea = va_u2f(r[VA_R_INVAREA]);
  float n0 = (float)b0, n1 = (float)b1, n2 = (float)b2;
  n0 *= invarea;
  n1 *= invarea;
  n2 *= invarea;
  *depth = n0 * va_u2f(r[VA_R_Z0]) + n1 * va_u2f(r[VA_R_Z0 + 1]) + n2 * va_u2f(r[VA_R_Z0 + 2]);
  return b0 >= 0 && b1 >= 0 && b2 >= 0;
}

template <int T>
VA_INLINE void va_raster_tile(const VaDraw *__restrict__ draws, const VaFrame *__restrict__ frames,
                              u32 nframes, const u32 *__restrict__ rec, const u32 *__restrict__ bins,
                              const u32 *__restrict__ bin_start, const u32 *__restrict__ bin_end,
                              u32 use_fused_clear, u64 *__restrict__ part, va_cmask *__restrict__ masks, u32 tile)
{
  constexpr u32 SB = T / 8, NW = SB * SB, NT = NW * 64u;
  constexpr u32 Q = VA_REC_STRIDE / 4u;
  constexpr u32 LS = VA_LDS_REC_WORDS;
  __shared__ __attribute__((aligned(16))) u32 s_recw[VA_CHUNK * LS];
  __shared__ u16 s_flags[VA_CHUNK];   // per-record fixed-function state (va_rec_flags)
  __shared__ u16 s_job[NW][64];       // jobs of one shading round: record << 6 | pixel
  __shared__ float4 s_res[NW][64];    // fragment program results of one shading round
  const u32 *s_rec = s_recw;

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
  u64 *row = part + (u64)(tile % VA_COUNTER_ROWS) * 4u;
  if(beg == en && !fused)
    return;

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
      c0 = va_gld(cp);
      c1 = va_gld(cp + 1);
    }
    else
      c0 = va_gld((const u32 *)fr.color + pix);
    if(has_depth)
      dep = (fused && (fr.clear_flags & VA_CLEAR_DEPTH)) ? fr.clear_depth : va_u2f(va_gld((const u32 *)fr.depth + pix));
  }

  u32 n_cov = 0, n_pass = 0, n_blend = 0, n_dw = 0;
  const u32 nchunks = (en - beg + VA_CHUNK - 1u) / VA_CHUNK;
  auto chunk_len = [&](u32 c) { return c < nchunks ? min(VA_CHUNK, en - (beg + c * VA_CHUNK)) : 0u; };

#if VA_FS_KILL
  // Fragment programs that may discard decide depth writes: one pass over the bin, shading in
  // place, in order.
  VaStager<Q, NT> stg;
  stg.load_tris(bins, beg, chunk_len(0), 0u);
  stg.load_recs(rec, chunk_len(0));
  stg.load_tris(bins, beg, chunk_len(1), 1u);
  for(u32 c = 0; c < nchunks; c++)
  {
    const u32 n = chunk_len(c);
    __syncthreads();  // the previous chunk is consumed by every wave
    stg.store(s_recw, n);
    __syncthreads();
    stg.load_recs(rec, chunk_len(c + 1u));  // addresses from the bin entries loaded last iteration
    stg.load_tris(bins, beg, chunk_len(c + 2u), c + 2u);
    for(u32 k = 0; k < n; ++k)
    {
      const u32 *r = s_rec + k * LS;
      const u32 bmin = r[VA_R_BBMIN], bmax = r[VA_R_BBMAX];
      const i32 minx = (i32)(bmin & 0xffffu), miny = (i32)(bmin >> 16);
      const i32 maxx = (i32)(bmax & 0xffffu), maxy = (i32)(bmax >> 16);
      if(maxx <= bx || minx >= bx + 8 || maxy <= by || miny >= by + 8) continue;
      float b[3][4];
      float pixdepth;
      u32 ux, uy;
      va_edges(r, px, py, ux, uy);
      bool cov = va_bary_uv(r, ux, uy, b[0], &pixdepth);
      cov = cov && inside && px >= minx && px < maxx && py >= miny && py < maxy;
      if(__ballot(cov) == 0ull || !cov) continue;
      const VaDraw &d = draws[r[VA_R_DRAW] & VA_DRAW_MASK];
      n_cov++;
      if(d.depth_op != 7u && has_depth && !va_depth_pass(d.depth_op, pixdepth, dep)) continue;
      // the right and lower neighbours: the edge functions stepped by one
      va_bary_uv(r, ux + r[VA_R_ACY], uy - r[VA_R_ABY], b[1], nullptr);
      va_bary_uv(r, ux - r[VA_R_ACX], uy + r[VA_R_ABX], b[2], nullptr);
      float o[4];
      if(!va_fs(d.slot, d, (const float *)(r + VA_REC_HDR), b, o)) continue;
      n_pass++;
      n_blend += d.blend_enable ? 1u : 0u;
      va_blend_store(d, hdr, o, c0, c1);
      if(d.depth_write && has_depth)
      {
        dep = pixdepth;
        n_dw++;
      }
    }
  }
#else
  // Two passes over the bin. Pass A (visibility) runs coverage, the early depth test and depth
  // writes for every fragment in primitive order, keeps each chunk's per-pixel pass mask in
  // `masks`, and finds per pixel the last passing fragment of a draw without blending. That
  // fragment's store replaces the pixel's B, G and R bytes (every half of an RGBA16F pixel)
  // whatever was there, and fragment programs here cannot discard, so the passing fragments
  // before it cannot change the final colour. Pass B (shading) runs the fragment program only
  // for that fragment and the ones after it, and blends them in order. The depth buffer and
  // every counter are final after pass A.
  va_cmask *mk = masks + (u64)(NW * 64u) * (beg / VA_CHUNK + tile) + wave * 64u + lane;
  u32 cut_c = 0xffffffffu, cut_k = 0u, first_c = 0xffffffffu;
  {
    // the leading header words of each record (all but 1/w) suffice for coverage and depth
    constexpr u32 HQ = (VA_REC_VIS + 3u) / 4u < Q ? (VA_REC_VIS + 3u) / 4u : Q;
    VaStager<HQ, NT> stg;
    stg.load_tris(bins, beg, chunk_len(0), 0u);
    stg.load_recs(rec, chunk_len(0));
    stg.load_tris(bins, beg, chunk_len(1), 1u);
    for(u32 c = 0; c < nchunks; c++)
    {
      const u32 n = chunk_len(c);
      __syncthreads();  // the previous chunk is consumed by every wave
      stg.store(s_recw, n);
      __syncthreads();
      stg.load_recs(rec, chunk_len(c + 1u));
      stg.load_tris(bins, beg, chunk_len(c + 2u), c + 2u);

      // 1. Coverage, triangle-parallel where cheap: triangle k is tested against this wave's
      //    block by its own lanes when the bbox overlap is at most VA_SMALL pixels (same integer
      //    arithmetic as the pixel-parallel test); larger overlaps are tested pixel-parallel, one
      //    ballot per triangle. The result is row k of the coverage matrix (bit p = pixel p
      //    covered by triangle k).
#if VA_CHUNK <= 32
      //    Lanes k and k + 32 share triangle k: lane k tests the block's even rows and lane k + 32
      //    its odd rows, each stepping two rows at a time; then each takes from the other the rows
      //    of its half of the block, so that lane k holds pixels 0..31 (rows 0..3) of row k and
      //    lane k + 32 pixels 32..63 (rows 4..7), the layout va_transpose32 takes.
      const u32 tk = lane & 31u, half = lane >> 5;
      u64 own = 0ull;
      bool big = false;
      u32 fl = 0u;
      if(tk < n)
      {
        const u32 *r = s_rec + tk * LS;
        fl = r[VA_R_DRAW] >> VA_DRAW_BITS;
        const u32 bmin = r[VA_R_BBMIN], bmax = r[VA_R_BBMAX];
        const i32 x0 = max((i32)(bmin & 0xffffu), bx), y0 = max((i32)(bmin >> 16), by);
        const i32 x1 = min(min((i32)(bmax & 0xffffu), bx + 8), (i32)W);
        const i32 y1 = min(min((i32)(bmax >> 16), by + 8), (i32)H);
        if(x0 < x1 && y0 < y1)
        {
          if((x1 - x0) * (y1 - y0) > VA_SMALL)
            big = half == 0u;
          else
          {
            // the edge functions at (x0, ys), ys the lane's first row, then stepped: x + 1 adds
            // ACY to ux and -ABY to uy, y + 2 adds -2 ACX and 2 ABX (the same values modulo 2^32
            // as evaluating each pixel)
            const u32 ABX = r[VA_R_ABX], ABY = r[VA_R_ABY], ACX = r[VA_R_ACX], ACY = r[VA_R_ACY], AR = r[VA_R_AREA];
            const i32 ys = y0 + (((y0 - by) ^ (i32)half) & 1);
            const u32 PAx = r[VA_R_X0] - (u32)x0, PAy = r[VA_R_Y0] - (u32)ys;
            u32 ux_row = ACX * PAy - ACY * PAx;
            u32 uy_row = PAx * ABY - PAy * ABX;
            for(i32 y = ys; y < y1; y += 2)
            {
              u32 ux = ux_row, uy = uy_row;
              for(i32 x = x0; x < x1; x++)
              {
                if((i32)(AR - (ux + uy)) >= 0 && (i32)ux >= 0 && (i32)uy >= 0)
                  own |= 1ull << (u32)((y - by) * 8 + (x - bx));
                ux += ACY;
                uy -= ABY;
              }
              ux_row -= ACX << 1;
              uy_row += ABX << 1;
            }
          }
        }
      }
      const u32 recv = (u32)__shfl_xor((int)(half ? (u32)own : (u32)(own >> 32)), 32);
      u32 rows = (half ? (u32)(own >> 32) : (u32)own) | recv;
      // records of draws that blend (bit k = record k)
      const va_cmask blends = (va_cmask)__ballot((fl & VA_F_BLEND) != 0u);
      for(u64 bigs = __ballot(big); bigs; bigs &= bigs - 1ull)
      {
        const u32 k = (u32)__builtin_ctzll(bigs);
        const u32 *r = s_rec + k * LS;
        const u32 bmin = r[VA_R_BBMIN], bmax = r[VA_R_BBMAX];
        const bool cov = inside && va_covered(r, px, py) && px >= (i32)(bmin & 0xffffu) && px < (i32)(bmax & 0xffffu) &&
                         py >= (i32)(bmin >> 16) && py < (i32)(bmax >> 16);
        const u64 m = __ballot(cov);
        if(tk == k)
          rows = half ? (u32)(m >> 32) : (u32)m;
      }
      // 2. Transpose: lane p now holds the set of records (bit k) covering pixel p, so the
      //    in-order loop below runs once per fragment of this pixel, not once per triangle.
      const va_cmask cov = va_transpose32(rows, lane);
#else
      //    Lane k tests triangle k and ends up holding row k.
      u64 row_mask = 0ull;
      bool big = false;
      u32 fl = 0u;
      if(lane < n)
      {
        const u32 *r = s_rec + lane * LS;
        fl = r[VA_R_DRAW] >> VA_DRAW_BITS;
        const u32 bmin = r[VA_R_BBMIN], bmax = r[VA_R_BBMAX];
        const i32 x0 = max((i32)(bmin & 0xffffu), bx), y0 = max((i32)(bmin >> 16), by);
        const i32 x1 = min(min((i32)(bmax & 0xffffu), bx + 8), (i32)W);
        const i32 y1 = min(min((i32)(bmax >> 16), by + 8), (i32)H);
        if(x0 < x1 && y0 < y1)
        {
          if((x1 - x0) * (y1 - y0) > VA_SMALL)
            big = true;
          else
          {
            // the edge functions at (x0, y0), then stepped: x + 1 adds ACY to ux and -ABY to uy,
            // y + 1 adds -ACX and ABX (the same values modulo 2^32 as evaluating each pixel)
            const u32 ABX = r[VA_R_ABX], ABY = r[VA_R_ABY], ACX = r[VA_R_ACX], ACY = r[VA_R_ACY], AR = r[VA_R_AREA];
            const u32 PAx = r[VA_R_X0] - (u32)x0, PAy = r[VA_R_Y0] - (u32)y0;
            u32 ux_row = ACX * PAy - ACY * PAx;
            u32 uy_row = PAx * ABY - PAy * ABX;
            for(i32 y = y0; y < y1; y++)
            {
              u32 ux = ux_row, uy = uy_row;
              for(i32 x = x0; x < x1; x++)
              {
                if((i32)(AR - (ux + uy)) >= 0 && (i32)ux >= 0 && (i32)uy >= 0)
                  row_mask |= 1ull << (u32)((y - by) * 8 + (x - bx));
                ux += ACY;
                uy -= ABY;
              }
              ux_row -= ACX;
              uy_row += ABX;
            }
          }
        }
      }
      // records of draws that blend (bit k = record k)
      const va_cmask blends = (va_cmask)__ballot((fl & VA_F_BLEND) != 0u);
      for(u64 bigs = __ballot(big); bigs; bigs &= bigs - 1ull)
      {
        const u32 k = (u32)__builtin_ctzll(bigs);
        const u32 *r = s_rec + k * LS;
        const u32 bmin = r[VA_R_BBMIN], bmax = r[VA_R_BBMAX];
        const bool cov = inside && va_covered(r, px, py) && px >= (i32)(bmin & 0xffffu) && px < (i32)(bmax & 0xffffu) &&
                         py >= (i32)(bmin >> 16) && py < (i32)(bmax >> 16);
        const u64 m = __ballot(cov);
        if(lane == k)
          row_mask = m;
      }
      // 2. Transpose: lane p now holds the set of records (bit k) covering pixel p, so the
      //    in-order loop below runs once per fragment of this pixel, not once per triangle.
      const va_cmask cov = va_transpose64(row_mask, lane);
#endif

      // 3. Early depth test and depth write per pixel in primitive order (ascending k).
      va_cmask pass = 0;
      for(va_cmask todo = cov; todo; todo &= todo - 1u)
      {
        const u32 k = va_cmask_ctz(todo);
        const u32 f = s_rec[k * LS + VA_R_DRAW] >> VA_DRAW_BITS;
        const u32 op = f & 7u;
        n_cov++;
        // straight-line code: the depth of every fragment is computed (ALWAYS without depth
        // writes is rare), and the test and its consequences are selects, not branches
        float z = 0.0f;
        if(has_depth)
          va_cov_depth(s_rec + k * LS, px, py, &z);
        const bool passed = !has_depth || va_depth_pass(op, z, dep);
        const bool dw = passed && has_depth && (f & VA_F_DWRITE) != 0u;
        pass |= (va_cmask)(passed ? 1u : 0u) << k;
        n_pass += passed ? 1u : 0u;
        n_blend += passed ? (f >> 4) & 1u : 0u;
        n_dw += dw ? 1u : 0u;
        dep = dw ? z : dep;
      }
      mk[(u64)c * (NW * 64u)] = pass;
      const va_cmask opaque = pass & ~blends;
      if(opaque)
      {
        cut_c = c;
        cut_k = va_cmask_top(opaque);
      }
      if(pass && first_c == 0xffffffffu)
        first_c = c;
    }
  }

  // Pass B. A pixel's shading starts at its cut (chunk cut_c, record cut_k) or, with no
  // fragment of a non-blending draw, at its first passing fragment; the tile's walk starts at
  // the earliest of these.
  u32 start_c = cut_c != 0xffffffffu ? cut_c : first_c;
  start_c = min(start_c, (u32)__shfl_xor((int)start_c, 1));
  start_c = min(start_c, (u32)__shfl_xor((int)start_c, 2));
  start_c = min(start_c, (u32)__shfl_xor((int)start_c, 4));
  start_c = min(start_c, (u32)__shfl_xor((int)start_c, 8));
  start_c = min(start_c, (u32)__shfl_xor((int)start_c, 16));
  start_c = min(start_c, (u32)__shfl_xor((int)start_c, 32));
  if(NW > 1u)
  {
    __shared__ u32 s_start;
    if(threadIdx.x == 0u)
      s_start = 0xffffffffu;
    __syncthreads();
    if(lane == 0u)
      atomicMin(&s_start, start_c);
    __syncthreads();
    start_c = s_start;
  }
  start_c = __builtin_amdgcn_readfirstlane(start_c);
  if(start_c < nchunks)
  {
    VaStager<Q, NT> stg;
    stg.load_tris(bins, beg, chunk_len(start_c), start_c);
    stg.load_recs(rec, chunk_len(start_c));
    stg.load_tris(bins, beg, chunk_len(start_c + 1u), start_c + 1u);
    va_cmask m_next = mk[(u64)start_c * (NW * 64u)];
    for(u32 c = start_c; c < nchunks; c++)
    {
      const u32 n = chunk_len(c);
      const va_cmask m = m_next;
      __syncthreads();  // the previous chunk is consumed by every wave
      stg.store(s_recw, n);
      __syncthreads();
      for(u32 k = threadIdx.x; k < n; k += NT)
        s_flags[k] = (u16)(s_rec[k * LS + VA_R_DRAW] >> VA_DRAW_BITS);
      __syncthreads();
      stg.load_recs(rec, chunk_len(c + 1u));
      stg.load_tris(bins, beg, chunk_len(c + 2u), c + 2u);
      if(c + 1u < nchunks)
        m_next = mk[(u64)(c + 1u) * (NW * 64u)];
      // this pixel's fragments to shade in this chunk
      const va_cmask alive = cut_c == 0xffffffffu || c > cut_c ? m
                             : c == cut_c                     ? m & (va_cmask)(~(va_cmask)0 << cut_k)
                                                              : (va_cmask)0;

      // 4. Shade every alive fragment in rounds of 64, jobs in pixel-major order (pixel p owns
      //    jobs [off_p, off_p + count_p), in ascending k): each pixel writes its jobs of the round
      //    to the job list, the wave shades the list 64 wide, then each pixel blends its own
      //    results of the round in order.
      const u32 cnt = va_cmask_pop(alive);
      const u32 off = va_wave_exclusive_sum(cnt);
      const u32 total = (u32)__builtin_amdgcn_readlane((int)(off + cnt), 63);
      va_cmask wrem = alive, brem = alive;  // write and blend cursors over this pixel's jobs
      u32 wnext = off, bnext = off;
      const u32 mine_end = off + cnt;
      for(u32 base = 0; base < total; base += 64u)
      {
        const u32 stop = min(mine_end, base + 64u);
        // each pixel lists its jobs of this round, in ascending record order
        for(; wnext < stop; ++wnext)
        {
          const u32 k = va_cmask_ctz(wrem);
          wrem &= wrem - 1u;
          s_job[wave][wnext - base] = (u16)((k << 6) | lane);
        }
        va_wave_sync();
        if(base + lane < total)
        {
          const u32 job = s_job[wave][lane];
          const u32 k = job >> 6, p = job & 63u;
          const u32 *r = s_rec + k * LS;
          const i32 x = bx + (i32)(p & 7u), y = by + (i32)(p >> 3);
          float b[3][4];
#if VA_FS_DERIV
          va_bary3(r, x, y, b, nullptr);
#else
          va_bary(r, x, y, b[0], nullptr);
#endif
          const VaDraw &d = draws[r[VA_R_DRAW] & VA_DRAW_MASK];
          float o[4];
          va_fs(d.slot, d, (const float *)(r + VA_REC_HDR), b, o);
          s_res[wave][lane] = make_float4(o[0], o[1], o[2], o[3]);
        }
        va_wave_sync();
        for(; bnext < stop; ++bnext)
        {
          const u32 k = va_cmask_ctz(brem);
          brem &= brem - 1u;
          const float4 v = s_res[wave][bnext - base];
          float o[4] = {v.x, v.y, v.z, v.w};
          va_blend_store_flags(s_flags[k], hdr, o, c0, c1);
        }
        // the next round rewrites s_job and s_res
        va_wave_sync();
      }
    }
  }
#endif

  if(inside)
  {
    if(hdr)
    {
      u32 *cp = (u32 *)fr.color + pix * 2u;
      va_gst(cp, c0);
      va_gst(cp + 1, c1);
    }
    else
      va_gst((u32 *)fr.color + pix, c0);
    if(has_depth)
      va_gst((u32 *)fr.depth + pix, va_f2u(dep));
  }
  // columns in VaCounters order from pixels_written
  va_block_partials(row, n_cov, n_pass, n_blend, n_dw);
}

// Workgroups are dispatched round-robin over the 8 XCDs (workgroup b to XCD b % 8), each with its
// own L2. Workgroup b takes tile va_xcd_tile(b, n), a bijection of [0, n) that deals runs of 16
// consecutive tiles to the XCDs in turn (the tail past the last whole round keeps b), so
// neighbouring tiles of a run, which share triangles, read those records through the same L2,
// while every XCD still gets work from everywhere on the target (one contiguous band per XCD
// unbalances the load).
VA_INLINE u32 va_xcd_tile(u32 b, u32 n)
{
  const u32 full = n & ~127u;
  if(b >= full)
    return b;
  const u32 x = b & 7u, s = b >> 3;  // s-th workgroup of XCD x
  return (((s >> 4) << 3) + x) * 16u + (s & 15u);
}

// pair_total / pair_cap: when the pairs overflowed the bins the launch writes nothing, so the
// host can grow the buffers and run the pass again on an untouched target. Workgroup 0 stamps
// the start of the raster stage.
#define VA_RASTER_KERNEL(T)                                                                                 \
  extern "C" __global__ __launch_bounds__((T / 8) * (T / 8) * 64) void va_raster_##T(                      \
      const VaDraw *__restrict__ draws, const VaFrame *__restrict__ frames, u32 nframes,                  \
      const u32 *__restrict__ rec, const u32 *__restrict__ bins, const u32 *__restrict__ bin_start,        \
      const u32 *__restrict__ bin_end, u32 ntiles, u32 use_fused_clear, u64 *__restrict__ part,            \
      const u32 *__restrict__ pair_total, u32 pair_cap, va_cmask *__restrict__ masks, u64 *__restrict__ stamp) \
  {                                                                                                         \
    const u32 tile = va_xcd_tile(blockIdx.x, ntiles);                                                       \
    if(blockIdx.x == 0u && threadIdx.x == 0u)                                                               \
      *stamp = (u64)wall_clock64();                                                                         \
    if(tile < ntiles && *pair_total <= pair_cap)                                                            \
      va_raster_tile<T>(draws, frames, nframes, rec, bins, bin_start, bin_end, use_fused_clear, part, masks, tile); \
  }

VA_RASTER_KERNEL(8)
VA_RASTER_KERNEL(16)
VA_RASTER_KERNEL(32)

// Persistent form: a grid that fills the GPU once; each workgroup takes tiles from a counter
// (zeroed before the launch) until none are left, so every wave reaches the exit.
#define VA_RASTER_PERSISTENT_KERNEL(T)                                                                        \
  extern "C" __global__ __launch_bounds__((T / 8) * (T / 8) * 64) void va_raster_p##T(                       \
      const VaDraw *__restrict__ draws, const VaFrame *__restrict__ frames, u32 nframes,                    \
      const u32 *__restrict__ rec, const u32 *__restrict__ bins, const u32 *__restrict__ bin_start,          \
      const u32 *__restrict__ bin_end, u32 ntiles, u32 use_fused_clear, u64 *__restrict__ part,              \
      const u32 *__restrict__ pair_total, u32 pair_cap, va_cmask *__restrict__ masks, u32 *__restrict__ tile_counter, \
      u64 *__restrict__ stamp)                                                                                \
  {                                                                                                           \
    __shared__ u32 s_next;                                                                                    \
    if(blockIdx.x == 0u && threadIdx.x == 0u)                                                                 \
      *stamp = (u64)wall_clock64();                                                                           \
    if(*pair_total > pair_cap)                                                                                \
      return;                                                                                                 \
    for(;;)                                                                                                   \
    {                                                                                                         \
      if(threadIdx.x == 0u)                                                                                   \
        s_next = atomicAdd(tile_counter, 1u);                                                                 \
      __syncthreads();                                                                                        \
      const u32 tile = s_next;                                                                                \
      __syncthreads();                                                                                        \
      if(tile >= ntiles)                                                                                      \
        break;                                                                                                \
      va_raster_tile<T>(draws, frames, nframes, rec, bins, bin_start, bin_end, use_fused_clear, part, masks, tile); \
    }                                                                                                         \
  }

VA_RASTER_PERSISTENT_KERNEL(8)
VA_RASTER_PERSISTENT_KERNEL(16)
VA_RASTER_PERSISTENT_KERNEL(32)
