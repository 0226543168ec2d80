// texel_decode.h - the four bytes the reference's texture fetch returns for one texel
// (reference texture_sampling.cpp:34-137, with its vendored BC2/BC3 block decoder
// 3rdparty/decompress.c). The transcode kernel (runtime_kernels.hip) writes them into an RGBA8
// shadow image, so the pipelines' samplers only ever read 4-byte texels; the tests compile this
// header for the host as well.
#pragma once

#include <stdint.h>

#ifndef VA_TD_FN
#define VA_TD_FN __device__ __forceinline__
#endif

#define VA_TEXFMT_PLAIN 0u  // 4 bytes at (y * width + x) * bpp, whatever the format
#define VA_TEXFMT_BC2 1u    // 4x4 blocks of 16 bytes: explicit 4-bit alpha, then a colour block
#define VA_TEXFMT_BC3 2u    // 4x4 blocks of 16 bytes: interpolated alpha, then a colour block

// n-bit endpoint channel to 8 bits: t = v * 255 + 2^(n-1), (t / 2^n + t) / 2^n.
VA_TD_FN uint32_t va_bc_widen(uint32_t v, uint32_t nbits)
{
  const uint32_t t = v * 255u + (1u << (nbits - 1u));
  return ((t >> nbits) + t) >> nbits;
}

// Palette entry `sel` between endpoints e0, e1: four colours, or (BC1's rule, which BC2 follows
// when c0 <= c1) three colours and black.
VA_TD_FN uint32_t va_bc_pick(uint32_t e0, uint32_t e1, uint32_t sel, bool four)
{
  if(sel == 0u) return e0;
  if(sel == 1u) return e1;
  if(four) return sel == 2u ? (2u * e0 + e1) / 3u : (e0 + 2u * e1) / 3u;
  return sel == 2u ? (e0 + e1) >> 1 : 0u;
}

// RGBA8 (r | g << 8 | b << 16 | a << 24) of texel px (0..15, row-major) of one block given as
// four little-endian words. BC3's colour block always uses four colours.
VA_TD_FN uint32_t va_bc_texel(const uint32_t q[4], uint32_t fmt, uint32_t px)
{
  uint32_t a;
  if(fmt == VA_TEXFMT_BC2)
    a = (uint32_t)(((((uint64_t)q[1] << 32) | q[0]) >> (4u * px)) & 15u) * 17u;
  else
  {
    const uint32_t a0 = q[0] & 0xffu, a1 = (q[0] >> 8) & 0xffu;
    const uint32_t k = (uint32_t)((((uint64_t)q[1] << 16) | (q[0] >> 16)) >> (3u * px)) & 7u;
    if(k == 0u) a = a0;
    else if(k == 1u) a = a1;
    else if(a0 > a1) a = ((8u - k) * a0 + (k - 1u) * a1) / 7u;
    else if(k == 6u) a = 0u;
    else if(k == 7u) a = 255u;
    else a = ((6u - k) * a0 + (k - 1u) * a1) / 5u;
  }
  const uint32_t c0 = q[2] & 0xffffu, c1 = q[2] >> 16;
  const uint32_t sel = (q[3] >> (2u * px)) & 3u;
  const bool four = fmt == VA_TEXFMT_BC3 || c0 > c1;
  const uint32_t r = va_bc_pick(va_bc_widen(c0 >> 11, 5u), va_bc_widen(c1 >> 11, 5u), sel, four);
  const uint32_t g = va_bc_pick(va_bc_widen((c0 >> 5) & 63u, 6u), va_bc_widen((c1 >> 5) & 63u, 6u), sel, four);
  const uint32_t b = va_bc_pick(va_bc_widen(c0 & 31u, 5u), va_bc_widen(c1 & 31u, 5u), sel, four);
  return r | g << 8 | b << 16 | a << 24;
}

// Little-endian word at byte `off` of an image of `bytes` bytes; bytes past the end read as 0.
VA_TD_FN uint32_t va_td_word(const uint8_t *img, uint64_t bytes, uint64_t off)
{
  if(off + 4u <= bytes && ((uintptr_t)(img + off) & 3u) == 0u)
    return *(const uint32_t *)(img + off);
  uint32_t w = 0u;
  for(uint32_t k = 0; k < 4u; k++)
    w |= (off + k < bytes ? (uint32_t)img[off + k] : 0u) << (8u * k);
  return w;
}

// The four bytes the reference's fetch returns for texel (x, y) of the subresource that starts at
// byte `sub` of an image of `bytes` bytes, `w` texels wide. Plain formats: the 4 bytes at
// (y * w + x) * bpp (texture_sampling.cpp:121-133). BC2/BC3: the texel of the 16-byte block at
// ((y / 4) * (w / 4) + x / 4) * 16, decoded (texture_sampling.cpp:95-118). Bytes past the end
// of the image read as 0 (the reference reads whatever follows).
VA_TD_FN uint32_t va_texel_bytes(const uint8_t *img, uint64_t bytes, uint64_t sub, uint32_t w, uint32_t bpp,
                                 uint32_t fmt, uint32_t x, uint32_t y)
{
  if(fmt == VA_TEXFMT_PLAIN)
    return va_td_word(img, bytes, sub + ((uint64_t)y * w + x) * bpp);
  const uint64_t blk = sub + ((uint64_t)(y >> 2) * (w >> 2) + (x >> 2)) * 16u;
  uint32_t q[4];
  for(uint32_t k = 0; k < 4u; k++)
    q[k] = va_td_word(img, bytes, blk + 4u * k);
  return va_bc_texel(q, fmt, ((y & 3u) << 2) | (x & 3u));
}
