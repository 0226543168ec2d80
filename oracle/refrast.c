/*
 * refrast.c - ORACLE (test infrastructure only; see refrast.h).
 *
 * Plain-C restatement of the reference draw path. Build with -ffp-contract=off so every
 * a*b+c stays two IEEE-rounded operations, as in the reference's x64 SSE code.
 */
#include "refrast.h"

#include <math.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>

static inline uint32_t f2u(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static inline float u2f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }

/* ---------------------------------------------------------------------------------------- */
/* Formats                                                                                   */

/* reference precompiled.cpp:3-36 (CalcSubresourceByteOffset) */
uint64_t rr_subresource_offset(const rr_image *img, uint32_t mip, uint32_t layer)
{
  uint64_t offs = 0;
  const uint32_t w = img->width, h = img->height, bpp = img->bpp;
  for(uint32_t m = 0; m < mip; m++)
  {
    uint32_t mw = w >> m; if(mw < 1) mw = 1;
    uint32_t mh = h >> m; if(mh < 1) mh = 1;
    offs += (uint64_t)mw * mh * bpp;
  }
  if(layer > 0)
  {
    uint32_t mw = w, mh = h;
    uint64_t slice = 0;
    for(uint32_t m = 0; m < img->mip_levels; m++)
    {
      slice += (uint64_t)mw * mh * bpp;
      mw = mw >> 1; if(mw < 1) mw = 1;
      mh = mh >> 1; if(mh < 1) mh = 1;
    }
    offs += slice * layer;
  }
  return offs;
}

/* IEEE binary32 -> binary16, round to nearest even (beyond the reference: RGBA16F targets). */
uint16_t rr_f32_to_f16(float f)
{
  uint32_t x = f2u(f);
  uint32_t sign = (x >> 16) & 0x8000u;
  uint32_t ax = x & 0x7fffffffu;
  if(ax >= 0x7f800000u)
    return (uint16_t)(sign | 0x7c00u | (ax > 0x7f800000u ? (0x200u | ((ax >> 13) & 0x3ffu)) : 0u));
  if(ax >= 0x477ff000u)
    return (uint16_t)(sign | 0x7c00u);
  if(ax < 0x38800000u)
  {
    uint32_t e = ax >> 23;
    if(e < 102)
      return (uint16_t)sign;
    uint32_t m = (ax & 0x7fffffu) | 0x800000u;
    uint32_t s = 126 - e;
    uint32_t q = m >> s;
    uint32_t rem = m & ((1u << s) - 1u);
    uint32_t half = 1u << (s - 1);
    if(rem > half || (rem == half && (q & 1u)))
      q++;
    return (uint16_t)(sign | q);
  }
  uint32_t r = ax - 0x38000000u;
  r = (r + 0xfffu + ((r >> 13) & 1u)) >> 13;
  return (uint16_t)(sign | r);
}

float rr_f16_to_f32(uint16_t h)
{
  uint32_t sign = ((uint32_t)h & 0x8000u) << 16;
  uint32_t e = (h >> 10) & 0x1fu;
  uint32_t m = h & 0x3ffu;
  if(e == 0)
  {
    if(m == 0)
      return u2f(sign);
    /* subnormal: m * 2^-24, exact in binary32 */
    float v = (float)m * u2f(0x33800000u);
    return sign ? -v : v;
  }
  if(e == 31)
    return u2f(sign | 0x7f800000u | (m << 13));
  return u2f(sign | ((e + 112u) << 23) | (m << 13));
}

/* log2 for LOD selection (beyond the reference, which never selects a mip). Deterministic:
 * exponent extraction plus the atanh series log2(m) = (2/ln2) * (s + s^3/3 + s^5/5 + s^7/7 + s^9/9),
 * s = (m-1)/(m+1), with m reduced to [sqrt(1/2), sqrt(2)); only +,-,*,/ in a fixed order so the
 * GPU path reproduces it bit for bit. Non-positive or NaN input returns -128. */
float rr_log2(float x)
{
  if(!(x > 0.0f))
    return -128.0f;
  if(x == INFINITY)
    return 128.0f;
  uint32_t b = f2u(x);
  int e = (int)((b >> 23) & 0xffu);
  if(e == 0)
    return -128.0f; /* subnormal: far below any mip level */
  e -= 127;
  float m = u2f((b & 0x7fffffu) | 0x3f800000u);
  if(m > 1.41421353816986083984375f)
  {
    m = m * 0.5f;
    e += 1;
  }
  float s = (m - 1.0f) / (m + 1.0f);
  float s2 = s * s;
  float p = 0.111111111f;       /* 1/9 */
  p = p * s2 + 0.142857143f;    /* 1/7 */
  p = p * s2 + 0.2f;            /* 1/5 */
  p = p * s2 + 0.333333343f;    /* 1/3 */
  p = p * s2 + 1.0f;
  p = p * s;
  return (float)e + p * 2.88539004f; /* 2/ln(2) */
}

/* ---------------------------------------------------------------------------------------- */
/* Texture sampling                                                                          */

/* BC2 / BC3 block decode, as the reference's vendored decoder does it (3rdparty/decompress.c,
 * Dobell / Chajdas, MIT; called from texture_sampling.cpp:100-103). Output is 16 RGBA8 texels in
 * row-major order. Two traits of that decoder are kept: BC2's colour block follows BC1's rule
 * (three colours and black when c0 <= c1, decompress.c:141-206, 330-350), while BC3's colour
 * block always uses four colours (decompress.c:300-318). */

/* n-bit endpoint channel to 8 bits (decompress.c:122-135): t = v * 255 + 2^(n-1), (t / 2^n + t) / 2^n. */
static uint32_t bc_widen(uint32_t v, uint32_t nbits)
{
  const uint32_t s = 1u << nbits;
  const uint32_t t = v * 255u + (s >> 1);
  return (t / s + t) / s;
}

static void bc_colour(const uint8_t *blk, int bc1_rule, const uint8_t alpha[16], uint8_t out[64])
{
  const uint32_t c0 = (uint32_t)blk[0] | (uint32_t)blk[1] << 8;
  const uint32_t c1 = (uint32_t)blk[2] | (uint32_t)blk[3] << 8;
  const uint32_t code = (uint32_t)blk[4] | (uint32_t)blk[5] << 8 | (uint32_t)blk[6] << 16 | (uint32_t)blk[7] << 24;
  uint32_t pal[4][3];
  pal[0][0] = bc_widen(c0 >> 11, 5); pal[0][1] = bc_widen((c0 >> 5) & 63u, 6); pal[0][2] = bc_widen(c0 & 31u, 5);
  pal[1][0] = bc_widen(c1 >> 11, 5); pal[1][1] = bc_widen((c1 >> 5) & 63u, 6); pal[1][2] = bc_widen(c1 & 31u, 5);
  const int four = !bc1_rule || c0 > c1;
  for(int k = 0; k < 3; k++)
  {
    pal[2][k] = four ? (2u * pal[0][k] + pal[1][k]) / 3u : (pal[0][k] + pal[1][k]) / 2u;
    pal[3][k] = four ? (pal[0][k] + 2u * pal[1][k]) / 3u : 0u;
  }
  for(int i = 0; i < 16; i++)
  {
    const uint32_t sel = (code >> (2 * i)) & 3u;
    out[i * 4 + 0] = (uint8_t)pal[sel][0];
    out[i * 4 + 1] = (uint8_t)pal[sel][1];
    out[i * 4 + 2] = (uint8_t)pal[sel][2];
    out[i * 4 + 3] = alpha[i];
  }
}

void rr_decode_bc_block(uint32_t format, const uint8_t blk[16], uint8_t out[64])
{
  uint8_t alpha[16];
  if(format == RR_FMT_BC2_UNORM_BLOCK)
  {
    /* 4-bit explicit alpha, row r in bytes 2r..2r+1, texel c in bits 4c.. (decompress.c:335-345) */
    for(int i = 0; i < 16; i++)
    {
      const uint32_t row = (uint32_t)blk[2 * (i >> 2)] | (uint32_t)blk[2 * (i >> 2) + 1] << 8;
      alpha[i] = (uint8_t)(((row >> (4 * (i & 3))) & 15u) * 17u);
    }
    bc_colour(blk + 8, 1, alpha, out);
    return;
  }
  /* BC3: two 8-bit endpoints and sixteen 3-bit selectors from byte 2 (decompress.c:245-298) */
  const uint32_t a0 = blk[0], a1 = blk[1];
  uint64_t sel = 0;
  for(int k = 0; k < 6; k++)
    sel |= (uint64_t)blk[2 + k] << (8 * k);
  for(int i = 0; i < 16; i++)
  {
    const uint32_t k = (uint32_t)(sel >> (3 * i)) & 7u;
    uint32_t a;
    if(k == 0) a = a0;
    else if(k == 1) a = a1;
    else if(a0 > a1) a = ((8u - k) * a0 + (k - 1u) * a1) / 7u;
    else if(k == 6) a = 0;
    else if(k == 7) a = 255;
    else a = ((6u - k) * a0 + (k - 1u) * a1) / 5u;
    alpha[i] = (uint8_t)a;
  }
  bc_colour(blk + 8, 0, alpha, out);
}

/* reference texture_sampling.cpp:5-137: an 8-entry per-thread LRU of 4x4 texel blocks. It is a
 * pure memo (a hit returns exactly what a direct fetch /255 returns); kept for the CPU timing. */
typedef struct tc_entry {
  struct tc_entry *prev, *next;
  int tlx, tly;
  const uint8_t *base;
  float px[4][4][4];
} tc_entry;

static _Thread_local tc_entry tc_data[8];
static _Thread_local tc_entry *tc_head;
/* Every public entry point bumps the generation, so no thread can hit on a block cached from
 * image memory that the caller has since freed or rewritten (the key is the subresource's base
 * address, which a later allocation can reuse). */
static _Atomic uint64_t tc_generation = 1;
static _Thread_local uint64_t tc_seen;

static void tc_init(void)
{
  memset(tc_data, 0, sizeof(tc_data));
  for(int i = 0; i < 8; i++)
  {
    tc_data[i].tlx = tc_data[i].tly = -1;
    tc_data[i].prev = i > 0 ? &tc_data[i - 1] : NULL;
    tc_data[i].next = i < 7 ? &tc_data[i + 1] : NULL;
  }
  tc_head = &tc_data[0];
}

static const float *tc_fetch(const rr_image *tex, uint64_t byte_offs, uint32_t row_w,
                            uint32_t row_h, int x, int y)
{
  uint64_t gen = tc_generation;
  if(!tc_head || tc_seen != gen)
  {
    tc_init();
    tc_seen = gen;
  }
  int tlx = x & ~3, tly = y & ~3;
  const uint8_t *base = tex->pixels + byte_offs;
  tc_entry *e = tc_head;
  for(;;)
  {
    if(e->base == base && e->tlx == tlx && e->tly == tly)
    {
      if(e->prev)
      {
        e->prev->next = e->next;
        if(e->next)
          e->next->prev = e->prev;
        tc_head->prev = e;
        e->prev = NULL;
        e->next = tc_head;
        tc_head = e;
      }
      return e->px[y & 3][x & 3];
    }
    if(!e->next)
      break;
    e = e->next;
  }
  e->prev->next = NULL;
  tc_head->prev = e;
  e->prev = NULL;
  e->next = tc_head;
  tc_head = e;
  e->base = base;
  e->tlx = tlx;
  e->tly = tly;
  /* Bytes past the end of the image read as 0 (the reference reads whatever follows). */
  const uint64_t img_bytes = rr_subresource_offset(tex, 0, tex->array_layers);
  const uint64_t boff = (uint64_t)(base - tex->pixels);
  if(tex->format == RR_FMT_BC2_UNORM_BLOCK || tex->format == RR_FMT_BC3_UNORM_BLOCK)
  {
    /* reference texture_sampling.cpp:95-118: decode the 4x4 block holding the quad; the row
     * pitch in blocks is the width >> 2 of the sampled subresource. */
    uint8_t blk[16], rgba[64];
    const uint64_t at = boff + ((uint64_t)(tly >> 2) * (row_w >> 2) + (uint32_t)(tlx >> 2)) * 16u;
    for(int k = 0; k < 16; k++)
      blk[k] = at + k < img_bytes ? tex->pixels[at + k] : 0;
    rr_decode_bc_block(tex->format, blk, rgba);
    for(int i = 0; i < 16; i++)
      for(int c = 0; c < 4; c++)
        e->px[i >> 2][i & 3][c] = (float)rgba[i * 4 + c] / 255.0f;
    return e->px[y & 3][x & 3];
  }
  /* reference texture_sampling.cpp:121-133: 4 bytes at (y * width + x) * bpp, whatever the
   * format. The reference fills the whole 4x4 block, reading past the row or image end at the
   * edges; here only in-bounds texels are read (the others are never returned for in-bounds x, y). */
  for(int r = 0; r < 4; r++)
    for(int c = 0; c < 4; c++)
    {
      uint32_t xx = (uint32_t)(tlx + c), yy = (uint32_t)(tly + r);
      if(xx < row_w && yy < row_h)
      {
        const uint64_t at = boff + ((uint64_t)yy * row_w + xx) * tex->bpp;
        for(int k = 0; k < 4; k++)
          e->px[r][c][k] = (float)(at + k < img_bytes ? tex->pixels[at + k] : 0) / 255.0f;
      }
      else
        e->px[r][c][0] = e->px[r][c][1] = e->px[r][c][2] = e->px[r][c][3] = 0.0f;
    }
  return e->px[y & 3][x & 3];
}

/* Direct texel read /255 in memory byte order (reference texture_sampling.cpp:121-133). */
static void texel(const rr_image *tex, uint64_t byte_offs, uint32_t row_w, uint32_t row_h, int x,
                  int y, float out[4])
{
  const float *t = tc_fetch(tex, byte_offs, row_w, row_h, x, y);
  out[0] = t[0];
  out[1] = t[1];
  out[2] = t[2];
  out[3] = t[3];
}

/* reference texture_sampling.cpp:139-184 (sample_tex_wrapped): bilinear, wrap, no texel-centre
 * offset. One guard beyond the reference: when u - floor(u) rounds to 1.0 the left index wraps
 * too (the reference reads one texel past the row there). */
void rr_sample_tex_wrapped(float u, float v, const rr_image *tex, uint64_t byte_offs, float out[4])
{
  atomic_fetch_add(&tc_generation, 1);
  const uint32_t W = tex->width, H = tex->height;
  u = u - floorf(u);
  v = v - floorf(v);
  u *= (float)W;
  v *= (float)H;
  int iu0 = (int)u, iv0 = (int)v;
  int iu1 = iu0 + 1, iv1 = iv0 + 1;
  if(iu1 >= (int)W)
    iu1 -= W;
  if(iv1 >= (int)H)
    iv1 -= H;
  float fu = u - (float)iu0, fv = v - (float)iv0;
  if(iu0 >= (int)W)
    iu0 -= W;
  if(iv0 >= (int)H)
    iv0 -= H;
  float inv_fu = 1.0f - fu, inv_fv = 1.0f - fv;
  float TL[4], TR[4], BL[4], BR[4];
  texel(tex, byte_offs, W, H, iu0, iv0, TL);
  texel(tex, byte_offs, W, H, iu1, iv0, TR);
  texel(tex, byte_offs, W, H, iu0, iv1, BL);
  texel(tex, byte_offs, W, H, iu1, iv1, BR);
  for(int c = 0; c < 4; c++)
  {
    float top = TL[c] * inv_fu + TR[c] * fu;
    float bottom = BL[c] * inv_fu + BR[c] * fu;
    out[c] = top * inv_fv + bottom * fv;
  }
}

/* One mip level, one layer. REPEAT + LINEAR is exactly sample_tex_wrapped above; CLAMP_TO_EDGE
 * and NEAREST extend it with the same no-half-texel convention (beyond the reference). */
static void sample_level(const rr_image *img, uint32_t level, uint32_t layer, float u, float v,
                         uint32_t filter, uint32_t addr_u, uint32_t addr_v, float out[4])
{
  uint32_t W = img->width >> level; if(W < 1) W = 1;
  uint32_t H = img->height >> level; if(H < 1) H = 1;
  uint64_t offs = rr_subresource_offset(img, level, layer);
  if(!(u == u))
    u = 0.0f;
  if(!(v == v))
    v = 0.0f;
  if(addr_u == RR_ADDR_REPEAT)
    u = u - floorf(u);
  else
  {
    if(!(u >= 0.0f)) u = 0.0f;
    if(u > 1.0f) u = 1.0f;
  }
  if(addr_v == RR_ADDR_REPEAT)
    v = v - floorf(v);
  else
  {
    if(!(v >= 0.0f)) v = 0.0f;
    if(v > 1.0f) v = 1.0f;
  }
  float U = u * (float)W, V = v * (float)H;
  int i0 = (int)U, j0 = (int)V;
  if(filter == RR_FILTER_NEAREST)
  {
    if(i0 >= (int)W) i0 = addr_u == RR_ADDR_REPEAT ? i0 - (int)W : (int)W - 1;
    if(j0 >= (int)H) j0 = addr_v == RR_ADDR_REPEAT ? j0 - (int)H : (int)H - 1;
    texel(img, offs, W, H, i0, j0, out);
    return;
  }
  int i1 = i0 + 1, j1 = j0 + 1;
  float fu = U - (float)i0, fv = V - (float)j0;
  if(addr_u == RR_ADDR_REPEAT)
  {
    if(i1 >= (int)W) i1 -= W;
    if(i0 >= (int)W) i0 -= W;
  }
  else
  {
    if(i1 > (int)W - 1) i1 = W - 1;
    if(i0 > (int)W - 1) i0 = W - 1;
  }
  if(addr_v == RR_ADDR_REPEAT)
  {
    if(j1 >= (int)H) j1 -= H;
    if(j0 >= (int)H) j0 -= H;
  }
  else
  {
    if(j1 > (int)H - 1) j1 = H - 1;
    if(j0 > (int)H - 1) j0 = H - 1;
  }
  float inv_fu = 1.0f - fu, inv_fv = 1.0f - fv;
  float TL[4], TR[4], BL[4], BR[4];
  texel(img, offs, W, H, i0, j0, TL);
  texel(img, offs, W, H, i1, j0, TR);
  texel(img, offs, W, H, i0, j1, BL);
  texel(img, offs, W, H, i1, j1, BR);
  for(int c = 0; c < 4; c++)
  {
    float top = TL[c] * inv_fu + TR[c] * fu;
    float bottom = BL[c] * inv_fu + BR[c] * fu;
    out[c] = top * inv_fv + bottom * fv;
  }
}

/* Full sampler (beyond the reference, which samples mip 0 of layer 0 only): LOD from the
 * coordinate's screen-space differences, Vulkan 1.0 level selection and filtering (15.6.7-8),
 * array layer = clamp(floor(z + 0.5)). */
static void sample_full(const rr_image *img, const rr_sampler *s, const float coord[3],
                        const float ddx[2], const float ddy[2], int use_lod, float out[4])
{
  float lambda = 0.0f;
  if(use_lod)
  {
    float W0 = (float)img->width, H0 = (float)img->height;
    float ax = ddx[0] * W0, ay = ddx[1] * H0;
    float bx = ddy[0] * W0, by = ddy[1] * H0;
    float rx = ax * ax + ay * ay;
    float ry = bx * bx + by * by;
    float r = rx > ry ? rx : ry;
    lambda = 0.5f * rr_log2(r);
  }
  lambda = lambda + s->mip_lod_bias;
  if(!(lambda >= s->min_lod))
    lambda = s->min_lod;
  if(lambda > s->max_lod)
    lambda = s->max_lod;
  uint32_t filter = lambda > 0.0f ? s->min_filter : s->mag_filter;
  float q = (float)(img->mip_levels - 1);
  float d = lambda;
  if(!(d >= 0.0f))
    d = 0.0f;
  if(d > q)
    d = q;
  int layer = (int)floorf(coord[2] + 0.5f);
  if(layer < 0)
    layer = 0;
  if(layer > (int)img->array_layers - 1)
    layer = (int)img->array_layers - 1;
  if(s->mipmap_mode == RR_MIPMAP_NEAREST)
  {
    int level = d <= 0.5f ? 0 : (int)ceilf(d + 0.5f) - 1;
    if(level > (int)q)
      level = (int)q;
    sample_level(img, level, layer, coord[0], coord[1], filter, s->address_u, s->address_v, out);
    return;
  }
  int dhi = (int)floorf(d);
  int dlo = dhi + 1 > (int)q ? (int)q : dhi + 1;
  float delta = d - (float)dhi;
  sample_level(img, dhi, layer, coord[0], coord[1], filter, s->address_u, s->address_v, out);
  if(delta != 0.0f)
  {
    float b[4];
    sample_level(img, dlo, layer, coord[0], coord[1], filter, s->address_u, s->address_v, b);
    float inv = 1.0f - delta;
    for(int c = 0; c < 4; c++)
      out[c] = out[c] * inv + b[c] * delta;
  }
}

/* reference texture_sampling.cpp:186-250 (sample_cube_wrapped) face selection: six tests in the
 * order +X, -X, +Y, -Y, +Z, -Z, each overriding the previous ones, so ties go to Z, then Y. The
 * face coordinate is 0.5 * (sc / |ma| + 1). When no test holds (a NaN component) the reference
 * reads uninitialised values; here face 0 with sc = tc = 0, |ma| = 1 (parity unpinned there). */
static void cube_face(float x, float y, float z, uint32_t *face, float *u, float *v)
{
  const float ax = fabsf(x), ay = fabsf(y), az = fabsf(z);
  const int px = x > 0.0f, py = y > 0.0f, pz = z > 0.0f;
  float axis = 1.0f, sc = 0.0f, tc = 0.0f;
  uint32_t f = 0;
  if(px && ax >= ay && ax >= az) { axis = ax; sc = -z; tc = -y; f = 0; }
  if(!px && ax >= ay && ax >= az) { axis = ax; sc = z; tc = -y; f = 1; }
  if(py && ay >= ax && ay >= az) { axis = ay; sc = x; tc = z; f = 2; }
  if(!py && ay >= ax && ay >= az) { axis = ay; sc = x; tc = -z; f = 3; }
  if(pz && az >= ax && az >= ay) { axis = az; sc = x; tc = -y; f = 4; }
  if(!pz && az >= ax && az >= ay) { axis = az; sc = -x; tc = -y; f = 5; }
  *face = f;
  *u = 0.5f * (sc / axis + 1.0f);
  *v = 0.5f * (tc / axis + 1.0f);
}

void rr_sample_cube_wrapped(float x, float y, float z, const rr_image *tex, float out[4])
{
  uint32_t face;
  float u, v;
  cube_face(x, y, z, &face, &u, &v);
  rr_sample_tex_wrapped(u, v, tex, rr_subresource_offset(tex, 0, face), out);
}

/* Cube sampling through the full sampler: the face is the array layer, no derivative LOD (the
 * level follows the sampler's bias and clamps). With REPEAT + LINEAR and a single level it is
 * sample_cube_wrapped exactly. */
static void sample_cube_full(const rr_image *img, const rr_sampler *s, float x, float y, float z, float out[4])
{
  uint32_t face;
  float c[3], zero[2] = {0.0f, 0.0f};
  cube_face(x, y, z, &face, &c[0], &c[1]);
  c[2] = (float)face;
  sample_full(img, s, c, zero, zero, 0, out);
}

void rr_sample_cube(const rr_image *img, const rr_sampler *s, const float dir[3], float out[4])
{
  atomic_fetch_add(&tc_generation, 1);
  sample_cube_full(img, s, dir[0], dir[1], dir[2], out);
}

/* ---------------------------------------------------------------------------------------- */
/* Vertex fetch and the restated shaders                                                     */

/* reference spirv_compile.cpp:572-627 (GetVertexAttributeData) */
void rr_get_vertex_attribute(const rr_state *st, uint32_t vtx, uint32_t attr, float out[4])
{
  const rr_vattr *a = &st->pipeline->vattrs[attr];
  out[0] = 0.0f; out[1] = 0.0f; out[2] = 0.0f; out[3] = 1.0f;
  const uint8_t *p = st->vbs[a->binding] + a->offset + (uint64_t)a->stride * vtx;
  float f[4];
  uint32_t u;
  switch(a->format)
  {
    case RR_FMT_R32G32B32A32_SFLOAT: case 107: case 108:
      memcpy(f, p, 16); out[3] = f[3]; out[2] = f[2]; out[1] = f[1]; out[0] = f[0]; break;
    case RR_FMT_R32G32B32_SFLOAT: case 104: case 105:
      memcpy(f, p, 12); out[2] = f[2]; out[1] = f[1]; out[0] = f[0]; break;
    case RR_FMT_R32G32_SFLOAT: case 101: case 102:
      memcpy(f, p, 8); out[1] = f[1]; out[0] = f[0]; break;
    case RR_FMT_R32_SFLOAT: case 98: case 99:
      memcpy(f, p, 4); out[0] = f[0]; break;
    case RR_FMT_R8G8B8A8_UNORM:
      memcpy(&u, p, 4);
      out[0] = (float)((u & 0x000000ffu) >> 0) / 255.0f;
      out[1] = (float)((u & 0x0000ff00u) >> 8) / 255.0f;
      out[2] = (float)((u & 0x00ff0000u) >> 16) / 255.0f;
      out[3] = (float)((u & 0xff000000u) >> 24) / 255.0f;
      break;
    default: break;
  }
}

static void pc_float(const rr_state *st, uint32_t off, float *out, int n)
{
  memcpy(out, st->pushconsts + off, 4 * n);
}

/* Uniform-buffer floats of set 0, binding `binding`: reference spirv_compile.cpp:552-558
 * (GetDescriptorBufferPointer: the bound buffer's bytes at the descriptor's offset). */
static void ubo_float(const rr_state *st, uint32_t binding, uint32_t off, float *out, int n)
{
  memcpy(out, st->ubo[binding] + off, 4 * n);
}

/* reference spirv_compile.cpp:423-431 (Float4x4TimesVec4), column-major m[col*4+row] */
static void mat4_times_vec4(const float *m, const float v[4], float out[4])
{
  for(int r = 0; r < 4; r++)
  {
    float acc = 0.0f;
    for(int c = 0; c < 4; c++)
      acc += m[c * 4 + r] * v[c];
    out[r] = acc;
  }
}

/* reference spirv_compile.cpp:433-441 (Float3x3TimesVec3); columns are 16 bytes apart */
static void mat3_times_vec3(const float *m, const float v[3], float out[3])
{
  for(int r = 0; r < 3; r++)
  {
    float acc = 0.0f;
    for(int c = 0; c < 3; c++)
      acc += m[c * 4 + r] * v[c];
    out[r] = acc;
  }
}

/* VS output padding: reference spirv_compile.cpp:2059-2074 (vec<4 repeats component 0). */
static void store_interp(rr_vertex *o, int loc, const float *v, int n)
{
  for(int i = 0; i < 4; i++)
    o->interps[loc][i] = i < n ? v[i] : v[0];
}

static void vs_program(const rr_state *st, uint32_t vtx, rr_vertex *o)
{
  memset(o, 0, sizeof(*o));
  float mvp[16];
  pc_float(st, 0, mvp, 16);
  switch(st->pipeline->vs_program)
  {
    case RR_VS_MVP_POS4_COL:
    {
      float pos[4], col[4];
      rr_get_vertex_attribute(st, vtx, 0, pos);
      rr_get_vertex_attribute(st, vtx, 1, col);
      mat4_times_vec4(mvp, pos, o->position);
      store_interp(o, 0, col, 4);
      break;
    }
    case RR_VS_UBO_POS4_COL:
    {
      float pos[4], col[4], umvp[16];
      rr_get_vertex_attribute(st, vtx, 0, pos);
      rr_get_vertex_attribute(st, vtx, 1, col);
      ubo_float(st, RR_UBO_BINDING, 0, umvp, 16);
      mat4_times_vec4(umvp, pos, o->position);
      store_interp(o, 0, col, 4);
      break;
    }
    case RR_VS_MVP_NORMAL_UV:
    {
      float pos[4], nrm[4], uv[4], nmat[12], p4[4], on[3];
      rr_get_vertex_attribute(st, vtx, 0, pos);
      rr_get_vertex_attribute(st, vtx, 1, nrm);
      rr_get_vertex_attribute(st, vtx, 2, uv);
      pc_float(st, 64, nmat, 12);
      p4[0] = pos[0]; p4[1] = pos[1]; p4[2] = pos[2]; p4[3] = 1.0f;
      mat4_times_vec4(mvp, p4, o->position);
      mat3_times_vec3(nmat, nrm, on);
      store_interp(o, 0, on, 3);
      store_interp(o, 1, uv, 2);
      break;
    }
    default: break;
  }
}

/* Interpolated input component: reference spirv_compile.cpp:2159-2240 with CreateDot
 * (:629-643) over (b0,b1,b2,b3=0) and (a0,a1,a2,0). */
static float interp(const float b[4], const rr_vertex tri[3], int loc, int comp)
{
  float acc = b[0] * tri[0].interps[loc][comp];
  acc = acc + b[1] * tri[1].interps[loc][comp];
  acc = acc + b[2] * tri[2].interps[loc][comp];
  acc = acc + b[3] * 0.0f;
  return acc;
}

static const float kLightDir[3] = {0.48f, 0.6f, 0.64f};

/* Lambert term shared by the lit programs: normalize (reference :1635-1647), dot (:629-643),
 * FMax as select(a > b) (:1535-1545), then light = ndl * 0.75 + 0.25. */
static float lambert(const float b[4], const rr_vertex tri[3])
{
  float n[3];
  for(int c = 0; c < 3; c++)
    n[c] = interp(b, tri, 0, c);
  float sq = n[0] * n[0];
  sq = sq + n[1] * n[1];
  sq = sq + n[2] * n[2];
  float len = sqrtf(sq);
  float inv = 1.0f / len;
  for(int c = 0; c < 3; c++)
    n[c] = n[c] * inv;
  float ndl = n[0] * kLightDir[0];
  ndl = ndl + n[1] * kLightDir[1];
  ndl = ndl + n[2] * kLightDir[2];
  ndl = ndl > 0.0f ? ndl : 0.0f;
  return ndl * 0.75f + 0.25f;
}

static const float kTapScale[8] = {1.0f, 2.0f, 0.5f, 4.0f, 1.0f, 3.0f, 0.25f, 8.0f};
static const float kTapOffX[8] = {0.0f, 0.125f, 0.25f, 0.5f, 0.75f, 0.375f, 0.0625f, 0.5f};
static const float kTapOffY[8] = {0.0f, 0.25f, 0.125f, 0.75f, 0.5f, 0.0625f, 0.375f, 0.25f};
static const float kTapLayer[8] = {0.0f, 1.0f, 2.0f, 3.0f, 0.0f, 1.0f, 2.0f, 3.0f};
static const float kTapWeight[8] = {0.25f, 0.125f, 0.125f, 0.125f, 0.125f, 0.125f, 0.0625f, 0.0625f};

static void tex_uv(const rr_state *st, int binding, const float b3[3][4], const rr_vertex tri[3],
                   int loc, float out[4])
{
  float c[3], ddx[2], ddy[2];
  float u0 = interp(b3[0], tri, loc, 0), v0 = interp(b3[0], tri, loc, 1);
  float u1 = interp(b3[1], tri, loc, 0), v1 = interp(b3[1], tri, loc, 1);
  float u2 = interp(b3[2], tri, loc, 0), v2 = interp(b3[2], tri, loc, 1);
  c[0] = u0; c[1] = v0; c[2] = 0.0f;
  ddx[0] = u1 - u0; ddx[1] = v1 - v0;
  ddy[0] = u2 - u0; ddy[1] = v2 - v0;
  sample_full(st->tex[binding], st->samp[binding], c, ddx, ddy, 1, out);
}

static void fs_program(const rr_state *st, float pixdepth, const float b3[3][4],
                       const rr_vertex tri[3], float out[4])
{
  (void)pixdepth;
  const float *b = b3[0];
  switch(st->pipeline->fs_program)
  {
    case RR_FS_COLOR:
      for(int c = 0; c < 4; c++)
        out[c] = interp(b, tri, 0, c);
      break;
    case RR_FS_TEX_ONLY:
      tex_uv(st, 0, b3, tri, 1, out);
      break;
    case RR_FS_UBO_TINT:
    {
      /* out = in0 * tint + bias, component-wise (OpFMul then OpFAdd, two roundings) */
      float tint[4], bias[4];
      ubo_float(st, RR_UBO_BINDING, 64, tint, 4);
      ubo_float(st, RR_UBO_BINDING, 80, bias, 4);
      for(int c = 0; c < 4; c++)
        out[c] = interp(b, tri, 0, c) * tint[c] + bias[c];
      break;
    }
    case RR_FS_CUBE_ENV:
      sample_cube_full(st->tex[0], st->samp[0], interp(b, tri, 0, 0), interp(b, tri, 0, 1), interp(b, tri, 0, 2), out);
      break;
    case RR_FS_LAMBERT_TEX:
    case RR_FS_LAMBERT_TEX_A:
    {
      float light = lambert(b, tri);
      float t[4];
      tex_uv(st, 0, b3, tri, 1, t);
      out[0] = t[0] * light;
      out[1] = t[1] * light;
      out[2] = t[2] * light;
      if(st->pipeline->fs_program == RR_FS_LAMBERT_TEX_A)
        pc_float(st, 112, &out[3], 1);
      else
        out[3] = 1.0f;
      break;
    }
    case RR_FS_MULTITAP8:
    {
      float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
      float uv[3][2];
      for(int p = 0; p < 3; p++)
      {
        uv[p][0] = interp(b3[p], tri, 1, 0);
        uv[p][1] = interp(b3[p], tri, 1, 1);
      }
      for(int k = 0; k < 8; k++)
      {
        float cx[3], cy[3];
        for(int p = 0; p < 3; p++)
        {
          cx[p] = uv[p][0] * kTapScale[k] + kTapOffX[k];
          cy[p] = uv[p][1] * kTapScale[k] + kTapOffY[k];
        }
        float c[3] = {cx[0], cy[0], kTapLayer[k]};
        float ddx[2] = {cx[1] - cx[0], cy[1] - cy[0]};
        float ddy[2] = {cx[2] - cx[0], cy[2] - cy[0]};
        float t[4];
        sample_full(st->tex[0], st->samp[0], c, ddx, ddy, 1, t);
        for(int i = 0; i < 4; i++)
          acc[i] = acc[i] + t[i] * kTapWeight[k];
      }
      for(int i = 0; i < 4; i++)
        out[i] = acc[i];
      break;
    }
    default:
      out[0] = out[1] = out[2] = out[3] = 0.0f;
      break;
  }
}

/* ---------------------------------------------------------------------------------------- */
/* Rasterizer                                                                                */

typedef struct tri_work {
  const rr_state *state;
  int ABx, ABy, ACx, ACy;
  const rr_vertex *vsout;
  const int32_t *tri; /* 3 x (x, y) window coords */
  int barymul, area2;
  float invarea;
  float invw[3], depth[3];
  int minx, miny, maxx, maxy;
} tri_work;

/* int32 subtraction and negation wrapping modulo 2^32, as the reference's x64 build computes
 * them (far-off vertices reach the wrapping range; C leaves signed overflow undefined). */
static inline int rr_wsub(int a, int b) { return (int)((uint32_t)a - (uint32_t)b); }
static inline int rr_wneg(int a) { return (int)(0u - (uint32_t)a); }

/* float -> int32 as the reference's x64 build converts (cvttss2si: truncation toward zero, and
 * the "integer indefinite" 0x80000000 for NaN and every value outside [-2^31, 2^31)). C leaves
 * the out-of-range case undefined, so it is spelled out here. */
int32_t rr_f2i(float f)
{
  if(f >= -2147483648.0f && f < 2147483648.0f)
    return (int32_t)f;
  return INT32_MIN;
}

/* reference rasterizer.cpp:239-253 (ToWindow): each op is one FP32 rounding, left to right. */
void rr_to_window(uint32_t w, uint32_t h, const float p[4], int32_t out[2])
{
  out[0] = rr_f2i((p[0] / p[3] + 1.0f) * 0.5f * (float)w);
  out[1] = rr_f2i((p[1] * -1.0f / p[3] + 1.0f) * 0.5f * (float)h);
}

/* reference rasterizer.cpp:100-119 (GetIndex). The draw's vertexOffset is recorded
 * (cmd_record.cpp:125-135, st->vertex_offset here) but GetIndex never adds it. */
static uint32_t get_index(const rr_state *st, uint32_t i, int indexed)
{
  if(!indexed)
    return i;
  if(st->index_type == 0)
  {
    uint16_t v;
    memcpy(&v, st->ib + 2ull * i, 2);
    return v;
  }
  uint32_t v;
  memcpy(&v, st->ib + 4ull * i, 4);
  return v;
}

/* reference rasterizer.cpp:121-237 (ShadeVerts): list = whole triangles only; strip keeps
 * winding with (N,N+1,N+2),(N+2,N+1,N+3),... */
static size_t shade_verts(const rr_state *st, int num_verts, uint32_t first, int indexed,
                          rr_vertex **out_buf)
{
  size_t cap = num_verts > 3 ? (size_t)num_verts * 3 : 12;
  rr_vertex *out = (rr_vertex *)malloc(cap * sizeof(rr_vertex));
  size_t n = 0;
  uint32_t vi = first;
  rr_vertex t[4];
  if(st->pipeline->topology == RR_TOPO_TRIANGLE_LIST)
  {
    for(int v = 0; v <= num_verts - 3; v += 3)
    {
      vs_program(st, get_index(st, vi++, indexed), &t[0]);
      vs_program(st, get_index(st, vi++, indexed), &t[1]);
      vs_program(st, get_index(st, vi++, indexed), &t[2]);
      out[n++] = t[0];
      out[n++] = t[1];
      out[n++] = t[2];
    }
  }
  else if(st->pipeline->topology == RR_TOPO_TRIANGLE_STRIP && num_verts >= 3)
  {
    vs_program(st, get_index(st, vi++, indexed), &t[0]);
    vs_program(st, get_index(st, vi++, indexed), &t[1]);
    vs_program(st, get_index(st, vi++, indexed), &t[2]);
    out[n++] = t[0];
    out[n++] = t[1];
    out[n++] = t[2];
    num_verts -= 3;
    if(num_verts > 0)
    {
      vs_program(st, get_index(st, vi++, indexed), &t[3]);
      num_verts--;
      out[n++] = t[2];
      out[n++] = t[1];
      out[n++] = t[3];
    }
    while(num_verts > 0)
    {
      t[0] = t[2];
      t[1] = t[3];
      vs_program(st, get_index(st, vi++, indexed), &t[2]);
      num_verts--;
      out[n++] = t[0];
      out[n++] = t[1];
      out[n++] = t[2];
      if(num_verts > 0)
      {
        vs_program(st, get_index(st, vi++, indexed), &t[3]);
        num_verts--;
        out[n++] = t[2];
        out[n++] = t[1];
        out[n++] = t[3];
      }
    }
  }
  *out_buf = out;
  return n;
}

/* Perspective-correct barycentrics at integer corner (x, y): reference rasterizer.cpp:282-310
 * (barycentric) and :545-588. Returns coverage; writes the normalised (pre-perspective) depth. */
static int bary_at(const tri_work *w, int x, int y, float out[4], float *pixdepth)
{
  const int PAx = rr_wsub(w->tri[0], x);
  const int PAy = rr_wsub(w->tri[1], y);
  const int ux = (int)((uint32_t)w->ACx * (uint32_t)PAy - (uint32_t)w->ACy * (uint32_t)PAx);
  const int uy = (int)((uint32_t)PAx * (uint32_t)w->ABy - (uint32_t)PAy * (uint32_t)w->ABx);
  int b0 = (int)((uint32_t)w->area2 - ((uint32_t)ux + (uint32_t)uy));
  int b1 = ux, b2 = uy;
  b0 = (int)((uint32_t)b0 * (uint32_t)w->barymul);
  b1 = (int)((uint32_t)b1 * (uint32_t)w->barymul);
  b2 = (int)((uint32_t)b2 * (uint32_t)w->barymul);
  int covered = b0 >= 0 && b1 >= 0 && b2 >= 0;
  float n0 = (float)b0, n1 = (float)b1, n2 = (float)b2;
  n0 *= w->invarea;
  n1 *= w->invarea;
  n2 *= w->invarea;
  if(pixdepth)
    *pixdepth = n0 * w->depth[0] + n1 * w->depth[1] + n2 * w->depth[2];
  n0 *= w->invw[0];
  n1 *= w->invw[1];
  n2 *= w->invw[2];
  float invlen = 1.0f / (n0 + n1 + n2);
  out[0] = n0 * invlen;
  out[1] = n1 * invlen;
  out[2] = n2 * invlen;
  out[3] = 0.0f;
  return covered;
}

static inline uint8_t unorm8(float v)
{
  /* reference rasterizer.cpp:277-280,674-676: clamp01 then truncate */
  v = v > 1.0f ? 1.0f : (v < 0.0f ? 0.0f : v);
  return (uint8_t)(v * 255.0f);
}

static float blend_factor(uint32_t f, float src_alpha)
{
  switch(f)
  {
    case RR_BF_ZERO: return 0.0f;
    case RR_BF_ONE: return 1.0f;
    case RR_BF_SRC_ALPHA: return src_alpha;
    case RR_BF_ONE_MINUS_SRC_ALPHA: return 1.0f - src_alpha;
    default: return 1.0f; /* reference prints "Unsupported blend factor" and keeps 1.0 */
  }
}

/* reference rasterizer.cpp:522-696 (ProcessTriangles) for one block. */
static void process_block(const tri_work *w, rr_counters *ctr)
{
  const rr_state *st = w->state;
  const rr_pipeline *pl = st->pipeline;
  rr_image *col = st->col0;
  const uint32_t W = col->width, bpp = col->bpp;
  uint8_t *bits = col->pixels;
  float *depthbits = st->depth ? (float *)st->depth->pixels : NULL;
  uint64_t written = 0, tested = 0, passed_n = 0;

  for(int y = w->miny; y < w->maxy; y++)
  {
    for(int x = w->minx; x < w->maxx; x++)
    {
      float b3[3][4];
      float pixdepth;
      int covered = bary_at(w, x, y, b3[0], &pixdepth);
      if(covered)
      {
        int passed = 1;
        const uint64_t pix = (uint64_t)y * W + x;
        if(pl->depth_compare_op != RR_CMP_ALWAYS && depthbits)
        {
          float cur = depthbits[pix];
          switch(pl->depth_compare_op)
          {
            case RR_CMP_NEVER: passed = 0; break;
            case RR_CMP_LESS: passed = pixdepth < cur; break;
            case RR_CMP_EQUAL: passed = pixdepth == cur; break;
            case RR_CMP_LEQUAL: passed = pixdepth <= cur; break;
            case RR_CMP_GREATER: passed = pixdepth > cur; break;
            case RR_CMP_NOTEQUAL: passed = pixdepth != cur; break;
            case RR_CMP_GEQUAL: passed = pixdepth >= cur; break;
          }
        }
        if(passed)
        {
          /* helper evaluations at (x+1, y) and (x, y+1) feed implicit-LOD derivatives */
          bary_at(w, x + 1, y, b3[1], NULL);
          bary_at(w, x, y + 1, b3[2], NULL);
          float pix4[4];
          fs_program(st, pixdepth, (const float(*)[4])b3, w->vsout, pix4);
          if(bpp == 8)
          {
            /* RGBA16F target (beyond the reference) */
            uint16_t *h = (uint16_t *)(bits + pix * 8);
            if(pl->blend_enable)
            {
              float dst[4];
              for(int c = 0; c < 4; c++)
                dst[c] = rr_f16_to_f32(h[c]);
              float sf = blend_factor(pl->src_color_factor, pix4[3]);
              float df = blend_factor(pl->dst_color_factor, pix4[3]);
              for(int c = 0; c < 4; c++)
                pix4[c] = sf * pix4[c] + df * dst[c];
            }
            for(int c = 0; c < 4; c++)
              h[c] = rr_f32_to_f16(pix4[c]);
          }
          else
          {
            if(pl->blend_enable)
            {
              float ex0 = (float)bits[pix * bpp + 2], ex1 = (float)bits[pix * bpp + 1],
                    ex2 = (float)bits[pix * bpp + 0], ex3 = 1.0f;
              ex0 /= 255.0f;
              ex1 /= 255.0f;
              ex2 /= 255.0f;
              float sf = blend_factor(pl->src_color_factor, pix4[3]);
              float df = blend_factor(pl->dst_color_factor, pix4[3]);
              float o0 = sf * pix4[0] + df * ex0;
              float o1 = sf * pix4[1] + df * ex1;
              float o2 = sf * pix4[2] + df * ex2;
              float o3 = sf * pix4[3] + df * ex3;
              pix4[0] = o0; pix4[1] = o1; pix4[2] = o2; pix4[3] = o3;
            }
            bits[pix * bpp + 2] = unorm8(pix4[0]);
            bits[pix * bpp + 1] = unorm8(pix4[1]);
            bits[pix * bpp + 0] = unorm8(pix4[2]);
          }
          passed_n++;
          if(pl->depth_write && depthbits)
            depthbits[pix] = pixdepth;
        }
        written++;
      }
      tested++;
    }
  }
  if(ctr)
  {
    ctr->blocks++;
    ctr->pixels_tested += tested;
    ctr->pixels_written += written;
    ctr->depth_passed += passed_n;
  }
}

/* reference rasterizer.cpp:312-330 */
void rr_clear_depth(rr_image *t, float depth)
{
  uint64_t n = (uint64_t)t->width * t->height;
  for(uint64_t i = 0; i < n; i++)
    memcpy(t->pixels + i * 4, &depth, 4);
}

/* reference rasterizer.cpp:332-361: bytes [2]=R, [1]=G, [0]=B, [3]=A, truncated, no clamp. */
void rr_clear_color(rr_image *t, const float rgba[4])
{
  uint64_t n = (uint64_t)t->width * t->height;
  if(t->bpp == 8)
  {
    uint16_t h[4];
    for(int c = 0; c < 4; c++)
      h[c] = rr_f32_to_f16(rgba[c]);
    for(uint64_t i = 0; i < n; i++)
      memcpy(t->pixels + i * 8, h, 8);
    return;
  }
  uint8_t e[4];
  e[2] = (uint8_t)rr_f2i(rgba[0] * 255.0f);
  e[1] = (uint8_t)rr_f2i(rgba[1] * 255.0f);
  e[0] = (uint8_t)rr_f2i(rgba[2] * 255.0f);
  e[3] = (uint8_t)rr_f2i(rgba[3] * 255.0f);
  if(t->bpp == 1)
  {
    memset(t->pixels, e[2], n);
    return;
  }
  for(uint64_t i = 0; i < n; i++)
    memcpy(t->pixels + i * 4, e, 4);
}

typedef void (*emit_fn)(void *ctx, const tri_work *w);

/* reference rasterizer.cpp:363-486 (DrawTriangles up to the queue flush): VS, ToWindow, per
 * triangle area / facing / cull / bbox / setup, then the 32x32 block split. */
static void draw_setup(const rr_state *st, int num_verts, uint32_t first, int indexed,
                       rr_counters *ctr, emit_fn emit, void *ctx, rr_vertex **vs_keep,
                       int32_t **win_keep)
{
  const uint32_t w = st->col0->width, h = st->col0->height;
  rr_vertex *sv;
  size_t nv = shade_verts(st, num_verts, first, indexed, &sv);
  int32_t *win = (int32_t *)malloc((nv ? nv : 1) * 2 * sizeof(int32_t));
  for(size_t i = 0; i < nv; i++)
    rr_to_window(w, h, sv[i].position, &win[2 * i]);
  uint64_t tris_in = 0, tris_out = 0;
  for(size_t i = 0; i + 2 < nv; i += 3)
  {
    const int32_t *t = &win[2 * i];
    const rr_vertex *vo = &sv[i];
    tris_in++;
    int area2 = (int)(((uint32_t)t[2] - (uint32_t)t[0]) * ((uint32_t)t[5] - (uint32_t)t[1]) -
                      ((uint32_t)t[3] - (uint32_t)t[1]) * ((uint32_t)t[4] - (uint32_t)t[0]));
    if(area2 == 0)
      continue;
    int flipped = area2, barymul = 1;
    if(st->pipeline->front_face == RR_FRONT_CW)
    {
      barymul = -barymul;
      flipped = rr_wneg(flipped);
    }
    if(flipped > 0 && (st->pipeline->cull_mode & RR_CULL_FRONT))
      continue;
    if(flipped < 0)
    {
      if(st->pipeline->cull_mode & RR_CULL_BACK)
        continue;
      barymul = -barymul;
      flipped = rr_wneg(flipped);
    }
    tris_out++;
    int minx = t[0], miny = t[1], maxx = t[0], maxy = t[1];
    for(int k = 1; k < 3; k++)
    {
      if(t[2 * k] < minx) minx = t[2 * k];
      if(t[2 * k] > maxx) maxx = t[2 * k];
      if(t[2 * k + 1] < miny) miny = t[2 * k + 1];
      if(t[2 * k + 1] > maxy) maxy = t[2 * k + 1];
    }
    if(minx < 0) minx = 0;
    if(miny < 0) miny = 0;
    if(maxx > (int)w - 1) maxx = (int)w - 1;
    if(maxy > (int)h - 1) maxy = (int)h - 1;
    tri_work wk;
    wk.state = st;
    wk.ABx = rr_wsub(t[2], t[0]);
    wk.ABy = rr_wsub(t[3], t[1]);
    wk.ACx = rr_wsub(t[4], t[0]);
    wk.ACy = rr_wsub(t[5], t[1]);
    wk.vsout = vo;
    wk.tri = t;
    wk.barymul = barymul;
    wk.area2 = area2;
    wk.invarea = 1.0f / (float)flipped;
    wk.invw[0] = 1.0f / vo[0].position[3];
    wk.invw[1] = 1.0f / vo[1].position[3];
    wk.invw[2] = 1.0f / vo[2].position[3];
    wk.depth[0] = vo[0].position[2] * wk.invw[0];
    wk.depth[1] = vo[1].position[2] * wk.invw[1];
    wk.depth[2] = vo[2].position[2] * wk.invw[2];
    const int bs = 32;
    int xblocks = 1 + (maxx - minx) / bs;
    int yblocks = 1 + (maxy - miny) / bs;
    for(int bx = 0; bx < xblocks; bx++)
      for(int by = 0; by < yblocks; by++)
      {
        tri_work b = wk;
        b.minx = minx + bs * bx;
        b.miny = miny + bs * by;
        b.maxx = maxx < b.minx + bs ? maxx : b.minx + bs;
        b.maxy = maxy < b.miny + bs ? maxy : b.miny + bs;
        emit(ctx, &b);
      }
    emit(ctx, NULL); /* the triangle's blocks are all queued */
  }
  if(ctr)
  {
    ctr->tris_in += tris_in;
    ctr->tris_out += tris_out;
    ctr->draws++;
  }
  *vs_keep = sv;
  *win_keep = win;
}

typedef struct fifo {
  tri_work *items;
  size_t head, tail, cap;
} fifo;

static void fifo_push(fifo *q, const tri_work *w)
{
  if(q->tail == q->cap)
  {
    q->cap = q->cap ? q->cap * 2 : 1024;
    q->items = (tri_work *)realloc(q->items, q->cap * sizeof(tri_work));
  }
  q->items[q->tail++] = *w;
}

static void emit_serial(void *ctx, const tri_work *w)
{
  if(w)
    fifo_push((fifo *)ctx, w);
}

void rr_draw(const rr_state *st, int num_verts, uint32_t first, int indexed, rr_counters *ctr)
{
  atomic_fetch_add(&tc_generation, 1);
  fifo q = {0};
  rr_vertex *sv;
  int32_t *win;
  draw_setup(st, num_verts, first, indexed, ctr, emit_serial, &q, &sv, &win);
  /* No worker threads: the submitter drains the FIFO in push order (reference :488-515 with
   * WORK_STEAL and an empty pool), which is API primitive order. */
  for(size_t i = 0; i < q.tail; i++)
    process_block(&q.items[i], ctr);
  free(q.items);
  free(sv);
  free(win);
}

/* ---------------------------------------------------------------------------------------- */
/* Reference-structured threaded mode (timing only): reference rasterizer.cpp:36-98,454-515.  */

/* Counters: one slot per thread (workers 0..n-1, the submitter n), each on its own cache line,
 * summed by the submitter once the draw's blocks are done (the reference's counters are
 * compiled-out microprofile counters, so it takes no lock for them either). */
typedef struct ctr_slot {
  rr_counters c;
  char pad[64];
} ctr_slot;

typedef struct pool {
  pthread_t threads[64];
  int nthreads;
  pthread_mutex_t mutex;
  pthread_cond_t wake;
  int kill;
  fifo q;
  atomic_int pending;
  ctr_slot slots[65];
} pool;

static pool g_pool;
static int g_pool_started;

static void *raster_loop(void *arg)
{
  rr_counters *mine = &g_pool.slots[(int)(intptr_t)arg].c;
  tc_init();
  for(;;)
  {
    tri_work w;
    pthread_mutex_lock(&g_pool.mutex);
    while(!g_pool.kill && g_pool.q.head == g_pool.q.tail)
      pthread_cond_wait(&g_pool.wake, &g_pool.mutex);
    if(g_pool.kill)
    {
      pthread_mutex_unlock(&g_pool.mutex);
      return NULL;
    }
    w = g_pool.q.items[g_pool.q.head++];
    pthread_mutex_unlock(&g_pool.mutex);
    process_block(&w, mine);
    atomic_fetch_sub(&g_pool.pending, 1);
  }
}

static void pool_stop(void)
{
  if(!g_pool_started)
    return;
  pthread_mutex_lock(&g_pool.mutex);
  g_pool.kill = 1;
  pthread_mutex_unlock(&g_pool.mutex);
  pthread_cond_broadcast(&g_pool.wake);
  for(int i = 0; i < g_pool.nthreads; i++)
    pthread_join(g_pool.threads[i], NULL);
  free(g_pool.q.items);
  memset(&g_pool.q, 0, sizeof(g_pool.q));
  g_pool_started = 0;
}

static void pool_start(int n)
{
  if(g_pool_started && g_pool.nthreads == n)
    return;
  pool_stop();
  if(n > 64)
    n = 64;
  /* the reference's std::mutex is an MSVC SRW lock, which spins briefly before it sleeps;
   * glibc's adaptive mutex behaves alike */
  pthread_mutexattr_t ma;
  pthread_mutexattr_init(&ma);
  pthread_mutexattr_settype(&ma, PTHREAD_MUTEX_ADAPTIVE_NP);
  pthread_mutex_init(&g_pool.mutex, &ma);
  pthread_mutexattr_destroy(&ma);
  pthread_cond_init(&g_pool.wake, NULL);
  g_pool.kill = 0;
  g_pool.nthreads = n;
  atomic_store(&g_pool.pending, 0);
  memset(g_pool.slots, 0, sizeof(g_pool.slots));
  for(int i = 0; i < n; i++)
    pthread_create(&g_pool.threads[i], NULL, raster_loop, (void *)(intptr_t)i);
  g_pool_started = 1;
}

typedef struct thr_ctx {
  int dummy;
} thr_ctx;

/* reference rasterizer.cpp:466-486: each block is queued under the lock; the workers are woken
 * once per triangle, after its blocks */
static void emit_threaded(void *ctx, const tri_work *w)
{
  (void)ctx;
  if(!w)
  {
    pthread_cond_broadcast(&g_pool.wake);
    return;
  }
  pthread_mutex_lock(&g_pool.mutex);
  fifo_push(&g_pool.q, w);
  atomic_fetch_add(&g_pool.pending, 1);
  pthread_mutex_unlock(&g_pool.mutex);
}

void rr_draw_threaded(const rr_state *st, int num_verts, uint32_t first, int indexed,
                      int nthreads, rr_counters *ctr)
{
  atomic_fetch_add(&tc_generation, 1);
  pool_start(nthreads);
  rr_counters *mine = &g_pool.slots[g_pool.nthreads].c;
  rr_vertex *sv;
  int32_t *win;
  rr_counters setup_ctr = {0};
  thr_ctx tc;
  draw_setup(st, num_verts, first, indexed, &setup_ctr, emit_threaded, &tc, &sv, &win);
  while(atomic_load(&g_pool.pending) > 0)
  {
    tri_work w;
    int have = 0;
    pthread_mutex_lock(&g_pool.mutex);
    if(g_pool.q.head != g_pool.q.tail)
    {
      w = g_pool.q.items[g_pool.q.head++];
      have = 1;
    }
    pthread_mutex_unlock(&g_pool.mutex);
    if(have)
    {
      process_block(&w, mine);
      atomic_fetch_sub(&g_pool.pending, 1);
    }
  }
  pthread_mutex_lock(&g_pool.mutex);
  g_pool.q.head = g_pool.q.tail = 0;
  pthread_mutex_unlock(&g_pool.mutex);
  if(ctr)
  {
    ctr->draws += setup_ctr.draws;
    ctr->tris_in += setup_ctr.tris_in;
    ctr->tris_out += setup_ctr.tris_out;
  }
  for(int i = 0; i <= g_pool.nthreads; i++)
  {
    rr_counters *c = &g_pool.slots[i].c;
    if(ctr)
    {
      ctr->blocks += c->blocks;
      ctr->pixels_tested += c->pixels_tested;
      ctr->pixels_written += c->pixels_written;
      ctr->depth_passed += c->depth_passed;
    }
    memset(c, 0, sizeof(*c));
  }
  free(sv);
  free(win);
}

void rr_sample(const rr_image *img, const rr_sampler *s, const float coord[3], const float ddx[2],
               const float ddy[2], int use_lod, float out[4])
{
  atomic_fetch_add(&tc_generation, 1);
  sample_full(img, s, coord, ddx, ddy, use_lod, out);
}
