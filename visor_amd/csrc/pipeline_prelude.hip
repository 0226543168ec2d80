// pipeline_prelude.hip - device helpers for hiprtc-built pipeline modules.
//
// A pipeline module is: device_abi.h + this prelude + the translated shaders (one VS/FS pair per
// slot) + pipeline_kernels.hip. Everything here follows the reference's arithmetic exactly (op
// order, no contraction: modules are built with -ffp-contract=off and correctly rounded
// division and square root), so results match the oracle bit for bit.

typedef unsigned int u32;
typedef int i32;
typedef unsigned long long u64;
typedef unsigned short u16;
typedef unsigned char u8;

#define VA_INLINE __device__ __forceinline__

// Loads and stores of buffer, texture and target memory through global-address-space pointers.
// These pointers come out of the draw and frame tables as integers, so the compiler would use
// flat instructions, which may also address LDS and therefore also count on the LDS wait counter
// (every wait for a texel then waits for outstanding LDS operations too). Every such buffer is
// device memory (host-visible memory is mirrored to the device around a submission).
#ifdef __AMDGCN__
#define VA_GLOBAL __attribute__((address_space(1)))
#else
#define VA_GLOBAL
#endif
VA_INLINE u32 va_gld(const void *p) { return *(const VA_GLOBAL u32 *)p; }
VA_INLINE void va_gst(void *p, u32 v) { *(VA_GLOBAL u32 *)p = v; }

VA_INLINE float va_u2f(u32 u) { return __uint_as_float(u); }
VA_INLINE u32 va_f2u(float f) { return __float_as_uint(f); }
// float -> int32 as the reference's x64 build converts (cvttss2si): truncation inside
// [-2^31, 2^31), 0x80000000 for NaN and everything outside (gfx950's conversion saturates instead)
VA_INLINE i32 va_f2i(float f) { return (f >= -2147483648.0f && f < 2147483648.0f) ? (i32)f : (i32)0x80000000u; }
// float -> uint32 as x64 code converts it (through a 64-bit signed conversion, low 32 bits kept;
// 0 for NaN and everything outside [-2^63, 2^63))
VA_INLINE u32 va_f2u32(float f)
{
  return (f >= -9223372036854775808.0f && f < 9223372036854775808.0f) ? (u32)(long long)f : 0u;
}

// byte / 255.0f, correctly rounded, without a division: 1/255 split into hi + lo (hi = RN(1/255),
// lo = RN(1/255 - hi)), then one product and one FMA, b * hi + RN(b * lo), rounded once. Equal to
// the IEEE quotient for every b in 0..255 (checked exhaustively by the test suite), so texel and
// attribute values match the reference's float(byte) / 255.0f bit for bit.
VA_INLINE float va_unorm8(u32 b)
{
  const float hi = 0.0039215688593685627f;  // RN(1/255)
  const float lo = -0x1.fdfdfep-33f;        // RN(1/255 - hi)
  const float x = (float)b;
  return __builtin_fmaf(x, hi, x * lo);
}

// ---------------------------------------------------------------------------------------------
// Triple-point floats: value at the pixel and at its right and lower neighbours. Fragment
// programs that take implicit-LOD samples or derivatives run on these; the differences give
// deterministic per-pixel derivatives (independent of quad or batch placement).
struct df {
  float c, x, y;
  VA_INLINE df() {}
  VA_INLINE df(float v) : c(v), x(v), y(v) {}
  VA_INLINE df(float a, float b, float e) : c(a), x(b), y(e) {}
};
VA_INLINE df operator+(df a, df b) { return df(a.c + b.c, a.x + b.x, a.y + b.y); }
VA_INLINE df operator-(df a, df b) { return df(a.c - b.c, a.x - b.x, a.y - b.y); }
VA_INLINE df operator*(df a, df b) { return df(a.c * b.c, a.x * b.x, a.y * b.y); }
VA_INLINE df operator/(df a, df b) { return df(a.c / b.c, a.x / b.x, a.y / b.y); }
VA_INLINE df operator-(df a) { return df(-a.c, -a.x, -a.y); }
VA_INLINE float cv(float a) { return a; }
VA_INLINE float cv(df a) { return a.c; }
VA_INLINE float va_sqrt(float a) { return sqrtf(a); }
VA_INLINE df va_sqrt(df a) { return df(sqrtf(a.c), sqrtf(a.x), sqrtf(a.y)); }
VA_INLINE float va_floor(float a) { return floorf(a); }
VA_INLINE df va_floor(df a) { return df(floorf(a.c), floorf(a.x), floorf(a.y)); }
VA_INLINE float va_ceil(float a) { return ceilf(a); }
VA_INLINE df va_ceil(df a) { return df(ceilf(a.c), ceilf(a.x), ceilf(a.y)); }
VA_INLINE float va_trunc(float a) { return truncf(a); }
VA_INLINE df va_trunc(df a) { return df(truncf(a.c), truncf(a.x), truncf(a.y)); }
// round to nearest, halfway cases to even (GLSL roundEven; also used for round)
VA_INLINE float va_rint(float a) { return rintf(a); }
VA_INLINE df va_rint(df a) { return df(rintf(a.c), rintf(a.x), rintf(a.y)); }
VA_INLINE float va_fabs(float a) { return fabsf(a); }
VA_INLINE df va_fabs(df a) { return df(fabsf(a.c), fabsf(a.x), fabsf(a.y)); }
VA_INLINE float va_sin(float a) { return sinf(a); }
VA_INLINE df va_sin(df a) { return df(sinf(a.c), sinf(a.x), sinf(a.y)); }
VA_INLINE float va_cos(float a) { return cosf(a); }
VA_INLINE df va_cos(df a) { return df(cosf(a.c), cosf(a.x), cosf(a.y)); }
VA_INLINE float va_pow(float a, float b) { return powf(a, b); }
VA_INLINE df va_pow(df a, df b) { return df(powf(a.c, b.c), powf(a.x, b.x), powf(a.y, b.y)); }
VA_INLINE float va_exp2(float a) { return exp2f(a); }
VA_INLINE df va_exp2(df a) { return df(exp2f(a.c), exp2f(a.x), exp2f(a.y)); }
VA_INLINE float va_log2f(float a) { return log2f(a); }
VA_INLINE df va_log2f(df a) { return df(log2f(a.c), log2f(a.x), log2f(a.y)); }
// select on the pixel's own comparison; the neighbour values follow the same branch
VA_INLINE float va_sel(bool c, float a, float b) { return c ? a : b; }
VA_INLINE df va_sel(bool c, df a, df b) { return c ? a : b; }
VA_INLINE float va_dpdx(float a) { return 0.0f; }
VA_INLINE df va_dpdx(df a) { return df(a.x - a.c); }
VA_INLINE float va_dpdy(float a) { return 0.0f; }
VA_INLINE df va_dpdy(df a) { return df(a.y - a.c); }
VA_INLINE float va_bits_to_f(u32 u, float) { return va_u2f(u); }
VA_INLINE df va_bits_to_f(u32 u, df) { return df(va_u2f(u)); }

// ---------------------------------------------------------------------------------------------
// Resource access (reference spirv_compile.cpp:552-627)

// Reads outside the 128-byte push block or the bound uniform range return 0 (never fault).
VA_INLINE u32 va_pc_u(const VaDraw &d, u32 off) { return off <= 124u ? *(const u32 *)(d.push + off) : 0u; }
VA_INLINE float va_pc_f(const VaDraw &d, u32 off) { return va_u2f(va_pc_u(d, off)); }
VA_INLINE u32 va_ubo_u(const VaDraw &d, u32 slot, u64 off)
{
  return off + 4ull <= d.ubo_bytes[slot] ? va_gld((const void *)(d.ubo[slot] + off)) : 0u;
}
VA_INLINE float va_ubo_f(const VaDraw &d, u32 slot, u64 off) { return va_u2f(va_ubo_u(d, slot, off)); }

// 16 zero bytes: the load address of an index or a vertex attribute outside its bound buffer.
__device__ u32 va_zero_words[4] __attribute__((aligned(16))) = {0u, 0u, 0u, 0u};  // never written

// GetVertexAttributeData: default (0,0,0,1); 32-bit formats copy raw bits with fall-through,
// R8G8B8A8_UNORM divides each byte by 255.
// In two steps, so that the loads of every attribute of a primitive's three vertices are issued
// together and waited for once: va_vload has no branch and issues one 16-byte load per attribute
// (an attribute past the end of its bound buffer loads va_zero_words instead), and va_vdecode
// turns the raw words into the attribute's four values, dropping the words beyond the format's.
// Device memory is allocated 16 bytes longer than asked, so the 16 bytes from an attribute that
// lies inside its buffer are always readable.
struct VaRaw
{
  u32 w[4];
  bool ok;  // the attribute lies inside its bound buffer
};

// four words at 4-byte alignment (vertex attributes are aligned to their 4-byte components)
struct __attribute__((packed, aligned(4))) VaWords4
{
  u32 x, y, z, w;
};

// words of a format: R32 .. R32G32B32A32 in UINT, SINT, SFLOAT (VK_FORMAT 98..109, three formats
// per width) take 1..4, R8G8B8A8_UNORM (37) and every other format 1
VA_INLINE u32 va_vfmt_words(u32 fmt) { return fmt >= 98u && fmt <= 109u ? (fmt - 98u) / 3u + 1u : 1u; }

VA_INLINE VaRaw va_vload(const VaDraw &d, u32 vtx, u32 loc)
{
  const u32 vb = d.vattr_vb[loc];
  const u64 off = d.vattr_off[loc] + (u64)d.vattr_stride[loc] * vtx;
  const u32 nw = va_vfmt_words(d.vattr_fmt[loc]);
  VaRaw r;
  r.ok = (d.vb[vb] != 0ull) & (off + 4ull * nw <= d.vb_bytes[vb]);
  const VA_GLOBAL u8 *base = (const VA_GLOBAL u8 *)d.vb[vb];
  const VA_GLOBAL VaWords4 *p = r.ok ? (const VA_GLOBAL VaWords4 *)(base + off) : (const VA_GLOBAL VaWords4 *)va_zero_words;
  const VaWords4 x = *p;
  r.w[0] = x.x;
  r.w[1] = x.y;
  r.w[2] = x.z;
  r.w[3] = x.w;
  return r;
}

VA_INLINE void va_vdecode(const VaDraw &d, u32 loc, const VaRaw &r, float out[4])
{
  const u32 fmt = d.vattr_fmt[loc];
  float o[4] = {0.0f, 0.0f, 0.0f, 1.0f};
  if(fmt >= 98u && fmt <= 109u)
  {
    const u32 nw = va_vfmt_words(fmt);
#pragma unroll
    for(u32 k = 0; k < 4u; k++)
      if(k < nw)
        o[k] = va_u2f(r.w[k]);
  }
  else if(fmt == 37u)
  {
#pragma unroll
    for(u32 k = 0; k < 4u; k++)
      o[k] = va_unorm8((r.w[0] >> (8u * k)) & 0xffu);
  }
  out[0] = r.ok ? o[0] : 0.0f;
  out[1] = r.ok ? o[1] : 0.0f;
  out[2] = r.ok ? o[2] : 0.0f;
  out[3] = r.ok ? o[3] : 1.0f;
}

// ---------------------------------------------------------------------------------------------
// Texture sampling. REPEAT + LINEAR at one level is the reference's sample_tex_wrapped
// (texture_sampling.cpp:139-184): u - floor(u), scale by the extent, truncate, no texel-centre
// offset, bilinear in the reference's order. Levels, clamp, nearest and layers extend it.

VA_INLINE void va_texel(const u8 *base, u32 W, i32 x, i32 y, float out[4])
{
  u32 u = va_gld(base + ((u64)(u32)y * W + (u32)x) * 4u);
  out[0] = va_unorm8(u & 0xffu);
  out[1] = va_unorm8((u >> 8) & 0xffu);
  out[2] = va_unorm8((u >> 16) & 0xffu);
  out[3] = va_unorm8(u >> 24);
}

VA_INLINE void va_sample_level(const VaTex &t, u32 level, u32 layer, float u, float v, u32 filter,
                               float out[4])
{
  u32 W = t.width >> level; if(W < 1u) W = 1u;
  u32 H = t.height >> level; if(H < 1u) H = 1u;
  const u8 *base = (const u8 *)t.base + t.slice_bytes * layer + t.level_off[level];
  if(!(u == u)) u = 0.0f;
  if(!(v == v)) v = 0.0f;
  if(t.addr_u == 0u) u = u - floorf(u);
  else { if(!(u >= 0.0f)) u = 0.0f; if(u > 1.0f) u = 1.0f; }
  if(t.addr_v == 0u) v = v - floorf(v);
  else { if(!(v >= 0.0f)) v = 0.0f; if(v > 1.0f) v = 1.0f; }
  float U = u * (float)W, V = v * (float)H;
  i32 i0 = (i32)U, j0 = (i32)V;
  if(filter == 0u)
  {
    if(i0 >= (i32)W) i0 = t.addr_u == 0u ? i0 - (i32)W : (i32)W - 1;
    if(j0 >= (i32)H) j0 = t.addr_v == 0u ? j0 - (i32)H : (i32)H - 1;
    va_texel(base, W, i0, j0, out);
    return;
  }
  i32 i1 = i0 + 1, j1 = j0 + 1;
  float fu = U - (float)i0, fv = V - (float)j0;
  if(t.addr_u == 0u) { if(i1 >= (i32)W) i1 -= (i32)W; if(i0 >= (i32)W) i0 -= (i32)W; }
  else { if(i1 > (i32)W - 1) i1 = (i32)W - 1; if(i0 > (i32)W - 1) i0 = (i32)W - 1; }
  if(t.addr_v == 0u) { if(j1 >= (i32)H) j1 -= (i32)H; if(j0 >= (i32)H) j0 -= (i32)H; }
  else { if(j1 > (i32)H - 1) j1 = (i32)H - 1; if(j0 > (i32)H - 1) j0 = (i32)H - 1; }
  float inv_fu = 1.0f - fu, inv_fv = 1.0f - fv;
  float TL[4], TR[4], BL[4], BR[4];
  va_texel(base, W, i0, j0, TL);
  va_texel(base, W, i1, j0, TR);
  va_texel(base, W, i0, j1, BL);
  va_texel(base, W, i1, j1, BR);
#pragma unroll
  for(int c = 0; c < 4; c++)
  {
    float top = TL[c] * inv_fu + TR[c] * fu;
    float bottom = BL[c] * inv_fu + BR[c] * fu;
    out[c] = top * inv_fv + bottom * fv;
  }
}

// Deterministic log2 for LOD: exponent + atanh series on m in [sqrt(1/2), sqrt(2)).
VA_INLINE float va_log2(float x)
{
  if(!(x > 0.0f)) return -128.0f;
  if(x == __builtin_huge_valf()) return 128.0f;
  u32 b = va_f2u(x);
  i32 e = (i32)((b >> 23) & 0xffu);
  if(e == 0) return -128.0f;
  e -= 127;
  float m = va_u2f((b & 0x7fffffu) | 0x3f800000u);
  if(m > 1.41421353816986083984375f) { m = m * 0.5f; e += 1; }
  float s = (m - 1.0f) / (m + 1.0f);
  float s2 = s * s;
  float p = 0.111111111f;
  p = p * s2 + 0.142857143f;
  p = p * s2 + 0.2f;
  p = p * s2 + 0.333333343f;
  p = p * s2 + 1.0f;
  p = p * s;
  return (float)e + p * 2.88539004f;
}

VA_INLINE void va_sample(const VaTex &t, float u, float v, float layerf, float dudx, float dvdx,
                         float dudy, float dvdy, int use_lod, float lod, float out[4])
{
  float lambda = 0.0f;
  if(use_lod == 1)
  {
    float W0 = (float)t.width, H0 = (float)t.height;
    float ax = dudx * W0, ay = dvdx * H0;
    float bx = dudy * W0, by = dvdy * H0;
    float rx = ax * ax + ay * ay;
    float ry = bx * bx + by * by;
    float r = rx > ry ? rx : ry;
    lambda = 0.5f * va_log2(r);
  }
  else if(use_lod == 2)
    lambda = lod;
  lambda = lambda + t.lod_bias;
  if(!(lambda >= t.min_lod)) lambda = t.min_lod;
  if(lambda > t.max_lod) lambda = t.max_lod;
  u32 filter = lambda > 0.0f ? t.min_filter : t.mag_filter;
  float q = (float)(t.mips - 1u);
  float dd = lambda;
  if(!(dd >= 0.0f)) dd = 0.0f;
  if(dd > q) dd = q;
  i32 layer = (i32)floorf(layerf + 0.5f);
  if(layer < 0) layer = 0;
  if(layer > (i32)t.layers - 1) layer = (i32)t.layers - 1;
  if(t.mipmap_mode == 0u)
  {
    i32 level = dd <= 0.5f ? 0 : (i32)ceilf(dd + 0.5f) - 1;
    if(level > (i32)q) level = (i32)q;
    va_sample_level(t, (u32)level, (u32)layer, u, v, filter, out);
    return;
  }
  i32 dhi = (i32)floorf(dd);
  i32 dlo = dhi + 1 > (i32)q ? (i32)q : dhi + 1;
  float delta = dd - (float)dhi;
  va_sample_level(t, (u32)dhi, (u32)layer, u, v, filter, out);
  if(delta != 0.0f)
  {
    float b[4];
    va_sample_level(t, (u32)dlo, (u32)layer, u, v, filter, b);
    float inv = 1.0f - delta;
#pragma unroll
    for(int c = 0; c < 4; c++)
      out[c] = out[c] * inv + b[c] * delta;
  }
}

// Shader-facing wrappers: plain floats sample with LOD 0 (vertex stage / no derivatives),
// triple-point floats take their derivatives from the neighbour values.
VA_INLINE void va_tex(const VaDraw &d, u32 slot, float u, float v, float layer, float out[4])
{
  va_sample(d.tex[slot], u, v, layer, 0.0f, 0.0f, 0.0f, 0.0f, 0, 0.0f, out);
}
VA_INLINE void va_tex(const VaDraw &d, u32 slot, df u, df v, df layer, float out[4])
{
  va_sample(d.tex[slot], u.c, v.c, layer.c, u.x - u.c, v.x - v.c, u.y - u.c, v.y - v.c, 1, 0.0f, out);
}
VA_INLINE void va_tex_lod(const VaDraw &d, u32 slot, float u, float v, float layer, float lod,
                          float out[4])
{
  va_sample(d.tex[slot], u, v, layer, 0.0f, 0.0f, 0.0f, 0.0f, 2, lod, out);
}
VA_INLINE void va_tex_lod(const VaDraw &d, u32 slot, df u, df v, df layer, df lod, float out[4])
{
  va_sample(d.tex[slot], u.c, v.c, layer.c, 0.0f, 0.0f, 0.0f, 0.0f, 2, lod.c, out);
}

// Cube maps (reference texture_sampling.cpp:186-250): the face tests run in the order +X, -X,
// +Y, -Y, +Z, -Z, each overriding the previous ones (ties go to Z, then Y); the face coordinate
// is 0.5 * (sc / |ma| + 1). The face is then sampled as array layer `face`, without derivative
// LOD. A NaN direction, where no test holds, samples face 0 at (0.5, 0.5).
VA_INLINE void va_cube_face(float x, float y, float z, float &u, float &v, float &face)
{
  const float ax = fabsf(x), ay = fabsf(y), az = fabsf(z);
  const bool px = x > 0.0f, py = y > 0.0f, pz = z > 0.0f;
  float axis = 1.0f, sc = 0.0f, tc = 0.0f, f = 0.0f;
  if(px && ax >= ay && ax >= az) { axis = ax; sc = -z; tc = -y; f = 0.0f; }
  if(!px && ax >= ay && ax >= az) { axis = ax; sc = z; tc = -y; f = 1.0f; }
  if(py && ay >= ax && ay >= az) { axis = ay; sc = x; tc = z; f = 2.0f; }
  if(!py && ay >= ax && ay >= az) { axis = ay; sc = x; tc = -z; f = 3.0f; }
  if(pz && az >= ax && az >= ay) { axis = az; sc = x; tc = -y; f = 4.0f; }
  if(!pz && az >= ax && az >= ay) { axis = az; sc = -x; tc = -y; f = 5.0f; }
  u = 0.5f * (sc / axis + 1.0f);
  v = 0.5f * (tc / axis + 1.0f);
  face = f;
}
VA_INLINE void va_tex_cube(const VaDraw &d, u32 slot, float x, float y, float z, int use_lod, float lod,
                           float out[4])
{
  float u, v, f;
  va_cube_face(x, y, z, u, v, f);
  va_sample(d.tex[slot], u, v, f, 0.0f, 0.0f, 0.0f, 0.0f, use_lod, lod, out);
}
VA_INLINE void va_tex_cube(const VaDraw &d, u32 slot, df x, df y, df z, int use_lod, df lod, float out[4])
{
  va_tex_cube(d, slot, x.c, y.c, z.c, use_lod, lod.c, out);
}

// Interpolated fragment input (reference spirv_compile.cpp:2159-2240 + CreateDot :629-643):
// (((b0*a0) + b1*a1) + b2*a2) + b3*0 with b3 = 0.
VA_INLINE float va_interp(const float *b, const float *a)
{
  float acc = b[0] * a[0];
  acc = acc + b[1] * a[1];
  acc = acc + b[2] * a[2];
  acc = acc + b[3] * 0.0f;
  return acc;
}
