/*
 * device_abi.h - data layouts shared by the host runtime and the pipeline kernels (which are
 * built with hiprtc, where <stdint.h> is not available: plain C integer types only).
 *
 * HBM layout of one submission ("batch"):
 *   VaFrame[nframes]   render-target sets (colour + depth), tile grid, fused clear values
 *   VaDraw[ndraws]     one per vkCmdDraw / vkCmdDrawIndexed, in API order, with its state snapshot
 *   records            one triangle record per primitive, global index = draw.tri_base + prim
 *   tile counts / scan / (tile, triangle) pairs / sorted bins / per-tile ranges
 */
#ifndef VISOR_DEVICE_ABI_H
#define VISOR_DEVICE_ABI_H

#define VA_MAX_VATTRS 16
#define VA_MAX_TEX 4
#define VA_MAX_UBO 4
#define VA_MAX_LEVELS 16
#define VA_REC_HDR 17   /* header words of a triangle record, varyings follow */
#define VA_REC_VIS 14   /* leading header words the raster's visibility pass reads */
#define VA_SETUP_PRIMS 256u /* primitives (threads) per setup workgroup (512: no faster) */
/* Counter rows per stage: workgroup b of setup (tile t of the raster) adds its four counters
 * into row b % VA_COUNTER_ROWS (t % VA_COUNTER_ROWS) of its stage's rows, 4 u64 each; the finish
 * kernel sums the rows and leaves them zero. */
#define VA_COUNTER_ROWS 64u

/* record header word indices */
#define VA_R_X0 0
#define VA_R_Y0 1
#define VA_R_ABX 2
#define VA_R_ABY 3
#define VA_R_ACX 4
#define VA_R_ACY 5
#define VA_R_AREA 6
#define VA_R_INVAREA 7
#define VA_R_Z0 8
#define VA_R_BBMIN 11
#define VA_R_BBMAX 12
#define VA_R_DRAW 13
#define VA_R_IW0 14     /* 1/w of the vertices: only perspective-correct shading reads them */

/* VaDraw.flags */
#define VA_DRAW_INDEXED 1u
#define VA_DRAW_INDEX32 2u
#define VA_DRAW_STRIP 4u

/* VaFrame.clear_flags */
#define VA_CLEAR_COLOR 1u
#define VA_CLEAR_DEPTH 2u

typedef struct VaTex {
  unsigned long long base;          /* device address of mip 0, layer 0 */
  unsigned long long slice_bytes;   /* bytes per array layer (all mips) */
  unsigned int width, height, mips, layers;
  unsigned int mag_filter, min_filter, mipmap_mode, addr_u;
  unsigned int addr_v, bpp, pad0, pad1;  /* texels are 4-byte RGBA (bpp 4) */
  float lod_bias, min_lod, max_lod, pad2;
  unsigned int level_off[VA_MAX_LEVELS]; /* byte offset of each mip inside a layer */
} VaTex;

typedef struct VaDraw {
  unsigned int tri_base;   /* global index of the first triangle of this draw */
  unsigned int ntris;
  unsigned int first;      /* firstVertex or firstIndex */
  unsigned int flags;
  unsigned long long ib;   /* index buffer device address + bound offset */
  unsigned long long ib_bytes;  /* bytes readable from ib (out-of-range indices read as 0) */
  unsigned long long vb[4];
  unsigned long long vb_bytes[4];  /* bytes readable from vb[i] (out-of-range vertices read defaults) */
  unsigned int vattr_fmt[VA_MAX_VATTRS];
  unsigned int vattr_stride[VA_MAX_VATTRS];
  unsigned int vattr_off[VA_MAX_VATTRS];
  unsigned int vattr_vb[VA_MAX_VATTRS];
  unsigned int cull_mode, front_face, depth_op, depth_write;
  unsigned int blend_enable, src_factor, dst_factor, blend_op;
  unsigned int frame;      /* index into the VaFrame table */
  unsigned int slot;       /* shader pair slot inside the pipeline module */
  unsigned int block_base; /* first setup workgroup of this draw (VA_SETUP_PRIMS primitives each) */
  unsigned int pad;
  unsigned char push[128];
  unsigned long long ubo[VA_MAX_UBO];
  unsigned long long ubo_bytes[VA_MAX_UBO];
  VaTex tex[VA_MAX_TEX];
} VaDraw;

typedef struct VaFrame {
  unsigned long long color;  /* device address of the colour attachment */
  unsigned long long depth;  /* device address of the D32 attachment, 0 if none */
  unsigned int width, height;
  unsigned int color_bpp;    /* 4 (8-bit UNORM, B,G,R byte order) or 8 (RGBA16F) */
  unsigned int clear_flags;  /* fused clear-on-first-touch (batched path) */
  float clear_color[4];
  float clear_depth;
  unsigned int tiles_x, tiles_y;
  unsigned int tile_base;    /* first global tile id of this frame */
} VaFrame;

typedef struct VaCounters {
  unsigned long long draws;
  unsigned long long tris_in, tris_out, blocks, pixels_tested;  /* setup: per-workgroup partials */
  unsigned long long pixels_written, depth_passed;                /* raster: per-tile partials     */
  unsigned long long shaded_blend;   /* shaded fragments of draws with blending on */
  unsigned long long shaded_dwrite;  /* shaded fragments that wrote depth */
  unsigned long long pad[3];
} VaCounters;

#endif
