// exec.cpp - command buffer execution on the GPU.
//
// Decodes the recorded stream (reference cmd_exec.cpp:15-185) into render-pass and copy
// operations with a per-draw state snapshot, then runs the draw-execution path:
//   plain    per draw: setup, bin, raster, then wait for the GPU (as the reference's blocking
//            DrawTriangles, rasterizer.cpp:363-520); BeginRenderPass clears with fill kernels.
//   batched  every draw of a render pass in one setup / bin / raster sequence with the clear
//            fused into the raster kernel's first touch of each tile.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <set>

#include "internal.h"
#include "runtime_kernels.h"
#include "texel_decode.h"

namespace visor
{
#define HIPCHK(x)                                                                     \
  do                                                                                  \
  {                                                                                   \
    hipError_t e_ = (x);                                                              \
    if(e_ != hipSuccess)                                                              \
    {                                                                                 \
      LogError("%s failed: %s (%s:%d)", #x, hipGetErrorString(e_), __FILE__, __LINE__); \
      return HipToVk(e_);                                                             \
    }                                                                                 \
  } while(0)

static uint32_t TileShiftFromEnv()
{
  const char *t = getenv("VISOR_TILE");
  if(t)
  {
    int v = atoi(t);
    if(v == 8) return 3;
    if(v == 16) return 4;
    if(v == 32) return 5;
  }
  return 0;
}

Queue::Queue()
{
  const char *m = getenv("VISOR_EXEC_MODE");
  if(m && !strcmp(m, "plain"))
    mode = ExecMode::Plain;
  // VISOR_TILE=8/16/32 fixes the tile size; otherwise the autotuner chooses (VISOR_AUTOTUNE=0:
  // fixed 8x8 tiles)
  if(uint32_t sh = TileShiftFromEnv())
  {
    tiles.tile_shift = sh;
    autotune = false;
  }
  if(const char *pp = getenv("VISOR_PERSISTENT"))
    tiles.persistent = !strcmp(pp, "1") ? 1u : 0u;
  if(const char *at = getenv("VISOR_AUTOTUNE"))
    if(!strcmp(at, "0"))
      autotune = false;
  // VISOR_CHUNK / VISOR_SMALL: raster module parameters (see TileConfig)
  if(const char *c = getenv("VISOR_CHUNK"))
  {
    int v = atoi(c);
    if(v >= 16 && v <= 64)
    {
      tiles.chunk = (uint32_t)v;
      autotune = false;
    }
  }
  if(const char *c = getenv("VISOR_SMALL"))
  {
    int v = atoi(c);
    if(v >= 0 && v <= 64)
      tiles.small = (uint32_t)v;
  }
  const char *a = getenv("VISOR_SUBMIT");
  async = a && !strcmp(a, "async");
  const char *g = getenv("VISOR_GRAPHS");
  graphs = !(g && !strcmp(g, "0"));
}

Queue::~Queue()
{
  if(worker.joinable())
  {
    {
      std::lock_guard<std::mutex> lk(qm);
      stop = true;
    }
    qcv.notify_all();
    worker.join();
  }
  void *bufs[] = {d_frames, d_draws, d_rec,  d_count, d_bpairs, d_keys,  d_vals,   d_keys2,
                  d_vals2,  d_start, d_blocks, d_part, d_stemp, d_einfo, d_stamps, d_masks};
  for(void *b : bufs)
    if(b)
      hipFree(b);
  if(h_pinned)
    hipHostFree(h_pinned);
  for(auto &g : graph_cache)
    for(hipGraphExec_t x : g.exec)
      if(x)
        hipGraphExecDestroy(x);

  if(stream)
    hipStreamDestroy(stream);
}

// Bytes of the raster's pass masks: per tile and chunk of its bin, one mask per pixel (4 bytes for
// chunks of at most 32 records, else 8). Tile t's masks start at chunk slot floor(start[t] /
// chunk) + t, which leaves every tile room for its ceil(len / chunk) chunks.
static size_t RasterMaskBytes(const TileConfig &cfg, uint32_t pair_cap, uint32_t tiles)
{
  const size_t px = (size_t)64 << (2 * (cfg.tile_shift - 3));
  const size_t elem = cfg.chunk <= 32 ? 4 : 8;
  return elem * px * ((size_t)pair_cap / cfg.chunk + tiles + 1);
}

bool Queue::Grow(void **p, size_t &cap, size_t need)
{
  if(need <= cap && *p)
    return true;
  size_t ncap = std::max(need + need / 2, (size_t)4096);
  if(*p)
  {
    hipStreamSynchronize(stream);
    hipFree(*p);
    *p = nullptr;
  }
  if(hipMalloc(p, ncap) != hipSuccess)
  {
    cap = 0;
    *p = nullptr;
    return false;
  }
  cap = ncap;
  return true;
}

static uint64_t ImageBytes(const VkImage_T *img)
{
  uint64_t slice = 0;
  uint32_t w = img->extent.width, h = img->extent.height;
  for(uint32_t m = 0; m < img->mipLevels; m++)
  {
    slice += (uint64_t)w * h * img->bytesPerPixel;
    w = std::max(1u, w >> 1);
    h = std::max(1u, h >> 1);
  }
  return slice * img->arrayLayers;
}

// reference precompiled.cpp:3-36 (CalcSubresourceByteOffset)
static uint64_t SubresourceOffset(const VkImage_T *img, uint32_t mip, uint32_t layer)
{
  uint64_t offs = 0;
  const uint32_t w = img->extent.width, h = img->extent.height, bpp = img->bytesPerPixel;
  for(uint32_t m = 0; m < mip; m++)
    offs += (uint64_t)std::max(1u, w >> m) * std::max(1u, h >> m) * bpp;
  if(layer > 0)
  {
    uint64_t slice = 0;
    uint32_t mw = w, mh = h;
    for(uint32_t m = 0; m < img->mipLevels; m++)
    {
      slice += (uint64_t)mw * mh * bpp;
      mw = std::max(1u, mw >> 1);
      mh = std::max(1u, mh >> 1);
    }
    offs += slice * layer;
  }
  return offs;
}

// Images whose texels are not 4-byte RGBA are sampled through an RGBA8 shadow holding, per
// texel, the 4 bytes the reference's fetch returns (BC2/BC3 decoded, R8 and RGBA16F as the
// reference reads them): the pipelines' samplers then read one aligned word per texel.
static bool NeedsShadow(const VkImage_T *img)
{
  return img->bytesPerPixel != 4 || img->format == VK_FORMAT_BC2_UNORM_BLOCK ||
         img->format == VK_FORMAT_BC3_UNORM_BLOCK;
}

static uint32_t ShadowLevels(const VkImage_T *img) { return std::max(1u, std::min<uint32_t>(img->mipLevels, VA_MAX_LEVELS)); }

// Layout of a linear image: levels packed per layer, layers back to back (precompiled.cpp:3-36).
static void LevelLayout(const VkImage_T *img, uint32_t bpp, uint32_t levels, uint64_t *level_off, uint64_t &slice)
{
  uint64_t off = 0;
  uint32_t w = img->extent.width, h = img->extent.height;
  for(uint32_t m = 0; m < levels; m++)
  {
    if(level_off && m < VA_MAX_LEVELS)
      level_off[m] = off;
    off += (uint64_t)w * h * bpp;
    w = std::max(1u, w >> 1);
    h = std::max(1u, h >> 1);
  }
  slice = off;
}

static bool EnsureShadow(VkImage_T *img, std::string &err)
{
  if(img->shadow)
    return true;
  uint64_t slice = 0;
  LevelLayout(img, 4, ShadowLevels(img), nullptr, slice);
  const uint64_t bytes = slice * std::max(1u, img->arrayLayers);
  if(bytes == 0 || hipMalloc((void **)&img->shadow, bytes) != hipSuccess)
  {
    img->shadow = nullptr;
    err = "out of device memory for a sampled image's RGBA8 copy";
    return false;
  }
  return true;
}

static hipError_t TranscodeAsync(const VkImage_T *img, hipStream_t stream)
{
  VaTranscode t;
  memset(&t, 0, sizeof(t));
  t.width = img->extent.width;
  t.height = img->extent.height;
  t.mips = ShadowLevels(img);
  t.layers = std::max(1u, img->arrayLayers);
  t.bpp = img->bytesPerPixel;
  t.fmt = img->format == VK_FORMAT_BC2_UNORM_BLOCK ? VA_TEXFMT_BC2
          : img->format == VK_FORMAT_BC3_UNORM_BLOCK ? VA_TEXFMT_BC3 : VA_TEXFMT_PLAIN;
  LevelLayout(img, img->bytesPerPixel, img->mipLevels, t.src_level_off, t.src_slice);
  LevelLayout(img, 4, t.mips, t.dst_level_off, t.dst_slice);
  t.src_bytes = t.src_slice * t.layers;
  return va_transcode_async(img->devPixels(), img->shadow, &t, stream);
}

static void FillTex(VaTex &t, const VkImage_T *img, const VkSampler_T *s)
{
  memset(&t, 0, sizeof(t));
  const bool sh = NeedsShadow(img);
  t.base = (unsigned long long)(uintptr_t)(sh ? img->shadow : img->devPixels());
  t.width = img->extent.width;
  t.height = img->extent.height;
  t.mips = ShadowLevels(img);
  t.layers = std::max(1u, img->arrayLayers);
  t.bpp = 4;
  uint64_t level_off[VA_MAX_LEVELS], slice = 0;
  LevelLayout(img, 4, sh ? t.mips : img->mipLevels, level_off, slice);
  for(uint32_t m = 0; m < t.mips; m++)
    t.level_off[m] = (unsigned)level_off[m];
  t.slice_bytes = slice;
  if(s)
  {
    const VkSamplerCreateInfo &i = s->info;
    t.mag_filter = i.magFilter;
    t.min_filter = i.minFilter;
    t.mipmap_mode = i.mipmapMode;
    t.addr_u = i.addressModeU == VK_SAMPLER_ADDRESS_MODE_REPEAT ? 0u : 2u;
    t.addr_v = i.addressModeV == VK_SAMPLER_ADDRESS_MODE_REPEAT ? 0u : 2u;
    t.lod_bias = i.mipLodBias;
    t.min_lod = i.minLod;
    t.max_lod = i.maxLod;
  }
  else
  {
    // the reference's only sampling mode: bilinear, wrap, level 0 (texture_sampling.cpp:139-184)
    t.mag_filter = t.min_filter = 1;
    t.mipmap_mode = 0;
    t.addr_u = t.addr_v = 0;
    t.min_lod = t.max_lod = 0.0f;
  }
}

// bytes of a buffer readable from offset, limited by both the buffer and its memory
static unsigned long long BoundBytes(const VkBuffer_T *b, VkDeviceSize off)
{
  if(!b || !b->mem)
    return 0;
  uint64_t end = std::min<uint64_t>(b->size, b->mem->size > b->offset ? b->mem->size - b->offset : 0);
  return off < end ? end - off : 0;
}

static void AddMem(std::vector<VkDeviceMemory_T *> &v, VkDeviceMemory_T *m)
{
  if(m && m->type == 1 && std::find(v.begin(), v.end(), m) == v.end())
    v.push_back(m);
}

std::atomic<uint64_t> g_descriptor_writes{0};

bool Queue::Decode(const VkCommandBuffer_T *cb, Submission &sub, std::string &err)
{
  const uint8_t *cur = cb->stream.data(), *end = cur + cb->stream.size();
  // bound state (reference gpu.h:3-23), reset per command buffer like cmd_exec.cpp:20
  VkPipeline_T *pipeline = nullptr;
  VkDescriptorSet_T *sets[8] = {};
  uint8_t push[128] = {};
  VkBuffer_T *ib = nullptr;
  VkDeviceSize ib_off = 0;
  VkIndexType ib_type = VK_INDEX_TYPE_UINT16;
  VkBuffer_T *vbs[4] = {};
  VkDeviceSize vb_off[4] = {};
  int pass = -1;
  std::set<const VkDeviceMemory_T *> touched;  // memories written earlier in this command buffer
  sub.cb = cb;
  sub.recording = cb->recording;
  const uint64_t desc_writes = g_descriptor_writes.load();
  sub.desc_writes = desc_writes;
  auto pull = [&](size_t n) {
    const uint8_t *p = cur;
    cur += n;
    return p;
  };
  while(cur < end)
  {
    Cmd id;
    memcpy(&id, pull(2), 2);
    switch(id)
    {
      case Cmd::PipelineBarrier: case Cmd::SetScissors: break;
      case Cmd::EndRenderPass: pass = -1; break;
      case Cmd::BeginRenderPass:
      {
        const cmd::BeginRenderPass *d = (const cmd::BeginRenderPass *)pull(sizeof(cmd::BeginRenderPass));
        VkRenderPass_T *rp = d->renderPass;
        if(!rp || rp->subpasses.empty() || rp->subpasses[0].colAttachments.empty())
        {
          err = "render pass without a colour attachment in subpass 0";
          return false;
        }
        Op op;
        op.kind = Op::Pass;
        // reference cmd_exec.cpp:35-64: subpass 0, colour attachment 0, depth, clears in order
        const auto &sp = rp->subpasses[0];
        const auto &c0 = sp.colAttachments[0];
        if(!d->framebuffer || c0.idx < 0 || (size_t)c0.idx >= d->framebuffer->attachments.size())
        {
          err = "framebuffer has no colour attachment at the render pass's index";
          return false;
        }
        op.frame.color = d->framebuffer->attachments[c0.idx]->image;
        if(!op.frame.color->mem)
        {
          err = "colour attachment has no memory bound";
          return false;
        }
        int ci = 0;
        if(c0.clear)
        {
          op.frame.clear_color = true;
          op.frame.clear_col = d->clearval[ci++];
        }
        if(sp.depthAttachment.idx >= 0)
        {
          if((size_t)sp.depthAttachment.idx >= d->framebuffer->attachments.size() ||
             !d->framebuffer->attachments[sp.depthAttachment.idx]->image->mem)
          {
            err = "depth attachment missing or without memory";
            return false;
          }
          op.frame.depth = d->framebuffer->attachments[sp.depthAttachment.idx]->image;
          if(sp.depthAttachment.clear)
          {
            op.frame.clear_depth = true;
            op.frame.clear_dep = d->clearval[ci++];
          }
        }
        // A target whose memory the pass overwrites completely (cleared, image covering the
        // whole allocation) and that nothing earlier in the submission touched needs no upload.
        auto overwritten = [&](const VkImage_T *img, bool cleared) {
          return cleared && img->offset == 0 && ImageBytes(img) >= img->mem->size &&
                 std::find(sub.uploads.begin(), sub.uploads.end(), img->mem) == sub.uploads.end() &&
                 !touched.count(img->mem);
        };
        if(!overwritten(op.frame.color, op.frame.clear_color))
          AddMem(sub.uploads, op.frame.color->mem);
        AddMem(sub.downloads, op.frame.color->mem);
        touched.insert(op.frame.color->mem);
        if(op.frame.depth)
        {
          if(!overwritten(op.frame.depth, op.frame.clear_depth))
            AddMem(sub.uploads, op.frame.depth->mem);
          AddMem(sub.downloads, op.frame.depth->mem);
          touched.insert(op.frame.depth->mem);
        }
        op.key = cb->recording && sub.ops.size() < 65536u ? (cb->recording << 16 | sub.ops.size()) : 0u;
        op.desc_writes = desc_writes;
        sub.ops.push_back(std::move(op));
        pass = (int)sub.ops.size() - 1;
        break;
      }
      case Cmd::BindPipeline: pipeline = ((const cmd::BindPipeline *)pull(sizeof(cmd::BindPipeline)))->pipeline; break;
      case Cmd::BindDescriptorSets:
      {
        const cmd::BindDescriptorSets *d = (const cmd::BindDescriptorSets *)pull(sizeof(cmd::BindDescriptorSets));
        if(d->idx < 8)
          sets[d->idx] = d->set;
        break;
      }
      case Cmd::BindVB:
      {
        const cmd::BindVB *d = (const cmd::BindVB *)pull(sizeof(cmd::BindVB));
        if(d->slot < 4)
        {
          vbs[d->slot] = d->buffer;
          vb_off[d->slot] = d->offset;
        }
        break;
      }
      case Cmd::BindIB:
      {
        const cmd::BindIB *d = (const cmd::BindIB *)pull(sizeof(cmd::BindIB));
        ib = d->buffer;
        ib_off = d->offset;
        ib_type = d->indexType;
        break;
      }
      case Cmd::SetViewport: pull(sizeof(cmd::SetViewport)); break;  // recorded, ignored (reference :108-114)
      case Cmd::PushConstants:
      {
        const cmd::PushConstants *d = (const cmd::PushConstants *)pull(sizeof(cmd::PushConstants));
        if(d->offset < 128)
          memcpy(push + d->offset, d->values, std::min<uint32_t>(d->size, 128 - d->offset));
        break;
      }
      case Cmd::Draw: case Cmd::DrawIndexed:
      {
        uint32_t count, first;
        bool indexed = id == Cmd::DrawIndexed;
        if(indexed)
        {
          const cmd::DrawIndexed *d = (const cmd::DrawIndexed *)pull(sizeof(cmd::DrawIndexed));
          count = d->indexCount;
          first = d->firstIndex;
        }
        else
        {
          const cmd::Draw *d = (const cmd::Draw *)pull(sizeof(cmd::Draw));
          count = d->vertexCount;
          first = d->firstVertex;
        }
        if(pass < 0)
        {
          err = "draw outside a render pass";
          return false;
        }
        if(!pipeline || !pipeline->pair)
        {
          err = "draw without a graphics pipeline";
          return false;
        }
        DrawDesc dd;
        memset(&dd.d, 0, sizeof(dd.d));
        dd.pipeline = pipeline;
        dd.frame = (uint32_t)pass;
        VaDraw &v = dd.d;
        // ShadeVerts (rasterizer.cpp:121-237): whole triangles of a list, n-2 of a strip
        if(pipeline->topology == VK_PRIMITIVE_TOPOLOGY_TRIANGLE_LIST)
          v.ntris = count / 3;
        else if(pipeline->topology == VK_PRIMITIVE_TOPOLOGY_TRIANGLE_STRIP)
        {
          v.ntris = count >= 3 ? count - 2 : 0;
          v.flags |= VA_DRAW_STRIP;
        }
        else
          v.ntris = 0;  // the reference prints "Unsupported primitive topology!" and draws nothing
        v.first = first;
        if(indexed)
        {
          if(!ib || !ib->mem)
          {
            err = "indexed draw without an index buffer";
            return false;
          }
          v.flags |= VA_DRAW_INDEXED | (ib_type == VK_INDEX_TYPE_UINT32 ? VA_DRAW_INDEX32 : 0u);
          v.ib = (unsigned long long)(uintptr_t)(ib->mem->dev + ib->offset + ib_off);
          v.ib_bytes = BoundBytes(ib, ib_off);
          AddMem(sub.uploads, ib->mem);
        }
        for(int s = 0; s < 4; s++)
          if(vbs[s] && vbs[s]->mem)
          {
            v.vb[s] = (unsigned long long)(uintptr_t)(vbs[s]->mem->dev + vbs[s]->offset + vb_off[s]);
            v.vb_bytes[s] = BoundBytes(vbs[s], vb_off[s]);
            AddMem(sub.uploads, vbs[s]->mem);
          }
        for(int a = 0; a < 16; a++)
        {
          v.vattr_fmt[a] = pipeline->vattrs[a].format;
          v.vattr_stride[a] = pipeline->vattrs[a].stride;
          v.vattr_off[a] = pipeline->vattrs[a].offset;
          v.vattr_vb[a] = std::min(pipeline->vattrs[a].vb, 3u);
        }
        v.cull_mode = pipeline->cullMode;
        v.front_face = pipeline->frontFace;
        v.depth_op = pipeline->depthCompareOp;
        v.depth_write = pipeline->depthWriteEnable ? 1u : 0u;
        v.blend_enable = pipeline->blend.blendEnable ? 1u : 0u;
        v.src_factor = pipeline->blend.srcColorBlendFactor;
        v.dst_factor = pipeline->blend.dstColorBlendFactor;
        v.blend_op = pipeline->blend.colorBlendOp;
        memcpy(v.push, push, 128);
        for(const ShaderResource &r : pipeline->pair->resources)
        {
          VkDescriptorSet_T *ds = r.set < 8 ? sets[r.set] : nullptr;
          if(!ds || r.binding >= ds->binds.size())
          {
            err = "draw references an unbound descriptor";
            return false;
          }
          const auto &b = ds->binds[r.binding];
          if(r.image)
          {
            if(!b.imageInfo.imageView || !b.imageInfo.imageView->image)
            {
              err = "combined image sampler descriptor without an image view";
              return false;
            }
            VkImage_T *img = b.imageInfo.imageView->image;
            if(!img->mem)
            {
              err = "sampled image has no memory bound";
              return false;
            }
            if(NeedsShadow(img))
            {
              if(!EnsureShadow(img, err))
                return false;
              dd.shadowed[r.binding] = img;
            }
            FillTex(v.tex[r.binding], img, b.imageInfo.sampler);
            AddMem(sub.uploads, img->mem);
          }
          else
          {
            VkBuffer_T *buf = b.bufferInfo.buffer;
            if(!buf || !buf->mem)
            {
              err = "uniform buffer descriptor without a buffer";
              return false;
            }
            v.ubo[r.binding] = (unsigned long long)(uintptr_t)(buf->mem->dev + buf->offset + b.bufferInfo.offset);
            v.ubo_bytes[r.binding] = BoundBytes(buf, b.bufferInfo.offset);
            AddMem(sub.uploads, buf->mem);
          }
        }
        if(v.ntris)
          sub.ops[pass].draws.push_back(dd);
        break;
      }
      case Cmd::CopyBuf2Img:
      {
        const cmd::CopyBuf2Img *d = (const cmd::CopyBuf2Img *)pull(sizeof(cmd::CopyBuf2Img));
        Op op;
        op.kind = Op::Copy;
        op.copy.to_image = true;
        op.copy.src = d->srcBuffer;
        op.copy.dstImg = d->dstImage;
        op.copy.dstBuf = nullptr;
        op.copy.bi = d->region;
        // the copy writes only its region: the rest of the destination's host-visible memory
        // must reach the device first, or the download after the submission would replace it
        AddMem(sub.uploads, d->srcBuffer->mem);
        AddMem(sub.uploads, d->dstImage->mem);
        AddMem(sub.downloads, d->dstImage->mem);
        sub.ops.push_back(op);
        break;
      }
      case Cmd::CopyBuf:
      {
        const cmd::CopyBuf *d = (const cmd::CopyBuf *)pull(sizeof(cmd::CopyBuf));
        Op op;
        op.kind = Op::Copy;
        op.copy.to_image = false;
        op.copy.src = d->srcBuffer;
        op.copy.dstBuf = d->dstBuffer;
        op.copy.dstImg = nullptr;
        op.copy.bb = d->region;
        // reference cmd_exec.cpp:174-182 copies without checks; here a region outside either
        // buffer (or its memory) is refused before any device work
        auto inside = [](const VkBuffer_T *b, VkDeviceSize off, VkDeviceSize size) {
          return b && b->mem && off <= b->size && size <= b->size - off && b->offset <= b->mem->size &&
                 b->size <= b->mem->size - b->offset;
        };
        if(!inside(d->srcBuffer, d->region.srcOffset, d->region.size) ||
           !inside(d->dstBuffer, d->region.dstOffset, d->region.size))
        {
          err = "buffer copy region outside its buffers";
          return false;
        }
        AddMem(sub.uploads, d->srcBuffer->mem);
        AddMem(sub.uploads, d->dstBuffer->mem);
        AddMem(sub.downloads, d->dstBuffer->mem);
        sub.ops.push_back(op);
        break;
      }
      default:
        err = "corrupt command stream";
        return false;
    }
  }
  return true;
}

VkResult Queue::RunCopy(const CopyDesc &c)
{
  if(c.to_image)
  {
    // reference cmd_exec.cpp:143-173 copies whole, tightly packed levels of one layer; here any
    // region of an uncompressed image (offsets, bufferRowLength / bufferImageHeight, several
    // layers) is copied as 2D strided DMA, one layer at a time. Block-compressed images keep
    // the whole-level rule (their linear layout has no rows of texels).
    const VkImage_T *img = c.dstImg;
    const VkBufferImageCopy &r = c.bi;
    const uint32_t mip = r.imageSubresource.mipLevel;
    const uint64_t W = std::max(1u, img->extent.width >> mip), H = std::max(1u, img->extent.height >> mip);
    const uint64_t bpp = img->bytesPerPixel;
    const bool block = img->format == VK_FORMAT_BC2_UNORM_BLOCK || img->format == VK_FORMAT_BC3_UNORM_BLOCK;
    const uint64_t x = (uint32_t)r.imageOffset.x, y = (uint32_t)r.imageOffset.y;
    const uint64_t w = r.imageExtent.width, h = r.imageExtent.height;
    const uint64_t row = r.bufferRowLength ? r.bufferRowLength : w;
    const uint64_t rows = r.bufferImageHeight ? r.bufferImageHeight : h;
    const uint32_t layers = r.imageSubresource.layerCount, layer0 = r.imageSubresource.baseArrayLayer;
    bool ok = mip < img->mipLevels && r.imageOffset.x >= 0 && r.imageOffset.y >= 0 && r.imageOffset.z == 0 &&
              r.imageExtent.depth <= 1 && x + w <= W && y + h <= H && row >= w && rows >= h && layers >= 1 &&
              (uint64_t)layer0 + layers <= img->arrayLayers;
    if(ok && block)
      ok = x == 0 && y == 0 && w == W && h == H && row == w && rows == h;
    const uint64_t layer_bytes = row * rows * bpp;
    const uint64_t src_need = r.bufferOffset + layer_bytes * (layers - 1) + ((h - 1) * row + w) * bpp;
    if(ok && w && h)
      ok = src_need <= c.src->size && c.src->offset + src_need <= c.src->mem->size &&
           img->offset + SubresourceOffset(img, mip, layer0 + layers - 1) +
                                                               ((y + h - 1) * W + x + w) * bpp <= img->mem->size;
    if(!ok)
    {
      LogError("vkCmdCopyBufferToImage: region outside the image or the buffer, or a partial copy of a "
               "block-compressed image");
      return VK_ERROR_FEATURE_NOT_PRESENT;
    }
    if(!w || !h)
      return VK_SUCCESS;
    for(uint32_t l = 0; l < layers; l++)
    {
      uint8_t *dst = img->mem->dev + img->offset + SubresourceOffset(img, mip, layer0 + l) + (y * W + x) * bpp;
      const uint8_t *src = c.src->mem->dev + c.src->offset + r.bufferOffset + layer_bytes * l;
      HIPCHK(hipMemcpy2DAsync(dst, W * bpp, src, row * bpp, w * bpp, h, hipMemcpyDeviceToDevice, stream));
    }
  }
  else
  {
    HIPCHK(hipMemcpyAsync(c.dstBuf->mem->dev + c.dstBuf->offset + c.bb.dstOffset,
                          c.src->mem->dev + c.src->offset + c.bb.srcOffset, c.bb.size, hipMemcpyDeviceToDevice, stream));
  }
  return VK_SUCCESS;
}

// reference rasterizer.cpp:312-361 (ClearTarget x2)
void AccumulateStats(ExecStats &acc, const ExecStats &st)
{
  acc.gpu_ms_total += st.gpu_ms_total;
  acc.raster_ms += st.raster_ms;
  acc.setup_ms += st.setup_ms;
  acc.bin_ms += st.bin_ms;
  acc.raster_launches += st.raster_launches;
  acc.total_tris += st.total_tris;
  acc.total_pairs += st.total_pairs;
  acc.total_tiles += st.total_tiles;
  acc.graph_launches += st.graph_launches;
  acc.ctr.pixels_written += st.ctr.pixels_written;
  acc.ctr.depth_passed += st.ctr.depth_passed;
  acc.ctr.tris_in += st.ctr.tris_in;
  acc.ctr.tris_out += st.ctr.tris_out;
  acc.ctr.draws += st.ctr.draws;
  acc.ctr.blocks += st.ctr.blocks;
  acc.ctr.pixels_tested += st.ctr.pixels_tested;
  acc.ctr.shaded_blend += st.ctr.shaded_blend;
  acc.ctr.shaded_dwrite += st.ctr.shaded_dwrite;
  acc.last_cfg = st.last_cfg;
  acc.last_tuning = st.last_tuning;
}

// float -> int32 as x64 converts it (truncation; 0x80000000 for NaN and outside [-2^31, 2^31))
static int32_t X64FloatToInt(float f)
{
  return (f >= -2147483648.0f && f < 2147483648.0f) ? (int32_t)f : INT32_MIN;
}

VkResult Queue::ClearFrame(const FrameDesc &f)
{
  if(f.clear_color)
  {
    VkImage_T *img = f.color;
    uint64_t n = (uint64_t)img->extent.width * img->extent.height;
    const float *c = f.clear_col.color.float32;
    if(img->bytesPerPixel == 8)
    {
      // RGBA16F (beyond the reference)
      uint32_t h[4];
      for(int i = 0; i < 4; i++)
        h[i] = HalfBits(c[i]);
      HIPCHK(va_fill_u64_async(img->devPixels(), n, h[0] | (h[1] << 16), h[2] | (h[3] << 16), stream));
    }
    else
    {
      // reference rasterizer.cpp:341-345: the low byte of the x64 float->int conversion
      uint8_t e[4];
      e[2] = (uint8_t)X64FloatToInt(c[0] * 255.0f);
      e[1] = (uint8_t)X64FloatToInt(c[1] * 255.0f);
      e[0] = (uint8_t)X64FloatToInt(c[2] * 255.0f);
      e[3] = (uint8_t)X64FloatToInt(c[3] * 255.0f);
      if(img->bytesPerPixel == 1)
        HIPCHK(hipMemsetD8Async((hipDeviceptr_t)img->devPixels(), e[2], n, stream));
      else
      {
        uint32_t v;
        memcpy(&v, e, 4);
        HIPCHK(va_fill_u32_async(img->devPixels(), n, v, stream));
      }
    }
  }
  if(f.clear_depth && f.depth)
  {
    uint64_t n = (uint64_t)f.depth->extent.width * f.depth->extent.height;
    uint32_t v;
    memcpy(&v, &f.clear_dep.depthStencil.depth, 4);
    HIPCHK(va_fill_u32_async(f.depth->devPixels(), n, v, stream));
  }
  return VK_SUCCESS;
}

static int BitsFor(uint32_t n)
{
  int b = 1;
  while(b < 32 && (1ull << b) < n)
    b++;
  return b;
}

// ---------------------------------------------------------------------------------------------
// Autotuner. Every launch shape (tile size and its workgroup of (T/8)^2 waves, LDS chunk) gives
// the same bytes, so a repeated pass shape is simply run with the next untried candidate and
// timed by the device stamps; after two rounds over the candidates the fastest is kept. The
// small-triangle threshold (third field) is the queue's for every candidate.

static const TileConfig kTuneCandidates[] = {
    {3, 64, 0, 0}, {3, 32, 0, 0}, {4, 64, 0, 0}, {4, 32, 0, 0}, {5, 64, 0, 0},
    {3, 32, 0, 1}, {4, 32, 0, 1},
};
static const int kTuneRounds = 2;

uint64_t Queue::PassSignature(const std::vector<Op *> &passes) const
{
  uint64_t h = 1469598103934665603ull;
  auto mix = [&](uint64_t v) {
    h ^= v;
    h *= 1099511628211ull;
  };
  for(Op *p : passes)
  {
    mix(p->frame.color->extent.width);
    mix(p->frame.color->extent.height);
    mix(p->frame.color->bytesPerPixel);
    mix(p->frame.depth ? 1 : 0);
    mix(p->draws.size());
    uint64_t tris = 0;
    for(const DrawDesc &dd : p->draws)
      tris += dd.d.ntris;
    mix(tris);
  }
  return h;
}

TileConfig Queue::ChooseConfig(uint64_t sig, bool *tuning)
{
  *tuning = false;
  if(!autotune || mode != ExecMode::Batched)
    return tiles;
  TuneState &t = tune[sig];
  if(t.done)
    return t.best;
  const size_t nc = sizeof(kTuneCandidates) / sizeof(kTuneCandidates[0]);
  *tuning = true;
  TileConfig c = kTuneCandidates[t.runs % nc];
  c.small = tiles.small;  // the candidates differ in tile size, chunk and grid form only
  return c;
}

void Queue::RecordTuning(uint64_t sig, const TileConfig &cfg, double ms)
{
  TuneState &t = tune[sig];
  const size_t nc = sizeof(kTuneCandidates) / sizeof(kTuneCandidates[0]);
  if(t.done)
    return;
  if(t.ms.size() < nc)
    t.ms.assign(nc, 1e30);
  size_t i = t.runs % nc;
  t.ms[i] = std::min(t.ms[i], ms);
  t.runs++;
  if(t.runs >= nc * kTuneRounds)
  {
    size_t b = 0;
    for(size_t k = 1; k < nc; k++)
      if(t.ms[k] < t.ms[b])
        b = k;
    t.best = kTuneCandidates[b];
    t.best.small = tiles.small;
    t.done = true;
    (void)cfg;
  }
}

VkResult Queue::RunPass(const std::vector<Op *> &passes, bool fused_clear, ExecStats &st, PassLaunch *deferred)
{
  // A pass that stops part way may leave counts in the counter rows: every early return makes
  // the next pass zero them again.
  struct ResetOnEarlyReturn
  {
    void *&zeroed;
    bool done = false;
    ~ResetOnEarlyReturn()
    {
      if(!done)
        zeroed = nullptr;
    }
  } reset_guard{rows_zeroed};
  // a single pass decoded from a recording reuses the tables built for it last time
  const bool reusable = passes.size() == 1 && passes[0]->key != 0;
  const bool have = reusable && tables_cache.key == passes[0]->key &&
                    tables_cache.desc_writes == passes[0]->desc_writes;
  PassTables local;
  PassTables &tb = reusable ? tables_cache : local;
  if(!have)
  {
    tb.key = 0;  // invalid until rebuilt below
    tb.sig = PassSignature(passes);
  }
  bool tuning = false;
  const uint64_t sig = tb.sig;
  const TileConfig cfg = ChooseConfig(sig, &tuning);
  const uint32_t ts = cfg.tile_shift, T = 1u << ts;
  std::vector<VaFrame> &frames = tb.frames;
  std::vector<VaDraw> &draws = tb.draws;
  std::vector<std::shared_ptr<ShaderPair>> &pairs = tb.pairs;
  std::vector<uint32_t> &block_draw = tb.block_draw;  // setup workgroup -> draw
  uint32_t &tile_total = tb.tile_total, &tri_total = tb.tri_total;
  if(!(have && tb.tile_shift == ts && tb.fused_clear == fused_clear))
  {
    frames.clear();
    draws.clear();
    pairs.clear();
    block_draw.clear();
    tile_total = 0;
    tri_total = 0;
    for(Op *p : passes)
    {
      VaFrame fr;
      memset(&fr, 0, sizeof(fr));
      VkImage_T *c = p->frame.color;
      fr.color = (unsigned long long)(uintptr_t)c->devPixels();
      fr.depth = p->frame.depth ? (unsigned long long)(uintptr_t)p->frame.depth->devPixels() : 0ull;
      fr.width = c->extent.width;
      fr.height = c->extent.height;
      fr.color_bpp = c->bytesPerPixel;
      if(fused_clear)
      {
        fr.clear_flags = (p->frame.clear_color ? VA_CLEAR_COLOR : 0u) | (p->frame.clear_depth && p->frame.depth ? VA_CLEAR_DEPTH : 0u);
        memcpy(fr.clear_color, p->frame.clear_col.color.float32, 16);
        fr.clear_depth = p->frame.clear_dep.depthStencil.depth;
      }
      fr.tiles_x = (fr.width + T - 1) >> ts;
      fr.tiles_y = (fr.height + T - 1) >> ts;
      fr.tile_base = tile_total;
      if(fr.width >= 65535 || fr.height >= 65535)
      {
        LogError("render target larger than 65534 pixels per side");
        return VK_ERROR_FEATURE_NOT_PRESENT;
      }
      tile_total += fr.tiles_x * fr.tiles_y;
      for(DrawDesc &dd : p->draws)
      {
        VaDraw v = dd.d;
        v.frame = (uint32_t)frames.size();
        auto it = std::find(pairs.begin(), pairs.end(), dd.pipeline->pair);
        if(it == pairs.end())
        {
          pairs.push_back(dd.pipeline->pair);
          it = pairs.end() - 1;
        }
        v.slot = (uint32_t)(it - pairs.begin());
        v.tri_base = tri_total;
        tri_total += v.ntris;
        v.block_base = (uint32_t)block_draw.size();
        for(uint32_t b = 0; b < (v.ntris + VA_SETUP_PRIMS - 1) / VA_SETUP_PRIMS; b++)
          block_draw.push_back((uint32_t)draws.size());
        draws.push_back(v);
      }
      frames.push_back(fr);
    }
    // the device's block table: block_draw, then each block's first primitive (and the total)
    tb.block_tab = block_draw;
    for(uint32_t b = 0; b < (uint32_t)block_draw.size(); b++)
    {
      const VaDraw &v = draws[block_draw[b]];
      tb.block_tab.push_back(v.tri_base + (b - v.block_base) * VA_SETUP_PRIMS);
    }
    tb.block_tab.push_back(tri_total);
    tb.version = 0;
    if(reusable)
    {
      tb.key = passes[0]->key;
      tb.desc_writes = passes[0]->desc_writes;
      tb.tile_shift = ts;
      tb.fused_clear = fused_clear;
      tb.version = ++table_versions;
    }
  }
  bool any_clear = false;
  for(auto &f : frames)
    any_clear |= f.clear_flags != 0;
  if(tri_total == 0 && !any_clear)
    return VK_SUCCESS;
  // Only a pass with primitives (whose emit kernel sees the overflow flag) and a chosen launch
  // shape is launched without waiting; another one waits for the passes still running first.
  const bool can_defer = !tuning && tri_total > 0;
  if(deferred && !can_defer && !inflight.empty())
  {
    deferred->drain_first = true;
    reset_guard.done = true;
    return VK_SUCCESS;
  }
  std::shared_ptr<PipelineModule> mod;
  if(!pairs.empty())
  {
    std::string err;
    mod = Device::Get().GetModule(pairs, err, cfg);
    if(!mod)
    {
      LogError("pipeline module: %s", err.c_str());
      return VK_ERROR_DEVICE_LOST;
    }
  }
  else
  {
    // a clear-only pass still needs raster kernels: any module will do; use a cached one
    LogError("clear-only render pass in batched mode: using fill kernels");
    for(Op *p : passes)
    {
      VkResult r = ClearFrame(p->frame);
      if(r != VK_SUCCESS)
        return r;
    }
    return VK_SUCCESS;
  }
  const uint32_t stride = mod->rec_stride;
  const uint32_t nframes = (uint32_t)frames.size(), ndraws = (uint32_t)draws.size();
  if(ndraws >= (1u << 23))
  {
    LogError("a render pass batch holds at most 8388607 draws");
    return VK_ERROR_FEATURE_NOT_PRESENT;
  }
  const uint32_t nblocks = (uint32_t)block_draw.size();
  if(!Grow(&d_frames, c_frames, sizeof(VaFrame) * nframes) || !Grow(&d_draws, c_draws, sizeof(VaDraw) * std::max(ndraws, 1u)) ||
     !Grow(&d_rec, c_rec, (size_t)stride * 4 * std::max(tri_total, 1u)) || !Grow(&d_count, c_count, 4ull * (tri_total + 1)) ||
     !Grow(&d_bpairs, c_bpairs, 4ull * (nblocks + 1)) || !Grow(&d_start, c_start, 4ull * (tile_total + 1)) ||
     !Grow(&d_einfo, c_einfo, 16ull * std::max(tri_total, 1u)) || !Grow(&d_stamps, c_stamps, 256) ||
     !Grow(&d_blocks, c_blocks, 4ull * (2ull * nblocks + 1)) ||
     !Grow(&d_part, c_part, 2ull * 32ull * VA_COUNTER_ROWS))
    return VK_ERROR_OUT_OF_DEVICE_MEMORY;
  // d_stamps: stage stamps (u64 0-3), persistent tile counter (u32 8) and overflow flag (u32 34).
  // The flag starts at zero and is zeroed only for a new allocation (afterwards the host clears it
  // when it has seen the overflow): a pass launched behind one still running must see its flag.
  // d_part: the counter rows of setup and raster, zero when a pass begins (the finish kernel
  // leaves them zero; a pass that stopped part way has them zeroed again).
  if(d_part != rows_zeroed)
  {
    HIPCHK(va_fill_u32_async(d_part, 2 * 8 * VA_COUNTER_ROWS, 0u, stream));
    rows_zeroed = d_part;
  }
  if(d_stamps != poison_zeroed)
  {
    // the overflow flag (word 34) and the pass sequence counter (word 38)
    HIPCHK(va_fill_u32_async((uint32_t *)d_stamps + 34, 6, 0u, stream));
    poison_zeroed = d_stamps;
    launch_seq = 0;
  }
  if(!h_pinned)  // written by the finish kernel, read by the host after the stream completes
    HIPCHK(hipHostMalloc((void **)&h_pinned, 4096, hipHostMallocCoherent));

  // sampled images read through an RGBA8 shadow get it refreshed from their current contents
  {
    std::vector<const VkImage_T *> shadows;
    for(Op *p : passes)
      for(const DrawDesc &dd : p->draws)
        for(const VkImage_T *im : dd.shadowed)
          if(im && std::find(shadows.begin(), shadows.end(), im) == shadows.end())
            shadows.push_back(im);
    for(const VkImage_T *im : shadows)
      HIPCHK(TranscodeAsync(im, stream));
  }

  // the per-pass tables go up only when they changed (a re-submitted command buffer reuses them;
  // tables reused from the cache and already resident are not even compared)
  const bool resident = tb.version != 0 && up_version == tb.version && d_frames == up_frames_ptr &&
                        d_draws == up_draws_ptr && d_blocks == up_blocks_ptr;
  if(!resident)
  {
  if(d_frames != up_frames_ptr || frames.size() != up_frames.size() ||
     memcmp(frames.data(), up_frames.data(), sizeof(VaFrame) * nframes) != 0)
  {
    HIPCHK(hipMemcpyAsync(d_frames, frames.data(), sizeof(VaFrame) * nframes, hipMemcpyHostToDevice, stream));
    up_frames = frames;
    up_frames_ptr = d_frames;
  }
  if(ndraws && (d_draws != up_draws_ptr || draws.size() != up_draws.size() ||
                memcmp(draws.data(), up_draws.data(), sizeof(VaDraw) * ndraws) != 0))
  {
    HIPCHK(hipMemcpyAsync(d_draws, draws.data(), sizeof(VaDraw) * ndraws, hipMemcpyHostToDevice, stream));
    up_draws = draws;
    up_draws_ptr = d_draws;
  }
  if(nblocks && (d_blocks != up_blocks_ptr || tb.block_tab != up_blocks))
  {
    HIPCHK(hipMemcpyAsync(d_blocks, tb.block_tab.data(), 4 * tb.block_tab.size(), hipMemcpyHostToDevice, stream));
    up_blocks = tb.block_tab;
    up_blocks_ptr = d_blocks;
  }
  up_version = tb.version;
  }

  // The pair buffers are sized by a capacity that only grows: the device work never waits for
  // the host. If a frame needs more pairs than the capacity, the raster launch writes nothing
  // and the pass runs again with larger buffers. The launches (sort blocks, histogram words) are
  // sized by the capacity of the pass shape (signature and tile size): its first pass runs at the
  // queue's capacity, later ones at the pairs it needed plus an eighth, rounded up to 64K.
  if(pair_cap == 0)
    pair_cap = std::max<uint32_t>(1u << 20, 2 * tri_total);
  if(shape_caps.size() > 4096)
    shape_caps.clear();
  uint32_t &shape_cap = shape_caps[sig * 8u + ts];
  uint32_t pass_cap = shape_cap ? shape_cap : pair_cap;
  pair_cap = std::max(pair_cap, pass_cap);
  uint32_t total_pairs = 0;
  int overflow_retries = 0;
  // output words of this launch in pinned host memory: two slots, so that a pass can be
  // launched while the previous one's results still wait to be read
  const int slot = inflight.empty() ? 0 : 1 - inflight.back().slot;
  const bool defer = deferred && can_defer;
  for(;;)
  {
    if(!Grow(&d_keys, c_keys, 4ull * pair_cap) || !Grow(&d_vals, c_vals, 4ull * pair_cap) ||
       !Grow(&d_keys2, c_keys2, 4ull * pair_cap) || !Grow(&d_vals2, c_vals2, 4ull * pair_cap) ||
       !Grow(&d_stemp, c_stemp, 4 * va_sort_temp_words(pair_cap)) ||
       !Grow(&d_masks, c_masks, RasterMaskBytes(cfg, pair_cap, tile_total)))
      return VK_ERROR_OUT_OF_DEVICE_MEMORY;
    VaDraw *dd = (VaDraw *)d_draws;
    VaFrame *df = (VaFrame *)d_frames;
    // block_pairs: (tile, primitive) pairs per setup block, then the pair total
    uint32_t *rec = (uint32_t *)d_rec, *cnt = (uint32_t *)d_count, *block_pairs = (uint32_t *)d_bpairs;
    uint32_t *bd = (uint32_t *)d_blocks, *total_dev = block_pairs + nblocks;
    // counter rows: setup's VA_COUNTER_ROWS rows, then the raster's
    uint64_t *part = (uint64_t *)d_part, *part_r = part + 4ull * VA_COUNTER_ROWS;
    uint64_t *h_out = (uint64_t *)h_pinned + 16 * slot;
    uint32_t *poison = (uint32_t *)d_stamps + 34;
    uint32_t *seq_dev = (uint32_t *)d_stamps + 38;  // passes that reached the finish kernel
    uint32_t cap = pass_cap;
    // The device steps of the pass in four stages, kernels only, captured together into one
    // graph: 0 setup, 1 binning, 2 raster, 3 finish (counter sums, pair total and the stage
    // stamps written into host memory). Setup, emit and raster stamp their own start.
    const int sort_passes = (BitsFor(tile_total) + 7) / 8;
    const bool in_alt = tri_total && (sort_passes & 1);
    uint32_t *bins = in_alt ? (uint32_t *)d_vals2 : (uint32_t *)d_vals;
    uint32_t *sorted_keys = in_alt ? (uint32_t *)d_keys2 : (uint32_t *)d_keys;
    auto enqueue = [&](int stage) -> hipError_t {
      hipError_t e;
#define VA_TRY(x)               \
  do                            \
  {                             \
    if((e = (x)) != hipSuccess) \
      return e;                 \
  } while(0)
      uint64_t *stamps = (uint64_t *)d_stamps, *stamp_r = stamps + 2;
      if(stage == 0)
      {
        if(tri_total)
        {
          // the setup kernel writes stamp 0 itself
          uint4 *einfo = (uint4 *)d_einfo;
          void *args[] = {&dd, &bd, &df, (void *)&ts, &rec, &cnt, &part, &einfo, &block_pairs, &stamps};
          VA_TRY(hipModuleLaunchKernel(mod->setup, nblocks, 1, 1, VA_SETUP_PRIMS, 1, 1, 0, stream, args, nullptr));
        }
        else
        {
          VA_TRY(va_stamp_async(stamps, stream));
          VA_TRY(va_fill_u32_async(total_dev, 1, 0u, stream));
        }
      }
      else if(stage == 1)
      {
        if(tri_total)
        {
          uint32_t *keys = (uint32_t *)d_keys, *vals = (uint32_t *)d_vals;
          uint4 *einfo = (uint4 *)d_einfo;
          uint32_t *block_first = bd + nblocks;
          // groups of four setup blocks; a group is split over several workgroups when the pass's
          // capacity allows more than 4 pairs per thread on average (few, large triangles)
          const uint32_t groups = (nblocks + 3) / 4;
          uint32_t split = (uint32_t)std::min<uint64_t>(32, std::max<uint64_t>(1, (uint64_t)cap / (4096ull * groups)));
          void *eargs[] = {&block_first, &block_pairs, (void *)&nblocks, &einfo,  &cnt,   &cap,
                           &keys,        &vals,        &total_dev,       &stamps, &poison, &split};
          VA_TRY(hipModuleLaunchKernel(mod->emit, groups * split, 1, 1, 1024, 1, 1, 0, stream, eargs, nullptr));
          int alt = 0;
          VA_TRY(va_sort_pairs((uint32_t *)d_stemp, keys, vals, (uint32_t *)d_keys2, (uint32_t *)d_vals2, total_dev, cap,
                               BitsFor(tile_total), &alt, stream));
          if((alt != 0) != in_alt)
            return hipErrorUnknown;
          uint32_t *st_ = (uint32_t *)d_start;
          void *rargs[] = {&sorted_keys, &total_dev, &cap, &tile_total, &st_};
          // a grid-stride pass over the (at most cap + 1) bin boundaries, one per thread up to 4M
          // pairs (a thread that strides runs its key loads one step after another); workgroups
          // past the pass's pair count leave at once
          const uint32_t rblocks = (uint32_t)std::min<uint64_t>(((uint64_t)cap + 256) / 256, 16384);
          VA_TRY(hipModuleLaunchKernel(mod->ranges, rblocks, 1, 1, 256, 1, 1, 0, stream, rargs, nullptr));
        }
        else
        {
          VA_TRY(va_stamp_async(stamps + 1, stream));
          VA_TRY(va_fill_u32_async(d_start, tile_total + 1, 0u, stream));
        }
      }
      else if(stage == 3)
        VA_TRY(va_finish_async(part, VA_COUNTER_ROWS, total_dev, stamps, h_out, seq_dev, stream));
      else
      {
        // tile t's bin is [start[t], start[t + 1])
        uint32_t *st_ = (uint32_t *)d_start, *en = (uint32_t *)d_start + 1;
        uint32_t use_fused = fused_clear ? 1u : 0u;
        void *masks = d_masks;
        void *args[] = {&dd,         &df,        (void *)&nframes, &rec,      &bins,  &st_,    &en,
                        &tile_total, &use_fused, &part_r,          &total_dev, &cap,  &masks, &stamp_r};
        uint32_t threads = (T / 8) * (T / 8) * 64;
        if(cfg.persistent)
        {
          uint32_t *tile_counter = (uint32_t *)d_stamps + 8;
          VA_TRY(va_fill_u32_async(tile_counter, 1, 0u, stream));
          void *pargs[] = {&dd,         &df,        (void *)&nframes, &rec,       &bins,  &st_,          &en,
                           &tile_total, &use_fused, &part_r,          &total_dev, &cap,  &masks, &tile_counter,
                           &stamp_r};
          const uint32_t grid = std::min(mod->raster_p_grid[ts - 3], std::max(tile_total, 1u));
          VA_TRY(hipModuleLaunchKernel(mod->raster_p[ts - 3], grid, 1, 1, threads, 1, 1, 0, stream, pargs, nullptr));
          return hipSuccess;
        }
        hipFunction_t k = mod->raster[ts - 3];
        VA_TRY(hipModuleLaunchKernel(k, tile_total, 1, 1, threads, 1, 1, 0, stream, args, nullptr));
      }
#undef VA_TRY
      return hipSuccess;
    };

    const GraphSet *gs = nullptr;
    if(graphs && mode == ExecMode::Batched)
    {
      // the graphs' kernel arguments are fixed by these values; anything else is table content
      GraphKey key;
      memset(&key, 0, sizeof(key));
      key.mod = mod.get();
      key.dims[0] = ts;
      key.dims[1] = tri_total;
      key.dims[2] = nblocks;
      key.dims[3] = tile_total;
      key.dims[4] = nframes;
      key.dims[5] = cap;
      key.dims[6] = fused_clear ? 1u : 0u;
      key.dims[7] = cfg.persistent;
      void *bufs[] = {d_frames, d_draws, d_rec,   d_count, d_bpairs,   d_keys,   d_vals,  d_keys2, d_vals2,
                      d_stemp,  d_start, d_einfo, d_blocks, d_part, d_stamps, d_masks, h_out};
      static_assert(sizeof(bufs) <= sizeof(key.bufs), "graph key holds every buffer");
      memcpy(key.bufs, bufs, sizeof(bufs));
      for(size_t i = 0; i < graph_cache.size(); i++)
        if(memcmp(&graph_cache[i].key, &key, sizeof(key)) == 0)
        {
          std::rotate(graph_cache.begin(), graph_cache.begin() + i, graph_cache.begin() + i + 1);
          gs = &graph_cache[0];
          break;
        }
      if(!gs)
      {
        GraphSet set;
        set.key = key;
        hipError_t err = hipSuccess;
        {
          hipError_t e = hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal);
          hipError_t er = e;
          for(int stage = 0; stage < 4 && er == hipSuccess; stage++)
            er = enqueue(stage);
          hipGraph_t g = nullptr;
          hipError_t ec = e == hipSuccess ? hipStreamEndCapture(stream, &g) : e;
          err = er != hipSuccess ? er : ec;
          if(err == hipSuccess && g)
            err = hipGraphInstantiate(&set.exec[0], g, nullptr, nullptr, 0);
          else if(err == hipSuccess)
            err = hipErrorUnknown;
          if(g)
            hipGraphDestroy(g);
        }
        if(err != hipSuccess)
        {
          for(hipGraphExec_t x : set.exec)
            if(x)
              hipGraphExecDestroy(x);
          (void)hipGetLastError();
          LogError("graph capture failed (%s); running passes kernel by kernel", hipGetErrorString(err));
          graphs = false;
        }
        else
        {
          graph_cache.insert(graph_cache.begin(), set);
          if(graph_cache.size() > 8)
          {
            for(hipGraphExec_t x : graph_cache.back().exec)
              hipGraphExecDestroy(x);
            graph_cache.pop_back();
          }
          gs = &graph_cache[0];
        }
      }
    }
    // the whole pass is one graph launch; the stage times come from the device stamps
    if(gs)
    {
      HIPCHK(hipGraphLaunch(gs->exec[0], stream));
      st.graph_launches++;
    }
    else
      for(int stage = 0; stage < 4; stage++)
        HIPCHK(enqueue(stage));
    // the finish kernel of this launch will publish this number in the slot's word 14
    const uint32_t seq = ++launch_seq;
    if(defer)
    {
      PassLaunch &pl = *deferred;
      pl.launched = true;
      pl.passes = passes;
      pl.fused_clear = fused_clear;
      pl.st = &st;
      pl.sig = sig;
      pl.cfg = cfg;
      pl.pass_cap = pass_cap;
      pl.ndraws = ndraws;
      pl.tri_total = tri_total;
      pl.tile_total = tile_total;
      pl.slot = slot;
      pl.seq = seq;
      reset_guard.done = true;
      return VK_SUCCESS;
    }
    HIPCHK(WaitPass(slot, seq));
    total_pairs = (uint32_t)h_out[0];
    if(total_pairs <= pass_cap)
      break;
    // the launch set the overflow flag; it is cleared before the pass runs again
    HIPCHK(va_fill_u32_async(poison, 1, 0u, stream));
    if(overflow_retries++ > 0)
    {
      LogError("bin capacity did not converge (%u pairs)", total_pairs);
      return VK_ERROR_OUT_OF_DEVICE_MEMORY;
    }
    if(total_pairs != 0xffffffffu)  // all ones: the launch only saw an earlier overflow's flag
    {
      pass_cap = (uint32_t)std::min<uint64_t>((uint64_t)total_pairs + total_pairs / 4 + 4096, 0xffff0000u);
      pair_cap = std::max(pair_cap, pass_cap);
      shape_cap = pass_cap;
    }
  }
  drained = true;  // everything enqueued so far has completed
  reset_guard.done = true;
  PassDone((const uint64_t *)h_pinned + 16 * slot, total_pairs, sig, cfg, tuning, ndraws, tri_total, tile_total, st);
  return VK_SUCCESS;
}

void Queue::PassDone(const uint64_t *h_out, uint32_t total_pairs, uint64_t sig, const TileConfig &cfg, bool tuning,
                     uint32_t ndraws, uint32_t tri_total, uint32_t tile_total, ExecStats &st)
{
  {
    // later passes of this shape launch with the pairs it needed plus an eighth
    const uint64_t want = ((uint64_t)total_pairs + total_pairs / 8 + 4096 + 0xffffu) & ~(uint64_t)0xffffu;
    if(shape_caps.size() > 4096)
      shape_caps.clear();
    shape_caps[sig * 8u + cfg.tile_shift] = (uint32_t)std::min<uint64_t>(std::max<uint64_t>(want, 1u << 16), 0xffff0000u);
  }
  VaCounters c;
  memset(&c, 0, sizeof(c));
  memcpy(&c.tris_in, h_out + 2, 8 * sizeof(uint64_t));  // tris_in .. shaded_dwrite
  c.draws = ndraws;
  st.last_cfg = cfg;
  st.last_tuning = tuning;
  {
    // stamps: 0 setup start, 1 binning start, 2 raster start, 3 finish start (constant-rate clock)
    const uint64_t s0 = h_out[10], s1 = h_out[11], s2 = h_out[12], s3 = h_out[13];
    const double khz = (double)Device::Get().wall_khz;
    if(khz > 0.0 && s0 <= s1 && s1 <= s2 && s2 <= s3)
    {
      st.gpu_ms_total += (double)(s3 - s0) / khz;
      st.setup_ms += (double)(s1 - s0) / khz;
      st.bin_ms += (double)(s2 - s1) / khz;
      st.raster_ms += (double)(s3 - s2) / khz;
      if(tuning)
        RecordTuning(sig, cfg, (double)(s3 - s0) / khz);
    }
    else
      LogError("stage timing unavailable (device clock stamps out of order)");
  }
  st.raster_launches++;
  st.total_tris += tri_total;
  st.total_pairs += total_pairs;
  st.total_tiles += tile_total;
  st.ctr.draws += c.draws;
  st.ctr.tris_in += c.tris_in;
  st.ctr.tris_out += c.tris_out;
  st.ctr.blocks += c.blocks;
  st.ctr.pixels_tested += c.pixels_tested;
  st.ctr.pixels_written += c.pixels_written;
  st.ctr.depth_passed += c.depth_passed;
  st.ctr.shaded_blend += c.shaded_blend;
  st.ctr.shaded_dwrite += c.shaded_dwrite;
}

void Queue::SignalFence(VkFence f, VkResult r)
{
  if(!f)
    return;
  std::lock_guard<std::mutex> lk(f->m);
  f->signaled = true;
  f->status = r;
  f->cv.notify_all();
}

hipError_t Queue::WaitPass(int slot, uint32_t seq)
{
  // The finish kernel writes the sequence number into fine-grained host memory after the rest of
  // the slot: spin on it (a pass is a fraction of a millisecond), and block on the stream only
  // when it takes longer than 50 ms.
  volatile const uint64_t *w = (volatile const uint64_t *)h_pinned + 16 * slot + 14;
  const double t0 = NowSeconds();
  for(uint32_t spin = 1; (uint32_t)*w != seq; spin++)
  {
    if((spin & 1023u) == 0u && NowSeconds() - t0 > 0.05)
    {
      const hipError_t e = hipStreamSynchronize(stream);
      if(e != hipSuccess)
        return e;
      return (uint32_t)*w == seq ? hipSuccess : hipErrorUnknown;
    }
#if defined(__x86_64__)
    __builtin_ia32_pause();
#endif
  }
  std::atomic_thread_fence(std::memory_order_acquire);
  return hipSuccess;
}

VkResult Queue::CompleteOldest(std::vector<int> &open, std::vector<VkResult> &sres)
{
  PassLaunch p = std::move(inflight.front());
  inflight.pop_front();
  auto fail = [&](VkResult r) {
    // nothing launched can be trusted: the pass and those behind it end without results
    (void)hipStreamSynchronize(stream);
    (void)hipGetLastError();
    auto drop = [&](const PassLaunch &q) {
      open[q.sub]--;
      if(sres[q.sub] == VK_SUCCESS)
        sres[q.sub] = r;
    };
    drop(p);
    for(; !inflight.empty(); inflight.pop_front())
      drop(inflight.front());
    return r;
  };
  hipError_t e = WaitPass(p.slot, p.seq);
  if(e != hipSuccess)
    return fail(HipToVk(e));
  const uint64_t *h_out = (const uint64_t *)h_pinned + 16 * p.slot;
  const uint32_t total = (uint32_t)h_out[0];
  if(total <= p.pass_cap)
  {
    PassDone(h_out, total, p.sig, p.cfg, false, p.ndraws, p.tri_total, p.tile_total, *p.st);
    open[p.sub]--;
    return VK_SUCCESS;
  }
  // Its pairs overflowed: its raster wrote nothing, and every pass launched behind it saw the
  // overflow flag and wrote nothing either. Once those have ended, the flag is cleared, the
  // pass's shape gets the capacity it needed, and all of them run again in launch order.
  std::deque<PassLaunch> redo;
  redo.push_back(std::move(p));
  while(!inflight.empty())
  {
    redo.push_back(std::move(inflight.front()));
    inflight.pop_front();
  }
  e = hipStreamSynchronize(stream);
  if(e == hipSuccess)
    e = va_fill_u32_async((uint32_t *)d_stamps + 34, 1, 0u, stream);
  if(total != 0xffffffffu)  // all ones: the pass only saw another's overflow flag
  {
    const PassLaunch &f = redo.front();
    const uint32_t need = (uint32_t)std::min<uint64_t>((uint64_t)total + total / 4 + 4096, 0xffff0000u);
    pair_cap = std::max(pair_cap, need);
    shape_caps[f.sig * 8u + f.cfg.tile_shift] = need;
  }
  VkResult r = e == hipSuccess ? VK_SUCCESS : HipToVk(e);
  for(PassLaunch &q : redo)
  {
    if(r == VK_SUCCESS)
      r = RunPass(q.passes, q.fused_clear, *q.st, nullptr);
    if(r != VK_SUCCESS && sres[q.sub] == VK_SUCCESS)
      sres[q.sub] = r;
    open[q.sub]--;
  }
  return r;
}

VkResult Queue::Execute(std::vector<Submission> &subs)
{
  std::lock_guard<std::mutex> lk(submit_m);
  if(!stream)
    HIPCHK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  VkResult res = VK_SUCCESS;
  // Asynchronous batches run consecutive batched passes back to back: each pass is launched
  // before the one ahead of it has been waited for, and completed (results read, an overflow
  // handled) while the next one runs. A submission made of such passes gets its stats and its
  // fence once all of them are complete; every other submission first waits for them.
  const bool pipelined = async && mode == ExecMode::Batched;
  std::vector<ExecStats> stats(subs.size());
  std::vector<int> open(subs.size(), 0);          // passes launched and not complete, per submission
  std::vector<VkResult> sres(subs.size(), VK_SUCCESS);
  std::vector<char> launched(subs.size(), 0);     // every operation of the submission launched
  size_t next_final = 0;
  auto finalize = [&](size_t i) {
    last_stats = stats[i];
    AccumulateStats(accum_stats, stats[i]);
    AccumulateStats(frame_stats, stats[i]);
    SignalFence(subs[i].fence, sres[i]);
  };
  // completes passes in flight (all of them, or until one is left) and finalizes, in order, the
  // submissions whose passes are all complete
  auto settle = [&](bool all) {
    VkResult r = VK_SUCCESS;
    while(!inflight.empty() && (all || inflight.size() > 1))
    {
      VkResult e = CompleteOldest(open, sres);
      if(r == VK_SUCCESS)
        r = e;
    }
    while(next_final < subs.size() && launched[next_final] && open[next_final] == 0)
      finalize(next_final++);
    return r;
  };
  for(size_t i = 0; i < subs.size(); i++)
  {
    Submission &sub = subs[i];
    ExecStats &st = stats[i];
    bool passes_only = pipelined && sub.downloads.empty();
    for(const Op &op : sub.ops)
      passes_only = passes_only && op.kind == Op::Pass;
    if(passes_only)
    {
      for(VkDeviceMemory_T *m : sub.uploads)
      {
        hipError_t e = res == VK_SUCCESS ? hipMemcpyAsync(m->dev, m->host, m->size, hipMemcpyHostToDevice, stream)
                                         : hipSuccess;
        if(e != hipSuccess)
          res = HipToVk(e);
      }
      for(Op &op : sub.ops)
      {
        if(res != VK_SUCCESS)
          break;
        PassLaunch pl;
        pl.sub = i;
        res = RunPass({&op}, true, st, &pl);
        if(res == VK_SUCCESS && pl.drain_first)
        {
          res = settle(true);
          pl = PassLaunch();
          pl.sub = i;
          if(res == VK_SUCCESS)
            res = RunPass({&op}, true, st, &pl);
        }
        if(res == VK_SUCCESS && pl.launched)
        {
          open[i]++;
          inflight.push_back(std::move(pl));
          // one pass runs, at most one more waits behind it
          VkResult e = settle(false);
          if(res == VK_SUCCESS)
            res = e;
        }
      }
      if(sres[i] == VK_SUCCESS)
        sres[i] = res;
      launched[i] = 1;
      VkResult e = settle(false);
      if(res == VK_SUCCESS)
        res = e;
      continue;
    }
    {
      VkResult e = settle(true);
      if(res == VK_SUCCESS)
        res = e;
    }
    for(VkDeviceMemory_T *m : sub.uploads)
      HIPCHK(hipMemcpyAsync(m->dev, m->host, m->size, hipMemcpyHostToDevice, stream));
    drained = false;
    for(Op &op : sub.ops)
    {
      if(res != VK_SUCCESS)
        break;
      drained = false;  // RunPass sets it again once its work has completed
      if(op.kind == Op::Copy)
        res = RunCopy(op.copy);
      else if(mode == ExecMode::Batched)
        res = RunPass({&op}, true, st);
      else
      {
        res = ClearFrame(op.frame);
        for(size_t k = 0; k < op.draws.size() && res == VK_SUCCESS; k++)
        {
          Op one;
          one.kind = Op::Pass;
          one.frame = op.frame;
          one.draws.assign(1, op.draws[k]);
          res = RunPass({&one}, false, st);
        }
      }
    }
    // one synchronization per submission: none more when the last operation was a pass that
    // has already waited for its work and nothing needs copying back
    bool need_sync = !drained;
    for(VkDeviceMemory_T *m : sub.downloads)
    {
      HIPCHK(hipMemcpyAsync(m->host, m->dev, m->size, hipMemcpyDeviceToHost, stream));
      need_sync = true;
    }
    if(need_sync)
    {
      hipError_t e = hipStreamSynchronize(stream);
      if(e != hipSuccess && res == VK_SUCCESS)
        res = HipToVk(e);
    }
    sres[i] = res;
    launched[i] = 1;
    finalize(next_final++);
  }
  VkResult e = settle(true);
  if(res == VK_SUCCESS)
    res = e;
  return res;
}

void Queue::WorkerLoop()
{
  for(;;)
  {
    std::vector<Submission> batch;
    {
      std::unique_lock<std::mutex> lk(qm);
      qcv.wait(lk, [&] { return stop || !pending.empty(); });
      if(stop && pending.empty())
        return;
      while(!pending.empty())
      {
        batch.push_back(std::move(pending.front()));
        pending.pop_front();
      }
      busy = true;
    }
    Execute(batch);
    {
      std::lock_guard<std::mutex> lk(qm);
      busy = false;
    }
    idle_cv.notify_all();
  }
}

VkResult Queue::Submit(uint32_t count, const VkSubmitInfo *infos, VkFence fence)
{
  if(!Device::Get().ok())
    return VK_ERROR_DEVICE_LOST;
  std::vector<Submission> subs;
  // a synchronous submit of one command buffer keeps its decoded form for the next submit of
  // the same recording (no decode, and RunPass reuses the pass tables built from it)
  const bool single = !async && count == 1 && infos[0].commandBufferCount == 1;
  for(uint32_t i = 0; i < count; i++)
    for(uint32_t c = 0; c < infos[i].commandBufferCount; c++)
    {
      Submission s;
      std::string err;
      const VkCommandBuffer_T *cbp = infos[i].pCommandBuffers[c];
      bool reused = false;
      if(single && cbp->recording != 0)
      {
        std::lock_guard<std::mutex> lk(decoded_m);
        if(decoded_valid && decoded.cb == cbp && decoded.recording == cbp->recording &&
           decoded.desc_writes == g_descriptor_writes.load())
        {
          s = std::move(decoded);
          decoded_valid = false;
          reused = true;
        }
      }
      if(!reused && !Decode(cbp, s, err))
      {
        LogError("vkQueueSubmit: %s", err.c_str());
        return VK_ERROR_FEATURE_NOT_PRESENT;
      }
      s.t_submit = NowSeconds();
      subs.push_back(std::move(s));
    }
  if(fence)
  {
    std::lock_guard<std::mutex> lk(fence->m);
    fence->signaled = false;
  }
  if(subs.empty())
  {
    SignalFence(fence, VK_SUCCESS);
    return VK_SUCCESS;
  }
  subs.back().fence = fence;
  if(!async)
  {
    // synchronous, as the reference's vkQueueSubmit (cmd_exec.cpp:187-201)
    VkResult r = Execute(subs);
    if(single && r == VK_SUCCESS && subs[0].recording != 0)
    {
      std::lock_guard<std::mutex> lk(decoded_m);
      decoded = std::move(subs[0]);
      decoded.fence = nullptr;
      decoded_valid = true;
    }
    return r;
  }
  {
    std::lock_guard<std::mutex> lk(qm);
    if(!worker.joinable())
      worker = std::thread([this] { WorkerLoop(); });
    for(auto &s : subs)
      pending.push_back(std::move(s));
  }
  qcv.notify_all();
  return VK_SUCCESS;
}

VkResult Queue::WaitIdle()
{
  std::unique_lock<std::mutex> lk(qm);
  idle_cv.wait(lk, [&] { return pending.empty() && !busy; });
  return VK_SUCCESS;
}
}  // namespace visor
