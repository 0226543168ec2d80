// internal.h - the ICD's object model and runtime (host side).
//
// Object model follows the reference (precompiled.h:22-203): dispatchable objects start with the
// loader magic word; non-dispatchable ones are plain structs created with new. Memory objects are
// backed by HBM; host-visible ones also keep a pinned host mirror that the application maps.
#pragma once

#include <hip/hip_runtime_api.h>

#include <atomic>
#include <condition_variable>
#include <cstdint>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/visor_vk.h"
#include "device_abi.h"
#include "spirv_hip.h"

namespace visor
{
void SetLoaderMagic(void *obj);
const char *HipErrorName(hipError_t e);
VkResult HipToVk(hipError_t e);
void LogError(const char *fmt, ...);
uint16_t HalfBits(float f);  // binary32 -> binary16, round to nearest even

// ---------------------------------------------------------------------------------------------
// command stream: {uint16 id}{POD}, as the reference's VkCommandBuffer_T::push (precompiled.h:63-69)
enum class Cmd : uint16_t
{
  PipelineBarrier,
  BeginRenderPass,
  EndRenderPass,
  BindPipeline,
  BindDescriptorSets,
  BindVB,
  BindIB,
  SetViewport,
  SetScissors,
  PushConstants,
  Draw,
  DrawIndexed,
  CopyBuf2Img,
  CopyBuf,
};

namespace cmd
{
struct BeginRenderPass { VkRenderPass renderPass; VkFramebuffer framebuffer; VkClearValue clearval[8]; };
struct BindPipeline { VkPipeline pipeline; };
struct BindDescriptorSets { uint32_t idx; VkDescriptorSet set; };
struct BindVB { uint32_t slot; VkBuffer buffer; VkDeviceSize offset; };
struct BindIB { VkBuffer buffer; VkDeviceSize offset; VkIndexType indexType; };
struct SetViewport { VkViewport view; };
struct PushConstants { uint32_t offset, size; uint8_t values[128]; };
struct Draw { uint32_t vertexCount, instanceCount, firstVertex, firstInstance; };
struct DrawIndexed { uint32_t indexCount, instanceCount, firstIndex, firstInstance; int32_t vertexOffset; };
struct CopyBuf2Img { VkBuffer srcBuffer; VkImage dstImage; VkBufferImageCopy region; };
struct CopyBuf { VkBuffer srcBuffer; VkBuffer dstBuffer; VkBufferCopy region; };
}  // namespace cmd

class Runtime;
class Queue;
struct PipelineModule;
struct ShaderPair;
}  // namespace visor

struct VkInstance_T
{
  uintptr_t loaderMagic;
};

struct VkPhysicalDevice_T
{
  uintptr_t loaderMagic;
};

struct VkQueue_T
{
  uintptr_t loaderMagic;
  visor::Queue *q;
};

struct VkDevice_T
{
  uintptr_t loaderMagic;
  std::vector<VkQueue_T *> queues;
};

struct VkDeviceMemory_T
{
  VkDeviceSize size = 0;
  uint32_t type = 0;          // 0 DEVICE_LOCAL, 1 HOST_VISIBLE|COHERENT|CACHED
  uint8_t *host = nullptr;    // pinned mirror (type 1)
  uint8_t *dev = nullptr;     // HBM
};

struct VkBuffer_T
{
  VkDeviceSize size = 0;
  VkDeviceMemory_T *mem = nullptr;
  VkDeviceSize offset = 0;
};

struct VkImage_T
{
  VkExtent3D extent = {0, 0, 0};
  VkImageType imageType = VK_IMAGE_TYPE_2D;
  VkFormat format = VK_FORMAT_UNDEFINED;
  uint32_t arrayLayers = 1;
  uint32_t mipLevels = 1;
  uint32_t bytesPerPixel = 4;
  VkDeviceMemory_T *mem = nullptr;
  VkDeviceSize offset = 0;
  // RGBA8 copy that the samplers read when the texels are not 4-byte RGBA (BC2/BC3, R8,
  // RGBA16F): refreshed by a transcode launch before every pass that samples the image
  uint8_t *shadow = nullptr;
  uint8_t *devPixels() const { return mem ? mem->dev + offset : nullptr; }
};

struct VkImageView_T
{
  VkImage_T *image = nullptr;
};

struct VkSampler_T
{
  VkSamplerCreateInfo info;
};

struct VkShaderModule_T
{
  std::vector<uint32_t> code;
};

struct VkPipeline_T
{
  struct
  {
    VkFormat format;
    uint32_t stride;
    uint32_t offset;
    uint32_t vb;
  } vattrs[16];
  VkPrimitiveTopology topology = VK_PRIMITIVE_TOPOLOGY_MAX_ENUM;
  VkFrontFace frontFace = VK_FRONT_FACE_COUNTER_CLOCKWISE;
  VkCullModeFlags cullMode = VK_CULL_MODE_NONE;
  VkCompareOp depthCompareOp = VK_COMPARE_OP_ALWAYS;
  bool depthWriteEnable = false;
  VkPipelineColorBlendAttachmentState blend;
  std::shared_ptr<visor::ShaderPair> pair;
  bool graphics = true;
};

struct VkDescriptorSetLayout_T
{
  uint32_t bindingCount = 0;
};

struct VkDescriptorSet_T
{
  struct Bind
  {
    VkDescriptorType type = VK_DESCRIPTOR_TYPE_MAX_ENUM;
    VkDescriptorImageInfo imageInfo;
    VkDescriptorBufferInfo bufferInfo;
  };
  std::vector<Bind> binds;
};

struct VkRenderPass_T
{
  struct Attachment
  {
    int32_t idx = -1;
    bool clear = false;
  };
  struct Subpass
  {
    std::vector<Attachment> colAttachments;
    Attachment depthAttachment;
  };
  std::vector<Subpass> subpasses;
};

struct VkFramebuffer_T
{
  std::vector<VkImageView> attachments;
};

struct VkCommandBuffer_T
{
  uintptr_t loaderMagic;
  bool live = false;
  uint64_t recording = 0;  // unique per vkBeginCommandBuffer (never reused)
  std::vector<uint8_t> stream;

  template <typename T>
  T *push(visor::Cmd id)
  {
    size_t at = stream.size();
    stream.resize(at + sizeof(uint16_t) + sizeof(T));
    memcpy(&stream[at], &id, sizeof(uint16_t));
    T *ret = (T *)&stream[at + sizeof(uint16_t)];
    memset((void *)ret, 0, sizeof(T));
    return ret;
  }
  void pushId(visor::Cmd id)
  {
    size_t at = stream.size();
    stream.resize(at + sizeof(uint16_t));
    memcpy(&stream[at], &id, sizeof(uint16_t));
  }
};

struct VkCommandPool_T
{
  std::vector<VkCommandBuffer> buffers;
};

struct VkFence_T
{
  std::mutex m;
  std::condition_variable cv;
  bool signaled = false;
  VkResult status = VK_SUCCESS;
};

struct VkSemaphore_T
{
  int dummy = 0;
};

struct VkSwapchainKHR_T
{
  VkExtent2D extent = {0, 0};
  VkFormat format = VK_FORMAT_B8G8R8A8_UNORM;
  std::vector<VkImage> images;
  std::vector<VkDeviceMemory> mems;
  uint32_t current = 0;       // last acquired image (reference wsi.cpp:185-195 advances, then returns it)
  uint32_t presented = 0;     // image shown by the last vkQueuePresentKHR
  uint64_t presents = 0;
};

namespace visor
{
// ---------------------------------------------------------------------------------------------
// Shader pairs and pipeline modules (hiprtc-built)

struct ShaderPair
{
  std::string key;  // hash of both SPIR-V modules and entry names
  TranslatedShader vs, fs;
  uint32_t nv = 0;                 // varying components
  std::vector<uint32_t> vary_src;  // per varying: index into the VS output array (loc*4+comp)
  std::vector<ShaderResource> resources;
};

struct PipelineModule
{
  std::string key;
  hipModule_t module = nullptr;
  hipFunction_t setup = nullptr, emit = nullptr, ranges = nullptr;
  hipFunction_t raster[3] = {nullptr, nullptr, nullptr};   // tile 8, 16, 32
  hipFunction_t raster_p[3] = {nullptr, nullptr, nullptr};  // persistent forms
  uint32_t raster_p_grid[3] = {0, 0, 0};                     // workgroups that fill the GPU
  uint32_t rec_stride = 0;  // words
  std::vector<std::string> slot_keys;
  double compile_seconds = 0.0;
};

// ---------------------------------------------------------------------------------------------
// Executor: decoded submissions and the device work

struct FrameDesc
{
  VkImage_T *color = nullptr;
  VkImage_T *depth = nullptr;
  bool clear_color = false, clear_depth = false;
  VkClearValue clear_col{}, clear_dep{};
};

struct DrawDesc
{
  uint32_t frame;          // index into the pass list
  VaDraw d;                // everything but tri_base / frame / slot
  VkPipeline_T *pipeline;
  VkImage_T *shadowed[VA_MAX_TEX] = {};  // sampled images read through their RGBA8 shadow
};

struct CopyDesc
{
  bool to_image;
  VkBuffer_T *src;
  VkBuffer_T *dstBuf;
  VkImage_T *dstImg;
  VkBufferImageCopy bi;
  VkBufferCopy bb;
};

struct Op
{
  enum Kind { Pass, Copy } kind;
  FrameDesc frame;
  std::vector<DrawDesc> draws;
  CopyDesc copy;
  // (recording << 16 | index) for a pass decoded from a command buffer recording, so the
  // tables built from it can be reused when the same recording is submitted again; 0 = none.
  // desc_writes: the descriptor-write count at decode time (descriptors feed the tables).
  uint64_t key = 0, desc_writes = 0;
};

// Count of vkUpdateDescriptorSets calls: decoded submissions and pass tables cached from a
// recording are reused only while it is unchanged.
extern std::atomic<uint64_t> g_descriptor_writes;

struct Submission
{
  const VkCommandBuffer_T *cb = nullptr;  // the command buffer and recording decoded into ops
  uint64_t recording = 0, desc_writes = 0;
  std::vector<Op> ops;
  std::vector<VkDeviceMemory_T *> uploads;    // host-visible memory read by the GPU
  std::vector<VkDeviceMemory_T *> downloads;  // host-visible memory written by the GPU
  VkFence fence = nullptr;
  double t_submit = 0.0;
};

// Launch shape of the raster stage. tile_shift selects the kernel at launch; chunk (bin
// entries staged in LDS per round, 16..64) and small (largest bbox-in-block overlap tested
// triangle-parallel; 64, the whole 8x8 block, tests every triangle that way) are compile-time
// parameters of the pipeline module.
struct TileConfig
{
  uint32_t tile_shift = 3;  // 3, 4, 5 -> 8, 16, 32 pixel tiles
  uint32_t chunk = 64;
  uint32_t small = 64;
  uint32_t persistent = 0;  // 1: one GPU-filling grid taking tiles from a counter
  bool operator==(const TileConfig &o) const
  {
    return tile_shift == o.tile_shift && chunk == o.chunk && small == o.small && persistent == o.persistent;
  }
};

struct ExecStats
{
  double gpu_ms_total = 0.0;    // device time of the submission's passes (setup start to finish start)
  double raster_ms = 0.0;       // raster start to finish start, device clock stamps
  double setup_ms = 0.0;
  double bin_ms = 0.0;
  uint64_t raster_launches = 0;
  uint64_t total_tris = 0, total_pairs = 0, total_tiles = 0;
  uint64_t graph_launches = 0;  // passes replayed from a captured graph
  TileConfig last_cfg;          // launch shape of the last pass
  bool last_tuning = false;     // the last pass ran one of the autotuner's trial shapes
  VaCounters ctr{};
};

// acc += st: counters and times summed, launch shape and tuning flag of the latest
void AccumulateStats(ExecStats &acc, const ExecStats &st);

enum class ExecMode
{
  Plain = 0,    // per draw: setup, bin, raster, host sync (the reference's blocking DrawTriangles)
  Batched = 1,  // every draw of a render pass in one setup/bin/raster sequence, fused clears
};


class Device
{
public:
  static Device &Get();
  bool Init(std::string &err);
  bool ok() const { return initialized; }
  int device_id = 0;
  hipDeviceProp_t props{};
  int wall_khz = 100000;  // device wall clock rate (wall_clock64 ticks per millisecond)
  std::string arch = "gfx950";

  std::shared_ptr<ShaderPair> MakePair(VkShaderModule_T *vs, const char *vs_entry, VkShaderModule_T *fs,
                                       const char *fs_entry, std::string &err);
  std::shared_ptr<PipelineModule> GetModule(const std::vector<std::shared_ptr<ShaderPair>> &pairs,
                                            std::string &err, const TileConfig &cfg = TileConfig());
  std::string cache_dir;

private:
  bool initialized = false;
  std::mutex m;
  std::map<std::string, std::shared_ptr<ShaderPair>> pairs;
  std::map<std::string, std::shared_ptr<PipelineModule>> modules;
};

class Queue
{
public:
  Queue();
  ~Queue();
  VkResult Submit(uint32_t count, const VkSubmitInfo *infos, VkFence fence);
  VkResult WaitIdle();
  // several submissions' render passes in one batched launch sequence (serving)
  VkResult ExecuteBatch(std::vector<Submission *> &subs, ExecStats &st);
  ExecStats last_stats;
  ExecStats accum_stats;
  ExecStats frame_stats;      // submissions since the last vkQueuePresentKHR
  ExecStats presented_stats;  // submissions of the frame the last present ended
  ExecMode mode = ExecMode::Batched;
  TileConfig tiles;
  bool async = false;
  bool graphs = true;  // batched passes run as a captured, replayed HIP graph
  bool autotune = true;  // batched passes pick their launch shape per pass signature
  bool Decode(const VkCommandBuffer_T *cb, Submission &sub, std::string &err);

private:
  VkResult Execute(std::vector<Submission> &subs);
  // A batched pass launched without waiting for its end: asynchronous batches queue the next
  // pass behind it and complete it afterwards, so the device does not idle between passes.
  struct PassLaunch
  {
    bool launched = false;     // set by RunPass when it returned without waiting
    bool drain_first = false;  // set by RunPass when the pass must wait for the ones in flight
    std::vector<Op *> passes;
    bool fused_clear = false;
    ExecStats *st = nullptr;
    size_t sub = 0;  // submission of the batch it belongs to
    uint64_t sig = 0;
    TileConfig cfg;
    uint32_t pass_cap = 0, ndraws = 0, tri_total = 0, tile_total = 0;
    int slot = 0;      // output words of the launch in h_pinned
    uint32_t seq = 0;  // the number its finish kernel publishes in the slot's word 14
  };
  // deferred == nullptr: the pass completes before RunPass returns; otherwise RunPass may
  // return right after the launch (deferred->launched) and CompleteOldest finishes it
  VkResult RunPass(const std::vector<Op *> &passes, bool fused_clear, ExecStats &st,
                   PassLaunch *deferred = nullptr);
  void PassDone(const uint64_t *h_out, uint32_t total_pairs, uint64_t sig, const TileConfig &cfg, bool tuning,
                uint32_t ndraws, uint32_t tri_total, uint32_t tile_total, ExecStats &st);
  // waits for the oldest launched pass and takes its results; a pass whose pairs overflowed is
  // run again after every pass queued behind it has ended, and those are run again in order
  VkResult CompleteOldest(std::vector<int> &open, std::vector<VkResult> &sres);
  std::deque<PassLaunch> inflight;  // launch order, at most two
  // passes launched on this queue since d_stamps was allocated (the finish kernel counts the same
  // on the device); WaitPass waits until slot `slot` holds sequence number seq
  uint32_t launch_seq = 0;
  hipError_t WaitPass(int slot, uint32_t seq);
  uint64_t PassSignature(const std::vector<Op *> &passes) const;
  TileConfig ChooseConfig(uint64_t sig, bool *tuning);
  void RecordTuning(uint64_t sig, const TileConfig &cfg, double ms);
  struct TuneState
  {
    std::vector<double> ms;  // best time per candidate
    size_t runs = 0;
    bool done = false;
    TileConfig best;
  };
  std::map<uint64_t, TuneState> tune;
  VkResult RunCopy(const CopyDesc &c);
  VkResult ClearFrame(const FrameDesc &f);
  bool Grow(void **p, size_t &cap, size_t need);
  void WorkerLoop();
  void SignalFence(VkFence f, VkResult r);

  hipStream_t stream = nullptr;
  std::mutex submit_m;  // one submission executes at a time
  // set by RunPass when every operation enqueued so far has completed
  bool drained = false;
  // growable device scratch
  // d_count: tiles per primitive; d_bpairs: pairs per setup block, then the pair total; d_start:
  // bin starts; d_blocks: setup block -> draw, then each block's first primitive
  void *d_frames = nullptr, *d_draws = nullptr, *d_rec = nullptr, *d_count = nullptr, *d_bpairs = nullptr;
  void *d_keys = nullptr, *d_vals = nullptr, *d_keys2 = nullptr, *d_vals2 = nullptr;
  void *d_start = nullptr, *d_blocks = nullptr;
  size_t c_blocks = 0;
  void *d_part = nullptr;  // counter rows of setup and raster (VA_COUNTER_ROWS each)
  size_t c_part = 0;
  void *d_stemp = nullptr;  // radix sort histograms
  size_t c_stemp = 0;
  void *d_einfo = nullptr;  // per-triangle pair-emission descriptors
  size_t c_einfo = 0;
  void *d_stamps = nullptr;  // device wall-clock stamps at stage boundaries (+ the tile counter)
  size_t c_stamps = 0;
  void *d_masks = nullptr;  // the raster's per-chunk pass masks (visibility pass -> shading pass)
  size_t c_masks = 0;
  uint32_t pair_cap = 0;    // (tile, primitive) pair capacity of the bin buffers
  std::unordered_map<uint64_t, uint32_t> shape_caps;  // launch capacity per pass shape
  // last uploaded per-pass tables (skipped when a submission repeats them)
  std::vector<VaFrame> up_frames;
  std::vector<VaDraw> up_draws;
  std::vector<uint32_t> up_blocks;
  void *up_frames_ptr = nullptr, *up_draws_ptr = nullptr, *up_blocks_ptr = nullptr;
  uint64_t up_version = 0;      // PassTables::version the device buffers hold (0 = compare)
  void *rows_zeroed = nullptr;    // d_part allocation whose counter rows were zeroed
  void *poison_zeroed = nullptr;  // d_stamps allocation whose overflow flag was zeroed
  uint64_t table_versions = 0;  // versions handed out to rebuilt reusable tables
  // device tables of one pass, built on the host
  struct PassTables
  {
    uint64_t key = 0, desc_writes = 0;  // Op::key / desc_writes they were built from
    uint64_t version = 0;               // changes whenever the tables are rebuilt
    uint32_t tile_shift = 0;
    bool fused_clear = false;
    uint64_t sig = 0;
    std::vector<VaFrame> frames;
    std::vector<VaDraw> draws;
    std::vector<std::shared_ptr<ShaderPair>> pairs;
    std::vector<uint32_t> block_draw;  // setup workgroup -> draw
    std::vector<uint32_t> block_tab;   // block_draw, then each block's first primitive, then the total
    uint32_t tile_total = 0, tri_total = 0;
  };
  PassTables tables_cache;  // the last reusable pass's tables
  // the last single-command-buffer submission, decoded; reused when the same recording is
  // submitted again (synchronous submits)
  std::mutex decoded_m;
  Submission decoded;
  bool decoded_valid = false;
  // the captured device sequence of the last batched pass, replayed while its key matches
  struct GraphKey
  {
    const void *mod;
    uint32_t dims[8];
    void *bufs[20];
  };
  struct GraphSet
  {
    GraphKey key;
    hipGraphExec_t exec[1] = {nullptr};  // setup, scan, binning, raster and finish
  };
  std::vector<GraphSet> graph_cache;  // most recent first, at most 8
  size_t c_frames = 0, c_draws = 0, c_rec = 0, c_count = 0, c_bpairs = 0, c_keys = 0, c_vals = 0, c_keys2 = 0,
         c_vals2 = 0, c_start = 0;
  uint32_t *h_pinned = nullptr;  // small pinned staging for readbacks

  // async worker
  std::thread worker;
  std::mutex qm;
  std::condition_variable qcv, idle_cv;
  std::deque<Submission> pending;
  bool stop = false;
  bool busy = false;
};

double NowSeconds();
}  // namespace visor
