/*
 * visor_icd.h - the C-ABI exported by libvisor_mi355x.so.
 *
 * The drop-in boundary is the reference's Vulkan ICD interface: the loader negotiates a version,
 * then resolves every Vulkan entry point by name. The same three loader symbols are exported here
 * (reference icd_interface.cpp:132-169), and the name tables cover every entry point of the
 * reference's tables (icd_interface.cpp:7-102). Unknown names resolve to NULL.
 *
 * The vh_* functions below are runtime controls with no counterpart in the reference.
 */
#ifndef VISOR_ICD_H
#define VISOR_ICD_H

#include "visor_vk.h"

#ifdef __cplusplus
extern "C" {
#endif

/* reference icd_interface.cpp:161-169 (vk_icdNegotiateLoaderICDInterfaceVersion):
 * returns min(4, *pSupportedVersion). */
VkResult vk_icdNegotiateLoaderICDInterfaceVersion(uint32_t *pSupportedVersion);

/* reference icd_interface.cpp:132-148 (vk_icdGetInstanceProcAddr): instance, physical-device and
 * device entry points by name; NULL for unknown names. */
PFN_vkVoidFunction vk_icdGetInstanceProcAddr(VkInstance instance, const char *pName);

/* reference icd_interface.cpp:150-159 (vk_icdGetPhysicalDeviceProcAddr) */
PFN_vkVoidFunction vk_icdGetPhysicalDeviceProcAddr(VkInstance instance, const char *pName);

/* reference icd_interface.cpp:115-123 (vkGetDeviceProcAddr) */
PFN_vkVoidFunction vkGetDeviceProcAddr(VkDevice device, const char *pName);

/* The hot entry points, also exported directly (reference cmd_record.cpp:114-135 and
 * cmd_exec.cpp:187-201). Every other entry point is reached through the name tables. */
void vkCmdDraw(VkCommandBuffer commandBuffer, uint32_t vertexCount, uint32_t instanceCount,
               uint32_t firstVertex, uint32_t firstInstance);
void vkCmdDrawIndexed(VkCommandBuffer commandBuffer, uint32_t indexCount, uint32_t instanceCount,
                      uint32_t firstIndex, int32_t vertexOffset, uint32_t firstInstance);
VkResult vkQueueSubmit(VkQueue queue, uint32_t submitCount, const VkSubmitInfo *pSubmits,
                       VkFence fence);

/* ---- runtime controls (no counterpart in the reference) ---------------------------------- */

typedef struct VhStats {
  double gpu_ms;            /* device time of the submission's work (HIP events)          */
  double raster_ms;         /* device time of the raster kernel(s) (device clock stamps)  */
  double setup_ms;          /* vertex program + triangle setup + count scan               */
  double bin_ms;            /* pair emit + stable sort + bin ranges                        */
  uint64_t raster_launches;
  uint64_t triangles, bin_pairs, tiles;
  uint64_t draws, tris_in, tris_out, blocks;
  uint64_t pixels_tested;   /* reference counter rasterizer/pixels/tested                  */
  uint64_t pixels_written;  /* covered fragments (rasterizer/pixels/written)               */
  uint64_t depth_passed;    /* shaded fragments (rasterizer/depth/passed)                  */
  uint64_t shaded_blend;    /* shaded fragments of blending draws                          */
  uint64_t shaded_dwrite;   /* shaded fragments that wrote depth                           */
  uint64_t graph_launches;  /* batched passes replayed from a captured HIP graph             */
  uint64_t tile_size;       /* launch shape of the last pass: tile side (8, 16, 32) ...      */
  uint64_t chunk;           /* ... bin entries staged in LDS per round ...                  */
  uint64_t persistent;      /* ... and 1 when the raster grid was the persistent form       */
  uint64_t tuning;          /* 1 when the last pass ran one of the autotuner's trial shapes */
} VhStats;

/* 1 when a HIP device is usable. */
int vh_device_available(void);
/* Text of the last error the driver reported. */
const char *vh_last_error(void);
/* 0 = plain per-draw execution (host-synchronised per draw), 1 = batched (default). */
int vh_set_exec_mode(VkQueue queue, int mode);
/* 1 (default) = a batched pass's device work is captured once into a HIP graph and replayed
 * while its shape (sizes, buffers, pipelines) repeats; 0 = launched kernel by kernel. */
int vh_set_graph_mode(VkQueue queue, int on);
/* Screen tile side for binning and rasterization: 8, 16 or 32 fixes the launch shape; 0 (the
 * default) lets the autotuner pick tile size and LDS chunk per pass shape, by timing each
 * candidate on the pass's first repetitions (every candidate renders identical bytes). */
int vh_set_tile_size(VkQueue queue, int tile);
/* 0 = vkQueueSubmit returns after the GPU finished (the reference's behaviour), 1 = returns at
 * once; completion through fences / vkQueueWaitIdle. */
int vh_set_async_submit(VkQueue queue, int on);
/* which = 0: last submission; 1: accumulated since vh_reset_stats; 2: the frame the last
 * vkQueuePresentKHR ended (submissions between the two latest presents); 3: the current frame
 * (submissions since the last present). A present ends and begins frame statistics as the
 * reference's does (wsi.cpp:215-217, stats.cpp:25-41). */
int vh_get_stats(VkQueue queue, int which, VhStats *out);
/* Clears the accumulated (1) and current-frame (3) statistics. */
int vh_reset_stats(VkQueue queue);
/* The headless present's output (the reference blits the presented image to a window,
 * wsi.cpp:201-213): index of the image the last vkQueuePresentKHR showed, the number of
 * presents, and the host address of that image's bytes (NULL before the first present). */
int vh_swapchain_frame(VkSwapchainKHR swapchain, uint32_t *image_index, uint64_t *presents, const void **pixels);
/* SPIR-V -> HIP C of one entry point "main" (stage 0 vertex, 4 fragment); returns the length
 * written, or -1 with the error text in out. Needs no GPU. */
int vh_translate_spirv(const uint32_t *code, size_t nwords, int stage, char *out, size_t out_size);
/* Builds (and caches) the code object of a vertex/fragment pair for arch without a GPU; returns
 * the code object size or -1 with the error text. */
int vh_build_pipeline_offline(const uint32_t *vs, size_t vs_words, const uint32_t *fs,
                              size_t fs_words, const char *arch, char *err_out, size_t err_size);

/* ---- serving (no counterpart in the reference) -------------------------------------------
 * Many clients hand recorded command buffers to one server. Pending requests are admitted in
 * arrival order while the device bytes they touch (targets, textures, triangle records and
 * bins) fit the HBM budget beside the requests already on the GPU; admitted requests' render
 * passes, whatever their target sizes, share one setup / bin / raster launch sequence; up to
 * `slots` batches run at once on separate streams. A request completes when its host-visible
 * targets hold the result (submit -> present). Requests writing the same memory never share a
 * batch or overlap, so every result equals the one the request gives alone. */
typedef struct VhServer_T *VhServer;
typedef struct VhServerConfig {
  uint64_t hbm_budget_bytes;  /* admission limit (0: 16 GiB)                                */
  uint32_t slots;             /* batches on the GPU at once, 1..4                           */
  uint32_t max_batch;         /* requests merged into one launch sequence                   */
} VhServerConfig;
typedef struct VhServerStats {
  uint64_t submitted, completed, batches;
  uint64_t fragments;         /* covered fragments of completed requests                    */
  uint64_t shaded;            /* shaded fragments of completed requests                     */
  uint64_t max_inflight_bytes;
  double busy_ms;             /* wall time with at least one batch executing               */
  double mean_batch;          /* requests per batch                                         */
  double p50_ms, p99_ms, mean_ms, max_ms;  /* submit -> present latency                     */
} VhServerStats;
VkResult vh_server_create(VkDevice device, const VhServerConfig *cfg, VhServer *out);
void vh_server_destroy(VhServer server);
/* Queues a request; returns at once with its ticket. */
VkResult vh_server_submit(VhServer server, VkCommandBuffer cb, uint64_t *ticket);
/* Waits for a request (timeout in ns, UINT64_MAX = forever); latency_ms may be NULL. */
VkResult vh_server_wait(VhServer server, uint64_t ticket, uint64_t timeout_ns, double *latency_ms);
int vh_server_stats(VhServer server, VhServerStats *out);
int vh_server_reset_stats(VhServer server);

#ifdef __cplusplus
}
#endif

#endif /* VISOR_ICD_H */
