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
  double raster_ms;         /* device time of the raster kernel(s)                        */
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
} VhStats;

/* 1 when a HIP device is usable. */
int vh_device_available(void);
/* Text of the last error the driver reported. */
const char *vh_last_error(void);
/* 0 = plain per-draw execution (host-synchronised per draw), 1 = batched (default). */
int vh_set_exec_mode(VkQueue queue, int mode);
/* Screen tile side for binning and rasterization: 8 (default), 16 or 32. */
int vh_set_tile_size(VkQueue queue, int tile);
/* 0 = vkQueueSubmit returns after the GPU finished (the reference's behaviour), 1 = returns at
 * once; completion through fences / vkQueueWaitIdle. */
int vh_set_async_submit(VkQueue queue, int on);
/* which = 0: last submission; 1: accumulated since vh_reset_stats. */
int vh_get_stats(VkQueue queue, int which, VhStats *out);
int vh_reset_stats(VkQueue queue);
/* SPIR-V -> HIP C of one entry point "main" (stage 0 vertex, 4 fragment); returns the length
 * written, or -1 with the error text in out. Needs no GPU. */
int vh_translate_spirv(const uint32_t *code, size_t nwords, int stage, char *out, size_t out_size);
/* Builds (and caches) the code object of a vertex/fragment pair for arch without a GPU; returns
 * the code object size or -1 with the error text. */
int vh_build_pipeline_offline(const uint32_t *vs, size_t vs_words, const uint32_t *fs,
                              size_t fs_words, const char *arch, char *err_out, size_t err_size);

#ifdef __cplusplus
}
#endif

#endif /* VISOR_ICD_H */
