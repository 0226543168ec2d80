// icd.cpp - the Vulkan ICD surface: loader negotiation, name tables, object creation, command
// recording and submission. Mirrors the reference's entry points (icd_interface.cpp:7-169,
// icd_stubs.cpp, query.cpp, memory.cpp, buffers.cpp, images.cpp, descriptors.cpp,
// renderpass.cpp, shaders.cpp, cmd_alloc.cpp, cmd_record.cpp, cmd_exec.cpp:187-201, wsi.cpp)
// with HBM-backed memory and GPU execution.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <thread>

#include "../../include/visor_icd.h"
#include "internal.h"

using namespace visor;

namespace visor
{
void SetLoaderMagic(void *obj) { *(uintptr_t *)obj = ICD_LOADER_MAGIC; }

uint16_t HalfBits(float f)
{
  uint32_t x;
  memcpy(&x, &f, 4);
  uint32_t sign = (x >> 16) & 0x8000u, ax = x & 0x7fffffffu;
  if(ax >= 0x7f800000u)
    return (uint16_t)(sign | 0x7c00u | (ax > 0x7f800000u ? (0x200u | ((ax >> 13) & 0x3ffu)) : 0u));
  if(ax >= 0x477ff000u)
    return (uint16_t)(sign | 0x7c00u);
  if(ax < 0x38800000u)
  {
    uint32_t e = ax >> 23;
    if(e < 102)
      return (uint16_t)sign;
    uint32_t m = (ax & 0x7fffffu) | 0x800000u, s = 126 - e, q = m >> s;
    uint32_t rem = m & ((1u << s) - 1u), half = 1u << (s - 1);
    if(rem > half || (rem == half && (q & 1u)))
      q++;
    return (uint16_t)(sign | q);
  }
  uint32_t r = ax - 0x38000000u;
  r = (r + 0xfffu + ((r >> 13) & 1u)) >> 13;
  return (uint16_t)(sign | r);
}
}  // namespace visor

static VkQueue_T *g_last_queue = nullptr;
static std::string g_last_error;

static void SetError(const std::string &e)
{
  g_last_error = e;
  LogError("%s", e.c_str());
}

extern "C" {

// ---------------------------------------------------------------------------------------------
// instance / device (reference icd_stubs.cpp:5-62, query.cpp)

VKAPI_ATTR VkResult VKAPI_CALL vkCreateInstance(const VkInstanceCreateInfo *, const VkAllocationCallbacks *,
                                                VkInstance *pInstance)
{
  VkInstance i = new VkInstance_T;
  SetLoaderMagic(i);
  *pInstance = i;
  return VK_SUCCESS;
}

VKAPI_ATTR void VKAPI_CALL vkDestroyInstance(VkInstance instance, const VkAllocationCallbacks *) { delete instance; }

VKAPI_ATTR VkResult VKAPI_CALL vkEnumerateInstanceVersion(uint32_t *pApiVersion)
{
  if(!pApiVersion)
    return VK_ERROR_INITIALIZATION_FAILED;
  *pApiVersion = VK_MAKE_VERSION(1, 0, 0);
  return VK_SUCCESS;
}

static VkResult FillExtensions(const VkExtensionProperties *src, uint32_t n, uint32_t *count, VkExtensionProperties *props)
{
  if(count && !props)
  {
    *count = n;
    return VK_SUCCESS;
  }
  if(count && props)
  {
    uint32_t space = *count;
    *count = std::min(n, space);
    memcpy(props, src, sizeof(VkExtensionProperties) * std::min(n, space));
    return space >= n ? VK_SUCCESS : VK_INCOMPLETE;
  }
  return VK_INCOMPLETE;
}

VKAPI_ATTR VkResult VKAPI_CALL vkEnumerateInstanceExtensionProperties(const char *, uint32_t *pPropertyCount,
                                                                      VkExtensionProperties *pProperties)
{
  VkExtensionProperties e[1];
  memset(e, 0, sizeof(e));
  strcpy(e[0].extensionName, "VK_KHR_surface");
  e[0].specVersion = 25;
  return FillExtensions(e, 1, pPropertyCount, pProperties);
}

VKAPI_ATTR VkResult VKAPI_CALL vkEnumerateDeviceExtensionProperties(VkPhysicalDevice, const char *,
                                                                    uint32_t *pPropertyCount,
                                                                    VkExtensionProperties *pProperties)
{
  VkExtensionProperties e[1];
  memset(e, 0, sizeof(e));
  strcpy(e[0].extensionName, "VK_KHR_swapchain");
  e[0].specVersion = 68;
  return FillExtensions(e, 1, pPropertyCount, pProperties);
}

static VkPhysicalDevice_T g_phys = {ICD_LOADER_MAGIC};

VKAPI_ATTR VkResult VKAPI_CALL vkEnumeratePhysicalDevices(VkInstance, uint32_t *pCount, VkPhysicalDevice *pDevices)
{
  if(pCount && !pDevices)
  {
    *pCount = 1;
    return VK_SUCCESS;
  }
  if(!pCount || *pCount < 1)
    return VK_INCOMPLETE;
  *pCount = 1;
  pDevices[0] = &g_phys;
  return VK_SUCCESS;
}

VKAPI_ATTR void VKAPI_CALL vkGetPhysicalDeviceFeatures(VkPhysicalDevice, VkPhysicalDeviceFeatures *f)
{
  memset(f, 0, sizeof(*f));
  f->fullDrawIndexUint32 = VK_TRUE;
}

VKAPI_ATTR void VKAPI_CALL vkGetPhysicalDeviceProperties(VkPhysicalDevice, VkPhysicalDeviceProperties *p)
{
  memset(p, 0, sizeof(*p));
  p->apiVersion = VK_MAKE_VERSION(1, 0, 47);
  p->driverVersion = VK_MAKE_VERSION(0, 1, 0);
  p->vendorID = 0x1002;
  p->deviceID = 0x75a3;
  p->deviceType = VK_PHYSICAL_DEVICE_TYPE_DISCRETE_GPU;
  snprintf(p->deviceName, sizeof(p->deviceName), "visor on %s", Device::Get().ok() ? Device::Get().arch.c_str() : "gfx950");
  VkPhysicalDeviceLimits &l = p->limits;
  l.maxImageDimension1D = l.maxImageDimension2D = 16384;
  l.maxImageDimension3D = 2048;
  l.maxImageDimensionCube = 16384;
  l.maxImageArrayLayers = 2048;
  l.maxUniformBufferRange = 65536;
  l.maxStorageBufferRange = 1u << 30;
  l.maxPushConstantsSize = 128;
  l.maxMemoryAllocationCount = 1u << 20;
  l.maxSamplerAllocationCount = 1u << 20;
  l.bufferImageGranularity = 1;
  l.maxBoundDescriptorSets = 1;
  l.maxPerStageDescriptorSamplers = 4;
  l.maxPerStageDescriptorUniformBuffers = 4;
  l.maxPerStageDescriptorSampledImages = 4;
  l.maxPerStageResources = 8;
  l.maxDescriptorSetSamplers = 4;
  l.maxDescriptorSetUniformBuffers = 4;
  l.maxDescriptorSetSampledImages = 4;
  l.maxVertexInputAttributes = 16;
  l.maxVertexInputBindings = 4;
  l.maxVertexInputAttributeOffset = 2047;
  l.maxVertexInputBindingStride = 2048;
  l.maxVertexOutputComponents = 40;
  l.maxFragmentInputComponents = 40;
  l.maxFragmentOutputAttachments = 1;
  l.maxFragmentCombinedOutputResources = 1;
  l.maxComputeSharedMemorySize = 0;
  l.subPixelPrecisionBits = 0;  // window coordinates are whole pixels (reference rasterizer.cpp:248-249)
  l.subTexelPrecisionBits = 8;
  l.mipmapPrecisionBits = 8;
  l.maxDrawIndexedIndexValue = 0xffffffffu;
  l.maxDrawIndirectCount = 1;
  l.maxSamplerLodBias = 16.0f;
  l.maxSamplerAnisotropy = 1.0f;
  l.maxViewports = 1;
  l.maxViewportDimensions[0] = l.maxViewportDimensions[1] = 16384;
  l.viewportBoundsRange[0] = -32768.0f;
  l.viewportBoundsRange[1] = 32767.0f;
  l.minMemoryMapAlignment = 64;
  l.minTexelBufferOffsetAlignment = l.minUniformBufferOffsetAlignment = l.minStorageBufferOffsetAlignment = 4;
  l.maxFramebufferWidth = l.maxFramebufferHeight = 16384;
  l.maxFramebufferLayers = 1;
  l.framebufferColorSampleCounts = l.framebufferDepthSampleCounts = 1;
  l.framebufferStencilSampleCounts = l.framebufferNoAttachmentsSampleCounts = 1;
  l.maxColorAttachments = 1;
  l.sampledImageColorSampleCounts = l.sampledImageIntegerSampleCounts = 1;
  l.sampledImageDepthSampleCounts = l.sampledImageStencilSampleCounts = l.storageImageSampleCounts = 1;
  l.maxSampleMaskWords = 1;
  l.timestampPeriod = 1;
  l.discreteQueuePriorities = 2;
  l.pointSizeRange[0] = l.pointSizeRange[1] = 1.0f;
  l.lineWidthRange[0] = l.lineWidthRange[1] = 1.0f;
  l.standardSampleLocations = VK_TRUE;
  l.optimalBufferCopyOffsetAlignment = l.optimalBufferCopyRowPitchAlignment = 1;
  l.nonCoherentAtomSize = 1;
}

VKAPI_ATTR void VKAPI_CALL vkGetPhysicalDeviceQueueFamilyProperties(VkPhysicalDevice, uint32_t *pCount,
                                                                    VkQueueFamilyProperties *pProps)
{
  if(pCount && !pProps)
  {
    *pCount = 1;
    return;
  }
  memset(pProps, 0, sizeof(*pProps));
  pProps->minImageTransferGranularity = {1, 1, 1};
  pProps->queueCount = 1;
  pProps->queueFlags = VK_QUEUE_GRAPHICS_BIT | VK_QUEUE_COMPUTE_BIT | VK_QUEUE_TRANSFER_BIT;
  pProps->timestampValidBits = 64;
  if(pCount)
    *pCount = 1;
}

// Two memory types as the reference (query.cpp:246-267): 0 = DEVICE_LOCAL in HBM, 1 =
// HOST_VISIBLE | COHERENT | CACHED, a pinned host mirror synchronised at submit and completion.
VKAPI_ATTR void VKAPI_CALL vkGetPhysicalDeviceMemoryProperties(VkPhysicalDevice, VkPhysicalDeviceMemoryProperties *p)
{
  memset(p, 0, sizeof(*p));
  uint64_t hbm = Device::Get().ok() ? (uint64_t)Device::Get().props.totalGlobalMem : (288ull << 30);
  p->memoryHeapCount = 2;
  p->memoryHeaps[0].flags = VK_MEMORY_HEAP_DEVICE_LOCAL_BIT;
  p->memoryHeaps[0].size = hbm;
  p->memoryHeaps[1].flags = 0;
  p->memoryHeaps[1].size = 64ull << 30;
  p->memoryTypeCount = 2;
  p->memoryTypes[0].propertyFlags = VK_MEMORY_PROPERTY_DEVICE_LOCAL_BIT;
  p->memoryTypes[0].heapIndex = 0;
  p->memoryTypes[1].propertyFlags =
      VK_MEMORY_PROPERTY_HOST_VISIBLE_BIT | VK_MEMORY_PROPERTY_HOST_COHERENT_BIT | VK_MEMORY_PROPERTY_HOST_CACHED_BIT;
  p->memoryTypes[1].heapIndex = 1;
}

VKAPI_ATTR void VKAPI_CALL vkGetPhysicalDeviceFormatProperties(VkPhysicalDevice, VkFormat format, VkFormatProperties *p)
{
  p->bufferFeatures = VK_FORMAT_FEATURE_UNIFORM_TEXEL_BUFFER_BIT | VK_FORMAT_FEATURE_VERTEX_BUFFER_BIT |
                      VK_FORMAT_FEATURE_TRANSFER_SRC_BIT_KHR | VK_FORMAT_FEATURE_TRANSFER_DST_BIT_KHR;
  p->linearTilingFeatures =
      VK_FORMAT_FEATURE_SAMPLED_IMAGE_BIT | VK_FORMAT_FEATURE_COLOR_ATTACHMENT_BIT | VK_FORMAT_FEATURE_COLOR_ATTACHMENT_BLEND_BIT;
  if(format == VK_FORMAT_D32_SFLOAT || format == VK_FORMAT_D32_SFLOAT_S8_UINT)
  {
    p->bufferFeatures = 0;
    p->linearTilingFeatures = VK_FORMAT_FEATURE_SAMPLED_IMAGE_BIT | VK_FORMAT_FEATURE_DEPTH_STENCIL_ATTACHMENT_BIT;
  }
  p->optimalTilingFeatures = p->linearTilingFeatures;
}

VKAPI_ATTR VkResult VKAPI_CALL vkCreateDevice(VkPhysicalDevice, const VkDeviceCreateInfo *info, const VkAllocationCallbacks *,
                                              VkDevice *pDevice)
{
  std::string err;
  if(!Device::Get().Init(err))
  {
    SetError("vkCreateDevice: " + err);
    return VK_ERROR_INITIALIZATION_FAILED;
  }
  if(!info || info->queueCreateInfoCount != 1)
    return VK_ERROR_INITIALIZATION_FAILED;
  VkDevice d = new VkDevice_T;
  SetLoaderMagic(d);
  for(uint32_t i = 0; i < info->pQueueCreateInfos[0].queueCount; i++)
  {
    VkQueue_T *q = new VkQueue_T;
    SetLoaderMagic(q);
    q->q = new Queue;
    d->queues.push_back(q);
  }
  *pDevice = d;
  return VK_SUCCESS;
}

VKAPI_ATTR void VKAPI_CALL vkDestroyDevice(VkDevice device, const VkAllocationCallbacks *)
{
  if(!device)
    return;
  for(VkQueue_T *q : device->queues)
  {
    if(g_last_queue == q)
      g_last_queue = nullptr;
    delete q->q;
    delete q;
  }
  delete device;
}

VKAPI_ATTR void VKAPI_CALL vkGetDeviceQueue(VkDevice device, uint32_t, uint32_t queueIndex, VkQueue *pQueue)
{
  *pQueue = queueIndex < device->queues.size() ? device->queues[queueIndex] : nullptr;
  g_last_queue = *pQueue;
}

// ---------------------------------------------------------------------------------------------
// synchronisation (fences really wait here; the reference's are no-ops, icd_stubs.cpp:64-155)

VKAPI_ATTR VkResult VKAPI_CALL vkCreateFence(VkDevice, const VkFenceCreateInfo *info, const VkAllocationCallbacks *,
                                             VkFence *pFence)
{
  VkFence f = new VkFence_T;
  f->signaled = info && (info->flags & VK_FENCE_CREATE_SIGNALED_BIT);
  *pFence = f;
  return VK_SUCCESS;
}

VKAPI_ATTR void VKAPI_CALL vkDestroyFence(VkDevice, VkFence fence, const VkAllocationCallbacks *) { delete fence; }

VKAPI_ATTR VkResult VKAPI_CALL vkResetFences(VkDevice, uint32_t n, const VkFence *fences)
{
  for(uint32_t i = 0; i < n; i++)
    if(fences[i])
    {
      std::lock_guard<std::mutex> lk(fences[i]->m);
      fences[i]->signaled = false;
    }
  return VK_SUCCESS;
}

VKAPI_ATTR VkResult VKAPI_CALL vkWaitForFences(VkDevice, uint32_t n, const VkFence *fences, VkBool32 waitAll, uint64_t timeout)
{
  auto deadline = std::chrono::steady_clock::now() +
                  std::chrono::nanoseconds(timeout > (uint64_t)INT64_MAX / 2 ? (uint64_t)INT64_MAX / 2 : timeout);
  if(!waitAll)
  {
    // any one fence: poll all of them until one is signalled or the deadline passes
    for(;;)
    {
      for(uint32_t i = 0; i < n; i++)
        if(VkFence f = fences[i])
        {
          std::lock_guard<std::mutex> lk(f->m);
          if(f->signaled)
            return f->status;
        }
      if(std::chrono::steady_clock::now() >= deadline)
        return VK_TIMEOUT;
      std::this_thread::sleep_for(std::chrono::microseconds(20));
    }
  }
  VkResult res = VK_SUCCESS;
  for(uint32_t i = 0; i < n; i++)
  {
    VkFence f = fences[i];
    if(!f)
      continue;
    std::unique_lock<std::mutex> lk(f->m);
    if(!f->cv.wait_until(lk, deadline, [&] { return f->signaled; }))
      return VK_TIMEOUT;
    if(f->status != VK_SUCCESS)
      res = f->status;
  }
  return res;
}

VKAPI_ATTR VkResult VKAPI_CALL vkCreateSemaphore(VkDevice, const VkSemaphoreCreateInfo *, const VkAllocationCallbacks *,
                                                 VkSemaphore *p)
{
  *p = new VkSemaphore_T;
  return VK_SUCCESS;
}

VKAPI_ATTR void VKAPI_CALL vkDestroySemaphore(VkDevice, VkSemaphore s, const VkAllocationCallbacks *) { delete s; }

VKAPI_ATTR VkResult VKAPI_CALL vkQueueWaitIdle(VkQueue queue) { return queue->q->WaitIdle(); }

VKAPI_ATTR VkResult VKAPI_CALL vkDeviceWaitIdle(VkDevice device)
{
  for(VkQueue_T *q : device->queues)
    q->q->WaitIdle();
  return VK_SUCCESS;
}

// ---------------------------------------------------------------------------------------------
// memory, buffers, images (reference memory.cpp, buffers.cpp, images.cpp)

VKAPI_ATTR VkResult VKAPI_CALL vkAllocateMemory(VkDevice, const VkMemoryAllocateInfo *info, const VkAllocationCallbacks *,
                                                VkDeviceMemory *pMemory)
{
  VkDeviceMemory m = new VkDeviceMemory_T;
  m->size = info->allocationSize;
  m->type = info->memoryTypeIndex == 1 ? 1 : 0;
  // 16 bytes beyond the allocation stay readable: the vertex fetch loads 16 bytes from the start
  // of every attribute that lies inside its buffer
  size_t bytes = std::max<size_t>(m->size, 4) + 16;
  if(hipMalloc((void **)&m->dev, bytes) != hipSuccess)
  {
    delete m;
    return VK_ERROR_OUT_OF_DEVICE_MEMORY;
  }
  if(m->type == 1)
  {
    if(hipHostMalloc((void **)&m->host, bytes, hipHostMallocDefault) != hipSuccess)
    {
      hipFree(m->dev);
      delete m;
      return VK_ERROR_OUT_OF_HOST_MEMORY;
    }
    memset(m->host, 0, bytes);
  }
  *pMemory = m;
  return VK_SUCCESS;
}

VKAPI_ATTR void VKAPI_CALL vkFreeMemory(VkDevice device, VkDeviceMemory memory, const VkAllocationCallbacks *)
{
  if(!memory)
    return;
  if(device)
    vkDeviceWaitIdle(device);
  hipDeviceSynchronize();
  hipFree(memory->dev);
  if(memory->host)
    hipHostFree(memory->host);
  delete memory;
}

VKAPI_ATTR VkResult VKAPI_CALL vkMapMemory(VkDevice, VkDeviceMemory memory, VkDeviceSize offset, VkDeviceSize, VkMemoryMapFlags,
                                           void **ppData)
{
  if(memory->type != 1)
    return VK_ERROR_MEMORY_MAP_FAILED;
  *ppData = memory->host + offset;
  return VK_SUCCESS;
}

VKAPI_ATTR void VKAPI_CALL vkUnmapMemory(VkDevice, VkDeviceMemory) {}

VKAPI_ATTR VkResult VKAPI_CALL vkFlushMappedMemoryRanges(VkDevice, uint32_t, const VkMappedMemoryRange *)
{
  return VK_SUCCESS;  // host-visible memory is uploaded at every submit that reads it
}

VKAPI_ATTR VkResult VKAPI_CALL vkCreateBuffer(VkDevice, const VkBufferCreateInfo *info, const VkAllocationCallbacks *, VkBuffer *p)
{
  VkBuffer b = new VkBuffer_T;
  b->size = info->size;
  *p = b;
  return VK_SUCCESS;
}

VKAPI_ATTR void VKAPI_CALL vkDestroyBuffer(VkDevice, VkBuffer buffer, const VkAllocationCallbacks *) { delete buffer; }

VKAPI_ATTR void VKAPI_CALL vkGetBufferMemoryRequirements(VkDevice, VkBuffer buffer, VkMemoryRequirements *r)
{
  r->alignment = 16;
  r->memoryTypeBits = 0x3;
  r->size = buffer->size;
}

VKAPI_ATTR VkResult VKAPI_CALL vkBindBufferMemory(VkDevice, VkBuffer buffer, VkDeviceMemory memory, VkDeviceSize offset)
{
  buffer->mem = memory;
  buffer->offset = offset;
  return VK_SUCCESS;
}

static uint32_t FormatBpp(VkFormat f)
{
  // reference images.cpp:26-33: 4 bytes except R8 and BC2/BC3 (1); RGBA16F (8) is beyond it
  if(f == VK_FORMAT_R8_UNORM || f == VK_FORMAT_BC2_UNORM_BLOCK || f == VK_FORMAT_BC3_UNORM_BLOCK)
    return 1;
  if(f == VK_FORMAT_R16G16B16A16_SFLOAT)
    return 8;
  return 4;
}

VKAPI_ATTR VkResult VKAPI_CALL vkCreateImage(VkDevice, const VkImageCreateInfo *info, const VkAllocationCallbacks *, VkImage *p)
{
  VkImage i = new VkImage_T;
  i->extent = info->extent;
  i->imageType = info->imageType;
  i->format = info->format;
  i->arrayLayers = std::max(1u, info->arrayLayers);
  i->mipLevels = std::max(1u, info->mipLevels);
  i->bytesPerPixel = FormatBpp(info->format);
  *p = i;
  return VK_SUCCESS;
}

VKAPI_ATTR void VKAPI_CALL vkDestroyImage(VkDevice, VkImage image, const VkAllocationCallbacks *)
{
  if(image && image->shadow)
    (void)hipFree(image->shadow);  // waits for work still reading it
  delete image;
}

VKAPI_ATTR VkResult VKAPI_CALL vkBindImageMemory(VkDevice, VkImage image, VkDeviceMemory memory, VkDeviceSize offset)
{
  image->mem = memory;
  image->offset = offset;
  return VK_SUCCESS;
}

// Exact size of the linear layout: mips packed per layer, layers back to back (the reference
// over-allocates 2x for any mip chain, images.cpp:301-315).
VKAPI_ATTR void VKAPI_CALL vkGetImageMemoryRequirements(VkDevice, VkImage image, VkMemoryRequirements *r)
{
  uint64_t slice = 0;
  uint32_t w = image->extent.width, h = image->extent.height;
  for(uint32_t m = 0; m < image->mipLevels; m++)
  {
    slice += (uint64_t)w * h * image->bytesPerPixel;
    w = std::max(1u, w >> 1);
    h = std::max(1u, h >> 1);
  }
  uint64_t depth = image->imageType == VK_IMAGE_TYPE_3D ? std::max(1u, image->extent.depth) : 1;
  r->size = slice * image->arrayLayers * depth;
  r->alignment = 256;
  r->memoryTypeBits = 0x3;
}

VKAPI_ATTR void VKAPI_CALL vkGetImageSubresourceLayout(VkDevice, VkImage image, const VkImageSubresource *, VkSubresourceLayout *l)
{
  l->offset = 0;
  l->rowPitch = (VkDeviceSize)image->extent.width * image->bytesPerPixel;
  l->arrayPitch = l->depthPitch = l->size = (VkDeviceSize)image->extent.width * image->extent.height * image->bytesPerPixel;
}

VKAPI_ATTR VkResult VKAPI_CALL vkCreateImageView(VkDevice, const VkImageViewCreateInfo *info, const VkAllocationCallbacks *,
                                                 VkImageView *p)
{
  VkImageView v = new VkImageView_T;
  v->image = info->image;
  *p = v;
  return VK_SUCCESS;
}

VKAPI_ATTR void VKAPI_CALL vkDestroyImageView(VkDevice, VkImageView view, const VkAllocationCallbacks *) { delete view; }

VKAPI_ATTR VkResult VKAPI_CALL vkCreateSampler(VkDevice, const VkSamplerCreateInfo *info, const VkAllocationCallbacks *, VkSampler *p)
{
  VkSampler s = new VkSampler_T;
  s->info = *info;
  s->info.pNext = nullptr;
  *p = s;
  return VK_SUCCESS;
}

VKAPI_ATTR void VKAPI_CALL vkDestroySampler(VkDevice, VkSampler s, const VkAllocationCallbacks *) { delete s; }

// ---------------------------------------------------------------------------------------------
// descriptors, layouts, render passes, framebuffers (reference descriptors.cpp, renderpass.cpp)

VKAPI_ATTR VkResult VKAPI_CALL vkCreateDescriptorSetLayout(VkDevice, const VkDescriptorSetLayoutCreateInfo *info,
                                                           const VkAllocationCallbacks *, VkDescriptorSetLayout *p)
{
  VkDescriptorSetLayout l = new VkDescriptorSetLayout_T;
  uint32_t n = 0;
  for(uint32_t i = 0; i < info->bindingCount; i++)
    n = std::max(n, info->pBindings[i].binding + 1);
  l->bindingCount = std::max(n, info->bindingCount);
  *p = l;
  return VK_SUCCESS;
}

VKAPI_ATTR void VKAPI_CALL vkDestroyDescriptorSetLayout(VkDevice, VkDescriptorSetLayout l, const VkAllocationCallbacks *) { delete l; }

struct VkPipelineLayout_T { int dummy; };
struct VkDescriptorPool_T { int dummy; };
struct VkPipelineCache_T { int dummy; };

VKAPI_ATTR VkResult VKAPI_CALL vkCreatePipelineLayout(VkDevice, const VkPipelineLayoutCreateInfo *, const VkAllocationCallbacks *,
                                                      VkPipelineLayout *p)
{
  *p = new VkPipelineLayout_T;
  return VK_SUCCESS;
}

VKAPI_ATTR void VKAPI_CALL vkDestroyPipelineLayout(VkDevice, VkPipelineLayout l, const VkAllocationCallbacks *) { delete l; }

VKAPI_ATTR VkResult VKAPI_CALL vkCreateDescriptorPool(VkDevice, const VkDescriptorPoolCreateInfo *, const VkAllocationCallbacks *,
                                                      VkDescriptorPool *p)
{
  *p = new VkDescriptorPool_T;
  return VK_SUCCESS;
}

VKAPI_ATTR void VKAPI_CALL vkDestroyDescriptorPool(VkDevice, VkDescriptorPool p, const VkAllocationCallbacks *) { delete p; }

// One set per requested layout (the reference writes every set to index 0, descriptors.cpp:58-63).
VKAPI_ATTR VkResult VKAPI_CALL vkAllocateDescriptorSets(VkDevice, const VkDescriptorSetAllocateInfo *info, VkDescriptorSet *sets)
{
  for(uint32_t i = 0; i < info->descriptorSetCount; i++)
  {
    sets[i] = new VkDescriptorSet_T;
    sets[i]->binds.resize(info->pSetLayouts[i]->bindingCount);
  }
  return VK_SUCCESS;
}

VKAPI_ATTR VkResult VKAPI_CALL vkFreeDescriptorSets(VkDevice, VkDescriptorPool, uint32_t n, const VkDescriptorSet *sets)
{
  for(uint32_t i = 0; i < n; i++)
    delete sets[i];
  return VK_SUCCESS;
}

VKAPI_ATTR void VKAPI_CALL vkUpdateDescriptorSets(VkDevice, uint32_t nw, const VkWriteDescriptorSet *writes, uint32_t nc,
                                                  const VkCopyDescriptorSet *copies)
{
  visor::g_descriptor_writes.fetch_add(1);
  for(uint32_t i = 0; i < nw; i++)
  {
    const VkWriteDescriptorSet &w = writes[i];
    if(w.dstBinding >= w.dstSet->binds.size())
      w.dstSet->binds.resize(w.dstBinding + 1);
    auto &b = w.dstSet->binds[w.dstBinding];
    b.type = w.descriptorType;
    switch(w.descriptorType)
    {
      case VK_DESCRIPTOR_TYPE_SAMPLER: case VK_DESCRIPTOR_TYPE_COMBINED_IMAGE_SAMPLER: case VK_DESCRIPTOR_TYPE_SAMPLED_IMAGE:
      case VK_DESCRIPTOR_TYPE_STORAGE_IMAGE: case VK_DESCRIPTOR_TYPE_INPUT_ATTACHMENT:
        b.imageInfo = w.pImageInfo[0];
        break;
      case VK_DESCRIPTOR_TYPE_UNIFORM_BUFFER: case VK_DESCRIPTOR_TYPE_STORAGE_BUFFER:
      case VK_DESCRIPTOR_TYPE_UNIFORM_BUFFER_DYNAMIC: case VK_DESCRIPTOR_TYPE_STORAGE_BUFFER_DYNAMIC:
        b.bufferInfo = w.pBufferInfo[0];
        break;
      default: break;
    }
  }
  for(uint32_t i = 0; i < nc; i++)
  {
    const VkCopyDescriptorSet &c = copies[i];
    if(c.dstBinding >= c.dstSet->binds.size())
      c.dstSet->binds.resize(c.dstBinding + 1);
    if(c.srcBinding < c.srcSet->binds.size())
      c.dstSet->binds[c.dstBinding] = c.srcSet->binds[c.srcBinding];
  }
}

VKAPI_ATTR VkResult VKAPI_CALL vkCreateFramebuffer(VkDevice, const VkFramebufferCreateInfo *info, const VkAllocationCallbacks *,
                                                   VkFramebuffer *p)
{
  VkFramebuffer f = new VkFramebuffer_T;
  f->attachments.assign(info->pAttachments, info->pAttachments + info->attachmentCount);
  *p = f;
  return VK_SUCCESS;
}

VKAPI_ATTR void VKAPI_CALL vkDestroyFramebuffer(VkDevice, VkFramebuffer f, const VkAllocationCallbacks *) { delete f; }

VKAPI_ATTR VkResult VKAPI_CALL vkCreateRenderPass(VkDevice, const VkRenderPassCreateInfo *info, const VkAllocationCallbacks *,
                                                  VkRenderPass *p)
{
  VkRenderPass r = new VkRenderPass_T;
  for(uint32_t i = 0; i < info->subpassCount; i++)
  {
    const VkSubpassDescription &sd = info->pSubpasses[i];
    VkRenderPass_T::Subpass sub;
    sub.colAttachments.resize(sd.colorAttachmentCount);
    for(uint32_t a = 0; a < sd.colorAttachmentCount; a++)
    {
      sub.colAttachments[a].idx = (int32_t)sd.pColorAttachments[a].attachment;
      sub.colAttachments[a].clear = info->pAttachments[sub.colAttachments[a].idx].loadOp == VK_ATTACHMENT_LOAD_OP_CLEAR;
    }
    if(sd.pDepthStencilAttachment && sd.pDepthStencilAttachment->attachment != ~0u)
    {
      sub.depthAttachment.idx = (int32_t)sd.pDepthStencilAttachment->attachment;
      sub.depthAttachment.clear = info->pAttachments[sub.depthAttachment.idx].loadOp == VK_ATTACHMENT_LOAD_OP_CLEAR;
    }
    r->subpasses.push_back(sub);
  }
  *p = r;
  return VK_SUCCESS;
}

VKAPI_ATTR void VKAPI_CALL vkDestroyRenderPass(VkDevice, VkRenderPass r, const VkAllocationCallbacks *) { delete r; }

// ---------------------------------------------------------------------------------------------
// shaders and pipelines (reference shaders.cpp)

VKAPI_ATTR VkResult VKAPI_CALL vkCreateShaderModule(VkDevice, const VkShaderModuleCreateInfo *info, const VkAllocationCallbacks *,
                                                    VkShaderModule *p)
{
  size_t n = info->codeSize / 4;
  if(!SpirvHeaderValid(info->pCode, n))
  {
    SetError("vkCreateShaderModule: not a SPIR-V module");
    return VK_ERROR_DEVICE_LOST;  // the reference's failure code (shaders.cpp:13-14)
  }
  VkShaderModule m = new VkShaderModule_T;
  m->code.assign(info->pCode, info->pCode + n);
  *p = m;
  return VK_SUCCESS;
}

VKAPI_ATTR void VKAPI_CALL vkDestroyShaderModule(VkDevice, VkShaderModule m, const VkAllocationCallbacks *) { delete m; }

VKAPI_ATTR VkResult VKAPI_CALL vkCreatePipelineCache(VkDevice, const VkPipelineCacheCreateInfo *, const VkAllocationCallbacks *,
                                                     VkPipelineCache *p)
{
  *p = new VkPipelineCache_T;  // code objects are cached on disk by the runtime
  return VK_SUCCESS;
}

VKAPI_ATTR void VKAPI_CALL vkDestroyPipelineCache(VkDevice, VkPipelineCache c, const VkAllocationCallbacks *) { delete c; }

VKAPI_ATTR VkResult VKAPI_CALL vkCreateGraphicsPipelines(VkDevice, VkPipelineCache, uint32_t n, const VkGraphicsPipelineCreateInfo *infos,
                                                         const VkAllocationCallbacks *, VkPipeline *out)
{
  for(uint32_t i = 0; i < n; i++)
  {
    const VkGraphicsPipelineCreateInfo &ci = infos[i];
    VkPipeline p = new VkPipeline_T;
    memset(p->vattrs, 0, sizeof(p->vattrs));
    uint32_t strides[16] = {0};
    const VkPipelineVertexInputStateCreateInfo *vi = ci.pVertexInputState;
    // Unsupported state is an error here, not a silent substitute: vertex buffers bind to slots
    // 0..3 (the reference's GPUState has vbs[4], gpu.h:3-23) and colour blending runs the ADD
    // op only (the reference leaves every other op's result undefined, rasterizer.cpp:657-669).
    const char *unsupported = nullptr;
    if(vi)
    {
      for(uint32_t b = 0; b < vi->vertexBindingDescriptionCount; b++)
        if(vi->pVertexBindingDescriptions[b].binding >= 4)
          unsupported = "vertex binding beyond slot 3";
      for(uint32_t a = 0; a < vi->vertexAttributeDescriptionCount; a++)
        if(vi->pVertexAttributeDescriptions[a].binding >= 4 || vi->pVertexAttributeDescriptions[a].location >= 16)
          unsupported = "vertex attribute on a binding beyond slot 3 or a location beyond 15";
    }
    if(ci.pColorBlendState && ci.pColorBlendState->attachmentCount > 0 &&
       ci.pColorBlendState->pAttachments[0].blendEnable && ci.pColorBlendState->pAttachments[0].colorBlendOp != VK_BLEND_OP_ADD)
      unsupported = "colour blend op other than ADD";
    if(unsupported)
    {
      SetError(std::string("vkCreateGraphicsPipelines: unsupported: ") + unsupported);
      delete p;
      return VK_ERROR_FEATURE_NOT_PRESENT;
    }
    if(vi)
    {
      for(uint32_t b = 0; b < vi->vertexBindingDescriptionCount; b++)
        strides[vi->pVertexBindingDescriptions[b].binding] = vi->pVertexBindingDescriptions[b].stride;
      for(uint32_t a = 0; a < vi->vertexAttributeDescriptionCount; a++)
      {
        const VkVertexInputAttributeDescription &ad = vi->pVertexAttributeDescriptions[a];
        p->vattrs[ad.location].vb = ad.binding;
        p->vattrs[ad.location].format = ad.format;
        p->vattrs[ad.location].offset = ad.offset;
        p->vattrs[ad.location].stride = strides[ad.binding];
      }
    }
    p->topology = ci.pInputAssemblyState ? ci.pInputAssemblyState->topology : VK_PRIMITIVE_TOPOLOGY_TRIANGLE_LIST;
    if(ci.pRasterizationState)
    {
      p->frontFace = ci.pRasterizationState->frontFace;
      p->cullMode = ci.pRasterizationState->cullMode;
    }
    if(ci.pDepthStencilState)
    {
      // the reference takes the compare op whether or not depthTestEnable is set (shaders.cpp:228-232)
      p->depthCompareOp = ci.pDepthStencilState->depthCompareOp;
      p->depthWriteEnable = ci.pDepthStencilState->depthWriteEnable == VK_TRUE;
    }
    if(ci.pColorBlendState && ci.pColorBlendState->attachmentCount > 0)
      p->blend = ci.pColorBlendState->pAttachments[0];
    else
      memset(&p->blend, 0, sizeof(p->blend));
    VkShaderModule vs = nullptr, fs = nullptr;
    const char *vsn = "main", *fsn = "main";
    for(uint32_t s = 0; s < ci.stageCount; s++)
    {
      if(ci.pStages[s].stage == VK_SHADER_STAGE_VERTEX_BIT)
      {
        vs = ci.pStages[s].module;
        vsn = ci.pStages[s].pName;
      }
      else if(ci.pStages[s].stage == VK_SHADER_STAGE_FRAGMENT_BIT)
      {
        fs = ci.pStages[s].module;
        fsn = ci.pStages[s].pName;
      }
    }
    if(!vs || !fs)
    {
      SetError("vkCreateGraphicsPipelines: a vertex and a fragment stage are required");
      delete p;
      return VK_ERROR_FEATURE_NOT_PRESENT;
    }
    std::string err;
    p->pair = Device::Get().MakePair(vs, vsn, fs, fsn, err);
    if(!p->pair)
    {
      SetError("vkCreateGraphicsPipelines: " + err);
      delete p;
      return VK_ERROR_DEVICE_LOST;
    }
    // build the single-slot module now so the first draw does not pay for hiprtc
    if(Device::Get().ok() && !getenv("VISOR_LAZY_PIPELINES"))
    {
      if(!Device::Get().GetModule({p->pair}, err))
      {
        SetError("vkCreateGraphicsPipelines: " + err);
        delete p;
        return VK_ERROR_DEVICE_LOST;
      }
    }
    out[i] = p;
  }
  return VK_SUCCESS;
}

VKAPI_ATTR VkResult VKAPI_CALL vkCreateComputePipelines(VkDevice, VkPipelineCache, uint32_t n, const VkComputePipelineCreateInfo *,
                                                        const VkAllocationCallbacks *, VkPipeline *out)
{
  for(uint32_t i = 0; i < n; i++)
  {
    out[i] = new VkPipeline_T;
    out[i]->graphics = false;  // empty, as the reference's (shaders.cpp:93-104)
  }
  return VK_SUCCESS;
}

VKAPI_ATTR void VKAPI_CALL vkDestroyPipeline(VkDevice, VkPipeline p, const VkAllocationCallbacks *) { delete p; }

// ---------------------------------------------------------------------------------------------
// command pools and recording (reference cmd_alloc.cpp, cmd_record.cpp)

VKAPI_ATTR VkResult VKAPI_CALL vkCreateCommandPool(VkDevice, const VkCommandPoolCreateInfo *, const VkAllocationCallbacks *,
                                                   VkCommandPool *p)
{
  *p = new VkCommandPool_T;
  return VK_SUCCESS;
}

VKAPI_ATTR void VKAPI_CALL vkDestroyCommandPool(VkDevice, VkCommandPool pool, const VkAllocationCallbacks *)
{
  for(VkCommandBuffer c : pool->buffers)
    delete c;
  delete pool;
}

VKAPI_ATTR VkResult VKAPI_CALL vkAllocateCommandBuffers(VkDevice, const VkCommandBufferAllocateInfo *info, VkCommandBuffer *out)
{
  for(uint32_t i = 0; i < info->commandBufferCount; i++)
  {
    VkCommandBuffer c = nullptr;
    for(VkCommandBuffer b : info->commandPool->buffers)
      if(!b->live)
      {
        c = b;
        break;
      }
    if(!c)
    {
      c = new VkCommandBuffer_T;
      info->commandPool->buffers.push_back(c);
    }
    SetLoaderMagic(c);
    c->live = true;
    c->stream.clear();
    out[i] = c;
  }
  return VK_SUCCESS;
}

VKAPI_ATTR void VKAPI_CALL vkFreeCommandBuffers(VkDevice, VkCommandPool, uint32_t n, const VkCommandBuffer *bufs)
{
  for(uint32_t i = 0; i < n; i++)
    if(bufs[i])
      bufs[i]->live = false;
}

VKAPI_ATTR VkResult VKAPI_CALL vkBeginCommandBuffer(VkCommandBuffer cb, const VkCommandBufferBeginInfo *)
{
  static std::atomic<uint64_t> next_recording{1};
  cb->stream.clear();
  cb->recording = next_recording.fetch_add(1);
  return VK_SUCCESS;
}

VKAPI_ATTR VkResult VKAPI_CALL vkEndCommandBuffer(VkCommandBuffer) { return VK_SUCCESS; }

VKAPI_ATTR void VKAPI_CALL vkCmdPipelineBarrier(VkCommandBuffer cb, VkPipelineStageFlags, VkPipelineStageFlags, VkDependencyFlags,
                                                uint32_t, const VkMemoryBarrier *, uint32_t, const VkBufferMemoryBarrier *,
                                                uint32_t, const VkImageMemoryBarrier *)
{
  cb->pushId(Cmd::PipelineBarrier);
}

VKAPI_ATTR void VKAPI_CALL vkCmdBeginRenderPass(VkCommandBuffer cb, const VkRenderPassBeginInfo *info, VkSubpassContents)
{
  cmd::BeginRenderPass *c = cb->push<cmd::BeginRenderPass>(Cmd::BeginRenderPass);
  c->renderPass = info->renderPass;
  c->framebuffer = info->framebuffer;
  memcpy(c->clearval, info->pClearValues, sizeof(VkClearValue) * std::min(info->clearValueCount, 8u));
}

VKAPI_ATTR void VKAPI_CALL vkCmdEndRenderPass(VkCommandBuffer cb) { cb->pushId(Cmd::EndRenderPass); }

VKAPI_ATTR void VKAPI_CALL vkCmdBindPipeline(VkCommandBuffer cb, VkPipelineBindPoint, VkPipeline pipeline)
{
  cb->push<cmd::BindPipeline>(Cmd::BindPipeline)->pipeline = pipeline;
}

VKAPI_ATTR void VKAPI_CALL vkCmdBindDescriptorSets(VkCommandBuffer cb, VkPipelineBindPoint, VkPipelineLayout, uint32_t firstSet,
                                                   uint32_t count, const VkDescriptorSet *sets, uint32_t, const uint32_t *)
{
  for(uint32_t i = 0; i < count; i++)
  {
    cmd::BindDescriptorSets *c = cb->push<cmd::BindDescriptorSets>(Cmd::BindDescriptorSets);
    c->idx = firstSet + i;
    c->set = sets[i];
  }
}

VKAPI_ATTR void VKAPI_CALL vkCmdBindVertexBuffers(VkCommandBuffer cb, uint32_t first, uint32_t count, const VkBuffer *bufs,
                                                  const VkDeviceSize *offs)
{
  for(uint32_t i = 0; i < count; i++)
  {
    cmd::BindVB *c = cb->push<cmd::BindVB>(Cmd::BindVB);
    c->slot = first + i;
    c->buffer = bufs[i];
    c->offset = offs[i];
  }
}

VKAPI_ATTR void VKAPI_CALL vkCmdBindIndexBuffer(VkCommandBuffer cb, VkBuffer buffer, VkDeviceSize offset, VkIndexType type)
{
  cmd::BindIB *c = cb->push<cmd::BindIB>(Cmd::BindIB);
  c->buffer = buffer;
  c->offset = offset;
  c->indexType = type;
}

VKAPI_ATTR void VKAPI_CALL vkCmdSetViewport(VkCommandBuffer cb, uint32_t, uint32_t, const VkViewport *v)
{
  cb->push<cmd::SetViewport>(Cmd::SetViewport)->view = v[0];
}

VKAPI_ATTR void VKAPI_CALL vkCmdSetScissor(VkCommandBuffer cb, uint32_t, uint32_t, const VkRect2D *) { cb->pushId(Cmd::SetScissors); }

VKAPI_ATTR void VKAPI_CALL vkCmdPushConstants(VkCommandBuffer cb, VkPipelineLayout, VkShaderStageFlags, uint32_t offset, uint32_t size,
                                              const void *values)
{
  cmd::PushConstants *c = cb->push<cmd::PushConstants>(Cmd::PushConstants);
  c->offset = offset;
  c->size = std::min(size, 128u);
  memcpy(c->values, values, c->size);
}

VKAPI_ATTR void VKAPI_CALL vkCmdDraw(VkCommandBuffer cb, uint32_t vertexCount, uint32_t instanceCount, uint32_t firstVertex,
                                     uint32_t firstInstance)
{
  cmd::Draw *c = cb->push<cmd::Draw>(Cmd::Draw);
  c->vertexCount = vertexCount;
  c->instanceCount = instanceCount;
  c->firstVertex = firstVertex;
  c->firstInstance = firstInstance;
}

VKAPI_ATTR void VKAPI_CALL vkCmdDrawIndexed(VkCommandBuffer cb, uint32_t indexCount, uint32_t instanceCount, uint32_t firstIndex,
                                            int32_t vertexOffset, uint32_t firstInstance)
{
  cmd::DrawIndexed *c = cb->push<cmd::DrawIndexed>(Cmd::DrawIndexed);
  c->indexCount = indexCount;
  c->instanceCount = instanceCount;
  c->firstIndex = firstIndex;
  c->vertexOffset = vertexOffset;
  c->firstInstance = firstInstance;
}

VKAPI_ATTR void VKAPI_CALL vkCmdCopyBufferToImage(VkCommandBuffer cb, VkBuffer src, VkImage dst, VkImageLayout, uint32_t n,
                                                  const VkBufferImageCopy *regions)
{
  for(uint32_t r = 0; r < n; r++)
  {
    cmd::CopyBuf2Img *c = cb->push<cmd::CopyBuf2Img>(Cmd::CopyBuf2Img);
    c->srcBuffer = src;
    c->dstImage = dst;
    c->region = regions[r];
  }
}

VKAPI_ATTR void VKAPI_CALL vkCmdCopyBuffer(VkCommandBuffer cb, VkBuffer src, VkBuffer dst, uint32_t n, const VkBufferCopy *regions)
{
  for(uint32_t r = 0; r < n; r++)
  {
    cmd::CopyBuf *c = cb->push<cmd::CopyBuf>(Cmd::CopyBuf);
    c->srcBuffer = src;
    c->dstBuffer = dst;
    c->region = regions[r];
  }
}

VKAPI_ATTR VkResult VKAPI_CALL vkQueueSubmit(VkQueue queue, uint32_t count, const VkSubmitInfo *infos, VkFence fence)
{
  g_last_queue = queue;
  return queue->q->Submit(count, infos, fence);
}

// ---------------------------------------------------------------------------------------------
// headless WSI: swapchain images are host-visible linear BGRA8 images; present counts frames

VKAPI_ATTR VkResult VKAPI_CALL vkGetPhysicalDeviceSurfaceSupportKHR(VkPhysicalDevice, uint32_t, VkSurfaceKHR, VkBool32 *s)
{
  *s = VK_TRUE;
  return VK_SUCCESS;
}

VKAPI_ATTR VkResult VKAPI_CALL vkGetPhysicalDeviceSurfaceCapabilitiesKHR(VkPhysicalDevice, VkSurfaceKHR, VkSurfaceCapabilitiesKHR *c)
{
  memset(c, 0, sizeof(*c));
  c->minImageCount = 1;
  c->maxImageCount = 2;
  c->currentExtent = {0xffffffffu, 0xffffffffu};
  c->minImageExtent = {1, 1};
  c->maxImageExtent = {16384, 16384};
  c->maxImageArrayLayers = 1;
  c->supportedTransforms = VK_SURFACE_TRANSFORM_IDENTITY_BIT_KHR;
  c->currentTransform = VK_SURFACE_TRANSFORM_IDENTITY_BIT_KHR;
  c->supportedCompositeAlpha = VK_COMPOSITE_ALPHA_OPAQUE_BIT_KHR;
  c->supportedUsageFlags = 0x10;
  return VK_SUCCESS;
}

VKAPI_ATTR VkResult VKAPI_CALL vkGetPhysicalDeviceSurfaceFormatsKHR(VkPhysicalDevice, VkSurfaceKHR, uint32_t *n, VkSurfaceFormatKHR *f)
{
  if(!f)
  {
    *n = 1;
    return VK_SUCCESS;
  }
  if(*n < 1)
    return VK_INCOMPLETE;
  *n = 1;
  f[0].format = VK_FORMAT_B8G8R8A8_UNORM;
  f[0].colorSpace = VK_COLOR_SPACE_SRGB_NONLINEAR_KHR;
  return VK_SUCCESS;
}

VKAPI_ATTR VkResult VKAPI_CALL vkGetPhysicalDeviceSurfacePresentModesKHR(VkPhysicalDevice, VkSurfaceKHR, uint32_t *n, VkPresentModeKHR *m)
{
  if(!m)
  {
    *n = 1;
    return VK_SUCCESS;
  }
  if(*n < 1)
    return VK_INCOMPLETE;
  *n = 1;
  m[0] = VK_PRESENT_MODE_FIFO_KHR;
  return VK_SUCCESS;
}

VKAPI_ATTR void VKAPI_CALL vkDestroySwapchainKHR(VkDevice device, VkSwapchainKHR s, const VkAllocationCallbacks *);

VKAPI_ATTR VkResult VKAPI_CALL vkCreateSwapchainKHR(VkDevice device, const VkSwapchainCreateInfoKHR *info, const VkAllocationCallbacks *,
                                                    VkSwapchainKHR *p)
{
  VkSwapchainKHR s = new VkSwapchainKHR_T;
  s->extent = info->imageExtent;
  s->format = info->imageFormat;
  // minImageCount back buffers (reference wsi.cpp:96), host-visible so a present can be read back
  uint32_t n = std::max(1u, std::min(info->minImageCount, 16u));
  for(uint32_t i = 0; i < n; i++)
  {
    VkImageCreateInfo ici;
    memset(&ici, 0, sizeof(ici));
    ici.imageType = VK_IMAGE_TYPE_2D;
    ici.format = info->imageFormat;
    ici.extent = {info->imageExtent.width, info->imageExtent.height, 1};
    ici.mipLevels = ici.arrayLayers = 1;
    VkImage img;
    vkCreateImage(device, &ici, nullptr, &img);
    VkMemoryRequirements req;
    vkGetImageMemoryRequirements(device, img, &req);
    VkMemoryAllocateInfo mai = {VK_STRUCTURE_TYPE_MEMORY_ALLOCATE_INFO, nullptr, req.size, 1};
    VkDeviceMemory mem;
    VkResult r = vkAllocateMemory(device, &mai, nullptr, &mem);
    if(r != VK_SUCCESS)
    {
      vkDestroyImage(device, img, nullptr);
      vkDestroySwapchainKHR(device, s, nullptr);
      return r;
    }
    vkBindImageMemory(device, img, mem, 0);
    s->images.push_back(img);
    s->mems.push_back(mem);
  }
  *p = s;
  return VK_SUCCESS;
}

VKAPI_ATTR void VKAPI_CALL vkDestroySwapchainKHR(VkDevice device, VkSwapchainKHR s, const VkAllocationCallbacks *)
{
  if(!s)
    return;
  for(size_t i = 0; i < s->images.size(); i++)
  {
    vkDestroyImage(device, s->images[i], nullptr);
    vkFreeMemory(device, s->mems[i], nullptr);
  }
  delete s;
}

VKAPI_ATTR VkResult VKAPI_CALL vkGetSwapchainImagesKHR(VkDevice, VkSwapchainKHR s, uint32_t *n, VkImage *imgs)
{
  if(!imgs)
  {
    *n = (uint32_t)s->images.size();
    return VK_SUCCESS;
  }
  uint32_t k = std::min<uint32_t>(*n, (uint32_t)s->images.size());
  for(uint32_t i = 0; i < k; i++)
    imgs[i] = s->images[i];
  *n = k;
  return k < s->images.size() ? VK_INCOMPLETE : VK_SUCCESS;
}

VKAPI_ATTR VkResult VKAPI_CALL vkAcquireNextImageKHR(VkDevice, VkSwapchainKHR s, uint64_t, VkSemaphore, VkFence fence, uint32_t *idx)
{
  // reference wsi.cpp:185-195: the next image in turn, no waiting (images are never in use by
  // a display here)
  s->current = (s->current + 1) % (uint32_t)s->images.size();
  *idx = s->current;
  if(fence)
  {
    std::lock_guard<std::mutex> lk(fence->m);
    fence->signaled = true;
    fence->cv.notify_all();
  }
  return VK_SUCCESS;
}

// Headless present (reference wsi.cpp:201-219 blits pImageIndices[i] to a window): once the
// queue's work is done, the presented image's host-visible memory holds the frame
// (vh_swapchain_frame), and the queue's frame statistics end and begin as the reference's
// EndFrameStats / BeginFrameStats do (stats.cpp:25-41; vh_get_stats 2 and 3).
VKAPI_ATTR VkResult VKAPI_CALL vkQueuePresentKHR(VkQueue queue, const VkPresentInfoKHR *info)
{
  queue->q->WaitIdle();
  VkResult res = VK_SUCCESS;
  for(uint32_t i = 0; i < info->swapchainCount; i++)
  {
    VkSwapchainKHR s = info->pSwapchains[i];
    VkResult r = VK_SUCCESS;
    if(!info->pImageIndices || info->pImageIndices[i] >= s->images.size())
      r = VK_ERROR_OUT_OF_DATE_KHR;
    else
    {
      s->presented = info->pImageIndices[i];
      s->presents++;
    }
    if(info->pResults)
      info->pResults[i] = r;
    if(r != VK_SUCCESS)
      res = r;
  }
  queue->q->presented_stats = queue->q->frame_stats;
  queue->q->frame_stats = ExecStats();
  return res;
}

// ---------------------------------------------------------------------------------------------
// loader interface (reference icd_interface.cpp:3-169)

// Entry points by name, sorted for binary search. Every name is visible through
// vk_icdGetInstanceProcAddr; the scope bits say which names vk_icdGetPhysicalDeviceProcAddr and
// vkGetDeviceProcAddr also return (the three lists of icd_interface.cpp:7-102).
enum : uint8_t
{
  kInst = 0,
  kPhys = 1,
  kDev = 2,
};

struct EntryPoint
{
  const char *name;
  PFN_vkVoidFunction fn;
  uint8_t scope;
};

static const EntryPoint kEntryPoints[] = {
    {"vkAcquireNextImageKHR", reinterpret_cast<PFN_vkVoidFunction>(&vkAcquireNextImageKHR), kDev},
    {"vkAllocateCommandBuffers", reinterpret_cast<PFN_vkVoidFunction>(&vkAllocateCommandBuffers), kDev},
    {"vkAllocateDescriptorSets", reinterpret_cast<PFN_vkVoidFunction>(&vkAllocateDescriptorSets), kDev},
    {"vkAllocateMemory", reinterpret_cast<PFN_vkVoidFunction>(&vkAllocateMemory), kDev},
    {"vkBeginCommandBuffer", reinterpret_cast<PFN_vkVoidFunction>(&vkBeginCommandBuffer), kDev},
    {"vkBindBufferMemory", reinterpret_cast<PFN_vkVoidFunction>(&vkBindBufferMemory), kDev},
    {"vkBindImageMemory", reinterpret_cast<PFN_vkVoidFunction>(&vkBindImageMemory), kDev},
    {"vkCmdBeginRenderPass", reinterpret_cast<PFN_vkVoidFunction>(&vkCmdBeginRenderPass), kDev},
    {"vkCmdBindDescriptorSets", reinterpret_cast<PFN_vkVoidFunction>(&vkCmdBindDescriptorSets), kDev},
    {"vkCmdBindIndexBuffer", reinterpret_cast<PFN_vkVoidFunction>(&vkCmdBindIndexBuffer), kDev},
    {"vkCmdBindPipeline", reinterpret_cast<PFN_vkVoidFunction>(&vkCmdBindPipeline), kDev},
    {"vkCmdBindVertexBuffers", reinterpret_cast<PFN_vkVoidFunction>(&vkCmdBindVertexBuffers), kDev},
    {"vkCmdCopyBuffer", reinterpret_cast<PFN_vkVoidFunction>(&vkCmdCopyBuffer), kDev},
    {"vkCmdCopyBufferToImage", reinterpret_cast<PFN_vkVoidFunction>(&vkCmdCopyBufferToImage), kDev},
    {"vkCmdDraw", reinterpret_cast<PFN_vkVoidFunction>(&vkCmdDraw), kDev},
    {"vkCmdDrawIndexed", reinterpret_cast<PFN_vkVoidFunction>(&vkCmdDrawIndexed), kDev},
    {"vkCmdEndRenderPass", reinterpret_cast<PFN_vkVoidFunction>(&vkCmdEndRenderPass), kDev},
    {"vkCmdPipelineBarrier", reinterpret_cast<PFN_vkVoidFunction>(&vkCmdPipelineBarrier), kDev},
    {"vkCmdPushConstants", reinterpret_cast<PFN_vkVoidFunction>(&vkCmdPushConstants), kDev},
    {"vkCmdSetScissor", reinterpret_cast<PFN_vkVoidFunction>(&vkCmdSetScissor), kDev},
    {"vkCmdSetViewport", reinterpret_cast<PFN_vkVoidFunction>(&vkCmdSetViewport), kDev},
    {"vkCreateBuffer", reinterpret_cast<PFN_vkVoidFunction>(&vkCreateBuffer), kDev},
    {"vkCreateCommandPool", reinterpret_cast<PFN_vkVoidFunction>(&vkCreateCommandPool), kDev},
    {"vkCreateComputePipelines", reinterpret_cast<PFN_vkVoidFunction>(&vkCreateComputePipelines), kDev},
    {"vkCreateDescriptorPool", reinterpret_cast<PFN_vkVoidFunction>(&vkCreateDescriptorPool), kDev},
    {"vkCreateDescriptorSetLayout", reinterpret_cast<PFN_vkVoidFunction>(&vkCreateDescriptorSetLayout), kDev},
    {"vkCreateDevice", reinterpret_cast<PFN_vkVoidFunction>(&vkCreateDevice), kInst},
    {"vkCreateFence", reinterpret_cast<PFN_vkVoidFunction>(&vkCreateFence), kDev},
    {"vkCreateFramebuffer", reinterpret_cast<PFN_vkVoidFunction>(&vkCreateFramebuffer), kDev},
    {"vkCreateGraphicsPipelines", reinterpret_cast<PFN_vkVoidFunction>(&vkCreateGraphicsPipelines), kDev},
    {"vkCreateImage", reinterpret_cast<PFN_vkVoidFunction>(&vkCreateImage), kDev},
    {"vkCreateImageView", reinterpret_cast<PFN_vkVoidFunction>(&vkCreateImageView), kDev},
    {"vkCreateInstance", reinterpret_cast<PFN_vkVoidFunction>(&vkCreateInstance), kInst},
    {"vkCreatePipelineCache", reinterpret_cast<PFN_vkVoidFunction>(&vkCreatePipelineCache), kDev},
    {"vkCreatePipelineLayout", reinterpret_cast<PFN_vkVoidFunction>(&vkCreatePipelineLayout), kDev},
    {"vkCreateRenderPass", reinterpret_cast<PFN_vkVoidFunction>(&vkCreateRenderPass), kDev},
    {"vkCreateSampler", reinterpret_cast<PFN_vkVoidFunction>(&vkCreateSampler), kDev},
    {"vkCreateSemaphore", reinterpret_cast<PFN_vkVoidFunction>(&vkCreateSemaphore), kDev},
    {"vkCreateShaderModule", reinterpret_cast<PFN_vkVoidFunction>(&vkCreateShaderModule), kDev},
    {"vkCreateSwapchainKHR", reinterpret_cast<PFN_vkVoidFunction>(&vkCreateSwapchainKHR), kDev},
    {"vkDestroyBuffer", reinterpret_cast<PFN_vkVoidFunction>(&vkDestroyBuffer), kDev},
    {"vkDestroyCommandPool", reinterpret_cast<PFN_vkVoidFunction>(&vkDestroyCommandPool), kDev},
    {"vkDestroyDescriptorPool", reinterpret_cast<PFN_vkVoidFunction>(&vkDestroyDescriptorPool), kDev},
    {"vkDestroyDescriptorSetLayout", reinterpret_cast<PFN_vkVoidFunction>(&vkDestroyDescriptorSetLayout), kDev},
    {"vkDestroyDevice", reinterpret_cast<PFN_vkVoidFunction>(&vkDestroyDevice), kDev},
    {"vkDestroyFence", reinterpret_cast<PFN_vkVoidFunction>(&vkDestroyFence), kDev},
    {"vkDestroyFramebuffer", reinterpret_cast<PFN_vkVoidFunction>(&vkDestroyFramebuffer), kDev},
    {"vkDestroyImage", reinterpret_cast<PFN_vkVoidFunction>(&vkDestroyImage), kDev},
    {"vkDestroyImageView", reinterpret_cast<PFN_vkVoidFunction>(&vkDestroyImageView), kDev},
    {"vkDestroyInstance", reinterpret_cast<PFN_vkVoidFunction>(&vkDestroyInstance), kInst},
    {"vkDestroyPipeline", reinterpret_cast<PFN_vkVoidFunction>(&vkDestroyPipeline), kDev},
    {"vkDestroyPipelineCache", reinterpret_cast<PFN_vkVoidFunction>(&vkDestroyPipelineCache), kDev},
    {"vkDestroyPipelineLayout", reinterpret_cast<PFN_vkVoidFunction>(&vkDestroyPipelineLayout), kDev},
    {"vkDestroyRenderPass", reinterpret_cast<PFN_vkVoidFunction>(&vkDestroyRenderPass), kDev},
    {"vkDestroySampler", reinterpret_cast<PFN_vkVoidFunction>(&vkDestroySampler), kDev},
    {"vkDestroySemaphore", reinterpret_cast<PFN_vkVoidFunction>(&vkDestroySemaphore), kDev},
    {"vkDestroyShaderModule", reinterpret_cast<PFN_vkVoidFunction>(&vkDestroyShaderModule), kDev},
    {"vkDestroySwapchainKHR", reinterpret_cast<PFN_vkVoidFunction>(&vkDestroySwapchainKHR), kDev},
    {"vkDeviceWaitIdle", reinterpret_cast<PFN_vkVoidFunction>(&vkDeviceWaitIdle), kDev},
    {"vkEndCommandBuffer", reinterpret_cast<PFN_vkVoidFunction>(&vkEndCommandBuffer), kDev},
    {"vkEnumerateDeviceExtensionProperties", reinterpret_cast<PFN_vkVoidFunction>(&vkEnumerateDeviceExtensionProperties), kInst},
    {"vkEnumerateInstanceExtensionProperties", reinterpret_cast<PFN_vkVoidFunction>(&vkEnumerateInstanceExtensionProperties), kInst},
    {"vkEnumerateInstanceVersion", reinterpret_cast<PFN_vkVoidFunction>(&vkEnumerateInstanceVersion), kInst},
    {"vkEnumeratePhysicalDevices", reinterpret_cast<PFN_vkVoidFunction>(&vkEnumeratePhysicalDevices), kInst},
    {"vkFlushMappedMemoryRanges", reinterpret_cast<PFN_vkVoidFunction>(&vkFlushMappedMemoryRanges), kDev},
    {"vkFreeCommandBuffers", reinterpret_cast<PFN_vkVoidFunction>(&vkFreeCommandBuffers), kDev},
    {"vkFreeDescriptorSets", reinterpret_cast<PFN_vkVoidFunction>(&vkFreeDescriptorSets), kDev},
    {"vkFreeMemory", reinterpret_cast<PFN_vkVoidFunction>(&vkFreeMemory), kDev},
    {"vkGetBufferMemoryRequirements", reinterpret_cast<PFN_vkVoidFunction>(&vkGetBufferMemoryRequirements), kDev},
    {"vkGetDeviceProcAddr", reinterpret_cast<PFN_vkVoidFunction>(&vkGetDeviceProcAddr), kInst},
    {"vkGetDeviceQueue", reinterpret_cast<PFN_vkVoidFunction>(&vkGetDeviceQueue), kDev},
    {"vkGetImageMemoryRequirements", reinterpret_cast<PFN_vkVoidFunction>(&vkGetImageMemoryRequirements), kDev},
    {"vkGetImageSubresourceLayout", reinterpret_cast<PFN_vkVoidFunction>(&vkGetImageSubresourceLayout), kDev},
    {"vkGetPhysicalDeviceFeatures", reinterpret_cast<PFN_vkVoidFunction>(&vkGetPhysicalDeviceFeatures), kPhys},
    {"vkGetPhysicalDeviceFormatProperties", reinterpret_cast<PFN_vkVoidFunction>(&vkGetPhysicalDeviceFormatProperties), kPhys},
    {"vkGetPhysicalDeviceMemoryProperties", reinterpret_cast<PFN_vkVoidFunction>(&vkGetPhysicalDeviceMemoryProperties), kPhys},
    {"vkGetPhysicalDeviceProperties", reinterpret_cast<PFN_vkVoidFunction>(&vkGetPhysicalDeviceProperties), kPhys},
    {"vkGetPhysicalDeviceQueueFamilyProperties", reinterpret_cast<PFN_vkVoidFunction>(&vkGetPhysicalDeviceQueueFamilyProperties), kPhys},
    {"vkGetPhysicalDeviceSurfaceCapabilitiesKHR", reinterpret_cast<PFN_vkVoidFunction>(&vkGetPhysicalDeviceSurfaceCapabilitiesKHR), kPhys},
    {"vkGetPhysicalDeviceSurfaceFormatsKHR", reinterpret_cast<PFN_vkVoidFunction>(&vkGetPhysicalDeviceSurfaceFormatsKHR), kPhys},
    {"vkGetPhysicalDeviceSurfacePresentModesKHR", reinterpret_cast<PFN_vkVoidFunction>(&vkGetPhysicalDeviceSurfacePresentModesKHR), kPhys},
    {"vkGetPhysicalDeviceSurfaceSupportKHR", reinterpret_cast<PFN_vkVoidFunction>(&vkGetPhysicalDeviceSurfaceSupportKHR), kPhys},
    {"vkGetSwapchainImagesKHR", reinterpret_cast<PFN_vkVoidFunction>(&vkGetSwapchainImagesKHR), kDev},
    {"vkMapMemory", reinterpret_cast<PFN_vkVoidFunction>(&vkMapMemory), kDev},
    {"vkQueuePresentKHR", reinterpret_cast<PFN_vkVoidFunction>(&vkQueuePresentKHR), kDev},
    {"vkQueueSubmit", reinterpret_cast<PFN_vkVoidFunction>(&vkQueueSubmit), kDev},
    {"vkQueueWaitIdle", reinterpret_cast<PFN_vkVoidFunction>(&vkQueueWaitIdle), kDev},
    {"vkResetFences", reinterpret_cast<PFN_vkVoidFunction>(&vkResetFences), kDev},
    {"vkUnmapMemory", reinterpret_cast<PFN_vkVoidFunction>(&vkUnmapMemory), kDev},
    {"vkUpdateDescriptorSets", reinterpret_cast<PFN_vkVoidFunction>(&vkUpdateDescriptorSets), kDev},
    {"vkWaitForFences", reinterpret_cast<PFN_vkVoidFunction>(&vkWaitForFences), kDev},
};

static PFN_vkVoidFunction LookupEntryPoint(const char *name, uint8_t need)
{
  if(!name)
    return nullptr;
  const EntryPoint *lo = kEntryPoints, *hi = kEntryPoints + sizeof(kEntryPoints) / sizeof(kEntryPoints[0]);
  while(lo < hi)
  {
    const EntryPoint *mid = lo + (hi - lo) / 2;
    int c = strcmp(mid->name, name);
    if(c == 0)
      return (mid->scope & need) == need ? mid->fn : nullptr;
    if(c < 0)
      lo = mid + 1;
    else
      hi = mid;
  }
  return nullptr;  // unknown names are NULL (the reference returns strdup(name), :119-122)
}

VKAPI_ATTR PFN_vkVoidFunction VKAPI_CALL vkGetDeviceProcAddr(VkDevice, const char *pName)
{
  return LookupEntryPoint(pName, kDev);
}

VKAPI_ATTR PFN_vkVoidFunction VKAPI_CALL vk_icdGetInstanceProcAddr(VkInstance, const char *pName)
{
  return LookupEntryPoint(pName, kInst);
}

VKAPI_ATTR PFN_vkVoidFunction VKAPI_CALL vk_icdGetPhysicalDeviceProcAddr(VkInstance, const char *pName)
{
  return LookupEntryPoint(pName, kPhys);
}

VKAPI_ATTR VkResult VKAPI_CALL vk_icdNegotiateLoaderICDInterfaceVersion(uint32_t *pSupportedVersion)
{
  if(!pSupportedVersion)
    return VK_ERROR_INITIALIZATION_FAILED;
  *pSupportedVersion = std::min<uint32_t>(VISOR_LOADER_ICD_INTERFACE_VERSION, *pSupportedVersion);
  return VK_SUCCESS;
}

// ---------------------------------------------------------------------------------------------
// vh_*: runtime controls and statistics (no counterpart in the reference)

int vh_device_available(void)
{
  std::string err;
  return Device::Get().Init(err) ? 1 : 0;
}

const char *vh_last_error(void) { return g_last_error.c_str(); }

int vh_set_exec_mode(VkQueue queue, int mode)
{
  if(!queue || (mode != 0 && mode != 1))
    return -1;
  queue->q->mode = mode == 0 ? ExecMode::Plain : ExecMode::Batched;
  return 0;
}

int vh_set_graph_mode(VkQueue queue, int on)
{
  if(!queue || (on != 0 && on != 1))
    return -1;
  queue->q->graphs = on != 0;
  return 0;
}

int vh_set_tile_size(VkQueue queue, int tile)
{
  if(!queue)
    return -1;
  if(tile == 0)
  {
    queue->q->autotune = true;
    return 0;
  }
  uint32_t shift = tile == 8 ? 3u : (tile == 16 ? 4u : (tile == 32 ? 5u : 0u));
  if(!shift)
    return -1;
  queue->q->tiles.tile_shift = shift;
  queue->q->autotune = false;  // an explicit launch shape replaces the autotuner's choice
  return 0;
}

int vh_set_async_submit(VkQueue queue, int on)
{
  if(!queue)
    return -1;
  queue->q->WaitIdle();
  queue->q->async = on != 0;
  return 0;
}

int vh_get_stats(VkQueue queue, int which, VhStats *out)
{
  if(!queue || !out)
    return -1;
  if(which < 0 || which > 3)
    return -1;
  const ExecStats &s = which == 0 ? queue->q->last_stats
                       : which == 1 ? queue->q->accum_stats
                       : which == 2 ? queue->q->presented_stats
                                    : queue->q->frame_stats;
  out->gpu_ms = s.gpu_ms_total;
  out->raster_ms = s.raster_ms;
  out->setup_ms = s.setup_ms;
  out->bin_ms = s.bin_ms;
  out->raster_launches = s.raster_launches;
  out->triangles = s.total_tris;
  out->bin_pairs = s.total_pairs;
  out->tiles = s.total_tiles;
  out->draws = s.ctr.draws;
  out->tris_in = s.ctr.tris_in;
  out->tris_out = s.ctr.tris_out;
  out->blocks = s.ctr.blocks;
  out->pixels_tested = s.ctr.pixels_tested;
  out->pixels_written = s.ctr.pixels_written;
  out->depth_passed = s.ctr.depth_passed;
  out->shaded_blend = s.ctr.shaded_blend;
  out->shaded_dwrite = s.ctr.shaded_dwrite;
  out->graph_launches = s.graph_launches;
  out->tile_size = 1ull << s.last_cfg.tile_shift;
  out->chunk = s.last_cfg.chunk;
  out->persistent = s.last_cfg.persistent;
  out->tuning = s.last_tuning ? 1u : 0u;
  return 0;
}

int vh_swapchain_frame(VkSwapchainKHR swapchain, uint32_t *image_index, uint64_t *presents, const void **pixels)
{
  if(!swapchain || swapchain->images.empty())
    return -1;
  if(image_index)
    *image_index = swapchain->presented;
  if(presents)
    *presents = swapchain->presents;
  if(pixels)
    *pixels = swapchain->presents ? swapchain->mems[swapchain->presented]->host : nullptr;
  return 0;
}

int vh_reset_stats(VkQueue queue)
{
  if(!queue)
    return -1;
  queue->q->accum_stats = ExecStats();
  queue->q->frame_stats = ExecStats();
  return 0;
}

}  // extern "C"

namespace visor
{
bool BuildModuleOffline(const std::vector<std::shared_ptr<ShaderPair>> &ps, const std::string &arch, const std::string &cache_dir,
                        std::string &err, size_t *code_size, std::string *src_out);
}

extern "C" int vh_translate_spirv(const uint32_t *code, size_t nwords, int stage, char *out, size_t out_size)
{
  TranslatedShader t;
  std::string err;
  bool ok = TranslateSpirv(code, nwords, "main", stage == 4 ? ShaderStage::Fragment : ShaderStage::Vertex, "va_shader", t, err);
  std::string s = ok ? t.code : err;
  if(out && out_size)
  {
    size_t n = std::min(out_size - 1, s.size());
    memcpy(out, s.data(), n);
    out[n] = 0;
  }
  return ok ? (int)s.size() : -1;
}

extern "C" int vh_build_pipeline_offline(const uint32_t *vs, size_t vs_words, const uint32_t *fs, size_t fs_words,
                                         const char *arch, char *err_out, size_t err_size)
{
  VkShaderModule_T v, f;
  v.code.assign(vs, vs + vs_words);
  f.code.assign(fs, fs + fs_words);
  std::string err;
  auto pair = Device::Get().MakePair(&v, "main", &f, "main", err);
  size_t size = 0;
  bool ok = pair && BuildModuleOffline({pair}, arch ? arch : "gfx950", "", err, &size, nullptr);
  if(err_out && err_size)
  {
    size_t n = std::min(err_size - 1, err.size());
    memcpy(err_out, err.data(), n);
    err_out[n] = 0;
  }
  return ok ? (int)size : -1;
}
