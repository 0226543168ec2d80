"""ctypes bindings of the Vulkan 1.0 subset the ICD implements (mirror of include/visor_vk.h).

Entry points are resolved the way the Vulkan loader does it: through vk_icdGetInstanceProcAddr
of libvisor_mi355x.so. `load()` fails loudly when the library has not been built.
"""

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libvisor_mi355x.so")

u32, i32, u64, f32, size_t = C.c_uint32, C.c_int32, C.c_uint64, C.c_float, C.c_size_t
vp, cp = C.c_void_p, C.c_char_p
Handle = C.c_void_p

# ---- enums used by the package and its tests -------------------------------------------------
VK_SUCCESS, VK_INCOMPLETE, VK_TIMEOUT = 0, 5, 2
ST = dict(APPLICATION_INFO=0, INSTANCE_CREATE_INFO=1, DEVICE_QUEUE_CREATE_INFO=2, DEVICE_CREATE_INFO=3,
          SUBMIT_INFO=4, MEMORY_ALLOCATE_INFO=5, FENCE_CREATE_INFO=8, SEMAPHORE_CREATE_INFO=9,
          BUFFER_CREATE_INFO=12, IMAGE_CREATE_INFO=14, IMAGE_VIEW_CREATE_INFO=15,
          SHADER_MODULE_CREATE_INFO=16, PIPELINE_SHADER_STAGE_CREATE_INFO=18,
          PIPELINE_VERTEX_INPUT_STATE_CREATE_INFO=19, PIPELINE_INPUT_ASSEMBLY_STATE_CREATE_INFO=20,
          PIPELINE_VIEWPORT_STATE_CREATE_INFO=22, PIPELINE_RASTERIZATION_STATE_CREATE_INFO=23,
          PIPELINE_MULTISAMPLE_STATE_CREATE_INFO=24, PIPELINE_DEPTH_STENCIL_STATE_CREATE_INFO=25,
          PIPELINE_COLOR_BLEND_STATE_CREATE_INFO=26, GRAPHICS_PIPELINE_CREATE_INFO=28,
          PIPELINE_LAYOUT_CREATE_INFO=30, SAMPLER_CREATE_INFO=31, DESCRIPTOR_SET_LAYOUT_CREATE_INFO=32,
          DESCRIPTOR_POOL_CREATE_INFO=33, DESCRIPTOR_SET_ALLOCATE_INFO=34, WRITE_DESCRIPTOR_SET=35,
          FRAMEBUFFER_CREATE_INFO=37, RENDER_PASS_CREATE_INFO=38, COMMAND_POOL_CREATE_INFO=39,
          COMMAND_BUFFER_ALLOCATE_INFO=40, COMMAND_BUFFER_BEGIN_INFO=42, RENDER_PASS_BEGIN_INFO=43,
          SWAPCHAIN_CREATE_INFO_KHR=1000001000, PRESENT_INFO_KHR=1000001001)

FMT = dict(R8_UNORM=9, R8G8B8A8_UNORM=37, B8G8R8A8_UNORM=44, R16G16B16A16_SFLOAT=97, R32_SFLOAT=100,
           R32G32_SFLOAT=103, R32G32B32_SFLOAT=106, R32G32B32A32_SFLOAT=109, D32_SFLOAT=126,
           BC2_UNORM_BLOCK=135, BC3_UNORM_BLOCK=137)
IMAGE_CREATE_CUBE_COMPATIBLE = 0x10
VIEW_TYPE_2D, VIEW_TYPE_CUBE, VIEW_TYPE_2D_ARRAY = 1, 3, 5


def texel_bytes(fmt):
    """Bytes per pixel of the linear image layout (1 for R8 and the 4x4-block BC2/BC3 formats)."""
    return {9: 1, 135: 1, 137: 1, 97: 8}.get(fmt, 4)
TOPO_LIST, TOPO_STRIP = 3, 4
CULL_NONE, CULL_FRONT, CULL_BACK = 0, 1, 2
FRONT_CCW, FRONT_CW = 0, 1
CMP = dict(NEVER=0, LESS=1, EQUAL=2, LEQUAL=3, GREATER=4, NOTEQUAL=5, GEQUAL=6, ALWAYS=7)
BF = dict(ZERO=0, ONE=1, SRC_ALPHA=6, ONE_MINUS_SRC_ALPHA=7)
LOAD_OP_LOAD, LOAD_OP_CLEAR = 0, 1
STAGE_VERTEX, STAGE_FRAGMENT = 1, 0x10
DESC_COMBINED_IMAGE_SAMPLER, DESC_UNIFORM_BUFFER = 1, 6
INDEX_U16, INDEX_U32 = 0, 1
FILTER_NEAREST, FILTER_LINEAR = 0, 1
MIPMAP_NEAREST, MIPMAP_LINEAR = 0, 1
ADDR_REPEAT, ADDR_CLAMP = 0, 2
MEM_DEVICE_LOCAL, MEM_HOST_VISIBLE = 0, 1


class S(C.Structure):
    """Structure with keyword construction and sType defaulting."""

    def __init__(self, **kw):
        super().__init__()
        for k, v in kw.items():
            setattr(self, k, v)


def _st(name, fields, stype=None):
    cls = type(name, (S,), {"_fields_": fields})
    if stype is not None:
        cls._stype = ST[stype]
        orig = cls.__init__

        def init(self, **kw):
            orig(self, **kw)
            self.sType = cls._stype
        cls.__init__ = init
    return cls


Offset2D = _st("Offset2D", [("x", i32), ("y", i32)])
Offset3D = _st("Offset3D", [("x", i32), ("y", i32), ("z", i32)])
Extent2D = _st("Extent2D", [("width", u32), ("height", u32)])
Extent3D = _st("Extent3D", [("width", u32), ("height", u32), ("depth", u32)])
Rect2D = _st("Rect2D", [("offset", Offset2D), ("extent", Extent2D)])
InstanceCreateInfo = _st("InstanceCreateInfo", [("sType", u32), ("pNext", vp), ("flags", u32), ("pApplicationInfo", vp),
                                                ("enabledLayerCount", u32), ("ppEnabledLayerNames", vp),
                                                ("enabledExtensionCount", u32), ("ppEnabledExtensionNames", vp)],
                         "INSTANCE_CREATE_INFO")
DeviceQueueCreateInfo = _st("DeviceQueueCreateInfo", [("sType", u32), ("pNext", vp), ("flags", u32), ("queueFamilyIndex", u32),
                                                      ("queueCount", u32), ("pQueuePriorities", C.POINTER(f32))],
                            "DEVICE_QUEUE_CREATE_INFO")
DeviceCreateInfo = _st("DeviceCreateInfo", [("sType", u32), ("pNext", vp), ("flags", u32), ("queueCreateInfoCount", u32),
                                            ("pQueueCreateInfos", C.POINTER(DeviceQueueCreateInfo)), ("enabledLayerCount", u32),
                                            ("ppEnabledLayerNames", vp), ("enabledExtensionCount", u32),
                                            ("ppEnabledExtensionNames", vp), ("pEnabledFeatures", vp)], "DEVICE_CREATE_INFO")
SwapchainCreateInfoKHR = _st("SwapchainCreateInfoKHR", [
    ("sType", u32), ("pNext", vp), ("flags", u32), ("surface", Handle), ("minImageCount", u32), ("imageFormat", u32),
    ("imageColorSpace", u32), ("imageExtent", Extent2D), ("imageArrayLayers", u32), ("imageUsage", u32),
    ("imageSharingMode", u32), ("queueFamilyIndexCount", u32), ("pQueueFamilyIndices", vp), ("preTransform", u32),
    ("compositeAlpha", u32), ("presentMode", u32), ("clipped", u32), ("oldSwapchain", Handle)], "SWAPCHAIN_CREATE_INFO_KHR")
PresentInfoKHR = _st("PresentInfoKHR", [
    ("sType", u32), ("pNext", vp), ("waitSemaphoreCount", u32), ("pWaitSemaphores", vp), ("swapchainCount", u32),
    ("pSwapchains", C.POINTER(Handle)), ("pImageIndices", C.POINTER(u32)), ("pResults", C.POINTER(i32))],
    "PRESENT_INFO_KHR")
SubmitInfo = _st("SubmitInfo", [("sType", u32), ("pNext", vp), ("waitSemaphoreCount", u32), ("pWaitSemaphores", vp),
                                ("pWaitDstStageMask", vp), ("commandBufferCount", u32), ("pCommandBuffers", C.POINTER(Handle)),
                                ("signalSemaphoreCount", u32), ("pSignalSemaphores", vp)], "SUBMIT_INFO")
MemoryAllocateInfo = _st("MemoryAllocateInfo", [("sType", u32), ("pNext", vp), ("allocationSize", u64), ("memoryTypeIndex", u32)],
                         "MEMORY_ALLOCATE_INFO")
MemoryRequirements = _st("MemoryRequirements", [("size", u64), ("alignment", u64), ("memoryTypeBits", u32)])
FenceCreateInfo = _st("FenceCreateInfo", [("sType", u32), ("pNext", vp), ("flags", u32)], "FENCE_CREATE_INFO")
BufferCreateInfo = _st("BufferCreateInfo", [("sType", u32), ("pNext", vp), ("flags", u32), ("size", u64), ("usage", u32),
                                            ("sharingMode", u32), ("queueFamilyIndexCount", u32), ("pQueueFamilyIndices", vp)],
                       "BUFFER_CREATE_INFO")
ImageCreateInfo = _st("ImageCreateInfo", [("sType", u32), ("pNext", vp), ("flags", u32), ("imageType", u32), ("format", u32),
                                          ("extent", Extent3D), ("mipLevels", u32), ("arrayLayers", u32), ("samples", u32),
                                          ("tiling", u32), ("usage", u32), ("sharingMode", u32), ("queueFamilyIndexCount", u32),
                                          ("pQueueFamilyIndices", vp), ("initialLayout", u32)], "IMAGE_CREATE_INFO")
ComponentMapping = _st("ComponentMapping", [("r", u32), ("g", u32), ("b", u32), ("a", u32)])
ImageSubresourceRange = _st("ImageSubresourceRange", [("aspectMask", u32), ("baseMipLevel", u32), ("levelCount", u32),
                                                      ("baseArrayLayer", u32), ("layerCount", u32)])
ImageViewCreateInfo = _st("ImageViewCreateInfo", [("sType", u32), ("pNext", vp), ("flags", u32), ("image", Handle),
                                                  ("viewType", u32), ("format", u32), ("components", ComponentMapping),
                                                  ("subresourceRange", ImageSubresourceRange)], "IMAGE_VIEW_CREATE_INFO")
SamplerCreateInfo = _st("SamplerCreateInfo", [("sType", u32), ("pNext", vp), ("flags", u32), ("magFilter", u32), ("minFilter", u32),
                                              ("mipmapMode", u32), ("addressModeU", u32), ("addressModeV", u32),
                                              ("addressModeW", u32), ("mipLodBias", f32), ("anisotropyEnable", u32),
                                              ("maxAnisotropy", f32), ("compareEnable", u32), ("compareOp", u32), ("minLod", f32),
                                              ("maxLod", f32), ("borderColor", u32), ("unnormalizedCoordinates", u32)],
                        "SAMPLER_CREATE_INFO")
ShaderModuleCreateInfo = _st("ShaderModuleCreateInfo", [("sType", u32), ("pNext", vp), ("flags", u32), ("codeSize", size_t),
                                                        ("pCode", vp)], "SHADER_MODULE_CREATE_INFO")
PipelineShaderStageCreateInfo = _st("PipelineShaderStageCreateInfo", [("sType", u32), ("pNext", vp), ("flags", u32), ("stage", u32),
                                                                      ("module", Handle), ("pName", cp),
                                                                      ("pSpecializationInfo", vp)], "PIPELINE_SHADER_STAGE_CREATE_INFO")
VertexInputBindingDescription = _st("VertexInputBindingDescription", [("binding", u32), ("stride", u32), ("inputRate", u32)])
VertexInputAttributeDescription = _st("VertexInputAttributeDescription", [("location", u32), ("binding", u32), ("format", u32),
                                                                          ("offset", u32)])
PipelineVertexInputStateCreateInfo = _st("PipelineVertexInputStateCreateInfo", [
    ("sType", u32), ("pNext", vp), ("flags", u32), ("vertexBindingDescriptionCount", u32),
    ("pVertexBindingDescriptions", C.POINTER(VertexInputBindingDescription)), ("vertexAttributeDescriptionCount", u32),
    ("pVertexAttributeDescriptions", C.POINTER(VertexInputAttributeDescription))], "PIPELINE_VERTEX_INPUT_STATE_CREATE_INFO")
PipelineInputAssemblyStateCreateInfo = _st("PipelineInputAssemblyStateCreateInfo", [
    ("sType", u32), ("pNext", vp), ("flags", u32), ("topology", u32), ("primitiveRestartEnable", u32)],
    "PIPELINE_INPUT_ASSEMBLY_STATE_CREATE_INFO")
Viewport = _st("Viewport", [("x", f32), ("y", f32), ("width", f32), ("height", f32), ("minDepth", f32), ("maxDepth", f32)])
PipelineViewportStateCreateInfo = _st("PipelineViewportStateCreateInfo", [
    ("sType", u32), ("pNext", vp), ("flags", u32), ("viewportCount", u32), ("pViewports", C.POINTER(Viewport)),
    ("scissorCount", u32), ("pScissors", C.POINTER(Rect2D))], "PIPELINE_VIEWPORT_STATE_CREATE_INFO")
PipelineRasterizationStateCreateInfo = _st("PipelineRasterizationStateCreateInfo", [
    ("sType", u32), ("pNext", vp), ("flags", u32), ("depthClampEnable", u32), ("rasterizerDiscardEnable", u32),
    ("polygonMode", u32), ("cullMode", u32), ("frontFace", u32), ("depthBiasEnable", u32), ("depthBiasConstantFactor", f32),
    ("depthBiasClamp", f32), ("depthBiasSlopeFactor", f32), ("lineWidth", f32)], "PIPELINE_RASTERIZATION_STATE_CREATE_INFO")
PipelineMultisampleStateCreateInfo = _st("PipelineMultisampleStateCreateInfo", [
    ("sType", u32), ("pNext", vp), ("flags", u32), ("rasterizationSamples", u32), ("sampleShadingEnable", u32),
    ("minSampleShading", f32), ("pSampleMask", vp), ("alphaToCoverageEnable", u32), ("alphaToOneEnable", u32)],
    "PIPELINE_MULTISAMPLE_STATE_CREATE_INFO")
StencilOpState = _st("StencilOpState", [("failOp", u32), ("passOp", u32), ("depthFailOp", u32), ("compareOp", u32),
                                        ("compareMask", u32), ("writeMask", u32), ("reference", u32)])
PipelineDepthStencilStateCreateInfo = _st("PipelineDepthStencilStateCreateInfo", [
    ("sType", u32), ("pNext", vp), ("flags", u32), ("depthTestEnable", u32), ("depthWriteEnable", u32), ("depthCompareOp", u32),
    ("depthBoundsTestEnable", u32), ("stencilTestEnable", u32), ("front", StencilOpState), ("back", StencilOpState),
    ("minDepthBounds", f32), ("maxDepthBounds", f32)], "PIPELINE_DEPTH_STENCIL_STATE_CREATE_INFO")
PipelineColorBlendAttachmentState = _st("PipelineColorBlendAttachmentState", [
    ("blendEnable", u32), ("srcColorBlendFactor", u32), ("dstColorBlendFactor", u32), ("colorBlendOp", u32),
    ("srcAlphaBlendFactor", u32), ("dstAlphaBlendFactor", u32), ("alphaBlendOp", u32), ("colorWriteMask", u32)])
PipelineColorBlendStateCreateInfo = _st("PipelineColorBlendStateCreateInfo", [
    ("sType", u32), ("pNext", vp), ("flags", u32), ("logicOpEnable", u32), ("logicOp", u32), ("attachmentCount", u32),
    ("pAttachments", C.POINTER(PipelineColorBlendAttachmentState)), ("blendConstants", f32 * 4)],
    "PIPELINE_COLOR_BLEND_STATE_CREATE_INFO")
GraphicsPipelineCreateInfo = _st("GraphicsPipelineCreateInfo", [
    ("sType", u32), ("pNext", vp), ("flags", u32), ("stageCount", u32), ("pStages", C.POINTER(PipelineShaderStageCreateInfo)),
    ("pVertexInputState", C.POINTER(PipelineVertexInputStateCreateInfo)),
    ("pInputAssemblyState", C.POINTER(PipelineInputAssemblyStateCreateInfo)), ("pTessellationState", vp),
    ("pViewportState", C.POINTER(PipelineViewportStateCreateInfo)),
    ("pRasterizationState", C.POINTER(PipelineRasterizationStateCreateInfo)),
    ("pMultisampleState", C.POINTER(PipelineMultisampleStateCreateInfo)),
    ("pDepthStencilState", C.POINTER(PipelineDepthStencilStateCreateInfo)),
    ("pColorBlendState", C.POINTER(PipelineColorBlendStateCreateInfo)), ("pDynamicState", vp), ("layout", Handle),
    ("renderPass", Handle), ("subpass", u32), ("basePipelineHandle", Handle), ("basePipelineIndex", i32)],
    "GRAPHICS_PIPELINE_CREATE_INFO")
PushConstantRange = _st("PushConstantRange", [("stageFlags", u32), ("offset", u32), ("size", u32)])
PipelineLayoutCreateInfo = _st("PipelineLayoutCreateInfo", [("sType", u32), ("pNext", vp), ("flags", u32), ("setLayoutCount", u32),
                                                            ("pSetLayouts", C.POINTER(Handle)), ("pushConstantRangeCount", u32),
                                                            ("pPushConstantRanges", C.POINTER(PushConstantRange))],
                               "PIPELINE_LAYOUT_CREATE_INFO")
DescriptorSetLayoutBinding = _st("DescriptorSetLayoutBinding", [("binding", u32), ("descriptorType", u32), ("descriptorCount", u32),
                                                                ("stageFlags", u32), ("pImmutableSamplers", vp)])
DescriptorSetLayoutCreateInfo = _st("DescriptorSetLayoutCreateInfo", [("sType", u32), ("pNext", vp), ("flags", u32),
                                                                      ("bindingCount", u32),
                                                                      ("pBindings", C.POINTER(DescriptorSetLayoutBinding))],
                                    "DESCRIPTOR_SET_LAYOUT_CREATE_INFO")
DescriptorPoolCreateInfo = _st("DescriptorPoolCreateInfo", [("sType", u32), ("pNext", vp), ("flags", u32), ("maxSets", u32),
                                                            ("poolSizeCount", u32), ("pPoolSizes", vp)], "DESCRIPTOR_POOL_CREATE_INFO")
DescriptorSetAllocateInfo = _st("DescriptorSetAllocateInfo", [("sType", u32), ("pNext", vp), ("descriptorPool", Handle),
                                                              ("descriptorSetCount", u32), ("pSetLayouts", C.POINTER(Handle))],
                                "DESCRIPTOR_SET_ALLOCATE_INFO")
DescriptorImageInfo = _st("DescriptorImageInfo", [("sampler", Handle), ("imageView", Handle), ("imageLayout", u32)])
DescriptorBufferInfo = _st("DescriptorBufferInfo", [("buffer", Handle), ("offset", u64), ("range", u64)])
WriteDescriptorSet = _st("WriteDescriptorSet", [("sType", u32), ("pNext", vp), ("dstSet", Handle), ("dstBinding", u32),
                                                ("dstArrayElement", u32), ("descriptorCount", u32), ("descriptorType", u32),
                                                ("pImageInfo", C.POINTER(DescriptorImageInfo)),
                                                ("pBufferInfo", C.POINTER(DescriptorBufferInfo)), ("pTexelBufferView", vp)],
                         "WRITE_DESCRIPTOR_SET")
FramebufferCreateInfo = _st("FramebufferCreateInfo", [("sType", u32), ("pNext", vp), ("flags", u32), ("renderPass", Handle),
                                                      ("attachmentCount", u32), ("pAttachments", C.POINTER(Handle)), ("width", u32),
                                                      ("height", u32), ("layers", u32)], "FRAMEBUFFER_CREATE_INFO")
AttachmentDescription = _st("AttachmentDescription", [("flags", u32), ("format", u32), ("samples", u32), ("loadOp", u32),
                                                      ("storeOp", u32), ("stencilLoadOp", u32), ("stencilStoreOp", u32),
                                                      ("initialLayout", u32), ("finalLayout", u32)])
AttachmentReference = _st("AttachmentReference", [("attachment", u32), ("layout", u32)])
SubpassDescription = _st("SubpassDescription", [("flags", u32), ("pipelineBindPoint", u32), ("inputAttachmentCount", u32),
                                                ("pInputAttachments", vp), ("colorAttachmentCount", u32),
                                                ("pColorAttachments", C.POINTER(AttachmentReference)), ("pResolveAttachments", vp),
                                                ("pDepthStencilAttachment", C.POINTER(AttachmentReference)),
                                                ("preserveAttachmentCount", u32), ("pPreserveAttachments", vp)])
RenderPassCreateInfo = _st("RenderPassCreateInfo", [("sType", u32), ("pNext", vp), ("flags", u32), ("attachmentCount", u32),
                                                    ("pAttachments", C.POINTER(AttachmentDescription)), ("subpassCount", u32),
                                                    ("pSubpasses", C.POINTER(SubpassDescription)), ("dependencyCount", u32),
                                                    ("pDependencies", vp)], "RENDER_PASS_CREATE_INFO")
CommandPoolCreateInfo = _st("CommandPoolCreateInfo", [("sType", u32), ("pNext", vp), ("flags", u32), ("queueFamilyIndex", u32)],
                            "COMMAND_POOL_CREATE_INFO")
CommandBufferAllocateInfo = _st("CommandBufferAllocateInfo", [("sType", u32), ("pNext", vp), ("commandPool", Handle), ("level", u32),
                                                              ("commandBufferCount", u32)], "COMMAND_BUFFER_ALLOCATE_INFO")
CommandBufferBeginInfo = _st("CommandBufferBeginInfo", [("sType", u32), ("pNext", vp), ("flags", u32), ("pInheritanceInfo", vp)],
                             "COMMAND_BUFFER_BEGIN_INFO")


class ClearValue(C.Union):
    _fields_ = [("color", f32 * 4), ("depthStencil", f32 * 2)]


RenderPassBeginInfo = _st("RenderPassBeginInfo", [("sType", u32), ("pNext", vp), ("renderPass", Handle), ("framebuffer", Handle),
                                                  ("renderArea", Rect2D), ("clearValueCount", u32),
                                                  ("pClearValues", C.POINTER(ClearValue))], "RENDER_PASS_BEGIN_INFO")
ImageSubresourceLayers = _st("ImageSubresourceLayers", [("aspectMask", u32), ("mipLevel", u32), ("baseArrayLayer", u32),
                                                        ("layerCount", u32)])
BufferImageCopy = _st("BufferImageCopy", [("bufferOffset", u64), ("bufferRowLength", u32), ("bufferImageHeight", u32),
                                          ("imageSubresource", ImageSubresourceLayers), ("imageOffset", Offset3D),
                                          ("imageExtent", Extent3D)])
BufferCopy = _st("BufferCopy", [("srcOffset", u64), ("dstOffset", u64), ("size", u64)])
ExtensionProperties = _st("ExtensionProperties", [("extensionName", C.c_char * 256), ("specVersion", u32)])


class VhStats(C.Structure):
    _fields_ = [("gpu_ms", C.c_double), ("raster_ms", C.c_double), ("setup_ms", C.c_double), ("bin_ms", C.c_double),
                ("raster_launches", u64), ("triangles", u64), ("bin_pairs", u64), ("tiles", u64), ("draws", u64),
                ("tris_in", u64), ("tris_out", u64), ("blocks", u64), ("pixels_tested", u64), ("pixels_written", u64),
                ("depth_passed", u64), ("shaded_blend", u64), ("shaded_dwrite", u64),
                ("graph_launches", u64), ("tile_size", u64), ("chunk", u64),
                ("persistent", u64), ("tuning", u64)]


class VhServerConfig(C.Structure):
    _fields_ = [("hbm_budget_bytes", u64), ("slots", u32), ("max_batch", u32)]


class VhServerStats(C.Structure):
    _fields_ = [("submitted", u64), ("completed", u64), ("batches", u64), ("fragments", u64), ("shaded", u64),
                ("max_inflight_bytes", u64), ("busy_ms", C.c_double), ("mean_batch", C.c_double),
                ("p50_ms", C.c_double), ("p99_ms", C.c_double), ("mean_ms", C.c_double), ("max_ms", C.c_double)]


P = C.POINTER
# (name, restype, argtypes); None restype = void
PROTOS = [
    ("vkCreateInstance", i32, [P(InstanceCreateInfo), vp, P(Handle)]),
    ("vkDestroyInstance", None, [Handle, vp]),
    ("vkEnumerateInstanceVersion", i32, [P(u32)]),
    ("vkEnumerateInstanceExtensionProperties", i32, [cp, P(u32), vp]),
    ("vkEnumerateDeviceExtensionProperties", i32, [Handle, cp, P(u32), vp]),
    ("vkEnumeratePhysicalDevices", i32, [Handle, P(u32), P(Handle)]),
    ("vkGetPhysicalDeviceFeatures", None, [Handle, vp]),
    ("vkGetPhysicalDeviceProperties", None, [Handle, vp]),
    ("vkGetPhysicalDeviceQueueFamilyProperties", None, [Handle, P(u32), vp]),
    ("vkGetPhysicalDeviceMemoryProperties", None, [Handle, vp]),
    ("vkGetPhysicalDeviceFormatProperties", None, [Handle, u32, vp]),
    ("vkCreateDevice", i32, [Handle, P(DeviceCreateInfo), vp, P(Handle)]),
    ("vkDestroyDevice", None, [Handle, vp]),
    ("vkGetDeviceQueue", None, [Handle, u32, u32, P(Handle)]),
    ("vkCreateFence", i32, [Handle, P(FenceCreateInfo), vp, P(Handle)]),
    ("vkDestroyFence", None, [Handle, Handle, vp]),
    ("vkResetFences", i32, [Handle, u32, P(Handle)]),
    ("vkWaitForFences", i32, [Handle, u32, P(Handle), u32, u64]),
    ("vkQueueWaitIdle", i32, [Handle]),
    ("vkDeviceWaitIdle", i32, [Handle]),
    ("vkAllocateMemory", i32, [Handle, P(MemoryAllocateInfo), vp, P(Handle)]),
    ("vkFreeMemory", None, [Handle, Handle, vp]),
    ("vkMapMemory", i32, [Handle, Handle, u64, u64, u32, P(vp)]),
    ("vkUnmapMemory", None, [Handle, Handle]),
    ("vkCreateBuffer", i32, [Handle, P(BufferCreateInfo), vp, P(Handle)]),
    ("vkDestroyBuffer", None, [Handle, Handle, vp]),
    ("vkGetBufferMemoryRequirements", None, [Handle, Handle, P(MemoryRequirements)]),
    ("vkBindBufferMemory", i32, [Handle, Handle, Handle, u64]),
    ("vkCreateImage", i32, [Handle, P(ImageCreateInfo), vp, P(Handle)]),
    ("vkDestroyImage", None, [Handle, Handle, vp]),
    ("vkGetImageMemoryRequirements", None, [Handle, Handle, P(MemoryRequirements)]),
    ("vkBindImageMemory", i32, [Handle, Handle, Handle, u64]),
    ("vkCreateImageView", i32, [Handle, P(ImageViewCreateInfo), vp, P(Handle)]),
    ("vkDestroyImageView", None, [Handle, Handle, vp]),
    ("vkCreateSampler", i32, [Handle, P(SamplerCreateInfo), vp, P(Handle)]),
    ("vkDestroySampler", None, [Handle, Handle, vp]),
    ("vkCreateShaderModule", i32, [Handle, P(ShaderModuleCreateInfo), vp, P(Handle)]),
    ("vkDestroyShaderModule", None, [Handle, Handle, vp]),
    ("vkCreateDescriptorSetLayout", i32, [Handle, P(DescriptorSetLayoutCreateInfo), vp, P(Handle)]),
    ("vkCreatePipelineLayout", i32, [Handle, P(PipelineLayoutCreateInfo), vp, P(Handle)]),
    ("vkCreateDescriptorPool", i32, [Handle, P(DescriptorPoolCreateInfo), vp, P(Handle)]),
    ("vkAllocateDescriptorSets", i32, [Handle, P(DescriptorSetAllocateInfo), P(Handle)]),
    ("vkUpdateDescriptorSets", None, [Handle, u32, P(WriteDescriptorSet), u32, vp]),
    ("vkCreateRenderPass", i32, [Handle, P(RenderPassCreateInfo), vp, P(Handle)]),
    ("vkCreateFramebuffer", i32, [Handle, P(FramebufferCreateInfo), vp, P(Handle)]),
    ("vkCreateGraphicsPipelines", i32, [Handle, Handle, u32, P(GraphicsPipelineCreateInfo), vp, P(Handle)]),
    ("vkDestroyPipeline", None, [Handle, Handle, vp]),
    ("vkCreateCommandPool", i32, [Handle, P(CommandPoolCreateInfo), vp, P(Handle)]),
    ("vkAllocateCommandBuffers", i32, [Handle, P(CommandBufferAllocateInfo), P(Handle)]),
    ("vkFreeCommandBuffers", None, [Handle, Handle, u32, P(Handle)]),
    ("vkBeginCommandBuffer", i32, [Handle, P(CommandBufferBeginInfo)]),
    ("vkEndCommandBuffer", i32, [Handle]),
    ("vkCmdBeginRenderPass", None, [Handle, P(RenderPassBeginInfo), u32]),
    ("vkCmdEndRenderPass", None, [Handle]),
    ("vkCmdBindPipeline", None, [Handle, u32, Handle]),
    ("vkCmdBindDescriptorSets", None, [Handle, u32, Handle, u32, u32, P(Handle), u32, vp]),
    ("vkCmdBindVertexBuffers", None, [Handle, u32, u32, P(Handle), P(u64)]),
    ("vkCmdBindIndexBuffer", None, [Handle, Handle, u64, u32]),
    ("vkCmdPushConstants", None, [Handle, Handle, u32, u32, u32, vp]),
    ("vkCmdDraw", None, [Handle, u32, u32, u32, u32]),
    ("vkCmdDrawIndexed", None, [Handle, u32, u32, u32, i32, u32]),
    ("vkCmdCopyBufferToImage", None, [Handle, Handle, Handle, u32, u32, P(BufferImageCopy)]),
    ("vkCmdCopyBuffer", None, [Handle, Handle, Handle, u32, P(BufferCopy)]),
    ("vkCmdPipelineBarrier", None, [Handle, u32, u32, u32, u32, vp, u32, vp, u32, vp]),
    ("vkQueueSubmit", i32, [Handle, u32, P(SubmitInfo), Handle]),
    ("vkCreateSwapchainKHR", i32, [Handle, P(SwapchainCreateInfoKHR), vp, P(Handle)]),
    ("vkDestroySwapchainKHR", None, [Handle, Handle, vp]),
    ("vkGetSwapchainImagesKHR", i32, [Handle, Handle, P(u32), P(Handle)]),
    ("vkAcquireNextImageKHR", i32, [Handle, Handle, u64, Handle, Handle, P(u32)]),
    ("vkQueuePresentKHR", i32, [Handle, P(PresentInfoKHR)]),
]

EXT_PROTOS = [
    ("vh_device_available", C.c_int, []),
    ("vh_last_error", cp, []),
    ("vh_set_exec_mode", C.c_int, [Handle, C.c_int]),
    ("vh_set_tile_size", C.c_int, [Handle, C.c_int]),
    ("vh_set_graph_mode", C.c_int, [Handle, C.c_int]),
    ("vh_server_create", i32, [Handle, P(VhServerConfig), P(vp)]),
    ("vh_server_destroy", None, [vp]),
    ("vh_server_submit", i32, [vp, Handle, P(u64)]),
    ("vh_server_wait", i32, [vp, u64, u64, P(C.c_double)]),
    ("vh_server_stats", C.c_int, [vp, P(VhServerStats)]),
    ("vh_server_reset_stats", C.c_int, [vp]),
    ("vh_set_async_submit", C.c_int, [Handle, C.c_int]),
    ("vh_get_stats", C.c_int, [Handle, C.c_int, P(VhStats)]),
    ("vh_reset_stats", C.c_int, [Handle]),
    ("vh_swapchain_frame", C.c_int, [Handle, P(u32), P(u64), P(vp)]),
    ("vh_translate_spirv", C.c_int, [P(u32), size_t, C.c_int, C.c_char_p, size_t]),
    ("vh_build_pipeline_offline", C.c_int, [P(u32), size_t, P(u32), size_t, cp, C.c_char_p, size_t]),
]


class VkError(RuntimeError):
    pass


class ICD:
    """The ICD library with every entry point resolved through vk_icdGetInstanceProcAddr."""

    def __init__(self, path=LIB_PATH):
        if not os.path.exists(path):
            raise RuntimeError("libvisor_mi355x.so is not built (run python -m visor_amd.build): " + path)
        self.lib = C.CDLL(path)
        self.lib.vk_icdGetInstanceProcAddr.restype = vp
        self.lib.vk_icdGetInstanceProcAddr.argtypes = [Handle, cp]
        self.lib.vk_icdNegotiateLoaderICDInterfaceVersion.restype = i32
        self.lib.vk_icdNegotiateLoaderICDInterfaceVersion.argtypes = [P(u32)]
        self.lib.vk_icdGetPhysicalDeviceProcAddr.restype = vp
        self.lib.vk_icdGetPhysicalDeviceProcAddr.argtypes = [Handle, cp]
        for name, res, args in PROTOS:
            addr = self.lib.vk_icdGetInstanceProcAddr(None, name.encode())
            if not addr:
                raise RuntimeError("ICD does not resolve " + name)
            fn = C.CFUNCTYPE(res, *args)(addr)
            setattr(self, name, fn)
        for name, res, args in EXT_PROTOS:
            f = getattr(self.lib, name)
            f.restype = res
            f.argtypes = args
            setattr(self, name, f)

    def check(self, r, what):
        if r != VK_SUCCESS:
            err = self.vh_last_error()
            raise VkError("%s returned %d %s" % (what, r, err.decode() if err else ""))


_icd = None


def load():
    global _icd
    if _icd is None:
        _icd = ICD()
    return _icd
