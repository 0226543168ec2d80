"""Host-side API behaviour that needs no GPU: pipeline state the draw path cannot honour is refused
at vkCreateGraphicsPipelines with VK_ERROR_FEATURE_NOT_PRESENT instead of being replaced silently,
and vkWaitForFences honours wait-any and wait-all (the reference's fences are no-ops,
icd_stubs.cpp:64-155; its vertex state has four buffer slots, gpu.h:3-23; its blend runs only
the ADD op, rasterizer.cpp:657-669)."""
import ctypes as C

import pytest

from visor_amd import shaders, vk

VK_SUCCESS, VK_TIMEOUT, VK_ERROR_FEATURE_NOT_PRESENT = 0, 2, -8


@pytest.fixture(scope="module")
def icd():
    return vk.load()


def _module(icd, prog):
    spv = shaders.PROGRAMS[prog]()
    buf = C.create_string_buffer(spv, len(spv))
    m = vk.Handle()
    sci = vk.ShaderModuleCreateInfo(codeSize=len(spv), pCode=C.cast(buf, vk.vp))
    assert icd.vkCreateShaderModule(None, C.byref(sci), None, C.byref(m)) == VK_SUCCESS
    return m, buf


def _create_pipeline(icd, blend_enable=0, blend_op=0, attr_binding=0, bind_binding=0):
    """vkCreateGraphicsPipelines of the colour-passthrough pair with one vertex binding and two
    attributes; returns the VkResult."""
    vs, vbuf = _module(icd, "vs_mvp_pos4_col")
    fs, fbuf = _module(icd, "fs_color")
    stages = (vk.PipelineShaderStageCreateInfo * 2)(
        vk.PipelineShaderStageCreateInfo(stage=vk.STAGE_VERTEX, module=vs, pName=b"main"),
        vk.PipelineShaderStageCreateInfo(stage=vk.STAGE_FRAGMENT, module=fs, pName=b"main"))
    bind = vk.VertexInputBindingDescription(binding=bind_binding, stride=20, inputRate=0)
    attrs = (vk.VertexInputAttributeDescription * 2)(
        vk.VertexInputAttributeDescription(location=0, binding=attr_binding, format=109, offset=0),
        vk.VertexInputAttributeDescription(location=1, binding=attr_binding, format=37, offset=16))
    vis = vk.PipelineVertexInputStateCreateInfo(vertexBindingDescriptionCount=1, pVertexBindingDescriptions=C.pointer(bind),
                                                vertexAttributeDescriptionCount=2, pVertexAttributeDescriptions=attrs)
    ias = vk.PipelineInputAssemblyStateCreateInfo(topology=3)
    rs = vk.PipelineRasterizationStateCreateInfo(cullMode=0, frontFace=0, lineWidth=1.0)
    ds = vk.PipelineDepthStencilStateCreateInfo(depthTestEnable=1, depthWriteEnable=1, depthCompareOp=1)
    att = vk.PipelineColorBlendAttachmentState(blendEnable=blend_enable, srcColorBlendFactor=6, dstColorBlendFactor=7,
                                               colorBlendOp=blend_op, srcAlphaBlendFactor=6, dstAlphaBlendFactor=7,
                                               colorWriteMask=0xF)
    cbs = vk.PipelineColorBlendStateCreateInfo(attachmentCount=1, pAttachments=C.pointer(att))
    gci = vk.GraphicsPipelineCreateInfo(stageCount=2, pStages=stages, pVertexInputState=C.pointer(vis),
                                        pInputAssemblyState=C.pointer(ias), pRasterizationState=C.pointer(rs),
                                        pDepthStencilState=C.pointer(ds), pColorBlendState=C.pointer(cbs))
    pl = vk.Handle()
    r = icd.vkCreateGraphicsPipelines(None, None, 1, C.byref(gci), None, C.byref(pl))
    if r == VK_SUCCESS:
        icd.vkDestroyPipeline(None, pl, None)
    icd.vkDestroyShaderModule(None, vs, None)
    icd.vkDestroyShaderModule(None, fs, None)
    return r


def test_supported_pipeline_is_created(icd):
    assert _create_pipeline(icd) == VK_SUCCESS
    assert _create_pipeline(icd, blend_enable=1, blend_op=0) == VK_SUCCESS


@pytest.mark.parametrize("op", [1, 2, 3, 4])  # SUBTRACT, REVERSE_SUBTRACT, MIN, MAX
def test_blend_ops_other_than_add_are_refused(icd, op):
    assert _create_pipeline(icd, blend_enable=1, blend_op=op) == VK_ERROR_FEATURE_NOT_PRESENT
    assert b"blend op" in icd.vh_last_error()
    # with blending off the op is never used
    assert _create_pipeline(icd, blend_enable=0, blend_op=op) == VK_SUCCESS


def test_vertex_bindings_beyond_slot_3_are_refused(icd):
    assert _create_pipeline(icd, attr_binding=3, bind_binding=3) == VK_SUCCESS
    assert _create_pipeline(icd, attr_binding=4, bind_binding=0) == VK_ERROR_FEATURE_NOT_PRESENT
    assert _create_pipeline(icd, attr_binding=0, bind_binding=5) == VK_ERROR_FEATURE_NOT_PRESENT
    assert b"binding" in icd.vh_last_error()


def _fence(icd, signalled):
    f = vk.Handle()
    info = vk.FenceCreateInfo(flags=1 if signalled else 0)
    assert icd.vkCreateFence(None, C.byref(info), None, C.byref(f)) == VK_SUCCESS
    return f


def test_wait_any_returns_when_a_later_fence_is_signalled(icd):
    unsignalled, signalled = _fence(icd, False), _fence(icd, True)
    fences = (vk.Handle * 2)(unsignalled, signalled)
    assert icd.vkWaitForFences(None, 2, fences, 0, 0) == VK_SUCCESS          # wait any
    assert icd.vkWaitForFences(None, 2, fences, 1, 0) == VK_TIMEOUT          # wait all
    assert icd.vkWaitForFences(None, 2, fences, 0, 2_000_000) == VK_SUCCESS
    assert icd.vkWaitForFences(None, 1, fences, 0, 1_000_000) == VK_TIMEOUT  # only the unsignalled one
    assert icd.vkResetFences(None, 1, (vk.Handle * 1)(signalled)) == VK_SUCCESS
    assert icd.vkWaitForFences(None, 2, fences, 0, 0) == VK_TIMEOUT
    for f in (unsignalled, signalled):
        icd.vkDestroyFence(None, f, None)
