"""A small Vulkan application over the ICD: renders a `scenes.Scene` through the Vulkan API.

This is the host side a user of the reference would write (instance, device, memory, buffers,
images, render pass, pipelines, descriptor sets, command buffers, vkQueueSubmit); every call goes
through the ICD's name tables, and the draws execute on the GPU.
"""

import ctypes as C
import time

import numpy as np

from . import shaders, vk
from .vk import Handle


class Context:
    """Instance, device, queue and command pool."""

    def __init__(self):
        self.icd = vk.load()
        i = self.icd
        self.instance = Handle()
        i.check(i.vkCreateInstance(C.byref(vk.InstanceCreateInfo()), None, C.byref(self.instance)), "vkCreateInstance")
        n = vk.u32(1)
        phys = (Handle * 1)()
        i.check(i.vkEnumeratePhysicalDevices(self.instance, C.byref(n), phys), "vkEnumeratePhysicalDevices")
        self.phys = phys[0]
        prio = (vk.f32 * 1)(1.0)
        qci = vk.DeviceQueueCreateInfo(queueFamilyIndex=0, queueCount=1, pQueuePriorities=prio)
        dci = vk.DeviceCreateInfo(queueCreateInfoCount=1, pQueueCreateInfos=C.pointer(qci))
        self.device = Handle()
        i.check(i.vkCreateDevice(self.phys, C.byref(dci), None, C.byref(self.device)), "vkCreateDevice")
        self.queue = Handle()
        i.vkGetDeviceQueue(self.device, 0, 0, C.byref(self.queue))
        self.pool = Handle()
        i.check(i.vkCreateCommandPool(self.device, C.byref(vk.CommandPoolCreateInfo()), None, C.byref(self.pool)),
                "vkCreateCommandPool")
        self._keep = []

    # -- memory ---------------------------------------------------------------------------------
    def alloc(self, size, host_visible):
        mem = Handle()
        mai = vk.MemoryAllocateInfo(allocationSize=max(int(size), 4),
                                    memoryTypeIndex=vk.MEM_HOST_VISIBLE if host_visible else vk.MEM_DEVICE_LOCAL)
        self.icd.check(self.icd.vkAllocateMemory(self.device, C.byref(mai), None, C.byref(mem)), "vkAllocateMemory")
        return mem

    def map(self, mem, size):
        p = vk.vp()
        self.icd.check(self.icd.vkMapMemory(self.device, mem, 0, size, 0, C.byref(p)), "vkMapMemory")
        return p.value

    def buffer(self, size, host_visible=True):
        b = Handle()
        self.icd.check(self.icd.vkCreateBuffer(self.device, C.byref(vk.BufferCreateInfo(size=max(int(size), 4))), None,
                                               C.byref(b)), "vkCreateBuffer")
        req = vk.MemoryRequirements()
        self.icd.vkGetBufferMemoryRequirements(self.device, b, C.byref(req))
        mem = self.alloc(req.size, host_visible)
        self.icd.check(self.icd.vkBindBufferMemory(self.device, b, mem, 0), "vkBindBufferMemory")
        return b, mem

    def upload_buffer(self, data, device_local=True):
        """A buffer holding data; device-local ones are filled through a staging copy."""
        data = bytes(data)
        stage, smem = self.buffer(len(data), host_visible=True)
        C.memmove(self.map(smem, len(data)), data, len(data))
        if not device_local:
            return stage, smem
        dst, dmem = self.buffer(len(data), host_visible=False)
        cb = self.begin()
        region = vk.BufferCopy(srcOffset=0, dstOffset=0, size=len(data))
        self.icd.vkCmdCopyBuffer(cb, stage, dst, 1, C.byref(region))
        self.end_submit(cb)
        self.free_buffer(stage, smem)
        return dst, dmem

    def free_buffer(self, b, mem):
        self.icd.vkDestroyBuffer(self.device, b, None)
        self.icd.vkFreeMemory(self.device, mem, None)

    def image(self, w, h, fmt, mips=1, layers=1, host_visible=False, cube=False):
        img = Handle()
        ici = vk.ImageCreateInfo(flags=vk.IMAGE_CREATE_CUBE_COMPATIBLE if cube else 0, imageType=1, format=fmt,
                                 extent=vk.Extent3D(width=w, height=h, depth=1), mipLevels=mips,
                                 arrayLayers=layers, samples=1, tiling=1)
        self.icd.check(self.icd.vkCreateImage(self.device, C.byref(ici), None, C.byref(img)), "vkCreateImage")
        req = vk.MemoryRequirements()
        self.icd.vkGetImageMemoryRequirements(self.device, img, C.byref(req))
        mem = self.alloc(req.size, host_visible)
        self.icd.check(self.icd.vkBindImageMemory(self.device, img, mem, 0), "vkBindImageMemory")
        return img, mem, self.view(img, fmt, mips, layers, cube), req.size

    def view(self, img, fmt, mips=1, layers=1, cube=False):
        view = Handle()
        view_type = vk.VIEW_TYPE_CUBE if cube else (vk.VIEW_TYPE_2D_ARRAY if layers > 1 else vk.VIEW_TYPE_2D)
        vci = vk.ImageViewCreateInfo(image=img, viewType=view_type, format=fmt,
                                     subresourceRange=vk.ImageSubresourceRange(aspectMask=1, levelCount=mips,
                                                                               layerCount=layers))
        self.icd.check(self.icd.vkCreateImageView(self.device, C.byref(vci), None, C.byref(view)), "vkCreateImageView")
        return view

    def texture(self, t):
        """Uploads a scenes.Texture (all levels and layers) with buffer-to-image copies."""
        img, mem, view, size = self.image(t.width, t.height, t.format, t.mips, t.layers, cube=t.cube)
        self.upload_texture(img, t)
        samp = Handle()
        sci = vk.SamplerCreateInfo(magFilter=t.mag_filter, minFilter=t.min_filter, mipmapMode=t.mipmap_mode,
                                   addressModeU=t.address, addressModeV=t.address, addressModeW=t.address,
                                   mipLodBias=t.lod_bias, minLod=t.min_lod, maxLod=t.max_lod)
        self.icd.check(self.icd.vkCreateSampler(self.device, C.byref(sci), None, C.byref(samp)), "vkCreateSampler")
        return img, mem, view, samp

    def upload_texture(self, img, t):
        """Copies t.data (every level and layer) into an existing image of t's shape."""
        bpp = vk.texel_bytes(t.format)
        stage, smem = self.buffer(len(t.data), host_visible=True)
        C.memmove(self.map(smem, len(t.data)), t.data, len(t.data))
        cb = self.begin()
        offs = 0
        regions = []
        for layer in range(t.layers):
            for m in range(t.mips):
                w, h = max(1, t.width >> m), max(1, t.height >> m)
                regions.append(vk.BufferImageCopy(
                    bufferOffset=offs, imageSubresource=vk.ImageSubresourceLayers(aspectMask=1, mipLevel=m,
                                                                                  baseArrayLayer=layer, layerCount=1),
                    imageExtent=vk.Extent3D(width=w, height=h, depth=1)))
                offs += w * h * bpp
        arr = (vk.BufferImageCopy * len(regions))(*regions)
        self.icd.vkCmdCopyBufferToImage(cb, stage, img, 7, len(regions), arr)
        self.end_submit(cb)
        self.free_buffer(stage, smem)

    # -- commands -------------------------------------------------------------------------------
    def begin(self):
        cb = Handle()
        cai = vk.CommandBufferAllocateInfo(commandPool=self.pool, level=0, commandBufferCount=1)
        self.icd.check(self.icd.vkAllocateCommandBuffers(self.device, C.byref(cai), C.byref(cb)), "vkAllocateCommandBuffers")
        self.icd.check(self.icd.vkBeginCommandBuffer(cb, C.byref(vk.CommandBufferBeginInfo())), "vkBeginCommandBuffer")
        return cb

    def submit(self, cbs, fence=None):
        arr = (Handle * len(cbs))(*cbs)
        si = vk.SubmitInfo(commandBufferCount=len(cbs), pCommandBuffers=arr)
        return self.icd.vkQueueSubmit(self.queue, 1, C.byref(si), fence)

    def end_submit(self, cb, free=True):
        self.icd.check(self.icd.vkEndCommandBuffer(cb), "vkEndCommandBuffer")
        self.icd.check(self.submit([cb]), "vkQueueSubmit")
        self.icd.check(self.icd.vkQueueWaitIdle(self.queue), "vkQueueWaitIdle")
        if free:
            arr = (Handle * 1)(cb)
            self.icd.vkFreeCommandBuffers(self.device, self.pool, 1, arr)

    def shader(self, spv):
        m = Handle()
        buf = C.create_string_buffer(spv, len(spv))
        self._keep.append(buf)
        sci = vk.ShaderModuleCreateInfo(codeSize=len(spv), pCode=C.cast(buf, vk.vp))
        self.icd.check(self.icd.vkCreateShaderModule(self.device, C.byref(sci), None, C.byref(m)), "vkCreateShaderModule")
        return m

    def set_mode(self, batched=True, tile=None, async_submit=None, graphs=None):
        self.icd.vh_set_exec_mode(self.queue, 1 if batched else 0)
        if graphs is not None:
            self.icd.vh_set_graph_mode(self.queue, 1 if graphs else 0)
        if tile is not None:
            self.icd.vh_set_tile_size(self.queue, tile)
        if async_submit is not None:
            self.icd.vh_set_async_submit(self.queue, 1 if async_submit else 0)

    def stats(self, which=0):
        s = vk.VhStats()
        self.icd.vh_get_stats(self.queue, which, C.byref(s))
        return {k: getattr(s, k) for k, _ in vk.VhStats._fields_}


class SceneRenderer:
    """Builds the Vulkan objects of a Scene once and records/replays its render pass."""

    def __init__(self, ctx, scene, host_visible_targets=True, target=None):
        """target: (image, view) of an existing colour image of the scene's size and format (a
        swapchain image, say) to render into instead of an image of its own."""
        self.ctx, self.scene = ctx, scene
        i, dev = ctx.icd, ctx.device
        s = scene
        self.vb, self.vbm = ctx.upload_buffer(s.vertex_data)
        self.ib = None
        if s.index_data is not None:
            self.ib, self.ibm = ctx.upload_buffer(s.index_data)
        self.textures = [ctx.texture(t) for t in s.textures]
        # render targets
        if target is not None:
            self.color, self.cview = target
            self.cmem = None
        else:
            self.color, self.cmem, self.cview, self.csize = ctx.image(s.width, s.height, s.color_format,
                                                                     host_visible=host_visible_targets)
        self.has_depth = s.clear_depth is not None
        views = [self.cview]
        if self.has_depth:
            self.depth, self.dmem, self.dview, self.dsize = ctx.image(s.width, s.height, vk.FMT["D32_SFLOAT"],
                                                                     host_visible=host_visible_targets)
            views.append(self.dview)
        def render_pass(load_op):
            atts = [vk.AttachmentDescription(format=s.color_format, samples=1, loadOp=load_op)]
            if self.has_depth:
                atts.append(vk.AttachmentDescription(format=vk.FMT["D32_SFLOAT"], samples=1, loadOp=load_op))
            att_arr = (vk.AttachmentDescription * len(atts))(*atts)
            cref = vk.AttachmentReference(attachment=0, layout=2)
            dref = vk.AttachmentReference(attachment=1, layout=3)
            sub = vk.SubpassDescription(colorAttachmentCount=1, pColorAttachments=C.pointer(cref),
                                        pDepthStencilAttachment=C.pointer(dref) if self.has_depth else None)
            rpci = vk.RenderPassCreateInfo(attachmentCount=len(atts), pAttachments=att_arr, subpassCount=1,
                                           pSubpasses=C.pointer(sub))
            rp = Handle()
            i.check(i.vkCreateRenderPass(dev, C.byref(rpci), None, C.byref(rp)), "vkCreateRenderPass")
            return rp

        self.rp = render_pass(vk.LOAD_OP_CLEAR)
        self.rp_load = render_pass(vk.LOAD_OP_LOAD)  # compatible: same formats, keeps the contents
        varr = (Handle * len(views))(*views)
        fci = vk.FramebufferCreateInfo(renderPass=self.rp, attachmentCount=len(views), pAttachments=varr,
                                       width=s.width, height=s.height, layers=1)
        self.fb = Handle()
        i.check(i.vkCreateFramebuffer(dev, C.byref(fci), None, C.byref(self.fb)), "vkCreateFramebuffer")
        # descriptors: set 0 binding 0 = combined image sampler, binding 1 = uniform buffer
        binding = (vk.DescriptorSetLayoutBinding * 2)(
            vk.DescriptorSetLayoutBinding(binding=0, descriptorType=vk.DESC_COMBINED_IMAGE_SAMPLER, descriptorCount=1,
                                          stageFlags=vk.STAGE_FRAGMENT),
            vk.DescriptorSetLayoutBinding(binding=shaders.UBO_BINDING, descriptorType=vk.DESC_UNIFORM_BUFFER,
                                          descriptorCount=1, stageFlags=vk.STAGE_VERTEX | vk.STAGE_FRAGMENT))
        self.dsl = Handle()
        i.check(i.vkCreateDescriptorSetLayout(dev, C.byref(vk.DescriptorSetLayoutCreateInfo(bindingCount=2, pBindings=binding)),
                                              None, C.byref(self.dsl)), "vkCreateDescriptorSetLayout")
        pcr = vk.PushConstantRange(stageFlags=vk.STAGE_VERTEX | vk.STAGE_FRAGMENT, offset=0, size=128)
        dsl_arr = (Handle * 1)(self.dsl)
        self.layout = Handle()
        i.check(i.vkCreatePipelineLayout(dev, C.byref(vk.PipelineLayoutCreateInfo(setLayoutCount=1, pSetLayouts=dsl_arr,
                                                                                  pushConstantRangeCount=1,
                                                                                  pPushConstantRanges=C.pointer(pcr))),
                                         None, C.byref(self.layout)), "vkCreatePipelineLayout")
        # uniform buffers, and one set per (texture, uniform buffer) pair that a draw binds with a
        # uniform buffer; self.sets[t] is the set of texture t alone
        self.uniform_bufs = [ctx.upload_buffer(u) for u in s.uniforms]
        ucombos = sorted({(d.texture, d.uniform) for d in s.draws if d.uniform is not None},
                         key=lambda k: (-1 if k[0] is None else k[0], k[1]))
        self.dpool = Handle()
        i.check(i.vkCreateDescriptorPool(dev, C.byref(vk.DescriptorPoolCreateInfo(
            maxSets=max(1, len(self.textures) + len(ucombos)))), None, C.byref(self.dpool)), "vkCreateDescriptorPool")

        def new_set(tex, ubo):
            ds = Handle()
            dai = vk.DescriptorSetAllocateInfo(descriptorPool=self.dpool, descriptorSetCount=1, pSetLayouts=dsl_arr)
            i.check(i.vkAllocateDescriptorSets(dev, C.byref(dai), C.byref(ds)), "vkAllocateDescriptorSets")
            if tex is not None:
                img, mem, view, samp = self.textures[tex]
                dii = vk.DescriptorImageInfo(sampler=samp, imageView=view, imageLayout=5)
                w = vk.WriteDescriptorSet(dstSet=ds, dstBinding=0, descriptorCount=1,
                                          descriptorType=vk.DESC_COMBINED_IMAGE_SAMPLER, pImageInfo=C.pointer(dii))
                i.vkUpdateDescriptorSets(dev, 1, C.byref(w), 0, None)
            if ubo is not None:
                dbi = vk.DescriptorBufferInfo(buffer=self.uniform_bufs[ubo][0], offset=0, range=len(s.uniforms[ubo]))
                w = vk.WriteDescriptorSet(dstSet=ds, dstBinding=shaders.UBO_BINDING, descriptorCount=1,
                                          descriptorType=vk.DESC_UNIFORM_BUFFER, pBufferInfo=C.pointer(dbi))
                i.vkUpdateDescriptorSets(dev, 1, C.byref(w), 0, None)
            return ds

        self.sets = [new_set(t, None) for t in range(len(self.textures))]
        self.usets = {k: new_set(*k) for k in ucombos}
        # pipelines
        self.modules = {}
        self.pipelines = {}
        for name, p in s.pipelines.items():
            self.pipelines[name] = self._pipeline(p)
        self.cb = None

    def _module(self, prog):
        if prog not in self.modules:
            self.modules[prog] = self.ctx.shader(shaders.PROGRAMS[prog]())
        return self.modules[prog]

    def _pipeline(self, p):
        i, s = self.ctx.icd, self.scene
        stages = (vk.PipelineShaderStageCreateInfo * 2)(
            vk.PipelineShaderStageCreateInfo(stage=vk.STAGE_VERTEX, module=self._module(p.vs), pName=b"main"),
            vk.PipelineShaderStageCreateInfo(stage=vk.STAGE_FRAGMENT, module=self._module(p.fs), pName=b"main"))
        bind = vk.VertexInputBindingDescription(binding=0, stride=s.vertex_stride, inputRate=0)
        attrs = (vk.VertexInputAttributeDescription * len(s.vattrs))(
            *[vk.VertexInputAttributeDescription(location=l, binding=0, format=f, offset=o) for (l, f, o) in s.vattrs])
        vis = vk.PipelineVertexInputStateCreateInfo(vertexBindingDescriptionCount=1, pVertexBindingDescriptions=C.pointer(bind),
                                                    vertexAttributeDescriptionCount=len(s.vattrs), pVertexAttributeDescriptions=attrs)
        ias = vk.PipelineInputAssemblyStateCreateInfo(topology=p.topology)
        vp_ = vk.Viewport(x=0, y=0, width=s.width, height=s.height, minDepth=0, maxDepth=1)
        sc = vk.Rect2D(extent=vk.Extent2D(width=s.width, height=s.height))
        vps = vk.PipelineViewportStateCreateInfo(viewportCount=1, pViewports=C.pointer(vp_), scissorCount=1, pScissors=C.pointer(sc))
        rs = vk.PipelineRasterizationStateCreateInfo(cullMode=p.cull, frontFace=p.front, lineWidth=1.0)
        ms = vk.PipelineMultisampleStateCreateInfo(rasterizationSamples=1)
        ds = vk.PipelineDepthStencilStateCreateInfo(depthTestEnable=1, depthWriteEnable=1 if p.depth_write else 0,
                                                    depthCompareOp=p.depth_op)
        att = vk.PipelineColorBlendAttachmentState(blendEnable=1 if p.blend else 0, srcColorBlendFactor=p.src,
                                                   dstColorBlendFactor=p.dst, colorBlendOp=0, srcAlphaBlendFactor=p.src,
                                                   dstAlphaBlendFactor=p.dst, colorWriteMask=0xF)
        cbs = vk.PipelineColorBlendStateCreateInfo(attachmentCount=1, pAttachments=C.pointer(att))
        gci = vk.GraphicsPipelineCreateInfo(stageCount=2, pStages=stages, pVertexInputState=C.pointer(vis),
                                            pInputAssemblyState=C.pointer(ias), pViewportState=C.pointer(vps),
                                            pRasterizationState=C.pointer(rs), pMultisampleState=C.pointer(ms),
                                            pDepthStencilState=C.pointer(ds), pColorBlendState=C.pointer(cbs),
                                            layout=self.layout, renderPass=self.rp)
        pl = Handle()
        i.check(i.vkCreateGraphicsPipelines(self.ctx.device, None, 1, C.byref(gci), None, C.byref(pl)), "vkCreateGraphicsPipelines")
        return pl

    def record(self, draws=None, passes=None):
        """Records one render pass with the scene's draws (or a subset) and returns the buffer.
        passes: a list of (draws, clear) recorded as consecutive render passes instead, the ones
        with clear False loading the previous contents."""
        ctx, i = self.ctx, self.ctx.icd
        cb = ctx.begin()
        for pd, clear in (passes if passes is not None else [(self.scene.draws if draws is None else draws, True)]):
            self._record_pass(cb, pd, clear)
        i.check(i.vkEndCommandBuffer(cb), "vkEndCommandBuffer")
        return cb

    def _record_pass(self, cb, draws, clear):
        i, s = self.ctx.icd, self.scene
        clears = (vk.ClearValue * 2)()
        for k in range(4):
            clears[0].color[k] = s.clear_color[k]
        clears[1].depthStencil[0] = s.clear_depth if s.clear_depth is not None else 1.0
        rbi = vk.RenderPassBeginInfo(renderPass=self.rp if clear else self.rp_load, framebuffer=self.fb,
                                     renderArea=vk.Rect2D(extent=vk.Extent2D(width=s.width, height=s.height)),
                                     clearValueCount=2 if self.has_depth else 1, pClearValues=clears)
        i.vkCmdBeginRenderPass(cb, C.byref(rbi), 0)
        vbs = (Handle * 1)(self.vb)
        offs = (vk.u64 * 1)(0)
        i.vkCmdBindVertexBuffers(cb, 0, 1, vbs, offs)
        if self.ib is not None:
            i.vkCmdBindIndexBuffer(cb, self.ib, 0, s.index_type)
        cur_pipe, cur_set = None, None
        for d in draws:
            if d.pipeline != cur_pipe:
                i.vkCmdBindPipeline(cb, 0, self.pipelines[d.pipeline])
                cur_pipe = d.pipeline
            key = (d.texture, d.uniform)
            if key != (None, None) and key != cur_set:
                sets = (Handle * 1)(self.usets[key] if d.uniform is not None else self.sets[d.texture])
                i.vkCmdBindDescriptorSets(cb, 0, self.layout, 0, 1, sets, 0, None)
                cur_set = key
            pc = C.create_string_buffer(d.push, 128)
            i.vkCmdPushConstants(cb, self.layout, vk.STAGE_VERTEX | vk.STAGE_FRAGMENT, 0, 128, pc)
            if d.indexed:
                i.vkCmdDrawIndexed(cb, d.count, 1, d.first, d.vertex_offset, 0)
            else:
                i.vkCmdDraw(cb, d.count, 1, d.first, 0)
        i.vkCmdEndRenderPass(cb)

    def render(self, draws=None):
        """Records, submits and waits; returns the wall time of the submit in seconds."""
        if self.cb is None or draws is not None:
            self.cb = self.record(draws)
        t0 = time.perf_counter()
        self.ctx.icd.check(self.ctx.submit([self.cb]), "vkQueueSubmit")
        self.ctx.icd.check(self.ctx.icd.vkQueueWaitIdle(self.ctx.queue), "vkQueueWaitIdle")
        return time.perf_counter() - t0

    def read_color(self):
        s = self.scene
        n = s.width * s.height * s.color_bpp
        p = self.ctx.map(self.cmem, n)
        a = np.frombuffer((C.c_uint8 * n).from_address(p), dtype=np.uint8).copy()
        if s.color_bpp == 8:
            return a.view(np.uint16).reshape(s.height, s.width, 4)
        return a.reshape(s.height, s.width, 4)

    def read_depth(self):
        s = self.scene
        n = s.width * s.height * 4
        p = self.ctx.map(self.dmem, n)
        return np.frombuffer((C.c_uint8 * n).from_address(p), dtype=np.uint8).copy().view(np.float32).reshape(s.height, s.width)
