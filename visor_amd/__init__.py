"""visor_amd: an MI355X-native (gfx950) implementation of visor's draw-execution path.

The product is the Vulkan ICD library libvisor_mi355x.so (built by visor_amd.build). This package
holds its sources (csrc/), the Python Vulkan bindings used to drive it (vk, render), the SPIR-V
builder and shader programs (spirv, shaders) and the seeded workloads (scenes).
"""
