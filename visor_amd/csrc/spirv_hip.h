// spirv_hip.h - SPIR-V to HIP C translation for the pipeline modules.
//
// Replaces the reference's SPIR-V -> LLVM IR -> x86 JIT (reference spirv_compile.cpp:645-2432,
// CompileFunction; :2434-2448, GetFuncPointer). Each entry point becomes one inlinable HIP C
// device function that keeps the reference's lowering rules (operation order of the matrix
// helpers :423-550, CreateDot :629-643, the GLSL.std.450 formulas :1527-1740, the vertex input
// fetch :572-627 and output padding :2059-2074, fragment input interpolation :2159-2279).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace visor
{
enum class ShaderStage
{
  Vertex = 0,
  Fragment = 4,
};

struct FsInput
{
  uint32_t location;
  uint32_t ncomp;  // flattened float components (1..4 per location slot)
  bool flat_int;   // integer input: vertex 0's bits, no interpolation
};

struct ShaderResource
{
  uint32_t set, binding;
  bool image;  // combined image sampler (else uniform buffer)
};

struct TranslatedShader
{
  ShaderStage stage = ShaderStage::Vertex;
  std::string fn_name;
  std::string code;
  std::vector<FsInput> fs_inputs;         // fragment inputs in varying order
  std::vector<uint32_t> vs_locations;     // vertex input locations the vertex program reads
  std::vector<ShaderResource> resources;  // resources referenced
  bool needs_deriv = false;               // implicit-LOD sampling or derivative ops
  bool has_kill = false;
};

// Parses and translates one entry point. On failure returns false and fills err.
bool TranslateSpirv(const uint32_t *words, size_t nwords, const char *entry, ShaderStage stage,
                    const std::string &fn_name, TranslatedShader &out, std::string &err);

// Quick validity check used at vkCreateShaderModule (magic, version, bound).
bool SpirvHeaderValid(const uint32_t *words, size_t nwords);
}  // namespace visor
