// device.cpp - HIP device bring-up, shader pairs and hiprtc-built pipeline modules.
//
// vkCreateGraphicsPipelines translates the stage modules to HIP C (reference shaders.cpp:27-91
// + spirv_compile.cpp:645-2448 become TranslateSpirv) and builds one code object per set of
// shader pairs with hiprtc for gfx950. Code objects are cached on disk, keyed by a hash of the
// full source and options, so a pipeline is compiled once per machine.
#include <dlfcn.h>
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>
#include <sys/stat.h>
#include <unistd.h>

#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <fstream>
#include <sstream>

#include "internal.h"

namespace visor
{
extern const char *kSrcDeviceAbi;
extern const char *kSrcPrelude;
extern const char *kSrcKernels;

double NowSeconds()
{
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

void LogError(const char *fmt, ...)
{
  if(getenv("VISOR_QUIET"))
    return;
  va_list ap;
  va_start(ap, fmt);
  fprintf(stderr, "[visor] ");
  vfprintf(stderr, fmt, ap);
  fprintf(stderr, "\n");
  va_end(ap);
}

const char *HipErrorName(hipError_t e) { return hipGetErrorString(e); }

VkResult HipToVk(hipError_t e)
{
  switch(e)
  {
    case hipSuccess: return VK_SUCCESS;
    case hipErrorOutOfMemory: return VK_ERROR_OUT_OF_DEVICE_MEMORY;
    case hipErrorNoDevice: case hipErrorInvalidDevice: return VK_ERROR_INITIALIZATION_FAILED;
    default: return VK_ERROR_DEVICE_LOST;
  }
}

static uint64_t Fnv1a(const std::string &s, uint64_t h = 1469598103934665603ull)
{
  for(unsigned char c : s)
  {
    h ^= c;
    h *= 1099511628211ull;
  }
  return h;
}

static std::string HexKey(uint64_t h)
{
  char b[32];
  snprintf(b, sizeof(b), "%016llx", (unsigned long long)h);
  return b;
}

Device &Device::Get()
{
  static Device d;
  return d;
}

static std::string LibraryDir()
{
  Dl_info info;
  if(dladdr((void *)&LibraryDir, &info) && info.dli_fname)
  {
    std::string p = info.dli_fname;
    size_t s = p.rfind('/');
    if(s != std::string::npos)
      return p.substr(0, s);
  }
  return ".";
}

static std::string DefaultCacheDir()
{
  const char *env = getenv("VISOR_CACHE_DIR");
  if(env && *env)
    return env;
  return LibraryDir() + "/_cocache";
}

bool Device::Init(std::string &err)
{
  std::lock_guard<std::mutex> lk(m);
  if(initialized)
    return true;
  if(cache_dir.empty())
    cache_dir = DefaultCacheDir();
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if(e != hipSuccess || n == 0)
  {
    err = "no HIP device available";
    return false;
  }
  e = hipSetDevice(0);
  if(e == hipSuccess)
    e = hipGetDeviceProperties(&props, 0);
  if(e != hipSuccess)
  {
    err = std::string("HIP device query failed: ") + hipGetErrorString(e);
    return false;
  }
  std::string name = props.gcnArchName;
  arch = name.substr(0, name.find(':'));
  int khz = 0;
  if(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, 0) == hipSuccess && khz > 0)
    wall_khz = khz;
  initialized = true;
  return true;
}

std::shared_ptr<ShaderPair> Device::MakePair(VkShaderModule_T *vsm, const char *vs_entry, VkShaderModule_T *fsm,
                                             const char *fs_entry, std::string &err)
{
  std::string raw;
  raw.append((const char *)vsm->code.data(), vsm->code.size() * 4);
  raw += "|";
  raw += vs_entry ? vs_entry : "main";
  raw += "|";
  raw.append((const char *)fsm->code.data(), fsm->code.size() * 4);
  raw += "|";
  raw += fs_entry ? fs_entry : "main";
  std::string key = HexKey(Fnv1a(raw));
  {
    std::lock_guard<std::mutex> lk(m);
    auto it = pairs.find(key);
    if(it != pairs.end())
      return it->second;
  }
  auto p = std::make_shared<ShaderPair>();
  p->key = key;
  if(!TranslateSpirv(vsm->code.data(), vsm->code.size(), vs_entry, ShaderStage::Vertex, "va_vs_" + key, p->vs, err))
  {
    err = "vertex program: " + err;
    return nullptr;
  }
  if(!TranslateSpirv(fsm->code.data(), fsm->code.size(), fs_entry, ShaderStage::Fragment, "va_fs_" + key, p->fs, err))
  {
    err = "fragment program: " + err;
    return nullptr;
  }
  // link: each fragment input component reads VS output slot loc*4+comp (reference gpu.h:53-57)
  for(const FsInput &in : p->fs.fs_inputs)
  {
    if(in.location >= 10)
    {
      err = "fragment input location beyond the 10 interpolants the vertex output holds";
      return nullptr;
    }
    for(uint32_t k = 0; k < in.ncomp; k++)
      p->vary_src.push_back(in.location * 4 + k);
  }
  p->nv = (uint32_t)p->vary_src.size();
  p->resources = p->vs.resources;
  for(auto &r : p->fs.resources)
  {
    bool dup = false;
    for(auto &q : p->resources)
      dup |= q.set == r.set && q.binding == r.binding;
    if(!dup)
      p->resources.push_back(r);
  }
  std::lock_guard<std::mutex> lk(m);
  pairs[key] = p;
  return p;
}

static const char *kRtcOptions[] = {
    "-O3",
    "-std=c++17",
    "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
    "-Wno-unused-label",
    "-Wno-unused-variable",
};

// Assembles the module source for a list of shader pairs (one slot each).
static bool ReadFile(const std::string &path, std::vector<char> &out);

static std::string ModuleSource(const std::vector<std::shared_ptr<ShaderPair>> &pairs, uint32_t &rec_stride,
                                const TileConfig &cfg)
{
  uint32_t maxnv = 0;
  bool deriv = false;
  for(auto &p : pairs)
  {
    maxnv = std::max(maxnv, p->nv);
    deriv |= p->fs.needs_deriv;
  }
  rec_stride = (VA_REC_HDR + 3 * maxnv + 3) & ~3u;
  std::ostringstream s;
  // VISOR_PRELUDE_SOURCE / VISOR_KERNEL_SOURCE: development overrides of the built-in prelude and
  // kernel files (same interface)
  std::vector<char> override_prelude;
  const char *ps = getenv("VISOR_PRELUDE_SOURCE");
  if(ps && ReadFile(ps, override_prelude))
    s << kSrcDeviceAbi << "\n" << std::string(override_prelude.begin(), override_prelude.end()) << "\n";
  else
    s << kSrcDeviceAbi << "\n" << kSrcPrelude << "\n";
  s << "#define VA_NSLOTS " << pairs.size() << "\n";
  s << "#define VA_MAXNV " << maxnv << "\n";
  s << "#define VA_REC_STRIDE " << rec_stride << "u\n";
  s << "#define VA_FS_DERIV " << (deriv ? 1 : 0) << "\n";
  bool kill = false;
  for(auto &p : pairs)
    kill |= p->fs.has_kill;
  s << "#define VA_FS_KILL " << (kill ? 1 : 0) << "\n";
  s << "#define VA_CHUNK " << cfg.chunk << "u\n";
  s << "#define VA_SMALL " << cfg.small << "\n";
  for(auto &p : pairs)
    s << p->vs.code << "\n" << p->fs.code << "\n";
  // the vertex program of a primitive's three vertices: the raw words of every attribute the
  // program reads are loaded for all three vertices first (one basic block, no branch, so the
  // loads are issued together), then the program runs per vertex on them
  s << "VA_INLINE void va_vs3(u32 slot, const VaDraw &d, const u32 vtx[3], float pos[3][4], float vary[3][VA_MAXNV > 0 ? "
       "VA_MAXNV : 1])\n{\n";
  s << "  float outs[3][40];\n  VaRaw raw[3][16];\n  switch(slot)\n  {\n";
  for(size_t i = 0; i < pairs.size(); i++)
  {
    const auto &p = pairs[i];
    s << "    case " << i << "u:\n";
    for(uint32_t loc : p->vs.vs_locations)
      for(int k = 0; k < 3; k++)
        s << "      raw[" << k << "][" << loc << "] = va_vload(d, vtx[" << k << "], " << loc << "u);\n";
    for(int k = 0; k < 3; k++)
      s << "      " << p->vs.fn_name << "(d, vtx[" << k << "], raw[" << k << "], pos[" << k << "], outs[" << k
        << "]);\n";
    for(int k = 0; k < 3; k++)
    {
      for(uint32_t j = 0; j < p->nv; j++)
        s << "      vary[" << k << "][" << j << "] = outs[" << k << "][" << p->vary_src[j] << "];\n";
      for(uint32_t j = p->nv; j < maxnv; j++)
        s << "      vary[" << k << "][" << j << "] = 0.0f;\n";
    }
    s << "      break;\n";
  }
  s << "    default: break;\n  }\n}\n";
  s << "VA_INLINE bool va_fs(u32 slot, const VaDraw &d, const float *vary, const float b[3][4], float out[4])\n{\n";
  s << "  switch(slot)\n  {\n";
  for(size_t i = 0; i < pairs.size(); i++)
    s << "    case " << i << "u: return " << pairs[i]->fs.fn_name << "(d, vary, b, out);\n";
  s << "    default: return false;\n  }\n}\n";
  std::vector<char> override_src;
  const char *ks = getenv("VISOR_KERNEL_SOURCE");
  if(ks && ReadFile(ks, override_src))
    s << std::string(override_src.begin(), override_src.end()) << "\n";
  else
    s << kSrcKernels << "\n";
  return s.str();
}

static bool ReadFile(const std::string &path, std::vector<char> &out)
{
  std::ifstream f(path, std::ios::binary);
  if(!f)
    return false;
  out.assign(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
  return !out.empty();
}

static void WriteFileAtomic(const std::string &dir, const std::string &name, const std::vector<char> &data)
{
  mkdir(dir.c_str(), 0755);
  std::string tmp = dir + "/." + name + "." + std::to_string(getpid());
  {
    std::ofstream f(tmp, std::ios::binary);
    if(!f)
      return;
    f.write(data.data(), (std::streamsize)data.size());
  }
  rename(tmp.c_str(), (dir + "/" + name).c_str());
}

// Compiles (or fetches from the cache) the code object for a module source. Usable without a GPU.
static bool BuildCodeObject(const std::string &src, const std::string &arch, const std::string &cache_dir,
                            std::vector<char> &code, std::string &key, double &seconds, std::string &err)
{
  std::string opts_joined = arch;
  for(const char *o : kRtcOptions)
    opts_joined += std::string(" ") + o;
  key = HexKey(Fnv1a(src, Fnv1a(opts_joined)));
  std::string path = cache_dir + "/" + key + ".hsaco";
  seconds = 0.0;
  if(const char *dump = getenv("VISOR_DUMP_SOURCE"))
  {
    // every module's HIP source, for reading the generated code and its ISA
    std::ofstream f(std::string(dump) + "/" + key + ".hip", std::ios::binary);
    f << src;
  }
  if(ReadFile(path, code))
    return true;
  double t0 = NowSeconds();
  hiprtcProgram prog;
  if(hiprtcCreateProgram(&prog, src.c_str(), "visor_pipeline.hip", 0, nullptr, nullptr) != HIPRTC_SUCCESS)
  {
    err = "hiprtcCreateProgram failed";
    return false;
  }
  std::vector<const char *> opts;
  std::string archopt = "--offload-arch=" + arch;
  opts.push_back(archopt.c_str());
  for(const char *o : kRtcOptions)
    opts.push_back(o);
  hiprtcResult r = hiprtcCompileProgram(prog, (int)opts.size(), opts.data());
  if(r != HIPRTC_SUCCESS)
  {
    size_t ls = 0;
    hiprtcGetProgramLogSize(prog, &ls);
    std::string log(ls + 1, '\0');
    hiprtcGetProgramLog(prog, &log[0]);
    err = "hiprtc compile failed: " + log.substr(0, 4000);
    hiprtcDestroyProgram(&prog);
    if(getenv("VISOR_DUMP_FAILED"))
    {
      std::ofstream f(std::string(getenv("VISOR_DUMP_FAILED")), std::ios::binary);
      f << src;
    }
    return false;
  }
  size_t cs = 0;
  hiprtcGetCodeSize(prog, &cs);
  code.resize(cs);
  hiprtcGetCode(prog, code.data());
  hiprtcDestroyProgram(&prog);
  seconds = NowSeconds() - t0;
  WriteFileAtomic(cache_dir, key + ".hsaco", code);
  return true;
}

std::shared_ptr<PipelineModule> Device::GetModule(const std::vector<std::shared_ptr<ShaderPair>> &ps, std::string &err,
                                                  const TileConfig &cfg)
{
  std::string mkey = "c" + std::to_string(cfg.chunk) + "s" + std::to_string(cfg.small) + ":";
  for(auto &p : ps)
    mkey += p->key + ",";
  {
    std::lock_guard<std::mutex> lk(m);
    auto it = modules.find(mkey);
    if(it != modules.end())
      return it->second;
  }
  auto pm = std::make_shared<PipelineModule>();
  pm->key = mkey;
  std::string src = ModuleSource(ps, pm->rec_stride, cfg);
  std::vector<char> code;
  std::string ckey;
  if(!BuildCodeObject(src, arch, cache_dir, code, ckey, pm->compile_seconds, err))
    return nullptr;
  if(!initialized)
  {
    err = "no HIP device: code object built but not loaded";
    return nullptr;
  }
  hipError_t e = hipModuleLoadData(&pm->module, code.data());
  if(e != hipSuccess)
  {
    err = std::string("hipModuleLoadData: ") + hipGetErrorString(e);
    return nullptr;
  }
  const char *names[] = {"va_setup",    "va_emit",      "va_ranges",    "va_raster_8", "va_raster_16",
                         "va_raster_32", "va_raster_p8", "va_raster_p16", "va_raster_p32"};
  hipFunction_t *fns[] = {&pm->setup,     &pm->emit,        &pm->ranges,      &pm->raster[0], &pm->raster[1],
                          &pm->raster[2], &pm->raster_p[0], &pm->raster_p[1], &pm->raster_p[2]};
  for(int i = 0; i < 9; i++)
  {
    e = hipModuleGetFunction(fns[i], pm->module, names[i]);
    if(e != hipSuccess)
    {
      err = std::string("hipModuleGetFunction(") + names[i] + "): " + hipGetErrorString(e);
      return nullptr;
    }
  }
  // persistent raster grids: resident workgroups per CU x CUs
  for(int i = 0; i < 3; i++)
  {
    int per_cu = 0;
    const int threads = (8 << i) / 8 * ((8 << i) / 8) * 64;
    if(hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, pm->raster_p[i], threads, 0) != hipSuccess ||
       per_cu < 1)
    {
      (void)hipGetLastError();
      per_cu = 1;
    }
    pm->raster_p_grid[i] = (uint32_t)per_cu * (uint32_t)std::max(props.multiProcessorCount, 1);
  }
  for(auto &p : ps)
    pm->slot_keys.push_back(p->key);
  std::lock_guard<std::mutex> lk(m);
  modules[mkey] = pm;
  return pm;
}

// Builds a module's code object without a device (used by the build step and CPU tests).
bool BuildModuleOffline(const std::vector<std::shared_ptr<ShaderPair>> &ps, const std::string &arch,
                        const std::string &cache_dir, std::string &err, size_t *code_size, std::string *src_out)
{
  // the module's LDS chunk is a compile-time parameter: build both sizes the autotuner tries
  for(uint32_t chunk : {64u, 32u})
  {
    TileConfig cfg;
    cfg.chunk = chunk;
    uint32_t stride;
    std::string src = ModuleSource(ps, stride, cfg);
    std::vector<char> code;
    std::string key;
    double secs;
    if(!BuildCodeObject(src, arch, cache_dir.empty() ? DefaultCacheDir() : cache_dir, code, key, secs, err))
      return false;
    if(chunk == TileConfig().chunk)
    {
      if(src_out)
        *src_out = src;
      if(code_size)
        *code_size = code.size();
    }
  }
  return true;
}
}  // namespace visor
