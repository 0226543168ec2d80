// spirv_hip.cpp - SPIR-V to HIP C translation (see spirv_hip.h).
//
// Every SPIR-V value is scalarised: a vec4 result %12 becomes four C locals v0_12_0..3, a mat4
// sixteen (column-major). Interface and function variables become small local arrays ("register
// storage"); push-constant and uniform blocks are read from memory with the Offset /
// MatrixStride / ArrayStride decorations. Control flow is emitted as labels and gotos (the
// module is structured, so this is always well formed), OpPhi through per-edge copies, and
// OpFunctionCall by inlining. Float values use the type F, which is float, or the triple-point
// df type when the fragment program needs derivatives.
#include "spirv_hip.h"

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <set>
#include <sstream>

namespace visor
{
namespace
{
// opcodes
enum Op : uint16_t
{
  OpNop = 0, OpSource = 3, OpSourceExtension = 4, OpName = 5, OpMemberName = 6, OpString = 7,
  OpLine = 8, OpExtension = 10, OpExtInstImport = 11, OpExtInst = 12, OpMemoryModel = 14,
  OpEntryPoint = 15, OpExecutionMode = 16, OpCapability = 17, OpTypeVoid = 19, OpTypeBool = 20,
  OpTypeInt = 21, OpTypeFloat = 22, OpTypeVector = 23, OpTypeMatrix = 24, OpTypeImage = 25,
  OpTypeSampler = 26, OpTypeSampledImage = 27, OpTypeArray = 28, OpTypeRuntimeArray = 29,
  OpTypeStruct = 30, OpTypePointer = 32, OpTypeFunction = 33, OpConstantTrue = 41,
  OpConstantFalse = 42, OpConstant = 43, OpConstantComposite = 44, OpConstantNull = 46,
  OpFunction = 54, OpFunctionParameter = 55, OpFunctionEnd = 56, OpFunctionCall = 57,
  OpVariable = 59, OpLoad = 61, OpStore = 62, OpAccessChain = 65, OpInBoundsAccessChain = 66,
  OpDecorate = 71, OpMemberDecorate = 72, OpVectorShuffle = 79, OpCompositeConstruct = 80,
  OpCompositeExtract = 81, OpCompositeInsert = 82, OpCopyObject = 83, OpTranspose = 84,
  OpImageSampleImplicitLod = 87, OpImageSampleExplicitLod = 88, OpConvertFToU = 109,
  OpConvertFToS = 110, OpConvertSToF = 111, OpConvertUToF = 112, OpBitcast = 124,
  OpSNegate = 126, OpFNegate = 127, OpIAdd = 128, OpFAdd = 129, OpISub = 130, OpFSub = 131,
  OpIMul = 132, OpFMul = 133, OpUDiv = 134, OpSDiv = 135, OpFDiv = 136, OpUMod = 137,
  OpSRem = 138, OpSMod = 139, OpFRem = 140, OpFMod = 141, OpVectorTimesScalar = 142,
  OpMatrixTimesScalar = 143, OpVectorTimesMatrix = 144, OpMatrixTimesVector = 145,
  OpMatrixTimesMatrix = 146, OpDot = 148, OpAny = 154, OpAll = 155, OpLogicalEqual = 164,
  OpLogicalNotEqual = 165, OpLogicalOr = 166, OpLogicalAnd = 167, OpLogicalNot = 168,
  OpSelect = 169, OpIEqual = 170, OpINotEqual = 171, OpUGreaterThan = 172, OpSGreaterThan = 173,
  OpUGreaterThanEqual = 174, OpSGreaterThanEqual = 175, OpULessThan = 176, OpSLessThan = 177,
  OpULessThanEqual = 178, OpSLessThanEqual = 179, OpFOrdEqual = 180, OpFUnordEqual = 181,
  OpFOrdNotEqual = 182, OpFUnordNotEqual = 183, OpFOrdLessThan = 184, OpFUnordLessThan = 185,
  OpFOrdGreaterThan = 186, OpFUnordGreaterThan = 187, OpFOrdLessThanEqual = 188,
  OpFUnordLessThanEqual = 189, OpFOrdGreaterThanEqual = 190, OpFUnordGreaterThanEqual = 191,
  OpShiftRightLogical = 194, OpShiftRightArithmetic = 195, OpShiftLeftLogical = 196,
  OpBitwiseOr = 197, OpBitwiseXor = 198, OpBitwiseAnd = 199, OpNot = 200, OpDPdx = 207,
  OpDPdy = 208, OpFwidth = 209, OpDPdxFine = 210, OpDPdyFine = 211, OpDPdxCoarse = 213,
  OpDPdyCoarse = 214, OpPhi = 245, OpLoopMerge = 246, OpSelectionMerge = 247, OpLabel = 248,
  OpBranch = 249, OpBranchConditional = 250, OpSwitch = 251, OpKill = 252, OpReturn = 253,
  OpReturnValue = 254, OpUnreachable = 255, OpNoLine = 317, OpModuleProcessed = 330,
};

enum Glsl
{
  G_Round = 1, G_RoundEven = 2, G_Trunc = 3, G_FAbs = 4, G_SAbs = 5, G_FSign = 6, G_SSign = 7,
  G_Floor = 8, G_Ceil = 9, G_Fract = 10, G_Sin = 13, G_Cos = 14, G_Pow = 26, G_Exp2 = 29,
  G_Log2 = 30, G_Sqrt = 31, G_InverseSqrt = 32, G_MatrixInverse = 34, G_FMin = 37, G_UMin = 38,
  G_SMin = 39, G_FMax = 40, G_UMax = 41, G_SMax = 42, G_FClamp = 43, G_UClamp = 44,
  G_SClamp = 45, G_FMix = 46, G_Step = 48, G_SmoothStep = 49, G_Length = 66, G_Distance = 67,
  G_Cross = 68, G_Normalize = 69, G_Reflect = 71,
};

constexpr uint32_t kDimCube = 3;  // SPIR-V Dim operand of OpTypeImage

enum Dec
{
  D_Block = 2, D_RowMajor = 4, D_ColMajor = 5, D_ArrayStride = 6, D_MatrixStride = 7,
  D_BuiltIn = 11, D_Flat = 14, D_Location = 30, D_Binding = 33, D_DescriptorSet = 34,
  D_Offset = 35,
};

enum StorageClass
{
  SC_UniformConstant = 0, SC_Input = 1, SC_Uniform = 2, SC_Output = 3, SC_Workgroup = 4,
  SC_Private = 6, SC_Function = 7, SC_PushConstant = 9, SC_StorageBuffer = 12,
};

enum Leaf
{
  LF,  // float
  LU,  // 32-bit integer (signedness is per operation)
  LB,  // bool
};

struct Type
{
  uint16_t op = 0;
  uint32_t width = 32, signedness = 0;
  uint32_t elem = 0;   // vector comp / matrix column / array element / pointee / image type
  uint32_t count = 0;  // vector N, matrix cols, array length
  uint32_t sc = 0;     // pointer storage class
  std::vector<uint32_t> members;
  uint32_t dim = 0, arrayed = 0;
};

struct Inst
{
  uint16_t op, wc;
  const uint32_t *w;  // w[0] is the opcode word
};

struct Function
{
  uint32_t id = 0, ret_type = 0;
  std::vector<uint32_t> params;
  std::vector<Inst> body;  // from the first OpLabel to OpFunctionEnd (exclusive)
};

struct Val
{
  uint32_t type = 0;
  std::vector<std::string> c;  // scalar components (C expressions)
};

struct Ptr
{
  uint32_t var = 0;      // base variable id (global) or local storage name key
  std::string storage;   // C array name for register storage
  uint32_t pointee = 0;  // type id
  uint32_t sc = 0;
  std::string off = "0";  // flattened index (register) or byte offset (memory)
  uint32_t mat_stride = 16;
  uint32_t slot = 0;  // descriptor slot for memory / images
};

std::string Hex(uint32_t v)
{
  char b[16];
  snprintf(b, sizeof(b), "0x%08xu", v);
  return b;
}

class Translator
{
public:
  Translator(const uint32_t *w, size_t n, ShaderStage st, TranslatedShader &o)
      : words(w), nwords(n), stage(st), out(o)
  {
  }

  bool Run(const char *entry, std::string &err);

private:
  const uint32_t *words;
  size_t nwords;
  ShaderStage stage;
  TranslatedShader &out;
  std::string error;

  uint32_t glsl_set = 0;
  std::map<uint32_t, Type> types;
  std::map<uint32_t, std::vector<std::pair<uint32_t, std::vector<uint32_t>>>> decos;
  std::map<std::pair<uint32_t, uint32_t>, std::vector<std::pair<uint32_t, std::vector<uint32_t>>>> mdecos;
  std::map<uint32_t, Inst> constants;
  std::map<uint32_t, std::pair<uint32_t, uint32_t>> globals;  // id -> (ptr type, sc)
  std::vector<uint32_t> global_order;
  std::map<uint32_t, Function> functions;
  uint32_t entry_fn = 0;
  std::vector<uint32_t> interface_ids;
  bool deriv = false;
  std::string entry_name;

  std::ostringstream decl, body;
  int inst_counter = 0;
  int tmp_counter = 0;

  bool Fail(const std::string &e)
  {
    if(error.empty())
      error = e;
    return false;
  }

  // ---------------------------------------------------------------------------------------
  // type helpers
  const Type &T(uint32_t id)
  {
    static Type none;
    auto it = types.find(id);
    return it == types.end() ? none : it->second;
  }

  uint32_t FlatCount(uint32_t t)
  {
    const Type &ty = T(t);
    switch(ty.op)
    {
      case OpTypeBool: case OpTypeInt: case OpTypeFloat: return 1;
      case OpTypeVector: return ty.count * FlatCount(ty.elem);
      case OpTypeMatrix: return ty.count * FlatCount(ty.elem);
      case OpTypeArray: return ty.count * FlatCount(ty.elem);
      case OpTypeStruct:
      {
        uint32_t s = 0;
        for(uint32_t m : ty.members)
          s += FlatCount(m);
        return s;
      }
      case OpTypeSampledImage: case OpTypeImage: case OpTypeSampler: return 1;
      default: return 0;
    }
  }

  Leaf LeafKind(uint32_t t, uint32_t k)
  {
    const Type &ty = T(t);
    switch(ty.op)
    {
      case OpTypeBool: return LB;
      case OpTypeInt: return LU;
      case OpTypeFloat: return LF;
      case OpTypeVector: case OpTypeMatrix: case OpTypeArray:
      {
        uint32_t e = FlatCount(ty.elem);
        return LeafKind(ty.elem, e ? k % e : 0);
      }
      case OpTypeStruct:
        for(uint32_t m : ty.members)
        {
          uint32_t n = FlatCount(m);
          if(k < n)
            return LeafKind(m, k);
          k -= n;
        }
        return LU;
      default: return LU;
    }
  }

  bool AllFloat(uint32_t t)
  {
    uint32_t n = FlatCount(t);
    for(uint32_t k = 0; k < n; k++)
      if(LeafKind(t, k) != LF)
        return false;
    return true;
  }

  const char *CType(Leaf l) { return l == LF ? "F" : (l == LU ? "u32" : "bool"); }

  std::vector<uint32_t> DecoParams(uint32_t id, uint32_t dec)
  {
    auto it = decos.find(id);
    if(it != decos.end())
      for(auto &d : it->second)
        if(d.first == dec)
          return d.second.empty() ? std::vector<uint32_t>{0} : d.second;
    return {};
  }

  std::vector<uint32_t> MemberDecoParams(uint32_t st, uint32_t m, uint32_t dec)
  {
    auto it = mdecos.find({st, m});
    if(it != mdecos.end())
      for(auto &d : it->second)
        if(d.first == dec)
          return d.second.empty() ? std::vector<uint32_t>{0} : d.second;
    return {};
  }

  // byte offset of flattened leaf k inside a memory-laid-out type
  bool LeafByteOffset(uint32_t t, uint32_t k, uint32_t mat_stride, bool row_major, uint64_t &off)
  {
    const Type &ty = T(t);
    switch(ty.op)
    {
      case OpTypeBool: case OpTypeInt: case OpTypeFloat: off = 0; return true;
      case OpTypeVector: off = 4ull * k; return true;
      case OpTypeMatrix:
      {
        uint32_t rows = T(ty.elem).count;
        uint32_t col = k / rows, row = k % rows;
        off = row_major ? (uint64_t)row * mat_stride + 4ull * col : (uint64_t)col * mat_stride + 4ull * row;
        return true;
      }
      case OpTypeArray:
      {
        auto as = DecoParams(t, D_ArrayStride);
        if(as.empty())
          return Fail("array in a block without ArrayStride");
        uint32_t e = FlatCount(ty.elem);
        uint64_t inner;
        if(!LeafByteOffset(ty.elem, k % e, mat_stride, row_major, inner))
          return false;
        off = (uint64_t)as[0] * (k / e) + inner;
        return true;
      }
      case OpTypeStruct:
        for(uint32_t m = 0; m < ty.members.size(); m++)
        {
          uint32_t n = FlatCount(ty.members[m]);
          if(k < n)
          {
            auto o = MemberDecoParams(t, m, D_Offset);
            if(o.empty())
              return Fail("block member without Offset");
            auto ms = MemberDecoParams(t, m, D_MatrixStride);
            bool rm = !MemberDecoParams(t, m, D_RowMajor).empty();
            uint64_t inner;
            if(!LeafByteOffset(ty.members[m], k, ms.empty() ? 16 : ms[0], rm, inner))
              return false;
            off = o[0] + inner;
            return true;
          }
          k -= n;
        }
        return Fail("struct leaf out of range");
      default: return Fail("unsupported type in a block");
    }
  }

  // ---------------------------------------------------------------------------------------
  // code-generation state for one inlined function instance
  struct Frame
  {
    int inst;
    std::map<uint32_t, Val> vals;
    std::map<uint32_t, Ptr> ptrs;
    std::map<uint32_t, std::vector<std::pair<uint32_t, uint32_t>>> phis;  // label -> (phi id, ...)
    std::map<uint32_t, std::vector<std::pair<uint32_t, uint32_t>>> phi_src; // phi id -> (value, pred)
    std::string ret_label;
    Val *ret_val = nullptr;
  };

  std::string Label(const Frame &f, uint32_t id) { return "L" + std::to_string(f.inst) + "_" + std::to_string(id); }

  Val NewVal(Frame &f, uint32_t id, uint32_t type)
  {
    Val v;
    v.type = type;
    uint32_t n = FlatCount(type);
    for(uint32_t k = 0; k < n; k++)
    {
      std::string name = "v" + std::to_string(f.inst) + "_" + std::to_string(id) + "_" + std::to_string(k);
      decl << "  " << CType(LeafKind(type, k)) << " " << name << ";\n";
      v.c.push_back(name);
    }
    return v;
  }

  std::string FloatConst(uint32_t bits) { return "F(va_u2f(" + Hex(bits) + "))"; }

  bool ConstVal(uint32_t id, Val &v)
  {
    auto it = constants.find(id);
    if(it == constants.end())
      return false;
    const Inst &in = it->second;
    v.type = in.w[1];
    v.c.clear();
    switch(in.op)
    {
      case OpConstantTrue: v.c.push_back("true"); return true;
      case OpConstantFalse: v.c.push_back("false"); return true;
      case OpConstant:
        if(T(v.type).op == OpTypeFloat)
          v.c.push_back(FloatConst(in.w[3]));
        else
          v.c.push_back("(u32)" + Hex(in.w[3]));
        return true;
      case OpConstantComposite:
        for(uint16_t i = 3; i < in.wc; i++)
        {
          Val s;
          if(!ConstVal(in.w[i], s))
            return false;
          v.c.insert(v.c.end(), s.c.begin(), s.c.end());
        }
        return true;
      case OpConstantNull:
      {
        uint32_t n = FlatCount(v.type);
        for(uint32_t k = 0; k < n; k++)
        {
          Leaf l = LeafKind(v.type, k);
          v.c.push_back(l == LF ? "F(0.0f)" : (l == LU ? "(u32)0u" : "false"));
        }
        return true;
      }
      default: return false;
    }
  }

  bool Get(Frame &f, uint32_t id, Val &v)
  {
    auto it = f.vals.find(id);
    if(it != f.vals.end())
    {
      v = it->second;
      return true;
    }
    if(ConstVal(id, v))
      return true;
    return Fail("unknown value id " + std::to_string(id));
  }

  bool GetPtr(Frame &f, uint32_t id, Ptr &p)
  {
    auto it = f.ptrs.find(id);
    if(it != f.ptrs.end())
    {
      p = it->second;
      return true;
    }
    auto g = global_ptrs.find(id);
    if(g != global_ptrs.end())
    {
      p = g->second;
      return true;
    }
    return Fail("unknown pointer id " + std::to_string(id));
  }

  std::map<uint32_t, Ptr> global_ptrs;

  // scalar index expression (u32) for an access-chain index operand
  bool IndexExpr(Frame &f, uint32_t id, std::string &e, bool &is_const, uint32_t &cval)
  {
    auto it = constants.find(id);
    if(it != constants.end() && it->second.op == OpConstant)
    {
      is_const = true;
      cval = it->second.w[3];
      e = std::to_string(cval);
      return true;
    }
    Val v;
    if(!Get(f, id, v))
      return false;
    is_const = false;
    e = "(u32)(" + v.c[0] + ")";
    return true;
  }

  std::string Load(const Ptr &p, uint32_t k, Leaf want)
  {
    if(p.sc == SC_PushConstant || p.sc == SC_Uniform || p.sc == SC_StorageBuffer)
    {
      uint64_t bo = 0;
      bool rm = false;
      LeafByteOffset(p.pointee, k, p.mat_stride, rm, bo);
      std::string addr = "(" + p.off + " + " + std::to_string(bo) + "ull)";
      std::string raw = p.sc == SC_PushConstant ? "va_pc_u(d, (u32)" + addr + ")"
                                                : "va_ubo_u(d, " + std::to_string(p.slot) + "u, " + addr + ")";
      if(want == LF)
        return "F(va_u2f(" + raw + "))";
      if(want == LB)
        return "(" + raw + " != 0u)";
      return raw;
    }
    std::string e = p.storage + "[" + p.off + " + " + std::to_string(k) + "]";
    if(storage_float[p.storage])
      return want == LF ? e : (want == LU ? "va_f2u(cv(" + e + "))" : "(cv(" + e + ") != 0.0f)");
    return want == LF ? "va_bits_to_f(" + e + ", F())" : (want == LU ? e : "(" + e + " != 0u)");
  }

  std::string StoreExpr(const Ptr &p, uint32_t k, Leaf have, const std::string &v)
  {
    std::string e = p.storage + "[" + p.off + " + " + std::to_string(k) + "]";
    if(storage_float[p.storage])
      return e + " = " + (have == LF ? v : (have == LU ? "F(va_u2f(" + v + "))" : "F((" + v + ") ? 1.0f : 0.0f)")) + ";";
    return e + " = " + (have == LF ? "va_f2u(cv(" + v + "))" : (have == LU ? v : "(u32)(" + v + ")")) + ";";
  }

  std::map<std::string, bool> storage_float;

  std::string DeclStorage(const std::string &name, uint32_t type)
  {
    uint32_t n = FlatCount(type);
    if(n == 0)
      n = 1;
    bool fl = AllFloat(type);
    storage_float[name] = fl;
    decl << "  " << (fl ? "F" : "u32") << " " << name << "[" << n << "];\n";
    return name;
  }

  // ---------------------------------------------------------------------------------------
  bool Parse();
  bool SetupGlobals();
  bool EmitFunction(uint32_t fn, const std::vector<uint32_t> &args, Frame &caller, Val *ret, const std::string &ret_label);
  bool EmitInst(Frame &f, const Inst &in, uint32_t cur_label);
  bool EmitBranch(Frame &f, uint32_t from, uint32_t to);
  bool EmitExt(Frame &f, const Inst &in, Val &r);
  void Assign(const Val &r, uint32_t k, const std::string &e) { body << "  " << r.c[k] << " = " << e << ";\n"; }
};

bool Translator::Parse()
{
  if(nwords < 5 || words[0] != 0x07230203u)
    return Fail("not a SPIR-V module");
  size_t i = 5;
  Function *curf = nullptr;
  bool in_body = false;
  while(i < nwords)
  {
    Inst in;
    in.op = (uint16_t)(words[i] & 0xffffu);
    in.wc = (uint16_t)(words[i] >> 16);
    in.w = &words[i];
    if(in.wc == 0 || i + in.wc > nwords)
      return Fail("truncated instruction");
    switch(in.op)
    {
      case OpExtInstImport:
        if(!strcmp((const char *)&in.w[2], "GLSL.std.450"))
          glsl_set = in.w[1];
        break;
      case OpEntryPoint:
      {
        const char *name = (const char *)&in.w[3];
        size_t nlen = strlen(name);
        size_t nw = (nlen + 4) / 4;
        if(in.w[1] == (uint32_t)stage && entry_name == name)
        {
          entry_fn = in.w[2];
          for(size_t k = 3 + nw; k < in.wc; k++)
            interface_ids.push_back(in.w[k]);
        }
        break;
      }
      case OpDecorate:
        decos[in.w[1]].push_back({in.w[2], std::vector<uint32_t>(in.w + 3, in.w + in.wc)});
        break;
      case OpMemberDecorate:
        mdecos[{in.w[1], in.w[2]}].push_back({in.w[3], std::vector<uint32_t>(in.w + 4, in.w + in.wc)});
        break;
      case OpTypeVoid: case OpTypeBool: case OpTypeSampler:
        types[in.w[1]].op = in.op;
        break;
      case OpTypeInt:
        types[in.w[1]].op = in.op;
        types[in.w[1]].width = in.w[2];
        types[in.w[1]].signedness = in.w[3];
        if(in.w[2] != 32)
          return Fail("only 32-bit integers are supported");
        break;
      case OpTypeFloat:
        types[in.w[1]].op = in.op;
        types[in.w[1]].width = in.w[2];
        if(in.w[2] != 32)
          return Fail("only 32-bit floats are supported");
        break;
      case OpTypeVector: case OpTypeMatrix:
        types[in.w[1]].op = in.op;
        types[in.w[1]].elem = in.w[2];
        types[in.w[1]].count = in.w[3];
        break;
      case OpTypeArray:
      {
        types[in.w[1]].op = in.op;
        types[in.w[1]].elem = in.w[2];
        auto c = constants.find(in.w[3]);
        if(c == constants.end())
          return Fail("array length must be a constant");
        types[in.w[1]].count = c->second.w[3];
        break;
      }
      case OpTypeStruct:
        types[in.w[1]].op = in.op;
        types[in.w[1]].members.assign(in.w + 2, in.w + in.wc);
        break;
      case OpTypePointer:
        types[in.w[1]].op = in.op;
        types[in.w[1]].sc = in.w[2];
        types[in.w[1]].elem = in.w[3];
        break;
      case OpTypeFunction:
        types[in.w[1]].op = in.op;
        break;
      case OpTypeImage:
        types[in.w[1]].op = in.op;
        types[in.w[1]].dim = in.w[3];
        types[in.w[1]].arrayed = in.w[5];
        break;
      case OpTypeSampledImage:
        types[in.w[1]].op = in.op;
        types[in.w[1]].elem = in.w[2];
        types[in.w[1]].arrayed = T(in.w[2]).arrayed;
        types[in.w[1]].dim = T(in.w[2]).dim;
        break;
      case OpConstant: case OpConstantTrue: case OpConstantFalse: case OpConstantComposite:
      case OpConstantNull:
        constants[in.w[2]] = in;
        break;
      case OpVariable:
        if(!curf)
        {
          globals[in.w[2]] = {in.w[1], in.w[3]};
          global_order.push_back(in.w[2]);
        }
        else
          curf->body.push_back(in);
        break;
      case OpFunction:
        curf = &functions[in.w[2]];
        curf->id = in.w[2];
        curf->ret_type = in.w[1];
        in_body = false;
        break;
      case OpFunctionParameter:
        if(curf)
          curf->params.push_back(in.w[2]);
        break;
      case OpFunctionEnd:
        curf = nullptr;
        break;
      case OpLabel:
        in_body = true;
        if(curf)
          curf->body.push_back(in);
        break;
      default:
        if(curf && in_body)
          curf->body.push_back(in);
        break;
    }
    i += in.wc;
  }
  if(curf)
    return Fail("function %" + std::to_string(curf->id) + " has no OpFunctionEnd");
  if(!entry_fn)
    return Fail("entry point '" + entry_name + "' not found for this stage");
  return true;
}

bool Translator::SetupGlobals()
{
  for(uint32_t id : global_order)
  {
    uint32_t ptype = globals[id].first, sc = globals[id].second;
    uint32_t pointee = T(ptype).elem;
    Ptr p;
    p.var = id;
    p.pointee = pointee;
    p.sc = sc;
    if(sc == SC_PushConstant)
    {
      p.off = "0";
    }
    else if(sc == SC_Uniform || sc == SC_StorageBuffer || sc == SC_UniformConstant)
    {
      auto set = DecoParams(id, D_DescriptorSet), bind = DecoParams(id, D_Binding);
      uint32_t s = set.empty() ? 0 : set[0], b = bind.empty() ? 0 : bind[0];
      if(s != 0 || b >= 4)
        return Fail("descriptor (set " + std::to_string(s) + ", binding " + std::to_string(b) +
                    ") is outside set 0 bindings 0..3");
      p.slot = b;
      bool image = sc == SC_UniformConstant;
      out.resources.push_back({s, b, image});
      p.off = "0";
    }
    else
    {
      p.storage = DeclStorage("s_" + std::to_string(id), pointee);
    }
    global_ptrs[id] = p;
  }
  return true;
}

bool Translator::EmitBranch(Frame &f, uint32_t from, uint32_t to)
{
  auto it = f.phis.find(to);
  if(it != f.phis.end())
  {
    for(auto &ph : it->second)
    {
      uint32_t phi_id = ph.first;
      for(auto &src : f.phi_src[phi_id])
        if(src.second == from)
        {
          Val v, in;
          if(!Get(f, src.first, v))
            return false;
          in = f.vals[phi_id + 0x40000000u];
          for(size_t k = 0; k < v.c.size(); k++)
            body << "  " << in.c[k] << " = " << v.c[k] << ";\n";
        }
    }
  }
  body << "  goto " << Label(f, to) << ";\n";
  return true;
}

bool Translator::EmitFunction(uint32_t fn, const std::vector<uint32_t> &args, Frame &caller, Val *ret,
                              const std::string &ret_label)
{
  auto fit = functions.find(fn);
  if(fit == functions.end())
    return Fail("call to an undefined function");
  Function &F = fit->second;
  Frame f;
  f.inst = inst_counter++;
  f.ret_label = ret_label;
  f.ret_val = ret;
  if(args.size() != F.params.size())
    return Fail("call argument count mismatch");
  for(size_t i = 0; i < args.size(); i++)
  {
    Ptr p;
    auto pit = caller.ptrs.find(args[i]);
    if(pit != caller.ptrs.end())
      f.ptrs[F.params[i]] = pit->second;
    else if(global_ptrs.count(args[i]))
      f.ptrs[F.params[i]] = global_ptrs[args[i]];
    else
    {
      Val v;
      if(!Get(caller, args[i], v))
        return false;
      f.vals[F.params[i]] = v;
    }
  }
  // pre-scan phis
  uint32_t lbl = 0;
  for(const Inst &in : F.body)
  {
    if(in.op == OpLabel)
      lbl = in.w[1];
    else if(in.op == OpPhi)
    {
      f.phis[lbl].push_back({in.w[2], in.w[1]});
      for(uint16_t k = 3; k + 1 < in.wc; k += 2)
        f.phi_src[in.w[2]].push_back({in.w[k], in.w[k + 1]});
    }
  }
  // declare phi values and their edge copies up front
  for(auto &lp : f.phis)
    for(auto &ph : lp.second)
    {
      f.vals[ph.first] = NewVal(f, ph.first, ph.second);
      Val in = NewVal(f, ph.first + 0x40000000u, ph.second);
      f.vals[ph.first + 0x40000000u] = in;
    }
  // function-scope variables are declared once (hoisted)
  for(const Inst &in : F.body)
    if(in.op == OpVariable)
    {
      Ptr p;
      p.var = in.w[2];
      p.pointee = T(in.w[1]).elem;
      p.sc = SC_Function;
      p.storage = DeclStorage("s" + std::to_string(f.inst) + "_" + std::to_string(in.w[2]), p.pointee);
      f.ptrs[in.w[2]] = p;
    }
  uint32_t cur = 0;
  for(const Inst &in : F.body)
  {
    if(in.op == OpLabel)
    {
      cur = in.w[1];
      body << Label(f, cur) << ":;\n";
      auto it = f.phis.find(cur);
      if(it != f.phis.end())
        for(auto &ph : it->second)
        {
          Val &a = f.vals[ph.first];
          Val &b = f.vals[ph.first + 0x40000000u];
          for(size_t k = 0; k < a.c.size(); k++)
            body << "  " << a.c[k] << " = " << b.c[k] << ";\n";
        }
      continue;
    }
    if(in.op == OpVariable)
    {
      if(in.wc > 4)
      {
        Val init;
        if(!Get(f, in.w[4], init))
          return false;
        Ptr &p = f.ptrs[in.w[2]];
        for(size_t k = 0; k < init.c.size(); k++)
          body << "  " << StoreExpr(p, (uint32_t)k, LeafKind(p.pointee, (uint32_t)k), init.c[k]) << "\n";
      }
      continue;
    }
    if(!EmitInst(f, in, cur))
      return false;
  }
  return true;
}

bool Translator::EmitExt(Frame &f, const Inst &in, Val &r)
{
  if(in.w[3] != glsl_set)
    return Fail("unsupported extended instruction set");
  uint32_t inst = in.w[4];
  auto arg = [&](int n, Val &v) { return Get(f, in.w[5 + n], v); };
  uint32_t rt = in.w[1];
  uint32_t n = FlatCount(rt);
  Val a, b, c;
  switch(inst)
  {
    case G_FMax: case G_FMin:
      // reference spirv_compile.cpp:1535-1545: select(a > b, a, b) / select(a < b, a, b)
      if(!arg(0, a) || !arg(1, b)) return false;
      for(uint32_t k = 0; k < n; k++)
        Assign(r, k, "va_sel(cv(" + a.c[k] + ") " + (inst == G_FMax ? ">" : "<") + " cv(" + b.c[k] + "), " + a.c[k] + ", " + b.c[k] + ")");
      return true;
    case G_UMax: case G_UMin: case G_SMax: case G_SMin:
    {
      if(!arg(0, a) || !arg(1, b)) return false;
      bool s = inst == G_SMax || inst == G_SMin;
      bool mx = inst == G_UMax || inst == G_SMax;
      for(uint32_t k = 0; k < n; k++)
      {
        std::string x = s ? "(i32)" + a.c[k] : a.c[k], y = s ? "(i32)" + b.c[k] : b.c[k];
        Assign(r, k, "(" + x + (mx ? " > " : " < ") + y + ") ? " + a.c[k] + " : " + b.c[k]);
      }
      return true;
    }
    case G_FClamp:
      // reference :1546-1560: up = val < hi ? val : hi; up > lo ? up : lo
      if(!arg(0, a) || !arg(1, b) || !arg(2, c)) return false;
      for(uint32_t k = 0; k < n; k++)
      {
        std::string up = "va_sel(cv(" + a.c[k] + ") < cv(" + c.c[k] + "), " + a.c[k] + ", " + c.c[k] + ")";
        Assign(r, k, up);
        Assign(r, k, "va_sel(cv(" + r.c[k] + ") > cv(" + b.c[k] + "), " + r.c[k] + ", " + b.c[k] + ")");
      }
      return true;
    case G_UClamp: case G_SClamp:
    {
      if(!arg(0, a) || !arg(1, b) || !arg(2, c)) return false;
      bool s = inst == G_SClamp;
      for(uint32_t k = 0; k < n; k++)
      {
        std::string cast = s ? "(i32)" : "";
        Assign(r, k, "(" + cast + a.c[k] + " < " + cast + c.c[k] + ") ? " + a.c[k] + " : " + c.c[k]);
        Assign(r, k, "(" + cast + r.c[k] + " > " + cast + b.c[k] + ") ? " + r.c[k] + " : " + b.c[k]);
      }
      return true;
    }
    case G_FMix:
      // reference :1561-1579: (1 - a) * x + a * y
      if(!arg(0, a) || !arg(1, b) || !arg(2, c)) return false;
      for(uint32_t k = 0; k < n; k++)
      {
        std::string t = c.c.size() == 1 ? c.c[0] : c.c[k];
        Assign(r, k, "(F(1.0f) - " + t + ") * " + a.c[k] + " + " + t + " * " + b.c[k]);
      }
      return true;
    case G_Sqrt:
      if(!arg(0, a)) return false;
      for(uint32_t k = 0; k < n; k++)
        Assign(r, k, "va_sqrt(" + a.c[k] + ")");
      return true;
    case G_InverseSqrt:
      // reference :1604-1634: 1 / sqrt(x) (component-wise for vectors)
      if(!arg(0, a)) return false;
      for(uint32_t k = 0; k < n; k++)
        Assign(r, k, "F(1.0f) / va_sqrt(" + a.c[k] + ")");
      return true;
    case G_Normalize:
    {
      // reference :1635-1647: sqrlen = CreateDot(a, a); inv = 1 / sqrt(sqrlen); a * inv
      if(!arg(0, a)) return false;
      std::string t = "t" + std::to_string(tmp_counter++);
      decl << "  F " << t << ";\n";
      body << "  " << t << " = " << a.c[0] << " * " << a.c[0] << ";\n";
      for(uint32_t k = 1; k < n; k++)
        body << "  " << t << " = " << t << " + " << a.c[k] << " * " << a.c[k] << ";\n";
      body << "  " << t << " = F(1.0f) / va_sqrt(" << t << ");\n";
      for(uint32_t k = 0; k < n; k++)
        Assign(r, k, a.c[k] + " * " + t);
      return true;
    }
    case G_Length: case G_Distance:
    {
      // reference :1648-1656: sqrt(CreateDot(a, a))
      if(!arg(0, a)) return false;
      uint32_t m = (uint32_t)a.c.size();
      std::vector<std::string> x = a.c;
      if(inst == G_Distance)
      {
        if(!arg(1, b)) return false;
        for(uint32_t k = 0; k < m; k++)
          x[k] = "(" + a.c[k] + " - " + b.c[k] + ")";
      }
      Assign(r, 0, x[0] + " * " + x[0]);
      for(uint32_t k = 1; k < m; k++)
        Assign(r, 0, r.c[0] + " + " + x[k] + " * " + x[k]);
      Assign(r, 0, "va_sqrt(" + r.c[0] + ")");
      return true;
    }
    case G_Cross:
      // the reference returns its first argument (:1657-1663); this is the cross product
      if(!arg(0, a) || !arg(1, b)) return false;
      Assign(r, 0, a.c[1] + " * " + b.c[2] + " - " + a.c[2] + " * " + b.c[1]);
      Assign(r, 1, a.c[2] + " * " + b.c[0] + " - " + a.c[0] + " * " + b.c[2]);
      Assign(r, 2, a.c[0] + " * " + b.c[1] + " - " + a.c[1] + " * " + b.c[0]);
      return true;
    case G_Reflect:
    {
      // reference :1690-1702: I - (dot(I, N) * 2) * N
      if(!arg(0, a) || !arg(1, b)) return false;
      std::string t = "t" + std::to_string(tmp_counter++);
      decl << "  F " << t << ";\n";
      body << "  " << t << " = " << a.c[0] << " * " << b.c[0] << ";\n";
      for(uint32_t k = 1; k < n; k++)
        body << "  " << t << " = " << t << " + " << a.c[k] << " * " << b.c[k] << ";\n";
      body << "  " << t << " = " << t << " * F(2.0f);\n";
      for(uint32_t k = 0; k < n; k++)
        Assign(r, k, a.c[k] + " - " + t + " * " + b.c[k]);
      return true;
    }
    case G_Floor: case G_Ceil: case G_Trunc: case G_FAbs: case G_Sin: case G_Cos: case G_Exp2:
    case G_Log2: case G_Fract: case G_Round: case G_RoundEven: case G_FSign:
    {
      if(!arg(0, a)) return false;
      for(uint32_t k = 0; k < n; k++)
      {
        const std::string &x = a.c[k];
        std::string e;
        switch(inst)
        {
          case G_Floor: e = "va_floor(" + x + ")"; break;
          case G_Ceil: e = "va_ceil(" + x + ")"; break;
          case G_Trunc: e = "va_trunc(" + x + ")"; break;
          case G_FAbs: e = "va_fabs(" + x + ")"; break;
          case G_Sin: e = "va_sin(" + x + ")"; break;
          case G_Cos: e = "va_cos(" + x + ")"; break;
          case G_Exp2: e = "va_exp2(" + x + ")"; break;
          case G_Log2: e = "va_log2f(" + x + ")"; break;
          case G_Fract: e = x + " - va_floor(" + x + ")"; break;
          case G_Round: case G_RoundEven: e = "va_rint(" + x + ")"; break;
          case G_FSign:
            e = "va_sel(cv(" + x + ") > 0.0f, F(1.0f), va_sel(cv(" + x + ") < 0.0f, F(-1.0f), F(0.0f)))";
            break;
        }
        Assign(r, k, e);
      }
      return true;
    }
    case G_Pow:
      if(!arg(0, a) || !arg(1, b)) return false;
      for(uint32_t k = 0; k < n; k++)
        Assign(r, k, "va_pow(" + a.c[k] + ", " + b.c[k] + ")");
      return true;
    case G_Step:
      if(!arg(0, a) || !arg(1, b)) return false;
      for(uint32_t k = 0; k < n; k++)
      {
        std::string e = a.c.size() == 1 ? a.c[0] : a.c[k];
        Assign(r, k, "va_sel(cv(" + b.c[k] + ") < cv(" + e + "), F(0.0f), F(1.0f))");
      }
      return true;
    case G_SmoothStep:
    {
      if(!arg(0, a) || !arg(1, b) || !arg(2, c)) return false;
      for(uint32_t k = 0; k < n; k++)
      {
        std::string e0 = a.c.size() == 1 ? a.c[0] : a.c[k], e1 = b.c.size() == 1 ? b.c[0] : b.c[k];
        Assign(r, k, "(" + c.c[k] + " - " + e0 + ") / (" + e1 + " - " + e0 + ")");
        Assign(r, k, "va_sel(cv(" + r.c[k] + ") < 0.0f, F(0.0f), va_sel(cv(" + r.c[k] + ") > 1.0f, F(1.0f), " + r.c[k] + "))");
        Assign(r, k, r.c[k] + " * " + r.c[k] + " * (F(3.0f) - F(2.0f) * " + r.c[k] + ")");
      }
      return true;
    }
    case G_SAbs:
      if(!arg(0, a)) return false;
      for(uint32_t k = 0; k < n; k++)
        Assign(r, k, "(u32)((i32)" + a.c[k] + " < 0 ? -(i32)" + a.c[k] + " : (i32)" + a.c[k] + ")");
      return true;
    case G_MatrixInverse:
    {
      // the reference lowers this to a transpose (:1703-1730); this is the 4x4 inverse, by the
      // cofactor formula of the reference's own Float4x4Inverse helper (:493-550)
      if(!arg(0, a)) return false;
      if(n != 16) return Fail("MatrixInverse supports 4x4 only");
      const std::vector<std::string> &m = a.c;
      auto d = [&](const char *nm, int i0, int i1, int i2, int i3) {
        decl << "  F " << nm << std::to_string(tmp_counter) << ";\n";
        body << "  " << nm << tmp_counter << " = " << m[i0] << " * " << m[i1] << " - " << m[i2] << " * " << m[i3] << ";\n";
      };
      int tc = tmp_counter;
      d("a0_", 0, 5, 1, 4); d("a1_", 0, 6, 2, 4); d("a2_", 0, 7, 3, 4); d("a3_", 1, 6, 2, 5);
      d("a4_", 1, 7, 3, 5); d("a5_", 2, 7, 3, 6); d("b0_", 8, 13, 9, 12); d("b1_", 8, 14, 10, 12);
      d("b2_", 8, 15, 11, 12); d("b3_", 9, 14, 10, 13); d("b4_", 9, 15, 11, 13); d("b5_", 10, 15, 11, 14);
      tmp_counter++;
      auto A = [&](int i) { return "a" + std::to_string(i) + "_" + std::to_string(tc); };
      auto B = [&](int i) { return "b" + std::to_string(i) + "_" + std::to_string(tc); };
      std::string det = "det_" + std::to_string(tc);
      decl << "  F " << det << ";\n";
      body << "  " << det << " = " << A(0) << " * " << B(5) << " - " << A(1) << " * " << B(4) << " + " << A(2) << " * " << B(3)
           << " + " << A(3) << " * " << B(2) << " - " << A(4) << " * " << B(1) << " + " << A(5) << " * " << B(0) << ";\n";
      const char *fm[16] = {
          "+m5*b5 - m6*b4 + m7*b3", "-m1*b5 + m2*b4 - m3*b3", "+m13*a5 - m14*a4 + m15*a3", "-m9*a5 + m10*a4 - m11*a3",
          "-m4*b5 + m6*b2 - m7*b1", "+m0*b5 - m2*b2 + m3*b1", "-m12*a5 + m14*a2 - m15*a1", "+m8*a5 - m10*a2 + m11*a1",
          "+m4*b4 - m5*b2 + m7*b0", "-m0*b4 + m1*b2 - m3*b0", "+m12*a4 - m13*a2 + m15*a0", "-m8*a4 + m9*a2 - m11*a0",
          "-m4*b3 + m5*b1 - m6*b0", "+m0*b3 - m1*b1 + m2*b0", "-m12*a3 + m13*a1 - m14*a0", "+m8*a3 - m9*a1 + m10*a0"};
      for(int k = 0; k < 16; k++)
      {
        std::string e;
        const char *p = fm[k];
        while(*p)
        {
          if(*p == 'm')
          {
            int idx = atoi(p + 1);
            e += m[idx];
            p++;
            while(*p >= '0' && *p <= '9') p++;
          }
          else if(*p == 'a' || *p == 'b')
          {
            int idx = p[1] - '0';
            e += (*p == 'a') ? A(idx) : B(idx);
            p += 2;
          }
          else if(*p == '+' && e.empty())
            p++;
          else if(*p == '-' && e.empty())
          {
            e += "-";
            p++;
          }
          else
          {
            e += std::string(" ") + *p + " ";
            p++;
          }
        }
        // the reference helper returns zeros unless |det| > FLT_EPSILON
        Assign(r, k, "va_sel(cv(va_fabs(" + det + ")) > 1.1920929e-7f, (" + e + ") * (F(1.0f) / " + det + "), F(0.0f))");
      }
      return true;
    }
    default:
      return Fail("unsupported GLSL.std.450 instruction " + std::to_string(inst));
  }
}

bool Translator::EmitInst(Frame &f, const Inst &in, uint32_t cur_label)
{
  const uint32_t *w = in.w;
  auto result = [&](uint32_t type_word, uint32_t id_word) -> Val & {
    f.vals[w[id_word]] = NewVal(f, w[id_word], w[type_word]);
    return f.vals[w[id_word]];
  };
  Val a, b, c;
  switch(in.op)
  {
    case OpNop: case OpLine: case OpNoLine: case OpSelectionMerge: case OpLoopMerge:
    case OpUnreachable: case OpPhi:
      return true;
    case OpLoad:
    {
      Ptr p;
      if(!GetPtr(f, w[3], p))
        return false;
      if(p.sc == SC_UniformConstant)
      {
        Val v;
        v.type = w[1];
        v.c.push_back(std::to_string(p.slot) + "u");
        f.vals[w[2]] = v;
        return true;
      }
      Val &r = result(1, 2);
      for(uint32_t k = 0; k < r.c.size(); k++)
        Assign(r, k, Load(p, k, LeafKind(w[1], k)));
      return true;
    }
    case OpStore:
    {
      Ptr p;
      if(!GetPtr(f, w[1], p) || !Get(f, w[2], a))
        return false;
      if(p.storage.empty())
        return Fail("store to a read-only storage class");
      for(uint32_t k = 0; k < a.c.size(); k++)
        body << "  " << StoreExpr(p, k, LeafKind(a.type, k), a.c[k]) << "\n";
      return true;
    }
    case OpAccessChain: case OpInBoundsAccessChain:
    {
      Ptr p;
      if(!GetPtr(f, w[3], p))
        return false;
      bool mem = p.storage.empty();
      uint32_t t = p.pointee;
      for(uint16_t i = 4; i < in.wc; i++)
      {
        std::string ie;
        bool isc;
        uint32_t cv;
        if(!IndexExpr(f, w[i], ie, isc, cv))
          return false;
        const Type &ty = T(t);
        if(ty.op == OpTypeStruct)
        {
          if(!isc || cv >= ty.members.size())
            return Fail("struct index must be constant");
          if(mem)
          {
            auto o = MemberDecoParams(t, cv, D_Offset);
            auto ms = MemberDecoParams(t, cv, D_MatrixStride);
            if(!MemberDecoParams(t, cv, D_RowMajor).empty())
              return Fail("RowMajor block members are not supported");
            p.off = "(" + p.off + " + " + std::to_string(o.empty() ? 0 : o[0]) + ")";
            if(!ms.empty())
              p.mat_stride = ms[0];
          }
          else
          {
            uint32_t s = 0;
            for(uint32_t m = 0; m < cv; m++)
              s += FlatCount(ty.members[m]);
            p.off = "(" + p.off + " + " + std::to_string(s) + ")";
          }
          t = ty.members[cv];
        }
        else if(ty.op == OpTypeArray || ty.op == OpTypeMatrix || ty.op == OpTypeVector)
        {
          uint64_t stride;
          if(mem)
          {
            if(ty.op == OpTypeArray)
            {
              auto as = DecoParams(t, D_ArrayStride);
              stride = as.empty() ? 16 : as[0];
            }
            else if(ty.op == OpTypeMatrix)
              stride = p.mat_stride;
            else
              stride = 4;
          }
          else
            stride = FlatCount(ty.elem);
          p.off = "(" + p.off + " + " + ie + " * " + std::to_string(stride) + "u)";
          t = ty.elem;
        }
        else
          return Fail("access chain into a scalar");
      }
      p.pointee = t;
      f.ptrs[w[2]] = p;
      return true;
    }
    case OpCompositeExtract:
    {
      if(!Get(f, w[3], a))
        return false;
      uint32_t t = a.type;
      uint32_t off = 0;
      for(uint16_t i = 4; i < in.wc; i++)
      {
        const Type &ty = T(t);
        uint32_t idx = w[i];
        if(ty.op == OpTypeStruct)
        {
          for(uint32_t m = 0; m < idx; m++)
            off += FlatCount(ty.members[m]);
          t = ty.members[idx];
        }
        else
        {
          off += idx * FlatCount(ty.elem);
          t = ty.elem;
        }
      }
      Val v;
      v.type = w[1];
      for(uint32_t k = 0; k < FlatCount(w[1]); k++)
        v.c.push_back(a.c[off + k]);
      f.vals[w[2]] = v;
      return true;
    }
    case OpCompositeInsert:
    {
      Val obj, comp;
      if(!Get(f, w[3], obj) || !Get(f, w[4], comp))
        return false;
      uint32_t t = comp.type;
      uint32_t off = 0;
      for(uint16_t i = 5; i < in.wc; i++)
      {
        const Type &ty = T(t);
        uint32_t idx = w[i];
        if(ty.op == OpTypeStruct)
        {
          for(uint32_t m = 0; m < idx; m++)
            off += FlatCount(ty.members[m]);
          t = ty.members[idx];
        }
        else
        {
          off += idx * FlatCount(ty.elem);
          t = ty.elem;
        }
      }
      Val &r = result(1, 2);
      for(uint32_t k = 0; k < r.c.size(); k++)
        Assign(r, k, comp.c[k]);
      for(uint32_t k = 0; k < obj.c.size(); k++)
        Assign(r, off + k, obj.c[k]);
      return true;
    }
    case OpCompositeConstruct:
    {
      Val v;
      v.type = w[1];
      for(uint16_t i = 3; i < in.wc; i++)
      {
        if(!Get(f, w[i], a))
          return false;
        v.c.insert(v.c.end(), a.c.begin(), a.c.end());
      }
      f.vals[w[2]] = v;
      return true;
    }
    case OpCopyObject:
      if(!Get(f, w[3], a))
        return false;
      f.vals[w[2]] = a;
      f.vals[w[2]].type = w[1];
      return true;
    case OpVectorShuffle:
    {
      if(!Get(f, w[3], a) || !Get(f, w[4], b))
        return false;
      Val v;
      v.type = w[1];
      for(uint16_t i = 5; i < in.wc; i++)
      {
        uint32_t s = w[i];
        if(s == 0xffffffffu)
          v.c.push_back("F(0.0f)");
        else
          v.c.push_back(s < a.c.size() ? a.c[s] : b.c[s - a.c.size()]);
      }
      f.vals[w[2]] = v;
      return true;
    }
    case OpFAdd: case OpFSub: case OpFMul: case OpFDiv:
    case OpIAdd: case OpISub: case OpIMul:
    case OpBitwiseAnd: case OpBitwiseOr: case OpBitwiseXor:
    case OpShiftLeftLogical: case OpShiftRightLogical: case OpShiftRightArithmetic:
    case OpUDiv: case OpSDiv: case OpUMod: case OpSRem: case OpSMod: case OpFMod: case OpFRem:
    {
      if(!Get(f, w[3], a) || !Get(f, w[4], b))
        return false;
      Val &r = result(1, 2);
      for(uint32_t k = 0; k < r.c.size(); k++)
      {
        const std::string &x = a.c[k], &y = b.c[k];
        std::string e;
        switch(in.op)
        {
          case OpFAdd: e = x + " + " + y; break;
          case OpFSub: e = x + " - " + y; break;
          case OpFMul: e = x + " * " + y; break;
          case OpFDiv: e = x + " / " + y; break;
          case OpIAdd: e = "(u32)(" + x + " + " + y + ")"; break;
          case OpISub: e = "(u32)(" + x + " - " + y + ")"; break;
          case OpIMul: e = "(u32)(" + x + " * " + y + ")"; break;
          case OpBitwiseAnd: e = x + " & " + y; break;
          case OpBitwiseOr: e = x + " | " + y; break;
          case OpBitwiseXor: e = x + " ^ " + y; break;
          // the reference lowers ShiftLeftLogical to a logical right shift (:1212-1216); this
          // is the left shift the opcode names
          case OpShiftLeftLogical: e = "(u32)(" + x + " << (" + y + " & 31u))"; break;
          case OpShiftRightLogical: e = "(u32)(" + x + " >> (" + y + " & 31u))"; break;
          case OpShiftRightArithmetic: e = "(u32)((i32)" + x + " >> (" + y + " & 31u))"; break;
          case OpUDiv: e = "(" + y + " ? " + x + " / " + y + " : 0u)"; break;
          case OpSDiv: e = "(" + y + " ? (u32)((i32)" + x + " / (i32)" + y + ") : 0u)"; break;
          case OpUMod: e = "(" + y + " ? " + x + " % " + y + " : 0u)"; break;
          case OpSRem: e = "(" + y + " ? (u32)((i32)" + x + " % (i32)" + y + ") : 0u)"; break;
          case OpSMod:
            e = "(" + y + " ? (u32)((((i32)" + x + " % (i32)" + y + ") + (i32)" + y + ") % (i32)" + y + ") : 0u)";
            break;
          case OpFMod: e = x + " - " + y + " * va_floor(" + x + " / " + y + ")"; break;
          case OpFRem: e = x + " - " + y + " * va_trunc(" + x + " / " + y + ")"; break;
        }
        Assign(r, k, e);
      }
      return true;
    }
    case OpFNegate: case OpSNegate: case OpNot: case OpLogicalNot:
    {
      if(!Get(f, w[3], a))
        return false;
      Val &r = result(1, 2);
      for(uint32_t k = 0; k < r.c.size(); k++)
        Assign(r, k, in.op == OpFNegate ? "-" + a.c[k]
                                        : (in.op == OpSNegate ? "(u32)(-(i32)" + a.c[k] + ")"
                                                              : (in.op == OpNot ? "~" + a.c[k] : "!" + a.c[k])));
      return true;
    }
    case OpVectorTimesScalar: case OpMatrixTimesScalar:
    {
      // reference :1470-1476 (splat multiply), :1417-1443 (Float4x4TimesFloat)
      if(!Get(f, w[3], a) || !Get(f, w[4], b))
        return false;
      Val &r = result(1, 2);
      for(uint32_t k = 0; k < r.c.size(); k++)
        Assign(r, k, a.c[k] + " * " + b.c[0]);
      return true;
    }
    case OpMatrixTimesVector:
    {
      // reference :1324-1386 -> Float4x4TimesVec4 / Float3x3TimesVec3 (:423-441):
      // out[r] = (((0 + M[0][r] v0) + M[1][r] v1) + ...)
      if(!Get(f, w[3], a) || !Get(f, w[4], b))
        return false;
      const Type &mt = T(a.type);
      uint32_t cols = mt.count, rows = T(mt.elem).count;
      Val &r = result(1, 2);
      for(uint32_t row = 0; row < rows; row++)
      {
        Assign(r, row, "F(0.0f)");
        for(uint32_t col = 0; col < cols; col++)
          Assign(r, row, r.c[row] + " + " + a.c[col * rows + row] + " * " + b.c[col]);
      }
      return true;
    }
    case OpVectorTimesMatrix:
    {
      // reference Vec4TimesFloat4x4 (:443-461): out[j] = 0 + sum_i M[j][i] * v[i]
      if(!Get(f, w[3], a) || !Get(f, w[4], b))
        return false;
      const Type &mt = T(b.type);
      uint32_t cols = mt.count, rows = T(mt.elem).count;
      Val &r = result(1, 2);
      for(uint32_t j = 0; j < cols; j++)
      {
        Assign(r, j, "F(0.0f)");
        for(uint32_t i = 0; i < rows; i++)
          Assign(r, j, r.c[j] + " + " + b.c[j * rows + i] + " * " + a.c[i]);
      }
      return true;
    }
    case OpMatrixTimesMatrix:
    {
      // reference Float4x4TimesFloat4x4 (:463-473): out[x][y] = b[x][0] a[0][y] + ... (no leading 0)
      if(!Get(f, w[3], a) || !Get(f, w[4], b))
        return false;
      const Type &at = T(a.type), &bt = T(b.type);
      uint32_t rows = T(at.elem).count, inner = at.count, cols = bt.count;
      Val &r = result(1, 2);
      for(uint32_t x = 0; x < cols; x++)
        for(uint32_t y = 0; y < rows; y++)
        {
          uint32_t k = x * rows + y;
          Assign(r, k, b.c[x * inner + 0] + " * " + a.c[0 * rows + y]);
          for(uint32_t i = 1; i < inner; i++)
            Assign(r, k, r.c[k] + " + " + b.c[x * inner + i] + " * " + a.c[i * rows + y]);
        }
      return true;
    }
    case OpTranspose:
    {
      if(!Get(f, w[3], a))
        return false;
      const Type &at = T(a.type);
      uint32_t cols = at.count, rows = T(at.elem).count;
      Val &r = result(1, 2);
      for(uint32_t x = 0; x < cols; x++)
        for(uint32_t y = 0; y < rows; y++)
          Assign(r, y * cols + x, a.c[x * rows + y]);
      return true;
    }
    case OpDot:
    {
      // reference CreateDot (:629-643): ((a0 b0 + a1 b1) + a2 b2) + a3 b3
      if(!Get(f, w[3], a) || !Get(f, w[4], b))
        return false;
      Val &r = result(1, 2);
      Assign(r, 0, a.c[0] + " * " + b.c[0]);
      for(uint32_t k = 1; k < a.c.size(); k++)
        Assign(r, 0, r.c[0] + " + " + a.c[k] + " * " + b.c[k]);
      return true;
    }
    case OpConvertSToF: case OpConvertUToF: case OpConvertFToS: case OpConvertFToU: case OpBitcast:
    {
      if(!Get(f, w[3], a))
        return false;
      Val &r = result(1, 2);
      for(uint32_t k = 0; k < r.c.size(); k++)
      {
        std::string e;
        Leaf src = LeafKind(a.type, k), dst = LeafKind(w[1], k);
        switch(in.op)
        {
          case OpConvertSToF: e = "F((float)(i32)" + a.c[k] + ")"; break;
          case OpConvertUToF: e = "F((float)" + a.c[k] + ")"; break;
          case OpConvertFToS: e = "(u32)va_f2i(cv(" + a.c[k] + "))"; break;
          case OpConvertFToU: e = "va_f2u32(cv(" + a.c[k] + "))"; break;
          default:
            if(src == dst) e = a.c[k];
            else if(src == LF) e = "va_f2u(cv(" + a.c[k] + "))";
            else e = "va_bits_to_f(" + a.c[k] + ", F())";
            break;
        }
        Assign(r, k, e);
      }
      return true;
    }
    case OpFOrdEqual: case OpFUnordEqual: case OpFOrdNotEqual: case OpFUnordNotEqual:
    case OpFOrdLessThan: case OpFUnordLessThan: case OpFOrdGreaterThan: case OpFUnordGreaterThan:
    case OpFOrdLessThanEqual: case OpFUnordLessThanEqual: case OpFOrdGreaterThanEqual:
    case OpFUnordGreaterThanEqual: case OpIEqual: case OpINotEqual: case OpUGreaterThan:
    case OpSGreaterThan: case OpUGreaterThanEqual: case OpSGreaterThanEqual: case OpULessThan:
    case OpSLessThan: case OpULessThanEqual: case OpSLessThanEqual: case OpLogicalEqual:
    case OpLogicalNotEqual: case OpLogicalOr: case OpLogicalAnd:
    {
      if(!Get(f, w[3], a) || !Get(f, w[4], b))
        return false;
      Val &r = result(1, 2);
      for(uint32_t k = 0; k < r.c.size(); k++)
      {
        std::string x = a.c[k], y = b.c[k], e;
        auto fx = "cv(" + x + ")", fy = "cv(" + y + ")";
        auto sx = "(i32)" + x, sy = "(i32)" + y;
        switch(in.op)
        {
          case OpFOrdEqual: e = fx + " == " + fy; break;
          case OpFUnordEqual: e = "!(" + fx + " != " + fy + ")"; break;
          case OpFOrdNotEqual: e = "(" + fx + " < " + fy + " || " + fx + " > " + fy + ")"; break;
          case OpFUnordNotEqual: e = fx + " != " + fy; break;
          case OpFOrdLessThan: e = fx + " < " + fy; break;
          case OpFUnordLessThan: e = "!(" + fx + " >= " + fy + ")"; break;
          case OpFOrdGreaterThan: e = fx + " > " + fy; break;
          case OpFUnordGreaterThan: e = "!(" + fx + " <= " + fy + ")"; break;
          case OpFOrdLessThanEqual: e = fx + " <= " + fy; break;
          case OpFUnordLessThanEqual: e = "!(" + fx + " > " + fy + ")"; break;
          case OpFOrdGreaterThanEqual: e = fx + " >= " + fy; break;
          case OpFUnordGreaterThanEqual: e = "!(" + fx + " < " + fy + ")"; break;
          case OpIEqual: case OpLogicalEqual: e = x + " == " + y; break;
          case OpINotEqual: case OpLogicalNotEqual: e = x + " != " + y; break;
          case OpUGreaterThan: e = x + " > " + y; break;
          case OpSGreaterThan: e = sx + " > " + sy; break;
          case OpUGreaterThanEqual: e = x + " >= " + y; break;
          case OpSGreaterThanEqual: e = sx + " >= " + sy; break;
          case OpULessThan: e = x + " < " + y; break;
          case OpSLessThan: e = sx + " < " + sy; break;
          case OpULessThanEqual: e = x + " <= " + y; break;
          case OpSLessThanEqual: e = sx + " <= " + sy; break;
          case OpLogicalOr: e = x + " || " + y; break;
          case OpLogicalAnd: e = x + " && " + y; break;
        }
        Assign(r, k, e);
      }
      return true;
    }
    case OpAny: case OpAll:
    {
      if(!Get(f, w[3], a))
        return false;
      Val &r = result(1, 2);
      std::string e = a.c[0];
      for(size_t k = 1; k < a.c.size(); k++)
        e += (in.op == OpAny ? " || " : " && ") + a.c[k];
      Assign(r, 0, e);
      return true;
    }
    case OpSelect:
    {
      Val cnd;
      if(!Get(f, w[3], cnd) || !Get(f, w[4], a) || !Get(f, w[5], b))
        return false;
      Val &r = result(1, 2);
      for(uint32_t k = 0; k < r.c.size(); k++)
      {
        const std::string &cc = cnd.c.size() == 1 ? cnd.c[0] : cnd.c[k];
        if(LeafKind(w[1], k) == LF)
          Assign(r, k, "va_sel(" + cc + ", " + a.c[k] + ", " + b.c[k] + ")");
        else
          Assign(r, k, "(" + cc + ") ? " + a.c[k] + " : " + b.c[k]);
      }
      return true;
    }
    case OpDPdx: case OpDPdy: case OpDPdxFine: case OpDPdyFine: case OpDPdxCoarse: case OpDPdyCoarse:
    case OpFwidth:
    {
      if(!Get(f, w[3], a))
        return false;
      Val &r = result(1, 2);
      for(uint32_t k = 0; k < r.c.size(); k++)
      {
        bool dx = in.op == OpDPdx || in.op == OpDPdxFine || in.op == OpDPdxCoarse;
        if(in.op == OpFwidth)
          Assign(r, k, "va_fabs(va_dpdx(" + a.c[k] + ")) + va_fabs(va_dpdy(" + a.c[k] + "))");
        else
          Assign(r, k, std::string(dx ? "va_dpdx(" : "va_dpdy(") + a.c[k] + ")");
      }
      return true;
    }
    case OpExtInst:
    {
      Val &r = result(1, 2);
      return EmitExt(f, in, r);
    }
    case OpImageSampleImplicitLod: case OpImageSampleExplicitLod:
    {
      Val img, coord;
      if(!Get(f, w[3], img) || !Get(f, w[4], coord))
        return false;
      Val &r = result(1, 2);
      const Type &it = T(img.type);
      // reference spirv_compile.cpp:896-899, 1290-1291, 1847-1881: a sampled image of a Cube
      // image type samples by direction (x, y, z); every other one by (u, v)
      const bool cube = it.dim == kDimCube;
      if(coord.c.size() < (cube ? 3u : 2u))
        return Fail("sampling coordinate has too few components");
      std::string layer = (it.arrayed && coord.c.size() > 2) ? coord.c[2] : "F(0.0f)";
      std::string t = "tx" + std::to_string(tmp_counter++);
      body << "  { float " << t << "[4];\n";
      std::string lod_arg;
      if(in.op == OpImageSampleExplicitLod)
      {
        uint32_t mask = in.wc > 5 ? w[5] : 0;
        if(!(mask & 2u))
          return Fail("explicit-LOD sampling needs the Lod operand");
        Val lod;
        if(!Get(f, w[6], lod))
          return false;
        lod_arg = lod.c[0];
      }
      if(cube)
        body << "  va_tex_cube(d, " << img.c[0] << ", " << coord.c[0] << ", " << coord.c[1] << ", " << coord.c[2] << ", "
             << (lod_arg.empty() ? "0, F(0.0f)" : "2, " + lod_arg) << ", " << t << ");\n";
      else if(!lod_arg.empty())
        body << "  va_tex_lod(d, " << img.c[0] << ", " << coord.c[0] << ", " << coord.c[1] << ", " << layer << ", "
             << lod_arg << ", " << t << ");\n";
      else
        body << "  va_tex(d, " << img.c[0] << ", " << coord.c[0] << ", " << coord.c[1] << ", " << layer << ", " << t << ");\n";
      for(uint32_t k = 0; k < r.c.size() && k < 4; k++)
        body << "  " << r.c[k] << " = F(" << t << "[" << k << "]);\n";
      body << "  }\n";
      return true;
    }
    case OpFunctionCall:
    {
      std::vector<uint32_t> args(w + 4, w + in.wc);
      Val *rv = nullptr;
      if(T(w[1]).op != OpTypeVoid)
      {
        result(1, 2);
        rv = &f.vals[w[2]];
      }
      std::string lbl = "R" + std::to_string(f.inst) + "_" + std::to_string(w[2]);
      if(!EmitFunction(w[3], args, f, rv, lbl))
        return false;
      body << lbl << ":;\n";
      return true;
    }
    case OpBranch:
      return EmitBranch(f, cur_label, w[1]);
    case OpBranchConditional:
    {
      Val cnd;
      if(!Get(f, w[1], cnd))
        return false;
      body << "  if(" << cnd.c[0] << ") {\n";
      if(!EmitBranch(f, cur_label, w[2]))
        return false;
      body << "  } else {\n";
      if(!EmitBranch(f, cur_label, w[3]))
        return false;
      body << "  }\n";
      return true;
    }
    case OpSwitch:
    {
      Val sel;
      if(!Get(f, w[1], sel))
        return false;
      for(uint16_t i = 3; i + 1 < in.wc; i += 2)
      {
        body << "  if(" << sel.c[0] << " == " << Hex(w[i]) << ") {\n";
        if(!EmitBranch(f, cur_label, w[i + 1]))
          return false;
        body << "  }\n";
      }
      return EmitBranch(f, cur_label, w[2]);
    }
    case OpReturn:
      body << "  goto " << f.ret_label << ";\n";
      return true;
    case OpReturnValue:
      if(!Get(f, w[1], a))
        return false;
      if(f.ret_val)
        for(size_t k = 0; k < a.c.size(); k++)
          body << "  " << f.ret_val->c[k] << " = " << a.c[k] << ";\n";
      body << "  goto " << f.ret_label << ";\n";
      return true;
    case OpKill:
      out.has_kill = true;
      body << "  return false;\n";
      return true;
    default:
      return Fail("unsupported SPIR-V opcode " + std::to_string(in.op));
  }
}

bool Translator::Run(const char *entry, std::string &err)
{
  entry_name = entry ? entry : "main";
  if(!Parse() || !SetupGlobals())
  {
    err = error;
    return false;
  }
  // derivative mode: any implicit-LOD sample or derivative op in a fragment program
  if(stage == ShaderStage::Fragment)
    for(auto &fn : functions)
      for(const Inst &in : fn.second.body)
        if(in.op == OpImageSampleImplicitLod || (in.op >= OpDPdx && in.op <= OpDPdyCoarse))
          deriv = true;
  out.needs_deriv = deriv;

  std::ostringstream pro, epi;
  Frame top;
  top.inst = inst_counter++;

  // interface variables
  uint32_t pos_var = 0;
  int pos_member = -1;
  struct OutLoc { uint32_t var; uint32_t loc; };
  std::vector<OutLoc> outs;
  for(uint32_t id : global_order)
  {
    uint32_t sc = globals[id].second;
    uint32_t pointee = T(globals[id].first).elem;
    auto loc = DecoParams(id, D_Location);
    auto bi = DecoParams(id, D_BuiltIn);
    const Ptr &p = global_ptrs[id];
    if(sc == SC_Input)
    {
      if(stage == ShaderStage::Vertex)
      {
        if(!bi.empty())
        {
          if(bi[0] == 42 || bi[0] == 5)  // VertexIndex / VertexId
            pro << "  " << StoreExpr(p, 0, LU, "vtx") << "\n";
          else if(bi[0] == 43 || bi[0] == 6)  // InstanceIndex / InstanceId: instancing is not applied
            pro << "  " << StoreExpr(p, 0, LU, "0u") << "\n";
          else
          {
            err = "unsupported vertex input builtin";
            return false;
          }
        }
        else if(!loc.empty())
        {
          // reference :1952-1986: GetVertexAttributeData, then the first N components
          uint32_t n = FlatCount(pointee);
          if(loc[0] >= 16)
          {
            err = "vertex input location beyond the 16 vertex attributes";
            return false;
          }
          if(std::find(out.vs_locations.begin(), out.vs_locations.end(), loc[0]) == out.vs_locations.end())
            out.vs_locations.push_back(loc[0]);
          std::string t = "va" + std::to_string(id);
          // raw[loc] holds the attribute's words, loaded before the program runs
          pro << "  { float " << t << "[4]; va_vdecode(d, " << loc[0] << "u, raw[" << loc[0] << "], " << t << ");\n";
          for(uint32_t k = 0; k < n && k < 4; k++)
          {
            Leaf l = LeafKind(pointee, k);
            std::string v = l == LF ? "F(" + t + "[" + std::to_string(k) + "])" : "va_f2u(" + t + "[" + std::to_string(k) + "])";
            pro << "  " << StoreExpr(p, k, l, v) << "\n";
          }
          pro << "  }\n";
        }
      }
      else
      {
        if(!bi.empty() || loc.empty())
        {
          err = "unsupported fragment input (only Location inputs are interpolated)";
          return false;
        }
        // reference :2159-2279: vectors/arrays interpolated with CreateDot, int scalars flat
        const Type &ty = T(pointee);
        uint32_t nslots = ty.op == OpTypeArray ? ty.count : 1;
        uint32_t et = ty.op == OpTypeArray ? ty.elem : pointee;
        uint32_t per = FlatCount(et);
        for(uint32_t a = 0; a < nslots; a++)
        {
          bool flat_int = LeafKind(et, 0) == LU;
          if(flat_int && per != 1)
          {
            err = "integer vector fragment inputs are not supported";
            return false;
          }
          uint32_t j0 = 0;
          for(auto &fi : out.fs_inputs)
            j0 += fi.ncomp;
          out.fs_inputs.push_back({loc[0] + a, per, flat_int});
          for(uint32_t k = 0; k < per; k++)
          {
            uint32_t j = j0 + k;
            std::string vj = "vary + " + std::to_string(3 * j);
            std::string v;
            if(flat_int)
              v = "va_f2u(vary[" + std::to_string(3 * j) + "])";
            else if(deriv)
              v = "F(va_interp(b[0], " + vj + "), va_interp(b[1], " + vj + "), va_interp(b[2], " + vj + "))";
            else
              v = "va_interp(b[0], " + vj + ")";
            pro << "  " << StoreExpr(p, a * per + k, flat_int ? LU : LF, v) << "\n";
          }
        }
      }
    }
    else if(sc == SC_Output)
    {
      if(!bi.empty())
      {
        if(bi[0] == 0)
          pos_var = id;
      }
      else if(!loc.empty())
        outs.push_back({id, loc[0]});
      else if(T(pointee).op == OpTypeStruct)
      {
        const Type &st = T(pointee);
        for(uint32_t m = 0; m < st.members.size(); m++)
        {
          auto mb = MemberDecoParams(pointee, m, D_BuiltIn);
          if(!mb.empty() && mb[0] == 0)
          {
            pos_var = id;
            pos_member = (int)m;
          }
        }
      }
    }
  }

  if(stage == ShaderStage::Vertex)
  {
    if(!pos_var)
    {
      err = "vertex program writes no Position";
      return false;
    }
    const Ptr &pp = global_ptrs[pos_var];
    uint32_t base = 0;
    if(pos_member >= 0)
    {
      const Type &st = T(pp.pointee);
      for(int m = 0; m < pos_member; m++)
        base += FlatCount(st.members[m]);
    }
    for(uint32_t k = 0; k < 4; k++)
      epi << "  pos[" << k << "] = cv(" << Load(pp, base + k, LF) << ");\n";
    // reference :2046-2083: vec<4 padded with component 0, scalars splatted, ints bitcast
    for(auto &o : outs)
    {
      const Ptr &op = global_ptrs[o.var];
      const Type &ty = T(op.pointee);
      uint32_t nslots = (ty.op == OpTypeArray || ty.op == OpTypeMatrix) ? ty.count : 1;
      uint32_t et = (ty.op == OpTypeArray || ty.op == OpTypeMatrix) ? ty.elem : op.pointee;
      uint32_t per = FlatCount(et);
      for(uint32_t a = 0; a < nslots; a++)
        for(uint32_t k = 0; k < 4; k++)
        {
          uint32_t src = a * per + (k < per ? k : 0);
          uint32_t dst = (o.loc + a) * 4 + k;
          if(dst >= 40)
            continue;
          Leaf l = LeafKind(et, src % per);
          std::string v = Load(op, src, l == LF ? LF : LU);
          epi << "  outs[" << dst << "] = " << (l == LF ? "cv(" + v + ")" : "va_u2f(" + v + ")") << ";\n";
        }
    }
  }
  else
  {
    bool found = false;
    for(auto &o : outs)
      if(o.loc == 0)
      {
        const Ptr &op = global_ptrs[o.var];
        uint32_t n = FlatCount(op.pointee);
        for(uint32_t k = 0; k < 4; k++)
          epi << "  out[" << k << "] = " << (k < n ? "cv(" + Load(op, k, LF) + ")" : std::string("0.0f")) << ";\n";
        found = true;
      }
    if(!found)
      epi << "  out[0] = 0.0f; out[1] = 0.0f; out[2] = 0.0f; out[3] = 0.0f;\n";
  }

  // the entry function body
  std::string exit_label = "va_exit";
  if(!EmitFunction(entry_fn, {}, top, nullptr, exit_label))
  {
    err = error;
    return false;
  }

  std::ostringstream src;
  if(stage == ShaderStage::Vertex)
  {
    src << "VA_INLINE void " << out.fn_name
        << "(const VaDraw &d, u32 vtx, const VaRaw *raw, float pos[4], float outs[40])\n{\n";
    src << "  typedef float F;\n";
    src << "#pragma unroll\n  for(int i_ = 0; i_ < 40; i_++) outs[i_] = 0.0f;\n";
  }
  else
  {
    src << "VA_INLINE bool " << out.fn_name << "(const VaDraw &d, const float *__restrict__ vary, const float b[3][4], float out[4])\n{\n";
    src << "  typedef " << (deriv ? "df" : "float") << " F;\n";
  }
  src << decl.str() << pro.str() << body.str() << exit_label << ":;\n" << epi.str();
  if(stage == ShaderStage::Fragment)
    src << "  return true;\n";
  src << "}\n";
  out.code = src.str();
  out.stage = stage;
  return true;
}
}  // namespace

bool SpirvHeaderValid(const uint32_t *words, size_t nwords)
{
  return nwords >= 5 && words[0] == 0x07230203u && words[1] <= 0x00010600u && words[3] > 0;
}

bool TranslateSpirv(const uint32_t *words, size_t nwords, const char *entry, ShaderStage stage,
                    const std::string &fn_name, TranslatedShader &out, std::string &err)
{
  out = TranslatedShader();
  out.fn_name = fn_name;
  Translator t(words, nwords, stage, out);
  return t.Run(entry, err);
}
}  // namespace visor
