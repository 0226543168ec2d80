"""A small SPIR-V 1.0 module builder.

The container has no glslang or SPIR-V tools, so the shaders the driver is exercised with are
written directly as SPIR-V through this builder. It emits the same instruction shapes a GLSL
front end does for straight-line code: `OpVariable` for every interface variable, `OpLoad` /
`OpAccessChain` for reads, and GLSL.std.450 extended instructions for the library calls.
"""

import struct

MAGIC = 0x07230203

# opcodes (SPIR-V 1.0 specification, section 3.32)
OP = dict(
    Nop=0, Source=3, Name=5, MemberName=6, ExtInstImport=11, ExtInst=12, MemoryModel=14,
    EntryPoint=15, ExecutionMode=16, Capability=17, TypeVoid=19, TypeBool=20, TypeInt=21,
    TypeFloat=22, TypeVector=23, TypeMatrix=24, TypeImage=25, TypeSampler=26,
    TypeSampledImage=27, TypeArray=28, TypeStruct=30, TypePointer=32, TypeFunction=33,
    ConstantTrue=41, ConstantFalse=42, Constant=43, ConstantComposite=44, Function=54,
    FunctionParameter=55, FunctionEnd=56, FunctionCall=57, Variable=59, Load=61, Store=62,
    AccessChain=65, Decorate=71, MemberDecorate=72, VectorShuffle=79, CompositeConstruct=80,
    CompositeExtract=81, Transpose=84, ImageSampleImplicitLod=87, ImageSampleExplicitLod=88,
    ConvertFToU=109, ConvertFToS=110, ConvertSToF=111, ConvertUToF=112, Bitcast=124,
    SNegate=126, FNegate=127, IAdd=128, FAdd=129, ISub=130, FSub=131, IMul=132, FMul=133,
    UDiv=134, SDiv=135, FDiv=136, UMod=137, SRem=138, SMod=139, FRem=140, FMod=141,
    VectorTimesScalar=142, MatrixTimesScalar=143, VectorTimesMatrix=144,
    MatrixTimesVector=145, MatrixTimesMatrix=146, Dot=148, LogicalOr=166, LogicalAnd=167,
    LogicalNot=168, Select=169, IEqual=170, INotEqual=171, UGreaterThan=172,
    SGreaterThan=173, UGreaterThanEqual=174, SGreaterThanEqual=175, ULessThan=176,
    SLessThan=177, ULessThanEqual=178, SLessThanEqual=179, FOrdEqual=180, FOrdNotEqual=182,
    FOrdLessThan=184, FOrdGreaterThan=186, FOrdLessThanEqual=188, FOrdGreaterThanEqual=190,
    ShiftRightLogical=194, ShiftRightArithmetic=195, ShiftLeftLogical=196, BitwiseOr=197,
    BitwiseXor=198, BitwiseAnd=199, Not=200, DPdx=207, DPdy=208, Fwidth=209, Phi=245,
    LoopMerge=246, SelectionMerge=247, Label=248, Branch=249, BranchConditional=250,
    Kill=252, Return=253, ReturnValue=254,
)

# GLSL.std.450 extended instruction numbers
GLSL = dict(
    Round=1, RoundEven=2, Trunc=3, FAbs=4, SAbs=5, FSign=6, Floor=8, Ceil=9, Fract=10,
    Sin=13, Cos=14, Pow=26, Exp2=29, Log2=30, Sqrt=31, InverseSqrt=32, MatrixInverse=34,
    FMin=37, UMin=38, SMin=39, FMax=40, UMax=41, SMax=42, FClamp=43, UClamp=44, SClamp=45,
    FMix=46, Step=48, SmoothStep=49, Length=66, Distance=67, Cross=68, Normalize=69,
    Reflect=71,
)

DEC = dict(Block=2, RowMajor=4, ColMajor=5, ArrayStride=6, MatrixStride=7, BuiltIn=11,
           Flat=14, Location=30, Binding=33, DescriptorSet=34, Offset=35)
BUILTIN = dict(Position=0, PointSize=1, VertexIndex=42, InstanceIndex=43)
SC = dict(UniformConstant=0, Input=1, Uniform=2, Output=3, Private=6, Function=7, PushConstant=9)
EXEC_MODEL = dict(Vertex=0, Fragment=4)


def _str_words(s):
    b = s.encode() + b"\0"
    b += b"\0" * ((4 - len(b) % 4) % 4)
    return list(struct.unpack("<%dI" % (len(b) // 4), b))


class Module:
    """Accumulates the sections of one SPIR-V module in the order the spec requires."""

    def __init__(self):
        self.bound = 1
        self.caps, self.extsec, self.entry, self.modes = [], [], [], []
        self.names, self.decos, self.types, self.funcs = [], [], [], []
        self._type_cache, self._const_cache = {}, {}
        self.glsl = self._id()
        self.extsec += self._inst("ExtInstImport", [self.glsl] + _str_words("GLSL.std.450"))
        self.caps += self._inst("Capability", [1])  # Shader
        self.body = None

    # -- helpers -----------------------------------------------------------------------------
    def _id(self):
        i = self.bound
        self.bound += 1
        return i

    @staticmethod
    def _inst(op, operands):
        return [((len(operands) + 1) << 16) | OP[op]] + list(operands)

    def _type(self, key, op, operands):
        if key not in self._type_cache:
            i = self._id()
            self.types += self._inst(op, [i] + operands)
            self._type_cache[key] = i
        return self._type_cache[key]

    # -- types -------------------------------------------------------------------------------
    def void(self): return self._type(("void",), "TypeVoid", [])
    def bool(self): return self._type(("bool",), "TypeBool", [])
    def float(self): return self._type(("float",), "TypeFloat", [32])
    def int(self, signed=True): return self._type(("int", signed), "TypeInt", [32, 1 if signed else 0])
    def vec(self, n, comp=None):
        comp = comp or self.float()
        return self._type(("vec", n, comp), "TypeVector", [comp, n])
    def mat(self, cols, rows):
        col = self.vec(rows)
        return self._type(("mat", cols, rows), "TypeMatrix", [col, cols])
    def array(self, elem, n):
        return self._type(("array", elem, n), "TypeArray", [elem, self.const_u(n)])
    def struct(self, members, key):
        return self._type(("struct", key), "TypeStruct", list(members))
    def ptr(self, sc, t): return self._type(("ptr", sc, t), "TypePointer", [SC[sc], t])
    def func_type(self, ret, params=()):
        return self._type(("fn", ret, tuple(params)), "TypeFunction", [ret] + list(params))
    def image2d(self, arrayed=False, cube=False):
        """Float image type, Dim 2D (1) or Cube (3)."""
        return self._type(("img", arrayed, cube), "TypeImage",
                          [self.float(), 3 if cube else 1, 0, 1 if arrayed else 0, 0, 1, 0])
    def sampled_image(self, arrayed=False, cube=False):
        img = self.image2d(arrayed, cube)
        return self._type(("simg", arrayed, cube), "TypeSampledImage", [img])

    # -- constants ---------------------------------------------------------------------------
    def const_f(self, v):
        key = ("f", struct.pack("<f", v))
        if key not in self._const_cache:
            i = self._id()
            self.types += self._inst("Constant", [self.float(), i, struct.unpack("<I", key[1])[0]])
            self._const_cache[key] = i
        return self._const_cache[key]

    def const_i(self, v, signed=True):
        key = ("i", signed, v)
        if key not in self._const_cache:
            i = self._id()
            self.types += self._inst("Constant", [self.int(signed), i, v & 0xFFFFFFFF])
            self._const_cache[key] = i
        return self._const_cache[key]

    def const_u(self, v): return self.const_i(v, signed=False)

    def const_vec(self, vals):
        comps = [self.const_f(v) for v in vals]
        key = ("cv", tuple(comps))
        if key not in self._const_cache:
            i = self._id()
            self.types += self._inst("ConstantComposite", [self.vec(len(vals)), i] + comps)
            self._const_cache[key] = i
        return self._const_cache[key]

    # -- decorations / interface ----------------------------------------------------------------
    def decorate(self, target, dec, *args):
        self.decos += self._inst("Decorate", [target, DEC[dec]] + list(args))

    def member_decorate(self, st, member, dec, *args):
        self.decos += self._inst("MemberDecorate", [st, member, DEC[dec]] + list(args))

    def variable(self, sc, t):
        i = self._id()
        self.types += self._inst("Variable", [self.ptr(sc, t), i, SC[sc]])
        return i

    def input(self, t, location=None, builtin=None):
        v = self.variable("Input", t)
        if location is not None:
            self.decorate(v, "Location", location)
        if builtin is not None:
            self.decorate(v, "BuiltIn", BUILTIN[builtin])
        return v

    def output(self, t, location=None, builtin=None):
        v = self.variable("Output", t)
        if location is not None:
            self.decorate(v, "Location", location)
        if builtin is not None:
            self.decorate(v, "BuiltIn", BUILTIN[builtin])
        return v

    def push_block(self, members, offsets, key, matrix_stride=16):
        """A push-constant block; members are type ids, offsets byte offsets (std430)."""
        st = self.struct(members, key)
        self.decorate(st, "Block")
        for m, (t, off) in enumerate(zip(members, offsets)):
            self.member_decorate(st, m, "Offset", off)
            if ("mat" in str(self._type_key(t))):
                self.member_decorate(st, m, "ColMajor")
                self.member_decorate(st, m, "MatrixStride", matrix_stride)
        return st, self.variable("PushConstant", st)

    def uniform_block(self, members, offsets, key, set_, binding, matrix_stride=16):
        """A uniform buffer block at (set_, binding); members are type ids, offsets byte offsets."""
        st = self.struct(members, key)
        self.decorate(st, "Block")
        for m, (t, off) in enumerate(zip(members, offsets)):
            self.member_decorate(st, m, "Offset", off)
            if ("mat" in str(self._type_key(t))):
                self.member_decorate(st, m, "ColMajor")
                self.member_decorate(st, m, "MatrixStride", matrix_stride)
        v = self.variable("Uniform", st)
        self.decorate(v, "DescriptorSet", set_)
        self.decorate(v, "Binding", binding)
        return st, v

    def _type_key(self, t):
        for k, v in self._type_cache.items():
            if v == t:
                return k
        return None

    def sampler(self, set_, binding, arrayed=False, cube=False):
        v = self.variable("UniformConstant", self.sampled_image(arrayed, cube))
        self.decorate(v, "DescriptorSet", set_)
        self.decorate(v, "Binding", binding)
        return v

    # -- function body -------------------------------------------------------------------------
    def begin_main(self):
        self.main = self._id()
        self.body = []
        self.body += self._inst("Function", [self.void(), self.main, 0, self.func_type(self.void())])
        self.body += self._inst("Label", [self._id()])
        return self.main

    def op(self, name, result_type, *operands):
        """Emit an instruction with a result id and return that id."""
        i = self._id()
        self.body += self._inst(name, [result_type, i] + list(operands))
        return i

    def op_void(self, name, *operands):
        self.body += self._inst(name, list(operands))

    def ext(self, result_type, inst, *args):
        return self.op("ExtInst", result_type, self.glsl, GLSL[inst], *args)

    def load(self, t, ptr): return self.op("Load", t, ptr)
    def store(self, ptr, val): self.op_void("Store", ptr, val)

    def access(self, sc, t, base, *indices):
        return self.op("AccessChain", self.ptr(sc, t), base, *indices)

    def end_main(self, model, name, interface, frag=False):
        self.body += self._inst("Return", [])
        self.body += self._inst("FunctionEnd", [])
        self.entry += self._inst("EntryPoint", [EXEC_MODEL[model], self.main] + _str_words(name) + list(interface))
        if frag:
            self.modes += self._inst("ExecutionMode", [self.main, 7])  # OriginUpperLeft
        self.funcs += self.body

    def words(self):
        mm = self._inst("MemoryModel", [0, 1])
        body = self.caps + self.extsec + mm + self.entry + self.modes + self.names + self.decos + self.types + self.funcs
        return [MAGIC, 0x00010000, 0, self.bound, 0] + body

    def to_bytes(self):
        w = self.words()
        return struct.pack("<%dI" % len(w), *w)
