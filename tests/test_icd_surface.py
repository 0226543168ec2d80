"""The drop-in boundary without a GPU: the library loads, exports every function that
include/*.h declares, resolves the reference's entry-point names (icd_interface.cpp:7-169), and
its ctypes mirror agrees with the C headers on every structure layout."""
import ctypes as C
import json
import os
import re
import subprocess

import pytest

from visor_amd import vk

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
NAMES = json.load(open(os.path.join(ROOT, "tests", "golden", "icd_entry_points.json")))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(vk.LIB_PATH):
        pytest.fail("libvisor_mi355x.so is not built (python -c 'import __graft_entry__ as g; g.build()')")
    L = C.CDLL(vk.LIB_PATH)
    L.vk_icdGetInstanceProcAddr.restype = C.c_void_p
    L.vk_icdGetInstanceProcAddr.argtypes = [C.c_void_p, C.c_char_p]
    L.vk_icdGetPhysicalDeviceProcAddr.restype = C.c_void_p
    L.vk_icdGetPhysicalDeviceProcAddr.argtypes = [C.c_void_p, C.c_char_p]
    L.vkGetDeviceProcAddr.restype = C.c_void_p
    L.vkGetDeviceProcAddr.argtypes = [C.c_void_p, C.c_char_p]
    return L


def _declared_functions():
    """Function names declared in include/*.h (prototypes at file scope)."""
    names = []
    for h in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if not h.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"//[^\n]*", "", text)
        for m in re.finditer(r"^[A-Za-z_][\w \*]*?\b(\w+)\s*\([^;{]*\)\s*;", text, flags=re.M):
            if m.group(0).lstrip().startswith(("typedef", "#")):
                continue
            names.append(m.group(1))
    return names


def test_every_declared_function_is_exported(lib):
    names = _declared_functions()
    assert "vk_icdGetInstanceProcAddr" in names and "vh_get_stats" in names and len(names) == 29
    for n in names:
        assert hasattr(lib, n), n


def test_instance_table_resolves_every_reference_name(lib):
    for n in NAMES["instance"]:
        assert lib.vk_icdGetInstanceProcAddr(None, n.encode()), n


def test_physical_device_and_device_tables(lib):
    for n in NAMES["physical_device"]:
        assert lib.vk_icdGetPhysicalDeviceProcAddr(None, n.encode()), n
    for n in NAMES["device"]:
        assert lib.vkGetDeviceProcAddr(None, n.encode()), n


def test_unknown_and_excluded_names_resolve_to_null(lib):
    # the reference returns NULL for these two and a string pointer for every other unknown name;
    # here every unknown name is NULL, which the loader treats as "not supported"
    for n in NAMES["null_names"] + ["vkCmdDispatchIndirectNV", "vkNotAFunction", ""]:
        assert not lib.vk_icdGetInstanceProcAddr(None, n.encode()), n
    assert not lib.vk_icdGetPhysicalDeviceProcAddr(None, b"vkCreateDevice")
    assert not lib.vkGetDeviceProcAddr(None, b"vkCreateInstance")


def test_direct_exports_match_table_entries(lib):
    for n in ("vkCmdDraw", "vkCmdDrawIndexed", "vkQueueSubmit", "vkGetDeviceProcAddr"):
        direct = C.cast(getattr(lib, n), C.c_void_p).value
        assert lib.vk_icdGetInstanceProcAddr(None, n.encode()) == direct, n


def test_negotiate_interface_version(lib):
    f = lib.vk_icdNegotiateLoaderICDInterfaceVersion
    f.restype = C.c_int32
    for req, expect in [(1, 1), (2, 2), (4, 4), (5, 4), (99, 4)]:
        v = C.c_uint32(req)
        assert f(C.byref(v)) == 0 and v.value == expect
    f.argtypes = [C.c_void_p]
    assert f(None) == -3  # VK_ERROR_INITIALIZATION_FAILED


def test_enumerate_instance_version(lib):
    f = C.CFUNCTYPE(C.c_int32, C.POINTER(C.c_uint32))(lib.vk_icdGetInstanceProcAddr(None, b"vkEnumerateInstanceVersion"))
    v = C.c_uint32(0)
    assert f(C.byref(v)) == 0 and v.value == (1 << 22)


def test_ctypes_layouts_match_headers(tmp_path):
    """Compiles a probe against include/visor_icd.h with gcc and compares sizeof/offsetof of every
    structure the ctypes mirror defines."""
    structs = {}
    for name in dir(vk):
        obj = getattr(vk, name)
        if isinstance(obj, type) and issubclass(obj, C.Structure) and obj not in (vk.S,) and obj._fields_:
            cname = name if name.startswith("Vh") else "Vk" + name
            structs[cname] = obj
    assert len(structs) > 40
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "visor_icd.h"', "int main(void){"]
    for cname, cls in sorted(structs.items()):
        lines.append('printf("%s %%zu\\n", sizeof(%s));' % (cname, cname))
        for fname, _ in cls._fields_:
            lines.append('printf("%s.%s %%zu\\n", offsetof(%s, %s));' % (cname, fname, cname, fname))
    lines.append("return 0;}")
    src = tmp_path / "probe.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = dict(l.split() for l in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n") if l)
    for cname, cls in structs.items():
        assert int(got[cname]) == C.sizeof(cls), cname
        for fname, _ in cls._fields_:
            assert int(got[cname + "." + fname]) == getattr(cls, fname).offset, (cname, fname)


def test_no_device_reports_unavailable(lib):
    """Without a usable HIP device the library says so instead of falling back to the CPU."""
    lib.vh_device_available.restype = C.c_int
    avail = lib.vh_device_available()
    import torch
    if not torch.cuda.is_available():
        assert avail == 0
