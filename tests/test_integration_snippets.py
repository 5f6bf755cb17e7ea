"""The reference-side binding shown in the repository's integration notes compiles and links against
include/mpflow_amd.h and libmpflow_amd.so (g++, C++17), and the ctypes argument types
of mpflow_amd/capi.py agree with the header's declarations parameter by parameter."""
import ctypes as C
import os
import re
import shutil
import subprocess

import pytest

from mpflow_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _cpp_blocks():
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    return re.findall(r"```cpp\n(.*?)```", text, flags=re.S)


@pytest.mark.skipif(shutil.which("g++") is None, reason="no g++")
def test_integration_snippets_compile_and_link(tmp_path):
    blocks = _cpp_blocks()
    assert len(blocks) >= 10
    src = tmp_path / "binding.cpp"
    # every inline function is referenced so that the linker resolves each library call
    names = sorted(set(re.findall(r"inline \w[\w\s:*]*?\b(\w+)\(", "\n".join(blocks))))
    uses = ",\n".join("        reinterpret_cast<fp>(&mpflow_binding::%s)" % n for n in names)
    src.write_text("\n".join(blocks) + "\nusing fp = void (*)();\nint main() {\n    fp volatile table[] = {\n"
                   + uses + "};\n    return table[0] == nullptr;\n}\n")
    lib = capi.LIB_PATH
    if not os.path.exists(lib):
        pytest.skip("libmpflow_amd.so not built")
    exe = tmp_path / "binding"
    cmd = ["g++", "-std=c++17", "-Wall", "-Werror", "-Wno-unused-function", "-I", os.path.join(ROOT, "include"),
           str(src), "-o", str(exe), "-L", os.path.dirname(lib), "-lmpflow_amd",
           "-Wl,-rpath-link,/opt/rocm/lib", "-Wl,--unresolved-symbols=report-all"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    # the executable references the library's entry points
    nm = subprocess.run(["nm", "-u", str(exe)], capture_output=True, text=True).stdout
    for f in ("mpf_eit_solve", "mpf_gn_step", "mpf_sparse_multiply", "mpf_pcg_solve", "mpf_normal_solve_noalloc"):
        assert f in nm


def _header_params():
    src = open(os.path.join(ROOT, "include", "mpflow_amd.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    out = {}
    for m in re.finditer(r"\b(?:int|int64_t|void\*|const char\*|void)\s+(mpf_[a-z0-9_]+)\s*\((.*?)\);", src,
                         flags=re.S):
        args = [a.strip() for a in m.group(2).replace("\n", " ").split(",")]
        out[m.group(1)] = [] if args == ["void"] else args
    return out


def _kind(decl):
    """'p' pointer, 'd' floating (double/float), 'i' integer: the C parameter's class."""
    d = re.sub(r"\s+", " ", decl)
    if "*" in d or "[" in d:
        return "p"
    if re.search(r"\b(double|float)\b", d):
        return "d"
    return "i"


def _ctype_kind(t):
    if t in (C.c_void_p, C.c_char_p) or (isinstance(t, type) and issubclass(t, C._Pointer)):
        return "p"
    if t in (C.c_double, C.c_float):
        return "d"
    return "i"


def test_ctypes_argtypes_match_header():
    L = capi.lib()
    params = _header_params()
    assert set(params) == set(capi.EXPORTED)
    checked = 0
    for name, decls in params.items():
        f = getattr(L, name)
        if f.argtypes is None:
            assert not decls, "%s: no argtypes but %d parameters" % (name, len(decls))
            continue
        assert len(f.argtypes) == len(decls), "%s: %d argtypes, %d parameters" % (name, len(f.argtypes), len(decls))
        for t, d in zip(f.argtypes, decls):
            assert _ctype_kind(t) == _kind(d), "%s: %s bound as %s" % (name, d, t)
        checked += 1
    assert checked >= 40
