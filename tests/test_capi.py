"""CPU tests of the C ABI: the library loads, exports every entry point declared in
include/mpflow_amd.h, and its host-side error paths return status codes (the
reference throws std::invalid_argument / runtime_error for the same conditions)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from mpflow_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "mpflow_amd.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(mpf_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    L = capi.lib()
    names = _declared()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert set(capi.EXPORTED) == set(names)


def test_version_and_no_error():
    L = capi.lib()
    assert L.mpf_version().decode().startswith("mpflow_amd")


def test_element_integrals_rejects_bad_order():
    L = capi.lib()
    xy = np.zeros(12)
    K = np.zeros(36)
    rc = L.mpf_element_integrals(4, xy.ctypes.data, K.ctypes.data, K.ctypes.data)
    assert rc == capi.MPF_EINVAL
    assert "k must be 3 or 6" in L.mpf_last_error().decode()


def test_null_arguments_are_einval():
    L = capi.lib()
    w = C.c_int32()
    rc = L.mpf_ports_circular(0, None, 0, 3, None, 0, 2, None, 16, 0.1, 0.0, 0, None, 0, C.byref(w))
    assert rc == capi.MPF_EINVAL
    assert L.mpf_model_create(None, None, None) == capi.MPF_EINVAL
    assert L.mpf_eit_solve(None, None, None, None, None, None, None) == capi.MPF_EINVAL
    assert L.mpf_gn_step(None, None, None, 1e-3, 0, None, None, None) == capi.MPF_EINVAL


def test_ports_query_then_fill():
    from mpflow_amd.eit import circular_ports
    from mpflow_amd.mesh import boundary_edges, ring_mesh
    n, e = ring_mesh(8, 8, 0.2)
    b = boundary_edges(e)
    p = circular_ports(n, e, b, 8, 0.2)
    assert p.shape[0] == 8 and p.dtype == np.int32
    assert np.all(p[p >= 0] < b.shape[0])


def test_context_without_device_fails_cleanly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a device is present")
    L = capi.lib()
    h = C.c_void_p()
    rc = L.mpf_ctx_create(0, None, C.byref(h))
    assert rc in (capi.MPF_EHIP, capi.MPF_EINVAL)
    assert L.mpf_last_error().decode() != ""


def test_struct_layouts_match_header(tmp_path):
    """The ctypes structures of mpflow_amd/capi.py have the header's field offsets and
    sizes: a C program compiled against include/mpflow_amd.h prints offsetof / sizeof."""
    import ctypes
    import os
    import shutil
    import subprocess
    from mpflow_amd import capi
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    inc = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include")
    structs = {"mpf_model_desc": capi.ModelDesc, "mpf_solve_opts": capi.SolveOpts,
               "mpf_solve_stats": capi.SolveStats}
    lines = ["#include <stddef.h>", "#include <stdio.h>", '#include "mpflow_amd.h"', "int main(void) {"]
    for cname, py in structs.items():
        lines.append('printf("%s %%zu\\n", sizeof(%s));' % (cname, cname))
        for f, _ in py._fields_:
            lines.append('printf("%s.%s %%zu\\n", offsetof(%s, %s));' % (cname, f, cname, f))
    lines += ["return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run([cc, "-I", inc, str(src), "-o", str(exe)], check=True, capture_output=True)
    out = dict(l.split() for l in subprocess.run([str(exe)], check=True, capture_output=True,
                                                  text=True).stdout.splitlines())
    for cname, py in structs.items():
        assert int(out[cname]) == ctypes.sizeof(py), cname
        for f, _ in py._fields_:
            assert int(out["%s.%s" % (cname, f)]) == getattr(py, f).offset, (cname, f)
