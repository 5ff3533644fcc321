"""The C-ABI library exports every function that include/*.h declares.

CPU-only: the library is loaded with ctypes and each extern "C" symbol is
looked up (nothing is called).  The __global__ kernels that the headers
declare for HIP hosts (lu_gpu.h, lu_gpu_par.h) have C++ linkage; they are
checked by demangled name in the dynamic symbol table.
"""
import ctypes
import glob
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "luspice_amd", "lib", "libluspice_amd.so")
HEADERS = sorted(glob.glob(os.path.join(ROOT, "include", "*.h")))


def _strip(src):
    src = re.sub(r"/\*.*?\*/", " ", src, flags=re.S)
    src = re.sub(r"//[^\n]*", " ", src)
    return src


def c_functions(path):
    """Names of the extern "C" functions a header declares."""
    s = _strip(open(path).read())
    s = re.sub(r"#ifdef __HIPCC__.*?#endif", " ", s, flags=re.S)
    s = re.sub(r"#ifdef __cplusplus.*?#endif", " ", s, flags=re.S)
    s = "\n".join(l for l in s.splitlines() if not l.strip().startswith("#"))
    while True:                                  # drop struct bodies
        t = re.sub(r"\{[^{}]*\}", " ", s)
        if t == s:
            break
        s = t
    names = []
    for st in s.split(";"):
        st = " ".join(st.split())
        if "(" not in st or st.startswith("typedef"):
            continue
        m = re.match(r"^[\w\s\*]*?\b(\w+)\s*\(", st)
        if m:
            names.append(m.group(1))
    return names


def kernels(path):
    """Names of the __global__ kernels a header declares for HIP hosts."""
    s = _strip(open(path).read())
    return re.findall(r"__global__\s+void\s+(\w+)\s*\(", s)


def test_headers_parse():
    per = {os.path.basename(h): c_functions(h) for h in HEADERS}
    # every header of the drop-in surface declares something
    for h, names in per.items():
        assert names, h
    assert "lsp_batch_refactor" in per["luspice.h"]
    assert "dsh_numeric" in per["lu_static.h"]
    assert "spmv_launch" in per["spmv.h"]


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
@pytest.mark.parametrize("header", [os.path.basename(h) for h in HEADERS])
def test_library_exports_header(header):
    lib = ctypes.CDLL(LIB)
    missing = [n for n in c_functions(os.path.join(ROOT, "include", header))
               if not hasattr(lib, n)]
    assert not missing, f"{header}: not exported: {missing}"


@pytest.mark.skipif(not os.path.exists(LIB) or shutil.which("nm") is None,
                    reason="library not built or nm missing")
def test_library_exports_kernels():
    out = subprocess.run(["nm", "-DC", "--defined-only", LIB], capture_output=True,
                         text=True, check=True).stdout
    exported = set()
    for line in out.splitlines():
        parts = line.split(None, 2)
        if len(parts) == 3:
            exported.add(parts[2].split("(")[0])
    want = [k for h in HEADERS for k in kernels(h)]
    assert want, "no kernels declared"
    missing = [k for k in want if k not in exported]
    assert not missing, f"kernels not exported: {missing}"
