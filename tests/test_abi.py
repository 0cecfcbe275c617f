"""The C ABI: every function include/g3a.h declares is exported by the built
library, and the argument checks that need no GPU behave as documented."""
from __future__ import annotations

import ctypes as C
import re
import subprocess

import g3sa_amd
from conftest import ROOT

HEADER = ROOT / "include" / "g3a.h"


def declared_functions() -> list[str]:
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    return sorted(set(re.findall(r"\b(g3a_\w+)\s*\(", text)))


def test_header_declares_the_boundary():
    names = declared_functions()
    for f in ["g3a_device_alloc", "g3a_memcpy_index", "g3a_memcpy_input", "g3a_bwa_align", "g3a_memcpy_output",
              "g3a_destroy", "g3a_align_batch"]:
        assert f in names, f


def test_library_exports_every_declared_function(built):
    lib = g3sa_amd.load_library()
    names = declared_functions()
    assert names
    for f in names:
        assert hasattr(lib, f), f"{f} declared in g3a.h but not exported"
    assert set(g3sa_amd.EXPORTS) <= set(names)
    # C linkage: the dynamic symbol table holds the plain names
    out = subprocess.run(["nm", "-D", "--defined-only", str(g3sa_amd.LIB_PATH)], capture_output=True, text=True,
                         check=True).stdout
    syms = {ln.split()[-1] for ln in out.splitlines() if ln.strip()}
    assert set(names) <= syms


def test_argument_errors_without_gpu(built, tmp_path):
    lib = g3sa_amd.load_library()
    ctx = C.c_void_p()
    # null prefix / bad sizes are rejected before any device call
    assert lib.g3a_device_alloc(0, None, None, 10, 100, C.byref(ctx)) == -1
    assert lib.g3a_device_alloc(0, b"x", None, 0, 100, C.byref(ctx)) == -1
    assert lib.g3a_memcpy_input(None, 1, None, None) == -1
    assert lib.g3a_bwa_align(None) == -1
    assert lib.g3a_set_form(None, 0) == -1
    assert lib.g3a_set_tuning(None, 16, 8, 8) == -1
    assert lib.g3a_last_error(None) == b"no context"
    lib.g3a_destroy(None)
    opt = g3sa_amd.default_opt()
    assert (opt.a, opt.b, opt.o_del, opt.e_del, opt.w, opt.T, opt.min_seed_len) == (1, 4, 6, 1, 100, 30, 19)
    assert list(opt.mat)[:5] == [1, -4, -4, -4, -1]
