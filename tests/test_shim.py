"""The reference-side shim (g3sa_amd/shim/g3a_shim.cpp) compiles against the
reference's own headers (bwa/src/cuda_wrapper.h, bwamem.h, pipeline.h,
bwa.h, macro.h) and this library's include/g3a.h.  Skipped where the
reference sources are absent (they are not shipped with this repository)."""
from __future__ import annotations

import shutil
import subprocess
from pathlib import Path

import pytest

from conftest import ROOT

REF_SRC = Path("/root/reference/bwa/src")


@pytest.mark.skipif(not (REF_SRC / "cuda_wrapper.h").exists() or shutil.which("g++") is None,
                    reason="the reference's headers are not present")
def test_shim_compiles_against_reference_headers():
    r = subprocess.run(["g++", "-fsyntax-only", "-std=c++17", "-Wall", "-I", str(ROOT / "include"), "-I",
                        str(REF_SRC), str(ROOT / "g3sa_amd" / "shim" / "g3a_shim.cpp")],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-3000:]


def test_shim_binds_only_declared_entry_points():
    """Every g3a_* call in the shim is declared in include/g3a.h."""
    import re
    shim = (ROOT / "g3sa_amd" / "shim" / "g3a_shim.cpp").read_text()
    header = (ROOT / "include" / "g3a.h").read_text()
    called = set(re.findall(r"\b(g3a_[a-z_]+)\s*\(", shim)) - {"g3a_shim_set_index_prefix", "g3a_shim_pestat"}
    declared = set(re.findall(r"\b(g3a_[a-z_]+)\s*\(", header))
    assert called and called <= declared, called - declared
