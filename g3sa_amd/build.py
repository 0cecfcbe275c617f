"""Build the native parts in-tree (no GPU needed: hipcc cross-compiles gfx950).

Outputs, all under g3sa_amd/_build/:
  libg3a.so   HIP kernels + C ABI (include/g3a.h) + index builder/loader
  g3a         command-line driver (`g3a index`, `g3a mem`)
  g3a-gen     synthetic reference/read generator
  g3a-gather  random-gather ceiling of the GPU (for bench.py's roofline)
The CPU oracle (test infrastructure) is built by oracle/Makefile into
oracle/_build/.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
SRC = PKG / "csrc"
OUT = PKG / "_build"
ORACLE = ROOT / "oracle"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]
CXX_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-Wno-unused-result"]

LIB_HIP = ["g3a_kernels.hip", "g3a_learned.hip", "g3a_index_gpu.hip", "g3a_text.hip", "g3a_pair.hip", "g3a_api.cpp"]
LIB_CXX = ["index_build.cpp", "index_io.cpp"]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {' '.join(cmd)}")


def build(verbose: bool = False) -> Path:
    """Compile every native component for gfx950; returns the output dir."""
    OUT.mkdir(exist_ok=True)
    jobs = []
    for f in LIB_HIP:
        obj = OUT / (Path(f).stem + ".o")
        extra = ["-x", "hip"] if f.endswith(".cpp") else []
        jobs.append([HIPCC, *HIP_FLAGS, *extra, "-I", str(ROOT / "include"), "-c", str(SRC / f), "-o", str(obj)])
    for f in LIB_CXX:
        obj = OUT / (Path(f).stem + ".o")
        jobs.append(["g++", *CXX_FLAGS, "-c", str(SRC / f), "-o", str(obj)])
    for f in ["g3a_main.cpp", "g3a_mem.cpp"]:
        obj = OUT / (Path(f).stem + ".o")
        jobs.append(["g++", *CXX_FLAGS, "-I", "/opt/rocm/include", "-c", str(SRC / f), "-o", str(obj)])
    jobs.append(["g++", "-O3", "-std=c++17", "-o", str(OUT / "g3a-gen"), str(SRC / "gen.cpp")])
    # the random-gather ceiling that bench.py reports beside the seeding roofline
    jobs.append([HIPCC, *HIP_FLAGS, "-x", "hip", "-o", str(OUT / "g3a-gather"), str(SRC / "g3a_gather.cpp")])
    with cf.ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        for fut in [ex.submit(_run, j) for j in jobs]:
            fut.result()
    objs = [str(OUT / (Path(f).stem + ".o")) for f in LIB_HIP + LIB_CXX]
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(OUT / "libg3a.so"), *objs])
    _run(["g++", "-o", str(OUT / "g3a"), str(OUT / "g3a_main.o"), str(OUT / "g3a_mem.o"),
          f"-L{OUT}", "-lg3a", "-Wl,-rpath,$ORIGIN", "-lpthread"])
    # the oracle is test infrastructure, built beside its sources
    _run(["make", "-s", "-C", str(ORACLE)])
    if verbose:
        print(f"built into {OUT}")
    return OUT


def clean() -> None:
    shutil.rmtree(OUT, ignore_errors=True)
    subprocess.run(["make", "-s", "-C", str(ORACLE), "clean"], check=False)


if __name__ == "__main__":
    build(verbose=True)
