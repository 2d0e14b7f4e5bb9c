"""Builds libhsolver_amd.so in-tree for gfx950 with hipcc (no JIT, no cache outside the repo).

Flags: -ffp-contract=off keeps every point update in its written operation order (the
oracle is built the same way, so sweeps compare bit for bit); packed FP32 instructions
are disabled (-packed-fp32-ops).
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
SRC = PKG / "csrc"
OBJ = PKG / "_obj"
LIB = PKG / "libhsolver_amd.so"
EXE = PKG / "hsolver_main"
SHIM_CHECK = PKG / "hsolver_shim_check"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("HS_OFFLOAD_ARCH", "gfx950")

COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", f"--offload-arch={ARCH}",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-Wall", "-Wno-unused-function", "-Wno-unused-variable", f"-I{PKG.parent / 'include'}"]

SOURCES = ["hs_kernels.hip", "hs_mg.hip", "hs_dense.hip", "hs_energy.hip", "hs_amg.hip", "hs_static.hip", "hs_api.cpp",
           "hs_mesh.cpp"]


def _compile(src: Path) -> Path:
    out = OBJ / (src.name + ".o")
    lang = ["-x", "hip"] if src.suffix == ".hip" else []
    cmd = [HIPCC, *COMMON, *lang, "-c", str(src), "-o", str(out)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return out


def build(verbose: bool = False) -> Path:
    OBJ.mkdir(exist_ok=True)
    srcs = [SRC / s for s in SOURCES if (SRC / s).exists()]
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    # host driver (main.cpp shape) and the shim check, linked against the library, found
    # next to it at run time
    for src, exe in ((PKG / "host" / "hs_main.cpp", EXE), (PKG / "host" / "hs_shim_check.cpp", SHIM_CHECK)):
        cmd = [HIPCC, "-O2", "-std=c++17", "-Wall", f"-I{PKG.parent / 'include'}", str(src),
               f"-L{PKG}", "-lhsolver_amd", "-Wl,-rpath,$ORIGIN", "-o", str(exe)]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"driver link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    if verbose:
        print(f"built {LIB} and {EXE}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
