"""Build libminife_hip.so (gfx950) in-tree with hipcc.

Every translation unit is compiled with -ffp-contract=off (each a*b+c of the
reference stays two rounded operations, which the bit-exact assembly, SpMV and
CG updates rely on) and with packed-FP32 instruction selection disabled.

    python -m minife_amd.build          # or minife_amd.build.build()
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJDIR = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libminife_hip.so")
SOURCES = ["mfe_api.hip", "mfe_assemble.hip", "mfe_solve.hip", "mfe_batch.hip", "mfe_multi.hip"]
ARCH = os.environ.get("MFE_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    for c in ("/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def compile_flags() -> list[str]:
    return [
        f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-I", os.path.join(ROOT, "include"), "-I", CSRC, "-Wall", "-Wno-unused-result",
    ]


def _compile(src: str) -> str:
    obj = os.path.join(OBJDIR, os.path.splitext(src)[0] + ".o")
    cmd = [hipcc()] + compile_flags() + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return obj


DRIVER_SRC = os.path.join(HERE, "host", "miniFE_hip_main.cpp")
DRIVER = os.path.join(HERE, "miniFE_hip.x")
DRIVER_ELL = os.path.join(HERE, "miniFE_hip_ell.x")  # the reference's MINIFE_ELL_MATRIX build


def build_driver() -> str:
    """The miniFE.x-compatible driver (host C++ over the C-ABI), with the CSR
    matrix type and, as miniFE_hip_ell.x, with the ELL one."""
    for exe, defs in ((DRIVER, []), (DRIVER_ELL, ["-DMINIFE_ELL_MATRIX"])):
        cmd = ["g++", "-O2", "-std=c++17", "-Wall", *defs, "-I", os.path.join(ROOT, "include"),
               "-I", os.path.join(HERE, "host"), DRIVER_SRC, "-o", exe + ".tmp", "-L", HERE,
               "-lminife_hip", "-L/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath,$ORIGIN",
               "-Wl,-rpath,/opt/rocm/lib"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"driver build failed:\n{r.stdout}\n{r.stderr}")
        os.replace(exe + ".tmp", exe)
    return DRIVER


def build(verbose: bool = False) -> str:
    """Compile all sources and link the shared library (and the driver);
    returns the library path."""
    os.makedirs(OBJDIR, exist_ok=True)
    with ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    build_driver()
    if verbose:
        print("built", LIB, "and", DRIVER)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
