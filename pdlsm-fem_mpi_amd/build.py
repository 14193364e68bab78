"""Build libpdx.so (HIP for gfx950 + host C++ set-up) in-tree.

    python pdlsm-fem_mpi_amd/build.py

Compiles every source under csrc/ with hipcc and links one shared library
next to this file.  Packed FP32 instructions are disabled for all device code.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libpdx.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

COMMON = ["-O3", "-std=c++17", "-fPIC", "-fopenmp", "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-result", "-Wno-unused-value"]
# -amdgpu-mfma-vgpr-form: matrix-core accumulators stay in VGPRs; with the default AGPR form the
# compiler copies a loop-carried accumulator to AGPRs and back around every MFMA and waits out its
# latency each time
DEVICE = ["--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-mllvm", "-amdgpu-mfma-vgpr-form=1"]
SOURCES = ["model.cpp", "assemble.hip", "solve.hip", "krylov.hip", "context.hip", "newmark.hip", "csr_solve.hip",
           "learned.hip", "family_batch.hip", "stress.hip", "vtk.cpp", "ilu0.hip", "setup.hip"]


def _compile(src: str, force: bool = False) -> str:
    path = os.path.join(CSRC, src)
    obj = os.path.join(OBJ, src + ".o")
    deps = [path, os.path.abspath(__file__)] + [os.path.join(CSRC, h) for h in os.listdir(CSRC) if h.endswith(".hpp")]
    deps.append(os.path.join(HERE, "..", "include", "pdx.h"))
    if not force and os.path.exists(obj) and all(os.path.getmtime(obj) >= os.path.getmtime(d) for d in deps):
        return obj
    cmd = [HIPCC] + COMMON
    if src.endswith(".hip"):
        cmd += ["-x", "hip"] + DEVICE
    cmd += ["-c", path, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = False, force: bool = False) -> str:
    """Compile the sources (all of them when force, else those older than their
    object) and link libpdx.so."""
    os.makedirs(OBJ, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(lambda src: _compile(src, force), srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fopenmp", "--offload-arch=gfx950", "-o", tmp] + objs + [
        "-Wl,-rpath,/opt/rocm/lib:/opt/rocm/llvm/lib", "-L/opt/rocm/llvm/lib", "-lomp"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True, force="--force" in sys.argv)
    sys.exit(0)
