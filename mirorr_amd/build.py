"""Builds the HIP library mirorr_amd/lib/libmirorr_gpu.so for gfx950.

hipcc compiles every source in mirorr_amd/csrc (device code for gfx950 only)
and links one shared library exporting the C ABI of include/mirorr_gpu.h and
include/mirorr_io.h (the file I/O part links zlib), then the `mirorr` command-line
executable mirorr_amd/bin/mirorr over that library.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "lib", "obj")
LIB = os.path.join(HERE, "lib", "libmirorr_gpu.so")
CLI = os.path.join(HERE, "bin", "mirorr")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", f"-I{os.path.join(HERE, '..', 'include')}"]
# Device code: gfx950 only; packed FP32 instructions disabled.
DEVICE = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics"]


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC)
                  if f.endswith((".hip", ".cpp")))


def _compile(src: str) -> str:
    obj = os.path.join(OBJ, os.path.basename(src) + ".o")
    cmd = [HIPCC, *COMMON, *DEVICE, "-c", src, "-o", obj]
    if src.endswith(".hip"):
        cmd[1:1] = ["-x", "hip"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{r.stderr}")
    if r.stderr.strip():
        sys.stderr.write(r.stderr)
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    srcs = sources()
    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs, "-lz"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    build_cli()
    if verbose:
        print("built", LIB, "and", CLI)
    return LIB


def build_cli() -> str:
    """The `mirorr` command-line executable (host code over the C ABI),
    linked to the library with a relative run path."""
    os.makedirs(os.path.dirname(CLI), exist_ok=True)
    tmp = CLI + ".tmp"
    cmd = [HIPCC, *COMMON, *DEVICE, os.path.join(HERE, "cli", "mirorr.cpp"), "-o", tmp,
           f"-L{os.path.dirname(LIB)}", "-lmirorr_gpu", "-Wl,-rpath,$ORIGIN/../lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on the command line tool:\n{r.stderr}")
    os.replace(tmp, CLI)
    return CLI


if __name__ == "__main__":
    build(verbose=True)
