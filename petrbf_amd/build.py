"""Build the HIP library petrbf_amd/libpetrbf.so for gfx950 (and the CPU oracle).

Every translation unit under petrbf_amd/csrc is compiled by hipcc for gfx950
into petrbf_amd/_build/, then linked into one shared library that exports
the C ABI declared in include/petrbf.h.
"""

from __future__ import annotations

import concurrent.futures as cf
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libpetrbf.so")
ARCHIVE = os.path.join(HERE, "libpetrbf.a")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

ARCH = "gfx950"
CFLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # packed FP32 VALU forms are kept out of the generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
    "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
]


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = [HIPCC, *CFLAGS, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return obj


def build_library(verbose: bool = True) -> str:
    os.makedirs(BUILD, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    # the same objects as a static archive, for callers that link the library
    # into their executable (the reference's 3D tree ships librbf.a,
    # 3d/Makefile:12-14); link it with hipcc so the HIP runtime comes along
    if os.path.exists(ARCHIVE):
        os.remove(ARCHIVE)
    r = subprocess.run(["ar", "rcs", ARCHIVE, *objs], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"ar failed:\n{r.stdout}\n{r.stderr}")
    if verbose:
        print(f"built {LIB} and {ARCHIVE}")
    return LIB


def build_oracle() -> None:
    subprocess.check_call(["sh", os.path.join(ROOT, "oracle", "build_oracle.sh")])


if __name__ == "__main__":
    build_library()
    if "--no-oracle" not in sys.argv:
        build_oracle()
