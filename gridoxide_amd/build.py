"""Builds the in-tree HIP library ``gridoxide_amd/libgridox_hip.so`` for gfx950 with hipcc."""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libgridox_hip.so")
SOURCES = ["gx_host.cpp", "gx_symbolic.cpp", "gx_kernels.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
COMMON = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
    # no packed FP32 (v_pk_*_f32) code generation anywhere
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"), "-I", CSRC,
]


def build(verbose: bool = False) -> str:
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)

    def compile_one(src):
        obj = os.path.join(objdir, src + ".o")
        lang = ["-x", "hip"] if src.endswith(".hip") else []
        cmd = [HIPCC] + COMMON + lang + ["-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
        return obj

    with ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        objs = list(ex.map(compile_one, SOURCES))
    tmp = OUT + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, OUT)
    return OUT


# Host-code sanitizer build (CPU checks only): the host sources compiled with AddressSanitizer and
# UndefinedBehaviorSanitizer (every -fsanitize applies to the host pass alone), linked with the
# normal build's kernel object.  Loaded through GRIDOX_HIP_LIB with the ASan runtime preloaded.
SAN_DIR = os.path.join(HERE, "_build_san")
SAN_OUT = os.path.join(SAN_DIR, "libgridox_hip_san.so")
SAN_FLAGS = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
             "-Xarch_host", "-fno-sanitize-recover=undefined", "-Xarch_host", "-fno-omit-frame-pointer", "-g"]


def sanitizer_runtime() -> str:
    """The compiler's ASan runtime (it carries the UBSan handlers), to preload into a process that
    loads the sanitized library."""
    import glob
    found = sorted(glob.glob(os.path.join(os.path.dirname(os.path.dirname(HIPCC)), "llvm", "lib", "clang", "*", "lib",
                                          "linux", "libclang_rt.asan-x86_64.so")))
    if not found:
        raise RuntimeError("no libclang_rt.asan-x86_64.so next to hipcc")
    return found[-1]


def build_sanitized(verbose: bool = False) -> str:
    kobj = os.path.join(HERE, "_build", "gx_kernels.hip.o")
    if not os.path.exists(kobj):
        build(verbose)
    os.makedirs(SAN_DIR, exist_ok=True)
    objs = []
    for src in SOURCES:
        if src.endswith(".hip"):
            continue
        obj = os.path.join(SAN_DIR, src + ".o")
        cmd = [HIPCC] + COMMON + SAN_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"hipcc (sanitized) failed for {src}:\n{r.stdout}\n{r.stderr}")
        objs.append(obj)
    tmp = SAN_OUT + ".tmp"
    r = subprocess.run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs + [kobj],
                       capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link (sanitized) failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, SAN_OUT)
    return SAN_OUT


if __name__ == "__main__":
    print(build_sanitized(verbose="-v" in sys.argv) if "--sanitize" in sys.argv else build(verbose="-v" in sys.argv))
