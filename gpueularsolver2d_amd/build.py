"""Builds libfvm_hip.so (gfx950) in-tree with hipcc, and the host driver.

    python -m gpueularsolver2d_amd.build        # library, fvm_solver driver, libmeshgen.so

FP64 translation units are compiled with -ffp-contract=off so every multiply
and add is rounded separately, matching the FP64 oracle bit for bit; the FP32
tiled kernels use fused multiply-adds within each expression,
reciprocal-multiply division and the hardware FP32 reciprocal and square root.  Packed FP32 instructions are
disabled for all device code.
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libfvm_hip.so")
DRIVER = os.path.join(HERE, "fvm_solver")

ARCH = os.environ.get("FVM_OFFLOAD_ARCH", "gfx950")
COMMON = ["-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}", "-Wall",
          "-Wno-unused-function", "-I", CSRC, "-I", os.path.join(ROOT, "include"),
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]

SOURCES = [
    ("fvm_mesh.cpp", ["-ffp-contract=off", "-x", "hip"]),
    ("fvm_plain.hip", ["-ffp-contract=off"]),
    ("fvm_tiled64.hip", ["-ffp-contract=off"]),
    # FP32: fused multiply-adds within each source expression
    # (-ffp-contract=on: the rounding of an expression is fixed by the source,
    # so every copy of it the compiler makes, peeled or not, gives the same
    # bits, and cells recomputed by several tiles agree), division as
    # reciprocal-multiply and the hardware reciprocal / square root (about
    # 1 ulp), FP32 denormals flushed; no reassociation and no finite-math
    # assumptions.  This removes the range-scaling sequences around every FP32
    # divide (about 40% of the flux kernel's VALU instructions).
    ("fvm_tiled32.hip", ["-ffp-contract=on", "-fno-hip-fp32-correctly-rounded-divide-sqrt",
                         "-freciprocal-math", "-fapprox-func", "-fgpu-flush-denormals-to-zero"]),
    ("fvm_engine.hip", ["-ffp-contract=off"]),
]


def hipcc():
    h = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(h):
        raise RuntimeError("hipcc not found: the HIP toolchain is required to build libfvm_hip.so")
    return h


def _compile(src, extra, build_dir=BUILD):
    obj = os.path.join(build_dir, os.path.splitext(src)[0] + ".o")
    cmd = [hipcc()] + COMMON + extra + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compiling {src} failed:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build_library(verbose=False, out=LIB, extra=None, build_dir=BUILD):
    """extra: {source name: [additional flags]} for variant builds."""
    os.makedirs(build_dir, exist_ok=True)
    extra = extra or {}
    with ThreadPoolExecutor(max_workers=min(len(SOURCES), 8)) as ex:
        objs = list(ex.map(lambda s: _compile(s[0], s[1] + extra.get(s[0], []), build_dir), SOURCES))
    tmp = out + ".tmp"
    cmd = [hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"linking libfvm_hip.so failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, out)
    if verbose:
        print("built", out)
    return out


def build_driver(verbose=False):
    """The reference-compatible host program (inputMesh ... Output) linked
    against libfvm_hip.so."""
    src = os.path.join(HERE, "host", "gpu_fvm_host.cpp")
    if not os.path.exists(src):
        return None
    cmd = ["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-I", os.path.join(ROOT, "include"),
           src, "-o", DRIVER, "-L", HERE, "-lfvm_hip", f"-Wl,-rpath,{HERE}", "-lm"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"building the host driver failed:\n{r.stdout}\n{r.stderr}")
    if verbose:
        print("built", DRIVER)
    return DRIVER


MESHGEN = os.path.join(HERE, "libmeshgen.so")


def build_meshgen(verbose=False):
    """The synthetic meshes' Delaunay triangulator (host code, g++)."""
    src = os.path.join(CSRC, "meshgen_delaunay.cpp")
    tmp = MESHGEN + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-Wall", src, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"building libmeshgen.so failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, MESHGEN)
    if verbose:
        print("built", MESHGEN)
    return MESHGEN


def build(verbose=False):
    build_library(verbose)
    build_driver(verbose)
    build_meshgen(verbose)


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
