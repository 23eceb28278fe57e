"""Build libsparrow_amd.so (HIP kernels + host C++ + C ABI) for gfx950 with hipcc.

Usage: python -m sparrow_amd.build   (or sparrow_amd.build.build())
The library is written in-tree (sparrow_amd/libsparrow_amd.so); objects go to sparrow_amd/_build/.
"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libsparrow_amd.so")
OBJDIR = os.path.join(HERE, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("SPARROW_AMD_ARCH", "gfx950")

SOURCES = ["integrals.hip", "fock.hip", "cis.hip", "linalg.hip", "tables.cpp", "params.cpp", "engine.cpp", "workspace.cpp",
           "server.cpp", "noalloc.cpp", "capi.cpp"]

COMMON = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", f"-I{os.path.join(ROOT, 'include')}",
          # packed-FP32 instructions are disabled for all device code of this library
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          # MFMA accumulators in VGPRs: keeps loop-carried accumulators out of AGPR<->VGPR copies
          "-mllvm", "-amdgpu-mfma-vgpr-form=1"]


def _compile(src):
    obj = os.path.join(OBJDIR, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + COMMON + ["-c", os.path.join(CSRC, src), "-o", obj]
    if src.endswith(".cpp"):
        cmd = [HIPCC] + COMMON + ["-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    errs = "\n".join(l for l in r.stderr.splitlines() if "unknown target feature" not in l and "packed-fp32" not in l
                     and l.strip() and not l.strip().startswith("1 warning"))
    if r.returncode != 0:
        raise RuntimeError(f"compiling {src} failed:\n{r.stderr}")
    return obj, errs


def build(verbose=False):
    os.makedirs(OBJDIR, exist_ok=True)
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 4)) as ex:
        results = list(ex.map(_compile, SOURCES))
    objs = []
    for obj, errs in results:
        objs.append(obj)
        if verbose and errs:
            print(errs, file=sys.stderr)
    tmp = OUT + ".tmp"
    cmd = [HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"linking libsparrow_amd.so failed:\n{r.stderr}")
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
