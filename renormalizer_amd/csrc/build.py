"""Build libreno_hip.so (gfx950) in-tree.

Compiles every ``*.hip`` file of this directory with hipcc for gfx950 and
links them, with rocSOLVER/rocBLAS, into ``renormalizer_amd/libreno_hip.so``.
Object files are cached under ``renormalizer_amd/csrc/_build`` and rebuilt
when their source or any header is newer.

Usage: ``python -m renormalizer_amd.csrc.build [-j N] [--force]``
"""
import argparse
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
ROOT = os.path.dirname(PKG)
OUT = os.path.join(PKG, "libreno_hip.so")
BUILD_DIR = os.path.join(HERE, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# Packed FP32 instructions are disabled for all device code (they are not
# needed: the contractions use MFMA) and the target is CDNA4 only.  Matrix-core
# accumulators are kept in ordinary VGPRs: with the AGPR form the compiler
# copied the GEMM accumulators between the two register files on every K step
# and the doubled register footprint halved the waves per SIMD.
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-mllvm", "-amdgpu-mfma-vgpr-form",
    "-Wno-unused-result", "-munsafe-fp-atomics",
    "-I" + os.path.join(ROOT, "include"),
]
LINK_FLAGS = ["-shared", "-fPIC", "--offload-arch=gfx950",
              "-L/opt/rocm/lib", "-lrocsolver", "-lrocblas",
              "-Wl,-rpath,/opt/rocm/lib"]


def _sources():
    return sorted(f for f in os.listdir(HERE) if f.endswith(".hip"))


def _headers():
    hs = [os.path.join(HERE, f) for f in os.listdir(HERE) if f.endswith(".h")]
    hs.append(os.path.join(ROOT, "include", "reno_hip.h"))
    return hs


def _compile(src, force):
    obj = os.path.join(BUILD_DIR, src.replace(".hip", ".o"))
    srcp = os.path.join(HERE, src)
    newest_dep = max([os.path.getmtime(srcp)] + [os.path.getmtime(h) for h in _headers()])
    if not force and os.path.exists(obj) and os.path.getmtime(obj) >= newest_dep:
        return obj
    cmd = [HIPCC] + HIP_FLAGS + ["-c", srcp, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr}")
    return obj


def build(jobs=8, force=False, verbose=False):
    os.makedirs(BUILD_DIR, exist_ok=True)
    srcs = _sources()
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda s: _compile(s, force), srcs))
    tmp = OUT + ".tmp"
    cmd = [HIPCC] + objs + LINK_FLAGS + ["-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, OUT)
    if verbose:
        print(f"built {OUT}")
    return OUT


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", type=int, default=8)
    ap.add_argument("--force", action="store_true")
    a = ap.parse_args()
    try:
        build(a.j, a.force, verbose=True)
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
