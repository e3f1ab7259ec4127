"""Build recipe for ``libgsx.so`` (the gfx950 kernels + C ABI).

Compiles every ``csrc/*.hip`` with hipcc for gfx950 and every host-only
``csrc/*.cpp`` with the host C++ compiler into ``_build/``, and links
``gstools_amd/libgsx.so`` in-tree.  Run ``python -m gstools_amd.build``.
"""

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libgsx.so")
ARCH = os.environ.get("GSX_OFFLOAD_ARCH", "gfx950")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    # packed FP32 instructions stay disabled for every HIP source
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


CXX = os.environ.get("CXX", "g++")
# host code: no contraction of multiply-adds (the sampler's arithmetic is
# restated from numpy's, which rounds every product)
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-ffp-contract=off", "-pthread"]


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith((".hip", ".cpp")))


def _compile(src):
    stem, ext = os.path.splitext(os.path.basename(src))
    obj = os.path.join(BUILD, stem + ("_host" if ext == ".cpp" else "") + ".o")
    if ext == ".cpp":
        cmd = [CXX, *HOST_FLAGS, "-c", src, "-o", obj]
    else:
        cmd = [HIPCC, *FLAGS, "-c", src, "-o", obj]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"compile failed for {src}:\n{res.stderr}")
    return obj


def build(verbose=False):
    """Compile and link; returns the library path."""
    os.makedirs(BUILD, exist_ok=True)
    srcs = sources()
    workers = min(len(srcs), int(os.environ.get("MAX_JOBS", "8")), 16)
    with cf.ThreadPoolExecutor(max_workers=max(1, workers)) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-pthread", *objs, "-o", tmp]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB} from {len(srcs)} sources")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
