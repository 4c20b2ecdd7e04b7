"""Build libsalsa_amd.so in-tree with hipcc for gfx950.

    python uclales-salsa_amd/build.py [--jobs N]

The translation units are compiled in parallel and linked into
uclales-salsa_amd/salsa_amd/libsalsa_amd.so.  Packed FP32 instructions are
disabled for all device code.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("SALSA_ARCH", "gfx950")
SRCS = ["csrc/salsa_layout.cpp", "csrc/salsa_kernels.hip", "csrc/salsa_gradient.hip", "csrc/salsa_adjoint.hip",
        "csrc/salsa_les.hip", "csrc/salsa_sed.hip", "csrc/salsa_capi.hip"]
OUT = os.path.join(HERE, "salsa_amd", "libsalsa_amd.so")
FLAGS = ["-O3", "--offload-arch=" + ARCH, "-fPIC", "-std=c++17",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-unused-result",
         "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(HERE, "csrc")]


def _compile(src):
    obj = os.path.join(HERE, "_build", os.path.basename(src) + ".o")
    cmd = [HIPCC] + FLAGS + ["-c", os.path.join(HERE, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    err = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s" % (" ".join(cmd), err))
    return obj


def build(jobs=3, verbose=False):
    os.makedirs(os.path.join(HERE, "_build"), exist_ok=True)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, SRCS))
    tmp = OUT + ".tmp"
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s" % (" ".join(cmd), r.stderr))
    os.replace(tmp, OUT)
    if verbose:
        print("built", OUT)
    return OUT


if __name__ == "__main__":
    j = 3
    if "--jobs" in sys.argv:
        j = int(sys.argv[sys.argv.index("--jobs") + 1])
    build(j, verbose=True)
