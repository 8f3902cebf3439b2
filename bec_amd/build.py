"""Build libbec_amd.so (gfx950) in-tree with hipcc, plus the bec_run driver.

Usage: python -m bec_amd.build [--jobs N]
"""
import argparse
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
ROOT = os.path.dirname(HERE)
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libbec_amd.so")
DRIVER = os.path.join(HERE, "bec_run")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# Packed FP32 instructions are disabled for every device compile.
DEVICE_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
COMMON = ["-std=c++17", "-O3", "-fPIC", "-Wall", "-Wno-unused-result", f"--offload-arch={ARCH}",
          "-I" + os.path.join(ROOT, "include")]

DEVICE_SOURCES = ["kernels.hip", "kernels_gs.hip", "gemm.hip", "fno.hip", "kernels2d.hip", "kernels1d.hip",
                  "static_scratch.hip", "multi.hip"]
HOST_SOURCES = ["bec_api.cpp", "params.cpp", "gs.cpp", "w2d.cpp", "w1d.cpp"]


def _run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in p.stdout.splitlines() if "packed-fp32-ops' is not a recognized" not in l
                    and "warning: feature flag" not in l)
    if p.returncode != 0:
        raise RuntimeError("build failed: " + " ".join(cmd) + "\n" + out)
    return out


def _compile(src):
    path = os.path.join(CSRC, src)
    obj = os.path.join(OBJ, src + ".o")
    # every translation unit is compiled as HIP (host files include device headers)
    cmd = [HIPCC] + COMMON + DEVICE_FLAGS + ["-x", "hip", "-c", path, "-o", obj]
    _run(cmd)
    return obj


def _stale(src, headers_mtime):
    """An object is stale when it is older than its source or than any header."""
    obj = os.path.join(OBJ, src + ".o")
    if not os.path.exists(obj):
        return True
    t = os.path.getmtime(obj)
    return t < os.path.getmtime(os.path.join(CSRC, src)) or t < headers_mtime


def build(jobs=4, verbose=True, incremental=False):
    """Compile every source (incremental=True: only those whose object is stale) and link."""
    os.makedirs(OBJ, exist_ok=True)
    srcs = [s for s in DEVICE_SOURCES + HOST_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    headers = [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith((".hpp", ".h"))]
    headers += [os.path.join(ROOT, "include", f) for f in os.listdir(os.path.join(ROOT, "include"))]
    hm = max(os.path.getmtime(h) for h in headers)
    todo = [s for s in srcs if not incremental or _stale(s, hm)]
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(_compile, todo))
    objs = [os.path.join(OBJ, s + ".o") for s in srcs]
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", LIB] + objs)
    drv = os.path.join(CSRC, "driver.cpp")
    if os.path.exists(drv):
        _run([HIPCC, "-std=c++17", "-O2", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", "-x", "c++", drv,
              "-x", "none", "-o", DRIVER,
              "-I" + os.path.join(ROOT, "include"), "-L" + HERE, "-lbec_amd",
              "-Wl,-rpath,$ORIGIN"])
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=4)
    ap.add_argument("--incremental", action="store_true", help="recompile only sources with stale objects")
    a = ap.parse_args()
    build(a.jobs, incremental=a.incremental)
    sys.exit(0)
