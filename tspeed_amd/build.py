"""Build libtspeed_amd.so in-tree: HIP kernels + C-ABI for gfx950 (hipcc) and the
host C++ layer (g++), linked into one shared library next to this file.

    python -m tspeed_amd.build            # or tspeed_amd.build.build()
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libtspeed_amd.so")
OBJ = os.path.join(HERE, "_build")

ARCH = os.environ.get("TSPEED_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")

DEVICE_SRC = ["device/assemble.hip", "device/step.hip", "device/spmm.hip", "device/multi.hip", "device/capi.hip"]
HOST_SRC = ["host/tspeed.cpp", "host/host_capi.cpp", "host/server.cpp"]

# Packed FP32 VALU forms are switched off for every device compile.
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wall", "-Wno-unused-function", "-Rpass-analysis=kernel-resource-usage"]

# Step kernels run every time step: a register spill there costs bandwidth (and
# one spilling variant was measured wrong), so the build refuses any.
HOT_KERNELS = ("k_step64", "k_step32c", "k_step32w", "k_face64", "k_face32", "k_multi32c", "k_multi32s",
               "k_multi32w", "k_multi_fused", "k_step_b32", "k_step_b16", "k_bmulti")
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall"]


def _check_scratch(out):
    """Raise if a hot kernel in the resource-usage remarks uses scratch."""
    name = None
    for line in out.splitlines():
        if "Function Name:" in line:
            name = line.split("Function Name:")[1].split()[0]
        elif "ScratchSize" in line and name and any(k in name for k in HOT_KERNELS):
            size = int(line.split("ScratchSize [bytes/lane]:")[1].split()[0])
            if size:
                raise RuntimeError(f"hot kernel {name} spills {size} bytes/lane to scratch")


def _run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if p.returncode == 0 and "-Rpass-analysis=kernel-resource-usage" in cmd:
        _check_scratch(p.stdout)
    out = "\n".join(l for l in p.stdout.splitlines() if "packed-fp32-ops" not in l and
                    "not a recognized feature for this target" not in l and "remark:" not in l)
    if p.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{out}")
    return out


def build(verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    jobs = []
    for src in DEVICE_SRC:
        obj = os.path.join(OBJ, os.path.basename(src) + ".o")
        jobs.append([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
    for src in HOST_SRC:
        obj = os.path.join(OBJ, os.path.basename(src) + ".o")
        jobs.append([CXX, *CXX_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
    with cf.ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        for out in ex.map(_run, jobs):
            if verbose and out.strip():
                print(out)
    objs = [j[-1] for j in jobs]
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    # example driver using the C++ facade (links the library just built)
    ex = os.path.join(os.path.dirname(HERE), "examples")
    _run([CXX, *CXX_FLAGS, "-I", os.path.join(CSRC, "host"), os.path.join(ex, "square_driver.cpp"),
          "-L", HERE, "-ltspeed_amd", "-Wl,-rpath,$ORIGIN/../tspeed_amd", "-o", os.path.join(ex, "square_driver")])
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
