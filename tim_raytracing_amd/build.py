"""In-tree build of the two native libraries.

  librtq_build.so  host builder + synthetic workloads (g++)
  librtq.so        gfx950 kernels + C ABI (hipcc --offload-arch=gfx950)

Floating-point flags are part of the contract with the reference arithmetic: no contraction into
FMA, IEEE division and square root, no denormal flushing, no fast-math. Packed fp32 instruction
selection is disabled for device code.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
HOST_SRC = [os.path.join(PKG, "csrc", "host", f) for f in ("bvh_builder.cpp", "scenes.cpp", "blob_cache.cpp",
                                                         "capi_build.cpp")]
DEVICE_SRC = [os.path.join(PKG, "csrc", "device", "rtq.hip")]
DEVICE_DEPS = [os.path.join(PKG, "csrc", "device", f)
               for f in ("kernels.h", "layout.h", "scene_layout.h", "stack_class.h", "batch.inc", "serve.inc", "serve_core.h")]

HOST_FLAGS = ["-std=c++17", "-O2", "-fPIC", "-shared", "-ffp-contract=off", "-fno-fast-math", "-Wall",
              "-Wextra", "-pthread"]
HIP_FLAGS = ["--offload-arch=gfx950", "-std=c++17", "-O3", "-fPIC", "-shared", "-ffp-contract=off",
             "-fno-fast-math", "-fhip-fp32-correctly-rounded-divide-sqrt", "-fno-gpu-flush-denormals-to-zero",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wall"]


def _hipcc() -> str:
    for c in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build failed: {' '.join(cmd[:3])} ...")
    return r


def build_host(out: str | None = None, extra=()) -> str:
    """extra: more compiler flags (the tests build planted-mistake variants with -DRTQ_BUILDER_PLANT=k
    into their own directories)."""
    out = out or os.path.join(PKG, "librtq_build.so")
    tmp = out + ".tmp"
    _run(["g++", *HOST_FLAGS, *extra, "-o", tmp, *HOST_SRC])
    os.replace(tmp, out)
    return out


def build_device(out: str | None = None, extra=()) -> str:
    out = out or os.path.join(PKG, "librtq.so")
    tmp = out + ".tmp"
    _run([_hipcc(), *HIP_FLAGS, *extra, "-o", tmp, *DEVICE_SRC])
    os.replace(tmp, out)
    return out


def build_oracle():
    oracle_dir = os.path.join(ROOT, "oracle")
    _run(["make", "-s", "-B", "-C", oracle_dir])


def build_all():
    build_host()
    build_device()
    build_oracle()


if __name__ == "__main__":
    build_all()
    print("built", os.path.join(PKG, "librtq_build.so"), os.path.join(PKG, "librtq.so"))
