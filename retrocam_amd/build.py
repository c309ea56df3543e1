"""Build recipes for the native libraries of retrocam_amd.

Produces, in-tree:
  retrocam_amd/lib/libretrocam_hip.so   HIP kernels + C ABI (gfx950 only)
  retrocam_amd/lib/librc_host.so        host scene assembly (g++)

Run ``python -m retrocam_amd.build`` (or ``__graft_entry__.build()``).
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "retrocam_amd")
CSRC = os.path.join(PKG, "csrc")
INC = os.path.join(ROOT, "include")
LIBDIR = os.path.join(PKG, "lib")
OBJDIR = os.path.join(PKG, "_build")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# Device code: gfx950 only, IEEE binary32 with no contraction (host/device
# bit-exactness), correctly rounded '/' and sqrt, packed-FP32 disabled.
HIP_FLAGS = [
    "-O3", "-std=c++17", "--offload-arch=gfx950", "-fPIC", "-fno-gpu-rdc",
    "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
    f"-I{INC}", f"-I{CSRC}",
]

HIP_SOURCES = ["rc_capi.hip", "rc_pathtrace.hip", "rc_pathtrace_lanes.hip", "rc_wavefront.hip", "rc_debug.hip", "rc_post.hip",
               "rc_sppm.hip", "rc_net.hip", "rc_batch.hip", "rc_tune.hip", "rc_export.hip", "rc_noalloc.hip"]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(cmd)}\n{r.stdout}")


def build_hip(jobs: int = 4) -> str:
    os.makedirs(LIBDIR, exist_ok=True)
    os.makedirs(OBJDIR, exist_ok=True)
    objs = []
    cmds = []
    for src in HIP_SOURCES:
        obj = os.path.join(OBJDIR, src.replace(".hip", ".o"))
        objs.append(obj)
        cmds.append([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(_run, cmds))
    out = os.path.join(LIBDIR, "libretrocam_hip.so")
    tmp = out + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, out)
    return out


def build_host() -> str:
    os.makedirs(LIBDIR, exist_ok=True)
    out = os.path.join(LIBDIR, "librc_host.so")
    tmp = out + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off",
          f"-I{INC}", os.path.join(CSRC, "host", "rc_host.cpp"),
          os.path.join(CSRC, "host", "rc_temporal.cpp"), os.path.join(CSRC, "host", "rc_scenejson.cpp"),
          "-o", tmp])
    os.replace(tmp, out)
    return out


def build_all(jobs: int = 4) -> list[str]:
    return [build_host(), build_hip(jobs)]


if __name__ == "__main__":
    for p in build_all(int(sys.argv[1]) if len(sys.argv) > 1 else 4):
        print(p)
