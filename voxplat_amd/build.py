"""Build the HIP library libvp_gpu.so in-tree for gfx950.

    python -m voxplat_amd.build          # or voxplat_amd.build.build_all()

Every .hip file under csrc/ is compiled with hipcc for gfx950 and linked,
together with the C host driver under host/, into voxplat_amd/libvp_gpu.so.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import pathlib
import shutil
import subprocess
import sys

PKG = pathlib.Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
HOST = PKG / "host"
BUILD = PKG / "_build"
LIB = PKG / "libvp_gpu.so"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")

HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-fPIC",
    "-std=c++17",
    "-ffp-contract=off",
    # packed fp32 instructions stay off: some of their forms have given wrong results
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{ROOT / 'include'}",
]
C_FLAGS = ["-O2", "-fPIC", "-std=c11", "-Wall", "-Wextra", "-ffp-contract=off",
           "-D__HIP_PLATFORM_AMD__", f"-I{ROCM}/include", f"-I{ROOT / 'include'}"]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {cmd[0]} {cmd[-1]}")


def build_all(jobs: int | None = None) -> pathlib.Path:
    BUILD.mkdir(exist_ok=True)
    hip_srcs = sorted(CSRC.glob("*.hip"))
    c_srcs = sorted(HOST.glob("*.c"))
    objs: list[pathlib.Path] = []
    cmds: list[list[str]] = []
    for s in hip_srcs:
        o = BUILD / (s.stem + ".o")
        objs.append(o)
        cmds.append([HIPCC, *HIP_FLAGS, "-c", str(s), "-o", str(o)])
    for s in c_srcs:
        o = BUILD / (s.stem + ".host.o")
        objs.append(o)
        cmds.append(["gcc", *C_FLAGS, "-c", str(s), "-o", str(o)])
    jobs = jobs or min(8, os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(jobs) as ex:
        list(ex.map(_run, cmds))
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp),
          *map(str, objs), f"-L{ROCM}/lib", "-lamdhip64"])
    shutil.move(str(tmp), str(LIB))
    return LIB


if __name__ == "__main__":
    print(build_all())
