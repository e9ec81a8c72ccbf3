"""Build the CPU oracle (test infrastructure): oracle/libvp_oracle.so.

The reference's own sources for this path cannot be built here without
stand-ins: src/chunkset/mesher.c calls chunk_open_ro and chunkset_edit_read
from chunkset.c and edit.c, which include <cglm/cglm.h> (absent); rle.c needs
mem.c, whose shell commands leave gfx_shell_exclusive_draw (OpenGL) and
ctx_terminate (GLFW) unresolved.  So there is no oracle/_ref build.
"""
from __future__ import annotations

import pathlib
import subprocess
import sys

HERE = pathlib.Path(__file__).resolve().parent
LIB = HERE / "libvp_oracle.so"

# -O2 as the reference build (src/CMakeLists.txt:70); x86-64-v3 instead of
# -march=native so the library also runs on the benchmark host's CPU.
FLAGS = ["-O2", "-march=x86-64-v3", "-fopenmp", "-ffp-contract=off", "-fPIC", "-shared",
         "-std=gnu11", "-Wall", "-Wextra"]


def build() -> pathlib.Path:
    tmp = LIB.with_suffix(".so.tmp")
    cmd = ["gcc", *FLAGS, "-o", str(tmp), str(HERE / "vp_oracle.c"), "-lm"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("oracle build failed")
    tmp.replace(LIB)
    return LIB


if __name__ == "__main__":
    print(build())
