"""TEST INFRASTRUCTURE ONLY: compile the shipped adapter (nrpy_amd/adapter/bhah_amd_adapter.c)
against the stand-in project header in this folder, and link it with a project main() and
libbhah_hip.so into tests/adapter/_build/adapter_main (gcc, no HIP compiler needed)."""

from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
OUT = os.path.join(HERE, "_build")
EXE = os.path.join(OUT, "adapter_main")
# the project of the test program: Spherical, symmetry axis phi (bit 2), no KO
DEFS = ["-DBHAH_AMD_SYMMETRY_AXES=4", "-DBHAH_AMD_ENABLE_KO=0"]
CFLAGS = ["-O2", "-std=gnu11", "-Wall", "-Wextra", "-Werror", "-I", HERE, "-I", os.path.join(ROOT, "include")]


def build(out: str = OUT) -> str:
    os.makedirs(out, exist_ok=True)
    objs = []
    for src in (os.path.join(ROOT, "nrpy_amd", "adapter", "bhah_amd_adapter.c"), os.path.join(HERE, "adapter_main.c")):
        obj = os.path.join(out, os.path.basename(src)[:-2] + ".o")
        subprocess.run(["gcc", *CFLAGS, *DEFS, "-c", src, "-o", obj], check=True, capture_output=True, text=True)
        objs.append(obj)
    exe = os.path.join(out, "adapter_main")
    libdir = os.path.join(ROOT, "nrpy_amd", "lib")
    rel = os.path.relpath(libdir, out)
    subprocess.run(["gcc", "-o", exe, *objs, "-L", libdir, "-lbhah_hip", "-Wl,-rpath-link,/opt/rocm/lib",
                    f"-Wl,-rpath,$ORIGIN/{rel}"], check=True, capture_output=True, text=True)
    return exe


if __name__ == "__main__":
    print(build(sys.argv[1] if len(sys.argv) > 1 else OUT))
