"""Build the gfx950 shared library ``nrpy_amd/lib/libbhah_hip.so`` in-tree.

Each generated kernel translation unit (``csrc/gen/*.hip``) and the host
runtime (``csrc/bhah_hip.hip``) are compiled by ``hipcc`` in parallel, then
linked.  ``python -m nrpy_amd.build [--regen]``; ``--regen`` first re-emits the
kernels from the SymPy transcription.  Schedules with early-issued loads are
also compiled to assembly and refused if they can keep more vector-memory
operations in flight than the 6-bit vmcnt counter tracks, or if a DMA-staged box
can be read before its loads have retired (``isa_check``).
"""

from __future__ import annotations

import concurrent.futures as cf
import os
import re
import shutil
import subprocess
import sys
from typing import List

from . import isa_check

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
GEN = os.path.join(CSRC, "gen")
LIBDIR = os.path.join(HERE, "lib")
OBJDIR = os.path.join(HERE, "_build")
LIB = os.path.join(LIBDIR, "libbhah_hip.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# Packed-FP32 instructions are disabled for all HIP code; the kernels are FP64 anyway.
HIPFLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-fPIC",
    "-std=c++17",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-variable",
]


def sources() -> List[str]:
    gen = sorted(os.path.join(GEN, f) for f in os.listdir(GEN) if f.endswith(".hip"))
    return gen + [os.path.join(CSRC, "bhah_hip.hip")]


def _flags(src: str) -> List[str]:
    """Per-source flags: schedules named with ``_mi`` are scheduled by LLVM's max-ILP
    strategy, which keeps fewer values in flight through the point algebra."""
    if re.search(r"_mi(_|\.hip$)", os.path.basename(src)):
        return [*HIPFLAGS, "-mllvm", "-amdgpu-sched-strategy=max-ilp"]
    return HIPFLAGS


def _compile(src: str) -> str:
    obj = os.path.join(OBJDIR, os.path.basename(src).replace(".hip", ".o"))
    cmd = [HIPCC, *_flags(src), "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr[-4000:]}")
    if re.search(r"_pf|_L\d", os.path.basename(src)):
        # early-issued stencil loads: verify the vmcnt range on the emitted assembly
        asm = obj[:-2] + ".s"
        r = subprocess.run([HIPCC, *_flags(src), "--cuda-device-only", "-S", src, "-o", asm], capture_output=True,
                           text=True)
        if r.returncode != 0:
            raise RuntimeError(f"hipcc -S failed for {src}:\n{r.stderr[-4000:]}")
        with open(asm) as f, open(src) as g:
            isa_check.check(f.read(), src, source_text=g.read())
    return obj


def build(regen: bool = False, jobs: int | None = None) -> str:
    """Compile every HIP source for gfx950 and link the shared library."""
    if regen:
        from .codegen import hip_gen

        hip_gen.generate(GEN)
    if not shutil.which(HIPCC) and not os.path.exists(HIPCC):
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    os.makedirs(OBJDIR, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    srcs = sources()
    jobs = jobs or min(len(srcs), max(1, (os.cpu_count() or 4)))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-4000:]}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(regen="--regen" in sys.argv))
