"""Static checks on the gfx950 assembly of the generated kernels.

On gfx9-family GPUs a wave counts its outstanding vector-memory operations
(loads, stores, scratch spills and reloads) in the 6-bit ``vmcnt`` field, so at
most 63 can be tracked; ``s_waitcnt vmcnt(N)`` then waits until at most N are
outstanding.  A schedule that keeps more than 63 in flight leaves the waits
with nothing to count against, so a value (including a spilled address) can be
consumed before it arrives.  Kernels whose stencil loads are issued early
(``_pf`` schedules) are built only if a linear scan of their assembly never
exceeds this limit.  The scan follows the listing order and counts both arms of
a branch, so it over-estimates and errs on the safe side.
"""

from __future__ import annotations

import re
from typing import Dict

VMCNT_MAX = 63

_VMEM = re.compile(r"^\s*(global_load|global_store|global_atomic|scratch_load|scratch_store|buffer_load|buffer_store"
                   r"|flat_load|flat_store)")
_WAIT = re.compile(r"^\s*s_waitcnt\b.*?vmcnt\((\d+)\)")
_LABEL = re.compile(r"^([A-Za-z_][\w.]*):")


def max_outstanding_vmem(asm: str) -> Dict[str, int]:
    """Per kernel symbol: the largest number of vector-memory operations in flight."""
    res: Dict[str, int] = {}
    kern, out, mx = None, 0, 0
    for line in asm.splitlines():
        m = _LABEL.match(line)
        if m and not m.group(1).startswith((".", "__hip")):
            if kern is not None:
                res[kern] = mx
            kern, out, mx = m.group(1), 0, 0
            continue
        if kern is None:
            continue
        if _VMEM.match(line):
            out += 1
            mx = max(mx, out)
            continue
        w = _WAIT.match(line)
        if w:
            out = min(out, int(w.group(1)))
    if kern is not None:
        res[kern] = mx
    return res


def check(asm: str, source: str = "") -> Dict[str, int]:
    counts = max_outstanding_vmem(asm)
    bad = {k: v for k, v in counts.items() if v > VMCNT_MAX}
    if bad:
        raise RuntimeError(f"{source}: kernels keep more than {VMCNT_MAX} vector-memory operations in flight "
                           f"(vmcnt is 6 bits): {bad}")
    return counts
