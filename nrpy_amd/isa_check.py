"""Static checks on the gfx950 assembly of the generated kernels.

A wave counts its outstanding vector-memory operations (loads, stores, global-to-LDS
loads, scratch spills and reloads) in ``vmcnt``; ``s_waitcnt vmcnt(N)`` waits until
at most N are outstanding, and the operations retire in issue order.  The field is
6 bits wide in the gfx9 ``s_waitcnt`` encoding (bits 3:0 and 15:14), so a wait can
name at most 63.  Two checks run on every schedule with early-issued loads:

1. **In-flight bound.**  No path through a kernel keeps more than 63 operations in
   flight.  The scan follows the control-flow graph of the listing: branches to
   both successors, backward branches of loops until the counts stop changing, the
   larger count where paths join.  A count that keeps growing around a loop is
   reported as unbounded.

2. **DMA waits** (schedules that fill LDS boxes with global-to-LDS loads).  The
   generator writes, per kernel, the phases of its DMA groups in issue order
   (``// bhah-dma-issue <kernel>: phase:loads ...``) and tags the barrier before the
   reads of phase p (``s_barrier ; bhah_read_phase p``).  The k-th
   ``global_load_lds`` of the listing is taken as the k-th DMA load issued (the
   listing has them in the generator's order; loads cannot cross the tagged
   barriers, which clobber memory, and their count between tags is checked).  A
   dataflow over the same graph keeps, per DMA load still possibly outstanding, the
   fewest vector-memory operations issued after it on any path; ``vmcnt(N)``
   retires it when that number is at least N.  At each tagged barrier every load
   of phase <= p must be retired on every path, otherwise the build is refused.
"""

from __future__ import annotations

import re
from typing import Callable, Dict, List, Optional, Tuple

VMCNT_MAX = 63
_SATURATE = 10 * VMCNT_MAX

_VMEM = re.compile(r"^\s*(global_load|global_store|global_atomic|scratch_load|scratch_store|buffer_load|buffer_store"
                   r"|buffer_atomic|flat_load|flat_store|flat_atomic)")
_DMA = re.compile(r"^\s*(global_load_lds|buffer_load_\w+\s.*\blds\b)")
_WAIT = re.compile(r"^\s*s_waitcnt\b.*?vmcnt\((\d+)\)")
_KLABEL = re.compile(r"^([A-Za-z_][\w.]*):")
_BLABEL = re.compile(r"^(\.L\w+):")
_BRANCH = re.compile(r"^\s*s_branch\s+(\.L\w+)")
_CBRANCH = re.compile(r"^\s*s_cbranch_\w+\s+(\.L\w+)")
_END = re.compile(r"^\s*(s_endpgm|s_setpc_b64|s_trap)\b")
_PHASE = re.compile(r"^\s*s_barrier\b.*bhah_read_phase\s+(\d+)")
_ISSUE = re.compile(r"//\s*bhah-dma-issue\s+(\w+):\s*([\d: ]+)")


class _Fn:
    """Basic blocks of one kernel's listing and their successors."""

    def __init__(self, lines: List[str]):
        self.blocks: List[List[str]] = [[]]
        labels: Dict[str, int] = {}
        ends: List[Optional[str]] = []
        for line in lines:
            code = line.split("//")[0]
            m = _BLABEL.match(code)
            if m:
                if self.blocks[-1]:
                    self.blocks.append([])
                labels[m.group(1)] = len(self.blocks) - 1
                continue
            if not code.strip() or code.strip().startswith((";", ".")):
                continue
            self.blocks[-1].append(line)
            if _BRANCH.match(code) or _CBRANCH.match(code) or _END.match(code):
                self.blocks.append([])
        self.succ: List[List[int]] = []
        n = len(self.blocks)
        for i, b in enumerate(self.blocks):
            last = b[-1] if b else ""
            if (m := _BRANCH.match(last)):
                self.succ.append([labels[m.group(1)]])
            elif (m := _CBRANCH.match(last)):
                self.succ.append([labels[m.group(1)]] + ([i + 1] if i + 1 < n else []))
            elif _END.match(last):
                self.succ.append([])
            else:
                self.succ.append([i + 1] if i + 1 < n else [])

    def dataflow(self, init, transfer: Callable, join: Callable, max_visits: int = 100000):
        """Forward fixed point; ``transfer(block_index, state) -> state``."""
        inn = {0: init}
        work = [0]
        visits = 0
        while work:
            b = work.pop()
            visits += 1
            if visits > max_visits:
                raise RuntimeError("assembly dataflow did not converge")
            out = transfer(b, inn[b])
            for s in self.succ[b]:
                new = out if s not in inn else join(inn[s], out)
                if s not in inn or new != inn[s]:
                    inn[s] = new
                    work.append(s)
        return inn


def _functions(asm: str) -> Dict[str, List[str]]:
    res: Dict[str, List[str]] = {}
    kern = None
    for line in asm.splitlines():
        m = _KLABEL.match(line)
        if m and not m.group(1).startswith((".", "__hip")):
            kern = m.group(1)
            res[kern] = []
            continue
        if kern is None:
            continue
        if line.startswith(".Lfunc_end"):
            kern = None
            continue
        res[kern].append(line)
    return res


def max_outstanding_vmem(asm: str) -> Dict[str, int]:
    """Per kernel symbol: the largest number of vector-memory operations in flight on any
    path (``_SATURATE`` or more: unbounded around a loop)."""
    res: Dict[str, int] = {}
    for kern, lines in _functions(asm).items():
        fn = _Fn(lines)
        peak = [0]

        def transfer(b, out):
            for line in fn.blocks[b]:
                if _VMEM.match(line):
                    out = min(out + 1, _SATURATE)
                    peak[0] = max(peak[0], out)
                elif (w := _WAIT.match(line)):
                    out = min(out, int(w.group(1)))
            return out

        fn.dataflow(0, transfer, max)
        res[kern] = peak[0]
    return res


def dma_issue_orders(source: str) -> Dict[str, List[Tuple[int, int]]]:
    """The generator's ``bhah-dma-issue`` notes: kernel -> [(phase, loads per lane), ...]."""
    out = {}
    for m in _ISSUE.finditer(source):
        out[m.group(1)] = [(int(a), int(b)) for a, b in (t.split(":") for t in m.group(2).split())]
    return out


def check_dma_waits(asm: str, issue: Dict[str, List[Tuple[int, int]]]) -> Dict[str, int]:
    """Prove for every kernel in ``issue`` that each tagged phase barrier is reached with
    all DMA loads of that phase and earlier ones retired.  Returns kernel -> number of
    tagged barriers checked; raises RuntimeError naming the first violation."""
    fns = _functions(asm)
    done: Dict[str, int] = {}
    for kern, groups in issue.items():
        if kern not in fns:
            raise RuntimeError(f"{kern}: listed in bhah-dma-issue but absent from the assembly")
        lines = fns[kern]
        phase_of: List[int] = [ph for ph, n in groups for _ in range(n)]
        ndma = sum(1 for line in lines if _DMA.match(line))
        if ndma != len(phase_of):
            raise RuntimeError(f"{kern}: {ndma} global-to-LDS loads in the listing, the schedule issues {len(phase_of)}")
        nphase = max(ph for ph, _ in groups) + 1
        seen = 0
        tags: List[Tuple[int, int]] = []  # (phase, DMA loads statically before the tag)
        for line in lines:
            if _DMA.match(line):
                seen += 1
            elif (m := _PHASE.match(line)):
                tags.append((int(m.group(1)), seen))
        if sorted(set(p for p, _ in tags)) != list(range(nphase)):
            raise RuntimeError(f"{kern}: tagged barriers {sorted(set(p for p, _ in tags))} do not cover phases 0..{nphase - 1}")
        for p, nbefore in tags:
            # every load of phase <= p must have been issued before the barrier of phase p
            need = max(i for i, ph in enumerate(phase_of) if ph <= p) + 1
            if nbefore < need:
                raise RuntimeError(f"{kern}: barrier of phase {p} precedes {need - nbefore} of that phase's DMA loads")
        fn = _Fn(lines)
        # static ordinal of each DMA instruction, per block
        ordinal: Dict[Tuple[int, int], int] = {}
        k = 0
        for bi, blk in enumerate(fn.blocks):
            for li, line in enumerate(blk):
                if _DMA.match(line):
                    ordinal[(bi, li)] = k
                    k += 1
        violations: List[str] = []
        checked = [0]

        def transfer(b, state):
            pend = dict(state)  # DMA ordinal -> fewest operations issued after it
            for li, line in enumerate(fn.blocks[b]):
                if _VMEM.match(line) or _DMA.match(line):
                    for o in pend:
                        pend[o] = min(pend[o] + 1, _SATURATE)
                    if (b, li) in ordinal:
                        pend[ordinal[(b, li)]] = 0
                elif (w := _WAIT.match(line)):
                    n = int(w.group(1))
                    pend = {o: c for o, c in pend.items() if c < n}
                elif (m := _PHASE.match(line)):
                    p = int(m.group(1))
                    checked[0] += 1
                    late = sorted(o for o in pend if phase_of[o] <= p)
                    if late:
                        violations.append(f"phase {p}: DMA loads {late[:4]}{'...' if len(late) > 4 else ''} "
                                          f"(phases {sorted({phase_of[o] for o in late})}) may still be outstanding")
            return tuple(sorted(pend.items()))

        def join(a, b):
            da, db = dict(a), dict(b)
            for o, c in db.items():
                da[o] = min(da.get(o, c), c)
            return tuple(sorted(da.items()))

        fn.dataflow((), transfer, join)
        if violations:
            raise RuntimeError(f"{kern}: DMA wait does not retire the box read next: {violations[0]}")
        done[kern] = checked[0]
    return done


def check(asm: str, source: str = "", source_text: str = "") -> Dict[str, int]:
    """Both checks; ``source_text`` is the kernel source holding the DMA issue notes."""
    counts = max_outstanding_vmem(asm)
    bad = {k: v for k, v in counts.items() if v > VMCNT_MAX}
    if bad:
        raise RuntimeError(f"{source}: kernels keep more than {VMCNT_MAX} vector-memory operations in flight "
                           f"(vmcnt is 6 bits): {bad}")
    if source_text:
        issue = dma_issue_orders(source_text)
        if issue:
            try:
                check_dma_waits(asm, issue)
            except RuntimeError as e:
                raise RuntimeError(f"{source}: {e}") from None
    return counts
