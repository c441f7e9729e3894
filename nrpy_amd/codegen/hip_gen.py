"""Emit gfx950 HIP kernels for the BSSN and scalar-wave evolution path.

Kernels are generated from this repository's SymPy transcription
(``nrpy_amd.symbolic``) with FP64 arithmetic, no MFMA (stencil work is not a
dense contraction).  One thread owns one gridpoint; a 64-lane wavefront spans
64 consecutive xx0 points, so every stencil load of a wavefront is one
coalesced 512-byte row.

* Interior BSSN right-hand side (``rhs_eval``,
  ``nrpy/infrastructures/BHaH/general_relativity/rhs_eval.py:336-379``) and
  Ricci tensor (``Ricci_eval.py:95-142``) come as sets of candidate schedules
  (``candidate_tu``): output partitions, register budgets, CSE over expanded
  or staged algebra, shared upwind/KO line loads, software-pipelined stencil
  loads and buffer-resource addressing; an on-device autotuner picks one per
  workload.  Upwinded derivatives select one-sided stencils per direction from
  the sign of beta^i (``nrpy/c_codegen.py:1085-1150``).
* The RK4 substep update (``MoLtimestepping/tests/register_all_MoL_step_forward_in_time.c:4-103``)
  and det(gammabar) enforcement (``enforce_detgammabar_equals_detgammahat.py:84-98``)
  run as one streaming pass over all points after the boundary conditions.
* Ensemble builds also get persistent batched forms of every schedule and of
  the RK pass (``_ensemble_kernel``).
"""

from __future__ import annotations

import os
import re
from typing import Dict, List, Set, Tuple

import sympy as sp

from ..symbolic import bssn
from . import emit, fd, specs

RFM_EXPR = {
    # rfm slot -> C expression reading the 1D precompute arrays
    0: "rfm0[i0]",
    1: "rfm0[p.Nx0 + i0]",
    2: "rfm0[2 * p.Nx0 + i0]",
    3: "rfm0[3 * p.Nx0 + i0]",
    4: "rfm1[i1]",
    5: "rfm1[p.Nx1 + i1]",
    6: "rfm1[2 * p.Nx1 + i1]",
}
PARAM_EXPR = {
    "eta": "p.eta",
    "KreissOliger_strength_gauge": "p.KO_gauge",
    "KreissOliger_strength_nongauge": "p.KO_nongauge",
    "wavespeed": "p.wavespeed",
}


def _dbl(r: sp.Rational) -> str:
    if r.q == 1:
        return f"{int(r.p)}.0"
    return f"({int(r.p)}.0 / {int(r.q)}.0)"


class StencilEmitter:
    """C expressions for derivative symbols, reading ``in`` around ``idx``."""

    def __init__(self, spec: specs.Spec, fd_order: int, array: str = "in", u32: bool = False):
        self.spec = spec
        self.fd_order = fd_order
        self.array = array
        # u32: buffer loads through a resource descriptor of the array, with one 32-bit
        # per-lane byte offset shared by every load of the thread and the stencil and
        # gridfunction offsets as uniform (scalar-register) offsets -- no 64-bit
        # per-lane address arithmetic or address registers.  PAD keeps every uniform
        # offset non-negative.
        self.u32 = u32
        self.info: Dict[str, tuple] = {}  # buffer-load expression -> (array, gridfunction, offset)
        self.R = fd.ghost_zones(fd_order)
        self._collect: List[Tuple[str, str]] | None = None
        self._nload = 0

    def split(self, dv: fd.Deriv) -> Tuple[List[Tuple[str, str]], str]:
        """(loads as (register name, load expression), combine expression over those names)."""
        self._collect = []
        try:
            e = self.expr(dv)
            return self._collect, e
        finally:
            self._collect = None

    def load(self, gf: str, off: Tuple[int, int, int]) -> str:
        if self._collect is not None:
            self._collect, coll = None, self._collect
            try:
                name = f"ld{self._nload}"
                self._nload += 1
                coll.append((name, self.load(gf, off)))
                return name
            finally:
                self._collect = coll
        g = self.spec.gfs.index(gf)
        if self.u32:
            t = [f"{g} * Ntot + PAD"]
            if off[0]:
                t.append(f"{off[0]}")
            if off[1]:
                t.append(f"{off[1]} * s1")
            if off[2]:
                t.append(f"{off[2]} * s2")
            e = f"BLD(rsrc_{self.array}, {' + '.join(t)})"
            self.info[e] = (self.array, g, tuple(off))
            return e
        terms = []
        if off[0]:
            terms.append(f"{off[0]}")
        if off[1]:
            terms.append(f"{off[1]} * s1")
        if off[2]:
            terms.append(f"{off[2]} * s2")
        o = " + ".join(terms) if terms else "0"
        base = f"{g} * Ntot + idx" if g else "idx"
        if o == "0":
            return f"{self.array}[{base}]"
        return f"{self.array}[{base} + ({o})]"

    def _pt(self, d: int, k: int) -> Tuple[int, int, int]:
        o = [0, 0, 0]
        o[d] = k
        return tuple(o)

    def expr(self, dv: fd.Deriv) -> str:
        gf, op = dv.gf, dv.op
        inv = [f"p.invdxx{d}" for d in range(3)]
        if op == "dD":
            d = dv.dirs[0]
            w = dict(fd.weights_1d("dD", self.fd_order))
            parts = []
            for k in sorted(o for o in w if o > 0):
                parts.append(f"{_dbl(w[k])} * ({self.load(gf, self._pt(d, k))} - {self.load(gf, self._pt(d, -k))})")
            return f"{inv[d]} * ({' + '.join(parts)})"
        if op == "dKOD" or (op == "dDD" and dv.dirs[0] == dv.dirs[1]):
            d = dv.dirs[0]
            w = dict(fd.weights_1d("dKOD" if op == "dKOD" else "dDD", self.fd_order))
            parts = [f"{_dbl(w[0])} * {self.load(gf, (0, 0, 0))}"]
            for k in sorted(o for o in w if o > 0):
                parts.append(f"{_dbl(w[k])} * ({self.load(gf, self._pt(d, k))} + {self.load(gf, self._pt(d, -k))})")
            scale = inv[d] if op == "dKOD" else f"{inv[d]} * {inv[d]}"
            return f"{scale} * ({' + '.join(parts)})"
        if op == "dDD":
            d0, d1 = dv.dirs
            w = dict(fd.weights_1d("dD", self.fd_order))
            parts = []
            for a in sorted(o for o in w if o > 0):
                for b in sorted(o for o in w if o > 0):
                    def L(sa, sb):
                        o = [0, 0, 0]
                        o[d0] += sa * a
                        o[d1] += sb * b
                        return self.load(gf, tuple(o))
                    parts.append(
                        f"{_dbl(w[a] * w[b])} * (({L(1, 1)} - {L(1, -1)}) - ({L(-1, 1)} - {L(-1, -1)}))"
                    )
            return f"{inv[d0]} * {inv[d1]} * ({' + '.join(parts)})"
        if op == "dupD":
            d = dv.dirs[0]
            terms = []
            for k in range(-self.R, self.R + 1):
                terms.append(f"uw{d}[{k + self.R}] * {self.load(gf, self._pt(d, k))}")
            return f"{inv[d]} * ({' + '.join(terms)})"
        raise ValueError(op)

    def upwind_tables(self) -> str:
        """Per-direction weight vectors selected by the sign of beta^d."""
        up = dict(fd.weights_1d("dupD", self.fd_order))
        dn = dict(fd.weights_1d("ddnD", self.fd_order))
        lines = []
        for d in range(3):
            lines.append(f"  double uw{d}[{2 * self.R + 1}];")
            for k in range(-self.R, self.R + 1):
                a = _dbl(sp.Rational(up.get(k, 0)))
                b = _dbl(sp.Rational(dn.get(k, 0)))
                lines.append(f"  uw{d}[{k + self.R}] = (betaU{d} > 0.0) ? {a} : {b};")
        return "\n".join(lines)


# Per-point FP64 operation counts of the last emitted body of each spec
# (add/sub/mul/div = 1, fma counted as its mul + add), filled by _ordered_body.
FLOPS: Dict[str, int] = {}


def stencil_flops(d: fd.Deriv, fd_order: int) -> int:
    """Operations of one emitted stencil sum (pairs of +-k points grouped)."""
    R = fd.ghost_zones(fd_order)
    if d.op == "dD":
        P = len([o for o, _ in fd.weights_1d("dD", fd_order) if o > 0])
        return 3 * P
    if d.op == "dKOD" or (d.op == "dDD" and d.dirs[0] == d.dirs[1]):
        P = len([o for o, _ in fd.weights_1d("dKOD" if d.op == "dKOD" else "dDD", fd_order) if o > 0])
        return 3 * P + (1 if d.op == "dKOD" else 2)
    if d.op == "dDD":
        P = len([o for o, _ in fd.weights_1d("dD", fd_order) if o > 0])
        return 5 * P * P + 1
    if d.op == "dupD":
        return 4 * R + 2
    raise ValueError(d.op)


def _split_one(args):
    g, e, lin_names = args
    lin = [sp.Symbol(n, real=True) for n in lin_names]
    core = e.xreplace({s: 0 for s in lin})
    terms = []
    for s in lin:
        if s in e.free_symbols:
            c = sp.diff(e, s)
            if c.free_symbols & set(lin):
                raise ValueError(f"{s} enters {g} nonlinearly")
            terms.append((str(s), c))
    return g, core, terms


_SPLIT_CACHE: Dict[str, tuple] = {}


def split_linear(spec: specs.Spec):
    """Split each output into core + sum(coef * s) over upwind/KO derivatives s.

    Every ``dupD`` and ``dKOD`` symbol enters the BSSN right-hand sides
    linearly (advection beta^k d_k X and dissipation eps ReU_k KO_k X), with a
    coefficient that depends only on the shift and the reference metric.
    Returns (cores, {output_gf: [(deriv, coef), ...]}).
    """
    if spec.name in _SPLIT_CACHE:
        return _SPLIT_CACHE[spec.name]
    _, _, derivs, _, _ = spec.classify()
    dmap = {d.symbol: d for d in derivs if d.op in ("dupD", "dKOD")}
    names = sorted(dmap)
    jobs = [(g, e, names) for g, e in spec.outputs]
    if names:
        import multiprocessing as mp

        with mp.get_context("fork").Pool(min(len(jobs), os.cpu_count() or 1)) as pool:
            res = pool.map(_split_one, jobs)
    else:
        res = [(g, e, []) for g, e, _ in jobs]
    cores, terms = [], {}
    for g, core, tl in res:
        cores.append(core)
        if tl:
            terms[g] = [(dmap[n], c) for n, c in tl]
    _SPLIT_CACHE[spec.name] = (cores, terms)
    return cores, terms


def _emit_defs(spec: specs.Spec, fd_order: int, u32: bool = False):
    centre, aux, derivs, rfm, params = spec.classify()
    st = StencilEmitter(spec, fd_order, u32=u32)
    defs: Dict[str, str] = {}
    for n in centre:
        defs[n] = st.load(n, (0, 0, 0))
    for n in aux:
        defs[n] = (f"BLD(rsrc_aux, {spec.aux.index(n)} * Ntot + PAD)" if u32
                   else f"aux[{spec.aux.index(n)} * Ntot + idx]")
    for d in derivs:
        defs[d.symbol] = st.expr(d)
    for n in rfm:
        defs[n] = RFM_EXPR[specs.RFM_SYMBOLS.index(n)]
    for n in params:
        defs[n] = PARAM_EXPR[n]
    return defs


def _ordered_body(spec: specs.Spec, fd_order: int, out_lvalues: List[str], rationals: Dict[str, str],
                  stream_linear: bool = True, order: str = "dfs", barriers: bool = True,
                  cse_scope: str = "joint", linear: str = "stream", prefetch: int = 0,
                  ahead: int = 12, u32: bool = False, chunk: int = 0, wide: bool = False,
                  defs_override: Dict[str, str] | None = None) -> str:
    """Stencils and CSE'd algebra for one pointwise spec.

    With ``stream_linear``, upwind and KO derivatives are consumed as soon as
    they are formed (accumulated into per-output registers), and only the
    remaining core algebra is CSE'd; centre values, centred derivatives and
    CSE temporaries are then emitted in first-use order.  Scheduling barriers
    after each stencil keep the compiler from hoisting every load to the top.

    ``linear="lines"`` evaluates an output's upwind and KO terms just before
    the output is stored, one 1D line of 2R points per (gridfunction,
    direction) loaded once and shared by the two one-sided upwind sums and the
    KO sum (the upwind stencils and the KO stencil cover the same line).

    ``prefetch=B`` issues each stencil's loads up to ``ahead`` barrier regions
    before the region that combines them, in consumption order, keeping at
    most B early-issued values in flight, so that load latency overlaps the
    arithmetic of earlier regions instead of stalling the single wave per SIMD.

    ``defs_override`` maps symbol names to C expressions that replace their
    loads or stencils (a caller that has already formed the derivatives).
    """
    defs = _emit_defs(spec, fd_order, u32)
    if defs_override:
        defs.update(defs_override)
    derivs = {d.symbol: d for d in spec.classify()[2]}
    centre_names = set(spec.classify()[0]) | set(spec.classify()[1])
    pr = emit.DoublePrinter(rationals)
    st = StencilEmitter(spec, fd_order, u32=u32)
    emitted: Set[str] = set()
    items: List[tuple] = []  # ("C", text) | ("L", text, group) | ("B",)
    flops = [0]
    gid = [0, 0]  # load-group counter, line counter
    split_loads = prefetch > 0
    wide = wide and u32 and split_loads

    def code(t):
        items.append(("C", t))

    def barrier():
        if barriers:
            items.append(("B",))

    def loads(pairs):
        gid[0] += 1
        if wide:
            # two values adjacent along xx0 (same array, gridfunction, xx1/xx2 offset) come
            # from one 16-byte buffer load
            info = {name: st.info.get(e) for name, e in pairs}
            by_key = {}
            for name, e in pairs:
                if info[name]:
                    a, g, off = info[name]
                    by_key[(a, g, off[1], off[2], off[0])] = name
            used = set()
            for name, e in pairs:
                if name in used:
                    continue
                inf = info[name]
                if inf:
                    a, g, off = inf
                    mate = by_key.get((a, g, off[1], off[2], off[0] + 1))
                    if mate and mate not in used:
                        used |= {name, mate}
                        so = e[e.index(",") + 1:-1].strip()
                        pr = f"w{name}"
                        items.append(("L", f"  const auto {pr} = __builtin_amdgcn_raw_buffer_load_b128(rsrc_{a}, vo, "
                                           f"(unsigned)(({so}) * 8), 0);", gid[0]))
                        items.append(("L", f"  const double {name} = BHAH_LO({pr});", gid[0]))
                        items.append(("L", f"  const double {mate} = BHAH_HI({pr});", gid[0]))
                        continue
                used.add(name)
                items.append(("L", f"  const double {name} = {e};", gid[0]))
            return
        for name, e in pairs:
            items.append(("L", f"  const double {name} = {e};", gid[0]))

    def need(expr, indent="  "):
        for sy in sorted(expr.free_symbols, key=str):
            n = str(sy)
            if n in defs and n not in emitted and linear == "fused" and n in derivs:
                dv = derivs[n]
                if dv.op in ("dD", "dDD") and len(set(dv.dirs)) == 1:
                    key = (dv.gf, dv.dirs[0])
                    if key in line_out:
                        if key not in lines_done:
                            emit_line(key)
                    elif key not in centred_done and dv.gf in spec.gfs:
                        emit_centred_line(key)
            if n in defs and n not in emitted:
                emitted.add(n)
                if n in derivs:
                    flops[0] += stencil_flops(derivs[n], fd_order)
                    if split_loads:
                        ld, comb = st.split(derivs[n])
                        loads(ld)
                        code(f"{indent}const double {n} = {comb};")
                    else:
                        code(f"{indent}const double {n} = {defs[n]};")
                    barrier()
                elif split_loads and n in centre_names:
                    loads([(n, defs[n])])
                else:
                    code(f"{indent}const double {n} = {defs[n]};")

    outs = [e for _, e in spec.outputs]
    acc_names: Dict[str, str] = {}
    lin_terms: Dict[str, list] = {}
    cname: Dict = {}
    if stream_linear:
        cores, terms = split_linear(spec)
        if terms and linear == "lds":
            # upwind/KO contributions were accumulated by the caller (workgroup-tiled kernels)
            for g in terms:
                acc_names[g] = f"VPT_lin_{g}"
        elif terms:
            coefs = sorted({c for tl in terms.values() for _, c in tl}, key=sp.default_sort_key)
            crep, cred = sp.cse(coefs, symbols=sp.numbered_symbols("lc"), order="none")
            for sy, e in crep:
                need(e)
                code(f"  const double {sy} = {pr.doprint(e)};")
            for k, (c, e) in enumerate(zip(coefs, cred)):
                need(e)
                code(f"  const double lcoef{k} = {pr.doprint(e)};")
                cname[c] = f"lcoef{k}"
            flops[0] += sum(emit._flops(e) for _, e in crep) + sum(emit._flops(e) for e in cred)
            if linear == "stream":
                for g, tl in terms.items():
                    acc = f"lin_{g}"
                    acc_names[g] = acc
                    code(f"  double {acc} = 0.0;")
                    for d, c in sorted(tl, key=lambda t: t[0].symbol):
                        code(f"  {acc} += {cname[c]} * ({defs[d.symbol]});")
                        flops[0] += 2 + stencil_flops(d, fd_order)
                    if barriers:
                        items.append(("B",))
            else:
                lin_terms = terms
        outs = cores

    R = fd.ghost_zones(fd_order)
    w_up = dict(fd.weights_1d("dupD", fd_order))
    w_dn = dict(fd.weights_1d("ddnD", fd_order))
    w_ko = dict(fd.weights_1d("dKOD", fd_order))

    w_dD = dict(fd.weights_1d("dD", fd_order))
    w_dDD = dict(fd.weights_1d("dDD", fd_order))
    line_cf: Dict[Tuple[str, int], Dict[str, str]] = {}
    line_out: Dict[Tuple[str, int], str] = {}
    for g_out, tl in lin_terms.items():
        for d, c in terms_sorted(tl):
            key = (d.gf, d.dirs[0])
            line_cf.setdefault(key, {})["up" if d.op == "dupD" else "ko"] = cname[c]
            line_out[key] = g_out
    lines_done: Set[Tuple[str, int]] = set()

    def emit_line(key):
        """Load one (gridfunction, direction) line of 2R+1 points once; form the output's
        upwind/KO contribution and, with linear == "fused", every centred derivative of that
        gridfunction along that direction the kernel uses."""
        lines_done.add(key)
        gf, dirn = key
        g = line_out[key]
        cf = line_cf[key]
        gid[1] += 1
        if gf in defs:
            need(sp.Symbol(gf, real=True))
            f0 = gf
        else:
            f0 = f"ctr{gid[1]}_{gf}"
            if split_loads:
                loads([(f0, st.load(gf, (0, 0, 0)))])
            else:
                code(f"  const double {f0} = {st.load(gf, (0, 0, 0))};")
        pts = {}
        ld = []
        for k in range(-R, R + 1):
            if k == 0:
                continue
            o = [0, 0, 0]
            o[dirn] = k
            name = f"ln{gid[1]}_{k + R}"
            ld.append((name, st.load(gf, tuple(o))))
            pts[k] = name
        pts[0] = f0
        if split_loads:
            loads(ld)
        else:
            for name, e in ld:
                code(f"  const double {name} = {e};")
        if linear == "fused":
            for n, dv in sorted(derivs.items()):
                if n in emitted or dv.gf != gf:
                    continue
                if dv.op == "dD" and dv.dirs[0] == dirn:
                    parts = [f"{_dbl(w_dD[k])} * ({pts[k]} - {pts[-k]})" for k in sorted(w_dD) if k > 0]
                    code(f"  const double {n} = p.invdxx{dirn} * ({' + '.join(parts)});")
                elif dv.op == "dDD" and dv.dirs == (dirn, dirn):
                    parts = [f"{_dbl(w_dDD[0])} * {pts[0]}"]
                    parts += [f"{_dbl(w_dDD[k])} * ({pts[k]} + {pts[-k]})" for k in sorted(w_dDD) if k > 0]
                    code(f"  const double {n} = p.invdxx{dirn} * p.invdxx{dirn} * ({' + '.join(parts)});")
                else:
                    continue
                emitted.add(n)
                flops[0] += stencil_flops(dv, fd_order)
        acc = f"lin_{g}"
        if g not in acc_names:
            acc_names[g] = acc
            code(f"  double {acc} = 0.0;")
        terms_c = []
        if "up" in cf:
            # both one-sided sums, then a select (no divergent branch)
            su = " + ".join(f"{_dbl(w)} * {pts[k]}" for k, w in sorted(w_up.items()))
            sd = " + ".join(f"{_dbl(w)} * {pts[k]}" for k, w in sorted(w_dn.items()))
            code(f"  const double su{gid[1]} = {su};")
            code(f"  const double sd{gid[1]} = {sd};")
            terms_c.append(f"{cf['up']} * (p.invdxx{dirn} * ((betaU{dirn} > 0.0) ? su{gid[1]} : sd{gid[1]}))")
            flops[0] += 2 * (2 * len(w_up) - 1) + 2
        if "ko" in cf:
            sk = [f"{_dbl(w_ko[0])} * {pts[0]}"]
            sk += [f"{_dbl(w_ko[k])} * ({pts[k]} + {pts[-k]})" for k in sorted(w_ko) if k > 0]
            terms_c.append(f"{cf['ko']} * (p.invdxx{dirn} * ({' + '.join(sk)}))")
            flops[0] += 3 * (len(w_ko) // 2) + 3
        code(f"  {acc} += {' + '.join(terms_c)};")
        barrier()

    centred_done: Set[Tuple[str, int]] = set()

    def emit_centred_line(key):
        """Load the 2*(fd/2)+1 centred points of a (gridfunction, direction) line once and
        form both its first and its unmixed second derivative where the kernel uses them."""
        centred_done.add(key)
        gf, dirn = key
        want = [(n, dv) for n, dv in sorted(derivs.items()) if n not in emitted and dv.gf == gf and
                ((dv.op == "dD" and dv.dirs[0] == dirn) or (dv.op == "dDD" and dv.dirs == (dirn, dirn)))]
        if len(want) < 2:
            return False
        gid[1] += 1
        pts = {0: gf}
        if any(dv.op == "dDD" for _, dv in want):
            if gf in defs:
                need(sp.Symbol(gf, real=True))
            else:
                pts[0] = f"clc{gid[1]}"
                code(f"  const double {pts[0]} = {st.load(gf, (0, 0, 0))};")
        ld = []
        for k in sorted(w_dD):
            o = [0, 0, 0]
            o[dirn] = k
            name = f"cl{gid[1]}_{k + R}"
            ld.append((name, st.load(gf, tuple(o))))
            pts[k] = name
        if split_loads:
            loads(ld)
        else:
            for name, e in ld:
                code(f"  const double {name} = {e};")
        for n, dv in want:
            if dv.op == "dD":
                parts = [f"{_dbl(w_dD[k])} * ({pts[k]} - {pts[-k]})" for k in sorted(w_dD) if k > 0]
                code(f"  const double {n} = p.invdxx{dirn} * ({' + '.join(parts)});")
            else:
                parts = [f"{_dbl(w_dDD[0])} * {pts[0]}"]
                parts += [f"{_dbl(w_dDD[k])} * ({pts[k]} + {pts[-k]})" for k in sorted(w_dDD) if k > 0]
                code(f"  const double {n} = p.invdxx{dirn} * p.invdxx{dirn} * ({' + '.join(parts)});")
            emitted.add(n)
            flops[0] += stencil_flops(dv, fd_order)
        barrier()
        return True

    def emit_lines(g):
        """All remaining lines of output g, just before it is stored."""
        for key in sorted(k for k, o in line_out.items() if o == g and k not in lines_done):
            emit_line(key)
        if g not in acc_names:
            acc_names[g] = f"lin_{g}"
            code(f"  double lin_{g} = 0.0;")

    sdefs = spec.used_defs()
    if sdefs and cse_scope == "local":
        # CSE within each staged definition and each output separately: temporaries
        # then live only inside one stage, at the cost of recomputing cheap
        # reference-metric products shared between stages.
        tsyms = sp.numbered_symbols("t")
        replaced, reduced_defs = [], []
        for s_, e_ in sdefs:
            r_, d_ = sp.cse([e_], symbols=tsyms, order="none")
            replaced += r_
            reduced_defs.append((s_, d_[0]))
        reduced = []
        for e_ in outs:
            r_, d_ = sp.cse([e_], symbols=tsyms, order="none")
            replaced += r_
            reduced.append(d_[0])
        replaced += reduced_defs
        order = "dfs"
    else:
        replaced, reduced = sp.cse([e for _, e in sdefs] + list(outs), symbols=sp.numbered_symbols("t"), order="none")
        staged_defs = [(s_, e_) for (s_, _), e_ in zip(sdefs, reduced[: len(sdefs)])]
        reduced = reduced[len(sdefs):]
        if staged_defs:
            replaced = list(replaced) + staged_defs
            order = "dfs"
    if order == "cse":
        for sy, e in replaced:
            need(e)
            code(f"  const double {sy} = {pr.doprint(e)};")
            flops[0] += emit._flops(e)
    else:
        # Demand-driven depth-first emission: each output pulls in the temporaries it
        # needs just before it is stored, so values used by one output die early.
        tdef = {sy: e for sy, e in replaced}
        done: Set = set()

        def emit_t(sym):
            stack = [(sym, False)]
            while stack:
                t, expanded = stack.pop()
                if t in done:
                    continue
                if expanded:
                    done.add(t)
                    need(tdef[t])
                    code(f"  const double {t} = {pr.doprint(tdef[t])};")
                    flops[0] += emit._flops(tdef[t])
                    continue
                stack.append((t, True))
                for d in sorted((x for x in tdef[t].free_symbols if x in tdef and x not in done), key=str, reverse=True):
                    stack.append((d, False))

    for (g, _), lv, e in zip(spec.outputs, out_lvalues, reduced):
        if g in lin_terms:
            emit_lines(g)
        if order != "cse":
            for d in sorted((x for x in e.free_symbols if x in tdef), key=str):
                emit_t(d)
        need(e)
        tail = f" + {acc_names[g]}" if g in acc_names else ""
        code(f"  {lv} = {pr.doprint(e)}{tail};")
        flops[0] += emit._flops(e) + (1 if tail else 0)
    FLOPS[f"{spec.name}_fd{fd_order}{_mode_suffix(linear, prefetch)}"] = flops[0]
    if prefetch:
        items = _schedule_prefetch(items, prefetch, ahead, chunk)
    out = []
    for it in items:
        out.append("  __builtin_amdgcn_sched_barrier(0);" if it[0] == "B" else it[1])
    return "\n".join(out)


def terms_sorted(tl):
    return sorted(tl, key=lambda t: t[0].symbol)


def _mode_suffix(linear: str, prefetch: int) -> str:
    return ("" if linear == "stream" else f"_{linear}") + (f"_pf{prefetch}" if prefetch else "")


def _schedule_prefetch(items: List[tuple], budget: int, ahead: int, chunk: int = 0) -> List[tuple]:
    """Move each group of loads to the earliest region allowed by the budget.

    Regions are the stretches between scheduling barriers.  Groups keep their
    consumption order (so a wait for one group never waits for a later one);
    a group consumed in region c may start in region s >= c - ahead when every
    region in [s, c) holds fewer than ``budget`` early values including it.
    With ``chunk``, groups are first cut into pieces of at most that many loads,
    so stencils wider than the budget (mixed second derivatives) are issued
    early piece by piece instead of all at their use.
    """
    regions: List[List[tuple]] = [[]]
    for it in items:
        if it[0] == "B":
            regions.append([])
        else:
            regions[-1].append(it)
    nreg = len(regions)
    inflight = [0] * nreg
    early: List[List[tuple]] = [[] for _ in range(nreg)]
    keep: List[List[tuple]] = [[] for _ in range(nreg)]
    own: List[List[tuple]] = [[] for _ in range(nreg)]
    prev_s = 0
    for c, reg in enumerate(regions):
        groups: Dict[int, List[tuple]] = {}
        order_g: List[int] = []
        for it in reg:
            if it[0] == "L":
                if it[2] not in groups:
                    groups[it[2]] = []
                    order_g.append(it[2])
                groups[it[2]].append(it)
            else:
                keep[c].append(it)
        # a region's own loads are issued first, then the early loads of later
        # regions, then its code: issue order equals consumption order
        placed_here: List[tuple] = []
        if chunk:
            pieces: Dict[tuple, List[tuple]] = {}
            order_p: List[tuple] = []
            for g in order_g:
                for k in range(0, len(groups[g]), chunk):
                    pieces[(g, k)] = groups[g][k:k + chunk]
                    order_p.append((g, k))
            groups, order_g = pieces, order_p
        for g in order_g:
            n = len(groups[g])
            sreg = c
            while sreg - 1 >= max(prev_s, c - ahead) and inflight[sreg - 1] + n <= budget:
                sreg -= 1
            for r in range(sreg, c):
                inflight[r] += n
            prev_s = sreg
            if sreg == c:
                placed_here.extend(groups[g])
            else:
                early[sreg].extend(groups[g])
        own[c] = placed_here
    out: List[tuple] = []
    for r in range(nreg):
        if r:
            out.append(("B",))
        out.extend(own[r])
        out.extend(early[r])
        out.extend(keep[r])
    return out


KPARAMS = """
#include "../bhah_device_types.h"
"""


XCD_REMAP = """  // Workgroups are dispatched round-robin over the 8 XCDs (each with its own L2):
  // renumber them so XCD k sweeps the k-th contiguous eighth of the tiles and
  // stencil neighbours share an L2.  Identity when the tile count is not a multiple of 8.
  unsigned bx = blockIdx.x, by = blockIdx.y, bz = blockIdx.z;
  {
    const unsigned gx = gridDim.x, gy = gridDim.y, n = gx * gy * gridDim.z;
    if ((n & 7u) == 0u) {
      const unsigned h = bx + gx * (by + gy * bz);
      const unsigned l = (h & 7u) * (n >> 3) + (h >> 3);
      bx = l % gx;
      by = (l / gx) % gy;
      bz = l / (gx * gy);
    }
  }
"""


U32_LINES = """  const long PAD = (long)p.NG * (1 + s1 + s2);
  const unsigned vo = (unsigned)((idx - PAD) * 8);
  const unsigned uix = (unsigned)(idx - PAD);
  const __amdgpu_buffer_rsrc_t rsrc_in = __builtin_amdgcn_make_buffer_rsrc((void *)in, 0, (int)(unsigned)(24 * Ntot * 8), 0x00020000);
  const __amdgpu_buffer_rsrc_t rsrc_aux = __builtin_amdgcn_make_buffer_rsrc((void *)aux, 0, (int)(unsigned)(6 * Ntot * 8), 0x00020000);
"""
# Raw buffer load of one double: per-lane byte offset vo, uniform element offset so.
BLD_MACRO = """#define BLD(r, so) __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64((r), vo, (unsigned)((so) * 8), 0))
#define BHAH_LO(v) __builtin_bit_cast(double, ((unsigned long long)(v)[1] << 32) | (unsigned long long)(v)[0])
#define BHAH_HI(v) __builtin_bit_cast(double, ((unsigned long long)(v)[3] << 32) | (unsigned long long)(v)[2])
"""


def _prologue_at(interior: bool, u32: bool = False) -> str:
    """Prologue of a per-tile device function: tile coordinates come in as bx, by, bz."""
    lo = "p.NG" if interior else "0"
    return f"""  const int i0 = {lo} + (int)(bx * blockDim.x + threadIdx.x);
  const int i1 = {lo} + (int)(by * blockDim.y + threadIdx.y);
  const int i2 = {lo} + bz;
  if (i0 >= p.Nx0 - {lo} || i1 >= p.Nx1 - {lo} || i2 >= p.Nx2 - {lo}) return;
  const long s1 = p.s1, s2 = p.s2, Ntot = p.Ntot;
  const long idx = i0 + s1 * i1 + s2 * i2;
{U32_LINES if u32 else ""}"""


ENS_SIG = ("dim3 g, dim3 b, hipStream_t s, const bhah_member *mem, const bhah_item *items, int nitems, "
           "int in_sel, int aux_sel, int out_sel")


def _ensemble_kernel(kname: str, upw: str, body: str, lb: str, u32: bool = False) -> str:
    """Persistent batched form of one interior kernel: a fixed grid of workgroups
    walks the ensemble's tile list; each tile runs the same per-point code on its
    member's arrays (selected by index from the member table)."""
    return f"""
static __device__ __forceinline__ void {kname}_tile(const double *__restrict__ in, const double *__restrict__ aux,
        double *__restrict__ out, const double *__restrict__ rfm0, const double *__restrict__ rfm1,
        const bhah_kparams p, const int bx, const int by, const int bz) {{
{_prologue_at(True, u32)}
{upw}
{body}
}}
extern "C" __global__ void __launch_bounds__({lb})
{kname}(const bhah_member *__restrict__ mem, const bhah_item *__restrict__ items, const int nitems,
        const int in_sel, const int aux_sel, const int out_sel) {{
  for (int it = blockIdx.x; it < nitems; it += gridDim.x) {{
    const bhah_item w = items[it];
    const bhah_member *M = &mem[w.m];
    {kname}_tile(M->arr[in_sel], M->arr[aux_sel], M->arr[out_sel], M->rfm0, M->rfm1, M->kp, w.bx, w.by, w.bz);
  }}
}}
"""


def _zsweep_remap(sz: int) -> str:
    """Workgroup order for L2 reuse along xx2: give XCD k the k-th contiguous eighth of a
    logical tile order in which xx2 runs fastest within slabs of ``sz`` planes, so the
    workgroups resident on one XCD at a time cover a few (xx0, xx1) tiles over ``sz``
    consecutive planes and share their stencil neighbour planes in that XCD's L2.
    Identity when the grid does not divide evenly."""
    return f"""  unsigned bx = blockIdx.x, by = blockIdx.y, bz = blockIdx.z;
  {{
    const unsigned gx = gridDim.x, gy = gridDim.y, gz = gridDim.z, n = gx * gy * gz;
    if ((n & 7u) == 0u && gz % {sz}u == 0u) {{
      const unsigned h = bx + gx * (by + gy * bz);
      const unsigned l = (h & 7u) * (n >> 3) + (h >> 3);
      const unsigned slab = gx * gy * {sz}u;
      const unsigned zb = l / slab, rem = l - zb * slab;
      const unsigned t = rem / {sz}u;
      bz = zb * {sz}u + (rem - t * {sz}u);
      bx = t % gx;
      by = t / gx;
    }}
  }}
"""


def _prologue(interior: bool, xcd: bool = False, u32: bool = False, zs: int = 0) -> str:
    lo = "p.NG" if interior else "0"
    tail = U32_LINES if u32 else ""
    if zs:
        return _zsweep_remap(zs) + f"""  const int i0 = {lo} + (int)(bx * blockDim.x + threadIdx.x);
  const int i1 = {lo} + (int)(by * blockDim.y + threadIdx.y);
  const int i2 = {lo} + (int)bz;
  if (i0 >= p.Nx0 - {lo} || i1 >= p.Nx1 - {lo}) return;
  const long s1 = p.s1, s2 = p.s2, Ntot = p.Ntot;
  const long idx = i0 + s1 * i1 + s2 * i2;
""" + tail
    if xcd:
        return XCD_REMAP + f"""  const int i0 = {lo} + (int)(bx * blockDim.x + threadIdx.x);
  const int i1 = {lo} + (int)(by * blockDim.y + threadIdx.y);
  const int i2 = {lo} + (int)bz;
  if (i0 >= p.Nx0 - {lo} || i1 >= p.Nx1 - {lo}) return;
  const long s1 = p.s1, s2 = p.s2, Ntot = p.Ntot;
  const long idx = i0 + s1 * i1 + s2 * i2;
""" + tail
    return f"""  const int i0 = {lo} + (int)(blockIdx.x * blockDim.x + threadIdx.x);
  const int i1 = {lo} + (int)(blockIdx.y * blockDim.y + threadIdx.y);
  const int i2 = {lo} + (int)blockIdx.z;
  if (i0 >= p.Nx0 - {lo} || i1 >= p.Nx1 - {lo}) return;
  const long s1 = p.s1, s2 = p.s2, Ntot = p.Ntot;
  const long idx = i0 + s1 * i1 + s2 * i2;
""" + tail


def _upwind_prologue(spec: specs.Spec, fd_order: int, rationals, tables: bool = True, u32: bool = False) -> str:
    st = StencilEmitter(spec, fd_order, u32=u32)
    pr = emit.DoublePrinter(rationals)
    lines = []
    used = emit.free_symbol_names(spec.upwind)
    for n in sorted(used):
        if n in spec.gfs:
            lines.append(f"  const double up_{n} = {st.load(n, (0, 0, 0))};")
        elif n in specs.RFM_SYMBOLS:
            lines.append(f"  const double up_{n} = {RFM_EXPR[specs.RFM_SYMBOLS.index(n)]};")
    sub = {sp.Symbol(n, real=True): sp.Symbol("up_" + n, real=True) for n in used}
    for d in range(3):
        lines.append(f"  const double betaU{d} = {pr.doprint(sp.sympify(spec.upwind[d]).xreplace(sub))};")
    if tables:
        lines.append(st.upwind_tables())
    return "\n".join(lines)


def detgamma_inline(v: specs.Variant, rationals, src: str) -> str:
    """Inline det(gammabar) enforcement on hDD values held in ``nxt[]``.

    ``src`` names the array of 24 next-stage values (EVOL order).
    """
    spec = specs.detgamma_spec(v)
    pr = emit.DoublePrinter(rationals)
    centre, _, _, rfm, _ = spec.classify()
    lines = ["  {"]
    for n in centre:
        lines.append(f"    const double {n} = {src}[{bssn.EVOL_GFS.index(n)}];")
    for n in rfm:
        lines.append(f"    const double {n} = {RFM_EXPR[specs.RFM_SYMBOLS.index(n)]};")
    replaced, reduced = sp.cse([e for _, e in spec.outputs], symbols=sp.numbered_symbols("dg"), order="none")
    for s, e in replaced:
        lines.append(f"    const double {s} = {pr.doprint(e)};")
    outs = []
    for (gf, _), e in zip(spec.outputs, reduced):
        lines.append(f"    const double new_{gf} = {pr.doprint(e)};")
        outs.append(gf)
    for gf in outs:
        lines.append(f"    {src}[{bssn.EVOL_GFS.index(gf)}] = new_{gf};")
    lines.append("  }")
    return "\n".join(lines)


RK_EPILOGUE = r"""
  // RK4 substep update (register_all_MoL_step_forward_in_time.c:4-103), in registers.
  const double dt = p.dt;
  double nxt[NGF];
  if (p.substep == 1) {
#pragma unroll
    for (int g = 0; g < NGF; g++) {
      const double k = kk[g];
      y_run[g * Ntot + idx] = (1.0 / 6.0) * dt * k;
      nxt[g] = (1.0 / 2.0) * dt * k + ctr[g];
    }
  } else if (p.substep == 4) {
#pragma unroll
    for (int g = 0; g < NGF; g++) {
      const double k = kk[g];
      nxt[g] = (1.0 / 6.0) * dt * k + y_n[g * Ntot + idx] + y_run[g * Ntot + idx];
    }
  } else {
    const double a = (p.substep == 2) ? (1.0 / 2.0) : 1.0;
#pragma unroll
    for (int g = 0; g < NGF; g++) {
      const double k = kk[g];
      y_run[g * Ntot + idx] = (1.0 / 3.0) * dt * k + y_run[g * Ntot + idx];
      nxt[g] = a * dt * k + y_n[g * Ntot + idx];
    }
  }
"""


# Output partitions: each group becomes one kernel with its own CSE, trading
# recomputation of shared subexpressions for a smaller live register set.
_A = ["aDD00", "aDD01", "aDD02", "aDD11", "aDD12", "aDD22"]
_H = ["hDD00", "hDD01", "hDD02", "hDD11", "hDD12", "hDD22"]
_L = ["lambdaU0", "lambdaU1", "lambdaU2"]
_B = ["betU0", "betU1", "betU2"]
_V = ["vetU0", "vetU1", "vetU2"]
RHS_PARTITIONS = {
    "p1": [list(bssn.EVOL_GFS)],
    "p2": [_A + ["alpha", "cf"] + _H + ["trK"], _L + _B + _V],
    "p3": [_H + ["cf", "alpha"] + _V, _A + ["trK"], _L + _B],
    "p4": [_H + ["cf", "alpha"] + _V, _A, ["trK"], _L + _B],
    "p5": [_H + ["cf"], ["alpha"] + _V, _A, ["trK"], _L + _B],
}
RICCI_PARTITIONS = {
    "r1": [list(bssn.AUXEVOL_GFS)],
    "r2": [["RbarDD00", "RbarDD01", "RbarDD02"], ["RbarDD11", "RbarDD12", "RbarDD22"]],
    "r3": [["RbarDD00", "RbarDD11"], ["RbarDD01", "RbarDD12"], ["RbarDD02", "RbarDD22"]],
    "r6": [[g] for g in bssn.AUXEVOL_GFS],
}
LAUNCH_BOUNDS = {"lb1": "256", "lb2": "256, 2"}
# "lbm": two waves per SIMD for every group but the last (the heaviest: the Lambdabar^i and
# B^i right-hand sides, or the last Ricci group), which keeps one wave and the full register file.


def _group_bounds(lb: str, k: int, ngroups: int) -> str:
    if lb == "lbm":
        return LAUNCH_BOUNDS["lb1" if k == ngroups - 1 else "lb2"]
    return LAUNCH_BOUNDS[lb]

# Candidate schedules per build: (kind, partition, launch bounds).
# Candidate schedules built per workload: the plain one-thread-per-point schedule first (the
# fallback for input that is not 16-byte aligned), then the tuned workgroup-tiled schedules
# and their nearest alternatives.  Schedules the autotuner measured slower on every workload
# are no longer built (their timings stay in the tuned tables under nrpy_amd/autotune/).
FULL_CANDIDATES = {
    "rhs": ["p1_lb1", "p1_lb1_sj_L16x4x4x1_d3_dma_rb1_io_pg_ub_mi_xcd", "p1_lb1_sj_L16x4x4x1_d3_dma_rb1_io_pg_ub_sa_lr40_di_mi_xcd_blk",
            "p1_lb1_sj_L16x4x4x1_d4_dma_rb1_io_pg_ub_sa_lr40_di_mi_xcd_blk"],
    "ricci": ["r1_lb1", "r1_lb1_sj_L16x4x4x1_d3_dma_rb1_io_pg2_mi_xcd", "r1_lb1_sj_L16x4x4x1_d3_dma_rb1_io_pg_lr40_mi_xcd"],
}
SMALL_CANDIDATES = {"rhs": ["p1_lb1", "p1_lb1_sl_lf_pf16_u32_w2", "p1_lb1_sj_L16x8x2x1_d3_dma_rb1",
                            "p1_lb1_sj_L16x8x2x1_d3_dma_rb1_pg8", "p1_lb1_sj_L16x8x2x1_d3_dma_rb1_pg8_ub",
                            "p1_lb1_sj_L16x8x2x1_d3_dma_rb1_pg8_ub_mi"],
                    "ricci": ["r1_lb1", "r1_lb1_pf16_u32", "r1_lb1_sj_L16x4x4x1_d2_dma_xcd",
                              "r1_lb1_sj_L16x8x2x1_d3_dma_rb1", "r1_lb1_sj_L16x8x2x1_d3_dma_rb1_mi"]}

CAND_SIG = ("dim3 g, dim3 b, hipStream_t s, const double *in, const double *aux, double *out, "
            "const double *rfm0, const double *rfm1, bhah_kparams p")


def _subspec(spec: specs.Spec, name: str, gfs: List[str]) -> specs.Spec:
    return specs.Spec(name, [(g, e) for g, e in spec.outputs if g in gfs], spec.gfs, spec.aux, spec.upwind,
                      defs=spec.defs)


# An empty register-constrained asm: the value must exist in a VGPR at this point.
LDS_PIN_MACRO = """#ifndef BHAH_PIN
#define BHAH_PIN(x) asm volatile("" : "+v"(x))
// several values at once: BHAH_PINS("+v"(a), "+v"(b), ...)
#define BHAH_PINS(...) asm volatile("" : __VA_ARGS__)
#endif
"""


class _LdsEmitter(StencilEmitter):
    """Stencil expressions reading a workgroup's LDS box ``sh`` around box index ``c``."""

    def __init__(self, spec: specs.Spec, fd_order: int, centre: str, BX: int, BY: int, base: int = 0):
        super().__init__(spec, fd_order)
        self.centre, self.BX, self.BY, self.base = centre, BX, BY, base

    def load(self, gf: str, off: Tuple[int, int, int]) -> str:
        o = self.base + off[0] + self.BX * off[1] + self.BX * self.BY * off[2]
        return f"sh[{self.centre}{f' + {o}' if o > 0 else (f' - {-o}' if o < 0 else '')}]"


def _lds_tile_order(xcd: bool, zfirst: bool, blocked: bool = False) -> str:
    """Tile of this workgroup.  Workgroups are dispatched in linear order, round-robin
    over the 8 XCDs; with ``xcd`` XCD k takes the k-th contiguous eighth of the tiles
    (neighbouring tiles, whose halo boxes overlap, then share an L2), and with
    ``zfirst`` consecutive tiles run along xx2 first (their boxes overlap in 2H of
    TZ + 2H planes).  With ``blocked`` each run of 32 consecutive tiles (an XCD's 32 CUs
    hold one workgroup each) is a 2 x 4 x 4-tile block, so the tiles in flight together on
    an XCD share their halo planes in every direction instead of along xx0 and xx1 only;
    grids that do not divide into such blocks keep the linear order."""
    if not (xcd or zfirst or blocked):
        return "  const unsigned bx = blockIdx.x, by = blockIdx.y, bz = blockIdx.z;"
    lin = ("((n & 7u) == 0u) ? (h & 7u) * (n >> 3) + (h >> 3) : h" if xcd else "h")
    dec = ("bz = l % gz;\n    bx = (l / gz) % gx;\n    by = l / (gz * gx);" if zfirst else
           "bx = l % gx;\n    by = (l / gx) % gy;\n    bz = l / (gx * gy);")
    if blocked:
        dec = ("if ((gx & 1u) == 0u && (gy & 3u) == 0u && (gz & 3u) == 0u) {\n"
               "      const unsigned w = l & 31u, b = l >> 5, nbx = gx >> 1, nby = gy >> 2;\n"
               "      bx = ((b % nbx) << 1) + (w & 1u);\n"
               "      by = (((b / nbx) % nby) << 2) + ((w >> 1) & 3u);\n"
               "      bz = ((b / (nbx * nby)) << 2) + (w >> 3);\n"
               "    } else {\n"
               "      " + dec.replace("\n    ", "\n      ") + "\n    }")
    return f"""  unsigned bx, by, bz;
  {{
    const unsigned gx = gridDim.x, gy = gridDim.y, gz = gridDim.z, n = gx * gy * gz;
    const unsigned h = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
    const unsigned l = {lin};
    {dec}
  }}"""


LDS_DMA_MACROS = """#ifndef BHAH_DMA16
// 16 bytes per lane straight from global memory into LDS at dst + 16 * lane (dst uniform per wave)
#define BHAH_DMA16(src, dst) __builtin_amdgcn_global_load_lds((const void *)(src), (__attribute__((address_space(3))) void *)(dst), 16, 0, 0)
// the same from a uniform byte base plus a 32-bit per-lane byte offset, widened at the load
// itself so it is addressed as scalar base + vector offset
#define BHAH_DMA16O(base, off, dst) do { unsigned o_ = (off); asm volatile("" : "+v"(o_)); BHAH_DMA16((base) + o_, dst); } while (0)
// wait until at most n vector-memory operations of this wave are outstanding (gfx9 encoding)
#define BHAH_VMWAIT(n) __builtin_amdgcn_s_waitcnt(((n) & 15) | (((n) >> 4) << 14) | 0x0F70)
// workgroup barrier without a memory fence (LDS reads of this wave retired first)
#define BHAH_BARRIER() asm volatile("s_waitcnt lgkmcnt(0)\\n\\ts_barrier" ::: "memory")
// the same before the LDS reads of DMA phase n; the comment marks it in the assembly listing
// for the build's wait check (nrpy_amd/isa_check.py)
#define BHAH_BARRIER_PHASE(n) asm volatile("s_waitcnt lgkmcnt(0)\\n\\ts_barrier ; bhah_read_phase " #n ::: "memory")
#endif
"""


def _lds_offsets(dma: bool, H: int, SH: int, NJ: int, NL: int) -> str:
    """Per-thread global offsets of the box elements it loads.  Indices are clamped into
    the array, so partial tiles at the upper interior edge read valid (unused) points."""
    if not dma:
        return f"""  int ofs[NL];
#pragma unroll
  for (int j = 0; j < NL; j++) {{
    const int e = tid + NT * j;
    const int ex = e % BX, ey = (e / BX) % BY, ez = e / (BX * BY);
    const int g0 = min(o0 - {H} + ex, p.Nx0 - 1), g1 = min(o1 - {H} + ey, p.Nx1 - 1), g2 = min(o2 - {H} + ez, p.Nx2 - 1);
    ofs[j] = g0 + (int)s1 * g1 + (int)s2 * g2;
  }}"""
    # pair j of this thread: box elements 2q, 2q + 1 (q = tid + NT j) are xx0 neighbours
    # (BX is even); a pair start past Nx0 - 2 only feeds points outside the interior
    # The wave index is made uniform so the LDS destination of every load is scalar
    # arithmetic, and the offsets are 32-bit byte offsets from a uniform gridfunction base:
    # each load is then one scalar-base + vector-offset instruction (the host keeps a
    # gridfunction under 4 GiB for these schedules).
    return f"""  const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
  unsigned pofs[{NJ}];
#pragma unroll
  for (int j = 0; j < {NJ}; j++) {{
    const int e = 2 * (tid + NT * j);
    const int ex = e % BX, ey = (e / BX) % BY, ez = e / (BX * BY);
    const int g0 = min(o0 - {H + SH} + ex, p.Nx0 - 2), g1 = min(o1 - {H} + ey, p.Nx1 - 1), g2 = min(o2 - {H} + ez, p.Nx2 - 1);
    pofs[j] = 8u * (unsigned)(g0 + (int)s1 * g1 + (int)s2 * g2);
  }}"""


def _int_sum(w: Dict[int, sp.Rational], D: int, pts: Dict[int, str]) -> Tuple[int, str]:
    """(sign, S) with D * sum_o w[o] pts[o] = sign * S and S a C expression with integer
    weights, led by a weight of one where the stencil has a weight of magnitude one (a
    plain value starts the multiply-add chain)."""
    W = {o: sp.Rational(v) * D for o, v in w.items()}
    order = sorted(W, key=lambda o: (abs(W[o]) != 1, -abs(o)))
    sgn = -1 if W[order[0]] < 0 else 1
    parts = []
    for o in order:
        c = int(sgn * W[o])
        parts.append(pts[o] if (c == 1 and not parts) else f"{c}.0 * {pts[o]}")
    return sgn, f"({' + '.join(parts)})"


def _int_pair_sum(w: Dict[int, sp.Rational], D: int, pts: Dict[int, str]) -> str:
    """D * sum_o w[o] pts[o] for a stencil symmetric about its centre, with integer weights
    on the sums of mirrored points, outermost pair first, the overall sign taken out so
    that a weight of magnitude one leads."""
    W = {o: sp.Rational(v) * D for o, v in w.items() if o >= 0}
    order = sorted(W, reverse=True)
    sgn = -1 if W[order[0]] < 0 else 1
    parts = []
    for o in order:
        c = int(sgn * W[o])
        v = f"({pts[o]} + {pts[-o]})" if o else pts[0]
        parts.append(v if (c == 1 and not parts) else f"{c}.0 * {v}")
    return f"{'-' if sgn < 0 else ''}({' + '.join(parts)})"


def _sym_line_weights(w_up, w_dn, w_ko):
    """Symmetric and antisymmetric parts of the one-sided upwind stencil along a line.

    With a_j = f_j + f_-j and b_j = f_j - f_-j, S = 2 w_0 f_0 + sum_j (w_j + w_-j) a_j and
    A = sum_j (w_j - w_-j) b_j (w the ``dupD`` weights): dupD = (S + A) / 2, and when the
    ``ddnD`` stencil is the mirror image with the sign flipped, ddnD = (A - S) / 2.  For
    NRPy's stencils the KO stencil is r S (upwinding is centred differencing plus exactly
    this dissipation), so one S serves the upwind sum of either sign and the KO sum.
    Returns (integer weights of S by j >= 0, of A by j >= 1, their common denominator DS,
    r), or None when the stencils do not have this form."""
    up = {k: sp.Rational(v) for k, v in w_up.items()}
    dn = {k: sp.Rational(v) for k, v in w_dn.items()}
    ko = {k: sp.Rational(v) for k, v in w_ko.items()}
    R = max(abs(k) for k in list(up) + list(dn) + list(ko))

    def g(w, k):
        return w.get(k, sp.Integer(0))

    if any(g(dn, k) != -g(up, -k) for k in range(-R, R + 1)) or any(g(ko, k) != g(ko, -k) for k in range(R + 1)):
        return None
    S = {0: 2 * g(up, 0), **{j: g(up, j) + g(up, -j) for j in range(1, R + 1)}}
    A = {j: g(up, j) - g(up, -j) for j in range(1, R + 1)}
    k0 = next(k for k in S if S[k] != 0)
    r = g(ko, k0) / S[k0]
    if any(g(ko, k) != r * S[k] for k in range(R + 1)):
        return None
    DS = int(sp.ilcm(*[v.q for v in list(S.values()) + list(A.values())]))
    IS = {k: int(v * DS) for k, v in S.items() if v != 0}
    IA = {k: int(v * DS) for k, v in A.items() if v != 0}
    return IS, IA, DS, r


def _int_line_sum(W: Dict[int, int], pts: Dict[int, str], mirror: str) -> Tuple[str, int]:
    """(C expression, flops) of sum_j W[j] (pts[j] mirror pts[-j]) (W[0]: pts[0] alone),
    outermost pair first, a weight of magnitude one leading as a plain (signed) value."""
    parts, fl = [], 0
    for o in sorted(W, reverse=True):
        v = pts[0] if o == 0 else f"({pts[o]} {mirror} {pts[-o]})"
        fl += 1 if o else 0
        if not parts and abs(W[o]) == 1:
            parts.append(v if W[o] == 1 else f"-{v}")
        else:
            parts.append(f"{W[o]}.0 * {v}")
            fl += 2 if len(parts) > 1 else 1
    return f"({' + '.join(parts)})", fl


_SH_REF = re.compile(r"sh\[c\d+(?: [+-] \d+)?\]")


def _batch_lds_reads(stmts: List[Tuple[str, List[str]]], nmax: int, prefix: str) -> List[Tuple[str, List[str]]]:
    """Group a gridfunction's stencil statements by the LDS values they read: statements
    reading few values (axis stencils, line sums) first, in their order, then the others
    (mixed-derivative planes), packed greedily into batches of at most ``nmax`` distinct
    values.  Each batch's values are read into registers and pinned there before its
    statements, so the reads of a batch are in flight together and the batch waits on LDS
    once.  Without this the compiler, short of registers, interleaves a few reads with the
    arithmetic that uses them, and each phase is a chain of LDS round trips."""
    small = [t for t in stmts if len(set(_SH_REF.findall(t[0]))) <= 9]
    big = [t for t in stmts if len(set(_SH_REF.findall(t[0]))) > 9]
    out: List[Tuple[str, List[str]]] = []
    batch: List[Tuple[str, List[str]]] = []
    refs: List[str] = []
    nb = [0]

    def flush():
        if not batch:
            return
        names = {r: f"{prefix}_{nb[0]}_{i}" for i, r in enumerate(refs)}
        nb[0] += 1
        out.append(("".join(f"  double {n} = {r};\n" for r, n in names.items()).rstrip("\n"), []))
        vals = list(names.values())
        for i in range(0, len(vals), 24):
            out.append((f"  BHAH_PINS({', '.join(f'{chr(34)}+v{chr(34)}({x})' for x in vals[i:i + 24])});", []))
        for text, formed in batch:
            out.append((_SH_REF.sub(lambda m: names[m.group(0)], text), formed))

    for text, formed in small + big:
        r = list(dict.fromkeys(_SH_REF.findall(text)))
        new = [x for x in r if x not in refs]
        if batch and len(refs) + len(new) > nmax:
            flush()
            batch, refs = [], []
            new = r
        refs += new
        batch.append((text, formed))
    flush()
    return out


def is_lds(cand: str) -> bool:
    """Whether a schedule name selects the workgroup-tiled (LDS) kernels."""
    return any(o[:1] == "L" and o[1:2].isdigit() for o in cand.split("_"))


def _lds_geometry(tok: str) -> Tuple[int, int, int, int]:
    """``L<TX>x<TY>x<TZ>x<P>``: a TX x TY x TZ tile of interior points per workgroup,
    P points per thread (spaced TZ/P apart along xx2)."""
    TX, TY, TZ, P = (int(x) for x in tok[1:].split("x"))
    if TZ % P or (TX * TY * TZ) % P:
        raise ValueError(f"bad LDS tile {tok}")
    return TX, TY, TZ, P


def lds_kernel(sub: specs.Spec, fd_order: int, kname: str, rats: Dict[str, str], geom: Tuple[int, int, int, int],
               min_waves: int, scope: str, lv: List[str], xcd: bool = False,
               depth: int = 1, zfirst: bool = False, group: int = 1, dma: bool = False,
               one_barrier: bool = False, ensemble: str = "", interleave: bool = False,
               pin_group: int = -1, branchless_upwind: bool = False, blocked: bool = False,
               sym_lines: bool = False, read_batch: int = 0, dma_spread: bool = False) -> Tuple[str, str, int]:
    """One workgroup-tiled kernel.  Each gridfunction with derivatives is staged, one at a
    time, from global memory into a halo box in LDS: ``group`` gridfunctions per
    barrier-separated phase, with the loads of the next ``depth`` phases in flight while
    the current box is read.  Every derivative of the thread's points is formed from the
    box, then the point algebra runs with all inputs in registers.

    With ``dma`` the boxes are filled by direct global-to-LDS loads (16 bytes, two xx0
    neighbours, per lane) into a ring of ``depth`` phase buffers, so prefetching takes no
    registers and no LDS store instructions; workgroup barriers are then plain s_barrier
    with explicit vmcnt waits counted per phase (``one_barrier``: one barrier per phase,
    refilling the slot the previous phase read).  The box's xx0 origin is rounded down
    to an even index (every pair 16-byte aligned), which needs an even Nx0 (checked by the
    host).

    With ``ensemble`` (a kernel name) the source also holds the batched-ensemble form: the
    per-tile code as a device function and a kernel walking (member, tile) records, tiles in
    this kernel's TX x TY x TZ units.

    Centred stencil sums are the same expressions, in the same order, as the global-load
    kernels; upwind and KO sums use integer weights with the denominator in a uniform scale.
    With ``sym_lines`` a line carrying both an upwind and a KO term is formed from its
    symmetric and antisymmetric sums (``_sym_line_weights``) with per-point coefficients
    that fold in the upwind sign, the KO factor and the scales: two multiply-adds into the
    output's accumulator and no select.  With ``read_batch`` N a gridfunction's LDS values are
    read in batches of at most N (the axis stencils first, then each mixed-derivative plane),
    each batch pinned to registers before its arithmetic, so a batch costs one LDS round trip
    (``_batch_lds_reads``).  With ``dma_spread`` the global-to-LDS loads a phase issues for a
    later phase are spread through its statements: issued all at once after the barrier, the
    four waves' loads queue for the address unit and each wave waits there before its first
    LDS read.
    Returns (kernel source, launch statement, flops per point)."""
    TX, TY, TZ, P = geom
    NT = TX * TY * TZ // P
    ZS = TZ // P
    centre, auxn, derivs, rfm, params = sub.classify()
    used: Set[str] = set()
    for _, e in sub.outputs:
        used |= {str(x) for x in e.free_symbols}
    for _, e in sub.used_defs():
        used |= {str(x) for x in e.free_symbols}
    dused = [d for d in derivs if d.symbol in used and d.op in ("dD", "dDD")]
    # upwind and KO derivatives enter linearly: accumulated per output while their
    # gridfunction is in LDS, with coefficients formed before the staging
    terms = split_linear(sub)[1] if any(d.op in ("dupD", "dKOD") for d in derivs) else {}
    coefs = sorted({c for tl in terms.values() for _, c in tl}, key=sp.default_sort_key)
    cidx = {c: j for j, c in enumerate(coefs)}
    line_cf: Dict[Tuple[str, int], Dict[str, int]] = {}
    line_out: Dict[Tuple[str, int], str] = {}
    for g_out, tl in terms.items():
        for d, c in terms_sorted(tl):
            key = (d.gf, d.dirs[0])
            if line_out.get(key, g_out) != g_out:
                raise ValueError(f"{d.symbol} enters two outputs")
            line_cf.setdefault(key, {})["up" if d.op == "dupD" else "ko"] = cidx[c]
            line_out[key] = g_out
    H = fd.ghost_zones(fd_order) if terms else fd_order // 2
    staged = [g for g in sub.gfs if any(d.gf == g for d in dused) or any(k[0] == g for k in line_out)]
    if interleave:
        # spread the gridfunctions with mixed second derivatives (several times the LDS reads
        # and arithmetic of the others) evenly over the phases, so that the loads running
        # ahead meet phases of similar length
        heavy = [g for g in staged if any(d.gf == g and d.op == "dDD" and d.dirs[0] != d.dirs[1] for d in dused)]
        light = [g for g in staged if g not in heavy]
        order: List[str] = []
        for i, g in enumerate(heavy):
            lo = (i * len(light)) // max(1, len(heavy))
            hi = ((i + 1) * len(light)) // max(1, len(heavy))
            order += [g] + light[lo:hi]
        staged = order if heavy else staged
    # xx0 shift of the box origin that makes it even (DMA pairs): NG - H is fixed per build
    SH = ((fd.ghost_zones(fd_order) - H) & 1) if dma else 0
    if dma and TX % 2:
        raise ValueError("DMA staging needs an even tile width")
    BX, BY, BZ = TX + 2 * H + 2 * SH, TY + 2 * H, TZ + 2 * H
    NB = BX * BY * BZ
    NL = (NB + NT - 1) // NT
    NJ = (NB // 2 + NT - 1) // NT  # 16-byte DMA loads per thread per box
    SLOT = 2 * NT * NJ if dma else NB  # doubles per box (DMA boxes padded so every lane loads)
    nslots = (max(1, min(depth, -(-len(staged) // group))) * group) if dma else min(group, len(staged))
    if 8 * SLOT * nslots > 160 * 1024:
        raise ValueError(f"{kname}: {8 * SLOT * nslots} bytes of LDS exceed the 160 KiB of a CU")
    # algebra for a placeholder point; each point substitutes its own register names
    ovr: Dict[str, str] = {}
    for n in centre:
        if n in used:
            ovr[n] = f"VPT_{n}" if n in staged else f"in[{sub.gfs.index(n)} * Ntot + idx]"
    for d in dused:
        ovr[d.symbol] = f"VPT_{d.symbol}"
    body = _ordered_body(sub, fd_order, lv, rats, cse_scope=scope, linear="lds", prefetch=0, barriers=False,
                         defs_override=ovr)
    flops = FLOPS[f"{sub.name}_fd{fd_order}{_mode_suffix('lds', 0)}"]
    pr = emit.DoublePrinter(rats)
    w_up = dict(fd.weights_1d("dupD", fd_order))
    w_dn = dict(fd.weights_1d("ddnD", fd_order))
    w_ko = dict(fd.weights_1d("dKOD", fd_order))
    sym = _sym_line_weights(w_up, w_dn, w_ko) if sym_lines and terms else None
    if sym_lines and terms and sym is None:
        raise ValueError(f"{kname}: the upwind and KO stencils have no shared symmetric part")
    both = {key: cf for key, cf in line_cf.items() if "up" in cf and "ko" in cf} if sym else {}
    ca_keys = sorted({(cf["up"], key[1]) for key, cf in both.items()})
    csk_keys = sorted({(cf["up"], cf["ko"], key[1]) for key, cf in both.items()})
    L: List[str] = []
    a = L.append
    a(f"""  // Tile of {TX}x{TY}x{TZ} interior points; halo box {BX}x{BY}x{BZ} = {NB} values per gridfunction.
  constexpr int BX = {BX}, BY = {BY}, NB = {NB}, NT = {NT}, NL = {NL};
  const int tid = (int)threadIdx.x;
  const int lx = tid % {TX}, ly = (tid / {TX}) % {TY}, lz = tid / {TX * TY};
  const long s1 = p.s1, s2 = p.s2, Ntot = p.Ntot;
  const int o0 = p.NG + (int)bx * {TX}, o1 = p.NG + (int)by * {TY}, o2 = p.NG + (int)bz * {TZ};
{_lds_offsets(dma, H, SH, NJ, NL)}""")
    for k in range(P):
        a(f"  const int c{k} = (lx + {H + SH}) + BX * (ly + {H}) + BX * BY * (lz + {k * ZS + H});")
    # Upwind sums are formed with integer weights (their common denominator DU moved into
    # one uniform scale per direction): an FP64 multiply-add takes an integer weight as an
    # instruction literal, where a weight such as 1/30 needs a register pair loaded first.
    DU = int(sp.ilcm(*[sp.Rational(v).q for v in list(w_up.values()) + list(w_dn.values())]))
    DK = int(sp.ilcm(*[sp.Rational(v).q for v in w_ko.values()]))
    for d in range(3):
        a(f"  const double ius{d} = p.invdxx{d} * (1.0 / {DU}.0), nius{d} = -ius{d}, "
          f"ikos{d} = p.invdxx{d} * (1.0 / {DK}.0);")
    if terms:
        # coefficients of the linear terms and the upwind control vector at each point
        crep, cred = sp.cse(coefs, symbols=sp.numbered_symbols("lc"), order="none")
        flops += sum(emit._flops(e) for _, e in crep) + sum(emit._flops(e) for e in cred)
        upw = [sp.sympify(sub.upwind[d]) for d in range(3)]
        names = set()
        for e in [e for _, e in crep] + list(cred) + upw:
            names |= {str(x) for x in e.free_symbols}
        for k in range(P):
            a(f"  double L{k}_coef[{len(coefs)}], L{k}_beta[3];")
            a("  {")
            a(f"    const int i0 = min(o0 + lx, p.Nx0 - p.NG - 1), i1 = min(o1 + ly, p.Nx1 - p.NG - 1);")
            a(f"    const long idx = i0 + s1 * i1 + s2 * min(o2 + lz + {k * ZS}, p.Nx2 - p.NG - 1);")
            for n in sorted(names):
                if n in sub.gfs:
                    a(f"    const double {n} = in[{sub.gfs.index(n)} * Ntot + idx];")
                elif n in specs.RFM_SYMBOLS:
                    a(f"    const double {n} = {RFM_EXPR[specs.RFM_SYMBOLS.index(n)]};")
                elif n in PARAM_EXPR:
                    a(f"    const double {n} = {PARAM_EXPR[n]};")
                elif not n.startswith("lc"):
                    raise ValueError(f"linear coefficient depends on {n}")
            for sy, e in crep:
                a(f"    const double {sy} = {pr.doprint(e)};")
            for j, e in enumerate(cred):
                a(f"    L{k}_coef[{j}] = {pr.doprint(e)};")
            for d in range(3):
                a(f"    L{k}_beta[{d}] = {pr.doprint(upw[d])};")
            a("  }")
            if ca_keys:
                # per line: acc += ca * A + csk * S, with ca = c_up * h / (2 DS) and
                # csk = sign(beta) * ca + c_ko * h * r / DS (h = 1/dxx of the line's direction)
                a(f"  double L{k}_ca[{len(ca_keys)}], L{k}_csk[{len(csk_keys)}];")
                for i, (cu, d) in enumerate(ca_keys):
                    a(f"  L{k}_ca[{i}] = L{k}_coef[{cu}] * (p.invdxx{d} * {_dbl(sp.Rational(1, 2 * sym[2]))});")
                for m, (cu, ck, d) in enumerate(csk_keys):
                    i = ca_keys.index((cu, d))
                    a(f"  L{k}_csk[{m}] = ((L{k}_beta[{d}] > 0.0) ? L{k}_ca[{i}] : -L{k}_ca[{i}]) + "
                      f"L{k}_coef[{ck}] * (p.invdxx{d} * {_dbl(sym[3] / sym[2])});")
            for g_out in sorted(terms):
                a(f"  double V{k}_lin_{g_out} = 0.0;")
    phases = [staged[i:i + group] for i in range(0, len(staged), group)]
    depth = max(1, min(depth, len(phases)))
    fl = [flops]

    def box(ph, q):
        """LDS offset (doubles) of box q of phase ph."""
        return ((ph % depth) * group + q) * SLOT if dma else q * NB

    def emit_phase(ph, gfs, dma_lines=()):
        """Derivatives and linear-term contributions of the phase's gridfunctions; with
        ``dma_lines`` (the global-to-LDS loads of a later phase) those are spread evenly
        through the phase's statements instead of all issued at its start."""
        lines: List[str] = []
        e = lines.append
        for q, g in enumerate(gfs):
            pins: List[str] = []

            def pin(x):
                # each value is pinned to a register once formed: otherwise the compiler
                # sinks the stencil sums into the guarded algebra below and keeps every raw
                # LDS value alive across the staging barriers instead.  A pin is a volatile
                # asm, which no memory access crosses; with pin_group >= 0 one pin covers
                # pin_group formed values (0: all of the staged gridfunction's), so that the
                # LDS reads of several stencils can be in flight together.
                if pin_group < 0:
                    e(f"  BHAH_PIN({x});")
                    return
                if x not in pins:
                    pins.append(x)
                if pin_group and len(pins) >= pin_group:
                    e(f"  BHAH_PINS({', '.join(f'{chr(34)}+v{chr(34)}({y})' for y in pins)});")
                    pins.clear()

            for k in range(P):
                st = _LdsEmitter(sub, fd_order, f"c{k}", BX, BY, box(ph, q))
                stmts: List[Tuple[str, List[str]]] = []  # (statement, values it forms)
                if g in ovr and g in used:
                    stmts.append((f"  double V{k}_{g} = {st.load(g, (0, 0, 0))};", [f"V{k}_{g}"]))
                for d in dused:
                    if d.gf == g:
                        stmts.append((f"  double V{k}_{d.symbol} = {st.expr(d)};", [f"V{k}_{d.symbol}"]))
                for dirn in range(3):
                    key = (g, dirn)
                    if key not in line_out:
                        continue
                    cf = line_cf[key]
                    acc = f"V{k}_lin_{line_out[key]}"
                    pts = {o: st.load(g, tuple(o if r == dirn else 0 for r in range(3))) for o in range(-H, H + 1)}
                    if key in both:
                        S, fs = _int_line_sum(sym[0], pts, "+")
                        A, fa = _int_line_sum(sym[1], pts, "-")
                        i = ca_keys.index((cf["up"], dirn))
                        m = csk_keys.index((cf["up"], cf["ko"], dirn))
                        stmts.append((f"  {acc} = fma(L{k}_ca[{i}], {A}, fma(L{k}_csk[{m}], {S}, {acc}));", [acc]))
                        if k == 0:
                            fl[0] += fs + fa + 4
                        continue
                    tc = []
                    if "up" in cf:
                        # the sign taken out of each sum goes into its scale (a free operand
                        # modifier of the multiply)
                        (gu, su), (gd, sd) = _int_sum(w_up, DU, pts), _int_sum(w_dn, DU, pts)
                        fu, fd_ = ("n" if gu < 0 else "") + f"ius{dirn}", ("n" if gd < 0 else "") + f"ius{dirn}"
                        if branchless_upwind:
                            # both one-sided sums formed (from the points the KO sum reads
                            # anyway) and one selected: no branch splits the scheduling region
                            stmts.append((f"  const double U{k}_{g}_{dirn} = {fu} * {su}, D{k}_{g}_{dirn} = {fd_} * {sd};",
                                          []))
                            tc.append(f"L{k}_coef[{cf['up']}] * ((L{k}_beta[{dirn}] > 0.0) ? U{k}_{g}_{dirn} : "
                                      f"D{k}_{g}_{dirn})")
                        else:
                            tc.append(f"L{k}_coef[{cf['up']}] * ((L{k}_beta[{dirn}] > 0.0) ? ({fu} * {su}) : "
                                      f"({fd_} * {sd}))")
                        if k == 0:
                            fl[0] += 2 * (2 * len(w_up) - 1) + 2
                    if "ko" in cf:
                        tc.append(f"L{k}_coef[{cf['ko']}] * (ikos{dirn} * {_int_pair_sum(w_ko, DK, pts)})")
                        if k == 0:
                            fl[0] += 3 * (len(w_ko) // 2) + 3
                    stmts.append((f"  {acc} += {' + '.join(tc)};", [acc]))
                if read_batch:
                    stmts = _batch_lds_reads(stmts, read_batch, f"R{k}_{g}")
                for text, formed in stmts:
                    e(text)
                    for x in formed:
                        pin(x)
            for i in range(0, len(pins), 24):
                e(f"  BHAH_PINS({', '.join(f'{chr(34)}+v{chr(34)}({x})' for x in pins[i:i + 24])});")
        if dma_lines:
            # load i goes where the statements' text passes (i + 1) / (loads + 1) of its
            # length; every load is placed exactly once, in order
            m = len(dma_lines)
            tot = sum(len(ln) for ln in lines) or 1
            merged: List[str] = []
            cum, i = 0, 0
            for ln in lines:
                while i < m and cum >= (i + 1) * tot / (m + 1):
                    merged.append(dma_lines[i])
                    i += 1
                merged.append(ln)
                cum += len(ln)
            merged += dma_lines[i:]
            lines = merged
        for ln in lines:
            a(ln)

    if dma:
        # every lane of every wave issues NJ loads per box (padded boxes), so the number of
        # loads issued after a phase's own is known at compile time: waiting until at most
        # that many are outstanding means this phase's boxes have landed (vmcnt counts in
        # order); the barrier then publishes all waves' boxes
        issue: List[Tuple[int, int]] = []  # (phase, loads per lane) in issue order

        def dma_phase(ph):
            issue.append((ph, len(phases[ph]) * NJ))
            for q, g in enumerate(phases[ph]):
                a(f"""  {{
    const char *src = (const char *)(in + {sub.gfs.index(g)} * Ntot);
#pragma unroll
    for (int j = 0; j < {NJ}; j++) BHAH_DMA16O(src, pofs[j], sh + {box(ph, q)} + 2 * (NT * j + 64 * wv));
  }}""")

        def dma_lines(ph):
            """The loads of phase ph one statement each, in the order dma_phase issues them."""
            issue.append((ph, len(phases[ph]) * NJ))
            return [f"  BHAH_DMA16O((const char *)(in + {sub.gfs.index(g)} * Ntot), pofs[{j}], "
                    f"sh + {box(ph, q)} + 2 * (NT * {j} + 64 * wv));"
                    for q, g in enumerate(phases[ph]) for j in range(NJ)]

        if one_barrier:
            # Refill, right after a phase's barrier, the slot the previous phase read: every
            # wave passed that barrier only after finishing the previous phase, so one barrier
            # per phase suffices; loads run depth - 1 phases ahead.
            if depth < 2:
                raise ValueError("single-barrier DMA staging needs at least two ring slots")
            for ph in range(depth - 1):
                dma_phase(ph)
            for ph, gfs in enumerate(phases):
                later = sum(len(phases[x]) for x in range(ph + 1, min(ph + depth - 1, len(phases)))) * NJ
                if later + group * NJ > 63:
                    raise ValueError("DMA prefetch deeper than the 6-bit vmcnt counter")
                a(f"  // ---- {', '.join(gfs)}")
                a(f"  BHAH_VMWAIT({later});")
                a(f"  BHAH_BARRIER_PHASE({ph});")
                if ph + depth - 1 < len(phases) and dma_spread:
                    emit_phase(ph, gfs, dma_lines(ph + depth - 1))
                    continue
                if ph + depth - 1 < len(phases):
                    dma_phase(ph + depth - 1)
                emit_phase(ph, gfs)
        else:
            for ph in range(depth):
                dma_phase(ph)
            for ph, gfs in enumerate(phases):
                later = sum(len(phases[x]) for x in range(ph + 1, min(ph + depth, len(phases)))) * NJ
                if later > 63:
                    raise ValueError("DMA prefetch deeper than the 6-bit vmcnt counter")
                a(f"  // ---- {', '.join(gfs)}")
                a(f"  BHAH_VMWAIT({later});")
                a(f"  BHAH_BARRIER_PHASE({ph});")
                emit_phase(ph, gfs)
                if ph + depth < len(phases):
                    a("  BHAH_BARRIER();")  # every wave is done with the ring slot about to be refilled
                    dma_phase(ph + depth)
    else:
        for q in range(depth * group):
            a(f"  double pf{q}[NL];")

        def fetch(ph):
            for q, g in enumerate(phases[ph]):
                a(f"""  {{
    const double *src = in + {sub.gfs.index(g)} * Ntot;
#pragma unroll
    for (int j = 0; j < NL; j++) pf{(ph % depth) * group + q}[j] = src[ofs[j]];
  }}""")

        for ph in range(depth):
            fetch(ph)
        for ph, gfs in enumerate(phases):
            a(f"  // ---- {', '.join(gfs)}")
            a("  __syncthreads();")
            a("#pragma unroll")
            a("  for (int j = 0; j < NL; j++) {")
            a("    const int e = tid + NT * j;")
            a("    if (e < NB) {")
            for q in range(len(gfs)):
                a(f"      sh[{q} * NB + e] = pf{(ph % depth) * group + q}[j];")
            a("    }")
            a("  }")
            a("  __syncthreads();")
            if ph + depth < len(phases):
                fetch(ph + depth)
            emit_phase(ph, gfs)
    flops = fl[0]
    for k in range(P):
        a(f"""  {{
    const int i0 = o0 + lx, i1 = o1 + ly, i2 = o2 + lz + {k * ZS};
    if (i0 < p.Nx0 - p.NG && i1 < p.Nx1 - p.NG && i2 < p.Nx2 - p.NG) {{
      const long idx = i0 + s1 * i1 + s2 * i2;
{body.replace("VPT_", f"V{k}_")}
    }}
  }}""")
    tile_body = "\n".join(L)
    bounds = f"{NT}{f', {min_waves}' if min_waves > 1 else ''}"
    dma_note = ""
    if dma:
        # read by nrpy_amd/isa_check.py: DMA groups (phase:loads per lane) in issue order
        groups = " ".join(f"{ph}:{n}" for ph, n in issue)
        dma_note = "".join(f"// bhah-dma-issue {k}: {groups}\n" for k in [kname] + ([ensemble] if ensemble else []))
    src = f"""
{dma_note}extern "C" __global__ void __launch_bounds__({bounds})
{kname}(const double *__restrict__ in, const double *__restrict__ aux, double *__restrict__ out,
        const double *__restrict__ rfm0, const double *__restrict__ rfm1, const bhah_kparams p) {{
  __shared__ double sh[{nslots * SLOT}];
{_lds_tile_order(xcd, zfirst, blocked)}
{tile_body}
}}
"""
    if ensemble:
        # Batched form: a fixed grid of workgroups walks the ensemble's list of (member, tile)
        # records; the per-tile code is the same, on the member's arrays.  The barrier before
        # each tile keeps a fast wave from refilling LDS the previous tile still reads.
        src += f"""
static __device__ __forceinline__ void {kname}_tile(const double *__restrict__ in, const double *__restrict__ aux,
        double *__restrict__ out, const double *__restrict__ rfm0, const double *__restrict__ rfm1,
        const bhah_kparams p, const unsigned bx, const unsigned by, const unsigned bz, double *sh) {{
{tile_body}
}}
extern "C" __global__ void __launch_bounds__({bounds})
{ensemble}(const bhah_member *__restrict__ mem, const bhah_item *__restrict__ items, const int nitems,
        const int in_sel, const int aux_sel, const int out_sel) {{
  __shared__ double sh[{nslots * SLOT}];
  for (int it = blockIdx.x; it < nitems; it += gridDim.x) {{
    const bhah_item w = items[it];
    const bhah_member *M = &mem[w.m];
    {"BHAH_BARRIER();" if dma else "__syncthreads();"}
    {kname}_tile(M->arr[in_sel], M->arr[aux_sel], M->arr[out_sel], M->rfm0, M->rfm1, M->kp, (unsigned)w.bx,
                 (unsigned)w.by, (unsigned)w.bz, sh);
  }}
}}
"""
    launch = (f"  hipLaunchKernelGGL({kname}, dim3((p.Nx0 - 2 * p.NG + {TX - 1}) / {TX}, "
              f"(p.Nx1 - 2 * p.NG + {TY - 1}) / {TY}, (p.Nx2 - 2 * p.NG + {TZ - 1}) / {TZ}), dim3({NT}), 0, s, "
              f"in, aux, out, rfm0, rfm1, p);")
    return src, launch, flops


def candidate_tu(v: specs.Variant, fd_order: int, kind: str, cand: str, ensemble: bool = False) -> Tuple[str, int]:
    """One candidate schedule: group kernels plus a launcher running them in order.

    ``cand`` is ``<partition>_<launch bounds>[_<mode>]``; mode ``x`` (default) CSEs the
    fully expanded expressions, ``sj`` evaluates named intermediate tensors with one
    joint CSE, ``sl`` with a separate CSE per intermediate (shortest live ranges).
    Returns (source, flops per point summed over the groups)."""
    tag = f"{v.tag}_fd{fd_order}"
    parts = cand.split("_")
    part, lb = parts[0], parts[1]
    opts = parts[2:]
    mode = next((o for o in opts if o in ("x", "sj", "sl")), "x")
    staged = mode in ("sj", "sl")
    scope = "local" if mode == "sl" else "joint"
    linear = "fused" if "lf" in opts else ("lines" if "ln" in opts else "stream")
    prefetch = next((int(o[2:]) for o in opts if o.startswith("pf")), 0)
    ahead = next((int(o[1:]) for o in opts if o[0] == "a" and o[1:].isdigit()), 12)
    xcd = "xcd" in opts
    u32 = "u32" in opts
    chunk = next((int(o[2:]) for o in opts if o.startswith("ck") and o[2:].isdigit()), 0)
    zs = next((int(o[2:]) for o in opts if o.startswith("zs") and o[2:].isdigit()), 0)
    wide = "w2" in opts
    lds = next((_lds_geometry(o) for o in opts if o[0] == "L" and o[1:2].isdigit()), None)
    rats: Dict[str, str] = {}
    kernels, calls, total = [], [], 0
    ens_kernels, ens_calls = [], []
    if kind == "rhs":
        spec = specs.bssn_rhs_spec(v, staged=staged)
        groups = RHS_PARTITIONS[part]
    else:
        spec = specs.ricci_spec(v, staged=staged)
        groups = RICCI_PARTITIONS[part]
    for k, gfs in enumerate(groups):
        sub = spec if len(groups) == 1 else _subspec(spec, f"{spec.name}_{part}g{k}", gfs)
        if scope == "local":
            sub = _subspec(sub, f"{sub.name}_local", [g for g, _ in sub.outputs])
        kname = f"bssn_{kind}_{tag}_{cand}_g{k}"
        if lds:
            outs = list(bssn.EVOL_GFS if kind == "rhs" else bssn.AUXEVOL_GFS)
            ename = f"bssn_ens_{kind}_{tag}_{cand}_g{k}" if ensemble else ""
            src, launch, fl = lds_kernel(sub, fd_order, kname, rats, lds, 2 if lb == "lb2" else 1, scope,
                                         [f"out[{outs.index(g)} * Ntot + idx]" for g, _ in sub.outputs], xcd=xcd,
                                         depth=next((int(o[1:]) for o in opts if o[0] == "d" and o[1:].isdigit()), 1),
                                         zfirst="zf" in opts, blocked="blk" in opts,
                                         group=next((int(o[1:]) for o in opts if o[0] == "g" and o[1:].isdigit()), 1),
                                         dma="dma" in opts, one_barrier="rb1" in opts, ensemble=ename,
                                         interleave="io" in opts,
                                         branchless_upwind="ub" in opts, sym_lines="sa" in opts,
                                         read_batch=next((int(o[2:]) for o in opts if o[:2] == "lr" and o[2:].isdigit()), 0),
                                         dma_spread="di" in opts,
                                         pin_group=next((int(o[2:] or 0) for o in opts if o[:2] == "pg" and o[2:].isdigit() or o == "pg"), -1))
            kernels.append(src)
            calls.append(launch)
            total += fl
            if ensemble:
                ens_calls.append(f"  hipLaunchKernelGGL({ename}, g, dim3({lds[0] * lds[1] * lds[2] // lds[3]}), 0, s, mem, "
                                 f"items, nitems, in_sel, aux_sel, out_sel);")
            continue
        if kind == "rhs":
            lv = [f"out[{bssn.EVOL_GFS.index(g)} * Ntot + idx]" for g, _ in sub.outputs]
            body = _ordered_body(sub, fd_order, lv, rats, cse_scope=scope, linear=linear, prefetch=prefetch,
                                 ahead=ahead, u32=u32, chunk=chunk, wide=wide)
            has_up = any(d.op == "dupD" for d in sub.classify()[2])
            upw = _upwind_prologue(sub, fd_order, rats, tables=(linear == "stream"), u32=u32) if has_up else ""
        else:
            lv = [f"out[{bssn.AUXEVOL_GFS.index(g)} * Ntot + idx]" for g, _ in sub.outputs]
            body = _ordered_body(sub, fd_order, lv, rats, cse_scope=scope, prefetch=prefetch, ahead=ahead, u32=u32,
                                 chunk=chunk, linear="fused" if linear == "fused" else "stream", wide=wide)
            upw = ""
        used_linear = linear if kind == "rhs" else ("fused" if linear == "fused" else "stream")
        total += FLOPS[f"{sub.name}_fd{fd_order}{_mode_suffix(used_linear, prefetch)}"]
        kernels.append(f"""
extern "C" __global__ void __launch_bounds__({_group_bounds(lb, k, len(groups))})
{kname}(const double *__restrict__ in, const double *__restrict__ aux, double *__restrict__ out,
        const double *__restrict__ rfm0, const double *__restrict__ rfm1, const bhah_kparams p) {{
{_prologue(True, xcd, u32, zs)}
{upw}
{body}
}}
""")
        calls.append(f"  hipLaunchKernelGGL({kname}, g, b, 0, s, in, aux, out, rfm0, rfm1, p);")
        if ensemble:
            ename = f"bssn_ens_{kind}_{tag}_{cand}_g{k}"
            ens_kernels.append(_ensemble_kernel(ename, upw, body, _group_bounds(lb, k, len(groups)), u32))
            ens_calls.append(f"  hipLaunchKernelGGL({ename}, g, b, 0, s, mem, items, nitems, in_sel, aux_sel, out_sel);")
    consts = "\n".join(f"#define {k} ({v})" for k, v in sorted(rats.items()))
    undefs = "\n".join(f"#undef {k}" for k in sorted(rats))
    if lds:
        consts = LDS_PIN_MACRO + LDS_DMA_MACROS + consts
    if u32:
        consts = BLD_MACRO + consts
        undefs += "\n#undef BLD\n#undef BHAH_LO\n#undef BHAH_HI"
    launcher = f"""
extern "C" hipError_t launch_bssn_{kind}_{tag}_{cand}({CAND_SIG}) {{
{chr(10).join(calls)}
  return hipGetLastError();
}}
"""
    if ensemble:
        launcher += f"""
extern "C" hipError_t launch_bssn_ens_{kind}_{tag}_{cand}({ENS_SIG}) {{
{chr(10).join(ens_calls)}
  return hipGetLastError();
}}
"""
    head = f"// ---- {kind} candidate {cand}: {len(groups)} kernel(s); {v.coord}, fd_order={fd_order}\n"
    return head + consts + "\n" + "".join(kernels) + "".join(ens_kernels) + undefs + "\n" + launcher, total


def bssn_kernels(v: specs.Variant, fd_order: int, ensemble: bool = False) -> str:
    """Kernels shared by every schedule: RK4+det pass, det enforcement, constraints
    (and, for ensemble builds, the batched RK4+det pass)."""
    tag = f"{v.tag}_fd{fd_order}"
    rats: Dict[str, str] = {}
    ngf = len(bssn.EVOL_GFS)
    # ---- RK4 substep update + det(gammabar) enforcement, all points, one streaming pass
    det = detgamma_inline(v, rats, "nxt")
    # One thread per point in array order (every lane of every wave on consecutive
    # addresses, no partial rows); every array is streamed once, so loads and stores
    # carry the non-temporal hint.
    rk_nt = RK_EPILOGUE
    for a, b in (("y_run[g * Ntot + idx] = (1.0 / 6.0) * dt * k;",
                  "__builtin_nontemporal_store((1.0 / 6.0) * dt * k, &y_run[g * Ntot + idx]);"),
                 ("(1.0 / 6.0) * dt * k + y_n[g * Ntot + idx] + y_run[g * Ntot + idx];",
                  "(1.0 / 6.0) * dt * k + __builtin_nontemporal_load(&y_n[g * Ntot + idx]) + "
                  "__builtin_nontemporal_load(&y_run[g * Ntot + idx]);"),
                 ("y_run[g * Ntot + idx] = (1.0 / 3.0) * dt * k + y_run[g * Ntot + idx];",
                  "__builtin_nontemporal_store((1.0 / 3.0) * dt * k + __builtin_nontemporal_load(&y_run[g * Ntot + idx]), "
                  "&y_run[g * Ntot + idx]);"),
                 ("a * dt * k + y_n[g * Ntot + idx];", "a * dt * k + __builtin_nontemporal_load(&y_n[g * Ntot + idx]);")):
        assert a in rk_nt, a
        rk_nt = rk_nt.replace(a, b)
    rkdet = f"""
extern "C" __global__ void __launch_bounds__(256)
bssn_rk_det_{tag}(const double *__restrict__ kin, const double *y_n, double *__restrict__ y_run, double *out,
                  const double *__restrict__ rfm0, const double *__restrict__ rfm1, const bhah_kparams p) {{
  const long Ntot = p.Ntot;
  const long idx = (long)blockIdx.x * 256 + threadIdx.x;
  if (idx >= Ntot) return;
  const int i0 = (int)(idx % p.Nx0), i1 = (int)((idx / p.Nx0) % p.Nx1);
  constexpr int NGF = {ngf};
  double kk[NGF], ctr[NGF];
#pragma unroll
  for (int g = 0; g < NGF; g++) {{
    kk[g] = __builtin_nontemporal_load(&kin[g * Ntot + idx]);
    ctr[g] = (p.substep == 1) ? __builtin_nontemporal_load(&y_n[g * Ntot + idx]) : 0.0;
  }}
{rk_nt}
{det}
#pragma unroll
  for (int g = 0; g < NGF; g++) __builtin_nontemporal_store(nxt[g], &out[g * Ntot + idx]);
}}
"""
    # Two consecutive points per thread with 16-byte loads and stores: twice the bytes per
    # request of every array stream.  Same arithmetic per point as the pass above; the
    # launcher takes it when Ntot is even and every array is 16-byte aligned.
    rkdet += f"""
typedef double bhah_d2 __attribute__((ext_vector_type(2)));
extern "C" __global__ void __launch_bounds__(256)
bssn_rk_det2_{tag}(const double *__restrict__ kin, const double *y_n, double *__restrict__ y_run, double *out,
                   const double *__restrict__ rfm0, const double *__restrict__ rfm1, const bhah_kparams p) {{
  const long Ntot = p.Ntot;
  const long idx2 = ((long)blockIdx.x * 256 + threadIdx.x) * 2;
  if (idx2 >= Ntot) return;
  constexpr int NGF = {ngf};
  const bhah_d2 *K2 = (const bhah_d2 *)(kin + idx2), *YN2 = (const bhah_d2 *)(y_n + idx2);
  bhah_d2 *YR2 = (bhah_d2 *)(y_run + idx2), *OUT2 = (bhah_d2 *)(out + idx2);
  const long S2 = Ntot / 2;  // gridfunction stride in pairs
  // RK4 substep update (register_all_MoL_step_forward_in_time.c:4-103), in registers.
  const double dt = p.dt;
  bhah_d2 nx2[NGF];
  if (p.substep == 1) {{
#pragma unroll
    for (int g = 0; g < NGF; g++) {{
      const bhah_d2 k = __builtin_nontemporal_load(&K2[g * S2]);
      __builtin_nontemporal_store((1.0 / 6.0) * dt * k, &YR2[g * S2]);
      nx2[g] = (1.0 / 2.0) * dt * k + __builtin_nontemporal_load(&YN2[g * S2]);
    }}
  }} else if (p.substep == 4) {{
#pragma unroll
    for (int g = 0; g < NGF; g++) {{
      const bhah_d2 k = __builtin_nontemporal_load(&K2[g * S2]);
      nx2[g] = (1.0 / 6.0) * dt * k + __builtin_nontemporal_load(&YN2[g * S2]) + __builtin_nontemporal_load(&YR2[g * S2]);
    }}
  }} else {{
    const double a = (p.substep == 2) ? (1.0 / 2.0) : 1.0;
#pragma unroll
    for (int g = 0; g < NGF; g++) {{
      const bhah_d2 k = __builtin_nontemporal_load(&K2[g * S2]);
      __builtin_nontemporal_store((1.0 / 3.0) * dt * k + __builtin_nontemporal_load(&YR2[g * S2]), &YR2[g * S2]);
      nx2[g] = a * dt * k + __builtin_nontemporal_load(&YN2[g * S2]);
    }}
  }}
#pragma unroll
  for (int h = 0; h < 2; h++) {{
    const long idx = idx2 + h;
    const int i0 = (int)(idx % p.Nx0), i1 = (int)((idx / p.Nx0) % p.Nx1);
    double nxt[NGF];
#pragma unroll
    for (int g = 0; g < NGF; g++) nxt[g] = nx2[g][h];
{det}
#pragma unroll
    for (int g = 0; g < NGF; g++) nx2[g][h] = nxt[g];
  }}
#pragma unroll
  for (int g = 0; g < NGF; g++) __builtin_nontemporal_store(nx2[g], &OUT2[g * S2]);
}}
"""
    # The same pass with the inner-parity boundary condition of the right-hand side folded in:
    # at an inner ghost point k is the value at its source point with the parity sign
    # (apply_bcs_inner_only__openmp.c:3-57).  kin is only read by this pass, so the value is the
    # one the separate copy would have stored there: the same k, bit for bit.  Each pair of points
    # reads its k from one 16-byte-aligned source pair (its own pair away from the inner ghost
    # zones), so a gridfunction costs one 16-byte load, as in the pass above.
    rkdet += f"""
extern "C" __global__ void __launch_bounds__(256)
bssn_rk_deti2_{tag}(const double *__restrict__ kin, const double *y_n, double *__restrict__ y_run, double *out,
                    const double *__restrict__ rfm0, const double *__restrict__ rfm1, const bhah_kparams p,
                    const int2 *__restrict__ pmap, const int *__restrict__ gfpar) {{
  const long Ntot = p.Ntot;
  const long q = (long)blockIdx.x * 256 + threadIdx.x;  // pair index
  const long idx2 = q * 2;
  if (idx2 >= Ntot) return;
  constexpr int NGF = {ngf};
  const bhah_d2 *K2 = (const bhah_d2 *)kin, *YN2 = (const bhah_d2 *)(y_n + idx2);
  bhah_d2 *YR2 = (bhah_d2 *)(y_run + idx2), *OUT2 = (bhah_d2 *)(out + idx2);
  const long S2 = Ntot / 2;  // gridfunction stride in pairs
  // source pair; code bit 0 / bit 1: point 0 / point 1 takes the source pair's second value;
  // bits 4-15 / 16-27: the parity-type masks of points 0 / 1 (bit t set: type t flips sign)
  const int2 pm = pmap[q];
  const long sp = pm.x;
  const unsigned code = (unsigned)pm.y;
#define BHAH_KPAIR(g)                                                                               \\
  ({{                                                                                               \\
    const bhah_d2 v_ = __builtin_nontemporal_load(&K2[(g) * S2 + sp]);                              \\
    const unsigned par_ = (unsigned)gfpar[g];                                                       \\
    const double a_ = (code & 1u) ? v_.y : v_.x, b_ = (code & 2u) ? v_.y : v_.x;                    \\
    bhah_d2 k_;                                                                                     \\
    k_.x = ((code >> (4 + par_)) & 1u) ? -a_ : a_;                                                  \\
    k_.y = ((code >> (16 + par_)) & 1u) ? -b_ : b_;                                                 \\
    k_;                                                                                             \\
  }})
  const double dt = p.dt;
  bhah_d2 nx2[NGF];
  if (p.substep == 1) {{
#pragma unroll
    for (int g = 0; g < NGF; g++) {{
      const bhah_d2 k = BHAH_KPAIR(g);
      __builtin_nontemporal_store((1.0 / 6.0) * dt * k, &YR2[g * S2]);
      nx2[g] = (1.0 / 2.0) * dt * k + __builtin_nontemporal_load(&YN2[g * S2]);
    }}
  }} else if (p.substep == 4) {{
#pragma unroll
    for (int g = 0; g < NGF; g++) {{
      const bhah_d2 k = BHAH_KPAIR(g);
      nx2[g] = (1.0 / 6.0) * dt * k + __builtin_nontemporal_load(&YN2[g * S2]) + __builtin_nontemporal_load(&YR2[g * S2]);
    }}
  }} else {{
    const double a = (p.substep == 2) ? (1.0 / 2.0) : 1.0;
#pragma unroll
    for (int g = 0; g < NGF; g++) {{
      const bhah_d2 k = BHAH_KPAIR(g);
      __builtin_nontemporal_store((1.0 / 3.0) * dt * k + __builtin_nontemporal_load(&YR2[g * S2]), &YR2[g * S2]);
      nx2[g] = a * dt * k + __builtin_nontemporal_load(&YN2[g * S2]);
    }}
  }}
#undef BHAH_KPAIR
#pragma unroll
  for (int h = 0; h < 2; h++) {{
    const long idx = idx2 + h;
    const int i0 = (int)(idx % p.Nx0), i1 = (int)((idx / p.Nx0) % p.Nx1);
    double nxt[NGF];
#pragma unroll
    for (int g = 0; g < NGF; g++) nxt[g] = nx2[g][h];
{det}
#pragma unroll
    for (int g = 0; g < NGF; g++) nx2[g][h] = nxt[g];
  }}
#pragma unroll
  for (int g = 0; g < NGF; g++) __builtin_nontemporal_store(nx2[g], &OUT2[g * S2]);
}}
"""
    ens_rk = ""
    if ensemble:
        ens_rk = f"""
// Records are (member, chunk): 256 consecutive points of the member's arrays, in array order
// (the single-patch pass's indexing and non-temporal hints).
static __device__ __forceinline__ void bssn_ens_rk_det_{tag}_tile(const double *__restrict__ kin, const double *y_n,
        double *__restrict__ y_run, double *out, const double *__restrict__ rfm0, const double *__restrict__ rfm1,
        const bhah_kparams p, const int chunk) {{
  const long Ntot = p.Ntot;
  const long idx = (long)chunk * 256 + (threadIdx.x + blockDim.x * threadIdx.y);
  if (idx >= Ntot) return;
  const int i0 = (int)(idx % p.Nx0), i1 = (int)((idx / p.Nx0) % p.Nx1);
  constexpr int NGF = {ngf};
  double kk[NGF], ctr[NGF];
#pragma unroll
  for (int g = 0; g < NGF; g++) {{
    kk[g] = __builtin_nontemporal_load(&kin[g * Ntot + idx]);
    ctr[g] = (p.substep == 1) ? __builtin_nontemporal_load(&y_n[g * Ntot + idx]) : 0.0;
  }}
{rk_nt}
{det}
#pragma unroll
  for (int g = 0; g < NGF; g++) __builtin_nontemporal_store(nxt[g], &out[g * Ntot + idx]);
}}
extern "C" __global__ void __launch_bounds__(256)
bssn_ens_rk_det_{tag}(const bhah_member *__restrict__ mem, const bhah_item *__restrict__ items, const int nitems,
                      const int k_sel, const int out_sel, const int substep) {{
  for (int it = blockIdx.x; it < nitems; it += gridDim.x) {{
    const bhah_item w = items[it];
    const bhah_member *M = &mem[w.m];
    bhah_kparams p = M->kp;
    p.substep = substep;
    bssn_ens_rk_det_{tag}_tile(M->arr[k_sel], M->arr[0], M->arr[1], M->arr[out_sel], M->rfm0, M->rfm1, p, w.bx);
  }}
}}
"""
    # ---- det enforcement over all points (standalone ABI call)
    dspec = specs.detgamma_spec(v)
    det_all = f"""
extern "C" __global__ void __launch_bounds__(256)
bssn_detgamma_{tag}(double *__restrict__ gfs, const double *__restrict__ rfm0, const double *__restrict__ rfm1,
                    const bhah_kparams p) {{
{_prologue(False)}
  double nxt[{ngf}];
#pragma unroll
  for (int g = 0; g < {ngf}; g++) nxt[g] = 0.0;
{"".join(f"  nxt[{bssn.EVOL_GFS.index(g)}] = gfs[{bssn.EVOL_GFS.index(g)} * Ntot + idx];{chr(10)}" for g, _ in dspec.outputs)}
{detgamma_inline(v, rats, "nxt")}
{"".join(f"  gfs[{bssn.EVOL_GFS.index(g)} * Ntot + idx] = nxt[{bssn.EVOL_GFS.index(g)}];{chr(10)}" for g, _ in dspec.outputs)}
}}
"""
    # ---- constraints H, MU^i, M^2 at interior points
    kspec = specs.constraints_spec(v)
    lv = [f"diag[{k} * Ntot + idx]" for k in range(len(kspec.outputs))]
    body = _ordered_body(kspec, fd_order, lv, rats)
    cons = f"""
extern "C" __global__ void __launch_bounds__(256)
bssn_constraints_{tag}(const double *__restrict__ in, const double *__restrict__ aux, double *__restrict__ diag,
                       const double *__restrict__ rfm0, const double *__restrict__ rfm1, const bhah_kparams p) {{
{_prologue(True)}
{body}
}}
"""
    # ---- lambda^i of converted initial data from finite differences of hDD, at interior
    # points, in place (reads hDD only, writes lambdaU)
    lspec = specs.initial_lambda_spec(v)
    lv = [f"gfs[{bssn.EVOL_GFS.index(g)} * Ntot + idx]" for g, _ in lspec.outputs]
    body = _ordered_body(lspec, fd_order, lv, rats)
    lam = f"""
extern "C" __global__ void __launch_bounds__(256)
bssn_initial_lambda_{tag}(double *gfs, const double *__restrict__ rfm0, const double *__restrict__ rfm1,
                          const bhah_kparams p) {{
  const double *in = gfs;
{_prologue(True)}
{body}
}}
"""
    cons += lam
    head = [f"// ---- BSSN kernels: {v.coord}, symmetry axes '{v.symmetry_axes}', KO={v.enable_KO}, fd_order={fd_order}"]
    consts = "\n".join(f"#define {k} ({v})" for k, v in sorted(rats.items()))
    undefs = "\n".join(f"#undef {k}" for k in sorted(rats))
    launch = f"""
// Host-side launchers (one translation unit per variant; no relocatable device code needed).
extern "C" hipError_t launch_bssn_rk_det_{tag}(dim3 g, dim3 b, hipStream_t s, const double *kin, const double *y_n,
    double *y_run, double *out, const double *rfm0, const double *rfm1, bhah_kparams p) {{
  (void)g;
  (void)b;
  const bool pairs = !(p.Ntot & 1) && !((reinterpret_cast<uintptr_t>(kin) | reinterpret_cast<uintptr_t>(y_n) |
                                         reinterpret_cast<uintptr_t>(y_run) | reinterpret_cast<uintptr_t>(out)) & 15);
  if (pairs)
    hipLaunchKernelGGL(bssn_rk_det2_{tag}, dim3((unsigned)((p.Ntot / 2 + 255) / 256)), dim3(256), 0, s, kin, y_n, y_run,
                       out, rfm0, rfm1, p);
  else
    hipLaunchKernelGGL(bssn_rk_det_{tag}, dim3((unsigned)((p.Ntot + 255) / 256)), dim3(256), 0, s, kin, y_n, y_run,
                       out, rfm0, rfm1, p);
  return hipGetLastError();
}}
// The RK pass with the inner boundary condition folded in: pairs only (even Ntot, 16-byte aligned
// arrays); a call that does not meet that returns hipErrorInvalidValue and launches nothing.
extern "C" hipError_t launch_bssn_rk_det_inner_{tag}(hipStream_t s, const double *kin, const double *y_n,
    double *y_run, double *out, const double *rfm0, const double *rfm1, bhah_kparams p, const int *pmap,
    const int *gfpar) {{
  if ((p.Ntot & 1) || ((reinterpret_cast<uintptr_t>(kin) | reinterpret_cast<uintptr_t>(y_n) |
                        reinterpret_cast<uintptr_t>(y_run) | reinterpret_cast<uintptr_t>(out) |
                        reinterpret_cast<uintptr_t>(pmap)) & 15))
    return hipErrorInvalidValue;
  hipLaunchKernelGGL(bssn_rk_deti2_{tag}, dim3((unsigned)((p.Ntot / 2 + 255) / 256)), dim3(256), 0, s, kin, y_n, y_run,
                     out, rfm0, rfm1, p, reinterpret_cast<const int2 *>(pmap), gfpar);
  return hipGetLastError();
}}
extern "C" hipError_t launch_bssn_constraints_{tag}(dim3 g, dim3 b, hipStream_t s, const double *in,
    const double *aux, double *diag, const double *rfm0, const double *rfm1, bhah_kparams p) {{
  hipLaunchKernelGGL(bssn_constraints_{tag}, g, b, 0, s, in, aux, diag, rfm0, rfm1, p);
  return hipGetLastError();
}}
extern "C" hipError_t launch_bssn_detgamma_{tag}(dim3 g, dim3 b, hipStream_t s, double *gfs, const double *rfm0,
    const double *rfm1, bhah_kparams p) {{
  hipLaunchKernelGGL(bssn_detgamma_{tag}, g, b, 0, s, gfs, rfm0, rfm1, p);
  return hipGetLastError();
}}
extern "C" hipError_t launch_bssn_initial_lambda_{tag}(dim3 g, dim3 b, hipStream_t s, double *gfs, const double *rfm0,
    const double *rfm1, bhah_kparams p) {{
  hipLaunchKernelGGL(bssn_initial_lambda_{tag}, g, b, 0, s, gfs, rfm0, rfm1, p);
  return hipGetLastError();
}}
"""
    if ensemble:
        launch += f"""extern "C" hipError_t launch_bssn_ens_rk_det_{tag}(dim3 g, dim3 b, hipStream_t s, const bhah_member *mem,
    const bhah_item *items, int nitems, int k_sel, int out_sel, int substep) {{
  hipLaunchKernelGGL(bssn_ens_rk_det_{tag}, g, b, 0, s, mem, items, nitems, k_sel, out_sel, substep);
  return hipGetLastError();
}}
"""
    return "\n".join(head) + "\n" + consts + "\n" + rkdet + ens_rk + det_all + cons + undefs + "\n" + launch


def wave_kernels(fd_order: int) -> str:
    spec = specs.wave_spec()
    rats: Dict[str, str] = {}
    tag = f"fd{fd_order}"
    lv = [f"rhs[{spec.gfs.index(g)} * Ntot + idx]" for g, _ in spec.outputs]
    body = _ordered_body(spec, fd_order, lv, rats)
    lvk = [f"kk[{spec.gfs.index(g)}]" for g, _ in spec.outputs]
    bodyk = _ordered_body(spec, fd_order, lvk, rats)
    consts = "\n".join(f"#define {k} ({v})" for k, v in sorted(rats.items()))
    undefs = "\n".join(f"#undef {k}" for k in sorted(rats))
    return f"""// ---- scalar wave kernels, Cartesian, fd_order={fd_order}
{consts}
extern "C" __global__ void __launch_bounds__(256)
wave_rhs_{tag}(const double *__restrict__ in, double *__restrict__ rhs, const bhah_kparams p) {{
{_prologue(True)}
{body}
}}

extern "C" __global__ void __launch_bounds__(256)
wave_rhs_rk_{tag}(const double *__restrict__ in, const double *y_n, double *__restrict__ y_run,
                  double *out, const bhah_kparams p) {{
{_prologue(True)}
  constexpr int NGF = 2;
  double kk[NGF];
{bodyk}
  double ctr[NGF];
#pragma unroll
  for (int g = 0; g < NGF; g++) ctr[g] = (p.substep == 1) ? in[g * Ntot + idx] : 0.0;
{RK_EPILOGUE}
#pragma unroll
  for (int g = 0; g < NGF; g++) out[g * Ntot + idx] = nxt[g];
}}
{undefs}

extern "C" hipError_t launch_wave_rhs_{tag}(dim3 g, dim3 b, hipStream_t s, const double *in, double *rhs, bhah_kparams p) {{
  hipLaunchKernelGGL(wave_rhs_{tag}, g, b, 0, s, in, rhs, p);
  return hipGetLastError();
}}
extern "C" hipError_t launch_wave_rhs_rk_{tag}(dim3 g, dim3 b, hipStream_t s, const double *in, const double *y_n,
    double *y_run, double *out, bhah_kparams p) {{
  hipLaunchKernelGGL(wave_rhs_rk_{tag}, g, b, 0, s, in, y_n, y_run, out, p);
  return hipGetLastError();
}}
"""


# (variant, fd order, candidate schedules, batched-ensemble kernels too)
BSSN_BUILDS = [
    (specs.VARIANTS[0], 4, SMALL_CANDIDATES, True),
    (specs.VARIANTS[1], 6, FULL_CANDIDATES, False),
    (specs.VARIANTS[1], 4, SMALL_CANDIDATES, False),
]
WAVE_BUILDS = [4]
COORD_ENUM = {"Cartesian": 0, "Spherical": 1, "SinhSpherical": 2}


def kernel_name_tag(v: specs.Variant, fd_order: int) -> str:
    return f"{v.tag}_fd{fd_order}"


def generate(outdir: str) -> List[str]:
    import json

    os.makedirs(outdir, exist_ok=True)
    for f in os.listdir(outdir):
        if f.endswith(".hip") or f.endswith(".inc"):
            os.remove(os.path.join(outdir, f))
    paths = []
    hdr = "// Generated by nrpy_amd/codegen/hip_gen.py -- do not edit.\n#include <hip/hip_runtime.h>\n#include <math.h>\n" + KPARAMS
    meta = {"flops_per_point": {}, "candidates": {}}
    table_decl, table_rows = [], []
    from .adjoint_gen import ADJOINT_BUILDS as adjoint_builds

    for v, order, cands, ens in BSSN_BUILDS:
        tag = kernel_name_tag(v, order)
        path = os.path.join(outdir, f"bssn_{tag}.hip")
        with open(path, "w", encoding="utf-8") as f:
            f.write(hdr + bssn_kernels(v, order, ensemble=ens))
        paths.append(path)
        arrays, ncand = {}, {}
        for kind in ("rhs", "ricci"):
            rows = []
            kcands = list(cands[kind])
            for cand in kcands:
                src, fl = candidate_tu(v, order, kind, cand, ensemble=ens)
                path = os.path.join(outdir, f"bssn_{tag}_{kind}_{cand}.hip")
                with open(path, "w", encoding="utf-8") as f:
                    f.write(hdr + src)
                paths.append(path)
                meta["candidates"][f"{kind}_{tag}_{cand}"] = fl
                table_decl.append(f'extern "C" hipError_t launch_bssn_{kind}_{tag}_{cand}({CAND_SIG});')
                geo = next((_lds_geometry(o) for o in cand.split("_") if o[:1] == "L" and o[1:2].isdigit()), None)
                tiles = f"{{{geo[0]}, {geo[1]}, {geo[2]}}}" if geo else "{0, 0, 0}"
                if ens:
                    table_decl.append(f'extern "C" hipError_t launch_bssn_ens_{kind}_{tag}_{cand}({ENS_SIG});')
                    rows.append(f'{{"{cand}", launch_bssn_{kind}_{tag}_{cand}, launch_bssn_ens_{kind}_{tag}_{cand}, '
                                f'{tiles}}}')
                else:
                    rows.append(f'{{"{cand}", launch_bssn_{kind}_{tag}_{cand}, nullptr, {tiles}}}')
            arrays[kind] = f"kCand_{kind}_{tag}"
            ncand[kind] = len(kcands)
            table_decl.append(f"static const BssnCandidate {arrays[kind]}[] = {{{', '.join(rows)}}};")
        sym_mask = sum(1 << int(c) for c in v.symmetry_axes)
        table_rows.append(
            f"    {{{COORD_ENUM[v.coord]}, {sym_mask}, {int(v.enable_KO)}, {order}, \"{tag}\", "
            f"{arrays['rhs']}, {ncand['rhs']}, {arrays['ricci']}, {ncand['ricci']}, "
            f"launch_bssn_rk_det_{tag}, launch_bssn_rk_det_inner_{tag}, launch_bssn_detgamma_{tag}, "
            f"launch_bssn_constraints_{tag}, "
            f"launch_bssn_initial_lambda_{tag}, "
            f"{f'launch_bssn_ens_rk_det_{tag}' if ens else 'nullptr'}, "
            + (f"adj_nslots_{tag}, launch_adj_rhs_{tag}, launch_adj_ricci_{tag}, launch_adj_final_{tag}, "
               f"launch_adj_stage_{tag}, launch_adj_gather_{tag}}},"
               if (v, order) in adjoint_builds else "nullptr, nullptr, nullptr, nullptr, nullptr, nullptr},")
        )
        if (v, order) in adjoint_builds:
            from . import adjoint_gen

            path = os.path.join(outdir, f"adjoint_{tag}.hip")
            with open(path, "w", encoding="utf-8") as f:
                f.write(hdr + adjoint_gen.adjoint_tu(v, order))
            paths.append(path)
            table_decl += [
                f'extern "C" int adj_nslots_{tag}(void);',
                f'extern "C" hipError_t launch_adj_rhs_{tag}(dim3, dim3, hipStream_t, const double *, const double *, '
                f"const double *, double *, const double *, const double *, bhah_kparams);",
                f'extern "C" hipError_t launch_adj_ricci_{tag}(dim3, dim3, hipStream_t, const double *, double *, '
                f"const double *, const double *, bhah_kparams);",
                f'extern "C" hipError_t launch_adj_final_{tag}(dim3, dim3, hipStream_t, const double *, '
                f"const double *, double, double, double, double, const double *, double *, double *, "
                f"const double *, const double *, bhah_kparams);",
                f'extern "C" hipError_t launch_adj_stage_{tag}(dim3, dim3, hipStream_t, const double *, '
                f"const double *, double, const double *, double *, double *, const double *, const double *, "
                f"bhah_kparams);",
                f'extern "C" hipError_t launch_adj_gather_{tag}(dim3, dim3, hipStream_t, const double *, double *, '
                f"bhah_kparams);",
            ]
        if ens:
            table_decl.append(
                f'extern "C" hipError_t launch_bssn_ens_rk_det_{tag}(dim3, dim3, hipStream_t, const bhah_member *, '
                f"const bhah_item *, int, int, int, int);"
            )
        table_decl.append(
            f'extern "C" hipError_t launch_bssn_rk_det_{tag}(dim3, dim3, hipStream_t, const double *, const double *, '
            f"double *, double *, const double *, const double *, bhah_kparams);"
        )
        table_decl.append(
            f'extern "C" hipError_t launch_bssn_rk_det_inner_{tag}(hipStream_t, const double *, const double *, '
            f"double *, double *, const double *, const double *, bhah_kparams, const int *, const int *);"
        )
        table_decl.append(
            f'extern "C" hipError_t launch_bssn_detgamma_{tag}(dim3, dim3, hipStream_t, double *, const double *, '
            f"const double *, bhah_kparams);"
        )
        table_decl.append(
            f'extern "C" hipError_t launch_bssn_constraints_{tag}(dim3, dim3, hipStream_t, const double *, '
            f"const double *, double *, const double *, const double *, bhah_kparams);"
        )
        table_decl.append(
            f'extern "C" hipError_t launch_bssn_initial_lambda_{tag}(dim3, dim3, hipStream_t, double *, '
            f"const double *, const double *, bhah_kparams);"
        )
    for order in WAVE_BUILDS:
        path = os.path.join(outdir, f"wave_fd{order}.hip")
        with open(path, "w", encoding="utf-8") as f:
            f.write(hdr + wave_kernels(order))
        paths.append(path)
    with open(os.path.join(outdir, "kernel_table.inc"), "w", encoding="utf-8") as f:
        f.write("// Generated by nrpy_amd/codegen/hip_gen.py -- do not edit.\n")
        f.write("\n".join(table_decl) + "\n")
        f.write("static const BssnKernels kBssnKernels[] = {\n" + "\n".join(table_rows) + "\n};\n")
    meta["flops_per_point"] = dict(sorted(FLOPS.items()))
    with open(os.path.join(outdir, "kernel_meta.json"), "w", encoding="utf-8") as f:
        json.dump(meta, f, indent=1, sort_keys=True)
    return paths


if __name__ == "__main__":
    here = os.path.dirname(os.path.abspath(__file__))
    for p in generate(os.path.join(os.path.dirname(here), "csrc", "gen")):
        print("wrote", p)
