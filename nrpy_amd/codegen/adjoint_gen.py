"""Emit reverse-mode (adjoint) HIP kernels of one BSSN RK4 step.

The forward step is Ricci_eval -> rhs_eval -> boundary conditions -> RK4 update
with det(gammabar) enforcement.  Its adjoint is assembled from:

* ``adj_rhs`` / ``adj_ricci``: per interior point, the transposed Jacobian of
  the pointwise algebra applied to the output adjoints.  The algebra is the same
  SymPy transcription the forward kernels are generated from; its CSE'd
  straight-line form is differentiated in reverse order (each temporary's
  adjoint is complete before it is pushed to its arguments), so the kernels
  compute exact derivatives of the discrete right-hand side.  Every input
  symbol (centre value, auxiliary Ricci value, FD derivative) gets an adjoint
  "slot" field.  Upwinded derivatives get two slots, split by the sign of
  beta^i at the point (the stencil choice is piecewise constant in beta, so it
  contributes no derivative of its own).
* ``adj_gather``: the transpose of every FD stencil as a gather: at each point
  q and gridfunction g, the sum over the slots of g's derivatives of
  weight x slot(q - offset).  Table-driven (slot, gridfunction, taps).
* ``adj_det``: the transposed Jacobian of det(gammabar) enforcement, pointwise.

Boundary-condition and RK-combination adjoints are hand-written in the host
runtime (they are linear maps with index lists).
"""

from __future__ import annotations

from typing import Dict, List, Tuple

import sympy as sp

from ..symbolic import bssn
from . import emit, fd, specs
from .hip_gen import RFM_EXPR, PARAM_EXPR, StencilEmitter, _dbl, _upwind_prologue, kernel_name_tag

# Builds with adjoint kernels: the axisymmetric Spherical 4th-order variant.
ADJOINT_BUILDS = [(specs.VARIANTS[0], 4)]


class SlotTable:
    """Adjoint slots: one field per input symbol of the pointwise algebra."""

    def __init__(self):
        self.slots: List[dict] = []

    def add(self, gf: int, taps: List[Tuple[Tuple[int, int, int], sp.Rational, Tuple[int, int, int]]]) -> int:
        """taps: (offset, rational weight, powers of invdxx0..2)."""
        self.slots.append({"gf": gf, "taps": taps})
        return len(self.slots) - 1


def _deriv_taps(d: fd.Deriv, fd_order: int, variant: str = "") -> List[tuple]:
    """Stencil taps of one derivative symbol (offset, weight, invdxx powers)."""
    if d.op == "dD":
        k = d.dirs[0]
        out = []
        for o, w in fd.weights_1d("dD", fd_order):
            off = [0, 0, 0]
            off[k] = o
            pw = [0, 0, 0]
            pw[k] = 1
            out.append((tuple(off), sp.Rational(w), tuple(pw)))
        return out
    if d.op == "dKOD":
        k = d.dirs[0]
        out = []
        for o, w in fd.weights_1d("dKOD", fd_order):
            off = [0, 0, 0]
            off[k] = o
            pw = [0, 0, 0]
            pw[k] = 1
            out.append((tuple(off), sp.Rational(w), tuple(pw)))
        return out
    if d.op == "dDD":
        a, b = d.dirs
        if a == b:
            out = []
            for o, w in fd.weights_1d("dDD", fd_order):
                off = [0, 0, 0]
                off[a] = o
                pw = [0, 0, 0]
                pw[a] = 2
                out.append((tuple(off), sp.Rational(w), tuple(pw)))
            return out
        out = []
        for oa, wa in fd.weights_1d("dD", fd_order):
            for ob, wb in fd.weights_1d("dD", fd_order):
                off = [0, 0, 0]
                off[a] += oa
                off[b] += ob
                pw = [0, 0, 0]
                pw[a] += 1
                pw[b] += 1
                out.append((tuple(off), sp.Rational(wa) * sp.Rational(wb), tuple(pw)))
        return out
    if d.op == "dupD":
        k = d.dirs[0]
        out = []
        for o, w in fd.weights_1d("dupD" if variant == "up" else "ddnD", fd_order):
            off = [0, 0, 0]
            off[k] = o
            pw = [0, 0, 0]
            pw[k] = 1
            out.append((tuple(off), sp.Rational(w), tuple(pw)))
        return out
    raise ValueError(d.op)


def _reverse_sweep(outputs: List[Tuple[str, sp.Expr]], lam: Dict[str, str], pr, prefix: str):
    """Forward temporaries, then the reverse accumulation.  Returns
    (forward lines, reverse lines, {input symbol name: adjoint variable})."""
    exprs = [e.replace(sp.Function("nrpyAbs"), sp.Abs) for _, e in outputs]
    replaced, reduced = sp.cse(exprs, symbols=sp.numbered_symbols(prefix), order="none")
    fwd = [f"  const double {t} = {pr.doprint(e)};" for t, e in replaced]
    temps = {t for t, _ in replaced}
    rev: List[str] = []
    acc: Dict[str, str] = {}

    def add(sym, term: str):
        n = str(sym)
        var = f"a_{n}"
        if n not in acc:
            acc[n] = var
            rev.append(f"  double {var} = {term};")
        else:
            rev.append(f"  {var} += {term};")

    for (g, _), e in zip(outputs, reduced):
        for a in sorted(e.free_symbols, key=str):
            dd = sp.diff(e, a)
            if dd != 0:
                add(a, f"{lam[g]} * ({pr.doprint(dd)})")
    for t, e in reversed(replaced):
        if str(t) not in acc:
            continue
        for a in sorted(e.free_symbols, key=str):
            dd = sp.diff(e, a)
            if dd != 0:
                add(a, f"a_{t} * ({pr.doprint(dd)})")
    inputs = {n: v for n, v in acc.items() if sp.Symbol(n) not in temps and n not in {str(t) for t in temps}}
    return fwd, rev, inputs


def _input_defs(spec: specs.Spec, fd_order: int, names) -> List[str]:
    st = StencilEmitter(spec, fd_order)
    centre, aux, derivs, rfm, params = spec.classify()
    dmap = {d.symbol: d for d in derivs}
    lines = []
    for n in sorted(names):
        if n in centre:
            lines.append(f"  const double {n} = {st.load(n, (0, 0, 0))};")
        elif n in aux:
            lines.append(f"  const double {n} = aux[{spec.aux.index(n)} * Ntot + idx];")
        elif n in dmap:
            lines.append(f"  const double {n} = {st.expr(dmap[n])};")
        elif n in specs.RFM_SYMBOLS:
            lines.append(f"  const double {n} = {RFM_EXPR[specs.RFM_SYMBOLS.index(n)]};")
        elif n in specs.PARAM_SYMBOLS:
            lines.append(f"  const double {n} = {PARAM_EXPR[n]};")
        else:
            raise ValueError(n)
    return lines


PROLOGUE = """  const int i0 = p.NG + (int)(blockIdx.x * blockDim.x + threadIdx.x);
  const int i1 = p.NG + (int)(blockIdx.y * blockDim.y + threadIdx.y);
  const int i2 = p.NG + (int)blockIdx.z;
  if (i0 >= p.Nx0 - p.NG || i1 >= p.Nx1 - p.NG) return;
  const long s1 = p.s1, s2 = p.s2, Ntot = p.Ntot;
  const long idx = i0 + s1 * i1 + s2 * i2;
"""


def _slots_for_inputs(spec: specs.Spec, fd_order: int, inputs: Dict[str, str], table: SlotTable,
                      skip_aux: bool) -> Tuple[List[str], Dict[str, int]]:
    """Store lines writing each input adjoint to its slot(s); returns (lines, aux slot map)."""
    centre, aux, derivs, rfm, params = spec.classify()
    dmap = {d.symbol: d for d in derivs}
    lines, aux_slots = [], {}
    for n in sorted(inputs):
        v = inputs[n]
        if n in centre:
            k = table.add(spec.gfs.index(n), [((0, 0, 0), sp.Integer(1), (0, 0, 0))])
            lines.append(f"  S[{k} * Ntot + idx] = {v};")
        elif n in aux:
            k = table.add(-1, [])
            aux_slots[n] = k
            lines.append(f"  S[{k} * Ntot + idx] = {v};")
        elif n in dmap:
            d = dmap[n]
            g = spec.gfs.index(d.gf)
            if d.op == "dupD":
                ku = table.add(g, _deriv_taps(d, fd_order, "up"))
                kd = table.add(g, _deriv_taps(d, fd_order, "dn"))
                c = f"(betaU{d.dirs[0]} > 0.0)"
                lines.append(f"  S[{ku} * Ntot + idx] = {c} ? {v} : 0.0;")
                lines.append(f"  S[{kd} * Ntot + idx] = {c} ? 0.0 : {v};")
            else:
                k = table.add(g, _deriv_taps(d, fd_order))
                lines.append(f"  S[{k} * Ntot + idx] = {v};")
        # rfm and parameters are constants of the step: no adjoint
    return lines, aux_slots


def adjoint_tu(v: specs.Variant, fd_order: int) -> str:
    tag = kernel_name_tag(v, fd_order)
    rats: Dict[str, str] = {}
    pr = emit.DoublePrinter(rats)
    table = SlotTable()

    # ---- rhs_eval adjoint
    rspec = specs.bssn_rhs_spec(v)
    lam = {g: f"lam[{bssn.EVOL_GFS.index(g)} * Ntot + idx]" for g, _ in rspec.outputs}
    fwd, rev, inputs = _reverse_sweep(rspec.outputs, lam, pr, "t")
    used = set()
    for _, e in rspec.outputs:
        used |= {str(s) for s in e.free_symbols}
    defs = _input_defs(rspec, fd_order, used)
    stores, aux_slots = _slots_for_inputs(rspec, fd_order, inputs, table, skip_aux=False)
    upw = _upwind_prologue(rspec, fd_order, rats, tables=True)
    adj_rhs = f"""
extern "C" __global__ void __launch_bounds__(256)
adj_rhs_{tag}(const double *__restrict__ in, const double *__restrict__ aux, const double *__restrict__ lam,
              double *__restrict__ S, const double *__restrict__ rfm0, const double *__restrict__ rfm1,
              const bhah_kparams p) {{
{PROLOGUE}
{upw}
{chr(10).join(defs)}
{chr(10).join(fwd)}
{chr(10).join(rev)}
{chr(10).join(stores)}
}}
"""
    # ---- Ricci_eval adjoint: output adjoints are the auxiliary slots written above
    qspec = specs.ricci_spec(v)
    qlam = {g: f"S[{aux_slots[g]} * Ntot + idx]" if g in aux_slots else "0.0" for g, _ in qspec.outputs}
    qfwd, qrev, qinputs = _reverse_sweep(qspec.outputs, qlam, pr, "u")
    qused = set()
    for _, e in qspec.outputs:
        qused |= {str(s) for s in e.free_symbols}
    qdefs = _input_defs(qspec, fd_order, qused)
    qstores, _ = _slots_for_inputs(qspec, fd_order, qinputs, table, skip_aux=True)
    adj_ricci = f"""
extern "C" __global__ void __launch_bounds__(256)
adj_ricci_{tag}(const double *__restrict__ in, double *__restrict__ S, const double *__restrict__ rfm0,
                const double *__restrict__ rfm1, const bhah_kparams p) {{
{PROLOGUE}
{chr(10).join(qdefs)}
{chr(10).join(qfwd)}
{chr(10).join(qrev)}
{chr(10).join(qstores)}
}}
"""
    # ---- det(gammabar) enforcement adjoint, all points, in place on lam
    dspec = specs.detgamma_spec(v)
    dlam = {g: f"l_{g}" for g, _ in dspec.outputs}
    dfwd, drev, dinputs = _reverse_sweep(dspec.outputs, dlam, pr, "w")
    dcentre, _, _, drfm, _ = dspec.classify()
    dlines = [f"  const double {n} = u[{bssn.EVOL_GFS.index(n)} * Ntot + idx];" for n in dcentre]
    dlines += [f"  const double {n} = {RFM_EXPR[specs.RFM_SYMBOLS.index(n)]};" for n in drfm]
    dlines += [f"  const double l_{g} = lam[{bssn.EVOL_GFS.index(g)} * Ntot + idx];" for g, _ in dspec.outputs]
    dstore = [f"  lam[{bssn.EVOL_GFS.index(g)} * Ntot + idx] = {dinputs.get(g, '0.0')};" for g, _ in dspec.outputs]
    adj_det = f"""
extern "C" __global__ void __launch_bounds__(256)
adj_det_{tag}(const double *__restrict__ u, double *__restrict__ lam, const double *__restrict__ rfm0,
              const double *__restrict__ rfm1, const bhah_kparams p) {{
  const int i0 = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  const int i1 = (int)(blockIdx.y * blockDim.y + threadIdx.y);
  const int i2 = (int)blockIdx.z;
  if (i0 >= p.Nx0 || i1 >= p.Nx1) return;
  const long Ntot = p.Ntot;
  const long idx = i0 + p.s1 * i1 + p.s2 * i2;
{chr(10).join(dlines)}
{chr(10).join(dfwd)}
{chr(10).join(drev)}
{chr(10).join(dstore)}
}}
"""
    # ---- transposed stencils: table of (gridfunction -> slots -> taps)
    by_gf: Dict[int, List[int]] = {}
    for k, s in enumerate(table.slots):
        if s["gf"] >= 0:
            by_gf.setdefault(s["gf"], []).append(k)
    gf_start, slot_ids, tap_start, tap_off, tap_w, tap_pw = [], [], [], [], [], []
    for g in range(len(bssn.EVOL_GFS)):
        gf_start.append(len(slot_ids))
        for k in by_gf.get(g, []):
            slot_ids.append(k)
            tap_start.append(len(tap_w))
            for off, w, pw in table.slots[k]["taps"]:
                tap_off.append(off)
                tap_w.append(w)
                tap_pw.append(pw)
    gf_start.append(len(slot_ids))
    tap_start.append(len(tap_w))

    def arr(name, ctype, vals):
        return f"__constant__ {ctype} {name}[{max(len(vals), 1)}] = {{{', '.join(vals) if vals else '0'}}};"

    tables = "\n".join([
        arr(f"adjg_gf_start_{tag}", "int", [str(x) for x in gf_start]),
        arr(f"adjg_slot_{tag}", "int", [str(x) for x in slot_ids]),
        arr(f"adjg_tap_start_{tag}", "int", [str(x) for x in tap_start]),
        arr(f"adjg_tap_off_{tag}", "short", [str(c) for o in tap_off for c in o]),
        arr(f"adjg_tap_w_{tag}", "double", [_dbl(w) for w in tap_w]),
        arr(f"adjg_tap_pw_{tag}", "signed char", [str(c) for o in tap_pw for c in o]),
    ])
    gather = f"""
{tables}
// lam_y[g](q) += sum over slots k of g, taps t: w_t * invdxx^pw_t * S_k(q - off_t), S_k being zero
// outside the interior.  One thread per (point, gridfunction); all points.
extern "C" __global__ void __launch_bounds__(256)
adj_gather_{tag}(const double *__restrict__ S, double *__restrict__ lam_y, const bhah_kparams p) {{
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int g = (int)(t / p.Ntot);
  if (g >= {len(bssn.EVOL_GFS)}) return;
  const long q = t - (long)g * p.Ntot;
  const int i0 = (int)(q % p.Nx0), i1 = (int)((q / p.Nx0) % p.Nx1), i2 = (int)(q / p.s2);
  const double inv[3] = {{p.invdxx0, p.invdxx1, p.invdxx2}};
  double acc = 0.0;
  for (int a = adjg_gf_start_{tag}[g]; a < adjg_gf_start_{tag}[g + 1]; a++) {{
    const int k = adjg_slot_{tag}[a];
    for (int b = adjg_tap_start_{tag}[a]; b < adjg_tap_start_{tag}[a + 1]; b++) {{
      const int j0 = i0 - adjg_tap_off_{tag}[3 * b], j1 = i1 - adjg_tap_off_{tag}[3 * b + 1],
                j2 = i2 - adjg_tap_off_{tag}[3 * b + 2];
      if (j0 < p.NG || j0 >= p.Nx0 - p.NG || j1 < p.NG || j1 >= p.Nx1 - p.NG || j2 < p.NG || j2 >= p.Nx2 - p.NG)
        continue;
      double w = adjg_tap_w_{tag}[b];
      for (int d = 0; d < 3; d++)
        for (int e = 0; e < adjg_tap_pw_{tag}[3 * b + d]; e++) w *= inv[d];
      acc += w * S[(long)k * p.Ntot + j0 + p.s1 * j1 + p.s2 * j2];
    }}
  }}
  lam_y[t] += acc;
}}
"""
    nslots = len(table.slots)
    consts = "\n".join(f"#define {k} ({val})" for k, val in sorted(rats.items()))
    undefs = "\n".join(f"#undef {k}" for k in sorted(rats))
    launch = f"""
extern "C" int adj_nslots_{tag}(void) {{ return {nslots}; }}
extern "C" hipError_t launch_adj_rhs_{tag}(dim3 g, dim3 b, hipStream_t s, const double *in, const double *aux,
    const double *lam, double *S, const double *rfm0, const double *rfm1, bhah_kparams p) {{
  hipLaunchKernelGGL(adj_rhs_{tag}, g, b, 0, s, in, aux, lam, S, rfm0, rfm1, p);
  return hipGetLastError();
}}
extern "C" hipError_t launch_adj_ricci_{tag}(dim3 g, dim3 b, hipStream_t s, const double *in, double *S,
    const double *rfm0, const double *rfm1, bhah_kparams p) {{
  hipLaunchKernelGGL(adj_ricci_{tag}, g, b, 0, s, in, S, rfm0, rfm1, p);
  return hipGetLastError();
}}
extern "C" hipError_t launch_adj_det_{tag}(dim3 g, dim3 b, hipStream_t s, const double *u, double *lam,
    const double *rfm0, const double *rfm1, bhah_kparams p) {{
  hipLaunchKernelGGL(adj_det_{tag}, g, b, 0, s, u, lam, rfm0, rfm1, p);
  return hipGetLastError();
}}
extern "C" hipError_t launch_adj_gather_{tag}(dim3 g, dim3 b, hipStream_t s, const double *S, double *lam_y,
    bhah_kparams p) {{
  hipLaunchKernelGGL(adj_gather_{tag}, g, b, 0, s, S, lam_y, p);
  return hipGetLastError();
}}
"""
    head = f"// ---- adjoint of the BSSN step: {v.coord}, symmetry axes '{v.symmetry_axes}', fd_order={fd_order}; " \
           f"{nslots} adjoint slots\n"
    return head + consts + "\n" + adj_rhs + adj_ricci + adj_det + gather + undefs + "\n" + launch
