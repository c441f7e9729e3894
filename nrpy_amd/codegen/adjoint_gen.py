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
  contributes no derivative of its own).  Each kernel is split into output
  groups that run as the waves of one workgroup and sum their input adjoints
  through LDS in a fixed order (``_grouped_adjoint``).
* ``adj_gather``: the transpose of every FD stencil as a gather: at each point
  q and gridfunction g, the sum over the slots of g's derivatives of
  weight x slot(q - offset), unrolled per gridfunction with literal offsets
  and weights.
* ``adj_final`` / ``adj_stage``: the RK combinations with det(gammabar) enforcement,
  reversed pointwise (the transposed Jacobian of the enforcement at the combination,
  then the scatter to y_n and the stage right-hand sides).

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
        self.aux_slots: set = set()

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


def _alloc_slots(spec: specs.Spec, fd_order: int, names: List[str], table: "SlotTable") -> Dict[str, tuple]:
    """One slot per input symbol (two for an upwinded derivative: up, down, split by the sign of
    beta^i at the point); auxiliary inputs are recorded in table.aux_slots."""
    centre, aux, derivs, rfm, params = spec.classify()
    dmap = {d.symbol: d for d in derivs}
    out: Dict[str, tuple] = {}
    for n in names:
        if n in centre:
            out[n] = (table.add(spec.gfs.index(n), [((0, 0, 0), sp.Integer(1), (0, 0, 0))]),)
        elif n in aux:
            k = table.add(-1, [])
            table.aux_slots.add(k)
            out[n] = (k,)
        elif n in dmap:
            d = dmap[n]
            g = spec.gfs.index(d.gf)
            if d.op == "dupD":
                out[n] = (table.add(g, _deriv_taps(d, fd_order, "up")), table.add(g, _deriv_taps(d, fd_order, "dn")))
            else:
                out[n] = (table.add(g, _deriv_taps(d, fd_order)),)
    return out


def _deriv_dir(name: str) -> int:
    return fd.parse_deriv(name).dirs[0]


# Output groups of the adjoint kernels.  A group's reverse sweep needs the forward temporaries
# of its own outputs only, so each is a fraction of the whole sweep (which keeps over a thousand
# values live and spills); the groups run side by side.  betU's right-hand side contains
# lambdaU's (the Gamma-driver), so the rhs groups are {betU}, {lambdaU, vetU, alpha}, {the rest}.
RHS_GROUPS = [("betU",), ("lambdaU", "vetU", "alpha"), ("aDD", "hDD", "cf", "trK")]
RICCI_GROUPS = [("RbarDD00", "RbarDD01"), ("RbarDD02", "RbarDD11"), ("RbarDD12", "RbarDD22")]


def _groups(spec: specs.Spec, prefixes) -> List[List[str]]:
    names = [g for g, _ in spec.outputs]
    groups = [[n for n in names if any(n == pre or (n.startswith(pre) and n[len(pre):].isdigit()) for pre in grp)]
              for grp in prefixes]
    if sorted(sum(groups, [])) != sorted(names):
        raise ValueError(f"adjoint groups {groups} do not partition the outputs {names}")
    return [g for g in groups if g]


def _grouped_adjoint(kname: str, params: str, spec: specs.Spec, groups: List[List[str]], lam: Dict[str, str],
                     fd_order: int, pr, table: "SlotTable", prologue: str) -> Tuple[str, Dict[str, tuple]]:
    """The adjoint of one pointwise kernel, split by output group.  A workgroup is G waves over
    the same 64 points of an xx0 row, wave g running group g's reverse sweep (each wave on its
    own SIMD).  Each wave leaves its input adjoints in LDS; after one barrier the sums over the
    groups, in group order, are stored to one slot per input symbol (two for an upwinded
    derivative, by the sign of beta^i), so the result is deterministic and the transposed
    stencils read each input's adjoint once.  Returns (kernel source, {input: slots})."""
    G = len(groups)
    swept = []
    all_inputs: set = set()
    for gi, names in enumerate(groups):
        outs = [(g, e) for g, e in spec.outputs if g in names]
        fwd, rev, inputs = _reverse_sweep(outs, lam, pr, f"g{gi}t")
        used = set()
        for _, e in outs:
            used |= {str(x) for x in e.free_symbols}
        swept.append((names, _input_defs(spec, fd_order, used), fwd, rev, inputs))
        all_inputs |= set(inputs)
    slot_of = _alloc_slots(spec, fd_order, sorted(all_inputs), table)
    contrib: Dict[int, List[int]] = {}  # slot -> LDS rows holding its partial sums, group order
    cases, row = [], 0
    for gi, (names, defs, fwd, rev, inputs) in enumerate(swept):
        writes = []
        for n in sorted(inputs):
            if n not in slot_of:  # reference-metric values and parameters: constants of the step
                continue
            ks, val = slot_of[n], inputs[n]
            if len(ks) == 2:
                c = f"(betaU{_deriv_dir(n)} > 0.0)"
                parts = [(ks[0], f"{c} ? {val} : 0.0"), (ks[1], f"{c} ? 0.0 : {val}")]
            else:
                parts = [(ks[0], val)]
            for k, e in parts:
                writes.append(f"  red[{row} * 64 + lane] = {e};")
                contrib.setdefault(k, []).append(row)
                row += 1
        body = "\n".join(defs + fwd + rev + writes)
        cases.append(f"  case {gi}: {{ // {', '.join(names)}\n{body}\n  break;\n  }}")
    slots = sorted(contrib)
    store_cases = []
    for gi in range(G):
        lines = [f"    S[{k} * Ntot + idx] = " + " + ".join(f"red[{r} * 64 + lane]" for r in contrib[k]) + ";"
                 for j, k in enumerate(slots) if j % G == gi]
        store_cases.append(f"  case {gi}:\n" + "\n".join(lines) + "\n    break;")
    src = f"""
// {G} output groups, one wave each over the same 64 points; {row} LDS rows of partial sums
extern "C" __global__ void __launch_bounds__({64 * G})
{kname}({params}) {{
  __shared__ double red[{row} * 64];
  const int lane = (int)threadIdx.x, group = (int)threadIdx.y;
  const int i0r = p.NG + (int)blockIdx.x * 64 + lane;
  const bool active = i0r < p.Nx0 - p.NG;
  const int i0 = active ? i0r : p.Nx0 - p.NG - 1;  // idle lanes repeat the row's last point
  const int i1 = p.NG + (int)blockIdx.y;
  const int i2 = p.NG + (int)blockIdx.z;
  const long s1 = p.s1, s2 = p.s2, Ntot = p.Ntot;
  const long idx = i0 + s1 * i1 + s2 * i2;
{prologue}
  switch (group) {{
{chr(10).join(cases)}
  default:
    break;
  }}
  __syncthreads();
  if (!active) return;
  switch (group) {{
{chr(10).join(store_cases)}
  default:
    break;
  }}
}}
// workgroups of 64 points of one xx0 row x {G} output groups
static void {kname}_dims(const bhah_kparams &p, dim3 &g, dim3 &b) {{
  g = dim3((unsigned)((p.Nx0 - 2 * p.NG + 63) / 64), (unsigned)(p.Nx1 - 2 * p.NG), (unsigned)(p.Nx2 - 2 * p.NG));
  b = dim3(64, {G});
}}
"""
    return src, slot_of


def adjoint_tu(v: specs.Variant, fd_order: int) -> str:
    tag = kernel_name_tag(v, fd_order)
    rats: Dict[str, str] = {}
    pr = emit.DoublePrinter(rats)
    table = SlotTable()

    # ---- rhs_eval adjoint, in output groups (see _grouped_adjoint)
    rspec = specs.bssn_rhs_spec(v)
    lam = {g: f"lam[{bssn.EVOL_GFS.index(g)} * Ntot + idx]" for g, _ in rspec.outputs}
    upw = _upwind_prologue(rspec, fd_order, rats, tables=True)
    adj_rhs, slot_of = _grouped_adjoint(
        f"adj_rhs_{tag}", "const double *__restrict__ in, const double *__restrict__ aux, const double *__restrict__ lam,\n"
        "    double *__restrict__ S, const double *__restrict__ rfm0, const double *__restrict__ rfm1, const bhah_kparams p",
        rspec, _groups(rspec, RHS_GROUPS), lam, fd_order, pr, table, upw)
    aux_slot = {n: k[0] for n, k in slot_of.items() if k[0] in table.aux_slots}
    # ---- Ricci_eval adjoint: its output adjoints are the auxiliary slots written above
    qspec = specs.ricci_spec(v)
    qlam = {g: f"S[{aux_slot[g]} * Ntot + idx]" if g in aux_slot else "0.0" for g, _ in qspec.outputs}
    adj_ricci, _ = _grouped_adjoint(
        f"adj_ricci_{tag}", "const double *__restrict__ in, double *__restrict__ S, const double *__restrict__ rfm0,\n"
        "    const double *__restrict__ rfm1, const bhah_kparams p",
        qspec, _groups(qspec, RICCI_GROUPS), qlam, fd_order, pr, table, "")
    # ---- the RK combinations with det(gammabar) enforcement, reversed, at every point.  A stage
    # input is Y = D(u), u = y_n + sum_j c_j K_j (D acts on hDD only), so the adjoint of Y goes
    # through the transposed Jacobian of D at u and is then scattered to y_n and the K_j.
    # adj_final: the last combination (y_{n+1}, weights b_j): lamYN = J_D^T lam, lamK_j = b_j lamYN.
    # adj_stage: stage input Y_s (weight a on K_{s-1}): lamYN += J_D^T lamY, lamK_{s-1} += a (...).
    # The operations are those of the separate combination, Jacobian and update kernels, in the
    # same order, so the results are theirs bit for bit.
    dspec = specs.detgamma_spec(v)
    dlam = {g: f"l_{g}" for g, _ in dspec.outputs}
    dfwd, drev, dinputs = _reverse_sweep(dspec.outputs, dlam, pr, "w")
    dcentre, _, _, drfm, _ = dspec.classify()
    det_out = {bssn.EVOL_GFS.index(g): g for g, _ in dspec.outputs}
    ngf = len(bssn.EVOL_GFS)

    def det_body(u_expr, lam_src: str) -> List[str]:
        lines = [f"  const double {n} = {u_expr(bssn.EVOL_GFS.index(n))};" for n in dcentre]
        lines += [f"  const double {n} = {RFM_EXPR[specs.RFM_SYMBOLS.index(n)]};" for n in drfm]
        lines += [f"  const double l_{g} = {lam_src}[{bssn.EVOL_GFS.index(g)} * Ntot + idx];" for g, _ in dspec.outputs]
        lines += dfwd + drev
        # the adjoint of every gridfunction after the transposed Jacobian
        lines += [f"  const double v{k} = " + (dinputs.get(det_out[k], "0.0") if k in det_out
                                              else f"{lam_src}[{k} * Ntot + idx]") + ";" for k in range(ngf)]
        return lines

    point = """  const int i0 = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  const int i1 = (int)(blockIdx.y * blockDim.y + threadIdx.y);
  const int i2 = (int)blockIdx.z;
  if (i0 >= p.Nx0 || i1 >= p.Nx1) return;
  const long Ntot = p.Ntot, nv = """ + str(ngf) + """ * Ntot;
  const long idx = i0 + p.s1 * i1 + p.s2 * i2;"""
    final_body = det_body(lambda k: (f"y[{k} * Ntot + idx] + b0 * K[{k} * Ntot + idx] + b1 * K[nv + {k} * Ntot + idx]"
                                     f" + b2 * K[2 * nv + {k} * Ntot + idx] + b3 * K[3 * nv + {k} * Ntot + idx]"),
                          "lam")
    final_body += [f"  lamYN[{k} * Ntot + idx] = v{k};" for k in range(ngf)]
    final_body += [f"  lamK[{j} * nv + {k} * Ntot + idx] = b{j} * v{k};" for j in range(4) for k in range(ngf)]
    stage_body = det_body(lambda k: f"y[{k} * Ntot + idx] + a * K[{k} * Ntot + idx]", "lamY")
    stage_body += [f"  lamYN[{k} * Ntot + idx] += v{k};\n  lamK[{k} * Ntot + idx] += a * v{k};" for k in range(ngf)]
    adj_rk = f"""
extern "C" __global__ void __launch_bounds__(256)
adj_final_{tag}(const double *__restrict__ y, const double *__restrict__ K, double b0, double b1, double b2,
                double b3, const double *__restrict__ lam, double *__restrict__ lamYN, double *__restrict__ lamK,
                const double *__restrict__ rfm0, const double *__restrict__ rfm1, const bhah_kparams p) {{
{point}
{chr(10).join(final_body)}
}}

extern "C" __global__ void __launch_bounds__(256)
adj_stage_{tag}(const double *__restrict__ y, const double *__restrict__ K, double a,
                const double *__restrict__ lamY, double *__restrict__ lamYN, double *__restrict__ lamK,
                const double *__restrict__ rfm0, const double *__restrict__ rfm1, const bhah_kparams p) {{
{point}
{chr(10).join(stage_body)}
}}
"""
    # ---- transposed stencils, unrolled per gridfunction: lam_y[g](q) += sum over the slots k
    # of g of scale_k * sum over taps t of w_t S_k(q - off_t).  S is zero outside the interior,
    # so a tap only needs q - off_t inside the array; its index is clamped into slot k's array
    # and the value selected away when it lies outside, so every load is unconditional.
    by_gf: Dict[int, List[int]] = {}
    for k, s in enumerate(table.slots):
        if s["gf"] >= 0:
            by_gf.setdefault(s["gf"], []).append(k)
    R = max(abs(c) for s in table.slots if s["gf"] >= 0 for off, _, _ in s["taps"] for c in off)
    cases = []
    for g in range(len(bssn.EVOL_GFS)):
        body = []
        for k in by_gf.get(g, []):
            taps = table.slots[k]["taps"]
            pw = taps[0][2]
            if any(t[2] != pw for t in taps):
                raise ValueError(f"slot {k}: taps with different grid-spacing powers")
            scale = " * ".join([f"inv{d}" for d in range(3) for _ in range(pw[d])]) or "1.0"
            terms = []
            for off, w, _ in taps:
                ok = " & ".join(f"(m{d} >> {off[d] + R})" for d in range(3))
                lin = f"{off[0]} + {off[1]} * s1 + {off[2]} * s2"
                terms.append(f"{_dbl(w)} * adjg_tap(S + {k} * Ntot, q - ({lin}), Ntot, ({ok}) & 1u)")
            body.append(f"      acc += {scale} * (" + "\n          + ".join(terms) + ");")
        cases.append(f"    case {g}: {{\n" + "\n".join(body) + "\n      break;\n    }")
    gather = f"""
__device__ __forceinline__ double adjg_tap(const double *__restrict__ Sk, long j, long Ntot, unsigned ok) {{
  j = j < 0 ? 0 : (j >= Ntot ? Ntot - 1 : j);
  const double v = Sk[j];
  return ok ? v : 0.0;
}}

// One thread per (point, gridfunction); all points.  Bit o + {R} of m<d> says that the point
// o steps below along direction d lies inside the array.
extern "C" __global__ void __launch_bounds__(256)
adj_gather_{tag}(const double *__restrict__ S, double *__restrict__ lam_y, const bhah_kparams p) {{
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int g = (int)(t / p.Ntot);
  if (g >= {len(bssn.EVOL_GFS)}) return;
  const long Ntot = p.Ntot, s1 = p.s1, s2 = p.s2;
  const long q = t - (long)g * Ntot;
  const int i[3] = {{(int)(q % p.Nx0), (int)((q / p.Nx0) % p.Nx1), (int)(q / s2)}};
  const int Nx[3] = {{p.Nx0, p.Nx1, p.Nx2}};
  unsigned m[3];
  for (int d = 0; d < 3; d++) {{
    m[d] = 0u;
    for (int o = -{R}; o <= {R}; o++) m[d] |= ((unsigned)(i[d] - o) < (unsigned)Nx[d] ? 1u : 0u) << (o + {R});
  }}
  const unsigned m0 = m[0], m1 = m[1], m2 = m[2];
  const double inv0 = p.invdxx0, inv1 = p.invdxx1, inv2 = p.invdxx2;
  double acc = 0.0;
  switch (g) {{
{chr(10).join(cases)}
    default:
      break;
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
  adj_rhs_{tag}_dims(p, g, b);
  hipLaunchKernelGGL(adj_rhs_{tag}, g, b, 0, s, in, aux, lam, S, rfm0, rfm1, p);
  return hipGetLastError();
}}
extern "C" hipError_t launch_adj_ricci_{tag}(dim3 g, dim3 b, hipStream_t s, const double *in, double *S,
    const double *rfm0, const double *rfm1, bhah_kparams p) {{
  adj_ricci_{tag}_dims(p, g, b);
  hipLaunchKernelGGL(adj_ricci_{tag}, g, b, 0, s, in, S, rfm0, rfm1, p);
  return hipGetLastError();
}}
extern "C" hipError_t launch_adj_final_{tag}(dim3 g, dim3 b, hipStream_t s, const double *y, const double *K,
    double b0, double b1, double b2, double b3, const double *lam, double *lamYN, double *lamK, const double *rfm0,
    const double *rfm1, bhah_kparams p) {{
  hipLaunchKernelGGL(adj_final_{tag}, g, b, 0, s, y, K, b0, b1, b2, b3, lam, lamYN, lamK, rfm0, rfm1, p);
  return hipGetLastError();
}}
extern "C" hipError_t launch_adj_stage_{tag}(dim3 g, dim3 b, hipStream_t s, const double *y, const double *K, double a,
    const double *lamY, double *lamYN, double *lamK, const double *rfm0, const double *rfm1, bhah_kparams p) {{
  hipLaunchKernelGGL(adj_stage_{tag}, g, b, 0, s, y, K, a, lamY, lamYN, lamK, rfm0, rfm1, p);
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
    return head + consts + "\n" + adj_rhs + adj_ricci + adj_rk + gather + undefs + "\n" + launch
