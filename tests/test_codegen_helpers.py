"""Generator helpers of the workgroup-tiled stencil kernels, on the CPU.

- The symmetric/antisymmetric split of the upwind stencils: with S and A from
  ``_sym_line_weights``, (A + S) / (2 DS) and (A - S) / (2 DS) must reproduce the dupD and ddnD
  sums and r S / DS the KO sum, for random line values (a wrong sign, a dropped weight or a
  transposed mirror would fail at O(1)).
- ``_batch_lds_reads``: every LDS value read is renamed to a batch register, no batch holds
  more values than asked (unless one statement alone needs more), and the statements keep
  their order within the axis group and within the mixed-plane group.
- ``phaseprobe``: the instrumented source carries a stamp at every phase boundary and an
  entry point; the analysis of synthetic stamps returns the shares they were built with.
"""

import os
import re

import numpy as np
import pytest

from nrpy_amd.codegen import fd
from nrpy_amd.codegen import hip_gen as H

GEN = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "nrpy_amd", "csrc", "gen")


@pytest.mark.parametrize("order", [4, 6, 8])
def test_line_sums_reproduce_upwind_and_ko(order):
    up = dict(fd.weights_1d("dupD", order))
    dn = dict(fd.weights_1d("ddnD", order))
    ko = dict(fd.weights_1d("dKOD", order))
    res = H._sym_line_weights(up, dn, ko)
    assert res is not None
    IS, IA, DS, r = res
    R = max(abs(k) for k in list(up) + list(dn) + list(ko))
    rng = np.random.default_rng(order)
    for _ in range(5):
        f = {k: float(v) for k, v in zip(range(-R, R + 1), rng.standard_normal(2 * R + 1))}
        S = sum(w * (f[j] + f[-j] if j else f[0]) for j, w in IS.items())
        A = sum(w * (f[j] - f[-j]) for j, w in IA.items())
        direct = {name: sum(float(w) * f[k] for k, w in wt.items()) for name, wt in (("up", up), ("dn", dn), ("ko", ko))}
        assert abs((A + S) / (2 * DS) - direct["up"]) < 1e-12 * (1 + abs(direct["up"]))
        assert abs((A - S) / (2 * DS) - direct["dn"]) < 1e-12 * (1 + abs(direct["dn"]))
        assert abs(float(r) * S / DS - direct["ko"]) < 1e-12 * (1 + abs(direct["ko"]))


def test_line_sums_refuse_other_stencils():
    up = dict(fd.weights_1d("dupD", 6))
    dn = dict(fd.weights_1d("ddnD", 6))
    ko = dict(fd.weights_1d("dKOD", 6))
    # a KO stencil that is not proportional to the symmetric part, and a backward stencil that
    # is not the mirror image of the forward one
    assert H._sym_line_weights(up, dn, {k: v * (1 + (k == 1)) for k, v in ko.items()}) is None
    assert H._sym_line_weights(up, {k: -v for k, v in dn.items()}, ko) is None


def _stmt(name, offs):
    return (f"  double {name} = " + " + ".join(f"sh[c0 + {o}]" if o > 0 else (f"sh[c0 - {-o}]" if o else "sh[c0]")
                                             for o in offs) + ";", [name])


def test_batch_lds_reads():
    axis = [_stmt(f"ax{i}", [i - 4 + j for j in range(9)]) for i in range(3)]
    plane = [_stmt(f"pl{i}", [100 * i + 1000 + j for j in range(36)]) for i in range(3)]
    stmts = [plane[0], axis[0], plane[1], axis[1], axis[2], plane[2]]
    out = H._batch_lds_reads(stmts, 40, "R")
    text = "\n".join(t for t, _ in out)
    body = [t for t, _ in out if not t.lstrip().startswith(("double R", "BHAH_PINS"))]
    # every statement's LDS reads were renamed, and only the batch loads read LDS
    assert not any(re.search(r"sh\[c", t) for t in body)
    loads = re.findall(r"double (R_\d+_\d+) = sh\[c0[^\]]*\];", text)
    assert len(loads) == len(set(loads))
    per_batch = {}
    for n in loads:
        b = n.split("_")[1]
        per_batch[b] = per_batch.get(b, 0) + 1
    assert max(per_batch.values()) <= 40
    # axis statements first, in order, then the planes in order
    names = [f[0] for _, f in out if f]
    assert names == ["ax0", "ax1", "ax2", "pl0", "pl1", "pl2"]
    # each batch: its loads, then the pins, then its statements
    kinds = "".join("L" if t.lstrip().startswith("double R") else ("P" if "BHAH_PINS" in t else "S") for t, _ in out)
    assert re.fullmatch(r"(LP+S+)+", kinds), kinds


def test_phaseprobe_instrument_and_analyse():
    from nrpy_amd import phaseprobe

    sched = phaseprobe._default_schedule()
    with open(os.path.join(GEN, f"bssn_{phaseprobe.TAG}_rhs_{sched}.hip")) as f:
        src = f.read()
    nph = src.count("BHAH_BARRIER_PHASE(") - 1  # one occurrence is the macro definition
    out = phaseprobe.instrument(src)
    assert out.count("TSMARK();") == 2 * nph + 3
    assert 'extern "C" int phaseprobe_launch(' in out and "launch_bssn_rhs_" not in out
    # synthetic stamps: 2 workgroups of 4 waves, prologue 10, every wait 5, every phase 20,
    # algebra 50 ticks
    n = 2 * nph + 3
    rows = []
    for w in range(8):
        t, st = 1000 * (w // 4), []
        st.append(t)
        t += 10
        for _ in range(nph):
            st.append(t)
            t += 5
            st.append(t)
            t += 20
        st.append(t)
        t += 50
        st.append(t)
        row = np.zeros(64, np.int64)
        row[:n] = st
        row[62] = 0
        row[63] = n
        rows.append(row)
    res = phaseprobe.analyse(np.array(rows))
    tile = 10 + 25 * nph + 50
    assert res["tile_ticks"] == tile
    assert abs(res["waits"] - 5 * nph / tile) < 1e-12 and abs(res["phases"] - 20 * nph / tile) < 1e-12
    assert abs(res["prologue"] - 10 / tile) < 1e-12 and abs(res["algebra"] - 50 / tile) < 1e-12
    assert res["phase_ticks"] == [20] * nph and res["barrier_arrival_spread_ticks"] == 0
