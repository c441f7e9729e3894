"""The assembly check that guards schedules with early-issued loads."""

import pytest

from nrpy_amd import isa_check


def _asm(n_loads, waits=()):
    lines = ["kern:", "  s_load_dwordx2 s[0:1], s[4:5], 0x0"]
    for i in range(n_loads):
        lines.append(f"  global_load_dwordx2 v[{2 * i}:{2 * i + 1}], v[200:201], off offset:{8 * i}")
        if i in waits:
            lines.append(f"  s_waitcnt vmcnt({waits[i]})")
    lines += ["  s_waitcnt vmcnt(0)", "  s_endpgm", ".Lfunc_end0:"]
    return "\n".join(lines)


def test_counts_outstanding_and_waits():
    assert isa_check.max_outstanding_vmem(_asm(10)) == {"kern": 10}
    # a wait down to 2 after the 6th load: 2 + 4 more = 6 at most
    assert isa_check.max_outstanding_vmem(_asm(10, {5: 2})) == {"kern": 6}


def test_rejects_more_than_vmcnt_can_track():
    isa_check.check(_asm(63))
    with pytest.raises(RuntimeError, match="vmcnt"):
        isa_check.check(_asm(64))
    isa_check.check(_asm(100, {40: 10, 80: 10}))


def test_lds_schedule_names_and_limits():
    """Schedule-name parsing of the workgroup-tiled kernels, and the generator's refusal of
    an LDS ring larger than a CU's 160 KiB or a DMA prefetch deeper than vmcnt tracks."""
    from nrpy_amd.codegen import hip_gen, specs

    assert hip_gen._lds_geometry("L16x4x4x1") == (16, 4, 4, 1)
    assert hip_gen.is_lds("p1_lb1_sj_L16x4x4x1_d3_dma_rb1_xcd")
    assert not hip_gen.is_lds("p2_lb1_sj_lf_pf16_u32_ck8_w2")
    with pytest.raises(ValueError):
        hip_gen._lds_geometry("L16x4x3x2")  # points per thread must divide TZ
    v = specs.VARIANTS[1]
    # three 3-gridfunction phases of 20 KiB boxes in a ring of three: 184 KiB
    with pytest.raises(ValueError, match="LDS"):
        hip_gen.candidate_tu(v, 6, "ricci", "r1_lb1_sj_L16x4x4x1_g3_d3_dma_rb1_xcd")
    src, flops = hip_gen.candidate_tu(v, 6, "ricci", "r1_lb1_sj_L16x4x4x1_d2_dma_xcd")
    assert "BHAH_DMA16" in src and "BHAH_VMWAIT(" in src and flops > 3000


def test_count_follows_loops():
    """A loop that issues loads and waits for none keeps more in flight each trip: the
    scan carries the count around the back edge and reports it; a wait inside the loop
    bounds it."""
    body = [f"  global_load_dwordx2 v[{2 * i}:{2 * i + 1}], v[200:201], off" for i in range(8)]
    loop = ["kern:", "  s_mov_b32 s0, 0", ".LBB0_1:", *body, "  s_add_i32 s0, s0, 1",
            "  s_cmp_lt_i32 s0, 100", "  s_cbranch_scc1 .LBB0_1", "  s_waitcnt vmcnt(0)", "  s_endpgm",
            ".Lfunc_end0:"]
    n = isa_check.max_outstanding_vmem("\n".join(loop))["kern"]
    assert n > isa_check.VMCNT_MAX
    with pytest.raises(RuntimeError, match="vmcnt"):
        isa_check.check("\n".join(loop))
    waited = loop[:3] + ["  s_waitcnt vmcnt(4)"] + loop[3:]
    assert isa_check.max_outstanding_vmem("\n".join(waited))["kern"] == 12
    # both arms of a branch: the larger count survives the join
    br = ["kern:", "  s_cbranch_scc1 .LBB0_2", *body[:5], "  s_branch .LBB0_3", ".LBB0_2:", *body[:2],
          ".LBB0_3:", *body[:3], "  s_waitcnt vmcnt(0)", "  s_endpgm", ".Lfunc_end0:"]
    assert isa_check.max_outstanding_vmem("\n".join(br))["kern"] == 8


def _dma_listing(wait0, wait1, spill_between=0):
    """Two phases of 2 DMA loads each, both issued up front; phase 0 is read after
    vmcnt(wait0), phase 1 after vmcnt(wait1)."""
    dma = "  global_load_lds_dwordx4 v1, s[4:5] offset:0"
    lines = ["kern:", dma, dma, dma, dma]
    lines += ["  scratch_store_dword off, v7, s32"] * spill_between
    lines += [f"  s_waitcnt vmcnt({wait0})", "  s_waitcnt lgkmcnt(0)", "  s_barrier ; bhah_read_phase 0",
              "  ds_read_b64 v[2:3], v9", f"  s_waitcnt vmcnt({wait1})", "  s_waitcnt lgkmcnt(0)",
              "  s_barrier ; bhah_read_phase 1", "  ds_read_b64 v[2:3], v9", "  s_endpgm", ".Lfunc_end0:"]
    return "\n".join(lines)


def test_dma_wait_proof():
    issue = {"kern": [(0, 2), (1, 2)]}
    assert isa_check.check_dma_waits(_dma_listing(2, 0), issue) == {"kern": 2}
    with pytest.raises(RuntimeError, match="phase 0"):
        isa_check.check_dma_waits(_dma_listing(3, 0), issue)  # one phase-0 load may be pending
    with pytest.raises(RuntimeError, match="phase 1"):
        isa_check.check_dma_waits(_dma_listing(2, 1), issue)
    # younger spills make a wait stronger: vmcnt(3) then retires both phase-0 loads
    assert isa_check.check_dma_waits(_dma_listing(3, 0, spill_between=1), issue) == {"kern": 2}
    # the note and the listing must agree on the number of DMA loads
    with pytest.raises(RuntimeError, match="global-to-LDS"):
        isa_check.check_dma_waits(_dma_listing(2, 0), {"kern": [(0, 2), (1, 3)]})


def test_planted_wrong_wait_in_a_generated_kernel_is_refused(tmp_path):
    """The generated DMA-staged Ricci kernel compiles to a listing the check accepts; the
    same source with the wait before its second phase raised by two (so two of that
    phase's own box loads may still be in flight) is refused."""
    import re
    import subprocess

    from nrpy_amd import build as nb
    from nrpy_amd.codegen import hip_gen, specs

    src, _ = hip_gen.candidate_tu(specs.VARIANTS[0], 4, "ricci", "r1_lb1_sj_L16x8x2x1_d3_dma_rb1")
    hdr = "#include <hip/hip_runtime.h>\n#include <math.h>\n" + hip_gen.KPARAMS
    waits = [m for m in re.finditer(r"BHAH_VMWAIT\((\d+)\);", src)]
    assert len(waits) >= 3
    m = waits[1]
    bad = src[:m.start()] + f"BHAH_VMWAIT({int(m.group(1)) + 2});" + src[m.end():]
    for name, text, ok in (("good", src, True), ("bad", bad, False)):
        f = tmp_path / f"{name}.hip"
        f.write_text(hdr + text)
        asm = tmp_path / f"{name}.s"
        r = subprocess.run([nb.HIPCC, *nb.HIPFLAGS, "-I", nb.GEN, "--cuda-device-only", "-S", str(f), "-o", str(asm)],
                           capture_output=True, text=True)
        assert r.returncode == 0, r.stderr[-2000:]
        if ok:
            isa_check.check(asm.read_text(), name, source_text=hdr + text)
        else:
            with pytest.raises(RuntimeError, match="DMA wait"):
                isa_check.check(asm.read_text(), name, source_text=hdr + text)
