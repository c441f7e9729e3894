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
