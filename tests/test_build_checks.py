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
