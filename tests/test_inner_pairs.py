"""The condition under which the MoL step takes the inner-parity copy inside the RK pass, on the
oracle's boundary structs (CPU): both points of every pair (2q, 2q+1) find the sources of their
right-hand sides in one 16-byte-aligned pair.

With four ghost zones (fd6) and an even Nx0 this holds on every spherical-like grid here, the
192x96x96 grid of the headline step among them: the theta and phi ghost zones keep xx0, and the
xx0 reflection maps the pairs (0,1) and (2,3) to (7,6) and (5,4).  With three ghost zones (fd4)
the pair (2,3) holds a ghost point and an interior point, and the step keeps the separate copy.
"""

import numpy as np
import pytest

from oracle import oracle

CASES = [
    # coord, Nxx, fd_order, size, SINHW, expected
    (2, (20, 12, 10), 6, 10.0, 0.4, True),
    (2, (192, 96, 96), 6, 7.5, 0.2, True),
    (1, (32, 16, 16), 6, 10.0, 0.2, True),
    (2, (21, 10, 12), 6, 10.0, 0.4, False),  # odd Nx0: no pairs at all
    (1, (24, 8, 4), 4, 10.0, 0.2, False),
    (2, (192, 96, 96), 4, 7.5, 0.2, False),
]


def pairs_fit(O) -> bool:
    dst, src, _ = O.inner_bc()
    n = O.Ntot
    if n % 2:
        return False
    s = np.arange(n, dtype=np.int64)
    s[dst] = src
    return bool(np.all((s[0::2] >> 1) == (s[1::2] >> 1)))


@pytest.fixture(scope="module")
def lib():
    oracle.build()
    return oracle


@pytest.mark.parametrize("coord,Nxx,fd,size,w,expected", CASES)
def test_inner_sources_share_an_aligned_pair(lib, coord, Nxx, fd, size, w, expected):
    O = lib.OracleGrid(coord, Nxx, fd, size, w)
    dst, src, _ = O.inner_bc()
    assert len(dst) > 0
    # sources are never inner ghost points themselves, so one gather per point suffices
    assert not np.isin(src, dst).any()
    assert pairs_fit(O) == expected


def test_pair_check_rejects_a_split_pair(lib):
    """The check is not vacuous: moving one source by one point splits its pair."""
    O = lib.OracleGrid(2, (20, 12, 10), 6, 10.0, 0.4)
    dst, src, _ = O.inner_bc()
    s = np.arange(O.Ntot, dtype=np.int64)
    s[dst] = src
    q = int(dst[0]) // 2
    s[2 * q] = s[2 * q + 1] + 2
    assert not bool(np.all((s[0::2] >> 1) == (s[1::2] >> 1)))
