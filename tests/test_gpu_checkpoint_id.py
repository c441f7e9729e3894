"""Device checkpoint payloads against the oracle's (byte-identical), a full file
round trip, corruption handling, and device Brill-Lindquist data against the numpy
initial data and the oracle's Hamiltonian constraint."""

import struct

import numpy as np
import pytest

from nrpy_amd import initial_data as ID

pytestmark = pytest.mark.gpu


def test_checkpoint_payload_matches_oracle_and_round_trips(hip, oracle_lib, tmp_path):
    import torch
    from nrpy_amd.checkpoint import read_checkpoint, write_checkpoint
    from nrpy_amd.runtime import BHaHError, Patch, PatchConfig

    P = Patch(PatchConfig(Nxx=(24, 8, 2)), tuned=False)
    rng = np.random.default_rng(3)
    y = rng.standard_normal((24, *P.Nx[::-1]))
    yd = torch.from_numpy(y).cuda()
    payload = P.checkpoint_payload(yd)
    ref = np.zeros(len(payload), np.uint8)
    L = oracle_lib.lib()
    assert L.og_checkpoint_pack(P.Ntot, 24, np.ascontiguousarray(y).ctypes.data, ref.ctypes.data) == len(payload)
    assert payload == ref.tobytes()
    cd, pr = b"C" * 40, b"P" * 24
    path = str(tmp_path / "checkpoint-conv_factor1.00.dat")
    write_checkpoint(path, cd, [(pr, P, yd)])
    back = torch.zeros_like(yd)
    cd2, pr2 = read_checkpoint(path, len(cd), len(pr), [(P, back)])
    assert cd2 == cd and pr2 == [pr]
    assert torch.equal(back, yd)
    bad = bytearray(payload)
    bad[8:12] = struct.pack("<i", 0)  # indices 0, 0: not increasing
    with pytest.raises(BHaHError, match="non-increasing"):
        P.restore_payload(bytes(bad), back)


@pytest.mark.parametrize("coord,N,size", [(1, (24, 8, 2), 7.5), (2, (16, 8, 8), 64.0)])
def test_device_brill_lindquist(hip, oracle_lib, coord, N, size):
    import torch
    from nrpy_amd.runtime import Patch, PatchConfig

    fd = 4
    P = Patch(PatchConfig(coord=coord, Nxx=N, fd_order=fd, symmetry_axes=4 if coord == 1 else 0,
                          enable_KO=coord == 2, grid_physical_size=size), tuned=False)
    yd = P.new_gfs()
    P.initial_data_brill_lindquist(yd, 0.6, (0.3, 0.2, 0.5), 0.4, (-0.3, -0.2, -0.5))
    ref = ID.brill_lindquist(coord, P.xx(), size, 0.2, (0.6, 0.4), (0.3, 0.2, 0.5), (-0.3, -0.2, -0.5))
    got = yd.cpu().numpy()
    assert np.abs(got - ref).max() < 1e-14 * np.abs(ref).max()
