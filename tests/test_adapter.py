"""The shipped adapter (nrpy_amd/adapter/bhah_amd_adapter.c) binds a generated BHaH project:
it compiles against a header with the reference's struct and member names and the
reference's function signatures, links with libbhah_hip.so, and (on a GPU) a project main()
with the reference's main loop runs through it and matches the oracle."""

import math
import os
import struct
import subprocess

import numpy as np
import pytest

from tests.adapter import build as adapter_build


def test_adapter_compiles_against_reference_names_and_links(tmp_path):
    exe = adapter_build.build(str(tmp_path))
    assert os.path.exists(exe)
    nm = subprocess.run(["nm", "-D", "--undefined-only", exe], capture_output=True, text=True, check=True).stdout
    for sym in ("bhah_hip_MoL_step_forward_in_time", "bhah_hip_diagnostics", "bhah_hip_rhs_eval"):
        assert sym in nm


@pytest.mark.gpu
def test_project_main_through_the_adapter_matches_oracle(hip, oracle_lib, tmp_path):
    from nrpy_amd import initial_data as ID

    exe = adapter_build.EXE
    assert os.path.exists(exe), "tests/adapter/_build/adapter_main missing: run __graft_entry__.build()"
    Nxx, t_final = (24, 8, 2), 0.6
    payload = tmp_path / "y.bin"
    r = subprocess.run([exe, *map(str, Nxx), str(t_final), str(payload)], cwd=tmp_path, capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    time, nn, dt = r.stdout.split()
    time, nn, dt = float(time), int(nn), float(dt)
    assert nn == math.ceil(t_final / dt) and time == nn * dt  # t_0 + (nn - nn_0) dt with t_0 = nn_0 = 0
    O = oracle_lib.OracleGrid(1, Nxx, 4, 7.5, 0.2, variant=0)
    assert abs(dt - O.cfl_dt(0.5)) <= 1e-15 * dt
    y = ID.brill_lindquist(1, O.xx, 7.5, 0.2, (0.5, 0.5), (0, 0, 0.5), (0, 0, -0.5))
    O.apply_bcs_outerextrap_and_inner(y)
    st = O.new_state(np.ascontiguousarray(y))
    ref0 = O.volume_integrals(O.constraints(st["y_n"], O.ricci(st["y_n"])))
    for _ in range(nn):
        O.bssn_mol_step(st, O.params(), dt)
    got = np.zeros_like(st["y_n"])
    buf = np.frombuffer(payload.read_bytes(), np.uint8).copy()
    assert oracle_lib.lib().og_checkpoint_unpack(O.Ntot, 24, buf.ctypes.data, got.ctypes.data) == 0
    num = np.abs(got - st["y_n"]).reshape(24, -1).max(axis=1)
    den = np.abs(st["y_n"]).reshape(24, -1).max(axis=1)
    assert (num / np.where(den > 0, den, 1)).max() < 1e-10
    # the diagnostics rows: one per output time 0, 0.25, 0.5; columns time, volume, then
    # (integral, L2, RMS) for H, M^2 and the Lambda constraint
    rows = [list(map(float, line.split())) for line in
            (tmp_path / "out3d-integrals-conv_factor1.00-whole_domain.txt").read_text().splitlines()
            if not line.startswith("#")]
    assert len(rows) == 3 and all(len(row) == 11 for row in rows)
    assert rows[0][0] == 0.0
    assert abs(rows[0][1] - ref0["whole_domain"]["volume"]) < 1e-12 * ref0["whole_domain"]["volume"]
    assert abs(rows[0][3] - ref0["whole_domain"]["H"]["L2"]) < 1e-11 * ref0["whole_domain"]["H"]["L2"]
    assert (tmp_path / "out3d-integrals-conv_factor1.00-sphere_R_gt_8.txt").exists()
