"""Device-side set-up (setup_on_device) against the host set-up, which the parity tests pin
to the oracle: the inner and pure-outer boundary lists (indices, parity signs, face normals,
order) must be identical.  The reference-metric arrays and the CFL timestep come from the
device's exp/sin/cos instead of the host library's: they agree to 1e-14 of each array's
magnitude and 1e-13 relative (near the origin f0 = A (e^{x/w} - e^{-x/w}) / ... subtracts
nearly equal terms, which turns last-bit differences of exp into ~1e-14 relative ones)."""

import dataclasses

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CONFIGS = {
    "spherical_72x12x2": dict(coord=1, Nxx=(72, 12, 2), fd_order=4, symmetry_axes=4, enable_KO=False,
                              grid_physical_size=7.5),
    "sinhspherical_192x96x96": dict(coord=2, Nxx=(192, 96, 96), fd_order=6, symmetry_axes=0, enable_KO=True,
                                    grid_physical_size=64.0, SINHW=0.2),
    # odd Nxx0 and Nxx1 (Nxx2 must be even: phi + pi has to be a grid point)
    "sinhspherical_fd4_odd": dict(coord=2, Nxx=(21, 9, 8), fd_order=4, symmetry_axes=0, enable_KO=True,
                                  grid_physical_size=20.0, SINHW=0.4),
    "ensemble_96x16x4": dict(coord=1, Nxx=(96, 16, 4), fd_order=4, symmetry_axes=4, enable_KO=False,
                             grid_physical_size=10.0),
    "ensemble_128x24x8": dict(coord=1, Nxx=(128, 24, 8), fd_order=4, symmetry_axes=4, enable_KO=False,
                              grid_physical_size=10.0),
    "wave_64": dict(system=1, coord=0, Nxx=(64, 64, 64), fd_order=4, symmetry_axes=0, grid_physical_size=10.0),
}


@pytest.mark.parametrize("name", list(CONFIGS))
def test_device_setup_matches_host(hip, name):
    from nrpy_amd.runtime import Patch, PatchConfig

    cfg = PatchConfig(**CONFIGS[name])
    H = Patch(cfg, tuned=False)
    D = Patch(dataclasses.replace(cfg, setup_on_device=True), tuned=False)
    assert H.Nx == D.Nx
    for a, b in zip(H.xx(), D.xx()):
        assert np.array_equal(a, b)
    if cfg.system == 0:
        for a, b in zip(H.inner_bc(), D.inner_bc()):
            assert a.shape == b.shape and np.array_equal(a, b)
        r0h, r1h = H.rfm()
        r0d, r1d = D.rfm()
        for a, b in ((r0d, r0h), (r1d, r1h)):
            scale = np.abs(b).max(axis=1, keepdims=True)
            assert (np.abs(a - b) <= 1e-14 * np.where(scale > 0, scale, 1)).all()
    for gz in range(H.NG):
        for d in range(3):
            assert np.array_equal(H.outer_bc(gz, d), D.outer_bc(gz, d)), (gz, d)
    assert abs(D.dt - H.dt) <= 1e-13 * H.dt
    print(f"\n{name}: {H.info.num_inner_boundary_points} inner, {H.info.num_pure_outer_boundary_points} outer "
          f"points identical; dt rel diff {abs(D.dt - H.dt) / H.dt:.1e}")
