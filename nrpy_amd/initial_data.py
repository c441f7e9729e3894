"""Host-side initial data for the evolution path (set-up, not the hot path).

* Brill-Lindquist punctures (``nrpy/equations/general_relativity/InitialData_Cartesian.py:111-141``):
  conformally flat gamma_ij = psi^4 delta_ij, K_ij = 0,
  psi = 1 + m1/(2 r1) + m2/(2 r2), pre-collapsed lapse alpha = psi^-2
  (``InitialData_Cartesian.py:86-108``), zero shift and B^i.
  Converted to rescaled BSSN variables in any spherical-like coordinates
  (``ADM_to_BSSN.py:60-170``): gammabar_ij = (det ghat/det gamma)^(1/3) gamma_ij =
  ghat_ij, hence h_ij = 0, a_ij = 0, trK = 0, W = cf = psi^-2.  lambda^i, which the
  reference obtains by finite differences of gammabar (``ADM_Initial_Data_Reader__BSSN_Converter.py:462-580``),
  is identically zero for conformally flat data and is set to zero here.
* Spherical Gaussian pulse for the scalar wave equation
  (``nrpy/equations/wave_equation/WaveEquation_Solutions_InitialData.py:83-140``).

Arrays have the BHaH layout (num_gfs, Nx2, Nx1, Nx0).
"""

from __future__ import annotations

from typing import Sequence, Tuple

import numpy as np

EVOL_INDEX = {
    g: i
    for i, g in enumerate(
        [
            "aDD00", "aDD01", "aDD02", "aDD11", "aDD12", "aDD22", "alpha",
            "betU0", "betU1", "betU2", "cf",
            "hDD00", "hDD01", "hDD02", "hDD11", "hDD12", "hDD22",
            "lambdaU0", "lambdaU1", "lambdaU2", "trK", "vetU0", "vetU1", "vetU2",
        ]
    )
}


def radius_of_xx0(coord: int, xx0: np.ndarray, size: float, sinhw: float) -> np.ndarray:
    """r(xx0) of the spherical-like coordinates (reference_metric.py:1201-1270)."""
    if coord == 1:  # Spherical
        return xx0
    if coord == 2:  # SinhSpherical
        return size * (np.exp(xx0 / sinhw) - np.exp(-xx0 / sinhw)) / (np.exp(1 / sinhw) - np.exp(-1 / sinhw))
    raise ValueError("spherical-like coordinates only")


def cartesian_coords(coord: int, xx: Sequence[np.ndarray], size: float, sinhw: float):
    """Cartesian (x, y, z) on the full grid, each of shape (Nx2, Nx1, Nx0)."""
    x0, x1, x2 = xx
    if coord == 0:
        Z, Y, X = np.meshgrid(x2, x1, x0, indexing="ij")
        return X, Y, Z
    r = radius_of_xx0(coord, x0, size, sinhw)
    P, T, Rr = np.meshgrid(x2, x1, r, indexing="ij")
    return Rr * np.sin(T) * np.cos(P), Rr * np.sin(T) * np.sin(P), Rr * np.cos(T)


def brill_lindquist(
    coord: int,
    xx: Sequence[np.ndarray],
    size: float,
    sinhw: float,
    masses: Tuple[float, float] = (0.5, 0.5),
    pos1: Tuple[float, float, float] = (0.0, 0.0, 0.5),
    pos2: Tuple[float, float, float] = (0.0, 0.0, -0.5),
) -> np.ndarray:
    X, Y, Z = cartesian_coords(coord, xx, size, sinhw)
    r1 = np.sqrt((X - pos1[0]) ** 2 + (Y - pos1[1]) ** 2 + (Z - pos1[2]) ** 2)
    r2 = np.sqrt((X - pos2[0]) ** 2 + (Y - pos2[1]) ** 2 + (Z - pos2[2]) ** 2)
    psi = 1.0 + masses[0] / (2.0 * r1) + masses[1] / (2.0 * r2)
    out = np.zeros((24,) + X.shape)
    out[EVOL_INDEX["cf"]] = psi ** -2
    out[EVOL_INDEX["alpha"]] = psi ** -2
    return out


def smooth_perturbation(
    base: np.ndarray, xx: Sequence[np.ndarray], seed: int, amplitude: float = 1e-2
) -> np.ndarray:
    """Seeded smooth perturbation of all 24 gridfunctions.

    Multiplicative (1 + amplitude * mode) on alpha and cf, additive amplitude * mode on the
    others, with one low-wavenumber product of sines per gridfunction, so that every term of
    the right-hand sides is exercised (pure Brill-Lindquist data zeros K, A, beta, B).
    """
    rng = np.random.default_rng(seed)
    x0, x1, x2 = xx
    out = base.copy()
    for g in range(base.shape[0]):
        k = rng.uniform(0.5, 2.0, 3)
        ph = rng.uniform(0, 2 * np.pi, 3)
        a = rng.uniform(-1, 1)
        P, T, R0 = np.meshgrid(x2, x1, x0, indexing="ij")
        mode = a * np.sin(k[0] * R0 + ph[0]) * np.sin(k[1] * T + ph[1]) * np.cos(k[2] * P + ph[2])
        if g in (EVOL_INDEX["alpha"], EVOL_INDEX["cf"]):
            out[g] = base[g] * (1.0 + amplitude * mode)
        else:
            out[g] = base[g] + amplitude * mode
    return out


def wave_spherical_gaussian(xx: Sequence[np.ndarray], t: float = 0.0, sigma: float = 1.0, c: float = 1.0) -> np.ndarray:
    """uu, vv of the exact outgoing+ingoing Gaussian, offset by 2 (Cartesian grid)."""
    X, Y, Z = cartesian_coords(0, xx, 0.0, 0.0)
    r = np.sqrt(X * X + Y * Y + Z * Z)
    s2 = 2.0 * sigma * sigma
    em = np.exp(-((r - c * t) ** 2) / s2)
    ep = np.exp(-((r + c * t) ** 2) / s2)
    uu = (r - c * t) / r * em + (r + c * t) / r * ep + 2.0
    # vv = d uu / dt
    dem = em * (-c / r + (r - c * t) / r * (2 * c * (r - c * t) / s2))
    dep = ep * (c / r + (r + c * t) / r * (-2 * c * (r + c * t) / s2))
    vv = dem + dep
    return np.stack([uu, vv])
