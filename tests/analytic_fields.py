"""Analytic BSSN-layout fields built from Cartesian scalar, vector and tensor fields.

In spherical-like coordinates the rescaled components of BHaH are components in the
orthonormal frame of the reference's unit vectors (nrpy/reference_metric.py:1362-1364):
r^ = (sin th cos ph, sin th sin ph, cos th), th^ = (cos th cos ph, cos th sin ph, -sin th),
ph^ = (-sin ph, cos ph, 0); vetU^i = e_i . V and hDD_ij = e_i . T . e_j.  A ghost point and
its inner-boundary source are the same Cartesian point, so for any Cartesian field the
ghost value must equal the source value times the product of the unit-vector dot products
of its indices: the inner boundary condition then reproduces the field exactly.  The
fields below are built from the gridfunction NAMES (aDD01 -> tensor (0, 1), ...), not from
a parity table, so such a check pins both the dot products and the name -> parity map.
"""

import numpy as np

from nrpy_amd import initial_data as ID

GFS = list(ID.EVOL_INDEX)


def unit_vectors(th, ph):
    """The reference's spherical-like unit vectors, each of shape (3,) + th.shape."""
    z = np.zeros_like(th + ph)
    return [
        np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th) + z]),
        np.stack([np.cos(th) * np.cos(ph), np.cos(th) * np.sin(ph), -np.sin(th) + z]),
        np.stack([-np.sin(ph) + 0 * th, np.cos(ph) + 0 * th, z]),
    ]


def bssn_field(coord, xx, size, sinhw):
    """24 gridfunctions (BHaH layout) of smooth Cartesian fields on every point, ghosts included."""
    P, T, R0 = np.meshgrid(xx[2], xx[1], xx[0], indexing="ij")
    r = ID.radius_of_xx0(coord, R0, size, sinhw)
    X, Y, Z = r * np.sin(T) * np.cos(P), r * np.sin(T) * np.sin(P), r * np.cos(T)
    e = unit_vectors(T, P)
    scal = 1.0 + 0.1 * X - 0.2 * Y + 0.3 * Z
    vec = np.stack([1.0 + 0.1 * Y, -0.5 + 0.2 * Z, 0.7 + 0.05 * X])
    ten = np.empty((3, 3) + X.shape)
    base = [[1.0, 0.2, -0.3], [0.2, 0.8, 0.1], [-0.3, 0.1, 1.3]]
    for a in range(3):
        for b in range(a, 3):
            ten[a, b] = ten[b, a] = base[a][b] + 0.02 * (a + 1) * X - 0.03 * (b + 1) * Z
    out = np.empty((24,) + X.shape)
    for k, name in enumerate(GFS):
        if name in ("alpha", "cf", "trK"):
            out[k] = scal
        elif name[:-1] in ("betU", "lambdaU", "vetU"):
            i = int(name[-1])
            out[k] = np.einsum("c...,c...->...", e[i], vec)
        else:  # aDD, hDD
            i, j = int(name[-2]), int(name[-1])
            out[k] = np.einsum("c...,cd...,d...->...", e[i], ten, e[j])
    return out


def adm_data(coord, xx, size, sinhw, axisymmetric=True, eps=0.1):
    """Non-conformally-flat ADM data in the Cartesian basis at every point, (19,) + grid shape:
    gammaDD (6), KDD (6), alpha, betaU (3), BU (3).  gamma_ij = psi^4 (delta_ij + eps A_ij) with
    psi the Brill-Lindquist factor of two punctures on the z axis and A_ij, K_ij, beta^i, B^i
    smooth and decaying; built from x_i and z^_i only when axisymmetric (so the fields do not
    depend on phi in spherical-like coordinates)."""
    P, T, R0 = np.meshgrid(xx[2], xx[1], xx[0], indexing="ij")
    r = ID.radius_of_xx0(coord, R0, size, sinhw)
    X, Y, Z = r * np.sin(T) * np.cos(P), r * np.sin(T) * np.sin(P), r * np.cos(T)
    psi = 1.0 + 0.5 / (2 * np.sqrt(X**2 + Y**2 + (Z - 0.5) ** 2)) + 0.4 / (2 * np.sqrt(X**2 + Y**2 + (Z + 0.5) ** 2))
    w = np.exp(-(X * X + Y * Y + Z * Z) / 8.0)
    x = [X, Y, Z]
    zh = [0.0 * X, 0.0 * X, 1.0 + 0.0 * X]
    A = np.empty((3, 3) + X.shape)
    Kt = np.empty((3, 3) + X.shape)
    for i in range(3):
        for j in range(3):
            d = 1.0 if i == j else 0.0
            A[i, j] = w * ((0.5 + 0.1 * Z) * d + 0.05 * x[i] * x[j] + 0.2 * zh[i] * zh[j]
                           + 0.03 * (x[i] * zh[j] + zh[i] * x[j]))
            Kt[i, j] = w * (0.04 * d + 0.02 * x[i] * x[j] - 0.03 * (x[i] * zh[j] + zh[i] * x[j]))
            if not axisymmetric:
                A[i, j] += w * 0.04 * X * (1.0 if (i, j) in ((0, 1), (1, 0)) else 0.5 * d) * (1 + 0.1 * Y)
                Kt[i, j] += w * 0.01 * Y * (1.0 if (i, j) in ((0, 2), (2, 0), (1, 1)) else 0.0)
    out = []
    for i, j in [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]:
        out.append(psi**4 * ((1.0 if i == j else 0.0) + eps * A[i, j]))
    for i, j in [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]:
        out.append(Kt[i, j])
    out.append(psi**-2)
    for i in range(3):
        out.append(w * (0.1 * x[i] + 0.05 * zh[i] * Z + (0.0 if axisymmetric else 0.02 * Y * (i == 0))))
    for i in range(3):
        out.append(w * 0.02 * x[i])
    return np.ascontiguousarray(np.stack(out))


def adm_from_bssn(O, y):
    """ADM data in the Cartesian basis rebuilt from rescaled BSSN variables (the inverse of the
    conversion, written independently of it): gammabar_rfm = h ReDD + ghat, basis change with
    the inverse Jacobian, gamma = gammabar / W^2, K_ij = Abar_ij / W^2 + gamma_ij K / 3."""
    E = ID.EVOL_INDEX
    J = O.jacobian()
    Jinv = np.moveaxis(np.linalg.inv(np.moveaxis(J, (0, 1), (-2, -1))), (-2, -1), (0, 1))
    P, T, R0 = np.meshgrid(O.xx[2], O.xx[1], O.xx[0], indexing="ij")
    f0 = np.broadcast_to(O.rfm0[: O.Nx[0]], R0.shape)
    f0p = np.broadcast_to(O.rfm0[O.Nx[0]: 2 * O.Nx[0]], R0.shape)
    s = [f0p, f0, f0 * np.sin(T)]
    gb, Ab = np.empty((3, 3) + R0.shape), np.empty((3, 3) + R0.shape)
    for i in range(3):
        for j in range(3):
            a, b = min(i, j), max(i, j)
            gb[i, j] = y[E[f"hDD{a}{b}"]] * s[i] * s[j] + (s[i] * s[i] if i == j else 0.0)
            Ab[i, j] = y[E[f"aDD{a}{b}"]] * s[i] * s[j]
    gC = np.einsum("ia...,ij...,jb...->ab...", Jinv, gb, Jinv)
    AC = np.einsum("ia...,ij...,jb...->ab...", Jinv, Ab, Jinv)
    W2 = y[E["cf"]] ** 2
    g = gC / W2
    K = AC / W2 + g * y[E["trK"]] / 3.0
    beta = np.einsum("ai...,i...->a...", J, np.stack([y[E[f"vetU{i}"]] / s[i] for i in range(3)]))
    return g, K, beta
