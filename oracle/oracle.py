"""ORACLE / TEST INFRASTRUCTURE ONLY: ctypes wrapper of ``liboracle_bhah.so``.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's CPU-baseline leg use
this module, always as the checker.  Arrays are numpy float64 in the BHaH
layout (num_gfs, Nx2, Nx1, Nx0).
"""

from __future__ import annotations

import ctypes
import os
import subprocess
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle_bhah.so")
# the same restatement built with the reference's project flags, with unrolled-stencil
# rhs_eval / Ricci_eval for the timed workloads (the CPU baseline; never the checker)
NATIVE_LIB_PATH = os.path.join(HERE, "liboracle_bhah_native.so")

CARTESIAN, SPHERICAL, SINHSPHERICAL = 0, 1, 2
PARITY = [4, 5, 6, 7, 8, 9, 0, 1, 2, 3, 0, 4, 5, 6, 7, 8, 9, 1, 2, 3, 0, 1, 2, 3]
_libs: dict = {}


class OGrid(ctypes.Structure):
    _fields_ = [
        ("coord", ctypes.c_int),
        ("Nxx", ctypes.c_int * 3),
        ("NG", ctypes.c_int),
        ("Nx", ctypes.c_int * 3),
        ("xxmin", ctypes.c_double * 3),
        ("xxmax", ctypes.c_double * 3),
        ("dxx", ctypes.c_double * 3),
        ("invdxx", ctypes.c_double * 3),
        ("p0", ctypes.c_double),
        ("p1", ctypes.c_double),
        ("fd_order", ctypes.c_int),
    ]


def build(native: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc); ``native`` builds the CPU-baseline form."""
    subprocess.run(["make", "-s", "-C", HERE] + (["-j8", "native"] if native else []), check=True)
    return NATIVE_LIB_PATH if native else LIB_PATH


def native_supported() -> bool:
    """The native form is built for AVX-512 (x86-64-v4): true when this host has it."""
    try:
        with open("/proc/cpuinfo") as f:
            flags = next((line for line in f if line.startswith("flags")), "").split()
    except OSError:
        return False
    return all(x in flags for x in ("avx512f", "avx512dq", "avx512cd", "avx512bw", "avx512vl"))


def lib(native: bool = False):
    path = NATIVE_LIB_PATH if native else LIB_PATH
    if path not in _libs:
        if not os.path.exists(path):
            build(native)
        _lib = ctypes.CDLL(path)
        P, I, D, L = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_long
        sig = {
            "og_grid_init": (None, [P, I, I, I, I, I, I, D, D]),
            "og_xx": (None, [P, P, P, P]),
            "og_rfm": (None, [P, P, P, P, P]),
            "og_bssn_rhs": (None, [P, I, P, P, P, P, P, P]),
            "og_ricci": (None, [P, I, P, P, P, P]),
            "og_enforce_detgamma": (None, [P, I, P, P, P]),
            "og_constraints": (None, [P, I, P, P, P, P, P]),
            "og_initial_lambda": (None, [P, I, P, P, P]),
            "og_wave_rhs": (None, [P, D, P, P]),
            "og_bcstruct_new": (P, [P, P, P, P]),
            "og_bc_num_inner": (I, [P]),
            "og_bc_inner": (None, [P, P, P, P]),
            "og_bc_num_outer": (I, [P, I, I]),
            "og_bc_outer": (None, [P, I, I, P]),
            "og_bcstruct_free": (None, [P, I]),
            "og_apply_bcs_inner_only": (None, [P, P, I, P, P]),
            "og_apply_bcs_outerextrap_and_inner": (None, [P, P, I, P, P]),
            "og_apply_bcs_outerradiation_and_inner": (None, [P, P, P, I, P, P, P, I, P, P]),
            "og_wave_apply_bcs": (None, [P, P]),
            "og_rk_substep": (None, [I, L, D, P, P, P]),
            "og_bssn_mol_step": (None, [P, I, P, P, P, P, P, I, D, P, P, P, P, P]),
            "og_wave_mol_step": (None, [P, D, D, P, P, P, P]),
            "og_cfl_dt": (D, [P, P, P, P, D]),
            "og_checkpoint_pack": (ctypes.c_long, [ctypes.c_long, I, P, P]),
            "og_checkpoint_unpack": (I, [ctypes.c_long, I, P, P]),
            "og_unrolled_available": (I, [I, I]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(_lib, name)
            fn.restype = res
            fn.argtypes = args
        _libs[path] = _lib
    return _libs[path]


def _p(a: Optional[np.ndarray]):
    if a is None:
        return None
    assert a.dtype in (np.float64, np.int64, np.int8, np.int16, np.int32) and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data


class OracleGrid:
    """Grid + rfm arrays + boundary structs of the CPU restatement."""

    def __init__(self, coord: int, Nxx, fd_order: int, size: float, sinhw: float = 0.2,
                 variant: int = 0, NG: Optional[int] = None, with_bc: bool = True, native: bool = False):
        L = self.L = lib(native)
        self.g = OGrid()
        if NG is None:
            NG = fd_order // 2 + (0 if coord == CARTESIAN else 1)
        L.og_grid_init(ctypes.byref(self.g), coord, int(Nxx[0]), int(Nxx[1]), int(Nxx[2]), NG, fd_order, size, sinhw)
        self.coord, self.variant, self.NG = coord, variant, NG
        self.Nx = tuple(self.g.Nx)
        self.Ntot = self.Nx[0] * self.Nx[1] * self.Nx[2]
        self.xx = [np.zeros(n) for n in self.Nx]
        L.og_xx(ctypes.byref(self.g), *[_p(a) for a in self.xx])
        self.rfm0 = np.zeros(4 * self.Nx[0])
        self.rfm1 = np.zeros(3 * self.Nx[1])
        L.og_rfm(ctypes.byref(self.g), _p(self.xx[0]), _p(self.xx[1]), _p(self.rfm0), _p(self.rfm1))
        self.bc = None
        if with_bc and coord != CARTESIAN:
            self.bc = L.og_bcstruct_new(ctypes.byref(self.g), *[_p(a) for a in self.xx])

    def __del__(self):
        if getattr(self, "bc", None):
            self.L.og_bcstruct_free(self.bc, self.NG)
            self.bc = None

    def gfs(self, n: int = 24) -> np.ndarray:
        return np.zeros((n, self.Nx[2], self.Nx[1], self.Nx[0]))

    def params(self, eta=2.0, ko_g=0.3, ko_ng=0.3, wavespeed=1.0) -> np.ndarray:
        return np.array([eta, ko_g, ko_ng, wavespeed])

    # ---- pointwise-kernel loops
    def rhs(self, in_gfs, aux, params) -> np.ndarray:
        out = self.gfs(24)
        self.L.og_bssn_rhs(ctypes.byref(self.g), self.variant, _p(params), _p(self.rfm0), _p(self.rfm1),
                          _p(aux), _p(in_gfs), _p(out))
        return out

    def ricci(self, in_gfs) -> np.ndarray:
        aux = self.gfs(6)
        self.L.og_ricci(ctypes.byref(self.g), self.variant, _p(self.rfm0), _p(self.rfm1), _p(in_gfs), _p(aux))
        return aux

    def enforce_detgamma(self, gfs) -> np.ndarray:
        out = gfs.copy()
        self.L.og_enforce_detgamma(ctypes.byref(self.g), self.variant, _p(self.rfm0), _p(self.rfm1), _p(out))
        return out

    def constraints(self, in_gfs, aux) -> np.ndarray:
        """H, MU0, MU1, MU2, MSQUARED, LAMBDA_CONSTRAINT at interior points (6 gridfunctions)."""
        out = self.gfs(6)
        self.L.og_constraints(ctypes.byref(self.g), self.variant, _p(self.rfm0), _p(self.rfm1), _p(aux),
                             _p(in_gfs), _p(out))
        return out

    def volume_integrals(self, diag) -> dict:
        """The two diagnostics recipes over interior points, restated in numpy
        (diagnostics_volume_integration.py:177-207; diagnostics_volume_integration_helpers.h:396-476):
        "whole_domain" and "sphere_R_gt_8" (|x| > 8), integrands H^2, MSQUARED^2 and
        LAMBDA_CONSTRAINT^2 against dV = sqrt(det gammahat) |dxx0 dxx1 dxx2|, with L2 = sqrt(I)
        and RMS = sqrt(I / V) (0 for an empty volume, helpers.h:332-345)."""
        NG = self.NG
        sl = (slice(NG, -NG),) * 3
        f0 = self.rfm0[: self.Nx[0]][NG:-NG]
        f0p = self.rfm0[self.Nx[0]: 2 * self.Nx[0]][NG:-NG]
        s1 = self.rfm1[: self.Nx[1]][NG:-NG]
        dxx = abs(self.g.dxx[0] * self.g.dxx[1] * self.g.dxx[2])
        dV = np.abs(f0p[None, None, :] * f0[None, None, :] ** 2 * s1[None, :, None]) * dxx
        dV = np.broadcast_to(dV, (self.Nx[2] - 2 * NG, self.Nx[1] - 2 * NG, self.Nx[0] - 2 * NG))
        th = self.xx[1][NG:-NG][None, :, None]
        ph = self.xx[2][NG:-NG][:, None, None]
        r = f0[None, None, :]
        X, Y, Z = r * np.sin(th) * np.cos(ph), r * np.sin(th) * np.sin(ph), r * np.cos(th) + 0 * ph
        outside = ~(np.sqrt(X * X + Y * Y + Z * Z) <= 8.0)
        fields = {"H": diag[0][sl], "MSQUARED": diag[4][sl], "LAMBDA_CONSTRAINT": diag[5][sl]}
        out = {}
        for name, keep in (("whole_domain", np.ones_like(outside)), ("sphere_R_gt_8", outside)):
            w = np.where(keep & (dV != 0), dV, 0.0)
            vol = float(w.sum())
            d = {"volume": vol}
            for f, v in fields.items():
                I = float((v * v * w).sum())
                d[f] = {"integral": I, "L2": np.sqrt(max(I, 0.0)), "RMS": np.sqrt(max(I, 0.0) / vol) if vol > 0 else 0.0}
            out[name] = d
        return out

    def jacobian(self):
        """dx_Cart^a / dxx^i of the spherical-like map at every point: shape (3, 3, Nx2, Nx1, Nx0)."""
        P, T, R0 = np.meshgrid(self.xx[2], self.xx[1], self.xx[0], indexing="ij")
        f0 = np.broadcast_to(self.rfm0[: self.Nx[0]], R0.shape)
        f0p = np.broadcast_to(self.rfm0[self.Nx[0]: 2 * self.Nx[0]], R0.shape)
        st, ct, sp_, cp = np.sin(T), np.cos(T), np.sin(P), np.cos(P)
        J = np.empty((3, 3) + R0.shape)
        J[:, 0] = f0p * np.stack([st * cp, st * sp_, ct])
        J[:, 1] = f0 * np.stack([ct * cp, ct * sp_, -st])
        J[:, 2] = f0 * np.stack([-st * sp_, st * cp, 0 * st])
        return J

    def adm_cartesian_to_bssn(self, adm) -> np.ndarray:
        """initial_data_reader__convert_ADM_Cartesian_to_BSSN followed by the outer
        extrapolation initial_data() applies (ADM_Initial_Data_Reader__BSSN_Converter.py:268-460,
        1019-1129; initial_data.py:252), restated with numpy and the oracle's own boundary
        conditions and stencils.  adm: (19, Nx2, Nx1, Nx0), Cartesian basis at every point:
        gammaDD00,01,02,11,12,22, KDD00..22, alpha, betaU0-2, BU0-2.  Returns the 24 BSSN
        gridfunctions.

        1. ADM -> BSSN in the Cartesian basis (ADM_to_BSSN.py:80-160 with det(ghat) = 1):
           gammabar = det(gamma)^(-1/3) gamma, cf = W = (det gamma / det gammabar)^(-1/6),
           K = gamma^ij K_ij, Abar = det(gamma)^(-1/3) (K_ij - gamma_ij K / 3).
        2. Cartesian -> reference basis with the Jacobian (covariant indices) and its inverse
           (contravariant), then rescaling: h = (gammabar - ghat) / ReDD, a = Abar / ReDD,
           vet = beta / ReU, bet = B / ReU (:344-440).
        3. lambda^i = 0, inner boundary conditions, then lambda^i by finite differences of h at
           interior points (:462-580, 986-1010); then extrapolation + inner BCs on everything."""
        NG = self.NG
        shp = adm.shape[1:]
        sym = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]

        def mat(first):
            m = np.empty((3, 3) + shp)
            for k, (i, j) in enumerate(sym):
                m[i, j] = m[j, i] = adm[first + k]
            return m

        g, K = mat(0), mat(6)
        gm = np.moveaxis(g, (0, 1), (-2, -1))
        detg = np.linalg.det(gm)
        gbar = g * np.cbrt(1.0 / detg)
        detgbar = np.linalg.det(np.moveaxis(gbar, (0, 1), (-2, -1)))
        cf = (detg / detgbar) ** (-1.0 / 6.0)
        gUU = np.moveaxis(np.linalg.inv(gm), (-2, -1), (0, 1))
        trK = np.einsum("ij...,ij...->...", gUU, K)
        Abar = np.cbrt(1.0 / detg) * (K - g * trK / 3.0)
        J = self.jacobian()
        Jinv = np.moveaxis(np.linalg.inv(np.moveaxis(J, (0, 1), (-2, -1))), (-2, -1), (0, 1))  # [i, a]
        gbar_r = np.einsum("ai...,ab...,bj...->ij...", J, gbar, J)
        Abar_r = np.einsum("ai...,ab...,bj...->ij...", J, Abar, J)
        beta_r = np.einsum("ia...,a...->i...", Jinv, adm[13:16])
        B_r = np.einsum("ia...,a...->i...", Jinv, adm[16:19])
        # scale factors s_i = 1 / ReU^i; ghat_ij = s_i^2 delta_ij; ReDD_ij = s_i s_j
        P, T, R0 = np.meshgrid(self.xx[2], self.xx[1], self.xx[0], indexing="ij")
        f0 = np.broadcast_to(self.rfm0[: self.Nx[0]], R0.shape)
        f0p = np.broadcast_to(self.rfm0[self.Nx[0]: 2 * self.Nx[0]], R0.shape)
        s = [f0p, f0, f0 * np.sin(T)]
        y = np.zeros((24,) + shp)
        E = {n: k for k, n in enumerate(["aDD00", "aDD01", "aDD02", "aDD11", "aDD12", "aDD22", "alpha", "betU0",
                                         "betU1", "betU2", "cf", "hDD00", "hDD01", "hDD02", "hDD11", "hDD12",
                                         "hDD22", "lambdaU0", "lambdaU1", "lambdaU2", "trK", "vetU0", "vetU1",
                                         "vetU2"])}
        for i, j in sym:
            ghat = s[i] * s[i] if i == j else 0.0
            y[E[f"hDD{i}{j}"]] = (gbar_r[i, j] - ghat) / (s[i] * s[j])
            y[E[f"aDD{i}{j}"]] = Abar_r[i, j] / (s[i] * s[j])
        for i in range(3):
            y[E[f"vetU{i}"]] = beta_r[i] * s[i]
            y[E[f"betU{i}"]] = B_r[i] * s[i]
        y[E["alpha"]], y[E["cf"]], y[E["trK"]] = adm[12], cf, trK
        y = np.ascontiguousarray(y)
        self.apply_bcs_inner_only(y)
        self.L.og_initial_lambda(ctypes.byref(self.g), self.variant, _p(self.rfm0), _p(self.rfm1), _p(y))
        self.apply_bcs_outerextrap_and_inner(y)
        return y

    def wave_rhs(self, in_gfs, wavespeed=1.0) -> np.ndarray:
        out = self.gfs(2)
        self.L.og_wave_rhs(ctypes.byref(self.g), wavespeed, _p(in_gfs), _p(out))
        return out

    # ---- boundary conditions
    def inner_bc(self):
        L = self.L
        n = L.og_bc_num_inner(self.bc)
        dst, src = np.zeros(n, np.int64), np.zeros(n, np.int64)
        par = np.zeros((n, 10), np.int8)
        if n:
            L.og_bc_inner(self.bc, _p(dst), _p(src), _p(par))
        return dst, src, par

    def outer_bc(self, gz, d):
        L = self.L
        n = L.og_bc_num_outer(self.bc, gz, d)
        out = np.zeros((n, 6), np.int16)
        if n:
            L.og_bc_outer(self.bc, gz, d, _p(out))
        return out

    def apply_bcs_inner_only(self, gfs, parity=None):
        parity = np.array(parity if parity is not None else PARITY[: gfs.shape[0]], dtype=np.int32)
        self.L.og_apply_bcs_inner_only(ctypes.byref(self.g), self.bc, gfs.shape[0], _p(parity), _p(gfs))

    def apply_bcs_outerextrap_and_inner(self, gfs, parity=None):
        parity = np.array(parity if parity is not None else PARITY[: gfs.shape[0]], dtype=np.int32)
        self.L.og_apply_bcs_outerextrap_and_inner(ctypes.byref(self.g), self.bc, gfs.shape[0], _p(parity), _p(gfs))

    def apply_bcs_outerradiation_and_inner(self, gfs, rhs, rad_order=4):
        parity = np.array(PARITY, dtype=np.int32)
        ws = np.ones(24)
        fi = np.zeros(24)
        ws[6] = np.sqrt(2.0)
        fi[6] = 1.0
        fi[10] = 1.0
        self.L.og_apply_bcs_outerradiation_and_inner(ctypes.byref(self.g), self.bc, _p(self.xx[0]), 24, _p(parity),
                                                    _p(ws), _p(fi), rad_order, _p(gfs), _p(rhs))

    def wave_apply_bcs(self, gfs):
        self.L.og_wave_apply_bcs(ctypes.byref(self.g), _p(gfs))

    def cfl_dt(self, CFL: float = 0.5) -> float:
        return self.L.og_cfl_dt(ctypes.byref(self.g), *[_p(a) for a in self.xx], CFL)

    # ---- time stepping
    def bssn_mol_step(self, state: dict, params, dt: float, rad_order: int = 4) -> None:
        self.L.og_bssn_mol_step(ctypes.byref(self.g), self.variant, _p(params), _p(self.rfm0), _p(self.rfm1),
                               self.bc, _p(self.xx[0]), rad_order, dt, _p(state["y_n"]), _p(state["y_run"]),
                               _p(state["k_odd"]), _p(state["k_even"]), _p(state["aux"]))

    def new_state(self, y_n: np.ndarray, nvars: int = 24) -> dict:
        return {
            "y_n": np.ascontiguousarray(y_n.copy()),
            "y_run": self.gfs(nvars),
            "k_odd": self.gfs(nvars),
            "k_even": self.gfs(nvars),
            "aux": self.gfs(6),
        }

    def wave_mol_step(self, state: dict, dt: float, wavespeed: float = 1.0) -> None:
        self.L.og_wave_mol_step(ctypes.byref(self.g), wavespeed, dt, _p(state["y_n"]), _p(state["y_run"]),
                               _p(state["k_odd"]), _p(state["k_even"]))
