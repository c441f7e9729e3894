"""ctypes binding of ``libbhah_hip.so`` with PyTorch-ROCm tensors as device memory.

``Patch`` mirrors the reference's generated-C call sequence for one grid
(``nrpy/examples/two_blackholes_collide.py:290-313``): ``Ricci_eval``,
``rhs_eval``, the curvilinear boundary conditions, det(gammabar) enforcement and
``MoL_step_forward_in_time``.  Gridfunction arrays are torch float64 CUDA
(HIP) tensors of shape (num_gfs, Nx2, Nx1, Nx0), i.e. the BHaH gf-major
layout; their ``data_ptr()`` goes through the C-ABI unchanged.

There is no CPU fallback: every call goes to the HIP library, and loading
fails with an error if it was not built.
"""

from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass
from typing import Dict, Optional

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libbhah_hip.so")

CARTESIAN, SPHERICAL, SINHSPHERICAL = 0, 1, 2
SYSTEM_BSSN, SYSTEM_WAVE = 0, 1
OUTER_BC_RADIATION, OUTER_BC_EXTRAPOLATION = 0, 1
KERNEL_RHS, KERNEL_RICCI = 0, 1

EVOL_GFS = [
    "aDD00", "aDD01", "aDD02", "aDD11", "aDD12", "aDD22", "alpha",
    "betU0", "betU1", "betU2", "cf",
    "hDD00", "hDD01", "hDD02", "hDD11", "hDD12", "hDD22",
    "lambdaU0", "lambdaU1", "lambdaU2", "trK", "vetU0", "vetU1", "vetU2",
]


class PatchDesc(ctypes.Structure):
    _fields_ = [
        ("system", ctypes.c_int),
        ("coord", ctypes.c_int),
        ("Nxx", ctypes.c_int * 3),
        ("fd_order", ctypes.c_int),
        ("symmetry_axes", ctypes.c_int),
        ("enable_KO", ctypes.c_int),
        ("outer_bc", ctypes.c_int),
        ("radiation_bc_fd_order", ctypes.c_int),
        ("grid_physical_size", ctypes.c_double),
        ("SINHW", ctypes.c_double),
        ("CFL_FACTOR", ctypes.c_double),
        ("eta", ctypes.c_double),
        ("KO_strength_gauge", ctypes.c_double),
        ("KO_strength_nongauge", ctypes.c_double),
        ("wavespeed", ctypes.c_double),
        ("setup_on_device", ctypes.c_int),
    ]


class MoLGridfunctions(ctypes.Structure):
    _fields_ = [
        ("y_n_gfs", ctypes.c_void_p),
        ("y_nplus1_running_total_gfs", ctypes.c_void_p),
        ("k_odd_gfs", ctypes.c_void_p),
        ("k_even_gfs", ctypes.c_void_p),
        ("auxevol_gfs", ctypes.c_void_p),
    ]


class Clock(ctypes.Structure):
    """bhah_clock: the time members of commondata_struct (time, t_0, dt, nn, nn_0)."""

    _fields_ = [
        ("time", ctypes.c_double),
        ("t_0", ctypes.c_double),
        ("dt", ctypes.c_double),
        ("nn", ctypes.c_int),
        ("nn_0", ctypes.c_int),
    ]


NUM_DIAG_GFS = 6  # H, MU0, MU1, MU2, MSQUARED, LAMBDA_CONSTRAINT
DIAG_RECIPES = ("whole_domain", "sphere_R_gt_8")
DIAG_INTEGRANDS = ("H", "MSQUARED", "LAMBDA_CONSTRAINT")


class VolumeIntegrals(ctypes.Structure):
    _fields_ = [
        ("volume", ctypes.c_double * 2),
        ("integral", (ctypes.c_double * 3) * 2),
        ("L2", (ctypes.c_double * 3) * 2),
        ("RMS", (ctypes.c_double * 3) * 2),
    ]

    def as_dict(self) -> dict:
        """{recipe: {"volume": V, f: {"integral": I, "L2": L, "RMS": R}}}"""
        out = {}
        for q, name in enumerate(DIAG_RECIPES):
            d = {"volume": self.volume[q]}
            for k, f in enumerate(DIAG_INTEGRANDS):
                d[f] = {"integral": self.integral[q][k], "L2": self.L2[q][k], "RMS": self.RMS[q][k]}
            out[name] = d
        return out


class PatchInfo(ctypes.Structure):
    _fields_ = [
        ("Nxx", ctypes.c_int * 3),
        ("nghosts", ctypes.c_int),
        ("Nxx_plus_2NGHOSTS", ctypes.c_int * 3),
        ("num_evol_gfs", ctypes.c_int),
        ("num_auxevol_gfs", ctypes.c_int),
        ("Ntot", ctypes.c_int64),
        ("xxmin", ctypes.c_double * 3),
        ("xxmax", ctypes.c_double * 3),
        ("dxx", ctypes.c_double * 3),
        ("invdxx", ctypes.c_double * 3),
        ("dt", ctypes.c_double),
        ("num_inner_boundary_points", ctypes.c_int),
        ("num_pure_outer_boundary_points", ctypes.c_int),
    ]


# Every entry point declared in include/bhah_hip.h, with its argument types.
_P = ctypes.c_void_p
_D = ctypes.c_double
_I = ctypes.c_int
SIGNATURES: Dict[str, tuple] = {
    "bhah_hip_last_error": (ctypes.c_char_p, []),
    "bhah_hip_abi_version": (_I, []),
    "bhah_hip_patch_create": (_I, [ctypes.POINTER(PatchDesc), ctypes.POINTER(_P)]),
    "bhah_hip_patch_destroy": (_I, [_P]),
    "bhah_hip_patch_get_info": (_I, [_P, ctypes.POINTER(PatchInfo)]),
    "bhah_hip_patch_set_stream": (_I, [_P, _P]),
    "bhah_hip_synchronize": (_I, [_P]),
    "bhah_hip_patch_get_xx": (_I, [_P, _P, _P, _P]),
    "bhah_hip_patch_get_inner_bc": (_I, [_P, _P, _P, _P]),
    "bhah_hip_patch_get_outer_bc_count": (_I, [_P, _I, _I]),
    "bhah_hip_patch_get_outer_bc": (_I, [_P, _I, _I, _P]),
    "bhah_hip_patch_get_rfm": (_I, [_P, _P, _P]),
    "bhah_hip_rhs_eval": (_I, [_P, _P, _P, _P]),
    "bhah_hip_Ricci_eval": (_I, [_P, _P, _P]),
    "bhah_hip_enforce_detgammabar_equals_detgammahat": (_I, [_P, _P]),
    "bhah_hip_apply_bcs_outerradiation_and_inner": (_I, [_P, _P, _P]),
    "bhah_hip_apply_bcs_outerextrap_and_inner": (_I, [_P, _P]),
    "bhah_hip_apply_bcs_inner_only": (_I, [_P, _P]),
    "bhah_hip_MoL_malloc_intermediate_stage_gfs": (_I, [_P, ctypes.POINTER(MoLGridfunctions)]),
    "bhah_hip_MoL_free": (_I, [_P, ctypes.POINTER(MoLGridfunctions)]),
    "bhah_hip_MoL_step_forward_in_time": (_I, [_P, ctypes.POINTER(MoLGridfunctions), ctypes.POINTER(Clock)]),
    "bhah_hip_MoL_step_n": (_I, [_P, ctypes.POINTER(MoLGridfunctions), ctypes.POINTER(Clock), _I, _I]),
    "bhah_hip_patch_num_schedules": (_I, [_P, _I]),
    "bhah_hip_patch_schedule_name": (ctypes.c_char_p, [_P, _I, _I]),
    "bhah_hip_patch_select_schedule": (_I, [_P, _I, _I]),
    "bhah_hip_patch_get_schedule": (_I, [_P, _I]),
    "bhah_hip_patch_set_tile": (_I, [_P, _I, _I]),
    "bhah_hip_wave_rhs_eval": (_I, [_P, _P, _P]),
    "bhah_hip_wave_apply_bcs": (_I, [_P, _P]),
    "bhah_hip_diagnostics_constraints": (_I, [_P, _P, _P, _P, ctypes.POINTER(VolumeIntegrals)]),
    "bhah_hip_diagnostics": (_I, [_P, ctypes.POINTER(Clock), _D, _P, _P, _P, ctypes.POINTER(VolumeIntegrals),
                                  ctypes.POINTER(_I)]),
    "bhah_hip_diagnostics_constraints_noalloc": (_I, [_P, _P, _P, _P, _P]),
    "bhah_hip_diagnostic_gfs_malloc": (_I, [_P, ctypes.POINTER(_P)]),
    "bhah_hip_diagnostic_gfs_free": (_I, [_P]),
    "bhah_hip_MoL_step_adjoint": (_I, [_P, _P, _D, _P]),
    "bhah_hip_MoL_step_tape_doubles": (ctypes.c_int64, [_P]),
    "bhah_hip_MoL_step_forward_taped": (_I, [_P, _P, _D, _P]),
    "bhah_hip_MoL_step_adjoint_taped": (_I, [_P, _P, _D, _P]),
    "bhah_hip_checkpoint_payload_bytes": (ctypes.c_int64, [_P]),
    "bhah_hip_checkpoint_pack": (_I, [_P, _P, _P]),
    "bhah_hip_checkpoint_unpack": (_I, [_P, _P, ctypes.c_int64, _P]),
    "bhah_hip_initial_data_brill_lindquist": (_I, [_P, ctypes.POINTER(_D), _P]),
    "bhah_hip_initial_data_convert_ADM_Cartesian_to_BSSN": (_I, [_P, _P, _P]),
    "bhah_hip_ensemble_create": (_I, [ctypes.POINTER(PatchDesc), _I, ctypes.POINTER(_P)]),
    "bhah_hip_ensemble_destroy": (_I, [_P]),
    "bhah_hip_ensemble_size": (_I, [_P]),
    "bhah_hip_ensemble_member": (_I, [_P, _I, ctypes.POINTER(_P)]),
    "bhah_hip_ensemble_bind": (_I, [_P, ctypes.POINTER(MoLGridfunctions), ctypes.POINTER(_D)]),
    "bhah_hip_ensemble_set_stream": (_I, [_P, _P]),
    "bhah_hip_ensemble_select_schedule": (_I, [_P, _I, _I]),
    "bhah_hip_ensemble_set_launch": (_I, [_P, _I, _I, _I]),
    "bhah_hip_ensemble_step_n": (_I, [_P, _I, _I]),
    "bhah_hip_ensemble_synchronize": (_I, [_P]),
}

_LIB: Optional[ctypes.CDLL] = None


class BHaHError(RuntimeError):
    pass


def load_library(path: str = LIB_PATH) -> ctypes.CDLL:
    """Load the HIP library (no fallback: raises if it is absent)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(path):
        raise BHaHError(f"{path} not found: build it with `python -m nrpy_amd.build`")
    lib = ctypes.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _LIB = lib
    return lib


def _check(rc: int) -> None:
    if rc != 0:
        msg = load_library().bhah_hip_last_error().decode()
        raise BHaHError(f"bhah_hip error {rc}: {msg}")


@dataclass
class PatchConfig:
    """Python view of ``bhah_patch_desc`` with the reference's defaults."""

    system: int = SYSTEM_BSSN
    coord: int = SPHERICAL
    Nxx: tuple = (72, 12, 2)
    fd_order: int = 4
    symmetry_axes: int = 4  # phi
    enable_KO: bool = False
    outer_bc: int = OUTER_BC_RADIATION
    radiation_bc_fd_order: int = 4
    grid_physical_size: float = 7.5
    SINHW: float = 0.2
    CFL_FACTOR: float = 0.5
    eta: float = 2.0
    KO_strength_gauge: float = 0.3
    KO_strength_nongauge: float = 0.3
    wavespeed: float = 1.0
    setup_on_device: bool = False  # reference-metric arrays, boundary structs and CFL on the device

    def to_c(self) -> PatchDesc:
        d = PatchDesc()
        d.system, d.coord = self.system, self.coord
        for i in range(3):
            d.Nxx[i] = int(self.Nxx[i])
        d.fd_order = self.fd_order
        d.symmetry_axes = self.symmetry_axes
        d.enable_KO = int(self.enable_KO)
        d.outer_bc = self.outer_bc
        d.radiation_bc_fd_order = self.radiation_bc_fd_order
        d.grid_physical_size = self.grid_physical_size
        d.SINHW = self.SINHW
        d.CFL_FACTOR = self.CFL_FACTOR
        d.eta = self.eta
        d.KO_strength_gauge = self.KO_strength_gauge
        d.KO_strength_nongauge = self.KO_strength_nongauge
        d.wavespeed = self.wavespeed
        d.setup_on_device = int(bool(self.setup_on_device))
        return d


def _ptr(t) -> int:
    if t is None:
        return None
    import torch

    if not (t.is_cuda and t.dtype == torch.float64 and t.is_contiguous()):
        raise BHaHError("gridfunction tensors must be contiguous float64 device tensors")
    return t.data_ptr()


class Patch:
    """One grid patch: geometry, boundary structs and kernels on the current device."""

    def __init__(self, cfg: PatchConfig, stream=None, tuned: bool = True, handle=None):
        import torch

        self.lib = load_library()
        self.cfg = cfg
        self._desc = cfg.to_c()
        self.owned = handle is None
        if handle is None:
            h = ctypes.c_void_p()
            _check(self.lib.bhah_hip_patch_create(ctypes.byref(self._desc), ctypes.byref(h)))
            handle = h
        self.handle = handle
        info = PatchInfo()
        _check(self.lib.bhah_hip_patch_get_info(self.handle, ctypes.byref(info)))
        self.info = info
        self.NG = info.nghosts
        self.Nx = tuple(info.Nxx_plus_2NGHOSTS)
        self.Ntot = int(info.Ntot)
        self.ngfs = info.num_evol_gfs
        self.naux = info.num_auxevol_gfs
        self.dt = info.dt
        self.dxx = tuple(info.dxx)
        s = stream if stream is not None else torch.cuda.current_stream()
        _check(self.lib.bhah_hip_patch_set_stream(self.handle, ctypes.c_void_p(s.cuda_stream)))
        self.tile = (64, 4)
        self.tuned = None
        if tuned:
            from .autotune import apply_tuned

            self.tuned = apply_tuned(self)

    def close(self) -> None:
        if getattr(self, "handle", None):
            if self.owned:
                self.lib.bhah_hip_patch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- allocation helpers
    def new_gfs(self, n: Optional[int] = None):
        import torch

        n = self.ngfs if n is None else n
        return torch.zeros((n, self.Nx[2], self.Nx[1], self.Nx[0]), dtype=torch.float64, device="cuda")

    def xx(self):
        import numpy as np

        x = [np.zeros(n) for n in self.Nx]
        _check(self.lib.bhah_hip_patch_get_xx(self.handle, *[a.ctypes.data for a in x]))
        return x

    def inner_bc(self):
        import numpy as np

        n = self.info.num_inner_boundary_points
        dst, src = np.zeros(n, np.int64), np.zeros(n, np.int64)
        par = np.zeros((n, 10), np.int8)
        if n:
            _check(self.lib.bhah_hip_patch_get_inner_bc(self.handle, dst.ctypes.data, src.ctypes.data, par.ctypes.data))
        return dst, src, par

    def outer_bc(self, gz: int, dirn: int):
        import numpy as np

        n = self.lib.bhah_hip_patch_get_outer_bc_count(self.handle, gz, dirn)
        if n < 0:
            _check(n)
        out = np.zeros((n, 6), np.int16)
        if n:
            _check(self.lib.bhah_hip_patch_get_outer_bc(self.handle, gz, dirn, out.ctypes.data))
        return out

    # ---- kernel schedules (autotuned)
    def schedules(self, kind: int):
        n = self.lib.bhah_hip_patch_num_schedules(self.handle, kind)
        if n < 0:
            _check(n)
        return [self.lib.bhah_hip_patch_schedule_name(self.handle, kind, i).decode() for i in range(n)]

    def select_schedule(self, kind: int, name_or_index):
        names = self.schedules(kind)
        idx = names.index(name_or_index) if isinstance(name_or_index, str) else int(name_or_index)
        _check(self.lib.bhah_hip_patch_select_schedule(self.handle, kind, idx))

    def schedule(self, kind: int) -> str:
        return self.schedules(kind)[self.lib.bhah_hip_patch_get_schedule(self.handle, kind)]

    def set_tile(self, tile_x: int, tile_y: int):
        _check(self.lib.bhah_hip_patch_set_tile(self.handle, tile_x, tile_y))
        self.tile = (tile_x, tile_y)

    # ---- reference API counterparts
    def rhs_eval(self, auxevol, in_gfs, rhs):
        _check(self.lib.bhah_hip_rhs_eval(self.handle, _ptr(auxevol), _ptr(in_gfs), _ptr(rhs)))

    def Ricci_eval(self, in_gfs, auxevol):
        _check(self.lib.bhah_hip_Ricci_eval(self.handle, _ptr(in_gfs), _ptr(auxevol)))

    def enforce_detgammabar_equals_detgammahat(self, gfs):
        _check(self.lib.bhah_hip_enforce_detgammabar_equals_detgammahat(self.handle, _ptr(gfs)))

    def apply_bcs_outerradiation_and_inner(self, gfs, rhs):
        _check(self.lib.bhah_hip_apply_bcs_outerradiation_and_inner(self.handle, _ptr(gfs), _ptr(rhs)))

    def apply_bcs_outerextrap_and_inner(self, gfs):
        _check(self.lib.bhah_hip_apply_bcs_outerextrap_and_inner(self.handle, _ptr(gfs)))

    def apply_bcs_inner_only(self, gfs):
        _check(self.lib.bhah_hip_apply_bcs_inner_only(self.handle, _ptr(gfs)))

    def wave_rhs_eval(self, in_gfs, rhs):
        _check(self.lib.bhah_hip_wave_rhs_eval(self.handle, _ptr(in_gfs), _ptr(rhs)))

    def wave_apply_bcs(self, gfs):
        _check(self.lib.bhah_hip_wave_apply_bcs(self.handle, _ptr(gfs)))

    def _check_diag(self, diag):
        if diag.numel() < NUM_DIAG_GFS * self.Ntot:
            raise BHaHError(f"diag must hold {NUM_DIAG_GFS} gridfunctions ({NUM_DIAG_GFS * self.Ntot} values), "
                            f"has {diag.numel()}")

    def rfm(self):
        """Host copies of the reference-metric precompute arrays: rfm0 (4, Nx0) = f0 = r(xx0)
        and its first three derivatives, rfm1 (3, Nx1) = sin(xx1) and its first two."""
        import numpy as np

        r0, r1 = np.zeros((4, self.Nx[0])), np.zeros((3, self.Nx[1]))
        _check(self.lib.bhah_hip_patch_get_rfm(self.handle, r0.ctypes.data, r1.ctypes.data))
        return r0, r1

    def diagnostics_constraints(self, auxevol, in_gfs, diag, integrals: bool = True):
        """Constraints H, MU^i, M^2, |Lambda constraint| into diag (6 gridfunctions); with
        ``integrals`` returns the two recipes' volume integrals as a dict."""
        self._check_diag(diag)
        out = VolumeIntegrals()
        _check(self.lib.bhah_hip_diagnostics_constraints(self.handle, _ptr(auxevol), _ptr(in_gfs), _ptr(diag),
                                                         ctypes.byref(out) if integrals else None))
        return out.as_dict() if integrals else None

    def diagnostics(self, clock: "Clock", every: float, y_n, auxevol, diag):
        """The reference's diagnostics() for this grid: None on non-output steps, else the
        volume integrals (Ricci and constraint gridfunctions are left in auxevol and diag)."""
        self._check_diag(diag)
        out = VolumeIntegrals()
        flag = ctypes.c_int(0)
        _check(self.lib.bhah_hip_diagnostics(self.handle, ctypes.byref(clock), every, _ptr(y_n), _ptr(auxevol),
                                             _ptr(diag), ctypes.byref(out), ctypes.byref(flag)))
        return out.as_dict() if flag.value else None

    def synchronize(self):
        _check(self.lib.bhah_hip_synchronize(self.handle))

    def checkpoint_payload(self, y_n) -> bytes:
        """The reference checkpoint's per-grid payload (after params_struct) for y_n."""
        n = self.lib.bhah_hip_checkpoint_payload_bytes(self.handle)
        if n < 0:
            _check(int(n))
        buf = ctypes.create_string_buffer(int(n))
        _check(self.lib.bhah_hip_checkpoint_pack(self.handle, _ptr(y_n), buf))
        return buf.raw

    def restore_payload(self, payload: bytes, y_n) -> int:
        """Scatter a per-grid checkpoint payload into y_n; returns the bytes consumed."""
        _check(self.lib.bhah_hip_checkpoint_unpack(self.handle, payload, len(payload), _ptr(y_n)))
        count = int.from_bytes(payload[:4], "little", signed=True)
        return 4 * (1 + count) + 8 * count * self.ngfs

    def initial_data_brill_lindquist(self, y_n, m1, pos1, m2, pos2):
        bh = (ctypes.c_double * 8)(m1, *pos1, m2, *pos2)
        _check(self.lib.bhah_hip_initial_data_brill_lindquist(self.handle, bh, _ptr(y_n)))

    def initial_data_from_adm(self, adm_cartesian, y_n):
        """initial_data_reader__convert_ADM_Cartesian_to_BSSN + outer extrapolation: adm_cartesian
        is a (19, Nx2, Nx1, Nx0) float64 device tensor of ADM data in the Cartesian basis
        (gammaDD00..22, KDD00..22, alpha, betaU0-2, BU0-2) at every point."""
        if adm_cartesian.numel() < 19 * self.Ntot:
            raise BHaHError(f"ADM data must hold 19 gridfunctions ({19 * self.Ntot} values)")
        _check(self.lib.bhah_hip_initial_data_convert_ADM_Cartesian_to_BSSN(self.handle, _ptr(adm_cartesian),
                                                                           _ptr(y_n)))

    def MoL_step_adjoint(self, y_n, dt: float, lam):
        """lam (adjoint of the state after one RK4 step from y_n) <- adjoint of y_n."""
        _check(self.lib.bhah_hip_MoL_step_adjoint(self.handle, _ptr(y_n), dt, _ptr(lam)))

    @property
    def tape_doubles(self) -> int:
        """Doubles of one step's tape (bhah_hip_MoL_step_tape_doubles)."""
        n = int(self.lib.bhah_hip_MoL_step_tape_doubles(self.handle))
        if n < 0:
            raise BHaHError("patch has no tape (not a BSSN patch)")
        return n

    def MoL_step_forward_taped(self, y_n, dt: float, tape):
        """One RK4 step of y_n in place, the step's stages recorded in tape (a float64 device
        tensor of at least tape_doubles values)."""
        if tape.numel() < self.tape_doubles or tape.dtype.itemsize != 8:
            raise BHaHError(f"tape must hold {self.tape_doubles} float64 values")
        _check(self.lib.bhah_hip_MoL_step_forward_taped(self.handle, _ptr(y_n), dt, _ptr(tape)))

    def MoL_step_adjoint_taped(self, tape, dt: float, lam):
        """lam (adjoint of the state after the taped step) <- adjoint of the state before it."""
        if tape.numel() < self.tape_doubles or tape.dtype.itemsize != 8:
            raise BHaHError(f"tape must hold {self.tape_doubles} float64 values")
        _check(self.lib.bhah_hip_MoL_step_adjoint_taped(self.handle, _ptr(tape), dt, _ptr(lam)))


class MoLState:
    """MoL gridfunctions of one patch (MoL_malloc_intermediate_stage_gfs), torch-owned."""

    def __init__(self, patch: Patch, y_n=None):
        self.patch = patch
        self.y_n = y_n if y_n is not None else patch.new_gfs()
        self.y_run = patch.new_gfs()
        self.k_odd = patch.new_gfs()
        self.k_even = patch.new_gfs()
        self.aux = patch.new_gfs(max(patch.naux, 1))
        self.c = MoLGridfunctions(
            _ptr(self.y_n), _ptr(self.y_run), _ptr(self.k_odd), _ptr(self.k_even), _ptr(self.aux)
        )
        self.clock = Clock(0.0, 0.0, patch.dt, 0, 0)

    @property
    def time(self) -> float:
        return self.clock.time

    @property
    def nn(self) -> int:
        return self.clock.nn

    def reset_clock(self, time: float = 0.0, nn: int = 0) -> None:
        """Set time and nn and restart the time formula there (t_0 = time, nn_0 = nn), as the
        reference does after reading a checkpoint."""
        self.clock.time, self.clock.t_0, self.clock.nn, self.clock.nn_0 = time, time, nn, nn

    def _set_dt(self, dt: Optional[float]) -> None:
        dt = self.patch.dt if dt is None else dt
        if dt != self.clock.dt:
            # a new dt restarts the time formula at the current time (as the reference's regrid)
            self.clock.t_0, self.clock.nn_0, self.clock.dt = self.clock.time, self.clock.nn, dt

    def step(self, dt: Optional[float] = None):
        """MoL_step_forward_in_time: one RK4 step, then time = t_0 + (nn - nn_0 + 1) dt, nn++."""
        self._set_dt(dt)
        _check(self.patch.lib.bhah_hip_MoL_step_forward_in_time(self.patch.handle, ctypes.byref(self.c),
                                                                ctypes.byref(self.clock)))

    def steps(self, n: int, dt: Optional[float] = None, use_graph: bool = True):
        self._set_dt(dt)
        _check(self.patch.lib.bhah_hip_MoL_step_n(self.patch.handle, ctypes.byref(self.c), ctypes.byref(self.clock), n,
                                                  int(use_graph)))


class Ensemble:
    """Many independent BSSN evolutions of one kernel build advanced together.

    Each member owns torch-allocated MoL arrays (``states[m]``, a ``MoLState``
    over the member's patch); ``steps(n)`` advances every member n RK4 steps
    with its own timestep through the batched kernels (one persistent launch per
    stage for the whole ensemble, a hipGraph per step).
    """

    def __init__(self, cfgs, stream=None, dts=None):
        import torch

        self.lib = load_library()
        self.cfgs = list(cfgs)
        descs = (PatchDesc * len(self.cfgs))(*[c.to_c() for c in self.cfgs])
        h = ctypes.c_void_p()
        _check(self.lib.bhah_hip_ensemble_create(descs, len(self.cfgs), ctypes.byref(h)))
        self.handle = h
        s = stream if stream is not None else torch.cuda.current_stream()
        _check(self.lib.bhah_hip_ensemble_set_stream(self.handle, ctypes.c_void_p(s.cuda_stream)))
        self.members = []
        for m, c in enumerate(self.cfgs):
            ph = ctypes.c_void_p()
            _check(self.lib.bhah_hip_ensemble_member(self.handle, m, ctypes.byref(ph)))
            self.members.append(Patch(c, stream=s, tuned=False, handle=ph))
        self.states = [MoLState(P) for P in self.members]
        self.dts = [P.dt for P in self.members] if dts is None else list(dts)
        self.choice = {KERNEL_RHS: 0, KERNEL_RICCI: 0}
        self.tile, self.max_wg = (64, 4), 0
        self.bind()

    def bind(self):
        arr = (MoLGridfunctions * len(self.states))(*[st.c for st in self.states])
        dts = (ctypes.c_double * len(self.dts))(*self.dts)
        _check(self.lib.bhah_hip_ensemble_bind(self.handle, arr, dts))

    def schedules(self, kind: int):
        return self.members[0].schedules(kind)

    def select_schedule(self, kind: int, name_or_index):
        names = self.schedules(kind)
        idx = names.index(name_or_index) if isinstance(name_or_index, str) else int(name_or_index)
        _check(self.lib.bhah_hip_ensemble_select_schedule(self.handle, kind, idx))
        self.choice[kind] = idx

    def set_launch(self, tile_x: int, tile_y: int, max_workgroups: int = 0):
        _check(self.lib.bhah_hip_ensemble_set_launch(self.handle, tile_x, tile_y, max_workgroups))
        self.tile, self.max_wg = (tile_x, tile_y), max_workgroups

    def steps(self, n: int, use_graph: bool = True):
        _check(self.lib.bhah_hip_ensemble_step_n(self.handle, n, int(use_graph)))
        for st, dt in zip(self.states, self.dts):
            st._set_dt(dt)
            for _ in range(n):
                st.clock.time = st.clock.t_0 + (st.clock.nn - st.clock.nn_0 + 1) * dt
                st.clock.nn += 1

    def synchronize(self):
        _check(self.lib.bhah_hip_ensemble_synchronize(self.handle))

    def close(self):
        if getattr(self, "handle", None):
            for P in self.members:
                P.handle = None
            self.lib.bhah_hip_ensemble_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
