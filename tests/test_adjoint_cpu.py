"""Generated adjoint kernels executed on the CPU, checked by the dot-product test
against central finite differences of the oracle:

    <lam, (F(y + h v) - F(y - h v)) / 2h>  ==  <F'(y)^T lam, v>

for F = rhs_eval o (y -> (y, Ricci_eval(y))) at interior points (the Ricci chain
included) and for det(gammabar) enforcement at all points.  A dropped term, a
wrong partial derivative or a mis-transposed stencil tap changes the right-hand
side at the 1e-3..1 level; the finite differences agree to ~1e-9.
"""

import os
import shutil
import subprocess
import tempfile

import numpy as np
import pytest

from nrpy_amd import initial_data as ID

HERE = os.path.dirname(os.path.abspath(__file__))
GEN = os.path.join(os.path.dirname(HERE), "nrpy_amd", "csrc", "gen")
SHIM = os.path.join(HERE, "kernel_cpu", "shim")
DRIVER = os.path.join(HERE, "kernel_cpu", "adj_driver.cpp")
TAG = "Spherical_sym2_KO0_fd4"


@pytest.fixture(scope="module")
def adj_exe():
    wd = tempfile.mkdtemp(prefix="kadj_")
    exe = os.path.join(wd, "adj")
    src = os.path.join(GEN, f"adjoint_{TAG}.hip")
    # threaded emulation: the rhs adjoint's output groups meet at a workgroup barrier
    subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address", "-fno-omit-frame-pointer", "-w",
                    "-DBHAH_CPU_THREADED", "-pthread", f"-I{SHIM}", f'-DKERNEL_FILE="{src}"', f"-DTAG={TAG}", DRIVER,
                    "-o", exe, "-lm"],
                   check=True, capture_output=True)
    yield exe
    shutil.rmtree(wd, ignore_errors=True)


def _state(O, seed):
    rng = np.random.default_rng(seed)
    y = ID.brill_lindquist(O.coord, O.xx, O.g.p0, O.g.p1, (0.5, 0.5), (0, 0, 0.5), (0, 0, -0.5))
    y = ID.smooth_perturbation(y, O.xx, seed, amplitude=3e-2)
    y = y + 1e-3 * rng.standard_normal(y.shape)
    O.apply_bcs_outerextrap_and_inner(y)
    return np.ascontiguousarray(y)


def _run(exe, O, mode, y, aux, lam):
    wd = os.path.dirname(exe)
    inp, outp = os.path.join(wd, f"in{mode}.bin"), os.path.join(wd, f"out{mode}.bin")
    with open(inp, "wb") as f:
        np.array([mode, *O.Nx, O.NG], dtype=np.int32).tofile(f)
        np.array([1.0 / (O.xx[d][1] - O.xx[d][0]) for d in range(3)] + [2.0, 0.3, 0.3, 1.0]).tofile(f)
        np.ascontiguousarray(y).tofile(f)
        np.ascontiguousarray(aux).tofile(f)
        np.ascontiguousarray(lam).tofile(f)
        O.rfm0.tofile(f)
        O.rfm1.tofile(f)
    r = subprocess.run([exe, inp, outp], capture_output=True, text=True,
                       env=dict(os.environ, ASAN_OPTIONS="detect_leaks=0"))
    assert r.returncode == 0, r.stderr[-3000:]
    return np.fromfile(outp).reshape(24, *O.Nx[::-1])


def _interior_mask(O):
    m = np.zeros(O.Nx[::-1], bool)
    NG = O.NG
    m[NG:-NG, NG:-NG, NG:-NG] = True
    return m


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
@pytest.mark.parametrize("seed", [1, 2])
def test_rhs_ricci_adjoint_dot_product(oracle_lib, adj_exe, seed):
    O = oracle_lib.OracleGrid(1, (12, 8, 2), 4, 7.5, 0.2, variant=0)
    y = _state(O, seed)
    aux = O.ricci(y)
    rng = np.random.default_rng(100 + seed)
    mask = _interior_mask(O)
    lam = rng.standard_normal((24, *O.Nx[::-1])) * mask
    v = rng.standard_normal((24, *O.Nx[::-1]))
    h = 1e-6

    def F(z):
        return O.rhs(z, O.ricci(z), O.params()) * mask

    lhs = np.sum(lam * (F(y + h * v) - F(y - h * v))) / (2 * h)
    adj = _run(adj_exe, O, 0, y, aux, lam)
    rhs = np.sum(adj * v)
    assert abs(lhs - rhs) < 1e-6 * abs(lhs), (lhs, rhs)
    # a planted mistake in the transposition (one gridfunction's adjoint dropped) is caught
    assert abs(lhs - (rhs - np.sum(adj[11] * v[11]))) > 1e-3 * abs(lhs)


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
@pytest.mark.parametrize("mode", [1, 2])
def test_det_enforcement_adjoint_dot_product(oracle_lib, adj_exe, mode):
    """Through both fused RK/det adjoints (1: adj_stage, 2: adj_final) with zero combination
    weights, so their Jacobian part alone is checked."""
    O = oracle_lib.OracleGrid(1, (12, 8, 2), 4, 7.5, 0.2, variant=0)
    u = _state(O, 5)
    rng = np.random.default_rng(7)
    lam = rng.standard_normal(u.shape)
    v = rng.standard_normal(u.shape)
    h = 1e-6
    lhs = np.sum(lam * (O.enforce_detgamma(u + h * v) - O.enforce_detgamma(u - h * v))) / (2 * h)
    adj = _run(adj_exe, O, mode, u, np.zeros((6, *O.Nx[::-1])), lam)
    rhs = np.sum(adj * v)
    assert abs(lhs - rhs) < 1e-7 * abs(lhs), (lhs, rhs)
