"""Host-side pieces of the lapse-loss gradient: the analytic d y0 / d theta of the
Brill-Lindquist data against central differences, and the sampling operator
(cubic Lagrange in r reproduces cubics exactly, theta samples sit on the grid)."""

import numpy as np

from nrpy_amd import adjoint


def test_initial_data_parameter_derivatives():
    xx = [np.linspace(0.1, 5, 9), np.linspace(0.2, 3.0, 5), np.linspace(-3, 3, 4)]
    theta = np.array([0.45, 0.55, 0.6, -0.4])
    y0, dy = adjoint.initial_data(xx, theta)
    for k in range(4):
        h = 1e-6
        tp, tm = theta.copy(), theta.copy()
        tp[k] += h
        tm[k] -= h
        fd = (adjoint.initial_data(xx, tp)[0] - adjoint.initial_data(xx, tm)[0]) / (2 * h)
        assert np.abs(fd - dy[k]).max() < 1e-8 * np.abs(dy[k]).max(), k
    assert np.all(dy[:, adjoint.ALPHA] == dy[:, adjoint.CF])


def test_lagrange_weights_reproduce_cubics():
    r = np.linspace(0.05, 9.95, 100)
    rs = 4.0
    i = int(np.searchsorted(r, rs)) - 1
    st = [i - 1, i, i + 1, i + 2]
    w = [np.prod([(rs - r[b]) / (r[a] - r[b]) for b in st if b != a]) for a in st]
    for p in range(4):
        assert abs(sum(wa * r[a] ** p for wa, a in zip(w, st)) - rs ** p) < 1e-12 * max(1, rs ** p)
