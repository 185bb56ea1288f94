"""Synthetic inputs: generators match their closed forms."""
import math

import numpy as np

from prandtl_amd import cases


def test_mt19937_64_known_answer():
    g = cases.MT19937_64(5489)
    for _ in range(9999):
        g()
    assert g() == 9981545732273789042   # the C++ standard's 10000th output


def test_node_coords_and_weights():
    X = cases.node_coords(2, 3, (4, 2), (-1.0, 0.0), (1.0, 1.0))
    assert X.shape == (2, 8, 16)
    assert X[0].min() == -1.0 and X[0].max() == 1.0
    WJ = cases.quadrature_weights(2, 3, (4, 2), (-1.0, 0.0), (1.0, 1.0))
    assert abs(WJ.sum() - 2.0) < 1e-14


def test_taylor_green_closed_form():
    X = np.array([[[0.3]], [[1.1]], [[2.0]]])
    U = cases.taylor_green(X)
    x, y, z = 0.3, 1.1, 2.0
    p = 1 / (1.4 * 0.01) + (math.cos(2 * x) + math.cos(2 * y)) * (math.cos(2 * z) + 2) / 16
    u = math.sin(x) * math.cos(y) * math.cos(z)
    v = -math.cos(x) * math.sin(y) * math.cos(z)
    assert abs(U[0, 0, 0] - 1.0) < 1e-15
    assert abs(U[1, 0, 0] - u) < 1e-15 and abs(U[2, 0, 0] - v) < 1e-15 and U[3, 0, 0] == 0
    assert abs(U[4, 0, 0] - (p / 0.4 + 0.5 * (u * u + v * v))) < 1e-12


def test_isentropic_vortex_is_isentropic():
    X = cases.node_coords(2, 3, (8, 8), (-5, -5), (5, 5))
    U = cases.isentropic_vortex(X)
    rho = U[0]
    vel = U[1:3] / rho
    p = 0.4 * (U[3] - 0.5 * rho * np.sum(vel ** 2, axis=0))
    assert np.abs(p / rho ** 1.4 - 1.0).max() < 1e-13
    assert abs(U[0].min() - (1 - 0.4 * 25 / (8 * 1.4 * math.pi ** 2) * math.e) ** 2.5) < 0.05


def test_kelvin_helmholtz_noise_is_deterministic():
    X = cases.node_coords(2, 4, (4, 4), (-1, -1), (1, 1))
    a = cases.kelvin_helmholtz(X)
    b = cases.kelvin_helmholtz(X)
    assert np.array_equal(a, b)
    v = a[2] / a[0]
    base = 0.1 * np.sin(2 * np.pi * X[0])
    assert np.abs(v - base).max() <= 1e-3 and np.abs(v - base).max() > 9e-4
