"""Pins for the oracle's interpolators (operators.py:134-209, steppers.py:362-367).

Each check is independent of the oracle's formula: a closed form, a brute-force
re-derivation with Catmull-Rom weights, or SciPy's map_coordinates.
"""
import itertools

import numpy as np
import pytest
import scipy.ndimage as ndi

from oracle import sisl as O


def _cr_weights(t):
    """Catmull-Rom cardinal weights (Keys a = -1/2), written from the kernel
    definition, not from the reference's polynomial-in-p form."""
    def k(s):
        s = abs(s)
        if s <= 1.0:
            return 1.5 * s ** 3 - 2.5 * s ** 2 + 1.0
        if s < 2.0:
            return -0.5 * s ** 3 + 2.5 * s ** 2 - 4.0 * s + 2.0
        return 0.0
    return [k(t + 1.0), k(t), k(1.0 - t), k(2.0 - t)]


def _brute_tricubic(field, x, y, z):
    n = field.shape
    c = [min(max(v, 0.0), n[a] - 1) for a, v in enumerate((x, y, z))]
    b = [int(np.floor(v)) for v in c]
    t = [c[a] - b[a] for a in range(3)]
    taps = [[max(b[a] - 1, 0), b[a], min(b[a] + 1, n[a] - 1), min(b[a] + 2, n[a] - 1)] for a in range(3)]
    w = [_cr_weights(t[a]) for a in range(3)]
    s = 0.0
    for i, j, k in itertools.product(range(4), repeat=3):
        s += w[0][i] * w[1][j] * w[2][k] * field[taps[0][i], taps[1][j], taps[2][k]]
    return s


def test_tricubic_matches_bruteforce_with_clamped_edges():
    rng = np.random.default_rng(0)
    f = rng.standard_normal((5, 4, 6))
    pts = rng.uniform(-2.0, 7.0, size=(3, 300))
    pts[:, :20] = np.round(pts[:, :20])  # integer coordinates and exact edges
    pts[0, 20:25] = 4.0
    out = O.tricubic(f, pts)
    ref = np.array([_brute_tricubic(f, *pts[:, i]) for i in range(pts.shape[1])])
    np.testing.assert_allclose(out, ref, rtol=1e-12, atol=1e-12)


def test_tricubic_reproduces_quadratics_in_the_interior():
    nx, ny, nz = 9, 8, 7
    X, Y, Z = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing='ij')

    def q(x, y, z):
        return 0.3 * x ** 2 - 0.2 * x * y + 0.7 * y ** 2 * z - 0.1 * z ** 2 * x ** 2 + 2.0 * z + 1.0
    f = q(X * 1.0, Y * 1.0, Z * 1.0)
    rng = np.random.default_rng(1)
    pts = np.stack([rng.uniform(1, nx - 3, 500), rng.uniform(1, ny - 3, 500), rng.uniform(1, nz - 3, 500)])
    np.testing.assert_allclose(O.tricubic(f, pts), q(*pts), rtol=1e-12, atol=1e-10)


def test_tricubic_interpolates_at_nodes():
    rng = np.random.default_rng(2)
    f = rng.standard_normal((6, 5, 4))
    X, Y, Z = np.meshgrid(np.arange(6.0), np.arange(5.0), np.arange(4.0), indexing='ij')
    np.testing.assert_array_equal(O.tricubic(f, np.stack([X, Y, Z])), f)


def test_limiter_bounds_and_nonnegativity():
    # reference tests/test_dynamical_core_3d.py:181-199 (blob shifted half a cell)
    nx, ny, nz = 32, 32, 15
    g = O.Grid(nx, ny, nz, 1000.0, 1000.0, 500.0, 45.0, 0.0)
    X, Y, Z = np.meshgrid(g.x_m, g.y_m, g.z_m, indexing='ij')
    blob = np.exp(-(X ** 2 + Y ** 2 + (Z - 2000.0) ** 2) / 1500.0 ** 2)
    I, J, K = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing='ij')
    co = np.stack([I - 0.5, J - 0.5, K - 0.5]).astype(np.float64)
    un = O.tricubic(blob, co, False)
    li = O.tricubic(blob, co, True)
    assert un.min() < 0.0
    assert li.min() >= 0.0
    rng = np.random.default_rng(3)
    f = rng.standard_normal((7, 6, 5))
    pts = rng.uniform(-1, 8, size=(3, 400))
    lim = O.tricubic(f, pts, True)
    for i in range(pts.shape[1]):
        c = [min(max(pts[a, i], 0), f.shape[a] - 1) for a in range(3)]
        b = [int(np.floor(v)) for v in c]
        corners = [f[min(b[0] + a, 6), min(b[1] + bb, 5), min(b[2] + cc, 4)]
                   for a in (0, 1) for bb in (0, 1) for cc in (0, 1)]
        assert min(corners) - 1e-14 <= lim[i] <= max(corners) + 1e-14


def test_trilinear_nearest_matches_scipy():
    rng = np.random.default_rng(4)
    f = rng.standard_normal((6, 7, 5))
    pts = rng.uniform(-3, 9, size=(3, 500))
    ours = O.map_coordinates_linear_nearest(f, list(pts))
    ref = ndi.map_coordinates(f, pts, order=1, mode='nearest')
    np.testing.assert_allclose(ours, ref, rtol=1e-13, atol=1e-13)
