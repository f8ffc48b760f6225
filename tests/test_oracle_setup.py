"""Pins of the oracle's restatement of the set-up and post-processing steps
(oracle/setup.py): properties the reference's formulas imply."""
import numpy as np

from oracle import setup


def test_corners_come_from_the_left_and_right_edges():
    """Top/bottom are written first and left/right overwrite the corners
    (VLPRectangleModel.cs:541-542); min / max cover every value written,
    including the overwritten corner values."""
    u, mn, mx = setup.boundary(5, 4, 0.5, 1.0, bottom=lambda x: 10 + x, top=lambda x: 20 + x,
                               left=lambda y: -1 - y, right=lambda y: 100 + y)
    assert u[0, 0] == -1 and u[0, -1] == -4 and u[-1, 0] == 100 and u[-1, -1] == 103
    assert np.array_equal(u[1:-1, 0], 10 + 0.5 * np.arange(1, 4)) and np.array_equal(u[1:-1, -1], 20 + 0.5 * np.arange(1, 4))
    assert mn == -4 and mx == 103
    assert np.all(u[1:-1, 1:-1] == 0)


def test_boundary_function_on_the_edges():
    g = lambda x, y: x * x - y * y
    u, mn, mx = setup.boundary(9, 5, 0.25, 0.5, border=g)
    x = 0.25 * np.arange(9)
    y = 0.5 * np.arange(5)
    assert np.array_equal(u[:, 0], g(x, 0 * x)) and np.array_equal(u[:, -1], g(x, 2.0 + 0 * x))
    assert np.array_equal(u[0, :], g(0 * y, y)) and np.array_equal(u[-1, :], g(2.0 + 0 * y, y))
    assert (mn, mx) == (-4.0, 4.0)


def test_linear_guesses_reproduce_bilinear_data():
    """For g = a + b x + c y + d x y both one-dimensional interpolations are
    exact, so Linear and WeightLinear reproduce g up to the FP32 weights."""
    g = lambda x, y: 1.5 - 2 * x + 0.75 * y + 3 * x * y
    nx, ny, h = 33, 17, 1 / 32
    u, _, _ = setup.boundary(nx, ny, h, h, border=g)
    x = h * np.arange(nx)[:, None]
    y = h * np.arange(ny)[None, :]
    ex = g(x, y)
    for kind in (2, 3):
        v = setup.init_guess(u, kind)
        assert np.abs(v - ex).max() <= 1e-6 * np.abs(ex).max()
        assert np.array_equal(v[0, :], u[0, :]) and np.array_equal(v[:, -1], u[:, -1])


def test_mean_guesses():
    u, mn, mx = setup.boundary(9, 9, 1 / 8, 1 / 8, border=lambda x, y: 2.0 + 0 * x)
    assert np.all(setup.init_guess(u, 1)[1:-1, 1:-1] == 2.0)     # arithmetic mean of a constant
    v = setup.init_guess(u, 0, (mn + mx) / 2)
    assert np.all(v[1:-1, 1:-1] == 2.0)
    w = setup.init_guess(u, 3)
    assert np.all(w[1:-1, 1:-1] == 2.0)                          # weights sum to one


def test_weight_linear_weights():
    """WeightLinear: weight = lngX / (lngX + lngY) with the distances to the
    nearer edge; at the centre column both distances are equal (0.5)."""
    nx = ny = 9
    u = np.zeros((nx, ny))
    u[0, :], u[-1, :] = 0.0, 0.0      # x-interpolation gives 0
    u[:, 0], u[:, -1] = 1.0, 1.0      # y-interpolation gives 1
    w = setup.init_guess(u, 3)
    i, j = 4, 4
    assert w[i, j] == 0.5             # lngX = lngY = 4
    i, j = 1, 4                        # lngX = 1, lngY = 4: weight 0.2 on the y-interpolation
    assert abs(w[i, j] - np.float32(1 / 5)) < 1e-7


def test_difference_and_min_max():
    nx, ny, h = 9, 5, 0.125
    a = lambda x, y: x + 2 * y
    x = h * np.arange(nx)[:, None]
    y = h * np.arange(ny)[None, :]
    u = a(x, y) + 0.0
    u[3, 2] += 0.5
    u[4, 1] = np.nan
    d, dmin, dmax = setup.difference(u, a, h, h)
    assert dmax == 0.5 and dmin == 0.0
    assert np.all(d[0, :] == 0) and np.all(d[:, 0] == 0) and np.isnan(d[4, 1])
    assert setup.points_min_max(u) == (0.0, float(np.nanmax(u)))
