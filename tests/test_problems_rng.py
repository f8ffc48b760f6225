"""Known-answer tests of the input generators."""
import numpy as np

from vlp2d_amd import problems


def test_splitmix64_known_answers():
    # Published SplitMix64 outputs for seed 1234567.
    r = problems.SplitMix64(1234567)
    got = [int(v) for v in r.next(5)]
    assert got == [6457827717110365317, 3203168211198807973, 9817491932198370423, 4593380528125082431,
                   16408922859458223821]
    r2 = problems.SplitMix64(1234567)
    assert [int(r2.next(1)[0]) for _ in range(5)] == got  # streaming == batched


def test_uniform_normal_ranges():
    r = problems.SplitMix64(9)
    u = r.uniform(100000)
    assert u.min() >= 0 and u.max() < 1 and abs(u.mean() - 0.5) < 0.01
    z = r.normal(100000)
    assert abs(z.mean()) < 0.02 and abs(z.std() - 1) < 0.02
    k = r.integers(1, 32, 10000)
    assert k.min() == 1 and k.max() == 32


def test_configs_shapes_and_boundaries():
    p = problems.config(2, 64)
    assert p.shape == (65, 33) and p.hx == p.hy == 1 / 32
    assert np.all(p.g[1:-1, 0] == 0) and np.all(p.g[1:-1, -1] == 0)
    p3 = problems.config(3, 64)
    x = np.arange(65) / 64
    np.testing.assert_allclose(p3.g[:, 0], np.sin(3 * x))
    np.testing.assert_allclose(p3.g[-1, :], np.sin(3.0) * np.cosh(3 * x))
    # separable assembly equals the direct sum
    rng = problems.SplitMix64(3)
    xc, yc = rng.uniform(64), rng.uniform(64)
    w = 0.01 + 0.04 * rng.uniform(64)
    A = 100.0 * rng.normal(64)
    X, Y = np.meshgrid(x, x, indexing="ij")
    direct = sum(A[c] * np.exp(-((X - xc[c]) ** 2 + (Y - yc[c]) ** 2) / (2 * w[c] ** 2)) for c in range(64))
    np.testing.assert_allclose(p3.f, direct, rtol=1e-12, atol=1e-9)
    p5 = problems.config(5, 64)
    assert p5.f.shape == (65, 65)
