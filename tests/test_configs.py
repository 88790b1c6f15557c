"""Synthetic model generators (CPU): the curved megathrust mesh and its
random-field friction."""
import numpy as np

from unicycle_amd import configs


def test_megathrust_mesh_geometry():
    m = configs.megathrust_mesh()
    assert m.n == 20000
    v = m.vertices
    # every vertex on z = 5 + 0.1 x + 0.0004 x^2 km (x = east, km)
    xk = v[:, 1] / 1e3
    assert np.allclose(v[:, 2], (5.0 + 0.1 * xk + 0.0004 * xk * xk) * 1e3, rtol=0, atol=1e-6)
    # boundary vertices on the 200 x 100 km outline, interior jitter below 0.1 edge
    assert v[:, 0].min() == 0.0 and v[:, 0].max() == 200e3
    assert v[:, 1].min() == 0.0 and v[:, 1].max() == 100e3
    grid_n = np.round(v[:, 0] / 2000.0) * 2000.0
    grid_e = np.round(v[:, 1] / 1000.0) * 1000.0
    assert np.abs(v[:, 0] - grid_n).max() <= 0.1 * 2000.0 + 1e-9
    assert np.abs(v[:, 1] - grid_e).max() <= 0.1 * 1000.0 + 1e-9
    # 1-based indices, every triangle non-degenerate in plan view
    for idx in (m.i1, m.i2, m.i3):
        assert idx.min() >= 1 and idx.max() <= v.shape[0]
    A, B, C = v[m.i1 - 1], v[m.i2 - 1], v[m.i3 - 1]
    area = 0.5 * np.abs((B[:, 0] - A[:, 0]) * (C[:, 1] - A[:, 1])
                        - (B[:, 1] - A[:, 1]) * (C[:, 0] - A[:, 0]))
    assert area.min() > 0.3 * 0.5 * 2000.0 * 1000.0


def test_megathrust_friction_field():
    m = configs.megathrust_mesh()
    Vpl, rake, mu0, sig, a, b, L, Vo, damping = m.friction()
    amb = a - b
    # one realisation of a field with covariance sd^2 exp(-r^2/(2 l^2)),
    # l = 20 km, over A = 200 x 100 km: the spatial mean has variance
    # ~ sd^2 2 pi l^2 / A, so it lies within 3 of its standard deviations
    sd_mean = 0.003 * np.sqrt(2 * np.pi * 20e3 ** 2 / (200e3 * 100e3))
    assert abs(amb.mean() + 0.002) < 3 * sd_mean
    assert 0.5 * 0.003 < amb.std() < 1.5 * 0.003
    assert abs(np.median(np.log(L)) - np.log(0.02)) < 0.05
    assert abs(np.log(L).std() - 0.5) < 0.05
    # neighbouring triangles are correlated, distant ones are not
    c = (m.vertices[m.i1 - 1] + m.vertices[m.i2 - 1] + m.vertices[m.i3 - 1]) / 3.0
    f = amb - amb.mean()
    near = np.hypot(*(c[1:, :2] - c[:-1, :2]).T) < 3e3
    assert np.corrcoef(f[1:][near], f[:-1][near])[0, 1] > 0.9


def test_gaussian_field_covariance():
    """Random Fourier features: the sample covariance at lag r approaches
    sd^2 exp(-r^2 / (2 corr^2)) over many seeds."""
    pts = np.array([[0.0, 0.0], [10e3, 0.0], [20e3, 0.0], [40e3, 0.0]])
    vals = np.array([configs.gaussian_field(pts, 0.0, 1.0, 20e3, seed=s) for s in range(2000)])
    cov = (vals[:, :1] * vals).mean(axis=0)
    r = pts[:, 0]
    assert np.allclose(cov, np.exp(-r ** 2 / (2 * 20e3 ** 2)), atol=0.08)
