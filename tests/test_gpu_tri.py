"""GPU parity of the triangular-dislocation path (Nikkhoo & Walter 2015) against
the CPU oracle and the reference's own outputs (golden fixtures), through the
C ABI.

Both the device kernel and the oracle differentiate the displacement field
exactly (forward-mode duals); the reference differentiates by hand.  Same
field, different rounding:
- GPU vs oracle: <= 1e-12 normwise per kernel column block, entries of the
  assembled K within 1e-10 of the row scale (libm and operation order).
- GPU vs reference fixture: <= 1e-11 normwise (the oracle's own bound).
- NaN where the reference yields NaN (receiver on a triangle side).
"""
import os

import numpy as np
import pytest

import orc
from unicycle_amd import configs

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _fixture():
    return dict(np.load(os.path.join(GOLD, "nikkhoo_traction.npz")))


def _oracle(fn, rcv, g, s, d, nout):
    x, sv, dv, nv = rcv[:3], rcv[3:6], rcv[6:9], rcv[9:12]
    ns = g["i1"].size
    out = [np.zeros(ns) for _ in range(nout)]
    getattr(orc.oracle(), fn)(
        orc.p(orc.f64(x)), orc.p(orc.f64(sv)), orc.p(orc.f64(dv)), orc.p(orc.f64(nv)), ns,
        orc.pi(g["i1"]), orc.pi(g["i2"]), orc.pi(g["i3"]), g["vertices"].shape[0],
        orc.p(orc.f64(g["vertices"])), orc.D(s), orc.D(d), orc.D(0.0), orc.D(30e3),
        orc.D(30e3), *[orc.p(o) for o in out])
    return np.stack(out)


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def test_nikkhoo_hooks_vs_reference_and_oracle(engine_factory):
    eng = engine_factory()
    g = _fixture()
    ns, nve = g["i1"].size, g["vertices"].shape[0]
    k = 0
    for rcv in g["receivers"]:
        x, sv, dv, nv = rcv[:3], rcv[3:6], rcv[6:9], rcv[9:12]
        for s, d in [(1.0, 0.0), (0.0, 1.0)]:
            t = np.stack(eng.computeTractionKernelsNikkhoo15(
                x, sv, dv, nv, ns, g["i1"], g["i2"], g["i3"], nve, g["vertices"], s, d, 0.0,
                30e3, 30e3))
            q = np.stack(eng.computeStressKernelsNikkhoo15(
                x, sv, dv, nv, ns, g["i1"], g["i2"], g["i3"], nve, g["vertices"], s, d, 0.0,
                30e3, 30e3))
            assert _rel(t, g["traction"][k]) < 1e-11
            assert _rel(q, g["stress"][k]) < 1e-11
            assert _rel(t, _oracle("orc_traction_kernels_nikkhoo15", rcv, g, s, d, 3)) < 1e-12
            assert _rel(q, _oracle("orc_stress_kernels_nikkhoo15", rcv, g, s, d, 6)) < 1e-12
            k += 1


def test_straininhalfspace_vs_reference(engine_factory):
    """The straininhalfspace_ entry point against the reference's own
    outputs (golden fixture: 300 random triangles and points, unit strike,
    dip and opening slip) with the oracle's bounds (strain by exact
    differentiation vs the reference's hand-derived strain: relative to the
    largest component of each tensor), and against the oracle itself."""
    eng = engine_factory()
    g = dict(np.load(os.path.join(GOLD, "nikkhoo_strain.npz")))
    rels, rels_orc = [], []
    for s in range(3):
        b = [0.0, 0.0, 0.0]
        b[s] = 1.0
        stress, strain = eng.straininhalfspace(g["points"], g["P1"], g["P2"], g["P3"], *b,
                                               30e3, 30e3)
        for i in range(stress.shape[0]):
            X, Y, Z = g["points"][i]
            sg, e = np.zeros(6), np.zeros(6)
            orc.oracle().orc_strain_in_halfspace(
                orc.p(sg), orc.p(e), orc.D(X), orc.D(Y), orc.D(Z), orc.p(orc.f64(g["P1"][i])),
                orc.p(orc.f64(g["P2"][i])), orc.p(orc.f64(g["P3"][i])), *[orc.D(v) for v in b],
                orc.D(30e3), orc.D(30e3))
            for got, want, mine in ((stress[i], g["stress"][i, s], sg),
                                    (strain[i], g["strain"][i, s], e)):
                assert np.array_equal(np.isnan(got), np.isnan(want)), (i, s)
                if np.isnan(want).any():
                    continue
                rels.append(np.abs(got - want).max() / np.abs(want).max())
                rels_orc.append(np.abs(got - mine).max() / np.abs(mine).max())
    rels, rels_orc = np.array(rels), np.array(rels_orc)
    assert np.median(rels) < 1e-12
    assert np.quantile(rels, 0.99) < 1e-10
    assert rels.max() < 1e-8
    assert rels_orc.max() < 1e-10


def test_nikkhoo_edge_cases(engine_factory):
    """Surface triangle (image flip of the reference), horizontal buried
    triangle (strike fallback), receiver on a side (NaN), receiver on the
    extension of a side and far away; opening slip included."""
    eng = engine_factory()
    v = np.array([[0.0, 0.0, 0.0], [2e3, 0.0, 0.0], [0.0, 2e3, 0.0],         # surface
                  [0.0, 0.0, 3e3], [2e3, 0.0, 3e3], [0.0, 2e3, 3e3],         # horizontal
                  [5e3, 0.0, 1e3], [7e3, 0.0, 1e3], [6e3, 0.0, 3e3],         # vertical
                  [0.0, 9e3, 2e3], [1.5e3, 9.5e3, 2.5e3], [0.5e3, 8e3, 4e3]])  # generic
    i1 = np.array([1, 4, 7, 10], dtype=np.int32)
    g = {"i1": i1, "i2": i1 + 1, "i3": i1 + 2, "vertices": v}
    frames = np.array([1.0, 0, 0, 0, -1.0, 0, 0, 0, -1.0])
    rcvs = [np.concatenate([[600.0, 700.0, 1500.0], frames]),
            np.concatenate([[1e3, 0.0, 3e3], frames]),           # on a side of triangle 2
            np.concatenate([[-1e3, 0.0, 3e3], frames]),          # on the extension of it
            np.concatenate([[6e3, 0.0, 2e3], frames]),           # inside the vertical one
            np.concatenate([[4e4, -3e4, 5e3], frames])]
    for rcv in rcvs:
        for s, d, t in [(1.0, 0.0, 0.0), (0.0, 1.0, 0.0), (0.0, 0.0, 1.0)]:
            x, sv, dv, nv = rcv[:3], rcv[3:6], rcv[6:9], rcv[9:12]
            got = np.stack(eng.computeStressKernelsNikkhoo15(
                x, sv, dv, nv, 4, g["i1"], g["i2"], g["i3"], 12, v, s, d, t, 30e3, 30e3))
            out = [np.zeros(4) for _ in range(6)]
            orc.oracle().orc_stress_kernels_nikkhoo15(
                orc.p(orc.f64(x)), orc.p(orc.f64(sv)), orc.p(orc.f64(dv)), orc.p(orc.f64(nv)), 4,
                orc.pi(g["i1"]), orc.pi(g["i2"]), orc.pi(g["i3"]), 12, orc.p(v), orc.D(s),
                orc.D(d), orc.D(t), orc.D(30e3), orc.D(30e3), *[orc.p(o) for o in out])
            want = np.stack(out)
            assert np.array_equal(np.isnan(got), np.isnan(want)), (rcv[:3], s, d, t)
            ok = ~np.isnan(want)
            scale = np.abs(want[ok]).max() if ok.any() else 1.0
            assert np.all(np.abs(got[ok] - want[ok]) <= 1e-11 * scale), (rcv[:3], s, d, t)


def _tri_engine(engine_factory, tm):
    eng = engine_factory()
    eng.set_elastic(tm.lambda_, tm.mu)
    eng.set_triangles(*tm.geometry())
    eng.set_friction(*tm.friction())
    return eng


def test_buildG_triangles_matches_oracle(engine_factory):
    tm = configs.triangulate(configs.thrust_fault(nx=16, nz=8, ellipse=None))
    eng = _tri_engine(engine_factory, tm)
    eng.buildG()
    K = eng.get_kernel()
    Kref = orc.buildG(orc.Mesh(tm.lambda_, tm.mu, tri=tm.geometry()), nthreads=8)
    assert K.shape == Kref.shape == (3 * tm.n, 2 * tm.n)
    assert not np.isnan(K).any()
    assert _rel(K, Kref) < 1e-12
    rowscale = np.abs(Kref).max(axis=1, keepdims=True)
    assert np.all(np.abs(K - Kref) <= 1e-10 * rowscale)


def test_buildG_megathrust_mesh_matches_oracle(engine_factory):
    """The curved, jittered megathrust mesh (non-planar, irregular
    triangles of both orientations) at 12 x 8 quads: K and the triangle
    frames against the oracle."""
    tm = configs.megathrust_mesh(nx=12, ny=8, seed=5)
    eng = _tri_engine(engine_factory, tm)
    eng.buildG()
    K = eng.get_kernel()
    Kref = orc.buildG(orc.Mesh(tm.lambda_, tm.mu, tri=tm.geometry()), nthreads=8)
    assert K.shape == Kref.shape == (3 * tm.n, 2 * tm.n)
    assert not np.isnan(K).any()
    assert _rel(K, Kref) < 1e-12
    rowscale = np.abs(Kref).max(axis=1, keepdims=True)
    assert np.all(np.abs(K - Kref) <= 1e-10 * rowscale)


def test_buildG_mixed_mesh_layout(engine_factory):
    """Rectangles first, then triangles (ratestate.f90:1174-1290 element
    order) in both the rows and the columns of K."""
    rm = configs.thrust_fault(nx=6, nz=3, ellipse=None)
    tm = configs.triangulate(configs.thrust_fault(nx=4, nz=2, ellipse=None, top=9000.0))
    eng = engine_factory()
    eng.set_elastic(rm.lambda_, rm.mu)
    eng.set_rectangles(*rm.geometry())
    eng.set_triangles(*tm.geometry())
    fr = tuple(np.concatenate([a, b]) for a, b in zip(rm.friction(), tm.friction()))
    eng.set_friction(*fr)
    eng.buildG()
    K = eng.get_kernel()
    Kref = orc.buildG(orc.Mesh(rm.lambda_, rm.mu, rect=rm.geometry(), tri=tm.geometry()),
                      nthreads=8)
    n = rm.n + tm.n
    assert K.shape == Kref.shape == (3 * n, 2 * n)
    rowscale = np.abs(Kref).max(axis=1, keepdims=True)
    # Okada entries carry binary32 DC3D rounding: compare at that resolution
    assert np.all(np.abs(K - Kref) <= 1e-6 * rowscale)
    tri_cols = slice(2 * rm.n, None)
    assert _rel(K[:, tri_cols], Kref[:, tri_cols]) < 1e-12


def test_triangle_odefun_and_steps_match_oracle(engine_factory):
    """RHS on a triangle mesh (rate-and-state per triangle), then 200
    device-resident adaptive steps against the oracle's."""
    tm = configs.triangulate(configs.thrust_fault(nx=8, nz=4, ellipse=None))
    eng = _tri_engine(engine_factory, tm)
    K = orc.buildG(orc.Mesh(tm.lambda_, tm.mu, tri=tm.geometry()), nthreads=8)
    eng.set_kernel(K)
    sysm = orc.System(tm.friction(), K)
    y0 = sysm.init_state()
    np.testing.assert_array_equal(eng.initStateVector(), y0)
    rng = np.random.default_rng(9)
    y = y0.copy()
    y[6::7] += 0.5 * rng.standard_normal(tm.n)
    y[5::7] += 0.3 * rng.standard_normal(tm.n)
    dydt = np.empty_like(y)
    vmax = eng.odefun(y.size, 0.0, y, dydt)
    want, vmax_ref, _, _ = sysm.odefun(y)
    np.testing.assert_allclose(dydt, want, rtol=1e-12, atol=1e-12 * np.abs(want).max())
    assert vmax == pytest.approx(vmax_ref, rel=1e-14)
    k = 200
    _, st, tl, dl, rc = sysm.integrate(y0, 10 * configs.YEAR, max_steps=k, log_cap=k)
    res = eng.integrate(y0, 10 * configs.YEAR, max_steps=k, log_cap=k)
    assert res.steps == st.steps == k
    assert res.attempts == st.attempts
    np.testing.assert_allclose(res.dt_log, dl, rtol=1e-5)
