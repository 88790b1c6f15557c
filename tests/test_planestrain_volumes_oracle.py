"""CPU pinning of the plane-strain viscouscycles oracle (strain volumes in
oracle/orc_planestrain.cpp) against the reference's own routines
(tests/golden/planestrain_volumes.npz, written from oracle/_ref; the live
checker when built) and against the mathematics.

- The kernel routines of planestrain.f90:221-475 (traction and volume-frame
  stress kernels of volumes, volume-frame stress kernels of segments, the
  displacement kernels of volumes): stresses within the bounds of the
  plane-strain oracle's own tests (1e-11 of the stress for volumes, 1e-7 for
  line dislocations, or 1e-13 of the terms that cancel), displacements bit
  for bit.
- buildG of segments and volumes (greens.f90:66-226) against the
  reference's kernels assembled in the same layout, including the volume
  rows of segment sources, which stressRows computes with G and lambda
  exchanged (it passes mu, lambda into computeStressKernelsPlaneStrain's
  lambda, G).
- The volume columns are strains of 1e-3: linear in the strain.
- The initial stress flows at the background rate (v = 0), for n = 1 and
  n = 3 with an Arrhenius factor; the s0 option replaces it.
- The odefun's strain rate is homogeneous of degree n in the stress.
- A lone Maxwell volume (n = 1) integrates as the matrix exponential of its
  linear system.
"""
import ctypes as C
import os

import numpy as np
import pytest
import scipy.linalg

import orc
from unicycle_amd import configs

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "planestrain_volumes.npz")


def _sv_tol(s, G, e):
    return 1e-11 * np.abs(s).max() + 1e-13 * G * np.abs(e).max()


def _fault_tol(s, G, W):
    return 1e-7 * np.abs(s).max() + 1e-13 * G / (W / 1e4)


def _frame(r):
    return r[0:3], r[3:6], r[6:9], r[9:12]


def test_fixture_kernels():
    g = np.load(GOLD)
    G, lam = float(g["G"]), float(g["lam"])
    y, T, W, dip = g["y"], g["T"], g["W"], g["dip"]
    ns = W.size
    for r, td, tn, vs, fs in zip(g["receivers"], g["td"], g["tn"], g["vstress"], g["fstress"]):
        e = vs[:3]
        a, b = orc.ps_kernel_call("orc_traction_kernels_sv_planestrain", *_frame(r), y,
                                  (T, W, dip), (*e, G, lam), 2)
        for got, want in ((a, td), (b, tn)):
            assert np.all(np.abs(got - want) <= _sv_tol(want, G, e))
        o = orc.ps_kernel_call("orc_stress_kernels_sv_planestrain", *_frame(r), y, (T, W, dip),
                               (*e, G, lam), 3)
        want = vs[3:].reshape(3, ns)
        for k in range(3):
            assert np.all(np.abs(o[k] - want[k]) <= _sv_tol(want[k], G, e)), k
        o = orc.ps_kernel_call("orc_stress_kernels_planestrain", *_frame(r), y, (W, dip),
                               (G, lam), 3)
        want = fs.reshape(3, ns)
        for k in range(3):
            assert np.all(np.abs(o[k] - want[k]) <= _fault_tol(want[k], G, W)), k
    # buildO's unit strains, including the zero third row of its eye
    eyes = ((1.0, 0.0, 0.0), (0.0, 1.0, 0.0), (0.0, 0.0, 0.0))
    d = g["disp"].reshape(len(g["points"]), 3, 2, ns)
    for q, dq in zip(g["points"], d):
        for j, eye in enumerate(eyes):
            for s in range(ns):
                u = orc.displacement_sv_planestrain(q[1], q[2], y[s, 1], y[s, 2], T[s], W[s],
                                                    dip[s], *eye, G, lam)
                assert u == (dq[j, 0, s], dq[j, 1, s])
    assert not d[:, 2].any()


def _mesh_tol(K, G, pW, nps):
    """1e-7 of the row for segment columns (a zone 1e-4 of the width thick)
    with 1e-13 G / T, 1e-11 of the row for volume columns with 1e-13 G 1e-3"""
    rowmax = np.abs(K).max(axis=1, keepdims=True)
    tol = np.empty_like(K)
    tol[:, :nps] = 1e-7 * rowmax + 1e-13 * G / (pW / 1e4)
    tol[:, nps:] = 1e-11 * rowmax + 1e-13 * G * 1e-3
    return tol


def test_fixture_buildG():
    g = np.load(GOLD)
    G, lam = float(g["G"]), float(g["lam"])
    K = orc.buildG_planestrain_volumes(g["px"], g["pW"], g["pdip"], g["vx"], g["vT"], g["vW"],
                                       g["vdip"], G, lam, nthreads=4)
    Kr = g["K"]
    assert K.shape == Kr.shape == (2 * 6 + 3 * 5, 6 + 3 * 5)
    err = np.abs(K - Kr)
    tol = _mesh_tol(Kr, G, g["pW"], 6)
    print(f"buildG vs the reference's kernels: max err / bound {(err / tol).max():.2e}", flush=True)
    assert np.all(err <= tol)


def test_live_reference_checker():
    ref = orc.ref()
    if ref is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    rng = np.random.default_rng(31)
    G, lam = 30e3, 45e3
    for _ in range(20):
        ns = 30
        y = np.stack([np.zeros(ns), rng.uniform(-2e4, 2e4, ns), rng.uniform(0, 1e4, ns)], 1)
        T, W = rng.uniform(200, 5e3, ns), rng.uniform(200, 5e3, ns)
        dip = rng.uniform(0.1, np.pi - 0.1, ns)
        d = rng.uniform(0.1, np.pi - 0.1)
        fr = (np.array([0.0, rng.uniform(-1e4, 1e4), rng.uniform(0, 1e4)]), np.array([1.0, 0, 0]),
              np.array([0.0, -np.cos(d), -np.sin(d)]), np.array([-0.0, np.sin(d), -np.cos(d)]))
        e = rng.standard_normal(3)
        for fo, fr_, nout in (("orc_traction_kernels_sv_planestrain",
                               "ref_traction_sv_planestrain", 2),
                              ("orc_stress_kernels_sv_planestrain",
                               "ref_stress_kernels_sv_planestrain", 3)):
            got = orc.ps_kernel_call(fo, *fr, y, (T, W, dip), (*e, G, lam), nout)
            want = orc.ps_kernel_call(fr_, *fr, y, (T, W, dip), (*e, G, lam), nout, lib=ref)
            for a, b in zip(got, want):
                assert np.all(np.abs(a - b) <= _sv_tol(b, G, e))
        # segments: the reference routine takes (lambda, G)
        got = orc.ps_kernel_call("orc_stress_kernels_planestrain", *fr, y, (W, dip), (G, lam), 3)
        want = orc.ps_kernel_call("ref_stress_kernels_planestrain", *fr, y, (W, dip), (lam, G), 3,
                                  lib=ref)
        for a, b in zip(got, want):
            assert np.all(np.abs(a - b) <= _fault_tol(b, G, W))


def _volume_rows_of_segments(m, G, lam):
    """orc_stress_kernels_planestrain on every volume receiver, moduli as given"""
    f = m.fault
    ne = m.n + m.nvol
    sv, dv, nv, xc = (np.zeros((m.nvol, 3)) for _ in range(4))
    orc.oracle().orc_refsys_antiplane(m.nvol, orc.p(orc.f64(m.vx)), orc.p(orc.f64(m.vwidth)),
                                      orc.p(orc.f64(m.vdip)), *[orc.p(a) for a in (sv, dv, nv, xc)])
    rows = np.zeros((3 * m.nvol, m.n))
    for k in range(m.nvol):
        o = orc.ps_kernel_call("orc_stress_kernels_planestrain", xc[k], sv[k], dv[k], nv[k], f.x,
                               (f.width, f.dip), (G, lam), 3)
        rows[3 * k:3 * k + 3] = o
    assert ne == m.n + m.nvol
    return rows


@pytest.mark.parametrize("lam", [30e3, 20e3])
def test_volume_rows_of_segments_exchange_the_moduli(lam):
    m = configs.planestrain_viscous(n=12, n2=3, n3=2, lambda_=lam)
    f = m.fault
    K = orc.buildG_planestrain_volumes(f.x, f.width, f.dip, m.vx, m.thickness, m.vwidth, m.vdip,
                                       f.mu, f.lambda_, nthreads=4)
    block = K[2 * m.n:, :m.n]
    exchanged = _volume_rows_of_segments(m, f.lambda_, f.mu)
    assert np.array_equal(block, exchanged)
    plain = _volume_rows_of_segments(m, f.mu, f.lambda_)
    if lam == f.mu:
        assert np.array_equal(block, plain)
    else:
        assert np.abs(block - plain).max() > 1e-3 * np.abs(plain).max()


def test_volume_columns_are_strains_of_1e_3():
    """K's volume columns against the kernel routines at unit strain"""
    m = configs.planestrain_viscous(n=8, n2=2, n3=2)
    f = m.fault
    K = orc.buildG_planestrain_volumes(f.x, f.width, f.dip, m.vx, m.thickness, m.vwidth, m.vdip,
                                       f.mu, f.lambda_)
    sv, dv, nv, xc = (np.zeros((f.n, 3)) for _ in range(4))
    orc.oracle().orc_refsys_antiplane(f.n, orc.p(orc.f64(f.x)), orc.p(orc.f64(f.width)),
                                      orc.p(orc.f64(f.dip)), *[orc.p(a) for a in (sv, dv, nv, xc)])
    for r in range(f.n):
        for j in range(3):
            e = [float(j == k) for k in range(3)]
            td, tn = orc.ps_kernel_call("orc_traction_kernels_sv_planestrain", xc[r], sv[r], dv[r],
                                        nv[r], m.vx, (m.thickness, m.vwidth, m.vdip),
                                        (*e, f.mu, f.lambda_), 2)
            # the corner terms round differently at strains 1 and 1e-3
            for got, want in ((K[2 * r, f.n + j::3], 1e-3 * td), (K[2 * r + 1, f.n + j::3], 1e-3 * tn)):
                assert np.all(np.abs(got - want) <= _sv_tol(want, f.mu, [1e-3]))
                assert np.abs(want).max() > 0


def _system(m, **kw):
    f = m.fault
    K = orc.buildG_planestrain_volumes(f.x, f.width, f.dip, m.vx, m.thickness, m.vwidth, m.vdip,
                                       f.mu, f.lambda_, nthreads=4)
    return orc.PSSystem(f.oracle_friction(), K, rheo=m.oracle_rheology(), **kw), K


@pytest.mark.parametrize("npower,Q", [(1.0, 0.0), (3.0, 1.5e5)])
def test_initial_stress_flows_at_the_background_rate(npower, Q):
    m = configs.planestrain_viscous(n=10, n2=3, n3=2, npower=npower, Q=Q,
                                    gammadot0=1e-12 if npower == 1 else 1e-5)
    m.e23[:] = 3e-16
    s, _ = _system(m)
    y = s.init_state()
    dydt, _ = s.odefun(y)
    q = dydt[5 * m.n:].reshape(-1, 6)
    ekk = m.e22 + m.e33
    want = np.stack([m.e22 - ekk / 2, m.e23, m.e33 - ekk / 2], 1)
    np.testing.assert_allclose(q[:, :3], want, rtol=1e-12)
    # the background here is deviatoric: the volumes' slip-rate entries vanish,
    # so the stress rates are the patches' contribution alone
    yv = y[5 * m.n:].reshape(-1, 6)
    assert not yv[:, :3].any()
    # s0 replaces it where not all zero
    s0 = np.zeros((m.nvol, 3))
    s0[1] = [0.5, -0.25, 0.125]
    s2, _ = _system(m, s0=s0.ravel())
    y2 = s2.init_state()
    yv2 = y2[5 * m.n:].reshape(-1, 6)
    np.testing.assert_array_equal(yv2[1, 3:], s0[1])
    np.testing.assert_array_equal(np.delete(yv2, 1, 0), np.delete(yv, 1, 0))


def test_strain_rate_is_homogeneous_of_degree_n():
    m = configs.planestrain_viscous(n=4, n2=2, n3=1, npower=3.0, gammadot0=1e-5, Q=1e5)
    s, _ = _system(m)
    y = s.init_state()
    rng = np.random.default_rng(3)
    yv = y[5 * m.n:].reshape(-1, 6)
    yv[:, 3:] = rng.standard_normal((m.nvol, 3))
    y[5 * m.n:] = yv.ravel()
    d1 = s.odefun(y)[0][5 * m.n:].reshape(-1, 6)[:, :3]
    y2 = y.copy()
    y2[5 * m.n:] = (yv * np.array([1, 1, 1, 2, 2, 2])).ravel()
    d2 = s.odefun(y2)[0][5 * m.n:].reshape(-1, 6)[:, :3]
    np.testing.assert_allclose(d2, 8.0 * d1, rtol=1e-13)
    # the 2-D deviator: adding pressure (s22 = s33 += p) changes nothing
    y3 = y.copy()
    y3[5 * m.n:] = (yv + np.array([0, 0, 0, 7.0, 0, 7.0])).ravel()
    d3 = s.odefun(y3)[0][5 * m.n:].reshape(-1, 6)[:, :3]
    np.testing.assert_allclose(d3, d1, rtol=1e-9)


def test_lone_maxwell_volume_matches_the_matrix_exponential():
    """No patches, one volume, n = 1, no background rate: the state is
    linear, y' = B y with ds/dt = 1e3 K g D s and de/dt = g D s (D the 2-D
    deviatoric projection), so y(t) = expm(B t) y(0)."""
    m = configs.planestrain_viscous(n=1, n2=1, n3=1, gammadot0=3e-11, e22=0.0, e33=0.0)
    f = m.fault
    K = orc.buildG_planestrain_volumes(np.zeros((0, 3)), np.zeros(0), np.zeros(0), m.vx,
                                       m.thickness, m.vwidth, m.vdip, f.mu, f.lambda_)
    assert K.shape == (3, 3)
    empty = [np.zeros(0)] * 8
    s0 = np.array([0.3, -0.2, 0.1])
    s = orc.PSSystem(empty, K, rheo=m.oracle_rheology(), s0=s0)
    y0 = s.init_state()
    np.testing.assert_array_equal(y0, [0, 0, 0, *s0])
    g = m.gammadot0[0]
    Dp = np.array([[0.5, 0, -0.5], [0, 1, 0], [-0.5, 0, 0.5]])
    B = np.zeros((6, 6))
    B[:3, 3:] = g * Dp
    B[3:, 3:] = 1e3 * K @ (g * Dp)
    t_end = 0.5 / np.abs(np.linalg.eigvals(B[3:, 3:])).max()   # half a relaxation time
    y, st, *_ = s.integrate(y0, t_end, max_steps=100000)
    # the last step overshoots t_end: compare at the time reached
    want = scipy.linalg.expm(B * st.time) @ y0
    print(f"lone Maxwell volume: {st.steps} steps to t = {st.time:.4g} s", flush=True)
    np.testing.assert_allclose(y, want, rtol=2e-5, atol=2e-5 * np.abs(want).max())
