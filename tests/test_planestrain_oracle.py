"""CPU pinning of the 2-D plane-strain oracle (oracle/orc_planestrain.cpp,
with the displacement integrals of unicycle_amd/csrc/ps_jint.h) against the
reference's own routines (tests/golden/planestrain.npz, written from
oracle/_ref; the live checker when built) and against the mathematics.

- The displacement integrals keep the reference's operand order: the
  displacement is bit for bit the reference's.
- The stress is Hooke's law on the exact (forward-mode) gradient of that
  displacement, less the eigenstrain inside.  The reference differentiates
  by hand, so the two agree to rounding: 1e-11 of the stress scale for a
  strain volume.  A line dislocation is a zone 1e-4 of its width thick, whose
  four corner terms cancel four digits, so there 1e-7 of the scale.
- Central differences of the (bit-exact) displacement reproduce the stress.
- The free surface carries no s23 or s33.
- Across a dislocation the displacement jumps by the slip along dip.
- The odefun of one patch by hand; the initial state with and without tau0.
"""
import ctypes as C
import os

import numpy as np
import pytest

import orc
from unicycle_amd import configs

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "planestrain.npz")


def _sv_tol(s, G, e):
    """1e-11 of the stress, or 1e-13 of G |e|, the size of the terms that
    cancel (the free surface's zero components, say)"""
    return 1e-11 * np.abs(s).max() + 1e-13 * G * np.abs(e).max()


def _fault_tol(s, G, slip, W):
    """1e-7 of the stress, or 1e-13 of G slip / T with T = W / 1e4"""
    return 1e-7 * np.abs(s).max() + 1e-13 * G * abs(slip) / (W / 1e4)


def test_fixture():
    g = np.load(GOLD)
    G, lam = float(g["G"]), float(g["lam"])
    for r, s, u in zip(g["inputs"], g["stress"], g["displacement"]):
        assert orc.displacement_sv_planestrain(*r, G, lam) == tuple(u)
        got = np.array(orc.stress_sv_planestrain(*r, G, lam))
        assert np.abs(got - s).max() <= _sv_tol(s, G, r[7:10]), (r, got, s)
    for r, s, u in zip(g["fault_inputs"], g["fault_stress"], g["fault_displacement"]):
        assert orc.displacement_planestrain(*r, G, lam) == tuple(u)
        got = np.array(orc.stress_planestrain(*r, G, lam))
        assert np.abs(got - s).max() <= _fault_tol(s, G, r[6], r[4]), (r, got, s)
    ns = g["W"].size
    lib = orc.oracle()
    for rcv, td, tn in zip(g["receivers"], g["td"], g["tn"]):
        a, b = np.zeros(ns), np.zeros(ns)
        rc = lib.orc_traction_kernels_planestrain(
            *[orc.p(orc.f64(rcv[3 * k:3 * k + 3])) for k in range(4)], ns, orc.p(orc.f64(g["y"])),
            orc.p(orc.f64(g["W"])), orc.p(orc.f64(g["dip"])), C.c_double(G), C.c_double(lam),
            orc.p(a), orc.p(b))
        assert rc == 0
        for got, want in ((a, td), (b, tn)):
            tol = 1e-7 * np.abs(want) + 1e-13 * G / (g["W"] / 1e4)
            assert np.all(np.abs(got - want) <= tol)


def test_live_reference_checker():
    ref = orc.ref()
    if ref is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    D = C.c_double
    rng = np.random.default_rng(21)
    G, lam = 30e3, 45e3
    for _ in range(300):
        q2, q3 = rng.uniform(-1e4, 1e4), rng.uniform(0, 1e4)
        T, W = rng.uniform(100, 5e3, 2)
        dip = rng.uniform(0.1, np.pi - 0.1)
        r = (q2 + rng.uniform(-2, 2) * W, max(0.0, q3 + rng.uniform(-1, 2) * W), q2, q3, T, W,
             dip, *rng.standard_normal(3))
        o = [D() for _ in range(5)]
        ref.ref_stress_sv_planestrain(*[D(v) for v in r], D(G), D(lam), *[C.byref(x) for x in o[:3]])
        ref.ref_displacement_sv_planestrain(*[D(v) for v in r], D(G), D(lam),
                                            *[C.byref(x) for x in o[3:]])
        s = np.array([x.value for x in o[:3]])
        assert orc.displacement_sv_planestrain(*r, G, lam) == (o[3].value, o[4].value)
        got = np.array(orc.stress_sv_planestrain(*r, G, lam))
        assert np.abs(got - s).max() <= _sv_tol(s, G, r[7:10])
    # frames: the reference's routine against the oracle's (the antiplane one)
    x = np.stack([np.zeros(40), rng.uniform(-1e4, 1e4, 40), rng.uniform(0, 1e4, 40)], 1)
    W, dip = rng.uniform(500, 3e3, 40), rng.uniform(0.2, 2.9, 40)
    outs = []
    for lib, fn in ((orc.oracle(), "orc_refsys_antiplane"), (ref, "ref_refsys_planestrain")):
        o = [np.zeros((40, 3)) for _ in range(4)]
        getattr(lib, fn)(40, orc.p(orc.f64(x)), orc.p(W), orc.p(dip), *[orc.p(a) for a in o])
        outs.append(o)
    for a, b in zip(*outs):
        assert np.array_equal(a, b)


def test_stress_is_hooke_of_the_displacement_gradient():
    """Central differences of the displacement (values only) against the
    forward-mode stress, away from the zone's edges."""
    rng = np.random.default_rng(5)
    G, lam = 30e3, 25e3
    checked = 0
    while checked < 40:
        q2, q3 = rng.uniform(-5e3, 5e3), rng.uniform(500, 5e3)
        T, W = rng.uniform(1e3, 4e3, 2)
        dip = rng.uniform(0.3, np.pi - 0.3)
        r2, x3 = q2 + rng.uniform(-1.5, 1.5) * W, q3 + rng.uniform(-0.5, 1.5) * W
        x2 = r2 - q2
        a = x2 * np.sin(dip) - (x3 - q3) * np.cos(dip)       # across the zone
        b = x2 * np.cos(dip) + (x3 - q3) * np.sin(dip)       # along dip
        if x3 < 100 or min(abs(abs(a) - T / 2), abs(b), abs(b - W)) < 100:
            continue
        e = rng.standard_normal(3)
        s = np.array(orc.stress_sv_planestrain(r2, x3, q2, q3, T, W, dip, *e, G, lam))
        h = 1e-2
        u = lambda p, q: np.array(orc.displacement_sv_planestrain(p, q, q2, q3, T, W, dip, *e, G, lam))
        d2 = (u(r2 + h, x3) - u(r2 - h, x3)) / (2 * h)
        d3 = (u(r2, x3 + h) - u(r2, x3 - h)) / (2 * h)
        inside = abs(a) < T / 2 and 0 < b < W
        c, sn = np.cos(dip), np.sin(dip)
        e22 = sn * (e[0] * sn + e[1] * c) + c * (e[1] * sn + e[2] * c)
        e23 = sn * (-e[0] * c + e[1] * sn) + c * (-e[1] * c + e[2] * sn)
        e33 = -c * (-e[0] * c + e[1] * sn) + sn * (-e[1] * c + e[2] * sn)
        E22 = d2[0] - (e22 if inside else 0.0)
        E23 = (d3[0] + d2[1]) / 2 - (e23 if inside else 0.0)
        E33 = d3[1] - (e33 if inside else 0.0)
        want = np.array([lam * (E22 + E33) + 2 * G * E22, 2 * G * E23, lam * (E22 + E33) + 2 * G * E33])
        assert np.abs(s - want).max() <= 1e-6 * G * np.abs(e).max(), (s, want)
        checked += 1


def test_free_surface_and_decay():
    rng = np.random.default_rng(6)
    G, lam = 30e3, 30e3
    for _ in range(100):
        q2, q3, W = rng.uniform(-5e3, 5e3), rng.uniform(100, 5e3), rng.uniform(500, 4e3)
        dip = rng.uniform(0.2, np.pi - 0.2)
        s22, s23, s33 = orc.stress_planestrain(rng.uniform(-3e4, 3e4), 0.0, q2, q3, W, dip, 1.0,
                                               G, lam)
        sc = abs(s22) + G * W / 1e8
        assert abs(s23) <= 1e-7 * sc and abs(s33) <= 1e-7 * sc
    # far field of a dislocation (a dipole of edge dislocations): stress
    # ~ r^-2, displacement ~ r^-1, at depth below it and along the surface
    # (the source's depth enters at order q3 / r)
    q2, q3, W, dip = 0.0, 5e3, 2e3, 0.6
    for at in (lambda r: (0.0, r), lambda r: (r, 0.0)):
        f = [np.array(orc.stress_planestrain(*at(r), q2, q3, W, dip, 1.0, G, lam)) for r in (4e6, 8e6)]
        u = [np.array(orc.displacement_planestrain(*at(r), q2, q3, W, dip, 1.0, G, lam)) for r in (4e6, 8e6)]
        assert np.linalg.norm(f[0]) / np.linalg.norm(f[1]) == pytest.approx(4.0, rel=5e-3)
        assert np.linalg.norm(u[0]) / np.linalg.norm(u[1]) == pytest.approx(2.0, rel=5e-3)


@pytest.mark.parametrize("dip", [0.4, np.pi / 2, 2.3])
def test_displacement_jumps_by_the_slip(dip):
    """Crossing the dislocation at its middle, the displacement changes by the
    slip along dip (the zone's simple shear 2 e23' T = -s, along the dip
    direction (cos dip, sin dip) of the frame)."""
    G, lam, q2, q3, W, s = 30e3, 30e3, 1e3, 2e3, 3e3, 0.7
    mid = np.array([q2 + W / 2 * np.cos(dip), q3 + W / 2 * np.sin(dip)])
    across = np.array([np.sin(dip), -np.cos(dip)])
    d = 1.0                        # 1 m either side; the zone is 0.3 m thick
    ua = np.array(orc.displacement_planestrain(*(mid + d * across), q2, q3, W, dip, s, G, lam))
    ub = np.array(orc.displacement_planestrain(*(mid - d * across), q2, q3, W, dip, s, G, lam))
    jump = ua - ub
    along = np.array([np.cos(dip), np.sin(dip)])
    assert jump @ along == pytest.approx(-s, rel=2e-3)
    assert abs(jump @ across) <= 2e-3 * s


def test_odefun_one_patch_and_initial_state():
    m = configs.planestrain_fault(n=1, length=2e3)
    K = orc.buildG_planestrain(m.x, m.width, m.dip, m.mu, m.lambda_)
    s = orc.PSSystem(m.oracle_friction(), K)
    y = s.init_state()
    lg = np.log(10.0)
    a, b, sig, L, eta, Vpl, Vo, mu0 = (m.a[0], m.b[0], m.sig[0], m.L[0], m.damping[0], m.Vpl[0],
                                       m.Vo[0], m.mu0[0])
    np.testing.assert_allclose(y, [0.0, sig * (mu0 + (a - b) * np.log(Vpl / Vo)) - eta * Vpl, 0.0,
                                   np.log10(L / Vpl), np.log10(0.98 * Vpl)], rtol=1e-15)
    s0 = orc.PSSystem(m.oracle_friction(), K, tau0=[37.5])
    assert s0.init_state()[1] == 37.5
    y = np.array([0.3, 50.0, -1.0, 8.5, -8.7])
    dydt, vmax = s.odefun(y)
    V = 10.0 ** y[4]
    t = K @ np.array([V - Vpl])
    dth = (10.0 ** -y[3] - V / L) / lg
    dv = (t[0] - b * sig * dth * lg) / (a * sig + eta * V) / lg
    np.testing.assert_allclose(dydt, [V, t[0] - eta * V * dv * lg, t[1], dth, dv], rtol=1e-12)
    assert vmax == pytest.approx(V, rel=1e-15)
