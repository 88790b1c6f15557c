"""CPU pinning of the 2-D antiplane oracle (oracle/orc_antiplane.c) against
the reference's own routines (tests/golden/antiplane.npz from oracle/_ref,
the live checker when built) and the mathematics: the stress is G times the
gradient of the reference's displacement; the free surface carries no s13;
the displacement jumps by the slip across the segment; the odefun of one
patch by hand."""
import ctypes as C
import math
import os

import numpy as np
import pytest

import orc

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "antiplane.npz")
ORC = orc.oracle()
D = C.c_double
ORC.orc_displacement_antiplane.restype = D


def stress(x2, x3, q2, q3, W, dip, G=30e3):
    a, b = D(), D()
    ORC.orc_stress_antiplane(*[D(v) for v in (x2, x3, q2, q3, W, dip, G)], C.byref(a), C.byref(b))
    return a.value, b.value


def disp(x2, x3, q2, q3, W, dip=0.5):
    return ORC.orc_displacement_antiplane(*[D(v) for v in (x2, x3, q2, q3, W, dip, 1.0)])


def test_stress_and_traction_bit_exact_vs_reference_fixture():
    g = np.load(GOLD)
    for r, want in zip(g["stress_inputs"], g["stress"]):
        assert stress(*r) == tuple(want)
    ns = g["W"].size
    for rcv, want in zip(g["receivers"], g["traction"]):
        out = np.zeros(ns)
        ORC.orc_traction_kernels_antiplane(*[orc.p(orc.f64(rcv[3 * k:3 * k + 3])) for k in range(4)],
                                           ns, orc.p(orc.f64(g["y"])), orc.p(orc.f64(g["W"])),
                                           orc.p(orc.f64(g["dip"])), D(30e3), orc.p(out))
        assert np.array_equal(out, want)


def test_live_reference_checker():
    ref = orc.ref()
    if ref is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    rng = np.random.default_rng(3)
    x = np.stack([np.zeros(50), rng.uniform(-1e4, 1e4, 50), rng.uniform(0, 1e4, 50)], 1)
    W, dip = rng.uniform(500, 3e3, 50), rng.uniform(0.2, 1.4, 50)
    outs = []
    for lib, fn in ((ORC, "orc_refsys_antiplane"), (ref, "ref_refsys_antiplane")):
        o = [np.zeros((50, 3)) for _ in range(4)]
        getattr(lib, fn)(50, orc.p(orc.f64(x)), orc.p(W), orc.p(dip), *[orc.p(a) for a in o])
        outs.append(o)
    for a, b in zip(*outs):
        assert np.array_equal(a, b)


def test_stress_is_shear_modulus_times_displacement_gradient():
    """s12 = G du/dx2, s13 = G du/dx3 for the reference's displacement of unit
    slip (antiplane.f90:178-191): a dropped image term or a sign fails it."""
    rng = np.random.default_rng(7)
    G = 30e3
    for _ in range(50):
        q2, q3, W = rng.uniform(-5e3, 5e3), rng.uniform(0, 5e3), rng.uniform(500, 5e3)
        x2, x3 = q2 + rng.choice([-1, 1]) * rng.uniform(500, 8e3), rng.uniform(200, 1e4)
        h = 1e-3
        du2 = (disp(x2 + h, x3, q2, q3, W) - disp(x2 - h, x3, q2, q3, W)) / (2 * h)
        du3 = (disp(x2, x3 + h, q2, q3, W) - disp(x2, x3 - h, q2, q3, W)) / (2 * h)
        s12, s13 = stress(x2, x3, q2, q3, W, 0.5, G)
        scale = max(abs(s12), abs(s13))
        assert abs(G * du2 - s12) <= 1e-5 * scale + 1e-12
        assert abs(G * du3 - s13) <= 1e-5 * scale + 1e-12


def test_free_surface_and_slip_jump():
    rng = np.random.default_rng(8)
    for _ in range(30):
        q2, q3, W = rng.uniform(-5e3, 5e3), rng.uniform(0, 5e3), rng.uniform(500, 5e3)
        # s13 vanishes on x3 = 0 (the image cancels it exactly)
        assert stress(rng.uniform(-2e4, 2e4), 0.0, q2, q3, W, 0.3)[1] == 0.0
        # across the segment the displacement jumps by the unit slip
        z = q3 + rng.uniform(0.05, 0.95) * W
        jump = disp(q2 - 1e-6, z, q2, q3, W) - disp(q2 + 1e-6, z, q2, q3, W)
        assert abs(abs(jump) - 1.0) < 1e-6
        # and not below it
        zb = q3 + W + rng.uniform(10, 1e3)
        assert abs(disp(q2 - 1e-6, zb, q2, q3, W) - disp(q2 + 1e-6, zb, q2, q3, W)) < 1e-6


def test_far_field_decay():
    rs = np.array([1e5, 2e5, 4e5, 8e5])
    vals = [abs(stress(r, 0.0, 0.0, 1e3, 2e3, 0.5)[0]) for r in rs]
    slope = np.polyfit(np.log(rs), np.log(vals), 1)[0]
    # a finite segment is a pair of screw dislocations: u ~ 1/r, stress ~ 1/r^2
    assert -2.1 < slope < -1.9


def test_odefun_one_patch_by_hand():
    """One patch, K = [k]: every entry of dydt from the formulas of
    ratestate.f90:655-722 written out by hand."""
    Vpl, mu0, sig, a, b, L, Vo, eta = 1e-9, 0.6, 100.0, 0.01, 0.014, 0.01, 1e-6, 5.0
    k = -2.5
    s = orc.APSystem([np.array([v]) for v in (Vpl, mu0, sig, a, b, L, Vo, eta)],
                     np.array([[k]]))
    y0 = s.init_state()
    ln10 = math.log(10.0)
    assert y0[0] == 0.0
    assert y0[1] == sig * (mu0 + (a - b) * math.log(Vpl / Vo)) - eta * Vpl
    assert y0[2] == math.log(L / Vpl) / ln10
    assert y0[3] == math.log(0.98 * Vpl) / ln10
    y = np.array([0.3, 60.0, 7.2, -8.5])
    dydt, vmax = s.odefun(y)
    V = math.exp(y[3] * ln10)
    dth = (math.exp(-y[2] * ln10) - V / L) / ln10
    dtau = k * (V - Vpl)
    dv = (dtau - b * sig * dth * ln10) / (a * sig + eta * V) / ln10
    np.testing.assert_allclose(dydt, [V, dtau - eta * V * dv * ln10, dth, dv], rtol=1e-14)
    assert vmax == V
