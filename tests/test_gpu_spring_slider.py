"""The whole single-instance path (odefun pre/post-pass, K v, Cash-Karp
RK45 with RKQSm, the device-resident main loop) on the 0-D spring-slider of
fortran/0d: one patch whose K is the self-stiffness k = 7 pi / 16 mu / R
(fortran/0d/src/ratestate.f90:261) in both slip directions, aging law and
radiation damping, through its first seismic event.

The reference solution is the same physics written as a scalar ODE in
(slip, shear traction, log10 theta, log10 V), the variables of
ratestate.f90:624-778 -- dtheta*/dt = (10^-theta* - V/L)/ln10,
dv*/dt = (-k (V - Vpl) - b sigma ln10 dtheta*/dt) / ((a sigma + eta V) ln10),
dtau/dt = -k (V - Vpl) - eta V ln10 dv*/dt -- integrated by scipy's DOP853
at rtol 1e-12: an independent integrator at a tolerance six orders tighter
than the engine's (epsilon 1e-6)."""
import numpy as np
import pytest
from scipy.integrate import solve_ivp

from unicycle_amd import configs

pytestmark = pytest.mark.gpu

MU, R, A, B, L, SIG, VPL, VO, MU0 = 30e3, 2000.0, 0.01, 0.014, 0.01, 100.0, 1e-9, 1e-6, 0.6
ETA = MU / (2.0 * 3000.0)
K = 7.0 * np.pi / 16.0 * MU / R
LN10 = np.log(10.0)
RAKE = np.deg2rad(89.0)


def _scalar(t, y):
    s, tau, th, v = y
    V = 10.0 ** v
    dth = (10.0 ** (-th) - V / L) / LN10
    dk = -K * (V - VPL)
    dv = (dk - B * SIG * LN10 * dth) / (A * SIG + ETA * V) / LN10
    return [V, dk - ETA * V * LN10 * dv, dth, dv]


def test_spring_slider_tracks_reference_solution(engine_factory):
    tau0 = SIG * (MU0 + (A - B) * np.log(VPL / VO)) - ETA * VPL
    th0, v0 = np.log10(L / VPL), np.log10(0.5 * VPL)
    T = 10 * configs.YEAR
    ev = lambda t, y: y[3] - np.log10(1e-3)  # noqa: E731
    ev.direction = 1
    ref = solve_ivp(_scalar, [0.0, 1.2 * T], [0.0, tau0, th0, v0], method="DOP853", rtol=1e-12,
                    atol=1e-14, events=ev, dense_output=True)
    assert ref.status == 0 and ref.t_events[0].size >= 1
    t_event = ref.t_events[0][0]
    assert 0.2 * T < t_event < 0.8 * T                       # the window holds one event

    eng = engine_factory()
    eng.set_elastic(MU, MU)
    eng.set_rectangles(np.zeros((1, 3)), [2 * R], [2 * R], [0.0], [np.pi / 2])
    one = np.ones(1)
    eng.set_friction(VPL * one, RAKE * one, MU0 * one, SIG * one, A * one, B * one, L * one,
                     VO * one, ETA * one)
    eng.set_kernel(np.array([[-K, 0.0], [0.0, -K], [0.0, 0.0]]))
    y0 = np.array([0.0, 0.0, tau0 * np.cos(RAKE), tau0 * np.sin(RAKE), 0.0, th0, v0])
    cap = 200000
    res = eng.integrate(y0, T, max_steps=cap, log_cap=cap)
    assert T <= res.time < 1.5 * T and res.steps < cap
    # first step starting above 1e-3 m/s (the step-start vMax is logged per step)
    k = int(np.argmax(res.vmax_log > 1e-3))
    assert res.vmax_log[k] > 1e-3
    t_gpu = res.t_log[k] - res.dt_log[k]
    rel_event = abs(t_gpu - t_event) / t_event
    y = res.y
    slip = np.hypot(y[0], y[1])
    want = ref.sol(res.time)
    rel_slip = abs(slip - want[0]) / want[0]
    rel_tau = abs(np.hypot(y[2], y[3]) - want[1]) / abs(want[1])
    print(f"event {t_gpu / configs.YEAR:.6f} yr vs {t_event / configs.YEAR:.6f} yr "
          f"(rel {rel_event:.2e}); slip {slip:.6f} vs {want[0]:.6f} m (rel {rel_slip:.2e}); "
          f"traction rel {rel_tau:.2e}; {res.steps} steps, {res.attempts} attempts")
    # measured: event 3.8e-8, slip 9.2e-7, traction 1.3e-9 (epsilon 1e-6)
    assert rel_event < 1e-6
    assert rel_slip < 2e-5
    assert rel_tau < 1e-7
    # the slip stays along the rake, as the isotropic stiffness requires
    assert abs(np.arctan2(y[1], y[0]) - RAKE) < 1e-9
