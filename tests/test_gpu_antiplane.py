"""The 2-D antiplane model (fortran/2d/antiplane) on the GPU against the
oracle: K (buildG), the traction hook on the reference's fixture, the
initial state, odefun, and device-resident runs through the first
earthquake of a vertical strike-slip fault; the reduced-precision K
tracks binary64."""
import os

import numpy as np
import pytest

import orc
from unicycle_amd import PREC_F64, PREC_MIXED, PREC_SPLIT_BF16, configs

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "antiplane.npz")


def _engine(engine_factory, m, precision=PREC_F64):
    eng = engine_factory()
    eng.set_antiplane(m.mu, m.x, m.width, m.dip)
    eng.set_friction(*m.friction())
    eng.buildG(precision)
    return eng


@pytest.mark.parametrize("n", [300, 301])
def test_buildG_and_hook_match_oracle_and_reference(engine_factory, n):
    """An odd segment count pads K's rows to an even stride inside the
    library; the caller sees n x n."""
    m = configs.antiplane_fault(n=n)
    eng = _engine(engine_factory, m)
    K = eng.get_kernel()
    Kref = orc.buildG_antiplane(m.x, m.width, m.dip, m.mu, nthreads=8)
    assert K.shape == (m.n, m.n)
    np.testing.assert_array_equal(K, Kref)
    v = np.random.default_rng(1).standard_normal(m.n)
    np.testing.assert_allclose(eng.matvec(v), Kref @ v, rtol=1e-12,
                               atol=1e-12 * np.abs(Kref).sum(axis=1).max())
    g = np.load(GOLD)
    for rcv, want in zip(g["receivers"], g["traction"]):
        ts = eng.computeTractionKernelsAntiplane(rcv[0:3], rcv[3:6], rcv[6:9], rcv[9:12],
                                                 g["W"].size, g["y"], g["W"], g["dip"], 30e3)
        np.testing.assert_array_equal(ts, want)


def test_state_odefun_and_run_through_first_event(engine_factory):
    m = configs.antiplane_fault(n=200)
    eng = _engine(engine_factory, m)
    sysm = orc.APSystem(m.oracle_friction(), eng.get_kernel())
    y0 = sysm.init_state()
    np.testing.assert_array_equal(eng.initStateVector(), y0)
    rng = np.random.default_rng(2)
    y = y0.copy()
    y[3::4] += 0.5 * rng.standard_normal(m.n)
    dydt = np.empty_like(y)
    vmax = eng.odefun(y.size, 0.0, y, dydt)
    want, vmax_ref = sysm.odefun(y)
    np.testing.assert_allclose(dydt, want, rtol=1e-12, atol=1e-12 * np.abs(want).max())
    assert vmax == vmax_ref
    # into the first event (vMax above 1e-3 m/s; the rupture then takes some
    # 3000 sub-second steps): same accept / reject sequence, step sizes and
    # onset time over the nucleation and the first 100 steps of the rupture
    _, _, _, _, vl, _ = sysm.integrate(y0, 300 * configs.YEAR, max_steps=1000, log_cap=1000)
    i1 = int(np.argmax(vl > 1e-3))
    assert vl[i1] > 1e-3
    k = i1 + 100
    yo, st, tl, dl, vlk, rc = sysm.integrate(y0, 300 * configs.YEAR, max_steps=k, log_cap=k)
    res = eng.integrate(y0, 300 * configs.YEAR, max_steps=k, log_cap=k)
    print(f"antiplane: event onset at step {i1}, attempts {res.attempts} / {st.attempts}, "
          f"vmax {vlk.max():.3g}", flush=True)
    assert vlk.max() > 1e-2
    assert res.steps == st.steps == k and res.attempts == st.attempts
    np.testing.assert_allclose(res.dt_log, dl, rtol=1e-5)
    assert int(np.argmax(res.vmax_log > 1e-3)) == i1
    assert res.t_log[i1] == pytest.approx(tl[i1], rel=1e-5)
    assert np.abs(res.y - yo).max() <= 1e-5 * np.abs(yo).max()


@pytest.mark.parametrize("prec,n", [(PREC_MIXED, 256), (PREC_MIXED, 255),
                                    (PREC_SPLIT_BF16, 256), (PREC_SPLIT_BF16, 255)])
def test_reduced_precision_tracks_f64(engine_factory, prec, n):
    """Mixed (binary64 step-start traction, split increments) stays at 1e-8;
    split bf16 carries its ~17-bit K in the step-start traction too, so the
    slip drifts by ~1e-5 per decade on this fault.  Odd and even segment
    counts (the padded and the unpadded K) behave alike."""
    m = configs.antiplane_fault(n=n)
    ref = _engine(engine_factory, m)
    red = _engine(engine_factory, m, prec)
    y0 = ref.initStateVector()
    # well before the first event (about 28 years in: nucleation amplifies
    # any difference), at a bounded step so that both runs end together
    T = 10 * configs.YEAR
    a = ref.integrate(y0, T, max_steps=200000, max_step=T / 4000)
    b = red.integrate(y0, T, max_steps=200000, max_step=T / 4000)
    assert a.time >= T and b.time >= T
    slip_a, slip_b = a.y[0::4], b.y[0::4]
    err = np.linalg.norm(slip_a - slip_b) / np.linalg.norm(slip_a)
    print(f"prec {prec}: slip rel err {err:.2e}, attempts {a.attempts} / {b.attempts}, "
          f"end times {a.time:.6e} / {b.time:.6e}", flush=True)
    assert b.attempts <= 1.2 * a.attempts + 20
    assert err < (1e-6 if prec == PREC_MIXED else 1e-4)


def test_buildO_and_displacement_match_oracle(engine_factory):
    """buildO of the 2-D model (greens.f90:275-383): one displacement
    component per point; d = O u with u the segments' slip, df = d, dl = 0."""
    m = configs.antiplane_fault(n=151)
    eng = _engine(engine_factory, m)
    rng = np.random.default_rng(4)
    pts = np.stack([np.zeros(30), rng.uniform(-5e4, 5e4, 30),
                    np.where(np.arange(30) % 2 == 0, 0.0, rng.uniform(0, 3e4, 30))], 1)
    eng.buildO(pts)
    O, Of, Ol = eng.get_displacement_kernels()
    Oref = orc.buildO_antiplane(m.x, m.width, m.dip, pts)
    assert O.shape == Oref.shape == (30, m.n)
    # four atan terms cancel to the far field: rounding-level differences of
    # the device's atan, relative to each row's size
    err = np.abs(O - Oref).max(axis=1) / np.abs(Oref).max(axis=1)
    assert err.max() < 1e-11, err.max()
    np.testing.assert_array_equal(Of, O)
    assert not Ol.any()
    y = eng.initStateVector()
    y[0::4] = rng.standard_normal(m.n)
    d, df, dl = eng.displacement(y)
    np.testing.assert_allclose(d, O @ y[0::4], rtol=1e-12, atol=1e-12 * np.abs(O).sum(1).max())
    np.testing.assert_array_equal(df, d)
    assert not dl.any()
