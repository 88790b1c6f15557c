"""Single-instance reduced-precision K: the split-bf16 pairs (4 bytes per
entry, power-of-two row scales) that ucy_integrate streams when K is built
with PREC_SPLIT_BF16 or PREC_MIXED, against the binary64 path on the same
fault + Maxwell-volume model (the viscoelastic model's layout at 8 x 4
patches and 4 x 4 x 2 cuboids).

- matvec: |K~ v - K v| <= 2^-16 sum_j |K_ij| |v_j| per row (hi and lo
  each round to 8 significant bits, so an entry carries at most 2^-16 of
  relative error; accumulation is binary64);
- integration over a fixed window (steps capped so that the final step
  overshoots little): the attempt count stays within 1.2x of binary64 (no
  controller stall from an error floor) and the slip at the window's end,
  corrected to the same time, stays within the stated bound."""
import numpy as np
import pytest

from unicycle_amd import PREC_F64, PREC_MIXED, PREC_SPLIT_BF16, configs

pytestmark = pytest.mark.gpu


def _model(npower=1.0):
    fault = configs.thrust_fault(nx=8, nz=4, ellipse=None)
    vol = configs.volume_grid(4, 4, 2, size=4000.0, top=20e3, origin=(0.0, 0.0), npower=npower)
    return fault, vol


def _engine(engine_factory, fault, vol, prec):
    eng = engine_factory()
    eng.set_elastic(fault.lambda_, fault.mu)
    eng.set_rectangles(*fault.geometry())
    eng.set_volumes(*vol.geometry())
    eng.set_friction(*fault.friction())
    eng.set_rheology(*vol.rheology())
    eng.buildG(prec)
    return eng


def test_matvec_reduced_error_bound(engine_factory):
    fault, vol = _model()
    eng = _engine(engine_factory, fault, vol, PREC_SPLIT_BF16)
    K = eng.get_kernel()
    rng = np.random.default_rng(0)
    n = K.shape[1]
    for v in (rng.standard_normal(n), 1e-9 * (1.0 + 1e-3 * rng.standard_normal(n)),
              np.where(rng.random(n) < 0.1, 1.0, 0.0)):
        t = eng.matvec_reduced(v)
        ref = K @ v
        bound = 2.0 ** -16 * (np.abs(K) @ np.abs(v)) * 1.0001
        assert np.all(np.abs(t - ref) <= bound + 1e-15 * np.abs(ref).max())
        assert np.abs(t - ref).max() > 0.0          # the stored K is the reduced one


def _slip_at(eng, res, T, n):
    """Slip of the fault patches at time T from a run that ended at
    res.time >= T: one step back along the final slip rates."""
    y = res.y
    dydt = np.empty_like(y)
    eng.odefun(y.size, 0.0, y, dydt)
    dt = res.time - T
    ss = y[0:7 * n:7] - dt * dydt[0:7 * n:7]
    sd = y[1:7 * n:7] - dt * dydt[1:7 * n:7]
    return np.hypot(ss, sd)


@pytest.mark.parametrize("prec,npower", [(PREC_SPLIT_BF16, 1.0), (PREC_MIXED, 1.0),
                                         (PREC_SPLIT_BF16, 3.0), (PREC_MIXED, 3.0)])
def test_reduced_integrate_tracks_f64(engine_factory, prec, npower):
    """Maxwell (n = 1) and power-law (n = 3) volumes: the reduced K against
    binary64 on the same run."""
    fault, vol = _model(npower)
    T = 5 * configs.YEAR
    hmax = T / 4000
    out = {}
    for p in (PREC_F64, prec):
        eng = _engine(engine_factory, fault, vol, p)
        y0 = eng.initStateVector(1)
        rng = np.random.default_rng(4)
        y0[6:7 * fault.n:7] += 0.3 * rng.standard_normal(fault.n)
        res = eng.integrate(y0, T, max_steps=200000, max_step=hmax)
        assert res.time >= T
        out[p] = (res, _slip_at(eng, res, T, fault.n))
    (r64, s64), (rr, sr) = out[PREC_F64], out[prec]
    err = np.linalg.norm(sr - s64) / np.linalg.norm(s64)
    print(f"prec {prec} n {npower}: attempts {rr.attempts} vs {r64.attempts}, "
          f"slip error {err:.3e}")
    assert rr.attempts <= 1.2 * r64.attempts + 20
    assert err < (1e-6 if prec == PREC_MIXED else 1e-5)
