"""The step-size network on the GPU: forward pass and gradient of the
training loss against the binary64 numpy oracle (oracle/mlp_oracle.py), and
training by the hand-written kernels.

Tolerances: the device computes in binary32 (tanhf, fmaf) and sums the
per-sample gradient terms in blocks of 64; against binary64 the relative
error of the loss and the normwise error of the gradient are <= 1e-4, the
outputs within 1e-4 of their scale."""
import os
import sys

import numpy as np
import pytest

from unicycle_amd import stepnet

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                "oracle"))
import mlp_oracle as M  # noqa: E402

pytestmark = pytest.mark.gpu


def _case(seed=0, n=3000):
    rng = np.random.default_rng(seed)
    th = (rng.normal(0, 0.25, M.NP)).astype(np.float32)
    X = rng.normal(0, 1, (n, 8)).astype(np.float32) * 3 + 1
    T = rng.normal(0, 1, n).astype(np.float32) * 2 - 5
    norm = (np.full(8, 1.0, np.float32), np.full(8, 3.0, np.float32), -5.0, 2.0)
    return th, X, T, norm


def test_forward_and_gradient_match_oracle(engine_factory):
    eng = engine_factory()
    net = stepnet.StepNet(eng, seed=1)
    th, X, T, norm = _case()
    net.params = th
    net.set_normalization(*norm)
    np.testing.assert_array_equal(net.params, th)
    y = net.forward(X)
    yo = M.predict(th.astype(np.float64), X, norm)
    assert np.abs(y - yo).max() <= 1e-4 * np.abs(yo).max()
    loss, g = net.loss_grad(X, T)
    lo, go = M.loss_grad(th.astype(np.float64), X, T, norm)
    assert abs(loss - lo) <= 1e-4 * lo
    assert np.linalg.norm(g - go) <= 1e-4 * np.linalg.norm(go)
    # every parameter block carries gradient (a dropped term shows here)
    for a, b in ((0, M.OFF_B1), (M.OFF_B1, M.OFF_W2), (M.OFF_W2, M.OFF_B2), (M.OFF_B2, M.OFF_W3),
                 (M.OFF_W3, M.OFF_B3), (M.OFF_B3, M.NP)):
        assert np.linalg.norm(g[a:b] - go[a:b]) <= 1e-3 * np.linalg.norm(go[a:b]) + 1e-12
    net.close()


def test_training_fits_a_smooth_map(engine_factory):
    eng = engine_factory()
    net = stepnet.StepNet(eng, seed=2)
    rng = np.random.default_rng(3)
    X = rng.uniform(-2, 2, (20000, 8)).astype(np.float32)
    T = (np.sin(X[:, 0]) + 0.5 * X[:, 2] * X[:, 3] - 0.3 * X[:, 6]).astype(np.float32)
    net.fit_normalization(X, T)
    l0, _ = net.loss_grad(X, T)
    log = net.train(X, T, epochs=25, batch=2048, lr=3e-3, seed=4)
    l1, _ = net.loss_grad(X, T)
    assert np.all(np.isfinite(log))
    assert l1 < 0.05 * l0, (l0, l1, log[-5:])
    # the trained parameters reproduce the oracle's prediction
    yo = M.predict(net.params.astype(np.float64), X[:500], net.norm)
    assert np.abs(net.forward(X[:500]) - yo).max() <= 1e-4 * np.abs(yo).max()
    net.close()


def _instances(engine_factory, nx, nz, seed):
    from unicycle_amd import configs
    m = configs.thrust_fault(nx=nx, nz=nz, ellipse=None)
    d = configs.ensemble_instances(m, 64, seed=seed)
    ic = engine_factory()
    ic.set_elastic(m.lambda_, m.mu)
    ic.set_rectangles(*m.geometry())
    y0 = []
    for i in range(64):
        ic.set_friction(*configs.instance_friction(m, d, i))
        y0.append(configs.perturbed_initial_state(m, ic.initStateVector(), int(d["seeds"][i])))
    return m, d, ic, np.array(y0)


def _group(eng, m, d, y0, idx, prec):
    from unicycle_amd.ensemble import Ensemble
    ens = Ensemble(eng, prec)
    ens.add_group(m, d["a"][idx], d["b"][idx], d["sig"][idx], d["L"][idx], y0[idx])
    ens.build()
    return ens


def test_ensemble_step_log(engine_factory):
    """The ensemble's per-step log: times are the running sums of the
    accepted steps, rejections add up to attempts - steps, the first steps
    of a binary64 instance follow the plain integrator's log (same RKQSm
    arithmetic, ode45.f90:51-108), and the log's features match the plain
    log's feature definitions."""
    from unicycle_amd import PREC_F64, configs
    m, d, ic, y0 = _instances(engine_factory, 8, 8, 3)
    eng = engine_factory()
    ens = _group(eng, m, d, y0, list(range(4)), PREC_F64)
    ens.set_step_log(300, 0)
    ens.reset()
    ens.run(1e30, max_attempts=400)
    logs = ens.step_log(0)
    y, t, dt, steps, att = ens.state(0)
    ic.set_friction(*configs.instance_friction(m, d, 0))
    ic.buildG()
    r = ic.integrate(y0[0], 1e30, max_steps=40, log_cap=40, history=True)
    for i, lg in enumerate(logs):
        n = len(lg["dt"])
        assert n == min(steps[i], 300) and n > 50
        assert np.allclose(np.cumsum(lg["dt"]), lg["t"], rtol=1e-12, atol=0)
        if steps[i] <= 300:
            assert lg["rejects"].sum() == att[i] - steps[i]
    # the error norm is a difference of stage rates that are themselves small
    # differences (K (v - Vpl) near steady state), so the GEMM's and the
    # GEMV's summation orders show in it at ~1e-5 after ten steps
    k = 20
    assert np.array_equal(logs[0]["dt"][:9], r.dt_log[:9])     # the x10 growth phase
    assert np.allclose(logs[0]["dt"][:k], r.dt_log[:k], rtol=1e-4)
    assert np.allclose(logs[0]["vmax"][:k], r.vmax_log[:k], rtol=1e-6)
    assert np.allclose(logs[0]["errmax"][:k], r.errmax_log[:k], rtol=1e-3)
    ens.close()


def test_ensemble_step_log_per_group_capacity(engine_factory):
    """Logs of two groups read after the first group's log is switched off:
    each group's arrays follow its own capacity."""
    from unicycle_amd import PREC_SPLIT_BF16
    from unicycle_amd.ensemble import Ensemble
    eng = engine_factory()
    ens = Ensemble(eng, PREC_SPLIT_BF16)
    for nx, nz, seed in ((8, 8, 3), (8, 4, 4)):
        m, d, ic, y0 = _instances(engine_factory, nx, nz, seed)
        ens.add_group(m, d["a"][:8], d["b"][:8], d["sig"][:8], d["L"][:8], y0[:8])
    ens.build()
    ens.set_step_log(200, 0)
    ens.set_step_log(120, 1)
    ens.reset()
    ens.run(1e30, max_attempts=300)
    l0 = ens.step_log(0)
    ens.set_step_log(0, 0)
    l1 = ens.step_log(1)
    assert all(len(r["dt"]) == 200 for r in l0)
    assert all(len(r["dt"]) == 120 for r in l1)
    ens.close()


def test_learned_proposal_gain_on_held_out_instances(engine_factory):
    """Train on the plain binary64 integrator's step logs of 16 seeded
    instances of a 256-patch fault, refine on the histories of the same
    instances run under the network (two closed-loop rounds), pick the
    output offset on those instances (fewest attempts to a 0.15-year
    window), then run 48 held-out
    instances as an ensemble over a fixed simulated window (every instance
    clamped to land on it) with the classic RKQSm growth rule, with the
    network trained on the plain logs only, and with the refined network.

    Measured gain (what this test records): fewer RK45 attempts (RHS
    evaluations) for the same simulated window; first-event times of the
    instances with an event agree with the classic run to < 0.1 %."""
    import time
    from unicycle_amd import PREC_F64, configs
    t_start = time.perf_counter()
    lap = lambda what: print(f"[{time.perf_counter() - t_start:6.1f} s] {what}", flush=True)  # noqa: E731
    m, d, ic, y0 = _instances(engine_factory, 16, 16, 7)
    eng = engine_factory()
    X, T = stepnet.training_data(ic, m, d, range(16), 3000, y0)
    lap("plain histories")
    net = stepnet.StepNet(eng, seed=0)
    net.fit_normalization(X, T)
    log = net.train(X, T, epochs=40, batch=4096, lr=2e-3, seed=1)
    assert log[-1] < 0.05 * log[0]
    rmse = float(np.sqrt(np.mean((net.forward(X) - T) ** 2)))
    rmse_classic = float(np.sqrt(np.mean((X[:, 6] - T) ** 2)))
    held = list(range(16, 64))
    W = 0.5 * configs.YEAR

    out = {}
    for name in ("classic", "open_loop", "closed_loop"):
        if name == "closed_loop":
            train = _group(eng, m, d, y0, list(range(16)), PREC_F64)
            X, T = stepnet.fit_closed_loop(net, X, T, train, 3000, rounds=2)
            lap("closed loop")
            off, tried = stepnet.calibrate_offset(net, train, window=0.15 * configs.YEAR)
            lap("offset")
            train.close()
        ens = _group(eng, m, d, y0, held, PREC_F64)
        if name != "classic":
            net.attach(ens)
        ens.set_stop_at_end(True)
        st = ens.run(W, max_attempts=10**6, chunk_attempts=5000)
        assert st.n_failed == 0, f"{name}: status {ens.status(0)}"
        assert st.min_time >= W
        y, t, dt, s, a = ens.state(0)
        out[name] = (a.sum(), s.sum(), ens.first_events(0))
        ens.close()
        lap(name)
    a0, s0, e0 = out["classic"]
    for name in ("open_loop", "closed_loop"):
        a1, s1, e1 = out[name]
        both = (e0 > 0) & (e1 > 0)
        rel = np.abs(e1[both] - e0[both]) / e0[both]
        print(f"{name}: attempts {a1} vs {a0} (gain {a0 / a1:.3f}); rejected "
              f"{1 - s1 / a1:.3f} vs {1 - s0 / a0:.3f}; events {both.sum()}, max rel time "
              f"diff {rel.max():.2e}")
        assert ((e0 > 0) ^ (e1 > 0)).sum() == 0 and rel.max() < 1e-3
    print(f"rmse {rmse:.3f} (classic rule {rmse_classic:.3f}) in log10 h; offset {off} "
          f"from {tried}")
    assert rmse < rmse_classic
    assert out["closed_loop"][0] < a0     # fewer RHS evaluations for the same window
    net.close()
