"""Learned RK45 step-size proposal (no counterpart in the reference).

An 8 -> 64 -> 64 -> 1 tanh perceptron (binary32, 4801 parameters; layout and
features in csrc/mlp_dev.h) proposes log10 of the next step of every
ensemble instance from its step history.  It is trained on the step logs of
the plain binary64 integrator (ratestate.f90:307-328 with the RKQSm of
ode45.f90:51-108) by the library's own forward/backward kernels and Adam.

Target of step k: the step that the accepted attempt of step k+1 shows to be
affordable at a fraction ERR_TARGET of the tolerance,
h_{k+1} (ERR_TARGET / errmax_{k+1})^(1/5), the error of an order-4 embedded
pair scaling as h^5; at most 10 h_{k+1}.  ERR_TARGET = 0.3 leaves room for
the network's prediction error (~0.08 in log10, a factor 1.2, i.e. 2.5 in
errmax).  The controller uses the proposal only where the classic rule's
error ratio exceeds ERRCON (ode45.f90:92-96); below it the classic x10
growth is already the largest allowed.

Trained on the classic controller's histories alone, the network meets
states of its own making that those histories never show (a step the classic
rule would not have taken); fit_closed_loop adds the histories of runs under
the network's own proposals, labelled the same way (DAgger, Ross, Gordon and
Bagnell 2011), and calibrate_offset picks the output offset on the training
instances: the furthest simulated time for a fixed attempt budget, or the
fewest attempts to a fixed window.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .engine import Engine, UcyError

N_IN = 8
N_PARAMS = 4801
ERR_TARGET = 0.3
_P = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))  # noqa: E731
_PD = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))  # noqa: E731


def classic_proposal(h, errmax, max_step=np.inf):
    """hnext of RKQSm after an accepted attempt (ode45.f90:92-99)."""
    hn = np.where(errmax > 1.89e-4, 0.7 * h * np.power(np.maximum(errmax, 1e-300), -0.2), 10.0 * h)
    return np.minimum(hn, max_step)


def features_from_history(dt, vmax, errmax, rejects):
    """Features (n-1, 8) of the proposals after steps 0..n-2 and their
    targets (n-1,), from one run's per-step logs (Engine.integrate with
    history=True).  Definitions as csrc/mlp_dev.h."""
    dt, vmax = np.asarray(dt, float), np.asarray(vmax, float)
    errmax, rejects = np.asarray(errmax, float), np.asarray(rejects, float)
    n = dt.size
    if n < 2:
        return np.zeros((0, N_IN), np.float32), np.zeros(0, np.float32)
    lv = np.log10(np.maximum(vmax, 1e-30))
    le = np.log10(np.maximum(errmax, 1e-10))
    lh = np.log10(dt)
    X = np.empty((n, N_IN))
    X[:, 0] = lv
    X[:, 1] = np.concatenate([[0.0], lv[1:] - lv[:-1]])
    X[:, 2] = lh
    X[:, 3] = le
    X[:, 4] = np.concatenate([[lh[0]], lh[:-1]])
    X[:, 5] = rejects
    X[:, 6] = np.log10(classic_proposal(dt, errmax))
    X[:, 7] = np.concatenate([[le[0]], le[:-1]])
    grow = np.minimum(0.2 * np.log10(ERR_TARGET / np.maximum(errmax[1:], 1e-10)), 1.0)
    T = lh[1:] + grow
    return X[:-1].astype(np.float32), T.astype(np.float32)


class StepNet:
    """The proposal network on one GPU context."""

    def __init__(self, engine: Engine, seed: int = 0):
        self.eng = engine
        self._L = engine._L
        self._h = self._L.ucy_mlp_create(engine._h, int(seed))
        if not self._h:
            raise UcyError(-1, "cannot create the step-size network")

    def close(self):
        if getattr(self, "_h", None):
            self._L.ucy_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, st):
        self.eng._chk(st)

    @property
    def params(self) -> np.ndarray:
        th = np.empty(N_PARAMS, np.float32)
        self._chk(self._L.ucy_mlp_get_params(self._h, _P(th)))
        return th

    @params.setter
    def params(self, th):
        th = np.ascontiguousarray(th, np.float32)
        assert th.size == N_PARAMS
        self._chk(self._L.ucy_mlp_set_params(self._h, _P(th)))

    def set_normalization(self, x_mean, x_scale, y_mean, y_scale):
        xm = np.ascontiguousarray(x_mean, np.float32)
        xs = np.ascontiguousarray(x_scale, np.float32)
        self.norm = (xm.copy(), xs.copy(), float(y_mean), float(y_scale))
        self._chk(self._L.ucy_mlp_set_normalization(self._h, _P(xm), _P(xs), float(y_mean),
                                                    float(y_scale)))

    def fit_normalization(self, X, T):
        X = np.asarray(X, np.float64)
        xs = X.std(axis=0)
        xs[xs < 1e-6] = 1.0
        ys = float(np.std(T)) or 1.0
        self.set_normalization(X.mean(axis=0), xs, float(np.mean(T)), ys)

    def forward(self, X) -> np.ndarray:
        X = np.ascontiguousarray(X, np.float32).reshape(-1, N_IN)
        y = np.empty(X.shape[0], np.float32)
        self._chk(self._L.ucy_mlp_forward(self._h, X.shape[0], _P(X), _P(y)))
        return y

    def loss_grad(self, X, T):
        X = np.ascontiguousarray(X, np.float32).reshape(-1, N_IN)
        T = np.ascontiguousarray(T, np.float32)
        g = np.empty(N_PARAMS, np.float32)
        loss = C.c_double()
        self._chk(self._L.ucy_mlp_loss_grad(self._h, X.shape[0], _P(X), _P(T), C.byref(loss),
                                            _P(g)))
        return loss.value, g

    def train(self, X, T, epochs=20, batch=4096, lr=2e-3, seed=0) -> np.ndarray:
        X = np.ascontiguousarray(X, np.float32).reshape(-1, N_IN)
        T = np.ascontiguousarray(T, np.float32)
        log = np.empty(epochs)
        self._chk(self._L.ucy_mlp_train(self._h, X.shape[0], _P(X), _P(T), int(epochs),
                                        int(batch), float(lr), int(seed), _PD(log)))
        return log

    def attach(self, ensemble):
        """Use this network for the step proposals of an ensemble."""
        self._chk(self._L.ucy_ensemble_set_mlp(ensemble._h, self._h))


def detach(ensemble):
    ensemble._chk(ensemble._L.ucy_ensemble_set_mlp(ensemble._h, None))


def training_data(engine: Engine, model, draws, instances, steps, y0s):
    """Step histories of the plain binary64 integrator for the given
    instances of one geometry group (same K, per-instance friction),
    converted to (features, targets)."""
    from . import configs
    engine.set_elastic(model.lambda_, model.mu)
    engine.set_rectangles(*model.geometry())
    Xs, Ts = [], []
    built = False
    for i in instances:
        engine.set_friction(*configs.instance_friction(model, draws, int(i)))
        if not built:
            engine.buildG()
            built = True
        r = engine.integrate(y0s[i], 1e30, max_steps=steps, log_cap=steps, history=True)
        X, T = features_from_history(r.dt_log, r.vmax_log, r.errmax_log, r.reject_log)
        Xs.append(X)
        Ts.append(T)
    return np.concatenate(Xs), np.concatenate(Ts)


def features_from_logs(logs):
    """Features and targets of several instances' step logs (dicts with
    dt, vmax, errmax, rejects, as Ensemble.step_log returns)."""
    Xs, Ts = [], []
    for r in logs:
        X, T = features_from_history(r["dt"], r["vmax"], r["errmax"], r["rejects"])
        Xs.append(X)
        Ts.append(T)
    return np.concatenate(Xs), np.concatenate(Ts)


def ensemble_histories(ens, steps, net=None, groups=None, progress=None):
    """Step histories of the instances of an ensemble's groups (all by
    default) over `steps` accepted steps per instance from their initial
    states: under the network's proposals when `net` is given, else under the
    classic rule."""
    groups = list(range(len(ens.groups))) if groups is None else list(groups)
    for g in groups:
        ens.set_step_log(steps, g)
    ens.reset()
    if net is not None:
        net.attach(ens)
    else:
        detach(ens)
    ens.set_stop_at_end(False)
    ens.run(1e30, max_attempts=int(2 * steps), chunk_attempts=2000, progress=progress)
    Xs, Ts = [], []
    for g in groups:
        X, T = features_from_logs(ens.step_log(g))
        Xs.append(X)
        Ts.append(T)
        ens.set_step_log(0, g)
    return np.concatenate(Xs), np.concatenate(Ts)


def closed_loop_data(net, ens, steps, group=None, progress=None):
    """Histories with the network's proposals (one group, or all)."""
    return ensemble_histories(ens, steps, net, None if group is None else [group], progress)


def fit_closed_loop(net, X, T, ens, steps, rounds=2, epochs=40, batch=4096, lr=1e-3,
                    seed=1, group=None, progress=None):
    """Rounds of: run the training instances (ensemble `ens`) under the
    current network, add their histories to (X, T), train on the union.
    Returns the grown data set."""
    for r in range(rounds):
        Xr, Tr = closed_loop_data(net, ens, steps, group, progress)
        X, T = np.concatenate([X, Xr]), np.concatenate([T, Tr])
        net.train(X, T, epochs=epochs, batch=batch, lr=lr, seed=seed + r + 1)
    return X, T


def calibrate_offset(net, ens, attempts=None, window=None,
                     offsets=(-0.1, -0.05, 0.0, 0.05, 0.1, 0.15, 0.2, 0.3), progress=None):
    """Shift of the network's output (log10 h) picked on the instances of
    `ens`, with one of two objectives:

    attempts=R   the furthest summed simulated time within R RK45 attempts
                 per instance from the initial states (every candidate costs
                 the same fixed budget);
    window=W     the fewest attempts for every instance to reach t = W (last
                 step clamped to land on it).  Offsets are tried from 0
                 outwards and a direction stops at the first worse one; every
                 run is capped at twice the per-instance attempts of offset
                 0, and a capped run is not eligible.

    An offset that fails an instance is not eligible.  Leaves the chosen
    shift in the network's normalisation and returns (offset, {offset:
    score}) with score = simulated seconds (attempts=R) or attempts
    (window=W)."""
    assert (attempts is None) != (window is None), "one objective"
    xm, xs, ym, ys = net.norm

    def run(d, cap):
        net.set_normalization(xm, xs, ym + d, ys)
        ens.reset()
        net.attach(ens)
        if attempts is not None:
            ens.set_stop_at_end(False)
            st = ens.run(1e30, max_attempts=int(attempts), chunk_attempts=max(1, int(attempts) // 2),
                         progress=progress)
            return (float(st.inst_time_sum) if st.n_failed == 0 else -float("inf")), st
        ens.set_stop_at_end(True)
        st = ens.run(window, max_attempts=int(cap), chunk_attempts=5000, progress=progress)
        ok = st.n_failed == 0 and st.min_time >= window
        return (-float(st.attempts) if ok else -float("inf")), st

    res = {}
    if attempts is not None:
        for d in offsets:
            res[d] = run(d, None)[0]
    else:
        base, st0 = run(0.0, 10**7)
        res[0.0] = base
        most = max(int(np.max(ens.state(g)[4])) for g in range(len(ens.groups)))
        cap = 2 * max(1, most)
        for side in (sorted(d for d in offsets if d > 0), sorted((d for d in offsets if d < 0),
                                                                 reverse=True)):
            prev = base
            for d in side:
                res[d] = run(d, cap)[0]
                if res[d] < prev:
                    break
                prev = res[d]
    best = max(res, key=res.get)
    net.set_normalization(xm, xs, ym + best, ys)
    score = {d: (v if attempts is not None else -v) for d, v in res.items()}
    return best, score
