"""Host-side mirror of the reference's fortran/3d call surface over libucy.

Names and argument meaning follow jdpmoore/unicycle fortran/3d/src:
``buildG`` (greens.f90:73), ``odefun`` (ratestate.f90:624), ``RKCKm`` /
``RKQSm`` (ode45.f90:116 / 51), ``initStateVector`` (ratestate.f90:548), the
Green's-function routines of okada92.f90 / nikkhoo15.f90 / strainvolume.f90,
and the main loop of ratestate.f90:307-328 (``integrate``).  Errors that the
reference turns into ``STOP`` raise :class:`UcyError` here.

Every method calls the HIP library; there is no CPU path.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib

PREC_F64 = 0
PREC_SPLIT_BF16 = 1
PREC_MIXED = 2        # ensembles: binary64 step-start matvec, split-bf16 increments
PREC_BF16_I8 = 3      # ensembles: split bf16 with K in 3 bytes (bf16 hi + 8-bit lo code)


class UcyError(RuntimeError):
    """A nonzero status from libucy (the reference would STOP)."""

    def __init__(self, status: int, message: str):
        super().__init__(f"[{status}] {message}")
        self.status = status


def _p(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _pi(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def _f64(a, shape=None) -> np.ndarray:
    out = np.ascontiguousarray(np.asarray(a, dtype=np.float64))
    if shape is not None:
        out = out.reshape(shape)
    return out


def _vec3(a) -> np.ndarray:
    return _f64(a).reshape(3)


@dataclass
class IntegrateResult:
    y: np.ndarray
    time: float
    dt_next: float
    steps: int
    attempts: int
    rhs_evals: int
    vmax: float
    wall_seconds: float
    t_log: np.ndarray | None = None
    dt_log: np.ndarray | None = None
    vmax_log: np.ndarray | None = None
    errmax_log: np.ndarray | None = None    # errmax/epsilon of each accepted attempt
    reject_log: np.ndarray | None = None    # rejected attempts before each accepted one
    obs_y: np.ndarray | None = None         # (steps, n_obs) step-start values of observed entries
    obs_dydt: np.ndarray | None = None      # (steps, n_obs) their step-start derivatives
    disp: np.ndarray | None = None          # (steps, 3 npts) observation-point displacement
    disp_f: np.ndarray | None = None        # its fault part (Of u) and ...
    disp_l: np.ndarray | None = None        # ... volume part (Ol u), at each step start
    emax_log: np.ndarray | None = None      # step-start eMax (largest volume strain-rate norm)


class Engine:
    """One GPU context (one MPI rank of the reference)."""

    def __init__(self, device: int = 0):
        self._L = _lib.lib()
        self.device = int(device)
        self._h = self._L.ucy_create(device, 0)
        if not self._h:
            raise UcyError(-2, f"cannot create a context on device {device}")
        self.np = 0
        self.nvol = 0
        self.antiplane = False
        self.planestrain = False
        self._ensembles = []

    # -- plumbing ---------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None):
            for e in getattr(self, "_ensembles", []):
                e.close()
            self._L.ucy_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _chk(self, status: int):
        if status != 0:
            msg = self._L.ucy_last_error(self._h)
            raise UcyError(status, msg.decode() if msg else "")

    # -- model definition -------------------------------------------------
    def set_elastic(self, lambda_: float, mu: float):
        self.lambda_, self.mu = float(lambda_), float(mu)
        self._chk(self._L.ucy_set_elastic(self._h, self.lambda_, self.mu))

    def set_antiplane(self, mu, x, width, dip):
        """The 2-D antiplane model (fortran/2d/antiplane): vertical segments
        with upper corners x (ns, 3; x2, x3 used), width, dip (radians),
        shear modulus mu.  State: 4 entries per patch."""
        x = _f64(x).reshape(-1, 3)
        self._ap = [x, _f64(width), _f64(dip)]
        self.mu = float(mu)
        self._chk(self._L.ucy_set_antiplane(self._h, self.mu, x.shape[0],
                                            *[_p(a) for a in self._ap]))
        self.antiplane = True
        self.planestrain = False
        self.nrect, self.ntri, self.nvol = x.shape[0], 0, 0
        self._update_np()

    def set_planestrain(self, lambda_, mu, x, width, dip):
        """The 2-D plane-strain model (fortran/2d/planestrain): segments with
        upper corners x (ns, 3; x2, x3 used), width, dip (radians), slipping
        along dip in a half space of Lame parameters lambda_, mu.  State: 5
        entries per patch (dip slip, dip and normal traction, log10 theta,
        log10 V)."""
        x = _f64(x).reshape(-1, 3)
        self._ps = [x, _f64(width), _f64(dip)]
        self.lambda_, self.mu = float(lambda_), float(mu)
        self._chk(self._L.ucy_set_planestrain(self._h, self.lambda_, self.mu, x.shape[0],
                                              *[_p(a) for a in self._ps]))
        self.antiplane, self.planestrain = False, True
        self.nrect, self.ntri, self.nvol = x.shape[0], 0, 0
        self._update_np()

    def set_initial_traction(self, tau0):
        """The plane-strain deck's tau0: a positive entry replaces the
        patch's initial dip traction (ratestate.f90 of 2d/planestrain:604-609)."""
        self._tau0 = _f64(tau0)
        self._chk(self._L.ucy_set_initial_traction(self._h, self._tau0.size, _p(self._tau0)))

    def set_planestrain_volumes(self, x, thickness, width, dip):
        """Strain volumes of the plane-strain viscouscycles program
        (fortran/2d/planestrain/src/viscouscycles.f90): tops x (nv, 3; x2, x3
        used), thickness, width, dip (radians); after set_planestrain.
        State: 6 entries per volume after the patches (e22, e23, e33, s22,
        s23, s33); set_rheology takes the background e22, e23, e33 in
        entries 3, 4, 5 of each edot_bg row (Gm unused)."""
        x = _f64(x).reshape(-1, 3)
        self._psv = [x, _f64(thickness), _f64(width), _f64(dip)]
        self._chk(self._L.ucy_set_planestrain_volumes(self._h, x.shape[0],
                                                      *[_p(a) for a in self._psv]))
        self.nvol = x.shape[0]

    def set_initial_volume_stress(self, s0):
        """The plane-strain volumes' s0 (nv, 3: s22, s23, s33): a volume
        whose three values are not all zero starts from them
        (viscouscycles.f90 of 2d/planestrain:779-810)."""
        self._s0 = _f64(s0).reshape(-1)
        self._chk(self._L.ucy_set_initial_volume_stress(self._h, self._s0.size // 3,
                                                        _p(self._s0)))

    def computeTractionKernelsPlaneStrain(self, x, sv, dv, nv, ns, y, W, dip, G, lambda_):
        """planestrain.f90:266-291: (td, tn) on one receiver from unit dip slip
        on ns sources."""
        td, tn = np.empty(ns), np.empty(ns)
        self._chk(self._L.ucy_planestrain_traction_kernels(
            self._h, _p(_f64(x)), _p(_f64(sv)), _p(_f64(dv)), _p(_f64(nv)), ns,
            _p(_f64(y).reshape(-1)), _p(_f64(W)), _p(_f64(dip)), float(G), float(lambda_),
            _p(td), _p(tn)))
        return td, tn

    def set_antiplane_volumes(self, x, thickness, width, dip):
        """Strain volumes of the 2-D antiplane viscouscycles model
        (fortran/2d/antiplane/src/viscouscycles.f90): upper centres x (nv, 3;
        x2, x3 used), thickness, width, dip (radians); after set_antiplane.
        State: 4 entries per volume (e12, e13, s12, s13); set_rheology takes
        the background e12, e13 in entries 1 and 2 of each edot_bg row."""
        x = _f64(x).reshape(-1, 3)
        self._apv = [x, _f64(thickness), _f64(width), _f64(dip)]
        self._chk(self._L.ucy_set_antiplane_volumes(self._h, x.shape[0],
                                                    *[_p(a) for a in self._apv]))
        self.nvol = x.shape[0]

    def set_rectangles(self, x, length, width, strike, dip):
        x = _f64(x).reshape(-1, 3)
        ns = x.shape[0]
        self.antiplane = self.planestrain = False
        self._rect = [x, _f64(length), _f64(width), _f64(strike), _f64(dip)]
        self.nrect = ns
        self._chk(self._L.ucy_set_rectangles(self._h, ns, *[_p(a) for a in self._rect]))
        self._update_np()

    def set_triangles(self, i1, i2, i3, vertices):
        """i1, i2, i3 are 1-based vertex indices (ratestate.f90:1252-1290)."""
        idx = [np.ascontiguousarray(np.asarray(a, dtype=np.int32)) for a in (i1, i2, i3)]
        v = _f64(vertices).reshape(-1, 3)
        self._tri = idx + [v]
        if self.antiplane or self.planestrain:
            self.nrect, self.antiplane, self.planestrain = 0, False, False
        self.ntri = idx[0].size
        self._chk(self._L.ucy_set_triangles(self._h, self.ntri, _pi(idx[0]), _pi(idx[1]),
                                            _pi(idx[2]), v.shape[0], _p(v)))
        self._update_np()

    def set_volumes(self, x, length, thickness, width, strike):
        x = _f64(x).reshape(-1, 3)
        self._vol = [x, _f64(length), _f64(thickness), _f64(width), _f64(strike)]
        if self.antiplane or self.planestrain:
            self.nrect, self.antiplane, self.planestrain = 0, False, False
            self._update_np()
        self.nvol = x.shape[0]
        self._chk(self._L.ucy_set_volumes(self._h, self.nvol, *[_p(a) for a in self._vol]))

    def _update_np(self):
        self.np = getattr(self, "nrect", 0) + getattr(self, "ntri", 0)

    def set_friction(self, Vpl, rake, mu0, sig, a, b, L, Vo, damping):
        arrs = [_f64(v) for v in (Vpl, rake, mu0, sig, a, b, L, Vo, damping)]
        n = arrs[0].size
        self._fric = arrs
        self._chk(self._L.ucy_set_friction(self._h, n, *[_p(v) for v in arrs]))

    def set_rheology(self, gammadot0, npower, edot_bg, Gm, Q, R, To):
        arrs = [_f64(v) for v in (gammadot0, npower)] + [_f64(edot_bg).reshape(-1)] + \
               [_f64(v) for v in (Gm, Q, R, To)]
        self._rheo = arrs
        self._chk(self._L.ucy_set_rheology(self._h, arrs[0].size, *[_p(v) for v in arrs]))

    # -- assembly ---------------------------------------------------------
    def buildG(self, precision: int = PREC_F64):
        """Assemble K (greens.f90:73-275) on the device.  PREC_SPLIT_BF16 /
        PREC_MIXED also keep K as split-bf16 pairs, which integrate() then
        streams (stage matvecs as increments over the step start)."""
        self._chk(self._L.ucy_build_kernel(self._h, precision))

    def matvec_reduced(self, v):
        """K~ v with the stored split-bf16 K (reduced-precision builds)."""
        v = _f64(v)
        t = np.empty(self.shape[0])
        self._chk(self._L.ucy_matvec_reduced(self._h, _p(v), _p(t)))
        return t

    def build_seconds(self):
        """Wall seconds of the last buildG by source type: (rectangles,
        triangles, volumes)."""
        r, t, v = C.c_double(), C.c_double(), C.c_double()
        self._chk(self._L.ucy_build_seconds(self._h, C.byref(r), C.byref(t), C.byref(v)))
        return r.value, t.value, v.value

    @property
    def shape(self):
        m, n = C.c_int64(), C.c_int64()
        self._chk(self._L.ucy_kernel_shape(self._h, C.byref(m), C.byref(n)))
        return m.value, n.value

    @property
    def dim(self) -> int:
        if self.planestrain:
            return 5 * self.np + 6 * self.nvol
        return 4 * (self.np + self.nvol) if self.antiplane else 7 * self.np + 12 * self.nvol

    def get_kernel(self) -> np.ndarray:
        m, n = self.shape
        K = np.empty((m, n))
        self._chk(self._L.ucy_get_kernel(self._h, _p(K)))
        return K

    def set_kernel(self, K):
        K = _f64(K)
        self._chk(self._L.ucy_set_kernel(self._h, _p(K), K.shape[0], K.shape[1]))

    def matvec(self, v) -> np.ndarray:
        v = _f64(v)
        m, _ = self.shape
        t = np.empty(m)
        self._chk(self._L.ucy_matvec_f64(self._h, _p(v), _p(t)))
        return t

    def kernel_dev_ptr(self) -> int:
        return self._L.ucy_kernel_dev(self._h) or 0

    def matvec_dev(self, v_ptr: int, t_ptr: int):
        self._chk(self._L.ucy_matvec_f64_dev(self._h, C.c_void_p(v_ptr), C.c_void_p(t_ptr)))

    # -- observation points (buildO, exportPoints) --------------------------
    def buildO(self, points):
        """Displacement kernels at observation points (npts, 3) north/east/down
        (greens.f90:302-462), built and kept on the device."""
        x = _f64(points).reshape(-1, 3)
        self._chk(self._L.ucy_build_displacement(self._h, x.shape[0], _p(x)))

    @property
    def displacement_shape(self):
        r, n = C.c_int64(), C.c_int64()
        self._chk(self._L.ucy_displacement_shape(self._h, C.byref(r), C.byref(n)))
        return r.value, n.value

    def get_displacement_kernels(self):
        """(O, Of, Ol) of buildO, each (3 npts, n) row-major."""
        r, n = self.displacement_shape
        O, Of, Ol = np.empty((r, n)), np.empty((r, n)), np.empty((r, n))
        self._chk(self._L.ucy_get_displacement_kernel(self._h, _p(O), _p(Of), _p(Ol)))
        return O, Of, Ol

    def displacement(self, y):
        """(d, df, dl) of exportPoints (ratestate.f90:414-486,
        viscouscycles.f90:523-625) for the state y."""
        r, _ = self.displacement_shape
        d, df, dl = np.empty(r), np.empty(r), np.empty(r)
        self._chk(self._L.ucy_displacement(self._h, _p(_f64(y)), _p(d), _p(df), _p(dl)))
        return d, df, dl

    # -- RHS and stepper --------------------------------------------------
    def initStateVector(self, variant: int = 0) -> np.ndarray:
        y = np.empty(self.dim)
        self._chk(self._L.ucy_init_state(self._h, variant, _p(y)))
        return y

    def odefun(self, n: int, time: float, y, dydt: np.ndarray) -> float:
        """Fill ``dydt`` in place (ratestate.f90:624); returns vMax."""
        y = _f64(y)
        vmax = C.c_double()
        self._chk(self._L.ucy_odefun(self._h, int(n), float(time), _p(y), _p(dydt),
                                     C.byref(vmax)))
        return vmax.value

    def RKCKm(self, n, t, yin, dydt, h, yout: np.ndarray, yerr: np.ndarray):
        self._chk(self._L.ucy_rkck(self._h, int(n), float(t), _p(_f64(yin)), _p(_f64(dydt)),
                                   float(h), _p(yout), _p(yerr)))

    def RKQSm(self, n, t, y: np.ndarray, dydt, yscal, htry, epsilon=1e-6,
              maximumTimeStep=1.7976931348623158e308):
        """Adaptive step (ode45.f90:51); updates ``y`` in place, returns (t, hdid, hnext)."""
        tt, hdid, hnext = C.c_double(t), C.c_double(), C.c_double()
        self._chk(self._L.ucy_rkqs(self._h, int(n), C.byref(tt), _p(y), _p(_f64(dydt)),
                                   _p(_f64(yscal)), float(htry), C.byref(hdid), C.byref(hnext),
                                   float(epsilon), float(maximumTimeStep)))
        return tt.value, hdid.value, hnext.value

    def integrate(self, y, interval: float, dt0: float = 1.0e-3, max_steps: int = 1000000,
                  epsilon: float = 1.0e-6, max_step: float = 1.7976931348623158e308,
                  log_cap: int = 0, history: bool = False, observe=None,
                  t0: float = 0.0, displacement: bool = False) -> IntegrateResult:
        """Main loop of ratestate.f90:307-328, device resident.  log_cap > 0
        logs time, dt and vMax per accepted step; history adds errmax/epsilon
        and the rejections of each step (the step-size proposal's data);
        observe (state indices, needs log_cap) adds the step-start values and
        derivatives of those entries per step (the export of
        ratestate.f90:488-536); displacement (needs log_cap and buildO) adds
        the observation-point displacement at each step start (exportPoints).
        t0 continues a run: the interval and the logged times are absolute."""
        y = _f64(y).copy()
        st = _lib.RunStats()
        logs = [np.empty(max(log_cap, 1)) for _ in range(6)]
        lg = _lib.StepLog()
        if log_cap:
            lg.t, lg.dt, lg.vmax = _p(logs[0]), _p(logs[1]), _p(logs[2])
            lg.emax = _p(logs[5])
            if history:
                lg.errmax, lg.rejects = _p(logs[3]), _p(logs[4])
        lg.cap = int(log_cap)
        obs = None
        if observe is not None and len(observe):
            idx = np.ascontiguousarray(np.asarray(observe, dtype=np.int64))
            obs = (idx, np.empty((max(log_cap, 1), idx.size)), np.empty((max(log_cap, 1), idx.size)))
            lg.n_obs = idx.size
            lg.obs_index = idx.ctypes.data_as(C.POINTER(C.c_int64))
            lg.obs_y, lg.obs_dydt = _p(obs[1]), _p(obs[2])
        dsp = None
        if displacement:
            r, _ = self.displacement_shape
            dsp = [np.empty((max(log_cap, 1), r)) for _ in range(3)]
            lg.disp, lg.disp_f, lg.disp_l = (_p(a) for a in dsp)
        self._chk(self._L.ucy_integrate_ex(self._h, _p(y), float(t0), float(interval), float(dt0),
                                           int(max_steps), float(epsilon), float(max_step),
                                           C.byref(lg), C.byref(st)))
        nlog = min(st.steps, log_cap)
        cut = lambda a, on=True: a[:nlog].copy() if (log_cap and on) else None  # noqa: E731
        res = IntegrateResult(y, st.time, st.dt_next, st.steps, st.attempts, st.rhs_evals,
                              st.vmax, st.wall_seconds, cut(logs[0]), cut(logs[1]),
                              cut(logs[2]), cut(logs[3], history), cut(logs[4], history))
        if obs is not None:
            res.obs_y, res.obs_dydt = obs[1][:nlog].copy(), obs[2][:nlog].copy()
        if dsp is not None:
            res.disp, res.disp_f, res.disp_l = (a[:nlog].copy() for a in dsp)
        res.emax_log = cut(logs[5])
        return res

    def time_rhs_kernel(self, iters: int = 50) -> float:
        """Average seconds of the fused matvec + post-pass kernel (HIP events)."""
        out = C.c_double()
        self._chk(self._L.ucy_time_rhs_kernel(self._h, int(iters), C.byref(out)))
        return out.value

    # -- Green's-function routines (one receiver, ns sources) -------------
    def computeTractionKernelsOkada92(self, x, sv, dv, nv, ns, y, L, W, strike, dip, s, d, t,
                                      G, lambda_):
        out = [np.empty(ns) for _ in range(3)]
        self._chk(self._L.ucy_okada92_traction_kernels(
            self._h, _p(_vec3(x)), _p(_vec3(sv)), _p(_vec3(dv)), _p(_vec3(nv)), int(ns),
            _p(_f64(y).reshape(-1)), _p(_f64(L)), _p(_f64(W)), _p(_f64(strike)), _p(_f64(dip)),
            float(s), float(d), float(t), float(G), float(lambda_), *[_p(o) for o in out]))
        return tuple(out)

    def computeStressKernelsOkada92(self, x, sv, dv, nv, ns, y, L, W, strike, dip, s, d, t,
                                    G, lambda_):
        out = [np.empty(ns) for _ in range(6)]
        self._chk(self._L.ucy_okada92_stress_kernels(
            self._h, _p(_vec3(x)), _p(_vec3(sv)), _p(_vec3(dv)), _p(_vec3(nv)), int(ns),
            _p(_f64(y).reshape(-1)), _p(_f64(L)), _p(_f64(W)), _p(_f64(strike)), _p(_f64(dip)),
            float(s), float(d), float(t), float(G), float(lambda_), *[_p(o) for o in out]))
        return tuple(out)

    def computeTractionKernelsNikkhoo15(self, x, sv, dv, nv, ns, i1, i2, i3, nVe, vertices,
                                        s, d, t, G, lambda_):
        idx = [np.ascontiguousarray(np.asarray(a, dtype=np.int32)) for a in (i1, i2, i3)]
        out = [np.empty(ns) for _ in range(3)]
        self._chk(self._L.ucy_nikkhoo15_traction_kernels(
            self._h, _p(_vec3(x)), _p(_vec3(sv)), _p(_vec3(dv)), _p(_vec3(nv)), int(ns),
            *[_pi(a) for a in idx], int(nVe), _p(_f64(vertices).reshape(-1)), float(s),
            float(d), float(t), float(G), float(lambda_), *[_p(o) for o in out]))
        return tuple(out)

    def computeTractionKernelsAntiplane(self, x, sv, dv, nv, ns, y, W, dip, G):
        ts = np.empty(ns)
        self._chk(self._L.ucy_antiplane_traction_kernels(
            self._h, _p(_vec3(x)), _p(_vec3(sv)), _p(_vec3(dv)), _p(_vec3(nv)), int(ns),
            _p(_f64(y)), _p(_f64(W)), _p(_f64(dip)), float(G), _p(ts)))
        return ts

    def straininhalfspace(self, points, P1, P2, P3, Ss, Ds, Ts, mu, lambda_):
        """straininhalfspace_ (StrainInHalfSpace.c:42) for n (point, triangle)
        pairs, east/north/up: returns (stress, strain), each n x 6
        (xx, xy, xz, yy, yz, zz)."""
        pts = _f64(points).reshape(-1, 3)
        n = pts.shape[0]
        X, Y, Z = (np.ascontiguousarray(pts[:, k]) for k in range(3))
        Ps = [_f64(P).reshape(n, 3) for P in (P1, P2, P3)]
        stress, strain = np.empty((n, 6)), np.empty((n, 6))
        self._chk(self._L.ucy_straininhalfspace(self._h, n, _p(stress), _p(strain), _p(X), _p(Y),
                                                _p(Z), *[_p(P) for P in Ps], float(Ss), float(Ds),
                                                float(Ts), float(mu), float(lambda_)))
        return stress, strain

    def computeStressKernelsNikkhoo15(self, x, sv, dv, nv, ns, i1, i2, i3, nVe, vertices,
                                      s, d, t, G, lambda_):
        idx = [np.ascontiguousarray(np.asarray(a, dtype=np.int32)) for a in (i1, i2, i3)]
        out = [np.empty(ns) for _ in range(6)]
        self._chk(self._L.ucy_nikkhoo15_stress_kernels(
            self._h, _p(_vec3(x)), _p(_vec3(sv)), _p(_vec3(dv)), _p(_vec3(nv)), int(ns),
            *[_pi(a) for a in idx], int(nVe), _p(_f64(vertices).reshape(-1)), float(s),
            float(d), float(t), float(G), float(lambda_), *[_p(o) for o in out]))
        return tuple(out)

    def computeTractionKernelsVerticalStrainVolume(self, x, sv, dv, nv, ns, y, L, T, W, strike,
                                                   e11p, e12p, e13p, e22p, e23p, e33p, G,
                                                   lambda_):
        out = [np.empty(ns) for _ in range(3)]
        self._chk(self._L.ucy_strainvolume_traction_kernels(
            self._h, _p(_vec3(x)), _p(_vec3(sv)), _p(_vec3(dv)), _p(_vec3(nv)), int(ns),
            _p(_f64(y).reshape(-1)), _p(_f64(L)), _p(_f64(T)), _p(_f64(W)), _p(_f64(strike)),
            float(e11p), float(e12p), float(e13p), float(e22p), float(e23p), float(e33p),
            float(G), float(lambda_), *[_p(o) for o in out]))
        return tuple(out)

    def computeStressKernelsVerticalStrainVolume(self, x, sv, dv, nv, ns, y, L, T, W, strike,
                                                 e11p, e12p, e13p, e22p, e23p, e33p, G, lambda_):
        out = [np.empty(ns) for _ in range(6)]
        self._chk(self._L.ucy_strainvolume_stress_kernels(
            self._h, _p(_vec3(x)), _p(_vec3(sv)), _p(_vec3(dv)), _p(_vec3(nv)), int(ns),
            _p(_f64(y).reshape(-1)), _p(_f64(L)), _p(_f64(T)), _p(_f64(W)), _p(_f64(strike)),
            float(e11p), float(e12p), float(e13p), float(e22p), float(e23p), float(e33p),
            float(G), float(lambda_), *[_p(o) for o in out]))
        return tuple(out)


# -- reference frames (host arithmetic in the library) ---------------------
def computeReferenceSystemOkada92(x, L, W, strike, dip):
    x = _f64(x).reshape(-1, 3)
    ns = x.shape[0]
    out = [np.empty((ns, 3)) for _ in range(4)]
    st = _lib.lib().ucy_reference_system_okada92(ns, _p(x), _p(_f64(L)), _p(_f64(W)),
                                                 _p(_f64(strike)), _p(_f64(dip)),
                                                 *[_p(o) for o in out])
    if st:
        raise UcyError(st, "computeReferenceSystemOkada92")
    return tuple(out)  # sv, dv, nv, xc


def computeReferenceSystemNikkhoo15(v, i1, i2, i3):
    v = _f64(v).reshape(-1, 3)
    idx = [np.ascontiguousarray(np.asarray(a, dtype=np.int32)) for a in (i1, i2, i3)]
    ns = idx[0].size
    out = [np.empty((ns, 3)) for _ in range(4)]
    st = _lib.lib().ucy_reference_system_nikkhoo15(v.shape[0], _p(v), ns, *[_pi(a) for a in idx],
                                                   *[_p(o) for o in out])
    if st:
        raise UcyError(st, "computeReferenceSystemNikkhoo15")
    return tuple(out)


def computeReferenceSystemVerticalStrainVolume(x, L, W, strike):
    x = _f64(x).reshape(-1, 3)
    ns = x.shape[0]
    out = [np.empty((ns, 3)) for _ in range(4)]
    st = _lib.lib().ucy_reference_system_strainvolume(ns, _p(x), _p(_f64(L)), _p(_f64(W)),
                                                      _p(_f64(strike)), *[_p(o) for o in out])
    if st:
        raise UcyError(st, "computeReferenceSystemVerticalStrainVolume")
    return tuple(out)
