"""ctypes binding of the CPU oracle (oracle/_build/libucyoracle.so) and of the
reference checker (oracle/_ref/libucyref.so) -- test infrastructure only.

The oracle is built on demand with ``make -C oracle oracle`` (gcc, seconds).
The reference checker exists only where the reference tree was available at
build time; tests that need it skip otherwise and rely on the committed
golden fixtures under tests/golden/.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "_build", "libucyoracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libucyref.so")

D = C.c_double
I = C.c_int
PD = C.POINTER(C.c_double)
PI = C.POINTER(C.c_int)


def p(a):
    return a.ctypes.data_as(PD)


def pi(a):
    return a.ctypes.data_as(PI)


def f64(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


class OrcMesh(C.Structure):
    _fields_ = [("lambda_", D), ("mu", D),
                ("nrect", I), ("rx", PD), ("rL", PD), ("rW", PD), ("rstrike", PD), ("rdip", PD),
                ("ntri", I), ("i1", PI), ("i2", PI), ("i3", PI), ("nve", I), ("vert", PD),
                ("nvol", I), ("vx", PD), ("vL", PD), ("vT", PD), ("vW", PD), ("vstrike", PD)]


class OrcSystem(C.Structure):
    _fields_ = [("n", I), ("Vpl", PD), ("rake", PD), ("mu0", PD), ("sig", PD), ("a", PD),
                ("b", PD), ("L", PD), ("Vo", PD), ("damping", PD),
                ("nvol", I), ("gammadot0", PD), ("npower", PD), ("edot_bg", PD),
                ("Gm", PD), ("Qa", PD), ("Rgas", PD), ("To", PD), ("ic_variant", I),
                ("K", PD), ("m", I), ("ncol", I), ("nthreads", I)]


class OrcStats(C.Structure):
    _fields_ = [("time", D), ("dt_next", D), ("steps", C.c_long), ("attempts", C.c_long),
                ("rhs_evals", C.c_long)]


_ORC = None


def oracle() -> C.CDLL:
    global _ORC
    if _ORC is None:
        if not os.path.exists(ORACLE_SO):
            subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "oracle"],
                           check=True)
        lib = C.CDLL(ORACLE_SO)
        lib.orc_dc3d.restype = I
        lib.orc_buildG.restype = I
        lib.orc_odefun.restype = D
        lib.orc_rkqs.restype = I
        lib.orc_integrate.restype = C.c_long
        lib.orc_stress_strainvolume.restype = I
        lib.orc_displacement_okada92.restype = I
        lib.orc_buildO.restype = I
        _ORC = lib
    return _ORC


def ref():
    if not os.path.exists(REF_SO):
        return None
    return C.CDLL(REF_SO)


class Mesh:
    """Keeps the numpy arrays alive behind an OrcMesh."""

    def __init__(self, lambda_, mu, rect=None, tri=None, vol=None):
        self.keep = []
        m = OrcMesh()
        m.lambda_, m.mu = lambda_, mu
        if rect is not None:
            x, L, W, st, dp = [f64(a) for a in rect]
            self.keep += [x, L, W, st, dp]
            m.nrect = L.size
            m.rx, m.rL, m.rW, m.rstrike, m.rdip = p(x), p(L), p(W), p(st), p(dp)
        if tri is not None:
            i1, i2, i3 = [np.ascontiguousarray(np.asarray(a, dtype=np.int32)) for a in tri[:3]]
            v = f64(tri[3])
            self.keep += [i1, i2, i3, v]
            m.ntri = i1.size
            m.i1, m.i2, m.i3, m.nve, m.vert = pi(i1), pi(i2), pi(i3), v.size // 3, p(v)
        if vol is not None:
            x, L, T, W, st = [f64(a) for a in vol]
            self.keep += [x, L, T, W, st]
            m.nvol = L.size
            m.vx, m.vL, m.vT, m.vW, m.vstrike = p(x), p(L), p(T), p(W), p(st)
        self.c = m
        self.np = m.nrect + m.ntri
        self.nvol = m.nvol
        self.shape = (3 * self.np + 6 * self.nvol, 2 * self.np + 6 * self.nvol)


def buildG(mesh: Mesh, nthreads: int = 1) -> np.ndarray:
    K = np.empty(mesh.shape)
    rc = oracle().orc_buildG(C.byref(mesh.c), p(K), nthreads)
    if rc:
        raise RuntimeError(f"orc_buildG failed {rc}")
    return K


def buildO(mesh: Mesh, points, nthreads: int = 1):
    """(O, Of, Ol) of greens.f90:302-462, each (3 npts x n) row-major."""
    x = f64(points)
    nobs = x.size // 3
    shp = (3 * nobs, mesh.shape[1])
    O, Of, Ol = np.empty(shp), np.empty(shp), np.empty(shp)
    rc = oracle().orc_buildO(C.byref(mesh.c), nobs, p(x), p(O), p(Of), p(Ol), nthreads)
    if rc:
        raise RuntimeError(f"orc_buildO failed {rc}")
    return O, Of, Ol


def slip_vector(y, np_, nvol):
    """u of exportPoints (ratestate.f90:431-447, viscouscycles.f90:539-580):
    strike and dip slip per patch, then the six strain entries per volume."""
    y = np.asarray(y)
    u = [y[:7 * np_].reshape(np_, 7)[:, :2].ravel()]
    if nvol:
        u.append(y[7 * np_:].reshape(nvol, 12)[:, :6].ravel())
    return np.concatenate(u)


AP_RHEO = ("gammadot0", "npower", "e12", "e13", "Gm", "Q", "R", "To")


class OrcAPSystem(C.Structure):
    _fields_ = [("n", C.c_int)] + [(k, C.POINTER(C.c_double)) for k in (
        "Vpl", "mu0", "sig", "a", "b", "L", "Vo", "damping", "K")] + [
        ("nvol", C.c_int)] + [(k, C.POINTER(C.c_double)) for k in AP_RHEO]


def buildG_antiplane(x, W, dip, mu, nthreads=1):
    x, W, dip = f64(x), f64(W), f64(dip)
    ns = W.size
    K = np.empty((ns, ns))
    oracle().orc_buildG_antiplane(ns, p(x), p(W), p(dip), C.c_double(mu), p(K), nthreads)
    return K


def buildG_antiplane_volumes(x, W, dip, vx, vT, vW, vdip, mu, nthreads=1):
    """buildG of the 2-D viscouscycles model: patches and strain volumes."""
    x, W, dip = f64(x), f64(W), f64(dip)
    vx, vT, vW, vdip = f64(vx), f64(vT), f64(vW), f64(vdip)
    np_, nv = W.size, vW.size
    n = np_ + 2 * nv
    K = np.empty((n, n))
    oracle().orc_buildG_antiplane_volumes(np_, p(x), p(W), p(dip), nv, p(vx), p(vT), p(vW),
                                          p(vdip), C.c_double(mu), p(K), nthreads)
    return K


def stress_sv_antiplane(x2, x3, q2, q3, T, W, dip, e12, e13, G):
    s12, s13 = C.c_double(), C.c_double()
    d = C.c_double
    oracle().orc_stress_sv_antiplane(d(x2), d(x3), d(q2), d(q3), d(T), d(W), d(dip), d(e12),
                                     d(e13), d(G), C.byref(s12), C.byref(s13))
    return s12.value, s13.value


def displacement_sv_antiplane(x2, x3, q2, q3, T, W, dip, e12, e13):
    lib = oracle()
    lib.orc_displacement_sv_antiplane.restype = C.c_double
    d = C.c_double
    return lib.orc_displacement_sv_antiplane(d(x2), d(x3), d(q2), d(q3), d(T), d(W), d(dip),
                                             d(e12), d(e13))


def buildO_antiplane(x, W, dip, points):
    x, W, dip, pts = f64(x), f64(W), f64(dip), f64(points)
    ns, npts = W.size, pts.size // 3
    O = np.empty((npts, ns))
    oracle().orc_buildO_antiplane(ns, p(x), p(W), p(dip), npts, p(pts), p(O))
    return O


class OrcPSSystem(C.Structure):
    _fields_ = [("n", C.c_int)] + [(k, C.POINTER(C.c_double)) for k in (
        "Vpl", "tau0", "mu0", "sig", "a", "b", "L", "Vo", "damping", "K")] + \
        [("nv", C.c_int)] + [(k, C.POINTER(C.c_double)) for k in (
            "gammadot0", "npower", "Q", "R", "To", "ebg", "s0")]


# the plane-strain volumes' rheology arrays, one entry per volume (ebg: 3
# per volume, the background e22, e23, e33)
PS_RHEO = ("gammadot0", "npower", "Q", "R", "To", "ebg")


def buildG_planestrain_volumes(x, W, dip, vx, vT, vW, vdip, G, lam, nthreads=1):
    """buildG of the plane-strain viscouscycles program: (2 np + 3 nv) x
    (np + 3 nv); rows td, tn per patch then s22', s23', s33' per volume;
    columns the dip slips then 3 unit strains (1e-3) per volume."""
    x, W, dip, vx, vT, vW, vdip = (f64(a) for a in (x, W, dip, vx, vT, vW, vdip))
    np_, nv = W.size, vW.size
    K = np.empty((2 * np_ + 3 * nv, np_ + 3 * nv))
    rc = oracle().orc_buildG_planestrain_volumes(np_, p(x), p(W), p(dip), nv, p(vx), p(vT), p(vW),
                                                 p(vdip), C.c_double(G), C.c_double(lam), p(K),
                                                 nthreads)
    if rc:
        raise RuntimeError(f"orc_buildG_planestrain_volumes failed {rc}")
    return K


def ps_kernel_call(fn, x, sv, dv, nv, y, arrays, scalars, nout, lib=None):
    """A vectorised plane-strain kernel routine (the oracle's orc_* or the
    reference build's ref_*): receiver frame, sources y (3, ns) and per-source
    arrays, then scalars; returns nout arrays of ns values."""
    y = f64(y)
    ns = y.size // 3
    outs = [np.empty(ns) for _ in range(nout)]
    args = [p(f64(a)) for a in (x, sv, dv, nv)] + [ns, p(y)] + [p(f64(a)) for a in arrays] + \
        [C.c_double(v) for v in scalars] + [p(o) for o in outs]
    rc = getattr(lib or oracle(), fn)(*args)
    if lib is None and rc:          # the reference's routines are subroutines
        raise RuntimeError(f"{fn} failed {rc}")
    return outs


def buildG_planestrain(x, W, dip, G, lam, nthreads=1):
    """buildG of the plane-strain model: (2 ns x ns), rows dip / normal traction."""
    x, W, dip = f64(x), f64(W), f64(dip)
    ns = W.size
    K = np.empty((2 * ns, ns))
    rc = oracle().orc_buildG_planestrain(ns, p(x), p(W), p(dip), C.c_double(G), C.c_double(lam),
                                         p(K), nthreads)
    if rc:
        raise RuntimeError(f"orc_buildG_planestrain failed {rc}")
    return K


def _ps_call(fn, args, nout):
    outs = [C.c_double() for _ in range(nout)]
    rc = getattr(oracle(), fn)(*[C.c_double(a) for a in args], *[C.byref(o) for o in outs])
    if rc:
        raise RuntimeError(f"{fn} failed {rc}")
    return tuple(o.value for o in outs)


def stress_sv_planestrain(r2, x3, q2, q3, T, W, dip, e22, e23, e33, G, lam):
    return _ps_call("orc_stress_sv_planestrain", (r2, x3, q2, q3, T, W, dip, e22, e23, e33, G, lam), 3)


def displacement_sv_planestrain(r2, x3, q2, q3, T, W, dip, e22, e23, e33, G, lam):
    return _ps_call("orc_displacement_sv_planestrain",
                    (r2, x3, q2, q3, T, W, dip, e22, e23, e33, G, lam), 2)


def stress_planestrain(x2, x3, q2, q3, W, dip, s, G, lam):
    return _ps_call("orc_stress_planestrain", (x2, x3, q2, q3, W, dip, s, G, lam), 3)


def displacement_planestrain(x2, x3, q2, q3, W, dip, s, G, lam):
    return _ps_call("orc_displacement_planestrain", (x2, x3, q2, q3, W, dip, s, G, lam), 2)


class PSSystem:
    """2-D plane-strain rate-and-state (fortran/2d/planestrain/src/ratestate.f90);
    tau0 (may be None) is the deck's initial-traction option.  With rheo (the
    PS_RHEO arrays) the strain volumes of that directory's viscouscycles
    program follow the patches; s0 (3 per volume, may be None) is their
    initial-stress option."""

    def __init__(self, fric, K, tau0=None, rheo=None, s0=None):
        arrs = [f64(a) for a in fric]
        n = arrs[0].size
        t0 = f64(tau0) if tau0 is not None else None
        self.keep = arrs + [f64(K)] + ([t0] if t0 is not None else [])
        s = OrcPSSystem()
        s.n = n
        for k, a in zip(("Vpl", "mu0", "sig", "a", "b", "L", "Vo", "damping"), arrs):
            setattr(s, k, p(a))
        s.K = p(self.keep[8])
        if t0 is not None:
            s.tau0 = p(t0)
        nv = 0
        if rheo is not None:
            ra = [f64(a) for a in rheo]
            nv = ra[0].size
            assert ra[5].size == 3 * nv
            self.keep += ra
            for k, a in zip(PS_RHEO, ra):
                setattr(s, k, p(a))
            if s0 is not None:
                z = f64(s0)
                assert z.size == 3 * nv
                self.keep.append(z)
                s.s0 = p(z)
        s.nv = nv
        assert self.keep[8].shape == (2 * n + 3 * nv, n + 3 * nv)
        self.c = s
        self.n = n
        self.nv = nv
        self.dim = 5 * n + 6 * nv

    def init_state(self):
        y = np.empty(self.dim)
        oracle().orc_ps_init_state(C.byref(self.c), p(y))
        return y

    def odefun(self, y):
        y = f64(y)
        dydt, v = np.empty(self.dim), np.empty(self.n + 3 * self.nv)
        t = np.empty(2 * self.n + 3 * self.nv)
        lib = oracle()
        lib.orc_ps_odefun.restype = C.c_double
        vmax = lib.orc_ps_odefun(C.byref(self.c), p(y), p(dydt), p(v), p(t))
        return dydt, vmax

    def integrate(self, y0, interval, dt0=1e-3, max_steps=10**6, epsilon=1e-6,
                  max_step=1.7976931348623158e308, log_cap=0):
        y = f64(y0).copy()
        logs = [np.empty(max(log_cap, 1)) for _ in range(3)]
        st = OrcStats()
        lib = oracle()
        lib.orc_ps_integrate.restype = C.c_long
        rc = lib.orc_ps_integrate(C.byref(self.c), p(y), C.c_double(interval), C.c_double(dt0),
                                  C.c_long(max_steps), C.c_double(epsilon),
                                  C.c_double(max_step), p(logs[0]), p(logs[1]), p(logs[2]),
                                  C.c_long(log_cap), C.byref(st))
        n = min(st.steps, log_cap)
        return y, st, logs[0][:n], logs[1][:n], logs[2][:n], rc


class APSystem:
    """2-D antiplane rate-and-state (fortran/2d/antiplane/src/ratestate.f90),
    with strain volumes when rheo (the AP_RHEO arrays, one entry per volume)
    is given (viscouscycles.f90 there); K is then (n + 2 nvol) square."""

    def __init__(self, fric, K, rheo=None):
        self.keep = [f64(a) for a in fric] + [f64(K)]
        s = OrcAPSystem()
        s.n = self.keep[0].size
        for k, a in zip(("Vpl", "mu0", "sig", "a", "b", "L", "Vo", "damping", "K"), self.keep):
            setattr(s, k, p(a))
        s.nvol = 0
        if rheo is not None:
            arrs = [f64(a) for a in rheo]
            self.keep += arrs
            s.nvol = arrs[0].size
            for k, a in zip(AP_RHEO, arrs):
                setattr(s, k, p(a))
        assert self.keep[8].shape == (s.n + 2 * s.nvol,) * 2
        self.c = s
        self.n = s.n
        self.nvol = s.nvol
        self.dim = 4 * (s.n + s.nvol)

    def init_state(self):
        y = np.empty(self.dim)
        oracle().orc_ap_init_state(C.byref(self.c), p(y))
        return y

    def odefun(self, y):
        y = f64(y)
        nc = self.n + 2 * self.nvol
        dydt, v, t = np.empty(self.dim), np.empty(nc), np.empty(nc)
        lib = oracle()
        lib.orc_ap_odefun.restype = C.c_double
        vmax = lib.orc_ap_odefun(C.byref(self.c), p(y), p(dydt), p(v), p(t))
        return dydt, vmax

    def integrate(self, y0, interval, dt0=1e-3, max_steps=10**6, epsilon=1e-6,
                  max_step=1.7976931348623158e308, log_cap=0):
        y = f64(y0).copy()
        logs = [np.empty(max(log_cap, 1)) for _ in range(3)]
        st = OrcStats()
        lib = oracle()
        lib.orc_ap_integrate.restype = C.c_long
        rc = lib.orc_ap_integrate(C.byref(self.c), p(y), C.c_double(interval), C.c_double(dt0),
                                  C.c_long(max_steps), C.c_double(epsilon),
                                  C.c_double(max_step), p(logs[0]), p(logs[1]), p(logs[2]),
                                  C.c_long(log_cap), C.byref(st))
        n = min(st.steps, log_cap)
        return y, st, logs[0][:n], logs[1][:n], logs[2][:n], rc


class System:
    def __init__(self, fric, K, rheo=None, ic_variant=0, nthreads=1):
        self.keep = []
        s = OrcSystem()
        arrs = [f64(a) for a in fric]
        self.keep += arrs
        s.n = arrs[0].size
        (s.Vpl, s.rake, s.mu0, s.sig, s.a, s.b, s.L, s.Vo, s.damping) = [p(a) for a in arrs]
        if rheo is not None:
            r = [f64(a) for a in rheo]
            self.keep += r
            s.nvol = r[0].size
            (s.gammadot0, s.npower, s.edot_bg, s.Gm, s.Qa, s.Rgas, s.To) = [p(a) for a in r]
        s.ic_variant = ic_variant
        self.K = f64(K)
        self.keep.append(self.K)
        s.K = p(self.K)
        s.m, s.ncol = self.K.shape
        s.nthreads = nthreads
        self.c = s
        self.dim = 7 * s.n + 12 * s.nvol

    def init_state(self):
        y = np.empty(self.dim)
        oracle().orc_init_state(C.byref(self.c), p(y))
        return y

    def odefun(self, y):
        dydt = np.empty(self.dim)
        v = np.empty(self.c.ncol)
        t = np.empty(self.c.m)
        vmax = oracle().orc_odefun(C.byref(self.c), p(f64(y)), p(dydt), p(v), p(t))
        return dydt, vmax, v, t

    def integrate(self, y0, interval, dt0=1e-3, max_steps=10**6, epsilon=1e-6,
                  max_step=1.7976931348623158e308, log_cap=0):
        y = f64(y0).copy()
        tl = np.empty(max(log_cap, 1))
        dl = np.empty(max(log_cap, 1))
        vl = np.empty(max(log_cap, 1))
        st = OrcStats()
        rc = oracle().orc_integrate(C.byref(self.c), p(y), D(interval), D(dt0), C.c_long(max_steps),
                                    D(epsilon), D(max_step), p(tl), p(dl), p(vl),
                                    C.c_long(log_cap), C.byref(st))
        n = min(st.steps, log_cap)
        self.vmax_log = vl[:n].copy()
        return y, st, tl[:n].copy(), dl[:n].copy(), rc


RHS_FN = C.CFUNCTYPE(None, C.c_void_p, PD, PD)


def rkqs(n, t, y, dydt, yscal, htry, f, epsilon=1e-6, max_step=1.7976931348623158e308):
    """Oracle RKQSm with a Python RHS f(y)->dydt (small problems only)."""
    def cb(ctx, yp, dp):
        yy = np.ctypeslib.as_array(yp, shape=(n,))
        dd = np.ctypeslib.as_array(dp, shape=(n,))
        dd[:] = f(yy.copy())
    cbf = RHS_FN(cb)
    w = np.empty(9 * n)
    tt = D(t)
    hdid, hnext = D(), D()
    att = I()
    y = f64(y).copy()
    rc = oracle().orc_rkqs(n, C.byref(tt), p(y), p(f64(dydt)), p(f64(yscal)), D(htry),
                           C.byref(hdid), C.byref(hnext), D(epsilon), D(max_step), p(w), cbf,
                           None, C.byref(att))
    return rc, tt.value, y, hdid.value, hnext.value, att.value


def rkck(n, y, dydt, h, f):
    def cb(ctx, yp, dp):
        yy = np.ctypeslib.as_array(yp, shape=(n,))
        dd = np.ctypeslib.as_array(dp, shape=(n,))
        dd[:] = f(yy.copy())
    cbf = RHS_FN(cb)
    yout = np.empty(n)
    yerr = np.empty(n)
    w = np.empty(7 * n)
    oracle().orc_rkck(n, p(f64(y)), p(f64(dydt)), D(h), p(yout), p(yerr), p(w), cbf, None)
    return yout, yerr
