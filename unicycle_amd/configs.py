"""Synthetic fault models (seeded) in the reference's units and layouts.

Units follow the reference input decks: lengths in m, stresses and moduli in
MPa, velocities in m/s, time in s.  Angles handed to the engine are radians
(the deck's degrees are converted at ratestate.f90:1174-1176).

thrust_fault
    planar dipping thrust of nx (along strike) x nz (down dip) Okada
    rectangles; aging-law rate-and-state with a velocity-weakening ellipse,
    radiation damping mu/(2 Vs).  nx=64, nz=32, 2 km patches is the
    2048-patch single-instance model; the ensemble groups use the same rules
    at 1024/2048/4096/8192 patches.
megathrust_mesh
    curved megathrust of ~20,000 jittered triangles with random-field
    friction (the triangle-mesh model).
volume_grid
    vertical strain-volume cuboids with Maxwell or power-law rheology (with
    the 4096-patch, 1.5 km thrust: the fault + viscoelastic model).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

YEAR = 3.15576e7  # s (Julian year)


@dataclass
class FaultModel:
    lambda_: float
    mu: float
    x: np.ndarray          # (N,3) upper-left corners, north/east/down [m]
    length: np.ndarray
    width: np.ndarray
    strike: np.ndarray     # rad
    dip: np.ndarray        # rad
    Vpl: np.ndarray
    rake: np.ndarray       # rad
    mu0: np.ndarray
    sig: np.ndarray
    a: np.ndarray
    b: np.ndarray
    L: np.ndarray
    Vo: np.ndarray
    damping: np.ndarray
    nx: int = 0
    nz: int = 0
    meta: dict = field(default_factory=dict)

    @property
    def n(self) -> int:
        return self.x.shape[0]

    def friction(self):
        return (self.Vpl, self.rake, self.mu0, self.sig, self.a, self.b, self.L, self.Vo,
                self.damping)

    def geometry(self):
        return (self.x, self.length, self.width, self.strike, self.dip)


def thrust_fault(nx: int = 64, nz: int = 32, patch: float = 2000.0, dip_deg: float = 30.0,
                 top: float = 500.0, mu: float = 30e3, nu: float = 0.25, vs: float = 3000.0,
                 a: float = 0.010, b_in: float = 0.014, b_out: float = 0.006,
                 ellipse: tuple | None = (40e3, 20e3), L: float = 0.01, sig: float = 100.0,
                 Vo: float = 1e-6, mu0: float = 0.6, Vpl: float = 1e-9,
                 rake_deg: float = 89.0) -> FaultModel:
    """Planar thrust, strike north, dipping east; patch (i,j) row-major with
    i along strike fastest.

    The default rake is 89 degrees (a thrust with a 1-degree oblique
    component).  At exactly 90 degrees the strike-slip and strike-traction
    state entries are pure rounding noise (cos(pi/2) ~ 6e-17), yet they enter
    the RK45 error norm with the 1e-30 floor of yscal (ratestate.f90:315), so
    the step-size sequence -- and the run length -- is set by round-off and
    differs between any two summation orders of the matvec."""
    lam = 2.0 * mu * nu / (1.0 - 2.0 * nu)
    dip = np.deg2rad(dip_deg)
    i = np.tile(np.arange(nx), nz)
    j = np.repeat(np.arange(nz), nx)
    N = nx * nz
    x = np.empty((N, 3))
    x[:, 0] = i * patch                                   # north
    x[:, 1] = j * patch * np.cos(dip)                     # east (down-dip)
    x[:, 2] = top + j * patch * np.sin(dip)               # depth
    # velocity-weakening ellipse centred on the fault
    sa = (i + 0.5) * patch - 0.5 * nx * patch
    sd = (j + 0.5) * patch - 0.5 * nz * patch
    if ellipse is None:
        ax, az = 0.625 * 0.5 * nx * patch, 0.625 * 0.5 * nz * patch
    else:
        ax, az = ellipse
    inside = (sa / ax) ** 2 + (sd / az) ** 2 < 1.0
    one = np.ones(N)
    return FaultModel(
        lambda_=lam, mu=mu, x=x, length=patch * one, width=patch * one,
        strike=0.0 * one, dip=dip * one, Vpl=Vpl * one, rake=np.deg2rad(rake_deg) * one,
        mu0=mu0 * one, sig=sig * one, a=a * one, b=np.where(inside, b_in, b_out),
        L=L * one, Vo=Vo * one, damping=mu / (2.0 * vs) * one, nx=nx, nz=nz,
        meta={"inside": inside})


@dataclass
class TriangleModel:
    """Triangular mesh in the layout of the reference's .tri/.ned decks
    (ratestate.f90:1252-1290): 1-based vertex indices, vertices north/east/down."""
    lambda_: float
    mu: float
    i1: np.ndarray
    i2: np.ndarray
    i3: np.ndarray
    vertices: np.ndarray   # (nVe,3)
    fric: tuple            # (Vpl, rake, mu0, sig, a, b, L, Vo, damping), per triangle

    @property
    def n(self) -> int:
        return self.i1.size

    def geometry(self):
        return (self.i1, self.i2, self.i3, self.vertices)

    def friction(self):
        return self.fric


def triangulate(model: FaultModel) -> TriangleModel:
    """Split every rectangle of ``model`` along its diagonal into the
    triangles (A,B,C) and (A,C,D), A the upper corner at the start of strike,
    B along strike, C,D one width down dip.  Vertices shared between patches
    are merged; friction is inherited by both halves."""
    st, dp = model.strike, model.dip
    sv = np.stack([np.cos(st), np.sin(st), 0.0 * st], 1)
    dd = np.stack([-np.sin(st) * np.cos(dp), np.cos(st) * np.cos(dp), np.sin(dp)], 1)
    A = model.x
    B = A + model.length[:, None] * sv
    C = B + model.width[:, None] * dd
    Dc = A + model.width[:, None] * dd
    corners = np.stack([A, B, C, Dc], 1).reshape(-1, 3)
    key = np.round(corners, 6)
    verts, inv = np.unique(key, axis=0, return_inverse=True)
    inv = inv.reshape(-1, 4) + 1
    i1 = np.stack([inv[:, 0], inv[:, 0]], 1).reshape(-1).astype(np.int32)
    i2 = np.stack([inv[:, 1], inv[:, 2]], 1).reshape(-1).astype(np.int32)
    i3 = np.stack([inv[:, 2], inv[:, 3]], 1).reshape(-1).astype(np.int32)
    fric = tuple(np.repeat(np.asarray(f, dtype=np.float64), 2) for f in model.friction())
    return TriangleModel(model.lambda_, model.mu, i1, i2, i3, verts.astype(np.float64), fric)


def gaussian_field(points: np.ndarray, mean: float, sd: float, corr: float, seed: int,
                   modes: int = 512) -> np.ndarray:
    """Stationary Gaussian random field with covariance sd^2 exp(-r^2 / (2 corr^2))
    at ``points`` (n, 2) [m], by random Fourier features (Rahimi and Recht
    2007): sqrt(2/M) sum_m cos(k_m . x + phi_m), k_m ~ N(0, I/corr^2),
    phi_m ~ U(0, 2 pi).  Seeded."""
    rng = np.random.default_rng(seed)
    k = rng.normal(0.0, 1.0 / corr, (modes, 2))
    ph = rng.uniform(0.0, 2.0 * np.pi, modes)
    f = np.sqrt(2.0 / modes) * np.cos(points @ k.T + ph).sum(axis=1)
    return mean + sd * f


def megathrust_mesh(nx: int = 100, ny: int = 100, length: float = 200e3, width: float = 100e3,
                    jitter: float = 0.1, seed: int = 11, mu: float = 30e3, nu: float = 0.25,
                    vs: float = 3000.0, a: float = 0.010, sig: float = 100.0, Vo: float = 1e-6,
                    mu0: float = 0.6, Vpl: float = 1e-9, rake_deg: float = 89.0) -> TriangleModel:
    """Curved megathrust z = 5 + 0.1 x + 0.0004 x^2 km (x across strike, km)
    over ``length`` along strike (north) x ``width`` across (east): an
    nx (strike) x ny (dip) grid of quadrilaterals split along diagonals into
    2 nx ny triangles (20,000 at the defaults); interior vertices jittered by
    U(-jitter, jitter) x edge length in plan view, depths then taken on the
    surface.  Friction: a - b from a seeded Gaussian random field (mean
    -0.002, sd 0.003, correlation 20 km) at the centroids, L ~ lognormal(ln
    0.02 m, 0.5), the rest as thrust_fault."""
    rng = np.random.default_rng(seed)
    dn, de = length / nx, width / ny
    gi, gj = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), indexing="ij")
    north = gi * dn
    east = gj * de
    interior = (gi > 0) & (gi < nx) & (gj > 0) & (gj < ny)
    north = north + interior * rng.uniform(-jitter, jitter, north.shape) * dn
    east = east + interior * rng.uniform(-jitter, jitter, east.shape) * de
    xk = east / 1e3
    depth = (5.0 + 0.1 * xk + 0.0004 * xk * xk) * 1e3
    verts = np.stack([north.ravel(), east.ravel(), depth.ravel()], 1)
    vid = lambda i, j: i * (ny + 1) + j + 1  # noqa: E731  (1-based, as the decks)
    qi, qj = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    qi, qj = qi.ravel(), qj.ravel()
    A, B, Cc, D = vid(qi, qj), vid(qi + 1, qj), vid(qi + 1, qj + 1), vid(qi, qj + 1)
    i1 = np.stack([A, A], 1).reshape(-1).astype(np.int32)
    i2 = np.stack([B, Cc], 1).reshape(-1).astype(np.int32)
    i3 = np.stack([Cc, D], 1).reshape(-1).astype(np.int32)
    n = i1.size
    cen = (verts[i1 - 1] + verts[i2 - 1] + verts[i3 - 1]) / 3.0
    amb = gaussian_field(cen[:, :2], -0.002, 0.003, 20e3, seed + 1)
    Lc = np.exp(np.log(0.02) + 0.5 * np.random.default_rng(seed + 2).normal(size=n))
    lam = 2.0 * mu * nu / (1.0 - 2.0 * nu)
    one = np.ones(n)
    fric = (Vpl * one, np.deg2rad(rake_deg) * one, mu0 * one, sig * one, a * one, a - amb,
            Lc, Vo * one, mu / (2.0 * vs) * one)
    return TriangleModel(lam, mu, i1, i2, i3, verts, fric)


def perturbed_initial_state(model: FaultModel, y0: np.ndarray, seed: int,
                            amplitude: float = 1e-3) -> np.ndarray:
    """Initial slip rate Vpl*(1 + amplitude*N(0,1)) with the state variable at
    steady state theta = L/V; tractions as the reference initial state."""
    rng = np.random.default_rng(seed)
    y = y0.copy()
    V = model.Vpl * (1.0 + amplitude * rng.standard_normal(model.n))
    y[6:7 * model.n:7] = np.log10(V)
    y[5:7 * model.n:7] = np.log10(model.L / V)
    return y


ENSEMBLE_GROUPS = ((32, 32), (64, 32), (64, 64), (128, 64))  # 1024, 2048, 4096, 8192 patches


def ensemble_instances(base: FaultModel, n_inst: int, seed: int):
    """Per-instance friction draws of the uneven-size ensemble: in the
    weakening zone a-b ~ U(-0.006,-0.002) (a kept, b = a - (a-b)); everywhere
    L ~ lognormal(ln 0.01 m, 0.4), sigma ~ U(50,150) MPa."""
    rng = np.random.default_rng(seed)
    N = base.n
    inside = base.meta["inside"]
    amb = rng.uniform(-0.006, -0.002, size=(n_inst, 1))
    a = np.broadcast_to(base.a, (n_inst, N)).copy()
    b = np.where(inside[None, :], a - amb, np.broadcast_to(base.b, (n_inst, N)))
    L = np.exp(np.log(0.01) + 0.4 * rng.standard_normal((n_inst, 1))) * np.ones((1, N))
    sig = rng.uniform(50.0, 150.0, size=(n_inst, 1)) * np.ones((1, N))
    seeds = rng.integers(0, 2**31 - 1, size=n_inst)
    return {"a": a, "b": b, "L": L, "sig": sig, "seeds": seeds}


def ensemble_setup(engine, precision: int, groups=ENSEMBLE_GROUPS, n_inst: int = 64,
                   seed: int = 100):
    """The uneven-size ensemble on ``engine``: one group per (nx, nz) of
    ``groups`` (2 km patches, the thrust_fault layout rules), n_inst
    instances each with the draws of ensemble_instances and the reference
    initial state (ratestate.f90:548-610) of its own friction, perturbed by
    perturbed_initial_state.  Returns (ensemble, [(model, draws, y0), ...])
    with the ensemble built."""
    from .engine import Engine
    from .ensemble import Ensemble
    ens = Ensemble(engine, precision)
    info = []
    for gi, (nx, nz) in enumerate(groups):
        m = thrust_fault(nx=nx, nz=nz, ellipse=None)
        d = ensemble_instances(m, n_inst, seed=seed + gi)
        ic = Engine(engine.device)
        try:
            ic.set_elastic(m.lambda_, m.mu)
            ic.set_rectangles(*m.geometry())
            y0 = []
            for i in range(n_inst):
                fr = list(m.friction())
                fr[3], fr[4], fr[5], fr[6] = d["sig"][i], d["a"][i], d["b"][i], d["L"][i]
                ic.set_friction(*fr)
                y0.append(perturbed_initial_state(m, ic.initStateVector(), int(d["seeds"][i])))
        finally:
            ic.close()
        y0 = np.array(y0)
        ens.add_group(m, d["a"], d["b"], d["sig"], d["L"], y0)
        info.append((m, d, y0))
    ens.build()
    return ens, info


def instance_friction(model: FaultModel, draws, i: int):
    """Friction tuple of instance i of an ensemble group."""
    fr = list(model.friction())
    fr[3], fr[4], fr[5], fr[6] = draws["sig"][i], draws["a"][i], draws["b"][i], draws["L"][i]
    return tuple(fr)


@dataclass
class VolumeModel:
    """Vertical strain volumes (strainvolume.f90 conventions: x = upper corner
    at the start of strike, north/east/down; length along strike, thickness
    across, width down) and their rheology (viscouscycles.f90:933-1000:
    power law with prefactor gammadot0 and exponent n; Maxwell is n = 1,
    gammadot0 = 1/(2 eta)) with the initial-stress law of
    viscouscycles.f90:769-803 (Gm, Q, R, To; Q = 0 makes it consistent with
    the RHS)."""
    x: np.ndarray
    length: np.ndarray
    thickness: np.ndarray
    width: np.ndarray
    strike: np.ndarray
    gammadot0: np.ndarray
    npower: np.ndarray
    edot_bg: np.ndarray    # (nvol, 6)
    Gm: np.ndarray
    Q: np.ndarray
    R: np.ndarray
    To: np.ndarray

    @property
    def n(self) -> int:
        return self.x.shape[0]

    def geometry(self):
        return (self.x, self.length, self.thickness, self.width, self.strike)

    def rheology(self):
        return (self.gammadot0, self.npower, self.edot_bg, self.Gm, self.Q, self.R, self.To)


def volume_grid(nx: int = 32, ny: int = 32, nz: int = 8, size: float = 4000.0,
                top: float = 20e3, origin=(0.0, -64e3), eta: float = 1e19,
                npower: float = 1.0, seed: int = 0, edot: float = 1e-15) -> VolumeModel:
    """nx (north) x ny (east) x nz (down) cubes of ``size`` m from depth
    ``top``; strike north.  eta in Pa s (stress unit MPa: gammadot0 =
    1/(2 eta 1e-6)); npower > 1 draws gammadot0 from a seeded lognormal about
    the Maxwell value.  Background strain rate: simple shear e12 = edot."""
    i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    i, j, k = i.ravel(), j.ravel(), k.ravel()
    n = i.size
    x = np.stack([origin[0] + i * size, origin[1] + j * size, top + k * size], 1).astype(float)
    one = np.ones(n)
    g0 = 1.0 / (2.0 * eta * 1e-6) * one
    if npower != 1.0:
        g0 = g0 * np.exp(np.random.default_rng(seed).normal(0.0, 0.5, n))
    eb = np.zeros((n, 6))
    eb[:, 1] = edot
    return VolumeModel(x=x, length=size * one, thickness=size * one, width=size * one,
                       strike=0.0 * one, gammadot0=g0, npower=npower * one, edot_bg=eb,
                       Gm=1.0 * one, Q=0.0 * one, R=8.314 * one, To=873.0 * one)


@dataclass
class AntiplaneModel:
    """Vertical strike-slip fault of the 2-D antiplane model
    (fortran/2d/antiplane): n segments of width W from the surface down,
    velocity-weakening between depths vw_top and vw_bottom."""
    mu: float
    x: np.ndarray          # (n, 3) upper corners; x2 = 0, x3 = depth
    width: np.ndarray
    dip: np.ndarray        # radians (the Green's function ignores it)
    Vpl: np.ndarray
    mu0: np.ndarray
    sig: np.ndarray
    a: np.ndarray
    b: np.ndarray
    L: np.ndarray
    Vo: np.ndarray
    damping: np.ndarray

    @property
    def n(self) -> int:
        return self.width.size

    def friction(self):
        """Engine.set_friction order (the rake is unused in 2-D)."""
        return (self.Vpl, np.zeros(self.n), self.mu0, self.sig, self.a, self.b, self.L, self.Vo,
                self.damping)

    def oracle_friction(self):
        return (self.Vpl, self.mu0, self.sig, self.a, self.b, self.L, self.Vo, self.damping)


@dataclass
class PlaneStrainModel(AntiplaneModel):
    """Dipping thrust of the 2-D plane-strain model (fortran/2d/planestrain):
    n segments of width W down dip from the surface; the Lame parameter
    lambda_ beside the antiplane model's fields."""
    lambda_: float = 30e3


def planestrain_fault(n: int = 200, dip_deg: float = 30.0, length: float = 60e3,
                      vw_top: float = 5e3, vw_bottom: float = 15e3, mu: float = 30e3,
                      lambda_: float = 30e3, vs: float = 3000.0, L: float = 0.01,
                      sig: float = 100.0, Vpl: float = 1e-9) -> PlaneStrainModel:
    """A thrust dipping dip_deg from the surface, length m down dip, velocity
    weakening (b = 0.014 against a = 0.010) between the depths vw_top and
    vw_bottom of the segments' centres."""
    W = length / n
    dip = np.deg2rad(dip_deg)
    k = np.arange(n)
    x = np.zeros((n, 3))
    x[:, 1] = k * W * np.cos(dip)
    x[:, 2] = k * W * np.sin(dip)
    one = np.ones(n)
    mid = x[:, 2] + W / 2 * np.sin(dip)
    b = np.where((mid > vw_top) & (mid < vw_bottom), 0.014, 0.006)
    return PlaneStrainModel(mu=mu, x=x, width=W * one, dip=dip * one, Vpl=Vpl * one,
                            mu0=0.6 * one, sig=sig * one, a=0.010 * one, b=b, L=L * one,
                            Vo=1e-6 * one, damping=mu / (2.0 * vs) * one, lambda_=lambda_)


@dataclass
class PlaneStrainViscousModel:
    """The plane-strain viscouscycles model (fortran/2d/planestrain/src/
    viscouscycles.f90): a dipping fault of segments (PlaneStrainModel) and a
    grid of strain volumes with the nonlinear Maxwell rheology
    eII = gammadot0 sII^(n-1) exp(-Q/(R To))."""
    fault: PlaneStrainModel
    vx: np.ndarray          # (nv, 3) tops; x2, x3 used
    thickness: np.ndarray
    vwidth: np.ndarray      # extent down dip
    vdip: np.ndarray        # radians
    e22: np.ndarray         # background strain rate (1/s)
    e23: np.ndarray
    e33: np.ndarray
    gammadot0: np.ndarray
    npower: np.ndarray
    Q: np.ndarray
    R: np.ndarray
    To: np.ndarray

    @property
    def n(self) -> int:
        return self.fault.n

    @property
    def nvol(self) -> int:
        return self.thickness.size

    def edot_bg(self):
        """Engine.set_rheology's background strain rate (6, nv): e22, e23,
        e33 in entries 3, 4, 5 of each volume."""
        e = np.zeros((self.nvol, 6))
        e[:, 3], e[:, 4], e[:, 5] = self.e22, self.e23, self.e33
        return e

    def rheology(self):
        """Engine.set_rheology order (Gm unused by this model)."""
        return (self.gammadot0, self.npower, self.edot_bg(), np.ones(self.nvol), self.Q, self.R,
                self.To)

    def oracle_rheology(self):
        """orc.PSSystem's rheo order."""
        return (self.gammadot0, self.npower, self.Q, self.R, self.To,
                np.stack([self.e22, self.e23, self.e33], 1).ravel())


def planestrain_viscous(n: int = 60, n2: int = 8, n3: int = 4, vsize: float = 4e3,
                        depth: float = 20e3, npower: float = 1.0, gammadot0: float = 1e-12,
                        Q: float = 0.0, To: float = 1200.0, e22: float = -1e-15,
                        e23: float = 0.0, e33: float = 1e-15, **fault) -> PlaneStrainViscousModel:
    """A thrust of n segments (planestrain_fault, 30 deg over 60 km down dip
    by default) above an n2 x n3 grid of vsize x vsize volumes (dip 90 deg:
    width down, thickness across) from depth, centred on x2 = 10 km (clear
    of the fault, which reaches that depth at x2 = 35 km); pure
    shear e22 = -e33 of 1e-15/s in the background, n = 1, gammadot0 = 1e-12,
    Q = 0 and R = 8.314 by default."""
    f = planestrain_fault(n=n, **fault)
    j2, j3 = np.meshgrid(np.arange(n2), np.arange(n3))
    nv = n2 * n3
    vx = np.zeros((nv, 3))
    vx[:, 1] = (10e3 + (j2 - (n2 - 1) / 2) * vsize).ravel()
    vx[:, 2] = (depth + j3 * vsize).ravel()
    one = np.ones(nv)
    return PlaneStrainViscousModel(fault=f, vx=vx, thickness=vsize * one, vwidth=vsize * one,
                                   vdip=np.pi / 2 * one, e22=e22 * one, e23=e23 * one,
                                   e33=e33 * one, gammadot0=gammadot0 * one,
                                   npower=npower * one, Q=Q * one, R=8.314 * one, To=To * one)


@dataclass
class AntiplaneViscousModel:
    """The 2-D antiplane viscouscycles model (fortran/2d/antiplane/src/
    viscouscycles.f90): a fault of segments (AntiplaneModel) above a grid of
    strain volumes with a power-law (Maxwell for n = 1) rheology."""
    fault: AntiplaneModel
    vx: np.ndarray          # (nv, 3) upper centres; x2, x3 used
    thickness: np.ndarray   # horizontal extent
    vwidth: np.ndarray      # depth extent
    vdip: np.ndarray        # radians
    e12: np.ndarray         # background strain rate (1/s)
    e13: np.ndarray
    Gm: np.ndarray
    gammadot0: np.ndarray
    npower: np.ndarray
    Q: np.ndarray
    R: np.ndarray
    To: np.ndarray

    @property
    def n(self) -> int:
        return self.fault.n

    @property
    def nvol(self) -> int:
        return self.thickness.size

    def edot_bg(self):
        """Engine.set_rheology's background strain rate (6, nv): e12 and e13 in
        entries 1 and 2 of each volume."""
        e = np.zeros((self.nvol, 6))
        e[:, 1], e[:, 2] = self.e12, self.e13
        return e

    def rheology(self):
        """Engine.set_rheology order."""
        return (self.gammadot0, self.npower, self.edot_bg(), self.Gm, self.Q, self.R, self.To)

    def oracle_rheology(self):
        """orc.APSystem's rheo order."""
        return (self.gammadot0, self.npower, self.e12, self.e13, self.Gm, self.Q, self.R, self.To)


def antiplane_viscous(n: int = 100, n2: int = 10, n3: int = 10, depth: float = 25e3,
                      gap: float = 10e3, vsize: float = 2e3, npower: float = 1.0,
                      gammadot0: float = 1e-12, Gm: float = 1.0, Q: float = 0.0,
                      To: float = 1200.0, e12: float = 1e-15, e13: float = 0.0,
                      **fault) -> AntiplaneViscousModel:
    """The layout of the reference's viscous1 tutorial
    (fortran/2d/antiplane/examples/tutorials/viscous1.sh) at a chosen size:
    n segments over [0, depth) and, below a gap, an n2 x n3 grid of volumes of
    vsize x vsize centred on x2 = 0 with e12 = 1e-15/s, Gm = 1,
    gammadot0 = 1e-12, n = 1, Q = 0 and R = 8.314 by default."""
    f = antiplane_fault(n=n, depth=depth, vw_top=5e3, vw_bottom=15e3, **fault)
    j2, j3 = np.meshgrid(np.arange(1, n2 + 1), np.arange(1, n3 + 1))
    nv = n2 * n3
    vx = np.zeros((nv, 3))
    vx[:, 1] = ((j2 - n2 / 2) * vsize).ravel()
    vx[:, 2] = (depth + gap + (j3 - 1) * vsize).ravel()
    one = np.ones(nv)
    return AntiplaneViscousModel(fault=f, vx=vx, thickness=vsize * one, vwidth=vsize * one,
                                 vdip=np.pi / 2 * one, e12=e12 * one, e13=e13 * one,
                                 Gm=Gm * one, gammadot0=gammadot0 * one, npower=npower * one,
                                 Q=Q * one, R=8.314 * one, To=To * one)


def antiplane_fault(n: int = 400, depth: float = 40e3, vw_top: float = 2e3,
                    vw_bottom: float = 15e3, mu: float = 30e3, vs: float = 3000.0,
                    L: float = 0.01, sig: float = 100.0, Vpl: float = 1e-9) -> AntiplaneModel:
    W = depth / n
    z = np.arange(n) * W
    x = np.zeros((n, 3))
    x[:, 2] = z
    one = np.ones(n)
    mid = z + W / 2
    b = np.where((mid > vw_top) & (mid < vw_bottom), 0.014, 0.006)
    return AntiplaneModel(mu=mu, x=x, width=W * one, dip=np.pi / 2 * one, Vpl=Vpl * one,
                          mu0=0.6 * one, sig=sig * one, a=0.010 * one, b=b, L=L * one,
                          Vo=1e-6 * one, damping=mu / (2.0 * vs) * one)
