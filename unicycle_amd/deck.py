"""The input decks of the reference's 3-D programs.

parse() reads the deck of ratestate.f90:958-1739 (`init`); parse_viscous()
the deck of viscouscycles.f90:1291-2428, which adds cuboid strain volumes
(after the triangles: the volumes, their nonlinear Maxwell rows and thermal
rows) and observed volumes (after the observed patches).  Both read section by
section, with its checks: comment lines start with '#' (getdata.f), values
are Fortran list-directed (blanks or commas, D or E exponents, quoted or bare
names), angles are in degrees and converted with the reference's DEG2RAD.

Sections, in order:
  output directory
  elastic moduli                       lambda mu
  time interval                        (s)
  number of rectangular patches        ns
    i Vpl x1 x2 x3 length width strike dip rake        (ns lines)
    number of frictional rectangular patches           (= ns)
    i mu0 sig a b L Vo G/(2Vs)                         (ns lines)
  number of triangular patches         nt
    i Vpl i1 i2 i3 rake                                (nt lines)
    number of vertices                                 nVe
    i x1 x2 x3                                         (nVe lines)
    number of frictional triangular patches            (= nt)
    i mu0 sig a b L Vo G/(2Vs)                         (nt lines)
  number of observation patches        no
    i patch                                            (no lines)
  number of observation points         np
    i name x1 x2 x3                                    (np lines)
  number of perturbations              ne
    time                                               (ne - 1 lines: the first
                                                        event is at t = 0)
The patch and friction sections of a source type are present only when its
count is positive.
"""
from __future__ import annotations

import shlex
from dataclasses import dataclass, field
from typing import ClassVar

import numpy as np

DEG2RAD = 0.01745329251994329547437168059786927  # input of ratestate.f90, as the reference


class DeckError(ValueError):
    """An input deck that the reference would reject (its STOPs)."""


@dataclass
class Deck:
    wdir: str
    lambda_: float
    mu: float
    interval: float
    # rectangles: upper corners (ns, 3) north/east/down, sizes, radians
    rect_x: np.ndarray = field(default_factory=lambda: np.zeros((0, 3)))
    rect_length: np.ndarray = field(default_factory=lambda: np.zeros(0))
    rect_width: np.ndarray = field(default_factory=lambda: np.zeros(0))
    rect_strike: np.ndarray = field(default_factory=lambda: np.zeros(0))
    rect_dip: np.ndarray = field(default_factory=lambda: np.zeros(0))
    # triangles: 1-based vertex indices, vertices (nVe, 3)
    tri_i1: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    tri_i2: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    tri_i3: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    vertices: np.ndarray = field(default_factory=lambda: np.zeros((0, 3)))
    # per element (rectangles then triangles): Vpl, rake (rad), mu0, sig, a, b, L, Vo, damping
    friction: dict = field(default_factory=dict)
    observation_states: list = field(default_factory=list)   # 1-based element indices
    observation_points: list = field(default_factory=list)   # (name, x1, x2, x3)
    events: list = field(default_factory=list)               # times, the first at 0

    @property
    def n_rect(self) -> int:
        return self.rect_length.size

    @property
    def n_tri(self) -> int:
        return self.tri_i1.size

    @property
    def nu(self) -> float:
        return self.lambda_ / 2.0 / (self.lambda_ + self.mu)

    def friction_arrays(self):
        """(Vpl, rake, mu0, sig, a, b, L, Vo, damping) in element order, the
        argument order of Engine.set_friction."""
        keys = ("Vpl", "rake", "mu0", "sig", "a", "b", "L", "Vo", "damping")
        return tuple(np.asarray(self.friction[k], dtype=np.float64) for k in keys)


def _fortran_tokens(line: str):
    """Values of one list-directed record: blanks and commas separate, a
    slash ends the record, quotes delimit strings."""
    cut = line
    # a slash outside quotes ends the record
    quote = None
    for i, ch in enumerate(line):
        if ch in "'\"":
            quote = None if quote == ch else (ch if quote is None else quote)
        elif ch == "/" and quote is None:
            cut = line[:i]
            break
    lex = shlex.shlex(cut.replace(",", " "), posix=True)
    lex.whitespace_split = True
    lex.commenters = ""
    return list(lex)


def _real(tok: str) -> float:
    try:
        return float(tok.replace("d", "e").replace("D", "e"))
    except ValueError as exc:
        raise DeckError(f"not a number: {tok!r}") from exc


def _int(tok: str) -> int:
    try:
        return int(tok)
    except ValueError as exc:
        raise DeckError(f"not an integer: {tok!r}") from exc


class _Lines:
    """getdata.f: skips lines whose first non-blank character is '#'
    (and blank lines)."""

    def __init__(self, text: str):
        self.lines = text.splitlines()
        self.k = 0

    def next(self, what: str) -> str:
        while self.k < len(self.lines):
            line = self.lines[self.k]
            self.k += 1
            s = line.strip()
            if s and not s.startswith("#"):
                return line
        raise DeckError(f"input file ends before the {what}")

    def values(self, what: str, n: int):
        toks = _fortran_tokens(self.next(what))
        if len(toks) < n:
            raise DeckError(f"{what}: {n} values expected, got {len(toks)}")
        return toks


def _header(L: _Lines, cls):
    """output directory, elastic moduli and interval (ratestate.f90:1073-1127,
    viscouscycles.f90:1410-1465: the same checks)"""
    wdir = L.next("output directory").strip()
    t = L.values("elastic moduli", 2)
    lam, mu = _real(t[0]), _real(t[1])
    if 0 > mu:
        raise DeckError("shear modulus must be positive")
    nu = lam / 2.0 / (lam + mu)
    if -1.0 > nu:
        raise DeckError("Poisson's ratio must be greater than -1")
    if 0.5 < nu:
        raise DeckError("Poisson's ratio must be lower than 0.5")
    interval = _real(L.values("time interval", 1)[0])
    if interval <= 0.0:
        raise DeckError("simulation time must be positive")
    return cls(wdir=wdir, lambda_=lam, mu=mu, interval=interval)


def _patches(L: _Lines, d: Deck):
    """rectangular and triangular patches with their friction
    (ratestate.f90:1128-1393; viscouscycles.f90:1466-1729 reads the same)"""
    fr = {k: [] for k in ("Vpl", "rake", "mu0", "sig", "a", "b", "L", "Vo", "damping")}

    def friction_rows(n, kind):
        dummy = _int(L.values(f"number of frictional {kind} patches", 1)[0])
        if dummy != n:
            raise DeckError(f"all {kind} patches require frictional properties")
        for k in range(1, n + 1):
            t = L.values(f"{kind} friction row {k}", 8)
            if _int(t[0]) != k:
                raise DeckError(f"invalid friction property definition for {kind} patch: "
                                "unexpected index")
            for key, tok in zip(("mu0", "sig", "a", "b", "L", "Vo", "damping"), t[1:8]):
                fr[key].append(_real(tok))

    ns = _int(L.values("number of rectangular patches", 1)[0])
    if ns > 0:
        rows = []
        for k in range(1, ns + 1):
            t = L.values(f"rectangular patch {k}", 10)
            if _int(t[0]) != k:
                raise DeckError("invalid rectangular patch definition: unexpected index")
            v = [_real(x) for x in t[1:10]]
            if max(v[4], v[5]) <= 0.0:
                raise DeckError("patch length and width must be positive")
            rows.append(v)
        r = np.array(rows)
        d.rect_x = r[:, 1:4].copy()
        d.rect_length, d.rect_width = r[:, 4].copy(), r[:, 5].copy()
        d.rect_strike, d.rect_dip = r[:, 6] * DEG2RAD, r[:, 7] * DEG2RAD
        fr["Vpl"] += list(r[:, 0])
        fr["rake"] += list(r[:, 8] * DEG2RAD)
        friction_rows(ns, "rectangular")
    nt = _int(L.values("number of triangular patches", 1)[0])
    if nt > 0:
        i1, i2, i3, vpl, rake = [], [], [], [], []
        for k in range(1, nt + 1):
            t = L.values(f"triangular patch {k}", 6)
            idx = [_int(x) for x in t[2:5]]
            if min(idx) < 0:
                raise DeckError("negative vertex index")
            if _int(t[0]) != k:
                raise DeckError("invalid triangular patch definition: unexpected index")
            vpl.append(_real(t[1]))
            i1.append(idx[0])
            i2.append(idx[1])
            i3.append(idx[2])
            rake.append(_real(t[5]) * DEG2RAD)
        nve = _int(L.values("number of vertices", 1)[0])
        if max(max(i1), max(i2), max(i3)) > nve:
            raise DeckError("not enough vertices")
        verts = []
        for k in range(1, nve + 1):
            t = L.values(f"vertex {k}", 4)
            if _int(t[0]) != k:
                raise DeckError("invalid vertex definition: unexpected index")
            verts.append([_real(x) for x in t[1:4]])
        d.tri_i1, d.tri_i2, d.tri_i3 = (np.array(a, np.int32) for a in (i1, i2, i3))
        d.vertices = np.array(verts, dtype=np.float64).reshape(-1, 3)
        fr["Vpl"] += vpl
        fr["rake"] += rake
        friction_rows(nt, "triangular")
    d.friction = fr


def _indices(L: _Lines, what: str):
    """a count, then (k, index) rows"""
    n = _int(L.values(f"number of {what}s", 1)[0])
    out = []
    for k in range(1, n + 1):
        t = L.values(f"{what} {k}", 2)
        if _int(t[0]) != k:
            raise DeckError(f"{what}s: unexpected index")
        out.append(_int(t[1]))
    return out


def _points_and_events(L: _Lines, d: Deck):
    """observation points (ratestate.f90:1421-1456, viscouscycles.f90:1986-2021)
    and perturbations (ratestate.f90:1457-1490, viscouscycles.f90:2022-2055)"""
    npnt = _int(L.values("number of observation points", 1)[0])
    for k in range(1, npnt + 1):
        t = L.values(f"observation point {k}", 5)
        if _int(t[0]) != k:
            raise DeckError("observation points: unexpected index")
        # the name is CHARACTER(LEN=10) (types.f90:88): a list-directed read
        # keeps its leftmost 10 characters
        d.observation_points.append((t[1][:10], _real(t[2]), _real(t[3]), _real(t[4])))
    # the first event is at t = 0 and has no line
    ne = _int(L.values("number of perturbations", 1)[0])
    for k in range(1, ne + 1):
        if k == 1:
            d.events.append(0.0)
            continue
        tk = _real(L.values(f"event {k}", 1)[0])
        if tk <= d.events[-1]:
            raise DeckError("timing of perturbations must increase")
        d.events.append(tk)


def parse(text: str) -> Deck:
    """Parse a ratestate input deck; raises DeckError where the reference stops."""
    L = _Lines(text)
    d = _header(L, Deck)
    _patches(L, d)
    # observation patches (ratestate.f90:1394-1420)
    d.observation_states = _indices(L, "observation patch")
    _points_and_events(L, d)
    if (d.n_rect == 0 and d.n_tri == 0) or d.interval <= 0.0:
        raise DeckError("nothing to do")
    nel = d.n_rect + d.n_tri
    for j in d.observation_states:
        if not 1 <= j <= nel:
            raise DeckError(f"observation patch {j} is not an element")
    return d


@dataclass
class ViscousDeck(Deck):
    """The input deck of viscouscycles.f90 (1291-2428): a ratestate deck with
    cuboid strain volumes, their nonlinear Maxwell and thermal properties and
    observed volumes between the patches and the observation points."""
    vol_x: np.ndarray = field(default_factory=lambda: np.zeros((0, 3)))
    vol_length: np.ndarray = field(default_factory=lambda: np.zeros(0))
    vol_width: np.ndarray = field(default_factory=lambda: np.zeros(0))
    vol_thickness: np.ndarray = field(default_factory=lambda: np.zeros(0))
    vol_strike: np.ndarray = field(default_factory=lambda: np.zeros(0))
    vol_dip: np.ndarray = field(default_factory=lambda: np.zeros(0))
    vol_ebg: np.ndarray = field(default_factory=lambda: np.zeros((0, 6)))  # e11..e33 loading
    n_maxwell: int = 0
    maxwell: dict = field(default_factory=dict)   # Gm, gammadot0, npower, Q, R (per volume)
    thermal: dict = field(default_factory=dict)   # rhoc, To (per volume), or empty
    observation_volumes: list = field(default_factory=list)   # 1-based volume indices

    @property
    def n_vol(self) -> int:
        return self.vol_length.size


def parse_viscous(text: str) -> ViscousDeck:
    """Parse a viscouscycles input deck; raises DeckError where the reference
    stops (viscouscycles.f90:1410-2065)."""
    L = _Lines(text)
    d = _header(L, ViscousDeck)
    _patches(L, d)
    # cuboid strain volumes (viscouscycles.f90:1730-1820): loading strain rate,
    # upper corner, length, width, thickness, strike, dip (must be 90)
    nv = _int(L.values("number of cuboid strain volumes", 1)[0])
    if nv > 0:
        rows = []
        for k in range(1, nv + 1):
            t = L.values(f"strain volume {k}", 15)
            v = [_real(x) for x in t[1:15]]
            if 90.0 != v[13]:
                raise DeckError("strain volume must be vertical (dip=90)")
            if 0 > v[8]:
                raise DeckError("strain volume depth must be positive")
            if _int(t[0]) != k:
                raise DeckError("invalid strain volume definition: unexpected index")
            if max(v[9], v[10], v[11]) <= 0.0:
                raise DeckError("strain volume dimension must be positive")
            rows.append(v)
        r = np.array(rows)
        d.vol_ebg = r[:, 0:6].copy()
        d.vol_x = r[:, 6:9].copy()
        d.vol_length, d.vol_width, d.vol_thickness = r[:, 9].copy(), r[:, 10].copy(), r[:, 11].copy()
        d.vol_strike, d.vol_dip = r[:, 12] * DEG2RAD, r[:, 13] * DEG2RAD
        # nonlinear Maxwell properties (:1821-1881), for none or all volumes
        d.n_maxwell = _int(L.values("number of nonlinear Maxwell strain volumes", 1)[0])
        if d.n_maxwell != 0:
            if d.n_maxwell != nv:
                raise DeckError("nonlinear Maxwell properties are for none or all strain volumes")
            keys = ("Gm", "gammadot0", "npower", "Q", "R")
            d.maxwell = {k: [] for k in keys}
            for k in range(1, nv + 1):
                t = L.values(f"Maxwell row {k}", 6)
                v = [_real(x) for x in t[1:6]]
                if 0 >= v[0]:
                    raise DeckError("sig0 must be positive")
                if 0 >= v[4]:
                    raise DeckError("R must be positive")
                if _int(t[0]) != k:
                    raise DeckError("invalid property definition for strain volume: "
                                    "unexpected index")
                for key, x in zip(keys, v):
                    d.maxwell[key].append(x)
        # thermal properties (:1882-1930), for none or all volumes
        nth = _int(L.values("number of thermal strain volumes", 1)[0])
        if nth != 0:
            if nth != nv:
                raise DeckError("thermal properties are for none or all strain volumes")
            d.thermal = {"rhoc": [], "To": []}
            for k in range(1, nv + 1):
                t = L.values(f"thermal row {k}", 3)
                rhoc, To = _real(t[1]), _real(t[2])
                if 0 >= To:
                    raise DeckError("To must be positive")
                if _int(t[0]) != k:
                    raise DeckError("invalid property definition for strain volume: "
                                    "unexpected index")
                d.thermal["rhoc"].append(rhoc)
                d.thermal["To"].append(To)
    d.observation_states = _indices(L, "observation patch")
    d.observation_volumes = _indices(L, "observation volume")
    _points_and_events(L, d)
    if (d.n_rect == 0 and d.n_tri == 0 and d.n_vol == 0) or d.interval <= 0.0:
        raise DeckError("nothing to do")
    nel = d.n_rect + d.n_tri
    for j in d.observation_states:
        if not 1 <= j <= nel:
            raise DeckError(f"observation patch {j} is not an element")
    for j in d.observation_volumes:
        if not 1 <= j <= d.n_vol:
            raise DeckError(f"observation volume {j} is not a volume")
    return d


def read(path: str) -> Deck:
    with open(path) as f:
        return parse(f.read())


def format_rectangles(wdir, lambda_, mu, interval, x, length, width, strike, dip, Vpl, rake,
                      mu0, sig, a, b, L, Vo, damping, observe=(), points=(), events=1) -> str:
    """An input deck of rectangular patches (angles in radians here, written in
    degrees), in the layout parse() reads."""
    R2D = 1.0 / DEG2RAD
    out = ["# output directory", str(wdir), "# elastic moduli", f"{lambda_!r} {mu!r}",
           "# time interval", f"{interval!r}", "# number of rectangular patches", str(len(length)),
           "# n Vpl x1 x2 x3 length width strike dip rake"]
    for k in range(len(length)):
        out.append(" ".join([str(k + 1)] + [repr(float(v)) for v in (
            Vpl[k], x[k][0], x[k][1], x[k][2], length[k], width[k], strike[k] * R2D,
            dip[k] * R2D, rake[k] * R2D)]))
    out += ["# number of frictional rectangular patches", str(len(length)),
            "# n mu0 sig a b L Vo G/(2Vs)"]
    for k in range(len(length)):
        out.append(" ".join([str(k + 1)] + [repr(float(v)) for v in (
            mu0[k], sig[k], a[k], b[k], L[k], Vo[k], damping[k])]))
    out += ["# number of triangular patches", "0", "# number of observation patches",
            str(len(observe))]
    out += [f"{k + 1} {j}" for k, j in enumerate(observe)]
    out += ["# number of observation points", str(len(points))]
    out += [f"{k + 1} {n} {x1!r} {x2!r} {x3!r}" for k, (n, x1, x2, x3) in enumerate(points)]
    out += ["# number of perturbations", str(events)]
    out += [f"{float(k)!r}" for k in range(1, events)]
    return "\n".join(out) + "\n"


@dataclass
class AntiplaneDeck:
    """The input deck of fortran/2d/antiplane/src/ratestate.f90 (init,
    958-1230): shear modulus, interval, vertical segments (Vpl, x2, x3,
    width, dip) with friction, observed patches, observation points (name,
    x2, x3).  The subclasses are the decks of the other 2-D programs; the
    volume fields are used by the viscouscycles ones."""
    wdir: str
    mu: float
    interval: float
    x: np.ndarray = field(default_factory=lambda: np.zeros((0, 3)))
    width: np.ndarray = field(default_factory=lambda: np.zeros(0))
    dip: np.ndarray = field(default_factory=lambda: np.zeros(0))
    friction: dict = field(default_factory=dict)
    observation_states: list = field(default_factory=list)
    observation_points: list = field(default_factory=list)   # (name, x1 = 0, x2, x3)
    # strain volumes: tops (x2, x3 used), width, thickness, dip; the
    # background strain rate in the 3-D component order (e11, e12, e13, e22,
    # e23, e33: the antiplane programs fill e12, e13, the plane-strain ones
    # e22, e23, e33); Maxwell rows (Gm, gammadot0, npower, Q, R) and thermal
    # rows (rhoc, To), each for none or all volumes
    vol_x: np.ndarray = field(default_factory=lambda: np.zeros((0, 3)))
    vol_width: np.ndarray = field(default_factory=lambda: np.zeros(0))
    vol_thickness: np.ndarray = field(default_factory=lambda: np.zeros(0))
    vol_dip: np.ndarray = field(default_factory=lambda: np.zeros(0))
    vol_ebg: np.ndarray = field(default_factory=lambda: np.zeros((0, 6)))
    n_maxwell: int = 0
    maxwell: dict = field(default_factory=dict)
    thermal: dict = field(default_factory=dict)
    observation_volumes: list = field(default_factory=list)   # 1-based volume indices
    planestrain: ClassVar[bool] = False
    viscous: ClassVar[bool] = False

    @property
    def n(self) -> int:
        return self.width.size

    @property
    def n_vol(self) -> int:
        return self.vol_width.size

    def friction_arrays(self):
        """Engine.set_friction order (the rake is unused in 2-D)."""
        keys = ("Vpl", "rake", "mu0", "sig", "a", "b", "L", "Vo", "damping")
        fr = dict(self.friction, rake=[0.0] * self.n)
        return tuple(np.asarray(fr[k], dtype=np.float64) for k in keys)


@dataclass
class AntiplaneViscousDeck(AntiplaneDeck):
    """The deck of fortran/2d/antiplane/src/viscouscycles.f90 (init,
    1124-1760): the ratestate deck with strain volumes (e12, e13, x2, x3,
    width, thickness, dip = 90), their Maxwell rows (Gm, gammadot0, n, Q, R)
    and thermal rows, and observed volumes after the observed patches."""
    viscous: ClassVar[bool] = True


@dataclass
class PlaneStrainDeck(AntiplaneDeck):
    """The deck of fortran/2d/planestrain/src/ratestate.f90 (init, 895-1374):
    elastic moduli (mu, lambda), interval, dipping segments (Vpl, x2, x3,
    width, dip) with friction rows that begin with tau0, observed patches,
    observation points."""
    lambda_: float = 0.0
    tau0: np.ndarray = field(default_factory=lambda: np.zeros(0))
    planestrain: ClassVar[bool] = True


@dataclass
class PlaneStrainViscousDeck(PlaneStrainDeck):
    """The deck of fortran/2d/planestrain/src/viscouscycles.f90 (init,
    1231-2040): the plane-strain ratestate deck with strain volumes (e22,
    e23, e33, x2, x3, width, thickness, dip), their nonlinear Maxwell rows
    (s22, s23, s33 of the initial stress, gammadot0, n, Q, R) and thermal
    rows, and observed volumes after the observed patches."""
    s0: np.ndarray = field(default_factory=lambda: np.zeros((0, 3)))
    viscous: ClassVar[bool] = True


def _volumes_2d(L: _Lines, d: AntiplaneDeck):
    """strain volumes, Maxwell rows and thermal rows of the 2-D viscouscycles
    decks (fortran/2d/antiplane/src/viscouscycles.f90:1410-1590,
    fortran/2d/planestrain/src/viscouscycles.f90:1522-1688)"""
    ps = d.planestrain
    nv = _int(L.values("number of strain volumes", 1)[0])
    if nv <= 0:
        return
    ne = 3 if ps else 2                      # background strain-rate components
    rows = []
    for k in range(1, nv + 1):
        t = L.values(f"strain volume {k}", 6 + ne)
        v = [_real(x) for x in t[1:6 + ne]]
        if not ps and 90.0 != v[6]:          # antiplane viscouscycles.f90:1457-1462
            raise DeckError("strain volume must be vertical (dip=90)")
        if 0 > v[ne + 1]:
            raise DeckError("depth must be positive")
        if _int(t[0]) != k:
            raise DeckError("invalid strain volume definition: unexpected index")
        if min(v[ne + 2], v[ne + 3]) <= 0.0:
            raise DeckError("strain volume dimension must be positive")
        rows.append(v)
    r = np.array(rows)
    d.vol_ebg = np.zeros((nv, 6))
    if ps:
        d.vol_ebg[:, 3:6] = r[:, 0:3]
    else:
        d.vol_ebg[:, 1:3] = r[:, 0:2]
    d.vol_x = np.stack([np.zeros(nv), r[:, ne], r[:, ne + 1]], 1)
    d.vol_width, d.vol_thickness = r[:, ne + 2].copy(), r[:, ne + 3].copy()
    d.vol_dip = r[:, ne + 4] * DEG2RAD
    d.n_maxwell = _int(L.values("number of nonlinear Maxwell strain volumes", 1)[0])
    if d.n_maxwell != 0:
        if d.n_maxwell != nv:
            raise DeckError("nonlinear Maxwell properties are for none or all strain volumes")
        keys = ("gammadot0", "npower", "Q", "R")
        d.maxwell = {k: [] for k in ("Gm",) + keys}
        s0 = []
        for k in range(1, nv + 1):
            if ps:                           # i s22 s23 s33 gammadot0 n Q R
                t = L.values(f"Maxwell row {k}", 8)
                v = [_real(x) for x in t[1:8]]
                s0.append(v[0:3])
                gm, rest = 1.0, v[3:7]
            else:                            # i Gm gammadot0 n Q R
                t = L.values(f"Maxwell row {k}", 6)
                v = [_real(x) for x in t[1:6]]
                gm, rest = v[0], v[1:5]
                if 0 >= gm:
                    raise DeckError("sig0 must be positive")
            if 0 >= rest[3]:
                raise DeckError("R must be positive")
            if _int(t[0]) != k:
                raise DeckError("invalid property definition for strain volume: unexpected index")
            d.maxwell["Gm"].append(gm)
            for key, x in zip(keys, rest):
                d.maxwell[key].append(x)
        if ps:
            d.s0 = np.array(s0)
    nth = _int(L.values("number of thermal strain volumes", 1)[0])
    if nth != 0:
        if nth != nv:
            raise DeckError("thermal properties are for none or all strain volumes")
        d.thermal = {"rhoc": [], "To": []}
        for k in range(1, nv + 1):
            t = L.values(f"thermal row {k}", 3)
            rhoc, To = _real(t[1]), _real(t[2])
            if 0 >= To:
                raise DeckError("To must be positive")
            if _int(t[0]) != k:
                raise DeckError("invalid property definition for strain volume: unexpected index")
            d.thermal["rhoc"].append(rhoc)
            d.thermal["To"].append(To)


def _parse_2d(text: str, cls):
    """A deck of one of the 2-D programs; raises DeckError where the
    reference stops (the `init` routines of fortran/2d/antiplane/src/
    ratestate.f90:1003-1230, viscouscycles.f90:1250-1760 and of
    fortran/2d/planestrain/src/ratestate.f90:1041-1240, viscouscycles.f90:
    1386-1790)."""
    L = _Lines(text)
    ps = cls.planestrain
    wdir = L.next("output directory").strip()
    if ps:
        t = L.values("elastic moduli", 2)
        mu, lam = _real(t[0]), _real(t[1])  # mu first in the plane-strain decks
    else:
        mu = _real(L.values("shear modulus", 1)[0])
    if 0 > mu:
        raise DeckError("shear modulus must be positive")
    interval = _real(L.values("time interval", 1)[0])
    if interval <= 0.0:
        raise DeckError("simulation time must be positive")
    d = cls(wdir=wdir, mu=mu, interval=interval)
    if ps:
        d.lambda_ = lam
    ns = _int(L.values("number of patches", 1)[0])
    fkeys = (("tau0",) if ps else ()) + ("mu0", "sig", "a", "b", "L", "Vo", "damping")
    fr = {k: [] for k in ("Vpl",) + fkeys}
    if ns > 0:
        rows = []
        for k in range(1, ns + 1):
            t = L.values(f"patch {k}", 6)
            if _int(t[0]) != k:
                raise DeckError("invalid patch definition: unexpected index")
            v = [_real(x) for x in t[1:6]]
            if v[3] <= 0.0:
                raise DeckError("patch width must be positive")
            rows.append(v)
        r = np.array(rows)
        d.x = np.stack([np.zeros(ns), r[:, 1], r[:, 2]], 1)
        d.width, d.dip = r[:, 3].copy(), r[:, 4] * DEG2RAD
        fr["Vpl"] = list(r[:, 0])
        dummy = _int(L.values("number of frictional patches", 1)[0])
        if dummy != ns:
            raise DeckError("all patches require frictional properties")
        for k in range(1, ns + 1):
            t = L.values(f"friction row {k}", 1 + len(fkeys))
            if _int(t[0]) != k:
                raise DeckError("invalid friction property definition: unexpected index")
            for key, tok in zip(fkeys, t[1:1 + len(fkeys)]):
                fr[key].append(_real(tok))
    if ps:
        d.tau0 = np.asarray(fr.pop("tau0"), dtype=np.float64)
    d.friction = fr
    if cls.viscous:
        _volumes_2d(L, d)
    d.observation_states = _indices(L, "observation patch")
    if cls.viscous:
        d.observation_volumes = _indices(L, "observation volume")
    npnt = _int(L.values("number of observation points", 1)[0])
    for k in range(1, npnt + 1):
        t = L.values(f"observation point {k}", 4)
        if _int(t[0]) != k:
            raise DeckError("observation points: unexpected index")
        # the along-strike coordinate is zero in 2-D (ratestate.f90:1193-1194)
        d.observation_points.append((t[1][:10], 0.0, _real(t[2]), _real(t[3])))
    # the references test (ns = 0 .AND. interval <= 0), which the interval
    # check excludes; a deck with no element has nothing to integrate here
    if d.n + d.n_vol == 0:
        raise DeckError("nothing to do")
    for j in d.observation_states:
        if not 1 <= j <= d.n:
            raise DeckError(f"observation patch {j} is not an element")
    for j in d.observation_volumes:
        if not 1 <= j <= d.n_vol:
            raise DeckError(f"observation volume {j} is not a volume")
    return d


def parse_antiplane(text: str) -> AntiplaneDeck:
    """Parse a 2-D antiplane deck (fortran/2d/antiplane/src/ratestate.f90)."""
    return _parse_2d(text, AntiplaneDeck)


def parse_antiplane_viscous(text: str) -> AntiplaneViscousDeck:
    """Parse a 2-D antiplane viscouscycles deck."""
    return _parse_2d(text, AntiplaneViscousDeck)


def parse_planestrain(text: str) -> PlaneStrainDeck:
    """Parse a 2-D plane-strain ratestate deck."""
    return _parse_2d(text, PlaneStrainDeck)


def parse_planestrain_viscous(text: str) -> PlaneStrainViscousDeck:
    """Parse a 2-D plane-strain viscouscycles deck."""
    return _parse_2d(text, PlaneStrainViscousDeck)


def format_antiplane(wdir, model, interval, observe=(), points=()) -> str:
    """A 2-D antiplane deck of an AntiplaneModel (configs) in the layout
    parse_antiplane() reads."""
    R2D = 1.0 / DEG2RAD
    f = lambda v: repr(float(v))  # noqa: E731
    out = ["# output directory", str(wdir), "# shear modulus", f(model.mu), "# time interval",
           f(interval), "# number of patches", str(model.n), "# n Vpl x2 x3 width dip"]
    for k in range(model.n):
        out.append(" ".join([str(k + 1)] + [f(v) for v in (
            model.Vpl[k], model.x[k][1], model.x[k][2], model.width[k], model.dip[k] * R2D)]))
    out += ["# number of frictional patches", str(model.n)]
    for k in range(model.n):
        out.append(" ".join([str(k + 1)] + [f(v) for v in (
            model.mu0[k], model.sig[k], model.a[k], model.b[k], model.L[k], model.Vo[k],
            model.damping[k])]))
    out += ["# number of observation patches", str(len(observe))]
    out += [f"{k + 1} {j}" for k, j in enumerate(observe)]
    out += ["# number of observation points", str(len(points))]
    out += [f"{k + 1} {n} {f(x2)} {f(x3)}" for k, (n, x2, x3) in enumerate(points)]
    return "\n".join(out) + "\n"


def format_2d(wdir, fault, interval, observe=(), points=(), vol=None, observe_vol=(),
              tau0=None, s0=None, maxwell=True, thermal=True) -> str:
    """A deck of one of the 2-D programs in the layout _parse_2d() reads:
    fault an AntiplaneModel or a PlaneStrainModel (configs: the plane-strain
    decks carry mu, lambda and tau0 per patch), vol an AntiplaneViscousModel
    or PlaneStrainViscousModel for the viscouscycles decks (None: a
    ratestate deck)."""
    R2D = 1.0 / DEG2RAD
    f = lambda v: repr(float(v))  # noqa: E731
    ps = hasattr(fault, "lambda_")
    out = ["# output directory", str(wdir)]
    out += (["# elastic moduli", f"{f(fault.mu)} {f(fault.lambda_)}"] if ps
            else ["# shear modulus", f(fault.mu)])
    out += ["# time interval", f(interval), "# number of patches", str(fault.n),
            "# n Vpl x2 x3 width dip"]
    for k in range(fault.n):
        out.append(" ".join([str(k + 1)] + [f(v) for v in (
            fault.Vpl[k], fault.x[k][1], fault.x[k][2], fault.width[k], fault.dip[k] * R2D)]))
    if fault.n:
        out += ["# number of frictional patches", str(fault.n)]
    for k in range(fault.n):
        t0 = [] if not ps else [0.0 if tau0 is None else tau0[k]]
        out.append(" ".join([str(k + 1)] + [f(v) for v in t0 + [
            fault.mu0[k], fault.sig[k], fault.a[k], fault.b[k], fault.L[k], fault.Vo[k],
            fault.damping[k]]]))
    if vol is not None:
        nv = vol.nvol
        out += ["# number of strain volumes", str(nv)]
        for k in range(nv):
            e = [vol.e22[k], vol.e23[k], vol.e33[k]] if ps else [vol.e12[k], vol.e13[k]]
            out.append(" ".join([str(k + 1)] + [f(v) for v in e + [
                vol.vx[k][1], vol.vx[k][2], vol.vwidth[k], vol.thickness[k], vol.vdip[k] * R2D]]))
        out += ["# number of nonlinear Maxwell strain volumes", str(nv if maxwell else 0)]
        for k in range(nv if maxwell else 0):
            head = (list(s0[k]) if s0 is not None else [0.0, 0.0, 0.0]) if ps else [vol.Gm[k]]
            out.append(" ".join([str(k + 1)] + [f(v) for v in head + [
                vol.gammadot0[k], vol.npower[k], vol.Q[k], vol.R[k]]]))
        out += ["# number of thermal strain volumes", str(nv if thermal else 0)]
        for k in range(nv if thermal else 0):
            out.append(f"{k + 1} {f(3.3)} {f(vol.To[k])}")
    out += ["# number of observation patches", str(len(observe))]
    out += [f"{k + 1} {j}" for k, j in enumerate(observe)]
    if vol is not None:
        out += ["# number of observation volumes", str(len(observe_vol))]
        out += [f"{k + 1} {j}" for k, j in enumerate(observe_vol)]
    out += ["# number of observation points", str(len(points))]
    out += [f"{k + 1} {n} {f(x2)} {f(x3)}" for k, (n, x2, x3) in enumerate(points)]
    return "\n".join(out) + "\n"


@dataclass
class PointDeck:
    """The input deck of fortran/0d/src/ratestate.f90 (init, 292-476): one
    circular patch (a spring-slider): output directory, shear modulus,
    interval, (Vpl, radius), (mu0, sig, a, b, L, Vo, G/(2Vs))."""
    wdir: str
    mu: float
    interval: float
    Vpl: float
    radius: float
    mu0: float
    sig: float
    a: float
    b: float
    L: float
    Vo: float
    damping: float


def parse_point(text: str) -> PointDeck:
    """Parse a 0-D deck; raises DeckError where the reference stops
    (fortran/0d/src/ratestate.f90:386-475)."""
    L = _Lines(text)
    wdir = L.next("output directory").strip()
    mu = _real(L.values("shear modulus", 1)[0])
    if 0 > mu:
        raise DeckError("shear modulus must be positive")
    interval = _real(L.values("time interval", 1)[0])
    if interval <= 0.0:
        raise DeckError("simulation time must be positive")
    g = [_real(t) for t in L.values("Vpl and radius", 2)[:2]]
    f = [_real(t) for t in L.values("friction properties", 7)[:7]]
    return PointDeck(wdir, mu, interval, g[0], g[1], *f)


def format_viscous(wdir, lambda_, mu, interval, fault=None, vol=None, maxwell=True,
                   thermal=True, observe=(), observe_vol=(), points=(), events=1) -> str:
    """An input deck of viscouscycles.f90 in the layout parse_viscous() reads:
    rectangles of a FaultModel (configs) and the cuboids of a VolumeModel with
    its Maxwell (and thermal) rows; angles written in degrees."""
    R2D = 1.0 / DEG2RAD
    f = lambda v: repr(float(v))  # noqa: E731
    out = ["# output directory", str(wdir), "# elastic moduli", f"{f(lambda_)} {f(mu)}",
           "# time interval", f(interval), "# number of rectangular patches"]
    if fault is not None and fault.n:
        Vpl, rake, mu0, sig, a, b, L, Vo, damping = fault.friction()
        out.append(str(fault.n))
        for k in range(fault.n):
            out.append(" ".join([str(k + 1)] + [f(v) for v in (
                Vpl[k], *fault.x[k], fault.length[k], fault.width[k], fault.strike[k] * R2D,
                fault.dip[k] * R2D, rake[k] * R2D)]))
        out += ["# number of frictional rectangular patches", str(fault.n)]
        for k in range(fault.n):
            out.append(" ".join([str(k + 1)] + [f(v) for v in (
                mu0[k], sig[k], a[k], b[k], L[k], Vo[k], damping[k])]))
    else:
        out.append("0")
    out += ["# number of triangular patches", "0", "# number of cuboid strain volumes"]
    if vol is not None and vol.n:
        out.append(str(vol.n))
        out.append("# n e11 e12 e13 e22 e23 e33 x1 x2 x3 length width thickness strike dip")
        for k in range(vol.n):
            out.append(" ".join([str(k + 1)] + [f(v) for v in (
                *vol.edot_bg[k], *vol.x[k], vol.length[k], vol.width[k], vol.thickness[k],
                vol.strike[k] * R2D, 90.0)]))
        out += ["# number of nonlinear Maxwell strain volumes", str(vol.n if maxwell else 0)]
        if maxwell:
            for k in range(vol.n):
                out.append(" ".join([str(k + 1)] + [f(v) for v in (
                    vol.Gm[k], vol.gammadot0[k], vol.npower[k], vol.Q[k], vol.R[k])]))
        out += ["# number of thermal strain volumes", str(vol.n if thermal else 0)]
        if thermal:
            out += [f"{k + 1} 1.0 {f(vol.To[k])}" for k in range(vol.n)]
    else:
        out.append("0")
    out += ["# number of observation patches", str(len(observe))]
    out += [f"{k + 1} {j}" for k, j in enumerate(observe)]
    out += ["# number of observation volumes", str(len(observe_vol))]
    out += [f"{k + 1} {j}" for k, j in enumerate(observe_vol)]
    out += ["# number of observation points", str(len(points))]
    out += [f"{k + 1} {n} {f(x1)} {f(x2)} {f(x3)}" for k, (n, x1, x2, x3) in enumerate(points)]
    out += ["# number of perturbations", str(events)]
    out += [f(k) for k in range(1, events)]
    return "\n".join(out) + "\n"


def _es(x: float, digits: int, ewidth: int) -> str:
    """Fortran ES(w).(digits)E(ewidth) of one value, without padding."""
    m = f"{x:.{digits}E}"
    mant, exp = m.split("E")
    e = int(exp)
    return f"{mant}E{'-' if e < 0 else '+'}{abs(e):0{ewidth}d}"


def es_field(x: float, width: int, digits: int, ewidth: int) -> str:
    return _es(float(x), digits, ewidth).rjust(width)
