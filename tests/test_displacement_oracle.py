"""CPU pinning of the oracle's displacement kernels (buildO, greens.f90:302-462):
computeDisplacementOkada92 (okada92.f90:230-298), computeDisplacementNikkhoo15
(nikkhoo15.f90:324-367 over DispInHalfSpace.c) and the strain-volume
displacement (strainvolume.f90:215-268, pinned in test_oracle.py).

Pins: the reference's own routines (golden fixture tests/golden/
displacement.npz from oracle/_ref, and the live checker when built); the
displacement jump across a dislocation equals its slip vector (sign, frame
and component order); a rectangle equals its two triangles (two independent
solutions); far-field decay; the row/column layout of O and the split into
Of and Ol with the reference's assignment of the volume's third component.
"""
import ctypes as C
import os

import numpy as np
import pytest

import orc
from unicycle_amd import configs

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "displacement.npz")
ORC = orc.oracle()
D = C.c_double
MU = LAM = 30e3


def okada(x, src, slip, lib=ORC, fn="orc_displacement_okada92"):
    u = np.zeros(3)
    getattr(lib, fn)(*[D(v) for v in list(x) + list(src) + list(slip) + [MU, LAM]], orc.p(u))
    return u


def nikkhoo(x, P, slip, lib=ORC, fn="orc_displacement_nikkhoo15"):
    u = np.zeros(3)
    getattr(lib, fn)(*[D(v) for v in x], *[orc.p(np.ascontiguousarray(q, dtype=np.float64))
                                            for q in P],
                     *[D(v) for v in slip], D(MU), D(LAM), orc.p(u))
    return u


UNIT = ([1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0])


def test_okada_displacement_bit_exact_vs_reference_fixture():
    g = np.load(GOLD)
    for r, u_ref in zip(g["rect_inputs"], g["rect_u"]):
        for s in range(3):
            assert np.array_equal(okada(r[:3], r[3:], UNIT[s]), u_ref[s])
    # the fault-edge receiver is DC3D-singular: zero displacement
    assert not g["rect_u"][0].any() and g["rect_u"][1].any()


def test_nikkhoo_displacement_vs_reference_fixture():
    """Same field as the reference, evaluated in a different order: normwise
    relative differences at rounding level (the sum of the main, harmonic and
    image parts cancels to a few digits near the surface)."""
    g = np.load(GOLD)
    rel = []
    for i in range(g["tri_points"].shape[0]):
        P = (g["P1"][i], g["P2"][i], g["P3"][i])
        for s in range(3):
            u, r = nikkhoo(g["tri_points"][i], P, UNIT[s]), g["tri_u"][i, s]
            if i == 0:
                # a surface triangle seen from a surface point: the main and the
                # sign-flipped image parts cancel exactly, leaving round-off
                assert np.abs(u).max() < 1e-15 and np.abs(r).max() < 1e-15
                continue
            rel.append(np.abs(u - r).max() / np.abs(r).max())
            if i == 1:  # horizontal buried triangle
                assert rel[-1] < 1e-11
    rel = np.array(rel)
    assert np.median(rel) < 1e-13 and np.quantile(rel, 0.99) < 1e-10 and rel.max() < 1e-9


def test_displacement_live_reference_checker():
    ref = orc.ref()
    if ref is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    rng = np.random.default_rng(11)
    for _ in range(300):
        x = [rng.uniform(-2e4, 2e4), rng.uniform(-2e4, 2e4), rng.choice([0.0, rng.uniform(0, 2e4)])]
        src = [rng.uniform(-2e4, 2e4), rng.uniform(-2e4, 2e4), rng.uniform(0, 2e4),
               rng.uniform(500, 5e3), rng.uniform(500, 5e3), rng.uniform(0, 6.28),
               rng.uniform(0.05, 1.57)]
        for sl in UNIT:
            assert np.array_equal(okada(x, src, sl),
                                  okada(x, src, sl, ref, "ref_displacement_okada92"))
        P1 = np.array([rng.uniform(-5e3, 5e3), rng.uniform(-5e3, 5e3), rng.uniform(500, 1e4)])
        P = (P1, P1 + rng.uniform(-3e3, 3e3, 3) * [1, 1, 0.1],
             P1 + rng.uniform(-3e3, 3e3, 3) * [1, 1, 0.1])
        for sl in UNIT:
            a = nikkhoo(x, P, sl)
            b = nikkhoo(x, P, sl, ref, "ref_displacement_nikkhoo15")
            assert np.abs(a - b).max() <= 1e-9 * np.abs(b).max()


def _plane(st, dp):
    sv = np.array([np.cos(st), np.sin(st), 0.0])
    dd = np.array([-np.sin(st) * np.cos(dp), np.cos(st) * np.cos(dp), np.sin(dp)])  # down dip
    nv = np.array([-np.sin(st) * np.sin(dp), np.cos(st) * np.sin(dp), -np.cos(dp)])
    return sv, dd, nv


@pytest.mark.parametrize("st,dp", [(0.7, 0.5), (3.5, 1.2), (0.0, np.pi / 2)])
def test_displacement_jump_equals_slip(st, dp):
    """Across the dislocation, u(hanging wall) - u(footwall) = slip: strike slip
    along strike, dip slip up dip (okada92.f90:59-87 frames), opening along the
    normal -- for the rectangle and for its two triangles.  A swapped NED/ENU
    component, a sign or a dropped image/harmonic term moves the jump."""
    x0, L, W = np.array([1000.0, -2000.0, 1500.0]), 2000.0, 2000.0
    sv, dd, nv = _plane(st, dp)
    A, B = x0, x0 + L * sv
    tris = ((A, B, B + W * dd), (A, B + W * dd, A + W * dd))
    eps = 0.05  # the smooth part changes by ~eps |grad u| ~ 5e-5 over the step
    for a, b in [(0.4, 0.55), (0.8, 0.2)]:
        p = x0 + a * L * sv + b * W * dd
        for sl, want in zip(UNIT, (sv, -dd, nv)):
            jr = okada(p + eps * nv, list(x0) + [L, W, st, dp], sl) - \
                okada(p - eps * nv, list(x0) + [L, W, st, dp], sl)
            jt = sum(nikkhoo(p + eps * nv, P, sl) - nikkhoo(p - eps * nv, P, sl) for P in tris)
            assert np.abs(jr - want).max() < 2e-4
            assert np.abs(jt - want).max() < 2e-4


def test_rectangle_equals_two_triangles_displacement():
    m = configs.thrust_fault(nx=1, nz=1, patch=2000.0, top=1000.0)
    st, dp, x0 = m.strike[0], m.dip[0], m.x[0]
    sv, dd, _ = _plane(st, dp)
    A, B = x0, x0 + 2000.0 * sv
    tris = ((A, B, B + 2000.0 * dd), (A, B + 2000.0 * dd, A + 2000.0 * dd))
    rng = np.random.default_rng(5)
    for _ in range(20):
        x = np.array([rng.uniform(-8e3, 8e3), rng.uniform(-8e3, 8e3),
                      rng.choice([0.0, rng.uniform(0, 8e3)])])
        for sl in UNIT[:2]:
            ur = okada(x, list(x0) + [2000.0, 2000.0, st, dp], sl)
            ut = sum(nikkhoo(x, P, sl) for P in tris)
            # DC3D's binary32 interface bounds the agreement
            assert np.abs(ur - ut).max() <= 2e-6 * np.abs(ur).max()


def test_displacement_far_field_decay():
    """A finite dislocation's surface displacement decays as r^-2."""
    rs = np.array([2e5, 4e5, 8e5, 1.6e6])
    src = [0.0, 0.0, 5e3, 2e3, 2e3, 0.3, 0.6]
    P = (np.array([0.0, 0.0, 5e3]), np.array([2e3, 0.0, 5e3]), np.array([0.0, 2e3, 6e3]))
    for f in (lambda x: okada(x, src, UNIT[1]), lambda x: nikkhoo(x, P, UNIT[1])):
        vals = [np.abs(f([r * 0.6, r * 0.8, 0.0])).max() for r in rs]
        slope = np.polyfit(np.log(rs), np.log(vals), 1)[0]
        assert -2.1 < slope < -1.9


def _mixed_mesh():
    m = configs.thrust_fault(nx=3, nz=2)
    t = configs.triangulate(configs.thrust_fault(nx=2, nz=1, top=9000.0))
    v = configs.volume_grid(nx=2, ny=1, nz=1, size=3000.0, top=15e3, origin=(0.0, 10e3))
    return m, t, v, orc.Mesh(m.lambda_, m.mu, rect=m.geometry(), tri=t.geometry(),
                              vol=v.geometry())


def test_buildO_layout_and_partial_kernels():
    """Row 3k+c (point k, NED component c), column as in buildG; O = Of + Ol with
    disjoint supports: Of the fault columns and the volumes' third component,
    Ol the volumes' first two (greens.f90:386-390, 411-415, 441-445)."""
    m, t, v, mesh = _mixed_mesh()
    pts = np.array([[5e3, -3e3, 0.0], [-2e4, 1e4, 0.0], [1e3, 2e4, 500.0]])
    O, Of, Ol = orc.buildO(mesh, pts)
    nr, nt, nv = m.n, t.n, v.n
    n = 2 * (nr + nt) + 6 * nv
    assert O.shape == (9, n)
    nu = m.lambda_ / (m.lambda_ + m.mu) / 2.0
    for k, x in enumerate(pts):
        for s in range(nr):
            src = list(m.x[s]) + [m.length[s], m.width[s], m.strike[s], m.dip[s]]
            for j in range(2):
                u = np.zeros(3)
                ORC.orc_displacement_okada92(*[D(a) for a in list(x) + src + UNIT[j]],
                                             D(m.mu), D(m.lambda_), orc.p(u))
                assert np.array_equal(O[3 * k:3 * k + 3, 2 * s + j], u)
        for s in range(nt):
            P = [t.vertices[i[s] - 1] for i in (t.i1, t.i2, t.i3)]
            for j in range(2):
                u = np.zeros(3)
                ORC.orc_displacement_nikkhoo15(*[D(a) for a in x],
                                               *[orc.p(np.ascontiguousarray(q)) for q in P],
                                               *[D(a) for a in UNIT[j]], D(m.mu), D(m.lambda_),
                                               orc.p(u))
                assert np.array_equal(O[3 * k:3 * k + 3, 2 * nr + 2 * s + j], u)
        for s in range(nv):
            for j in range(6):
                e = np.zeros(6)
                e[j] = 1.0
                u = np.zeros(3)
                ORC.orc_displacement_strainvolume(*[D(a) for a in x],
                                                  *[D(a) for a in v.x[s]], D(v.length[s]),
                                                  D(v.thickness[s]), D(v.width[s]),
                                                  D(v.strike[s]), orc.p(e), D(m.mu), D(nu),
                                                  orc.p(u))
                assert np.array_equal(O[3 * k:3 * k + 3, 2 * (nr + nt) + 6 * s + j], u)
    nf = 2 * (nr + nt)
    assert np.array_equal(Of[:, :nf], O[:, :nf]) and not Ol[:, :nf].any()
    assert np.array_equal(Of[2::3, nf:], O[2::3, nf:]) and not Ol[2::3, nf:].any()
    for c in (0, 1):
        assert np.array_equal(Ol[c::3, nf:], O[c::3, nf:]) and not Of[c::3, nf:].any()
    assert np.array_equal(Of + Ol, O) and np.abs(O[:, nf:]).max() > 0


def test_slip_vector_gather():
    """u of exportPoints: (slip strike, slip dip) per patch, then e11..e33 per
    volume (state offsets of macros.f90:57-79)."""
    y = np.arange(7 * 3 + 12 * 2, dtype=float)
    u = orc.slip_vector(y, 3, 2)
    assert u.tolist() == [0, 1, 7, 8, 14, 15] + list(range(21, 27)) + list(range(33, 39))
