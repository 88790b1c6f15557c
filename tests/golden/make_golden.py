"""Write the golden fixtures of tests/golden/ from the reference's own
Green's-function routines, compiled from jdpmoore/unicycle fortran/3d/src by
oracle/Makefile into oracle/_ref/libucyref.so (run `make -C oracle ref`).

The fixtures are inputs and the reference's outputs only:
  okada_dc3d.npz        DC3D (dc3d.f:1-199) on random binary32 inputs, plus
                        singular-edge and negative-extension cases
  okada_traction.npz    computeTractionKernelsOkada92 / computeStressKernels-
                        Okada92 (okada92.f90:368-473) for random sources
  nikkhoo_strain.npz    straininhalfspace_ (StrainInHalfSpace.c:42) at random
                        points, three unit slips, incl. a surface triangle and
                        a horizontal triangle
  nikkhoo_traction.npz  computeTractionKernelsNikkhoo15 / computeStressKernels-
                        Nikkhoo15 (nikkhoo15.f90:187-304)
  strainvolume.npz      computeTractionKernels / computeStressKernels-
                        VerticalStrainVolume (strainvolume.f90:4671-4799), incl.
                        a receiver inside a volume and one at the surface, and
                        computeDisplacementVerticalStrainVolume (215-268)
  antiplane.npz         computeStressAntiplane / computeTractionKernels-
                        Antiplane (fortran/2d/antiplane/src/antiplane.f90)
  antiplane_volumes.npz computeStressStrainVolumeAntiplane, computeDisplacement-
                        StrainVolumeAntiplane, the traction and stress kernels
                        of volumes and segments (antiplane.f90:213-255,
                        401-635) and buildG of a small mesh of segments and
                        volumes assembled from those kernels in the layout of
                        greens.f90:93-248 (fortran/2d/antiplane)
  planestrain.npz       computeStressStrainVolumePlaneStrain, computeDisplacement-
                        StrainVolumePlaneStrain, computeStressPlaneStrain,
                        computeDisplacementPlaneStrain and computeTraction-
                        KernelsPlaneStrain (fortran/2d/planestrain/src)
  planestrain_volumes.npz  the traction and stress kernels of plane-strain
                        strain volumes and the volume-frame stress kernels of
                        segments (planestrain.f90:314-475), the volumes'
                        displacement kernels with buildO's unit strains
                        (:221-264), and buildG of a small mesh of segments and
                        volumes assembled from those kernels in the layout of
                        greens.f90:66-226 (fortran/2d/planestrain)
  displacement.npz      computeDisplacementOkada92 (okada92.f90:230-298) and
                        computeDisplacementNikkhoo15 (nikkhoo15.f90:324-367,
                        DispInHalfSpace.c:42-152) at random points, at the
                        surface and at depth, three unit slips, incl. a fault
                        edge (DC3D singular), a surface triangle and a
                        horizontal triangle

Usage: python tests/golden/make_golden.py [name ...]   (default: all)

"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = C.CDLL(os.path.join(ROOT, "oracle", "_ref", "libucyref.so"))
D, F, I = C.c_double, C.c_float, C.c_int


def P(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def PI(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def dc3d_cases(rng, n=400):
    rows = []
    for _ in range(n):
        rows.append([2 / 3, rng.uniform(-5e3, 5e3), rng.uniform(-5e3, 5e3), -rng.uniform(0, 5e3),
                     rng.uniform(0, 5e3), rng.uniform(1, 90), 0.0, rng.uniform(500, 3000),
                     -rng.uniform(500, 3000), 0.0])
    # receiver on the fault edge (singular) and on a negative edge extension
    rows.append([2 / 3, 0.0, 0.0, -1000.0, 1000.0, 90.0, 0.0, 1000.0, -1000.0, 0.0])
    rows.append([2 / 3, -500.0, 0.0, -1000.0, 1000.0, 90.0, 0.0, 1000.0, -1000.0, 0.0])
    rows.append([2 / 3, 500.0, 0.0, -500.0, 1000.0, 90.0, 0.0, 1000.0, -1000.0, 0.0])
    rows.append([2 / 3, 1500.0, 300.0, 0.0, 1000.0, 45.0, 0.0, 1000.0, -1000.0, 0.0])
    inp = np.array(rows, dtype=np.float32)
    out = np.zeros((len(rows), 3, 12), dtype=np.float32)
    iret = np.zeros((len(rows), 3), dtype=np.int32)
    for i, r in enumerate(inp):
        for s in range(3):
            u = (F * 12)()
            ir = I()
            disl = [0.0, 0.0, 0.0]
            disl[s] = 1.0
            REF.ref_dc3d(*[F(v) for v in r], *[F(v) for v in disl], u, C.byref(ir))
            out[i, s] = np.array(u[:], dtype=np.float32)
            iret[i, s] = ir.value
    np.savez_compressed(os.path.join(HERE, "okada_dc3d.npz"), inputs=inp, u=out, iret=iret)


def rect_sources(rng, ns):
    y = np.zeros((ns, 3))
    y[:, 0] = rng.uniform(-20e3, 20e3, ns)
    y[:, 1] = rng.uniform(-20e3, 20e3, ns)
    y[:, 2] = rng.uniform(0, 20e3, ns)
    return (y, rng.uniform(500, 5000, ns), rng.uniform(500, 5000, ns),
            rng.uniform(0, 2 * np.pi, ns), rng.uniform(0.05, np.pi / 2, ns))


def receiver(rng):
    x = np.array([rng.uniform(-5e3, 5e3), rng.uniform(-5e3, 5e3), rng.uniform(100, 10e3)])
    st, dp = rng.uniform(0, 2 * np.pi), rng.uniform(0.1, np.pi / 2)
    sv = np.array([np.cos(st), np.sin(st), 0.0])
    dv = np.array([np.sin(st) * np.cos(dp), -np.cos(st) * np.cos(dp), -np.sin(dp)])
    nv = np.array([-np.sin(st) * np.sin(dp), np.cos(st) * np.sin(dp), -np.cos(dp)])
    return x, sv, dv, nv


def okada_kernels(rng, ns=300, nrcv=4):
    y, L, W, st, dp = rect_sources(rng, ns)
    rcv, trac, stress = [], [], []
    for _ in range(nrcv):
        x, sv, dv, nv = receiver(rng)
        rcv.append(np.concatenate([x, sv, dv, nv]))
        for s, d in [(1.0, 0.0), (0.0, 1.0)]:
            t = [np.zeros(ns) for _ in range(3)]
            REF.ref_traction_okada92(P(x), P(sv), P(dv), P(nv), ns, P(np.ascontiguousarray(y)), P(L),
                                     P(W), P(st), P(dp), D(s), D(d), D(0.0), D(30e3), D(30e3),
                                     *[P(a) for a in t])
            trac.append(np.stack(t))
            q = [np.zeros(ns) for _ in range(6)]
            REF.ref_stress_okada92(P(x), P(sv), P(dv), P(nv), ns, P(np.ascontiguousarray(y)), P(L),
                                   P(W), P(st), P(dp), D(s), D(d), D(0.0), D(30e3), D(30e3),
                                   *[P(a) for a in q])
            stress.append(np.stack(q))
    np.savez_compressed(os.path.join(HERE, "okada_traction.npz"), y=y, L=L, W=W, strike=st, dip=dp,
                        receivers=np.array(rcv), traction=np.array(trac), stress=np.array(stress))


def triangles(rng, n):
    P1 = np.stack([rng.uniform(-5, 5, n), rng.uniform(-5, 5, n), -rng.uniform(0.5, 10, n)], 1)
    P2 = P1 + rng.uniform(-3, 3, (n, 3))
    P3 = P1 + rng.uniform(-3, 3, (n, 3))
    P2[:, 2] = np.minimum(P2[:, 2], -0.1)
    P3[:, 2] = np.minimum(P3[:, 2], -0.1)
    # a triangle on the surface and a horizontal buried triangle
    P1[0], P2[0], P3[0] = [0, 0, 0], [2, 0, 0], [0, 2, 0]
    P1[1], P2[1], P3[1] = [0, 0, -3], [2, 0, -3], [0, 2, -3]
    return P1, P2, P3


def nikkhoo_strain(rng, n=300):
    P1, P2, P3 = triangles(rng, n)
    pts = np.stack([rng.uniform(-10, 10, n), rng.uniform(-10, 10, n), -rng.uniform(0, 12, n)], 1)
    stress = np.zeros((n, 3, 6))
    strain = np.zeros((n, 3, 6))
    for i in range(n):
        for s in range(3):
            b = [0.0, 0.0, 0.0]
            b[s] = 1.0
            sg, e = np.zeros(6), np.zeros(6)
            REF.straininhalfspace_(P(sg), P(e), D(pts[i, 0]), D(pts[i, 1]), D(pts[i, 2]),
                                   P(np.ascontiguousarray(P1[i])), P(np.ascontiguousarray(P2[i])),
                                   P(np.ascontiguousarray(P3[i])), *[D(v) for v in b],
                                   D(30e3), D(30e3))
            stress[i, s], strain[i, s] = sg, e
    np.savez_compressed(os.path.join(HERE, "nikkhoo_strain.npz"), P1=P1, P2=P2, P3=P3, points=pts,
                        stress=stress, strain=strain)


def nikkhoo_kernels(rng, ntri=200, nrcv=4):
    # vertices in NED metres, sources below the receivers' depth range
    nve = 3 * ntri
    v = np.zeros((nve, 3))
    c = np.stack([rng.uniform(-20e3, 20e3, ntri), rng.uniform(-20e3, 20e3, ntri),
                  rng.uniform(1e3, 20e3, ntri)], 1)
    for k in range(3):
        v[k::3] = c + rng.uniform(-2e3, 2e3, (ntri, 3))
    v[:, 2] = np.abs(v[:, 2]) + 100.0
    i1 = np.arange(1, nve, 3, dtype=np.int32)
    i2 = i1 + 1
    i3 = i1 + 2
    rcv, trac, stress = [], [], []
    for _ in range(nrcv):
        x, sv, dv, nv = receiver(rng)
        rcv.append(np.concatenate([x, sv, dv, nv]))
        for s, d in [(1.0, 0.0), (0.0, 1.0)]:
            t = [np.zeros(ntri) for _ in range(3)]
            REF.ref_traction_nikkhoo15(P(x), P(sv), P(dv), P(nv), ntri, PI(i1), PI(i2), PI(i3), nve,
                                       P(np.ascontiguousarray(v)), D(s), D(d), D(0.0), D(30e3),
                                       D(30e3), *[P(a) for a in t])
            trac.append(np.stack(t))
            q = [np.zeros(ntri) for _ in range(6)]
            REF.ref_stress_nikkhoo15(P(x), P(sv), P(dv), P(nv), ntri, PI(i1), PI(i2), PI(i3), nve,
                                     P(np.ascontiguousarray(v)), D(s), D(d), D(0.0), D(30e3),
                                     D(30e3), *[P(a) for a in q])
            stress.append(np.stack(q))
    np.savez_compressed(os.path.join(HERE, "nikkhoo_traction.npz"), vertices=v, i1=i1, i2=i2,
                        i3=i3, receivers=np.array(rcv), traction=np.array(trac),
                        stress=np.array(stress))


def strainvolume_kernels(rng, ns=60, nrcv=4):
    y = np.zeros((ns, 3))
    y[:, 0] = rng.uniform(-20e3, 20e3, ns)
    y[:, 1] = rng.uniform(-20e3, 20e3, ns)
    y[:, 2] = rng.uniform(1e3, 20e3, ns)
    L, T, W = rng.uniform(1e3, 5e3, ns), rng.uniform(1e3, 5e3, ns), rng.uniform(1e3, 5e3, ns)
    st = rng.uniform(0, 2 * np.pi, ns)
    rcv, trac, stress = [], [], []
    rlist = [receiver(rng) for _ in range(nrcv)]
    # special receivers: inside volume 0 (eigenstrain removal) and at the
    # free surface above volume 1
    x, sv, dv, nv = receiver(rng)
    c0 = y[0] + np.array([0.3 * L[0] * np.cos(st[0]), 0.3 * L[0] * np.sin(st[0]), 0.4 * W[0]])
    rlist.append((c0, sv, dv, nv))
    rlist.append((np.array([y[1, 0], y[1, 1], 0.0]), sv, dv, nv))
    for x, sv, dv, nv in rlist:
        rcv.append(np.concatenate([x, sv, dv, nv]))
        for j in range(6):
            e = [0.0] * 6
            e[j] = 1.0
            t = [np.zeros(ns) for _ in range(3)]
            REF.ref_traction_strainvolume(P(x), P(sv), P(dv), P(nv), ns, P(np.ascontiguousarray(y)),
                                          P(L), P(T), P(W), P(st), *[D(v) for v in e], D(30e3),
                                          D(30e3), *[P(a) for a in t])
            trac.append(np.stack(t))
            q = [np.zeros(ns) for _ in range(6)]
            REF.ref_stress_strainvolume(P(x), P(sv), P(dv), P(nv), ns, P(np.ascontiguousarray(y)),
                                        P(L), P(T), P(W), P(st), *[D(v) for v in e], D(30e3),
                                        D(30e3), *[P(a) for a in q])
            stress.append(np.stack(q))
    # displacement routine (strainvolume.f90:215) at random points, unit eigenstrains
    pts, disp, src = [], [], []
    for _ in range(120):
        q = np.array([rng.uniform(-5e3, 5e3), rng.uniform(-5e3, 5e3), rng.uniform(1e3, 1e4)])
        g = np.array([rng.uniform(1e3, 5e3), rng.uniform(1e3, 5e3), rng.uniform(1e3, 5e3),
                      rng.uniform(0, 2 * np.pi)])
        xp = np.array([rng.uniform(-2e4, 2e4), rng.uniform(-2e4, 2e4), rng.uniform(0, 2e4)])
        for j in range(6):
            e = [0.0] * 6
            e[j] = 1.0
            u = [D() for _ in range(3)]
            REF.ref_displacement_strainvolume(*[D(v) for v in xp], *[D(v) for v in q],
                                              *[D(v) for v in g], *[D(v) for v in e], D(30e3),
                                              D(0.25), *[C.byref(v) for v in u])
            disp.append([v.value for v in u])
        pts.append(xp)
        src.append(np.concatenate([q, g]))
    np.savez_compressed(os.path.join(HERE, "strainvolume.npz"), y=y, L=L, T=T, W=W, strike=st,
                        receivers=np.array(rcv), traction=np.array(trac), stress=np.array(stress),
                        disp_points=np.array(pts), disp_sources=np.array(src),
                        disp=np.array(disp).reshape(len(pts), 6, 3))


def displacement(rng, nrect=400, ntri=400):
    # rectangles: (x1, x2, x3, q1, q2, q3, L, W, strike, dip), every third
    # point at the surface
    rin = np.zeros((nrect, 10))
    for i in range(nrect):
        rin[i] = [rng.uniform(-2e4, 2e4), rng.uniform(-2e4, 2e4),
                  0.0 if i % 3 == 0 else rng.uniform(0, 2e4), rng.uniform(-2e4, 2e4),
                  rng.uniform(-2e4, 2e4), rng.uniform(0, 2e4), rng.uniform(500, 5e3),
                  rng.uniform(500, 5e3), rng.uniform(0, 2 * np.pi), rng.uniform(0.05, np.pi / 2)]
    # a fault reaching the surface seen from its top corner (singular edge)
    # and from the surface trace's extension
    rin[0] = [1e3, 2e3, 0.0, 1e3, 2e3, 0.0, 4e3, 3e3, 0.3, 1.2]
    rin[1] = [1e3 - 2e3 * np.cos(0.3), 2e3 - 2e3 * np.sin(0.3), 0.0, 1e3, 2e3, 0.0, 4e3, 3e3,
              0.3, 1.2]
    ru = np.zeros((nrect, 3, 3))
    for i, r in enumerate(rin):
        for s in range(3):
            b = [0.0, 0.0, 0.0]
            b[s] = 1.0
            u = np.zeros(3)
            REF.ref_displacement_okada92(*[D(v) for v in r], *[D(v) for v in b], D(30e3),
                                         D(30e3), P(u))
            ru[i, s] = u
    # triangles in NED (positive depth), points at the surface and at depth
    P1 = np.stack([rng.uniform(-5e3, 5e3, ntri), rng.uniform(-5e3, 5e3, ntri),
                   rng.uniform(500, 1e4, ntri)], 1)
    P2 = P1 + rng.uniform(-3e3, 3e3, (ntri, 3))
    P3 = P1 + rng.uniform(-3e3, 3e3, (ntri, 3))
    P2[:, 2] = np.maximum(P2[:, 2], 100.0)
    P3[:, 2] = np.maximum(P3[:, 2], 100.0)
    P1[0], P2[0], P3[0] = [0, 0, 0], [2e3, 0, 0], [0, 2e3, 0]
    P1[1], P2[1], P3[1] = [0, 0, 3e3], [2e3, 0, 3e3], [0, 2e3, 3e3]
    pts = np.stack([rng.uniform(-1e4, 1e4, ntri), rng.uniform(-1e4, 1e4, ntri),
                    rng.uniform(0, 1.2e4, ntri)], 1)
    pts[::3, 2] = 0.0
    tu = np.zeros((ntri, 3, 3))
    for i in range(ntri):
        for s in range(3):
            b = [0.0, 0.0, 0.0]
            b[s] = 1.0
            u = np.zeros(3)
            REF.ref_displacement_nikkhoo15(*[D(v) for v in pts[i]],
                                           P(np.ascontiguousarray(P1[i])),
                                           P(np.ascontiguousarray(P2[i])),
                                           P(np.ascontiguousarray(P3[i])), *[D(v) for v in b],
                                           D(30e3), D(30e3), P(u))
            tu[i, s] = u
    np.savez_compressed(os.path.join(HERE, "displacement.npz"), rect_inputs=rin, rect_u=ru,
                        P1=P1, P2=P2, P3=P3, tri_points=pts, tri_u=tu)


def antiplane(rng, n=400, ns=120):
    """computeStressAntiplane at random points (some at the surface) and
    computeTractionKernelsAntiplane of dipping receivers on random segments
    (fortran/2d/antiplane/src/antiplane.f90:131-156, 352-379)."""
    inp = np.stack([rng.uniform(-2e4, 2e4, n), rng.uniform(0, 2e4, n), rng.uniform(-2e4, 2e4, n),
                    rng.uniform(0, 1e4, n), rng.uniform(100, 5e3, n), rng.uniform(0.1, 1.5, n)],
                   1)
    inp[::4, 1] = 0.0
    s = np.zeros((n, 2))
    for i, r in enumerate(inp):
        a, b = D(), D()
        REF.ref_stress_antiplane(*[D(v) for v in r], D(30e3), C.byref(a), C.byref(b))
        s[i] = a.value, b.value
    y = np.stack([np.zeros(ns), rng.uniform(-2e4, 2e4, ns), rng.uniform(0, 1e4, ns)], 1)
    W, dip = rng.uniform(100, 5e3, ns), rng.uniform(0.1, 1.5, ns)
    rcv, ts = [], []
    for _ in range(4):
        d = rng.uniform(0.1, 1.5)
        x = np.array([0.0, rng.uniform(-1e4, 1e4), rng.uniform(100, 1e4)])
        sv = np.array([1.0, 0.0, 0.0])
        dv = np.array([0.0, -np.cos(d), -np.sin(d)])
        nv = np.array([-0.0, np.sin(d), -np.cos(d)])
        out = np.zeros(ns)
        REF.ref_traction_antiplane(P(x), P(sv), P(dv), P(nv), ns, P(np.ascontiguousarray(y)),
                                   P(W), P(dip), D(30e3), P(out))
        rcv.append(np.concatenate([x, sv, dv, nv]))
        ts.append(out)
    np.savez_compressed(os.path.join(HERE, "antiplane.npz"), stress_inputs=inp, stress=s,
                        y=y, W=W, dip=dip, receivers=np.array(rcv), traction=np.array(ts))


def _ap_frames(x, W, dip):
    n = W.size
    o = [np.zeros((n, 3)) for _ in range(4)]
    REF.ref_refsys_antiplane(n, P(np.ascontiguousarray(x)), P(W), P(dip), *[P(a) for a in o])
    return o


def antiplane_volumes(rng, n=400):
    """Strain volumes of the 2-D antiplane model: the stress and displacement
    of unit and random eigenstrains at random points (a quarter at the
    surface, a quarter inside a volume), and buildG of 7 dipping segments and
    5 volumes assembled from the reference's kernel routines."""
    G = 30e3
    q = np.stack([rng.uniform(-2e4, 2e4, n), rng.uniform(0, 1e4, n), rng.uniform(200, 5e3, n),
                  rng.uniform(200, 5e3, n)], 1)            # q2, q3, T, W
    x = np.stack([rng.uniform(-2e4, 2e4, n), rng.uniform(0, 2e4, n)], 1)
    x[::4, 1] = 0.0
    inside = np.arange(1, n, 4)
    x[inside, 0] = q[inside, 0] + rng.uniform(-0.45, 0.45, inside.size) * q[inside, 2]
    x[inside, 1] = q[inside, 1] + rng.uniform(0.05, 0.95, inside.size) * q[inside, 3]
    e = rng.standard_normal((n, 2))
    e[::3] = [1.0, 0.0]
    e[1::3] = [0.0, 1.0]
    dip = np.full(n, np.pi / 2)
    inp = np.concatenate([x, q, dip[:, None], e], 1)      # x2 x3 q2 q3 T W dip e12 e13
    s, u = np.zeros((n, 2)), np.zeros(n)
    for i, r in enumerate(inp):
        a, b, c = D(), D(), D()
        REF.ref_stress_sv_antiplane(*[D(v) for v in r], D(G), C.byref(a), C.byref(b))
        REF.ref_displacement_sv_antiplane(*[D(v) for v in r], C.byref(c))
        s[i] = a.value, b.value
        u[i] = c.value
    # a mesh: segments (x, W, dip) and volumes (x, T, W, dip = 90 degrees)
    nps, nv = 7, 5
    px = np.stack([np.zeros(nps), rng.uniform(-1e4, 1e4, nps), rng.uniform(0, 5e3, nps)], 1)
    pW, pdip = rng.uniform(500, 3e3, nps), rng.uniform(0.3, 1.5, nps)
    vx = np.stack([np.zeros(nv), rng.uniform(-1e4, 1e4, nv), rng.uniform(1e4, 2e4, nv)], 1)
    vT, vW, vdip = rng.uniform(1e3, 4e3, nv), rng.uniform(1e3, 4e3, nv), np.full(nv, np.pi / 2)
    pf, vf = _ap_frames(px, pW, pdip), _ap_frames(vx, vW, vdip)
    m = nps + 2 * nv
    K = np.zeros((m, m))
    pxc, vxc = np.ascontiguousarray(px), np.ascontiguousarray(vx)
    for r in range(nps + nv):
        sv, dv, nvv, xc = [f[r] for f in pf] if r < nps else [f[r - nps] for f in vf]
        fr = [P(np.ascontiguousarray(a)) for a in (xc, sv, dv, nvv)]
        if r < nps:      # tractionRows: segments, then (e12, e13) per volume
            ts = np.zeros(nps)
            REF.ref_traction_antiplane(*fr, nps, P(pxc), P(pW), P(pdip), D(G), P(ts))
            K[r, :nps] = ts
            for j in range(2):
                tv = np.zeros(nv)
                REF.ref_traction_sv_antiplane(*fr, nv, P(vxc), P(vT), P(vW), P(vdip),
                                              D(float(j == 0)), D(float(j == 1)), D(G), P(tv))
                K[r, nps + j::2] = tv
        else:            # stressRows: s12 row, then s13 row
            row = nps + 2 * (r - nps)
            a, b = np.zeros(nps), np.zeros(nps)
            REF.ref_stress_kernels_antiplane(*fr, nps, P(pxc), P(pW), P(pdip), D(G), P(a), P(b))
            K[row, :nps], K[row + 1, :nps] = a, b
            for j in range(2):
                a, b = np.zeros(nv), np.zeros(nv)
                REF.ref_stress_kernels_sv_antiplane(*fr, nv, P(vxc), P(vT), P(vW), P(vdip),
                                                    D(float(j == 0)), D(float(j == 1)), D(G),
                                                    P(a), P(b))
                K[row, nps + j::2], K[row + 1, nps + j::2] = a, b
    np.savez_compressed(os.path.join(HERE, "antiplane_volumes.npz"), inputs=inp, stress=s,
                        displacement=u, px=px, pW=pW, pdip=pdip, vx=vx, vT=vT, vW=vW, vdip=vdip,
                        G=G, K=K)


def planestrain(rng, n=400, nf=300, ns=60):
    """The plane-strain routines on random inputs: dipping strain volumes
    with random strain (a quarter of the points at the surface, a quarter
    inside the zone), line dislocations, and the traction kernels of 4
    dipping receivers on ns random segments."""
    G, lam = 30e3, 20e3
    inp = []
    for i in range(n):
        q2, q3 = rng.uniform(-1e4, 1e4), rng.uniform(0, 1e4)
        T, W = rng.uniform(100, 5e3, 2)
        dip = rng.uniform(0.1, np.pi - 0.1)
        if i % 4 == 1:     # inside: a point of the zone's interior
            a, b = rng.uniform(-0.45, 0.45) * T, rng.uniform(0.05, 0.95) * W
            r2 = q2 + a * np.sin(dip) + b * np.cos(dip)
            x3 = q3 - a * np.cos(dip) + b * np.sin(dip)
        else:
            r2, x3 = q2 + rng.uniform(-2, 2) * W, q3 + rng.uniform(-1, 2) * W
        x3 = 0.0 if i % 4 == 0 or x3 < 0 else x3
        inp.append([r2, x3, q2, q3, T, W, dip, *rng.standard_normal(3)])
    inp = np.array(inp)
    s, u = np.zeros((n, 3)), np.zeros((n, 2))
    for i, r in enumerate(inp):
        o = [D() for _ in range(5)]
        REF.ref_stress_sv_planestrain(*[D(v) for v in r], D(G), D(lam), *[C.byref(x) for x in o[:3]])
        REF.ref_displacement_sv_planestrain(*[D(v) for v in r], D(G), D(lam), *[C.byref(x) for x in o[3:]])
        s[i] = [x.value for x in o[:3]]
        u[i] = [x.value for x in o[3:]]
    finp = []
    for i in range(nf):
        q2, q3 = rng.uniform(-1e4, 1e4), rng.uniform(0, 1e4)
        W, dip = rng.uniform(100, 5e3), rng.uniform(0.1, np.pi - 0.1)
        x3 = 0.0 if i % 4 == 0 else max(0.0, q3 + rng.uniform(-1, 2) * W)
        finp.append([q2 + rng.uniform(-3, 3) * W, x3, q2, q3, W, dip, rng.uniform(-2, 2)])
    finp = np.array(finp)
    fs, fu = np.zeros((nf, 3)), np.zeros((nf, 2))
    for i, r in enumerate(finp):
        o = [D() for _ in range(5)]
        REF.ref_stress_planestrain(*[D(v) for v in r], D(G), D(lam), *[C.byref(x) for x in o[:3]])
        REF.ref_displacement_planestrain(*[D(v) for v in r], D(G), D(lam), *[C.byref(x) for x in o[3:]])
        fs[i] = [x.value for x in o[:3]]
        fu[i] = [x.value for x in o[3:]]
    y = np.stack([np.zeros(ns), rng.uniform(-2e4, 2e4, ns), rng.uniform(0, 1e4, ns)], 1)
    W, dip = rng.uniform(100, 5e3, ns), rng.uniform(0.1, np.pi - 0.1, ns)
    rcv, td, tn = [], [], []
    for _ in range(4):
        d = rng.uniform(0.1, np.pi - 0.1)
        x = np.array([0.0, rng.uniform(-1e4, 1e4), rng.uniform(100, 1e4)])
        sv = np.array([1.0, 0.0, 0.0])
        dv = np.array([0.0, -np.cos(d), -np.sin(d)])
        nv = np.array([-0.0, np.sin(d), -np.cos(d)])
        a, b = np.zeros(ns), np.zeros(ns)
        REF.ref_traction_planestrain(P(x), P(sv), P(dv), P(nv), ns, P(np.ascontiguousarray(y)),
                                     P(W), P(dip), D(G), D(lam), P(a), P(b))
        rcv.append(np.concatenate([x, sv, dv, nv]))
        td.append(a)
        tn.append(b)
    np.savez_compressed(os.path.join(HERE, "planestrain.npz"), G=G, lam=lam, inputs=inp,
                        stress=s, displacement=u, fault_inputs=finp, fault_stress=fs,
                        fault_displacement=fu, y=y, W=W, dip=dip, receivers=np.array(rcv),
                        td=np.array(td), tn=np.array(tn))


def _ps_frames(x, W, dip):
    n = W.size
    o = [np.zeros((n, 3)) for _ in range(4)]
    REF.ref_refsys_planestrain(n, P(np.ascontiguousarray(x)), P(W), P(dip), *[P(a) for a in o])
    return o


def planestrain_volumes(rng, nrcv=4, ns=40, npts=12):
    """Strain volumes of the plane-strain viscouscycles program: the volume
    kernel routines of planestrain.f90:221-475 (traction and volume-frame
    stress kernels of volumes, volume-frame stress kernels of segments,
    displacement kernels of volumes with the unit strains of buildO) on
    random receivers and sources, and buildG of 6 segments and 5 volumes
    assembled from those routines in the layout of greens.f90:66-226, each
    called with the arguments greens.f90 passes (G = 30e3, lambda = 20e3)."""
    G, lam = 30e3, 20e3
    y = np.stack([np.zeros(ns), rng.uniform(-2e4, 2e4, ns), rng.uniform(0, 1e4, ns)], 1)
    T, W = rng.uniform(200, 5e3, ns), rng.uniform(200, 5e3, ns)
    dip = rng.uniform(0.1, np.pi - 0.1, ns)
    yc = np.ascontiguousarray(y)
    rcv, out = [], {k: [] for k in ("td", "tn", "vs", "fs")}
    for i in range(nrcv):
        d = rng.uniform(0.1, np.pi - 0.1)
        x = np.array([0.0, rng.uniform(-1e4, 1e4), rng.uniform(100, 1e4)])
        if i == 1:      # inside source 0's zone
            a, b = 0.2 * T[0], 0.4 * W[0]
            x[1] = y[0, 1] + a * np.sin(dip[0]) + b * np.cos(dip[0])
            x[2] = y[0, 2] - a * np.cos(dip[0]) + b * np.sin(dip[0])
        if i == 2:
            x[2] = 0.0
        fr = [x, np.array([1.0, 0.0, 0.0]), np.array([0.0, -np.cos(d), -np.sin(d)]),
              np.array([-0.0, np.sin(d), -np.cos(d)])]
        rcv.append(np.concatenate(fr))
        fp = [P(np.ascontiguousarray(a)) for a in fr]
        e = rng.standard_normal(3)
        a, b = np.zeros(ns), np.zeros(ns)
        REF.ref_traction_sv_planestrain(*fp, ns, P(yc), P(T), P(W), P(dip), *[D(v) for v in e],
                                        D(G), D(lam), P(a), P(b))
        out["td"].append(a)
        out["tn"].append(b)
        o = [np.zeros(ns) for _ in range(3)]
        REF.ref_stress_kernels_sv_planestrain(*fp, ns, P(yc), P(T), P(W), P(dip),
                                              *[D(v) for v in e], D(G), D(lam), *[P(a) for a in o])
        out["vs"].append(np.concatenate([[*e], *o]))
        o = [np.zeros(ns) for _ in range(3)]
        REF.ref_stress_kernels_planestrain(*fp, ns, P(yc), P(W), P(dip), D(lam), D(G),
                                           *[P(a) for a in o])
        out["fs"].append(np.concatenate(o))
    pts = np.stack([np.zeros(npts), rng.uniform(-2e4, 2e4, npts),
                    np.where(np.arange(npts) % 2 == 0, 0.0, rng.uniform(0, 1e4, npts))], 1)
    disp = []
    for q in pts:
        for eye in ((1.0, 0.0, 0.0), (0.0, 1.0, 0.0), (0.0, 0.0, 0.0)):
            u2, u3 = np.zeros(ns), np.zeros(ns)
            REF.ref_displacement_kernels_sv_planestrain(P(np.ascontiguousarray(q)), ns, P(yc), P(T),
                                                        P(W), P(dip), *[D(v) for v in eye], D(G),
                                                        D(lam), P(u2), P(u3))
            disp.append(np.concatenate([u2, u3]))
    # the mesh: 6 dipping segments and 5 volumes; K rows td, tn per segment,
    # then s22', s23', s33' per volume; columns the dip slips, then e22',
    # e23', e33' (strains of 1e-3) per volume
    nps, nv = 6, 5
    px = np.stack([np.zeros(nps), rng.uniform(-1e4, 1e4, nps), rng.uniform(0, 5e3, nps)], 1)
    pW, pdip = rng.uniform(500, 3e3, nps), rng.uniform(0.3, 1.5, nps)
    vx = np.stack([np.zeros(nv), rng.uniform(-1e4, 1e4, nv), rng.uniform(8e3, 2e4, nv)], 1)
    vT, vW = rng.uniform(1e3, 4e3, nv), rng.uniform(1e3, 4e3, nv)
    vdip = rng.uniform(0.3, np.pi - 0.3, nv)
    pf, vf = _ps_frames(px, pW, pdip), _ps_frames(vx, vW, vdip)
    K = np.zeros((2 * nps + 3 * nv, nps + 3 * nv))
    pxc, vxc = np.ascontiguousarray(px), np.ascontiguousarray(vx)
    for r in range(nps + nv):
        sv, dv, nvv, xc = [f[r] for f in pf] if r < nps else [f[r - nps] for f in vf]
        fr = [P(np.ascontiguousarray(a)) for a in (xc, sv, dv, nvv)]
        if r < nps:      # tractionRows
            a, b = np.zeros(nps), np.zeros(nps)
            REF.ref_traction_planestrain(*fr, nps, P(pxc), P(pW), P(pdip), D(G), D(lam), P(a), P(b))
            K[2 * r, :nps], K[2 * r + 1, :nps] = a, b
            for j in range(3):
                a, b = np.zeros(nv), np.zeros(nv)
                u = [1e-3 * float(j == k) for k in range(3)]
                REF.ref_traction_sv_planestrain(*fr, nv, P(vxc), P(vT), P(vW), P(vdip),
                                                *[D(v) for v in u], D(G), D(lam), P(a), P(b))
                K[2 * r, nps + j::3], K[2 * r + 1, nps + j::3] = a, b
        else:            # stressRows: (mu, lambda) into (lambda, G), as greens.f90 passes them
            row = 2 * nps + 3 * (r - nps)
            o = [np.zeros(nps) for _ in range(3)]
            REF.ref_stress_kernels_planestrain(*fr, nps, P(pxc), P(pW), P(pdip), D(G), D(lam),
                                               *[P(a) for a in o])
            for c in range(3):
                K[row + c, :nps] = o[c]
            for j in range(3):
                o = [np.zeros(nv) for _ in range(3)]
                u = [1e-3 * float(j == k) for k in range(3)]
                REF.ref_stress_kernels_sv_planestrain(*fr, nv, P(vxc), P(vT), P(vW), P(vdip),
                                                      *[D(v) for v in u], D(G), D(lam),
                                                      *[P(a) for a in o])
                for c in range(3):
                    K[row + c, nps + j::3] = o[c]
    np.savez_compressed(os.path.join(HERE, "planestrain_volumes.npz"), G=G, lam=lam, y=y, T=T,
                        W=W, dip=dip, receivers=np.array(rcv), td=np.array(out["td"]),
                        tn=np.array(out["tn"]), vstress=np.array(out["vs"]),
                        fstress=np.array(out["fs"]), points=pts, disp=np.array(disp), px=px,
                        pW=pW, pdip=pdip, vx=vx, vT=vT, vW=vW, vdip=vdip, K=K)


FIXTURES = {
    "okada_dc3d": (20260101, [dc3d_cases, okada_kernels, nikkhoo_strain, nikkhoo_kernels,
                              strainvolume_kernels]),
    "displacement": (20261202, [displacement]),
    "antiplane": (20261203, [antiplane]),
    "antiplane_volumes": (20261204, [antiplane_volumes]),
    "planestrain": (20261205, [planestrain]),
    "planestrain_volumes": (20261206, [planestrain_volumes]),
}

if __name__ == "__main__":
    import sys
    # the first five fixtures share one random stream (one seed, in order)
    names = sys.argv[1:] or list(FIXTURES)
    for name in names:
        seed, fns = FIXTURES["okada_dc3d" if name in ("okada_traction", "nikkhoo_strain",
                                                       "nikkhoo_traction", "strainvolume")
                             else name]
        rng = np.random.default_rng(seed)
        for fn in fns:
            fn(rng)
    for f in sorted(os.listdir(HERE)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(HERE, f)))
