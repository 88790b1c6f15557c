"""The drop-in driver end to end: an input deck in the reference's layout
(8 x 4 thrust, two observed patches) through unicycle_amd.ratestate on the
GPU, its time.txt and patch-*.dat against the oracle's run of the same
model (ratestate.f90:307-328 restated in C), and the outputs independent of
how the device run is split into calls (the start time continues the
reference's running sum)."""
import io
import os

import numpy as np
import pytest

import orc
from unicycle_amd import configs, deck, ratestate

pytestmark = pytest.mark.gpu


def _read(path, widths):
    """Fixed-width fields, as the reference's formats write them (a negative
    ES20.12E3 value fills its field and touches the previous one)."""
    out = []
    for line in open(path).read().splitlines():
        if line.startswith("#"):
            continue
        row, p = [], 0
        for w in widths:
            row.append(float(line[p:p + w]))
            p += w
        assert p == len(line)
        out.append(row)
    return np.array(out)


def test_deck_run_matches_oracle(tmp_path, engine_factory):
    m = configs.thrust_fault(nx=8, nz=4)
    wdir = str(tmp_path / "out")
    text = deck.format_rectangles(wdir, m.lambda_, m.mu, 1.0e9, m.x, m.length, m.width,
                                  m.strike, m.dip, *m.friction(), observe=[1, 17])
    d = deck.parse(text)
    steps = 300
    log = io.StringIO()
    it, time, y = ratestate.run(d, max_iterations=steps, chunk=steps, out=log)
    assert it == steps
    tt = _read(os.path.join(wdir, "time.txt"), [19, 19])
    p1 = _read(os.path.join(wdir, "patch-00000001.dat"), [19] + [20] * 14)
    p17 = _read(os.path.join(wdir, "patch-00000017.dat"), [19] + [20] * 14)
    assert tt.shape == (steps + 1, 2) and p1.shape == (steps, 15) and p17.shape == (steps, 15)
    assert log.getvalue().splitlines()[-1] == f"{steps + 1:09d} time steps."
    # the oracle on the deck's own model (its K from the oracle's buildG)
    fr = d.friction_arrays()
    K = orc.buildG(orc.Mesh(d.lambda_, d.mu, rect=(d.rect_x, d.rect_length, d.rect_width,
                                                   d.rect_strike, d.rect_dip)), nthreads=8)
    sysm = orc.System(fr, K)
    y0 = sysm.init_state()
    yref, st, tl, dl, rc = sysm.integrate(y0, d.interval, max_steps=steps, log_cap=steps)
    assert st.steps == steps
    # time.txt: (0, 1e-3), then per iteration (time at its start, previous step)
    assert tt[0].tolist() == [0.0, 1.0e-3]
    t_start = np.concatenate([[0.0], tl[:-1]])
    prev = np.concatenate([[0.0], dl[:-1]])
    np.testing.assert_allclose(tt[1:, 0], t_start, rtol=1e-6, atol=0)
    np.testing.assert_allclose(tt[1:, 1], prev, rtol=1e-6, atol=0)
    # patch files: first row = initial state and its derivative
    dydt0 = sysm.odefun(y0)[0]
    for p, e in ((p1, 0), (p17, 16)):
        assert p[0, 0] == 0.0
        np.testing.assert_allclose(p[0, 1:8], y0[7 * e:7 * e + 7], rtol=1e-12, atol=0)
        np.testing.assert_allclose(p[0, 8:15], dydt0[7 * e:7 * e + 7], rtol=1e-9,
                                   atol=1e-12 * np.abs(dydt0[7 * e:7 * e + 7]).max())
        # the time column of the patch file is the iteration start time
        np.testing.assert_array_equal(p[:, 0], tt[1:, 0])
    np.testing.assert_allclose(y, yref, rtol=1e-6, atol=1e-9 * np.abs(yref).max())
    # split into many calls, the run writes the same files
    wdir2 = str(tmp_path / "out2")
    d2 = deck.parse(text.replace(wdir, wdir2))
    ratestate.run(d2, max_iterations=steps, chunk=37, out=io.StringIO())
    for name in ("time.txt", "patch-00000001.dat", "patch-00000017.dat"):
        assert open(os.path.join(wdir, name)).read() == open(os.path.join(wdir2, name)).read()


def test_deck_run_writes_observation_points(tmp_path, engine_factory):
    """opts-<name>.dat (ratestate.f90:291-304, 414-486): per iteration the time
    and the displacement of each observation point at the iteration's start.
    With every patch observed, each row equals the oracle's O (buildO) times
    the slip written in the patch files for the same iteration."""
    m = configs.thrust_fault(nx=8, nz=4)
    wdir = str(tmp_path / "pts")
    pts = [("GPS1", 3000.0, -2000.0, 0.0), ("STATION_LONGNAME", 8000.0, 20000.0, 0.0),
           ("deep", 12000.0, 4000.0, 1500.0)]
    text = deck.format_rectangles(wdir, m.lambda_, m.mu, 1.0e9, m.x, m.length, m.width,
                                  m.strike, m.dip, *m.friction(),
                                  observe=list(range(1, m.n + 1)), points=pts)
    d = deck.parse(text)
    assert [p[0] for p in d.observation_points] == ["GPS1", "STATION_LO", "deep"]
    steps = 150
    ratestate.run(d, max_iterations=steps, chunk=steps, out=io.StringIO())
    tt = _read(os.path.join(wdir, "time.txt"), [19, 19])
    O, _, _ = orc.buildO(orc.Mesh(d.lambda_, d.mu, rect=(d.rect_x, d.rect_length, d.rect_width,
                                                         d.rect_strike, d.rect_dip)),
                         [p[1:] for p in d.observation_points])
    slip = np.zeros((steps, 2 * m.n))
    for e in range(m.n):
        p = _read(os.path.join(wdir, f"patch-{e + 1:08d}.dat"), [19] + [20] * 14)
        slip[:, 2 * e:2 * e + 2] = p[:, 1:3]
    want = slip @ O.T
    scale = np.abs(want).max()
    assert scale > 0
    for j, name in enumerate(["GPS1", "STATION_LO", "deep"]):
        q = _read(os.path.join(wdir, f"opts-{name}.dat"), [19] * 4)
        assert q.shape == (steps, 4)
        np.testing.assert_array_equal(q[:, 0], tt[1:, 0])
        np.testing.assert_allclose(q[:, 1:], want[:, 3 * j:3 * j + 3], rtol=1e-6,
                                   atol=1e-9 * scale)
    # the run split into calls writes the same files
    wdir2 = str(tmp_path / "pts2")
    ratestate.run(deck.parse(text.replace(wdir, wdir2)), max_iterations=steps, chunk=41,
                  out=io.StringIO())
    for name in ("GPS1", "STATION_LO", "deep"):
        f = f"opts-{name}.dat"
        assert open(os.path.join(wdir, f)).read() == open(os.path.join(wdir2, f)).read()


def test_viscous_deck_run_matches_oracle(tmp_path, engine_factory):
    """viscouscycles deck (faults + Maxwell volumes, every element observed,
    two observation points) through the driver against the oracle's run of the
    same model: time.txt, the first rows of the patch and volume files, eMax
    of the progress lines against the volume files' strain rates, and the
    10-column opts rows against the oracle's O, Of, Ol times the slips and
    strains the same run wrote."""
    fault = configs.thrust_fault(nx=6, nz=3, ellipse=None)
    vol = configs.volume_grid(3, 3, 2, size=4000.0, top=20e3, origin=(0.0, 0.0))
    wdir = str(tmp_path / "visc")
    pts = [("GPS1", 5e3, 2e3, 0.0), ("GPS2", 2e4, 8e3, 0.0)]
    text = deck.format_viscous(wdir, fault.lambda_, fault.mu, 50 * configs.YEAR, fault, vol,
                               observe=list(range(1, fault.n + 1)),
                               observe_vol=list(range(1, vol.n + 1)), points=pts)
    d = deck.parse_viscous(text)
    steps = 100
    log = io.StringIO()
    it, _, y = ratestate.run(d, max_iterations=steps, chunk=steps, out=log)
    assert it == steps
    mesh = orc.Mesh(fault.lambda_, fault.mu, rect=fault.geometry(), vol=vol.geometry())
    K = orc.buildG(mesh, nthreads=8)
    sysm = orc.System(fault.friction(), K, rheo=vol.rheology(), ic_variant=1)
    y0 = sysm.init_state()
    yref, st, tl, dl, rc = sysm.integrate(y0, d.interval, max_steps=steps, log_cap=steps)
    tt = _read(os.path.join(wdir, "time.txt"), [19, 19])
    np.testing.assert_allclose(tt[1:, 0], np.concatenate([[0.0], tl[:-1]]), rtol=1e-5, atol=0)
    np.testing.assert_allclose(tt[2:, 1], dl[:-1], rtol=1e-5)
    dydt0 = sysm.odefun(y0)[0]
    npch = fault.n
    slip = np.zeros((steps, 2 * npch))
    for e in range(npch):
        p = _read(os.path.join(wdir, f"patch-{e + 1:08d}.dat"), [19] + [20] * 14)
        np.testing.assert_allclose(p[0, 1:8], y0[7 * e:7 * e + 7], rtol=1e-12, atol=0)
        slip[:, 2 * e:2 * e + 2] = p[:, 1:3]
    strain = np.zeros((steps, 6 * vol.n))
    erate = np.zeros((steps, vol.n))
    for v in range(vol.n):
        q = _read(os.path.join(wdir, f"volume-{v + 1:08d}.dat"), [19] + [20] * 24)
        o = 7 * npch + 12 * v
        np.testing.assert_allclose(q[0, 1:13], y0[o:o + 12], rtol=1e-12, atol=0)
        np.testing.assert_allclose(q[0, 13:25], dydt0[o:o + 12], rtol=1e-9,
                                   atol=1e-12 * np.abs(dydt0[o:o + 12]).max())
        np.testing.assert_array_equal(q[:, 0], tt[1:, 0])
        strain[:, 6 * v:6 * v + 6] = q[:, 1:7]
        r = q[:, 13:19]
        erate[:, v] = np.sqrt((r[:, 0] ** 2 + 2 * r[:, 1] ** 2 + 2 * r[:, 2] ** 2 + r[:, 3] ** 2
                               + 2 * r[:, 4] ** 2 + r[:, 5] ** 2) / 2)
    # progress lines: iteration, time, dt, vMax, eMax (viscouscycles.f90:690-695)
    lines = [ln for ln in log.getvalue().splitlines() if not ln.startswith("#")]
    prog = [ln for ln in lines[1:-1]]
    assert [int(ln[:9]) for ln in prog] == [50, 100]
    for ln in prog:
        i = int(ln[:9])
        emax = float(ln[9 + 38 + 11:9 + 38 + 22])
        assert emax == pytest.approx(erate[i - 1].max(), rel=1e-4)
    # opts: time, d, df, dl (viscouscycles.f90:612-629)
    O, Of, Ol = orc.buildO(mesh, [p[1:] for p in pts])
    u = np.concatenate([slip, strain], axis=1)
    for j, (name, *_x) in enumerate(pts):
        q = _read(os.path.join(wdir, f"opts-{name}.dat"), [19] * 10)
        assert q.shape == (steps, 10)
        for c, A in enumerate((O, Of, Ol)):
            want = u @ A[3 * j:3 * j + 3].T
            np.testing.assert_allclose(q[:, 1 + 3 * c:4 + 3 * c], want, rtol=1e-6,
                                       atol=1e-9 * np.abs(want).max() + 1e-300)
    assert os.path.getsize(os.path.join(wdir, "volume.vtp")) > 0
    assert np.abs(y - yref).max() <= 1e-5 * np.abs(yref).max()


def test_antiplane_deck_run_matches_oracle(tmp_path, engine_factory):
    """fortran/2d/antiplane deck through the driver: time.txt and the patch
    files (4 state entries) against the oracle's run of the same model, and
    the one-column opts rows against the oracle's buildO times the slips the
    run wrote (every patch observed)."""
    m = configs.antiplane_fault(n=100)
    wdir = str(tmp_path / "ap")
    pts = [("GPS1", 2e3, 0.0), ("GPS2", 2e4, 0.0), ("BH", 500.0, 3e3)]
    d = deck.parse_antiplane(deck.format_antiplane(wdir, m, 100 * configs.YEAR,
                                                   observe=list(range(1, m.n + 1)), points=pts))
    steps = 120
    log = io.StringIO()
    it, _, y = ratestate.run(d, max_iterations=steps, chunk=50, out=log)
    assert it == steps
    K = orc.buildG_antiplane(d.x, d.width, d.dip, d.mu, nthreads=8)
    sysm = orc.APSystem(m.oracle_friction(), K)
    y0 = sysm.init_state()
    yref, st, tl, dl, vl, rc = sysm.integrate(y0, d.interval, max_steps=steps, log_cap=steps)
    tt = _read(os.path.join(wdir, "time.txt"), [19, 19])
    np.testing.assert_allclose(tt[1:, 0], np.concatenate([[0.0], tl[:-1]]), rtol=1e-6, atol=0)
    slip = np.zeros((steps, m.n))
    for e in range(m.n):
        p = _read(os.path.join(wdir, f"patch-{e + 1:08d}.dat"), [19] + [20] * 8)
        np.testing.assert_allclose(p[0, 1:5], y0[4 * e:4 * e + 4], rtol=1e-12, atol=0)
        slip[:, e] = p[:, 1]
    O = orc.buildO_antiplane(d.x, d.width, d.dip, [p[1:] for p in d.observation_points])
    want = slip @ O.T
    for j, (name, *_x) in enumerate(pts):
        q = _read(os.path.join(wdir, f"opts-{name}.dat"), [19] * 2)
        assert q.shape == (steps, 2)
        np.testing.assert_array_equal(q[:, 0], tt[1:, 0])
        np.testing.assert_allclose(q[:, 1], want[:, j], rtol=1e-6,
                                   atol=1e-9 * np.abs(want).max() + 1e-300)
    assert not os.path.exists(os.path.join(wdir, "patch-rectangular.vtp"))
    assert np.abs(y - yref).max() <= 1e-6 * np.abs(yref).max()


def test_springslider_deck_run(tmp_path, engine_factory):
    """fortran/0d deck through the driver: the step sequence of the oracle's
    one-patch model with the circular-crack stiffness (0d ratestate.f90:
    230-277 is the 2-D antiplane odefun with K = -7 pi mu / (16 r)), and the
    slip against scipy's DOP853 on the scalar ODE through the first event."""
    from scipy.integrate import solve_ivp
    from unicycle_amd import springslider
    wdir = str(tmp_path / "p")
    text = (f"{wdir}\n30e3\n{30 * configs.YEAR!r}\n1e-9 2000\n"
            "0.6 100 0.01 0.014 0.01 1e-6 5\n")
    d = deck.parse_point(text)
    log = io.StringIO()
    it, time, y = springslider.run(d, out=log, chunk=300)
    k = springslider.stiffness(d)
    sysm = orc.APSystem([np.array([v]) for v in (d.Vpl, d.mu0, d.sig, d.a, d.b, d.L, d.Vo,
                                                 d.damping)], np.array([[k]]))
    y0 = sysm.init_state()
    yref, st, tl, dl, vl, rc = sysm.integrate(y0, d.interval, max_steps=10**6, log_cap=10**6)
    assert it == st.steps and time == pytest.approx(st.time, rel=1e-9)
    assert vl.max() > 1e-3   # the first event is at about 23 years
    tt = _read(os.path.join(wdir, "time.txt"), [19, 19])
    np.testing.assert_allclose(tt[1:, 0], np.concatenate([[0.0], tl[:-1]]), rtol=1e-6, atol=0)
    rows = _read(os.path.join(wdir, "patch.dat"), [19] + [21] * 8)
    assert rows.shape == (it, 9)
    np.testing.assert_array_equal(rows[:, 0], tt[1:, 0])
    np.testing.assert_allclose(rows[0, 1:5], y0, rtol=1e-12)
    LN10 = np.log(10.0)

    def f(t, s):
        V = 10.0 ** s[3]
        dth = (10.0 ** (-s[2]) - V / d.L) / LN10
        dk = k * (V - d.Vpl)
        dv = (dk - d.b * d.sig * LN10 * dth) / (d.a * d.sig + d.damping * V) / LN10
        return [V, dk - d.damping * V * LN10 * dv, dth, dv]

    ref = solve_ivp(f, [0.0, time], y0, method="DOP853", rtol=1e-12, atol=1e-14)
    # the global error of RK45 at epsilon 1e-6 over 30 years and an event
    # (the DOP853 solution is six orders tighter): 3.6e-5 here
    assert y[0] == pytest.approx(ref.y[0, -1], rel=1e-4)
    assert np.abs(y - yref).max() <= 1e-6 * np.abs(yref).max()
