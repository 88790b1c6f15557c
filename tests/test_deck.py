"""The input-deck reader and the output formats of the drop-in driver (CPU).
Deck layout and checks of ratestate.f90:958-1739; comment handling of
getdata.f; Fortran list-directed values (commas, D exponents, quoted
names); ES formats of the time and patch files (ratestate.f90:270-271, 530,
505-521)."""
import numpy as np
import pytest

from unicycle_amd import configs, deck

DECK = """# a hand-written deck in the reference's layout
# output directory
  output/run1
# elastic moduli (lambda, mu)
   30d3, 3.0E4
# time interval
 3.15576d8
# number of rectangular patches
 2
#  n  Vpl  x1 x2 x3 length width strike dip rake
   1 1d-9 0 0 500 2000 2000 0 30 90
   2, 1e-9, 2000, 0, 500, 2000, 2000, 10, 30, 89
# number of frictional rectangular patches
 2
#  n mu0 sig a b L Vo G/(2Vs)
   1 0.6 100 0.01 0.014 0.01 1d-6 5
   2 0.6 100 0.01 0.006 0.01 1d-6 5
# number of triangular patches
 1
   1 1e-9 1 2 3 45
# number of vertices
 3
   1 0 0 1000
   2 1000 0 1000
   3 0 1000 2000
# number of frictional triangular patches
 1
   1 0.6 50 0.01 0.015 0.02 1e-6 5
# number of observation patches
 2
   1 1
   2 3
# number of observation points
 1
   1 'GPS1' 100 200 0
# number of events
 3
 10
 20
"""


def test_parse_hand_written_deck():
    d = deck.parse(DECK)
    assert d.wdir == "output/run1"
    assert (d.lambda_, d.mu, d.interval) == (30e3, 3.0e4, 3.15576e8)
    assert d.n_rect == 2 and d.n_tri == 1
    np.testing.assert_array_equal(d.rect_x, [[0, 0, 500], [2000, 0, 500]])
    # degrees to radians with the reference's constant
    assert d.rect_strike[1] == 10 * deck.DEG2RAD and d.rect_dip[0] == 30 * deck.DEG2RAD
    Vpl, rake, mu0, sig, a, b, L, Vo, damping = d.friction_arrays()
    assert rake.tolist() == [90 * deck.DEG2RAD, 89 * deck.DEG2RAD, 45 * deck.DEG2RAD]
    assert b.tolist() == [0.014, 0.006, 0.015] and sig.tolist() == [100, 100, 50]
    assert Vo[0] == 1e-6 and L[2] == 0.02 and damping.tolist() == [5, 5, 5]
    assert d.tri_i1.tolist() == [1] and d.tri_i3.tolist() == [3]
    np.testing.assert_array_equal(d.vertices[2], [0, 1000, 2000])
    assert d.observation_states == [1, 3]
    assert d.observation_points == [("GPS1", 100.0, 200.0, 0.0)]
    assert d.events == [0.0, 10.0, 20.0]      # the first event has no line
    assert d.nu == pytest.approx(0.25)


@pytest.mark.parametrize("bad,needle", [
    (("   2, 1e-9, 2000", "   3, 1e-9, 2000"), "unexpected index"),
    (("# number of frictional rectangular patches\n 2", "# number of frictional rectangular patches\n 1"),
     "frictional properties"),
    (("# number of vertices\n 3", "# number of vertices\n 2"), "not enough vertices"),
    (("   30d3, 3.0E4", "   -4d4, 3.0E4"), "Poisson"),
    ((" 10\n 20", " 20\n 10"), "must increase"),
    (("   1 1e-9 1 2 3 45", "   1 1e-9 1 -2 3 45"), "negative"),
])
def test_deck_errors(bad, needle):
    with pytest.raises(deck.DeckError, match=needle):
        deck.parse(DECK.replace(*bad))


def test_truncated_deck():
    with pytest.raises(deck.DeckError, match="ends before"):
        deck.parse(DECK[:DECK.index("# number of triangular")])


def test_written_deck_round_trip():
    m = configs.thrust_fault(nx=4, nz=2)
    text = deck.format_rectangles("out", m.lambda_, m.mu, 1e9, m.x, m.length, m.width, m.strike,
                                  m.dip, *m.friction(), observe=[1, 5], events=2)
    d = deck.parse(text)
    assert d.n_rect == 8 and d.observation_states == [1, 5] and d.events == [0.0, 1.0]
    np.testing.assert_array_equal(d.rect_x, m.x)
    np.testing.assert_allclose(d.rect_dip, m.dip, rtol=1e-15)
    for got, want in zip(d.friction_arrays(), m.friction()):
        np.testing.assert_allclose(got, want, rtol=1e-15)


def test_fortran_es_formats():
    f = deck.es_field
    assert f(1234.5678901234, 19, 12, 2) == " 1.234567890123E+03"
    assert f(-0.001, 19, 12, 2) == "-1.000000000000E-03"
    assert f(0.0, 19, 12, 2) == " 0.000000000000E+00"
    assert f(6.02e23, 20, 12, 3) == " 6.020000000000E+023"
    assert f(-1.5e-10, 20, 12, 3) == "-1.500000000000E-010"
    assert f(2.5e-7, 11, 4, 2) == " 2.5000E-07"


# ------------------------------------------------------------ viscouscycles --
VDECK = """# viscouscycles deck: one rectangle, two volumes
out/v1
# elastic moduli
30e3 30e3
# time interval
1e10
# number of rectangular patches
1
1 1e-9 0 0 500 2000 2000 0 30 90
# number of frictional rectangular patches
1
1 0.6 100 0.01 0.014 0.01 1e-6 5
# number of triangular patches
0
# number of cuboid strain volumes
2
# n e11 e12 e13 e22 e23 e33 x1 x2 x3 length width thickness strike dip
1 0 1d-15 0 0 0 0   0 0 20e3 4000 3000 2000 0 90
2 0 1e-15 0 0 0 0 4000 0 20e3 4000 3000 2000 10 90
# number of nonlinear Maxwell strain volumes
2
# n Gm gammadot0m n Q R
1 1 5e-14 1 0 8.314
2 1 5e-14 3 0 8.314
# number of thermal strain volumes
2
1 1 873
2 1 900
# number of observation patches
1
1 1
# number of observation volumes
1
1 2
# number of observation points
1
1 GPS 1e3 2e3 0
# number of perturbations
1
"""


def test_parse_viscous_deck():
    d = deck.parse_viscous(VDECK)
    assert d.n_rect == 1 and d.n_tri == 0 and d.n_vol == 2
    np.testing.assert_array_equal(d.vol_x, [[0, 0, 20e3], [4000, 0, 20e3]])
    # the deck's order is length, width, thickness
    np.testing.assert_array_equal(d.vol_length, [4000, 4000])
    np.testing.assert_array_equal(d.vol_width, [3000, 3000])
    np.testing.assert_array_equal(d.vol_thickness, [2000, 2000])
    assert d.vol_strike[1] == 10 * deck.DEG2RAD and d.vol_dip[0] == 90 * deck.DEG2RAD
    np.testing.assert_array_equal(d.vol_ebg[:, 1], [1e-15, 1e-15])
    assert d.n_maxwell == 2 and d.maxwell["npower"] == [1.0, 3.0]
    assert d.thermal["To"] == [873.0, 900.0]
    assert d.observation_states == [1] and d.observation_volumes == [2]
    assert d.observation_points == [("GPS", 1e3, 2e3, 0.0)] and d.events == [0.0]
    from unicycle_amd import ratestate
    g0, npow, ebg, Gm, Q, R, To = ratestate.rheology(d)
    np.testing.assert_array_equal(npow, [1.0, 3.0])
    np.testing.assert_array_equal(To, [873.0, 900.0])


@pytest.mark.parametrize("old,new,needle", [
    ("2 0 1e-15 0 0 0 0 4000 0 20e3 4000 3000 2000 10 90",
     "2 0 1e-15 0 0 0 0 4000 0 20e3 4000 3000 2000 10 80", "vertical"),
    ("1 0 1d-15 0 0 0 0   0 0 20e3", "1 0 1d-15 0 0 0 0   0 0 -2e3", "depth must be positive"),
    ("2 0 1e-15 0 0 0 0 4000", "3 0 1e-15 0 0 0 0 4000", "unexpected index"),
    ("1 0 1d-15 0 0 0 0   0 0 20e3 4000 3000 2000 0 90",
     "1 0 1d-15 0 0 0 0   0 0 20e3 0 0 0 0 90", "dimension must be positive"),
    ("# number of nonlinear Maxwell strain volumes\n2",
     "# number of nonlinear Maxwell strain volumes\n1", "none or all"),
    ("2 1 5e-14 3 0 8.314", "2 1 5e-14 3 0 0", "R must be positive"),
    ("1 1 5e-14 1 0 8.314", "1 0 5e-14 1 0 8.314", "sig0 must be positive"),
    ("2 1 900", "2 1 0", "To must be positive"),
    ("# number of thermal strain volumes\n2", "# number of thermal strain volumes\n1",
     "none or all"),
    ("# number of observation volumes\n1\n1 2", "# number of observation volumes\n1\n1 3",
     "not a volume"),
])
def test_viscous_deck_errors(old, new, needle):
    assert old in VDECK
    with pytest.raises(deck.DeckError, match=needle):
        deck.parse_viscous(VDECK.replace(old, new, 1))


def test_viscous_deck_without_thermal_rows():
    """No thermal rows: allowed with Q = 0 (To does not enter), refused with
    Q != 0 (the reference would read an unset temperature)."""
    from unicycle_amd import ratestate
    text = VDECK.replace("# number of thermal strain volumes\n2\n1 1 873\n2 1 900",
                         "# number of thermal strain volumes\n0")
    d = deck.parse_viscous(text)
    assert not d.thermal
    np.testing.assert_array_equal(ratestate.rheology(d)[6], [1.0, 1.0])
    d = deck.parse_viscous(text.replace("2 1 5e-14 3 0 8.314", "2 1 5e-14 3 1e5 8.314"))
    with pytest.raises(deck.DeckError, match="thermal"):
        ratestate.rheology(d)


def test_viscous_deck_round_trip():
    fault = configs.thrust_fault(nx=3, nz=2, ellipse=None)
    vol = configs.volume_grid(2, 2, 1, size=4000.0, top=20e3, origin=(0.0, 0.0))
    text = deck.format_viscous("w", fault.lambda_, fault.mu, 1e9, fault, vol, observe=[1, 6],
                               observe_vol=[4], points=[("A", 1.0, 2.0, 0.0)], events=3)
    d = deck.parse_viscous(text)
    assert d.n_rect == fault.n and d.n_vol == vol.n
    np.testing.assert_array_equal(d.vol_x, vol.x)
    np.testing.assert_array_equal(d.vol_thickness, vol.thickness)
    np.testing.assert_array_equal(d.vol_ebg, vol.edot_bg)
    np.testing.assert_allclose(d.vol_strike, vol.strike, rtol=1e-15)
    from unicycle_amd import ratestate
    for got, want in zip(ratestate.rheology(d), vol.rheology()):
        np.testing.assert_allclose(got, want, rtol=1e-15)
    for got, want in zip(d.friction_arrays(), fault.friction()):
        np.testing.assert_allclose(got, want, rtol=1e-15)
    assert d.observation_volumes == [4] and d.events == [0.0, 1.0, 2.0]
    # the volume-free deck with no elements at all is refused
    with pytest.raises(deck.DeckError, match="nothing to do"):
        deck.parse_viscous(deck.format_viscous("w", 30e3, 30e3, 1e9))


# --------------------------------------------------------------- 2-D antiplane --
APDECK = """# fortran/2d/antiplane deck
out/ap
# shear modulus
30e3
# time interval
3.15e9
# number of patches
2
# n Vpl x2 x3 width dip
1 1e-9 0 0 1000 90
2 1d-9, 0, 1000, 1000, 90
# number of frictional patches
2
1 0.6 100 0.01 0.014 0.01 1e-6 5
2 0.6 100 0.01 0.006 0.01 1e-6 5
# number of observation patches
1
1 2
# number of observation points
1
1 GPS01 5000 0
"""


def test_parse_antiplane_deck():
    d = deck.parse_antiplane(APDECK)
    assert d.n == 2 and d.mu == 30e3 and d.interval == 3.15e9
    np.testing.assert_array_equal(d.x, [[0, 0, 0], [0, 0, 1000]])
    assert d.dip[0] == 90 * deck.DEG2RAD
    assert d.observation_states == [2]
    assert d.observation_points == [("GPS01", 0.0, 5000.0, 0.0)]
    Vpl, rake, mu0, sig, a, b, L, Vo, eta = d.friction_arrays()
    np.testing.assert_array_equal(b, [0.014, 0.006])
    assert not rake.any()


@pytest.mark.parametrize("old,new,needle", [
    ("30e3\n# time", "-1\n# time", "shear modulus"),
    ("3.15e9", "0", "simulation time"),
    ("2 1d-9, 0, 1000, 1000, 90", "3 1d-9, 0, 1000, 1000, 90", "unexpected index"),
    ("1 1e-9 0 0 1000 90", "1 1e-9 0 0 0 90", "width must be positive"),
    ("# number of frictional patches\n2", "# number of frictional patches\n1",
     "frictional properties"),
    ("1 GPS01 5000 0", "2 GPS01 5000 0", "unexpected index"),
    ("# number of observation patches\n1\n1 2", "# number of observation patches\n1\n1 3",
     "not an element"),
])
def test_antiplane_deck_errors(old, new, needle):
    assert old in APDECK
    with pytest.raises(deck.DeckError, match=needle):
        deck.parse_antiplane(APDECK.replace(old, new, 1))


def test_antiplane_deck_round_trip():
    m = configs.antiplane_fault(n=40)
    d = deck.parse_antiplane(deck.format_antiplane("w", m, 1e9, observe=[1, 40],
                                                   points=[("A", 1e3, 0.0)]))
    np.testing.assert_array_equal(d.x, m.x)
    np.testing.assert_array_equal(d.width, m.width)
    np.testing.assert_allclose(d.dip, m.dip, rtol=1e-15)
    for got, want in zip(d.friction_arrays(), m.friction()):
        np.testing.assert_allclose(got, want, rtol=1e-15)
    assert d.observation_states == [1, 40] and d.observation_points == [("A", 0.0, 1e3, 0.0)]


# ---------------------------------------------------------------------- 0-D --
PDECK = """# fortran/0d deck
out/p
# shear modulus
30e3
# time interval
3.15e8
# Vpl radius
1d-9 2000
# mu0 sig a b L Vo G/(2Vs)
0.6 100 0.01 0.014 0.01 1e-6 5
"""


def test_parse_point_deck():
    d = deck.parse_point(PDECK)
    assert (d.wdir, d.mu, d.interval, d.Vpl, d.radius) == ("out/p", 30e3, 3.15e8, 1e-9, 2000.0)
    assert (d.mu0, d.sig, d.a, d.b, d.L, d.Vo, d.damping) == (0.6, 100, 0.01, 0.014, 0.01, 1e-6, 5)
    from unicycle_amd import springslider
    assert springslider.stiffness(d) == -7.0 * springslider.PI / 16.0 * 30e3 / 2000.0
    for old, new, needle in (("30e3\n# time", "-1\n# time", "shear modulus"),
                             ("3.15e8", "-1", "simulation time"),
                             ("0.01 1e-6 5", "0.01 1e-6", "7 values")):
        with pytest.raises(deck.DeckError, match=needle):
            deck.parse_point(PDECK.replace(old, new, 1))


PSDECK = """# output directory
out/ps
# elastic moduli (mu, lambda)
30d3, 20e3
# time interval
3.15e9
# number of patches
2
# n Vpl x2 x3 width dip
1 1e-9 0 0 1000 30
2 1d-9, 866.0254, 500, 1000, 30
# number of frictional patches
2
# n tau0 mu0 sig a b L Vo G/(2Vs)
1 0 0.6 100 0.01 0.014 0.01 1e-6 5
2 42.5 0.6 100 0.01 0.006 0.01 1e-6 5
# number of observation patches
1
1 2
# number of observation points
1
1 GPS01 5000 0
"""

PSVDECK = PSDECK.replace("""# number of observation patches
1
1 2
""", """# number of strain volumes
2
# n e22 e23 e33 x2 x3 width thickness dip
1 -1e-15 0 1e-15 -5e3 2e4 4e3 4e3 90
2 -1d-15, 2e-16, 1d-15, -1e3, 2e4, 4e3, 4e3, 60
# number of nonlinear Maxwell strain volumes
2
# n s22 s23 s33 gammadot0m n Q R
1 0 0 0 1e-12 1 0 8.314
2 0.5 -0.25 0 1e-12 3 1.2e5 8.314
# number of thermal strain volumes
2
1 3.3 1200
2 3.3 1300
# number of observation patches
1
1 2
# number of observation volumes
1
1 2
""")

APVDECK = APDECK.replace("""# number of observation patches
1
1 2
""", """# number of strain volumes
1
# n e12 e13 x2 x3 width thickness dip
1 1e-15 0 0 2e4 4e3 4e3 90
# number of nonlinear Maxwell strain volumes
1
# n Gm gammadot0m n Q R
1 1.0 1e-12 1 0 8.314
# number of thermal strain volumes
0
# number of observation patches
1
1 2
# number of observation volumes
1
1 1
""")


def test_parse_planestrain_deck():
    d = deck.parse_planestrain(PSDECK)
    assert isinstance(d, deck.PlaneStrainDeck) and d.planestrain and not d.viscous
    assert d.mu == 30e3 and d.lambda_ == 20e3 and d.n == 2
    np.testing.assert_array_equal(d.x, [[0, 0, 0], [0, 866.0254, 500]])
    assert d.dip[1] == 30 * deck.DEG2RAD
    np.testing.assert_array_equal(d.tau0, [0.0, 42.5])
    Vpl, rake, mu0, sig, a, b, L, Vo, eta = d.friction_arrays()
    np.testing.assert_array_equal(b, [0.014, 0.006])
    assert d.observation_points == [("GPS01", 0.0, 5000.0, 0.0)]


def test_parse_planestrain_viscous_deck():
    d = deck.parse_planestrain_viscous(PSVDECK)
    assert d.viscous and d.planestrain and d.n_vol == 2
    np.testing.assert_array_equal(d.vol_ebg[:, 3:], [[-1e-15, 0, 1e-15], [-1e-15, 2e-16, 1e-15]])
    assert not d.vol_ebg[:, :3].any()
    np.testing.assert_array_equal(d.vol_x, [[0, -5e3, 2e4], [0, -1e3, 2e4]])
    assert d.vol_dip[1] == 60 * deck.DEG2RAD          # any dip in plane strain
    np.testing.assert_array_equal(d.s0, [[0, 0, 0], [0.5, -0.25, 0]])
    assert d.maxwell["npower"] == [1.0, 3.0] and d.maxwell["Q"] == [0.0, 1.2e5]
    assert d.thermal["To"] == [1200.0, 1300.0]
    assert d.observation_states == [2] and d.observation_volumes == [2]


def test_parse_antiplane_viscous_deck():
    d = deck.parse_antiplane_viscous(APVDECK)
    assert d.viscous and not d.planestrain and d.n_vol == 1
    np.testing.assert_array_equal(d.vol_ebg[0], [0, 1e-15, 0, 0, 0, 0])
    assert d.maxwell["Gm"] == [1.0] and not d.thermal
    assert d.observation_volumes == [1]


@pytest.mark.parametrize("text,old,new,needle", [
    ("ps", "30d3, 20e3", "-1, 20e3", "shear modulus"),
    ("ps", "1 1e-9 0 0 1000 30", "1 1e-9 0 0 0 30", "width must be positive"),
    ("ps", "2 42.5 0.6", "3 42.5 0.6", "unexpected index"),
    ("psv", "1 -1e-15 0 1e-15 -5e3 2e4", "1 -1e-15 0 1e-15 -5e3 -2e4", "depth must be positive"),
    ("psv", "-1e3, 2e4, 4e3, 4e3, 60", "-1e3, 2e4, 4e3, 0, 60", "dimension must be positive"),
    ("psv", "# number of nonlinear Maxwell strain volumes\n2",
     "# number of nonlinear Maxwell strain volumes\n1", "none or all"),
    ("psv", "1 0 0 0 1e-12 1 0 8.314", "1 0 0 0 1e-12 1 0 0", "R must be positive"),
    ("psv", "2 3.3 1300", "2 3.3 0", "To must be positive"),
    ("psv", "# number of observation volumes\n1\n1 2", "# number of observation volumes\n1\n1 3",
     "not a volume"),
    ("apv", "4e3 4e3 90", "4e3 4e3 60", "vertical"),
    ("apv", "1 1.0 1e-12 1 0 8.314", "1 0 1e-12 1 0 8.314", "sig0 must be positive"),
])
def test_2d_viscous_deck_errors(text, old, new, needle):
    src, parse = {"ps": (PSDECK, deck.parse_planestrain),
                  "psv": (PSVDECK, deck.parse_planestrain_viscous),
                  "apv": (APVDECK, deck.parse_antiplane_viscous)}[text]
    assert old in src
    with pytest.raises(deck.DeckError, match=needle):
        parse(src.replace(old, new, 1))


def test_2d_decks_round_trip():
    m = configs.planestrain_viscous(n=7, n2=2, n3=2)
    s0 = np.arange(12.0).reshape(4, 3) * 1e-3
    d = deck.parse_planestrain_viscous(deck.format_2d(
        "w", m.fault, 1e9, vol=m, s0=s0, tau0=np.full(7, 3.0), observe=[1, 7], observe_vol=[4],
        points=[("A", 1e3, 0.0)]))
    np.testing.assert_array_equal(d.x, m.fault.x)
    np.testing.assert_allclose(d.dip, m.fault.dip, rtol=1e-15)
    np.testing.assert_array_equal(d.vol_x, m.vx)
    np.testing.assert_array_equal(d.vol_ebg, m.edot_bg())
    np.testing.assert_array_equal(d.s0, s0)
    np.testing.assert_array_equal(d.tau0, np.full(7, 3.0))
    assert (d.mu, d.lambda_) == (m.fault.mu, m.fault.lambda_)
    for k, key in enumerate(("gammadot0", "npower", "Q", "R")):
        np.testing.assert_array_equal(d.maxwell[key], getattr(m, key))
    np.testing.assert_array_equal(d.thermal["To"], m.To)
    ma = configs.antiplane_viscous(n=6, n2=2, n3=1)
    da = deck.parse_antiplane_viscous(deck.format_2d("w", ma.fault, 1e9, vol=ma))
    np.testing.assert_array_equal(da.vol_ebg, ma.edot_bg())
    np.testing.assert_array_equal(da.maxwell["Gm"], ma.Gm)
