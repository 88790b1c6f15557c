"""The VTP geometry writers (exportvtk_rfaults.f90, exportvtk_volumes.f90):
Fortran ES11.2 numbers, a hand-written one-patch file, well-formed XML with
one piece per patch and two per volume, and corners that span the patch and
the cuboid (edge lengths L, W, T within the files' three digits)."""
import math
import xml.etree.ElementTree as ET

import numpy as np
import pytest

from unicycle_amd import configs, vtp


def _approx(x):
    return pytest.approx(x, rel=6e-3, abs=1e-6)


def test_es11_2():
    assert vtp.es(1234.5) == "   1.23E+03"
    assert vtp.es(-0.5) == "  -5.00E-01"
    assert vtp.es(0.0) == "   0.00E+00"
    assert vtp.es(0.006) == "   6.00E-03"
    assert vtp.es(1.0e-300) == "   1.00-300"


def test_one_rectangle_file():
    """strike 0, dip 90 degrees: the corners (0,0,0), (L,0,0), (L, W cos dip, W),
    (0, W cos dip, W) with cos(pi/2) ~ 6.12e-17 printed as the reference prints it."""
    fr = tuple(np.array([v]) for v in (1e-9, 0.5, 0.6, 100.0, 0.01, 0.014, 0.01))
    text = vtp.rectangles_vtp(np.zeros((1, 3)), [1000.0], [1000.0], [0.0], [math.pi / 2], fr)
    want = """<?xml version="1.0"?>
<VTKFile type="PolyData" version="0.1">
  <PolyData>
    <Piece NumberOfPoints="4" NumberOfPolys="1">
      <Points>
        <DataArray type="Float32" Name="Fault Patch" NumberOfComponents="3" format="ascii">
   0.00E+00   0.00E+00   0.00E+00   1.00E+03   0.00E+00   0.00E+00   1.00E+03   6.12E-14   1.00E+03   0.00E+00   6.12E-14   1.00E+03
        </DataArray>
      </Points>
      <Polys>
        <DataArray type="Int32" Name="connectivity" format="ascii" RangeMin="0" RangeMax="3">
0 1 2 3
        </DataArray>
        <DataArray type="Int32" Name="offsets" format="ascii" RangeMin="4" RangeMax="4">
          4
        </DataArray>
      </Polys>
      <CellData Normals="Normals">
        <DataArray type="Float32" Name="a" NumberOfComponents="1" format="ascii">
   1.00E-02
        </DataArray>
        <DataArray type="Float32" Name="b" NumberOfComponents="1" format="ascii">
   1.40E-02
        </DataArray>
        <DataArray type="Float32" Name="Vpl" NumberOfComponents="1" format="ascii">
   1.00E-09
        </DataArray>
        <DataArray type="Float32" Name="rake" NumberOfComponents="1" format="ascii">
   5.00E-01
        </DataArray>
        <DataArray type="Float32" Name="mu0" NumberOfComponents="1" format="ascii">
   6.00E-01
        </DataArray>
        <DataArray type="Float32" Name="sig" NumberOfComponents="1" format="ascii">
   1.00E+02
        </DataArray>
        <DataArray type="Float32" Name="L" NumberOfComponents="1" format="ascii">
   1.00E-02
        </DataArray>
      </CellData>
    </Piece>
  </PolyData>
</VTKFile>
"""
    assert text == want


def _points(piece):
    arr = piece.find("Points/DataArray")
    return np.array([float(v) for v in arr.text.split()]).reshape(-1, 3)


def test_rectangle_mesh_geometry():
    m = configs.thrust_fault(nx=3, nz=2)
    root = ET.fromstring(vtp.rectangles_vtp(m.x, m.length, m.width, m.strike, m.dip,
                                            m.friction()))
    pieces = root.findall("PolyData/Piece")
    assert len(pieces) == m.n
    for k, pc in enumerate(pieces):
        p = _points(pc)
        np.testing.assert_allclose(p[0], m.x[k], rtol=5e-3, atol=1.0)
        assert math.dist(p[0], p[1]) == _approx(m.length[k])
        assert math.dist(p[1], p[2]) == _approx(m.width[k])
        # the down-dip corner is deeper by W sin(dip)
        assert p[3][2] - p[0][2] == _approx(m.width[k] * math.sin(m.dip[k]))
        cells = {a.get("Name"): float(a.text) for a in pc.findall("CellData/DataArray")}
        assert cells["b"] == _approx(m.b[k]) and cells["sig"] == _approx(m.sig[k])


def test_volume_cuboids():
    v = configs.volume_grid(2, 2, 1, size=4000.0, top=20e3, origin=(0.0, 0.0))
    L, T, W = v.length.copy(), v.thickness.copy(), v.width.copy()
    L[1], T[2], W[3] = 2000.0, 1500.0, 3000.0
    root = ET.fromstring(vtp.volumes_vtp(v.x, L, T, W, v.strike, np.full(v.n, math.pi / 2)))
    pieces = root.findall("PolyData/Piece")
    assert len(pieces) == 2 * v.n
    for k in range(v.n):
        a, b = _points(pieces[2 * k]), _points(pieces[2 * k + 1])
        np.testing.assert_array_equal(a, b)
        assert pieces[2 * k].find("Points/DataArray").get("Name") == "Fault Patch"
        assert pieces[2 * k + 1].find("Points/DataArray").get("Name") == "Weak Zone"
        ext = a.max(axis=0) - a.min(axis=0)   # strike north, vertical: (L, T, W)
        np.testing.assert_allclose(ext, [L[k], T[k], W[k]], rtol=1e-2)
