"""VTK PolyData (.vtp) geometry files of the reference's drivers.

  rectangles_vtp   exportvtk_rfaults (exportvtk_rfaults.f90:10-159): one Piece
                   per rectangular patch, its four corners and the cell data
                   a, b, Vpl, rake, mu0, sig, L; written by ratestate.f90:374-399
                   and viscouscycles.f90:473-495 as <wdir>/patch-rectangular.vtp
  volumes_vtp      exportvtk_volumes (exportvtk_volumes.f90): two Pieces per
                   strain volume ("Fault Patch" and "Weak Zone", the eight
                   corners of the cuboid each); viscouscycles.f90:497-512,
                   <wdir>/volume.vtp

The text is the reference's, line for line: Fortran ES11.2 numbers (12 per
line for a patch, 24 for a volume, one record each) and its XML layout, so
that post-processing that reads these files keeps working.
"""
from __future__ import annotations

import math

Q = '"'


def es(x: float, w: int = 11, d: int = 2) -> str:
    """Fortran ESw.d of one value (exponent E+dd; +ddd without the E beyond 99)."""
    x = float(x)
    m = f"{x:.{d}E}"
    mant, exp = m.split("E")
    e = int(exp)
    if abs(e) <= 99:
        s = f"{mant}E{'-' if e < 0 else '+'}{abs(e):02d}"
    else:
        s = f"{mant}{'-' if e < 0 else '+'}{abs(e):03d}"
    return s.rjust(w)


def _header(out):
    out.append(f"<?xml version={Q}1.0{Q}?>")
    out.append(f"<VTKFile type={Q}PolyData{Q} version={Q}0.1{Q}>")
    out.append("  <PolyData>")


def _frame(strike, dip):
    cs, ss, cd, sd = math.cos(strike), math.sin(strike), math.cos(dip), math.sin(dip)
    s = (cs, ss, 0.0)
    d = (+ss * cd, -cs * cd, -sd)
    n = (-sd * ss, +sd * cs, -cd)
    return s, d, n


def _points_array(out, name):
    out.append("      <Points>")
    out.append(f"        <DataArray type={Q}Float32{Q} Name={Q}{name}{Q} "
               f"NumberOfComponents={Q}3{Q} format={Q}ascii{Q}>")


def _polys(out, conn, rmax, offset):
    out.append("        </DataArray>")
    out.append("      </Points>")
    out.append("      <Polys>")
    out.append(f"        <DataArray type={Q}Int32{Q} Name={Q}connectivity{Q} format={Q}ascii{Q} "
               f"RangeMin={Q}0{Q} RangeMax={Q}{rmax}{Q}>")
    out.append(conn)
    out.append("        </DataArray>")
    out.append(f"        <DataArray type={Q}Int32{Q} Name={Q}offsets{Q} format={Q}ascii{Q} "
               f"RangeMin={Q}{offset}{Q} RangeMax={Q}{offset}{Q}>")
    out.append(f"          {offset}")
    out.append("        </DataArray>")
    out.append("      </Polys>")


def rectangles_vtp(x, length, width, strike, dip, friction) -> str:
    """exportvtk_rfaults.  x (ns, 3) upper corners (north, east, down), angles
    in radians, friction = (Vpl, rake, mu0, sig, a, b, L, ...) per patch in
    the order of the deck (FaultModel.friction())."""
    Vpl, rake, mu0, sig, a, b, L = (friction[k] for k in range(7))
    out = []
    _header(out)
    for k in range(len(length)):
        Lk, Wk = float(length[k]), float(width[k])
        x1, x2, x3 = (float(v) for v in x[k])
        s, d, _ = _frame(float(strike[k]), float(dip[k]))
        out.append(f"    <Piece NumberOfPoints={Q}4{Q} NumberOfPolys={Q}1{Q}>")
        _points_array(out, "Fault Patch")
        pts = [x1, x2, x3,
               x1 + s[0] * Lk, x2 + s[1] * Lk, x3 + s[2] * Lk,
               x1 - d[0] * Wk + s[0] * Lk, x2 - d[1] * Wk + s[1] * Lk, x3 - d[2] * Wk + s[2] * Lk,
               x1 - d[0] * Wk, x2 - d[1] * Wk, x3 - d[2] * Wk]
        out.append("".join(es(v) for v in pts))
        _polys(out, "0 1 2 3", 3, 4)
        out.append(f"      <CellData Normals={Q}Normals{Q}>")
        for name, arr in (("a", a), ("b", b), ("Vpl", Vpl), ("rake", rake), ("mu0", mu0),
                          ("sig", sig), ("L", L)):
            out.append(f"        <DataArray type={Q}Float32{Q} Name={Q}{name}{Q} "
                       f"NumberOfComponents={Q}1{Q} format={Q}ascii{Q}>")
            out.append(es(arr[k]))
            out.append("        </DataArray>")
        out.append("      </CellData>")
        out.append("    </Piece>")
    out.append("  </PolyData>")
    out.append("</VTKFile>")
    return "\n".join(out) + "\n"


def volumes_vtp(x, length, thickness, width, strike, dip) -> str:
    """exportvtk_volumes: the cuboid of each volume, as a "Fault Patch" piece
    and a "Weak Zone" piece with the reference's two face lists."""
    out = []
    _header(out)
    for k in range(len(length)):
        Lk, Tk, Wk = float(length[k]), float(thickness[k]), float(width[k])
        x1, x2, x3 = (float(v) for v in x[k])
        s, d, n = _frame(float(strike[k]), float(dip[k]))
        h = [n[i] * Tk / 2.0 for i in range(3)]
        c = []
        for sign in (-1.0, 1.0):
            base = [(x1, x2, x3), (x1 + s[0] * Lk, x2 + s[1] * Lk, x3 + s[2] * Lk),
                    (x1 - d[0] * Wk + s[0] * Lk, x2 - d[1] * Wk + s[1] * Lk,
                     x3 - d[2] * Wk + s[2] * Lk),
                    (x1 - d[0] * Wk, x2 - d[1] * Wk, x3 - d[2] * Wk)]
            c.append([tuple(p[i] + sign * h[i] if sign > 0 else p[i] - h[i] for i in range(3))
                      for p in base])
        lo, hi = c
        # corner order of exportvtk_volumes: four on the -n side, then the +n
        # side starting from the down-dip corner at the start of strike
        corners = [lo[0], lo[1], lo[2], lo[3], hi[3], hi[0], hi[1], hi[2]]
        row = "".join(es(v) for p in corners for v in p)
        for name, conn in (("Fault Patch", "7 4 5 6 7 4 3 2 7 2 1 6"),
                           ("Weak Zone", "0 1 2 3 0 5 4 3 0 1 6 5")):
            out.append(f"    <Piece NumberOfPoints={Q}8{Q} NumberOfPolys={Q}1{Q}>")
            _points_array(out, name)
            out.append(row)
            _polys(out, conn, 7, 12)
            out.append("    </Piece>")
    out.append("  </PolyData>")
    out.append("</VTKFile>")
    return "\n".join(out) + "\n"
