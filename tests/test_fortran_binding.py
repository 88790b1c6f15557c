"""The Fortran host binding (unicycle_amd/fortran/ucy_fortran.f90): it
compiles with the image's Fortran compiler and links against libucy.so; on a
GPU, a host program in the shape of ratestate.f90:307-328 (tests/
fortran_driver.f90) steps through the library and matches the oracle's
main loop, and its buildO / exportPoints calls match the oracle's
displacement kernels."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import orc
from unicycle_amd import configs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FC = shutil.which("amdflang") or shutil.which("flang") or shutil.which("gfortran")
LIB = os.path.join(ROOT, "unicycle_amd")


def _build(tmp, driver="fortran_driver"):
    if FC is None:
        pytest.skip("no Fortran compiler")
    if not os.path.exists(os.path.join(LIB, "libucy.so")):
        pytest.skip("libucy.so not built")
    exe = os.path.join(tmp, driver)
    cmd = [FC, "-O1", "-J" if "gfortran" in FC else "-module-dir", str(tmp),
           os.path.join(ROOT, "unicycle_amd", "fortran", "ucy_fortran.f90"),
           os.path.join(ROOT, "tests", driver + ".f90"), "-o", exe,
           "-L" + LIB, "-lucy", "-Wl,-rpath," + LIB]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    return exe


POINTS = np.array([[3e3, -2e3, 0.0], [8e3, 2e4, 0.0], [1.2e4, 4e3, 0.0]])


def _write_mesh(path, model, nsteps, points=POINTS):
    with open(path, "wb") as f:
        np.array([model.n, nsteps], dtype=np.int32).tofile(f)
        np.array([model.lambda_, model.mu], dtype=np.float64).tofile(f)
        np.ascontiguousarray(model.x, dtype=np.float64).tofile(f)  # x(3,ns) column-major
        for arr in (model.length, model.width, model.strike, model.dip) + tuple(model.friction()):
            np.asarray(arr, dtype=np.float64).tofile(f)
        np.array([len(points)], dtype=np.int32).tofile(f)
        np.ascontiguousarray(points, dtype=np.float64).tofile(f)   # x(3,nobs)


def test_fortran_binding_links(tmp_path):
    """Compile + link; without a GPU the program stops with the library's
    message (the library itself never aborts)."""
    exe = _build(tmp_path)
    from conftest import gpu_available
    if gpu_available():
        pytest.skip("GPU present: covered by the run test")
    model = configs.thrust_fault(nx=4, nz=2, ellipse=None)
    _write_mesh(tmp_path / "mesh.bin", model, 5)
    r = subprocess.run([exe, str(tmp_path / "mesh.bin"), str(tmp_path / "out.bin")],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert "no GPU context" in r.stderr


@pytest.mark.gpu
def test_fortran_host_matches_oracle(tmp_path):
    from conftest import gpu_available
    if not gpu_available():
        pytest.skip("no GPU")
    exe = _build(tmp_path)
    model = configs.thrust_fault(nx=8, nz=4, ellipse=None)
    nsteps = 60
    _write_mesh(tmp_path / "mesh.bin", model, nsteps)
    r = subprocess.run([exe, str(tmp_path / "mesh.bin"), str(tmp_path / "out.bin")],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    out = np.fromfile(tmp_path / "out.bin", dtype=np.float64)
    nd = 3 * len(POINTS)
    tl, dl, y = out[:nsteps], out[nsteps:2 * nsteps], out[2 * nsteps:2 * nsteps + 7 * model.n]
    disp = out[2 * nsteps + 7 * model.n:].reshape(3, nd)
    K = orc.buildG(orc.Mesh(model.lambda_, model.mu, rect=model.geometry()), nthreads=8)
    sysm = orc.System(model.friction(), K)
    y0 = sysm.init_state()
    yo, st, tlo, dlo, rc = sysm.integrate(y0, 1e30, max_steps=nsteps, log_cap=nsteps)
    np.testing.assert_allclose(dl, dlo, rtol=1e-5)
    np.testing.assert_allclose(tl, tlo, rtol=1e-5)
    np.testing.assert_allclose(y, yo, rtol=1e-6, atol=1e-9 * np.abs(yo).max())
    # exportPoints through the binding: d = O u of the final state; no volumes,
    # so the fault part is all of it
    O, _, _ = orc.buildO(orc.Mesh(model.lambda_, model.mu, rect=model.geometry()), POINTS)
    want = O @ orc.slip_vector(y, model.n, 0)
    np.testing.assert_allclose(disp[0], want, rtol=1e-6, atol=1e-9 * np.abs(want).max())
    np.testing.assert_array_equal(disp[1], disp[0])
    assert not disp[2].any() and np.abs(want).max() > 0


def _mixed_mesh(nsteps):
    """8 rectangles, 8 triangles (a second thrust 20 km north, split) and 4
    Maxwell volumes below both"""
    rect = configs.thrust_fault(nx=4, nz=2, ellipse=None)
    other = configs.thrust_fault(nx=2, nz=2, ellipse=None)
    other.x = other.x + np.array([20e3, 0.0, 0.0])
    tm = configs.triangulate(other)
    vol = configs.volume_grid(2, 2, 1, size=4000.0, top=15e3, origin=(0.0, 0.0))
    return rect, tm, vol


def _write_mixed(path, rect, tm, vol, nsteps):
    with open(path, "wb") as f:
        np.array([rect.n, tm.n, tm.vertices.shape[0], vol.n, nsteps], dtype=np.int32).tofile(f)
        np.array([rect.lambda_, rect.mu], dtype=np.float64).tofile(f)
        np.ascontiguousarray(rect.x, dtype=np.float64).tofile(f)
        for arr in (rect.length, rect.width, rect.strike, rect.dip):
            np.asarray(arr, dtype=np.float64).tofile(f)
        for arr in (tm.i1, tm.i2, tm.i3):
            np.asarray(arr, dtype=np.int32).tofile(f)
        np.ascontiguousarray(tm.vertices, dtype=np.float64).tofile(f)
        for a, b in zip(rect.friction(), tm.friction()):
            np.concatenate([a, b]).astype(np.float64).tofile(f)
        np.ascontiguousarray(vol.x, dtype=np.float64).tofile(f)
        for arr in (vol.length, vol.thickness, vol.width, vol.strike, vol.gammadot0, vol.npower):
            np.asarray(arr, dtype=np.float64).tofile(f)
        np.ascontiguousarray(vol.edot_bg, dtype=np.float64).tofile(f)   # e(6,nv)
        for arr in (vol.Gm, vol.Q, vol.R, vol.To):
            np.asarray(arr, dtype=np.float64).tofile(f)


def test_fortran_mixed_mesh_binding_links(tmp_path):
    """The mesh, rheology and frame wrappers compile and link"""
    exe = _build(tmp_path, "fortran_driver_viscous")
    from conftest import gpu_available
    if gpu_available():
        pytest.skip("GPU present: covered by the run test")
    rect, tm, vol = _mixed_mesh(5)
    _write_mixed(tmp_path / "mesh.bin", rect, tm, vol, 5)
    r = subprocess.run([exe, str(tmp_path / "mesh.bin"), str(tmp_path / "out.bin")],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert "no GPU context" in r.stderr


@pytest.mark.gpu
def test_fortran_mixed_mesh_host_matches_oracle(tmp_path):
    """Rectangles, triangles and Maxwell volumes set up from Fortran in the
    reference's array shapes: the library's frames (computeReferenceSystem*)
    bit for bit against the oracle's, and the viscouscycles main loop against
    the oracle's run of the same mesh."""
    from conftest import gpu_available
    if not gpu_available():
        pytest.skip("no GPU")
    exe = _build(tmp_path, "fortran_driver_viscous")
    nsteps = 40
    rect, tm, vol = _mixed_mesh(nsteps)
    _write_mixed(tmp_path / "mesh.bin", rect, tm, vol, nsteps)
    r = subprocess.run([exe, str(tmp_path / "mesh.bin"), str(tmp_path / "out.bin")],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    out = np.fromfile(tmp_path / "out.bin", dtype=np.float64)
    nr, nt, nv = rect.n, tm.n, vol.n
    o = 0
    frames = []
    for k in (nr, nt, nv):
        frames.append(out[o:o + 12 * k].reshape(4, k, 3))
        o += 12 * k
    tl, dl = out[o:o + nsteps], out[o + nsteps:o + 2 * nsteps]
    y = out[o + 2 * nsteps:]
    want = [np.zeros((4, k, 3)) for k in (nr, nt, nv)]
    lib = orc.oracle()
    lib.orc_refsys_okada92(nr, orc.p(orc.f64(rect.x)), orc.p(orc.f64(rect.length)),
                           orc.p(orc.f64(rect.width)), orc.p(orc.f64(rect.strike)),
                           orc.p(orc.f64(rect.dip)), *[orc.p(a) for a in want[0]])
    i1, i2, i3 = (np.ascontiguousarray(a, dtype=np.int32) for a in (tm.i1, tm.i2, tm.i3))
    lib.orc_refsys_nikkhoo15(tm.vertices.shape[0], orc.p(orc.f64(tm.vertices)), nt, orc.pi(i1),
                             orc.pi(i2), orc.pi(i3), *[orc.p(a) for a in want[1]])
    lib.orc_refsys_strainvolume(nv, orc.p(orc.f64(vol.x)), orc.p(orc.f64(vol.length)),
                                orc.p(orc.f64(vol.width)), orc.p(orc.f64(vol.strike)),
                                *[orc.p(a) for a in want[2]])
    for got, w in zip(frames, want):
        np.testing.assert_array_equal(got, w)
    mesh = orc.Mesh(rect.lambda_, rect.mu, rect=rect.geometry(), tri=tm.geometry(),
                    vol=vol.geometry())
    K = orc.buildG(mesh, nthreads=8)
    fric = [np.concatenate([a, b]) for a, b in zip(rect.friction(), tm.friction())]
    sysm = orc.System(fric, K, rheo=vol.rheology(), ic_variant=1)
    y0 = sysm.init_state()
    yo, st, tlo, dlo, rc = sysm.integrate(y0, 1e30, max_steps=nsteps, log_cap=nsteps)
    assert y.size == y0.size == 7 * (nr + nt) + 12 * nv
    np.testing.assert_allclose(dl, dlo, rtol=1e-5)
    np.testing.assert_allclose(tl, tlo, rtol=1e-5)
    np.testing.assert_allclose(y, yo, rtol=1e-6, atol=1e-9 * np.abs(yo).max())
