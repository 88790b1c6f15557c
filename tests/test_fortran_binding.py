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


def _build(tmp):
    if FC is None:
        pytest.skip("no Fortran compiler")
    if not os.path.exists(os.path.join(LIB, "libucy.so")):
        pytest.skip("libucy.so not built")
    exe = os.path.join(tmp, "fortran_driver")
    cmd = [FC, "-O1", "-J" if "gfortran" in FC else "-module-dir", str(tmp),
           os.path.join(ROOT, "unicycle_amd", "fortran", "ucy_fortran.f90"),
           os.path.join(ROOT, "tests", "fortran_driver.f90"), "-o", exe,
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
