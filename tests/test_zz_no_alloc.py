"""ucy_odefun_dev_static, the odefun entry point that allocates nothing (its
slip-rate vector and vMax accumulator live in a static device array per
GPU), against ucy_odefun on the same context and state: the same kernels
on the same inputs, so the results are bit-identical.  Three calls in a
row on three contexts, each smaller than the one before (the static array
keeps the last call's contents): fault + strain volumes, a thrust fault,
and a 2-D antiplane fault with an odd segment count (whose kernel carries a
zero padding column inside the static array).  Device buffers come from the
HIP runtime that libucy.so itself links."""
import ctypes as C

import numpy as np
import pytest

from unicycle_amd import configs

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def _hip():
    """The HIP runtime already loaded by libucy.so (same handle)."""
    path = None
    with open("/proc/self/maps") as f:
        for line in f:
            if "libamdhip64.so" in line:
                path = line.split()[-1]
                break
    assert path, "libamdhip64 not loaded"
    hip = C.CDLL(path)
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
    hip.hipStreamSynchronize.argtypes = [C.c_void_p]
    hip.hipStreamDestroy.argtypes = [C.c_void_p]
    return hip


class DevBuf:
    def __init__(self, hip, n):
        self.hip, self.n, self.p = hip, n, C.c_void_p()
        assert hip.hipMalloc(C.byref(self.p), 8 * n) == 0

    def put(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert a.size == self.n
        assert self.hip.hipMemcpy(self.p, a.ctypes.data, 8 * self.n, H2D) == 0

    def get(self):
        a = np.empty(self.n)
        assert self.hip.hipMemcpy(a.ctypes.data, self.p, 8 * self.n, D2H) == 0
        return a

    def free(self):
        self.hip.hipFree(self.p)


def _viscous(eng):
    fault = configs.thrust_fault(nx=6, nz=3, ellipse=None)
    vol = configs.volume_grid(3, 3, 2, size=4000.0, top=20e3, origin=(0.0, 0.0), npower=3.0)
    eng.set_elastic(fault.lambda_, fault.mu)
    eng.set_rectangles(*fault.geometry())
    eng.set_volumes(*vol.geometry())
    eng.set_friction(*fault.friction())
    eng.set_rheology(*vol.rheology())
    eng.buildG()
    y = eng.initStateVector(1)
    rng = np.random.default_rng(4)
    y[6:7 * fault.n:7] += 0.3 * rng.standard_normal(fault.n)
    y[7 * fault.n:] *= 1.0 + 0.1 * rng.standard_normal(12 * vol.n)
    return y


def _thrust(eng):
    m = configs.thrust_fault(nx=8, nz=4, ellipse=None)
    eng.set_elastic(m.lambda_, m.mu)
    eng.set_rectangles(*m.geometry())
    eng.set_friction(*m.friction())
    eng.buildG()
    y = eng.initStateVector()
    rng = np.random.default_rng(5)
    y[6::7] += 0.5 * rng.standard_normal(m.n)
    y[5::7] += 0.3 * rng.standard_normal(m.n)
    return y


def _antiplane(eng):
    m = configs.antiplane_fault(n=7)
    eng.set_antiplane(m.mu, m.x, m.width, m.dip)
    eng.set_friction(*m.friction())
    eng.buildG()
    y = eng.initStateVector()
    rng = np.random.default_rng(6)
    y[3::4] += 0.5 * rng.standard_normal(m.n)
    return y


def test_static_odefun_matches_odefun(engine_factory):
    hip = _hip()
    sizes = []
    for setup in (_viscous, _thrust, _antiplane):
        eng = engine_factory()
        y = setup(eng)
        n = y.size
        sizes.append(n)
        want = np.empty(n)
        vmax_want = eng.odefun(n, 0.0, y, want)
        L, h = eng._L, eng._h
        yd, dd, vd = DevBuf(hip, n), DevBuf(hip, n), DevBuf(hip, 1)
        yd.put(y)
        dd.put(np.full(n, np.nan))
        # NULL stream: the context's stream, the call waits for the result
        rc = L.ucy_odefun_dev_static(h, n, yd.p, dd.p, vd.p, None)
        assert rc == 0, L.ucy_last_error(h)
        got, vmax = dd.get(), vd.get()[0]
        assert np.array_equal(got, want), f"{setup.__name__}: max diff {np.abs(got - want).max()}"
        assert vmax == vmax_want
        # a caller's stream: the call only enqueues; no vMax output
        st = C.c_void_p()
        assert hip.hipStreamCreate(C.byref(st)) == 0
        dd.put(np.full(n, np.nan))
        assert L.ucy_odefun_dev_static(h, n, yd.p, dd.p, None, st) == 0
        assert hip.hipStreamSynchronize(st) == 0
        hip.hipStreamDestroy(st)
        assert np.array_equal(dd.get(), want)
        # a wrong state length is an argument error, not a launch
        assert L.ucy_odefun_dev_static(h, n + 1, yd.p, dd.p, None, None) < 0
        for b in (yd, dd, vd):
            b.free()
    assert sizes[0] > sizes[1] > sizes[2]
