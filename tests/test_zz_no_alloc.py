"""mahyco_compute_delta_t_noalloc (the time step from caller-owned device arrays, with its
scratch in a static device array and no allocation) against the handle's own ComputeDeltaT
(mahyco_compute_delta_t, the step the cycle computes at the end of an iteration): bitwise
the same step on two calls in a row, the second on a smaller mesh; MH_ERR_DT_MIN below
deltat-min; meshes beyond the scratch capacity refused before any launch."""
import ctypes as C

import numpy as np
import pytest

from mahyco_amd import cases
from mahyco_amd import types as T

pytestmark = pytest.mark.gpu


def _hip():
    hip = C.CDLL("libamdhip64.so")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipFree.argtypes = [C.c_void_p]
    return hip


class _Dev:
    """Device copies of host arrays, made and freed by the test (the entry point allocates nothing)."""

    def __init__(self, hip):
        self.hip, self.ptrs = hip, []

    def put(self, arr):
        a = np.ascontiguousarray(arr)
        p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(p), max(a.nbytes, 8)) == 0
        assert self.hip.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # host to device
        self.ptrs.append(p)
        return p

    def get(self, p, dtype, n=1):
        out = np.empty(n, dtype=dtype)
        assert self.hip.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0  # device to host
        return out

    def free(self):
        for p in self.ptrs:
            self.hip.hipFree(p)


def _bind(lib):
    f = lib.mahyco_compute_delta_t_noalloc
    f.argtypes = [C.POINTER(T.MhMesh), C.POINTER(T.MhOptions), C.c_double] + [C.c_void_p] * 8
    f.restype = C.c_int
    return f


def _noalloc_step(case, n):
    """(step of the handle after n iterations, step of the entry point on its L, c, u, status)."""
    from mahyco_amd.lib import GpuCase
    hip = _hip()
    g = GpuCase(case)
    fn = _bind(g.lib)
    assert g.run(n - 1) == T.MH_OK
    old = g.deltat  # the step iteration n runs with
    assert g.run(1) == T.MH_OK
    ref = g.deltat  # computed at the end of iteration n by the cycle's own kernels
    L = g.get(T.MH_F_CARACTERISTIC_LENGTH)
    c = g.get(T.MH_F_SOUND_SPEED)
    u = [g.get(T.MH_N_VELOCITY, -1, k) for k in range(case.dim)]
    d = _Dev(hip)
    try:
        pL, pc = d.put(L), d.put(c)
        pu = [d.put(x) for x in u] + ([None] if case.dim == 2 else [])
        pdt, pst = d.put(np.zeros(1)), d.put(np.full(1, 7, dtype=np.int32))
        mesh, opt = case.mesh_c(), case.options_c()
        st = fn(C.byref(mesh), C.byref(opt), old, pL, pc, pu[0], pu[1], pu[2], pdt, pst, None)
        assert st == T.MH_OK
        assert hip.hipDeviceSynchronize() == 0
        dt = float(d.get(pdt, np.float64)[0])
        status = int(d.get(pst, np.int32)[0])
        # the same arrays with deltat-min above the step
        opt.deltat_min = 2.0 * ref
        st = fn(C.byref(mesh), C.byref(opt), old, pL, pc, pu[0], pu[1], pu[2], pdt, pst, None)
        assert st == T.MH_OK and hip.hipDeviceSynchronize() == 0
        status_min = int(d.get(pst, np.int32)[0])
    finally:
        d.free()
        g.close()
    return ref, dt, status, status_min


def test_noalloc_delta_t_equals_handle():
    print("[noalloc] 3D 64x32x32 then 2D-like 30x4x3", flush=True)
    for case, n in ((cases.config2_bisod3d(64, 32, 32), 13), (cases.bisod_x(30, 4, 3), 7),
                    (cases.lag2d_sod_x(), 9)):
        ref, dt, status, status_min = _noalloc_step(case, n)
        assert float(dt).hex() == float(ref).hex(), (case.name, dt, ref)
        assert status == T.MH_OK and status_min == T.MH_ERR_DT_MIN, case.name


def test_noalloc_refuses_beyond_capacity():
    """A mesh of more cells than the scratch array holds partials for is refused on the host
    (the pointers are never read)."""
    from mahyco_amd.lib import load
    fn = _bind(load())
    case = cases.config2_bisod3d(512, 256, 256)  # 33.5 M cells
    mesh, opt = case.mesh_c(), case.options_c()
    dummy = C.c_void_p(8)
    st = fn(C.byref(mesh), C.byref(opt), 1e-3, dummy, dummy, dummy, dummy, dummy, dummy, dummy, None)
    assert st == T.MH_ERR_INVALID
