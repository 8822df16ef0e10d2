"""ctypes binding of the CPU oracle (test infrastructure only).

Used by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg, for parity
checks and the CPU timing; never by the product path.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from mahyco_amd import types as T

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libmahyco_oracle.so")
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    lib = C.CDLL(LIB_PATH)
    lib.oracle_create.restype = C.c_void_p
    lib.oracle_create.argtypes = [C.POINTER(T.MhMesh), C.c_void_p, C.c_int, C.POINTER(T.MhOptions)]
    lib.oracle_destroy.argtypes = [C.c_void_p]
    lib.oracle_run.argtypes = [C.c_void_p, C.c_int]
    lib.oracle_run.restype = C.c_int
    lib.oracle_continue.argtypes = [C.c_void_p]
    lib.oracle_continue.restype = None
    for n in ("oracle_time", "oracle_deltat"):
        getattr(lib, n).restype = C.c_double
        getattr(lib, n).argtypes = [C.c_void_p]
    lib.oracle_iteration.argtypes = [C.c_void_p]
    lib.oracle_iteration.restype = C.c_int
    for n in ("oracle_nb_cells", "oracle_nb_nodes"):
        getattr(lib, n).restype = C.c_long
        getattr(lib, n).argtypes = [C.c_void_p]
    lib.oracle_get_field.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p]
    lib.oracle_get_field.restype = C.c_int
    lib.oracle_history.argtypes = [C.c_void_p, C.c_void_p, C.c_int]
    lib.oracle_history.restype = C.c_int
    lib.oracle_set_field.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p]
    lib.oracle_set_field.restype = C.c_int
    lib.oracle_get_scalars.argtypes = [C.c_void_p] + [C.c_void_p] * 5
    lib.oracle_get_scalars.restype = None
    lib.oracle_set_scalars.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int]
    lib.oracle_set_scalars.restype = None
    lib.oracle_apply_eos.argtypes = [C.c_void_p, C.c_int]
    lib.oracle_apply_eos.restype = None
    lib.oracle_compute_pression_moyenne.argtypes = [C.c_void_p]
    lib.oracle_compute_pression_moyenne.restype = None
    lib.oracle_set_threads.argtypes = [C.c_int]
    lib.oracle_set_pow_mode.argtypes = [C.c_int]
    _lib = lib
    return lib


class OracleCase:
    """CPU restatement of one case; same accessors as mahyco_amd.lib.GpuCase."""

    def __init__(self, case, pow_mode=0, threads=0):
        self.lib = load()
        self.case = case
        self.lib.oracle_set_pow_mode(pow_mode)
        if threads:
            self.lib.oracle_set_threads(threads)
        self._mesh, self._envs, self._opt = case.mesh_c(), case.envs_c(), case.options_c()
        self.h = self.lib.oracle_create(C.byref(self._mesh), self._envs, len(case.envs), C.byref(self._opt))
        if not self.h:
            raise RuntimeError("oracle_create rejected the case")
        self.nc = self.lib.oracle_nb_cells(self.h)
        self.nn = self.lib.oracle_nb_nodes(self.h)

    def close(self):
        if getattr(self, "h", None):
            self.lib.oracle_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, ncycles):
        return self.lib.oracle_run(self.h, ncycles)

    def continue_after_restart(self):
        """From here on the run behaves as one continued from a dump (hydroContinueInit)."""
        self.lib.oracle_continue(self.h)

    @property
    def time(self):
        return self.lib.oracle_time(self.h)

    @property
    def deltat(self):
        return self.lib.oracle_deltat(self.h)

    @property
    def iteration(self):
        return self.lib.oracle_iteration(self.h)

    def get(self, field, env=-1, comp=0):
        n = self.nn if field >= 100 else self.nc
        out = np.empty(n, dtype=np.float64)
        st = self.lib.oracle_get_field(self.h, field, env, comp, out.ctypes.data)
        if st != T.MH_OK:
            raise RuntimeError(f"oracle_get_field({field}) -> {st}")
        return out

    def set(self, field, values, env=-1, comp=0):
        a = np.ascontiguousarray(values, dtype=np.float64)
        st = self.lib.oracle_set_field(self.h, field, env, comp, a.ctypes.data)
        if st != T.MH_OK:
            raise RuntimeError(f"oracle_set_field({field}) -> {st}")

    def scalars(self):
        """(time, deltat, old deltat, iteration, sens_projection)."""
        t, dt, dto = C.c_double(), C.c_double(), C.c_double()
        it, sens = C.c_int(), C.c_int()
        self.lib.oracle_get_scalars(self.h, C.byref(t), C.byref(dt), C.byref(dto), C.byref(it), C.byref(sens))
        return t.value, dt.value, dto.value, it.value, sens.value

    def set_scalars(self, time, deltat, old_deltat, iteration, sens_projection):
        self.lib.oracle_set_scalars(self.h, time, deltat, old_deltat, iteration, sens_projection)

    def apply_eos(self, env):
        self.lib.oracle_apply_eos(self.h, env)

    def compute_pression_moyenne(self):
        self.lib.oracle_compute_pression_moyenne(self.h)

    def history(self):
        rows = self.lib.oracle_history(self.h, None, 0)
        w = 1 + 5 * len(self.case.envs)
        buf = np.zeros(max(rows, 1) * w)
        self.lib.oracle_history(self.h, buf.ctypes.data, rows)
        return buf[:rows * w].reshape(rows, w)
