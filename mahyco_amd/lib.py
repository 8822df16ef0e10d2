"""ctypes binding of the HIP library (include/mahyco.h).

The library is built in-tree by mahyco_amd/build.py.  Loading fails loudly when it is
missing: there is no CPU fallback for the hydro path.
"""
import ctypes as C
import os

import numpy as np

from . import types as T

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("MAHYCO_HIP_LIB") or os.path.join(_HERE, "_build", "libmahyco_hip.so")
_lib = None


class MahycoError(RuntimeError):
    def __init__(self, status, what=""):
        self.status = status
        super().__init__(f"{what}: {T.STATUS_NAMES.get(status, status)}")


def load():
    """Loads libmahyco_hip.so (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"HIP library not built: {LIB_PATH} (run python -m mahyco_amd.build)")
    lib = C.CDLL(LIB_PATH)
    H = C.c_void_p
    lib.mahyco_create.argtypes = [C.POINTER(T.MhMesh), C.c_void_p, C.c_int, C.POINTER(T.MhOptions),
                                  C.POINTER(C.c_void_p)]
    lib.mahyco_create.restype = C.c_int
    lib.mahyco_create_on_stream.argtypes = [C.POINTER(T.MhMesh), C.c_void_p, C.c_int, C.POINTER(T.MhOptions),
                                            C.c_void_p, C.POINTER(C.c_void_p)]
    lib.mahyco_create_on_stream.restype = C.c_int
    lib.mahyco_apply_one_cell_eos.argtypes = [H, C.c_int, C.c_long]
    lib.mahyco_apply_one_cell_eos.restype = C.c_int
    lib.mahyco_ensemble_poll_async.argtypes = [C.c_void_p]
    lib.mahyco_ensemble_poll_async.restype = C.c_int
    lib.mahyco_ensemble_poll_result.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    lib.mahyco_ensemble_poll_result.restype = C.c_int
    lib.mahyco_autotune.argtypes = [H, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_double)]
    lib.mahyco_autotune.restype = C.c_int
    lib.mahyco_get_variant.argtypes = [H]
    lib.mahyco_get_variant.restype = C.c_int
    lib.mahyco_set_variant.argtypes = [H, C.c_int]
    lib.mahyco_set_variant.restype = C.c_int
    lib.mahyco_set_staged.argtypes = [H, C.c_int]
    lib.mahyco_set_staged.restype = C.c_int
    lib.mahyco_ensemble_autotune.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_double)]
    lib.mahyco_ensemble_autotune.restype = C.c_int
    lib.mahyco_ensemble_graph_captures.argtypes = [C.c_void_p]
    lib.mahyco_ensemble_graph_captures.restype = C.c_long
    lib.mahyco_set_field.argtypes = [H, C.c_int, C.c_int, C.c_int, C.c_void_p]
    lib.mahyco_set_field.restype = C.c_int
    lib.mahyco_set_scalars.argtypes = [H, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int]
    lib.mahyco_set_scalars.restype = C.c_int
    lib.mahyco_destroy.argtypes = [H]
    lib.mahyco_destroy.restype = None
    for name in ("mahyco_cycle",):
        getattr(lib, name).argtypes = [H, C.c_int]
        getattr(lib, name).restype = C.c_int
    lib.mahyco_cycle_graph.argtypes = [H, C.c_int, C.c_int]
    lib.mahyco_cycle_graph.restype = C.c_int
    for name in ("mahyco_synchronize", "mahyco_status", "mahyco_iteration",
                 "mahyco_save_values_at_n", "mahyco_update_velocity", "mahyco_apply_boundary_condition",
                 "mahyco_update_position", "mahyco_compute_geometric_values", "mahyco_update_density",
                 "mahyco_compute_artificial_viscosity", "mahyco_update_elasticity_and_plasticity",
                 "mahyco_update_energy_and_pressure", "mahyco_remap", "mahyco_compute_delta_t"):
        getattr(lib, name).argtypes = [H]
        getattr(lib, name).restype = C.c_int
    for name in ("mahyco_time", "mahyco_deltat"):
        getattr(lib, name).argtypes = [H]
        getattr(lib, name).restype = C.c_double
    for name in ("mahyco_nb_cells", "mahyco_nb_nodes", "mahyco_graph_captures"):
        getattr(lib, name).argtypes = [H]
        getattr(lib, name).restype = C.c_long
    lib.mahyco_get_field.argtypes = [H, C.c_int, C.c_int, C.c_int, C.c_void_p]
    lib.mahyco_get_field.restype = C.c_int
    lib.mahyco_profile_enable.argtypes = [H, C.c_int]
    lib.mahyco_profile_enable.restype = C.c_int
    lib.mahyco_profile_read.argtypes = [H, C.c_char_p, C.c_int, C.c_void_p, C.c_void_p, C.c_int]
    lib.mahyco_profile_read.restype = C.c_int
    lib.mahyco_case_footprint.argtypes = [C.POINTER(T.MhMesh), C.c_int, C.POINTER(T.MhOptions)]
    lib.mahyco_case_footprint.restype = C.c_size_t
    lib.mahyco_ensemble_create.argtypes = [C.c_int, C.c_double, C.POINTER(C.c_void_p)]
    lib.mahyco_ensemble_create.restype = C.c_int
    lib.mahyco_ensemble_destroy.argtypes = [H]
    lib.mahyco_ensemble_destroy.restype = None
    lib.mahyco_ensemble_add.argtypes = [H, C.POINTER(T.MhMesh), C.c_void_p, C.c_int, C.POINTER(T.MhOptions),
                                        C.c_int, C.POINTER(C.c_int)]
    lib.mahyco_ensemble_add.restype = C.c_int
    lib.mahyco_ensemble_remove.argtypes = [H, C.c_int]
    lib.mahyco_ensemble_remove.restype = C.c_int
    lib.mahyco_ensemble_step.argtypes = [H, C.c_int, C.c_int]
    lib.mahyco_ensemble_step.restype = C.c_int
    lib.mahyco_ensemble_synchronize.argtypes = [H]
    lib.mahyco_ensemble_synchronize.restype = C.c_int
    lib.mahyco_ensemble_poll.argtypes = [H, C.c_void_p, C.c_void_p, C.c_void_p]
    lib.mahyco_ensemble_poll.restype = C.c_int
    lib.mahyco_ensemble_get_field.argtypes = [H, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p]
    lib.mahyco_ensemble_get_field.restype = C.c_int
    lib.mahyco_ensemble_used_bytes.argtypes = [H]
    lib.mahyco_ensemble_used_bytes.restype = C.c_double
    lib.mahyco_history_enable.argtypes = [H, C.c_int, C.c_int]
    lib.mahyco_history_enable.restype = C.c_int
    lib.mahyco_history_read.argtypes = [H, C.c_void_p, C.c_int]
    lib.mahyco_history_read.restype = C.c_int
    lib.mahyco_ensight_write.argtypes = [H, C.c_char_p, C.POINTER(C.c_char_p), C.c_int]
    lib.mahyco_ensight_write.restype = C.c_int
    lib.mahyco_history_write_gnuplot.argtypes = [H, C.c_char_p, C.POINTER(C.c_char_p)]
    lib.mahyco_history_write_gnuplot.restype = C.c_int
    for name in ("mahyco_checkpoint_write", "mahyco_checkpoint_read"):
        getattr(lib, name).argtypes = [H, C.c_char_p]
        getattr(lib, name).restype = C.c_int
    lib.mahyco_continue_init.argtypes = [H]
    lib.mahyco_continue_init.restype = C.c_int
    lib.mahyco_apply_eos.argtypes = [H, C.c_int]
    lib.mahyco_apply_eos.restype = C.c_int
    lib.mahyco_compute_pression_moyenne.argtypes = [H]
    lib.mahyco_compute_pression_moyenne.restype = C.c_int
    _lib = lib
    return lib


class GpuCase:
    """One case resident on the GPU (the MahycoModule state of one mesh)."""

    def __init__(self, case, stream=None):
        """stream: a hipStream_t (as an int or c_void_p) the case's work is ordered on."""
        self.lib = load()
        self.case = case
        self._mesh, self._envs, self._opt = case.mesh_c(), case.envs_c(), case.options_c()
        h = C.c_void_p()
        if stream is None:
            st = self.lib.mahyco_create(C.byref(self._mesh), self._envs, len(case.envs), C.byref(self._opt),
                                        C.byref(h))
        else:
            st = self.lib.mahyco_create_on_stream(C.byref(self._mesh), self._envs, len(case.envs),
                                                  C.byref(self._opt), C.c_void_p(int(stream)), C.byref(h))
        if st != T.MH_OK:
            raise MahycoError(st, "mahyco_create")
        self.h = h
        self.nc = self.lib.mahyco_nb_cells(h)
        self.nn = self.lib.mahyco_nb_nodes(h)

    def close(self):
        if getattr(self, "h", None):
            self.lib.mahyco_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, ncycles, graph=0):
        """Queues ncycles iterations (graph > 0: replay hipGraphs of that many cycles)."""
        st = (self.lib.mahyco_cycle_graph(self.h, ncycles, graph) if graph
              else self.lib.mahyco_cycle(self.h, ncycles))
        if st != T.MH_OK:
            raise MahycoError(st, "mahyco_cycle")
        return self.synchronize()

    def queue(self, ncycles, graph=0):
        st = (self.lib.mahyco_cycle_graph(self.h, ncycles, graph) if graph
              else self.lib.mahyco_cycle(self.h, ncycles))
        if st != T.MH_OK:
            raise MahycoError(st, "mahyco_cycle")

    def synchronize(self):
        return self.lib.mahyco_synchronize(self.h)

    @property
    def time(self):
        return self.lib.mahyco_time(self.h)

    @property
    def deltat(self):
        return self.lib.mahyco_deltat(self.h)

    @property
    def iteration(self):
        return self.lib.mahyco_iteration(self.h)

    @property
    def graph_captures(self):
        return self.lib.mahyco_graph_captures(self.h)

    def get(self, field, env=-1, comp=0):
        n = self.nn if field >= 100 else self.nc
        out = np.empty(n, dtype=np.float64)
        st = self.lib.mahyco_get_field(self.h, field, env, comp, out.ctypes.data)
        if st != T.MH_OK:
            raise MahycoError(st, "mahyco_get_field")
        return out

    def set(self, field, values, env=-1, comp=0):
        """Writes a field of the state (mahyco_set_field)."""
        a = np.ascontiguousarray(values, dtype=np.float64)
        n = self.nn if field >= 100 else self.nc
        if a.shape != (n,):
            raise ValueError(f"field {field}: expected {n} values, got {a.shape}")
        st = self.lib.mahyco_set_field(self.h, field, env, comp, a.ctypes.data)
        if st != T.MH_OK:
            raise MahycoError(st, "mahyco_set_field")

    def autotune(self, cycles=12):
        """Times the kernel variants on the case's own iterations and keeps the fastest
        (mahyco_autotune); returns (variant, ms per iteration)."""
        v, ms = C.c_int(-1), C.c_double(0.)
        st = self.lib.mahyco_autotune(self.h, cycles, C.byref(v), C.byref(ms))
        if st != T.MH_OK:
            raise MahycoError(st, "mahyco_autotune")
        return v.value, ms.value

    @property
    def variant(self):
        """Kernel-variant code of the case (mahyco_get_variant)."""
        return self.lib.mahyco_get_variant(self.h)

    def set_variant(self, variant):
        st = self.lib.mahyco_set_variant(self.h, int(variant))
        if st != T.MH_OK:
            raise MahycoError(st, "mahyco_set_variant")

    def set_staged(self, on=True):
        """Staged Lagrange step (mahyco_set_staged): the state of UpdateEnergyAndPressure is
        stored and the remap vector is formed from it at the remap, so that host code run
        between the two (step(hook=...)) can read and change it."""
        st = self.lib.mahyco_set_staged(self.h, 1 if on else 0)
        if st != T.MH_OK:
            raise MahycoError(st, "mahyco_set_staged")

    ENTRY_POINTS = ("mahyco_save_values_at_n", "mahyco_update_velocity", "mahyco_apply_boundary_condition",
                    "mahyco_update_position", "mahyco_compute_geometric_values", "mahyco_update_density",
                    "mahyco_compute_artificial_viscosity", "mahyco_update_elasticity_and_plasticity",
                    "mahyco_update_energy_and_pressure")

    def step(self, hook=None):
        """One iteration through the module's entry points in the time loop's order; hook(self)
        runs between UpdateEnergyAndPressure and Remap (in staged mode it may read and change
        the cell state with get / set)."""
        for fn in self.ENTRY_POINTS:
            st = getattr(self.lib, fn)(self.h)
            if st != T.MH_OK:
                raise MahycoError(st, fn)
        if hook is not None:
            hook(self)
        if self.case.with_projection:
            st = self.lib.mahyco_remap(self.h)
            if st != T.MH_OK:
                raise MahycoError(st, "mahyco_remap")
        st = self.lib.mahyco_compute_delta_t(self.h)
        if st != T.MH_OK:
            raise MahycoError(st, "mahyco_compute_delta_t")

    def set_scalars(self, time, deltat, old_deltat, iteration, sens_projection):
        st = self.lib.mahyco_set_scalars(self.h, time, deltat, old_deltat, iteration, sens_projection)
        if st != T.MH_OK:
            raise MahycoError(st, "mahyco_set_scalars")

    def history_enable(self, period=10, max_rows=4096):
        st = self.lib.mahyco_history_enable(self.h, period, max_rows)
        if st != T.MH_OK:
            raise MahycoError(st, "mahyco_history_enable")
        self._hist_max = max_rows

    def history(self):
        """EnvSummation rows, same layout as the oracle's."""
        w = 1 + 5 * len(self.case.envs)
        n = self.lib.mahyco_history_read(self.h, None, 0)
        buf = np.zeros(max(n, 1) * w)
        self.lib.mahyco_history_read(self.h, buf.ctypes.data, n)
        return buf[:n * w].reshape(n, w)

    def ensight_write(self, directory, variables):
        """Appends the current state as the next EnSight Gold step under directory."""
        arr = (C.c_char_p * len(variables))(*[v.encode() for v in variables])
        st = self.lib.mahyco_ensight_write(self.h, str(directory).encode(), arr, len(variables))
        if st != T.MH_OK:
            raise MahycoError(st, "mahyco_ensight_write")

    def history_write_gnuplot(self, directory, env_names):
        """Writes the EnvSummation histories as <Var>_<environment> gnuplot files."""
        arr = (C.c_char_p * len(env_names))(*[v.encode() for v in env_names])
        st = self.lib.mahyco_history_write_gnuplot(self.h, str(directory).encode(), arr)
        if st != T.MH_OK:
            raise MahycoError(st, "mahyco_history_write_gnuplot")

    def checkpoint_write(self, path):
        st = self.lib.mahyco_checkpoint_write(self.h, str(path).encode())
        if st != T.MH_OK:
            raise MahycoError(st, "mahyco_checkpoint_write")

    def checkpoint_read(self, path):
        st = self.lib.mahyco_checkpoint_read(self.h, str(path).encode())
        if st != T.MH_OK:
            raise MahycoError(st, "mahyco_checkpoint_read")

    def continue_init(self):
        """Continue as the reference does after a restart (hydroContinueInit)."""
        st = self.lib.mahyco_continue_init(self.h)
        if st != T.MH_OK:
            raise MahycoError(st, "mahyco_continue_init")

    def profile(self, on=True):
        self.lib.mahyco_profile_enable(self.h, int(on))

    def profile_read(self):
        names = C.create_string_buffer(1024)
        ms = np.zeros(16)
        cnt = np.zeros(16, dtype=np.int32)
        n = self.lib.mahyco_profile_read(self.h, names, 1024, ms.ctypes.data, cnt.ctypes.data, 16)
        lst = names.raw.split(b"\0")[:n]
        return {lst[i].decode(): (float(ms[i]), int(cnt[i])) for i in range(n)}


def footprint(case):
    """Device bytes a case needs (the admission measure of ensembles)."""
    lib = load()
    m, o = case.mesh_c(), case.options_c()
    return int(lib.mahyco_case_footprint(C.byref(m), len(case.envs), C.byref(o)))


class Ensemble:
    """Many independent cases advancing together (batched launches, per-case time step)."""

    def __init__(self, max_slots, hbm_budget_bytes):
        self.lib = load()
        e = C.c_void_p()
        st = self.lib.mahyco_ensemble_create(max_slots, float(hbm_budget_bytes), C.byref(e))
        if st != T.MH_OK:
            raise MahycoError(st, "mahyco_ensemble_create")
        self.e = e
        self.max_slots = max_slots
        self._keep = {}
        self.cases = {}

    def close(self):
        if getattr(self, "e", None):
            self.lib.mahyco_ensemble_destroy(self.e)
            self.e = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add(self, case, max_iterations=0):
        """Returns the slot, or None when the budget or the slots are exhausted."""
        m, envs, o = case.mesh_c(), case.envs_c(), case.options_c()
        slot = C.c_int(-1)
        st = self.lib.mahyco_ensemble_add(self.e, C.byref(m), envs, len(case.envs), C.byref(o),
                                          int(max_iterations), C.byref(slot))
        if st == T.MH_ERR_NOMEM:
            return None
        if st != T.MH_OK:
            raise MahycoError(st, "mahyco_ensemble_add")
        self._keep[slot.value] = (m, envs, o)
        self.cases[slot.value] = case
        return slot.value

    def remove(self, slot):
        st = self.lib.mahyco_ensemble_remove(self.e, slot)
        if st != T.MH_OK:
            raise MahycoError(st, "mahyco_ensemble_remove")
        self._keep.pop(slot, None)
        self.cases.pop(slot, None)

    def step(self, ncycles, graph=True):
        st = self.lib.mahyco_ensemble_step(self.e, ncycles, int(graph))
        if st != T.MH_OK:
            raise MahycoError(st, "mahyco_ensemble_step")

    def synchronize(self):
        return self.lib.mahyco_ensemble_synchronize(self.e)

    @property
    def graph_captures(self):
        return self.lib.mahyco_ensemble_graph_captures(self.e)

    def autotune(self, cycles=12):
        """Times the persistent grid sizes on the ensemble's own iterations and keeps the
        fastest (mahyco_ensemble_autotune); returns (blocks per CU, ms per iteration)."""
        b, ms = C.c_int(0), C.c_double(0.)
        st = self.lib.mahyco_ensemble_autotune(self.e, cycles, C.byref(b), C.byref(ms))
        if st != T.MH_OK:
            raise MahycoError(st, "mahyco_ensemble_autotune")
        return b.value, ms.value

    def poll(self):
        st = np.zeros(self.max_slots, dtype=np.int32)
        tm = np.zeros(self.max_slots)
        it = np.zeros(self.max_slots, dtype=np.int32)
        n = self.lib.mahyco_ensemble_poll(self.e, st.ctypes.data, tm.ctypes.data, it.ctypes.data)
        return n, st, tm, it

    def poll_async(self):
        """Queues a snapshot of every slot's scalars behind the queued work (no wait)."""
        st = self.lib.mahyco_ensemble_poll_async(self.e)
        if st != T.MH_OK:
            raise MahycoError(st, "mahyco_ensemble_poll_async")

    def poll_result(self):
        """Waits for the last poll_async snapshot only; same result as poll()."""
        st = np.zeros(self.max_slots, dtype=np.int32)
        tm = np.zeros(self.max_slots)
        it = np.zeros(self.max_slots, dtype=np.int32)
        n = self.lib.mahyco_ensemble_poll_result(self.e, st.ctypes.data, tm.ctypes.data, it.ctypes.data)
        if n < 0:
            raise MahycoError(n, "mahyco_ensemble_poll_result")
        return n, st, tm, it

    def get(self, slot, field, env=-1, comp=0):
        case = self.cases[slot]
        n = case.nb_nodes if field >= 100 else case.nb_cells
        out = np.empty(n)
        st = self.lib.mahyco_ensemble_get_field(self.e, slot, field, env, comp, out.ctypes.data)
        if st != T.MH_OK:
            raise MahycoError(st, "mahyco_ensemble_get_field")
        return out

    @property
    def used_bytes(self):
        return self.lib.mahyco_ensemble_used_bytes(self.e)
