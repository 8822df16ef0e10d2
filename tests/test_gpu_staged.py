"""Staged Lagrange step (mahyco_set_staged), GPU tests: the cell kernel stores the state of
UpdateEnergyAndPressure and the remap forms its vector from it (computeVariablesForRemap,
PrepareRemap.cc:100-155), so host code between the two entry points sees and may change the
state, as a module entry point placed there in the reference's time loop does.

* Unchanged state: the same bits as the fused kernels, through mahyco_cycle, graph replays and
  the entry points one by one.
* The state seen between the energy update and the remap is the Lagrange step's: on the first
  iteration it is bitwise the state a Lagrange-only run of the same case holds after one
  iteration (both start from the same state with the same time step).
* A change made there reaches the remap: scaling every environment's internal energy by
  (1 + eps) scales the remapped energies by (1 + eps) (the remap is linear in the energy and
  its masses and volumes do not read it) and leaves density, masses and velocity bitwise alone.
"""
import copy

import numpy as np
import pytest

from mahyco_amd import cases
from mahyco_amd import types as T
from tests.helpers import bitwise_equal, fields

pytestmark = pytest.mark.gpu

CELL_STATE = (T.MH_F_DENSITY, T.MH_F_INTERNAL_ENERGY, T.MH_F_PRESSURE, T.MH_F_SOUND_SPEED,
              T.MH_F_PSEUDO_VISCOSITY, T.MH_F_FRACVOL)


def _pb3d():
    c = cases.riemann3d(12, np.random.default_rng(5))
    c.pente_borne = True
    c.debar_fix = 1
    return c


def _newton2d():
    c = cases.bitriplepoint()
    c.with_newton = True
    return c


MAKES = {
    "bisod3d": lambda: cases.bisod_x(30, 4, 3),
    "bitriplepoint": cases.bitriplepoint,
    "sedov2d": cases.sedov_2d,
    "riemann3d_pente_borne": _pb3d,
    "bitriplepoint_newton": _newton2d,
    "bisod3d_ordre3": lambda: cases.ordre3(lambda: cases.bisod_x(20, 5, 4), "b3d_o3"),
}


@pytest.mark.parametrize("name", sorted(MAKES))
def test_staged_equals_fused(name):
    from mahyco_amd.lib import GpuCase
    case = MAKES[name]()
    print(f"[staged] {name}", flush=True)
    n = 18
    ref = GpuCase(case)
    ref.run(n, graph=6)
    fr = fields(ref, len(case.envs), case.dim)
    runs = {}
    for how in ("cycle", "graph", "steps"):
        g = GpuCase(case)
        g.set_staged(True)
        if how == "cycle":
            g.run(n)
        elif how == "graph":
            g.run(n, graph=6)
        else:
            for _ in range(n):
                g.step()
            g.synchronize()
        runs[how] = (fields(g, len(case.envs), case.dim), g.time)
        g.close()
    for how, (f, t) in runs.items():
        assert t == ref.time, how
        for k in fr:
            assert bitwise_equal(fr[k], f[k]), (how, k)
    ref.close()


@pytest.mark.parametrize("name", ["bisod3d", "bitriplepoint", "sedov2d"])
def test_staged_state_is_the_lagrange_step(name):
    """First iteration: the state between UpdateEnergyAndPressure and Remap equals a
    Lagrange-only run's state after its first iteration, per cell and per environment."""
    from mahyco_amd.lib import GpuCase
    case = MAKES[name]()
    lag = copy.deepcopy(case)
    lag.with_projection = False
    gl = GpuCase(lag)
    gl.run(1)
    m = len(case.envs)
    want = {(f, k): gl.get(f, k) for f in CELL_STATE for k in range(-1, m)}
    gl.close()
    g = GpuCase(case)
    g.set_staged(True)
    seen = {}

    def hook(gc):
        for f in CELL_STATE:
            for k in range(-1, m):
                seen[(f, k)] = gc.get(f, k)

    g.step(hook)
    g.close()
    for key in want:
        assert bitwise_equal(want[key], seen[key]), key


def _env_sum(g, m, field_a, field_b):
    return sum(float(np.sum(g.get(field_a, k) * g.get(field_b, k))) for k in range(m))


@pytest.mark.parametrize("name", ["bisod3d", "bitriplepoint"])
def test_staged_hook_reaches_the_remap(name):
    from mahyco_amd.lib import GpuCase
    case = MAKES[name]()
    m = len(case.envs)
    eps = 1e-3
    out = {}
    for scale in (False, True):
        g = GpuCase(case)
        g.set_staged(True)
        g.run(5)  # a flow with mixed cells and moving interfaces

        def hook(gc):
            if not scale:
                return
            for k in range(m):
                gc.set(T.MH_F_INTERNAL_ENERGY, gc.get(T.MH_F_INTERNAL_ENERGY, k) * (1. + eps), k)

        g.step(hook)
        g.synchronize()
        out[scale] = {
            "E": _env_sum(g, m, T.MH_F_CELL_MASS, T.MH_F_INTERNAL_ENERGY),
            "e": [g.get(T.MH_F_INTERNAL_ENERGY, k) for k in range(m)],
            "f": [g.get(T.MH_F_FRACVOL, k) for k in range(m)],
            "keep": {(f, k): g.get(f, k) for f in (T.MH_F_DENSITY, T.MH_F_CELL_MASS, T.MH_F_FRACVOL)
                     for k in range(-1, m)},
            "u": [g.get(T.MH_N_VELOCITY, -1, d) for d in range(case.dim)],
        }
        g.close()
    a, b = out[False], out[True]
    assert abs(b["E"] / a["E"] - (1. + eps)) < 1e-12, (a["E"], b["E"])
    for k in range(m):  # per cell where the environment holds a share the fluxes cannot cancel
        sel = (a["f"][k] > 1e-3) & (a["e"][k] > 0.)
        assert sel.any()
        r = b["e"][k][sel] / a["e"][k][sel]
        assert np.max(np.abs(r - (1. + eps))) < 1e-11, k
    for key in a["keep"]:
        assert bitwise_equal(a["keep"][key], b["keep"][key]), key
    for d in range(case.dim):
        assert bitwise_equal(a["u"][d], b["u"][d]), d


def test_staged_refusals():
    from mahyco_amd.lib import GpuCase, MahycoError
    case = MAKES["bisod3d"]()
    g = GpuCase(case)
    lib, h = g.lib, g.h
    e = g.get(T.MH_F_INTERNAL_ENERGY)
    for fn in g.ENTRY_POINTS:  # not staged: no state to change before the remap
        assert getattr(lib, fn)(h) == T.MH_OK
    assert lib.mahyco_set_field(h, T.MH_F_INTERNAL_ENERGY, -1, 0, e.ctypes.data) == T.MH_ERR_INVALID
    assert lib.mahyco_set_staged(h, 1) == T.MH_ERR_INVALID  # inside an iteration
    assert lib.mahyco_remap(h) == T.MH_OK
    assert lib.mahyco_compute_delta_t(h) == T.MH_OK
    g.set_staged(True)
    for fn in g.ENTRY_POINTS:
        assert getattr(lib, fn)(h) == T.MH_OK
    u = g.get(T.MH_N_VELOCITY, -1, 0)
    assert lib.mahyco_set_field(h, T.MH_N_VELOCITY, -1, 0, u.ctypes.data) == T.MH_ERR_INVALID  # node values
    assert lib.mahyco_remap(h) == T.MH_OK
    assert lib.mahyco_compute_delta_t(h) == T.MH_OK
    g.close()
    for make, why in ((lambda: cases.config1_sod2d_lagrange(), "no remap"),
                      (lambda: _csts(), "schema-csts"),
                      (lambda: cases.rider("r_ty", 32, .5), "advection")):
        g = GpuCase(make())
        with pytest.raises(MahycoError) as ex:
            g.set_staged(True)
        assert ex.value.status == T.MH_ERR_UNSUPPORTED, why
        g.close()


def _csts():
    c = cases.bisod_x(12, 3, 3)
    c.schema_csts = True
    return c
