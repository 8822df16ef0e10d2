"""Ensembles (GPU tests): every case gives the same bits alone (its own handle, eager
single-mesh launches) and inside a batch of uneven meshes, with or without hipGraphs, while
other cases join and leave; admission respects the HBM budget."""
import numpy as np
import pytest

from mahyco_amd import cases
from mahyco_amd import types as T

pytestmark = pytest.mark.gpu


def _mix(seed, n=10, stiff=False, dim3=False, o3=False, pb=False):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        if dim3 and i % 2:
            side = int(rng.integers(8, 16))
            c = cases.riemann3d(side, rng, stiffened=True, t_end=1e30, name=f"m{i}")
        else:
            nx = int(np.exp(rng.uniform(np.log(16), np.log(48))))
            c = cases.riemann2d(nx, nx, rng, stiffened=stiff, name=f"m{i}")
        if o3 and i % 3 == 0:
            c.ordre_projection = 3  # order-3 remap: a batch group of its own
        if pb and i % 3 == 1:  # pente-borne and generalised limiters: another group
            c.pente_borne, c.debar_fix, c.pente_borne_mixte = True, 1 + i % 2, True
            c.limiter = T.MH_SUPERBEE_G if i % 2 else T.MH_MINMOD
        out.append(c)
    return out


def _alone(case, ncycles):
    from mahyco_amd.lib import GpuCase
    case.max_cycles = ncycles
    g = GpuCase(case)
    g.run(ncycles)
    res = {f: g.get(f) for f in (T.MH_F_DENSITY, T.MH_F_PRESSURE, T.MH_F_INTERNAL_ENERGY, T.MH_F_FRACVOL)}
    res["u"] = g.get(T.MH_N_VELOCITY, -1, 0)
    res["t"] = g.time
    g.close()
    return res


@pytest.mark.parametrize("graph,stiff,dim3,o3", [(True, False, False, False), (False, True, True, False),
                                                (True, True, True, True)])
def test_alone_equals_mixed(graph, stiff, dim3, o3):
    from mahyco_amd.lib import Ensemble
    print(f"[ensemble] graph={graph} stiff={stiff} dim3={dim3} o3={o3}", flush=True)
    mix = _mix((3 if graph else 4) + (2 if o3 else 0), stiff=stiff, dim3=dim3, o3=o3, pb=o3)
    ncyc = 30
    E = Ensemble(16, 4e9)
    slots = [E.add(c, max_iterations=ncyc) for c in mix]
    assert all(s is not None for s in slots)
    E.step(ncyc + 6, graph=graph)  # extra iterations are no-ops once max_iterations is reached
    n, st, tm, it = E.poll()
    for s, c in zip(slots, mix):
        assert st[s] == T.MH_STOPPED and it[s] == ncyc
        ref = _alone(c, ncyc)
        assert tm[s] == ref["t"]
        for f in (T.MH_F_DENSITY, T.MH_F_PRESSURE, T.MH_F_INTERNAL_ENERGY, T.MH_F_FRACVOL):
            assert np.array_equal(E.get(s, f), ref[f]), (c.name, f)
        assert np.array_equal(E.get(s, T.MH_N_VELOCITY, -1, 0), ref["u"])
    E.close()


def _two_d(c, n):
    c.dim, c.n, c.length = 2, (n, n, 1), (1., 1., 0.)
    return c


def _variant_mix():
    """CSTS runs (2D and 3D) and 2D advection runs (RIDER bubble, ADVECTION square, with and
    without time reversal) beside ordinary ones: four batch groups."""
    rng = np.random.default_rng(7)
    out = []
    for i in range(3):
        c = cases.riemann2d(int(rng.integers(16, 40)), int(rng.integers(16, 40)), rng, stiffened=True, name=f"c{i}")
        c.schema_csts = i < 2
        out.append(c)
    c = cases.riemann3d(10, rng, stiffened=True, t_end=1e30, name="c3d")
    c.schema_csts = True
    out.append(c)
    out.append(_two_d(cases.rider("r2", 34, .2), 40))
    out.append(_two_d(cases.rider("r2rev", 38, .2, 1.), 32))
    out.append(_two_d(cases.advection("a2", 21, .2, n=36), 36))
    return out


@pytest.mark.parametrize("graph", [False, True])
def test_variant_groups_alone_equal_mixed(graph):
    """CSTS and advection-mode cases form their own batch groups (forward velocity step,
    CSTS cell kernel; prescribed node velocity, no dual sweeps) and give the same bits as
    alone."""
    from mahyco_amd.lib import Ensemble
    print(f"[ensemble] variants graph={graph}", flush=True)
    mix = _variant_mix()
    ncyc = 30
    E = Ensemble(16, 4e9)
    slots = [E.add(c, max_iterations=ncyc) for c in mix]
    assert all(s is not None for s in slots)
    E.step(ncyc + 6, graph=graph)
    n, st, tm, it = E.poll()
    for s, c in zip(slots, mix):
        assert st[s] == T.MH_STOPPED and it[s] == ncyc, (c.name, st[s], it[s])
        ref = _alone(c, ncyc)
        assert tm[s] == ref["t"], c.name
        for f in (T.MH_F_DENSITY, T.MH_F_PRESSURE, T.MH_F_INTERNAL_ENERGY, T.MH_F_FRACVOL):
            assert np.array_equal(E.get(s, f), ref[f]), (c.name, f)
        assert np.array_equal(E.get(s, T.MH_N_VELOCITY, -1, 0), ref["u"]), c.name
    E.close()


def test_join_and_leave():
    """A case admitted while others run, after one left, still equals its solo run."""
    from mahyco_amd.lib import Ensemble
    mix = _mix(5, n=6)
    E = Ensemble(8, 4e9)
    s0 = [E.add(c, max_iterations=40) for c in mix[:4]]
    E.step(12)
    E.remove(s0[1])
    late = E.add(mix[4], max_iterations=18)
    E.step(30)
    n, st, tm, it = E.poll()
    assert st[late] == T.MH_STOPPED and it[late] == 18
    ref = _alone(mix[4], 18)
    assert np.array_equal(E.get(late, T.MH_F_DENSITY), ref[T.MH_F_DENSITY])
    E.close()


def test_admission_by_footprint():
    from mahyco_amd.lib import Ensemble, footprint
    c = cases.riemann2d(64, 64, np.random.default_rng(0))
    fp = footprint(c)
    E = Ensemble(8, 2.5 * fp)
    assert E.add(c) is not None
    assert E.add(c) is not None
    assert E.add(c) is None  # a third would exceed the budget
    assert E.used_bytes <= 2.5 * fp
    E.close()
