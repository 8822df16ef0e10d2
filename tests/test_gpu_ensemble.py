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
    """CSTS runs (2D and 3D), 2D advection runs (RIDER bubble, ADVECTION square, with and
    without time reversal) and runs with total-energy conservation (2D, and 3D with CSTS)
    beside ordinary ones: six batch groups."""
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
    c = cases.riemann2d(28, 20, rng, stiffened=True, name="cons2d")
    c.conservation_energie_totale = True
    out.append(c)
    c = cases.riemann3d(9, np.random.default_rng(3), stiffened=True, t_end=1e30, name="cons3d_csts")
    c.conservation_energie_totale = c.schema_csts = True
    out.append(c)
    return out


@pytest.mark.parametrize("graph", [False, True])
def test_variant_groups_alone_equal_mixed(graph):
    """CSTS, advection-mode and total-energy-conservation cases form their own batch groups
    (forward velocity step, CSTS cell kernel; prescribed node velocity, no dual sweeps; dual
    kinetic energy and its return) and give the same bits as alone."""
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


class _SlotView:
    """One ensemble slot seen through get(field, env, comp), as GpuCase / OracleCase."""

    def __init__(self, E, slot):
        self.E, self.slot = E, slot

    def get(self, field, env=-1, comp=0):
        return self.E.get(self.slot, field, env, comp)


def test_batched_workload_full_size():
    """The batched workload at its size: 256 two-dimensional bi-material tubes (sides 64-512
    drawn log-uniform, seed 20261202, 17.4 M cells) in one ensemble, 300 iterations each with
    its own time step.  Every run reaches its horizon with finite positive density and
    pressure; eight members (the smallest and the largest among them) are bitwise their solo
    runs; the smallest and the largest match the oracle's digests over the whole mesh."""
    import json
    import os
    from mahyco_amd import serve
    from mahyco_amd.lib import Ensemble
    from tests.helpers import summarize
    batch = serve.batch_cases(256, seed=20261202)
    sides = [c.n[0] for c in batch]
    i_min, i_max = int(np.argmin(sides)), int(np.argmax(sides))
    print(f"[ensemble] batched 256 runs, sides {min(sides)}..{max(sides)}", flush=True)
    E = Ensemble(256, 200e9)
    slots = [E.add(c, max_iterations=300) for c in batch]
    assert all(s is not None for s in slots)
    E.step(306, graph=True)
    n, st, tm, it = E.poll()
    assert all(st[s] == T.MH_STOPPED and it[s] == 300 for s in slots)
    for s, c in zip(slots, batch):
        for f in (T.MH_F_DENSITY, T.MH_F_PRESSURE):
            v = E.get(s, f)
            assert np.all(np.isfinite(v)) and np.all(v > 0), (c.name, f)
    rng = np.random.default_rng(1)
    picks = sorted({i_min, i_max, *rng.choice(len(batch), 6, replace=False).tolist()})
    for i in picks:
        c = batch[i]
        print(f"[ensemble] solo {c.name}", flush=True)
        ref = _alone(c, 300)
        assert tm[slots[i]] == ref["t"], c.name
        for f in (T.MH_F_DENSITY, T.MH_F_PRESSURE, T.MH_F_INTERNAL_ENERGY, T.MH_F_FRACVOL):
            assert np.array_equal(E.get(slots[i], f), ref[f]), (c.name, f)
        assert np.array_equal(E.get(slots[i], T.MH_N_VELOCITY, -1, 0), ref["u"]), c.name
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    for i in (i_min, i_max):
        with open(os.path.join(root, "tests", "golden", f"oracle_batch{i}_full.json")) as fh:
            ref = json.load(fh)
        assert float(tm[slots[i]]).hex() == ref["time"] and int(it[slots[i]]) == ref["iteration"]
        got = summarize(_SlotView(E, slots[i]), batch[i])
        bad = [k for k, r in ref["fields"].items() if got[k]["sha256"] != r["sha256"]]
        assert not bad, (batch[i].name, bad)
    E.close()


def _small_stream(n=24, seed=99, rate=400.0):
    """A short Poisson stream of small stiffened-gas runs (2D 24-64 cells a side, 3D 8-16)
    arriving faster than they finish, so that admissions overlap running work."""
    from mahyco_amd import serve
    reqs = serve.serve_requests(n, seed=seed, rate=rate, t_end=0.02, lo2=24, hi2=64, lo3=8, hi3=16)
    return reqs


def test_served_stream_alone_equals_mixed():
    """The continuous-batching serve loop: requests admitted while others run, every result
    bitwise the same run solved alone (time, iteration and fields)."""
    from mahyco_amd import serve
    from mahyco_amd.lib import GpuCase
    reqs = _small_stream()
    got = {}

    def keep(i, E, slot):
        got[i] = {f: E.get(slot, f) for f in (T.MH_F_DENSITY, T.MH_F_PRESSURE, T.MH_F_INTERNAL_ENERGY,
                                              T.MH_F_FRACVOL)}
        got[i]["u"] = E.get(slot, T.MH_N_VELOCITY, -1, 0)

    print(f"[serve] {len(reqs)} requests", flush=True)
    recs = serve.serve(reqs, budget=8e9, max_slots=8, on_done=keep)
    assert all(r.status == T.MH_STOPPED for r in recs), [r.status for r in recs]
    assert sum(r.live_at_admission > 0 for r in recs) >= len(recs) // 4  # admissions overlapped running work
    for i, (_, c) in enumerate(reqs):
        g = GpuCase(c)
        while g.run(60, graph=6) == T.MH_OK:
            pass
        assert g.iteration == recs[i].iterations, c.name
        for f in (T.MH_F_DENSITY, T.MH_F_PRESSURE, T.MH_F_INTERNAL_ENERGY, T.MH_F_FRACVOL):
            assert np.array_equal(g.get(f), got[i][f], equal_nan=True), (c.name, f)
        assert np.array_equal(g.get(T.MH_N_VELOCITY, -1, 0), got[i]["u"]), c.name
        g.close()


def test_admission_keeps_the_captured_graph():
    """A case joining a batch group that already runs changes only the device-side work lists:
    no new graph is captured, and the late case still equals its solo run."""
    from mahyco_amd.lib import Ensemble
    mix = _mix(9, n=6)
    E = Ensemble(8, 4e9)
    for c in mix[:3]:
        assert E.add(c, max_iterations=48) is not None
    E.step(12)
    captures = E.graph_captures
    assert captures >= 1
    late = E.add(mix[3], max_iterations=24)
    E.step(12)
    E.remove(0)
    E.step(18)
    assert E.graph_captures == captures
    n, st, tm, it = E.poll()
    assert st[late] == T.MH_STOPPED and it[late] == 24
    ref = _alone(mix[3], 24)
    assert tm[late] == ref["t"]
    for f in (T.MH_F_DENSITY, T.MH_F_PRESSURE, T.MH_F_INTERNAL_ENERGY, T.MH_F_FRACVOL):
        assert np.array_equal(E.get(late, f), ref[f]), f
    E.close()


def test_ensemble_autotune_same_bits():
    """The ensemble autotuner times its candidate grids on the cases' own iterations; the runs
    end with the bits of their solo runs."""
    from mahyco_amd.lib import Ensemble
    mix = _mix(13, n=5, dim3=True, stiff=True)
    ncyc = 96
    E = Ensemble(8, 4e9)
    slots = [E.add(c, max_iterations=ncyc) for c in mix]
    bpc, ms = E.autotune(6)
    assert bpc in (2, 4, 8, 16) and ms > 0
    E.step(ncyc)
    n, st, tm, it = E.poll()
    for s, c in zip(slots, mix):
        assert st[s] == T.MH_STOPPED and it[s] == ncyc, c.name
        ref = _alone(c, ncyc)
        assert tm[s] == ref["t"], c.name
        for f in (T.MH_F_DENSITY, T.MH_F_PRESSURE, T.MH_F_INTERNAL_ENERGY):
            assert np.array_equal(E.get(s, f), ref[f]), (c.name, f)
    E.close()


def test_case_autotune_same_bits():
    """The single-case autotuner (kernel variants and tile depths, timed on the case's own
    iterations) leaves the results bit for bit those of the default kernels."""
    from mahyco_amd.lib import GpuCase
    from tests.helpers import bitwise_equal, fields
    case = cases.config2_bisod3d(40, 20, 18, cycles=100000)
    a, b = GpuCase(case), GpuCase(case)
    v, ms = b.autotune(6)
    n = b.iteration + 12
    assert a.run(n, graph=6) == T.MH_OK
    assert b.run(12, graph=6) == T.MH_OK
    assert a.iteration == b.iteration == n
    fa, fb = fields(a, 2, 3), fields(b, 2, 3)
    for k in fa:
        assert bitwise_equal(fa[k], fb[k]), (v, k)
    a.close()
    b.close()
