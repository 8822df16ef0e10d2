"""The GPU path through its different doors gives identical results (GPU tests):
per-entry-point calls vs fused cycles vs hipGraph replays, repeated runs, the C++ host
interface (MahycoModule over the C ABI), properties of the full-size 256x128x128
bi-material case (planar symmetry, positivity), and the full-size workloads against the
oracle's digests, bitwise over the whole mesh."""
import json
import os
import subprocess

import numpy as np
import pytest

from mahyco_amd import cases
from mahyco_amd import types as T
from tests.helpers import bitwise_equal, fields, summarize

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _entry_point_cycle(g):
    lib, h = g.lib, g.h
    for fn in ("mahyco_save_values_at_n", "mahyco_update_velocity", "mahyco_apply_boundary_condition",
               "mahyco_update_position", "mahyco_compute_geometric_values", "mahyco_update_density",
               "mahyco_compute_artificial_viscosity", "mahyco_update_elasticity_and_plasticity",
               "mahyco_update_energy_and_pressure"):
        assert getattr(lib, fn)(h) == T.MH_OK, fn
    if g.case.with_projection:
        assert lib.mahyco_remap(h) == T.MH_OK
    assert lib.mahyco_compute_delta_t(h) == T.MH_OK


@pytest.mark.parametrize("make", [lambda: cases.bisod_x(30, 4, 3), cases.lag2d_sod_x,
                                  lambda: cases.bitriplepoint()])
def test_entry_points_equal_fused_and_graph(make):
    from mahyco_amd.lib import GpuCase
    case = make()
    print(f"[paths] {case.name}", flush=True)
    a, b, c = GpuCase(case), GpuCase(case), GpuCase(case)
    n = 24
    for _ in range(n):
        _entry_point_cycle(a)
    a.synchronize()
    b.run(n)
    c.run(n, graph=6)
    fa, fb, fc = (fields(x, len(case.envs), case.dim) for x in (a, b, c))
    for k in fa:
        assert bitwise_equal(fa[k], fb[k]), ("entry points vs fused", k)
        assert bitwise_equal(fb[k], fc[k]), ("fused vs graph", k)
    assert a.time == b.time == c.time
    # out-of-order entry point is refused
    assert a.lib.mahyco_update_velocity(a.h) == T.MH_ERR_INVALID
    for x in (a, b, c):
        x.close()


def test_graph_replays_are_stable():
    """Several replays of the same captured graph, with eager work and synchronisations in
    between, match an eager run cycle for cycle."""
    from mahyco_amd.lib import GpuCase
    case = cases.bisod_x(24, 4, 4)
    g, e = GpuCase(case), GpuCase(case)
    for _ in range(4):
        g.run(6, graph=6)
        e.run(6)
        assert g.time == e.time
        assert bitwise_equal(g.get(T.MH_F_DENSITY), e.get(T.MH_F_DENSITY))
        g.run(1)
        e.run(1)
    g.close()
    e.close()


def test_repeatable():
    from mahyco_amd.lib import GpuCase
    case = cases.config2_bisod3d(48, 24, 24)
    r = []
    for _ in range(2):
        g = GpuCase(case)
        g.run(30, graph=6)
        r.append(fields(g, 2, 3))
        g.close()
    for k in r[0]:
        assert bitwise_equal(r[0][k], r[1][k]), k


def test_cpp_host_interface():
    """MahycoModule (C++) through the entry points and through fused cycles matches the
    C ABI, and reproduces the fixture's final time of CAS_BiSodCaseX."""
    exe = os.path.join(ROOT, "mahyco_amd", "_build", "mahyco_run")
    assert os.path.exists(exe), "C++ driver not built"
    out1 = subprocess.run([exe, "bisod_x"], check=True, capture_output=True, text=True, timeout=120).stdout
    out2 = subprocess.run([exe, "bisod_x", "--fused", "--graph", "6"], check=True, capture_output=True,
                          text=True, timeout=120).stdout
    assert out1 == out2, (out1, out2)
    kv = dict(x.split("=") for x in out1.split())
    assert int(kv["iteration"]) == 183
    assert abs(float(kv["time"]) - 0.198843560989769) <= 1e-14
    bad = subprocess.run([exe, "nosuchcase"], capture_output=True, text=True, timeout=60)
    assert bad.returncode == 2


@pytest.mark.parametrize("name,make", [
    ("monorider_rotation", lambda: cases.rider("mr_rot", 26, .5)),
    ("rider_deformation_reverse", lambda: cases.rider("r_def_rev", 38, 1., 1.)),
    ("bisod_x_csts", lambda: _csts(cases.bisod_x())),
])
def test_cpp_host_interface_variants(name, make):
    """The C++ MahycoModule with the Fictif EOS and the RIDER case service in advection mode,
    and with CSTS: entry points and fused cycles agree with each other and with the C ABI
    driven from Python (same iteration, time and field sums to the last bit)."""
    from mahyco_amd.lib import GpuCase
    exe = os.path.join(ROOT, "mahyco_amd", "_build", "mahyco_run")
    print(f"[cpp host] {name}", flush=True)
    out1 = subprocess.run([exe, name], check=True, capture_output=True, text=True, timeout=120).stdout
    out2 = subprocess.run([exe, name, "--fused", "--graph", "6"], check=True, capture_output=True,
                          text=True, timeout=120).stdout
    assert out1 == out2, (out1, out2)
    kv = dict(x.split("=") for x in out1.split())
    g = GpuCase(make())
    st = T.MH_OK
    while st == T.MH_OK:
        st = g.run(500, graph=6)
    assert st == T.MH_STOPPED
    assert int(kv["iteration"]) == g.iteration
    assert float(kv["time"]) == g.time
    assert float(kv["sum_rho"]) == float(np.cumsum(g.get(T.MH_F_DENSITY))[-1])
    assert float(kv["sum_e"]) == float(np.cumsum(g.get(T.MH_F_INTERNAL_ENERGY))[-1])
    g.close()


def _csts(c):
    c.schema_csts = True
    return c


def test_full_size_bisod_properties():
    """256x128x128 bi-material tube, 100 iterations: the solution stays planar (every y-z
    slice equal to the first to round-off) and positive."""
    from mahyco_amd.lib import GpuCase
    case = cases.config2_bisod3d()
    g = GpuCase(case)
    st = g.run(100, graph=6)
    assert st == T.MH_OK
    assert g.iteration == 100
    rho = g.get(T.MH_F_DENSITY).reshape(case.cell_shape)
    p = g.get(T.MH_F_PRESSURE).reshape(case.cell_shape)
    assert np.all(rho > 0) and np.all(p > 0)
    dev = np.max(np.abs(rho - rho[:1, :1, :]) / rho[:1, :1, :])
    assert dev < 1e-9, dev
    g.close()


def _served(i):
    from mahyco_amd import serve
    return serve.serve_requests(256, seed=20261202, rate=8.0)[i][1]


FULL_SIZE = {
    "bisod3d": (lambda: cases.config2_bisod3d(256, 128, 128), 100),
    "sedov3d": (lambda: cases.config3_sedov3d(192, 200), 200),
    "served114": (lambda: _served(114), 400),  # cells that empty out: kept mass fractions
    "served162": (lambda: _served(162), 400),
    "bisod3d_csts": (lambda: _csts(cases.config2_bisod3d(256, 128, 128)), 100),  # the CSTS scheme
    "rider512": (lambda: _rider512(), 300),  # advection, reversed deformation, ~2e4 mixed cells
}


def _rider512():
    c = cases.rider("rider_deformation_reverse_512", 38, 1., 1.)
    c.n = (512, 512, 1)
    return c


@pytest.mark.parametrize("name", sorted(FULL_SIZE))
def test_full_size_bitwise_vs_oracle(name):
    """Full-size single-mesh workloads and two served runs with emptying cells against the
    oracle's digests (SHA-256 of every parity field over the whole mesh, exact time and
    time step; tests/golden/make_oracle_full_size.py)."""
    from mahyco_amd.lib import GpuCase
    path = os.path.join(ROOT, "tests", "golden", f"oracle_{name}_full.json")
    if not os.path.exists(path):
        pytest.skip(f"no oracle digest {path}")
    with open(path) as f:
        ref = json.load(f)
    make, cycles = FULL_SIZE[name]
    case = make()
    print(f"[paths] full size {case.name} {cycles} iterations", flush=True)
    g = GpuCase(case)
    st = g.run(cycles, graph=6)
    assert st == ref["status"]
    assert g.iteration == ref["iteration"]
    assert float(g.time).hex() == ref["time"] and float(g.deltat).hex() == ref["deltat"]
    got = summarize(g, case)
    g.close()
    bad = {}
    for k, r in ref["fields"].items():
        if got[k]["sha256"] != r["sha256"]:
            row = [i for i, (a, b) in enumerate(zip(got[k]["row"], r["row"])) if a != b]
            bad[k] = {"sum": (got[k]["sum"], r["sum"]), "row_mismatches": row[:8]}
    assert not bad, bad


@pytest.mark.parametrize("shape", [(37, 19, 11), (16, 16, 8), (5, 33, 17)])
def test_node_kernel_tiled_equals_flat(monkeypatch, shape):
    """The z-marching tiled node kernel and the flat one give identical bits (odd sizes
    exercise partial tiles and z chunks)."""
    from mahyco_amd.lib import GpuCase
    res = {}
    for mode in ("flat", "tiled"):
        monkeypatch.setenv("MAHYCO_NODE_KERNEL", mode)
        c = cases.config2_bisod3d(*shape)
        g = GpuCase(c)
        g.run(20)
        res[mode] = fields(g, 2, 3)
        g.close()
    for k in res["flat"]:
        assert bitwise_equal(res["flat"][k], res["tiled"][k]), k


@pytest.mark.parametrize("shape", [(70, 9, 19), (64, 4, 8), (5, 33, 17)])
def test_cell_kernel_tiled_equals_flat(monkeypatch, shape):
    """The LDS-staged z-marching cell kernel (an autotuner candidate) and the one-thread-per-
    cell kernel give identical bits (sizes leave partial tiles in x, y and z)."""
    from mahyco_amd.lib import GpuCase
    res = {}
    for mode in ("flat", "tiled"):
        monkeypatch.setenv("MAHYCO_CELL_KERNEL", mode)
        c = cases.config2_bisod3d(*shape)
        g = GpuCase(c)
        g.run(20)
        res[mode] = fields(g, 2, 3)
        g.close()
    for k in res["flat"]:
        assert bitwise_equal(res["flat"][k], res["tiled"][k]), k


@pytest.mark.parametrize("make", [lambda: cases.config2_bisod3d(13, 37, 21), lambda: cases.config2_bisod3d(9, 2, 40),
                                  lambda: cases.config2_bisod3d(7, 17, 3), lambda: cases.config2_bisod3d(130, 5, 4),
                                  lambda: cases.config2_bisod3d(70, 35, 18), lambda: cases.bitriplepoint(),
                                  lambda: cases.sedov_2d(),
                                  lambda: cases.ordre3(lambda: cases.config2_bisod3d(70, 35, 18), "b3d_o3"),
                                  lambda: cases.config2_bisod3d(256, 4, 128), lambda: cases.config2_bisod3d(256, 128, 3)])
def test_sweep_rows_and_tiles_equal_flat(monkeypatch, make):
    """The row-segment x sweeps (cell and dual parts), the LDS-tiled y/z cell sweeps and the
    marching y/z dual sweeps give the same bits as the one-thread-per-item sweeps (sizes cover
    partial tiles and row segments, several segments per row and tiles per line, sweeps 2 or 3
    cells long, and marching rows of 257 x 129 nodes, more than 128 blocks each)."""
    from mahyco_amd.lib import GpuCase
    res = {}
    for mode in ("flat", "tile"):
        monkeypatch.setenv("MAHYCO_SWEEP", mode)
        c = make()
        print(f"[paths] sweep {mode} {c.name}", flush=True)
        g = GpuCase(c)
        g.run(18)
        res[mode] = fields(g, len(c.envs), c.dim)
        g.close()
    ref = res["flat"]
    for key, r in res.items():
        for k in ref:
            assert bitwise_equal(ref[k], r[k]), (key, k)


@pytest.mark.parametrize("length", [1, 3, 7, 10, 200])
def test_march_segment_lengths_equal_flat(monkeypatch, length):
    """The marching y/z dual sweeps split each node line into segments of at most the target
    length, balanced to within one node; every target (one node, lengths that leave uneven
    remainders, and one longer than the line) gives the bits of the one-thread-per-node sweep."""
    from mahyco_amd.lib import GpuCase
    res = {}
    for mode, ml in (("flat", None), ("tile", str(length))):
        monkeypatch.setenv("MAHYCO_SWEEP", mode)
        if ml:
            monkeypatch.setenv("MAHYCO_MARCH_LEN", ml)
        c = cases.config2_bisod3d(33, 29, 22)
        print(f"[paths] march length {length} {mode}", flush=True)
        g = GpuCase(c)
        g.run(18)
        res[mode] = fields(g, len(c.envs), c.dim)
        g.close()
    for k in res["flat"]:
        assert bitwise_equal(res["flat"][k], res["tile"][k]), k


def test_set_variant_same_bits_and_reported():
    """mahyco_set_variant applies a kernel-variant code (as the autotuner reports it) and
    mahyco_get_variant returns it; every variant gives the same bits."""
    from mahyco_amd.lib import GpuCase, MahycoError
    res = {}
    for v in (0, 2, 15, 16 | 5, 2 | 32 * 3, 1 | 32 * 4):
        g = GpuCase(cases.config2_bisod3d(33, 17, 19))
        g.set_variant(v)
        assert g.variant == v
        g.run(18, graph=6)
        res[v] = fields(g, 2, 3)
        g.close()
    for v in res:
        for k in res[0]:
            assert bitwise_equal(res[0][k], res[v][k]), (v, k)
    g = GpuCase(cases.config2_bisod3d(8, 8, 8))
    with pytest.raises(MahycoError):
        g.set_variant(32 * 5)
    g.close()
