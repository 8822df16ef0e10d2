"""State handed over through the boundary, the standalone EOS calls, and the status codes
that replace the reference's exit(1) sites.

* A state read from one run (get_field + scalars) and written into a fresh case (set_field +
  set_scalars) continues bit for bit as the run it came from: first oracle -> oracle (CPU),
  which shows that the listed fields are the whole state of the loop, then oracle -> GPU.
* IEquationOfState::applyEOS and MahycoModule::computePressionMoyenne as standalone calls on a
  mixed perfect-gas / stiffened-gas / Fictif state, against the oracle.
* MH_ERR_DT_MIN (MahycoModule.cc:1816-1833), MH_ERR_NAN_REMAP (RemapADIFinal.cc:274-284) and
  MH_ERR_NEG_VOLUME (MahycoModule.cc:1164-1191) at the same iteration as the oracle, forced
  from the host through the state (no kernel is made to fault); later iterations are no-ops.
"""
import ctypes as C

import numpy as np
import pytest

from mahyco_amd import cases
from mahyco_amd import types as T
from oracle.pyoracle import OracleCase
from tests.helpers import bitwise_equal, fields

STATE_CELL = [T.MH_F_DENSITY, T.MH_F_PRESSURE, T.MH_F_INTERNAL_ENERGY, T.MH_F_SOUND_SPEED,
              T.MH_F_PSEUDO_VISCOSITY, T.MH_F_CELL_MASS, T.MH_F_FRACVOL, T.MH_F_MASS_FRACTION]
STATE_ENV = [T.MH_F_DENSITY, T.MH_F_PRESSURE, T.MH_F_INTERNAL_ENERGY, T.MH_F_CELL_MASS,
             T.MH_F_FRACVOL, T.MH_F_MASS_FRACTION]


def transfer(src, dst, case):
    """Writes src's state (the fields the loop carries from one iteration to the next) into dst."""
    dst.set_scalars(*src.scalars())
    dst.set(T.MH_F_ENV_MASK, src.get(T.MH_F_ENV_MASK))
    for f in STATE_CELL:
        dst.set(f, src.get(f))
    for k in range(len(case.envs)):
        for f in STATE_ENV:
            dst.set(f, src.get(f, k), env=k)
    for c in range(case.dim):
        dst.set(T.MH_N_VELOCITY, src.get(T.MH_N_VELOCITY, -1, c), comp=c)
        dst.set(T.MH_N_COORD_LAGRANGE, src.get(T.MH_N_COORD_LAGRANGE, -1, c), comp=c)


TRANSFER_CASES = [lambda: cases.bisod_x(30, 4, 3, limiter=T.MH_SUPERBEE), cases.lag2d_sod_x,
                  cases.bitriplepoint]


def _same(a, b, case):
    fa, fb = fields(a, len(case.envs), case.dim), fields(b, len(case.envs), case.dim)
    for k in fa:
        assert bitwise_equal(fa[k], fb[k]), k
    assert a.time == b.time and a.deltat == b.deltat and a.iteration == b.iteration


@pytest.mark.parametrize("make", TRANSFER_CASES)
def test_oracle_state_transfer_resumes_bitwise(make):
    """Oracle -> oracle: the fields listed above are all the loop carries (13 + 11 iterations
    of a fresh case fed the state equal 24 uninterrupted ones)."""
    case = make()
    a = OracleCase(case, pow_mode=1, threads=2)
    assert a.run(13) == T.MH_OK
    b = OracleCase(case, pow_mode=1, threads=2)
    transfer(a, b, case)
    assert a.run(11) == T.MH_OK and b.run(11) == T.MH_OK
    _same(a, b, case)


def test_oracle_status_dt_min():
    """A step below deltat-min: at iteration 1 (deltat-init below it), and after a state whose
    time step is set below deltat-min / 1.05 (the 1.05 growth cap then keeps it there)."""
    case = cases.bisod_x(20, 3, 3, deltat_init=4e-9)
    o = OracleCase(case, pow_mode=1)
    assert o.run(5) == T.MH_ERR_DT_MIN and o.iteration == 1
    case = cases.bisod_x(20, 3, 3)
    o = OracleCase(case, pow_mode=1)
    assert o.run(7) == T.MH_OK
    t, dt, dto, it, sens = o.scalars()
    o.set_scalars(t, 0.9e-8, dt, it, sens)
    assert o.run(3) == T.MH_ERR_DT_MIN and o.iteration == 8


def _poke_energy(sim, case, cell):
    e = sim.get(T.MH_F_INTERNAL_ENERGY)
    e[cell] = np.nan
    sim.set(T.MH_F_INTERNAL_ENERGY, e)
    mask = sim.get(T.MH_F_ENV_MASK)
    for k in range(len(case.envs)):
        if int(mask[cell]) >> k & 1:
            ek = sim.get(T.MH_F_INTERNAL_ENERGY, k)
            ek[cell] = np.nan
            sim.set(T.MH_F_INTERNAL_ENERGY, ek, env=k)


def _poke_velocity(sim, case, node, amount):
    u = sim.get(T.MH_N_VELOCITY, -1, 0)
    u[node] += amount
    sim.set(T.MH_N_VELOCITY, u, comp=0)


def test_oracle_status_nan_and_negative_volume():
    case = cases.bisod_x(20, 3, 3)
    o = OracleCase(case, pow_mode=1)
    assert o.run(6) == T.MH_OK
    _poke_energy(o, case, 7 + 20 * (1 + 3 * 1))
    assert o.run(4) == T.MH_ERR_NAN_REMAP and o.iteration == 7
    case = cases.lag2d_sod_x()
    o = OracleCase(case, pow_mode=1)
    assert o.run(6) == T.MH_OK
    nx = case.n[0]
    node = 50 + (nx + 1) * 2
    _poke_velocity(o, case, node, 5.0 * (1.0 / nx) / o.deltat)
    assert o.run(4) == T.MH_ERR_NEG_VOLUME and o.iteration == 7


def _mixed_eos_case():
    """Three environments: perfect gas, stiffened gas (pi = 0.3) and Fictif (pressure
    p0 + dpdt t after tdebut)."""
    c = cases.bitriplepoint()
    c.envs[1].eos, c.envs[1].limit_tension, c.envs[1].adiabatic_cst = T.MH_EOS_STIFFENED_GAS, 0.3, 2.8
    c.envs[2].eos, c.envs[2].fictif_tdebut, c.envs[2].fictif_tfin = T.MH_EOS_FICTIF, 0.0, 10.0
    c.envs[2].fictif_p0, c.envs[2].fictif_dpdt = 0.25, 3.0
    return c


def _random_mixed_state(sim, case, rng):
    """Every membership pattern of three environments, positive densities and energies,
    volume fractions that sum to one over the members."""
    nc, m = case.nb_cells, len(case.envs)
    mask = rng.integers(1, 1 << m, nc).astype(np.float64)
    sim.set_scalars(0.0125, 1e-4, 1e-4, 9, 0)
    sim.set(T.MH_F_ENV_MASK, mask)
    f = rng.uniform(0.05, 1.0, (m, nc)) * np.array([(mask.astype(int) >> k) & 1 for k in range(m)])
    f /= f.sum(axis=0)
    for k in range(m):
        # rho e > pi: the stiffened gas keeps p + pi > 0 and a real sound speed
        sim.set(T.MH_F_DENSITY, rng.uniform(0.7, 3.0, nc), env=k)
        sim.set(T.MH_F_INTERNAL_ENERGY, rng.uniform(0.5, 4.0, nc), env=k)
        sim.set(T.MH_F_PRESSURE, rng.uniform(0.1, 2.0, nc), env=k)
        sim.set(T.MH_F_FRACVOL, f[k], env=k)


def _eos_fields(sim, case):
    out = {"p": sim.get(T.MH_F_PRESSURE), "c": sim.get(T.MH_F_SOUND_SPEED), "e": sim.get(T.MH_F_INTERNAL_ENERGY)}
    for k in range(len(case.envs)):
        out[f"p{k}"] = sim.get(T.MH_F_PRESSURE, k)
        out[f"e{k}"] = sim.get(T.MH_F_INTERNAL_ENERGY, k)
    return out


def test_oracle_standalone_eos_closed_forms():
    """The oracle's standalone applyEOS / computePressionMoyenne: PerfectGasEOSService.cc:45-67,
    StiffenedGasEOSService.cc:43-63, FictifEOSService.cc:45-75, MahycoModule.cc:1708-1735,
    checked against the formulas on a random mixed state."""
    case = _mixed_eos_case()
    o = OracleCase(case, pow_mode=1)
    rng = np.random.default_rng(5)
    _random_mixed_state(o, case, rng)
    rho = [o.get(T.MH_F_DENSITY, k) for k in range(3)]
    e = [o.get(T.MH_F_INTERNAL_ENERGY, k) for k in range(3)]
    f = [o.get(T.MH_F_FRACVOL, k) for k in range(3)]
    mask = o.get(T.MH_F_ENV_MASK).astype(int)
    for k in range(3):
        o.apply_eos(k)
    o.compute_pression_moyenne()
    g0, g1, pi1 = case.envs[0].adiabatic_cst, case.envs[1].adiabatic_cst, case.envs[1].limit_tension
    p_k = [(g0 - 1) * rho[0] * e[0], (g1 - 1) * rho[1] * e[1] - g1 * pi1, np.full_like(e[2], 0.25 + 3.0 * 0.0125)]
    with np.errstate(divide="ignore", invalid="ignore"):  # non-member cells read as 0
        c_k = [np.sqrt(g0 * p_k[0] / rho[0]), np.sqrt(g1 / rho[1] * (p_k[1] + pi1)), np.ones_like(e[2])]
    p, c = o.get(T.MH_F_PRESSURE), o.get(T.MH_F_SOUND_SPEED)
    for k in range(3):
        mem = (mask >> k) & 1 == 1
        assert np.allclose(o.get(T.MH_F_PRESSURE, k)[mem], p_k[k][mem], rtol=1e-14, atol=0)
        if k == 2:
            assert np.all(o.get(T.MH_F_INTERNAL_ENERGY, k)[mem] == 0.)
    mixed = np.array([bin(x).count("1") > 1 for x in mask])
    members = [((mask >> k) & 1) == 1 for k in range(3)]
    p_mix = sum(np.where(members[k], f[k] * p_k[k], 0.) for k in range(3))
    c_max = np.max([np.where(members[k], c_k[k], 0.) for k in range(3)], axis=0)
    assert np.allclose(p[mixed], p_mix[mixed], rtol=1e-13, atol=1e-15)
    assert np.allclose(c[mixed], c_max[mixed], rtol=1e-14, atol=0)
    pure = ~mixed
    kp = np.log2(mask).astype(int)  # the one environment of a pure cell
    for k in range(3):
        sel = pure & (kp == k)
        assert np.allclose(p[sel], p_k[k][sel], rtol=1e-14, atol=0)
        assert np.allclose(c[sel], c_k[k][sel], rtol=1e-14, atol=0)


# ------------------------------------------------------------------------------------- GPU
@pytest.mark.gpu
@pytest.mark.parametrize("make", TRANSFER_CASES)
def test_gpu_set_field_resumes_oracle_state(make):
    """Oracle state after 13 iterations uploaded with set_field / set_scalars into a new GPU
    case, 11 more iterations on the GPU: bitwise the oracle's 24-iteration state."""
    from mahyco_amd.lib import GpuCase
    case = make()
    print(f"[state] {case.name}", flush=True)
    o = OracleCase(case, pow_mode=1, threads=4)
    assert o.run(13) == T.MH_OK
    g = GpuCase(case)
    transfer(o, g, case)
    assert o.run(11) == T.MH_OK
    assert g.run(11) == T.MH_OK
    _same(g, o, case)
    g.close()


@pytest.mark.gpu
def test_gpu_standalone_eos_and_mean_pressure_vs_oracle():
    """mahyco_apply_eos / mahyco_compute_pression_moyenne on a random mixed PG / SG / Fictif
    state, bitwise against the oracle's applyEOS / computePressionMoyenne."""
    from mahyco_amd.lib import GpuCase
    case = _mixed_eos_case()
    o, g = OracleCase(case, pow_mode=1), GpuCase(case)
    for sim in (o, g):
        _random_mixed_state(sim, case, np.random.default_rng(11))
    for k in range(3):
        o.apply_eos(k)
        assert g.lib.mahyco_apply_eos(g.h, k) == T.MH_OK
    a, b = _eos_fields(g, case), _eos_fields(o, case)
    for key in ("p0", "p1", "p2", "e2"):
        assert bitwise_equal(a[key], b[key]), key
    o.compute_pression_moyenne()
    assert g.lib.mahyco_compute_pression_moyenne(g.h) == T.MH_OK
    g.synchronize()
    a, b = _eos_fields(g, case), _eos_fields(o, case)
    for key in a:
        assert bitwise_equal(a[key], b[key]), key
    g.close()


@pytest.mark.gpu
def test_gpu_apply_one_cell_eos_vs_oracle():
    """mahyco_apply_one_cell_eos (IEquationOfState::applyOneCellEOS) on chosen cells of each
    environment of a random mixed PG / SG / Fictif state: those cells take the oracle's applyEOS
    values, every other cell keeps its state."""
    from mahyco_amd.lib import GpuCase
    case = _mixed_eos_case()
    o, g = OracleCase(case, pow_mode=1), GpuCase(case)
    for sim in (o, g):
        _random_mixed_state(sim, case, np.random.default_rng(17))
    before = _eos_fields(g, case)
    mask = g.get(T.MH_F_ENV_MASK).astype(int)
    for k in range(3):
        o.apply_eos(k)
    after_o = _eos_fields(o, case)
    picks = {}
    for k in range(3):
        members = np.flatnonzero((mask >> k) & 1)
        picks[k] = members[:: max(1, len(members) // 7)][:7]
        for c in picks[k]:
            assert g.lib.mahyco_apply_one_cell_eos(g.h, k, int(c)) == T.MH_OK
    g.synchronize()
    after_g = _eos_fields(g, case)
    for k in range(3):
        for key in (f"p{k}",) + ((f"e{k}",) if k == 2 else ()):
            sel = np.zeros(case.nb_cells, bool)
            sel[picks[k]] = True
            assert bitwise_equal(after_g[key][sel], after_o[key][sel]), key
            # cells outside the picks keep their values (the picks of other environments aside)
            others = ~sel
            for kk in range(3):
                if kk != k:
                    others[picks[kk]] = False
            assert bitwise_equal(after_g[key][others], before[key][others]), key
    assert g.lib.mahyco_apply_one_cell_eos(g.h, 0, case.nb_cells) == T.MH_ERR_INVALID
    g.close()


@pytest.mark.gpu
def test_gpu_status_codes_match_oracle():
    """Each exit site of the reference as a device status word, at the oracle's iteration;
    further iterations leave the state, the time and the iteration as they are."""
    from mahyco_amd.lib import GpuCase

    def frozen(g, case):
        before = fields(g, len(case.envs), case.dim)
        t, it = g.time, g.iteration
        g.run(3)
        after = fields(g, len(case.envs), case.dim)
        assert g.time == t and g.iteration == it
        for k in before:
            assert bitwise_equal(before[k], after[k]), k

    print("[status] dt_min at iteration 1", flush=True)
    case = cases.bisod_x(20, 3, 3, deltat_init=4e-9)
    g = GpuCase(case)
    assert g.run(5) == T.MH_ERR_DT_MIN and g.iteration == 1
    frozen(g, case)
    g.close()

    print("[status] dt_min after set_scalars", flush=True)
    case = cases.bisod_x(20, 3, 3)
    o, g = OracleCase(case, pow_mode=1), GpuCase(case)
    assert o.run(7) == T.MH_OK and g.run(7) == T.MH_OK
    t, dt, dto, it, sens = o.scalars()
    o.set_scalars(t, 0.9e-8, dt, it, sens)
    g.set_scalars(t, 0.9e-8, dt, it, sens)
    assert o.run(3) == T.MH_ERR_DT_MIN and g.run(3) == T.MH_ERR_DT_MIN
    assert g.iteration == o.iteration == 8 and g.time == o.time
    frozen(g, case)
    g.close()

    print("[status] NaN energy -> remap", flush=True)
    o, g = OracleCase(case, pow_mode=1), GpuCase(case)
    assert o.run(6) == T.MH_OK and g.run(6) == T.MH_OK
    for sim in (o, g):
        _poke_energy(sim, case, 7 + 20 * (1 + 3 * 1))
    assert o.run(4) == T.MH_ERR_NAN_REMAP and g.run(4) == T.MH_ERR_NAN_REMAP
    assert g.iteration == o.iteration == 7
    frozen(g, case)
    g.close()

    print("[status] inverted cell", flush=True)
    case = cases.lag2d_sod_x()
    o, g = OracleCase(case, pow_mode=1), GpuCase(case)
    assert o.run(6) == T.MH_OK and g.run(6) == T.MH_OK
    nx = case.n[0]
    for sim in (o, g):
        _poke_velocity(sim, case, 50 + (nx + 1) * 2, 5.0 * (1.0 / nx) / o.deltat)
    assert o.run(4) == T.MH_ERR_NEG_VOLUME and g.run(4) == T.MH_ERR_NEG_VOLUME
    assert g.iteration == o.iteration == 7
    frozen(g, case)
    g.close()


@pytest.mark.gpu
def test_gpu_case_on_caller_stream():
    """mahyco_create_on_stream: the case's work on a stream the caller created; same bits as
    a case on its own stream."""
    from mahyco_amd.lib import GpuCase
    hip = C.CDLL("libamdhip64.so")
    stream = C.c_void_p()
    assert hip.hipStreamCreate(C.byref(stream)) == 0
    try:
        case = cases.bisod_x(24, 4, 4)
        a, b = GpuCase(case), GpuCase(case, stream=stream.value)
        a.run(18, graph=6)
        b.run(18, graph=6)
        fa, fb = fields(a, 2, 3), fields(b, 2, 3)
        for k in fa:
            assert bitwise_equal(fa[k], fb[k]), k
        a.close()
        b.close()
    finally:
        hip.hipStreamDestroy(stream)
