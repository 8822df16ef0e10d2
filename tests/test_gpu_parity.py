"""HIP path against the CPU oracle on the same inputs (GPU tests; call the C ABI).

Two bars:
  * bit-exact against the oracle in pow mode 1 (the oracle then uses the same correctly
    rounded V^(1/3) as the kernels; every other operation follows the reference order);
  * max relative error <= 1e-10 per cell (rho, e, p, volume fractions; velocity relative
    to its maximum) against the oracle in pow mode 0, which calls std::pow like the
    reference and is pinned to the reference's non-regression fixtures.
"""
import numpy as np
import pytest

from mahyco_amd import cases
from mahyco_amd import types as T
from tests.helpers import bitwise_equal, fields, max_rel, rel_err_scaled

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _riemann2d(stiff, seed, n=40):
    rng = np.random.default_rng(seed)
    return cases.riemann2d(n, n // 2, rng, stiffened=stiff, name=f"riemann2d_{seed}")


def _riemann3d(seed, n=14):
    rng = np.random.default_rng(seed)
    return cases.riemann3d(n, rng, stiffened=True, t_end=1e30, name=f"riemann3d_{seed}")


def _two_d(c, n):
    c.dim, c.n, c.length = 2, (n, n, 1), (1., 1., 0.)
    return c


def _sedov2d_mono():
    c = cases.sedov_2d()
    c.envs = c.envs[:1]
    return c


def _variant(c, **kw):
    for key, v in kw.items():
        setattr(c, key, v)
    c.name += "_" + "_".join(f"{key}{v}" for key, v in kw.items())
    return c


CASES = {
    "bisod_x_fixture": (cases.bisod_x, 400),       # runs to the final time (183 iterations)
    "bisod_y_fixture": (cases.bisod_y, 80),
    "bisod_z_fixture": (cases.bisod_z, 80),
    "sod2d_lagrange_200x10": (cases.config1_sod2d_lagrange, 3000),  # to t = 0.2
    "lag_bisod_x_fixture": (cases.lag_bisod_x, 300),
    "sedov2d_fixture": (cases.sedov_2d, 300),
    "bitriplepoint_fixture": (cases.bitriplepoint, 400),
    "bisod3d_superbee_32x16x16": (lambda: cases.config2_bisod3d(32, 16, 16), 100),
    "sedov3d_minmod_20": (lambda: cases.config3_sedov3d(20), 100),
    "riemann2d_perfect_gas": (lambda: _riemann2d(False, 11), 150),
    "riemann2d_stiffened_gas": (lambda: _riemann2d(True, 12), 150),
    "riemann3d_stiffened_gas": (lambda: _riemann3d(13), 60),
    # order-3 remap (Remap-ordre-3.cc): the reference's two- and one-material tubes, a 3D
    # tube in all directions, and Riemann problems with mixed cells (mass / energy face
    # values rebuilt from the upwind ratios)
    "bisod_x_ordre3_fixture": (lambda: cases.ordre3(cases.bisod_x, "bisod_x_o3"), 400),
    "sod_z_ordre3_fixture": (lambda: cases.ordre3(lambda: cases.sod_mono("Z"), "sod_z_o3"), 120),
    "bisod3d_ordre3_32x16x16": (lambda: cases.ordre3(lambda: cases.config2_bisod3d(32, 16, 16), "b3d_o3"), 100),
    "riemann2d_stiffened_ordre3": (lambda: cases.ordre3(lambda: _riemann2d(True, 14), "r2d_o3"), 150),
    "riemann3d_stiffened_ordre3": (lambda: cases.ordre3(lambda: _riemann3d(15), "r3d_o3"), 60),
    # pente-borne fluxes (UtilesRemap.cc:425-832) and generalised limiters (:130-183)
    "bisod_x_pente_borne_debar1": (lambda: cases.pente_borne(cases.bisod_x, "pb1", 1), 400),
    "bisod_y_pente_borne_superbee_debar2": (lambda: cases.pente_borne(cases.bisod_y, "pb2", 2, T.MH_SUPERBEE), 120),
    "bisod_z_pente_borne_vanleer_debar0": (lambda: cases.pente_borne(cases.bisod_z, "pb0", 0, T.MH_VANLEER), 120),
    "pointtriple_pb_superbeeG_fixture": (lambda: cases.mono_pointtriple(False), 400),
    "pointtriple_pb_sp_superbeeG_fixture": (lambda: cases.mono_pointtriple(True), 400),
    "riemann3d_pb_mixte_purelim1_vanleerG": (lambda: _variant(_riemann3d(16), pente_borne=True, debar_fix=1,
                                                              pente_borne_mixte=True, limiter=T.MH_VANLEER_G,
                                                              limiter_pure=T.MH_SUPERBEE), 60),
    "riemann2d_minmodG": (lambda: _variant(_riemann2d(True, 17), limiter=T.MH_MINMOD_G), 150),
    "riemann2d_vanleerG_simple_pente": (lambda: _variant(_riemann2d(True, 18), limiter=T.MH_VANLEER_G,
                                                         simple_pente=True), 150),
    "riemann2d_arithmeticG_pb": (lambda: _variant(_riemann2d(False, 19), limiter=T.MH_ARITHMETIC_G,
                                                  pente_borne=True, debar_fix=2), 150),
    "riemann3d_ultrabeeG": (lambda: _variant(_riemann3d(20), limiter=T.MH_ULTRABEE_G), 60),
    # energy of the Lagrange step by Newton iterations or with explicit pressure work
    "sod2d_lagrange_newton_fixture": (cases.lag2d_sod_x_newton, 3000),
    "riemann3d_stiffened_newton": (lambda: _variant(_riemann3d(21), with_newton=True), 60),
    "riemann2d_explicit_pressure": (lambda: _variant(_riemann2d(True, 22), pression_explicite=True), 150),
    "sod2d_lagrange_explicit_pressure": (lambda: _variant(cases.config1_sod2d_lagrange(), pression_explicite=True), 800),
    # advection mode (sans-lagrange) with the RIDER bubble: one and two environments, the
    # reverse option, pente-borne with the pure-neighbourhood limiter
    "monorider_rotation_fixture": (lambda: cases.rider("mr_rot", 26, .5), 500),
    "rider_deformation_reverse_fixture": (lambda: cases.rider("r_def_rev", 38, 1., 1.), 1000),
    "rider_vortex_reverse_fixture": (lambda: cases.rider("r_vort_rev", 37, 4., .25), 600),
    "rider_ty_pb_mixte_fixture": (lambda: _variant(cases.rider("r_ty_pb", 32, .5), pente_borne=True, debar_fix=2,
                                                   limiter=T.MH_SUPERBEE, pente_borne_mixte=True,
                                                   limiter_pure=T.MH_VANLEER), 500),
    # advection mode with the ADVECTION square: the reversed rotation, the linear density at
    # Courant number just below 1 (cells emptied of both environments), order 3, pente-borne
    # with a generalised limiter
    "monoadvection_rotation_reverse": (lambda: cases.advection("ma_rot", 10, .25, n=64, reverse=2.), 250),
    "advection_t45_lineaire_courant1": (lambda: cases.advection("a_t45", 20, 1., n=50, dt=(1. - 2. ** -30) / 50,
                                                                ordre=1, lineaire=True), 8),
    "advection_rotation_ordre3": (lambda: cases.advection("a_rot_o3", 21, .1, n=48, ordre=3), 100),
    "advection_ty_pb_superbeeG": (lambda: _variant(cases.advection("a_ty_pbg", 19, .2, n=40, dt=2e-3),
                                                   pente_borne=True, debar_fix=1, limiter=T.MH_SUPERBEE_G), 100),
    # 2D advection meshes (the datasets are 3D with one cell in z)
    "rider2d_rotation_reverse": (lambda: _two_d(cases.rider("r2rev", 34, .3, 2.), 48), 300),
    "advection2d_t45_ordre3": (lambda: _variant(_two_d(cases.advection("a2", 20, .2, n=40), 40), ordre_projection=3), 200),
    # characteristic lengths other than V^(1/d) (k_cell_length before the cell kernel)
    "bisod3d_faces_opposees": (lambda: _variant(cases.config2_bisod3d(32, 16, 16), longueur_caracteristique=1), 100),
    "riemann2d_monodimX": (lambda: _variant(_riemann2d(True, 26), longueur_caracteristique=2), 150),
    "riemann3d_monodimY_csts": (lambda: _variant(_riemann3d(27), longueur_caracteristique=3, schema_csts=True), 60),
    # CSTS (single environment): backward / forward velocity steps, corner-work energy, mean
    # face velocities; 2D and 3D, order 2 and 3, with and without the tiled kernels' paths
    "sedov3d_csts_20": (lambda: _variant(cases.config3_sedov3d(20), schema_csts=True), 100),
    "sod2d_remap_csts": (lambda: _variant(cases.config1_sod2d_lagrange(), with_projection=True, schema_csts=True), 600),
    "sedov2d_csts_ordre3": (lambda: _variant(_sedov2d_mono(), schema_csts=True, ordre_projection=3), 200),
    # CSTS with mixed cells: per-environment pseudo-viscosity history through membership changes
    "bisod3d_csts_32x16x16": (lambda: _variant(cases.config2_bisod3d(32, 16, 16), schema_csts=True), 100),
    "riemann2d_stiffened_csts": (lambda: _variant(_riemann2d(True, 23), schema_csts=True), 150),
    "bitriplepoint_csts": (lambda: _variant(cases.bitriplepoint(), schema_csts=True), 300),
    # CSTS without remap: no backward step, each environment's own pseudo-viscosity history
    "sod2d_lagrange_csts": (lambda: _variant(cases.config1_sod2d_lagrange(), schema_csts=True), 800),
    "riemann2d_lagrange_csts": (lambda: _variant(_riemann2d(True, 24), with_projection=False, schema_csts=True), 150),
    "riemann3d_lagrange_csts": (lambda: _variant(_riemann3d(25), with_projection=False, schema_csts=True), 60),
    # conservation-energie-totale: the dual kinetic energy through every sweep and its return
    # to the cells (RemapADIFinal.cc:318-345); one and two sweeps, orders 1-3, pente-borne,
    # generalised limiters, CSTS, mixed cells
    "bisod_x_conservation_fixture": (lambda: _variant(cases.bisod_x(), conservation_energie_totale=True), 400),
    "bisod_x_one_sweep_conservation": (lambda: _variant(cases.bisod_x(40, 1, 1), conservation_energie_totale=True), 120),
    "riemann2d_stiffened_conservation": (lambda: _variant(_riemann2d(True, 28), conservation_energie_totale=True), 150),
    "riemann2d_conservation_ordre1": (lambda: _variant(_riemann2d(False, 31), conservation_energie_totale=True,
                                                       ordre_projection=1), 150),
    "bisod3d_conservation_ordre3": (lambda: _variant(cases.ordre3(lambda: cases.config2_bisod3d(32, 16, 16), "b3d_o3"),
                                                     conservation_energie_totale=True), 100),
    "riemann2d_conservation_pb_vanleerG": (lambda: _variant(_riemann2d(False, 30), conservation_energie_totale=True,
                                                            pente_borne=True, debar_fix=1, limiter=T.MH_VANLEER_G), 150),
    "riemann3d_conservation_csts": (lambda: _variant(_riemann3d(29), conservation_energie_totale=True,
                                                     schema_csts=True), 60),
}


def _run_pair(name, pow_mode):
    from mahyco_amd.lib import GpuCase
    from oracle.pyoracle import OracleCase
    make, ncycles = CASES[name]
    case = make()
    print(f"[gpu parity] {name} cells={case.nb_cells} cycles<={ncycles} pow_mode={pow_mode}", flush=True)
    g = GpuCase(case)
    st_g = g.run(ncycles)
    o = OracleCase(case, pow_mode=pow_mode)
    st_o = o.run(ncycles)
    return case, g, o, st_g, st_o


@pytest.mark.parametrize("name", list(CASES))
def test_bitexact_vs_oracle(name):
    case, g, o, st_g, st_o = _run_pair(name, pow_mode=1)
    assert st_g == st_o, (st_g, st_o)
    assert g.iteration == o.iteration
    assert g.time == o.time, (g.time, o.time)
    assert g.deltat == o.deltat, (g.deltat, o.deltat)
    fg = fields(g, len(case.envs), case.dim)
    fo = fields(o, len(case.envs), case.dim)
    bad = [k for k in fo if not bitwise_equal(fg[k], fo[k])]
    msg = {k: (max_rel(fg[k], fo[k]), int(np.sum(fg[k] != fo[k]))) for k in bad}
    assert not bad, msg
    g.close()


@pytest.mark.parametrize("name", ["bisod_x_fixture", "sedov2d_fixture", "bitriplepoint_fixture",
                                  "bisod3d_superbee_32x16x16", "riemann2d_stiffened_gas"])
def test_tolerance_vs_reference_pow(name):
    case, g, o, st_g, st_o = _run_pair(name, pow_mode=0)
    assert st_g == st_o
    assert g.iteration == o.iteration
    assert abs(g.time - o.time) <= TOL * abs(o.time)
    for f in (T.MH_F_DENSITY, T.MH_F_INTERNAL_ENERGY, T.MH_F_PRESSURE):
        assert max_rel(g.get(f), o.get(f)) <= TOL, f
        for k in range(len(case.envs)):
            assert max_rel(g.get(f, k), o.get(f, k)) <= TOL, (f, k)
    for k in range(len(case.envs)):
        assert rel_err_scaled(g.get(T.MH_F_FRACVOL, k), o.get(T.MH_F_FRACVOL, k)) <= TOL
    ug = np.stack([g.get(T.MH_N_VELOCITY, -1, c) for c in range(case.dim)])
    uo = np.stack([o.get(T.MH_N_VELOCITY, -1, c) for c in range(case.dim)])
    assert rel_err_scaled(ug, uo) <= TOL  # relative to the largest velocity component
    g.close()


def test_csts_unsupported_combinations_refused():
    """CSTS runs with the closed-form energy; Newton, explicit pressure and advection mode are
    refused at creation."""
    from mahyco_amd.lib import GpuCase
    bad = [_variant(cases.config3_sedov3d(8), schema_csts=True, with_newton=True),
           _variant(cases.config3_sedov3d(8), schema_csts=True, pression_explicite=True),
           _variant(cases.advection("a", 7, .1, n=10), schema_csts=True)]
    for c in bad:
        with pytest.raises(Exception):
            GpuCase(c)


def test_conservation_with_advection_refused():
    """Total-energy conservation reads the dual remap's kinetic energy, which advection mode
    does not run: the combination is refused at creation."""
    from mahyco_amd.lib import GpuCase
    c = _variant(cases.rider("r", 34, .2), conservation_energie_totale=True)
    with pytest.raises(Exception):
        GpuCase(c)


def test_bubble_cases_need_two_environments():
    """RIDER and ADVECTION (Mono tests included) give the bubble environment 1's density: a
    single environment is refused at creation."""
    from mahyco_amd.lib import GpuCase
    for c in (cases.advection("adv", 7, .1, n=10), cases.rider("r", 23, .1)):
        c.envs = c.envs[:1]
        with pytest.raises(Exception):
            GpuCase(c)
