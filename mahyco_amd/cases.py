"""Case descriptions: the datasets of the reference's non-regression fixtures and the
benchmark workloads, as MhMesh / MhEnvironment / MhOptions triples.

Defaults are the reference option defaults: Mahyco.axl:263-557 (module),
RemapADI.axl:165-209 (remap), PerfectGasEOS.axl / StiffenedGasEOS.axl (EOS).
"""
from dataclasses import dataclass, field
from typing import List, Optional, Tuple

import numpy as np

from . import types as T


@dataclass
class Env:
    densite_initiale: float = 0.0
    pression_initiale: float = 1.0e5
    adiabatic_cst: float = 1.4
    eos: int = T.MH_EOS_PERFECT_GAS
    limit_tension: float = 0.0
    specific_heat: float = 0.0
    linear_pseudo_coeff: float = 0.5
    quadratic_pseudo_coeff: float = 1.2
    vitesse_initiale: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    energie_initiale: float = 0.0
    fictif_tdebut: float = 0.0
    fictif_tfin: float = 0.0
    fictif_p0: float = 0.0
    fictif_dpdt: float = 0.0

    def to_c(self) -> T.MhEnvironment:
        e = T.MhEnvironment()
        e.eos = self.eos
        e.adiabatic_cst = self.adiabatic_cst
        e.limit_tension = self.limit_tension
        e.specific_heat = self.specific_heat
        e.linear_pseudo_coeff = self.linear_pseudo_coeff
        e.quadratic_pseudo_coeff = self.quadratic_pseudo_coeff
        e.densite_initiale = self.densite_initiale
        e.pression_initiale = self.pression_initiale
        for i in range(3):
            e.vitesse_initiale[i] = self.vitesse_initiale[i]
        e.energie_initiale = self.energie_initiale
        e.fictif_tdebut, e.fictif_tfin = self.fictif_tdebut, self.fictif_tfin
        e.fictif_p0, e.fictif_dpdt = self.fictif_p0, self.fictif_dpdt
        return e


@dataclass
class Case:
    name: str
    dim: int
    n: Tuple[int, int, int]
    origin: Tuple[float, float, float]
    length: Tuple[float, float, float]
    envs: List[Env]
    cas_model: int
    cas_test: int = 0
    with_projection: bool = True
    cfl: float = 0.3
    threshold: float = 1.0e-10
    deltat_init: float = 1.0e-5
    deltat_min: float = 1.0e-8
    deltat_max: float = 0.01
    final_time: float = 0.2
    calcul_vitson_max: bool = True
    bcs: List[Tuple[str, str, float]] = field(default_factory=list)
    ordre_projection: int = 2
    limiter: int = T.MH_MINMOD
    remap_threshold: float = 1.0e-10
    interface_position: float = 0.5
    interface_axis: int = 0
    sedov_deposit: float = 0.244816
    coord_rule: int = T.MH_COORD_DIV
    max_cycles: Optional[int] = None
    pente_borne: bool = False
    debar_fix: int = 0
    pente_borne_mixte: bool = False
    limiter_pure: int = T.MH_MINMOD
    simple_pente: bool = False
    with_newton: bool = False
    pression_explicite: bool = False
    schema_csts: bool = False
    longueur_caracteristique: int = 0  # MhCaracteristicLength
    conservation_energie_totale: bool = False
    sans_lagrange: bool = False
    rider_reverse: bool = False
    rider_parametre: float = 0.0
    rider_rayon_bulle: float = 0.15
    rider_densite_lineaire: bool = False

    @property
    def nb_cells(self) -> int:
        nx, ny, nz = self.n
        return nx * ny * (nz if self.dim == 3 else 1)

    @property
    def nb_nodes(self) -> int:
        nx, ny, nz = self.n
        return (nx + 1) * (ny + 1) * ((nz + 1) if self.dim == 3 else 1)

    @property
    def cell_shape(self):
        nx, ny, nz = self.n
        return (nz, ny, nx) if self.dim == 3 else (ny, nx)

    @property
    def node_shape(self):
        nx, ny, nz = self.n
        return (nz + 1, ny + 1, nx + 1) if self.dim == 3 else (ny + 1, nx + 1)

    def mesh_c(self) -> T.MhMesh:
        m = T.MhMesh()
        m.dim = self.dim
        m.nx, m.ny, m.nz = self.n[0], self.n[1], (self.n[2] if self.dim == 3 else 1)
        for i in range(3):
            m.origin[i] = self.origin[i]
            m.length[i] = self.length[i]
        m.coord_rule = self.coord_rule
        return m

    def envs_c(self):
        arr = (T.MhEnvironment * len(self.envs))()
        for i, e in enumerate(self.envs):
            arr[i] = e.to_c()
        return arr

    def options_c(self) -> T.MhOptions:
        o = T.MhOptions()
        o.cfl = self.cfl
        o.threshold = self.threshold
        o.deltat_init = self.deltat_init
        o.deltat_min = self.deltat_min
        o.deltat_max = self.deltat_max
        o.final_time = self.final_time
        o.with_projection = int(self.with_projection)
        o.calcul_vitson_max = int(self.calcul_vitson_max)
        o.nb_bc = len(self.bcs)
        for i, (surf, typ, val) in enumerate(self.bcs):
            o.bc[i].surface = T.SURFACES[surf]
            o.bc[i].type = T.BC_TYPES[typ]
            o.bc[i].value = val
        o.ordre_projection = self.ordre_projection
        o.projection_limiteur_id = self.limiter
        o.remap_threshold = self.remap_threshold
        o.cas_model = self.cas_model
        o.cas_test = self.cas_test
        o.interface_position = self.interface_position
        o.interface_axis = self.interface_axis
        o.sedov_deposit = self.sedov_deposit
        o.projection_pente_borne = int(self.pente_borne)
        o.pente_borne_debar_fix = self.debar_fix
        o.pente_borne_mixte = int(self.pente_borne_mixte)
        o.projection_limiteur_pure_id = self.limiter_pure
        o.projection_simple_pente = int(self.simple_pente)
        o.with_newton = int(self.with_newton)
        o.pression_explicite = int(self.pression_explicite)
        o.schema_csts = int(self.schema_csts)
        o.longueur_caracteristique = int(self.longueur_caracteristique)
        o.conservation_energie_totale = int(self.conservation_energie_totale)
        o.sans_lagrange = int(self.sans_lagrange)
        o.rider_reverse = int(self.rider_reverse)
        o.rider_parametre = self.rider_parametre
        o.rider_rayon_bulle = self.rider_rayon_bulle
        o.rider_densite_lineaire = int(self.rider_densite_lineaire)
        return o


WALLS_3D = [("XMIN", "Vx", 0.0), ("XMAX", "Vx", 0.0), ("YMIN", "Vy", 0.0), ("YMAX", "Vy", 0.0),
            ("ZMIN", "Vz", 0.0), ("ZMAX", "Vz", 0.0)]
WALLS_2D = WALLS_3D  # in 2D every face also lies in ZMIN/ZMAX; Vz=0 there is harmless


def _pg(rho, p, gamma=1.4, cv=2.4):
    return Env(densite_initiale=rho, pression_initiale=p, adiabatic_cst=gamma, specific_heat=cv)


# ----------------------------------------------------------------------------- fixtures
def bisod_x(nx=100, ny=5, nz=5, **kw) -> Case:
    """NONREGRESSION/CAS_BiSodCaseX/Donnees.arc (3D ADI remap, 2 PerfectGas envs, minmod)."""
    c = Case("CAS_BiSodCaseX", 3, (nx, ny, nz), (0., 0., 0.), (1., .1, .1),
             [_pg(1., 1.), _pg(.125, .1)], T.MH_CASE_SOD, 13,
             deltat_init=1e-5, deltat_min=1e-8, deltat_max=0.01, final_time=.2, bcs=list(WALLS_3D))
    for k, v in kw.items():
        setattr(c, k, v)
    return c


def bisod_y() -> Case:
    """NONREGRESSION/CAS_BiSodCaseY/Donnees.arc (5x100x5)."""
    c = bisod_x()
    c.name, c.n, c.length, c.cas_test = "CAS_BiSodCaseY", (5, 100, 5), (.1, 1., .1), 14
    return c


def bisod_z() -> Case:
    """NONREGRESSION/CAS_BiSodCaseZ/Donnees.arc (5x5x100)."""
    c = bisod_x()
    c.name, c.n, c.length, c.cas_test = "CAS_BiSodCaseZ", (5, 5, 100), (.1, .1, 1.), 15
    return c


def sod_mono(axis="Y") -> Case:
    """NONREGRESSION/CAS_SodCase{Y,Z}/Donnees.arc (3D ADI remap, mono Sod, 2 declared envs)."""
    n = {"X": (100, 5, 5), "Y": (5, 100, 5), "Z": (5, 5, 100)}[axis]
    ln = {"X": (1., .1, .1), "Y": (.1, 1., .1), "Z": (.1, .1, 1.)}[axis]
    t = {"X": 3, "Y": 4, "Z": 5}[axis]
    return Case("CAS_SodCase" + axis, 3, n, (0., 0., 0.), ln, [_pg(1., 1.), _pg(.125, .1)],
                T.MH_CASE_SOD, t, deltat_init=1e-5, deltat_min=1e-8, deltat_max=0.01, final_time=.2,
                bcs=list(WALLS_3D))


def ordre3(make, name) -> Case:
    """The NONREGRESSION/CAS_*_ordre_3 variants: the same case with <ordre-projection>3."""
    c = make()
    c.name, c.ordre_projection = name, 3
    return c


def lag_bisod_x() -> Case:
    """NONREGRESSION/CAS_Lag_BiSodCaseX/Donnees.arc (3D Lagrange only)."""
    c = bisod_x()
    c.name, c.with_projection, c.deltat_max = "CAS_Lag_BiSodCaseX", False, 1e-4
    return c


def lag2d_sod_x() -> Case:
    """NONREGRESSION/CAS_2D_Lag_SodCaseX/Donnees.arc (2D Lagrange only, mono)."""
    return Case("CAS_2D_Lag_SodCaseX", 2, (100, 5, 1), (0., 0., 0.), (1., .1, 0.),
                [_pg(1., 1.), _pg(.125, .1)], T.MH_CASE_SOD, 3, with_projection=False,
                deltat_init=1e-5, deltat_min=1e-8, deltat_max=1e-4, final_time=.2, bcs=list(WALLS_2D))


def pente_borne(make, name, debar, limiter=T.MH_MINMOD) -> Case:
    """The NONREGRESSION/CAS_BiSodCase*_pente_borne_* variants: fluxes integrated over the cell
    (projection-pente-borne) with the given debar fix and limiter."""
    c = make()
    c.name, c.pente_borne, c.debar_fix, c.limiter = name, True, debar, limiter
    return c


def lag2d_sod_x_newton() -> Case:
    """NONREGRESSION/CAS_2D_Lag_SodCaseX_newton/Donnees.arc (energy by Newton iterations)."""
    c = lag2d_sod_x()
    c.name, c.with_newton = "CAS_2D_Lag_SodCaseX_newton", True
    return c


def lag_bisod(axis) -> Case:
    """NONREGRESSION/CAS_Lag_BiSodCase{X,Y,Z}/Donnees.arc (3D Lagrange only, two materials)."""
    c = {"X": bisod_x, "Y": bisod_y, "Z": bisod_z}[axis]()
    c.name, c.with_projection, c.deltat_max = "CAS_Lag_BiSodCase" + axis, False, 1e-4
    return c


def lag_sod(axis) -> Case:
    """NONREGRESSION/CAS_Lag_SodCase{X,Y,Z}/Donnees.arc (3D Lagrange only, mono Sod)."""
    c = sod_mono(axis)
    c.name, c.with_projection, c.deltat_max = "CAS_Lag_SodCase" + axis, False, 1e-4
    return c


def lag2d(axis, bi) -> Case:
    """NONREGRESSION/CAS_2D_Lag_{Bi,}SodCase{X,Y,Z}/Donnees.arc: Lagrange only on a 2D mesh
    along X or Y (100x5, 5x100), and on a 5x1x100 3D mesh for Z."""
    t = {"X": 3, "Y": 4, "Z": 5}[axis] + (10 if bi else 0)
    name = "CAS_2D_Lag_" + ("BiSod" if bi else "Sod") + "Case" + axis
    if axis == "Z":
        dim, n, ln = 3, (5, 1, 100), (.1, .1, 1.)
    else:
        dim, n, ln = 2, ((100, 5, 1) if axis == "X" else (5, 100, 1)), ((1., .1, 0.) if axis == "X" else (.1, 1., 0.))
    return Case(name, dim, n, (0., 0., 0.), ln, [_pg(1., 1.), _pg(.125, .1)], T.MH_CASE_SOD, t,
                with_projection=False, deltat_init=1e-5, deltat_min=1e-8, deltat_max=1e-4, final_time=.2,
                bcs=list(WALLS_3D))


def sedov_3d_fixture() -> Case:
    """NONREGRESSION/CAS_Sedov/Donnees.arc (3D ADI remap, 60^3 on [0,1.2]^3, to t = 1e-3)."""
    return Case("CAS_Sedov", 3, (60, 60, 60), (0., 0., 0.), (1.2, 1.2, 1.2),
                [_pg(1., 0.0979264e-5), _pg(1., 1.)], T.MH_CASE_SEDOV, 1,
                deltat_init=1e-5, deltat_min=1e-8, deltat_max=0.01, final_time=1e-3,
                bcs=[("XMIN", "Vx", 0.), ("YMIN", "Vy", 0.), ("ZMIN", "Vz", 0.)])


def sedov_2d() -> Case:
    """NONREGRESSION/CAS_2D_Sedov/Donnees.arc (2D ADI remap, 60x60 on [0,1.2]^2)."""
    return Case("CAS_2D_Sedov", 2, (60, 60, 1), (0., 0., 0.), (1.2, 1.2, 0.),
                [_pg(1., 0.0979264e-5), _pg(1., 1.)], T.MH_CASE_SEDOV, 1,
                deltat_init=1e-5, deltat_min=1e-8, deltat_max=0.01, final_time=1.,
                bcs=[("XMIN", "Vx", 0.), ("YMIN", "Vy", 0.)])


def bitriplepoint() -> Case:
    """NONREGRESSION/CAS_bitriplepoint_minmod/Donnees.arc (2D, 3 envs, vanLeer)."""
    return Case("CAS_bitriplepoint_minmod", 2, (70, 30, 1), (0., 0., 0.), (.07, .03, 0.),
                [_pg(1., 1.), _pg(1., .1), _pg(.1, .1)], T.MH_CASE_POINTTRIPLE, 12,
                threshold=1e-16, deltat_init=1e-9, deltat_min=1e-10, deltat_max=1e-5,
                final_time=.041, limiter=T.MH_VANLEER,
                bcs=[("XMIN", "Vx", 0.), ("XMAX", "Vx", 0.), ("YMIN", "Vy", 0.), ("YMAX", "Vy", 0.)])


def mono_pointtriple(simple_pente=False) -> Case:
    """NONREGRESSION/CAS_MonoPointtriple_pb{,_sp}_mixte_superbeeG_superbeeG/Donnees.arc: the
    triple point with pente-borne fluxes, the pure-neighbourhood variant, debar fix 1 and the
    generalised superbee limiter (optionally in simple-pente form), remap threshold 1e-16."""
    c = bitriplepoint()
    c.name = "CAS_MonoPointtriple_pb_" + ("sp_" if simple_pente else "") + "mixte_superbeeG_superbeeG"
    c.pente_borne, c.debar_fix, c.pente_borne_mixte, c.simple_pente = True, 1, True, simple_pente
    c.limiter = c.limiter_pure = T.MH_SUPERBEE_G
    c.remap_threshold = 1e-16
    return c


RIDER_TESTS = {"MonoRiderTx": 23, "MonoRiderTy": 24, "MonoRiderT45": 25, "MonoRiderRotation": 26,
               "MonoRiderVortex": 27, "MonoRiderDeformation": 28, "MonoRiderVortexTimeReverse": 29,
               "MonoRiderDeformationTimeReverse": 30, "RiderTx": 31, "RiderTy": 32, "RiderT45": 33,
               "RiderRotation": 34, "RiderVortex": 35, "RiderDeformation": 36, "RiderVortexTimeReverse": 37,
               "RiderDeformationTimeReverse": 38}


def _fictif(rho):
    return Env(densite_initiale=rho, pression_initiale=0., eos=T.MH_EOS_FICTIF, specific_heat=2.4, fictif_tdebut=0.,
               fictif_tfin=10.)


def rider(name, cas_test, final_time, reverse=None, bcs=None) -> Case:
    """NONREGRESSION/CAS_*Rider*/Donnees.arc: pure advection (sans-lagrange) of a bubble on a
    50x50x1 mesh of [0,1]^2 x [0,0.1], two Fictif environments (Vide rho 0, Bulle rho 1),
    constant time step 1e-3; reverse = parametre of the reverse option."""
    c = Case(name, 3, (50, 50, 1), (0., 0., 0.), (1., 1., .1), [_fictif(0.), _fictif(1.)], T.MH_CASE_RIDER,
             cas_test, deltat_init=1e-3, deltat_min=1e-8, deltat_max=1e-4, final_time=final_time,
             bcs=list(bcs or []))
    c.sans_lagrange = True
    if reverse is not None:
        c.rider_reverse, c.rider_parametre = True, reverse
    return c


ADVECTION_TESTS = {"MonoAdvectionTx": 7, "MonoAdvectionTy": 8, "MonoAdvectionT45": 9, "MonoAdvectionRotation": 10,
                   "AdvectionTx": 18, "AdvectionTy": 19, "AdvectionT45": 20, "AdvectionRotation": 21}


def advection(name, cas_test, final_time, n=50, dt=1e-3, ordre=2, reverse=None, lineaire=False, bcs=None) -> Case:
    """ADVECTIONService (casTest/ADVECTION): pure advection (sans-lagrange) of a square of side
    2 rayon-bulle (0.25) centred at (0.5, 0.5), or (0.5, 0.75) for the rotations, on an n x n x 1
    mesh of [0,1]^2 x [0,0.1]; two Fictif environments (rho 0 outside, 1 inside).  The
    reference holds no dataset for this service; the layout is the RIDER datasets'."""
    c = Case(name, 3, (n, n, 1), (0., 0., 0.), (1., 1., .1), [_fictif(0.), _fictif(1.)], T.MH_CASE_ADVECTION,
             cas_test, deltat_init=dt, deltat_min=1e-8, deltat_max=1e-4, final_time=final_time,
             bcs=list(bcs or []))
    c.ordre_projection = ordre
    c.sans_lagrange = True
    c.rider_rayon_bulle = .25
    c.rider_densite_lineaire = lineaire
    if reverse is not None:
        c.rider_reverse, c.rider_parametre = True, reverse
    return c


# ----------------------------------------------------------------------------- workloads
def config1_sod2d_lagrange() -> Case:
    """2D Sod 200x10 on [0,1]x[0,0.1], single material, Lagrange only, CFL 0.3, t=0.2."""
    return Case("sod2d_lagrange_200x10", 2, (200, 10, 1), (0., 0., 0.), (1., .1, 0.),
                [_pg(1., 1.)], T.MH_CASE_SOD, 3, with_projection=False,
                deltat_init=1e-5, deltat_min=1e-8, deltat_max=1e-4, final_time=.2, bcs=list(WALLS_2D))


def config2_bisod3d(nx=256, ny=128, nz=128, cycles=100) -> Case:
    """3D bi-material shock tube [0,1]^3, 256x128x128, env B gamma 1.67, superbee, 100 cycles."""
    return Case(f"bisod3d_{nx}x{ny}x{nz}", 3, (nx, ny, nz), (0., 0., 0.), (1., 1., 1.),
                [_pg(1., 1., 1.4), _pg(.125, .1, 1.67)], T.MH_CASE_SOD, 13,
                deltat_init=1e-5, deltat_min=1e-12, deltat_max=0.01, final_time=1e30,
                limiter=T.MH_SUPERBEE, bcs=list(WALLS_3D), max_cycles=cycles)


def config3_sedov3d(n=192, cycles=200) -> Case:
    """3D Sedov blast 192^3 on [0,1.2]^3, rho=1, p=1e-6, minmod, symmetry walls, 200 cycles."""
    return Case(f"sedov3d_{n}", 3, (n, n, n), (0., 0., 0.), (1.2, 1.2, 1.2),
                [_pg(1., 1e-6)], T.MH_CASE_SEDOV, 1,
                deltat_init=1e-5, deltat_min=1e-12, deltat_max=0.01, final_time=1e30,
                limiter=T.MH_MINMOD, bcs=[("XMIN", "Vx", 0.), ("YMIN", "Vy", 0.), ("ZMIN", "Vz", 0.)],
                max_cycles=cycles)


def riemann2d(nx, ny, rng: np.random.Generator, stiffened=False, t_end=None, cycles=None,
              name="riemann2d") -> Case:
    """Two-state 2D shock tube with a seeded interface and per-side state (ensemble member)."""
    envs = []
    for _ in range(2):
        if stiffened:
            g = float(rng.uniform(1.4, 4.4))
            pi = 0.0 if rng.random() < 0.5 else float(rng.uniform(1e-3, 0.6))
            e = Env(densite_initiale=float(rng.uniform(0.1, 2.0)), pression_initiale=float(rng.uniform(0.05, 2.0)),
                    adiabatic_cst=g, eos=T.MH_EOS_STIFFENED_GAS, limit_tension=pi)
        else:
            e = _pg(float(rng.uniform(0.1, 2.0)), float(rng.uniform(0.05, 2.0)), float(rng.uniform(1.2, 1.67)))
        envs.append(e)
    x0 = float(rng.uniform(0.3, 0.7))
    return Case(name, 2, (nx, ny, 1), (0., 0., 0.), (1., 1., 0.), envs, T.MH_CASE_RIEMANN, 0,
                deltat_init=1e-5, deltat_min=1e-12, deltat_max=0.01,
                final_time=(t_end if t_end is not None else 1e30), bcs=list(WALLS_2D),
                interface_position=x0, interface_axis=0, max_cycles=cycles)


def riemann3d(n, rng: np.random.Generator, stiffened=True, t_end=0.1, name="riemann3d") -> Case:
    c = riemann2d(n, n, rng, stiffened=stiffened, t_end=t_end, name=name)
    c.dim, c.n, c.length, c.bcs = 3, (n, n, n), (1., 1., 1.), list(WALLS_3D)
    return c
