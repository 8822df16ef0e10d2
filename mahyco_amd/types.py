"""ctypes mirrors of include/mahyco_types.h (the case description that crosses the C ABI)."""
import ctypes as C

MH_MAX_ENV = 4
MH_MAX_BC = 8

# MhStatus
MH_OK = 0
MH_STOPPED = 1
MH_ERR_NEG_VOLUME = -1
MH_ERR_DT_MIN = -2
MH_ERR_NAN_REMAP = -3
MH_ERR_INVALID = -10
MH_ERR_UNSUPPORTED = -11
MH_ERR_DEVICE = -20
MH_ERR_NOMEM = -21

STATUS_NAMES = {
    MH_OK: "MH_OK", MH_STOPPED: "MH_STOPPED", MH_ERR_NEG_VOLUME: "MH_ERR_NEG_VOLUME",
    MH_ERR_DT_MIN: "MH_ERR_DT_MIN", MH_ERR_NAN_REMAP: "MH_ERR_NAN_REMAP",
    MH_ERR_INVALID: "MH_ERR_INVALID", MH_ERR_UNSUPPORTED: "MH_ERR_UNSUPPORTED",
    MH_ERR_DEVICE: "MH_ERR_DEVICE", MH_ERR_NOMEM: "MH_ERR_NOMEM",
}

MH_EOS_PERFECT_GAS = 0
MH_EOS_STIFFENED_GAS = 1
MH_EOS_FICTIF = 2

MH_XMIN, MH_XMAX, MH_YMIN, MH_YMAX, MH_ZMIN, MH_ZMAX = range(6)
SURFACES = {"XMIN": MH_XMIN, "XMAX": MH_XMAX, "YMIN": MH_YMIN, "YMAX": MH_YMAX,
            "ZMIN": MH_ZMIN, "ZMAX": MH_ZMAX}
MH_BC_VX, MH_BC_VY, MH_BC_VZ = range(3)
BC_TYPES = {"Vx": MH_BC_VX, "Vy": MH_BC_VY, "Vz": MH_BC_VZ}

MH_MINMOD, MH_SUPERBEE, MH_VANLEER = 0, 1, 2
MH_MINMOD_G, MH_SUPERBEE_G, MH_VANLEER_G, MH_ARITHMETIC_G, MH_ULTRABEE_G = 3, 4, 5, 6, 7

MH_CASE_SOD, MH_CASE_SEDOV, MH_CASE_RIEMANN, MH_CASE_POINTTRIPLE, MH_CASE_RIDER, MH_CASE_ADVECTION = range(6)
# MhCaracteristicLength
MH_L_RACINE_CUBIQUE_VOLUME, MH_L_FACES_OPPOSEES, MH_L_MONODIM_X, MH_L_MONODIM_Y, MH_L_MONODIM_Z = range(5)
MH_COORD_MUL, MH_COORD_DIV = 0, 1

# MhField
MH_F_DENSITY = 0
MH_F_PRESSURE = 1
MH_F_INTERNAL_ENERGY = 2
MH_F_SOUND_SPEED = 3
MH_F_PSEUDO_VISCOSITY = 4
MH_F_CELL_MASS = 5
MH_F_CELL_VOLUME = 6
MH_F_FRACVOL = 7
MH_F_MASS_FRACTION = 8
MH_F_CARACTERISTIC_LENGTH = 9
MH_F_ENV_MASK = 10
MH_F_MATERIAU = 11
MH_N_VELOCITY = 100
MH_N_NODE_MASS = 101
MH_N_COORD = 102
MH_N_COORD_LAGRANGE = 103

CELL_FIELDS = {
    "Density": MH_F_DENSITY, "Pressure": MH_F_PRESSURE, "InternalEnergy": MH_F_INTERNAL_ENERGY,
    "SoundSpeed": MH_F_SOUND_SPEED, "PseudoViscosity": MH_F_PSEUDO_VISCOSITY,
    "CellMass": MH_F_CELL_MASS, "CellVolume": MH_F_CELL_VOLUME, "FracVol": MH_F_FRACVOL,
    "MassFraction": MH_F_MASS_FRACTION, "CaracteristicLength": MH_F_CARACTERISTIC_LENGTH,
    "EnvMask": MH_F_ENV_MASK, "Materiau": MH_F_MATERIAU,
}


class MhMesh(C.Structure):
    _fields_ = [("dim", C.c_int), ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("origin", C.c_double * 3), ("length", C.c_double * 3), ("coord_rule", C.c_int)]


class MhEnvironment(C.Structure):
    _fields_ = [("eos", C.c_int), ("adiabatic_cst", C.c_double), ("limit_tension", C.c_double),
                ("specific_heat", C.c_double), ("linear_pseudo_coeff", C.c_double),
                ("quadratic_pseudo_coeff", C.c_double), ("densite_initiale", C.c_double),
                ("pression_initiale", C.c_double), ("vitesse_initiale", C.c_double * 3),
                ("energie_initiale", C.c_double), ("fictif_tdebut", C.c_double), ("fictif_tfin", C.c_double),
                ("fictif_p0", C.c_double), ("fictif_dpdt", C.c_double)]


class MhBoundaryCondition(C.Structure):
    _fields_ = [("surface", C.c_int), ("type", C.c_int), ("value", C.c_double)]


class MhOptions(C.Structure):
    _fields_ = [("cfl", C.c_double), ("threshold", C.c_double), ("deltat_init", C.c_double),
                ("deltat_min", C.c_double), ("deltat_max", C.c_double), ("final_time", C.c_double),
                ("with_projection", C.c_int), ("calcul_vitson_max", C.c_int), ("nb_bc", C.c_int),
                ("bc", MhBoundaryCondition * MH_MAX_BC), ("ordre_projection", C.c_int),
                ("projection_limiteur_id", C.c_int), ("remap_threshold", C.c_double),
                ("cas_model", C.c_int), ("cas_test", C.c_int), ("interface_position", C.c_double),
                ("interface_axis", C.c_int), ("sedov_deposit", C.c_double),
                ("projection_pente_borne", C.c_int), ("pente_borne_debar_fix", C.c_int),
                ("pente_borne_mixte", C.c_int), ("projection_limiteur_pure_id", C.c_int),
                ("projection_simple_pente", C.c_int), ("with_newton", C.c_int),
                ("pression_explicite", C.c_int), ("sans_lagrange", C.c_int), ("rider_reverse", C.c_int),
                ("rider_parametre", C.c_double), ("rider_rayon_bulle", C.c_double),
                ("rider_densite_lineaire", C.c_int), ("schema_csts", C.c_int), ("longueur_caracteristique", C.c_int), ("conservation_energie_totale", C.c_int)]
