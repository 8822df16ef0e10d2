/*
 * mahyco_types.h -- plain-C description of one MaHyCo hydro case on a structured
 * Cartesian mesh: mesh, environments (EOS + pseudo-viscosity coefficients),
 * module options, remap options, boundary conditions and the initial-condition
 * model.  The layout is POD so it crosses a C ABI (ctypes, cgo, JNI) unchanged.
 *
 * Every field names the reference option it carries:
 *   Mahyco.axl            src/Mahyco.axl:263-557        (module options)
 *   RemapADI.axl          src/Remap/RemapADI.axl:165-209 (remap options)
 *   PerfectGasEOS.axl     src/eos/perfectgas/PerfectGasEOS.axl:157-183
 *   StiffenedGasEOS.axl   src/eos/stiffenedgas/StiffenedGasEOS.axl:158-184
 *   <cartesian> block of the dataset (Arcane Cartesian mesh generator)
 */
#ifndef MAHYCO_TYPES_H
#define MAHYCO_TYPES_H

#ifdef __cplusplus
extern "C" {
#endif

#define MH_MAX_ENV 4 /* environments per case (bitmask per cell fits in one byte) */
#define MH_MAX_BC 8

/* Status codes.  The reference calls exit(1) at these sites; the C ABI returns them. */
enum MhStatus {
  MH_OK = 0,
  MH_STOPPED = 1,           /* final time reached (MahycoModule.cc:1841-1857) */
  MH_ERR_NEG_VOLUME = -1,   /* MahycoModule.cc:1164-1191 */
  MH_ERR_DT_MIN = -2,       /* MahycoModule.cc:1816-1833 */
  MH_ERR_NAN_REMAP = -3,    /* RemapADIFinal.cc:274-284 */
  MH_ERR_INVALID = -10,     /* bad case description or argument */
  MH_ERR_UNSUPPORTED = -11, /* option outside the implemented subset */
  MH_ERR_DEVICE = -20,      /* HIP runtime error */
  MH_ERR_NOMEM = -21        /* allocation failed / admission refused */
};

/* Equation-of-state services (IEquationOfState implementations). */
enum MhEos { MH_EOS_PERFECT_GAS = 0, MH_EOS_STIFFENED_GAS = 1, MH_EOS_FICTIF = 2 /* eos/fictif/FictifEOSService.cc */ };

/* Face groups built by PrepareFaceGroup (MahycoAnnexe.cc:149-235). */
enum MhSurface { MH_XMIN = 0, MH_XMAX = 1, MH_YMIN = 2, MH_YMAX = 3, MH_ZMIN = 4, MH_ZMAX = 5 };

/* TypesMahyco::eBoundaryCondition (TypesMahyco.h:16-31), velocity subset. */
enum MhBcType { MH_BC_VX = 0, MH_BC_VY = 1, MH_BC_VZ = 2 };

/* TypesMahyco limiteur (TypesMahyco.h:96-105): classical limiters (fluxLimiter,
 * UtilesRemap.cc:108-120) and the generalised ones for unequal cell widths (fluxLimiterG,
 * UtilesRemap.cc:130-183). */
enum MhLimiter {
  MH_MINMOD = 0,
  MH_SUPERBEE = 1,
  MH_VANLEER = 2,
  MH_MINMOD_G = 3,
  MH_SUPERBEE_G = 4,
  MH_VANLEER_G = 5,
  MH_ARITHMETIC_G = 6,
  MH_ULTRABEE_G = 7
};

/* Case models (IInitialisations services). */
enum MhCaseModel {
  MH_CASE_SOD = 0,         /* casTest/SOD/SODService.cc; cas_test 3,4,5 (mono) 13,14,15 (bi) */
  MH_CASE_SEDOV = 1,       /* casTest/SEDOV/SEDOVService.cc; cas_test 1 (mono) */
  MH_CASE_RIEMANN = 2,     /* two-state shock tube with geometric-overlap volume fractions */
  MH_CASE_POINTTRIPLE = 3, /* casTest/POINTTRIPLE/POINTTRIPLEService.cc; cas_test 12 (3 envs) */
  MH_CASE_RIDER = 4,       /* casTest/RIDER/RIDERService.cc; cas_test 23-30 (mono), 31-38 (bubble in 2 envs) */
  MH_CASE_ADVECTION = 5    /* casTest/ADVECTION/ADVECTIONService.cc; cas_test 7-10 (mono), 18-21 (square in 2 envs) */
};

/* characteristic length of a cell (computeGeometricValues, MahycoModule.cc:1195-1236) */
typedef enum {
  MH_L_RACINE_CUBIQUE_VOLUME = 0, /* V^(1/d) */
  MH_L_FACES_OPPOSEES = 1,        /* d1 d2 d3 / (d1 d2 + d1 d3 + d2 d3), medians between opposite faces (3D) */
  MH_L_MONODIM_X = 2,             /* |x(node 2) - x(node 0)| */
  MH_L_MONODIM_Y = 3,
  MH_L_MONODIM_Z = 4
} MhCaracteristicLength;

/* Node coordinate convention of the Cartesian generator. */
enum MhCoordRule { MH_COORD_MUL = 0 /* x_i = o + i*(l/n) */, MH_COORD_DIV = 1 /* x_i = o + (l*i)/n */ };

typedef struct MhMesh {
  int dim;           /* 2 or 3 */
  int nx, ny, nz;    /* cells per direction (nz ignored in 2D) */
  double origin[3];  /* <origine> */
  double length[3];  /* <lx>, <ly>, <lz> */
  int coord_rule;    /* MhCoordRule */
} MhMesh;

typedef struct MhEnvironment {
  int eos;                       /* MhEos */
  double adiabatic_cst;          /* gamma (adiabatic-cst, default 1.4) */
  double limit_tension;          /* stiffened-gas pi (limit-tension) */
  double specific_heat;          /* cv (specific-heat) */
  double linear_pseudo_coeff;    /* a1 (Mahyco.axl:322-327, default 0.5) */
  double quadratic_pseudo_coeff; /* a2 (Mahyco.axl:328-333, default 1.2) */
  double densite_initiale;       /* densite-initiale */
  double pression_initiale;      /* pression-initiale */
  double vitesse_initiale[3];    /* vitesse-initiale */
  double energie_initiale;       /* energie-initiale (Mahyco.axl:295, default 0) */
  /* Fictif EOS (eos/fictif/FictifEOS.axl): p = valeur-debut-pression +
   * valeur-dependance-temps t after tdebut-pression, 0 after tfin-pression */
  double fictif_tdebut, fictif_tfin, fictif_p0, fictif_dpdt;
} MhEnvironment;

typedef struct MhBoundaryCondition {
  int surface; /* MhSurface */
  int type;    /* MhBcType */
  double value;
} MhBoundaryCondition;

typedef struct MhOptions {
  /* module (Mahyco.axl) */
  double cfl;               /* cfl, default 0.3 */
  double threshold;         /* threshold, default 1e-10 (face groups) */
  double deltat_init;
  double deltat_min;
  double deltat_max;
  double final_time;
  int with_projection;      /* 1: MahycoLoop (Lagrange + ADI remap); 0: MahycoLagrangeLoop */
  int calcul_vitson_max;    /* default 1 */
  int nb_bc;
  MhBoundaryCondition bc[MH_MAX_BC];
  /* remap (RemapADI.axl) */
  int ordre_projection;       /* 1, 2 or 3 */
  int projection_limiteur_id; /* MhLimiter */
  double remap_threshold;     /* threshold, default 1e-10 */
  /* case model */
  int cas_model;            /* MhCaseModel */
  int cas_test;             /* reference cas-test id (TypesMahyco.h enum Test) */
  double interface_position;/* MH_CASE_RIEMANN: interface coordinate */
  int interface_axis;       /* MH_CASE_RIEMANN: 0,1,2 */
  double sedov_deposit;     /* SEDOV total energy deposit (reference constant 0.244816) */
  /* remap variants (RemapADI.axl:181-205; all default 0) */
  int projection_pente_borne;      /* projection-pente-borne: fluxes integrated over the cell (order 2) */
  int pente_borne_debar_fix;       /* projection-pente-borne-debar-fix: 0, 1 (mean value), 2 (cell value) */
  int pente_borne_mixte;           /* projection-pente-borne-mixte: pure-neighbourhood variant */
  int projection_limiteur_pure_id; /* projection-limiteur-pure-id (MhLimiter) */
  int projection_simple_pente;     /* projection-simple-pente */
  /* Lagrange energy update variants (Mahyco.axl:476-488; default 0) */
  int with_newton;        /* with-newton: Newton iterations on the energy equation */
  int pression_explicite; /* pression-explicite: explicit pressure work */
  /* advection mode and the RIDER case service */
  int sans_lagrange;         /* sans-lagrange: prescribed velocity, remap without the dual phase */
  /* options of the RIDER and ADVECTION case services (the same names in RIDER.axl and ADVECTION.axl) */
  int rider_reverse;         /* reverse-option: velocity x cos(pi t parametre) */
  double rider_parametre;    /* parametre */
  double rider_rayon_bulle;  /* rayon-bulle (reference default 0.15 for RIDER, 0.25 for ADVECTION) */
  int rider_densite_lineaire;/* densite-interieure-initiale-lineaire */
  /* schema-csts (Mahyco.axl:396-399): velocity taken back to n-1/2 before the Lagrange step and on to
     n+1 before the remap, energy from the corner-force work at n and n+1 (MahycoModule.cc:451-453,
     574-690, 1661-1697; PrepareRemap.cc:210, 236-240) */
  int schema_csts;
  int longueur_caracteristique;  /* MhCaracteristicLength (Mahyco.axl longueur-caracteristique) */
  /* conservation-energie-totale (RemapADI.axl:178): the kinetic energy the dual remap loses is put
     back into the cells' internal energy (RemapADIFinal.cc:318-345); not with sans_lagrange */
  int conservation_energie_totale;
} MhOptions;

/* Field identifiers for get/set.  Cell fields take an environment index
 * (-1 = cell value, k = value of environment k, 0 for non-member cells). */
enum MhField {
  MH_F_DENSITY = 0,
  MH_F_PRESSURE = 1,
  MH_F_INTERNAL_ENERGY = 2,
  MH_F_SOUND_SPEED = 3,
  MH_F_PSEUDO_VISCOSITY = 4,
  MH_F_CELL_MASS = 5,
  MH_F_CELL_VOLUME = 6,
  MH_F_FRACVOL = 7,
  MH_F_MASS_FRACTION = 8,
  MH_F_CARACTERISTIC_LENGTH = 9,
  MH_F_ENV_MASK = 10,       /* membership bitmask, as double */
  MH_F_MATERIAU = 11,
  /* node fields; component selects x/y/z */
  MH_N_VELOCITY = 100,
  MH_N_NODE_MASS = 101,
  MH_N_COORD = 102,
  /* node coordinates at the end of the last Lagrange phase: the positions the corner vectors
     m_cell_cqs of the next velocity update are formed at (MahycoModule.cc:438-562, 1097-1162);
     equal to MH_N_COORD without remap */
  MH_N_COORD_LAGRANGE = 103
};

#ifdef __cplusplus
}
#endif
#endif /* MAHYCO_TYPES_H */
