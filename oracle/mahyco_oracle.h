/*
 * mahyco_oracle.h -- C ABI of the CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY.  The oracle is a CPU restatement of MaHyCo's
 * Lagrange + ADI-remap cycle used as the parity checker for the HIP path and as
 * the CPU baseline timing.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  The product library never links it.
 */
#ifndef MAHYCO_ORACLE_H
#define MAHYCO_ORACLE_H

#include "../include/mahyco_types.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct OracleCase OracleCase;

/* Builds the mesh and runs the init entry points (HydroStartInit, ComputeCellMass,
 * ComputeNodeMass).  Returns NULL on an invalid description. */
OracleCase* oracle_create(const MhMesh* mesh, const MhEnvironment* envs, int nb_env,
                          const MhOptions* opt);
void oracle_destroy(OracleCase* h);

/* Runs up to max_cycles iterations of the time loop (MahycoLoop or MahycoLagrangeLoop).
 * Returns MH_OK, MH_STOPPED (final time reached) or a negative MhStatus. */
int oracle_run(OracleCase* h, int max_cycles);
/* The run continues as after a restart from a dump (hydroContinueInit, MahycoModule.cc:273-306). */
void oracle_continue(OracleCase* h);

double oracle_time(const OracleCase* h);
double oracle_deltat(const OracleCase* h);
int oracle_iteration(const OracleCase* h);
long oracle_nb_cells(const OracleCase* h);
long oracle_nb_nodes(const OracleCase* h);

/* Copies a field (MhField) to out.  Cell fields: env = -1 (cell value) or env index.
 * Node fields: comp = 0,1,2 for vectors.  ORACLE_N_DUAL_U (the oracle's own id) reads the
 * dual state the last remap left, comp 0-2 momentum, 3 mass, 4 kinetic energy. */
#define ORACLE_N_DUAL_U 200
int oracle_get_field(const OracleCase* h, int field, int env, int comp, double* out);

/* Writes one field of the state with mahyco_set_field's meaning (env = -1 cell value, env = k
 * environment k's value); MH_N_COORD_LAGRANGE re-forms the corner vectors of the next step. */
int oracle_set_field(OracleCase* h, int field, int env, int comp, const double* in);
/* m_global_time, m_global_deltat, m_global_old_deltat, m_global_iteration, m_sens_projection. */
void oracle_get_scalars(const OracleCase* h, double* t, double* dt, double* dt_old, int* it, int* sens);
void oracle_set_scalars(OracleCase* h, double t, double dt, double dt_old, int it, int sens);
/* IEquationOfState::applyEOS(env) / MahycoModule::computePressionMoyenne between iterations. */
void oracle_apply_eos(OracleCase* h, int env);
void oracle_compute_pression_moyenne(OracleCase* h);

/* Time history rows (EnvSummation, every 10 iterations): per row
 * [time, then for env 0..m-1: CellMass avg, Pressure avg, Density vol-avg,
 * InternalEnergy mass-avg, nb_envcells]. Returns number of rows; fills up to max_rows. */
int oracle_history(const OracleCase* h, double* out, int max_rows);

/* 0 (default): characteristic length with std::pow as the reference; 1: with the correctly
 * rounded power used by the HIP kernels (for bit-exact comparisons). */
void oracle_set_pow_mode(int mode);

/* RemapADIService::fluxLimiter (UtilesRemap.cc:108-120) as restated by the oracle. */
double oracle_flux_limiter(int id, double r);

/* Number of OpenMP threads used by the oracle loops (0 = runtime default). */
void oracle_set_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
