/*
 * mahyco.h -- C ABI of the MI355X (gfx950) MaHyCo hydro path.
 *
 * One handle holds one structured-mesh case resident in HBM; every call is
 * ordered on the handle's own HIP stream.  Plain pointers and sizes only, so the
 * header binds from C, ctypes, cgo or JNI.  Status codes are MhStatus values
 * (mahyco_types.h): the reference's exit(1) sites become negative returns.
 *
 * Each declaration names the reference interface it replaces (paths relative to
 * the MaHyCo tree).
 */
#ifndef MAHYCO_H
#define MAHYCO_H

#include <stddef.h>
#include "mahyco_types.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct MhHandle MhHandle;

/* Init entry points HydroStartInit + ComputeCellMass + ComputeNodeMass
 * (MahycoModule.cc:21-262, MahycoAnnexe.cc:11-235, casModel()->initMat/initVar,
 * IEquationOfState::initEOS): builds the mesh, the initial state and uploads it. */
int mahyco_create(const MhMesh* mesh, const MhEnvironment* envs, int nb_env,
                  const MhOptions* options, MhHandle** out);
void mahyco_destroy(MhHandle* h);
/* Same, with every call of the handle ordered on the caller's HIP stream (a hipStream_t of the
 * current device, default priority, passed as void*; the caller keeps it alive until
 * mahyco_destroy).  Replaces running the module inside the caller's own execution context. */
int mahyco_create_on_stream(const MhMesh* mesh, const MhEnvironment* envs, int nb_env,
                            const MhOptions* options, void* stream, MhHandle** out);

/* Runs ncycles iterations of the time loop (Mahyco.config:23-36 MahycoLoop, or
 * :57-69 MahycoLagrangeLoop when with_projection = 0), asynchronously on the handle's
 * stream.  Iterations after the final-time stop or an error are no-ops on the device.
 * Returns MH_OK once queued; errors surface in mahyco_status(). */
int mahyco_cycle(MhHandle* h, int ncycles);

/* Same as mahyco_cycle but replays a captured hipGraph of `cycles_per_graph`
 * iterations (captured once per handle and graph size). */
int mahyco_cycle_graph(MhHandle* h, int ncycles, int cycles_per_graph);

/* Waits for the handle's stream; returns the device status word: MH_OK, MH_STOPPED
 * (final time reached, MahycoModule.cc:1841-1857) or a negative MhStatus. */
int mahyco_synchronize(MhHandle* h);
int mahyco_status(MhHandle* h);

/* Time-loop scalars (Arcane m_global_time, m_global_deltat, m_global_iteration). */
double mahyco_time(MhHandle* h);
double mahyco_deltat(MhHandle* h);
int mahyco_iteration(MhHandle* h);
long mahyco_nb_cells(MhHandle* h);
long mahyco_nb_nodes(MhHandle* h);
/* Number of hipGraphs mahyco_cycle_graph has captured and instantiated on this handle
   (no counterpart in the reference: lets a timing loop show that no capture fell inside it). */
long mahyco_graph_captures(MhHandle* h);

/* Copies one field (MhField) to host memory `out` (nb_cells or nb_nodes doubles).
 * Cell fields: env = -1 (cell value) or environment index (0 outside the env).
 * Node fields: comp = 0,1,2.  Replaces reading the Arcane variables by name. */
int mahyco_get_field(MhHandle* h, int field, int env, int comp, double* out);

/* Writes one field of the state from host memory (nb_cells or nb_nodes doubles); replaces a
 * module or case service writing the named Arcane variables that the module and its services
 * share (casTest/IInitialisations.h:26-56 initMat/initVar, MahycoModule.cc:110-140).
 * Cell fields: env = -1 writes the cell value; env = k writes environment k's value (its
 * partial value in a multi-environment cell, the cell value in a cell of k alone, nothing
 * outside k), so MH_F_ENV_MASK goes first.  Settable: density, pressure, internal energy,
 * cell mass, volume fraction, mass fraction (cell and environment values), sound speed and
 * pseudo-viscosity (cell values: a mixed cell's environment values are re-formed by the
 * kernels), the environment mask, node velocity and MH_N_COORD_LAGRANGE (MH_N_COORD without
 * remap).  Volume, characteristic length, Materiau and node mass are derived: MH_ERR_INVALID.
 * Between iterations, and for cell fields also between mahyco_update_energy_and_pressure and
 * mahyco_remap in staged mode (mahyco_set_staged); MH_ERR_INVALID inside an entry-point
 * sequence otherwise. */
int mahyco_set_field(MhHandle* h, int field, int env, int comp, const double* in);
/* The scalars that go with the state: m_global_time, m_global_deltat, m_global_old_deltat,
 * m_global_iteration and m_sens_projection (Mahyco.axl:11, RemapADIService.cc:23-25, 59-60).
 * Clears the stop flag and the status word; the next iteration is number iteration + 1. */
int mahyco_set_scalars(MhHandle* h, double time, double deltat, double old_deltat, int iteration,
                       int sens_projection);

/* ---- entry points of MahycoModule (Mahyco.axl:244-261), one call each ----------
 * They run the same kernels as mahyco_cycle, grouped as the kernels fuse them:
 *   mahyco_save_values_at_n        SaveValuesAtN        :311-382  (time t += dt)
 *   mahyco_update_velocity         UpdateVelocity       :438-562  + ApplyBoundaryCondition
 *                                  :756-807 + UpdatePosition (nodes) :730-751
 *   mahyco_compute_geometric_values  ComputeGeometricValues :1097-1253 + UpdateDensity
 *                                  :1267-1316 + ComputeArtificialViscosity :386-426 +
 *                                  UpdateEnergyAndPressure :1430-1480 (+ remap vector build,
 *                                  PrepareRemap.cc:94-193)
 *   mahyco_remap                   Remap  PrepareRemap.cc:203-251 (IRemap::appliRemap
 *                                  RemapADIService.cc:15-90, remapVariables
 *                                  RemapADIFinal.cc:21-346, applyEOS, computePressionMoyenne)
 *   mahyco_compute_delta_t         ComputeDeltaT :1739-1864
 * The calls must come in time-loop order; an out-of-order call returns MH_ERR_INVALID.
 * ApplyBoundaryCondition, UpdatePosition, UpdateDensity, ComputeArtificialViscosity,
 * UpdateElasticityAndPlasticity (NoModel) and UpdateEnergyAndPressure are folded
 * into the call that precedes them; their functions below only check the order. */
int mahyco_save_values_at_n(MhHandle* h);
int mahyco_update_velocity(MhHandle* h);
int mahyco_apply_boundary_condition(MhHandle* h);
int mahyco_update_position(MhHandle* h);
int mahyco_compute_geometric_values(MhHandle* h);
int mahyco_update_density(MhHandle* h);
int mahyco_compute_artificial_viscosity(MhHandle* h);
int mahyco_update_elasticity_and_plasticity(MhHandle* h);
int mahyco_update_energy_and_pressure(MhHandle* h);
int mahyco_remap(MhHandle* h);
int mahyco_compute_delta_t(MhHandle* h);

/* ComputeDeltaT on caller-owned device arrays, allocating nothing.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP graph capture,
 * serving loops).  L and c are the characteristic length and sound speed per cell, u0/u1/u2 the
 * node velocity (u2 unused in 2D); the step, with the rules of MahycoModule.cc:1739-1833 (cfl,
 * deltat_min and deltat_max of options, growth capped at 1.05 old_deltat), and the status
 * (MH_OK or MH_ERR_DT_MIN) go to the device words dt_out and status_out, ordered on `stream`
 * (a hipStream_t as void*, null for the null stream).  The per-block minima live in a static
 * device array of 65536 blocks: meshes over 16,777,216 cells are refused with MH_ERR_INVALID.
 * Calls that overlap on one GPU share that array and race. */
int mahyco_compute_delta_t_noalloc(const MhMesh* mesh, const MhOptions* options, double old_deltat,
                                   const double* L, const double* c, const double* u0, const double* u1,
                                   const double* u2, double* dt_out, int* status_out, void* stream);

/* IEquationOfState::applyEOS(env) (IEquationOfState.h:45-49; PerfectGasEOSService.cc:45-67,
 * StiffenedGasEOSService.cc:43-63): pressure and sound speed of environment env from its
 * density and internal energy.  Also run inside mahyco_remap / the Lagrange-only cycle. */
int mahyco_apply_eos(MhHandle* h, int env);
/* IEquationOfState::applyOneCellEOS(env, EnvCell) (IEquationOfState.h:50-52,
 * PerfectGasEOSService.cc:71-90, StiffenedGasEOSService.cc:67-83, FictifEOSService.cc:80): the same for one cell of the
 * environment (cell index i + nx (j + ny k)); nothing for a cell outside it. */
int mahyco_apply_one_cell_eos(MhHandle* h, int env, long cell);
/* MahycoModule::computePressionMoyenne (MahycoModule.cc:1708-1735): cell pressure and sound
 * speed of multi-environment cells.  Also run inside mahyco_remap / the Lagrange-only cycle. */
int mahyco_compute_pression_moyenne(MhHandle* h);

/* TimeHistory module with the EnvSummation service (time_history/EnvSummationService.cc:77-169):
 * every `period` iterations the device appends a row [time, then per environment: mean CellMass,
 * mean Pressure, volume-weighted Density, mass-weighted InternalEnergy, member count] without
 * host synchronisation.  Enable before the first iteration; read returns the number of rows
 * written and copies up to max_rows rows. */
int mahyco_history_enable(MhHandle* h, int period, int max_rows);
int mahyco_history_read(MhHandle* h, double* rows, int max_rows);

/* Post-processing output of the case's <output> variables (Donnees.arc <output> block,
 * written by Arcane's EnSight Gold post-processor into output/depouillement and compared
 * byte for byte by NONREGRESSION/launch_test.sh:74-83): appends the current state as the
 * next EnSight Gold ASCII step under dir (ensight.case, ensight.geo/ensight.geoNNNNNN,
 * <Var>/<Var>NNNNNN).  variables are reference names (CellMass, Pressure, Density, Velocity,
 * NodeMass, InternalEnergy, PseudoViscosity, Materiau, SoundSpeed, FracVol, CellVolume,
 * CaracteristicLength); every step of a case uses the same list. */
int mahyco_ensight_write(MhHandle* h, const char* dir, const char* const* variables, int nb_variables);
/* The EnvSummation time history as Arcane writes it into output/courbes/gnuplot
 * (compared by NONREGRESSION/launch_test.sh:85-91): files <Var>_<environment name> for
 * CellMass, Pressure, Density and InternalEnergy, one "%.16E %.16E" line per row.
 * env_names holds one name per environment. */
int mahyco_history_write_gnuplot(MhHandle* h, const char* dir, const char* const* env_names);

/* Protection / reprise (Arcane checkpoint service of the <arcane-checkpoint> block, restart
 * entry point MahycoModule::hydroContinueInit, MahycoModule.cc:273-307): write saves the
 * whole device state of the case between two iterations; read restores it into a case
 * created from the same description (mesh, options, environments, and the same
 * mahyco_history_enable call), after which the run continues bit for bit as if it had not
 * stopped.  A file of another case is refused with MH_ERR_INVALID and the case left as is. */
int mahyco_checkpoint_write(MhHandle* h, const char* path);
int mahyco_checkpoint_read(MhHandle* h, const char* path);

/* Restart entry point MahycoModule::hydroContinueInit (MahycoModule.cc:273-306), called after
 * mahyco_checkpoint_read to continue as the reference does: the continued run projects only
 * 3m+5 remap slots (:288), so with two or more environments the pseudo-viscosity slot 7m is
 * not swept and V q keeps its Lagrangian value through every later remap. */
int mahyco_continue_init(MhHandle* h);

/* ---- ensembles (no reference counterpart: many independent cases on one GPU) ---------
 * Cases of any size advance together, each with its own time step and stop rule: one
 * launch per kernel phase covers every active case of a (dimension, environment count)
 * group.  Admission is by HBM footprint: mahyco_ensemble_add returns MH_ERR_NOMEM when the
 * case would exceed the budget or no slot is free.  An added case becomes active at the next
 * iteration multiple of 6 of the ensemble clock (so all active cases share the sweep
 * rotation).  max_iterations > 0 stops a case after that many iterations.
 * mahyco_ensemble_poll fills per-slot status (MH_OK running, MH_STOPPED done, 2 pending,
 * -100 empty, or a negative error), time and iteration; returns the number of active slots. */
typedef struct MhEnsemble MhEnsemble;
size_t mahyco_case_footprint(const MhMesh* mesh, int nb_env, const MhOptions* options);
int mahyco_ensemble_create(int max_slots, double hbm_budget_bytes, MhEnsemble** out);
void mahyco_ensemble_destroy(MhEnsemble* e);
int mahyco_ensemble_add(MhEnsemble* e, const MhMesh* mesh, const MhEnvironment* envs, int nb_env,
                        const MhOptions* options, int max_iterations, int* slot_out);
int mahyco_ensemble_remove(MhEnsemble* e, int slot);
int mahyco_ensemble_step(MhEnsemble* e, int ncycles, int use_graph);
int mahyco_ensemble_synchronize(MhEnsemble* e);
int mahyco_ensemble_poll(MhEnsemble* e, int* status, double* time, int* iteration);
/* The same in two halves, so that the host can queue more iterations in between: _poll_async
 * enqueues a copy of every slot's scalars behind the queued work (pinned host memory, no wait);
 * _poll_result waits for that copy only and fills the arrays as mahyco_ensemble_poll does. */
int mahyco_ensemble_poll_async(MhEnsemble* e);
int mahyco_ensemble_poll_result(MhEnsemble* e, int* status, double* time, int* iteration);
int mahyco_ensemble_get_field(MhEnsemble* e, int slot, int field, int env, int comp, double* out);
double mahyco_ensemble_used_bytes(MhEnsemble* e);
/* Graphs the ensemble has captured (a case joining or leaving an existing batch group changes
 * only device-side work lists, so it captures nothing). */
long mahyco_ensemble_graph_captures(MhEnsemble* e);
/* Autotuner of the ensemble's launch shape (no reference counterpart): times `cycles`
 * iterations of the ensemble's own cases (hipGraph replays) with each candidate persistent grid
 * (2, 4, 8, 16 blocks per compute unit) and keeps the fastest for every batch group, and
 * for later groups of the same kind in this process.  The iterations count as the cases' own
 * (results do not depend on the launch shape).  Returns the choice and its time per iteration. */
int mahyco_ensemble_autotune(MhEnsemble* e, int cycles, int* blocks_per_cu_out, double* ms_per_cycle_out);

/* Autotuner of a single case's kernel variants and tile shapes (no reference counterpart):
 * times `cycles` iterations of the case itself per candidate, one coordinate at a time: the
 * sweeps (row segments and LDS tiles, or one thread per item), then in 3D the node kernel
 * (flat, or z-marching with 8, 12 or 16 node planes per block) and the cell kernel (flat, or
 * LDS-staged with 8, 12 or 16 cell planes per block), then the segment length of the marching
 * y / z dual sweeps, and keeps the fastest, also for later cases of the same dimensions and
 * options in this process.  variant = node + 4 cell + 16 sweep + 32 march with those codes.  The iterations count as the case's own (every variant gives the same
 * bits); tuning stops early at the final time or an error. */
int mahyco_autotune(MhHandle* h, int cycles, int* variant_out, double* ms_per_cycle_out);

/* The kernel-variant code of a case (the code mahyco_autotune returns: node 0 flat, 1-3 the
 * z-marching node kernel with 12/8/16 planes; + 4 x cell 0 flat, 1-3 the LDS-staged cell kernel
 * with 12/8/16 planes; + 16 for one thread per item in every sweep; + 32 x march 0-4 for
 * segments of 10/6/8/4/12 nodes in the marching y / z dual sweeps) and its application, so that a
 * tuned choice can be kept and applied to later cases.  Every variant gives the same bits.  No
 * counterpart in the reference, which has one kernel per loop. */
int mahyco_get_variant(const MhHandle* h);
int mahyco_set_variant(MhHandle* h, int variant);

/* Staged Lagrange step (on = 1; 0 returns to the fused kernels).  The reference's
 * UpdateEnergyAndPressure stores the Lagrangian state (MahycoModule.cc:1430-1480, with applyEOS
 * and computePressionMoyenne after it) and its Remap forms the remap vector from that state
 * (computeVariablesForRemap, PrepareRemap.cc:100-155), so a module entry point placed between
 * the two in the time loop can read and change the state.  By default the cell kernel forms
 * the remap vector itself and stores no Lagrangian state.  In staged mode the cell kernel
 * stores density, internal energy, pressure, sound speed and pseudo-viscosity (per environment
 * and per cell) and mahyco_remap forms the remap vector from them: between
 * mahyco_update_energy_and_pressure and mahyco_remap, mahyco_get_field reads that state and
 * mahyco_set_field may change cell values (node values only between iterations).  The remap
 * reads density, internal energy, volume fraction, pseudo-viscosity and the environment mask;
 * pressure and sound speed are formed again after it.  Same bits as the fused kernels when the
 * state is not changed.  Between iterations only; MH_ERR_UNSUPPORTED without remap, with
 * schema-csts, in advection mode or with a Fictif environment, MH_ERR_INVALID for a case of an
 * ensemble. */
int mahyco_set_staged(MhHandle* h, int on);

/* Kernel timing: when enabled, every launch is bracketed by HIP events on the
 * handle's stream.  mahyco_profile_read fills up to `max` entries of kernel
 * names (NUL-separated into `names`, `names_len` bytes), total milliseconds and
 * launch counts; returns the number of kernel kinds. */
int mahyco_profile_enable(MhHandle* h, int on);
int mahyco_profile_read(MhHandle* h, char* names, int names_len, double* total_ms, int* counts, int max);

#ifdef __cplusplus
}
#endif
#endif /* MAHYCO_H */
