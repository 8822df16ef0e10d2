// mh_state.h -- device-resident state of one structured-mesh case and its HBM layout.
//
// Cells are numbered i + nx (j + ny k) and nodes i + (nx+1)(j + (ny+1) k), x fastest,
// like the Arcane Cartesian generator.  Every array is structure-of-arrays FP64 so that a
// wavefront of 64 consecutive cells (or nodes) reads 512 contiguous bytes per field.
//
// Material storage follows Arcane's split between global and partial values: the
// global arrays hold the value of every cell; a cell with exactly one environment has no
// other storage (its environment value is the global value).  The per-environment
// arrays env[k][cell] are meaningful only for cells with two or more environments;
// kernels read them only for such cells, so a mostly pure mesh streams the global
// arrays alone.
#ifndef MH_STATE_H
#define MH_STATE_H

#include <cstdint>

#include "../../include/mahyco_types.h"

#define MH_EST_MIXED 99  // MhDev::est of a cell with more than one positive volume fraction

struct MhGrid {
  int dim, nx, ny, nz;  // nz = 1 in 2D
  int nnx, nny, nnz;    // node counts (nnz = 1 in 2D)
  long nc, nn;
};

struct MhEnvParam {
  double gamma, pinf, a1, a2;
  int stiffened;
  int fictif;                  // Fictif EOS (advection tests): p(t), c = 1, e = 0
  double ft0, ft1, fp0, fdpdt; // its tdebut / tfin-pression, valeur-debut-pression, valeur-dependance-temps
};

// Time-loop scalars kept on the device so that cycles run without host round trips.
struct MhScalars {
  double t;       // m_global_time of the last completed iteration
  double dt;      // m_global_deltat used by the next iteration
  double dt_old;  // m_global_old_deltat seen by the next iteration (0 before the first)
  double dxmin;   // last CFL minimum (diagnostics)
  int iteration;
  int stopped;  // final time reached: later iterations are no-ops
  int status;   // 0 or a negative MhStatus: later iterations are no-ops
  int hist_rows;  // time-history rows written
  int hist_last;  // iteration of the last row
};

struct MhDev {
  MhGrid g;
  int m;   // environments
  int nv;  // remap slots: V_k, M_k, E_k per env and V*q
  MhEnvParam env[MH_MAX_ENV];
  double cfl, thr_remap, dtmin, dtmax, tfinal;
  int order2, limiter, with_proj, vitson_max;
  int nbc;
  int bc_surf[MH_MAX_BC];
  int bc_type[MH_MAX_BC];
  double bc_val[MH_MAX_BC];
  // Eulerian grid: node coordinate along each axis (0 bytes per node to rebuild x0)
  const double* ax;
  const double* ay;
  const double* az;
  const double* VE;  // Euler volume per cell (m_euler_volume)
  // node state
  double* u[3];    // velocity
  double* xlp[3];  // Lagrangian coordinates of the previous cycle (CQS_n, MahycoModule.cc:1097)
  double* xlc[3];  // Lagrangian coordinates of this cycle
  double* mn;      // node mass used this cycle
  // cell state
  double *rho, *p, *c, *q, *M, *e, *fg;
  uint8_t* mask;                              // environment membership bits
  double *fk, *chik, *rhok, *ek, *pk, *Mk;  // [m][nc], mixed cells only
  double* chig;                               // global mass fraction (single-environment cells)
  double *L, *Vl;                             // characteristic length, Lagrangian volume
  double* partial;                            // per-block CFL minima
  int* err;                                   // error word raised inside a cycle
  MhScalars* sc;
  // remap buffers: U ping-pong [nv][nc], mass fluxes [m][nc], dual state [4][nn] (in
  // place), dual phi ping-pong [3][nn]
  double *Ua, *Ub, *dflux, *Ud, *Pa, *Pb;
  // conservation-energie-totale: the dual kinetic energy [nn] (in place), its phi ping-pong
  // [nn] each, and the energy each cell receives back [nc]; null otherwise
  double *ke_U, *ke_Pa, *ke_Pb, *dec;
  // cell widths per direction [dim][nc] (order 3, pente-borne, generalised limiters), and
  // the mixed-cell flag remapVariables leaves per cell (order 3, pente-borne-mixte):
  // MH_EST_MIXED, else the index of the one environment with a positive volume fraction,
  // -1 for none.  Null when unused.
  double* hc;
  int8_t* est;
  // pente-borne remap: phi of the first sweep [3m+1][nc] (the environments' own values for
  // members, the last sweep's values otherwise)
  double* phi0;
  int pb;            // projection-pente-borne
  int debar;         // projection-pente-borne-debar-fix (0, 1, 2)
  int pb_mixte;      // projection-pente-borne-mixte
  int lim_pure;      // projection-limiteur-pure-id
  int simple_pente;  // projection-simple-pente
  int nsweep;        // directions with more than one cell
  int newton;        // with-newton: energy of the Lagrange step by Newton iterations
  int p_explicit;    // pression-explicite
  double thr_module; // module threshold (Newton tolerance)
  int q_frozen;      // run continued from a dump with m >= 2: V q keeps its Lagrangian value
                     // through the remap (hydroContinueInit, MahycoModule.cc:288)
  // advection mode (sans-lagrange): prescribed node velocity u0 [3][nn] scaled by
  // rev_cos[iteration] = cos(pi t parametre) with the reverse option, constant step
  int sans_lagrange;
  int rider_reverse;
  // schema-csts: the velocity saved at the start of the cycle [3][nn]; per cell, the
  // pseudo-viscosity of the previous cycle (qm1), the remap's U_q / V of the last and the
  // previous remap (qe, qem1: the partial pseudo-viscosity of every environment of a mixed
  // cell) and the environment mask before the last remap (mkm1)
  int lc;  // MhCaracteristicLength
  int csts;
  double* un[3];
  double *qm1, *qe, *qem1;
  uint8_t* mkm1;
  double *qkn, *qknm1;  // without remap: each environment's pseudo-viscosity at n and n-1 [m][nc]
  double dt_init;
  const double* u0[3];
  const double* rev_cos;
  long rev_n;
  int nblk_c, nblk_n;  // blocks of MH_BLOCK threads over cells / nodes
  int max_iter;        // > 0: stop after this many iterations (fixed-horizon runs)
  int hist_period;     // > 0: EnvSummation rows every hist_period iterations
  double* hist;        // rows of 1 + 5m doubles (time, then per environment: mean M_k,
                       // mean p_k, volume-weighted rho_k, mass-weighted e_k, member count)
  double* hist_part;   // per-block partial sums, 6m doubles per block
  int hist_max;        // capacity in rows
  int nt_kc;           // node planes per block of the z-marching node kernel (autotuned)
  int ct_k;            // cell planes per block of the LDS-staged cell kernel (autotuned)
  int staged;          // remap loop with the Lagrangian state written before the remap (mahyco_set_staged)
  int nm_len;          // target nodes per segment of the marching y / z dual sweeps
};

#endif
