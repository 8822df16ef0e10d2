// mh_kernels.h -- gfx950 kernels of the Lagrange + ADI remap cycle (FP64 on the vector ALUs).
//
// Per cycle (Mahyco.config:23-36):
//   k_node_lag      SaveValuesAtN (nodes) + UpdateVelocity + ApplyBoundaryCondition +
//                   UpdatePosition (nodes): node-centric force gather in cell order
//   k_cell_lag      UpdatePosition (cells) + ComputeGeometricValues + UpdateDensity +
//                   ComputeArtificialViscosity + UpdateEnergyAndPressure, then the remap
//                   vector U (PrepareRemap.cc:94-193) or, without remap, EOS + mean pressure
//                   and the CFL candidate
//   k_sweep_cell    one ADI direction, cell part (RemapADIService.cc:119-575)
//   k_sweep_node    the same direction, dual part (DualPhaseRemap.cc:38-258)
//   k_finalize      remapVariables (RemapADIFinal.cc:21-316) + applyEOS + computePressionMoyenne
//                   + the CFL candidate of computeDeltaT
//   k_dt            computeDeltaT reduction and time-step rules (MahycoModule.cc:1739-1864)
//
// Every value is produced by one thread with the reference's operation order, so the
// results do not depend on launch shape or scheduling.
#ifndef MH_KERNELS_H
#define MH_KERNELS_H

#include <hip/hip_runtime.h>

#include "mh_geom.h"
#include "mh_state.h"

#define MH_BLOCK 256
#define MH_DT_BLOCK 1024

__device__ __forceinline__ bool mh_halted(const MhScalars* sc) { return sc->status != 0 || sc->stopped != 0; }

// Blocks are dispatched round-robin over the 8 XCDs, each with its own L2.  Single-mesh
// grids are a multiple of 8 blocks and block b works on block (b % 8) * (grid / 8) + b / 8,
// so each XCD sweeps one contiguous slab and the neighbours of a stencil stay in its L2.
__device__ __forceinline__ int xcd_block(int b, int nb) { return (b & 7) * (nb >> 3) + (b >> 3); }

// Arcane math::min / math::max
__device__ __forceinline__ double amin(double a, double b) { return (a < b) ? a : b; }
__device__ __forceinline__ double amax(double a, double b) { return (a < b) ? b : a; }

__device__ __forceinline__ D3 load3(double* const* a, long n) { return d3(a[0][n], a[1][n], a[2][n]); }


// The node coordinates of the cell whose node 0 is b, in Arcane's local node order.
template <int DIM>
__device__ __forceinline__ void cell_nodes(const MhDev& S, long b, double* const* xs, D3* x) {
  const long nnx = S.g.nnx, nnxy = (long)S.g.nnx * S.g.nny;
  if (DIM == 3) {
    x[0] = load3(xs, b);
    x[1] = load3(xs, b + 1);
    x[2] = load3(xs, b + 1 + nnx);
    x[3] = load3(xs, b + nnx);
    x[4] = load3(xs, b + nnxy);
    x[5] = load3(xs, b + 1 + nnxy);
    x[6] = load3(xs, b + 1 + nnx + nnxy);
    x[7] = load3(xs, b + nnx + nnxy);
  } else {
    x[0] = d3(xs[0][b], xs[1][b], 0.);
    x[1] = d3(xs[0][b + 1], xs[1][b + 1], 0.);
    x[2] = d3(xs[0][b + 1 + nnx], xs[1][b + 1 + nnx], 0.);
    x[3] = d3(xs[0][b + nnx], xs[1][b + nnx], 0.);
  }
}

// Block-wide minimum (Arcane math::min semantics on non-NaN input), written by thread 0.
// CFL minimum with the reference's sequential semantics.  computeDeltaT folds
// min = math::min(min, dx) over the cells in order (MahycoModule.cc:1777), and math::min(a,
// b) = a < b ? a : b turns NaN at a NaN candidate (an emptied cell's sound speed) and
// restarts at the next cell: the result is the minimum over the cells after the last NaN
// candidate, NaN if the last cell is one.  A run of cells therefore reduces to (nan: a NaN
// was seen, m: minimum after the last NaN, NaN if there is no value there), combined in
// cell order; (NaN, no NaN seen) is the empty run.  The reference's start value (the
// largest double) is applied once, at the end.
struct MhCflMin {
  double m;
  int nan;
};
__device__ __forceinline__ double min_skip_nan(double a, double b) { return (b < a || a != a) ? b : a; }
__device__ __forceinline__ MhCflMin cfl_leaf(double x, bool active) {
  MhCflMin r;
  r.nan = active && x != x;
  r.m = active ? x : __longlong_as_double(0x7ff8000000000000LL);  // inactive: the empty run
  return r;
}
__device__ __forceinline__ MhCflMin cfl_combine(MhCflMin l, MhCflMin r) {
  if (r.nan) return r;
  MhCflMin o;
  o.nan = l.nan;
  o.m = min_skip_nan(l.m, r.m);
  return o;
}
// in-order reduction over a wavefront (lane 0 ends with lanes 0..63 combined); without a
// NaN in the wavefront it is a plain minimum, in any order
__device__ __forceinline__ MhCflMin cfl_wave(MhCflMin v) {
  const int lane = threadIdx.x & 63;
  if (__ballot(v.nan) == 0) {
    for (int off = 32; off > 0; off >>= 1) v.m = min_skip_nan(v.m, __shfl_xor(v.m, off, 64));
    return v;
  }
  for (int off = 1; off < 64; off <<= 1) {
    MhCflMin o;
    o.m = __shfl_down(v.m, off, 64);
    o.nan = __shfl_down(v.nan, off, 64);
    if (lane + off < 64) v = cfl_combine(v, o);
  }
  return v;
}

// Per-block CFL partial of the cells of block lb, in cell order: out[2 lb] = minimum after
// the last NaN, out[2 lb + 1] = 1 if a NaN was seen.
__device__ __forceinline__ void block_min_store(double v, bool active, double* out, int lb) {
  __shared__ double rm[MH_BLOCK / 64];
  __shared__ int rn[MH_BLOCK / 64];
  MhCflMin r = cfl_wave(cfl_leaf(v, active));
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) {
    rm[wid] = r.m;
    rn[wid] = r.nan;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    MhCflMin t;
    t.m = rm[0];
    t.nan = rn[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) {
      MhCflMin o;
      o.m = rm[w];
      o.nan = rn[w];
      t = cfl_combine(t, o);
    }
    out[2 * lb] = t.m;
    out[2 * lb + 1] = t.nan ? 1. : 0.;
  }
}

// ----------------------------------------------------------------------------- nodes
// Corner L of cell (ii,jj,kk) seen from node (i,j,k); accumulates mass and force in the
// order of a cell-order scatter (MahycoModule.cc:247-262 and :455-462).
// FWD: the corner vectors at this cycle's Lagrangian positions (updateVelocityForward,
// :638-690) instead of the previous cycle's.
template <int DIM, int DI, int DJ, int DK, int FWD = 0>
__device__ __forceinline__ void node_gather(const MhDev& S, int i, int j, int k, double& mass,
                                            double& fx, double& fy, double& fz) {
  double* const* X = FWD ? S.xlc : S.xlp;
  const int ii = i - DI, jj = j - DJ, kk = k - DK;
  if (ii < 0 || ii >= S.g.nx || jj < 0 || jj >= S.g.ny) return;
  if (DIM == 3 && (kk < 0 || kk >= S.g.nz)) return;
  const long c = ii + (long)S.g.nx * (jj + (long)S.g.ny * kk);
  const double w = DIM == 3 ? 0.125 : 0.25;
  mass += w * S.M[c];
  const double pr = S.p[c] + S.q[c];
  const long nnx = S.g.nnx, nnxy = (long)S.g.nnx * S.g.nny;
  const long b = ii + nnx * (jj + (long)S.g.nny * kk);
  constexpr int L = 4 * DK + (DJ ? (DI ? 2 : 3) : (DI ? 1 : 0));
  D3 q;
  if (DIM == 3) {
    D3 x[8];
    x[0] = load3(X, b);
    x[1] = load3(X, b + 1);
    x[2] = load3(X, b + 1 + nnx);
    x[3] = load3(X, b + nnx);
    x[4] = load3(X, b + nnxy);
    x[5] = load3(X, b + 1 + nnxy);
    x[6] = load3(X, b + 1 + nnx + nnxy);
    x[7] = load3(X, b + nnx + nnxy);
    q = cqs_hex_corner<L>(x);
  } else {
    D3 x[4];
    x[0] = d3(X[0][b], X[1][b], 0.);
    x[1] = d3(X[0][b + 1], X[1][b + 1], 0.);
    x[2] = d3(X[0][b + 1 + nnx], X[1][b + 1 + nnx], 0.);
    x[3] = d3(X[0][b + nnx], X[1][b + nnx], 0.);
    q = cqs_quad_corner<L & 3>(x);
  }
  fx += pr * q.x;
  fy += pr * q.y;
  if (DIM == 3) fz += pr * q.z;
}

// UpdateVelocity tail (u += dt_half / m F), ApplyBoundaryCondition (Vx/Vy/Vz on the
// XMIN..ZMAX node groups, in option order) and UpdatePosition (x += dt u) for node n.
template <int DIM>
__device__ __forceinline__ void node_finish(const MhDev& S, long n, int i, int j, int k, double dt,
                                            double dth, double mass, double fx, double fy, double fz) {
  const double s = dth / mass;
  double ux = S.u[0][n], uy = S.u[1][n], uz = DIM == 3 ? S.u[2][n] : 0.;
  if (S.csts) {  // SaveValuesAtN of the velocity
    S.un[0][n] = ux;
    S.un[1][n] = uy;
    if (DIM == 3) S.un[2][n] = uz;
  }
  if (S.csts && S.with_proj) {  // updateVelocityBackward (:451-453, 574-627)
    const double sb = (-0.5 * S.sc->dt_old) / mass;
    ux = ux + sb * fx;
    uy = uy + sb * fy;
    if (DIM == 3) uz = uz + sb * fz;
  }
  ux = ux + s * fx;
  uy = uy + s * fy;
  if (DIM == 3) uz = uz + s * fz;
  for (int b = 0; b < S.nbc; ++b) {
    const int sf = S.bc_surf[b];
    bool on;
    switch (sf) {
      case MH_XMIN: on = (i == 0); break;
      case MH_XMAX: on = (i == S.g.nx); break;
      case MH_YMIN: on = (j == 0); break;
      case MH_YMAX: on = (j == S.g.ny); break;
      case MH_ZMIN: on = DIM == 3 ? (k == 0) : true; break;
      default: on = DIM == 3 ? (k == S.g.nz) : true; break;
    }
    if (on) {
      const int t = S.bc_type[b];
      if (t == MH_BC_VX) ux = S.bc_val[b];
      else if (t == MH_BC_VY) uy = S.bc_val[b];
      else if (DIM == 3) uz = S.bc_val[b];
    }
  }
  double x0, y0, z0;
  if (S.with_proj) {
    x0 = S.ax[i];
    y0 = S.ay[j];
    z0 = DIM == 3 ? S.az[k] : 0.;
  } else {
    x0 = S.xlp[0][n];
    y0 = S.xlp[1][n];
    z0 = DIM == 3 ? S.xlp[2][n] : 0.;
  }
  S.u[0][n] = ux;
  S.u[1][n] = uy;
  S.xlc[0][n] = x0 + dt * ux;
  S.xlc[1][n] = y0 + dt * uy;
  if (DIM == 3) {
    S.u[2][n] = uz;
    S.xlc[2][n] = z0 + dt * uz;
  }
  S.mn[n] = mass;
}


// SaveValuesAtN + UpdateVelocity (:438-562) + ApplyBoundaryCondition (:756-807) +
// UpdatePosition (:730-751) for one node.
template <int DIM>
__device__ __forceinline__ void k_node_lag_body(const MhDev& S, int lb) {
  const long n = lb * (long)blockDim.x + threadIdx.x;
  if (n >= S.g.nn) return;
  if (mh_halted(S.sc)) return;
  const int nnx = S.g.nnx, nny = S.g.nny;
  const int i = (int)(n % nnx);
  const long r = n / nnx;
  const int j = (int)(r % nny);
  const int k = (int)(r / nny);
  const double dt = S.sc->dt;
  const double dth = 0.5 * (S.sc->dt_old + dt);
  double mass = 0., fx = 0., fy = 0., fz = 0.;
  if (DIM == 3) {
    node_gather<3, 1, 1, 1>(S, i, j, k, mass, fx, fy, fz);
    node_gather<3, 0, 1, 1>(S, i, j, k, mass, fx, fy, fz);
    node_gather<3, 1, 0, 1>(S, i, j, k, mass, fx, fy, fz);
    node_gather<3, 0, 0, 1>(S, i, j, k, mass, fx, fy, fz);
  }
  node_gather<DIM, 1, 1, 0>(S, i, j, k, mass, fx, fy, fz);
  node_gather<DIM, 0, 1, 0>(S, i, j, k, mass, fx, fy, fz);
  node_gather<DIM, 1, 0, 0>(S, i, j, k, mass, fx, fy, fz);
  node_gather<DIM, 0, 0, 0>(S, i, j, k, mass, fx, fy, fz);
  node_finish<DIM>(S, n, i, j, k, dt, dth, mass, fx, fy, fz);
}


// updateVelocityForward (:638-690), before a CSTS remap: u += (dt/2) / m F with the force of
// this cycle's pressure array, pseudo-viscosity and corner vectors (no boundary condition).
template <int DIM>
__device__ __forceinline__ void k_node_fwd_body(const MhDev& S, int lb) {
  const long n = lb * (long)blockDim.x + threadIdx.x;
  if (n >= S.g.nn) return;
  if (mh_halted(S.sc)) return;
  const int nnx = S.g.nnx, nny = S.g.nny;
  const int i = (int)(n % nnx);
  const long r = n / nnx;
  const int j = (int)(r % nny);
  const int k = (int)(r / nny);
  double mass = 0., fx = 0., fy = 0., fz = 0.;
  if (DIM == 3) {
    node_gather<3, 1, 1, 1, 1>(S, i, j, k, mass, fx, fy, fz);
    node_gather<3, 0, 1, 1, 1>(S, i, j, k, mass, fx, fy, fz);
    node_gather<3, 1, 0, 1, 1>(S, i, j, k, mass, fx, fy, fz);
    node_gather<3, 0, 0, 1, 1>(S, i, j, k, mass, fx, fy, fz);
  }
  node_gather<DIM, 1, 1, 0, 1>(S, i, j, k, mass, fx, fy, fz);
  node_gather<DIM, 0, 1, 0, 1>(S, i, j, k, mass, fx, fy, fz);
  node_gather<DIM, 1, 0, 0, 1>(S, i, j, k, mass, fx, fy, fz);
  node_gather<DIM, 0, 0, 0, 1>(S, i, j, k, mass, fx, fy, fz);
  const double s = (0.5 * S.sc->dt) / S.mn[n];
  S.u[0][n] = S.u[0][n] + s * fx;
  S.u[1][n] = S.u[1][n] + s * fy;
  if (DIM == 3) S.u[2][n] = S.u[2][n] + s * fz;
}

// Advection mode (sans-lagrange), node part: updateVelocityWithoutLagrange
// (MahycoModule.cc:702-718), u = u0 (1 - r) + (r u0) cos(pi t parametre) with r = 1 under
// the reverse option, the boundary conditions, and updatePosition (:730-751), which moves
// only the nodes shared by exactly four cells.
template <int DIM>
__device__ __forceinline__ void k_node_adv_body(const MhDev& S, int lb) {
  const long n = lb * (long)blockDim.x + threadIdx.x;
  if (n >= S.g.nn) return;
  if (mh_halted(S.sc)) return;
  const int nnx = S.g.nnx, nny = S.g.nny;
  const int i = (int)(n % nnx);
  const long r = n / nnx;
  const int j = (int)(r % nny);
  const int k = (int)(r / nny);
  const double dt = S.sc->dt;
  const long it = S.sc->iteration;
  const double cs = S.rev_cos[it < S.rev_n ? it : S.rev_n - 1];
  const double opt = S.rider_reverse ? 1. : 0.;
  double u[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const double a = S.u0[d][n];
    u[d] = a * (1. - opt) + opt * a * cs;
  }
  for (int b = 0; b < S.nbc; ++b) {
    const int sf = S.bc_surf[b];
    bool on;
    switch (sf) {
      case MH_XMIN: on = (i == 0); break;
      case MH_XMAX: on = (i == S.g.nx); break;
      case MH_YMIN: on = (j == 0); break;
      case MH_YMAX: on = (j == S.g.ny); break;
      case MH_ZMIN: on = DIM == 3 ? (k == 0) : true; break;
      default: on = DIM == 3 ? (k == S.g.nz) : true; break;
    }
    if (on) u[S.bc_type[b]] = S.bc_val[b];
  }
  const int cells = ((i > 0) + (i < S.g.nx)) * ((j > 0) + (j < S.g.ny)) * (DIM == 3 ? (k > 0) + (k < S.g.nz) : 1);
  const bool moves = cells == 4;
  const double x0[3] = {S.ax[i], S.ay[j], DIM == 3 ? S.az[k] : 0.};
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    S.u[d][n] = u[d];
    S.xlc[d][n] = moves ? x0[d] + dt * u[d] : x0[d];
  }
  if (DIM == 2) S.u[2][n] = u[2];
}

// 3D node kernel, z-marching tiles: a block owns MH_NTX x MH_NTY node columns over
// MH_NT_KC node planes, so each cell plane of the tile is (MH_NTX+1) x (MH_NTY+1) = 256
// cells, one per thread.  Each cell plane
// is processed once: its corner vectors (cqs_hex, computed once per cell instead of once
// per corner) are multiplied by p+q and kept in LDS with M/8 -- the bottom corners for
// the node plane it sits on, the top ones (held in registers meanwhile) for the next
// plane.  Every node then adds its 8 cells in cell order, as k_node_lag does, so both
// kernels give the same bits.
#ifndef MH_NTX
#define MH_NTX 31  // node columns of a tile along x
#endif
#define MH_NTY (MH_BLOCK / (MH_NTX + 1) - 1)  // along y: (MH_NTX + 1)(MH_NTY + 1) cells = MH_BLOCK
#define MH_NT_KC 8  // node planes per tile by default (MhDev::nt_kc)
#define MH_NT_CELLS ((MH_NTX + 1) * (MH_NTY + 1))  // = MH_BLOCK
// What one cell of a column needs beyond its lower face: the 4 nodes of its upper face
// (previous Lagrangian positions) and p, q, M.  The lower face is the upper face of the
// cell below, carried in registers.
struct NtCellIn {
  D3 top[4];
  double p, q, M;
};
__device__ __forceinline__ void nt_load_face(const MhDev& S, int ii, int jj, int kn, D3* f) {
  const long nnx = S.g.nnx;
  const long b = ii + nnx * (jj + (long)S.g.nny * kn);
  f[0] = load3(S.xlp, b);
  f[1] = load3(S.xlp, b + 1);
  f[2] = load3(S.xlp, b + 1 + nnx);
  f[3] = load3(S.xlp, b + nnx);
}
__device__ __forceinline__ void nt_load_cell(const MhDev& S, int ii, int jj, int kk, NtCellIn& in) {
  const long c = ii + (long)S.g.nx * (jj + (long)S.g.ny * kk);
  nt_load_face(S, ii, jj, kk + 1, in.top);
  in.p = S.p[c];
  in.q = S.q[c];
  in.M = S.M[c];
}
// corner force contributions (p + q) CQS_l and the node-mass share M/8 of one cell: lower
// corners to LDS slot s, upper corners to keep
__device__ __forceinline__ void nt_cell(const D3* bot, const NtCellIn& in, double* lds, int s, double* keep) {
  D3 x[8];
  x[0] = bot[0];
  x[1] = bot[1];
  x[2] = bot[2];
  x[3] = bot[3];
  x[4] = in.top[0];
  x[5] = in.top[1];
  x[6] = in.top[2];
  x[7] = in.top[3];
  const double pr = in.p + in.q;
  const double wm = 0.125 * in.M;
  cqs_hex_each(x, [&](int l, const D3& q) {
    if (l < 4) {
      lds[(3 * l + 0) * MH_NT_CELLS + s] = pr * q.x;
      lds[(3 * l + 1) * MH_NT_CELLS + s] = pr * q.y;
      lds[(3 * l + 2) * MH_NT_CELLS + s] = pr * q.z;
    } else {
      keep[3 * (l - 4) + 0] = pr * q.x;
      keep[3 * (l - 4) + 1] = pr * q.y;
      keep[3 * (l - 4) + 2] = pr * q.z;
    }
  });
  lds[12 * MH_NT_CELLS + s] = wm;
  keep[12] = wm;
}

__device__ __forceinline__ void nt_add(const double* P, int s, int corner, bool ok, double& mass, double& fx,
                                       double& fy, double& fz) {
  if (!ok) return;
  mass += P[12 * MH_NT_CELLS + s];
  fx += P[(3 * corner + 0) * MH_NT_CELLS + s];
  fy += P[(3 * corner + 1) * MH_NT_CELLS + s];
  fz += P[(3 * corner + 2) * MH_NT_CELLS + s];
}

__device__ __forceinline__ void k_node_lag3_tiled_body(const MhDev& S, int bx, int by, int bz) {
  static_assert(MH_NT_CELLS == MH_BLOCK, "one cell of the tile plane per thread");
  __shared__ double lo[13 * MH_NT_CELLS];  // top corners (4..7) of the cell plane below
  __shared__ double up[13 * MH_NT_CELLS];  // bottom corners (0..3) of the cell plane above
  if (mh_halted(S.sc)) return;
  const int tid = threadIdx.x;
  const int i0 = bx * MH_NTX, j0 = by * MH_NTY, k0 = bz * S.nt_kc;
  const int nx = S.g.nx, ny = S.g.ny, nz = S.g.nz;
  const int kend = min(k0 + S.nt_kc, S.g.nnz);
  const double dt = S.sc->dt;
  const double dth = 0.5 * (S.sc->dt_old + dt);
  // this thread's cell column of the tile
  const int ci = i0 - 1 + tid % (MH_NTX + 1), cj = j0 - 1 + tid / (MH_NTX + 1);
  const bool cell_ok = ci >= 0 && ci < nx && cj >= 0 && cj < ny;
  double keep[13];
  // the thread's cell column: lower face carried from plane to plane, the next cell's
  // loads issued one plane ahead
  D3 bot[4];
  NtCellIn cur, nxt;
  const int kc_end = min(kend, nz);  // cell planes k0 .. kc_end-1
  if (cell_ok) {
    if (k0 >= 1) {  // prologue: top corners of the cell plane k0 - 1
      nt_load_face(S, ci, cj, k0 - 1, bot);
      nt_load_cell(S, ci, cj, k0 - 1, cur);
      nt_cell(bot, cur, up, tid, keep);
      for (int v = 0; v < 13; ++v) lo[v * MH_NT_CELLS + tid] = keep[v];
      for (int l = 0; l < 4; ++l) bot[l] = cur.top[l];
    } else {
      nt_load_face(S, ci, cj, k0, bot);
    }
    if (k0 < kc_end) nt_load_cell(S, ci, cj, k0, nxt);
  }
  // this thread's node column (threads below MH_NTX * MH_NTY)
  const int i = i0 + tid % MH_NTX, j = j0 + tid / MH_NTX;
  const bool node_ok = tid < MH_NTX * MH_NTY && i < S.g.nnx && j < S.g.nny;
  const int tx = tid % MH_NTX, ty = tid / MH_NTX;
  for (int k = k0; k < kend; ++k) {
    // cell plane k: bottom corners to LDS now, top corners kept for node plane k + 1
    if (k < nz && cell_ok) {
      cur = nxt;
      if (k + 1 < kc_end) nt_load_cell(S, ci, cj, k + 1, nxt);
      nt_cell(bot, cur, up, tid, keep);
      for (int l = 0; l < 4; ++l) bot[l] = cur.top[l];
    }
    __syncthreads();
    if (node_ok) {
      double mass = 0., fx = 0., fy = 0., fz = 0.;
      const int s00 = ty * (MH_NTX + 1) + tx;  // cell (i-1, j-1)
      const bool okx0 = i >= 1, okx1 = i < nx, oky0 = j >= 1, oky1 = j < ny;
      if (k >= 1) {  // cells k-1, corners 4..7 -> slots 0..3 of lo
        nt_add(lo, s00, 2, okx0 && oky0, mass, fx, fy, fz);
        nt_add(lo, s00 + 1, 3, okx1 && oky0, mass, fx, fy, fz);
        nt_add(lo, s00 + MH_NTX + 1, 1, okx0 && oky1, mass, fx, fy, fz);
        nt_add(lo, s00 + MH_NTX + 2, 0, okx1 && oky1, mass, fx, fy, fz);
      }
      if (k < nz) {  // cells k, corners 0..3
        nt_add(up, s00, 2, okx0 && oky0, mass, fx, fy, fz);
        nt_add(up, s00 + 1, 3, okx1 && oky0, mass, fx, fy, fz);
        nt_add(up, s00 + MH_NTX + 1, 1, okx0 && oky1, mass, fx, fy, fz);
        nt_add(up, s00 + MH_NTX + 2, 0, okx1 && oky1, mass, fx, fy, fz);
      }
      const long n = i + (long)S.g.nnx * (j + (long)S.g.nny * k);
      node_finish<3>(S, n, i, j, k, dt, dth, mass, fx, fy, fz);
    }
    __syncthreads();
    if (k < nz && cell_ok)
      for (int v = 0; v < 13; ++v) lo[v * MH_NT_CELLS + tid] = keep[v];
  }
}

// ----------------------------------------------------------------------------- cells

// Volume of the hexahedron whose node 0 is b (hex_volume_lowreg's three component passes):
// the node coordinates are read again for each pass, and a compiler fence between the passes
// keeps their values from being shared, so one component's terms are live at a time; each pass
// adds its corner sums face by face (cqs_hex_comp_fm), so only the running sums and one face's
// terms are live and the cell kernel fits 128 registers (4 waves per SIMD).  Same bits as
// cqs_hex + cell_volume<8>.
__device__ __forceinline__ double hex_volume_passes(const MhDev& S, long b) {
  D3 x[8];
  double q[8], d[8];
  cell_nodes<3>(S, b, S.xlc, x);
  cqs_hex_comp_fm<0>(x, q);
#pragma unroll
  for (int l = 0; l < 8; ++l) d[l] = x[l].x * q[l];
  asm volatile("" ::: "memory");
  cell_nodes<3>(S, b, S.xlc, x);
  cqs_hex_comp_fm<1>(x, q);
#pragma unroll
  for (int l = 0; l < 8; ++l) d[l] = d[l] + x[l].y * q[l];
  asm volatile("" ::: "memory");
  cell_nodes<3>(S, b, S.xlc, x);
  cqs_hex_comp_fm<2>(x, q);
  double v = 0.;
#pragma unroll
  for (int l = 0; l < 8; ++l) v += d[l] + x[l].z * q[l];
  return v / 3;
}

__device__ __forceinline__ double node_speed(const MhDev& S, long n, int dim) {
  const double ux = S.u[0][n], uy = S.u[1][n];
  const double uz = dim == 3 ? S.u[2][n] : 0.;
  return sqrt(ux * ux + uy * uy + uz * uz);
}

// max over the cell's nodes of |u| (MahycoModule.cc:1771-1774)
template <int DIM>
__device__ __forceinline__ double cell_vmax(const MhDev& S, long b) {
  const long nnx = S.g.nnx, nnxy = (long)S.g.nnx * S.g.nny;
  double v = 0.;
  v = amax(node_speed(S, b, DIM), v);
  v = amax(node_speed(S, b + 1, DIM), v);
  v = amax(node_speed(S, b + 1 + nnx, DIM), v);
  v = amax(node_speed(S, b + nnx, DIM), v);
  if (DIM == 3) {
    v = amax(node_speed(S, b + nnxy, DIM), v);
    v = amax(node_speed(S, b + 1 + nnxy, DIM), v);
    v = amax(node_speed(S, b + 1 + nnx + nnxy, DIM), v);
    v = amax(node_speed(S, b + nnx + nnxy, DIM), v);
  }
  return v;
}

// updateEnergyAndPressureforGP, MahycoModule.cc:1633-1659 (pseudo-centree off)
__device__ __forceinline__ double gp_energy(double e_n, double p, double rho, double rho_n, double q,
                                            double gamma) {
  const double dv = 1.0 / rho - 1.0 / rho_n;
  double pseudo = 0.;
  if (q * (1.0 / rho - 1.0 / rho_n) < 0.) pseudo = q;
  const double denom = 1 + 0.5 * (gamma - 1.0) * rho * dv;
  const double numer = e_n - (0.5 * p + pseudo) * dv;
  return numer / denom;
}

// Energy of one environment after the Lagrange step: updateEnergyAndPressureforGP (closed
// form above), updateEnergyAndPressurebyNewton (MahycoModule.cc:1506-1540: Newton on fvnr,
// MahycoModule.h:378-386, with the EOS of applyOneCellEOS at each iterate) or
// updateEnergyAndPressureExplicite (:1484-1502).  Newton's last pressure and sound speed are
// replaced by the EOS call that follows the update (or the remap), so only e is returned.
__device__ __forceinline__ double lag_energy(const MhDev& S, const MhEnvParam& ep, double e_n, double p_n, double rho,
                                             double rho_n, double q) {
  if (S.p_explicit) return e_n - (p_n + q) * (1.0 / rho - 1.0 / rho_n);
  if (!S.newton) return gp_energy(e_n, p_n, rho, rho_n, q, ep.gamma);
  double pseudo = 0.;
  if (q * (1.0 / rho - 1.0 / rho_n) < 0.) pseudo = q;
  const double g = ep.gamma;
  double e = e_n, p = p_n;
  for (int i = 0; i < 50; ++i) {
    if (!(fabs(e - e_n + 0.5 * (p + p_n + 2. * pseudo) * (1. / rho - 1. / rho_n)) >= S.thr_module)) break;
    p = ep.stiffened ? ((g - 1.) * rho * e) - (g * ep.pinf) : (g - 1.) * rho * e;
    const double dpde = (g - 1.) * rho;
    const double f = e - e_n + 0.5 * (p + p_n + 2. * pseudo) * (1. / rho - 1. / rho_n);
    e = e - f / (1. + 0.5 * dpde * (1. / rho - 1. / rho_n));
  }
  return e;
}

// applyEOS (PerfectGasEOSService.cc:45-67, StiffenedGasEOSService.cc:43-63)
__device__ __forceinline__ void eos(const MhEnvParam& ep, double rho, double e, double& p, double& c) {
  const double g = ep.gamma;
  if (ep.stiffened) {
    p = ((g - 1.) * rho * e) - (g * ep.pinf);
    c = sqrt((g / rho) * (p + ep.pinf));
  } else {
    p = (g - 1.) * rho * e;
    c = sqrt(g * p / rho);
  }
}

// FictifEOSService::applyEOS (eos/fictif/FictifEOSService.cc:45-75): the pressure only
// depends on the time t of the iteration, the sound speed is 1 and the caller zeroes the
// internal energy.
__device__ __forceinline__ void fictif_eos(const MhEnvParam& ep, double t, double pprev, double& p, double& c) {
  p = pprev;
  if (t > ep.ft0) p = ep.fp0 + ep.fdpdt * t;
  if (t > ep.ft1) p = 0.;
  c = 1.;
}

// updateEnergyAndPressureforGP, CSTS branch (MahycoModule.cc:1661-1697): the work of the
// corner forces at n (previous cycle's corner vectors, saved velocity) and n+1 (this cycle's
// corner vectors qv, velocity at n+1/2), one set per cell ...
template <int DIM>
__device__ __forceinline__ void csts_work(const MhDev& S, long b, const D3* qv, double& cn1, double& cn, double& cdn,
                                          double& cvon) {
  constexpr int NN = DIM == 3 ? 8 : 4;
  const long nnx = S.g.nnx, nnxy = (long)S.g.nnx * S.g.nny;
  const long off[8] = {0, 1, 1 + nnx, nnx, nnxy, 1 + nnxy, 1 + nnx + nnxy, nnx + nnxy};
  D3 xp[8], qn[8];
  for (int l = 0; l < NN; ++l)
    xp[l] = d3(S.xlp[0][b + off[l]], S.xlp[1][b + off[l]], DIM == 3 ? S.xlp[2][b + off[l]] : 0.);
  if (DIM == 3) cqs_hex(xp, qn);
  else cqs_quad(xp, qn);
  const double dtn = S.sc->dt, dto = S.sc->dt_old;
  cn1 = 0.;
  cn = 0.;
  cdn = 0.;
  cvon = 0.;
  for (int l = 0; l < NN; ++l) {
    const long n = b + off[l];
    const D3 v = d3(S.u[0][n], S.u[1][n], DIM == 3 ? S.u[2][n] : 0.);
    const D3 vn = d3(S.un[0][n], S.un[1][n], DIM == 3 ? S.un[2][n] : 0.);
    const D3 dv = v - vn;
    cn1 += (v.x * qv[l].x + v.y * qv[l].y + v.z * qv[l].z) * dtn;
    cn += (v.x * qn[l].x + v.y * qn[l].y + v.z * qn[l].z) * dtn;
    cdn += (dv.x * qn[l].x + dv.y * qn[l].y + dv.z * qn[l].z) * (dtn - dto);
    cvon += (vn.x * qn[l].x + vn.y * qn[l].y + vn.z * qn[l].z) * dto;
  }
}
// ... priced per environment with its own mass, density, pressure and pseudo-viscosities
__device__ __forceinline__ double csts_e(const MhEnvParam& ep, double e, double M, double rho, double pn, double qn,
                                         double qnm1, double q, double cn1, double cn, double cdn, double cvon) {
  const double denom = 1 + 0.5 * (ep.gamma - 1.0) * rho * cn1 / M;
  const double numer = e - (0.5 * (pn + qn) * cn / M) - (0.5 * q * cn1 / M) + (0.25 * (pn + qn) * cdn / M) -
                       (0.5 * (qn - qnm1) * cvon / M);
  return numer / denom;
}
// q^(n-1) of environment k in cell c (m_pseudo_viscosity_nmoins1, MahycoModule.cc:347, 360),
// through Arcane's membership changes at the last remap (mask before it: mkm1): an
// environment that entered has a new partial value (0); one that was alone in the cell takes
// the cell value (qm1); one of a mixed cell keeps its partial value, the remap-before-last's
// U_q / V (qem1).  A cell alone in one environment now reads the cell value.
__device__ __forceinline__ double csts_qnm1(const MhDev& S, long c, int k, bool mixed_now) {
  const uint8_t pm = S.mkm1[c];
  const bool was = (pm >> k) & 1;
  if (__popc(pm) >= 2) return was ? S.qem1[c] : 0.;
  if (mixed_now) return was ? S.qm1[c] : 0.;
  return S.qm1[c];
}

// The characteristic lengths other than V^(1/d) (computeGeometricValues, MahycoModule.cc:
// 1195-1236): the median lengths between opposite faces (3D), or one side of the cell.
template <int DIM>
__device__ __forceinline__ double char_length_alt(int lc, const D3* x) {
  if (DIM == 3 && lc == MH_L_FACES_OPPOSEES) {
    auto quarter_sum = [](D3 a, D3 b, D3 c, D3 d) {
      const D3 s = a + b + c + d;
      return d3(0.25 * s.x, 0.25 * s.y, 0.25 * s.z);
    };
    auto norm = [](D3 v) { return sqrt(v.x * v.x + v.y * v.y + v.z * v.z); };
    const D3 f0 = quarter_sum(x[0], x[3], x[2], x[1]), f1 = quarter_sum(x[0], x[4], x[7], x[3]);
    const D3 f2 = quarter_sum(x[0], x[1], x[5], x[4]), f3 = quarter_sum(x[4], x[5], x[6], x[7]);
    const D3 f4 = quarter_sum(x[1], x[2], x[6], x[5]), f5 = quarter_sum(x[2], x[3], x[7], x[6]);
    const double d1 = norm(f0 - f3), d2 = norm(f2 - f5), d3v = norm(f1 - f4);
    return (d1 * d2 * d3v) / (d1 * d2 + d1 * d3v + d2 * d3v);
  }
  if (lc == MH_L_MONODIM_X) return fabs(x[2].x - x[0].x);
  if (lc == MH_L_MONODIM_Y) return fabs(x[2].y - x[0].y);
  return fabs(x[2].z - x[0].z);
}

// A cell that no environment occupies gets no remap values from computeVariablesForRemap
// (PrepareRemap.cc:100-166 loop over environment cells): its U is 0 and its Phi is the one
// the last sweep of the previous remap left, U / sum V_k (RemapADIService.cc:493-538).
// Without pente-borne the remap vector written for such a cell holds that Phi, read by the
// first sweep with a unit denominator, and the sweep starts its own U from 0.
template <int M>
__device__ __forceinline__ void empty_cell_phi(const MhDev& S, long c, double* u_out) {
  const long nc = S.g.nc;
  const double* __restrict__ Ulast = (S.nsweep & 1) ? S.Ub : S.Ua;
  double sv = 0.;
  for (int kk = 0; kk < M; ++kk) sv += Ulast[kk * nc + c];
  for (int v = 0; v <= 3 * M; ++v) u_out[v] = Ulast[v * nc + c] / sv;
}

// ComputeGeometricValues (:1097-1253) + UpdateDensity (:1267-1316) +
// ComputeArtificialViscosity (:386-426) + UpdateEnergyAndPressure (:1430-1480) for one cell.
// With remap: writes U (V_k, V_k rho_k, V_k rho_k e_k, V q; PrepareRemap.cc:100-136), V, L.
// Without: applies the EOS and the mean pressure and writes the CFL candidate.
// The cell's work once its node coordinates x (at this cycle's Lagrangian positions) are
// in registers; returns the CFL candidate of the Lagrange-only loop.
// The cell's own state read by the Lagrange step, loaded ahead of the geometry so that its
// latency is hidden behind the corner-vector arithmetic.
struct CellPre {
  uint8_t mk;
  double cs, rho_n, Mc, e_c, p_c;
};
__device__ __forceinline__ void cell_pre_load(const MhDev& S, long c, CellPre& pr) {
  pr.mk = S.mask[c];
  pr.cs = S.c[c];
  pr.rho_n = S.rho[c];
  pr.Mc = S.M[c];
  pr.e_c = S.e[c];
  pr.p_c = S.p[c];
}

// VGIVEN: the volume Vgiven was formed by the caller (the LDS-tiled kernel); x is unused.
template <int DIM, int M, int CSTS = 0, int VGIVEN = 0, int REMAP = 0>
__device__ __forceinline__ double cell_lag_core(const MhDev& S, long c, long b, const D3* x, const CellPre& pre,
                                                double Vgiven = 0.) {
  // REMAP: known at compile time to be the Lagrange step of the remap loop (with_proj set,
  // sans_lagrange not), so the other branches are not compiled.  Staged mode takes the
  // Lagrange-only branch (state and EOS written back) and k_prepare_remap forms U later.
  const bool with_proj = REMAP ? true : (S.with_proj != 0 && !S.staged);
  const bool sans_lagrange = REMAP ? false : (S.sans_lagrange != 0);
  double* __restrict__ U = S.Ua;
  constexpr int NV = 3 * M + 1;
  double dxs = 1.79769313486231570815e+308;
  {
    const uint8_t mk = pre.mk;
    const double cs = pre.cs;
    const double rho_n = pre.rho_n;
    const double Mc = pre.Mc;
    const double e_c = pre.e_c;
    const double p_c = pre.p_c;
    D3 qv[8];
    double V;
    if constexpr (VGIVEN) {
      V = Vgiven;
    } else if constexpr (DIM == 3 && !CSTS) {
      V = hex_volume_passes(S, b);  // the corner vectors themselves are not needed (x unused)
    } else if (DIM == 3) {
      cqs_hex(x, qv);
      V = cell_volume<8>(x, qv, 3);
    } else {
      cqs_quad(x, qv);
      V = cell_volume<4>(x, qv, 2);
    }
    if (V < 0.) S.err[0] = MH_ERR_NEG_VOLUME;
    // the other characteristic lengths come from k_cell_length, launched just before
    const double L = S.lc == MH_L_RACINE_CUBIQUE_VOLUME ? (DIM == 3 ? mh_pow_third(V) : mh_pow_half(V)) : S.L[c];
    if (sans_lagrange) {
      // advection mode: no density, pseudo or energy update (MahycoModule.cc:1271, 389,
      // 1433); the remap vector of PrepareRemap.cc:100-155 from the stored state
      const long nc = S.g.nc;
      double u_out[NV];
      for (int v = 0; v < NV; ++v) u_out[v] = 0.;
      if (__popc(mk) == 1) {
        const int ke = __ffs(mk) - 1;
        const double rho = rho_n, e = e_c;
        u_out[ke] = V;
        u_out[M + ke] = V * rho;
        u_out[2 * M + ke] = V * rho * e;
        if (S.pb) {
          S.phi0[ke * nc + c] = S.fg[c];
          S.phi0[(M + ke) * nc + c] = rho;
          S.phi0[(2 * M + ke) * nc + c] = e;
        }
      } else {
        for (int kk = 0; kk < M; ++kk) {
          if (!((mk >> kk) & 1)) continue;
          const long o = kk * nc + c;
          const double Vk = S.fk[o] * V;
          u_out[kk] = Vk;
          u_out[M + kk] = Vk * S.rhok[o];
          u_out[2 * M + kk] = Vk * S.rhok[o] * S.ek[o];
          if (S.pb) {
            S.phi0[o] = S.fk[o];
            S.phi0[(M + kk) * nc + c] = S.rhok[o];
            S.phi0[(2 * M + kk) * nc + c] = S.ek[o];
          }
        }
      }
      u_out[3 * M] = V * S.q[c];
      if (S.pb && mk) S.phi0[3 * M * nc + c] = S.q[c];
      if (!mk && !S.pb && S.nsweep > 0) empty_cell_phi<M>(S, c, u_out);
      for (int v = 0; v < NV; ++v) U[v * nc + c] = u_out[v];
      S.Vl[c] = V;
      S.L[c] = L;
      return dxs;
    }
    const double dt = S.sc->dt;
    const double rho = Mc / V;
    const double tau = 0.5 * (1.0 / rho_n + 1.0 / rho);
    const double divu = 1.0 / dt * (1.0 / rho - 1.0 / rho_n) / tau;
    const long nc = S.g.nc;
    double u_out[NV];
    for (int v = 0; v < NV; ++v) u_out[v] = 0.;
    if (__popc(mk) == 1) {
      const int ke = __ffs(mk) - 1;
      const MhEnvParam& ep = S.env[ke];
      double q = 0.;
      if (divu < 0.0) q = 1. / tau * (-ep.a1 * L * cs * divu + ep.a2 * L * L * divu * divu);
      double e;
      if constexpr (CSTS) {
        double cn1, cn, cdn, cvon;
        csts_work<DIM>(S, b, qv, cn1, cn, cdn, cvon);
        e = csts_e(ep, e_c, Mc, rho, p_c, S.q[c], csts_qnm1(S, c, ke, false), q, cn1, cn, cdn, cvon);
      } else {
        e = lag_energy(S, ep, e_c, p_c, rho, rho_n, q);
      }
      if (with_proj) {
        for (int kk = 0; kk < M; ++kk)
          if (kk == ke) {
            u_out[kk] = V;
            u_out[M + kk] = V * rho;
            u_out[2 * M + kk] = V * rho * e;
          }
        u_out[3 * M] = V * q;
        if (S.q_frozen) S.q[c] = q;  // the Lagrangian pseudo, for the remap's unswept V q slot
        if (CSTS) {  // SaveValuesAtN of the pseudo; this cycle's for the forward velocity step
          S.qm1[c] = S.q[c];
          S.q[c] = q;
        }
        if (S.pb) {  // pente-borne phi of the first sweep (PrepareRemap.cc:138-155)
          S.phi0[ke * nc + c] = S.fg[c];
          S.phi0[(M + ke) * nc + c] = rho;
          S.phi0[(2 * M + ke) * nc + c] = e;
          S.phi0[3 * M * nc + c] = q;
        }
      } else {
        double p, cc;
        eos(ep, rho, e, p, cc);
        if (CSTS) S.qm1[c] = S.q[c];
        S.rho[c] = rho;
        S.e[c] = e;
        S.p[c] = p;
        S.c[c] = cc;
        S.q[c] = q;
        if (!S.staged) dxs = L / (cc + cell_vmax<DIM>(S, b));
      }
    } else {
      double fk[M], rk[M], qk[M], ek[M], Vk[M];
      double qsum = 0., esum = 0.;
      double cn1 = 0., cn = 0., cdn = 0., cvon = 0.;
      if constexpr (CSTS) csts_work<DIM>(S, b, qv, cn1, cn, cdn, cvon);
      for (int kk = 0; kk < M; ++kk) {
        if (!((mk >> kk) & 1)) continue;
        const MhEnvParam& ep = S.env[kk];
        const long o = kk * nc + c;
        fk[kk] = S.fk[o];
        Vk[kk] = fk[kk] * V;
        const double rkn = S.rhok[o];
        rk[kk] = S.Mk[o] / Vk[kk];
        const double tk = 0.5 * (1.0 / rkn + 1.0 / rk[kk]);
        qk[kk] = 0.;
        if (divu < 0.0) qk[kk] = 1. / tk * (-ep.a1 * L * cs * divu + ep.a2 * L * L * divu * divu);
        if constexpr (CSTS) {
          // with remap: the history of the remap (qe, csts_qnm1); without: the environment's own
          const double qn = with_proj ? S.qe[c] : S.qkn[o];
          const double qnm1 = with_proj ? csts_qnm1(S, c, kk, true) : S.qknm1[o];
          ek[kk] = csts_e(ep, S.ek[o], S.Mk[o], rk[kk], S.pk[o], qn, qnm1, qk[kk], cn1, cn, cdn, cvon);
          if (!with_proj) {
            S.qknm1[o] = qn;
            S.qkn[o] = qk[kk];
          }
        } else
          ek[kk] = lag_energy(S, ep, S.ek[o], S.pk[o], rk[kk], rkn, qk[kk]);
      }
      for (int kk = 0; kk < M; ++kk)
        if ((mk >> kk) & 1) qsum += qk[kk] * fk[kk];
      for (int kk = 0; kk < M; ++kk)
        if ((mk >> kk) & 1) esum += S.chik[kk * nc + c] * ek[kk];
      if (with_proj) {
        for (int kk = 0; kk < M; ++kk) {
          if (!((mk >> kk) & 1)) continue;
          u_out[kk] = Vk[kk];
          u_out[M + kk] = Vk[kk] * rk[kk];
          u_out[2 * M + kk] = Vk[kk] * rk[kk] * ek[kk];
        }
        u_out[3 * M] = V * qsum;
        if (S.q_frozen) S.q[c] = qsum;
        if (CSTS) {  // as for a single-environment cell
          S.qm1[c] = S.q[c];
          S.q[c] = qsum;
        }
        if (S.pb) {
          for (int kk = 0; kk < M; ++kk) {
            if (!((mk >> kk) & 1)) continue;
            S.phi0[kk * nc + c] = fk[kk];
            S.phi0[(M + kk) * nc + c] = rk[kk];
            S.phi0[(2 * M + kk) * nc + c] = ek[kk];
          }
          if (mk) S.phi0[3 * M * nc + c] = qsum;  // an empty cell keeps the last sweep's
        }
      } else {
        double ps = 0., cmix = 0.;
        const double opt = S.vitson_max ? 1. : 0.;
        for (int kk = 0; kk < M; ++kk) {
          if (!((mk >> kk) & 1)) continue;
          double p, cc;
          eos(S.env[kk], rk[kk], ek[kk], p, cc);
          const long o = kk * nc + c;
          S.rhok[o] = rk[kk];
          S.ek[o] = ek[kk];
          S.pk[o] = p;
          ps += fk[kk] * p;
          cmix = opt * ((cc < cmix) ? cmix : cc);  // std::max(c_k, c)
          cmix += (1. - opt) * fk[kk] * cc;
        }
        if (CSTS) S.qm1[c] = S.q[c];
        S.rho[c] = rho;
        S.e[c] = esum;
        S.p[c] = ps;
        S.c[c] = cmix;
        S.q[c] = qsum;
        if (!S.staged) dxs = L / (cmix + cell_vmax<DIM>(S, b));
      }
    }
    if (with_proj && !mk && !S.pb && S.nsweep > 0) empty_cell_phi<M>(S, c, u_out);
    if (with_proj)
      for (int v = 0; v < NV; ++v) U[v * nc + c] = u_out[v];
    S.Vl[c] = V;
    S.L[c] = L;
  }
  return dxs;
}

template <int DIM, int M, int CSTS = 0>
__device__ __forceinline__ void k_cell_lag_body(const MhDev& S, int lb) {
  const long c = lb * (long)blockDim.x + threadIdx.x;
  const bool active = c < S.g.nc && !mh_halted(S.sc);
  double dxs = 1.79769313486231570815e+308;
  if (active) {
    const int nx = S.g.nx, ny = S.g.ny;
    const int i = (int)(c % nx);
    const long r = c / nx;
    const int j = (int)(r % ny);
    const int k = (int)(r / ny);
    const long b = i + (long)S.g.nnx * (j + (long)S.g.nny * k);
    CellPre pre;
    cell_pre_load(S, c, pre);
    if constexpr (DIM == 3 && !CSTS) {
      dxs = cell_lag_core<DIM, M, CSTS>(S, c, b, nullptr, pre);  // the volume reads its nodes itself
    } else {
      D3 x[8];
      cell_nodes<DIM>(S, b, S.xlc, x);
      dxs = cell_lag_core<DIM, M, CSTS>(S, c, b, x, pre);
    }
  }
  if (!S.with_proj) block_min_store(dxs, active, S.partial, lb);
}
// computeVariablesForRemap (PrepareRemap.cc:100-155) from the stored Lagrangian state, for
// staged mode: the cell kernel wrote the state of UpdateEnergyAndPressure and the caller may
// have changed it.  U = (V_k, V_k rho_k, V_k rho_k e_k, V q) with V_k = f_k V for a cell of
// several environments and V for a single-environment cell, the pente-borne phi of the first
// sweep, and an empty cell's phi as the fused cell kernel forms them (same operations, so an
// unchanged state gives the fused kernel's bits).
template <int DIM, int M>
__device__ __forceinline__ void k_prepare_remap_body(const MhDev& S, int lb) {
  const long c = lb * (long)blockDim.x + threadIdx.x;
  if (c >= S.g.nc || mh_halted(S.sc)) return;
  constexpr int NV = 3 * M + 1;
  const long nc = S.g.nc;
  const uint8_t mk = S.mask[c];
  const double V = S.Vl[c];
  const double q = S.q[c];
  double u_out[NV];
  for (int v = 0; v < NV; ++v) u_out[v] = 0.;
  if (__popc(mk) == 1) {
    const int ke = __ffs(mk) - 1;
    const double rho = S.rho[c], e = S.e[c];
    for (int kk = 0; kk < M; ++kk)
      if (kk == ke) {
        u_out[kk] = V;
        u_out[M + kk] = V * rho;
        u_out[2 * M + kk] = V * rho * e;
      }
    if (S.pb) {
      S.phi0[ke * nc + c] = S.fg[c];
      S.phi0[(M + ke) * nc + c] = rho;
      S.phi0[(2 * M + ke) * nc + c] = e;
      S.phi0[3 * M * nc + c] = q;
    }
  } else {
    for (int kk = 0; kk < M; ++kk) {
      if (!((mk >> kk) & 1)) continue;
      const long o = kk * nc + c;
      const double Vk = S.fk[o] * V;
      u_out[kk] = Vk;
      u_out[M + kk] = Vk * S.rhok[o];
      u_out[2 * M + kk] = Vk * S.rhok[o] * S.ek[o];
      if (S.pb) {
        S.phi0[o] = S.fk[o];
        S.phi0[(M + kk) * nc + c] = S.rhok[o];
        S.phi0[(2 * M + kk) * nc + c] = S.ek[o];
      }
    }
    if (S.pb && mk) S.phi0[3 * M * nc + c] = q;
  }
  u_out[3 * M] = V * q;
  if (!mk && !S.pb && S.nsweep > 0) empty_cell_phi<M>(S, c, u_out);
  for (int v = 0; v < NV; ++v) S.Ua[v * nc + c] = u_out[v];
}

// 3D cell kernel of the remap loop with the node neighbourhood staged in LDS: a block owns
// MH_CT_X x MH_CT_Y cells of MH_CT_K consecutive z planes and marches up in z.  The node
// coordinates of a tile plane ((MH_CT_X+1) x (MH_CT_Y+1) nodes, every node read from HBM once
// per tile instead of by each of its cells) go to a rolling set of three LDS planes: the two a
// cell plane needs, and the next one, whose global loads are issued before the cell plane's
// arithmetic and written to LDS after it.  The cell's own state is also loaded one plane ahead.
// The volume takes its three component passes from LDS; everything else is cell_lag_core.
#define MH_CT_X 64
#define MH_CT_Y (MH_BLOCK / MH_CT_X)
#define MH_CT_K 8  // cell planes per tile by default (MhDev::ct_k)
#define MH_CT_NX (MH_CT_X + 1)
#define MH_CT_NP (MH_CT_NX * (MH_CT_Y + 1))  // nodes of a tile plane
#define MH_CT_PER ((MH_CT_NP + MH_BLOCK - 1) / MH_BLOCK)  // node entries loaded per thread
#ifndef MH_CT_MIN_WAVES
#define MH_CT_MIN_WAVES 3
#endif

// Volume of the cell whose node 0 is entry s of the bottom plane (hex_volume_passes with the
// nodes read from LDS; same bits).
__device__ __forceinline__ double hex_volume_lds(const double* bot, const double* top, int s) {
  const int o[4] = {s, s + 1, s + 1 + MH_CT_NX, s + MH_CT_NX};
  auto load = [&](D3* x) {
#pragma unroll
    for (int l = 0; l < 4; ++l) {
      x[l] = d3(bot[o[l]], bot[MH_CT_NP + o[l]], bot[2 * MH_CT_NP + o[l]]);
      x[l + 4] = d3(top[o[l]], top[MH_CT_NP + o[l]], top[2 * MH_CT_NP + o[l]]);
    }
  };
  D3 x[8];
  double q[8], d[8];
  load(x);
  cqs_hex_comp<0>(x, q);
#pragma unroll
  for (int l = 0; l < 8; ++l) d[l] = x[l].x * q[l];
  asm volatile("" ::: "memory");
  load(x);
  cqs_hex_comp<1>(x, q);
#pragma unroll
  for (int l = 0; l < 8; ++l) d[l] = d[l] + x[l].y * q[l];
  asm volatile("" ::: "memory");
  load(x);
  cqs_hex_comp<2>(x, q);
  double v = 0.;
#pragma unroll
  for (int l = 0; l < 8; ++l) v += d[l] + x[l].z * q[l];
  return v / 3;
}

template <int M>
__device__ __forceinline__ void k_cell_lag3_tiled_body(const MhDev& S, int bx, int by, int bz) {
  __shared__ double pl[3][3 * MH_CT_NP];
  if (mh_halted(S.sc)) return;
  const int tid = threadIdx.x, tx = tid % MH_CT_X, ty = tid / MH_CT_X;
  const int nx = S.g.nx, ny = S.g.ny, nz = S.g.nz;
  const int i0 = bx * MH_CT_X, j0 = by * MH_CT_Y, k0 = bz * S.ct_k;
  const int kend = min(k0 + S.ct_k, nz);
  const int i = i0 + tx, j = j0 + ty;
  const bool ok = i < nx && j < ny;
  const long nnx = S.g.nnx, nnxy = (long)S.g.nnx * S.g.nny;
  // the node entries of a tile plane this thread loads
  double nv[MH_CT_PER][3];
  auto fetch = [&](int kn) {
#pragma unroll
    for (int r = 0; r < MH_CT_PER; ++r) {
      const int e = tid + r * MH_BLOCK;
      const int gi = i0 + e % MH_CT_NX, gj = j0 + e / MH_CT_NX;
      if (e < MH_CT_NP && gi <= nx && gj <= ny) {
        const long n = gi + nnx * gj + nnxy * kn;
        nv[r][0] = S.xlc[0][n];
        nv[r][1] = S.xlc[1][n];
        nv[r][2] = S.xlc[2][n];
      }
    }
  };
  auto put = [&](double* P) {
#pragma unroll
    for (int r = 0; r < MH_CT_PER; ++r) {
      const int e = tid + r * MH_BLOCK;
      if (e < MH_CT_NP) {
        P[e] = nv[r][0];
        P[MH_CT_NP + e] = nv[r][1];
        P[2 * MH_CT_NP + e] = nv[r][2];
      }
    }
  };
  fetch(k0);
  put(pl[0]);
  fetch(k0 + 1);
  put(pl[1]);
  const long cplane = (long)nx * ny;
  const long c0 = i + (long)nx * j;
  CellPre cur, nxt;
  if (ok && k0 < kend) cell_pre_load(S, c0 + cplane * k0, nxt);
  __syncthreads();
  for (int k = k0; k < kend; ++k) {
    const int sl = k - k0;
    const bool more = k + 1 < kend;
    if (more) fetch(k + 2);  // node plane of the next step, into registers now
    cur = nxt;
    if (ok && more) cell_pre_load(S, c0 + cplane * (k + 1), nxt);
    if (ok) {
      const long c = c0 + cplane * k;
      const long b = i + nnx * j + nnxy * k;
      const double V = hex_volume_lds(pl[sl % 3], pl[(sl + 1) % 3], tx + MH_CT_NX * ty);
      (void)cell_lag_core<3, M, 0, 1, 1>(S, c, b, nullptr, cur, V);
    }
    if (more) put(pl[(sl + 2) % 3]);
    __syncthreads();
  }
}

// The characteristic length of the other options, written to L before the cell kernel
// (which then reads it instead of forming V^(1/d)): one thread per cell.
template <int DIM>
__device__ __forceinline__ void k_cell_length_body(const MhDev& S, int lb) {
  const long c = lb * (long)blockDim.x + threadIdx.x;
  if (c >= S.g.nc || mh_halted(S.sc)) return;
  const int nx = S.g.nx, ny = S.g.ny;
  const int i = (int)(c % nx);
  const long r = c / nx;
  const int j = (int)(r % ny);
  const int k = (int)(r / ny);
  const long b = i + (long)S.g.nnx * (j + (long)S.g.nny * k);
  D3 x[8];
  cell_nodes<DIM>(S, b, S.xlc, x);
  S.L[c] = char_length_alt<DIM>(S.lc, x);
}

// CSTS (with remap): its own instance, so that the other cell kernels do not carry the
// corner-work registers
template <int DIM, int M>
__device__ __forceinline__ void k_cell_lag_csts_body(const MhDev& S, int lb) {
  k_cell_lag_body<DIM, M, 1>(S, lb);
}

// ----------------------------------------------------------------------------- remap sweeps
// Lagrangian cell centre, direction-D component (UpdatePosition, MahycoModule.cc:743-750).
template <int DIM, int D>
__device__ __forceinline__ double cell_centre_d(const MhDev& S, long b) {
  const long nnx = S.g.nnx, nnxy = (long)S.g.nnx * S.g.nny;
  const double* xd = S.xlc[D];
  double s = 0.;
  s += xd[b];
  s += xd[b + 1];
  s += xd[b + 1 + nnx];
  s += xd[b + nnx];
  if (DIM == 3) {
    s += xd[b + nnxy];
    s += xd[b + 1 + nnxy];
    s += xd[b + 1 + nnx + nnxy];
    s += xd[b + nnx + nnxy];
    return 0.125 * s;
  }
  return 0.25 * s;
}

// Lagrangian quantities of the face of direction D whose first node is f0
// (computeFaceQuantitesForRemap, PrepareRemap.cc:18-59): projected area on D,
// centre coordinate along D, normal velocity.  Face nodes in Arcane order.
template <int DIM, int D>
__device__ __forceinline__ void face_quant(const MhDev& S, long f0, double& area, double& xf, double& vf) {
  const long nnx = S.g.nnx, nnxy = (long)S.g.nnx * S.g.nny;
  if (DIM == 3) {
    long nd[4];
    if (D == 0) { nd[0] = f0; nd[1] = f0 + nnx; nd[2] = f0 + nnx + nnxy; nd[3] = f0 + nnxy; }
    else if (D == 1) { nd[0] = f0; nd[1] = f0 + 1; nd[2] = f0 + 1 + nnxy; nd[3] = f0 + nnxy; }
    else { nd[0] = f0; nd[1] = f0 + 1; nd[2] = f0 + 1 + nnx; nd[3] = f0 + nnx; }
    const D3 x0 = load3(S.xlc, nd[0]), x1 = load3(S.xlc, nd[1]);
    const D3 x2 = load3(S.xlc, nd[2]), x3 = load3(S.xlc, nd[3]);
    const D3 v1 = x2 - x0, v2 = x3 - x1;
    if (D == 0) area = 0.5 * fabs(v1.y * v2.z - v1.z * v2.y);
    else if (D == 1) area = 0.5 * fabs(v2.x * v1.z - v2.z * v1.x);
    else area = 0.5 * fabs(v1.x * v2.y - v1.y * v2.x);
    const double xa = D == 0 ? x0.x : D == 1 ? x0.y : x0.z;
    const double xb = D == 0 ? x1.x : D == 1 ? x1.y : x1.z;
    const double xc = D == 0 ? x2.x : D == 1 ? x2.y : x2.z;
    const double xd = D == 0 ? x3.x : D == 1 ? x3.y : x3.z;
    double s = 0.;
    s += 0.25 * xa;
    s += 0.25 * xb;
    s += 0.25 * xc;
    s += 0.25 * xd;
    xf = s;
    const double* ud = S.u[D];
    double us = 0.;
    if (S.csts) {  // the mean of n and n+1 (PrepareRemap.cc:236-240)
      const double* vd = S.un[D];
      for (int l = 0; l < 4; ++l) us += 0.5 * (ud[nd[l]] + vd[nd[l]]);
    } else {
      us += ud[nd[0]];
      us += ud[nd[1]];
      us += ud[nd[2]];
      us += ud[nd[3]];
    }
    vf = 0.25 * us;
  } else {
    const long n1 = D == 0 ? f0 + nnx : f0 + 1;
    const double dxv = S.xlc[0][n1] - S.xlc[0][f0];
    const double dyv = S.xlc[1][n1] - S.xlc[1][f0];
    area = D == 0 ? fabs(dyv) : fabs(dxv);
    const double* xd = S.xlc[D];
    double s = 0.;
    s += 0.5 * xd[f0];
    s += 0.5 * xd[n1];
    xf = s;
    const double* ud = S.u[D];
    double us = 0.;
    if (S.csts) {
      const double* vd = S.un[D];
      us += 0.5 * (ud[f0] + vd[f0]);
      us += 0.5 * (ud[n1] + vd[n1]);
    } else {
      us += ud[f0];
      us += ud[n1];
    }
    vf = 0.5 * us;
  }
}

// One ADI direction, cell part: computeGradPhiFace (RemapADIService.cc:119-149),
// computeGradPhiCell + computeAndLimitGradPhi (:161-267, UtilesRemap.cc:325-357),
// computeUpwindFaceQuantitiesForProjection (:279-383) and computeUremap (:395-575).
// Phi of the stencil is rebuilt from U (U / V before the first sweep, U / sum V_k after).
template <int DIM, int D, int M, int LIM>
__device__ __forceinline__ void k_sweep_cell_body(const MhDev& S, int lb, int usel, int first) {
  const double* __restrict__ Uin = usel ? S.Ub : S.Ua;
  double* __restrict__ Uout = usel ? S.Ua : S.Ub;
  double* __restrict__ dflux = S.dflux;
  constexpr int NV = 3 * M + 1;
  const long c = lb * (long)blockDim.x + threadIdx.x;
  if (c >= S.g.nc) return;
  if (mh_halted(S.sc)) return;
  const bool own = !(first && S.mask[c] == 0);  // an empty cell's U starts from 0 (its Phi is stale)
  const int nx = S.g.nx, ny = S.g.ny;
  const int i = (int)(c % nx);
  const long r = c / nx;
  const int j = (int)(r % ny);
  const int k = (int)(r / ny);
  const int a = D == 0 ? i : D == 1 ? j : k;
  const int na = D == 0 ? nx : D == 1 ? ny : S.g.nz;
  const long sD = D == 0 ? 1 : D == 1 ? (long)nx : (long)nx * ny;
  const long nD = D == 0 ? 1 : D == 1 ? (long)S.g.nnx : (long)S.g.nnx * S.g.nny;
  const long b = i + (long)S.g.nnx * (j + (long)S.g.nny * k);
  const long nc = S.g.nc;
  const double dt = S.sc->dt;
  // per stencil cell (offsets -2..2 along D): existence, Phi denominator, centre along D
  double den[5], xc[5];
  bool valid[5];
#pragma unroll
  for (int o = -2; o <= 2; ++o) {
    const int bb = a + o;
    valid[o + 2] = (bb >= 0 && bb < na);
    den[o + 2] = 1.;
    xc[o + 2] = 0.;
    if (valid[o + 2]) {
      const long cc = c + o * sD;
      if (first) {
        den[o + 2] = S.mask[cc] ? S.Vl[cc] : 1.;  // an empty cell holds Phi itself
      } else {
        double sv = 0.;
        for (int kk = 0; kk < M; ++kk) sv += Uin[kk * nc + cc];
        den[o + 2] = sv;
      }
      xc[o + 2] = cell_centre_d<DIM, D>(S, b + o * nD);
    }
  }
  // face f lies between offsets f-2 and f-1; Delta x along D (boundary faces: none)
  double dxf[4];
#pragma unroll
  for (int f = 0; f < 4; ++f) dxf[f] = (valid[f] && valid[f + 1]) ? xc[f + 1] - xc[f] : 0.;
  // the two faces of the cell along D: s = 0 back (outer normal -e_D), s = 1 front (+e_D)
  bool inner[2], upprev[2];
  double area[2], vf[2], hup[2];
#pragma unroll
  for (int s = 0; s < 2; ++s) {
    inner[s] = s == 0 ? (a > 0) : (a < na - 1);
    area[s] = 0.;
    vf[s] = 0.;
    hup[s] = 0.;
    upprev[s] = false;
    if (!inner[s]) continue;
    double xf;
    face_quant<DIM, D>(S, b + s * nD, area[s], xf, vf[s]);
    upprev[s] = vf[s] * dxf[s + 1] > 0.0;  // upwind cell: the one behind the face
    hup[s] = (double)S.order2 * (xf - (upprev[s] ? xc[s + 1] : xc[s + 2]));
  }
  // cell faces in Arcane local order: 3D min face first; 2D x: max face first
  const int s0 = (DIM == 2 && D == 0) ? 1 : 0, s1 = 1 - s0;
  const double od[2] = {-1., 1.};
  for (int v = 0; v < NV; ++v) {
    double phi[5];
#pragma unroll
    for (int o = 0; o < 5; ++o) phi[o] = valid[o] ? Uin[v * nc + c + (o - 2) * sD] / den[o] : 0.;
    double gface[4];
#pragma unroll
    for (int f = 0; f < 4; ++f) gface[f] = (valid[f] && valid[f + 1]) ? (phi[f + 1] - phi[f]) / dxf[f] : 0.;
    double grad[3];
#pragma unroll
    for (int o = 0; o < 3; ++o) {
      double g = 0.;
      if (S.order2 && valid[o + 1]) {
        const double gb = gface[o], gf = gface[o + 1];
        if (gb != 0.) g += 0.5 * (mh_limiter<LIM>(S.limiter, gf / gb) * gb);
        if (gf != 0.) g += 0.5 * (mh_limiter<LIM>(S.limiter, gb / gf) * gf);
      }
      grad[o] = g;
    }
    double flux[2] = {0., 0.};
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      if (!inner[s]) continue;
      // Phi_face = Phi_up + h grad_up, up = offset s-1 (behind) or s (ahead)
      const double pu = upprev[s] ? phi[s + 1] : phi[s + 2];
      const double gu = upprev[s] ? grad[s] : grad[s + 1];
      const double pf = pu + hup[s] * gu;
      flux[s] = od[s] * vf[s] * area[s] * dt * pf;
    }
    double ff = 0.;
    if (inner[s0]) ff += flux[s0];
    if (inner[s1]) ff += flux[s1];
    double un = (own ? Uin[v * nc + c] : 0.) - ff;
    if (v >= M && v < 2 * M) {
      double dd = 0.;
      if (inner[s0]) dd += 0.5 * flux[s0] * od[s0];
      if (inner[s1]) dd += 0.5 * flux[s1] * od[s1];
      dflux[(v - M) * nc + c] = dd;
    }
    if (v >= M && v < 3 * M && un < 0.) un = 0.;  // computeUremap clamps, :465-485
    Uout[v * nc + c] = un;
  }
}

// ----------------------------------------------------------------------------- order-3 remap
// Lagrangian centre of the face of direction D whose first node is f0, all three components,
// summed over the face nodes in Arcane order as computeFaceQuantitesForRemap does
// (PrepareRemap.cc:18-59): 0 + w x_0 + w x_1 (+ w x_2 + w x_3), w = 1 / nodes.
template <int DIM, int D>
__device__ __forceinline__ D3 face_centre(const MhDev& S, long f0) {
  const long nnx = S.g.nnx, nnxy = (long)S.g.nnx * S.g.nny;
  double r[3] = {0., 0., 0.};
  if (DIM == 3) {
    long nd[4];
    if (D == 0) { nd[0] = f0; nd[1] = f0 + nnx; nd[2] = f0 + nnx + nnxy; nd[3] = f0 + nnxy; }
    else if (D == 1) { nd[0] = f0; nd[1] = f0 + 1; nd[2] = f0 + 1 + nnxy; nd[3] = f0 + nnxy; }
    else { nd[0] = f0; nd[1] = f0 + 1; nd[2] = f0 + 1 + nnx; nd[3] = f0 + nnx; }
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      double s = 0.;
#pragma unroll
      for (int l = 0; l < 4; ++l) s += 0.25 * S.xlc[q][nd[l]];
      r[q] = s;
    }
  } else {
    const long n1 = D == 0 ? f0 + nnx : f0 + 1;
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      double s = 0.;
      s += 0.5 * S.xlc[q][f0];
      s += 0.5 * S.xlc[q][n1];
      r[q] = s;
    }
  }
  return d3(r[0], r[1], r[2]);
}

__device__ __forceinline__ double norm3(D3 d) { return sqrt(d.x * d.x + d.y * d.y + d.z * d.z); }

// Cell widths of the order-3 remap, every direction at once (computeGradPhiFace,
// RemapADIService.cc:119-149): h_D(c) = sum over the inner faces of c normal to D of
// |x_f - x_c|, with the Lagrangian face and cell centres (full 3-vectors).  The Lagrangian
// mesh does not move during the remap, so one launch before the first sweep serves all of them.
template <int DIM>
__device__ __forceinline__ void k_hcell_body(const MhDev& S, int lb) {
  const long c = lb * (long)blockDim.x + threadIdx.x;
  if (c >= S.g.nc) return;
  if (mh_halted(S.sc)) return;
  const int nx = S.g.nx, ny = S.g.ny;
  const int i = (int)(c % nx);
  const long r = c / nx;
  const int j = (int)(r % ny);
  const int k = (int)(r / ny);
  const long b = i + (long)S.g.nnx * (j + (long)S.g.nny * k);
  const long nnx = S.g.nnx, nnxy = (long)S.g.nnx * S.g.nny;
  const D3 xc = d3(cell_centre_d<DIM, 0>(S, b), cell_centre_d<DIM, 1>(S, b),
                   DIM == 3 ? cell_centre_d<DIM, (DIM == 3 ? 2 : 0)>(S, b) : 0.);
  const long nc = S.g.nc;
  {
    double h = 0.;
    if (i > 0) h += norm3(face_centre<DIM, 0>(S, b) - xc);
    if (i < nx - 1) h += norm3(face_centre<DIM, 0>(S, b + 1) - xc);
    S.hc[c] = h;
  }
  {
    double h = 0.;
    if (j > 0) h += norm3(face_centre<DIM, 1>(S, b) - xc);
    if (j < ny - 1) h += norm3(face_centre<DIM, 1>(S, b + nnx) - xc);
    S.hc[nc + c] = h;
  }
  if (DIM == 3) {
    double h = 0.;
    if (k > 0) h += norm3(face_centre<DIM, (DIM == 3 ? 2 : 0)>(S, b) - xc);
    if (k < S.g.nz - 1) h += norm3(face_centre<DIM, (DIM == 3 ? 2 : 0)>(S, b + nnxy) - xc);
    S.hc[2 * nc + c] = h;
  }
}

// Remap-ordre-3.cc:16-112: evaluate_grad, evaluate_ystar, evaluate_fm / evaluate_fp (the
// piecewise-parabolic reconstruction on a non-uniform 1D stencil).
__device__ __forceinline__ double o3_grad(double hm, double h0, double hp, double ym, double y0, double yp) {
  return h0 / (hm + h0 + hp) * ((2. * hm + h0) / (h0 + hp) * (yp - y0) + (h0 + 2. * hp) / (hm + h0) * (y0 - ym));
}
__device__ __forceinline__ double o3_ystar(double hmm, double hm, double hp, double hpp, double ymm, double ym,
                                           double yp, double ypp, double gradm, double gradp) {
  const double t1 = (2. * hp * hm) / (hm + hp) * ((hmm + hm) / (2. * hm + hp) - (hpp + hp) / (2. * hp + hm)) * (yp - ym);
  const double t2 = -hm * (hmm + hm) / (2. * hm + hp) * gradp + hp * (hp + hpp) / (hm + 2. * hp) * gradm;
  return ym + hm / (hm + hp) * (yp - ym) + 1. / (hmm + hm + hp + hpp) * (t1 + t2);
}
__device__ __forceinline__ double o3_fm(double x, double dx, double up, double du, double u6) {
  return up - 0.5 * x / dx * (du - (1. - 2. / 3. * x / dx) * u6);
}
__device__ __forceinline__ double o3_fp(double x, double dx, double um, double du, double u6) {
  return um + 0.5 * x / dx * (du - (1. - 2. / 3. * x / dx) * u6);
}

// ComputeFluxOrdre3 (Remap-ordre-3.cc:127-189): face value from the six cells y[0..5]
// around the face (y[2] behind, y[3] ahead) of widths h[0..5], for the displacement
// vdt = u_face dt, clipped to the interval the upwind cell and its neighbours allow.
__device__ __forceinline__ double flux_ordre3(const double* y, const double* h, double vdt) {
  const double gmm = o3_grad(h[0], h[1], h[2], y[0], y[1], y[2]);
  const double gm = o3_grad(h[1], h[2], h[3], y[1], y[2], y[3]);
  const double gp = o3_grad(h[2], h[3], h[4], y[2], y[3], y[4]);
  const double gpp = o3_grad(h[3], h[4], h[5], y[3], y[4], y[5]);
  const double ysm = o3_ystar(h[0], h[1], h[2], h[3], y[0], y[1], y[2], y[3], gmm, gm);
  const double ys = o3_ystar(h[1], h[2], h[3], h[4], y[1], y[2], y[3], y[4], gm, gp);
  const double ysp = o3_ystar(h[2], h[3], h[4], h[5], y[2], y[3], y[4], y[5], gp, gpp);
  const double ym = y[2], yp = y[3], hm = h[2], hp = h[3];
  double flux;
  if (vdt > 0.) flux = o3_fm(vdt, hm, ys, ys - ysm, 6. * (ym - 0.5 * (ysm + ys)));
  else if (vdt < 0.) flux = o3_fp(-vdt, hp, ys, ysp - ys, 6. * (yp - 0.5 * (ys + ysp)));
  else return 0.;
  const double num = vdt / hm, nup = vdt / hp;
  const double ym_ym = ym + (1. - num) / num * (ym - y[1]);
  const double yp_ym = yp - (1. + nup) / nup * (yp - y[4]);
  const double a0 = amin(ym, yp), a1 = amax(ym, yp);
  double b0, b1;
  if (vdt >= 0.) {
    b0 = amin(ym, ym_ym);
    b1 = amax(ym, ym_ym);
  } else {
    b0 = amin(yp, yp_ym);
    b1 = amax(yp, yp_ym);
  }
  double lo, hi;
  if ((a1 < b0) || (b1 < a0)) {
    lo = 0.;
    hi = 0.;
  } else {
    lo = amax(a0, b0);
    hi = amin(a1, b1);
  }
  if (flux < lo) flux = lo;
  if (flux > hi) flux = hi;
  return flux;
}

// One ADI direction, cell part, order 3 (computeUpwindFaceQuantitiesForProjection,
// RemapADIService.cc:319-379, then computeUremap :395-575): the face value of each
// variable comes from the six cells around the face along D, clamped at the mesh boundary
// (ComputeFluxOrdre3).  Across a face whose two cells are not pure cells of one
// environment, the mass and energy face values are rebuilt from the upwind cell's
// ratios.  The flags of the two cells come from the sweep input (volume fractions > 0,
// :559-573), or in the first sweep of a cycle from the flags remapVariables left
// (RemapADIFinal.cc:124-145).  Phi as in k_sweep_cell_body.
template <int DIM, int D, int M>
__device__ __forceinline__ void k_sweep_cell3_body(const MhDev& S, int lb, int usel, int first) {
  const double* __restrict__ Uin = usel ? S.Ub : S.Ua;
  double* __restrict__ Uout = usel ? S.Ua : S.Ub;
  constexpr int NV = 3 * M + 1;
  const long c = lb * (long)blockDim.x + threadIdx.x;
  if (c >= S.g.nc) return;
  if (mh_halted(S.sc)) return;
  const bool own = !(first && S.mask[c] == 0);  // an empty cell's U starts from 0 (its Phi is stale)
  const int nx = S.g.nx, ny = S.g.ny;
  const int i = (int)(c % nx);
  const long r = c / nx;
  const int j = (int)(r % ny);
  const int k = (int)(r / ny);
  const int a = D == 0 ? i : D == 1 ? j : k;
  const int na = D == 0 ? nx : D == 1 ? ny : S.g.nz;
  const long sD = D == 0 ? 1 : D == 1 ? (long)nx : (long)nx * ny;
  const long nD = D == 0 ? 1 : D == 1 ? (long)S.g.nnx : (long)S.g.nnx * S.g.nny;
  const long b = i + (long)S.g.nnx * (j + (long)S.g.nny * k);
  const long nc = S.g.nc;
  const double dt = S.sc->dt;
  // stencil: offsets -3..3 along D, clamped to the mesh (index o + 3)
  long cs[7];
  double den[7], hs[7];
#pragma unroll
  for (int o = 0; o < 7; ++o) {
    const int ao = min(max(a + o - 3, 0), na - 1);
    cs[o] = c + (ao - a) * sD;
    if (first) {
      den[o] = S.mask[cs[o]] ? S.Vl[cs[o]] : 1.;  // an empty cell holds Phi itself
    } else {
      double sv = 0.;
      for (int kk = 0; kk < M; ++kk) sv += Uin[kk * nc + cs[o]];
      den[o] = sv;
    }
    hs[o] = S.hc[D * nc + cs[o]];
  }
  // the two faces of the cell along D: s = 0 back (cells at offsets -1, 0), s = 1 front (0, +1)
  bool inner[2];
  double area[2], vf[2], vdt[2];
#pragma unroll
  for (int s = 0; s < 2; ++s) {
    inner[s] = s == 0 ? (a > 0) : (a < na - 1);
    area[s] = 0.;
    vf[s] = 0.;
    vdt[s] = 0.;
    if (!inner[s]) continue;
    double xf;
    face_quant<DIM, D>(S, b + s * nD, area[s], xf, vf[s]);
    vdt[s] = vf[s] * dt;
  }
  // flags of the cells at offsets -1, 0, +1
  int code[3] = {0, 0, 0}, cnt[3] = {0, 0, 0}, pur[3] = {-1, -1, -1};
  if (first) {
#pragma unroll
    for (int o = 0; o < 3; ++o) code[o] = S.est[cs[o + 2]];
  }
  bool vpure[2] = {true, true};
  double fv[2][M], pv[2][M], fm[2][M], pm[2][M];
  bool cor[2][M];
  const int s0 = (DIM == 2 && D == 0) ? 1 : 0, s1 = 1 - s0;
  const double od[2] = {-1., 1.};
  for (int v = 0; v < NV; ++v) {
    if (v == M) {
      if (!first) {
#pragma unroll
        for (int o = 0; o < 3; ++o) code[o] = cnt[o] > 1 ? MH_EST_MIXED : pur[o];
      }
#pragma unroll
      for (int s = 0; s < 2; ++s) vpure[s] = code[s] == code[s + 1] && code[s] != MH_EST_MIXED;
    }
    double phi[7];
#pragma unroll
    for (int o = 0; o < 7; ++o) phi[o] = Uin[v * nc + cs[o]] / den[o];
    if (v < M && !first) {
#pragma unroll
      for (int o = 0; o < 3; ++o)
        if (phi[o + 2] > 0.) {
          cnt[o]++;
          pur[o] = v;
        }
    }
    double flux[2] = {0., 0.};
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      if (!inner[s]) continue;
      double pf = flux_ordre3(&phi[s], &hs[s], vdt[s]);
      // upwind cell of the face: offset s-1 (vdt > 0) or s (vdt < 0)
      const double pu = vdt[s] > 0. ? phi[s + 2] : phi[s + 3];
      if (v < M) {
        fv[s][v] = pf;
        pv[s][v] = pu;
      } else if (v < 2 * M) {
        const int e = v - M;
        cor[s][e] = !vpure[s] && (vdt[s] > 0. || vdt[s] < 0.) && pv[s][e] != 0. && pu != 0.;
        if (cor[s][e]) pf = (pu / pv[s][e]) * fv[s][e];
        fm[s][e] = pf;
        pm[s][e] = pu;
      } else if (v < 3 * M) {
        const int e = v - 2 * M;
        if (cor[s][e]) pf = (pu / pm[s][e]) * fm[s][e];
      }
      flux[s] = od[s] * vf[s] * area[s] * dt * pf;
    }
    double ff = 0.;
    if (inner[s0]) ff += flux[s0];
    if (inner[s1]) ff += flux[s1];
    double un = (own ? Uin[v * nc + c] : 0.) - ff;
    if (v >= M && v < 2 * M) {
      double dd = 0.;
      if (inner[s0]) dd += 0.5 * flux[s0] * od[s0];
      if (inner[s1]) dd += 0.5 * flux[s1] * od[s1];
      S.dflux[(v - M) * nc + c] = dd;
    }
    if (v >= M && v < 3 * M && un < 0.) un = 0.;  // computeUremap clamps, :465-485
    Uout[v * nc + c] = un;
  }
}


// ----------------------------------------------------------------------------- generalised
// limiters and pente-borne fluxes (UtilesRemap.cc), in the reference's operation order.
// std::min / std::max as the reference's host code has them.
__device__ __forceinline__ double smin(double a, double b) { return (b < a) ? b : a; }
__device__ __forceinline__ double smax(double a, double b) { return (a < b) ? b : a; }

// fluxLimiterG (UtilesRemap.cc:130-183): slope from the front / back gradients gp, gm and
// the neighbour values yp, ym for widths h0 (cell), hp / hm (front / back neighbour).
__device__ __forceinline__ double flux_limiter_g(int id, double gp, double gm, double yp, double ym, double h0,
                                                 double hp, double hm, int simple_pente) {
  double grady = 0.;
  if (gp * gm < 0.0) return 0.;
  if (id == MH_MINMOD_G) {
    grady = (yp - ym) > 0. ? smin(fabs(gp), fabs(gm)) : -smin(fabs(gp), fabs(gm));
  } else if (id == MH_SUPERBEE_G) {
    grady = (yp - ym) > 0. ? smax(fabs(gp), fabs(gm)) : -smax(fabs(gp), fabs(gm));
  } else if (id == MH_VANLEER_G) {
    const double lp = (h0 / 2. + hp) / (h0 + hp + hm);
    const double lm = (h0 / 2. + hm) / (h0 + hp + hm);
    const double den = lp * gp + lm * gm;
    grady = den != 0. ? gp * gm / den : 0.;
  } else if (id == MH_ULTRABEE_G) {
    grady = (yp - ym) / h0;
  } else if (id == MH_ARITHMETIC_G) {
    const double lp = (h0 / 2. + hp) / (h0 + hp + hm);
    const double lm = (h0 / 2. + hm) / (h0 + hp + hm);
    grady = lm * gp + lp * gm;
  }
  if (simple_pente) {
    const double gMp = gp * (h0 + hp) / h0;
    const double gMm = gm * (h0 + hm) / h0;
    const double gM = smin(fabs(gMp), fabs(gMm));
    grady = (yp - ym) > 0. ? smin(fabs(gM), fabs(grady)) : -smin(fabs(gM), fabs(grady));
  }
  return grady;
}

// computeAndLimitGradPhi (UtilesRemap.cc:325-357): classical limiters from the back / front
// face gradients gb, gf; generalised ones (id >= MH_MINMOD_G) from values and widths.
__device__ __forceinline__ double limited_grad(int id, double gb, double gf, double yp, double ym, double h0,
                                               double hp, double hm, int simple_pente) {
  if (id < MH_MINMOD_G) {
    double g = 0.;
    if (gb != 0.) g += 0.5 * (mh_flux_limiter(id, gf / gb) * gb);
    if (gf != 0.) g += 0.5 * (mh_flux_limiter(id, gb / gf) * gf);
    return g;
  }
  return flux_limiter_g(id, gf, gb, yp, ym, h0, hp, hm, simple_pente);
}

// INTY (UtilesRemap.cc:86-97): integral from x0 to X of the segment (x0, y0)-(x1, y1).
__device__ __forceinline__ double inty(double X, double x0, double y0, double x1, double y1, double thr) {
  double flux = 0.;
  if (fabs(x1 - x0) > thr) {
    const double eta = amin(amax(0., (X - x0) / (x1 - x0)), 1.);
    flux = (y0 + 0.5 * eta * (y1 - y0)) * (x1 - x0) * eta;
  }
  return flux;
}

// computeY0 (UtilesRemap.cc:194-258): monotonicity thresholds y0plus, y0moins.
__device__ __forceinline__ void compute_y0(int id, double yplus, double ymoins, double h0, double hplus, double hmoins,
                                           int simple_pente, double& y0plus, double& y0moins) {
  y0plus = 0.;
  y0moins = 0.;
  if (id == MH_MINMOD_G || id == MH_MINMOD) {
    y0plus = yplus;
    y0moins = ymoins;
  } else if (id == MH_SUPERBEE_G || id == MH_SUPERBEE) {
    y0plus = ((h0 + hmoins) * yplus + h0 * ymoins) / (2 * h0 + hmoins);
    y0moins = ((h0 + hplus) * ymoins + h0 * yplus) / (2 * h0 + hplus);
  } else if (id == MH_VANLEER_G || id == MH_VANLEER) {
    const double a = amin(yplus, ymoins);
    const double b = amax(yplus, ymoins);
    double xplus = (h0 * h0 + 3 * h0 * hmoins + 2 * hmoins * hmoins) * yplus;
    double xmoins = (h0 * h0 + 3 * h0 * hplus + 2 * hplus * hplus) * ymoins;
    xplus += (h0 * h0 - h0 * hplus - 2 * hplus * hplus + 2 * h0 * hmoins) * ymoins;
    xmoins += (h0 * h0 - h0 * hmoins - 2 * hmoins * hmoins + 2 * h0 * hplus) * yplus;
    xplus /= (2 * h0 * h0 + 5 * h0 * hmoins + 2 * hmoins * hmoins - h0 * hplus - 2 * hplus * hplus);
    xmoins /= (2 * h0 * h0 + 5 * h0 * hplus + 2 * hplus * hplus - h0 * hmoins - 2 * hmoins * hmoins);
    y0plus = amin(amax(xplus, a), b);
    y0moins = amin(amax(xmoins, a), b);
  } else if (id == MH_ULTRABEE_G) {
    y0plus = (yplus + ymoins) / 2.;
    y0moins = (yplus + ymoins) / 2.;
  } else if (id == MH_ARITHMETIC_G) {
    y0plus = ((h0 + hmoins + hplus) * yplus + h0 * ymoins) / (2 * h0 + hmoins + hplus);
    y0moins = ((h0 + hmoins + hplus) * ymoins + h0 * yplus) / (2 * h0 + hmoins + hplus);
  }
  if (simple_pente) {
    y0plus = yplus;
    y0moins = ymoins;
  }
}

// One pente-borne flux of one variable (computeFluxPP / computeFluxPPPure,
// UtilesRemap.cc:425-832): the three-piece reconstruction of computexgxd (:269-290) and
// computeygyd (:302-318) integrated over the part of the cell that leaves through its back
// (type 0) or front (type 1) face; ppv = positive part of the outgoing velocity x dt.
__device__ __forceinline__ double pb_flux(int type, double y0, double yplus, double ymoins, double h0, double hplus,
                                          double hmoins, double grady, double ppv, int limY, double thr,
                                          int simple_pente) {
  double y0plus, y0moins;
  compute_y0(limY, yplus, ymoins, h0, hplus, hmoins, simple_pente, y0plus, y0moins);
  double xplus = 1.;
  if (fabs(y0plus - yplus) > thr) xplus = (y0 - yplus) / (y0plus - yplus) - 1. / 2.;
  double xmoins = 1.;
  if (fabs(y0moins - ymoins) > thr) xmoins = (y0 - ymoins) / (y0moins - ymoins) - 1. / 2.;
  const double xd = +h0 * amin(amax(xplus, -1. / 2.), 1. / 2.);
  const double xg = -h0 * amin(amax(xmoins, -1. / 2.), 1. / 2.);
  const double xtd = y0 + h0 / 2 * grady;
  const double xtg = y0 - h0 / 2 * grady;
  const double ad = amin(yplus, 2. * y0moins - ymoins);
  const double bd = amax(yplus, 2. * y0moins - ymoins);
  const double ag = amin(ymoins, 2. * y0plus - yplus);
  const double bg = amax(ymoins, 2. * y0plus - yplus);
  const double yd = amin(amax(xtd, ad), bd);
  const double yg = amin(amax(xtg, ag), bg);
  double f1, f1m, f2, f2m, f3, f3m;
  if (type == 0) {
    f1m = inty(-h0 / 2. + ppv, -h0 / 2., ymoins, xg, yg, thr);
    f1 = inty(-h0 / 2., -h0 / 2., ymoins, xg, yg, thr);
    f2m = inty(-h0 / 2. + ppv, xg, yg, xd, yd, thr);
    f2 = inty(-h0 / 2., xg, yg, xd, yd, thr);
    f3m = inty(-h0 / 2. + ppv, xd, yd, h0 / 2., yplus, thr);
    f3 = inty(-h0 / 2., xd, yd, h0 / 2., yplus, thr);
  } else {
    f1 = inty(h0 / 2. - ppv, -h0 / 2., ymoins, xg, yg, thr);
    f1m = inty(h0 / 2., -h0 / 2., ymoins, xg, yg, thr);
    f2 = inty(h0 / 2. - ppv, xg, yg, xd, yd, thr);
    f2m = inty(h0 / 2., xg, yg, xd, yd, thr);
    f3 = inty(h0 / 2. - ppv, xd, yd, h0 / 2., yplus, thr);
    f3m = inty(h0 / 2., xd, yd, h0 / 2., yplus, thr);
  }
  double F = smax(((f1m - f1) + (f2m - f2) + (f3m - f3)), 0.);
  if (((yplus - y0) * (ymoins - y0)) >= 0.) F = y0 * ppv;  // formula 16
  return F;
}

// One ADI direction, cell part, for the remap variants at order <= 2: runtime classical or
// generalised limiters (the limiter of a pure neighbourhood when projection-limiteur-pure-id
// is 1, RemapADIService.cc:205-209), and pente-borne fluxes (:213-264, :312-317, :424-428,
// :441-463, :493-538): phi is (volume fraction, density, specific energy, pseudo), each
// face carries the flux leaving the cell behind it minus the flux leaving the cell ahead,
// each computed from that cell and its two neighbours, then adjusted by the debar fix
// (UtilesRemap.cc:577-623, :782-826).  Boundary cells are their own missing neighbour.
template <int DIM, int D, int M>
__device__ __forceinline__ void k_sweep_cell_gen_body(const MhDev& S, int lb, int usel, int first) {
  const double* __restrict__ Uin = usel ? S.Ub : S.Ua;
  double* __restrict__ Uout = usel ? S.Ua : S.Ub;
  constexpr int NV = 3 * M + 1;
  const long c = lb * (long)blockDim.x + threadIdx.x;
  if (c >= S.g.nc) return;
  if (mh_halted(S.sc)) return;
  const bool own = !(first && S.mask[c] == 0);  // an empty cell's U starts from 0 (its Phi is stale)
  const int nx = S.g.nx, ny = S.g.ny;
  const int i = (int)(c % nx);
  const long r = c / nx;
  const int j = (int)(r % ny);
  const int k = (int)(r / ny);
  const int a = D == 0 ? i : D == 1 ? j : k;
  const int na = D == 0 ? nx : D == 1 ? ny : S.g.nz;
  const long sD = D == 0 ? 1 : D == 1 ? (long)nx : (long)nx * ny;
  const long nD = D == 0 ? 1 : D == 1 ? (long)S.g.nnx : (long)S.g.nnx * S.g.nny;
  const long b = i + (long)S.g.nnx * (j + (long)S.g.nny * k);
  const long nc = S.g.nc;
  const double dt = S.sc->dt;
  const bool pb = S.pb != 0;
  const double thr = S.thr_remap;
  const int sp = S.simple_pente;
  // stencil offsets -2..2 (index o + 2), clamped to the mesh
  long cs[5];
  bool valid[5];
  double den[5], xc[5], hs[5];
#pragma unroll
  for (int o = 0; o < 5; ++o) {
    const int bb = a + o - 2;
    valid[o] = bb >= 0 && bb < na;
    const int ao = min(max(bb, 0), na - 1);
    cs[o] = c + (ao - a) * sD;
    hs[o] = S.hc[D * nc + cs[o]];
    xc[o] = valid[o] ? cell_centre_d<DIM, D>(S, b + (o - 2) * nD) : 0.;
    if (first && !pb) {
      den[o] = S.mask[cs[o]] ? S.Vl[cs[o]] : 1.;  // an empty cell holds Phi itself
    } else {
      double sv = 0.;
      if (!first)
        for (int kk = 0; kk < M; ++kk) sv += Uin[kk * nc + cs[o]];
      den[o] = sv;
    }
  }
  double dxf[4];
#pragma unroll
  for (int f = 0; f < 4; ++f) dxf[f] = (valid[f] && valid[f + 1]) ? xc[f + 1] - xc[f] : 0.;
  // the two faces of the cell along D: s = 0 back (outer normal -e_D), s = 1 front (+e_D)
  bool inner[2], upprev[2];
  double area[2], vf[2], hup[2];
#pragma unroll
  for (int s = 0; s < 2; ++s) {
    inner[s] = s == 0 ? (a > 0) : (a < na - 1);
    area[s] = 0.;
    vf[s] = 0.;
    hup[s] = 0.;
    upprev[s] = false;
    if (!inner[s]) continue;
    double xf;
    face_quant<DIM, D>(S, b + s * nD, area[s], xf, vf[s]);
    upprev[s] = vf[s] * dxf[s + 1] > 0.0;
    hup[s] = (double)S.order2 * (xf - (upprev[s] ? xc[s + 1] : xc[s + 2]));
  }
  // pure neighbourhoods of the cells at offsets -1, 0, +1 (index x = 1, 2, 3)
  bool vpure[5] = {false, false, false, false, false};
  if (S.pb_mixte) {
    int code[5];
#pragma unroll
    for (int o = 0; o < 5; ++o) {
      if (first) {
        code[o] = S.est[cs[o]];
      } else {
        int cnt = 0, pu = -1;
        for (int kk = 0; kk < M; ++kk)
          if (Uin[kk * nc + cs[o]] / den[o] > 0.) {
            cnt++;
            pu = kk;
          }
        code[o] = cnt > 1 ? MH_EST_MIXED : pu;
      }
    }
#pragma unroll
    for (int x = 1; x <= 3; ++x)
      vpure[x] = code[x] != MH_EST_MIXED && code[x - 1] != MH_EST_MIXED && code[x + 1] != MH_EST_MIXED &&
                 code[x] == code[x + 1] && code[x] == code[x - 1];
  }
  int lim[5], limY[5];
#pragma unroll
  for (int x = 0; x < 5; ++x) {
    lim[x] = (S.lim_pure == 1 && vpure[x]) ? S.lim_pure : S.limiter;
    limY[x] = vpure[x] ? S.lim_pure : S.limiter;
  }
  // pente-borne combinations q: (cell -1, front), (cell 0, back), (cell 0, front), (cell +1, back)
  const int qx[4] = {1, 2, 2, 3}, qt[4] = {1, 0, 1, 0};
  double ppv[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int s = q < 2 ? 0 : 1;
    const double vout = qt[q] == 1 ? vf[s] : -vf[s];
    ppv[q] = 0.5 * (vout + fabs(vout)) * dt;
  }
  const bool order2 = S.order2 != 0;
  // phi and limited slopes of variable v at the five stencil cells
  auto load_phi = [&](int v, double* phi) {
#pragma unroll
    for (int o = 0; o < 5; ++o) {
      const long cc = cs[o];
      double p;
      if (!pb) {
        p = Uin[v * nc + cc] / den[o];
      } else if (first || v == 3 * M) {
        p = S.phi0[v * nc + cc];
      } else if (v < M) {
        p = Uin[v * nc + cc] / den[o];
      } else {
        const double d = Uin[(v - M) * nc + cc];
        p = d != 0. ? Uin[v * nc + cc] / d : 0.;
      }
      phi[o] = p;
    }
  };
  auto slopes = [&](const double* phi, double* gr) {
    double gface[4];
#pragma unroll
    for (int f = 0; f < 4; ++f) gface[f] = (valid[f] && valid[f + 1]) ? (phi[f + 1] - phi[f]) / dxf[f] : 0.;
#pragma unroll
    for (int x = 1; x <= 3; ++x)
      gr[x] = order2 ? limited_grad(lim[x], gface[x - 1], gface[x], phi[x + 1], phi[x - 1], hs[x], hs[x + 1], hs[x - 1], sp)
                     : 0.;
  };
  auto pb_raw = [&](const double* phi, const double* gr, int q) {
    const int x = qx[q];
    return pb_flux(qt[q], phi[x], phi[x + 1], phi[x - 1], hs[x], hs[x + 1], hs[x - 1], gr[x], ppv[q], limY[x], thr, sp);
  };
  // pente-borne volume fluxes first: the debar fix and the zero tests need their sums
  double Fv[4][M], Fm[4][M], sfv[4] = {0., 0., 0., 0.};
  bool zero[4] = {true, true, true, true};
  if (pb && order2) {
    for (int v = 0; v < M; ++v) {
      double phi[5], gr[5];
      load_phi(v, phi);
      slopes(phi, gr);
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        Fv[q][v] = inner[q < 2 ? 0 : 1] ? pb_raw(phi, gr, q) : 0.;
        sfv[q] += Fv[q][v];
      }
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) zero[q] = ppv[q] == 0. || (S.debar == 1 && !(fabs(sfv[q]) > thr));
  }
  const int s0 = (DIM == 2 && D == 0) ? 1 : 0, s1 = 1 - s0;
  const double od[2] = {-1., 1.};
  double un[NV];
  for (int v = 0; v < NV; ++v) {
    double phi[5], gr[5];
    double face[2] = {0., 0.};
    if (pb) {
      double F[4] = {0., 0., 0., 0.};
      if (order2) {
        if (v >= M) {
          load_phi(v, phi);
          slopes(phi, gr);
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          if (!inner[q < 2 ? 0 : 1] || zero[q]) continue;
          const int x = qx[q];
          if (v < M) {
            F[q] = Fv[q][v];
          } else if (v < 2 * M) {
            const int e = v - M;
            double f = pb_raw(phi, gr, q);
            if (!vpure[x]) {
              if (S.debar == 1) f = (f / sfv[q]) * Fv[q][e];
              else if (S.debar == 2) f = phi[x] * Fv[q][e];
            }
            F[q] = Fm[q][e] = f;
          } else if (v < 3 * M) {
            const int e = v - 2 * M;
            double f = pb_raw(phi, gr, q);
            if (S.debar == 1) f = (f / sfv[q]) * Fm[q][e];
            else if (S.debar == 2) f = phi[x] * Fm[q][e];
            F[q] = f;
          } else {
            F[q] = pb_raw(phi, gr, q);
          }
        }
      }
      face[0] = F[0] - F[1];
      face[1] = F[2] - F[3];
    } else {
      load_phi(v, phi);
      slopes(phi, gr);
#pragma unroll
      for (int s = 0; s < 2; ++s) {
        const int up = upprev[s] ? s + 1 : s + 2;
        face[s] = phi[up] + hup[s] * gr[up];
      }
    }
    double flux[2] = {0., 0.};
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      if (!inner[s]) continue;
      flux[s] = pb ? od[s] * area[s] * face[s] : od[s] * vf[s] * area[s] * dt * face[s];
    }
    double ff = 0.;
    if (inner[s0]) ff += flux[s0];
    if (inner[s1]) ff += flux[s1];
    un[v] = (own ? Uin[v * nc + c] : 0.) - ff;
    if (v >= M && v < 2 * M) {
      double dd = 0.;
      if (inner[s0]) dd += 0.5 * flux[s0] * od[s0];
      if (inner[s1]) dd += 0.5 * flux[s1] * od[s1];
      S.dflux[(v - M) * nc + c] = dd;
    }
  }
  if (pb && !S.sans_lagrange)  // with the dual phase, tiny partial volumes, masses or energies drop the environment (:441-463)
    for (int kk = 0; kk < M; ++kk) {
      if (un[kk] < 1. * thr) un[kk] = un[M + kk] = un[2 * M + kk] = 0.;
      if (un[M + kk] < 1. * thr) un[kk] = un[M + kk] = un[2 * M + kk] = 0.;
      if (un[2 * M + kk] < 1. * thr) un[kk] = un[M + kk] = un[2 * M + kk] = 0.;
    }
  for (int v = 0; v < NV; ++v) {
    double w = un[v];
    if (v >= M && v < 3 * M && w < 0.) w = 0.;  // computeUremap clamps, :465-485
    Uout[v * nc + c] = w;
  }
}

// The x sweep with one wavefront per row segment: lane l holds cell i = 60 s - 2 + l of a
// row, so the stencil neighbours along x are the neighbouring lanes.  Phi, the gradient of
// the back face, the limited slope and the flux of the back face are formed once per cell
// and passed to the neighbour with a lane shuffle (the front-face value of cell i is the
// back-face value of cell i+1).  Lanes 2..61 write their cell; the two lanes at each end
// only feed the stencil.  Same formulas in the same order as k_sweep_cell_body.
#define MH_ROW_OUT 60
// Order of the cell rows (j, k) of the x sweep: slabs of MH_ROW_KT cell planes, k fastest
// inside a slab, so the four cell rows around a node row (j-1..j, k-1..k) run within
// 2 MH_ROW_KT rows of each other and find that row's coordinates and velocity in L2.
#ifndef MH_ROW_KT
#define MH_ROW_KT 4
#endif
__device__ __forceinline__ void row_jk(const MhGrid& g, long r, int& j, int& k) {
  const long slab = (long)g.ny * MH_ROW_KT;
  const int kt = (int)(r / slab);
  const int h = min(MH_ROW_KT, g.nz - kt * MH_ROW_KT);  // planes of this slab (1 in 2D)
  const long rem = r - kt * slab;
  j = (int)(rem / h);
  k = kt * MH_ROW_KT + (int)(rem % h);
}
template <int DIM, int M, int LIM>
__device__ __forceinline__ void k_sweep_cell_row_body(const MhDev& S, int lb, int usel, int first) {
  constexpr int NV = 3 * M + 1;
  const int nx = S.g.nx;
  const long rows = (long)S.g.ny * S.g.nz;
  const int wpr = (nx + MH_ROW_OUT - 1) / MH_ROW_OUT;
  const int lane = threadIdx.x & 63;
  const long gw = (long)lb * (MH_BLOCK / 64) + (threadIdx.x >> 6);
  if (gw >= rows * wpr) return;  // whole wavefront
  if (mh_halted(S.sc)) return;
  const double* __restrict__ Uin = usel ? S.Ub : S.Ua;
  double* __restrict__ Uout = usel ? S.Ua : S.Ub;
  double* __restrict__ dflux = S.dflux;
  int j, k;
  row_jk(S.g, gw / wpr, j, k);
  const long row = j + (long)S.g.ny * k;
  const int i = (int)(gw % wpr) * MH_ROW_OUT - 2 + lane;
  const bool valid = i >= 0 && i < nx;
  const bool validp = i - 1 >= 0 && i - 1 < nx;
  const bool out = valid && lane >= 2 && lane < 2 + MH_ROW_OUT;
  const long nc = S.g.nc;
  const long c = row * nx + i;
  const bool own = !(first && valid && S.mask[c] == 0);  // an empty cell's U starts from 0
  const long b = i + (long)S.g.nnx * (j + (long)S.g.nny * k);
  const double dt = S.sc->dt;
  double den = 1., xc = 0.;
  if (valid) {
    if (first) {
      den = S.mask[c] ? S.Vl[c] : 1.;  // an empty cell holds Phi itself
    } else {
      double sv = 0.;
      for (int kk = 0; kk < M; ++kk) sv += Uin[kk * nc + c];
      den = sv;
    }
    xc = cell_centre_d<DIM, 0>(S, b);
  }
  const double xcp = __shfl_up(xc, 1);
  const double dxb = (valid && validp) ? xc - xcp : 0.;
  // back face (node column i): inner when 0 < i < nx
  const bool inb = valid && i > 0;
  const bool inf = valid && i < nx - 1;
  double area = 0., xf = 0., vf = 0.;
  if (inb) face_quant<DIM, 0>(S, b, area, xf, vf);
  const bool upprev = vf * dxb > 0.0;
  const double hb = (double)S.order2 * (xf - (upprev ? xcp : xc));
  double unext = valid ? Uin[c] : 0.;  // slot loads run one slot ahead
  for (int v = 0; v < NV; ++v) {
    const double uv = unext;
    if (v + 1 < NV) unext = valid ? Uin[(v + 1) * nc + c] : 0.;
    const double phi = valid ? uv / den : 0.;
    const double phip = __shfl_up(phi, 1);
    const double gb = (valid && validp) ? (phi - phip) / dxb : 0.;
    const double gf = __shfl_down(gb, 1);
    double g = 0.;
    if (S.order2 && valid) {
      if (gb != 0.) g += 0.5 * (mh_limiter<LIM>(S.limiter, gf / gb) * gb);
      if (gf != 0.) g += 0.5 * (mh_limiter<LIM>(S.limiter, gb / gf) * gf);
    }
    const double gp = __shfl_up(g, 1);
    double F = 0.;  // flux of the back face with outer normal +e_x (for cell i-1)
    if (inb) {
      const double pf = (upprev ? phip : phi) + hb * (upprev ? gp : g);
      F = 1. * vf * area * dt * pf;
    }
    const double Ff = __shfl_down(F, 1);
    const double fb = -F;  // the back face seen from cell i (outer normal -e_x)
    double ff = 0.;
    double dd = 0.;
    if (DIM == 2) {  // 2D: the max-x face comes first in the cell's face order
      if (inf) ff += Ff;
      if (inb) ff += fb;
      if (inf) dd += 0.5 * Ff * 1.;
      if (inb) dd += 0.5 * fb * -1.;
    } else {
      if (inb) ff += fb;
      if (inf) ff += Ff;
      if (inb) dd += 0.5 * fb * -1.;
      if (inf) dd += 0.5 * Ff * 1.;
    }
    if (out) {
      double un = (own ? uv : 0.) - ff;
      if (v >= M && v < 2 * M) dflux[(v - M) * nc + c] = dd;
      if (v >= M && v < 3 * M && un < 0.) un = 0.;
      Uout[v * nc + c] = un;
    }
  }
}

// Mass flux through the dual faces between node position a and a+1 along D: the sum, in
// the order of k_sweep_node_body, of the cell mass fluxes of the 2^(d-1) cells at
// position a around the node line (it1, it2) -- the "front" flux of node a and the "back"
// flux of node a+1 (DualPhaseRemap.cc:75-108).  0 when position a holds no cells.
template <int DIM, int D>
__device__ __forceinline__ double dual_edge_flux(const MhDev& S, int a, int it1, int it2) {
  const int na = D == 0 ? S.g.nx : D == 1 ? S.g.ny : S.g.nz;
  if (a < 0 || a >= na) return 0.;
  const double w = DIM == 3 ? 0.25 : 0.5;
  const long nc = S.g.nc;
  const int n1 = D == 0 ? S.g.ny : S.g.nx;
  const int n2 = D == 2 ? S.g.ny : S.g.nz;
  const long st1 = D == 0 ? (long)S.g.nx : 1;
  const long st2 = D == 2 ? (long)S.g.nx : (long)S.g.nx * S.g.ny;
  const long cbase = D == 0 ? (long)a : D == 1 ? (long)a * S.g.nx : (long)a * S.g.nx * S.g.ny;
  const int lo2 = DIM == 3 ? it2 - 1 : 0, hi2 = DIM == 3 ? it2 : 0;
  double f = 0.;
  for (int p2 = lo2; p2 <= hi2; ++p2) {
    if (DIM == 3 && (p2 < 0 || p2 >= n2)) continue;
    for (int p1 = it1 - 1; p1 <= it1; ++p1) {
      if (p1 < 0 || p1 >= n1) continue;
      const long cf = cbase + p1 * st1 + (DIM == 3 ? p2 * st2 : 0);
      for (int kk = 0; kk < S.m; ++kk) f += w * S.dflux[kk * nc + cf];
    }
  }
  return f;
}

// Limited nodal gradient from the edge gradients behind (gb) and ahead (gf)
// (computeAndLimitGradPhiDual, UtilesRemap.cc:369-400), in the order of k_sweep_node_body.
template <int LIM>
__device__ __forceinline__ double dual_limited(const MhDev& S, double gb, double gf) {
  double g = 0.;
  if (gb != 0.) g += 0.5 * mh_limiter<LIM>(S.limiter, gf / gb) * gb;
  if (gf != 0.) g += 0.5 * mh_limiter<LIM>(S.limiter, gb / gf) * gf;
  return g;
}

// Dual update of one node from its neighbourhood along D (DualPhaseRemap.cc:110-257):
// Pm/P0/Pp, Xm/X0/Xp, Gm/G0/Gp are phi, coordinate and limited gradient of the nodes a-1,
// a, a+1; back/front the dual mass fluxes.  Writes Ud, phi and, on the last sweep, u.
template <int DIM, int D>
__device__ __forceinline__ void dual_node_update(const MhDev& S, long n, double* U, const double* Pm,
                                                 const double* P0, const double* Pp, double Xm, double X0,
                                                 double Xp, const double* Gm, const double* G0,
                                                 const double* Gp, double back, double front, double* Pout,
                                                 int last) {
  constexpr int NC = DIM;
  const long nn = S.g.nn;
  const double dt = S.sc->dt;
  U[NC] += back - front;
  const double o2 = (double)S.order2;
  const double ufront = 0.5 * (P0[D] + Pp[D]);
  const int sf = ufront > 0 ? 1 : -1;
  const bool fo = front < 0.;
  const double uback = 0.5 * (Pm[D] + P0[D]);
  const int sb = uback > 0 ? 1 : -1;
  const bool bo = back > 0.;
#pragma unroll
  for (int cc = 0; cc < NC; ++cc) {
    const double FU = (fo ? Pp[cc] : P0[cc]) + o2 * (0.5 * (fo ? Gp[cc] : G0[cc]) * (sf * (Xp - X0) - dt * ufront));
    const double BU = (bo ? Pm[cc] : P0[cc]) + o2 * (0.5 * (bo ? Gm[cc] : G0[cc]) * (sb * (X0 - Xm) - dt * uback));
    U[cc] += back * BU - front * FU;
  }
#pragma unroll
  for (int cc = 0; cc < NC; ++cc)
    if (fabs(U[cc]) < 1.e-300) U[cc] = 0.;
#pragma unroll
  for (int cc = 0; cc <= NC; ++cc) S.Ud[cc * nn + n] = U[cc];
#pragma unroll
  for (int cc = 0; cc < NC; ++cc) {
    const double ph = U[cc] / U[NC];
    Pout[cc * nn + n] = ph;
    if (last && U[NC] != 0.) S.u[cc][n] = ph;
  }
}

// A node with no neighbour along D keeps its dual state (k_sweep_node_body).
template <int DIM>
__device__ __forceinline__ void dual_node_keep(const MhDev& S, long n, const double* U, const double* P,
                                               double* Pout, int last) {
  constexpr int NC = DIM;
  const long nn = S.g.nn;
#pragma unroll
  for (int cc = 0; cc <= NC; ++cc) S.Ud[cc * nn + n] = U[cc];
#pragma unroll
  for (int cc = 0; cc < NC; ++cc) Pout[cc * nn + n] = P[cc];
  if (last && U[NC] != 0.)
#pragma unroll
    for (int cc = 0; cc < NC; ++cc) S.u[cc][n] = U[cc] / U[NC];
}

// Dual x sweep with one wavefront per 60 nodes of a node row (lane shuffles carry the
// neighbours' phi, coordinates, edge gradients and limited gradients; the front flux of
// node i is the back flux of node i+1).  Lanes 2..61 write their node.
template <int DIM, int LIM>
__device__ __forceinline__ void k_sweep_node_row_body(const MhDev& S, int lb, int psel, int first, int last) {
  constexpr int NC = DIM;
  const int nnx = S.g.nnx, na = S.g.nx;
  const long rows = (long)S.g.nny * S.g.nnz;
  const int wpr = (nnx + MH_ROW_OUT - 1) / MH_ROW_OUT;
  const int lane = threadIdx.x & 63;
  const long gw = (long)lb * (MH_BLOCK / 64) + (threadIdx.x >> 6);
  if (gw >= rows * wpr) return;  // whole wavefront
  if (mh_halted(S.sc)) return;
  const double* __restrict__ Pin = psel ? S.Pa : S.Pb;
  double* __restrict__ Pout = psel ? S.Pb : S.Pa;
  const long nn = S.g.nn;
  const long row = gw / wpr;
  const int i = (int)(gw % wpr) * MH_ROW_OUT - 2 + lane;
  const bool valid = i >= 0 && i <= na;
  const bool validp = i - 1 >= 0 && i - 1 <= na;
  const bool out = valid && lane >= 2 && lane < 2 + MH_ROW_OUT;
  const int j = (int)(row % S.g.nny);
  const int k = (int)(row / S.g.nny);
  const long n = row * nnx + i;
  double P[NC], X = 0.;
#pragma unroll
  for (int cc = 0; cc < NC; ++cc) P[cc] = valid ? (first ? S.u[cc][n] : Pin[cc * nn + n]) : 0.;
  if (valid) X = S.xlc[0][n];
  const bool inner = valid && i >= 1 && i <= na - 1;
  const double Xm = __shfl_up(X, 1), Xp = __shfl_down(X, 1);
  double Pm[NC], Pp[NC], G0[NC], Gm[NC], Gp[NC];
#pragma unroll
  for (int cc = 0; cc < NC; ++cc) {
    Pm[cc] = __shfl_up(P[cc], 1);
    Pp[cc] = __shfl_down(P[cc], 1);
    const double eb = (valid && validp) ? (P[cc] - Pm[cc]) / (X - Xm) : 0.;  // edge i-1/2
    const double ef = __shfl_down(eb, 1);                                    // edge i+1/2
    G0[cc] = (S.order2 && inner) ? dual_limited<LIM>(S, eb, ef) : 0.;
    Gm[cc] = __shfl_up(G0[cc], 1);
    Gp[cc] = __shfl_down(G0[cc], 1);
  }
  const double front = valid ? dual_edge_flux<DIM, 0>(S, i, j, k) : 0.;
  const double back = __shfl_up(front, 1);
  if (!out) return;
  double U[NC + 1];
  if (first) {
    const double mass = S.mn[n];
#pragma unroll
    for (int cc = 0; cc < NC; ++cc) U[cc] = mass * S.u[cc][n];
    U[NC] = mass;
  } else {
#pragma unroll
    for (int cc = 0; cc <= NC; ++cc) U[cc] = S.Ud[cc * nn + n];
  }
  if (!inner) {
    dual_node_keep<DIM>(S, n, U, P, Pout, last);
    return;
  }
  dual_node_update<DIM, 0>(S, n, U, Pm, P, Pp, Xm, X, Xp, Gm, G0, Gp, back, front, Pout, last);
}

// The same sweep for D = y or z, LDS tiles: a block of 64 x 4 threads owns 64 cells along
// x by MH_TILE_L cells along D (plus two halo cells on each side).  Every cell's Phi, every
// face gradient, every limited slope and every face flux of the tile is formed once, by
// the thread owning that row, and shared through LDS (four barriers per remap slot).
// Wavefront w holds rows w, w+4, ... of the tile, 64 consecutive cells in x, so every
// global access is coalesced.  Same formulas in the same order as k_sweep_cell_body.
#ifndef MH_TILE_L
#define MH_TILE_L 16  // cells along the sweep direction written per tile (a multiple of 4)
#endif
#define MH_TILE_R (MH_TILE_L + 4)  // rows held, halos included
#ifndef MH_TILE_3SYNC
#define MH_TILE_3SYNC 1  // one environment: three barriers per remap slot instead of four
#endif
#define MH_TILE_Q (MH_TILE_R / 4)  // rows per thread
#ifndef MH_TILE_MIN_WAVES
#define MH_TILE_MIN_WAVES 4  // waves per SIMD the register allocation must allow
#endif
template <int DIM, int D, int M, int LIM>
__device__ __forceinline__ void k_sweep_cell_tile_body(const MhDev& S, int lb_, int usel, int first) {
  constexpr int NV = 3 * M + 1;
  __shared__ double sphi[MH_TILE_R][64];
  __shared__ double sgf[MH_TILE_R][64];  // face gradients, then face fluxes
  __shared__ double ssl[MH_TILE_R][64];  // cell centres, then limited slopes
  const int nx = S.g.nx, ny = S.g.ny, nz = S.g.nz;
  const int na = D == 1 ? ny : nz;
  const int nt = (D == 1) ? (DIM == 3 ? nz : 1) : ny;
  const int tilesx = (nx + 63) / 64;
  const long ntiles = (long)tilesx * nt * ((na + MH_TILE_L - 1) / MH_TILE_L);
  const long lb = lb_;
  if (lb >= ntiles) return;  // whole block
  if (mh_halted(S.sc)) return;
  const double* __restrict__ Uin = usel ? S.Ub : S.Ua;
  double* __restrict__ Uout = usel ? S.Ua : S.Ub;
  double* __restrict__ dflux = S.dflux;
  const int lx = threadIdx.x & 63, ly = threadIdx.x >> 6;
  // tiles along D vary fastest: the tiles sharing a halo run side by side on one XCD
  const int ntd = (na + MH_TILE_L - 1) / MH_TILE_L;
  const int abase = (int)(lb % ntd) * MH_TILE_L - 2;  // cell along D of tile row 0
  const long rest = lb / ntd;
  const int i = (int)(rest % tilesx) * 64 + lx;
  const int t = (int)(rest / tilesx);
  const long nc = S.g.nc;
  const long nnx = S.g.nnx, nny = S.g.nny;
  const long sD = D == 1 ? (long)nx : (long)nx * ny;
  const long c_base = D == 1 ? i + (long)nx * ny * t : i + (long)nx * t;
  const long nD = D == 1 ? nnx : nnx * nny;
  const long n_base = D == 1 ? i + nnx * nny * t : i + nnx * t;
  const double dt = S.sc->dt;
  const double o2 = (double)S.order2;
  // rows r = ly + 4q of this thread: cell validity, Phi denominator, Delta x of face r
  // (between rows r-1 and r) and the upwind data of that face
  bool val[MH_TILE_Q], inner[MH_TILE_Q], up[MH_TILE_Q];
  double den[MH_TILE_Q], dx[MH_TILE_Q], w[MH_TILE_Q], hh[MH_TILE_Q], xcq[MH_TILE_Q];
  // one environment: the face gradients are formed where the slopes are (three barriers per
  // slot); with more, the registers this takes cost more than the barrier saves
  constexpr bool THREE = MH_TILE_3SYNC != 0 && M == 1;
  bool innern[THREE ? MH_TILE_Q : 1];  // face r+1 (rows r and r+1) and its Delta x
  double dxn[THREE ? MH_TILE_Q : 1];
  unsigned ownm = 0;  // bit q: the cell's own U enters its update
#pragma unroll
  for (int q = 0; q < MH_TILE_Q; ++q) {
    const int a = abase + ly + 4 * q;
    val[q] = i < nx && a >= 0 && a < na;
    den[q] = 1.;
    xcq[q] = 0.;
    if (val[q]) {
      const long cc = c_base + a * sD;
      if (first) {
        const bool empty = !S.mask[cc];  // an empty cell holds Phi itself and its U starts from 0
        den[q] = empty ? 1. : S.Vl[cc];
        if (!empty) ownm |= 1u << q;
      } else {
        ownm |= 1u << q;
        double sv = 0.;
        for (int k = 0; k < M; ++k) sv += Uin[k * nc + cc];
        den[q] = sv;
      }
      xcq[q] = cell_centre_d<DIM, D>(S, n_base + a * nD);
    }
    ssl[ly + 4 * q][lx] = xcq[q];
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < MH_TILE_Q; ++q) {
    const int r = ly + 4 * q, a = abase + r;
    const bool valp = r >= 1 && i < nx && a - 1 >= 0 && a - 1 < na;
    const double xcp = r >= 1 ? ssl[r - 1][lx] : 0.;
    dx[q] = (val[q] && valp) ? xcq[q] - xcp : 0.;
    inner[q] = val[q] && valp;  // face between cells a-1 and a
    if constexpr (THREE) {  // face r+1 as the thread of row r+1 forms it
      const bool valn = r + 1 < MH_TILE_R && i < nx && a + 1 >= 0 && a + 1 < na;
      innern[q] = valn && val[q];
      dxn[q] = innern[q] ? ssl[r + 1][lx] - xcq[q] : 0.;
    }
    up[q] = false;
    w[q] = 0.;
    hh[q] = 0.;
    if (inner[q] && r >= 2 && r <= MH_TILE_R - 2) {
      double area, xf, vf;
      face_quant<DIM, D>(S, n_base + a * nD, area, xf, vf);
      up[q] = vf * dx[q] > 0.0;
      hh[q] = o2 * (xf - (up[q] ? xcp : xcq[q]));
      w[q] = 1. * vf * area * dt;
    }
  }
  double un_next[MH_TILE_Q];  // slot loads run one slot ahead
#pragma unroll
  for (int q = 0; q < MH_TILE_Q; ++q) un_next[q] = val[q] ? Uin[c_base + (abase + ly + 4 * q) * sD] : 0.;
  for (int v = 0; v < NV; ++v) {
    // barriers: the previous slot read sphi and ssl before its last barrier, and reads
    // sgf (fluxes) before the first barrier of this slot
    double uq[MH_TILE_Q];
#pragma unroll
    for (int q = 0; q < MH_TILE_Q; ++q) {
      const int a = abase + ly + 4 * q;
      uq[q] = un_next[q];
      if (v + 1 < NV) un_next[q] = val[q] ? Uin[(v + 1) * nc + c_base + a * sD] : 0.;
      sphi[ly + 4 * q][lx] = val[q] ? uq[q] / den[q] : 0.;
    }
    __syncthreads();
    if constexpr (!THREE) {
#pragma unroll
      for (int q = 0; q < MH_TILE_Q; ++q) {
        const int r = ly + 4 * q;
        sgf[r][lx] = (r >= 1 && inner[q]) ? (sphi[r][lx] - sphi[r - 1][lx]) / dx[q] : 0.;
      }
      __syncthreads();
    }
#pragma unroll
    for (int q = 0; q < MH_TILE_Q; ++q) {
      const int r = ly + 4 * q;
      double g = 0.;
      if (S.order2 && val[q] && r >= 1 && r <= MH_TILE_R - 2) {
        double gb, gf;
        if constexpr (THREE) {  // both face gradients of the row from Phi (the front one is also
                                // formed by the next row's thread, with the same operands)
          gb = inner[q] ? (sphi[r][lx] - sphi[r - 1][lx]) / dx[q] : 0.;
          gf = innern[q] ? (sphi[r + 1][lx] - sphi[r][lx]) / dxn[q] : 0.;
        } else {
          gb = sgf[r][lx];
          gf = sgf[r + 1][lx];
        }
        if (gb != 0.) g += 0.5 * (mh_limiter<LIM>(S.limiter, gf / gb) * gb);
        if (gf != 0.) g += 0.5 * (mh_limiter<LIM>(S.limiter, gb / gf) * gf);
      }
      ssl[r][lx] = g;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < MH_TILE_Q; ++q) {  // the gradients were all read before the barrier
      const int r = ly + 4 * q;
      double f = 0.;
      if (inner[q] && r >= 2 && r <= MH_TILE_R - 2) {
        const double pf = up[q] ? sphi[r - 1][lx] + hh[q] * ssl[r - 1][lx] : sphi[r][lx] + hh[q] * ssl[r][lx];
        f = w[q] * pf;
      }
      sgf[r][lx] = f;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < MH_TILE_Q; ++q) {
      const int r = ly + 4 * q, a = abase + r;
      if (!val[q] || r < 2 || r >= 2 + MH_TILE_L) continue;
      const bool inb = a > 0, inf = a < na - 1;
      const double fb = -sgf[r][lx];
      const double Ff = sgf[r + 1][lx];
      double ff = 0.;
      if (inb) ff += fb;
      if (inf) ff += Ff;
      const long c = c_base + a * sD;
      double un = (((ownm >> q) & 1u) ? uq[q] : 0.) - ff;
      if (v >= M && v < 2 * M) {
        double dd = 0.;
        if (inb) dd += 0.5 * fb * -1.;
        if (inf) dd += 0.5 * Ff * 1.;
        dflux[(v - M) * nc + c] = dd;
      }
      if (v >= M && v < 3 * M && un < 0.) un = 0.;
      Uout[v * nc + c] = un;
    }
  }
}

// One ADI direction, dual part for one node: limited nodal velocity gradients
// (computeDualGradPhi UtilesRemap.cc:19-72, computeAndLimitGradPhiDual :369-400), back/front
// mass fluxes gathered from the cells behind/ahead in face order, env order
// (DualPhaseRemap.cc:75-108), upwind momentum update (:110-239), filter and phi (:241-257).
// Nodes with no neighbour along D keep their values.  On the last sweep the node velocity
// u = mu / m is formed for every node (RemapADIFinal.cc:286-297).
// The mass fluxes through an inner node's back and front dual faces along D
// (DualPhaseRemap.cc:75-108): 1/2 (2D) or 1/4 (3D) of the dual mass fluxes of the cells
// behind / ahead of the node, over those cells in face order and the environments in order.
template <int DIM, int D, bool BACK = true>
__device__ __forceinline__ void dual_mass_fluxes(const MhDev& S, int i, int j, int k, int a, double& back,
                                                 double& front) {
  const double* __restrict__ dflux = S.dflux;
  const long nc = S.g.nc;
  const double w = DIM == 3 ? 0.25 : 0.5;
  const int M = S.m;
  back = 0.;
  front = 0.;
  const long sD = D == 0 ? 1 : D == 1 ? (long)S.g.nx : (long)S.g.nx * S.g.ny;
  // transverse axes (t1 fastest, t2 slowest) and their cell counts
  const int t1 = D == 0 ? 1 : 0;
  const int t2 = D == 2 ? 1 : 2;
  const int it1 = t1 == 0 ? i : j;
  const int it2 = t2 == 1 ? j : k;
  const int n1 = t1 == 0 ? S.g.nx : S.g.ny;
  const int n2 = t2 == 1 ? S.g.ny : S.g.nz;
  const long st1 = t1 == 0 ? 1 : (long)S.g.nx;
  const long st2 = t2 == 1 ? (long)S.g.nx : (long)S.g.nx * S.g.ny;
  const int lo2 = DIM == 3 ? it2 - 1 : 0, hi2 = DIM == 3 ? it2 : 0;
  // cell at position a along D and (p1, p2) along the transverse axes
  const long cbase = D == 0 ? (long)a : D == 1 ? (long)a * S.g.nx : (long)a * S.g.nx * S.g.ny;
  for (int p2 = lo2; p2 <= hi2; ++p2) {
    if (DIM == 3 && (p2 < 0 || p2 >= n2)) continue;
    for (int p1 = it1 - 1; p1 <= it1; ++p1) {
      if (p1 < 0 || p1 >= n1) continue;
      const long cf = cbase + p1 * st1 + (DIM == 3 ? p2 * st2 : 0);
      const long cb = cf - sD;
      if (BACK)
        for (int kk = 0; kk < M; ++kk) back += w * dflux[kk * nc + cb];
      for (int kk = 0; kk < M; ++kk) front += w * dflux[kk * nc + cf];
    }
  }
}

template <int DIM, int D, int LIM>
__device__ __forceinline__ void k_sweep_node_at(const MhDev& S, long n, int psel, int first, int last) {
  double* __restrict__ Ud = S.Ud;
  const double* __restrict__ Pin = psel ? S.Pa : S.Pb;
  double* __restrict__ Pout = psel ? S.Pb : S.Pa;
  constexpr int NC = DIM;  // velocity components carried
  if (mh_halted(S.sc)) return;
  const int nnx = S.g.nnx, nny = S.g.nny;
  const int i = (int)(n % nnx);
  const long r = n / nnx;
  const int j = (int)(r % nny);
  const int k = (int)(r / nny);
  const int a = D == 0 ? i : D == 1 ? j : k;
  const int na = D == 0 ? S.g.nx : D == 1 ? S.g.ny : S.g.nz;
  const long nD = D == 0 ? 1 : D == 1 ? (long)nnx : (long)nnx * nny;
  const long nn = S.g.nn;
  double U[NC + 1];
  if (first) {
    const double mass = S.mn[n];
    for (int cc = 0; cc < NC; ++cc) U[cc] = mass * S.u[cc][n];
    U[NC] = mass;
  } else {
    for (int cc = 0; cc <= NC; ++cc) U[cc] = Ud[cc * nn + n];
  }
  const bool inner = a >= 1 && a <= na - 1;
  if (!inner) {
    for (int cc = 0; cc <= NC; ++cc) Ud[cc * nn + n] = U[cc];
    for (int cc = 0; cc < NC; ++cc) Pout[cc * nn + n] = first ? S.u[cc][n] : Pin[cc * nn + n];
    if (last && U[NC] != 0.)
      for (int cc = 0; cc < NC; ++cc) S.u[cc][n] = U[cc] / U[NC];
    return;
  }
  const double dt = S.sc->dt;
  double P[5][NC], X[5];
#pragma unroll
  for (int o = -2; o <= 2; ++o) {
    const int bb = a + o;
    if (bb >= 0 && bb <= na) {
      const long m2 = n + o * nD;
      for (int cc = 0; cc < NC; ++cc) P[o + 2][cc] = first ? S.u[cc][m2] : Pin[cc * nn + m2];
      X[o + 2] = S.xlc[D][m2];
    } else {
      for (int cc = 0; cc < NC; ++cc) P[o + 2][cc] = 0.;
      X[o + 2] = 0.;
    }
  }
  double G[3][NC];
  // coordinates two and three nodes away, for the dual widths of generalised limiters
  double Xmm = 0., Xpp = 0.;
  if (LIM == MH_LIM_ANY) {
    if (a - 3 >= 0) Xmm = S.xlc[D][n - 3 * nD];
    if (a + 3 <= na) Xpp = S.xlc[D][n + 3 * nD];
  }
#pragma unroll
  for (int o = -1; o <= 1; ++o) {
    const int bb = a + o;
    const bool in = S.order2 && bb >= 1 && bb <= na - 1;
    double h0 = 0., hm = 0., hp = 0.;
    if (LIM == MH_LIM_ANY && in) {  // dual cell widths (UtilesRemap.cc:50-61)
      const double xbb = o == -1 ? Xmm : X[o], xff = o == 1 ? Xpp : X[o + 4];
      h0 = 0.5 * (X[o + 3] - X[o + 1]);
      if (bb - 2 < 0) {
        if (bb + 2 <= na) hp = 0.5 * (xff - X[o + 2]);
      } else if (bb + 2 > na) {
        hm = 0.5 * (X[o + 2] - xbb);
      } else {
        hm = 0.5 * (X[o + 2] - xbb);
        hp = 0.5 * (xff - X[o + 2]);
      }
    }
    for (int cc = 0; cc < NC; ++cc) {
      double g = 0.;
      if (in) {
        const double gf = (P[o + 3][cc] - P[o + 2][cc]) / (X[o + 3] - X[o + 2]);
        const double gb = (P[o + 2][cc] - P[o + 1][cc]) / (X[o + 2] - X[o + 1]);
        if (LIM == MH_LIM_ANY && S.limiter >= MH_MINMOD_G) {
          g = flux_limiter_g(S.limiter, gf, gb, P[o + 3][cc], P[o + 1][cc], h0, hp, hm, S.simple_pente);
        } else {
          if (gb != 0.) g += 0.5 * mh_limiter<LIM>(S.limiter, gf / gb) * gb;
          if (gf != 0.) g += 0.5 * mh_limiter<LIM>(S.limiter, gb / gf) * gf;
        }
      }
      G[o + 1][cc] = g;
    }
  }
  double back, front;
  dual_mass_fluxes<DIM, D>(S, i, j, k, a, back, front);
  U[NC] += back - front;
  const double o2 = (double)S.order2;
  const double ufront = 0.5 * (P[2][D] + P[3][D]);
  const int sf = ufront > 0 ? 1 : -1;
  const int fo = (front < 0.) ? 1 : 0;
  const double uback = 0.5 * (P[1][D] + P[2][D]);
  const int sb = uback > 0 ? 1 : -1;
  const int bo = (back > 0.) ? -1 : 0;
  for (int cc = 0; cc < NC; ++cc) {
    const double FU = P[fo + 2][cc] + o2 * (0.5 * G[fo + 1][cc] * (sf * (X[3] - X[2]) - dt * ufront));
    const double BU = P[bo + 2][cc] + o2 * (0.5 * G[bo + 1][cc] * (sb * (X[2] - X[1]) - dt * uback));
    U[cc] += back * BU - front * FU;
  }
  for (int cc = 0; cc < NC; ++cc)
    if (fabs(U[cc]) < 1.e-300) U[cc] = 0.;
  for (int cc = 0; cc <= NC; ++cc) Ud[cc * nn + n] = U[cc];
  for (int cc = 0; cc < NC; ++cc) {
    const double ph = U[cc] / U[NC];
    Pout[cc * nn + n] = ph;
    if (last && U[NC] != 0.) S.u[cc][n] = ph;
  }
}

template <int DIM, int D, int LIM>
__device__ __forceinline__ void k_sweep_node_body(const MhDev& S, int lb, int psel, int first, int last) {
  const long n = lb * (long)blockDim.x + threadIdx.x;
  if (n >= S.g.nn) return;
  k_sweep_node_at<DIM, D, LIM>(S, n, psel, first, last);
}

// The y or z dual sweep of one mesh with each thread marching along D over a segment of
// MH_NM_L nodes: the window of phi and coordinates, the edge gradients, the limited gradients
// and the mass flux through the node's front dual face are carried to the next node (they are
// its back-side values), so each is formed once per node instead of by each of the three nodes
// whose update reads it.  Threads of a wave are consecutive along x.  Same formulas and orders
// as k_sweep_node_at (classical limiters; the generalised ones keep k_sweep_node).  The segment
// length is per case (the autotuner tries 4-12 nodes): 10 by default from 2.6 M nodes on (fastest
// on the 256x128x128 tube and the 192^3 Sedov blast), 4 on smaller meshes, which need the threads.
#ifndef MH_NM_L
#define MH_NM_L 10  // default target segment length (MhDev::nm_len; variant march code 0, kMarchLen[0])
#endif
// Segments of one node line: the n nodes split into ceil(n / len) segments whose lengths differ
// by at most one, so that no line ends in a short segment that costs a whole pass of threads.
MH_HD int march_nseg(int n, int len) { return (n + len - 1) / len; }
#ifndef MH_NM_MIN_WAVES
#define MH_NM_MIN_WAVES 1
#endif
template <int DIM, int D, int LIM>
__device__ __forceinline__ void k_sweep_node_march_body(const MhDev& S, int lb, int psel, int first, int last) {
  constexpr int NC = DIM;
  const int nnx = S.g.nnx, nny = S.g.nny;
  const int na = D == 1 ? S.g.ny : S.g.nz;
  const int nt2 = D == 1 ? S.g.nnz : nny;  // the other transverse node count
  const long cols = (long)nnx * nt2;
  const long g = lb * (long)blockDim.x + threadIdx.x;
  const int nseg = march_nseg(na + 1, S.nm_len);  // na + 1 nodes along D
  if (g >= cols * nseg) return;
  if (mh_halted(S.sc)) return;
  const int seg = (int)(g / cols);
  const long rem = g - seg * cols;
  const int i = (int)(rem % nnx), t2 = (int)(rem / nnx);
  const int a0 = (int)((long)seg * (na + 1) / nseg), a1 = (int)((long)(seg + 1) * (na + 1) / nseg);  // nodes a0 .. a1-1
  const long nn = S.g.nn;
  const long nD = D == 1 ? (long)nnx : (long)nnx * nny;
  const long base = D == 1 ? i + (long)nnx * nny * t2 : i + (long)nnx * t2;  // node at a = 0
  const double* __restrict__ Pin = psel ? S.Pa : S.Pb;
  double* __restrict__ Pout = psel ? S.Pb : S.Pa;
  double* __restrict__ Ud = S.Ud;
  const double* __restrict__ Xd = S.xlc[D];
  const double dt = S.sc->dt;
  const double o2 = (double)S.order2;
  auto loadP = [&](int b, double* P, double& X) {
    if (b >= 0 && b <= na) {
      const long m2 = base + b * nD;
#pragma unroll
      for (int cc = 0; cc < NC; ++cc) P[cc] = first ? S.u[cc][m2] : Pin[cc * nn + m2];
      X = Xd[m2];
    } else {
#pragma unroll
      for (int cc = 0; cc < NC; ++cc) P[cc] = 0.;
      X = 0.;
    }
  };
  // limited gradient of node b from the edge gradients behind (gb) and ahead (gf)
  auto limited = [&](int b, const double* gb, const double* gf, double* G) {
    const bool in = S.order2 && b >= 1 && b <= na - 1;
#pragma unroll
    for (int cc = 0; cc < NC; ++cc) {
      double gg = 0.;
      if (in) {
        if (gb[cc] != 0.) gg += 0.5 * mh_limiter<LIM>(S.limiter, gf[cc] / gb[cc]) * gb[cc];
        if (gf[cc] != 0.) gg += 0.5 * mh_limiter<LIM>(S.limiter, gb[cc] / gf[cc]) * gf[cc];
      }
      G[cc] = gg;
    }
  };
  // edge gradient between nodes b and b+1 (formed only where a limited gradient reads it)
  auto edge = [&](int b, const double* P0, double X0, const double* P1, double X1, double* E) {
    const bool used = S.order2 && b >= 0 && b + 1 <= na && (b >= 1 || b + 1 <= na - 1);
#pragma unroll
    for (int cc = 0; cc < NC; ++cc) E[cc] = used ? (P1[cc] - P0[cc]) / (X1 - X0) : 0.;
  };
  // window: P and X of nodes a-1 .. a+2 (slots 0..3), edges (a-1/2, a+1/2, a+3/2), G of a-1, a, a+1
  double P[4][NC], X[4], E[3][NC], G[3][NC];
  {
    double Pm2[NC], Xm2;
    loadP(a0 - 2, Pm2, Xm2);
    for (int o = 0; o < 4; ++o) loadP(a0 - 1 + o, P[o], X[o]);
    double Em[NC];  // edge a0 - 3/2
    edge(a0 - 2, Pm2, Xm2, P[0], X[0], Em);
    for (int o = 0; o < 3; ++o) edge(a0 - 1 + o, P[o], X[o], P[o + 1], X[o + 1], E[o]);
    limited(a0 - 1, Em, E[0], G[0]);
    limited(a0, E[0], E[1], G[1]);
    limited(a0 + 1, E[1], E[2], G[2]);
  }
  const int j0 = D == 1 ? 0 : t2, k0 = D == 1 ? t2 : 0;  // transverse node position (the D one is a)
  double back = 0., front = 0.;
  for (int a = a0; a < a1; ++a) {
    const long n = base + a * nD;
    double U[NC + 1];
    if (first) {
      const double mass = S.mn[n];
#pragma unroll
      for (int cc = 0; cc < NC; ++cc) U[cc] = mass * S.u[cc][n];
      U[NC] = mass;
    } else {
#pragma unroll
      for (int cc = 0; cc <= NC; ++cc) U[cc] = Ud[cc * nn + n];
    }
    if (!(a >= 1 && a <= na - 1)) {
      dual_node_keep<DIM>(S, n, U, P[1], Pout, last);
    } else {
      const int jj = D == 1 ? a : j0, kk = D == 1 ? k0 : a;
      if (a > a0 && a >= 2) {  // the back face of node a is the front face of inner node a-1
        back = front;
        double unused = 0.;
        dual_mass_fluxes<DIM, D, false>(S, i, jj, kk, a, unused, front);
      } else {
        dual_mass_fluxes<DIM, D>(S, i, jj, kk, a, back, front);
      }
      U[NC] += back - front;
      const double ufront = 0.5 * (P[1][D] + P[2][D]);
      const int sf = ufront > 0 ? 1 : -1;
      const bool fo = front < 0.;
      const double uback = 0.5 * (P[0][D] + P[1][D]);
      const int sb = uback > 0 ? 1 : -1;
      const bool bo = back > 0.;
#pragma unroll
      for (int cc = 0; cc < NC; ++cc) {
        const double FU = (fo ? P[2][cc] : P[1][cc]) + o2 * (0.5 * (fo ? G[2][cc] : G[1][cc]) * (sf * (X[2] - X[1]) - dt * ufront));
        const double BU = (bo ? P[0][cc] : P[1][cc]) + o2 * (0.5 * (bo ? G[0][cc] : G[1][cc]) * (sb * (X[1] - X[0]) - dt * uback));
        U[cc] += back * BU - front * FU;
      }
#pragma unroll
      for (int cc = 0; cc < NC; ++cc)
        if (fabs(U[cc]) < 1.e-300) U[cc] = 0.;
#pragma unroll
      for (int cc = 0; cc <= NC; ++cc) Ud[cc * nn + n] = U[cc];
#pragma unroll
      for (int cc = 0; cc < NC; ++cc) {
        const double ph = U[cc] / U[NC];
        Pout[cc * nn + n] = ph;
        if (last && U[NC] != 0.) S.u[cc][n] = ph;
      }
    }
    if (a + 1 < a1) {  // slide the window one node along D
#pragma unroll
      for (int o = 0; o < 3; ++o) {
        X[o] = X[o + 1];
#pragma unroll
        for (int cc = 0; cc < NC; ++cc) P[o][cc] = P[o + 1][cc];
      }
      loadP(a + 3, P[3], X[3]);
#pragma unroll
      for (int cc = 0; cc < NC; ++cc) {
        E[0][cc] = E[1][cc];
        E[1][cc] = E[2][cc];
        G[0][cc] = G[1][cc];
        G[1][cc] = G[2][cc];
      }
      edge(a + 2, P[2], X[2], P[3], X[3], E[2]);
      limited(a + 2, E[1], E[2], G[2]);
    }
  }
}


// u = mu / m for every node with a non-zero projected mass (RemapADIFinal.cc:286-297), for
// cycles whose remap has a single sweep (the sweep kernel still reads u) or none.
template <int DIM>
__device__ __forceinline__ void k_node_velocity_body(const MhDev& S, int lb, int materialize) {
  const double* __restrict__ Ud = S.Ud;
  const long n = lb * (long)blockDim.x + threadIdx.x;
  if (n >= S.g.nn) return;
  if (mh_halted(S.sc)) return;
  const long nn = S.g.nn;
  double U[DIM + 1];
  if (materialize) {
    const double mass = S.mn[n];
    for (int cc = 0; cc < DIM; ++cc) U[cc] = mass * S.u[cc][n];
    U[DIM] = mass;
  } else {
    for (int cc = 0; cc <= DIM; ++cc) U[cc] = Ud[cc * nn + n];
  }
  if (U[DIM] != 0.)
    for (int cc = 0; cc < DIM; ++cc) S.u[cc][n] = U[cc] / U[DIM];
}

// ------------------------------------------------- conservation-energie-totale (RemapADI.axl:178)
// The dual kinetic energy of one sweep along D, run after the node sweep of that direction:
// dual U = 1/2 m |u|^2 and phi = 1/2 |u|^2 at the first sweep (PrepareRemap.cc:179-190), then
// for inner nodes the upwind phi through the back and front dual faces with the node sweep's
// mass fluxes and upwind choices (DualPhaseRemap.cc:132-238).  Its limited gradient is never
// formed (UtilesRemap.cc:369-400 covers the velocity only), so the slope term multiplies 0.
// phi = U / the node mass the node sweep has just written (:251-257).
template <int DIM, int D>
__device__ __forceinline__ void k_sweep_ke_body(const MhDev& S, int lb, int psel, int first) {
  const long n = lb * (long)blockDim.x + threadIdx.x;
  if (n >= S.g.nn) return;
  if (mh_halted(S.sc)) return;
  const double* __restrict__ Pin = psel ? S.Pa : S.Pb;
  const double* __restrict__ Kin = psel ? S.ke_Pa : S.ke_Pb;
  double* __restrict__ Kout = psel ? S.ke_Pb : S.ke_Pa;
  const long nn = S.g.nn;
  const int nnx = S.g.nnx, nny = S.g.nny;
  const int i = (int)(n % nnx);
  const long r = n / nnx;
  const int j = (int)(r % nny);
  const int k = (int)(r / nny);
  const int a = D == 0 ? i : D == 1 ? j : k;
  const int na = D == 0 ? S.g.nx : D == 1 ? S.g.ny : S.g.nz;
  const long nD = D == 0 ? 1 : D == 1 ? (long)nnx : (long)nnx * nny;
  auto sq = [&](long m2) {
    const double ux = S.u[0][m2], uy = S.u[1][m2];
    double v = ux * ux + uy * uy;
    if (DIM == 3) v = v + S.u[2][m2] * S.u[2][m2];
    return v;
  };
  auto phi = [&](long m2) { return first ? 0.5 * sq(m2) : Kin[m2]; };
  double U = first ? 0.5 * S.mn[n] * sq(n) : S.ke_U[n];
  if (!(a >= 1 && a <= na - 1)) {
    S.ke_U[n] = U;
    Kout[n] = phi(n);
    return;
  }
  double back, front;
  dual_mass_fluxes<DIM, D>(S, i, j, k, a, back, front);
  const double dt = S.sc->dt;
  const double o2 = (double)S.order2;
  const double zero_slope = 0.;
  auto vel = [&](long m2) { return first ? S.u[D][m2] : Pin[D * nn + m2]; };
  const double xb = S.xlc[D][n - nD], xn = S.xlc[D][n], xf = S.xlc[D][n + nD];
  const double ufront = 0.5 * (vel(n) + vel(n + nD));
  const int sf = ufront > 0 ? 1 : -1;
  const double FE = phi(front < 0. ? n + nD : n) + o2 * (0.5 * zero_slope * (sf * (xf - xn) - dt * ufront));
  const double uback = 0.5 * (vel(n - nD) + vel(n));
  const int sb = uback > 0 ? 1 : -1;
  const double BE = phi(back > 0. ? n - nD : n) + o2 * (0.5 * zero_slope * (sb * (xn - xb) - dt * uback));
  U += back * BE - front * FE;
  S.ke_U[n] = U;
  Kout[n] = U / S.Ud[DIM * nn + n];
}

// The dual state of a remap with no sweep (every direction one cell wide): U = (m u, m) and
// the kinetic energy 1/2 m |u|^2, untouched by any sweep (PrepareRemap.cc:169-190).
template <int DIM>
__device__ __forceinline__ void k_dual_init_body(const MhDev& S, int lb) {
  const long n = lb * (long)blockDim.x + threadIdx.x;
  if (n >= S.g.nn) return;
  if (mh_halted(S.sc)) return;
  const long nn = S.g.nn;
  const double mass = S.mn[n];
  double v = 0.;
  for (int cc = 0; cc < DIM; ++cc) {
    const double uc = S.u[cc][n];
    S.Ud[cc * nn + n] = mass * uc;
    v = cc == 0 ? uc * uc : v + uc * uc;
  }
  S.Ud[DIM * nn + n] = mass;
  S.ke_U[n] = 0.5 * mass * v;
}

// The kinetic energy each cell gets back (RemapADIFinal.cc:318-339), once the node velocity
// is final: over the cell's nodes in cell order, 0.25 (also in 3D) times the projected
// minus the reconstructed kinetic energy per unit mass, for nodes with a non-zero projected
// mass; kept if positive (std::max(0., .)).  k_finalize adds it to the internal energies.
template <int DIM>
__device__ __forceinline__ void k_cons_ec_body(const MhDev& S, int lb) {
  const long c = lb * (long)blockDim.x + threadIdx.x;
  if (c >= S.g.nc) return;
  if (mh_halted(S.sc)) return;
  const long nn = S.g.nn;
  const int i = (int)(c % S.g.nx);
  const long r = c / S.g.nx;
  const int j = (int)(r % S.g.ny);
  const int k = (int)(r / S.g.ny);
  const long nnx = S.g.nnx, nnxy = (long)S.g.nnx * S.g.nny;
  const long b = i + nnx * (j + (long)S.g.nny * k);
  const long off[8] = {0, 1, 1 + nnx, nnx, nnxy, 1 + nnxy, 1 + nnx + nnxy, nnx + nnxy};
  double d = 0.;
#pragma unroll
  for (int l = 0; l < (DIM == 3 ? 8 : 4); ++l) {
    const long n = b + off[l];
    const double mp = S.Ud[DIM * nn + n];
    if (mp != 0.) {
      const double ux = S.u[0][n], uy = S.u[1][n];
      double v = ux * ux + uy * uy;
      if (DIM == 3) v = v + S.u[2][n] * S.u[2][n];
      const double ec_proj = S.ke_U[n] / mp;
      const double ec_reconst = 0.5 * v;
      d += 0.25 * (ec_proj - ec_reconst);
    }
  }
  S.dec[c] = (0. < d) ? d : 0.;
}

// ----------------------------------------------------------------------------- finalize
// remapVariables (RemapADIFinal.cc:21-316), the remap tail (PrepareRemap.cc:220-249:
// applyEOS per environment, computePressionMoyenne) and the per-cell CFL candidate of
// computeDeltaT (MahycoModule.cc:1763-1779).  Pure cells write through the global arrays
// only (their environment value is the global value).
template <int DIM, int M>
__device__ __forceinline__ void k_finalize_body(const MhDev& S, int lb, int usel) {
  const double* __restrict__ Uin = usel ? S.Ub : S.Ua;
  const long c = lb * (long)blockDim.x + threadIdx.x;
  const bool active = c < S.g.nc && !mh_halted(S.sc);
  double dxs = 1.79769313486231570815e+308;
  if (active) {
    const long nc = S.g.nc;
    const double thr = S.thr_remap;
    double U[3 * M + 1];
    for (int v = 0; v <= 3 * M; ++v) U[v] = Uin[v * nc + c];
    if (S.pb && S.nsweep > 0) {  // phi the last sweep left (RemapADIService.cc:493-538): what the
                                  // next cycle's first sweep sees for non-member environments
      double sv = 0.;
      for (int kk = 0; kk < M; ++kk) sv += U[kk];
      for (int kk = 0; kk < M; ++kk) {
        S.phi0[kk * nc + c] = U[kk] / sv;
        S.phi0[(M + kk) * nc + c] = U[kk] != 0. ? U[M + kk] / U[kk] : 0.;
      }
      for (int kk = 0; kk < M; ++kk) S.phi0[(2 * M + kk) * nc + c] = U[M + kk] != 0. ? U[2 * M + kk] / U[M + kk] : 0.;
    }
    const double VE = S.VE[c];
    // environment membership (:29-78)
    const uint8_t om = S.mask[c];
    uint8_t nm = om;
    for (int kk = 0; kk < M; ++kk) {
      const bool was = (om >> kk) & 1;
      if (((U[kk] / VE) < thr || (U[M + kk] == 0. && !S.sans_lagrange)) && was) {
        nm &= (uint8_t)~(1u << kk);
      } else if ((U[kk] / VE) > thr && !was) {
        double volt = 0.;
        for (int k2 = 0; k2 < M; ++k2) volt += U[k2];
        const double vol_ev = U[kk] * VE / volt;
        if ((vol_ev / VE) > thr && U[M + kk] != 0.) nm |= (uint8_t)(1u << kk);
      }
    }
    const bool pure = __popc(nm) == 1;
    // (:85-216)
    const double vol = VE;
    double volt = 0., masset = 0., vn1[M];
    for (int kk = 0; kk < M; ++kk) {
      vn1[kk] = U[kk];
      volt += vn1[kk];
      masset += U[M + kk];
    }
    const double unsurvolt = 1. / volt;
    for (int kk = 0; kk < M; ++kk) vn1[kk] *= vol * unsurvolt;
    double f[M], chi[M], dk[M], ek[M], pk[M], ck[M];
    double somme_frac = 0.;
    const double unsurvol = 1. / vol;
    for (int kk = 0; kk < M; ++kk) {
      f[kk] = 0.;
      if (!((nm >> kk) & 1)) continue;
      f[kk] = vn1[kk] * unsurvol;
      if (f[kk] < thr) f[kk] = 0.;
      somme_frac += f[kk];
    }
    double unsursomme = 0.;
    if (somme_frac != 0.) unsursomme = 1. / somme_frac;
    for (int kk = 0; kk < M; ++kk)
      if ((nm >> kk) & 1) f[kk] *= unsursomme;
    if (S.est) {  // mixed-cell flag for the order-3 remap of the next cycle (:124-145)
      int cnt = 0, pu = -1;
      for (int kk = 0; kk < M; ++kk)
        if (f[kk] > 0.) {
          cnt++;
          pu = kk;
        }
      S.est[c] = (int8_t)(cnt > 1 ? MH_EST_MIXED : pu);
    }
    if (masset > thr * thr) {
      const double unsurmasset = 1. / masset;
      double fm = 0.;
      for (int kk = 0; kk < M; ++kk) {
        chi[kk] = 0.;
        if (!((nm >> kk) & 1)) continue;
        chi[kk] = U[M + kk] * unsurmasset;
        if (f[kk] < thr) chi[kk] = 0.;
        fm += chi[kk];
      }
      if (fm != 0.) {
        const double unsurfm = 1. / fm;
        for (int kk = 0; kk < M; ++kk)
          if ((nm >> kk) & 1) chi[kk] *= unsurfm;
      }
    } else {
      // (near-)empty cell: the reference keeps the mass fractions it holds after the
      // membership change -- the global value for a single-environment cell (taken from
      // the remaining environment's partial value when the cell was mixed), the partial
      // values for a mixed cell (0 for environments that have just entered; all partial
      // values restart from 0 and the global value when a pure cell turns mixed)
      const int on = __popc(om), nnb = __popc(nm);
      for (int kk = 0; kk < M; ++kk) {
        chi[kk] = 0.;
        if (!((nm >> kk) & 1)) continue;
        if (nnb == 1) chi[kk] = on == 1 ? S.chig[c] : S.chik[kk * nc + c];
        else if (on == 1) chi[kk] = ((om >> kk) & 1) ? S.chig[c] : 0.;
        else chi[kk] = ((om >> kk) & 1) ? S.chik[kk * nc + c] : 0.;
      }
    }
    double density = 0.;
    for (int kk = 0; kk < M; ++kk) {
      dk[kk] = 0.;
      if (!((nm >> kk) & 1)) continue;
      if (f[kk] > thr && vn1[kk] != 0.) dk[kk] = U[M + kk] / vn1[kk];
      density += f[kk] * dk[kk];
    }
    for (int kk = 0; kk < M; ++kk) {
      ek[kk] = 0.;
      if (!((nm >> kk) & 1)) continue;
      if (f[kk] > thr && U[M + kk] != 0.) ek[kk] = U[2 * M + kk] / U[M + kk];
    }
    // writes, with the aliasing of single-environment cells (:221-253)
    double rho_g = density;
    double M_g = VE * density;
    double e_g, q_g;
    double Mk[M], qk[M];
    double energie = 0., pseudo = 0.;
    for (int kk = 0; kk < M; ++kk) {
      if (!((nm >> kk) & 1)) continue;
      Mk[kk] = chi[kk] * M_g;
      if (pure) M_g = Mk[kk];
      qk[kk] = (S.q_frozen ? S.Vl[c] * S.q[c] : U[3 * M]) * unsurvol;
      if (pure) rho_g = dk[kk];
      energie += chi[kk] * ek[kk];
      pseudo += f[kk] * qk[kk];
    }
    e_g = energie;
    q_g = pseudo;
    for (int kk = 0; kk < M; ++kk) {
      if (!((nm >> kk) & 1)) continue;
      if (pure) ek[kk] = e_g;  // the cell value overwrites the aliased environment value
      if (dk[kk] < 0. || ek[kk] < 0.) {
        ek[kk] = 0.;
        dk[kk] = 0.;
        f[kk] = 0.;
        chi[kk] = 0.;
        if (pure) {
          e_g = 0.;
          rho_g = 0.;
        }
      }
      if (dk[kk] != dk[kk] || ek[kk] != ek[kk]) S.err[0] = MH_ERR_NAN_REMAP;
    }
    if (S.dec) {  // conservation-energie-totale (RemapADIFinal.cc:340-344): the energy k_cons_ec
                  // found, to each environment by its mass fraction (for a single-environment
                  // cell that is the cell value) and to the value of any other cell
      const double dec = S.dec[c];
      for (int kk = 0; kk < M; ++kk) {
        if (!((nm >> kk) & 1)) continue;
        if (pure) {
          e_g += chi[kk] * dec;
          ek[kk] = e_g;
        } else {
          ek[kk] += chi[kk] * dec;
        }
      }
      if (!pure) e_g += dec;
    }
    // applyEOS per environment, computePressionMoyenne (MahycoModule.cc:1708-1735)
    double p_g = 0., c_g = 0.;
    const double opt = S.vitson_max ? 1. : 0.;
    const double tnow = S.sc->t + S.sc->dt;  // m_global_time of this iteration
    for (int kk = 0; kk < M; ++kk) {
      if (!((nm >> kk) & 1)) continue;
      if (S.env[kk].fictif) {
        fictif_eos(S.env[kk], tnow, pure ? S.p[c] : S.pk[kk * nc + c], pk[kk], ck[kk]);
        ek[kk] = 0.;
        if (pure) e_g = 0.;
      } else {
        eos(S.env[kk], pure ? rho_g : dk[kk], ek[kk], pk[kk], ck[kk]);
      }
      if (pure) {
        p_g = pk[kk];
        c_g = ck[kk];
      }
    }
    if (!pure && S.sans_lagrange) {  // computePressionMoyenne returns at once (:1711)
      p_g = S.p[c];
      c_g = S.c[c];
    } else if (!pure) {
      p_g = 0.;
      c_g = 0.;
      for (int kk = 0; kk < M; ++kk) {
        if (!((nm >> kk) & 1)) continue;
        p_g += f[kk] * pk[kk];
        c_g = opt * ((ck[kk] < c_g) ? c_g : ck[kk]);
        c_g += (1. - opt) * f[kk] * ck[kk];
      }
    }
    if (!pure) {
      for (int kk = 0; kk < M; ++kk) {
        if (!((nm >> kk) & 1)) continue;
        const long o = kk * nc + c;
        S.fk[o] = f[kk];
        S.chik[o] = chi[kk];
        S.rhok[o] = dk[kk];
        S.ek[o] = ek[kk];
        S.pk[o] = pk[kk];
        S.Mk[o] = Mk[kk];
      }
    }
    // the global volume and mass fractions are the environment values of a
    // single-environment cell; a mixed cell keeps the global values it had (Arcane leaves
    // them untouched).  Partial mass fractions follow Arcane's membership change: all
    // restart from 0 when a pure cell turns mixed, and the environment that replaces the
    // only one of a pure cell starts from 0
    if (pure) {
      double fg = 0., cg = 0.;
      for (int kk = 0; kk < M; ++kk)
        if ((nm >> kk) & 1) {
          fg = f[kk];
          cg = chi[kk];
        }
      S.fg[c] = fg;
      S.chig[c] = cg;
    }
    if (nm != om && __popc(om) == 1) {  // a mixed cell's old partial values are kept
      for (int kk = 0; kk < M; ++kk) {
        const bool now = (nm >> kk) & 1, was = (om >> kk) & 1;
        if (now && !pure) continue;  // written below with the cell's partial values
        if (pure ? (now && !was) : !now) S.chik[kk * nc + c] = 0.;
      }
    }
    S.mask[c] = nm;
    if (S.csts) {  // CSTS history: mixed cells' partial pseudo-viscosity, the mask before this remap
      S.qem1[c] = S.qe[c];
      S.qe[c] = (S.q_frozen ? S.Vl[c] * S.q[c] : U[3 * M]) * unsurvol;
      S.mkm1[c] = om;
    }
    S.rho[c] = rho_g;
    S.e[c] = e_g;
    S.p[c] = p_g;
    S.c[c] = c_g;
    S.q[c] = q_g;
    S.M[c] = M_g;
    const int nx = S.g.nx, ny = S.g.ny;
    const int i = (int)(c % nx);
    const long r = c / nx;
    const int j = (int)(r % ny);
    const int k = (int)(r / ny);
    const long b = i + (long)S.g.nnx * (j + (long)S.g.nny * k);
    dxs = S.L[c] / (c_g + cell_vmax<DIM>(S, b));
  }
  block_min_store(dxs, active, S.partial, lb);
}

// applyEOS of one environment on its member cells (PerfectGasEOSService.cc:45-67,
// StiffenedGasEOSService.cc:43-63), as a standalone call.
__device__ __forceinline__ void apply_eos_at(const MhDev& S, int env, long c) {
  const uint8_t mk = S.mask[c];
  if (!((mk >> env) & 1)) return;
  const double tnow = S.sc->t;  // between iterations: m_global_time of the last one
  if (__popc(mk) == 1) {
    double p, cs;
    if (S.env[env].fictif) {
      fictif_eos(S.env[env], tnow, S.p[c], p, cs);
      S.e[c] = 0.;
    } else {
      eos(S.env[env], S.rho[c], S.e[c], p, cs);
    }
    S.p[c] = p;
    S.c[c] = cs;
  } else {
    const long o = env * S.g.nc + c;
    double p, cs;
    if (S.env[env].fictif) {
      fictif_eos(S.env[env], tnow, S.pk[o], p, cs);
      S.ek[o] = 0.;
    } else {
      eos(S.env[env], S.rhok[o], S.ek[o], p, cs);
    }
    S.pk[o] = p;
  }
}
template <int U = 0>
__global__ __launch_bounds__(MH_BLOCK) void k_apply_eos(MhDev S, int env) {
  const long c = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (c >= S.g.nc) return;
  apply_eos_at(S, env, c);
}
// applyOneCellEOS (IEquationOfState.h:50-52): the same for one cell
template <int U = 0>
__global__ __launch_bounds__(64) void k_apply_eos_cell(MhDev S, int env, long c) {
  if (threadIdx.x == 0) apply_eos_at(S, env, c);
}

// computePressionMoyenne (MahycoModule.cc:1708-1735) on multi-environment cells; the
// environment sound speeds are those applyEOS stores for the stored density and energy
// (1 for a Fictif environment, FictifEOSService.cc:62).  Advection mode returns at once
// (:1711): the host launches nothing then.
template <int M>
__global__ __launch_bounds__(MH_BLOCK) void k_pression_moyenne(MhDev S) {
  const long c = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (c >= S.g.nc || S.sans_lagrange) return;
  const uint8_t mk = S.mask[c];
  if (__popc(mk) == 1) return;
  const double opt = S.vitson_max ? 1. : 0.;
  double ps = 0., cm = 0.;
  for (int kk = 0; kk < M; ++kk) {
    if (!((mk >> kk) & 1)) continue;
    const long o = kk * S.g.nc + c;
    double p, cs = 1.;
    if (!S.env[kk].fictif) eos(S.env[kk], S.rhok[o], S.ek[o], p, cs);
    ps += S.fk[o] * S.pk[o];
    cm = opt * ((cs < cm) ? cm : cs);
    cm += (1. - opt) * S.fk[o] * cs;
  }
  S.p[c] = ps;
  S.c[c] = cm;
}

// The CFL minimum of computeDeltaT (MahycoModule.cc:1767-1792) from the per-block minima
// partial[2b] (and NaN flags partial[2b+1]) of nblocks blocks, in cell order.  Called by every
// thread of a one-block launch; returns true on thread 0 only, with the minimum in *mn.
__device__ __forceinline__ bool cfl_fold(const double* __restrict__ partial, int nblocks, double* mn) {
  __shared__ double rm[MH_DT_BLOCK / 64];
  __shared__ int rn[MH_DT_BLOCK / 64];
  const int nw = (int)(blockDim.x >> 6), lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __shared__ int any_nan;
  // without a NaN candidate anywhere the fold is a plain minimum, in any order
  if (threadIdx.x == 0) any_nan = 0;
  __syncthreads();
  // four independent running minima keep four loads in flight per thread (the minimum of
  // non-NaN values does not depend on the order)
  const double qnan = __longlong_as_double(0x7ff8000000000000LL);
  double pm = qnan, pm1 = qnan, pm2 = qnan, pm3 = qnan;
  int pn = 0;
  const int bd = (int)blockDim.x;
  int b = threadIdx.x;
  for (; b + 3 * bd < nblocks; b += 4 * bd) {
    const double a0 = partial[2 * b], a1 = partial[2 * (b + bd)];
    const double a2 = partial[2 * (b + 2 * bd)], a3 = partial[2 * (b + 3 * bd)];
    pn |= (partial[2 * b + 1] != 0.) | (partial[2 * (b + bd) + 1] != 0.) |
          (partial[2 * (b + 2 * bd) + 1] != 0.) | (partial[2 * (b + 3 * bd) + 1] != 0.);
    pm = min_skip_nan(pm, a0);
    pm1 = min_skip_nan(pm1, a1);
    pm2 = min_skip_nan(pm2, a2);
    pm3 = min_skip_nan(pm3, a3);
  }
  for (; b < nblocks; b += bd) {
    pm = min_skip_nan(pm, partial[2 * b]);
    pn |= partial[2 * b + 1] != 0.;
  }
  pm = min_skip_nan(min_skip_nan(pm, pm1), min_skip_nan(pm2, pm3));
  for (int off = 32; off > 0; off >>= 1) pm = min_skip_nan(pm, __shfl_xor(pm, off, 64));
  if (__ballot(pn) != 0 && lane == 0) any_nan = 1;
  if (lane == 0) rm[wid] = pm;
  __syncthreads();
  MhCflMin tot;
  if (!any_nan) {
    if (threadIdx.x != 0) return false;
    tot.nan = 0;
    tot.m = rm[0];
    for (int w = 1; w < nw; ++w) tot.m = min_skip_nan(tot.m, rm[w]);
  } else {
    // block partials in block (= cell) order: wavefront w folds a contiguous run of them,
    // 64 at a time, then the wavefronts are combined in order
    __syncthreads();
    const int chunk = ((nblocks + nw - 1) / nw + 63) / 64 * 64;
    const int bbeg = wid * chunk, bend = min(nblocks, bbeg + chunk);
    MhCflMin acc = cfl_leaf(0., false);
    for (int b0 = bbeg; b0 < bend; b0 += 64) {
      const int b = b0 + lane;
      MhCflMin o = cfl_leaf(0., false);
      if (b < bend) {
        o.m = partial[2 * b];
        o.nan = partial[2 * b + 1] != 0.;
      }
      o = cfl_wave(o);
      acc = cfl_combine(acc, o);  // lane 0 holds the run's value; other lanes are ignored
    }
    if (lane == 0) {
      rm[wid] = acc.m;
      rn[wid] = acc.nan;
    }
    __syncthreads();
    if (threadIdx.x != 0) return false;
    tot.m = rm[0];
    tot.nan = rn[0];
    for (int w = 1; w < nw; ++w) {
      MhCflMin o;
      o.m = rm[w];
      o.nan = rn[w];
      tot = cfl_combine(tot, o);
    }
  }
  *mn = tot.nan ? tot.m : amin(1.79769313486231570815e+308, tot.m);
  return true;
}

// computeDeltaT (MahycoModule.cc:1739-1864) on the block minima, plus the time increment
// of the iteration and the final-time stop rule.  One block.
__device__ __forceinline__ void k_dt_body(const MhDev& S) {
  MhScalars* sc = S.sc;
  if (mh_halted(sc)) return;
  double mn;
  if (!cfl_fold(S.partial, S.nblk_c, &mn)) return;
  const double dt = sc->dt;
  const double t = sc->t + dt;
  double new_dt = S.cfl * mn;
  new_dt = amin(new_dt, 1.05 * dt);
  new_dt = amin(new_dt, S.dtmax);
  int status = S.err[0];
  if (S.sans_lagrange) new_dt = S.dt_init;  // advection mode keeps deltat-init (:1751-1753)
  else if (status == 0 && new_dt < S.dtmin) status = MH_ERR_DT_MIN;
  sc->t = t;
  sc->iteration = sc->iteration + 1;
  sc->dxmin = mn;
  if (status != 0) {
    sc->status = status;
    return;
  }
  sc->dt_old = dt;
  sc->dt = new_dt;
  const double T = S.tfinal;
  if ((t < T && t + new_dt > T) || t > T) sc->stopped = 1;
  if (S.max_iter > 0 && sc->iteration >= S.max_iter) sc->stopped = 1;
}


// ----------------------------------------------------------------------------- time history
// EnvSummationService::write (time_history/EnvSummationService.cc:77-169) every hist_period
// iterations: per environment the plain mean of M_k and p_k over its cells, the
// volume-weighted density and the mass-weighted internal energy.  Runs after k_dt; both
// kernels do nothing unless the iteration just completed is a new multiple of the period.
__device__ __forceinline__ bool hist_due(const MhDev& S) {
  const MhScalars* sc = S.sc;
  return S.hist_period > 0 && sc->status == 0 && sc->iteration > 0 &&
         (sc->iteration % S.hist_period) == 0 && sc->iteration != sc->hist_last && sc->hist_rows < S.hist_max;
}

template <int M>
__device__ __forceinline__ void k_hist_part_body(const MhDev& S, int lb) {
  __shared__ double red[6 * M][MH_BLOCK / 64];
  if (!hist_due(S)) return;
  const long c = lb * (long)blockDim.x + threadIdx.x;
  double acc[6 * M];
  for (int v = 0; v < 6 * M; ++v) acc[v] = 0.;
  if (c < S.g.nc) {
    const uint8_t mk = S.mask[c];
    const bool pure = __popc(mk) == 1;
    const double V = S.with_proj ? S.VE[c] : S.Vl[c];
    const long nc = S.g.nc;
    for (int k = 0; k < M; ++k) {
      if (!((mk >> k) & 1)) continue;
      const long o = k * nc + c;
      const double Mk = pure ? S.M[c] : S.Mk[o];
      const double pk = pure ? S.p[c] : S.pk[o];
      const double rk = pure ? S.rho[c] : S.rhok[o];
      const double ek = pure ? S.e[c] : S.ek[o];
      const double Vk = pure ? V : S.fk[o] * V;
      acc[6 * k + 0] = 1.;
      acc[6 * k + 1] = Mk;
      acc[6 * k + 2] = pk;
      acc[6 * k + 3] = rk * Vk;
      acc[6 * k + 4] = Vk;
      acc[6 * k + 5] = ek * Mk;
    }
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  for (int v = 0; v < 6 * M; ++v) {
    double x = acc[v];
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    if (lane == 0) red[v][wid] = x;
  }
  __syncthreads();
  if (threadIdx.x < 6 * M) {
    double x = 0.;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) x += red[threadIdx.x][w];
    S.hist_part[(long)lb * 6 * M + threadIdx.x] = x;
  }
}

template <int M>
__device__ __forceinline__ void k_hist_final_body(const MhDev& S) {
  if (!hist_due(S)) return;
  MhScalars* sc = S.sc;
  const int v = threadIdx.x;
  __shared__ double tot[6 * M];
  if (v < 6 * M) {
    double x = 0.;
    for (int b = 0; b < S.nblk_c; ++b) x += S.hist_part[(long)b * 6 * M + v];
    tot[v] = x;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  double* row = S.hist + (long)sc->hist_rows * (1 + 5 * M);
  row[0] = sc->t;
  for (int k = 0; k < M; ++k) {
    const double cnt = tot[6 * k + 0];
    row[1 + 5 * k + 0] = cnt > 0. ? tot[6 * k + 1] / cnt : 0.;
    row[1 + 5 * k + 1] = cnt > 0. ? tot[6 * k + 2] / cnt : 0.;
    row[1 + 5 * k + 2] = cnt > 0. ? tot[6 * k + 3] / tot[6 * k + 4] : 0.;
    row[1 + 5 * k + 3] = cnt > 0. ? tot[6 * k + 5] / tot[6 * k + 1] : 0.;
    row[1 + 5 * k + 4] = cnt;
  }
  sc->hist_rows = sc->hist_rows + 1;
  sc->hist_last = sc->iteration;
}

template <int M>
__global__ __launch_bounds__(MH_BLOCK) void k_hist_part(MhDev S) { k_hist_part_body<M>(S, (int)blockIdx.x); }
template <int M>
__global__ __launch_bounds__(MH_BLOCK) void k_hist_final(MhDev S) { k_hist_final_body<M>(S); }

// ----------------------------------------------------------------------------- wrappers
// Single mesh: the state travels by value in the kernel arguments.  Batch: one launch
// covers many meshes; block b works on mesh bmap[b].x, local block bmap[b].y, whose state
// is tab[bmap[b].x] (meshes of one launch share DIM, M and the sweep direction).
// Single-mesh grids use the XCD-contiguous block order of xcd_block.

#ifndef MH_CELL_MIN_WAVES
#define MH_CELL_MIN_WAVES 4  // register budget of k_cell_lag, in waves per SIMD
#endif
#ifndef MH_NODE_MIN_WAVES
#define MH_NODE_MIN_WAVES 1  // register budget of k_node_lag3_tiled, in waves per SIMD
#endif
// Batched variants are persistent: a fixed grid walks the work list bmap[0 .. *nwork) (one
// entry per (case, block) of every case of a batch group) with a grid stride, so the list and
// its length can change between launches of a captured graph without re-capture.
#define MH_WRAP(NAME, WAVES, TPL, TARGS, PARAMS, ARGS)                                      \
  TPL __global__ __launch_bounds__(MH_BLOCK, WAVES) void NAME(MhDev S PARAMS) {              \
    NAME##_body TARGS(S, xcd_block((int)blockIdx.x, (int)gridDim.x) ARGS);                   \
  }                                                                                          \
  TPL __global__ __launch_bounds__(MH_BLOCK, WAVES) void NAME##_batch(                        \
      const MhDev* __restrict__ tab, const int2* __restrict__ bmap, const int* __restrict__ nwork PARAMS) { \
    const int n = *nwork;                                                                    \
    for (int w = (int)blockIdx.x; w < n; w += (int)gridDim.x) {                              \
      const int2 bm = bmap[w];                                                               \
      NAME##_body TARGS(tab[bm.x], bm.y ARGS);                                               \
      __syncthreads();                                                                       \
    }                                                                                        \
  }
#define MH_COMMA ,
MH_WRAP(k_node_lag, 1, template <int DIM>, <DIM>, , )
MH_WRAP(k_cell_lag, MH_CELL_MIN_WAVES, template <int DIM MH_COMMA int M>, <DIM MH_COMMA M>, , )
MH_WRAP(k_cell_lag_csts, 1, template <int DIM MH_COMMA int M>, <DIM MH_COMMA M>, , )
MH_WRAP(k_cell_length, 1, template <int DIM>, <DIM>, , )
MH_WRAP(k_sweep_cell, 1, template <int DIM MH_COMMA int D MH_COMMA int M MH_COMMA int LIM>, <DIM MH_COMMA D MH_COMMA M MH_COMMA LIM>,
        MH_COMMA int usel MH_COMMA int first, MH_COMMA usel MH_COMMA first)
MH_WRAP(k_sweep_node, 1, template <int DIM MH_COMMA int D MH_COMMA int LIM>, <DIM MH_COMMA D MH_COMMA LIM>,
        MH_COMMA int psel MH_COMMA int first MH_COMMA int last, MH_COMMA psel MH_COMMA first MH_COMMA last)
MH_WRAP(k_node_velocity, 1, template <int DIM>, <DIM>, MH_COMMA int materialize, MH_COMMA materialize)
MH_WRAP(k_finalize, 1, template <int DIM MH_COMMA int M>, <DIM MH_COMMA M>, MH_COMMA int usel, MH_COMMA usel)
MH_WRAP(k_sweep_cell_row, 1, template <int DIM MH_COMMA int M MH_COMMA int LIM>, <DIM MH_COMMA M MH_COMMA LIM>,
        MH_COMMA int usel MH_COMMA int first, MH_COMMA usel MH_COMMA first)
MH_WRAP(k_sweep_cell_tile, MH_TILE_MIN_WAVES,
        template <int DIM MH_COMMA int D MH_COMMA int M MH_COMMA int LIM>, <DIM MH_COMMA D MH_COMMA M MH_COMMA LIM>,
        MH_COMMA int usel MH_COMMA int first, MH_COMMA usel MH_COMMA first)
MH_WRAP(k_sweep_node_row, 1, template <int DIM MH_COMMA int LIM>, <DIM MH_COMMA LIM>,
        MH_COMMA int psel MH_COMMA int first MH_COMMA int last, MH_COMMA psel MH_COMMA first MH_COMMA last)
MH_WRAP(k_hcell, 1, template <int DIM>, <DIM>, , )
MH_WRAP(k_node_adv, 1, template <int DIM>, <DIM>, , )
MH_WRAP(k_node_fwd, 1, template <int DIM>, <DIM>, , )
MH_WRAP(k_sweep_cell_gen, 1, template <int DIM MH_COMMA int D MH_COMMA int M>, <DIM MH_COMMA D MH_COMMA M>,
        MH_COMMA int usel MH_COMMA int first, MH_COMMA usel MH_COMMA first)
MH_WRAP(k_sweep_cell3, 1, template <int DIM MH_COMMA int D MH_COMMA int M>, <DIM MH_COMMA D MH_COMMA M>,
        MH_COMMA int usel MH_COMMA int first, MH_COMMA usel MH_COMMA first)
MH_WRAP(k_sweep_ke, 1, template <int DIM MH_COMMA int D>, <DIM MH_COMMA D>, MH_COMMA int psel MH_COMMA int first,
        MH_COMMA psel MH_COMMA first)
MH_WRAP(k_dual_init, 1, template <int DIM>, <DIM>, , )
MH_WRAP(k_cons_ec, 1, template <int DIM>, <DIM>, , )
MH_WRAP(k_prepare_remap, 1, template <int DIM MH_COMMA int M>, <DIM MH_COMMA M>, , )
MH_WRAP(k_sweep_node_march, MH_NM_MIN_WAVES, template <int DIM MH_COMMA int D MH_COMMA int LIM>, <DIM MH_COMMA D MH_COMMA LIM>,
        MH_COMMA int psel MH_COMMA int first MH_COMMA int last, MH_COMMA psel MH_COMMA first MH_COMMA last)

// (the plain kernels are templates with an unused parameter, so that every translation unit
// that launches one may instantiate it)
template <int U = 0>
__global__ __launch_bounds__(MH_BLOCK, MH_NODE_MIN_WAVES) void k_node_lag3_tiled(MhDev S) {
  k_node_lag3_tiled_body(S, blockIdx.x, blockIdx.y, blockIdx.z);
}

template <int U = 0>
__global__ __launch_bounds__(MH_DT_BLOCK) void k_dt(MhDev S) { k_dt_body(S); }
template <int M>
__global__ __launch_bounds__(MH_BLOCK, MH_CT_MIN_WAVES) void k_cell_lag3_tiled(MhDev S) {
  k_cell_lag3_tiled_body<M>(S, blockIdx.x, blockIdx.y, blockIdx.z);
}
template <int U = 0>
__global__ __launch_bounds__(MH_DT_BLOCK) void k_dt_batch(const MhDev* __restrict__ tab,
                                                          const int* __restrict__ list,
                                                          const int* __restrict__ nlist) {
  const int n = *nlist;
  for (int w = (int)blockIdx.x; w < n; w += (int)gridDim.x) {
    k_dt_body(tab[list[w]]);
    __syncthreads();
  }
}
// The z-marching node kernel over a batch group: work item (case, tile), tile index
// bx + ntx (by + nty bz) with ntx, nty the case's tile counts along x and y.
template <int U = 0>
__global__ __launch_bounds__(MH_BLOCK, MH_NODE_MIN_WAVES) void k_node_lag3_tiled_batch(
    const MhDev* __restrict__ tab, const int2* __restrict__ bmap, const int* __restrict__ nwork) {
  const int n = *nwork;
  for (int w = (int)blockIdx.x; w < n; w += (int)gridDim.x) {
    const int2 bm = bmap[w];
    const MhDev& S = tab[bm.x];
    const int ntx = (S.g.nnx + MH_NTX - 1) / MH_NTX, nty = (S.g.nny + MH_NTY - 1) / MH_NTY;
    k_node_lag3_tiled_body(S, bm.y % ntx, (bm.y / ntx) % nty, bm.y / (ntx * nty));
    __syncthreads();
  }
}

#endif
