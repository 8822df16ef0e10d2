// mahyco_oracle.cpp -- CPU restatement of MaHyCo's Lagrange + ADI-remap cycle.
//
// TEST INFRASTRUCTURE ONLY: this file is the parity checker for the HIP path and
// the CPU-baseline timing ("port").  The product library never links it.
//
// It follows the reference operation by operation, in the reference's evaluation
// order (no FMA contraction: build with -ffp-contract=off), so that its results
// can be compared with the reference's non-regression fixtures and, to the last
// bit, with the HIP kernels.  Each function cites the reference file:line it
// restates (paths relative to the MaHyCo tree).
//
// Arcane semantics restated here:
//  * material variables (MaterialVariableCellReal) hold one global value per cell
//    plus one partial value per (environment, cell) for cells with several
//    environments; for a cell with exactly one environment the environment value
//    IS the global value (same storage).  MatVar/ev() below model that aliasing.
//  * ENUMERATE_ENV / ENUMERATE_ENVCELL visit the member cells of one environment in
//    cell order; ENUMERATE_CELL visits cells in local-id order, which for the
//    Cartesian generator is x fastest, then y, then z.
//  * Scatter loops of the reference (node force, node mass, dual mass fluxes) are
//    written here as per-node gathers that visit the contributing cells/faces in
//    the same order as the scatter does, so every node sees the identical sequence
//    of additions while the loop can run on several threads.
//  * Face node order follows Arcane's reordering (smallest unique id first, then
//    its smaller neighbour), which makes every Cartesian face normal +e_d.
//
// Computations whose results no parity field ever reads are omitted and listed
// where they would sit: temperature (cv), phase fractions, stress deviators and
// plastic variables (all remap slots other than V_k, M_k, E_k and V*q: each slot
// is swept independently, so dropping a slot changes no other slot), the
// elasto-plastic force term of a NoModel run (adds +-0), gravity (0), and
// Endommagement (inert at its default thresholds -1e12 / 0).
//
// Every function is pinned by the reference's fixtures or by the mathematical checks of
// tests/test_oracle_*.py, tests/test_conservation_energie.py and tests/test_geom_host.py
// (the thresholds of computeY0 on unequal widths by the face values they are defined by).

#include "mahyco_oracle.h"
#include "../mahyco_amd/csrc/mh_math.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// 0: L = std::pow(V, 1/dim) as the reference (glibc); 1: the correctly rounded power of
// mh_math.h, which the HIP kernels use (glibc misrounds ~0.08% of arguments by one ulp).
int g_pow_mode = 0;

struct R3 {
  double x, y, z;
};
inline R3 operator+(R3 a, R3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline R3 operator-(R3 a, R3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline R3 operator*(double s, R3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline R3& operator+=(R3& a, R3 b) { a.x += b.x; a.y += b.y; a.z += b.z; return a; }
inline double dot(R3 a, R3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }  // Arcane math::dot
inline R3 vecMul(R3 u, R3 v) {                                                 // Arcane math::vecMul
  return {u.y * v.z - u.z * v.y, u.z * v.x - u.x * v.z, u.x * v.y - u.y * v.x};
}
inline double normL2(R3 a) { return std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
inline double squareNormL2(R3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
inline double comp(R3 a, int d) { return d == 0 ? a.x : (d == 1 ? a.y : a.z); }
inline double& compr(R3& a, int d) { return d == 0 ? a.x : (d == 1 ? a.y : a.z); }
inline double amax(double a, double b) { return (a < b) ? b : a; }  // Arcane math::max
inline double amin(double a, double b) { return (a < b) ? a : b; }  // Arcane math::min

// Arcane MaterialVariableCellReal: global [nc] + partial [nc][MH_MAX_ENV].
struct MatVar {
  std::vector<double> g, p;
  void init(long nc) { g.assign(nc, 0.0); p.assign(nc * MH_MAX_ENV, 0.0); }
};

}  // namespace

struct OracleCase {
  // ---------------------------------------------------------------- mesh
  int dim = 3, nx = 1, ny = 1, nz = 1;
  long nc = 0, nn = 0, nf = 0;
  int nnpc = 8;  // nodes per cell
  int nfpc = 6;  // faces per cell
  std::vector<int> cell_nodes;   // [nc][nnpc]
  std::vector<int> cell_faces;   // [nc][nfpc], Arcane local face order
  std::vector<int> face_nodes;   // [nf][4]
  std::vector<int> face_nbnode;  // [nf]
  std::vector<int> face_dir;     // structural normal direction
  std::vector<int> face_prev, face_next;  // DirFace previousCell / nextCell (-1 = none)
  long face_off[3] = {0, 0, 0}, face_cnt[3] = {0, 0, 0};
  std::vector<int> node_prev[3], node_next[3];  // DirNode previous / next (-1 = none)
  std::vector<int> nc_off, nc_cell, nc_local;  // node -> cells in cell order, local index
  std::vector<int> nf_off[3], nf_face;         // node -> faces of dir d in face order
  std::vector<int> nf_list[3];
  int cells_in_dir[3] = {1, 1, 1};

  // ---------------------------------------------------------------- options
  int m = 1;
  MhEnvironment env[MH_MAX_ENV];
  MhOptions opt;

  // ---------------------------------------------------------------- state
  std::vector<R3> node_coord, node_coord_0, velocity, velocity_n, force;
  std::vector<R3> node_coord_lag;  // with remap: the Lagrangian positions before x <- x0
  std::vector<double> node_mass;
  std::vector<R3> cell_coord, face_coord, face_normal;
  std::vector<R3> cqs, cqs_n;            // [nc][nnpc]
  std::vector<R3> outer_face_normal;     // [nc][nfpc]
  std::vector<double> caracteristic_length, div_u, euler_volume, materiau;
  std::vector<uint8_t> mask;             // environment membership bits
  std::vector<int8_t> nbenv;
  MatVar fracvol, mass_fraction, density, density_n, internal_energy, internal_energy_n,
      pressure, pressure_n, sound_speed, pseudo, pseudo_n, pseudo_nm1, cell_volume,
      cell_volume_n, cell_mass, tau_density;
  std::vector<MatVar*> matvars;

  // remap service state (RemapADIService.h); slots: k V_k, m+k M_k, 2m+k E_k, 3m V*q
  int nv = 4;
  std::vector<double> u_lagrange, phi_lagrange, grad_phi, grad_phi_face, phi_face, dual_phi_flux;
  std::vector<double> u_dual, phi_dual, dual_grad_phi;  // [nn][5]
  std::vector<double> deltax_lagrange, face_normal_velocity;
  std::vector<double> h_cell_lagrange;       // cell width along the sweep
  std::vector<int> est_mixte, est_pure;      // mixed-cell indicators; Arcane starts them at 0
  std::vector<double> dphi_ar, dphi_av;      // pente-borne fluxes through the back / front face [nc][nv]
  // A run continued from a dump projects only its first 3m+5 remap slots
  // (hydroContinueInit, MahycoModule.cc:288): with m >= 2 the pseudo-viscosity slot 7m is
  // not swept, so V q keeps its Lagrangian value through the remap.
  bool q_frozen = false;
  std::vector<R3> face_length_lagrange;
  std::vector<double> back_flux_mass, front_flux_mass;
  int sens_projection = 0;

  // boundary face groups -> node lists (XMIN..ZMAX)
  std::vector<int> group_nodes[6];
  std::vector<double> riemann_f1;  // MH_CASE_RIEMANN: volume fraction of env 1

  // time loop
  double global_time = 0.0, global_deltat = 0.0, global_old_deltat = 0.0, old_deltat = 0.0;
  int global_iteration = 0;
  bool stopped = false;
  std::vector<double> history;  // rows of 1 + 5m values

  // ---------------------------------------------------------------- helpers
  int node_id(int i, int j, int k) const { return i + (nx + 1) * (j + (ny + 1) * k); }
  int cell_id(int i, int j, int k) const { return i + nx * (j + ny * k); }
  double& ev(MatVar& v, int k, long c) { return nbenv[c] == 1 ? v.g[c] : v.p[c * MH_MAX_ENV + k]; }
  bool member(int k, long c) const { return (mask[c] >> k) & 1; }
  int nnodes_face() const { return dim == 3 ? 4 : 2; }

  void build_mesh(const MhMesh& mesh);
  void hydroStartInit();
  void hydroStartInitEnvAndMat();
  void initMatAndVar();
  void computeCellMass();
  void computeNodeMass();
  void InitGeometricValues();
  void PrepareFaceGroup();

  void saveValuesAtN();
  void updateVelocity();
  void updateVelocityForward();
  void applyBoundaryCondition();
  void updatePosition();
  int computeGeometricValues();
  void computeCQs(const R3* coord, const R3* face_coord, R3* out);
  void updateDensity();
  void computeArtificialViscosity();
  void updateEnergyAndPressure();
  void applyEOS(int k);
  void computePressionMoyenne();
  int remap();
  void computeVariablesForRemap();
  void computeFaceQuantitesForRemap();
  void computeGradPhiFace(int idir);
  void computeGradPhiCell(int idir);
  void computeUpwindFaceQuantitiesForProjection(int idir);
  void computeUremap(int idir);
  void computeDualUremap(int idir);
  int dirNeighbor(long c, int idir, bool back) const;
  void riderRadii(long c, bool low_centre, double& rmin, double& rmax) const;
  void updateVelocityWithoutLagrange();
  void applyFictif(int k);
  void setMixedFlags(long c, const double* f);
  int remapVariables();
  void setMembership(const std::vector<uint8_t>& newmask);
  int computeDeltaT();
  void timeHistoryWrite();
  int cycle();
};

// ============================================================================ mesh
// Arcane Cartesian generator restated: node unique id i + (nx+1)(j + (ny+1)k),
// hexahedron nodes (i,j,k),(i+1,j,k),(i+1,j+1,k),(i,j+1,k), then the same at k+1
// (NONREGRESSION/CAS_BiSodCaseX/output/depouillement/ensight.geo000000 connectivity).
void OracleCase::build_mesh(const MhMesh& mesh) {
  dim = mesh.dim;
  nx = mesh.nx;
  ny = mesh.ny;
  nz = dim == 3 ? mesh.nz : 1;
  nnpc = dim == 3 ? 8 : 4;
  nfpc = dim == 3 ? 6 : 4;
  nc = (long)nx * ny * nz;
  const int nnz = dim == 3 ? nz + 1 : 1;
  nn = (long)(nx + 1) * (ny + 1) * nnz;
  cells_in_dir[0] = nx;
  cells_in_dir[1] = ny;
  cells_in_dir[2] = dim == 3 ? nz : 1;

  auto coord1 = [&](int d, int i, int n) {
    double o = mesh.origin[d], l = mesh.length[d];
    if (mesh.coord_rule == MH_COORD_DIV) return o + (l * (double)i) / (double)n;
    return o + (double)i * (l / (double)n);
  };
  node_coord.resize(nn);
  for (int k = 0; k < nnz; ++k)
    for (int j = 0; j <= ny; ++j)
      for (int i = 0; i <= nx; ++i) {
        R3 x{coord1(0, i, nx), coord1(1, j, ny), dim == 3 ? coord1(2, k, nz) : 0.0};
        node_coord[node_id(i, j, k)] = x;
      }

  cell_nodes.resize(nc * nnpc);
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        int* cn = &cell_nodes[(long)cell_id(i, j, k) * nnpc];
        cn[0] = node_id(i, j, k);
        cn[1] = node_id(i + 1, j, k);
        cn[2] = node_id(i + 1, j + 1, k);
        cn[3] = node_id(i, j + 1, k);
        if (dim == 3) {
          cn[4] = node_id(i, j, k + 1);
          cn[5] = node_id(i + 1, j, k + 1);
          cn[6] = node_id(i + 1, j + 1, k + 1);
          cn[7] = node_id(i, j + 1, k + 1);
        }
      }

  // Faces, one family per normal direction, each ordered k, j, i (i fastest).
  int fdims[3][3];
  for (int d = 0; d < dim; ++d) {
    for (int e = 0; e < 3; ++e) fdims[d][e] = (e == 0 ? nx : e == 1 ? ny : nz);
    fdims[d][d] += 1;
    if (dim == 2) fdims[d][2] = 1;
  }
  nf = 0;
  for (int d = 0; d < dim; ++d) {
    face_off[d] = nf;
    face_cnt[d] = (long)fdims[d][0] * fdims[d][1] * fdims[d][2];
    nf += face_cnt[d];
  }
  face_nodes.assign(nf * 4, -1);
  face_nbnode.assign(nf, dim == 3 ? 4 : 2);
  face_dir.assign(nf, 0);
  face_prev.assign(nf, -1);
  face_next.assign(nf, -1);
  auto fid = [&](int d, int i, int j, int k) -> long {
    return face_off[d] + i + (long)fdims[d][0] * (j + (long)fdims[d][1] * k);
  };
  for (int d = 0; d < dim; ++d)
    for (int k = 0; k < fdims[d][2]; ++k)
      for (int j = 0; j < fdims[d][1]; ++j)
        for (int i = 0; i < fdims[d][0]; ++i) {
          long f = fid(d, i, j, k);
          face_dir[f] = d;
          int* fn = &face_nodes[f * 4];
          if (dim == 3) {
            if (d == 0) {
              fn[0] = node_id(i, j, k); fn[1] = node_id(i, j + 1, k);
              fn[2] = node_id(i, j + 1, k + 1); fn[3] = node_id(i, j, k + 1);
            } else if (d == 1) {
              fn[0] = node_id(i, j, k); fn[1] = node_id(i + 1, j, k);
              fn[2] = node_id(i + 1, j, k + 1); fn[3] = node_id(i, j, k + 1);
            } else {
              fn[0] = node_id(i, j, k); fn[1] = node_id(i + 1, j, k);
              fn[2] = node_id(i + 1, j + 1, k); fn[3] = node_id(i, j + 1, k);
            }
          } else {
            if (d == 0) { fn[0] = node_id(i, j, 0); fn[1] = node_id(i, j + 1, 0); }
            else { fn[0] = node_id(i, j, 0); fn[1] = node_id(i + 1, j, 0); }
          }
          int idx[3] = {i, j, k};
          int lim = d == 0 ? nx : d == 1 ? ny : nz;
          int c = idx[d];
          if (c - 1 >= 0) { int t[3] = {i, j, k}; t[d] = c - 1; face_prev[f] = cell_id(t[0], t[1], t[2]); }
          if (c < lim) face_next[f] = cell_id(i, j, k);
        }
  cell_faces.resize(nc * nfpc);
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        int* cf = &cell_faces[(long)cell_id(i, j, k) * nfpc];
        if (dim == 3) {  // Hexaedron8 local faces: (0,3,2,1) (0,4,7,3) (0,1,5,4) (4,5,6,7) (1,2,6,5) (2,3,7,6)
          cf[0] = (int)fid(2, i, j, k);
          cf[1] = (int)fid(0, i, j, k);
          cf[2] = (int)fid(1, i, j, k);
          cf[3] = (int)fid(2, i, j, k + 1);
          cf[4] = (int)fid(0, i + 1, j, k);
          cf[5] = (int)fid(1, i, j + 1, k);
        } else {  // Quad4 local faces: (0,1) (1,2) (2,3) (3,0)
          cf[0] = (int)fid(1, i, j, 0);
          cf[1] = (int)fid(0, i + 1, j, 0);
          cf[2] = (int)fid(1, i, j + 1, 0);
          cf[3] = (int)fid(0, i, j, 0);
        }
      }

  // Node direction neighbours (NodeDirectionMng).
  for (int d = 0; d < dim; ++d) {
    node_prev[d].assign(nn, -1);
    node_next[d].assign(nn, -1);
  }
  for (int k = 0; k < nnz; ++k)
    for (int j = 0; j <= ny; ++j)
      for (int i = 0; i <= nx; ++i) {
        int n = node_id(i, j, k);
        if (i > 0) node_prev[0][n] = node_id(i - 1, j, k);
        if (i < nx) node_next[0][n] = node_id(i + 1, j, k);
        if (j > 0) node_prev[1][n] = node_id(i, j - 1, k);
        if (j < ny) node_next[1][n] = node_id(i, j + 1, k);
        if (dim == 3) {
          if (k > 0) node_prev[2][n] = node_id(i, j, k - 1);
          if (k < nz) node_next[2][n] = node_id(i, j, k + 1);
        }
      }

  // node -> cells in cell order (the order in which a cell-order scatter reaches the node)
  std::vector<int> cnt(nn, 0);
  for (long c = 0; c < nc; ++c)
    for (int l = 0; l < nnpc; ++l) cnt[cell_nodes[c * nnpc + l]]++;
  nc_off.assign(nn + 1, 0);
  for (long n = 0; n < nn; ++n) nc_off[n + 1] = nc_off[n] + cnt[n];
  nc_cell.resize(nc_off[nn]);
  nc_local.resize(nc_off[nn]);
  std::vector<int> pos(nc_off.begin(), nc_off.end() - 1);
  for (long c = 0; c < nc; ++c)
    for (int l = 0; l < nnpc; ++l) {
      int n = cell_nodes[c * nnpc + l];
      nc_cell[pos[n]] = (int)c;
      nc_local[pos[n]++] = l;
    }
  // node -> faces of direction d in face iteration order
  for (int d = 0; d < dim; ++d) {
    std::fill(cnt.begin(), cnt.end(), 0);
    for (long f = face_off[d]; f < face_off[d] + face_cnt[d]; ++f)
      for (int l = 0; l < face_nbnode[f]; ++l) cnt[face_nodes[f * 4 + l]]++;
    nf_off[d].assign(nn + 1, 0);
    for (long n = 0; n < nn; ++n) nf_off[d][n + 1] = nf_off[d][n] + cnt[n];
    nf_list[d].resize(nf_off[d][nn]);
    std::vector<int> p2(nf_off[d].begin(), nf_off[d].end() - 1);
    for (long f = face_off[d]; f < face_off[d] + face_cnt[d]; ++f)
      for (int l = 0; l < face_nbnode[f]; ++l) nf_list[d][p2[face_nodes[f * 4 + l]]++] = (int)f;
  }
}

// ============================================================================ init
// MahycoModule::hydroStartInit, MahycoModule.cc:21-210.
void OracleCase::hydroStartInit() {
  global_deltat = opt.deltat_init;  // :64-65
  hydroStartInitEnvAndMat();        // :69
  computeGeometricValues();         // :71
  PrepareFaceGroup();               // :74
  // :88-108 fills: fraction 1, mass fraction 1, pseudo 0, e_n 0 (global and partial)
  std::fill(fracvol.g.begin(), fracvol.g.end(), 1.0);
  std::fill(fracvol.p.begin(), fracvol.p.end(), 1.0);
  std::fill(mass_fraction.g.begin(), mass_fraction.g.end(), 1.0);
  std::fill(mass_fraction.p.begin(), mass_fraction.p.end(), 1.0);
  std::fill(pseudo.g.begin(), pseudo.g.end(), 0.0);
  std::fill(pseudo.p.begin(), pseudo.p.end(), 0.0);
  initMatAndVar();  // casModel()->initVar :110-140 (membership was set by initMat)
  // :143-152 initEOS per environment, then the mixed-cell averages (not in advection mode)
  for (int k = 0; k < m && !opt.sans_lagrange; ++k) {
    const MhEnvironment& e = env[k];
    const double gamma = e.adiabatic_cst;
    for (long c = 0; c < nc; ++c) {
      if (!member(k, c)) continue;
      double p = ev(pressure, k, c), rho = ev(density, k, c);
      if (e.eos == MH_EOS_STIFFENED_GAS) {  // StiffenedGasEOSService.cc:14-33
        const double pi = e.limit_tension;
        ev(internal_energy, k, c) = (p + (gamma * pi)) / ((gamma - 1.) * rho);
        ev(sound_speed, k, c) = std::sqrt((gamma / rho) * (p + pi));
      } else {  // PerfectGasEOSService.cc:12-37
        ev(internal_energy, k, c) = p / ((gamma - 1.) * rho);
        ev(sound_speed, k, c) = std::sqrt(gamma * p / rho);
      }
    }
  }
  // :154-170 mixed-cell averages (density and pressure are not zeroed first)
  for (long c = 0; c < nc && !opt.sans_lagrange; ++c) {
    if (nbenv[c] == 1) continue;
    internal_energy.g[c] = 0.;
    for (int k = 0; k < m; ++k) {
      if (!member(k, c)) continue;
      internal_energy.g[c] += ev(internal_energy, k, c) * ev(mass_fraction, k, c);
      sound_speed.g[c] = std::max(ev(sound_speed, k, c), sound_speed.g[c]);
      density.g[c] += ev(density, k, c) * ev(fracvol, k, c);
      pressure.g[c] += ev(pressure, k, c) * ev(fracvol, k, c);
    }
  }
  euler_volume = cell_volume.g;  // :174-178
  InitGeometricValues();         // :195
  // :199-209 volume recomputed from the CQS
  for (long c = 0; c < nc; ++c) {
    double volume = 0.0;
    for (int l = 0; l < nnpc; ++l) volume += dot(node_coord[cell_nodes[c * nnpc + l]], cqs[c * nnpc + l]);
    volume /= dim;
    cell_volume.g[c] = volume;
  }
}

// MahycoAnnexe.cc:11-142 (environment registration, cell centres, initMat, membership).
void OracleCase::hydroStartInitEnvAndMat() {
  const double one_over_nbnode = dim == 2 ? .25 : .125;
  for (long c = 0; c < nc; ++c) {  // :75-82
    R3 somme{0., 0., 0.};
    for (int l = 0; l < nnpc; ++l) somme += node_coord[cell_nodes[c * nnpc + l]];
    cell_coord[c] = one_over_nbnode * somme;
  }
  // casModel()->initMat: sets Materiau (SODService.cc:41-69, SEDOVService.cc:11-38,
  // POINTTRIPLEService.cc:44-62); RIEMANN: geometric overlap
  riemann_f1.assign(nc, 0.0);
  std::vector<double>& f1 = riemann_f1;
  for (long c = 0; c < nc; ++c) {
    double mat = 0.;
    if (opt.cas_model == MH_CASE_SOD) {
      int t = opt.cas_test;
      if (t == 13 || t == 14 || t == 15) {
        double r = t == 13 ? cell_coord[c].x : t == 14 ? cell_coord[c].y : cell_coord[c].z;
        mat = (r < 0.5) ? 0. : 1.;
      }
    } else if (opt.cas_model == MH_CASE_POINTTRIPLE) {
      if (opt.cas_test != 2) {
        if (cell_coord[c].x <= 0.01) mat = 0.;
        else mat = (cell_coord[c].y <= 0.015) ? 1. : 2.;
      }
    } else if (opt.cas_model == MH_CASE_RIDER) {  // RIDERService::initMat :19-53
      if (opt.cas_test >= 31) {  // the Mono* tests (23-30) put every cell in environment 0
        double rmin, rmax;
        riderRadii(c, opt.cas_test < 34, rmin, rmax);
        const double rb = opt.rider_rayon_bulle;
        if (rmax < rb) mat = 1.;
        else if ((rmax >= rb) && (rmin < rb)) mat = (rb - rmin) / (rmax - rmin);
      }
    } else if (opt.cas_model == MH_CASE_ADVECTION) {  // ADVECTIONService::initMat :11-65
      if (opt.cas_test >= 18) {  // the Mono* tests (7-10) put every cell in environment 0
        const R3 Xb = opt.cas_test < 21 ? R3{0.50, 0.50, 0.} : R3{0.50, 0.75, 0.};
        const double rb = 0.25;  // a constant here, the rayon-bulle option in initVar
        if (std::fabs(cell_coord[c].x - Xb.x) < rb && std::fabs(cell_coord[c].y - Xb.y) < rb) mat = 1.;
      }
    } else if (opt.cas_model == MH_CASE_RIEMANN) {
      int a = opt.interface_axis;
      int n0 = cell_nodes[c * nnpc + 0];
      int n1 = a == 0 ? cell_nodes[c * nnpc + 1] : a == 1 ? cell_nodes[c * nnpc + 3] : cell_nodes[c * nnpc + 4];
      double lo = comp(node_coord[n0], a), hi = comp(node_coord[n1], a);
      double fr = (hi - opt.interface_position) / (hi - lo);
      fr = fr < 0. ? 0. : (fr > 1. ? 1. : fr);
      f1[c] = fr;
      mat = fr;
    }
    materiau[c] = mat;
  }
  // :97-139 membership from Materiau
  for (long c = 0; c < nc; ++c) {
    double v = materiau[c];
    uint8_t bits;
    if (v == 0.) bits = 1;
    else if (v > 0. && v < 1.) bits = 3;
    else if (v == 1.) bits = 2;
    else if (v > 1. && v < 2.) bits = 6;
    else if (v == 2.) bits = 4;
    else if (v > 2. && v < 3.) bits = 12;
    else if (v == 3.) bits = 8;
    else bits = 15;
    bits &= (uint8_t)((1u << m) - 1u);
    mask[c] = bits;
    nbenv[c] = (int8_t)__builtin_popcount(bits);
  }
}

// Case initial conditions: initVar of the case services.
void OracleCase::initMatAndVar() {
  if (opt.cas_model == MH_CASE_SOD) {
    int t = opt.cas_test;
    if (t == 3 || t == 4 || t == 5) {  // SODService::initVarMono :11-40
      for (long c = 0; c < nc; ++c) {
        double r = t == 3 ? cell_coord[c].x : t == 4 ? cell_coord[c].y : cell_coord[c].z;
        double pInit, rhoInit;
        if (r < 0.5) {
          pInit = env[0].pression_initiale;
          rhoInit = env[0].densite_initiale;
        } else {
          pInit = env[0].pression_initiale / 10.;
          rhoInit = env[0].densite_initiale / 8.;
        }
        density.g[c] = rhoInit;
        pressure.g[c] = pInit;
      }
    } else {  // SODService::initVar :71-121
      for (long c = 0; c < nc; ++c) {
        int s = (materiau[c] == 0) ? 0 : 1;
        density.g[c] = env[s].densite_initiale;
        pressure.g[c] = env[s].pression_initiale;
        if (nbenv[c] != 1)
          for (int k = 0; k < m; ++k)
            if (member(k, c)) {
              ev(density, k, c) = env[s].densite_initiale;
              ev(pressure, k, c) = env[s].pression_initiale;
            }
      }
    }
  } else if (opt.cas_model == MH_CASE_SEDOV) {  // SEDOVService::initVarMono :39-86
    const double rhoInit = env[0].densite_initiale, pInit = env[0].pression_initiale;
    const double e1 = 0.244816e-5;
    const double total_energy_deposit = opt.sedov_deposit;
    const double rmin = 1.e-10;
    for (long c = 0; c < nc; ++c) {
      bool isCenterCell = false;
      internal_energy.g[c] = e1;
      density.g[c] = rhoInit;
      pressure.g[c] = pInit;
      double vol = cell_volume.g[c];
      for (int l = 0; l < nnpc; ++l) {
        R3 x = node_coord[cell_nodes[c * nnpc + l]];
        double rnode;
        if (dim == 3) rnode = std::sqrt((x.x - 0.) * (x.x - 0.) + (x.y - 0.) * (x.y - 0.) + (x.z - 0.) * (x.z - 0.));
        else rnode = std::sqrt((x.x - 0.) * (x.x - 0.) + (x.y - 0.) * (x.y - 0.));
        if (rnode < rmin) isCenterCell = true;
      }
      if (isCenterCell) {
        internal_energy.g[c] += total_energy_deposit / vol;
        pressure.g[c] = 0.4 * rhoInit * internal_energy.g[c];
      }
    }
  } else if (opt.cas_model == MH_CASE_POINTTRIPLE) {  // POINTTRIPLEService::initVar :64-128
    for (long c = 0; c < nc; ++c) {
      int s = cell_coord[c].x <= 0.01 ? 0 : (cell_coord[c].y <= 0.015 ? 1 : 2);
      density.g[c] = env[s].densite_initiale;
      pressure.g[c] = env[s].pression_initiale;
      if (nbenv[c] != 1)
        for (int k = 0; k < m; ++k)
          if (member(k, c) && k == s) {
            ev(density, k, c) = env[s].densite_initiale;
            ev(pressure, k, c) = env[s].pression_initiale;
          }
    }
  } else if (opt.cas_model == MH_CASE_RIEMANN) {
    // Two-state tube with geometric-overlap fractions (no reference counterpart):
    // env k takes (densite_initiale, pression_initiale) of env k; mass fractions are
    // f_k rho_k / sum f_j rho_j; the cell density, pressure and sound speed of mixed
    // cells start at 0 so that the averaging of hydroStartInit yields sums.
    for (long c = 0; c < nc; ++c) {
      if (nbenv[c] != 1) {
        fracvol.p[c * MH_MAX_ENV + 0] = 1.0 - riemann_f1[c];
        fracvol.p[c * MH_MAX_ENV + 1] = riemann_f1[c];
      }
      if (nbenv[c] == 1) {
        int k = mask[c] == 1 ? 0 : 1;
        density.g[c] = env[k].densite_initiale;
        pressure.g[c] = env[k].pression_initiale;
      } else {
        double ft = 0.;
        for (int k = 0; k < 2; ++k)
          ft += fracvol.p[c * MH_MAX_ENV + k] * env[k].densite_initiale;
        for (int k = 0; k < 2; ++k) {
          ev(density, k, c) = env[k].densite_initiale;
          ev(pressure, k, c) = env[k].pression_initiale;
          ev(mass_fraction, k, c) = fracvol.p[c * MH_MAX_ENV + k] * env[k].densite_initiale / ft;
        }
        density.g[c] = 0.;
        pressure.g[c] = 0.;
        sound_speed.g[c] = 0.;
      }
    }
  }
  if (opt.cas_model == MH_CASE_RIDER) {  // RIDERService::initVarMono :58-158, initVar :163-290
    const bool mono = opt.cas_test <= 30;
    const bool lin = opt.rider_densite_lineaire != 0;
    const double rb = opt.rider_rayon_bulle;
    const R3 Xb = (opt.cas_test < (mono ? 26 : 34)) ? R3{0.20, 0.20, 0.} : R3{0.50, 0.75, 0.};
    const double croissance = lin ? 1. / rb : 1.;
    for (long c = 0; c < nc; ++c) {
      double rmin, rmax;
      riderRadii(c, opt.cas_test < (mono ? 26 : 34), rmin, rmax);
      density.g[c] = env[0].densite_initiale;
      pressure.g[c] = env[0].pression_initiale;
      internal_energy.g[c] = env[0].energie_initiale;
      const double r = std::sqrt((cell_coord[c].x - Xb.x) * (cell_coord[c].x - Xb.x) +
                                 (cell_coord[c].y - Xb.y) * (cell_coord[c].y - Xb.y));
      const double r_cell = lin ? r : rmax;
      if (r_cell < rb) {
        density.g[c] = env[1].densite_initiale;
        if (lin) density.g[c] = env[1].densite_initiale * std::min(croissance * 2 * (rb - r), 1.);
        pressure.g[c] = env[1].pression_initiale;
      } else if ((rmax >= rb) && (rmin < rb)) {
        if (!lin) {
          const double frac_b = (rb - rmin) / (rmax - rmin);
          if (!mono)
            for (int k = 0; k < m && k < 2; ++k) {
              if (!member(k, c)) continue;
              ev(density, k, c) = env[k == 0 ? 1 : 0].densite_initiale;
              ev(fracvol, k, c) = k == 0 ? frac_b : 1. - frac_b;
              ev(mass_fraction, k, c) = k == 0 ? frac_b : 1. - frac_b;
              ev(pressure, k, c) = 0.;
            }
          density.g[c] = frac_b * env[1].densite_initiale;
          pressure.g[c] = env[1].pression_initiale;
        }
      }
    }
  }
  if (opt.cas_model == MH_CASE_ADVECTION) {  // ADVECTIONService::initVarMono :67-145, initVar :146-246
    // a square of half side rayon-bulle; no mixed cells (the cut-cell branch is commented
    // out in the reference); the internal energy is left as allocated (0)
    const bool mono = opt.cas_test <= 10;
    const bool lin = opt.rider_densite_lineaire != 0;
    const double rb = opt.rider_rayon_bulle;
    const R3 Xb = (opt.cas_test < (mono ? 10 : 21)) ? R3{0.50, 0.50, 0.} : R3{0.50, 0.75, 0.};
    const double croissance = lin ? 1. / rb : 1.;
    for (long c = 0; c < nc; ++c) {
      density.g[c] = env[0].densite_initiale;
      pressure.g[c] = env[0].pression_initiale;
      const double r = std::sqrt((cell_coord[c].x - Xb.x) * (cell_coord[c].x - Xb.x) +
                                 (cell_coord[c].y - Xb.y) * (cell_coord[c].y - Xb.y));
      const double rx = std::fabs(cell_coord[c].x - Xb.x), ry = std::fabs(cell_coord[c].y - Xb.y);
      if ((rx < rb) && (ry < rb)) {
        density.g[c] = env[1].densite_initiale;
        if (lin) density.g[c] = env[1].densite_initiale * croissance * r;
        pressure.g[c] = env[1].pression_initiale;
      }
    }
  }
  if (opt.cas_model == MH_CASE_RIDER || opt.cas_model == MH_CASE_ADVECTION) {
    const R3 cc{0.5, 0.5, 0.};
    const double Pi = 3.14159265359;  // TypesMahyco.h:42
    // one numbering of the velocity fields: 23 Tx, 24 Ty, 25 T45, 26 rotation, 27/29 vortex,
    // 28/30 deformation (RIDER Mono/bi pairs share a field; ADVECTION 7-10 and 18-21 are
    // its first four, ADVECTIONService.cc:131-144, 233-246)
    const int t = opt.cas_model == MH_CASE_ADVECTION ? (opt.cas_test <= 10 ? opt.cas_test : opt.cas_test - 11) + 16
                  : opt.cas_test <= 30                ? opt.cas_test
                                                      : opt.cas_test - 8;
    for (long n = 0; n < nn; ++n) {
      R3 u{0., 0., 0.};
      const R3 x = node_coord[n];
      if (t == 23) u.x = 1.;
      if (t == 24) u.y = 1.;
      if (t == 25) u = R3{1.0, 1.0, 0.0};
      // sin and cos of one argument come from one sincos call, as GCC compiles the pairs of
      // RIDERService.cc:141-160 at -O1 and above (glibc's sincos and sin can differ by an ulp)
      double s0, c0, s1, c1;
      if (t == 26) {
        const R3 dd = x - cc;
        const double theta = std::atan2(dd.y, dd.x);
        const double r = std::sqrt(dd.x * dd.x + dd.y * dd.y);
        const double omega = 4. * Pi;
        ::sincos(omega * 0. + theta, &s0, &c0);
        u.x = -r * omega * s0;
        u.y = r * omega * c0;
      }
      if (t == 27 || t == 29) {
        const R3 dd = x;
        ::sincos(Pi * dd.x, &s0, &c0);
        ::sincos(Pi * dd.y, &s1, &c1);
        u.x = -2. * c1 * s1 * s0 * s0;
        u.y = 2. * c0 * s0 * s1 * s1;
      }
      if (t == 28 || t == 30) {
        const R3 dd = x + cc;
        ::sincos(4. * Pi * dd.x, &s0, &c0);
        ::sincos(4. * Pi * dd.y, &s1, &c1);
        u.x = s0 * s1;
        u.y = c0 * c1;
      }
      velocity[n] = u;
    }
    velocity_n = velocity;
    return;
  }
  for (long n = 0; n < nn; ++n) {
    const double* v0 = env[0].vitesse_initiale;
    velocity[n] = (opt.cas_model == MH_CASE_SEDOV) ? R3{0., 0., 0.} : R3{v0[0], v0[1], v0[2]};
  }
}

// Smallest and largest distance from the bubble centre (0.2, 0.2) or (0.5, 0.75) to the
// nodes of cell c, in the plane (RIDERService.cc:33-41).
void OracleCase::riderRadii(long c, bool low_centre, double& rmin, double& rmax) const {
  const R3 Xb = low_centre ? R3{0.20, 0.20, 0.} : R3{0.50, 0.75, 0.};
  rmin = 10.;
  rmax = 0.;
  for (int l = 0; l < nnpc; ++l) {
    const R3 x = node_coord[cell_nodes[c * nnpc + l]];
    const double rnode = std::sqrt((x.x - Xb.x) * (x.x - Xb.x) + (x.y - Xb.y) * (x.y - Xb.y));
    rmin = std::min(rmin, rnode);
    rmax = std::max(rmax, rnode);
  }
}

// MahycoModule::InitGeometricValues, MahycoModule.cc:1041-1093.
void OracleCase::InitGeometricValues() {
  node_coord_0 = node_coord;
  const double one_over_nbnode = dim == 2 ? .5 : .25;
  for (long f = 0; f < nf; ++f) {
    const int* fn = &face_nodes[f * 4];
    R3 nrm;
    if (dim == 3) {
      R3 v1 = node_coord[fn[2]] - node_coord[fn[0]];
      R3 v2 = node_coord[fn[3]] - node_coord[fn[1]];
      nrm.x = v1.y * v2.z - v1.z * v2.y;
      nrm.y = -(v2.x * v1.z - v2.z * v1.x);
      nrm.z = v1.x * v2.y - v1.y * v2.x;
    } else {
      nrm.x = node_coord[fn[1]].y - node_coord[fn[0]].y;
      nrm.y = node_coord[fn[1]].x - node_coord[fn[0]].x;
      nrm.z = 0.;
    }
    double nl = normL2(nrm);
    face_normal[f] = {nrm.x / nl, nrm.y / nl, nrm.z / nl};
    R3 fc{0., 0., 0.};
    for (int l = 0; l < face_nbnode[f]; ++l) fc += one_over_nbnode * node_coord[fn[l]];
    face_coord[f] = fc;
  }
  for (long c = 0; c < nc; ++c)
    for (int l = 0; l < nfpc; ++l) {
      R3 d = face_coord[cell_faces[c * nfpc + l]] - cell_coord[c];
      double nl = normL2(d);
      outer_face_normal[c * nfpc + l] = {d.x / nl, d.y / nl, d.z / nl};
    }
}

// MahycoModule::PrepareFaceGroup, MahycoAnnexe.cc:149-235 -> node lists per group.
void OracleCase::PrepareFaceGroup() {
  R3 maxC{0., 0., 0.}, minC{100., 100., 100.};
  for (long n = 0; n < nn; ++n) {
    maxC.x = std::max(maxC.x, node_coord[n].x);
    maxC.y = std::max(maxC.y, node_coord[n].y);
    maxC.z = std::max(maxC.z, node_coord[n].z);
    minC.x = std::min(minC.x, node_coord[n].x);
    minC.y = std::min(minC.y, node_coord[n].y);
    minC.z = std::min(minC.z, node_coord[n].z);
  }
  std::vector<char> in(nn);
  for (int g = 0; g < 6; ++g) {
    std::fill(in.begin(), in.end(), 0);
    int d = g / 2;
    double ref = (g % 2 == 0) ? comp(minC, d) : comp(maxC, d);
    for (long f = 0; f < nf; ++f) {
      bool flag = true;
      for (int l = 0; l < face_nbnode[f]; ++l)
        if (std::fabs(comp(node_coord[face_nodes[f * 4 + l]], d) - ref) > opt.threshold) flag = false;
      if (flag)
        for (int l = 0; l < face_nbnode[f]; ++l) in[face_nodes[f * 4 + l]] = 1;
    }
    group_nodes[g].clear();
    for (long n = 0; n < nn; ++n)
      if (in[n]) group_nodes[g].push_back((int)n);
  }
}

// MahycoModule::computeCellMass :221-237 and computeNodeMass :247-262.
void OracleCase::computeCellMass() {
  for (long c = 0; c < nc; ++c) cell_mass.g[c] = density.g[c] * cell_volume.g[c];
  for (int k = 0; k < m; ++k)
    for (long c = 0; c < nc; ++c)
      if (member(k, c)) ev(cell_mass, k, c) = ev(mass_fraction, k, c) * cell_mass.g[c];
}
void OracleCase::computeNodeMass() {
  const double one_over_nbnode = dim == 2 ? .25 : .125;
#pragma omp parallel for schedule(static)
  for (long n = 0; n < nn; ++n) {
    double s = 0.;
    for (int a = nc_off[n]; a < nc_off[n + 1]; ++a) s += one_over_nbnode * cell_mass.g[nc_cell[a]];
    node_mass[n] = s;
  }
}

// ============================================================================ Lagrange
// MahycoModule::saveValuesAtN, MahycoModule.cc:311-382.
void OracleCase::saveValuesAtN() {
  global_old_deltat = old_deltat;
  pseudo_nm1 = pseudo_n;
  pseudo_n = pseudo;
  pressure_n = pressure;
  cell_volume_n = cell_volume;
  density_n = density;
  internal_energy_n = internal_energy;
  cqs_n = cqs;
  if (!opt.sans_lagrange) velocity_n = velocity;  // :373-375
}

// MahycoModule::updateVelocityWithoutLagrange, MahycoModule.cc:702-718: the prescribed
// velocity, optionally reversed in time by cos(pi t parametre).
void OracleCase::updateVelocityWithoutLagrange() {
  const double option_real = opt.rider_reverse ? 1. : 0.;
  const double factor = opt.rider_parametre;
  const double Pi = 3.14159265359;  // TypesMahyco.h:42
  const double cs = std::cos(Pi * global_time * factor);
  for (long n = 0; n < nn; ++n) {  // u_n (1 - option) + (option u_n) cos, per component
    const R3 u = velocity_n[n];
    velocity[n] = R3{u.x * (1. - option_real) + option_real * u.x * cs, u.y * (1. - option_real) + option_real * u.y * cs,
                     u.z * (1. - option_real) + option_real * u.z * cs};
  }
}

// MahycoModule::updateVelocity, MahycoModule.cc:438-562 (elasto force of NoModel adds +-0;
// gravity 0; anti-drift off with projection).  With CSTS and a remap, updateVelocityBackward
// (:574-627) first takes the projected velocity back to n-1/2 with the same force.
void OracleCase::updateVelocity() {
  const double dt = 0.5 * (global_old_deltat + global_deltat);
  const bool backward = opt.schema_csts && opt.with_projection;  // :451-453
  const double dtb = -0.5 * global_old_deltat;
#pragma omp parallel for schedule(static)
  for (long n = 0; n < nn; ++n) {
    R3 F{0., 0., 0.};
    for (int a = nc_off[n]; a < nc_off[n + 1]; ++a) {
      int c = nc_cell[a];
      double pr = pressure_n.g[c] + pseudo_n.g[c];
      F += pr * cqs_n[(long)c * nnpc + nc_local[a]];
    }
    if (backward) velocity[n] += (dtb / node_mass[n]) * F;
    velocity[n] += (dt / node_mass[n]) * F;
  }
}

// MahycoModule::updateVelocityForward, MahycoModule.cc:638-690: before a CSTS remap, the
// velocity goes on from n+1/2 to n+1 with the force of this cycle's pressure array (not yet
// updated by the EOS when remapping), pseudo-viscosity and corner vectors.
void OracleCase::updateVelocityForward() {
  const double dt = 0.5 * global_deltat;
#pragma omp parallel for schedule(static)
  for (long n = 0; n < nn; ++n) {
    R3 F{0., 0., 0.};
    for (int a = nc_off[n]; a < nc_off[n + 1]; ++a) {
      int c = nc_cell[a];
      double pr = pressure.g[c] + pseudo.g[c];
      F += pr * cqs[(long)c * nnpc + nc_local[a]];
    }
    velocity[n] += (dt / node_mass[n]) * F;
  }
}

// MahycoModule::applyBoundaryCondition, MahycoModule.cc:756-807 (Vx/Vy/Vz).
void OracleCase::applyBoundaryCondition() {
  for (int b = 0; b < opt.nb_bc; ++b) {
    const MhBoundaryCondition& bc = opt.bc[b];
    for (int n : group_nodes[bc.surface]) compr(velocity[n], bc.type) = bc.value;
  }
}

// MahycoModule::updatePosition, MahycoModule.cc:730-751.
void OracleCase::updatePosition() {
  const double deltat = global_deltat;
#pragma omp parallel for schedule(static)
  for (long n = 0; n < nn; ++n)
    if (!opt.sans_lagrange || nc_off[n + 1] - nc_off[n] == 4)  // :737 (advection: 4-cell nodes only)
      node_coord[n] += deltat * velocity[n];
  const double one_over_nbnode = dim == 2 ? .25 : .125;
#pragma omp parallel for schedule(static)
  for (long c = 0; c < nc; ++c) {
    R3 somme{0., 0., 0.};
    for (int l = 0; l < nnpc; ++l) somme += node_coord[cell_nodes[c * nnpc + l]];
    cell_coord[c] = one_over_nbnode * somme;
  }
}

// MahycoModule::computeCQs, MahycoModule.cc:1869-1932.
void OracleCase::computeCQs(const R3* x, const R3* fc, R3* q) {
  const R3 c0 = fc[0], c1 = fc[1], c2 = fc[2], c3 = fc[3], c4 = fc[4], c5 = fc[5];
  const R3 n1a04 = 0.5 * vecMul(x[0] - c0, x[3] - c0);
  const R3 n1a03 = 0.5 * vecMul(x[3] - c0, x[2] - c0);
  const R3 n1a02 = 0.5 * vecMul(x[2] - c0, x[1] - c0);
  const R3 n1a01 = 0.5 * vecMul(x[1] - c0, x[0] - c0);
  const R3 n2a05 = 0.5 * vecMul(x[0] - c1, x[4] - c1);
  const R3 n2a12 = 0.5 * vecMul(x[4] - c1, x[7] - c1);
  const R3 n2a08 = 0.5 * vecMul(x[7] - c1, x[3] - c1);
  const R3 n2a04 = 0.5 * vecMul(x[3] - c1, x[0] - c1);
  const R3 n3a01 = 0.5 * vecMul(x[0] - c2, x[1] - c2);
  const R3 n3a06 = 0.5 * vecMul(x[1] - c2, x[5] - c2);
  const R3 n3a09 = 0.5 * vecMul(x[5] - c2, x[4] - c2);
  const R3 n3a05 = 0.5 * vecMul(x[4] - c2, x[0] - c2);
  const R3 n4a09 = 0.5 * vecMul(x[4] - c3, x[5] - c3);
  const R3 n4a10 = 0.5 * vecMul(x[5] - c3, x[6] - c3);
  const R3 n4a11 = 0.5 * vecMul(x[6] - c3, x[7] - c3);
  const R3 n4a12 = 0.5 * vecMul(x[7] - c3, x[4] - c3);
  const R3 n5a02 = 0.5 * vecMul(x[1] - c4, x[2] - c4);
  const R3 n5a07 = 0.5 * vecMul(x[2] - c4, x[6] - c4);
  const R3 n5a10 = 0.5 * vecMul(x[6] - c4, x[5] - c4);
  const R3 n5a06 = 0.5 * vecMul(x[5] - c4, x[1] - c4);
  const R3 n6a03 = 0.5 * vecMul(x[2] - c5, x[3] - c5);
  const R3 n6a08 = 0.5 * vecMul(x[3] - c5, x[7] - c5);
  const R3 n6a11 = 0.5 * vecMul(x[7] - c5, x[6] - c5);
  const R3 n6a07 = 0.5 * vecMul(x[6] - c5, x[2] - c5);
  const double t = 1. / 12.;
  q[0] = t * (5. * (n1a01 + n1a04 + n2a04 + n2a05 + n3a05 + n3a01) + (n1a02 + n1a03 + n2a08 + n2a12 + n3a06 + n3a09));
  q[1] = t * (5. * (n1a01 + n1a02 + n3a01 + n3a06 + n5a06 + n5a02) + (n1a04 + n1a03 + n3a09 + n3a05 + n5a10 + n5a07));
  q[2] = t * (5. * (n1a02 + n1a03 + n5a07 + n5a02 + n6a07 + n6a03) + (n1a01 + n1a04 + n5a06 + n5a10 + n6a11 + n6a08));
  q[3] = t * (5. * (n1a03 + n1a04 + n2a08 + n2a04 + n6a08 + n6a03) + (n1a01 + n1a02 + n2a05 + n2a12 + n6a07 + n6a11));
  q[4] = t * (5. * (n2a05 + n2a12 + n3a05 + n3a09 + n4a09 + n4a12) + (n2a08 + n2a04 + n3a01 + n3a06 + n4a10 + n4a11));
  q[5] = t * (5. * (n3a06 + n3a09 + n4a09 + n4a10 + n5a10 + n5a06) + (n3a01 + n3a05 + n4a12 + n4a11 + n5a07 + n5a02));
  q[6] = t * (5. * (n4a11 + n4a10 + n5a10 + n5a07 + n6a07 + n6a11) + (n4a12 + n4a09 + n5a06 + n5a02 + n6a03 + n6a08));
  q[7] = t * (5. * (n2a08 + n2a12 + n4a12 + n4a11 + n6a11 + n6a08) + (n2a04 + n2a05 + n4a09 + n4a10 + n6a07 + n6a03));
}

// MahycoModule::computeGeometricValues, MahycoModule.cc:1097-1253 ("racine-cubique-volume").
int OracleCase::computeGeometricValues() {
  const double racine = dim == 2 ? .5 : 1. / 3.;
  int err = 0;
#pragma omp parallel for schedule(static) reduction(| : err)
  for (long c = 0; c < nc; ++c) {
    R3 coord[8];
    for (int l = 0; l < nnpc; ++l) coord[l] = node_coord[cell_nodes[c * nnpc + l]];
    R3* q = &cqs[c * nnpc];
    if (dim == 3) {
      R3 fc[6];
      fc[0] = 0.25 * (coord[0] + coord[3] + coord[2] + coord[1]);
      fc[1] = 0.25 * (coord[0] + coord[4] + coord[7] + coord[3]);
      fc[2] = 0.25 * (coord[0] + coord[1] + coord[5] + coord[4]);
      fc[3] = 0.25 * (coord[4] + coord[5] + coord[6] + coord[7]);
      fc[4] = 0.25 * (coord[1] + coord[2] + coord[6] + coord[5]);
      fc[5] = 0.25 * (coord[2] + coord[3] + coord[7] + coord[6]);
      computeCQs(coord, fc, q);
    } else {
      R3 npc[5];
      coord[4] = coord[0];
      for (int l = 0; l < 4; ++l) {
        npc[l + 1].x = 0.5 * (coord[l + 1].y - coord[l].y);
        npc[l + 1].y = 0.5 * (coord[l].x - coord[l + 1].x);
        npc[l + 1].z = 0.;
      }
      npc[0] = npc[4];
      for (int l = 0; l < 4; ++l) q[l] = npc[l + 1] + npc[l];
    }
    double volume = 0.;
    for (int l = 0; l < nnpc; ++l) volume += dot(node_coord[cell_nodes[c * nnpc + l]], q[l]);
    volume /= dim;
    cell_volume.g[c] = volume;
    if (volume < 0.) err = 1;
    const int lc = opt.longueur_caracteristique;  // MahycoModule.cc:1195-1236
    const R3* cn = nullptr;
    R3 cl[8];
    for (int l = 0; l < nnpc; ++l) cl[l] = node_coord[cell_nodes[c * nnpc + l]];
    cn = cl;
    if (lc == MH_L_FACES_OPPOSEES) {
      const R3 f0 = 0.25 * (cn[0] + cn[3] + cn[2] + cn[1]), f1 = 0.25 * (cn[0] + cn[4] + cn[7] + cn[3]);
      const R3 f2 = 0.25 * (cn[0] + cn[1] + cn[5] + cn[4]), f3 = 0.25 * (cn[4] + cn[5] + cn[6] + cn[7]);
      const R3 f4 = 0.25 * (cn[1] + cn[2] + cn[6] + cn[5]), f5 = 0.25 * (cn[2] + cn[3] + cn[7] + cn[6]);
      const double d1 = normL2(f0 - f3), d2 = normL2(f2 - f5), d3 = normL2(f1 - f4);
      caracteristic_length[c] = (d1 * d2 * d3) / (d1 * d2 + d1 * d3 + d2 * d3);
    } else if (lc == MH_L_MONODIM_X) {
      caracteristic_length[c] = std::fabs(cn[2].x - cn[0].x);
    } else if (lc == MH_L_MONODIM_Y) {
      caracteristic_length[c] = std::fabs(cn[2].y - cn[0].y);
    } else if (lc == MH_L_MONODIM_Z) {
      caracteristic_length[c] = std::fabs(cn[2].z - cn[0].z);
    } else {
      caracteristic_length[c] = g_pow_mode == 0 ? std::pow(cell_volume.g[c], racine)
                                : (dim == 3 ? mh_pow_third(cell_volume.g[c]) : mh_pow_half(cell_volume.g[c]));
    }
  }
  // mixed cells: V_k = f_k V
  for (long c = 0; c < nc; ++c) {
    if (nbenv[c] == 1) continue;
    for (int k = 0; k < m; ++k)
      if (member(k, c)) ev(cell_volume, k, c) = ev(fracvol, k, c) * cell_volume.g[c];
  }
  return err ? MH_ERR_NEG_VOLUME : MH_OK;
}

// MahycoModule::updateDensity, MahycoModule.cc:1267-1316 (density-damage threshold 0: no clamp).
void OracleCase::updateDensity() {
  for (int k = 0; k < m; ++k) {
#pragma omp parallel for schedule(static)
    for (long c = 0; c < nc; ++c) {
      if (!member(k, c)) continue;
      double new_density = ev(cell_mass, k, c) / ev(cell_volume, k, c);
      ev(density, k, c) = new_density;
      ev(tau_density, k, c) = 0.5 * (1.0 / ev(density_n, k, c) + 1.0 / ev(density, k, c));
    }
  }
#pragma omp parallel for schedule(static)
  for (long c = 0; c < nc; ++c) {
    double new_density = cell_mass.g[c] / cell_volume.g[c];
    density.g[c] = new_density;
    tau_density.g[c] = 0.5 * (1.0 / density_n.g[c] + 1.0 / density.g[c]);
    div_u[c] = 1.0 / global_deltat * (1.0 / density.g[c] - 1.0 / density_n.g[c]) / tau_density.g[c];
  }
}

// MahycoModule::computeArtificialViscosity, MahycoModule.cc:386-426.
void OracleCase::computeArtificialViscosity() {
  for (int k = 0; k < m; ++k) {
    const double a1 = env[k].linear_pseudo_coeff, a2 = env[k].quadratic_pseudo_coeff;
#pragma omp parallel for schedule(static)
    for (long c = 0; c < nc; ++c) {
      if (!member(k, c)) continue;
      ev(pseudo, k, c) = 0.;
      if (div_u[c] < 0.0) {
        const double L = caracteristic_length[c];
        ev(pseudo, k, c) = 1. / ev(tau_density, k, c) *
                           (-a1 * L * sound_speed.g[c] * div_u[c] + a2 * L * L * div_u[c] * div_u[c]);
      }
    }
  }
#pragma omp parallel for schedule(static)
  for (long c = 0; c < nc; ++c) {
    if (nbenv[c] == 1) continue;
    pseudo.g[c] = 0.;
    for (int k = 0; k < m; ++k)
      if (member(k, c)) pseudo.g[c] += ev(pseudo, k, c) * ev(fracvol, k, c);
  }
}

// MahycoModule::updateEnergyAndPressure :1430-1480 -> updateEnergyAndPressureforGP :1623-1659
// (no CSTS, pseudo-centree off; the gamma of the env EOS for StiffenedGas too).
// MahycoModule::fvnr / fvnrderiv, MahycoModule.h:378-386: the energy equation of the
// Lagrange step and its derivative in e.
static inline double fvnr(double e, double p, double en, double qnn1, double pn, double rn1, double rn) {
  return e - en + 0.5 * (p + pn + 2. * qnn1) * (1. / rn1 - 1. / rn);
}
static inline double fvnrderiv(double dpde, double rn1, double rn) { return 1. + 0.5 * dpde * (1. / rn1 - 1. / rn); }

// updateEnergyAndPressure :1430-1480 with updateEnergyAndPressureforGP :1622-1660 (non-CSTS,
// pseudo not centred), updateEnergyAndPressurebyNewton :1506-1540 and
// updateEnergyAndPressureExplicite :1484-1502.  Newton leaves the pressure and sound speed
// of its last EOS call, which the EOS call after the update (or after the remap) replaces,
// so only the energy is kept here.
void OracleCase::updateEnergyAndPressure() {
  for (int k = 0; k < m; ++k) {
    const double adiabatic_cst = env[k].adiabatic_cst;
    const double pi = env[k].limit_tension;
    const bool sg = env[k].eos == MH_EOS_STIFFENED_GAS;
#pragma omp parallel for schedule(static)
    for (long c = 0; c < nc; ++c) {
      if (!member(k, c)) continue;
      const double rho = ev(density, k, c), rhon = ev(density_n, k, c);
      const double q = ev(pseudo, k, c);
      if (opt.pression_explicite) {
        ev(internal_energy, k, c) -= (ev(pressure_n, k, c) + q) * (1.0 / rho - 1.0 / rhon);
        continue;
      }
      if (opt.schema_csts) {  // updateEnergyAndPressureforGP, CSTS branch (:1661-1697)
        const double dtn = global_deltat, dto = global_old_deltat;
        double cn1 = 0., cn = 0., cdn = 0., cvon = 0.;
        for (int l = 0; l < nnpc; ++l) {
          const long n = cell_nodes[c * nnpc + l];
          const R3 C = cqs[c * nnpc + l], Cn = cqs_n[c * nnpc + l];
          cn1 += dot(velocity[n], C) * dtn;
          cn += dot(velocity[n], Cn) * dtn;
          cdn += dot(velocity[n] - velocity_n[n], Cn) * (dtn - dto);
          cvon += dot(velocity_n[n], Cn) * dto;
        }
        const double M = ev(cell_mass, k, c);
        const double pn = ev(pressure_n, k, c), qn = ev(pseudo_n, k, c), qnm1 = ev(pseudo_nm1, k, c);
        const double denom = 1 + 0.5 * (adiabatic_cst - 1.0) * rho * cn1 / M;
        const double numer = ev(internal_energy, k, c) - (0.5 * (pn + qn) * cn / M) - (0.5 * q * cn1 / M) +
                             (0.25 * (pn + qn) * cdn / M) - (0.5 * (qn - qnm1) * cvon / M);
        ev(internal_energy, k, c) = numer / denom;
        continue;
      }
      double pseudo_v = 0.;
      if (q * (1.0 / rho - 1.0 / rhon) < 0.) pseudo_v = q;
      if (opt.with_newton) {
        const double rn = rhon, pn = ev(pressure_n, k, c), en = ev(internal_energy_n, k, c);
        double e = ev(internal_energy, k, c), p = pn, dpde = 0.;
        int i = 0;
        while (i < 50 && std::fabs(fvnr(e, p, en, pseudo_v, pn, rho, rn)) >= opt.threshold) {
          // applyOneCellEOS (PerfectGasEOSService.cc:71-88, StiffenedGasEOSService.cc:67-83)
          p = sg ? ((adiabatic_cst - 1.) * rho * e) - (adiabatic_cst * pi) : (adiabatic_cst - 1.) * rho * e;
          dpde = (adiabatic_cst - 1.) * rho;
          const double enew = e - fvnr(e, p, en, pseudo_v, pn, rho, rn) / fvnrderiv(dpde, rho, rn);
          e = enew;
          i++;
        }
        ev(internal_energy, k, c) = e;
        continue;
      }
      const double denom = 1 + 0.5 * (adiabatic_cst - 1.0) * rho * (1.0 / rho - 1.0 / rhon);
      const double numer = ev(internal_energy_n, k, c) - (0.5 * ev(pressure, k, c) + pseudo_v) * (1.0 / rho - 1.0 / rhon);
      ev(internal_energy, k, c) = numer / denom;
    }
  }
#pragma omp parallel for schedule(static)
  for (long c = 0; c < nc; ++c) {
    if (nbenv[c] == 1) continue;
    internal_energy.g[c] = 0.;
    for (int k = 0; k < m; ++k)
      if (member(k, c)) internal_energy.g[c] += ev(mass_fraction, k, c) * ev(internal_energy, k, c);
  }
  if (!opt.with_projection) {  // :1465-1479
    for (int k = 0; k < m; ++k) applyEOS(k);
    computePressionMoyenne();
  }
}

// PerfectGasEOSService::applyEOS :45-67, StiffenedGasEOSService::applyEOS :43-63.
void OracleCase::applyEOS(int k) {
  const double gamma = env[k].adiabatic_cst;
  const double pi = env[k].limit_tension;
  const bool sg = env[k].eos == MH_EOS_STIFFENED_GAS;
#pragma omp parallel for schedule(static)
  for (long c = 0; c < nc; ++c) {
    if (!member(k, c)) continue;
    const double e = ev(internal_energy, k, c), rho = ev(density, k, c);
    if (sg) {
      const double p = ((gamma - 1.) * rho * e) - (gamma * pi);
      ev(pressure, k, c) = p;
      ev(sound_speed, k, c) = std::sqrt((gamma / rho) * (p + pi));
    } else {
      const double p = (gamma - 1.) * rho * e;
      ev(pressure, k, c) = p;
      ev(sound_speed, k, c) = std::sqrt(gamma * p / rho);
    }
  }
}

// MahycoModule::computePressionMoyenne, MahycoModule.cc:1708-1735.
void OracleCase::computePressionMoyenne() {
  const double option = opt.calcul_vitson_max ? 1. : 0.;
#pragma omp parallel for schedule(static)
  for (long c = 0; c < nc; ++c) {
    if (nbenv[c] == 1) continue;
    pressure.g[c] = 0.;
    sound_speed.g[c] = 0.;
    for (int k = 0; k < m; ++k) {
      if (!member(k, c)) continue;
      pressure.g[c] += ev(fracvol, k, c) * ev(pressure, k, c);
      sound_speed.g[c] = option * std::max(ev(sound_speed, k, c), sound_speed.g[c]);
      sound_speed.g[c] += (1. - option) * ev(fracvol, k, c) * ev(sound_speed, k, c);
    }
  }
}

// ============================================================================ remap
// MahycoModule::remap, PrepareRemap.cc:203-251.
int OracleCase::remap() {
  if (opt.schema_csts) updateVelocityForward();  // PrepareRemap.cc:210
  computeVariablesForRemap();
  computeFaceQuantitesForRemap();
  // RemapADIService::appliRemap, RemapADIService.cc:15-90
  for (int i = 0; i < dim; ++i) {
    int idir = (i + sens_projection) % dim;
    if (cells_in_dir[idir] == 1) continue;
    computeGradPhiFace(idir);
    computeGradPhiCell(idir);
    computeUpwindFaceQuantitiesForProjection(idir);
    computeUremap(idir);
    if (!opt.sans_lagrange) computeDualUremap(idir);  // withDualProjection (PrepareRemap.cc:211)
  }
  sens_projection = (sens_projection + 1) % dim;
  node_coord_lag.swap(node_coord);  // kept for the state accessors (cqs holds what the loop needs)
  node_coord = node_coord_0;  // isEuler() == "Full"
  int st = remapVariables();
  if (st != MH_OK) return st;
  // PrepareRemap.cc:227-240 Materiau
  for (long c = 0; c < nc; ++c) {
    double mt = 0.;
    for (int k = 0; k < m; ++k)
      if (member(k, c)) mt += k * ev(fracvol, k, c);
    materiau[c] = mt;
  }
  for (int k = 0; k < m; ++k) {
    if (env[k].eos == MH_EOS_FICTIF) applyFictif(k);
    else applyEOS(k);
  }
  if (!opt.sans_lagrange) computePressionMoyenne();  // returns at once in advection mode (:1711)
  return MH_OK;
}

// FictifEOSService::applyEOS, eos/fictif/FictifEOSService.cc:45-75: a pressure that only
// depends on time, a unit sound speed (no influence on the step) and no internal energy.
void OracleCase::applyFictif(int k) {
  const MhEnvironment& e = env[k];
  for (long c = 0; c < nc; ++c) {
    if (!member(k, c)) continue;
    if (global_time > e.fictif_tdebut) ev(pressure, k, c) = e.fictif_p0 + e.fictif_dpdt * global_time;
    if (global_time > e.fictif_tfin) ev(pressure, k, c) = 0.;
    ev(sound_speed, k, c) = 1.;
    ev(internal_energy, k, c) = 0.;
  }
}

// PrepareRemap.cc:94-193 (slots V_k, M_k, E_k and V*q; dual slots u, m, KE).
void OracleCase::computeVariablesForRemap() {
  std::fill(u_lagrange.begin(), u_lagrange.end(), 0.0);
  for (int k = 0; k < m; ++k) {
#pragma omp parallel for schedule(static)
    for (long c = 0; c < nc; ++c) {
      if (!member(k, c)) continue;
      double* U = &u_lagrange[c * nv];
      const double V = ev(cell_volume, k, c), rho = ev(density, k, c);
      U[k] = V;
      U[m + k] = V * rho;
      U[2 * m + k] = V * rho * ev(internal_energy, k, c);
      U[3 * m] = cell_volume.g[c] * pseudo.g[c];
      if (opt.projection_pente_borne) {  // :138-155: the environment's own values
        double* P = &phi_lagrange[c * nv];
        P[k] = ev(fracvol, k, c);
        P[m + k] = rho;
        P[2 * m + k] = ev(internal_energy, k, c);
        P[3 * m] = pseudo.g[c];
      } else {
        for (int v = 0; v < nv; ++v) phi_lagrange[c * nv + v] = U[v] / cell_volume.g[c];
      }
    }
  }
#pragma omp parallel for schedule(static)
  for (long n = 0; n < nn; ++n) {
    double* Ud = &u_dual[n * 5];
    double* Pd = &phi_dual[n * 5];
    const R3 u = velocity[n];
    const double mn = node_mass[n];
    Ud[0] = mn * u.x;
    Ud[1] = mn * u.y;
    Ud[2] = mn * u.z;
    Ud[3] = mn;
    Ud[4] = 0.5 * mn * squareNormL2(u);
    Pd[0] = u.x;
    Pd[1] = u.y;
    Pd[2] = u.z;
    Pd[3] = mn;
    Pd[4] = 0.5 * squareNormL2(u);
  }
}

// MahycoModule::computeFaceQuantitesForRemap, PrepareRemap.cc:18-59 (non-CSTS).
void OracleCase::computeFaceQuantitesForRemap() {
  const double one_over_nbnode = dim == 2 ? .5 : .25;
#pragma omp parallel for schedule(static)
  for (long f = 0; f < nf; ++f) {
    const int* fn = &face_nodes[f * 4];
    if (dim == 3) {
      R3 v1 = node_coord[fn[2]] - node_coord[fn[0]];
      R3 v2 = node_coord[fn[3]] - node_coord[fn[1]];
      face_length_lagrange[f].x = 0.5 * std::fabs(v1.y * v2.z - v1.z * v2.y);
      face_length_lagrange[f].y = 0.5 * std::fabs(v2.x * v1.z - v2.z * v1.x);
      face_length_lagrange[f].z = 0.5 * std::fabs(v1.x * v2.y - v1.y * v2.x);
    } else {
      R3 t = node_coord[fn[1]] - node_coord[fn[0]];
      face_length_lagrange[f] = {std::fabs(t.y), std::fabs(t.x), 0.};
    }
    R3 vm{0., 0., 0.}, fc{0., 0., 0.};
    for (int l = 0; l < face_nbnode[f]; ++l) {
      fc += one_over_nbnode * node_coord[fn[l]];
      if (opt.schema_csts) vm += 0.5 * (velocity[fn[l]] + velocity_n[fn[l]]);  // PrepareRemap.cc:236-240
      else vm += velocity[fn[l]];
    }
    face_coord[f] = fc;
    face_normal_velocity[f] = dot(one_over_nbnode * vm, face_normal[f]);
  }
}

// RemapADIService::computeGradPhiFace, RemapADIService.cc:119-149 (inner faces only).
// h_cell is the sum over the cell's inner faces of the sweep of |x_f - x_c|; with at most
// two terms per cell starting from 0 the face order of the sum does not matter.
void OracleCase::computeGradPhiFace(int idir) {
#pragma omp parallel for schedule(static)
  for (long f = face_off[idir]; f < face_off[idir] + face_cnt[idir]; ++f) {
    int cb = face_prev[f], cf = face_next[f];
    if (cb < 0 || cf < 0) continue;
    deltax_lagrange[f] = dot(cell_coord[cf] - cell_coord[cb], face_normal[f]);
    for (int v = 0; v < nv; ++v)
      grad_phi_face[f * nv + v] = (phi_lagrange[(long)cf * nv + v] - phi_lagrange[(long)cb * nv + v]) / deltax_lagrange[f];
  }
  {
#pragma omp parallel for schedule(static)
    for (long c = 0; c < nc; ++c) {
      double h = 0.;
      for (int l = 0; l < nfpc; ++l) {
        const int f = cell_faces[c * nfpc + l];
        if (face_dir[f] != idir || face_prev[f] < 0 || face_next[f] < 0) continue;
        h += normL2(face_coord[f] - cell_coord[c]);
      }
      h_cell_lagrange[c] = h;
    }
  }
}

// Remap-ordre-3.cc:16-112: evaluate_grad, evaluate_ystar, evaluate_fm / evaluate_fp,
// define_interval, intersection.
static inline double o3_grad(double hm, double h0, double hp, double ym, double y0, double yp) {
  return h0 / (hm + h0 + hp) * ((2. * hm + h0) / (h0 + hp) * (yp - y0) + (h0 + 2. * hp) / (hm + h0) * (y0 - ym));
}
static inline double o3_ystar(double hmm, double hm, double hp, double hpp, double ymm, double ym, double yp, double ypp,
                              double gradm, double gradp) {
  const double tmp1 = (2. * hp * hm) / (hm + hp) * ((hmm + hm) / (2. * hm + hp) - (hpp + hp) / (2. * hp + hm)) * (yp - ym);
  const double tmp2 = -hm * (hmm + hm) / (2. * hm + hp) * gradp + hp * (hp + hpp) / (hm + 2. * hp) * gradm;
  return ym + hm / (hm + hp) * (yp - ym) + 1. / (hmm + hm + hp + hpp) * (tmp1 + tmp2);
}
static inline double o3_fm(double x, double dx, double up, double du, double u6) {
  return up - 0.5 * x / dx * (du - (1. - 2. / 3. * x / dx) * u6);
}
static inline double o3_fp(double x, double dx, double um, double du, double u6) {
  return um + 0.5 * x / dx * (du - (1. - 2. / 3. * x / dx) * u6);
}
// RemapADIService::ComputeFluxOrdre3, Remap-ordre-3.cc:127-189 (face value with TVD clip)
static double flux_ordre3(double ymmm, double ymm, double ym, double yp, double ypp, double yppp, double hmmm,
                          double hmm, double hm, double hp, double hpp, double hppp, double vdt) {
  const double gradmm = o3_grad(hmmm, hmm, hm, ymmm, ymm, ym);
  const double gradm = o3_grad(hmm, hm, hp, ymm, ym, yp);
  const double gradp = o3_grad(hm, hp, hpp, ym, yp, ypp);
  const double gradpp = o3_grad(hp, hpp, hppp, yp, ypp, yppp);
  const double ystarm = o3_ystar(hmmm, hmm, hm, hp, ymmm, ymm, ym, yp, gradmm, gradm);
  const double ystar = o3_ystar(hmm, hm, hp, hpp, ymm, ym, yp, ypp, gradm, gradp);
  const double ystarp = o3_ystar(hm, hp, hpp, hppp, ym, yp, ypp, yppp, gradp, gradpp);
  const double ym_m = ystarm, ym_p = ystar, yp_m = ystar, yp_p = ystarp;
  const double grad_m = ym_p - ym_m, grad_p = yp_p - yp_m;
  const double ym6 = 6. * (ym - 0.5 * (ym_m + ym_p));
  const double yp6 = 6. * (yp - 0.5 * (yp_m + yp_p));
  double flux;
  if (vdt > 0.) flux = o3_fm(vdt, hm, ym_p, grad_m, ym6);
  else if (vdt < 0.) flux = o3_fp(-vdt, hp, yp_m, grad_p, yp6);
  else return 0.;  // vdt == 0 (a NaN vdt leaves flux undefined in the reference; not reached)
  const double num = vdt / hm, nup = vdt / hp;
  const double ym_ym = ym + (1. - num) / num * (ym - ymm);
  const double yp_ym = yp - (1. + nup) / nup * (yp - ypp);
  double I1[2], I2[2];
  I1[0] = amin(ym, yp);
  I1[1] = amax(ym, yp);
  if (vdt >= 0.) {
    I2[0] = amin(ym, ym_ym);
    I2[1] = amax(ym, ym_ym);
  } else {
    I2[0] = amin(yp, yp_ym);
    I2[1] = amax(yp, yp_ym);
  }
  double lo, hi;
  if ((I1[1] < I2[0]) || (I2[1] < I1[0])) {
    lo = 0.;
    hi = 0.;
  } else {
    lo = amax(I1[0], I2[0]);
    hi = amin(I1[1], I2[1]);
  }
  if (flux < lo) flux = lo;
  if (flux > hi) flux = hi;
  return flux;
}

// Neighbour of cell c across its back (back = true) or front face along idir, -1 at the boundary.
int OracleCase::dirNeighbor(long c, int idir, bool back) const {
  for (int l = 0; l < nfpc; ++l) {
    const int f = cell_faces[c * nfpc + l];
    if (face_dir[f] != idir) continue;
    if (back && face_next[f] == c) return face_prev[f];
    if (!back && face_prev[f] == c) return face_next[f];
  }
  return -1;
}

// RemapADIService::fluxLimiter, UtilesRemap.cc:108-120.
static inline double fluxLimiter(int id, double r) {
  if (id == MH_MINMOD) return std::max(0.0, std::min(1.0, r));
  if (id == MH_SUPERBEE) return std::max(0.0, std::max(std::min(2.0 * r, 1.0), std::min(r, 2.0)));
  if (id == MH_VANLEER) {
    if (r <= 0.0) return 0.0;
    return 2.0 * r / (1.0 + r);
  }
  return 0.0;
}

// RemapADIService::fluxLimiterG, UtilesRemap.cc:130-183: limited slope for unequal widths
// h0 (cell), hplus / hmoins (front / back neighbour); ids MH_MINMOD_G..MH_ULTRABEE_G.
static double fluxLimiterG(int id, double gradplus, double gradmoins, double y0, double yplus, double ymoins,
                           double h0, double hplus, double hmoins, bool simple_pente) {
  (void)y0;
  double grady = 0.;
  if (gradplus * gradmoins < 0.0) return 0.;
  if (id == MH_MINMOD_G) {
    if ((yplus - ymoins) > 0.) grady = std::min(std::fabs(gradplus), std::fabs(gradmoins));
    else grady = -std::min(std::fabs(gradplus), std::fabs(gradmoins));
  } else if (id == MH_SUPERBEE_G) {
    if ((yplus - ymoins) > 0.) grady = std::max(std::fabs(gradplus), std::fabs(gradmoins));
    else grady = -std::max(std::fabs(gradplus), std::fabs(gradmoins));
  } else if (id == MH_VANLEER_G) {
    const double lambdaplus = (h0 / 2. + hplus) / (h0 + hplus + hmoins);
    const double lambdamoins = (h0 / 2. + hmoins) / (h0 + hplus + hmoins);
    if ((lambdaplus * gradplus + lambdamoins * gradmoins) != 0.)
      grady = gradplus * gradmoins / (lambdaplus * gradplus + lambdamoins * gradmoins);
    else
      grady = 0.;
  } else if (id == MH_ULTRABEE_G) {
    grady = (yplus - ymoins) / h0;
  } else if (id == MH_ARITHMETIC_G) {
    const double lambdaplus = (h0 / 2. + hplus) / (h0 + hplus + hmoins);
    const double lambdamoins = (h0 / 2. + hmoins) / (h0 + hplus + hmoins);
    grady = lambdamoins * gradplus + lambdaplus * gradmoins;
  }
  if (simple_pente) {  // :166-180
    const double gradMplus = gradplus * (h0 + hplus) / h0;
    const double gradMmoins = gradmoins * (h0 + hmoins) / h0;
    const double gradM = std::min(std::fabs(gradMplus), std::fabs(gradMmoins));
    double gradyl;
    if ((yplus - ymoins) > 0.) gradyl = std::min(std::fabs(gradM), std::fabs(grady));
    else gradyl = -std::min(std::fabs(gradM), std::fabs(grady));
    grady = gradyl;
  }
  return grady;
}

// computeAndLimitGradPhi, UtilesRemap.cc:325-357: classical limiters from the two face
// gradients, generalised ones (id >= MH_MINMOD_G) from the three cell values and widths.
static double limitedGrad(int id, double gb, double gf, double y0, double yplus, double ymoins, double h0,
                          double hplus, double hmoins, bool simple_pente) {
  if (id < MH_MINMOD_G) {
    double g = 0.;
    if (gb != 0.) g += 0.5 * (fluxLimiter(id, gf / gb) * gb);
    if (gf != 0.) g += 0.5 * (fluxLimiter(id, gb / gf) * gf);
    return g;
  }
  return fluxLimiterG(id, gf, gb, y0, yplus, ymoins, h0, hplus, hmoins, simple_pente);
}

// ---------------------------------------------------------------- pente-borne (UtilesRemap.cc)
// `abs` of a Real in computeFluxPP / INTY / computexgxd.
static inline double ref_abs(double x) { return std::fabs(x); }

// INTY, UtilesRemap.cc:86-97: integral from x0 to X of the segment (x0, y0)-(x1, y1).
static inline double inty(double X, double x0, double y0, double x1, double y1, double thr) {
  double flux = 0.;
  if (ref_abs(x1 - x0) > thr) {
    const double eta = amin(amax(0., (X - x0) / (x1 - x0)), 1.);
    flux = (y0 + 0.5 * eta * (y1 - y0)) * (x1 - x0) * eta;
  }
  return flux;
}

// computeY0, UtilesRemap.cc:194-258: monotonicity thresholds y0plus (type 0), y0moins (type 1):
// the middle values at which the reconstruction reaches a neighbour value at the face, with
// superbee's slope, the centred slope between the neighbour centres (arithmetic) or the
// harmonic mean 1/s = l+/g+ + l-/g- (vanLeer, clamped to the neighbour values).
static double computeY0(int id, double y0, double yplus, double ymoins, double h0, double hplus, double hmoins,
                        int type, bool simple_pente) {
  (void)y0;
  double y0plus = 0., y0moins = 0.;
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
  return type == 0 ? y0plus : y0moins;
}

// One pente-borne flux (computeFluxPP / computeFluxPPPure, UtilesRemap.cc:425-832, per
// variable): the three-piece reconstruction of computexgxd (:269-290) and computeygyd
// (:302-318) integrated over the part of the cell that leaves through the back (type 0)
// or front (type 1) face during the step, ppv = positive part of the outgoing velocity x dt.
static double pbFlux(int type, double y0, double yplus, double ymoins, double h0, double hplus, double hmoins,
                     double grady, double ppv, int limY, double thr, bool simple_pente) {
  const double y0plus = computeY0(limY, y0, yplus, ymoins, h0, hplus, hmoins, 0, simple_pente);
  const double y0moins = computeY0(limY, y0, yplus, ymoins, h0, hplus, hmoins, 1, simple_pente);
  double xplus = 1.;
  if (ref_abs(y0plus - yplus) > thr) xplus = (y0 - yplus) / (y0plus - yplus) - 1. / 2.;
  double xmoins = 1.;
  if (ref_abs(y0moins - ymoins) > thr) xmoins = (y0 - ymoins) / (y0moins - ymoins) - 1. / 2.;
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
  double F = std::max(((f1m - f1) + (f2m - f2) + (f3m - f3)), 0.);
  if (((yplus - y0) * (ymoins - y0)) >= 0.) F = y0 * ppv;  // formula 16
  return F;
}

// RemapADIService::computeGradPhiCell :161-267 + computeAndLimitGradPhi UtilesRemap.cc:325-357:
// limited cell slopes (boundary cells are their own missing neighbour) and, with
// pente-borne, the fluxes through the back and front faces (computeFluxPP, or
// computeFluxPPPure for a pure neighbourhood with projection-pente-borne-mixte).
void OracleCase::computeGradPhiCell(int idir) {
  std::fill(grad_phi.begin(), grad_phi.end(), 0.0);
  const bool pb = opt.projection_pente_borne != 0;
  if (pb) {
    std::fill(dphi_ar.begin(), dphi_ar.end(), 0.0);
    std::fill(dphi_av.begin(), dphi_av.end(), 0.0);
  }
  if (opt.ordre_projection <= 1) return;
  const R3 dirproj{0.5 * (1 - idir) * (2 - idir), 1.0 * idir * (2 - idir), -0.5 * idir * (1 - idir)};
  const double deltat = global_deltat;
  const double thr = opt.remap_threshold;
  const bool sp = opt.projection_simple_pente != 0;
#pragma omp parallel for schedule(static)
  for (long c = 0; c < nc; ++c) {
    long backFace = -1, frontFace = -1;
    int lb = -1, lf = -1;
    long bc = c, fc = c;
    for (int l = 0; l < nfpc; ++l) {
      int f = cell_faces[c * nfpc + l];
      if (face_dir[f] != idir) continue;
      if (face_prev[f] == c) {
        frontFace = f;
        lf = l;
        if (face_next[f] >= 0) fc = face_next[f];
      } else if (face_next[f] == c) {
        backFace = f;
        lb = l;
        if (face_prev[f] >= 0) bc = face_prev[f];
      }
    }
    const bool vpure = opt.pente_borne_mixte && est_mixte[c] == 0 && est_mixte[fc] == 0 && est_mixte[bc] == 0 &&
                       est_pure[c] == est_pure[fc] && est_pure[c] == est_pure[bc];
    int lim = opt.projection_limiteur_id;
    if (opt.projection_limiteur_pure_id == 1 && vpure) lim = opt.projection_limiteur_pure_id;
    const double* P = &phi_lagrange[c * nv];
    const double* Pf = &phi_lagrange[fc * nv];
    const double* Pb = &phi_lagrange[bc * nv];
    const double h0 = h_cell_lagrange[c], hp = h_cell_lagrange[fc], hm = h_cell_lagrange[bc];
    for (int v = 0; v < nv; ++v)
      grad_phi[c * nv + v] = limitedGrad(lim, grad_phi_face[backFace * nv + v], grad_phi_face[frontFace * nv + v],
                                         P[v], Pf[v], Pb[v], h0, hp, hm, sp);
    if (!pb) continue;
    const int limY = vpure ? opt.projection_limiteur_pure_id : opt.projection_limiteur_id;
    for (int type = 0; type < 2; ++type) {
      const long face = type == 0 ? backFace : frontFace;
      const double vout = dot(outer_face_normal[c * nfpc + (type == 0 ? lb : lf)], dirproj) * face_normal_velocity[face];
      const double ppv = 0.5 * (vout + ref_abs(vout)) * deltat;
      double* F = type == 0 ? &dphi_ar[c * nv] : &dphi_av[c * nv];
      for (int v = 0; v < nv; ++v)
        F[v] = pbFlux(type, P[v], Pf[v], Pb[v], h0, hp, hm, grad_phi[c * nv + v], ppv, limY, thr, sp);
      // debar fixes (:577-623 mixed neighbourhood, :782-826 pure)
      if (opt.pente_borne_debar_fix == 1) {
        double sfv = 0.;
        for (int k = 0; k < m; ++k) sfv += F[k];
        if (std::fabs(sfv) > thr) {
          for (int k = 0; k < m; ++k) {
            if (!vpure) F[m + k] = (F[m + k] / sfv) * F[k];
            F[2 * m + k] = (F[2 * m + k] / sfv) * F[m + k];
          }
        } else {
          for (int v = 0; v < nv; ++v) F[v] = 0.;
        }
      } else if (opt.pente_borne_debar_fix == 2) {
        for (int k = 0; k < m; ++k) {
          if (!vpure) F[m + k] = P[m + k] * F[k];
          F[2 * m + k] = P[2 * m + k] * F[m + k];
        }
      }
      if (ppv == 0.)
        for (int v = 0; v < nv; ++v) F[v] = 0.;
    }
  }
}

// RemapADIService::computeUpwindFaceQuantitiesForProjection :279-383 (no pente-borne):
// order <= 2 upwind values; order 3 (:319-379) the 6-cell flux of ComputeFluxOrdre3 with
// neighbours clamped at the boundary, then, when the two cells are not pure cells of one
// environment, M_k and E_k face values rebuilt from the upwind cell's ratios.
void OracleCase::computeUpwindFaceQuantitiesForProjection(int idir) {
  std::fill(phi_face.begin(), phi_face.end(), 0.0);
  if (opt.ordre_projection == 3) {
#pragma omp parallel for schedule(static)
    for (long f = face_off[idir]; f < face_off[idir] + face_cnt[idir]; ++f) {
      const int cb = face_prev[f], cf = face_next[f];
      if (cb < 0 || cf < 0) continue;
      int cbb = cb, cbbb = cb, cff = cf, cfff = cf;
      const int pb = dirNeighbor(cb, idir, true);
      if (pb >= 0) {
        cbb = pb;
        cbbb = cbb;
        const int pbb = dirNeighbor(cbb, idir, true);
        if (pbb >= 0) cbbb = pbb;
      }
      const int nf_ = dirNeighbor(cf, idir, false);
      if (nf_ >= 0) {
        cff = nf_;
        cfff = cff;
        const int nff = dirNeighbor(cff, idir, false);
        if (nff >= 0) cfff = nff;
      }
      const double vdt = face_normal_velocity[f] * global_deltat;
      const double* P[6] = {&phi_lagrange[(long)cbbb * nv], &phi_lagrange[(long)cbb * nv], &phi_lagrange[(long)cb * nv],
                            &phi_lagrange[(long)cf * nv], &phi_lagrange[(long)cff * nv], &phi_lagrange[(long)cfff * nv]};
      for (int v = 0; v < nv; ++v)
        phi_face[f * nv + v] = flux_ordre3(P[0][v], P[1][v], P[2][v], P[3][v], P[4][v], P[5][v], h_cell_lagrange[cbbb],
                                           h_cell_lagrange[cbb], h_cell_lagrange[cb], h_cell_lagrange[cf],
                                           h_cell_lagrange[cff], h_cell_lagrange[cfff], vdt);
      const bool voisinage_pure = est_mixte[cb] == 0 && est_mixte[cf] == 0 && est_pure[cb] == est_pure[cf];
      if (!voisinage_pure) {
        double* F = &phi_face[f * nv];
        for (int k = 0; k < m; ++k) {
          if (vdt > 0. && P[2][k] != 0. && P[2][m + k] != 0.) {
            F[m + k] = (P[2][m + k] / P[2][k]) * F[k];
            F[2 * m + k] = (P[2][2 * m + k] / P[2][m + k]) * F[m + k];
          } else if (vdt < 0. && P[3][k] != 0. && P[3][m + k] != 0.) {
            F[m + k] = (P[3][m + k] / P[3][k]) * F[k];
            F[2 * m + k] = (P[3][2 * m + k] / P[3][m + k]) * F[m + k];
          }
        }
      }
    }
    return;
  }
  if (opt.projection_pente_borne) {  // :312-317: flux out of cb minus flux out of cf
#pragma omp parallel for schedule(static)
    for (long f = face_off[idir]; f < face_off[idir] + face_cnt[idir]; ++f) {
      int cb = face_prev[f], cf = face_next[f];
      if (cb < 0 || cf < 0) continue;
      for (int v = 0; v < nv; ++v) phi_face[f * nv + v] = dphi_av[(long)cb * nv + v] - dphi_ar[(long)cf * nv + v];
    }
    return;
  }
  const int order2 = opt.ordre_projection - 1;
#pragma omp parallel for schedule(static)
  for (long f = face_off[idir]; f < face_off[idir] + face_cnt[idir]; ++f) {
    int cb = face_prev[f], cf = face_next[f];
    if (cb < 0 || cf < 0) continue;
    int up = (face_normal_velocity[f] * deltax_lagrange[f] > 0.0) ? cb : cf;
    const double h = order2 * dot(face_coord[f] - cell_coord[up], face_normal[f]);
    for (int v = 0; v < nv; ++v)
      phi_face[f * nv + v] = phi_lagrange[(long)up * nv + v] + h * grad_phi[(long)up * nv + v];
  }
}

// RemapADIService::computeUremap :395-575 (pente-borne off).
void OracleCase::computeUremap(int idir) {
  const R3 dirproj{0.5 * (1 - idir) * (2 - idir), 1.0 * idir * (2 - idir), -0.5 * idir * (1 - idir)};
  const double deltat = global_deltat;
  const bool pb = opt.projection_pente_borne != 0;
  const double thr = opt.remap_threshold;
  std::fill(dual_phi_flux.begin(), dual_phi_flux.end(), 0.0);
#pragma omp parallel for schedule(static)
  for (long c = 0; c < nc; ++c) {
    double flux_face[4 * MH_MAX_ENV + 1];
    for (int v = 0; v < nv; ++v) flux_face[v] = 0.;
    for (int l = 0; l < nfpc; ++l) {
      int f = cell_faces[c * nfpc + l];
      if (std::fabs(dot(face_normal[f], dirproj)) >= 1.0E-10) {
        const double fnv = face_normal_velocity[f];
        const double len = comp(face_length_lagrange[f], idir);
        const R3 outer = outer_face_normal[c * nfpc + l];
        const double od = dot(outer, dirproj);
        for (int v = 0; v < nv; ++v) {
          // pente-borne: phi_face already holds the integrated flux
          double flux = pb ? od * len * phi_face[(long)f * nv + v] : od * fnv * len * deltat * phi_face[(long)f * nv + v];
          if (q_frozen && v == 3 * m) flux = 0.;
          flux_face[v] += flux;
          dual_phi_flux[c * nv + v] += 0.5 * flux * comp(outer, idir);
        }
      }
    }
    double* U = &u_lagrange[c * nv];
    for (int v = 0; v < nv; ++v) U[v] -= flux_face[v];
    if (pb && !opt.sans_lagrange)  // :441-463 (with the dual phase): tiny partial volumes, masses or energies drop the environment's three
      for (int k = 0; k < m; ++k) {
        if (U[k] < 1. * thr) U[k] = U[m + k] = U[2 * m + k] = 0.;
        if (U[m + k] < 1. * thr) U[k] = U[m + k] = U[2 * m + k] = 0.;
        if (U[2 * m + k] < 1. * thr) U[k] = U[m + k] = U[2 * m + k] = 0.;
      }
    for (int k = 0; k < m; ++k) {
      if (U[m + k] < 0.) U[m + k] = 0.;
      if (U[2 * m + k] < 0.) U[2 * m + k] = 0.;
    }
    double somme_volume = 0.;
    for (int k = 0; k < m; ++k) somme_volume += U[k];
    double* P = &phi_lagrange[c * nv];
    if (pb) {  // :493-538: volume fraction, density, specific energy; the pseudo keeps its value
      for (int k = 0; k < m; ++k) {
        P[k] = U[k] / somme_volume;
        P[m + k] = U[k] != 0. ? U[m + k] / U[k] : 0.;
      }
      for (int k = 0; k < m; ++k) P[2 * m + k] = U[m + k] != 0. ? U[2 * m + k] / U[m + k] : 0.;
    } else {
      for (int v = 0; v < nv; ++v)
        if (!(q_frozen && v == 3 * m)) P[v] = U[v] / somme_volume;
    }
    setMixedFlags(c, P);  // :559-573
  }
}

// est_mixte / est_pure from the volume fractions (RemapADIService.cc:559-573,
// RemapADIFinal.cc:124-145): mixed if more than one is positive, else the index of the
// positive one (-1 if none).
void OracleCase::setMixedFlags(long c, const double* f) {
  int cnt = 0, pure = -1;
  for (int k = 0; k < m; ++k)
    if (f[k] > 0.) {
      cnt++;
      pure = k;
    }
  est_mixte[c] = cnt > 1 ? 1 : 0;
  est_pure[c] = cnt > 1 ? -1 : pure;
}

// RemapADIService::computeDualUremap, DualPhaseRemap.cc:38-258, with
// computeDualGradPhi UtilesRemap.cc:19-72 and computeAndLimitGradPhiDual :369-400.
void OracleCase::computeDualUremap(int idir) {
  const double deltat = global_deltat;
  const int lim = opt.projection_limiteur_id;
  const std::vector<int>& prv = node_prev[idir];
  const std::vector<int>& nxt = node_next[idir];
  std::fill(dual_grad_phi.begin(), dual_grad_phi.end(), 0.0);
  if (opt.ordre_projection > 1) {
#pragma omp parallel for schedule(static)
    for (long n = 0; n < nn; ++n) {
      int back = prv[n], front = nxt[n];
      if (back < 0 || front < 0) continue;
      const double xn = comp(node_coord[n], idir);
      const double xf = comp(node_coord[front], idir), xb = comp(node_coord[back], idir);
      // dual cell widths (UtilesRemap.cc:50-61); with neither second neighbour (two
      // cells along the sweep) both outer widths are 0
      const int bb = prv[back], ff = nxt[front];
      const double h0 = 0.5 * (xf - xb);
      double hm = 0., hp = 0.;
      if (bb < 0) {
        if (ff >= 0) hp = 0.5 * (comp(node_coord[ff], idir) - xn);
      } else if (ff < 0) {
        hm = 0.5 * (xn - comp(node_coord[bb], idir));
      } else {
        hm = 0.5 * (xn - comp(node_coord[bb], idir));
        hp = 0.5 * (comp(node_coord[ff], idir) - xn);
      }
      for (int v = 0; v < 3; ++v) {
        double gf = (phi_dual[front * 5 + v] - phi_dual[n * 5 + v]) / (xf - xn);
        double gb = (phi_dual[n * 5 + v] - phi_dual[back * 5 + v]) / (xn - xb);
        double g;
        if (lim < MH_MINMOD_G) {  // computeAndLimitGradPhiDual :369-400
          g = 0.;
          if (gb != 0.) g += 0.5 * fluxLimiter(lim, gf / gb) * gb;
          if (gf != 0.) g += 0.5 * fluxLimiter(lim, gb / gf) * gf;
        } else {
          g = fluxLimiterG(lim, gf, gb, phi_dual[n * 5 + v], phi_dual[front * 5 + v], phi_dual[back * 5 + v], h0, hp,
                           hm, opt.projection_simple_pente != 0);
        }
        dual_grad_phi[n * 5 + v] = g;
      }
    }
  }
  // back / front mass fluxes: faces of direction idir in face order, env order
  const double OneOverNbcell = dim == 2 ? .5 : .25;
#pragma omp parallel for schedule(static)
  for (long n = 0; n < nn; ++n) {
    double bsum = 0., fsum = 0.;
    for (int a = nf_off[idir][n]; a < nf_off[idir][n + 1]; ++a) {
      int f = nf_list[idir][a];
      int cb = face_prev[f], cf = face_next[f];
      if (cb >= 0)
        for (int k = 0; k < m; ++k) bsum += OneOverNbcell * dual_phi_flux[(long)cb * nv + m + k];
      if (cf >= 0)
        for (int k = 0; k < m; ++k) fsum += OneOverNbcell * dual_phi_flux[(long)cf * nv + m + k];
    }
    back_flux_mass[n] = bsum;
    front_flux_mass[n] = fsum;
  }
  const int order2 = opt.ordre_projection - 1;
  std::vector<double> newU(u_dual);
#pragma omp parallel for schedule(static)
  for (long n = 0; n < nn; ++n) {
    int back = prv[n], front = nxt[n];
    if (back < 0 || front < 0) continue;
    double* Ud = &newU[n * 5];
    const double bf = back_flux_mass[n], ff = front_flux_mass[n];
    Ud[3] += bf - ff;
    const double xn = comp(node_coord[n], idir);
    const double xf = comp(node_coord[front], idir), xb = comp(node_coord[back], idir);
    double FU[3], BU[3], FE, BE;
    const double ufront = 0.5 * (phi_dual[n * 5 + idir] + phi_dual[front * 5 + idir]);
    const int signfront = ufront > 0 ? 1 : -1;
    const int fsrc = (ff < 0.) ? front : (int)n;
    for (int v = 0; v < 3; ++v)
      FU[v] = phi_dual[fsrc * 5 + v] + order2 * (0.5 * dual_grad_phi[fsrc * 5 + v] * (signfront * (xf - xn) - deltat * ufront));
    FE = phi_dual[fsrc * 5 + 4] + order2 * (0.5 * dual_grad_phi[fsrc * 5 + 4] * (signfront * (xf - xn) - deltat * ufront));
    const double uback = 0.5 * (phi_dual[back * 5 + idir] + phi_dual[n * 5 + idir]);
    const int signback = uback > 0 ? 1 : -1;
    const int bsrc = (bf > 0.) ? back : (int)n;
    for (int v = 0; v < 3; ++v)
      BU[v] = phi_dual[bsrc * 5 + v] + order2 * (0.5 * dual_grad_phi[bsrc * 5 + v] * (signback * (xn - xb) - deltat * uback));
    BE = phi_dual[bsrc * 5 + 4] + order2 * (0.5 * dual_grad_phi[bsrc * 5 + 4] * (signback * (xn - xb) - deltat * uback));
    for (int v = 0; v < 3; ++v) Ud[v] += bf * BU[v] - ff * FU[v];
    Ud[4] += bf * BE - ff * FE;
  }
  u_dual.swap(newU);
#pragma omp parallel for schedule(static)
  for (long n = 0; n < nn; ++n) {
    if (prv[n] < 0 || nxt[n] < 0) continue;
    double* Ud = &u_dual[n * 5];
    double* Pd = &phi_dual[n * 5];
    const double arithmetic_thresold = 1.e-300;
    for (int v = 0; v < 3; ++v)
      if (std::fabs(Ud[v]) < arithmetic_thresold) Ud[v] = 0.;
    Pd[0] = Ud[0] / Ud[3];
    Pd[1] = Ud[1] / Ud[3];
    Pd[2] = Ud[2] / Ud[3];
    Pd[3] = Ud[3];
    Pd[4] = Ud[4] / Ud[3];
  }
}

// Arcane environment membership change (env_cells.addItems/removeItems + forceRecompute).
// Values kept across the change: a cell that becomes multi-environment copies its global
// value into the partial slot of the environment it already had; new partial slots start
// at 0; a cell that becomes single-environment copies that environment's partial value
// into the global value.  remapVariables overwrites every value the cycle reads next.
void OracleCase::setMembership(const std::vector<uint8_t>& newmask) {
  for (long c = 0; c < nc; ++c) {
    uint8_t o = mask[c], nw = newmask[c];
    if (o == nw) continue;
    int on = __builtin_popcount(o), nnb = __builtin_popcount(nw);
    for (MatVar* v : matvars) {
      double* part = &v->p[c * MH_MAX_ENV];
      if (on == 1 && nnb != 1) {
        int k0 = __builtin_ctz(o);
        for (int k = 0; k < MH_MAX_ENV; ++k) part[k] = 0.;
        part[k0] = v->g[c];
      } else if (on != 1 && nnb == 1) {
        int k1 = __builtin_ctz(nw);
        v->g[c] = part[k1];
      } else {
        for (int k = 0; k < MH_MAX_ENV; ++k)
          if (((nw >> k) & 1) && !((o >> k) & 1)) part[k] = 0.;
      }
    }
    mask[c] = nw;
    nbenv[c] = (int8_t)nnb;
  }
}

// RemapADIService::remapVariables, RemapADIFinal.cc:21-346 (with dual projection).
int OracleCase::remapVariables() {
  const double thr = opt.remap_threshold;
  std::vector<uint8_t> newmask(mask);
  for (int k = 0; k < m; ++k) {  // :29-78
    for (long c = 0; c < nc; ++c) {
      const double* U = &u_lagrange[c * nv];
      const bool was = member(k, c);
      if (((U[k] / euler_volume[c]) < thr || (U[m + k] == 0. && !opt.sans_lagrange)) && was) {
        newmask[c] &= (uint8_t)~(1u << k);
      } else if ((U[k] / euler_volume[c]) > thr && !was) {
        double volt = 0.;
        for (int kk = 0; kk < m; ++kk) volt += U[kk];
        double vol_ev = U[k] * euler_volume[c] / volt;
        if ((vol_ev / euler_volume[c]) > thr && U[m + k] != 0.) newmask[c] |= (uint8_t)(1u << k);
      }
    }
  }
  setMembership(newmask);
  int err = 0;
#pragma omp parallel for schedule(static) reduction(| : err)
  for (long c = 0; c < nc; ++c) {
    const double* U = &u_lagrange[c * nv];
    double vol_nplus1[MH_MAX_ENV], density_env[MH_MAX_ENV], energy_env[MH_MAX_ENV];
    const double vol = euler_volume[c];
    cell_volume.g[c] = euler_volume[c];
    double volt = 0., masset = 0.;
    for (int k = 0; k < m; ++k) {
      vol_nplus1[k] = U[k];
      volt += vol_nplus1[k];
      masset += U[m + k];
    }
    const double unsurvolt = 1. / volt;
    for (int k = 0; k < m; ++k) vol_nplus1[k] *= vol * unsurvolt;
    double somme_frac = 0.;
    const double unsurvol = 1. / vol;
    for (int k = 0; k < m; ++k) {
      if (!member(k, c)) continue;
      double& f = ev(fracvol, k, c);
      f = vol_nplus1[k] * unsurvol;
      if (f < thr) f = 0.;
      somme_frac += f;
    }
    double unsursomme_frac = 0.;
    if (somme_frac != 0.) unsursomme_frac = 1. / somme_frac;
    for (int k = 0; k < m; ++k)
      if (member(k, c)) ev(fracvol, k, c) *= unsursomme_frac;
    {  // RemapADIFinal.cc:124-145
      double fr[MH_MAX_ENV];
      for (int k = 0; k < m; ++k) fr[k] = member(k, c) ? ev(fracvol, k, c) : 0.;
      setMixedFlags(c, fr);
    }
    if (masset > thr * thr) {
      const double unsurmasset = 1. / masset;
      double fmasset = 0.;
      for (int k = 0; k < m; ++k) {
        if (!member(k, c)) continue;
        double& mf = ev(mass_fraction, k, c);
        mf = U[m + k] * unsurmasset;
        if (ev(fracvol, k, c) < thr) mf = 0.;
        fmasset += mf;
      }
      if (fmasset != 0.) {
        const double unsurfmasset = 1. / fmasset;
        for (int k = 0; k < m; ++k)
          if (member(k, c)) ev(mass_fraction, k, c) *= unsurfmasset;
      }
    }
    double density_nplus1 = 0.;
    for (int k = 0; k < m; ++k) {
      if (!member(k, c)) continue;
      density_env[k] = 0.;
      if (ev(fracvol, k, c) > thr && vol_nplus1[k] != 0.) density_env[k] = U[m + k] / vol_nplus1[k];
      density_nplus1 += ev(fracvol, k, c) * density_env[k];
    }
    for (int k = 0; k < m; ++k) {
      if (!member(k, c)) continue;
      energy_env[k] = 0.;
      if (ev(fracvol, k, c) > thr && U[m + k] != 0.) energy_env[k] = U[2 * m + k] / U[m + k];
    }
    density.g[c] = density_nplus1;
    cell_mass.g[c] = euler_volume[c] * density_nplus1;
    double energie_nplus1 = 0., pseudo_nplus1 = 0.;
    for (int k = 0; k < m; ++k) {
      if (!member(k, c)) continue;
      ev(cell_mass, k, c) = ev(mass_fraction, k, c) * cell_mass.g[c];
      ev(pseudo, k, c) = U[3 * m] * unsurvol;
      ev(density, k, c) = density_env[k];
      ev(internal_energy, k, c) = energy_env[k];
      energie_nplus1 += ev(mass_fraction, k, c) * ev(internal_energy, k, c);
      pseudo_nplus1 += ev(fracvol, k, c) * ev(pseudo, k, c);
    }
    internal_energy.g[c] = energie_nplus1;
    pseudo.g[c] = pseudo_nplus1;
    for (int k = 0; k < m; ++k) {
      if (!member(k, c)) continue;
      if (ev(density, k, c) < 0. || ev(internal_energy, k, c) < 0.) {
        ev(internal_energy, k, c) = 0.;
        ev(density, k, c) = 0.;
        ev(fracvol, k, c) = 0.;
        ev(mass_fraction, k, c) = 0.;
      }
      if (ev(density, k, c) != ev(density, k, c) || ev(internal_energy, k, c) != ev(internal_energy, k, c)) err = 1;
    }
  }
  if (err) return MH_ERR_NAN_REMAP;
  // :286-297 node velocity from the dual projection (none in advection mode)
  if (!opt.sans_lagrange) {
#pragma omp parallel for schedule(static)
    for (long n = 0; n < nn; ++n) {
      const double mp = u_dual[n * 5 + 3];
      if (mp != 0.) {
        velocity[n].x = u_dual[n * 5 + 0] / mp;
        velocity[n].y = u_dual[n * 5 + 1] / mp;
        velocity[n].z = u_dual[n * 5 + 2] / mp;
      }
    }
  }
  computeNodeMass();  // :299-306
  for (long c = 0; c < nc; ++c) {  // :308-316
    if (nbenv[c] == 1) continue;
    for (int k = 0; k < m; ++k)
      if (member(k, c)) ev(cell_volume, k, c) = ev(fracvol, k, c) * cell_volume.g[c];
  }
  if (opt.conservation_energie_totale) {  // :318-345
    // per cell, a quarter of (projected minus reconstructed) kinetic energy per unit mass
    // over its nodes in cell order (0.25 in 3D as well), kept if positive, given to each
    // environment by its mass fraction; a cell with other than one environment also adds
    // it to its global value (a single-environment cell's value is the global one)
#pragma omp parallel for schedule(static)
    for (long c = 0; c < nc; ++c) {
      double delta_ec = 0.;
      for (int l = 0; l < nnpc; ++l) {
        const long n = cell_nodes[c * nnpc + l];
        if (u_dual[n * 5 + 3] != 0.) {
          const double ec_proj = u_dual[n * 5 + 4] / u_dual[n * 5 + 3];
          const double ec_reconst = 0.5 * squareNormL2(velocity[n]);
          delta_ec += 0.25 * (ec_proj - ec_reconst);
        }
      }
      delta_ec = std::max(0., delta_ec);
      for (int k = 0; k < m; ++k)
        if (member(k, c)) ev(internal_energy, k, c) += ev(mass_fraction, k, c) * delta_ec;
      if (nbenv[c] != 1) internal_energy.g[c] += delta_ec;
    }
  }
  return MH_OK;
}

// ============================================================================ time step
// MahycoModule::computeDeltaT, MahycoModule.cc:1739-1864 (isEuler "Full": vmax included).
int OracleCase::computeDeltaT() {
  global_old_deltat = global_deltat;
  old_deltat = global_old_deltat;
  if (opt.sans_lagrange) {  // :1751-1753: the step stays deltat-init
    global_deltat = opt.deltat_init;
    const double stop_time = opt.final_time;
    if ((global_time < stop_time && (global_time + global_deltat) > stop_time) || global_time > stop_time) stopped = true;
    return MH_OK;
  }
  double minimum_aux = 1.79769313486231570815e+308;
  std::vector<double> dxs(nc);
#pragma omp parallel for schedule(static)
  for (long c = 0; c < nc; ++c) {
    double vmax = 0.;
    for (int l = 0; l < nnpc; ++l) vmax = amax(normL2(velocity[cell_nodes[c * nnpc + l]]), vmax);
    dxs[c] = caracteristic_length[c] / (sound_speed.g[c] + vmax);
  }
  for (long c = 0; c < nc; ++c) minimum_aux = amin(minimum_aux, dxs[c]);
  double new_dt = opt.cfl * minimum_aux;
  new_dt = amin(new_dt, 1.05 * global_old_deltat);
  new_dt = amin(new_dt, opt.deltat_max);
  if (new_dt < opt.deltat_min) return MH_ERR_DT_MIN;
  global_deltat = new_dt;
  const double stop_time = opt.final_time;
  bool not_yet_finish = global_time < stop_time;
  bool finish = global_time > stop_time;
  bool too_much = (global_time + new_dt) > stop_time;
  if (not_yet_finish && too_much) stopped = true;
  if (finish) stopped = true;
  return MH_OK;
}

// EnvSummationService::write, time_history/EnvSummationService.cc:77-169 (period 10).
void OracleCase::timeHistoryWrite() {
  if (global_iteration % 10 != 0) return;
  history.push_back(global_time);
  for (int k = 0; k < m; ++k) {
    long cnt = 0;
    double sm = 0., sp = 0., sr = 0., sv = 0., se = 0., smm = 0.;
    for (long c = 0; c < nc; ++c) {
      if (!member(k, c)) continue;
      cnt++;
      sm += ev(cell_mass, k, c);
      sp += ev(pressure, k, c);
      sr += ev(density, k, c) * ev(cell_volume, k, c);
      sv += ev(cell_volume, k, c);
      se += ev(internal_energy, k, c) * ev(cell_mass, k, c);
      smm += ev(cell_mass, k, c);
    }
    history.push_back(cnt ? sm / cnt : 0.);
    history.push_back(cnt ? sp / cnt : 0.);
    history.push_back(cnt ? sr / sv : 0.);
    history.push_back(cnt ? se / smm : 0.);
    history.push_back((double)cnt);
  }
}

// One iteration of MahycoLoop / MahycoLagrangeLoop (Mahyco.config:23-36, 57-69).
int OracleCase::cycle() {
  global_iteration += 1;
  global_time += global_deltat;
  saveValuesAtN();
  if (opt.sans_lagrange) updateVelocityWithoutLagrange();
  else updateVelocity();
  applyBoundaryCondition();
  updatePosition();
  int st = computeGeometricValues();
  if (st != MH_OK) return st;
  if (!opt.sans_lagrange) {  // skipped in advection mode (:1271, :389, :1433)
    updateDensity();
    computeArtificialViscosity();
    updateEnergyAndPressure();
  }
  if (opt.with_projection) {
    st = remap();
    if (st != MH_OK) return st;
  }
  st = computeDeltaT();
  if (st != MH_OK) return st;
  timeHistoryWrite();
  return MH_OK;
}

// ============================================================================ C ABI
extern "C" {

OracleCase* oracle_create(const MhMesh* mesh, const MhEnvironment* envs, int nb_env, const MhOptions* opt) {
  if (!mesh || !envs || !opt || nb_env < 1 || nb_env > MH_MAX_ENV) return nullptr;
  if (mesh->dim != 2 && mesh->dim != 3) return nullptr;
  if (mesh->nx < 1 || mesh->ny < 1 || (mesh->dim == 3 && mesh->nz < 1)) return nullptr;
  if (opt->ordre_projection < 1 || opt->ordre_projection > 3) return nullptr;
  if (opt->projection_limiteur_id < 0 || opt->projection_limiteur_id > MH_ULTRABEE_G) return nullptr;
  if (opt->projection_pente_borne && (opt->pente_borne_debar_fix < 0 || opt->pente_borne_debar_fix > 2)) return nullptr;
  if (opt->projection_pente_borne && opt->ordre_projection == 3) return nullptr;
  // CSTS with Newton reads p before setting it (MahycoModule.cc:1595-1601); with explicit
  // pressure the EOS runs before the forward velocity step; advection mode has no Lagrange step
  if (opt->schema_csts && (opt->with_newton || opt->pression_explicite || opt->sans_lagrange)) return nullptr;
  if (opt->longueur_caracteristique < 0 || opt->longueur_caracteristique > MH_L_MONODIM_Z) return nullptr;
  if (opt->longueur_caracteristique == MH_L_FACES_OPPOSEES && mesh->dim != 3) return nullptr;  // 8 nodes, 6 faces
  // RIDER / ADVECTION give the bubble environment 1's density, Mono tests included
  if ((opt->cas_model == MH_CASE_RIDER || opt->cas_model == MH_CASE_ADVECTION) && nb_env < 2) return nullptr;
  OracleCase* h = new OracleCase();
  h->m = nb_env;
  for (int k = 0; k < nb_env; ++k) h->env[k] = envs[k];
  h->opt = *opt;
  h->build_mesh(*mesh);
  const long nc = h->nc, nn = h->nn, nf = h->nf;
  h->nv = 3 * h->m + 1;
  h->velocity.assign(nn, {0., 0., 0.});
  h->force.assign(nn, {0., 0., 0.});
  h->node_mass.assign(nn, 0.);
  h->cell_coord.assign(nc, {0., 0., 0.});
  h->face_coord.assign(nf, {0., 0., 0.});
  h->face_normal.assign(nf, {0., 0., 0.});
  h->cqs.assign(nc * h->nnpc, {0., 0., 0.});
  h->outer_face_normal.assign(nc * h->nfpc, {0., 0., 0.});
  h->caracteristic_length.assign(nc, 0.);
  h->div_u.assign(nc, 0.);
  h->euler_volume.assign(nc, 0.);
  h->materiau.assign(nc, 0.);
  h->mask.assign(nc, 1);
  h->nbenv.assign(nc, 1);
  h->matvars = {&h->fracvol, &h->mass_fraction, &h->density, &h->density_n, &h->internal_energy,
                &h->internal_energy_n, &h->pressure, &h->pressure_n, &h->sound_speed, &h->pseudo,
                &h->pseudo_n, &h->pseudo_nm1, &h->cell_volume, &h->cell_volume_n, &h->cell_mass,
                &h->tau_density};
  for (MatVar* v : h->matvars) v->init(nc);
  const int nv = h->nv;
  h->u_lagrange.assign(nc * nv, 0.);
  h->phi_lagrange.assign(nc * nv, 0.);
  h->grad_phi.assign(nc * nv, 0.);
  h->grad_phi_face.assign(nf * nv, 0.);
  h->phi_face.assign(nf * nv, 0.);
  h->dual_phi_flux.assign(nc * nv, 0.);
  h->u_dual.assign(nn * 5, 0.);
  h->phi_dual.assign(nn * 5, 0.);
  h->dual_grad_phi.assign(nn * 5, 0.);
  h->deltax_lagrange.assign(nf, 0.);
  h->h_cell_lagrange.assign(h->nc, 0.);
  if (opt->projection_pente_borne) {
    h->dphi_ar.assign(h->nc * h->nv, 0.);
    h->dphi_av.assign(h->nc * h->nv, 0.);
  }
  h->est_mixte.assign(h->nc, 0);
  h->est_pure.assign(h->nc, 0);
  h->face_normal_velocity.assign(nf, 0.);
  h->face_length_lagrange.assign(nf, {0., 0., 0.});
  h->back_flux_mass.assign(nn, 0.);
  h->front_flux_mass.assign(nn, 0.);
  h->hydroStartInit();
  h->computeCellMass();
  h->computeNodeMass();
  return h;
}

void oracle_destroy(OracleCase* h) { delete h; }

// hydroContinueInit (MahycoModule.cc:273-306): the state continues as dumped; the remap
// then carries 3m+5 slots, which leaves the pseudo-viscosity slot out when m >= 2.
void oracle_continue(OracleCase* h) { h->q_frozen = 7 * h->m >= 3 * h->m + 5; }

int oracle_run(OracleCase* h, int max_cycles) {
  for (int i = 0; i < max_cycles; ++i) {
    if (h->stopped) return MH_STOPPED;
    int st = h->cycle();
    if (st != MH_OK) return st;
  }
  return h->stopped ? MH_STOPPED : MH_OK;
}

double oracle_time(const OracleCase* h) { return h->global_time; }
double oracle_deltat(const OracleCase* h) { return h->global_deltat; }
int oracle_iteration(const OracleCase* h) { return h->global_iteration; }
long oracle_nb_cells(const OracleCase* h) { return h->nc; }
long oracle_nb_nodes(const OracleCase* h) { return h->nn; }

int oracle_get_field(const OracleCase* hc, int field, int env, int comp_, double* out) {
  OracleCase* h = const_cast<OracleCase*>(hc);
  if (field >= 100) {
    for (long n = 0; n < h->nn; ++n) {
      if (field == MH_N_VELOCITY) out[n] = comp(h->velocity[n], comp_);
      else if (field == MH_N_NODE_MASS) out[n] = h->node_mass[n];
      else if (field == MH_N_COORD) out[n] = comp(h->node_coord[n], comp_);
      else if (field == MH_N_COORD_LAGRANGE)
        out[n] = comp(h->opt.with_projection && !h->node_coord_lag.empty() ? h->node_coord_lag[n] : h->node_coord[n], comp_);
      // the last remap's dual state (oracle only, for tests): comp 0-2 momentum, 3 mass, 4 kinetic energy
      else if (field == ORACLE_N_DUAL_U && comp_ >= 0 && comp_ < 5)
        out[n] = h->u_dual.empty() ? 0. : h->u_dual[n * 5 + comp_];
      else return MH_ERR_INVALID;
    }
    return MH_OK;
  }
  if (env >= h->m) return MH_ERR_INVALID;
  MatVar* v = nullptr;
  switch (field) {
    case MH_F_DENSITY: v = &h->density; break;
    case MH_F_PRESSURE: v = &h->pressure; break;
    case MH_F_INTERNAL_ENERGY: v = &h->internal_energy; break;
    case MH_F_SOUND_SPEED: v = &h->sound_speed; break;
    case MH_F_PSEUDO_VISCOSITY: v = &h->pseudo; break;
    case MH_F_CELL_MASS: v = &h->cell_mass; break;
    case MH_F_CELL_VOLUME: v = &h->cell_volume; break;
    case MH_F_FRACVOL: v = &h->fracvol; break;
    case MH_F_MASS_FRACTION: v = &h->mass_fraction; break;
    default: break;
  }
  for (long c = 0; c < h->nc; ++c) {
    if (field == MH_F_CARACTERISTIC_LENGTH) out[c] = h->caracteristic_length[c];
    else if (field == MH_F_ENV_MASK) out[c] = h->mask[c];
    else if (field == MH_F_MATERIAU) out[c] = h->materiau[c];
    else if (!v) return MH_ERR_INVALID;
    else if (env < 0) out[c] = v->g[c];
    else out[c] = h->member(env, c) ? h->ev(*v, env, c) : 0.;
  }
  return MH_OK;
}

// Writes one field of the state, with mahyco_set_field's meaning: env = -1 the cell value,
// env = k environment k's value (partial value of a mixed cell, cell value of a cell of k
// alone, nothing outside k).  MH_N_COORD_LAGRANGE re-forms the corner vectors the next
// velocity update reads (MahycoModule.cc:1097-1162).
int oracle_set_field(OracleCase* h, int field, int env, int comp_, const double* in) {
  if (field >= 100) {
    if (comp_ < 0 || comp_ >= h->dim) return MH_ERR_INVALID;
    if (field == MH_N_VELOCITY) {
      for (long n = 0; n < h->nn; ++n) compr(h->velocity[n], comp_) = in[n];
      return MH_OK;
    }
    if (field != MH_N_COORD_LAGRANGE && !(field == MH_N_COORD && !h->opt.with_projection)) return MH_ERR_INVALID;
    std::vector<R3>& x = h->opt.with_projection ? h->node_coord_lag : h->node_coord;
    if (x.empty()) x = h->node_coord;
    for (long n = 0; n < h->nn; ++n) compr(x[n], comp_) = in[n];
    for (long c = 0; c < h->nc; ++c) {  // the corner vectors of computeGeometricValues at x
      R3 coord[8];
      for (int l = 0; l < h->nnpc; ++l) coord[l] = x[h->cell_nodes[c * h->nnpc + l]];
      R3* q = &h->cqs[c * h->nnpc];
      if (h->dim == 3) {
        R3 fc[6];
        fc[0] = 0.25 * (coord[0] + coord[3] + coord[2] + coord[1]);
        fc[1] = 0.25 * (coord[0] + coord[4] + coord[7] + coord[3]);
        fc[2] = 0.25 * (coord[0] + coord[1] + coord[5] + coord[4]);
        fc[3] = 0.25 * (coord[4] + coord[5] + coord[6] + coord[7]);
        fc[4] = 0.25 * (coord[1] + coord[2] + coord[6] + coord[5]);
        fc[5] = 0.25 * (coord[2] + coord[3] + coord[7] + coord[6]);
        h->computeCQs(coord, fc, q);
      } else {
        R3 npc[5];
        coord[4] = coord[0];
        for (int l = 0; l < 4; ++l) {
          npc[l + 1].x = 0.5 * (coord[l + 1].y - coord[l].y);
          npc[l + 1].y = 0.5 * (coord[l].x - coord[l + 1].x);
          npc[l + 1].z = 0.;
        }
        npc[0] = npc[4];
        for (int l = 0; l < 4; ++l) q[l] = npc[l + 1] + npc[l];
      }
    }
    return MH_OK;
  }
  if (env < -1 || env >= h->m) return MH_ERR_INVALID;
  if (field == MH_F_ENV_MASK) {
    if (env != -1) return MH_ERR_INVALID;
    for (long c = 0; c < h->nc; ++c) {
      h->mask[c] = (uint8_t)(int)in[c];
      h->nbenv[c] = (int8_t)__builtin_popcount(h->mask[c]);
    }
    return MH_OK;
  }
  MatVar* v = nullptr;
  switch (field) {
    case MH_F_DENSITY: v = &h->density; break;
    case MH_F_PRESSURE: v = &h->pressure; break;
    case MH_F_INTERNAL_ENERGY: v = &h->internal_energy; break;
    case MH_F_SOUND_SPEED: v = &h->sound_speed; break;
    case MH_F_PSEUDO_VISCOSITY: v = &h->pseudo; break;
    case MH_F_CELL_MASS: v = &h->cell_mass; break;
    case MH_F_FRACVOL: v = &h->fracvol; break;
    case MH_F_MASS_FRACTION: v = &h->mass_fraction; break;
    default: return MH_ERR_INVALID;
  }
  for (long c = 0; c < h->nc; ++c) {
    if (env < 0) v->g[c] = in[c];
    else if (h->member(env, c)) h->ev(*v, env, c) = in[c];
  }
  // the node mass the next velocity update divides by is sum M / 2^d (computeNodeMass,
  // MahycoModule.cc:247-262, and remapVariables, RemapADIFinal.cc:299-306)
  if (field == MH_F_CELL_MASS) h->computeNodeMass();
  return MH_OK;
}

// m_global_time, m_global_deltat, m_global_old_deltat, m_global_iteration, m_sens_projection.
void oracle_get_scalars(const OracleCase* h, double* t, double* dt, double* dt_old, int* it, int* sens) {
  *t = h->global_time;
  *dt = h->global_deltat;
  *dt_old = h->old_deltat;
  *it = h->global_iteration;
  *sens = h->sens_projection;
}
void oracle_set_scalars(OracleCase* h, double t, double dt, double dt_old, int it, int sens) {
  h->global_time = t;
  h->global_deltat = dt;
  h->global_old_deltat = dt_old;
  h->old_deltat = dt_old;
  h->global_iteration = it;
  h->sens_projection = sens;
  h->stopped = false;
}

// IEquationOfState::applyEOS(env) and MahycoModule::computePressionMoyenne as standalone calls
// (between iterations: m_global_time is the last completed iteration's time).
void oracle_apply_eos(OracleCase* h, int k) {
  if (h->env[k].eos == MH_EOS_FICTIF) h->applyFictif(k);
  else h->applyEOS(k);
}
void oracle_compute_pression_moyenne(OracleCase* h) {
  if (!h->opt.sans_lagrange) h->computePressionMoyenne();  // :1711
}

int oracle_history(const OracleCase* h, double* out, int max_rows) {
  const int w = 1 + 5 * h->m;
  int rows = (int)(h->history.size() / w);
  int n = rows < max_rows ? rows : max_rows;
  if (out) std::memcpy(out, h->history.data(), sizeof(double) * (size_t)n * w);
  return rows;
}

void oracle_set_pow_mode(int mode) { g_pow_mode = mode; }

double oracle_flux_limiter(int id, double r) { return fluxLimiter(id, r); }
double oracle_flux_limiter_g(int id, double gradplus, double gradmoins, double y0, double yplus, double ymoins,
                             double h0, double hplus, double hmoins, int simple_pente) {
  return fluxLimiterG(id, gradplus, gradmoins, y0, yplus, ymoins, h0, hplus, hmoins, simple_pente != 0);
}
double oracle_compute_y0(int id, double y0, double yplus, double ymoins, double h0, double hplus, double hmoins,
                         int type, int simple_pente) {
  return computeY0(id, y0, yplus, ymoins, h0, hplus, hmoins, type, simple_pente != 0);
}

void oracle_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

}  // extern "C"
