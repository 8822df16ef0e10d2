// mahyco.hip -- host side of the MI355X hydro path: case initialisation, device memory,
// kernel sequencing of one time-loop iteration, hipGraph replay and the C ABI of
// include/mahyco.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/mahyco.h"
#include "mh_kernels.h"
#include "mh_output.h"

namespace {

// profiled kernel groups (sweeps separately per direction: they run different code)
const char* kKernelNames[] = {"k_node_lag",     "k_cell_lag",     "k_sweep_cell_x",  "k_sweep_cell_y",
                              "k_sweep_cell_z", "k_sweep_node_x", "k_sweep_node_y",  "k_sweep_node_z",
                              "k_node_velocity", "k_finalize",    "k_dt",            "k_hcell",
                              "k_node_fwd"};
enum {
  KN_NODE = 0,
  KN_CELL,
  KN_SWEEP_CELL,  // + direction
  KN_SWEEP_NODE = KN_SWEEP_CELL + 3,
  KN_NODE_VEL = KN_SWEEP_NODE + 3,
  KN_FINAL,
  KN_DT,
  KN_HCELL,  // cell widths of the order-3 remap
  KN_NODE_FWD,  // CSTS forward velocity step before the remap
  KN_COUNT
};

// sweep variants: one thread per cell / node; row segments along x and LDS tiles along y / z
enum { MH_SWEEP_FLAT = 0, MH_SWEEP_TILE = 1 };

struct ProfRec {
  int kind;
  hipEvent_t a, b;
};

}  // namespace

struct MhEnsemble;
struct MhHandle {
  MhMesh mesh;
  MhOptions opt;
  MhEnvironment env[MH_MAX_ENV];
  int m = 1;
  MhGrid g;
  MhDev dev;
  hipStream_t stream = nullptr;
  std::vector<void*> allocs;
  std::vector<size_t> alloc_bytes;  // size of each allocation (checkpoints)
  double* xl[2][3] = {{nullptr, nullptr, nullptr}, {nullptr, nullptr, nullptr}};
  double *Ua = nullptr, *Ub = nullptr, *dflux = nullptr, *Ud = nullptr, *Pa = nullptr, *Pb = nullptr;
  double* u0[3] = {nullptr, nullptr, nullptr};  // advection mode: prescribed velocity
  double* rev_cos = nullptr;                     // advection mode: reverse factor per iteration
  long launched = 0;  // iterations queued by the host
  int sens = 0;       // m_sens_projection of the next iteration
  int nblk_c = 0, nblk_n = 0;
  int phase = 0;  // entry-point order check
  bool prof = false;
  std::vector<ProfRec> prof_pending;
  double prof_ms[KN_COUNT] = {0};
  int prof_cnt[KN_COUNT] = {0};
  std::map<long, hipGraphExec_t> graphs;  // key: (cycles << 8) | (parity << 4) | sens
  std::vector<hipGraph_t> graph_defs;
  long graph_captures = 0;  // graphs captured and instantiated so far
  bool owns_stream = true;
  bool node_tiled = true;   // 3D: z-marching node kernel (MAHYCO_NODE_KERNEL=flat selects the flat one)
  int sweep_mode = MH_SWEEP_TILE;  // MAHYCO_SWEEP=flat selects one thread per cell / node
  std::vector<const mh_output::VarSpec*> ens_vars;  // EnSight variables of this case's outputs
  std::vector<double> ens_times;                    // times of the EnSight steps written
  MhScalars* ext_sc = nullptr;  // scalars living in an ensemble's array
  size_t bytes = 0;             // device bytes of the case
};

namespace {

int hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) {
    fprintf(stderr, "mahyco: %s failed: %s\n", what, hipGetErrorString(e));
    return MH_ERR_DEVICE;
  }
  return MH_OK;
}
#define HIPCHK(call)                                     \
  do {                                                   \
    int _st = hip_check((call), #call);                  \
    if (_st != MH_OK) return _st;                        \
  } while (0)

template <class T>
int dalloc(MhHandle* h, T** p, size_t n) {
  void* q = nullptr;
  if (n == 0) n = 1;
  hipError_t e = hipMalloc(&q, n * sizeof(T));
  if (e != hipSuccess) {
    fprintf(stderr, "mahyco: hipMalloc(%zu bytes) failed: %s\n", n * sizeof(T), hipGetErrorString(e));
    return MH_ERR_NOMEM;
  }
  h->allocs.push_back(q);
  h->alloc_bytes.push_back(n * sizeof(T));
  h->bytes += n * sizeof(T);
  *p = (T*)q;
  return MH_OK;
}

template <class T>
int upload(MhHandle* h, T* d, const std::vector<T>& v) {
  HIPCHK(hipMemcpyAsync(d, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, h->stream));
  return MH_OK;
}

// Node coordinate along one axis of the Cartesian generator.
double axis_coord(const MhMesh& mesh, int d, int i, int n) {
  const double o = mesh.origin[d], l = mesh.length[d];
  if (mesh.coord_rule == MH_COORD_DIV) return o + (l * (double)i) / (double)n;
  return o + (double)i * (l / (double)n);
}

// Host-side case initialisation: restates the init entry points HydroStartInit
// (MahycoModule.cc:21-210, MahycoAnnexe.cc:11-142), the case services' initMat/initVar
// (SODService.cc, SEDOVService.cc, POINTTRIPLEService.cc), IEquationOfState::initEOS and
// ComputeCellMass (:221-237).  The node mass is formed by the first node kernel.
struct HostInit {
  std::vector<double> ax, ay, az, VE, rho, p, c, q, M, e, fg, fk, chik, rhok, ek, pk, Mk;
  std::vector<double> chig;  // global mass fraction: 1 everywhere, as Arcane starts it
  std::vector<uint8_t> mask;
  std::vector<double> u[3];
  std::vector<double> rev_cos;  // advection mode: cos(pi t_n parametre) per iteration
};

// RIDERService (casTest/RIDER/RIDERService.cc) constants: the bubble centre of the
// translations (0.2, 0.2) or of the rotations and later tests (0.5, 0.75), and the
// reference's Pi (TypesMahyco.h:42).
const double kRiderPi = 3.14159265359;
void rider_radii(const D3* x, int nnode, bool low_centre, double& rmin, double& rmax) {
  const double xb = low_centre ? 0.20 : 0.50, yb = low_centre ? 0.20 : 0.75;
  rmin = 10.;
  rmax = 0.;
  for (int l = 0; l < nnode; ++l) {
    const double rnode = std::sqrt((x[l].x - xb) * (x[l].x - xb) + (x[l].y - yb) * (x[l].y - yb));
    rmin = std::min(rmin, rnode);
    rmax = std::max(rmax, rnode);
  }
}

int host_init(MhHandle* h, HostInit& I) {
  const MhGrid& g = h->g;
  const MhOptions& opt = h->opt;
  const int m = h->m, dim = g.dim;
  const long nc = g.nc, nn = g.nn;
  I.ax.resize(g.nnx);
  I.ay.resize(g.nny);
  I.az.resize(dim == 3 ? g.nnz : 1, 0.);
  for (int i = 0; i < g.nnx; ++i) I.ax[i] = axis_coord(h->mesh, 0, i, g.nx);
  for (int j = 0; j < g.nny; ++j) I.ay[j] = axis_coord(h->mesh, 1, j, g.ny);
  if (dim == 3)
    for (int k = 0; k < g.nnz; ++k) I.az[k] = axis_coord(h->mesh, 2, k, g.nz);
  auto node = [&](int i, int j, int k) { return d3(I.ax[i], I.ay[j], dim == 3 ? I.az[k] : 0.); };
  const size_t ne = (size_t)m * nc;
  I.VE.assign(nc, 0.);
  I.rho.assign(nc, 0.);
  I.p.assign(nc, 0.);
  I.c.assign(nc, 0.);
  I.q.assign(nc, 0.);
  I.M.assign(nc, 0.);
  I.e.assign(nc, 0.);
  I.fg.assign(nc, 1.);
  I.fk.assign(ne, 1.);
  I.chik.assign(ne, 1.);
  I.chig.assign(nc, 1.);
  I.rhok.assign(ne, 0.);
  I.ek.assign(ne, 0.);
  I.pk.assign(ne, 0.);
  I.Mk.assign(ne, 0.);
  I.mask.assign(nc, 1);
  for (int d = 0; d < 3; ++d) I.u[d].assign(nn, 0.);
  std::vector<double> xcx(nc), xcy(nc), xcz(nc), f1(nc, 0.);
  std::vector<double> rider_frac(nc, -1.);  // RIDER cells cut by the bubble: its fraction
  const double w = dim == 3 ? .125 : .25;
  for (int k = 0; k < g.nz; ++k)
    for (int j = 0; j < g.ny; ++j)
      for (int i = 0; i < g.nx; ++i) {
        const long c = i + (long)g.nx * (j + (long)g.ny * k);
        D3 x[8], qv[8];
        x[0] = node(i, j, k);
        x[1] = node(i + 1, j, k);
        x[2] = node(i + 1, j + 1, k);
        x[3] = node(i, j + 1, k);
        if (dim == 3) {
          x[4] = node(i, j, k + 1);
          x[5] = node(i + 1, j, k + 1);
          x[6] = node(i + 1, j + 1, k + 1);
          x[7] = node(i, j + 1, k + 1);
        }
        const int nnode = dim == 3 ? 8 : 4;
        D3 s = d3(0., 0., 0.);
        for (int l = 0; l < nnode; ++l) s = s + x[l];
        const D3 xc = w * s;
        xcx[c] = xc.x;
        xcy[c] = xc.y;
        xcz[c] = xc.z;
        double V;
        if (dim == 3) {
          cqs_hex(x, qv);
          V = cell_volume<8>(x, qv, 3);
        } else {
          cqs_quad(x, qv);
          V = cell_volume<4>(x, qv, 2);
        }
        I.VE[c] = V;
        // Materiau (initMat) and membership (MahycoAnnexe.cc:97-139)
        double mat = 0.;
        if (opt.cas_model == MH_CASE_SOD) {
          const int t = opt.cas_test;
          if (t == 13 || t == 14 || t == 15) {
            const double rr = t == 13 ? xc.x : t == 14 ? xc.y : xc.z;
            mat = (rr < 0.5) ? 0. : 1.;
          }
        } else if (opt.cas_model == MH_CASE_POINTTRIPLE) {
          if (opt.cas_test != 2) {
            if (xc.x <= 0.01) mat = 0.;
            else mat = (xc.y <= 0.015) ? 1. : 2.;
          }
        } else if (opt.cas_model == MH_CASE_RIDER) {  // RIDERService::initMat :19-53
          if (opt.cas_test >= 31) {
            double rmin, rmax;
            rider_radii(x, nnode, opt.cas_test < 34, rmin, rmax);
            const double rb = opt.rider_rayon_bulle;
            if (rmax < rb) mat = 1.;
            else if ((rmax >= rb) && (rmin < rb)) mat = (rb - rmin) / (rmax - rmin);
          }
          // initVar (initVarMono :58-158, initVar :163-290), per cell
          const bool mono = opt.cas_test <= 30;
          const bool lin = opt.rider_densite_lineaire != 0;
          const double rb = opt.rider_rayon_bulle;
          double rmin, rmax;
          rider_radii(x, nnode, opt.cas_test < (mono ? 26 : 34), rmin, rmax);
          const double xb = opt.cas_test < (mono ? 26 : 34) ? 0.20 : 0.50;
          const double yb = opt.cas_test < (mono ? 26 : 34) ? 0.20 : 0.75;
          const double croissance = lin ? 1. / rb : 1.;
          I.rho[c] = h->env[0].densite_initiale;
          I.p[c] = h->env[0].pression_initiale;
          I.e[c] = h->env[0].energie_initiale;
          const double r = std::sqrt((xc.x - xb) * (xc.x - xb) + (xc.y - yb) * (xc.y - yb));
          const double r_cell = lin ? r : rmax;
          if (r_cell < rb) {
            I.rho[c] = h->env[1].densite_initiale;
            if (lin) I.rho[c] = h->env[1].densite_initiale * std::min(croissance * 2 * (rb - r), 1.);
            I.p[c] = h->env[1].pression_initiale;
          } else if ((rmax >= rb) && (rmin < rb) && !lin) {
            rider_frac[c] = (rb - rmin) / (rmax - rmin);
          }
        } else if (opt.cas_model == MH_CASE_ADVECTION) {  // ADVECTIONService::initMat :11-65
          if (opt.cas_test >= 18) {  // the Mono* tests (7-10) put every cell in environment 0
            const double yb = opt.cas_test < 21 ? 0.50 : 0.75;
            const double rb = 0.25;  // a constant here, the rayon-bulle option in initVar
            if (std::fabs(xc.x - 0.50) < rb && std::fabs(xc.y - yb) < rb) mat = 1.;
          }
        } else if (opt.cas_model == MH_CASE_RIEMANN) {
          const int a = opt.interface_axis;
          const D3 lo3 = x[0], hi3 = a == 0 ? x[1] : a == 1 ? x[3] : x[4];
          const double lo = a == 0 ? lo3.x : a == 1 ? lo3.y : lo3.z;
          const double hi = a == 0 ? hi3.x : a == 1 ? hi3.y : hi3.z;
          double fr = (hi - opt.interface_position) / (hi - lo);
          fr = fr < 0. ? 0. : (fr > 1. ? 1. : fr);
          f1[c] = fr;
          mat = fr;
        }
        uint8_t bits;
        if (mat == 0.) bits = 1;
        else if (mat > 0. && mat < 1.) bits = 3;
        else if (mat == 1.) bits = 2;
        else if (mat > 1. && mat < 2.) bits = 6;
        else if (mat == 2.) bits = 4;
        else if (mat > 2. && mat < 3.) bits = 12;
        else if (mat == 3.) bits = 8;
        else bits = 15;
        bits &= (uint8_t)((1u << m) - 1u);
        if (bits == 0) return MH_ERR_INVALID;
        I.mask[c] = bits;
        // Sedov deposit uses the cell volume of the initial geometry
        if (opt.cas_model == MH_CASE_SEDOV) {
          const double rhoInit = h->env[0].densite_initiale, pInit = h->env[0].pression_initiale;
          bool center = false;
          for (int l = 0; l < nnode; ++l) {
            double rn;
            if (dim == 3)
              rn = std::sqrt((x[l].x - 0.) * (x[l].x - 0.) + (x[l].y - 0.) * (x[l].y - 0.) + (x[l].z - 0.) * (x[l].z - 0.));
            else
              rn = std::sqrt((x[l].x - 0.) * (x[l].x - 0.) + (x[l].y - 0.) * (x[l].y - 0.));
            if (rn < 1.e-10) center = true;
          }
          I.e[c] = 0.244816e-5;
          I.rho[c] = rhoInit;
          I.p[c] = pInit;
          if (center) {
            I.e[c] += opt.sedov_deposit / V;
            I.p[c] = 0.4 * rhoInit * I.e[c];
          }
        }
      }
  // initVar of the other case models (global values; partial values of multi-env cells)
  for (long c = 0; c < nc; ++c) {
    const uint8_t bits = I.mask[c];
    const bool pure = __builtin_popcount(bits) == 1;
    if (opt.cas_model == MH_CASE_SOD) {
      const int t = opt.cas_test;
      if (t == 3 || t == 4 || t == 5) {  // SODService::initVarMono :11-40
        const double rr = t == 3 ? xcx[c] : t == 4 ? xcy[c] : xcz[c];
        if (rr < 0.5) {
          I.p[c] = h->env[0].pression_initiale;
          I.rho[c] = h->env[0].densite_initiale;
        } else {
          I.p[c] = h->env[0].pression_initiale / 10.;
          I.rho[c] = h->env[0].densite_initiale / 8.;
        }
      } else {  // SODService::initVar :71-121
        const int s = bits == 1 ? 0 : 1;
        I.rho[c] = h->env[s].densite_initiale;
        I.p[c] = h->env[s].pression_initiale;
        if (!pure)
          for (int k = 0; k < m; ++k)
            if ((bits >> k) & 1) {
              I.rhok[k * nc + c] = h->env[s].densite_initiale;
              I.pk[k * nc + c] = h->env[s].pression_initiale;
            }
      }
    } else if (opt.cas_model == MH_CASE_POINTTRIPLE) {  // POINTTRIPLEService::initVar :64-128
      const int s = xcx[c] <= 0.01 ? 0 : (xcy[c] <= 0.015 ? 1 : 2);
      I.rho[c] = h->env[s].densite_initiale;
      I.p[c] = h->env[s].pression_initiale;
      if (!pure && ((bits >> s) & 1)) {
        I.rhok[s * nc + c] = h->env[s].densite_initiale;
        I.pk[s * nc + c] = h->env[s].pression_initiale;
      }
    } else if (opt.cas_model == MH_CASE_RIDER) {
      const double frac_b = rider_frac[c];
      if (frac_b >= 0.) {  // environment 0 takes the bubble density and fraction (RIDERService.cc:252-266)
        if (opt.cas_test >= 31)
          for (int k = 0; k < m && k < 2; ++k) {
            if (!((bits >> k) & 1)) continue;
            const double rk = h->env[k == 0 ? 1 : 0].densite_initiale;
            const double fr = k == 0 ? frac_b : 1. - frac_b;
            if (pure) {  // the environment value is the cell value
              I.rho[c] = rk;
              I.fg[c] = fr;
              I.chig[c] = fr;
              I.p[c] = 0.;
            } else {
              I.rhok[k * nc + c] = rk;
              I.fk[k * nc + c] = fr;
              I.chik[k * nc + c] = fr;
              I.pk[k * nc + c] = 0.;
            }
          }
        I.rho[c] = frac_b * h->env[1].densite_initiale;
        I.p[c] = h->env[1].pression_initiale;
      }
    } else if (opt.cas_model == MH_CASE_ADVECTION) {  // ADVECTIONService::initVarMono :67-145, initVar :146-246
      // a square of half side rayon-bulle, no mixed cells, internal energy left at 0
      const bool lin = opt.rider_densite_lineaire != 0;
      const double rb = opt.rider_rayon_bulle;
      const double xb = 0.50, yb = opt.cas_test < (opt.cas_test <= 10 ? 10 : 21) ? 0.50 : 0.75;
      const double croissance = lin ? 1. / rb : 1.;
      I.rho[c] = h->env[0].densite_initiale;
      I.p[c] = h->env[0].pression_initiale;
      const double r = std::sqrt((xcx[c] - xb) * (xcx[c] - xb) + (xcy[c] - yb) * (xcy[c] - yb));
      if (std::fabs(xcx[c] - xb) < rb && std::fabs(xcy[c] - yb) < rb) {
        I.rho[c] = h->env[1].densite_initiale;
        if (lin) I.rho[c] = h->env[1].densite_initiale * croissance * r;
        I.p[c] = h->env[1].pression_initiale;
      }
    } else if (opt.cas_model == MH_CASE_RIEMANN) {
      if (pure) {
        const int k = bits == 1 ? 0 : 1;
        I.rho[c] = h->env[k].densite_initiale;
        I.p[c] = h->env[k].pression_initiale;
      } else {
        I.fk[0 * nc + c] = 1.0 - f1[c];
        I.fk[1 * nc + c] = f1[c];
        double ft = 0.;
        for (int k = 0; k < 2; ++k) ft += I.fk[k * nc + c] * h->env[k].densite_initiale;
        for (int k = 0; k < 2; ++k) {
          I.rhok[k * nc + c] = h->env[k].densite_initiale;
          I.pk[k * nc + c] = h->env[k].pression_initiale;
          I.chik[k * nc + c] = I.fk[k * nc + c] * h->env[k].densite_initiale / ft;
        }
        I.rho[c] = 0.;
        I.p[c] = 0.;
        I.c[c] = 0.;
      }
    }
  }
  const double* v0 = h->env[0].vitesse_initiale;
  if (opt.cas_model == MH_CASE_RIDER || opt.cas_model == MH_CASE_ADVECTION) {
    // RIDERService.cc:134-157, 267-289 and ADVECTIONService.cc:131-144, 233-246, numbered as
    // RIDER's fields: 23 Tx, 24 Ty, 25 T45, 26 rotation, 27/29 vortex, 28/30 deformation
    const int t = opt.cas_model == MH_CASE_ADVECTION ? (opt.cas_test <= 10 ? opt.cas_test : opt.cas_test - 11) + 16
                  : opt.cas_test <= 30                ? opt.cas_test
                                                      : opt.cas_test - 8;
    const double Pi = kRiderPi;
    for (long n = 0; n < nn; ++n) {
      const int i = (int)(n % g.nnx);
      const long r = n / g.nnx;
      const int j = (int)(r % g.nny), k = (int)(r / g.nny);
      const D3 x = node(i, j, k);
      double ux = 0., uy = 0.;
      if (t == 23) ux = 1.;
      if (t == 24) uy = 1.;
      if (t == 25) ux = uy = 1.;
      // one sincos per argument: GCC merges the sin/cos pairs of RIDERService.cc into sincos
      // calls at -O1 and above, and glibc's sincos can differ from its sin by an ulp
      double s0, c0, s1, c1;
      if (t == 26) {
        const double dx = x.x - 0.5, dy = x.y - 0.5;
        const double theta = std::atan2(dy, dx);
        const double rr = std::sqrt(dx * dx + dy * dy);
        const double omega = 4. * Pi;
        ::sincos(omega * 0. + theta, &s0, &c0);
        ux = -rr * omega * s0;
        uy = rr * omega * c0;
      }
      if (t == 27 || t == 29) {
        ::sincos(Pi * x.x, &s0, &c0);
        ::sincos(Pi * x.y, &s1, &c1);
        ux = -2. * c1 * s1 * s0 * s0;
        uy = 2. * c0 * s0 * s1 * s1;
      }
      if (t == 28 || t == 30) {
        const double dx = x.x + 0.5, dy = x.y + 0.5;
        ::sincos(4. * Pi * dx, &s0, &c0);
        ::sincos(4. * Pi * dy, &s1, &c1);
        ux = s0 * s1;
        uy = c0 * c1;
      }
      I.u[0][n] = ux;
      I.u[1][n] = uy;
      I.u[2][n] = 0.;
    }
  } else if (opt.cas_model != MH_CASE_SEDOV)
    for (int d = 0; d < 3; ++d) std::fill(I.u[d].begin(), I.u[d].end(), dim == 2 && d == 2 ? 0. : v0[d]);
  if (opt.sans_lagrange) {  // cos(pi t parametre) at t_n = t_(n-1) + deltat-init, as the device accumulates it
    const long nmax = (long)std::ceil(opt.final_time / opt.deltat_init) + 8;
    I.rev_cos.resize(nmax);
    double t = 0.;
    for (long it = 0; it < nmax; ++it) {
      t = t + opt.deltat_init;
      I.rev_cos[it] = std::cos(kRiderPi * t * opt.rider_parametre);
    }
  }
  // initEOS per environment (PerfectGasEOSService.cc:12-37, StiffenedGasEOSService.cc:14-33),
  // not in advection mode (MahycoModule.cc:143)
  for (long c = 0; c < nc && !opt.sans_lagrange; ++c) {
    const uint8_t bits = I.mask[c];
    const bool pure = __builtin_popcount(bits) == 1;
    for (int k = 0; k < m; ++k) {
      if (!((bits >> k) & 1)) continue;
      const double gamma = h->env[k].adiabatic_cst;
      double& pp = pure ? I.p[c] : I.pk[k * nc + c];
      double& rr = pure ? I.rho[c] : I.rhok[k * nc + c];
      double& ee = pure ? I.e[c] : I.ek[k * nc + c];
      double cs;
      if (h->env[k].eos == MH_EOS_STIFFENED_GAS) {
        const double pi = h->env[k].limit_tension;
        ee = (pp + (gamma * pi)) / ((gamma - 1.) * rr);
        cs = std::sqrt((gamma / rr) * (pp + pi));
      } else {
        ee = pp / ((gamma - 1.) * rr);
        cs = std::sqrt(gamma * pp / rr);
      }
      if (pure) I.c[c] = cs;
      else {
        // mixed-cell averages, MahycoModule.cc:154-170 (density and pressure accumulate
        // onto the values left by initVar)
        if (k == __builtin_ctz(bits)) I.e[c] = 0.;
        I.e[c] += ee * I.chik[k * nc + c];
        I.c[c] = (cs < I.c[c]) ? I.c[c] : cs;
        I.rho[c] += rr * I.fk[k * nc + c];
        I.p[c] += pp * I.fk[k * nc + c];
      }
    }
  }
  // ComputeCellMass (:221-237); m_euler_volume = V (:174-178)
  for (long c = 0; c < nc; ++c) {
    I.M[c] = I.rho[c] * I.VE[c];
    const uint8_t bits = I.mask[c];
    if (__builtin_popcount(bits) != 1)
      for (int k = 0; k < m; ++k)
        if ((bits >> k) & 1) I.Mk[k * nc + c] = I.chik[k * nc + c] * I.M[c];
  }
  return MH_OK;
}

void prof_begin(MhHandle* h, ProfRec& r, int kind) {
  r.kind = kind;
  (void)hipEventCreate(&r.a);
  (void)hipEventCreate(&r.b);
  (void)hipEventRecord(r.a, h->stream);
}
void prof_end(MhHandle* h, ProfRec& r) {
  (void)hipEventRecord(r.b, h->stream);
  h->prof_pending.push_back(r);
}

// Runs the statement with LIM, a compile-time limiter id: 0 minmod, 1 superbee, 2 vanLeer,
// 3 none (first order, or an id without a classical limiter: the slopes are 0).
#define MH_LIMSW(h, ...)                                                  \
  do {                                                                    \
    const int _lim = (h)->dev.order2 ? (h)->dev.limiter : 3;              \
    if (_lim == 0) { constexpr int LIM = 0; __VA_ARGS__; }                \
    else if (_lim == 1) { constexpr int LIM = 1; __VA_ARGS__; }           \
    else if (_lim == 2) { constexpr int LIM = 2; __VA_ARGS__; }           \
    else { constexpr int LIM = 3; __VA_ARGS__; }                          \
  } while (0)

#define LAUNCH(h, kind, kern, grid, block, ...)                          \
  do {                                                                   \
    ProfRec _r;                                                          \
    if ((h)->prof) prof_begin((h), _r, (kind));                          \
    hipLaunchKernelGGL(kern, dim3(grid), dim3(block), 0, (h)->stream, __VA_ARGS__); \
    if ((h)->prof) prof_end((h), _r);                                    \
  } while (0)

// a dual (node) sweep kernel: none in advection mode (withDualProjection, PrepareRemap.cc:211)
#define LAUNCH_DUAL(h, ...)                           \
  do {                                                \
    if (!(h)->dev.sans_lagrange) LAUNCH(h, __VA_ARGS__); \
  } while (0)

// blocks of the row-segment x sweep (one wavefront per MH_ROW_OUT cells of a row), a multiple of 8
int row_blocks(MhHandle* h) {
  const long waves = (long)h->g.ny * h->g.nz * ((h->g.nx + MH_ROW_OUT - 1) / MH_ROW_OUT);
  const long wpb = MH_BLOCK / 64;
  return (int)(((waves + wpb - 1) / wpb + 7) / 8 * 8);
}

// blocks of the row-segment dual x sweep (one wavefront per MH_ROW_OUT nodes of a node row), a multiple of 8
int node_row_blocks(MhHandle* h) {
  const long waves = (long)h->g.nny * h->g.nnz * ((h->g.nnx + MH_ROW_OUT - 1) / MH_ROW_OUT);
  const long wpb = MH_BLOCK / 64;
  return (int)(((waves + wpb - 1) / wpb + 7) / 8 * 8);
}

// blocks of the LDS-tiled y / z sweep (64 cells along x by MH_TILE_L along D), a multiple of 8
// k_sweep_node_z: one block per (row segment of an xy node plane, z plane)
int node_z_blocks(MhHandle* h) {
  const long segs = ((long)h->g.nnx * h->g.nny + MH_BLOCK - 1) / MH_BLOCK;
  return (int)((segs * h->g.nnz + 7) / 8 * 8);
}
int tile_blocks(MhHandle* h, int D) {
  const long na = D == 1 ? h->g.ny : h->g.nz;
  const long nt = D == 1 ? (h->g.dim == 3 ? h->g.nz : 1) : h->g.ny;
  const long tiles = (long)((h->g.nx + 63) / 64) * nt * ((na + MH_TILE_L - 1) / MH_TILE_L);
  return (int)((tiles + 7) / 8 * 8);
}

// cell and dual parts of a y or z sweep in the selected variant (the dual part is one
// thread per node in both)
template <int DIM, int D, int M>
void launch_sweep_yz(MhHandle* h, int sel, int first, int last) {
  const MhDev& S = h->dev;
  if (h->sweep_mode == MH_SWEEP_FLAT)
    MH_LIMSW(h, LAUNCH(h, KN_SWEEP_CELL + D, (k_sweep_cell<DIM, D, M, LIM>), h->nblk_c, MH_BLOCK, S, sel, first));
  else
    MH_LIMSW(h, LAUNCH(h, KN_SWEEP_CELL + D, (k_sweep_cell_tile<DIM, D, M, LIM>), tile_blocks(h, D), MH_BLOCK, S, sel, first));
  if constexpr (DIM == 3 && D == 2) {
    if (h->sweep_mode != MH_SWEEP_FLAT) {  // z dual sweep: blocks walk z fastest
      MH_LIMSW(h, LAUNCH_DUAL(h, KN_SWEEP_NODE + 2, (k_sweep_node_z<LIM>), node_z_blocks(h), MH_BLOCK, S, sel, first, last));
      return;
    }
  }
  MH_LIMSW(h, LAUNCH_DUAL(h, KN_SWEEP_NODE + D, (k_sweep_node<DIM, D, LIM>), h->nblk_n, MH_BLOCK, S, sel, first, last));
}

template <int DIM, int M>
void launch_sweeps_t(MhHandle* h) {
  const MhDev& S = h->dev;
  const int dim = DIM;
  int dirs[3], nd = 0;
  for (int i = 0; i < dim; ++i) {
    const int idir = (i + h->sens) % dim;
    const int ncell = idir == 0 ? h->g.nx : idir == 1 ? h->g.ny : h->g.nz;
    if (ncell == 1) continue;  // RemapADIService.cc:27
    dirs[nd++] = idir;
  }
  const bool o3 = S.order2 == 2;
  // pente-borne or generalised limiters: one thread per cell and node, limiter read at run time
  const bool gen = !o3 && (S.pb || (S.order2 && S.limiter >= MH_MINMOD_G));
  if (S.csts) LAUNCH(h, KN_NODE_FWD, (k_node_fwd<DIM>), h->nblk_n, MH_BLOCK, S);  // PrepareRemap.cc:210
  if (S.hc && nd > 0) LAUNCH(h, KN_HCELL, (k_hcell<DIM>), h->nblk_c, MH_BLOCK, S);
  for (int s = 0; s < nd; ++s) {
    const int first = s == 0, last = (s == nd - 1) && !first, sel = s & 1;
    const int D = dirs[s];
    const int mode = h->sweep_mode;
    if (gen) {
      if (D == 0) {
        LAUNCH(h, KN_SWEEP_CELL, (k_sweep_cell_gen<DIM, 0, M>), h->nblk_c, MH_BLOCK, S, sel, first);
        LAUNCH_DUAL(h, KN_SWEEP_NODE, (k_sweep_node<DIM, 0, MH_LIM_ANY>), h->nblk_n, MH_BLOCK, S, sel, first, last);
      } else if (D == 1) {
        LAUNCH(h, KN_SWEEP_CELL + 1, (k_sweep_cell_gen<DIM, 1, M>), h->nblk_c, MH_BLOCK, S, sel, first);
        LAUNCH_DUAL(h, KN_SWEEP_NODE + 1, (k_sweep_node<DIM, 1, MH_LIM_ANY>), h->nblk_n, MH_BLOCK, S, sel, first, last);
      } else if constexpr (DIM == 3) {
        LAUNCH(h, KN_SWEEP_CELL + 2, (k_sweep_cell_gen<3, 2, M>), h->nblk_c, MH_BLOCK, S, sel, first);
        LAUNCH_DUAL(h, KN_SWEEP_NODE + 2, (k_sweep_node<3, 2, MH_LIM_ANY>), h->nblk_n, MH_BLOCK, S, sel, first, last);
      }
    } else if (o3) {  // order 3: one thread per cell for the cell part, node part as selected
      if (D == 0) LAUNCH(h, KN_SWEEP_CELL, (k_sweep_cell3<DIM, 0, M>), h->nblk_c, MH_BLOCK, S, sel, first);
      else if (D == 1) LAUNCH(h, KN_SWEEP_CELL + 1, (k_sweep_cell3<DIM, 1, M>), h->nblk_c, MH_BLOCK, S, sel, first);
      else if constexpr (DIM == 3) LAUNCH(h, KN_SWEEP_CELL + 2, (k_sweep_cell3<3, 2, M>), h->nblk_c, MH_BLOCK, S, sel, first);
      if (D == 0 && mode != MH_SWEEP_FLAT)
        MH_LIMSW(h, LAUNCH_DUAL(h, KN_SWEEP_NODE, (k_sweep_node_row<DIM, LIM>), node_row_blocks(h), MH_BLOCK, S, sel, first, last));
      else if (D == 0)
        MH_LIMSW(h, LAUNCH_DUAL(h, KN_SWEEP_NODE, (k_sweep_node<DIM, 0, LIM>), h->nblk_n, MH_BLOCK, S, sel, first, last));
      else if (D == 1)
        MH_LIMSW(h, LAUNCH_DUAL(h, KN_SWEEP_NODE + 1, (k_sweep_node<DIM, 1, LIM>), h->nblk_n, MH_BLOCK, S, sel, first, last));
      else if constexpr (DIM == 3)
        MH_LIMSW(h, LAUNCH_DUAL(h, KN_SWEEP_NODE + 2, (k_sweep_node<3, 2, LIM>), h->nblk_n, MH_BLOCK, S, sel, first, last));
    } else if (D == 0) {
      if (mode == MH_SWEEP_FLAT) {
        MH_LIMSW(h, LAUNCH(h, KN_SWEEP_CELL, (k_sweep_cell<DIM, 0, M, LIM>), h->nblk_c, MH_BLOCK, S, sel, first));
        MH_LIMSW(h, LAUNCH_DUAL(h, KN_SWEEP_NODE, (k_sweep_node<DIM, 0, LIM>), h->nblk_n, MH_BLOCK, S, sel, first, last));
      } else {
        MH_LIMSW(h, LAUNCH(h, KN_SWEEP_CELL, (k_sweep_cell_row<DIM, M, LIM>), row_blocks(h), MH_BLOCK, S, sel, first));
        MH_LIMSW(h, LAUNCH_DUAL(h, KN_SWEEP_NODE, (k_sweep_node_row<DIM, LIM>), node_row_blocks(h), MH_BLOCK, S, sel, first, last));
      }
    } else if (D == 1) {
      launch_sweep_yz<DIM, 1, M>(h, sel, first, last);
    } else {
      if constexpr (DIM == 3) launch_sweep_yz<3, 2, M>(h, sel, first, last);
    }
  }
  if (nd <= 1 && !S.sans_lagrange) LAUNCH(h, KN_NODE_VEL, (k_node_velocity<DIM>), h->nblk_n, MH_BLOCK, S, nd == 0 ? 1 : 0);
  LAUNCH(h, KN_FINAL, (k_finalize<DIM, M>), h->nblk_c, MH_BLOCK, S, nd & 1);
  h->sens = (h->sens + 1) % dim;
}

void set_parity(MhHandle* h) {
  const int par = (int)(h->launched & 1);
  for (int d = 0; d < 3; ++d) {
    h->dev.xlp[d] = h->xl[par][d];
    h->dev.xlc[d] = h->xl[par ^ 1][d];
  }
}

template <int DIM, int M>
void launch_node_t(MhHandle* h) {
  if (h->dev.sans_lagrange) {
    LAUNCH(h, KN_NODE, (k_node_adv<DIM>), h->nblk_n, MH_BLOCK, h->dev);
  } else if (DIM == 3 && h->node_tiled) {
    const dim3 grid((h->g.nnx + MH_NTX - 1) / MH_NTX, (h->g.nny + MH_NTY - 1) / MH_NTY,
                    (h->g.nnz + MH_NT_KC - 1) / MH_NT_KC);
    LAUNCH(h, KN_NODE, k_node_lag3_tiled, grid, MH_BLOCK, h->dev);
  } else {
    LAUNCH(h, KN_NODE, (k_node_lag<DIM>), h->nblk_n, MH_BLOCK, h->dev);
  }
}
template <int DIM, int M>
void launch_cell_t(MhHandle* h) {
  if (h->dev.lc != MH_L_RACINE_CUBIQUE_VOLUME) LAUNCH(h, KN_CELL, (k_cell_length<DIM>), h->nblk_c, MH_BLOCK, h->dev);
  if (h->dev.csts) LAUNCH(h, KN_CELL, (k_cell_lag_csts<DIM, M>), h->nblk_c, MH_BLOCK, h->dev);
  else LAUNCH(h, KN_CELL, (k_cell_lag<DIM, M>), h->nblk_c, MH_BLOCK, h->dev);
}
#define MH_DISPATCH(fn, h)                                      \
  do {                                                          \
    const int _d = (h)->g.dim, _m = (h)->m;                     \
    if (_d == 3) {                                              \
      if (_m == 1) fn<3, 1>(h);                                 \
      else if (_m == 2) fn<3, 2>(h);                            \
      else if (_m == 3) fn<3, 3>(h);                            \
      else fn<3, 4>(h);                                         \
    } else {                                                    \
      if (_m == 1) fn<2, 1>(h);                                 \
      else if (_m == 2) fn<2, 2>(h);                            \
      else if (_m == 3) fn<2, 3>(h);                            \
      else fn<2, 4>(h);                                         \
    }                                                           \
  } while (0)

void launch_dt(MhHandle* h) { LAUNCH(h, KN_DT, k_dt, 1, MH_DT_BLOCK, h->dev); }

// Phase functions shared by mahyco_cycle and the per-entry-point calls.
void phase_node(MhHandle* h) {
  set_parity(h);
  MH_DISPATCH(launch_node_t, h);
}
void phase_cell(MhHandle* h) { MH_DISPATCH(launch_cell_t, h); }
void phase_remap(MhHandle* h) {
  if (h->opt.with_projection) MH_DISPATCH(launch_sweeps_t, h);
}
template <int DIM, int M>
void launch_hist_t(MhHandle* h) {
  hipLaunchKernelGGL(k_hist_part<M>, dim3(h->nblk_c), dim3(MH_BLOCK), 0, h->stream, h->dev);
  hipLaunchKernelGGL(k_hist_final<M>, dim3(1), dim3(MH_BLOCK), 0, h->stream, h->dev);
}

void phase_dt(MhHandle* h) {
  launch_dt(h);
  if (h->dev.hist_period > 0) MH_DISPATCH(launch_hist_t, h);
  h->launched += 1;
}

void enqueue_cycle(MhHandle* h) {
  phase_node(h);
  phase_cell(h);
  phase_remap(h);
  phase_dt(h);
}

int read_scalars(MhHandle* h, MhScalars* out) {
  HIPCHK(hipMemcpyAsync(out, h->dev.sc, sizeof(MhScalars), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  return MH_OK;
}

int validate(const MhMesh* mesh, const MhEnvironment* envs, int nb_env, const MhOptions* o) {
  if (!mesh || !envs || !o || nb_env < 1 || nb_env > MH_MAX_ENV) return MH_ERR_INVALID;
  if (mesh->dim != 2 && mesh->dim != 3) return MH_ERR_INVALID;
  if (mesh->nx < 1 || mesh->ny < 1 || (mesh->dim == 3 && mesh->nz < 1)) return MH_ERR_INVALID;
  if (o->ordre_projection < 1 || o->ordre_projection > 3) return MH_ERR_UNSUPPORTED;
  if (o->projection_limiteur_id < 0 || o->projection_limiteur_id > MH_ULTRABEE_G) return MH_ERR_INVALID;
  if (o->projection_limiteur_pure_id < 0 || o->projection_limiteur_pure_id > MH_ULTRABEE_G) return MH_ERR_INVALID;
  if (o->projection_pente_borne && (o->pente_borne_debar_fix < 0 || o->pente_borne_debar_fix > 2))
    return MH_ERR_INVALID;
  if (o->projection_pente_borne && o->ordre_projection == 3) return MH_ERR_UNSUPPORTED;
  if (o->nb_bc < 0 || o->nb_bc > MH_MAX_BC) return MH_ERR_INVALID;
  for (int b = 0; b < o->nb_bc; ++b)
    if (o->bc[b].surface < 0 || o->bc[b].surface > 5 || o->bc[b].type < 0 || o->bc[b].type > 2) return MH_ERR_INVALID;
  if (o->cas_model < 0 || o->cas_model > 5) return MH_ERR_UNSUPPORTED;
  if (o->cas_model == MH_CASE_ADVECTION && !((o->cas_test >= 7 && o->cas_test <= 10) || (o->cas_test >= 18 && o->cas_test <= 21)))
    return MH_ERR_UNSUPPORTED;
  if (o->cas_model == MH_CASE_ADVECTION && nb_env < 2) return MH_ERR_INVALID;  // the square takes env 1's density
  if (o->cas_model == MH_CASE_RIDER && (o->cas_test < 23 || o->cas_test > 38)) return MH_ERR_UNSUPPORTED;
  if (o->cas_model == MH_CASE_RIDER && nb_env < 2) return MH_ERR_INVALID;  // the bubble takes env 1's density
  if (o->sans_lagrange && (!o->with_projection || !(o->deltat_init > 0.) ||
                           !(o->final_time / o->deltat_init <= 1e7)))
    return MH_ERR_UNSUPPORTED;  // the reverse factors are tabulated per iteration up to the final time
  // CSTS: the closed-form energy (with Newton the reference reads p before setting it,
  // MahycoModule.cc:1595-1601; explicit pressure applies the EOS before the forward step)
  if (o->schema_csts && (o->with_newton || o->pression_explicite || o->sans_lagrange)) return MH_ERR_UNSUPPORTED;
  if (o->longueur_caracteristique < 0 || o->longueur_caracteristique > MH_L_MONODIM_Z) return MH_ERR_INVALID;
  if (o->longueur_caracteristique == MH_L_FACES_OPPOSEES && mesh->dim != 3) return MH_ERR_UNSUPPORTED;  // 8 nodes
  if (o->cas_model == MH_CASE_RIEMANN && nb_env < 2) return MH_ERR_INVALID;
  if (o->cas_model == MH_CASE_POINTTRIPLE && nb_env < 3) return MH_ERR_INVALID;
  for (int k = 0; k < nb_env; ++k)
    if (envs[k].eos != MH_EOS_PERFECT_GAS && envs[k].eos != MH_EOS_STIFFENED_GAS && envs[k].eos != MH_EOS_FICTIF)
      return MH_ERR_UNSUPPORTED;
  const long nc = (long)mesh->nx * mesh->ny * (mesh->dim == 3 ? mesh->nz : 1);
  const long nn = (long)(mesh->nx + 1) * (mesh->ny + 1) * (mesh->dim == 3 ? mesh->nz + 1 : 1);
  if (nc > (1L << 31) || nn > (1L << 31)) return MH_ERR_INVALID;
  return MH_OK;
}

}  // namespace

template <int M>
static void launch_eos_t(MhHandle* h, int env) {
  hipLaunchKernelGGL(k_apply_eos<M>, dim3(h->nblk_c), dim3(MH_BLOCK), 0, h->stream, h->dev, env);
}
template <int M>
static void launch_pm_t(MhHandle* h) {
  hipLaunchKernelGGL(k_pression_moyenne<M>, dim3(h->nblk_c), dim3(MH_BLOCK), 0, h->stream, h->dev);
}


// ============================================================================ ensembles
// Many independent cases advance together: one launch per kernel phase covers the blocks of
// every active case of a (dimension, environment count) group.  A case is admitted only at
// a global iteration multiple of 6, so all active cases share the coordinate parity and the
// sweep rotation (RemapADIService.cc:25, 59-60), hence the sweep direction of each launch.
size_t case_footprint(const MhMesh& mesh, int m, const MhOptions& o) {
  const long nc = (long)mesh.nx * mesh.ny * (mesh.dim == 3 ? mesh.nz : 1);
  const long nn = (long)(mesh.nx + 1) * (mesh.ny + 1) * (mesh.dim == 3 ? mesh.nz + 1 : 1);
  const long nv = 3 * m + 1;
  long words = (mesh.nx + mesh.ny + mesh.nz + 3) + nc /*VE*/ + 10 * nn /*u, xl x2, mn*/ + 7 * nc +
               6L * m * nc + 3 * nc + 2 * ((nc + nn) / MH_BLOCK + 8);
  long bytes = nc;  // mask
  if (o.with_projection) {
    words += 2 * nv * nc + m * nc + 10 * nn;
    const bool gen_lim = o.ordre_projection == 2 && o.projection_limiteur_id >= MH_MINMOD_G;
    if (o.ordre_projection == 3 || o.projection_pente_borne || gen_lim) words += (long)mesh.dim * nc;  // hc
    if (o.ordre_projection == 3 || o.pente_borne_mixte) bytes += nc;                                  // est
    if (o.projection_pente_borne) words += nv * nc;                                                   // phi0
    if (o.sans_lagrange && o.deltat_init > 0.) words += 3 * nn + (long)(o.final_time / o.deltat_init) + 8;
  }
  if (o.schema_csts) {  // saved velocity, pseudo-viscosity history
    words += 3 * nn + 3 * nc + (o.with_projection ? 0 : 2L * m * nc);
    bytes += nc;
  }
  return (size_t)words * 8 + (size_t)bytes;
}

// cases whose sweeps take the pente-borne / generalised-limiter kernels
bool ens_gen(const MhHandle* h) {
  const MhDev& S = h->dev;
  return S.order2 != 2 && (S.pb || (S.order2 && S.limiter >= MH_MINMOD_G));
}

struct EnsGroup {
  int dim = 0, m = 0;
  bool o3 = false;   // order-3 remap
  bool gen = false;  // pente-borne or generalised limiters
  bool csts = false; // schema-csts
  bool adv = false;  // advection mode (sans-lagrange): prescribed node velocity, no dual sweeps
  bool lc = false;   // characteristic length other than V^(1/d): k_cell_length first
  std::vector<int> slots;
  int2* d_cmap = nullptr;
  int2* d_nmap = nullptr;
  int* d_list = nullptr;
  int ncb = 0, nnb = 0;
  // block maps of the row-segment x sweeps and the tiled y / z cell sweeps
  bool rows = false;
  int2* d_rcmap = nullptr;
  int2* d_rnmap = nullptr;
  int2* d_tmap[2] = {nullptr, nullptr};
  int nrcb = 0, nrnb = 0, ntb[2] = {0, 0};
};

struct MhEnsembleImpl {
  hipStream_t stream = nullptr;
  int max_slots = 0;
  double budget = 0.;
  size_t used = 0;
  std::vector<MhHandle*> slot;     // active cases
  std::vector<MhHandle*> pending;  // admitted, activated at the next phase boundary
  std::vector<int> pending_slot;
  MhScalars* d_sc = nullptr;       // scalars of every slot
  MhDev* d_tab[2] = {nullptr, nullptr};  // case states by coordinate parity
  std::vector<EnsGroup> groups;
  long launched = 0;
  bool dirty = true;
  long version = 0;
  std::map<long, hipGraphExec_t> graphs;
  std::vector<hipGraph_t> graph_defs;
};

namespace {

void ens_free_groups(MhEnsembleImpl* E) {
  for (auto& g : E->groups) {
    (void)hipFree(g.d_cmap);
    (void)hipFree(g.d_nmap);
    (void)hipFree(g.d_list);
    (void)hipFree(g.d_rcmap);
    (void)hipFree(g.d_rnmap);
    (void)hipFree(g.d_tmap[0]);
    (void)hipFree(g.d_tmap[1]);
  }
  E->groups.clear();
}

void ens_drop_graphs(MhEnsembleImpl* E) {
  for (auto& kv : E->graphs) (void)hipGraphExecDestroy(kv.second);
  for (auto g : E->graph_defs) (void)hipGraphDestroy(g);
  E->graphs.clear();
  E->graph_defs.clear();
}

// Rebuilds the state tables and the block maps after a change of composition.
int ens_rebuild(MhEnsembleImpl* E) {
  HIPCHK(hipStreamSynchronize(E->stream));
  ens_drop_graphs(E);
  ens_free_groups(E);
  std::vector<MhDev> tab[2];
  for (int p = 0; p < 2; ++p) tab[p].resize(E->max_slots);
  for (int sl = 0; sl < E->max_slots; ++sl) {
    MhHandle* h = E->slot[sl];
    if (!h) continue;
    for (int p = 0; p < 2; ++p) {
      MhDev d = h->dev;
      for (int k = 0; k < 3; ++k) {
        d.xlp[k] = h->xl[p][k];
        d.xlc[k] = h->xl[p ^ 1][k];
      }
      tab[p][sl] = d;
    }
    bool found = false;
    for (auto& g : E->groups)
      if (g.dim == h->g.dim && g.m == h->m && g.o3 == (h->dev.order2 == 2) && g.gen == ens_gen(h) &&
          g.csts == (h->dev.csts != 0) && g.adv == (h->dev.sans_lagrange != 0) && g.lc == (h->dev.lc != 0)) {
        g.slots.push_back(sl);
        found = true;
      }
    if (!found) {
      EnsGroup g;
      g.dim = h->g.dim;
      g.m = h->m;
      g.o3 = h->dev.order2 == 2;
      g.gen = ens_gen(h);
      g.csts = h->dev.csts != 0;
      g.adv = h->dev.sans_lagrange != 0;
      g.lc = h->dev.lc != 0;
      g.slots.push_back(sl);
      E->groups.push_back(g);
    }
  }
  for (int p = 0; p < 2; ++p)
    HIPCHK(hipMemcpy(E->d_tab[p], tab[p].data(), sizeof(MhDev) * E->max_slots, hipMemcpyHostToDevice));
  for (auto& g : E->groups) {
    std::vector<int2> cm, nm;
    for (int sl : g.slots) {
      MhHandle* h = E->slot[sl];
      for (int b = 0; b < h->nblk_c; ++b) cm.push_back(make_int2(sl, b));
      for (int b = 0; b < h->nblk_n; ++b) nm.push_back(make_int2(sl, b));
    }
    g.ncb = (int)cm.size();
    g.nnb = (int)nm.size();
    HIPCHK(hipMalloc(&g.d_cmap, sizeof(int2) * cm.size()));
    HIPCHK(hipMalloc(&g.d_nmap, sizeof(int2) * nm.size()));
    HIPCHK(hipMalloc(&g.d_list, sizeof(int) * g.slots.size()));
    HIPCHK(hipMemcpy(g.d_cmap, cm.data(), sizeof(int2) * cm.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(g.d_nmap, nm.data(), sizeof(int2) * nm.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(g.d_list, g.slots.data(), sizeof(int) * g.slots.size(), hipMemcpyHostToDevice));
    // rows and tiles when every case of the group runs them
    g.rows = !g.o3 && !g.gen;
    for (int sl : g.slots) g.rows = g.rows && E->slot[sl]->sweep_mode == MH_SWEEP_TILE;
    if (g.rows) {
      std::vector<int2> rc, rn, t[2];
      for (int sl : g.slots) {
        MhHandle* h = E->slot[sl];
        for (int b = 0, n = row_blocks(h); b < n; ++b) rc.push_back(make_int2(sl, b));
        for (int b = 0, n = node_row_blocks(h); b < n; ++b) rn.push_back(make_int2(sl, b));
        for (int D = 1; D < g.dim; ++D)
          for (int b = 0, n = tile_blocks(h, D); b < n; ++b) t[D - 1].push_back(make_int2(sl, b));
      }
      auto put = [](const std::vector<int2>& v, int2** d, int* n) -> int {
        *n = (int)v.size();
        if (v.empty()) return MH_OK;
        HIPCHK(hipMalloc(d, sizeof(int2) * v.size()));
        HIPCHK(hipMemcpy(*d, v.data(), sizeof(int2) * v.size(), hipMemcpyHostToDevice));
        return MH_OK;
      };
      int st;
      if ((st = put(rc, &g.d_rcmap, &g.nrcb)) != MH_OK) return st;
      if ((st = put(rn, &g.d_rnmap, &g.nrnb)) != MH_OK) return st;
      for (int q = 0; q < 2; ++q)
        if ((st = put(t[q], &g.d_tmap[q], &g.ntb[q])) != MH_OK) return st;
    }
  }
  E->dirty = false;
  E->version++;
  return MH_OK;
}

template <int DIM, int M>
void ens_group_cycle(MhEnsembleImpl* E, const EnsGroup& g) {
  const int par = (int)(E->launched & 1);
  const MhDev* tab = E->d_tab[par];
  hipStream_t st = E->stream;
  if (g.adv) hipLaunchKernelGGL((k_node_adv_batch<DIM>), dim3(g.nnb), dim3(MH_BLOCK), 0, st, tab, g.d_nmap);
  else hipLaunchKernelGGL((k_node_lag_batch<DIM>), dim3(g.nnb), dim3(MH_BLOCK), 0, st, tab, g.d_nmap);
  if (g.lc) hipLaunchKernelGGL((k_cell_length_batch<DIM>), dim3(g.ncb), dim3(MH_BLOCK), 0, st, tab, g.d_cmap);
  if (g.csts) hipLaunchKernelGGL((k_cell_lag_csts_batch<DIM, M>), dim3(g.ncb), dim3(MH_BLOCK), 0, st, tab, g.d_cmap);
  else hipLaunchKernelGGL((k_cell_lag_batch<DIM, M>), dim3(g.ncb), dim3(MH_BLOCK), 0, st, tab, g.d_cmap);
  if (g.csts) hipLaunchKernelGGL((k_node_fwd_batch<DIM>), dim3(g.nnb), dim3(MH_BLOCK), 0, st, tab, g.d_nmap);
  const int sens = (int)(E->launched % DIM);
  if (g.o3 || g.gen) hipLaunchKernelGGL((k_hcell_batch<DIM>), dim3(g.ncb), dim3(MH_BLOCK), 0, st, tab, g.d_cmap);
  for (int s = 0; s < DIM; ++s) {
    const int D = (s + sens) % DIM, first = s == 0, last = s == DIM - 1, sel = s & 1;
    if (g.gen) {
      if (D == 0) {
        hipLaunchKernelGGL((k_sweep_cell_gen_batch<DIM, 0, M>), dim3(g.ncb), dim3(MH_BLOCK), 0, st, tab, g.d_cmap, sel, first);
        if (!g.adv) hipLaunchKernelGGL((k_sweep_node_batch<DIM, 0, MH_LIM_ANY>), dim3(g.nnb), dim3(MH_BLOCK), 0, st, tab, g.d_nmap, sel,
                           first, last);
      } else if (D == 1) {
        hipLaunchKernelGGL((k_sweep_cell_gen_batch<DIM, 1, M>), dim3(g.ncb), dim3(MH_BLOCK), 0, st, tab, g.d_cmap, sel, first);
        if (!g.adv) hipLaunchKernelGGL((k_sweep_node_batch<DIM, 1, MH_LIM_ANY>), dim3(g.nnb), dim3(MH_BLOCK), 0, st, tab, g.d_nmap, sel,
                           first, last);
      } else if constexpr (DIM == 3) {
        hipLaunchKernelGGL((k_sweep_cell_gen_batch<3, 2, M>), dim3(g.ncb), dim3(MH_BLOCK), 0, st, tab, g.d_cmap, sel, first);
        if (!g.adv) hipLaunchKernelGGL((k_sweep_node_batch<3, 2, MH_LIM_ANY>), dim3(g.nnb), dim3(MH_BLOCK), 0, st, tab, g.d_nmap, sel,
                           first, last);
      }
      continue;
    }
    if (g.o3) {
      if (D == 0)
        hipLaunchKernelGGL((k_sweep_cell3_batch<DIM, 0, M>), dim3(g.ncb), dim3(MH_BLOCK), 0, st, tab, g.d_cmap, sel, first);
      else if (D == 1)
        hipLaunchKernelGGL((k_sweep_cell3_batch<DIM, 1, M>), dim3(g.ncb), dim3(MH_BLOCK), 0, st, tab, g.d_cmap, sel, first);
      else if constexpr (DIM == 3)
        hipLaunchKernelGGL((k_sweep_cell3_batch<3, 2, M>), dim3(g.ncb), dim3(MH_BLOCK), 0, st, tab, g.d_cmap, sel, first);
    }
    if (g.o3 && D == 0) {
      if (!g.adv) hipLaunchKernelGGL((k_sweep_node_batch<DIM, 0, -1>), dim3(g.nnb), dim3(MH_BLOCK), 0, st, tab, g.d_nmap, sel, first, last);
    } else if (g.o3 && D == 1) {
      if (!g.adv) hipLaunchKernelGGL((k_sweep_node_batch<DIM, 1, -1>), dim3(g.nnb), dim3(MH_BLOCK), 0, st, tab, g.d_nmap, sel, first, last);
    } else if (g.o3) {
      if (!g.adv) hipLaunchKernelGGL((k_sweep_node_batch<DIM, (DIM == 3 ? 2 : 0), -1>), dim3(g.nnb), dim3(MH_BLOCK), 0, st, tab, g.d_nmap, sel, first, last);
    } else if (g.rows && D == 0) {
      hipLaunchKernelGGL((k_sweep_cell_row_batch<DIM, M, -1>), dim3(g.nrcb), dim3(MH_BLOCK), 0, st, tab, g.d_rcmap, sel, first);
      if (!g.adv) hipLaunchKernelGGL((k_sweep_node_row_batch<DIM, -1>), dim3(g.nrnb), dim3(MH_BLOCK), 0, st, tab, g.d_rnmap, sel, first, last);
    } else if (g.rows && D == 1) {
      hipLaunchKernelGGL((k_sweep_cell_tile_batch<DIM, 1, M, -1>), dim3(g.ntb[0]), dim3(MH_BLOCK), 0, st, tab, g.d_tmap[0], sel, first);
      if (!g.adv) hipLaunchKernelGGL((k_sweep_node_batch<DIM, 1, -1>), dim3(g.nnb), dim3(MH_BLOCK), 0, st, tab, g.d_nmap, sel, first, last);
    } else if (g.rows) {
      if constexpr (DIM == 3)
        hipLaunchKernelGGL((k_sweep_cell_tile_batch<3, 2, M, -1>), dim3(g.ntb[1]), dim3(MH_BLOCK), 0, st, tab, g.d_tmap[1], sel, first);
      if (!g.adv) hipLaunchKernelGGL((k_sweep_node_batch<DIM, (DIM == 3 ? 2 : 0), -1>), dim3(g.nnb), dim3(MH_BLOCK), 0, st, tab, g.d_nmap, sel, first, last);
    } else if (D == 0) {
      hipLaunchKernelGGL((k_sweep_cell_batch<DIM, 0, M, -1>), dim3(g.ncb), dim3(MH_BLOCK), 0, st, tab, g.d_cmap, sel, first);
      if (!g.adv) hipLaunchKernelGGL((k_sweep_node_batch<DIM, 0, -1>), dim3(g.nnb), dim3(MH_BLOCK), 0, st, tab, g.d_nmap, sel, first, last);
    } else if (D == 1) {
      hipLaunchKernelGGL((k_sweep_cell_batch<DIM, 1, M, -1>), dim3(g.ncb), dim3(MH_BLOCK), 0, st, tab, g.d_cmap, sel, first);
      if (!g.adv) hipLaunchKernelGGL((k_sweep_node_batch<DIM, 1, -1>), dim3(g.nnb), dim3(MH_BLOCK), 0, st, tab, g.d_nmap, sel, first, last);
    } else {
      hipLaunchKernelGGL((k_sweep_cell_batch<DIM, (DIM == 3 ? 2 : 0), M, -1>), dim3(g.ncb), dim3(MH_BLOCK), 0, st, tab, g.d_cmap, sel, first);
      if (!g.adv) hipLaunchKernelGGL((k_sweep_node_batch<DIM, (DIM == 3 ? 2 : 0), -1>), dim3(g.nnb), dim3(MH_BLOCK), 0, st, tab, g.d_nmap, sel, first, last);
    }
  }
  hipLaunchKernelGGL((k_finalize_batch<DIM, M>), dim3(g.ncb), dim3(MH_BLOCK), 0, st, tab, g.d_cmap, DIM & 1);
  hipLaunchKernelGGL(k_dt_batch, dim3((unsigned)g.slots.size()), dim3(MH_DT_BLOCK), 0, st, tab, g.d_list);
}

void ens_cycle(MhEnsembleImpl* E) {
  for (const auto& g : E->groups) {
    if (g.dim == 3) {
      if (g.m == 1) ens_group_cycle<3, 1>(E, g);
      else if (g.m == 2) ens_group_cycle<3, 2>(E, g);
      else if (g.m == 3) ens_group_cycle<3, 3>(E, g);
      else ens_group_cycle<3, 4>(E, g);
    } else {
      if (g.m == 1) ens_group_cycle<2, 1>(E, g);
      else if (g.m == 2) ens_group_cycle<2, 2>(E, g);
      else if (g.m == 3) ens_group_cycle<2, 3>(E, g);
      else ens_group_cycle<2, 4>(E, g);
    }
  }
  E->launched += 1;
}

// Activates the pending cases (only at a phase boundary).
int ens_activate(MhEnsembleImpl* E) {
  if (E->pending.empty() || (E->launched % 6) != 0) return MH_OK;
  for (size_t i = 0; i < E->pending.size(); ++i) E->slot[E->pending_slot[i]] = E->pending[i];
  E->pending.clear();
  E->pending_slot.clear();
  E->dirty = true;
  return MH_OK;
}

}  // namespace

// ============================================================================ C ABI
extern "C" {

static int create_case(const MhMesh* mesh, const MhEnvironment* envs, int nb_env, const MhOptions* options,
                       hipStream_t stream, MhScalars* ext_sc, MhHandle** out);

int mahyco_create(const MhMesh* mesh, const MhEnvironment* envs, int nb_env, const MhOptions* options,
                  MhHandle** out) {
  return create_case(mesh, envs, nb_env, options, nullptr, nullptr, out);
}

int mahyco_create_on_stream(const MhMesh* mesh, const MhEnvironment* envs, int nb_env, const MhOptions* options,
                            void* stream, MhHandle** out) {
  if (!stream) return MH_ERR_INVALID;
  return create_case(mesh, envs, nb_env, options, (hipStream_t)stream, nullptr, out);
}

static int create_case(const MhMesh* mesh, const MhEnvironment* envs, int nb_env, const MhOptions* options,
                       hipStream_t stream, MhScalars* ext_sc, MhHandle** out) {
  if (!out) return MH_ERR_INVALID;
  *out = nullptr;
  int st = validate(mesh, envs, nb_env, options);
  if (st != MH_OK) return st;
  MhHandle* h = new MhHandle();
  h->mesh = *mesh;
  h->opt = *options;
  h->m = nb_env;
  for (int k = 0; k < nb_env; ++k) h->env[k] = envs[k];
  MhGrid& g = h->g;
  g.dim = mesh->dim;
  g.nx = mesh->nx;
  g.ny = mesh->ny;
  g.nz = g.dim == 3 ? mesh->nz : 1;
  g.nnx = g.nx + 1;
  g.nny = g.ny + 1;
  g.nnz = g.dim == 3 ? g.nz + 1 : 1;
  g.nc = (long)g.nx * g.ny * g.nz;
  g.nn = (long)g.nnx * g.nny * g.nnz;
  // block counts are multiples of 8 (the XCD count) for the XCD-contiguous block order
  h->nblk_c = (int)(((g.nc + MH_BLOCK - 1) / MH_BLOCK + 7) / 8 * 8);
  h->nblk_n = (int)(((g.nn + MH_BLOCK - 1) / MH_BLOCK + 7) / 8 * 8);
  if (stream) {
    h->stream = stream;
    h->owns_stream = false;
  } else {
    st = hip_check(hipStreamCreate(&h->stream), "hipStreamCreate");
    if (st != MH_OK) {
      delete h;
      return st;
    }
  }
  h->ext_sc = ext_sc;
  {
    const char* nk = getenv("MAHYCO_NODE_KERNEL");
    h->node_tiled = !(nk && strcmp(nk, "flat") == 0);
    const char* sw = getenv("MAHYCO_SWEEP");
    h->sweep_mode = (sw && strcmp(sw, "flat") == 0) ? MH_SWEEP_FLAT : MH_SWEEP_TILE;
  }
  HostInit I;
  st = host_init(h, I);
  if (st != MH_OK) {
    mahyco_destroy(h);
    return st;
  }
  MhDev& S = h->dev;
  memset(&S, 0, sizeof(S));
  S.g = g;
  S.m = nb_env;
  S.nv = 3 * nb_env + 1;
  for (int k = 0; k < nb_env; ++k) {
    S.env[k].gamma = envs[k].adiabatic_cst;
    S.env[k].pinf = envs[k].limit_tension;
    S.env[k].a1 = envs[k].linear_pseudo_coeff;
    S.env[k].a2 = envs[k].quadratic_pseudo_coeff;
    S.env[k].stiffened = envs[k].eos == MH_EOS_STIFFENED_GAS;
    S.env[k].fictif = envs[k].eos == MH_EOS_FICTIF;
    S.env[k].ft0 = envs[k].fictif_tdebut;
    S.env[k].ft1 = envs[k].fictif_tfin;
    S.env[k].fp0 = envs[k].fictif_p0;
    S.env[k].fdpdt = envs[k].fictif_dpdt;
  }
  S.cfl = options->cfl;
  S.thr_remap = options->remap_threshold;
  S.dtmin = options->deltat_min;
  S.dtmax = options->deltat_max;
  S.tfinal = options->final_time;
  S.order2 = options->ordre_projection - 1;
  S.limiter = options->projection_limiteur_id;
  S.pb = options->projection_pente_borne ? 1 : 0;
  S.debar = options->pente_borne_debar_fix;
  S.pb_mixte = options->pente_borne_mixte ? 1 : 0;
  S.lim_pure = options->projection_limiteur_pure_id;
  S.simple_pente = options->projection_simple_pente ? 1 : 0;
  S.nsweep = (g.nx > 1) + (g.ny > 1) + (g.dim == 3 && g.nz > 1);
  S.newton = options->with_newton ? 1 : 0;
  S.p_explicit = options->pression_explicite ? 1 : 0;
  S.thr_module = options->threshold;
  S.sans_lagrange = options->sans_lagrange ? 1 : 0;
  S.rider_reverse = options->rider_reverse ? 1 : 0;
  S.csts = options->schema_csts ? 1 : 0;
  S.lc = options->longueur_caracteristique;
  S.dt_init = options->deltat_init;
  S.with_proj = options->with_projection ? 1 : 0;
  S.vitson_max = options->calcul_vitson_max ? 1 : 0;
  S.nbc = options->nb_bc;
  for (int b = 0; b < options->nb_bc; ++b) {
    S.bc_surf[b] = options->bc[b].surface;
    S.bc_type[b] = options->bc[b].type;
    S.bc_val[b] = options->bc[b].value;
  }
  const long nc = g.nc, nn = g.nn;
  const size_t ne = (size_t)nb_env * nc;
  double *ax, *ay, *az, *VE;
  int ok = MH_OK;
  ok |= dalloc(h, &ax, g.nnx);
  ok |= dalloc(h, &ay, g.nny);
  ok |= dalloc(h, &az, I.az.size());
  ok |= dalloc(h, &VE, nc);
  for (int d = 0; d < 3; ++d) {
    ok |= dalloc(h, &S.u[d], nn);
    ok |= dalloc(h, &h->xl[0][d], nn);
    ok |= dalloc(h, &h->xl[1][d], nn);
  }
  ok |= dalloc(h, &S.mn, nn);
  ok |= dalloc(h, &S.rho, nc);
  ok |= dalloc(h, &S.p, nc);
  ok |= dalloc(h, &S.c, nc);
  ok |= dalloc(h, &S.q, nc);
  ok |= dalloc(h, &S.M, nc);
  ok |= dalloc(h, &S.e, nc);
  ok |= dalloc(h, &S.fg, nc);
  ok |= dalloc(h, &S.mask, nc);
  ok |= dalloc(h, &S.fk, ne);
  ok |= dalloc(h, &S.chik, ne);
  ok |= dalloc(h, &S.chig, nc);
  ok |= dalloc(h, &S.rhok, ne);
  ok |= dalloc(h, &S.ek, ne);
  ok |= dalloc(h, &S.pk, ne);
  ok |= dalloc(h, &S.Mk, ne);
  ok |= dalloc(h, &S.L, nc);
  ok |= dalloc(h, &S.Vl, nc);
  ok |= dalloc(h, &S.partial, 2 * (size_t)std::max(h->nblk_c, h->nblk_n));  // CFL partials: 2 per block
  ok |= dalloc(h, &S.err, 4);
  if (ext_sc) S.sc = ext_sc;
  else ok |= dalloc(h, &S.sc, 1);
  if (options->with_projection) {
    const size_t nv = 3 * nb_env + 1;
    ok |= dalloc(h, &h->Ua, nv * nc);
    ok |= dalloc(h, &h->Ub, nv * nc);
    ok |= dalloc(h, &h->dflux, (size_t)nb_env * nc);
    ok |= dalloc(h, &h->Ud, (size_t)4 * nn);
    ok |= dalloc(h, &h->Pa, (size_t)3 * nn);
    ok |= dalloc(h, &h->Pb, (size_t)3 * nn);
    const bool gen_lim = options->ordre_projection == 2 && options->projection_limiteur_id >= MH_MINMOD_G;
    if (options->ordre_projection == 3 || options->projection_pente_borne || gen_lim)
      ok |= dalloc(h, &S.hc, (size_t)g.dim * nc);
    if (options->ordre_projection == 3 || options->pente_borne_mixte) ok |= dalloc(h, &S.est, (size_t)nc);
    if (options->projection_pente_borne) ok |= dalloc(h, &S.phi0, (size_t)nv * nc);
    if (options->sans_lagrange) {  // prescribed velocity and the reverse factors
      for (int d = 0; d < 3; ++d) ok |= dalloc(h, &h->u0[d], nn);
      ok |= dalloc(h, &h->rev_cos, I.rev_cos.size());
    }
  } else {
    ok |= dalloc(h, &h->Ua, 1);
  }
  if (options->schema_csts) {  // saved velocity; pseudo-viscosity history (per cell with remap, per environment without)
    for (int d = 0; d < 3; ++d) ok |= dalloc(h, &S.un[d], nn);
    ok |= dalloc(h, &S.qm1, nc);
    ok |= dalloc(h, &S.qe, nc);
    ok |= dalloc(h, &S.qem1, nc);
    ok |= dalloc(h, &S.mkm1, nc);
    if (!options->with_projection) {
      ok |= dalloc(h, &S.qkn, ne);
      ok |= dalloc(h, &S.qknm1, ne);
    }
  }
  if (ok != MH_OK) {
    mahyco_destroy(h);
    return MH_ERR_NOMEM;
  }
  S.ax = ax;
  S.ay = ay;
  S.az = az;
  S.VE = VE;
  S.Ua = h->Ua;
  S.Ub = h->Ub;
  S.dflux = h->dflux;
  S.Ud = h->Ud;
  S.Pa = h->Pa;
  S.Pb = h->Pb;
  S.nblk_c = h->nblk_c;
  S.nblk_n = h->nblk_n;
  S.max_iter = 0;
  // initial Lagrangian coordinates = Eulerian grid
  std::vector<double> xs[3];
  for (int d = 0; d < 3; ++d) xs[d].assign(nn, 0.);
  for (int k = 0; k < g.nnz; ++k)
    for (int j = 0; j < g.nny; ++j)
      for (int i = 0; i < g.nnx; ++i) {
        const long n = i + (long)g.nnx * (j + (long)g.nny * k);
        xs[0][n] = I.ax[i];
        xs[1][n] = I.ay[j];
        xs[2][n] = g.dim == 3 ? I.az[k] : 0.;
      }
  MhScalars sc;
  memset(&sc, 0, sizeof(sc));
  sc.t = 0.;
  sc.dt = options->deltat_init;
  sc.dt_old = 0.;
  std::vector<int> err(4, 0);
  ok = MH_OK;
  ok |= upload(h, ax, I.ax);
  ok |= upload(h, ay, I.ay);
  ok |= upload(h, az, I.az);
  ok |= upload(h, VE, I.VE);
  for (int d = 0; d < 3; ++d) {
    ok |= upload(h, S.u[d], I.u[d]);
    ok |= upload(h, h->xl[0][d], xs[d]);
    ok |= upload(h, h->xl[1][d], xs[d]);
  }
  std::vector<double> zc, zn, ze;  // kept until the stream is synchronised below (asynchronous uploads)
  if (S.csts) {  // the pseudo-viscosities at n and n-1 start at 0; the saved velocity is written before use
    zc.assign(nc, 0.);
    zn.assign(nn, 0.);
    ok |= upload(h, S.qm1, zc);
    ok |= upload(h, S.qe, zc);
    ok |= upload(h, S.qem1, zc);
    ok |= upload(h, S.mkm1, I.mask);
    for (int d = 0; d < 3; ++d) ok |= upload(h, S.un[d], zn);
    if (!options->with_projection) {
      ze.assign(ne, 0.);
      ok |= upload(h, S.qkn, ze);
      ok |= upload(h, S.qknm1, ze);
    }
  }
  ok |= upload(h, S.rho, I.rho);
  ok |= upload(h, S.p, I.p);
  ok |= upload(h, S.c, I.c);
  ok |= upload(h, S.q, I.q);
  ok |= upload(h, S.M, I.M);
  ok |= upload(h, S.e, I.e);
  ok |= upload(h, S.fg, I.fg);
  ok |= upload(h, S.mask, I.mask);
  ok |= upload(h, S.fk, I.fk);
  ok |= upload(h, S.chik, I.chik);
  ok |= upload(h, S.chig, I.chig);
  ok |= upload(h, S.rhok, I.rhok);
  ok |= upload(h, S.ek, I.ek);
  ok |= upload(h, S.pk, I.pk);
  ok |= upload(h, S.Mk, I.Mk);
  ok |= upload(h, S.err, err);
  std::vector<int8_t> est0;  // mixed-cell flags: Arcane starts est_mixte and est_pure at 0
  if (S.est) {
    est0.assign(nc, 0);
    ok |= upload(h, S.est, est0);
  }
  if (h->rev_cos) {
    for (int d = 0; d < 3; ++d) {
      ok |= upload(h, h->u0[d], I.u[d]);
      S.u0[d] = h->u0[d];
    }
    ok |= upload(h, h->rev_cos, I.rev_cos);
    S.rev_cos = h->rev_cos;
    S.rev_n = (long)I.rev_cos.size();
  }
  std::vector<double> phi00;  // pente-borne phi: Arcane starts m_phi_lagrange at 0
  if (S.phi0) {
    phi00.assign((size_t)S.nv * nc, 0.);
    ok |= upload(h, S.phi0, phi00);
  }
  ok |= hip_check(hipMemcpyAsync(S.sc, &sc, sizeof(sc), hipMemcpyHostToDevice, h->stream), "upload scalars");
  ok |= hip_check(hipStreamSynchronize(h->stream), "hipStreamSynchronize");
  if (ok != MH_OK) {
    mahyco_destroy(h);
    return MH_ERR_DEVICE;
  }
  *out = h;
  return MH_OK;
}

void mahyco_destroy(MhHandle* h) {
  if (!h) return;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  for (auto& kv : h->graphs) (void)hipGraphExecDestroy(kv.second);
  for (auto gdef : h->graph_defs) (void)hipGraphDestroy(gdef);
  for (auto& r : h->prof_pending) {
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
  }
  for (void* p : h->allocs) (void)hipFree(p);
  if (h->stream && h->owns_stream) (void)hipStreamDestroy(h->stream);
  delete h;
}

int mahyco_cycle(MhHandle* h, int ncycles) {
  if (!h || ncycles < 0) return MH_ERR_INVALID;
  if (h->phase != 0) return MH_ERR_INVALID;
  for (int i = 0; i < ncycles; ++i) enqueue_cycle(h);
  return hip_check(hipGetLastError(), "kernel launch");
}

int mahyco_cycle_graph(MhHandle* h, int ncycles, int cycles_per_graph) {
  if (!h || ncycles < 0 || cycles_per_graph < 1) return MH_ERR_INVALID;
  if (h->phase != 0) return MH_ERR_INVALID;
  if (h->prof) return mahyco_cycle(h, ncycles);
  // a replayable graph returns the coordinate parity and the sweep rotation to their
  // starting values: its length is a multiple of lcm(2, dim)
  const int period = h->g.dim == 3 ? 6 : 2;
  const int P = ((cycles_per_graph + period - 1) / period) * period;
  int done = 0;
  while (ncycles - done >= P) {
    const long key = ((long)P << 8) | ((h->launched & 1) << 4) | h->sens;
    auto it = h->graphs.find(key);
    if (it == h->graphs.end()) {
      hipGraph_t gdef = nullptr;
      HIPCHK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
      const long l0 = h->launched;
      const int s0 = h->sens;
      for (int i = 0; i < P; ++i) enqueue_cycle(h);
      const hipError_t ec = hipStreamEndCapture(h->stream, &gdef);
      h->launched = l0;  // capture enqueues nothing: roll the host counters back,
      h->sens = s0;      // also when the capture failed
      if (ec != hipSuccess) {
        if (gdef) (void)hipGraphDestroy(gdef);
        return hip_check(ec, "hipStreamEndCapture");
      }
      hipGraphExec_t ex;
      const hipError_t ei = hipGraphInstantiate(&ex, gdef, nullptr, nullptr, 0);
      if (ei != hipSuccess) {
        (void)hipGraphDestroy(gdef);
        return hip_check(ei, "hipGraphInstantiate");
      }
      h->graph_defs.push_back(gdef);
      it = h->graphs.emplace(key, ex).first;
      h->graph_captures++;
    }
    HIPCHK(hipGraphLaunch(it->second, h->stream));
    h->launched += P;
    done += P;
  }
  for (; done < ncycles; ++done) enqueue_cycle(h);
  return hip_check(hipGetLastError(), "graph launch");
}

int mahyco_synchronize(MhHandle* h) {
  if (!h) return MH_ERR_INVALID;
  HIPCHK(hipStreamSynchronize(h->stream));
  return mahyco_status(h);
}

int mahyco_status(MhHandle* h) {
  if (!h) return MH_ERR_INVALID;
  MhScalars sc;
  int st = read_scalars(h, &sc);
  if (st != MH_OK) return st;
  if (sc.status != 0) return sc.status;
  return sc.stopped ? MH_STOPPED : MH_OK;
}

double mahyco_time(MhHandle* h) {
  MhScalars sc;
  if (!h || read_scalars(h, &sc) != MH_OK) return NAN;
  return sc.t;
}
double mahyco_deltat(MhHandle* h) {
  MhScalars sc;
  if (!h || read_scalars(h, &sc) != MH_OK) return NAN;
  return sc.dt;
}
int mahyco_iteration(MhHandle* h) {
  MhScalars sc;
  if (!h || read_scalars(h, &sc) != MH_OK) return -1;
  return sc.iteration;
}
long mahyco_nb_cells(MhHandle* h) { return h ? h->g.nc : -1; }
long mahyco_graph_captures(MhHandle* h) { return h ? h->graph_captures : -1; }
long mahyco_nb_nodes(MhHandle* h) { return h ? h->g.nn : -1; }

int mahyco_get_field(MhHandle* h, int field, int env, int comp, double* out) {
  if (!h || !out) return MH_ERR_INVALID;
  MhScalars sc;
  int st = read_scalars(h, &sc);
  if (st != MH_OK) return st;
  const long nc = h->g.nc, nn = h->g.nn;
  MhDev& S = h->dev;
  auto d2h = [&](const double* d, double* hst, long n) -> int {
    HIPCHK(hipMemcpy(hst, d, n * sizeof(double), hipMemcpyDeviceToHost));
    return MH_OK;
  };
  if (field >= 100) {
    if (comp < 0 || comp > 2) return MH_ERR_INVALID;
    if (field == MH_N_VELOCITY) return d2h(S.u[comp], out, nn);
    if (field == MH_N_NODE_MASS) {
      // m_node_mass after the iteration: sum over cells of M / 2^dim, in cell order
      std::vector<double> M(nc);
      st = d2h(S.M, M.data(), nc);
      if (st != MH_OK) return st;
      const MhGrid& g = h->g;
      const double w = g.dim == 3 ? .125 : .25;
      for (long n = 0; n < nn; ++n) {
        const int i = (int)(n % g.nnx);
        const long r = n / g.nnx;
        const int j = (int)(r % g.nny), k = (int)(r / g.nny);
        double s = 0.;
        for (int kk = (g.dim == 3 ? k - 1 : 0); kk <= (g.dim == 3 ? k : 0); ++kk)
          for (int jj = j - 1; jj <= j; ++jj)
            for (int ii = i - 1; ii <= i; ++ii) {
              if (ii < 0 || ii >= g.nx || jj < 0 || jj >= g.ny || kk < 0 || kk >= g.nz) continue;
              s += w * M[ii + (long)g.nx * (jj + (long)g.ny * kk)];
            }
        out[n] = s;
      }
      return MH_OK;
    }
    if (field == MH_N_COORD) {
      if (h->opt.with_projection) {
        const MhGrid& g = h->g;
        for (long n = 0; n < nn; ++n) {
          const int i = (int)(n % g.nnx);
          const long r = n / g.nnx;
          const int j = (int)(r % g.nny), k = (int)(r / g.nny);
          out[n] = comp == 0 ? axis_coord(h->mesh, 0, i, g.nx)
                             : comp == 1 ? axis_coord(h->mesh, 1, j, g.ny)
                                         : (g.dim == 3 ? axis_coord(h->mesh, 2, k, g.nz) : 0.);
        }
        return MH_OK;
      }
      return d2h(h->xl[sc.iteration & 1][comp], out, nn);
    }
    if (field == MH_N_COORD_LAGRANGE) return d2h(h->xl[sc.iteration & 1][comp], out, nn);
    return MH_ERR_INVALID;
  }
  if (env >= h->m || env < -1) return MH_ERR_INVALID;
  std::vector<uint8_t> mask(nc);
  HIPCHK(hipMemcpy(mask.data(), S.mask, nc, hipMemcpyDeviceToHost));
  const double* gsrc = nullptr;
  const double* esrc = nullptr;
  switch (field) {
    case MH_F_DENSITY: gsrc = S.rho; esrc = S.rhok; break;
    case MH_F_PRESSURE: gsrc = S.p; esrc = S.pk; break;
    case MH_F_INTERNAL_ENERGY: gsrc = S.e; esrc = S.ek; break;
    case MH_F_SOUND_SPEED: gsrc = S.c; break;
    case MH_F_PSEUDO_VISCOSITY: gsrc = S.q; break;
    case MH_F_CELL_MASS: gsrc = S.M; esrc = S.Mk; break;
    case MH_F_FRACVOL: gsrc = S.fg; esrc = S.fk; break;
    case MH_F_MASS_FRACTION: gsrc = nullptr; esrc = S.chik; break;
    case MH_F_CARACTERISTIC_LENGTH: gsrc = S.L; break;
    case MH_F_CELL_VOLUME: gsrc = S.VE; break;
    default: break;
  }
  if (field == MH_F_ENV_MASK) {
    for (long c = 0; c < nc; ++c) out[c] = mask[c];
    return MH_OK;
  }
  if (field == MH_F_MATERIAU || field == MH_F_CELL_VOLUME || field == MH_F_FRACVOL ||
      field == MH_F_MASS_FRACTION) {
    std::vector<double> fg(nc), fk((size_t)h->m * nc), V(nc);
    if ((st = d2h(S.fg, fg.data(), nc)) != MH_OK) return st;
    if ((st = d2h(S.fk, fk.data(), (long)h->m * nc)) != MH_OK) return st;
    const bool lagv = !h->opt.with_projection && sc.iteration > 0;
    if ((st = d2h(lagv ? S.Vl : S.VE, V.data(), nc)) != MH_OK) return st;
    std::vector<double> chik, chig;
    if (field == MH_F_MASS_FRACTION) {
      chik.resize((size_t)h->m * nc);
      chig.resize(nc);
      if ((st = d2h(S.chik, chik.data(), (long)h->m * nc)) != MH_OK) return st;
      if ((st = d2h(S.chig, chig.data(), nc)) != MH_OK) return st;
    }
    for (long c = 0; c < nc; ++c) {
      const uint8_t b = mask[c];
      const bool pure = __builtin_popcount(b) == 1;
      const int kp = __builtin_ctz(b ? b : 1);
      double val = 0.;
      if (field == MH_F_MATERIAU) {
        if (pure) val = kp * fg[c];
        else
          for (int k = 0; k < h->m; ++k)
            if ((b >> k) & 1) val += k * fk[k * nc + c];
      } else if (field == MH_F_CELL_VOLUME) {
        if (env < 0) val = V[c];
        else if ((b >> env) & 1) val = pure ? V[c] : fk[env * nc + c] * V[c];
      } else if (field == MH_F_FRACVOL) {
        if (env < 0) val = fg[c];
        else if ((b >> env) & 1) val = pure ? fg[c] : fk[env * nc + c];
      } else {
        if (env < 0) val = 0.;
        else if ((b >> env) & 1) val = pure ? chig[c] : chik[env * nc + c];
      }
      out[c] = val;
    }
    return MH_OK;
  }
  if (!gsrc) return MH_ERR_INVALID;
  if ((st = d2h(gsrc, out, nc)) != MH_OK) return st;
  if (env < 0) return MH_OK;
  std::vector<double> ev;
  if (esrc) {
    ev.resize(nc);
    if ((st = d2h(esrc + (size_t)env * nc, ev.data(), nc)) != MH_OK) return st;
  }
  for (long c = 0; c < nc; ++c) {
    const uint8_t b = mask[c];
    if (!((b >> env) & 1)) out[c] = 0.;
    else if (__builtin_popcount(b) != 1) out[c] = esrc ? ev[c] : out[c];
  }
  return MH_OK;
}

static void drop_graphs(MhHandle* h);

// Writes one field of the state from host memory: the reverse of mahyco_get_field, for a
// caller that owns the state as the reference's module does (the case services and the
// module write the named Arcane variables, src/casTest/IInitialisations.h:26-56,
// MahycoModule.cc:110-140).  Cell values with env = -1 go to the cell (global) arrays; with
// env = k to environment k's value: its partial value in a cell of two or more environments,
// the cell value in a cell of k alone (Arcane's single-environment aliasing), nothing in a
// cell outside k.  Set MH_F_ENV_MASK before the environment values it selects.
int mahyco_set_field(MhHandle* h, int field, int env, int comp, const double* in) {
  if (!h || !in || h->phase != 0 || h->ext_sc) return MH_ERR_INVALID;
  HIPCHK(hipStreamSynchronize(h->stream));
  const long nc = h->g.nc, nn = h->g.nn;
  MhDev& S = h->dev;
  auto h2d = [&](double* d, const double* src, long n) -> int {
    HIPCHK(hipMemcpy(d, src, n * sizeof(double), hipMemcpyHostToDevice));
    return MH_OK;
  };
  if (field >= 100) {
    if (comp < 0 || comp >= h->g.dim) return MH_ERR_INVALID;
    for (long n = 0; n < nn; ++n)
      if (!std::isfinite(in[n]) && field != MH_N_VELOCITY) return MH_ERR_INVALID;
    if (field == MH_N_VELOCITY) return h2d(S.u[comp], in, nn);
    // the positions the next iteration starts from (set_parity: xlp = xl[launched & 1])
    if (field == MH_N_COORD_LAGRANGE || (field == MH_N_COORD && !h->opt.with_projection))
      return h2d(h->xl[h->launched & 1][comp], in, nn);
    return MH_ERR_INVALID;  // node mass follows the cell masses; Euler coordinates are the mesh's
  }
  if (env < -1 || env >= h->m) return MH_ERR_INVALID;
  std::vector<uint8_t> mask(nc);
  if (field == MH_F_ENV_MASK) {
    if (env != -1) return MH_ERR_INVALID;
    for (long c = 0; c < nc; ++c) {
      const double v = in[c];
      if (!(v >= 0. && v < (double)(1 << h->m)) || v != (double)(int)v) return MH_ERR_INVALID;
      mask[c] = (uint8_t)(int)v;
    }
    HIPCHK(hipMemcpy(S.mask, mask.data(), nc, hipMemcpyHostToDevice));
    return MH_OK;
  }
  double* gdst = nullptr;
  double* edst = nullptr;
  bool env_derived = false;  // recomputed per environment by the kernels (not stored)
  switch (field) {
    case MH_F_DENSITY: gdst = S.rho; edst = S.rhok; break;
    case MH_F_PRESSURE: gdst = S.p; edst = S.pk; break;
    case MH_F_INTERNAL_ENERGY: gdst = S.e; edst = S.ek; break;
    case MH_F_SOUND_SPEED: gdst = S.c; env_derived = true; break;
    case MH_F_PSEUDO_VISCOSITY: gdst = S.q; env_derived = true; break;
    case MH_F_CELL_MASS: gdst = S.M; edst = S.Mk; break;
    case MH_F_FRACVOL: gdst = S.fg; edst = S.fk; break;
    case MH_F_MASS_FRACTION: gdst = S.chig; edst = S.chik; break;
    default: return MH_ERR_INVALID;  // volume, length and Materiau are derived from the state
  }
  if (env < 0) return h2d(gdst, in, nc);
  HIPCHK(hipMemcpy(mask.data(), S.mask, nc, hipMemcpyDeviceToHost));
  std::vector<double> g(nc), e;
  int st = MH_OK;
  HIPCHK(hipMemcpy(g.data(), gdst, nc * sizeof(double), hipMemcpyDeviceToHost));
  if (edst) {
    e.resize(nc);
    HIPCHK(hipMemcpy(e.data(), edst + (size_t)env * nc, nc * sizeof(double), hipMemcpyDeviceToHost));
  }
  for (long c = 0; c < nc; ++c) {
    const uint8_t b = mask[c];
    if (!((b >> env) & 1)) continue;
    if (__builtin_popcount(b) == 1) g[c] = in[c];
    else if (edst) e[c] = in[c];
  }
  (void)env_derived;  // a mixed cell's environment sound speed and pseudo-viscosity are re-formed
  if ((st = h2d(gdst, g.data(), nc)) != MH_OK) return st;
  if (edst) st = h2d(edst + (size_t)env * nc, e.data(), nc);
  return st;
}

// Time-loop and module scalars of the state (m_global_time, m_global_deltat,
// m_global_old_deltat, m_global_iteration, m_sens_projection; Mahyco.axl:11): the next
// iteration runs as iteration + 1 of a loop that reached these values.
int mahyco_set_scalars(MhHandle* h, double time, double deltat, double old_deltat, int iteration,
                       int sens_projection) {
  if (!h || h->phase != 0 || h->ext_sc || iteration < 0 || sens_projection < 0 ||
      sens_projection >= h->g.dim || !std::isfinite(time) || !(deltat > 0.) || !(old_deltat >= 0.))
    return MH_ERR_INVALID;
  MhScalars sc;
  int st = read_scalars(h, &sc);
  if (st != MH_OK) return st;
  if ((h->launched ^ (long)iteration) & 1)  // the Lagrangian positions follow the parity
    for (int d = 0; d < 3; ++d) std::swap(h->xl[0][d], h->xl[1][d]);
  sc.t = time;
  sc.dt = deltat;
  sc.dt_old = old_deltat;
  sc.iteration = iteration;
  sc.stopped = 0;
  sc.status = 0;
  HIPCHK(hipMemcpy(h->dev.sc, &sc, sizeof sc, hipMemcpyHostToDevice));
  HIPCHK(hipMemset(h->dev.err, 0, sizeof(int)));
  h->launched = iteration;
  h->sens = sens_projection;
  drop_graphs(h);  // they hold the coordinate buffers of the old parity
  return MH_OK;
}

// ---- entry points -------------------------------------------------------------------
static int step_phase(MhHandle* h, int from, int to) {
  if (!h) return MH_ERR_INVALID;
  if (h->phase != from) return MH_ERR_INVALID;
  h->phase = to;
  return MH_OK;
}
int mahyco_save_values_at_n(MhHandle* h) { return step_phase(h, 0, 1); }
int mahyco_update_velocity(MhHandle* h) {
  int st = step_phase(h, 1, 2);
  if (st == MH_OK) phase_node(h);
  return st;
}
int mahyco_apply_boundary_condition(MhHandle* h) { return step_phase(h, 2, 3); }
int mahyco_update_position(MhHandle* h) { return step_phase(h, 3, 4); }
int mahyco_compute_geometric_values(MhHandle* h) {
  int st = step_phase(h, 4, 5);
  if (st == MH_OK) phase_cell(h);
  return st;
}
int mahyco_update_density(MhHandle* h) { return step_phase(h, 5, 6); }
int mahyco_compute_artificial_viscosity(MhHandle* h) { return step_phase(h, 6, 7); }
int mahyco_update_elasticity_and_plasticity(MhHandle* h) { return step_phase(h, 7, 8); }
int mahyco_update_energy_and_pressure(MhHandle* h) { return step_phase(h, 8, 9); }
int mahyco_remap(MhHandle* h) {
  if (!h || !h->opt.with_projection) return MH_ERR_INVALID;
  int st = step_phase(h, 9, 10);
  if (st == MH_OK) phase_remap(h);
  return st;
}
int mahyco_compute_delta_t(MhHandle* h) {
  if (!h) return MH_ERR_INVALID;
  const int expect = h->opt.with_projection ? 10 : 9;
  int st = step_phase(h, expect, 0);
  if (st == MH_OK) phase_dt(h);
  return st;
}

int mahyco_apply_eos(MhHandle* h, int env) {
  if (!h || env < 0 || env >= h->m) return MH_ERR_INVALID;
  switch (h->m) {
    case 1: launch_eos_t<1>(h, env); break;
    case 2: launch_eos_t<2>(h, env); break;
    case 3: launch_eos_t<3>(h, env); break;
    default: launch_eos_t<4>(h, env); break;
  }
  return hip_check(hipGetLastError(), "k_apply_eos");
}

int mahyco_compute_pression_moyenne(MhHandle* h) {
  if (!h) return MH_ERR_INVALID;
  if (h->opt.sans_lagrange) return MH_OK;  // MahycoModule.cc:1711
  switch (h->m) {
    case 1: launch_pm_t<1>(h); break;
    case 2: launch_pm_t<2>(h); break;
    case 3: launch_pm_t<3>(h); break;
    default: launch_pm_t<4>(h); break;
  }
  return hip_check(hipGetLastError(), "k_pression_moyenne");
}

// ---- ensembles ------------------------------------------------------------------------
size_t mahyco_case_footprint(const MhMesh* mesh, int nb_env, const MhOptions* options) {
  if (!mesh || !options || nb_env < 1 || nb_env > MH_MAX_ENV) return 0;
  return case_footprint(*mesh, nb_env, *options);
}

int mahyco_ensemble_create(int max_slots, double hbm_budget_bytes, MhEnsemble** out) {
  if (!out || max_slots < 1) return MH_ERR_INVALID;
  MhEnsembleImpl* E = new MhEnsembleImpl();
  E->max_slots = max_slots;
  E->budget = hbm_budget_bytes;
  E->slot.assign(max_slots, nullptr);
  int st = hip_check(hipStreamCreate(&E->stream), "hipStreamCreate");
  if (st == MH_OK) st = hip_check(hipMalloc(&E->d_sc, sizeof(MhScalars) * max_slots), "hipMalloc");
  if (st == MH_OK) st = hip_check(hipMalloc(&E->d_tab[0], sizeof(MhDev) * max_slots), "hipMalloc");
  if (st == MH_OK) st = hip_check(hipMalloc(&E->d_tab[1], sizeof(MhDev) * max_slots), "hipMalloc");
  if (st != MH_OK) {
    mahyco_ensemble_destroy((MhEnsemble*)E);
    return st;
  }
  *out = (MhEnsemble*)E;
  return MH_OK;
}

void mahyco_ensemble_destroy(MhEnsemble* e) {
  MhEnsembleImpl* E = (MhEnsembleImpl*)e;
  if (!E) return;
  if (E->stream) (void)hipStreamSynchronize(E->stream);
  ens_drop_graphs(E);
  ens_free_groups(E);
  for (MhHandle* h : E->slot) mahyco_destroy(h);
  for (MhHandle* h : E->pending) mahyco_destroy(h);
  (void)hipFree(E->d_sc);
  (void)hipFree(E->d_tab[0]);
  (void)hipFree(E->d_tab[1]);
  if (E->stream) (void)hipStreamDestroy(E->stream);
  delete E;
}

int mahyco_ensemble_add(MhEnsemble* e, const MhMesh* mesh, const MhEnvironment* envs, int nb_env,
                        const MhOptions* options, int max_iterations, int* slot_out) {
  MhEnsembleImpl* E = (MhEnsembleImpl*)e;
  if (!E || !slot_out || !mesh || !options) return MH_ERR_INVALID;
  *slot_out = -1;
  if (!options->with_projection) return MH_ERR_UNSUPPORTED;
  if (mesh->nx < 2 || mesh->ny < 2 || (mesh->dim == 3 && mesh->nz < 2)) return MH_ERR_UNSUPPORTED;
  const size_t need = case_footprint(*mesh, nb_env, *options);
  if ((double)(E->used + need) > E->budget) return MH_ERR_NOMEM;  // admission by HBM footprint
  int sl = -1;
  for (int i = 0; i < E->max_slots && sl < 0; ++i) {
    if (E->slot[i]) continue;
    bool taken = false;
    for (int q : E->pending_slot) taken |= (q == i);
    if (!taken) sl = i;
  }
  if (sl < 0) return MH_ERR_NOMEM;
  MhHandle* h = nullptr;
  int st = create_case(mesh, envs, nb_env, options, E->stream, E->d_sc + sl, &h);
  if (st != MH_OK) return st;
  h->dev.max_iter = max_iterations > 0 ? max_iterations : 0;
  h->bytes = need;
  E->used += need;
  E->pending.push_back(h);
  E->pending_slot.push_back(sl);
  *slot_out = sl;
  return ens_activate(E);
}

int mahyco_ensemble_remove(MhEnsemble* e, int slot) {
  MhEnsembleImpl* E = (MhEnsembleImpl*)e;
  if (!E || slot < 0 || slot >= E->max_slots) return MH_ERR_INVALID;
  for (size_t i = 0; i < E->pending.size(); ++i)
    if (E->pending_slot[i] == slot) {
      E->used -= E->pending[i]->bytes;
      mahyco_destroy(E->pending[i]);
      E->pending.erase(E->pending.begin() + i);
      E->pending_slot.erase(E->pending_slot.begin() + i);
      return MH_OK;
    }
  if (!E->slot[slot]) return MH_ERR_INVALID;
  HIPCHK(hipStreamSynchronize(E->stream));
  E->used -= E->slot[slot]->bytes;
  mahyco_destroy(E->slot[slot]);
  E->slot[slot] = nullptr;
  E->dirty = true;
  return MH_OK;
}

int mahyco_ensemble_step(MhEnsemble* e, int ncycles, int use_graph) {
  MhEnsembleImpl* E = (MhEnsembleImpl*)e;
  if (!E || ncycles < 0) return MH_ERR_INVALID;
  int done = 0;
  while (done < ncycles) {
    int st = ens_activate(E);
    if (st != MH_OK) return st;
    if (E->dirty) {
      st = ens_rebuild(E);
      if (st != MH_OK) return st;
    }
    // a 6-iteration graph when no admission can happen inside it
    if (use_graph && (E->launched % 6) == 0 && ncycles - done >= 6 && E->pending.empty()) {
      const long key = E->version;
      auto it = E->graphs.find(key);
      if (it == E->graphs.end()) {
        hipGraph_t gdef;
        HIPCHK(hipStreamBeginCapture(E->stream, hipStreamCaptureModeThreadLocal));
        const long l0 = E->launched;
        for (int i = 0; i < 6; ++i) ens_cycle(E);
        HIPCHK(hipStreamEndCapture(E->stream, &gdef));
        E->launched = l0;
        hipGraphExec_t ex;
        HIPCHK(hipGraphInstantiate(&ex, gdef, nullptr, nullptr, 0));
        E->graph_defs.push_back(gdef);
        it = E->graphs.emplace(key, ex).first;
      }
      HIPCHK(hipGraphLaunch(it->second, E->stream));
      E->launched += 6;
      done += 6;
      continue;
    }
    ens_cycle(E);
    done += 1;
  }
  return hip_check(hipGetLastError(), "ensemble launch");
}

int mahyco_ensemble_synchronize(MhEnsemble* e) {
  MhEnsembleImpl* E = (MhEnsembleImpl*)e;
  if (!E) return MH_ERR_INVALID;
  HIPCHK(hipStreamSynchronize(E->stream));
  return MH_OK;
}

int mahyco_ensemble_poll(MhEnsemble* e, int* status, double* time, int* iteration) {
  MhEnsembleImpl* E = (MhEnsembleImpl*)e;
  if (!E) return MH_ERR_INVALID;
  std::vector<MhScalars> sc(E->max_slots);
  HIPCHK(hipMemcpyAsync(sc.data(), E->d_sc, sizeof(MhScalars) * E->max_slots, hipMemcpyDeviceToHost, E->stream));
  HIPCHK(hipStreamSynchronize(E->stream));
  int active = 0;
  for (int i = 0; i < E->max_slots; ++i) {
    bool live = E->slot[i] != nullptr;
    bool pend = false;
    for (int q : E->pending_slot) pend |= (q == i);
    int stv = -100;  // empty slot
    if (pend) stv = 2;  // admitted, not yet active
    else if (live) stv = sc[i].status != 0 ? sc[i].status : (sc[i].stopped ? MH_STOPPED : MH_OK);
    if (live) active++;
    if (status) status[i] = stv;
    if (time) time[i] = live ? sc[i].t : 0.;
    if (iteration) iteration[i] = live ? sc[i].iteration : 0;
  }
  return active;
}

int mahyco_ensemble_get_field(MhEnsemble* e, int slot, int field, int env, int comp, double* out) {
  MhEnsembleImpl* E = (MhEnsembleImpl*)e;
  if (!E || slot < 0 || slot >= E->max_slots || !E->slot[slot]) return MH_ERR_INVALID;
  HIPCHK(hipStreamSynchronize(E->stream));
  return mahyco_get_field(E->slot[slot], field, env, comp, out);
}

double mahyco_ensemble_used_bytes(MhEnsemble* e) {
  MhEnsembleImpl* E = (MhEnsembleImpl*)e;
  return E ? (double)E->used : 0.;
}

int mahyco_history_enable(MhHandle* h, int period, int max_rows) {
  if (!h || period < 0 || max_rows < 0) return MH_ERR_INVALID;
  if (h->launched != 0) return MH_ERR_INVALID;  // before the first iteration
  if (period == 0) {
    h->dev.hist_period = 0;
    return MH_OK;
  }
  const int w = 1 + 5 * h->m;
  int ok = MH_OK;
  ok |= dalloc(h, &h->dev.hist, (size_t)max_rows * w);
  ok |= dalloc(h, &h->dev.hist_part, (size_t)h->nblk_c * 6 * h->m);
  if (ok != MH_OK) return MH_ERR_NOMEM;
  h->dev.hist_period = period;
  h->dev.hist_max = max_rows;
  return MH_OK;
}

int mahyco_history_read(MhHandle* h, double* rows, int max_rows) {
  if (!h) return MH_ERR_INVALID;
  MhScalars sc;
  int st = read_scalars(h, &sc);
  if (st != MH_OK) return st;
  if (h->dev.hist_period <= 0) return 0;
  const int n = std::min(sc.hist_rows, max_rows);
  if (rows && n > 0)
    HIPCHK(hipMemcpy(rows, h->dev.hist, sizeof(double) * (size_t)n * (1 + 5 * h->m), hipMemcpyDeviceToHost));
  return sc.hist_rows;
}

int mahyco_ensight_write(MhHandle* h, const char* dir, const char* const* variables, int nb_variables) {
  if (!h || !dir || (!variables && nb_variables > 0) || nb_variables < 0) return MH_ERR_INVALID;
  std::vector<const mh_output::VarSpec*> vars;
  for (int i = 0; i < nb_variables; ++i) {
    const mh_output::VarSpec* s = variables[i] ? mh_output::find_var(variables[i]) : nullptr;
    if (!s) return MH_ERR_INVALID;
    vars.push_back(s);
  }
  if (!h->ens_times.empty() && vars != h->ens_vars) return MH_ERR_INVALID;  // one variable list per case
  const std::string d(dir);
  if (!mh_output::make_dir(d) || !mh_output::make_dir(d + "/ensight.geo")) return MH_ERR_INVALID;
  for (const auto* s : vars)
    if (!mh_output::make_dir(d + "/" + s->name)) return MH_ERR_INVALID;
  const int step = (int)h->ens_times.size();
  const std::string sn = mh_output::step_name(step);
  const int n[3] = {h->g.nx, h->g.ny, h->g.nz};
  int st = mh_output::write_geometry(h, h->g.dim, n, d + "/ensight.geo/ensight.geo" + sn);
  for (size_t i = 0; st == MH_OK && i < vars.size(); ++i)
    st = mh_output::write_variable(h, h->g.dim, *vars[i], d + "/" + vars[i]->name + "/" + vars[i]->name + sn);
  if (st != MH_OK) return st;
  h->ens_vars = vars;
  h->ens_times.push_back(mahyco_time(h));
  return mh_output::write_case(d, h->ens_vars, h->ens_times);
}

int mahyco_history_write_gnuplot(MhHandle* h, const char* dir, const char* const* env_names) {
  if (!h || !dir || !env_names) return MH_ERR_INVALID;
  for (int k = 0; k < h->m; ++k)
    if (!env_names[k]) return MH_ERR_INVALID;
  const int n = mahyco_history_read(h, nullptr, 0);
  if (n < 0) return n;
  std::vector<double> rows((size_t)std::max(n, 1) * (1 + 5 * h->m));
  const int got = mahyco_history_read(h, rows.data(), n);
  if (got < 0) return got;
  if (!mh_output::make_dir(dir)) return MH_ERR_INVALID;
  return mh_output::write_gnuplot(dir, rows, n, h->m, env_names);
}

// Checkpoint file: a header (magic, layout version, the case's mesh, options and
// environments, counts and host-side cycle state) followed by every device allocation of
// the case in allocation order.  Restoring needs a case created with the same description
// (and the same time-history capacity), whose allocations therefore match one for one.
static const char kCkptMagic[8] = {'M', 'H', 'C', 'K', 'P', 'T', '0', '1'};
// field-by-field equality of case descriptions (struct padding is not compared)
static bool same_mesh(const MhMesh& a, const MhMesh& b) {
  bool eq = a.dim == b.dim && a.nx == b.nx && a.ny == b.ny && a.nz == b.nz && a.coord_rule == b.coord_rule;
  for (int d = 0; d < 3; ++d) eq = eq && a.origin[d] == b.origin[d] && a.length[d] == b.length[d];
  return eq;
}
static bool same_env(const MhEnvironment& a, const MhEnvironment& b) {
  bool eq = a.eos == b.eos && a.adiabatic_cst == b.adiabatic_cst && a.limit_tension == b.limit_tension &&
            a.specific_heat == b.specific_heat && a.linear_pseudo_coeff == b.linear_pseudo_coeff &&
            a.quadratic_pseudo_coeff == b.quadratic_pseudo_coeff && a.densite_initiale == b.densite_initiale &&
            a.pression_initiale == b.pression_initiale;
  for (int d = 0; d < 3; ++d) eq = eq && a.vitesse_initiale[d] == b.vitesse_initiale[d];
  return eq;
}
static bool same_options(const MhOptions& a, const MhOptions& b) {
  bool eq = a.cfl == b.cfl && a.threshold == b.threshold && a.deltat_init == b.deltat_init &&
            a.deltat_min == b.deltat_min && a.deltat_max == b.deltat_max && a.final_time == b.final_time &&
            a.with_projection == b.with_projection && a.calcul_vitson_max == b.calcul_vitson_max &&
            a.nb_bc == b.nb_bc && a.ordre_projection == b.ordre_projection &&
            a.projection_pente_borne == b.projection_pente_borne && a.pente_borne_debar_fix == b.pente_borne_debar_fix &&
            a.pente_borne_mixte == b.pente_borne_mixte &&
            a.projection_limiteur_pure_id == b.projection_limiteur_pure_id &&
            a.projection_simple_pente == b.projection_simple_pente && a.with_newton == b.with_newton &&
            a.pression_explicite == b.pression_explicite &&
            a.projection_limiteur_id == b.projection_limiteur_id && a.remap_threshold == b.remap_threshold &&
            a.cas_model == b.cas_model && a.cas_test == b.cas_test && a.interface_position == b.interface_position &&
            a.interface_axis == b.interface_axis && a.sedov_deposit == b.sedov_deposit &&
            a.sans_lagrange == b.sans_lagrange && a.rider_reverse == b.rider_reverse &&
            a.rider_parametre == b.rider_parametre && a.rider_rayon_bulle == b.rider_rayon_bulle &&
            a.rider_densite_lineaire == b.rider_densite_lineaire && a.schema_csts == b.schema_csts &&
            a.longueur_caracteristique == b.longueur_caracteristique;
  for (int i = 0; eq && i < a.nb_bc && i < MH_MAX_BC; ++i)
    eq = a.bc[i].surface == b.bc[i].surface && a.bc[i].type == b.bc[i].type && a.bc[i].value == b.bc[i].value;
  return eq;
}
struct CkptHead {
  char magic[8];
  int version, m;
  MhMesh mesh;
  MhOptions opt;
  MhEnvironment env[MH_MAX_ENV];
  long nc, nn, launched;
  int sens, hist_period, hist_max, max_iter;
  long nallocs;
};

int mahyco_checkpoint_write(MhHandle* h, const char* path) {
  if (!h || !path || h->ext_sc || h->phase != 0) return MH_ERR_INVALID;
  HIPCHK(hipStreamSynchronize(h->stream));
  CkptHead hd;
  memset(&hd, 0, sizeof hd);
  memcpy(hd.magic, kCkptMagic, 8);
  hd.version = 1;
  hd.m = h->m;
  hd.mesh = h->mesh;
  hd.opt = h->opt;
  for (int k = 0; k < h->m; ++k) hd.env[k] = h->env[k];
  hd.nc = h->g.nc;
  hd.nn = h->g.nn;
  hd.launched = h->launched;
  hd.sens = h->sens;
  hd.hist_period = h->dev.hist_period;
  hd.hist_max = h->dev.hist_max;
  hd.max_iter = h->dev.max_iter;
  hd.nallocs = (long)h->allocs.size();
  FILE* f = fopen(path, "wb");
  if (!f) return MH_ERR_INVALID;
  bool ok = fwrite(&hd, sizeof hd, 1, f) == 1;
  std::vector<char> buf;
  for (size_t i = 0; ok && i < h->allocs.size(); ++i) {
    const uint64_t n = h->alloc_bytes[i];
    buf.resize(n);
    ok = hipMemcpy(buf.data(), h->allocs[i], n, hipMemcpyDeviceToHost) == hipSuccess && fwrite(&n, sizeof n, 1, f) == 1 &&
         fwrite(buf.data(), 1, n, f) == n;
  }
  ok = (fclose(f) == 0) && ok;
  return ok ? MH_OK : MH_ERR_INVALID;
}

// Captured graphs hold the kernel arguments (the MhDev copy) of their capture.
static void drop_graphs(MhHandle* h) {
  for (auto& kv : h->graphs) (void)hipGraphExecDestroy(kv.second);
  for (auto gdef : h->graph_defs) (void)hipGraphDestroy(gdef);
  h->graphs.clear();
  h->graph_defs.clear();
}

int mahyco_checkpoint_read(MhHandle* h, const char* path) {
  if (!h || !path || h->ext_sc || h->phase != 0) return MH_ERR_INVALID;
  FILE* f = fopen(path, "rb");
  if (!f) return MH_ERR_INVALID;
  CkptHead hd;
  bool ok = fread(&hd, sizeof hd, 1, f) == 1 && memcmp(hd.magic, kCkptMagic, 8) == 0 && hd.version == 1 &&
            hd.m == h->m && same_mesh(hd.mesh, h->mesh) && same_options(hd.opt, h->opt) && hd.nc == h->g.nc &&
            hd.nn == h->g.nn &&
            hd.hist_period == h->dev.hist_period && hd.hist_max == h->dev.hist_max &&
            hd.nallocs == (long)h->allocs.size();
  for (int k = 0; ok && k < h->m; ++k) ok = same_env(hd.env[k], h->env[k]);
  std::vector<std::vector<char>> data(ok ? h->allocs.size() : 0);
  for (size_t i = 0; ok && i < h->allocs.size(); ++i) {
    uint64_t n = 0;
    ok = fread(&n, sizeof n, 1, f) == 1 && n == h->alloc_bytes[i];
    if (ok) {
      data[i].resize(n);
      ok = fread(data[i].data(), 1, n, f) == n;
    }
  }
  fclose(f);
  if (!ok) return MH_ERR_INVALID;  // nothing of the case has been changed
  HIPCHK(hipStreamSynchronize(h->stream));
  for (size_t i = 0; i < h->allocs.size(); ++i)
    HIPCHK(hipMemcpy(h->allocs[i], data[i].data(), data[i].size(), hipMemcpyHostToDevice));
  h->launched = hd.launched;
  h->sens = hd.sens;
  h->dev.max_iter = hd.max_iter;
  drop_graphs(h);  // they hold the old kernel arguments
  return MH_OK;
}

int mahyco_continue_init(MhHandle* h) {
  if (!h || h->ext_sc || h->phase != 0) return MH_ERR_INVALID;
  h->dev.q_frozen = (h->dev.with_proj && 7 * h->m >= 3 * h->m + 5) ? 1 : 0;
  drop_graphs(h);
  return MH_OK;
}

int mahyco_profile_enable(MhHandle* h, int on) {
  if (!h) return MH_ERR_INVALID;
  h->prof = on != 0;
  return MH_OK;
}

int mahyco_profile_read(MhHandle* h, char* names, int names_len, double* total_ms, int* counts, int max) {
  if (!h) return MH_ERR_INVALID;
  HIPCHK(hipStreamSynchronize(h->stream));
  for (auto& r : h->prof_pending) {
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, r.a, r.b);
    h->prof_ms[r.kind] += ms;
    h->prof_cnt[r.kind] += 1;
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
  }
  h->prof_pending.clear();
  int pos = 0;
  for (int k = 0; k < KN_COUNT && k < max; ++k) {
    if (total_ms) total_ms[k] = h->prof_ms[k];
    if (counts) counts[k] = h->prof_cnt[k];
    if (names) {
      const int len = (int)strlen(kKernelNames[k]) + 1;
      if (pos + len <= names_len) {
        memcpy(names + pos, kKernelNames[k], len);
        pos += len;
      }
    }
  }
  return KN_COUNT;
}

}  // extern "C"
