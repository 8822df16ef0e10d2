// mh_init.h -- initial state of a case, formed on the device, one thread per cell: the init
// entry points HydroStartInit (MahycoModule.cc:21-210, MahycoAnnexe.cc:11-142), the case
// services' initMat / initVar (SODService.cc, SEDOVService.cc, POINTTRIPLEService.cc,
// RIDERService.cc, ADVECTIONService.cc), IEquationOfState::initEOS and ComputeCellMass
// (:221-237).  Every quantity of a cell depends on that cell alone, so each thread runs the
// reference's loops for its cell in their order.  The node velocities (RIDER / ADVECTION
// fields use glibc's sincos and atan2, which the reference calls) and the time-reversal table
// are formed on the host.
#ifndef MH_INIT_H
#define MH_INIT_H

#include "mh_geom.h"
#include "mh_state.h"

struct MhInitArgs {
  MhGrid g;
  int m;
  MhOptions opt;
  MhEnvironment env[MH_MAX_ENV];
  const double *ax, *ay, *az;  // node coordinates of the generator along each axis
  double *VE, *rho, *p, *c, *q, *M, *e, *fg, *chig;
  double *fk, *chik, *rhok, *ek, *pk, *Mk;  // [m][nc]
  uint8_t* mask;
  int* bad;  // set when a cell has no environment among the case's m
};

// The cell's smallest and largest node distance from the bubble centre (0.2, 0.2) or
// (0.5, 0.75), in the plane (RIDERService.cc:33-41).
__device__ inline void mh_rider_radii(const D3* x, int nnode, bool low_centre, double& rmin, double& rmax) {
  const double xb = low_centre ? 0.20 : 0.50, yb = low_centre ? 0.20 : 0.75;
  rmin = 10.;
  rmax = 0.;
  for (int l = 0; l < nnode; ++l) {
    const double rnode = sqrt((x[l].x - xb) * (x[l].x - xb) + (x[l].y - yb) * (x[l].y - yb));
    rmin = (rnode < rmin) ? rnode : rmin;  // std::min(rmin, rnode)
    rmax = (rmax < rnode) ? rnode : rmax;  // std::max(rmax, rnode)
  }
}

template <int DIM>
__global__ __launch_bounds__(256) void k_init_cells(MhInitArgs P) {
  const MhGrid& g = P.g;
  const MhOptions& opt = P.opt;
  const int m = P.m;
  const long nc = g.nc;
  for (long c = blockIdx.x * (long)blockDim.x + threadIdx.x; c < nc; c += (long)gridDim.x * blockDim.x) {
    const int i = (int)(c % g.nx);
    const long r0 = c / g.nx;
    const int j = (int)(r0 % g.ny);
    const int k = (int)(r0 / g.ny);
    // the values every cell starts from
    double rho = 0., p = 0., cs = 0., e = 0., fg = 1., chig = 1.;
    double fk[MH_MAX_ENV], chik[MH_MAX_ENV], rhok[MH_MAX_ENV], ek[MH_MAX_ENV], pk[MH_MAX_ENV], Mk[MH_MAX_ENV];
    for (int kk = 0; kk < MH_MAX_ENV; ++kk) {
      fk[kk] = chik[kk] = 1.;
      rhok[kk] = ek[kk] = pk[kk] = Mk[kk] = 0.;
    }
    double rider_frac = -1., f1 = 0.;
    auto node = [&](int a, int b, int cc) { return d3(P.ax[a], P.ay[b], DIM == 3 ? P.az[cc] : 0.); };
    D3 x[8], qv[8];
    x[0] = node(i, j, k);
    x[1] = node(i + 1, j, k);
    x[2] = node(i + 1, j + 1, k);
    x[3] = node(i, j + 1, k);
    if (DIM == 3) {
      x[4] = node(i, j, k + 1);
      x[5] = node(i + 1, j, k + 1);
      x[6] = node(i + 1, j + 1, k + 1);
      x[7] = node(i, j + 1, k + 1);
    }
    constexpr int nnode = DIM == 3 ? 8 : 4;
    const double w = DIM == 3 ? .125 : .25;
    D3 s = d3(0., 0., 0.);
    for (int l = 0; l < nnode; ++l) s = s + x[l];
    const D3 xc = w * s;
    double V;
    if (DIM == 3) {
      cqs_hex(x, qv);
      V = cell_volume<8>(x, qv, 3);
    } else {
      cqs_quad(x, qv);
      V = cell_volume<4>(x, qv, 2);
    }
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
        mh_rider_radii(x, nnode, opt.cas_test < 34, rmin, rmax);
        const double rb = opt.rider_rayon_bulle;
        if (rmax < rb) mat = 1.;
        else if ((rmax >= rb) && (rmin < rb)) mat = (rb - rmin) / (rmax - rmin);
      }
      // initVar (initVarMono :58-158, initVar :163-290)
      const bool mono = opt.cas_test <= 30;
      const bool lin = opt.rider_densite_lineaire != 0;
      const double rb = opt.rider_rayon_bulle;
      double rmin, rmax;
      mh_rider_radii(x, nnode, opt.cas_test < (mono ? 26 : 34), rmin, rmax);
      const double xb = opt.cas_test < (mono ? 26 : 34) ? 0.20 : 0.50;
      const double yb = opt.cas_test < (mono ? 26 : 34) ? 0.20 : 0.75;
      const double croissance = lin ? 1. / rb : 1.;
      rho = P.env[0].densite_initiale;
      p = P.env[0].pression_initiale;
      e = P.env[0].energie_initiale;
      const double r = sqrt((xc.x - xb) * (xc.x - xb) + (xc.y - yb) * (xc.y - yb));
      const double r_cell = lin ? r : rmax;
      if (r_cell < rb) {
        rho = P.env[1].densite_initiale;
        if (lin) {
          const double a = croissance * 2 * (rb - r);
          rho = P.env[1].densite_initiale * ((1. < a) ? 1. : a);  // std::min(a, 1.)
        }
        p = P.env[1].pression_initiale;
      } else if ((rmax >= rb) && (rmin < rb) && !lin) {
        rider_frac = (rb - rmin) / (rmax - rmin);
      }
    } else if (opt.cas_model == MH_CASE_ADVECTION) {  // ADVECTIONService::initMat :11-65
      if (opt.cas_test >= 18) {  // the Mono* tests (7-10) put every cell in environment 0
        const double yb = opt.cas_test < 21 ? 0.50 : 0.75;
        const double rb = 0.25;  // a constant here, the rayon-bulle option in initVar
        if (fabs(xc.x - 0.50) < rb && fabs(xc.y - yb) < rb) mat = 1.;
      }
    } else if (opt.cas_model == MH_CASE_RIEMANN) {
      const int a = opt.interface_axis;
      const D3 lo3 = x[0], hi3 = a == 0 ? x[1] : a == 1 ? x[3] : x[4];
      const double lo = a == 0 ? lo3.x : a == 1 ? lo3.y : lo3.z;
      const double hi = a == 0 ? hi3.x : a == 1 ? hi3.y : hi3.z;
      double fr = (hi - opt.interface_position) / (hi - lo);
      fr = fr < 0. ? 0. : (fr > 1. ? 1. : fr);
      f1 = fr;
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
    if (bits == 0) {
      *P.bad = 1;
      bits = 1;
    }
    // Sedov deposit on the cell at the origin, with the cell volume of the initial geometry
    if (opt.cas_model == MH_CASE_SEDOV) {
      const double rhoInit = P.env[0].densite_initiale, pInit = P.env[0].pression_initiale;
      bool center = false;
      for (int l = 0; l < nnode; ++l) {
        double rn;
        if (DIM == 3)
          rn = sqrt((x[l].x - 0.) * (x[l].x - 0.) + (x[l].y - 0.) * (x[l].y - 0.) + (x[l].z - 0.) * (x[l].z - 0.));
        else
          rn = sqrt((x[l].x - 0.) * (x[l].x - 0.) + (x[l].y - 0.) * (x[l].y - 0.));
        if (rn < 1.e-10) center = true;
      }
      e = 0.244816e-5;
      rho = rhoInit;
      p = pInit;
      if (center) {
        e += opt.sedov_deposit / V;
        p = 0.4 * rhoInit * e;
      }
    }
    // initVar of the other case models (global values; partial values of multi-env cells)
    const bool pure = __popc(bits) == 1;
    if (opt.cas_model == MH_CASE_SOD) {
      const int t = opt.cas_test;
      if (t == 3 || t == 4 || t == 5) {  // SODService::initVarMono :11-40
        const double rr = t == 3 ? xc.x : t == 4 ? xc.y : xc.z;
        if (rr < 0.5) {
          p = P.env[0].pression_initiale;
          rho = P.env[0].densite_initiale;
        } else {
          p = P.env[0].pression_initiale / 10.;
          rho = P.env[0].densite_initiale / 8.;
        }
      } else {  // SODService::initVar :71-121
        const int sn = bits == 1 ? 0 : 1;
        rho = P.env[sn].densite_initiale;
        p = P.env[sn].pression_initiale;
        if (!pure)
          for (int kk = 0; kk < m; ++kk)
            if ((bits >> kk) & 1) {
              rhok[kk] = P.env[sn].densite_initiale;
              pk[kk] = P.env[sn].pression_initiale;
            }
      }
    } else if (opt.cas_model == MH_CASE_POINTTRIPLE) {  // POINTTRIPLEService::initVar :64-128
      const int sn = xc.x <= 0.01 ? 0 : (xc.y <= 0.015 ? 1 : 2);
      rho = P.env[sn].densite_initiale;
      p = P.env[sn].pression_initiale;
      if (!pure && ((bits >> sn) & 1)) {
        rhok[sn] = P.env[sn].densite_initiale;
        pk[sn] = P.env[sn].pression_initiale;
      }
    } else if (opt.cas_model == MH_CASE_RIDER) {
      const double frac_b = rider_frac;
      if (frac_b >= 0.) {  // environment 0 takes the bubble density and fraction (RIDERService.cc:252-266)
        if (opt.cas_test >= 31)
          for (int kk = 0; kk < m && kk < 2; ++kk) {
            if (!((bits >> kk) & 1)) continue;
            const double rk = P.env[kk == 0 ? 1 : 0].densite_initiale;
            const double fr = kk == 0 ? frac_b : 1. - frac_b;
            if (pure) {  // the environment value is the cell value
              rho = rk;
              fg = fr;
              chig = fr;
              p = 0.;
            } else {
              rhok[kk] = rk;
              fk[kk] = fr;
              chik[kk] = fr;
              pk[kk] = 0.;
            }
          }
        rho = frac_b * P.env[1].densite_initiale;
        p = P.env[1].pression_initiale;
      }
    } else if (opt.cas_model == MH_CASE_ADVECTION) {  // ADVECTIONService::initVarMono :67-145, initVar :146-246
      // a square of half side rayon-bulle, no mixed cells, internal energy left at 0
      const bool lin = opt.rider_densite_lineaire != 0;
      const double rb = opt.rider_rayon_bulle;
      const double xb = 0.50, yb = opt.cas_test < (opt.cas_test <= 10 ? 10 : 21) ? 0.50 : 0.75;
      const double croissance = lin ? 1. / rb : 1.;
      rho = P.env[0].densite_initiale;
      p = P.env[0].pression_initiale;
      const double r = sqrt((xc.x - xb) * (xc.x - xb) + (xc.y - yb) * (xc.y - yb));
      if (fabs(xc.x - xb) < rb && fabs(xc.y - yb) < rb) {
        rho = P.env[1].densite_initiale;
        if (lin) rho = P.env[1].densite_initiale * croissance * r;
        p = P.env[1].pression_initiale;
      }
    } else if (opt.cas_model == MH_CASE_RIEMANN) {
      if (pure) {
        const int kk = bits == 1 ? 0 : 1;
        rho = P.env[kk].densite_initiale;
        p = P.env[kk].pression_initiale;
      } else {
        fk[0] = 1.0 - f1;
        fk[1] = f1;
        double ft = 0.;
        for (int kk = 0; kk < 2; ++kk) ft += fk[kk] * P.env[kk].densite_initiale;
        for (int kk = 0; kk < 2; ++kk) {
          rhok[kk] = P.env[kk].densite_initiale;
          pk[kk] = P.env[kk].pression_initiale;
          chik[kk] = fk[kk] * P.env[kk].densite_initiale / ft;
        }
        rho = 0.;
        p = 0.;
        cs = 0.;
      }
    }
    // initEOS per environment (PerfectGasEOSService.cc:12-37, StiffenedGasEOSService.cc:14-33),
    // not in advection mode (MahycoModule.cc:143)
    if (!opt.sans_lagrange) {
      for (int kk = 0; kk < m; ++kk) {
        if (!((bits >> kk) & 1)) continue;
        const double gamma = P.env[kk].adiabatic_cst;
        double& pp = pure ? p : pk[kk];
        double& rr = pure ? rho : rhok[kk];
        double& ee = pure ? e : ek[kk];
        double ck;
        if (P.env[kk].eos == MH_EOS_STIFFENED_GAS) {
          const double pi = P.env[kk].limit_tension;
          ee = (pp + (gamma * pi)) / ((gamma - 1.) * rr);
          ck = sqrt((gamma / rr) * (pp + pi));
        } else {
          ee = pp / ((gamma - 1.) * rr);
          ck = sqrt(gamma * pp / rr);
        }
        if (pure) {
          cs = ck;
        } else {
          // mixed-cell averages, MahycoModule.cc:154-170 (density and pressure accumulate onto
          // the values left by initVar)
          if (kk == __ffs((int)bits) - 1) e = 0.;
          e += ee * chik[kk];
          cs = (ck < cs) ? cs : ck;
          rho += rr * fk[kk];
          p += pp * fk[kk];
        }
      }
    }
    // ComputeCellMass (:221-237); m_euler_volume = V (:174-178)
    const double Mc = rho * V;
    if (!pure)
      for (int kk = 0; kk < m; ++kk)
        if ((bits >> kk) & 1) Mk[kk] = chik[kk] * Mc;
    P.VE[c] = V;
    P.rho[c] = rho;
    P.p[c] = p;
    P.c[c] = cs;
    P.q[c] = 0.;
    P.M[c] = Mc;
    P.e[c] = e;
    P.fg[c] = fg;
    P.chig[c] = chig;
    P.mask[c] = bits;
    for (int kk = 0; kk < m; ++kk) {
      const long o = kk * nc + c;
      P.fk[o] = fk[kk];
      P.chik[o] = chik[kk];
      P.rhok[o] = rhok[kk];
      P.ek[o] = ek[kk];
      P.pk[o] = pk[kk];
      P.Mk[o] = Mk[kk];
    }
  }
}

#endif
