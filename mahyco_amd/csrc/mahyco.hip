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
#include <mutex>
#include <string>
#include <vector>

#include "../../include/mahyco.h"
#include "mh_init.h"
#include "mh_host.h"
#include "mh_kernels.h"
#include "mh_output.h"

namespace {

// profiled kernel groups (sweeps separately per direction: they run different code)
const char* kKernelNames[] = {"k_node_lag",     "k_cell_lag",     "k_sweep_cell_x",  "k_sweep_cell_y",
                              "k_sweep_cell_z", "k_sweep_node_x", "k_sweep_node_y",  "k_sweep_node_z",
                              "k_node_velocity", "k_finalize",    "k_dt",            "k_hcell",
                              "k_node_fwd",     "k_sweep_ke",     "k_cons_ec",      "k_prepare_remap"};

}  // namespace



namespace {


template <class T>
int dalloc(MhHandle* h, T** p, size_t n) {
  if (n == 0) n = 1;
  const size_t bytes = n * sizeof(T);
  if (h->planning) {  // a piece of the case's block, handed out by commit_arena
    h->plan.emplace_back((void**)p, bytes);
    *p = nullptr;
    return MH_OK;
  }
  void* q = nullptr;
  // a case of an ensemble allocates in the order of the ensemble's stream, so that admitting
  // or retiring a case does not wait for the work in flight
  hipError_t e = h->async_alloc ? hipMallocAsync(&q, bytes, h->stream) : hipMalloc(&q, bytes);
  if (e != hipSuccess) {
    fprintf(stderr, "mahyco: hipMalloc(%zu bytes) failed: %s\n", bytes, hipGetErrorString(e));
    return MH_ERR_NOMEM;
  }
  h->owned.push_back(q);
  h->allocs.push_back(q);
  h->alloc_bytes.push_back(bytes);
  h->bytes += bytes;
  *p = (T*)q;
  return MH_OK;
}

// One device block for every array dalloc recorded since `planning` was set: one allocation
// (and one release) per case instead of one per array.
int commit_arena(MhHandle* h) {
  h->planning = false;
  auto round = [](size_t b) { return (b + 255) & ~(size_t)255; };
  size_t total = 0;
  for (auto& s : h->plan) total += round(s.second);
  void* base = nullptr;
  hipError_t e = h->async_alloc ? hipMallocAsync(&base, total, h->astream) : hipMalloc(&base, total);
  if (e != hipSuccess) {
    fprintf(stderr, "mahyco: hipMalloc(%zu bytes) failed: %s\n", total, hipGetErrorString(e));
    h->plan.clear();
    return MH_ERR_NOMEM;
  }
  h->owned.push_back(base);
  h->bytes += total;
  size_t off = 0;
  for (auto& s : h->plan) {
    void* q = (char*)base + off;
    *s.first = q;
    h->allocs.push_back(q);
    h->alloc_bytes.push_back(s.second);
    off += round(s.second);
  }
  h->plan.clear();
  return MH_OK;
}

template <class T>
__global__ void k_fill_value(T* __restrict__ d, long n, T v) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) d[i] = v;
}
template <class T>
__global__ void k_scatter_values(T* __restrict__ d, const long* __restrict__ idx, const T* __restrict__ val, long k) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < k; i += (long)gridDim.x * blockDim.x) d[idx[i]] = val[i];
}
// both copies of the initial node coordinates from the generator's axis coordinates
__global__ void k_init_coords(MhGrid g, const double* __restrict__ ax, const double* __restrict__ ay,
                              const double* __restrict__ az, double* x0, double* y0, double* z0, double* x1,
                              double* y1, double* z1) {
  for (long n = blockIdx.x * (long)blockDim.x + threadIdx.x; n < g.nn; n += (long)gridDim.x * blockDim.x) {
    const int i = (int)(n % g.nnx);
    const long r = n / g.nnx;
    const int j = (int)(r % g.nny);
    const int k = (int)(r / g.nny);
    x0[n] = x1[n] = ax[i];
    y0[n] = y1[n] = ay[j];
    z0[n] = z1[n] = g.dim == 3 ? az[k] : 0.;
  }
}
__global__ void k_set_scalars(MhScalars* d, MhScalars v) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *d = v;
}
inline unsigned fill_blocks(long n) { return (unsigned)std::min<long>((n + MH_BLOCK - 1) / MH_BLOCK, 4096); }

// The initial uploads of a case, on h->astream.  An array that holds one value everywhere but
// in a few places (zeros and ones outside mixed cells, a uniform velocity) is filled on the
// device and its exceptions scattered, instead of crossing PCIe whole.  Host buffers handed
// to the copies are kept until the stream is synchronised.
struct Uploader {
  MhHandle* h;
  std::vector<std::vector<char>> keep;
  std::vector<void*> temp;  // device buffers of scattered exceptions, released after use
  template <class T>
  int operator()(T* d, const std::vector<T>& v) {
    const size_t n = v.size();
    if (n == 0) return MH_OK;
    if (n >= 4096) {
      const T v0 = v[0];
      std::vector<long> idx;
      std::vector<T> val;
      bool dense = false;
      for (size_t i = 1; i < n; ++i)
        if (memcmp(&v[i], &v0, sizeof(T)) != 0) {
          if (idx.size() >= n / 8) {
            dense = true;
            break;
          }
          idx.push_back((long)i);
          val.push_back(v[i]);
        }
      if (!dense) {
        hipLaunchKernelGGL(k_fill_value<T>, dim3(fill_blocks((long)n)), dim3(MH_BLOCK), 0, h->astream, d, (long)n, v0);
        const long k = (long)idx.size();
        if (k > 0) {
          std::vector<char> buf(k * sizeof(long) + k * sizeof(T));
          memcpy(buf.data(), idx.data(), k * sizeof(long));
          memcpy(buf.data() + k * sizeof(long), val.data(), k * sizeof(T));
          void* q = nullptr;
          HIPCHK(hipMallocAsync(&q, buf.size(), h->astream));
          temp.push_back(q);
          HIPCHK(hipMemcpyAsync(q, buf.data(), buf.size(), hipMemcpyHostToDevice, h->astream));
          hipLaunchKernelGGL(k_scatter_values<T>, dim3(fill_blocks(k)), dim3(MH_BLOCK), 0, h->astream, d,
                             (const long*)q, (const T*)((char*)q + k * sizeof(long)), k);
          keep.push_back(std::move(buf));
        }
        return hip_check(hipGetLastError(), "fill");
      }
    }
    HIPCHK(hipMemcpyAsync(d, v.data(), n * sizeof(T), hipMemcpyHostToDevice, h->astream));
    return MH_OK;
  }
  int release() {
    for (void* q : temp) HIPCHK(hipFreeAsync(q, h->astream));
    temp.clear();
    return MH_OK;
  }
};

// Node coordinate along one axis of the Cartesian generator.
double axis_coord(const MhMesh& mesh, int d, int i, int n) {
  const double o = mesh.origin[d], l = mesh.length[d];
  if (mesh.coord_rule == MH_COORD_DIV) return o + (l * (double)i) / (double)n;
  return o + (double)i * (l / (double)n);
}

// The host's part of a case's initial state: the generator's axis coordinates, the node
// velocities (the RIDER and ADVECTION fields call glibc's sincos and atan2, as the reference
// does) and the time-reversal table.  The cells are initialised on the device (mh_init.h); the
// node mass is formed by the first node kernel.
struct HostInit {
  std::vector<double> ax, ay, az;  // node coordinates of the generator along each axis
  std::vector<double> u[3];     // node velocities of the RIDER / ADVECTION fields (else empty)
  double u_uniform[3] = {0., 0., 0.};  // the velocity of every node otherwise
  std::vector<double> rev_cos;  // advection mode: cos(pi t_n parametre) per iteration
};

// The reference's Pi of the RIDER velocity fields (TypesMahyco.h:42).
const double kRiderPi = 3.14159265359;

int host_init(MhHandle* h, HostInit& I) {
  const MhGrid& g = h->g;
  const MhOptions& opt = h->opt;
  const int dim = g.dim;
  const long nn = g.nn;
  I.ax.resize(g.nnx);
  I.ay.resize(g.nny);
  I.az.resize(dim == 3 ? g.nnz : 1, 0.);
  for (int i = 0; i < g.nnx; ++i) I.ax[i] = axis_coord(h->mesh, 0, i, g.nx);
  for (int j = 0; j < g.nny; ++j) I.ay[j] = axis_coord(h->mesh, 1, j, g.ny);
  if (dim == 3)
    for (int k = 0; k < g.nnz; ++k) I.az[k] = axis_coord(h->mesh, 2, k, g.nz);
  auto node = [&](int i, int j, int k) { return d3(I.ax[i], I.ay[j], dim == 3 ? I.az[k] : 0.); };
  const double* v0 = h->env[0].vitesse_initiale;
  if (opt.cas_model == MH_CASE_RIDER || opt.cas_model == MH_CASE_ADVECTION) {
    for (int d = 0; d < 3; ++d) I.u[d].assign(nn, 0.);
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
  } else if (opt.cas_model != MH_CASE_SEDOV) {
    for (int d = 0; d < 3; ++d) I.u_uniform[d] = dim == 2 && d == 2 ? 0. : v0[d];
  }
  if (opt.sans_lagrange) {  // cos(pi t parametre) at t_n = t_(n-1) + deltat-init, as the device accumulates it
    const long nmax = (long)std::ceil(opt.final_time / opt.deltat_init) + 8;
    I.rev_cos.resize(nmax);
    double t = 0.;
    for (long it = 0; it < nmax; ++it) {
      t = t + opt.deltat_init;
      I.rev_cos[it] = std::cos(kRiderPi * t * opt.rider_parametre);
    }
  }
  return MH_OK;
}

void set_parity(MhHandle* h) {
  const int par = (int)(h->launched & 1);
  for (int d = 0; d < 3; ++d) {
    h->dev.xlp[d] = h->xl[par][d];
    h->dev.xlc[d] = h->xl[par ^ 1][d];
  }
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

void launch_dt(MhHandle* h) { LAUNCH(h, KN_DT, k_dt<>, 1, MH_DT_BLOCK, h->dev); }

// Phase functions shared by mahyco_cycle and the per-entry-point calls.
void phase_node(MhHandle* h) {
  set_parity(h);
  MH_DISPATCH(launch_node_t, h);
}
void phase_cell(MhHandle* h) { MH_DISPATCH(launch_cell_t, h); }
void phase_remap(MhHandle* h) {
  if (!h->opt.with_projection) return;
  if (h->staged) MH_DISPATCH(launch_prepare_t, h);
  MH_DISPATCH(launch_sweeps_t, h);
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
  // total-energy conservation reads the dual remap's kinetic energy, which advection mode does not run
  if (o->conservation_energie_totale && o->sans_lagrange) return MH_ERR_UNSUPPORTED;
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

static void launch_eos(MhHandle* h, int env) {
  hipLaunchKernelGGL(k_apply_eos<>, dim3(h->nblk_c), dim3(MH_BLOCK), 0, h->stream, h->dev, env);
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
    if (o.conservation_energie_totale) words += 3 * nn + nc;                                          // ke_*, dec
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


namespace {

void ens_drop_graphs(MhEnsembleImpl* E) {
  for (auto& kv : E->graphs) (void)hipGraphExecDestroy(kv.second);
  for (auto g : E->graph_defs) (void)hipGraphDestroy(g);
  E->graphs.clear();
  E->graph_defs.clear();
}

void ens_free_list(EnsList& l) {
  (void)hipFree(l.d);
  (void)hipFree(l.d_n);
  l.d = nullptr;
  l.d_n = nullptr;
  l.cap = 0;
}

void ens_free_groups(MhEnsembleImpl* E) {
  for (auto& g : E->groups) {
    for (EnsList* l : {&g.cmap, &g.nmap, &g.rcmap, &g.rnmap, &g.tmap[0], &g.tmap[1], &g.ntmap, &g.mmap[0], &g.mmap[1]}) ens_free_list(*l);
    (void)hipFree(g.d_list);
    (void)hipFree(g.d_nlist);
  }
  E->groups.clear();
}

// tiles of the z-marching node kernel over one case
int node_tile_blocks(const MhHandle* h) {
  return ((h->g.nnx + MH_NTX - 1) / MH_NTX) * ((h->g.nny + MH_NTY - 1) / MH_NTY) *
         ((h->g.nnz + h->dev.nt_kc - 1) / h->dev.nt_kc);
}

// Launch shapes chosen by the autotuners, per kind of work (process-wide, so that a later
// case or ensemble of the same kind starts with them).
std::mutex g_tune_mu;
std::map<std::string, int> g_tune;  // key -> choice

std::string group_key(const EnsGroup& g) {
  char b[128];
  snprintf(b, sizeof b, "ens:%d,%d,%d,%d,%d,%d,%d,%d,%d,%d,%d", g.dim, g.m, g.o3, g.gen, g.csts, g.adv, g.lc, g.rows,
           g.ntile, g.lim, g.cons);
  return b;
}
int tune_lookup(const std::string& k) {
  std::lock_guard<std::mutex> lock(g_tune_mu);
  auto it = g_tune.find(k);
  return it == g_tune.end() ? -1 : it->second;
}
void tune_store(const std::string& k, int v) {
  std::lock_guard<std::mutex> lock(g_tune_mu);
  g_tune[k] = v;
}

// Kernel variants of a single case as one number v = node + 4 cell + 16 sweep, every one of
// which gives the same bits: node 0 the flat node kernel, 1 / 2 / 3 the z-marching one with 12 /
// 8 / 16 node planes per block; cell 0 the flat cell kernel, 1 / 2 / 3 the LDS-staged one with
// 12 / 8 / 16 cell planes per block; sweep 0 row segments and tiles, 1 one thread per item.
const int kTileDepth[4] = {0, 12, 8, 16};
// target segment lengths of the marching y / z dual sweeps, by variant code bits 5-7
constexpr int kMarchLen[5] = {10, 6, 8, 4, 12};
void set_variant(MhHandle* h, int v) {
  const int node = v & 3, cell = (v >> 2) & 3, mc = (v >> 5) & 7;
  h->nm_len = kMarchLen[mc < 5 ? mc : 0];
  h->dev.nm_len = h->nm_len;
  h->node_tiled = node != 0;
  if (node) h->nt_kc = kTileDepth[node];
  h->cell_tiled = cell != 0;
  if (cell) h->ct_k = kTileDepth[cell];
  h->sweep_mode = (v >> 4) & 1 ? MH_SWEEP_FLAT : MH_SWEEP_TILE;
  h->dev.nt_kc = h->nt_kc;
  h->dev.ct_k = h->ct_k;
}
int tile_code(int depth) { return depth == 12 ? 1 : depth == 16 ? 3 : 2; }
int march_code(int len) {
  for (int c = 0; c < 5; ++c)
    if (kMarchLen[c] == len) return c;
  return 0;
}
int get_variant(const MhHandle* h) {
  return (h->node_tiled ? tile_code(h->nt_kc) : 0) | ((h->cell_tiled ? tile_code(h->ct_k) : 0) << 2) |
         (h->sweep_mode == MH_SWEEP_FLAT ? 16 : 0) | (march_code(h->nm_len) << 5);
}
// the key of a case before its device state exists (dimensions and the options that select kernels)
std::string case_key_dims(const MhHandle* h) {
  char b[160];
  snprintf(b, sizeof b, "case:%d,%d,%d,%d,%d,%d,%d,%d,%d,%d", h->g.dim, h->g.nx, h->g.ny, h->g.nz, h->m,
           h->opt.with_projection, h->opt.ordre_projection, h->opt.projection_limiteur_id,
           h->opt.projection_pente_borne, h->opt.schema_csts);
  return b;
}

// the batch group a case belongs to (created on first use: the launch sequence changes)
int ens_group_of(MhEnsembleImpl* E, MhHandle* h) {
  EnsGroup k;
  k.dim = h->g.dim;
  k.m = h->m;
  k.o3 = h->dev.order2 == 2;
  k.gen = ens_gen(h);
  k.csts = h->dev.csts != 0;
  k.adv = h->dev.sans_lagrange != 0;
  k.lc = h->dev.lc != 0;
  k.rows = !k.o3 && !k.gen && h->sweep_mode == MH_SWEEP_TILE;
  k.ntile = k.dim == 3 && h->node_tiled && !k.adv;
  k.lim = (k.o3 || k.gen) ? -1 : (h->dev.order2 ? h->dev.limiter : 3);
  k.cons = h->dev.dec != nullptr;
  for (size_t i = 0; i < E->groups.size(); ++i) {
    const EnsGroup& g = E->groups[i];
    if (g.dim == k.dim && g.m == k.m && g.o3 == k.o3 && g.gen == k.gen && g.csts == k.csts && g.adv == k.adv &&
        g.lc == k.lc && g.rows == k.rows && g.ntile == k.ntile && g.lim == k.lim && g.cons == k.cons)
      return (int)i;
  }
  const int cached = tune_lookup(group_key(k));
  k.bpc = cached > 0 ? cached : (E->tuned_bpc > 0 ? E->tuned_bpc : 4);
  E->groups.push_back(k);
  E->version++;
  return (int)E->groups.size() - 1;
}

void ens_list_add(EnsList& l, int slot, int nblocks) {
  for (int b = 0; b < nblocks; ++b) l.h.push_back(make_int2(slot, b));
  l.dirty = true;
}
void ens_list_drop(EnsList& l, int slot) {
  size_t w = 0;
  for (size_t r = 0; r < l.h.size(); ++r)
    if (l.h[r].x != slot) l.h[w++] = l.h[r];
  if (w != l.h.size()) l.dirty = true;
  l.h.resize(w);
}

// The blocks of case h (slot sl) in every list of its group.
void ens_lists_add(EnsGroup& g, MhHandle* h, int sl) {
  ens_list_add(g.cmap, sl, h->nblk_c);
  ens_list_add(g.nmap, sl, h->nblk_n);
  if (g.rows) {
    ens_list_add(g.rcmap, sl, row_blocks(h));
    ens_list_add(g.rnmap, sl, node_row_blocks(h));
    for (int D = 1; D < g.dim; ++D) ens_list_add(g.tmap[D - 1], sl, tile_blocks(h, D));
    for (int D = 1; D < g.dim; ++D) ens_list_add(g.mmap[D - 1], sl, node_march_blocks(h, D));
  }
  if (g.ntile) ens_list_add(g.ntmap, sl, node_tile_blocks(h));
  g.slots.push_back(sl);
  g.list_dirty = true;
}
void ens_lists_drop(EnsGroup& g, int sl) {
  for (EnsList* l : {&g.cmap, &g.nmap, &g.rcmap, &g.rnmap, &g.tmap[0], &g.tmap[1], &g.ntmap, &g.mmap[0], &g.mmap[1]}) ens_list_drop(*l, sl);
  g.slots.erase(std::remove(g.slots.begin(), g.slots.end(), sl), g.slots.end());
  g.list_dirty = true;
}

// Uploads what changed (state-table entries of new cases, work lists and their lengths) with
// copies ordered on the ensemble stream, through pinned staging that is reused once the copies
// of the previous update have run.  A list that outgrows its buffer gets a larger one, and the
// captured graphs, which hold the old address, are dropped.
int ens_flush(MhEnsembleImpl* E) {
  size_t need = 0;
  for (int sl = 0; sl < E->max_slots; ++sl)
    if (E->tab_dirty[sl]) need += 2 * sizeof(MhDev);
  for (auto& g : E->groups) {
    for (EnsList* l : {&g.cmap, &g.nmap, &g.rcmap, &g.rnmap, &g.tmap[0], &g.tmap[1], &g.ntmap, &g.mmap[0], &g.mmap[1]})
      if (l->dirty) need += l->h.size() * sizeof(int2) + 64;
    if (g.list_dirty) need += g.slots.size() * sizeof(int) + 64;
  }
  if (need == 0) return MH_OK;
  if (E->stage_pending) {
    HIPCHK(hipEventSynchronize(E->stage_done));
    E->stage_pending = false;
  }
  if (need > E->stage_cap) {
    if (E->stage) (void)hipHostFree(E->stage);
    E->stage = nullptr;
    E->stage_cap = 0;
    size_t cap = 1 << 20;
    while (cap < need) cap *= 2;
    HIPCHK(hipHostMalloc((void**)&E->stage, cap, hipHostMallocDefault));
    E->stage_cap = cap;
  }
  size_t off = 0;
  auto put = [&](void* dst, const void* src, size_t bytes) -> int {
    off = (off + 15) & ~(size_t)15;
    memcpy(E->stage + off, src, bytes);
    HIPCHK(hipMemcpyAsync(dst, E->stage + off, bytes, hipMemcpyHostToDevice, E->stream));
    off += bytes;
    return MH_OK;
  };
  int st;
  for (int sl = 0; sl < E->max_slots; ++sl) {
    if (!E->tab_dirty[sl]) continue;
    MhHandle* h = E->slot[sl];
    for (int p = 0; p < 2; ++p) {
      MhDev d = h->dev;
      for (int k = 0; k < 3; ++k) {
        d.xlp[k] = h->xl[p][k];
        d.xlc[k] = h->xl[p ^ 1][k];
      }
      if ((st = put(E->d_tab[p] + sl, &d, sizeof d)) != MH_OK) return st;
    }
    E->tab_dirty[sl] = 0;
  }
  for (auto& g : E->groups) {
    for (EnsList* l : {&g.cmap, &g.nmap, &g.rcmap, &g.rnmap, &g.tmap[0], &g.tmap[1], &g.ntmap, &g.mmap[0], &g.mmap[1]}) {
      if (!l->dirty) continue;
      const int n = (int)l->h.size();
      if (n > l->cap || !l->d_n) {
        int cap = 1024;
        while (cap < n) cap *= 2;
        ens_free_list(*l);
        HIPCHK(hipMalloc(&l->d, sizeof(int2) * cap));
        HIPCHK(hipMalloc(&l->d_n, sizeof(int)));
        l->cap = cap;
        E->version++;  // graphs hold the old buffer
      }
      if (n > 0 && (st = put(l->d, l->h.data(), sizeof(int2) * n)) != MH_OK) return st;
      if ((st = put(l->d_n, &n, sizeof(int))) != MH_OK) return st;
      l->dirty = false;
    }
    if (g.list_dirty) {
      const int n = (int)g.slots.size();
      if (n > g.list_cap || !g.d_nlist) {
        int cap = 64;
        while (cap < n) cap *= 2;
        (void)hipFree(g.d_list);
        (void)hipFree(g.d_nlist);
        HIPCHK(hipMalloc(&g.d_list, sizeof(int) * cap));
        HIPCHK(hipMalloc(&g.d_nlist, sizeof(int)));
        g.list_cap = cap;
        E->version++;
      }
      if (n > 0 && (st = put(g.d_list, g.slots.data(), sizeof(int) * n)) != MH_OK) return st;
      if ((st = put(g.d_nlist, &n, sizeof(int))) != MH_OK) return st;
      g.list_dirty = false;
    }
  }
  HIPCHK(hipEventRecord(E->stage_done, E->stream));
  E->stage_pending = true;
  return MH_OK;
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

// Activates the pending cases (only at a phase boundary: every active case shares the
// coordinate parity and the sweep rotation of the ensemble's iteration count).
int ens_activate(MhEnsembleImpl* E) {
  if (E->pending.empty() || (E->launched % 6) != 0) return MH_OK;
  for (size_t i = 0; i < E->pending.size(); ++i) {
    const int sl = E->pending_slot[i];
    MhHandle* h = E->pending[i];
    E->slot[sl] = h;
    E->tab_dirty[sl] = 1;
    ens_lists_add(E->groups[ens_group_of(E, h)], h, sl);
  }
  E->pending.clear();
  E->pending_slot.clear();
  return MH_OK;
}

// Runs ncycles iterations of every active case, through 6-iteration graphs when asked (a
// graph is captured once per launch sequence and reused while cases join and leave).
int ens_step(MhEnsembleImpl* E, int ncycles, int use_graph) {
  int done = 0;
  while (done < ncycles) {
    int st = ens_activate(E);
    if (st != MH_OK) return st;
    if ((st = ens_flush(E)) != MH_OK) return st;
    if (use_graph && (E->launched % 6) == 0 && ncycles - done >= 6 && E->pending.empty()) {
      const long key = E->version;
      auto it = E->graphs.find(key);
      if (it == E->graphs.end()) {
        ens_drop_graphs(E);  // older launch sequences are not replayed again
        hipGraph_t gdef = nullptr;
        HIPCHK(hipStreamBeginCapture(E->stream, hipStreamCaptureModeThreadLocal));
        const long l0 = E->launched;
        for (int i = 0; i < 6; ++i) ens_cycle(E);
        const hipError_t ec = hipStreamEndCapture(E->stream, &gdef);
        E->launched = l0;
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
        E->graph_defs.push_back(gdef);
        it = E->graphs.emplace(key, ex).first;
        E->graph_captures++;
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

}  // namespace

// ============================================================================ C ABI
extern "C" {

static int create_case(const MhMesh* mesh, const MhEnvironment* envs, int nb_env, const MhOptions* options,
                       hipStream_t stream, MhScalars* ext_sc, MhHandle** out, hipStream_t copy_stream = nullptr,
                       hipEvent_t copy_done = nullptr);

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
                       hipStream_t stream, MhScalars* ext_sc, MhHandle** out, hipStream_t copy_stream,
                       hipEvent_t copy_done) {
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
  h->async_alloc = ext_sc != nullptr;  // a case of an ensemble
  h->astream = copy_stream ? copy_stream : h->stream;
  {
    const char* nk = getenv("MAHYCO_NODE_KERNEL");
    h->node_tiled = !(nk && strcmp(nk, "flat") == 0);
    const char* ck = getenv("MAHYCO_CELL_KERNEL");
    h->cell_tiled = ck && strcmp(ck, "tiled") == 0;
    const char* sw = getenv("MAHYCO_SWEEP");
    h->sweep_mode = (sw && strcmp(sw, "flat") == 0) ? MH_SWEEP_FLAT : MH_SWEEP_TILE;
    const char* ml = getenv("MAHYCO_MARCH_LEN");  // segment length of the marching dual sweeps
    // 10-node segments where the mesh still fills the GPU with them (about 2.6 M nodes, one round of
    // waves), 4 on smaller meshes, whose sweeps need the threads
    const long nn = (long)h->g.nnx * h->g.nny * (h->g.dim == 3 ? h->g.nnz : 1);
    h->nm_len = ml && atoi(ml) >= 1 ? atoi(ml) : nn >= 10L * 4096 * 64 ? MH_NM_L : 4;
  }
  {
    const int tuned = tune_lookup(case_key_dims(h));
    if (tuned >= 0) set_variant(h, tuned);
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
  int* d_bad;  // set by the initialisation kernel for a cell outside the case's environments
  int ok = MH_OK;
  h->planning = true;  // one device block for the arrays below (commit_arena)
  ok |= dalloc(h, &d_bad, 1);
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
    if (options->conservation_energie_totale) {  // dual kinetic energy, its phi, the energy returned
      ok |= dalloc(h, &S.ke_U, nn);
      ok |= dalloc(h, &S.ke_Pa, nn);
      ok |= dalloc(h, &S.ke_Pb, nn);
      ok |= dalloc(h, &S.dec, nc);
    }
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
  if (ok == MH_OK) ok = commit_arena(h);
  h->planning = false;
  h->plan.clear();
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
  S.nt_kc = h->nt_kc;
  S.ct_k = h->ct_k;
  S.nm_len = h->nm_len;
  S.nblk_n = h->nblk_n;
  S.max_iter = 0;
  MhScalars sc;
  memset(&sc, 0, sizeof(sc));
  sc.t = 0.;
  sc.dt = options->deltat_init;
  sc.dt_old = 0.;
  std::vector<int> err(4, 0);
  ok = MH_OK;
  Uploader upload{h};
  ok |= upload(ax, I.ax);
  ok |= upload(ay, I.ay);
  ok |= upload(az, I.az);
  const std::vector<int> zero1(1, 0);
  ok |= upload(d_bad, zero1);
  {  // every cell's initial state, on the device
    MhInitArgs A;
    A.g = g;
    A.m = nb_env;
    A.opt = *options;
    for (int k = 0; k < MH_MAX_ENV; ++k) A.env[k] = h->env[k];
    A.ax = ax;
    A.ay = ay;
    A.az = az;
    A.VE = VE;
    A.rho = S.rho;
    A.p = S.p;
    A.c = S.c;
    A.q = S.q;
    A.M = S.M;
    A.e = S.e;
    A.fg = S.fg;
    A.chig = S.chig;
    A.fk = S.fk;
    A.chik = S.chik;
    A.rhok = S.rhok;
    A.ek = S.ek;
    A.pk = S.pk;
    A.Mk = S.Mk;
    A.mask = S.mask;
    A.bad = d_bad;
    if (g.dim == 3) hipLaunchKernelGGL(k_init_cells<3>, dim3(fill_blocks(nc)), dim3(256), 0, h->astream, A);
    else hipLaunchKernelGGL(k_init_cells<2>, dim3(fill_blocks(nc)), dim3(256), 0, h->astream, A);
    ok |= hip_check(hipGetLastError(), "k_init_cells");
  }
  for (int d = 0; d < 3; ++d) {
    if (!I.u[d].empty()) ok |= upload(S.u[d], I.u[d]);
    else hipLaunchKernelGGL(k_fill_value<double>, dim3(fill_blocks(nn)), dim3(MH_BLOCK), 0, h->astream, S.u[d], nn,
                            I.u_uniform[d]);
  }
  // initial Lagrangian coordinates = Eulerian grid, both copies, formed on the device
  hipLaunchKernelGGL(k_init_coords, dim3(fill_blocks(nn)), dim3(MH_BLOCK), 0, h->astream, g, ax, ay, az,
                     h->xl[0][0], h->xl[0][1], h->xl[0][2], h->xl[1][0], h->xl[1][1], h->xl[1][2]);
  auto fill = [&](auto* d, long n, auto v) {
    hipLaunchKernelGGL(k_fill_value<decltype(v)>, dim3(fill_blocks(n)), dim3(MH_BLOCK), 0, h->astream, d, n, v);
  };
  if (S.csts) {  // the pseudo-viscosities at n and n-1 start at 0; the saved velocity is written before use
    fill(S.qm1, nc, 0.);
    fill(S.qe, nc, 0.);
    fill(S.qem1, nc, 0.);
    ok |= hip_check(hipMemcpyAsync(S.mkm1, S.mask, (size_t)nc, hipMemcpyDeviceToDevice, h->astream), "copy mask");
    for (int d = 0; d < 3; ++d) fill(S.un[d], nn, 0.);
    if (!options->with_projection) {
      fill(S.qkn, (long)ne, 0.);
      fill(S.qknm1, (long)ne, 0.);
    }
  }
  ok |= upload(S.err, err);
  if (S.est) fill(S.est, nc, (int8_t)0);  // mixed-cell flags: Arcane starts est_mixte and est_pure at 0
  if (h->rev_cos) {
    for (int d = 0; d < 3; ++d) {
      ok |= hip_check(hipMemcpyAsync(h->u0[d], S.u[d], nn * sizeof(double), hipMemcpyDeviceToDevice, h->astream),
                      "copy velocity");
      S.u0[d] = h->u0[d];
    }
    ok |= upload(h->rev_cos, I.rev_cos);
    S.rev_cos = h->rev_cos;
    S.rev_n = (long)I.rev_cos.size();
  }
  if (S.phi0) fill(S.phi0, (long)S.nv * nc, 0.);  // pente-borne phi: Arcane starts m_phi_lagrange at 0
  // the scalars may sit in an ensemble's array, which the iterations' stream writes: set them in
  // that stream's order (a kernel argument, no host buffer)
  hipLaunchKernelGGL(k_set_scalars, dim3(1), dim3(64), 0, h->stream, S.sc, sc);
  ok |= hip_check(hipGetLastError(), "upload scalars");
  ok |= upload.release();
  if (h->astream != h->stream) {  // the case's first launches wait for its uploads, not the host
    ok |= hip_check(hipEventRecord(copy_done, h->astream), "hipEventRecord");
    ok |= hip_check(hipStreamWaitEvent(h->stream, copy_done, 0), "hipStreamWaitEvent");
  }
  int bad = 0;
  ok |= hip_check(hipMemcpyAsync(&bad, d_bad, sizeof(int), hipMemcpyDeviceToHost, h->astream), "read init flag");
  ok |= hip_check(hipStreamSynchronize(h->astream), "hipStreamSynchronize");
  if (ok != MH_OK || bad) {
    mahyco_destroy(h);
    return ok != MH_OK ? MH_ERR_DEVICE : MH_ERR_INVALID;  // a cell in none of the case's environments
  }
  *out = h;
  return MH_OK;
}

void mahyco_destroy(MhHandle* h) {
  if (!h) return;
  // an ensemble's case frees its memory in stream order, behind the launches that read it
  if (h->stream && !h->async_alloc) (void)hipStreamSynchronize(h->stream);
  for (auto& kv : h->graphs) (void)hipGraphExecDestroy(kv.second);
  for (auto gdef : h->graph_defs) (void)hipGraphDestroy(gdef);
  for (auto& r : h->prof_pending) {
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
  }
  for (void* p : h->owned) (void)(h->async_alloc ? hipFreeAsync(p, h->stream) : hipFree(p));
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
  if (!h || !in || h->ext_sc) return MH_ERR_INVALID;
  // between iterations; in staged mode also cell values between the energy update and the remap
  if (h->phase != 0 && !(h->staged && h->phase == 9 && field < 100)) return MH_ERR_INVALID;
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
  launch_eos(h, env);
  return hip_check(hipGetLastError(), "k_apply_eos");
}

int mahyco_apply_one_cell_eos(MhHandle* h, int env, long cell) {
  if (!h || env < 0 || env >= h->m || cell < 0 || cell >= h->g.nc) return MH_ERR_INVALID;
  hipLaunchKernelGGL(k_apply_eos_cell<>, dim3(1), dim3(64), 0, h->stream, h->dev, env, cell);
  return hip_check(hipGetLastError(), "k_apply_eos_cell");
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
  E->tab_dirty.assign(max_slots, 0);
  int dev = 0;
  if (hipGetDevice(&dev) == hipSuccess) {
    int ncu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && ncu > 0) E->ncu = ncu;
  }
  int st = hip_check(hipStreamCreate(&E->stream), "hipStreamCreate");
  if (st == MH_OK) st = hip_check(hipStreamCreate(&E->copy_stream), "hipStreamCreate");
  if (st == MH_OK) st = hip_check(hipEventCreateWithFlags(&E->copy_done, hipEventDisableTiming), "hipEventCreate");
  if (st == MH_OK) st = hip_check(hipEventCreateWithFlags(&E->stage_done, hipEventDisableTiming), "hipEventCreate");
  if (st == MH_OK) {
    // released cases' memory stays in the device pool up to the budget, for the next admissions
    hipMemPool_t pool;
    if (hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess) {
      uint64_t thr = 0;
      (void)hipMemPoolGetAttribute(pool, hipMemPoolAttrReleaseThreshold, &thr);
      uint64_t want = (uint64_t)std::min(hbm_budget_bytes, 1.8e19);
      if (want > thr) (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &want);
    }
  }
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
  if (E->stream) (void)hipStreamSynchronize(E->stream);  // the cases' stream-ordered frees
  (void)hipFree(E->d_sc);
  (void)hipFree(E->d_tab[0]);
  (void)hipFree(E->d_tab[1]);
  if (E->stage) (void)hipHostFree(E->stage);
  if (E->stage_done) (void)hipEventDestroy(E->stage_done);
  if (E->snap) (void)hipHostFree(E->snap);
  if (E->snap_done) (void)hipEventDestroy(E->snap_done);
  if (E->copy_stream) {
    (void)hipStreamSynchronize(E->copy_stream);
    (void)hipStreamDestroy(E->copy_stream);
  }
  if (E->copy_done) (void)hipEventDestroy(E->copy_done);
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
  int st = create_case(mesh, envs, nb_env, options, E->stream, E->d_sc + sl, &h, E->copy_stream, E->copy_done);
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
  // out of its group's work lists for the launches enqueued from now on; its memory is
  // released in stream order behind the launches already enqueued
  for (auto& g : E->groups)
    if (std::find(g.slots.begin(), g.slots.end(), slot) != g.slots.end()) ens_lists_drop(g, slot);
  int st = ens_flush(E);
  E->used -= E->slot[slot]->bytes;
  mahyco_destroy(E->slot[slot]);
  E->slot[slot] = nullptr;
  return st;
}

int mahyco_ensemble_step(MhEnsemble* e, int ncycles, int use_graph) {
  MhEnsembleImpl* E = (MhEnsembleImpl*)e;
  if (!E || ncycles < 0) return MH_ERR_INVALID;
  return ens_step(E, ncycles, use_graph);
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
// Enqueues a copy of every slot's scalars behind the work already queued, into pinned host
// memory, without waiting: the host can queue the next iterations before reading the result
// with mahyco_ensemble_poll_result.
int mahyco_ensemble_poll_async(MhEnsemble* e) {
  MhEnsembleImpl* E = (MhEnsembleImpl*)e;
  if (!E) return MH_ERR_INVALID;
  if (!E->snap) {
    HIPCHK(hipHostMalloc((void**)&E->snap, sizeof(MhScalars) * E->max_slots, hipHostMallocDefault));
    HIPCHK(hipEventCreateWithFlags(&E->snap_done, hipEventDisableTiming));
  }
  if (E->snap_pending) HIPCHK(hipEventSynchronize(E->snap_done));  // the buffer is reused
  E->snap_state.assign(E->max_slots, -100);
  for (int i = 0; i < E->max_slots; ++i)
    if (E->slot[i]) E->snap_state[i] = 1;
  for (int q : E->pending_slot) E->snap_state[q] = 2;
  HIPCHK(hipMemcpyAsync(E->snap, E->d_sc, sizeof(MhScalars) * E->max_slots, hipMemcpyDeviceToHost, E->stream));
  HIPCHK(hipEventRecord(E->snap_done, E->stream));
  E->snap_pending = true;
  return MH_OK;
}

// Waits for the snapshot of the last mahyco_ensemble_poll_async (not for later work) and fills
// status, time and iteration as mahyco_ensemble_poll does; returns the number of active slots.
int mahyco_ensemble_poll_result(MhEnsemble* e, int* status, double* time, int* iteration) {
  MhEnsembleImpl* E = (MhEnsembleImpl*)e;
  if (!E || !E->snap_pending) return MH_ERR_INVALID;
  HIPCHK(hipEventSynchronize(E->snap_done));
  int active = 0;
  for (int i = 0; i < E->max_slots; ++i) {
    const MhScalars& sc = E->snap[i];
    const int state = E->snap_state[i];
    int stv = -100;
    if (state == 2) stv = 2;
    else if (state == 1) stv = sc.status != 0 ? sc.status : (sc.stopped ? MH_STOPPED : MH_OK);
    if (state == 1) active++;
    if (status) status[i] = stv;
    if (time) time[i] = state == 1 ? sc.t : 0.;
    if (iteration) iteration[i] = state == 1 ? sc.iteration : 0;
  }
  return active;
}

long mahyco_ensemble_graph_captures(MhEnsemble* e) {
  MhEnsembleImpl* E = (MhEnsembleImpl*)e;
  return E ? E->graph_captures : -1;
}

// Times `cycles` iterations (hipGraph replays, after one untimed graph) of the ensemble's own
// cases with each candidate persistent grid (2, 4, 8, 16 blocks per CU) and keeps the
// fastest for every group, also for later groups of the same kind.  The iterations are the
// cases' own: the results do not depend on the launch shape.
int mahyco_ensemble_autotune(MhEnsemble* e, int cycles, int* bpc_out, double* ms_per_cycle_out) {
  MhEnsembleImpl* E = (MhEnsembleImpl*)e;
  if (!E || cycles < 1) return MH_ERR_INVALID;
  cycles = (cycles + 5) / 6 * 6;
  const int cand[] = {2, 4, 8, 16};
  int best = 0;
  float best_ms = 0.f;
  hipEvent_t a, b;
  HIPCHK(hipEventCreate(&a));
  HIPCHK(hipEventCreate(&b));
  int st = MH_OK;
  for (int c : cand) {
    for (auto& g : E->groups) g.bpc = c;
    E->version++;
    if ((st = ens_step(E, 6 + (int)((6 - E->launched % 6) % 6), 1)) != MH_OK) break;  // capture
    if ((st = hip_check(hipEventRecord(a, E->stream), "hipEventRecord")) != MH_OK) break;
    if ((st = ens_step(E, cycles, 1)) != MH_OK) break;
    if ((st = hip_check(hipEventRecord(b, E->stream), "hipEventRecord")) != MH_OK) break;
    if ((st = hip_check(hipEventSynchronize(b), "hipEventSynchronize")) != MH_OK) break;
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, a, b);
    if (best == 0 || ms < best_ms) {
      best = c;
      best_ms = ms;
    }
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  if (st != MH_OK) return st;
  for (auto& g : E->groups) {
    g.bpc = best;
    tune_store(group_key(g), best);
  }
  E->tuned_bpc = best;
  E->version++;
  if (bpc_out) *bpc_out = best;
  if (ms_per_cycle_out) *ms_per_cycle_out = best_ms / cycles;
  return MH_OK;
}

// Times `cycles` iterations (hipGraph replays of 6, after one untimed graph) of the case
// itself with each kernel variant (tiled or flat node kernel, tiled or flat cell kernel,
// row/tile or one-thread-per-item sweeps; only the sweep variants in 2D) and keeps the
// fastest, also for later cases of the same dimensions and options.  The iterations are the
// case's own: every variant gives the same bits.  Stops at the case's final time or an error.
int mahyco_autotune(MhHandle* h, int cycles, int* variant_out, double* ms_per_cycle_out) {
  if (!h || cycles < 1 || h->phase != 0 || h->ext_sc) return MH_ERR_INVALID;
  cycles = (cycles + 5) / 6 * 6;
  const int keep = get_variant(h);
  int best = -1;
  float best_ms = 0.f;
  hipEvent_t a, b;
  HIPCHK(hipEventCreate(&a));
  HIPCHK(hipEventCreate(&b));
  int st = MH_OK;
  bool stopped = false;
  auto timed = [&](int v) -> int {  // the variant's time per `cycles` iterations, best kept
    set_variant(h, v);
    drop_graphs(h);
    int s = mahyco_cycle_graph(h, 6, 6);  // capture and first replay, untimed
    if (s == MH_OK) s = hip_check(hipEventRecord(a, h->stream), "hipEventRecord");
    if (s == MH_OK) s = mahyco_cycle_graph(h, cycles, 6);
    if (s == MH_OK) s = hip_check(hipEventRecord(b, h->stream), "hipEventRecord");
    if (s == MH_OK) s = hip_check(hipEventSynchronize(b), "hipEventSynchronize");
    if (s != MH_OK) return s;
    if (mahyco_status(h) != MH_OK) {  // stopped or failed: the timings would not be comparable
      stopped = true;
      return MH_OK;
    }
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, a, b);
    if (best < 0 || ms < best_ms) {
      best = v;
      best_ms = ms;
    }
    return MH_OK;
  };
  // one coordinate at a time: the sweep variant, then (3D) the node kernel and its tile depth,
  // then the cell kernel and its tile depth, each with the best of the others so far
  const int dflt = h->g.dim == 3 ? 2 : 0;  // tiled node kernel with 8 planes, flat cell kernel, tiled sweeps
  for (int sw = 0; sw < 2 && st == MH_OK && !stopped; ++sw) st = timed((dflt & 15) | (sw << 4));
  if (h->g.dim == 3 && h->opt.with_projection) {
    for (int node = 0; node < 4 && st == MH_OK && !stopped; ++node) {
      const int v = (best & ~3) | node;
      if (v != best) st = timed(v);
    }
    for (int cell = 0; cell < 4 && st == MH_OK && !stopped; ++cell) {
      const int v = (best & ~12) | (cell << 2);
      if (v != best) st = timed(v);
    }
  }
  // the segment length of the marching y / z dual sweeps, where they run (tiled sweeps, dual
  // phase, order <= 2)
  if (h->opt.with_projection && !h->opt.sans_lagrange && h->dev.order2 != 2 && !(best & 16)) {
    for (int mc = 0; mc < 5 && st == MH_OK && !stopped; ++mc) {
      const int v = (best & ~(7 << 5)) | (mc << 5);
      if (v != best) st = timed(v);
    }
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  if (st != MH_OK || best < 0) {
    set_variant(h, keep);
    drop_graphs(h);
    return st != MH_OK ? st : mahyco_status(h);
  }
  set_variant(h, best);
  drop_graphs(h);
  tune_store(case_key_dims(h), best);
  if (variant_out) *variant_out = best;
  if (ms_per_cycle_out) *ms_per_cycle_out = best_ms / cycles;
  return MH_OK;
}

// The kernel-variant code of a case, as mahyco_autotune returns it, and its application: a
// caller can keep a tuned choice and apply it to later cases of the same shape.  Every variant
// gives the same bits; setting one drops the captured graphs.
int mahyco_get_variant(const MhHandle* h) { return h ? get_variant(h) : MH_ERR_INVALID; }

// Staged Lagrange step.  The reference stores the state of UpdateEnergyAndPressure (density,
// energy, pressure, sound speed, pseudo-viscosity; MahycoModule.cc:1430-1480, applyEOS and
// computePressionMoyenne after it) and forms the remap vector from that state at the remap
// (computeVariablesForRemap, PrepareRemap.cc:100-155), so a module can change the state in
// between.  The fused cell kernel writes the remap vector directly; staged mode writes the
// state instead and k_prepare_remap forms U at mahyco_remap.  Same bits when nothing changes.
int mahyco_set_staged(MhHandle* h, int on) {
  if (!h || h->phase != 0 || h->ext_sc) return MH_ERR_INVALID;
  if (on) {
    if (!h->opt.with_projection || h->opt.schema_csts || h->opt.sans_lagrange) return MH_ERR_UNSUPPORTED;
    for (int k = 0; k < h->m; ++k)
      if (h->dev.env[k].fictif) return MH_ERR_UNSUPPORTED;  // its EOS follows the remap's time only
  }
  h->staged = on != 0;
  h->dev.staged = h->staged ? 1 : 0;
  drop_graphs(h);  // they hold the other cell kernel
  return MH_OK;
}
int mahyco_set_variant(MhHandle* h, int variant) {
  if (!h || variant < 0 || (variant >> 5) > 4 || h->phase != 0) return MH_ERR_INVALID;
  set_variant(h, variant);
  drop_graphs(h);
  return MH_OK;
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
