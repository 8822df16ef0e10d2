// mh_noalloc.hip -- ComputeDeltaT on caller-owned device arrays with no device allocation:
// for callers that cannot allocate (code under HIP graph capture, serving loops).  The
// per-block CFL minima live in a file-scope device array, `scratch`, whose address is taken
// once per device with hipGetSymbolAddress; every call on a device shares that device's
// array, so calls that overlap on one GPU race.
#include <hip/hip_runtime.h>

#include <mutex>

#include "mh_host.h"
#include "mh_kernels.h"

// largest mesh: MH_NOALLOC_MAX_BLOCKS blocks of MH_BLOCK cells (16.7 M cells)
#define MH_NOALLOC_MAX_BLOCKS 65536

static __device__ __align__(256) double scratch[2 * MH_NOALLOC_MAX_BLOCKS];  // minimum, NaN flag per block

namespace {

constexpr int kMaxDevices = 64;
double* g_scratch[kMaxDevices] = {nullptr};  // device address of `scratch`, per device
std::mutex g_scratch_mu;

// The device address of `scratch` on the current device (the host-side name is a stand-in,
// not a device address).
int scratch_ptr(double** out) {
  int dev = -1;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return MH_ERR_DEVICE;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch[dev] = (double*)p;
  }
  *out = g_scratch[dev];
  return MH_OK;
}

// The CFL candidate of every cell, L / (c + max over the cell's nodes of |u|), in the node
// order and with the formula of cell_vmax (MahycoModule.cc:1767-1790), folded per block of
// MH_BLOCK cells into partial[2b], partial[2b+1] as the cycle's kernels do (block_min_store).
template <int DIM>
__global__ __launch_bounds__(MH_BLOCK) void k_cfl_cells(MhGrid g, const double* __restrict__ L,
                                                        const double* __restrict__ cs,
                                                        const double* __restrict__ u0,
                                                        const double* __restrict__ u1,
                                                        const double* __restrict__ u2,
                                                        double* __restrict__ partial) {
  const int lb = (int)blockIdx.x;
  const long c = lb * (long)MH_BLOCK + threadIdx.x;
  const bool active = c < g.nc;
  double dxs = 1.79769313486231570815e+308;
  if (active) {
    const int i = (int)(c % g.nx);
    const long r = c / g.nx;
    const int j = (int)(r % g.ny);
    const int k = (int)(r / g.ny);
    const long nnx = g.nnx, nnxy = (long)g.nnx * g.nny;
    const long b = i + nnx * (j + (long)g.nny * k);
    auto speed = [&](long n) {
      const double ux = u0[n], uy = u1[n];
      const double uz = DIM == 3 ? u2[n] : 0.;
      return sqrt(ux * ux + uy * uy + uz * uz);
    };
    double v = 0.;
    v = amax(speed(b), v);
    v = amax(speed(b + 1), v);
    v = amax(speed(b + 1 + nnx), v);
    v = amax(speed(b + nnx), v);
    if (DIM == 3) {
      v = amax(speed(b + nnxy), v);
      v = amax(speed(b + 1 + nnxy), v);
      v = amax(speed(b + 1 + nnx + nnxy), v);
      v = amax(speed(b + nnx + nnxy), v);
    }
    dxs = L[c] / (cs[c] + v);
  }
  block_min_store(dxs, active, partial, lb);
}

// The step from the block minima with the rules of computeDeltaT (MahycoModule.cc:1792-1833):
// cfl times the in-order minimum, capped by 1.05 old_deltat and deltat_max; MH_ERR_DT_MIN
// below deltat_min.  One block.
template <int U = 0>
__global__ __launch_bounds__(MH_DT_BLOCK) void k_dt_noalloc(const double* __restrict__ partial, int nblocks,
                                                            double cfl, double old_deltat, double dtmax,
                                                            double dtmin, double* dt_out, int* status_out) {
  double mn;
  if (!cfl_fold(partial, nblocks, &mn)) return;
  double new_dt = cfl * mn;
  new_dt = amin(new_dt, 1.05 * old_deltat);
  new_dt = amin(new_dt, dtmax);
  *dt_out = new_dt;
  *status_out = new_dt < dtmin ? MH_ERR_DT_MIN : MH_OK;
}

}  // namespace

extern "C" {

int mahyco_compute_delta_t_noalloc(const MhMesh* mesh, const MhOptions* options, double old_deltat,
                                   const double* L, const double* c, const double* u0, const double* u1,
                                   const double* u2, double* dt_out, int* status_out, void* stream) {
  if (!mesh || !options || !L || !c || !u0 || !u1 || !dt_out || !status_out) return MH_ERR_INVALID;
  if (mesh->dim != 2 && mesh->dim != 3) return MH_ERR_INVALID;
  if (mesh->dim == 3 && !u2) return MH_ERR_INVALID;
  if (mesh->nx < 1 || mesh->ny < 1 || (mesh->dim == 3 && mesh->nz < 1)) return MH_ERR_INVALID;
  MhGrid g;
  g.dim = mesh->dim;
  g.nx = mesh->nx;
  g.ny = mesh->ny;
  g.nz = g.dim == 3 ? mesh->nz : 1;
  g.nnx = g.nx + 1;
  g.nny = g.ny + 1;
  g.nnz = g.dim == 3 ? g.nz + 1 : 1;
  g.nc = (long)g.nx * g.ny * g.nz;
  g.nn = (long)g.nnx * g.nny * g.nnz;
  const long nblocks = (g.nc + MH_BLOCK - 1) / MH_BLOCK;
  if (nblocks > MH_NOALLOC_MAX_BLOCKS) {
    fprintf(stderr, "mahyco: mahyco_compute_delta_t_noalloc: %ld cells exceed the %d-cell maximum\n", g.nc,
            MH_NOALLOC_MAX_BLOCKS * MH_BLOCK);
    return MH_ERR_INVALID;
  }
  double* partial = nullptr;
  int st = scratch_ptr(&partial);
  if (st != MH_OK) return st;
  hipStream_t s = (hipStream_t)stream;
  if (g.dim == 3)
    hipLaunchKernelGGL(k_cfl_cells<3>, dim3((unsigned)nblocks), dim3(MH_BLOCK), 0, s, g, L, c, u0, u1, u2, partial);
  else
    hipLaunchKernelGGL(k_cfl_cells<2>, dim3((unsigned)nblocks), dim3(MH_BLOCK), 0, s, g, L, c, u0, u1, u1, partial);
  hipLaunchKernelGGL(k_dt_noalloc<>, dim3(1), dim3(MH_DT_BLOCK), 0, s, partial, (int)nblocks, options->cfl,
                     old_deltat, options->deltat_max, options->deltat_min, dt_out, status_out);
  return hip_check(hipGetLastError(), "mahyco_compute_delta_t_noalloc");
}

}  // extern "C"
