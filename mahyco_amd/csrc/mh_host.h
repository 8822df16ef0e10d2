// mh_host.h -- host-side declarations shared by the translation units of the HIP library:
// the case handle, the ensemble tables, launch helpers with optional HIP-event timing, and
// the per-(dimension, environment count) launch sequences, each compiled in its own unit
// (mh_lag.hip: Lagrange kernels; mh_sweep.hip, once per environment count: remap sweeps,
// finalize and the batched ensemble cycle).
#ifndef MH_HOST_H
#define MH_HOST_H

#include <hip/hip_runtime.h>

#include <cstdio>
#include <map>
#include <vector>

#include "../../include/mahyco.h"
#include "mh_kernels.h"
#include "mh_output.h"

// profiled kernel groups (sweeps separately per direction: they run different code)
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
  KN_SWEEP_KE,  // dual kinetic energy of a sweep (conservation-energie-totale)
  KN_CONS_EC,   // the kinetic energy each cell receives back (conservation-energie-totale)
  KN_PREPARE,   // remap vector from the stored Lagrangian state (staged mode)
  KN_COUNT
};

// sweep variants: one thread per cell / node; row segments along x and LDS tiles along y / z
enum { MH_SWEEP_FLAT = 0, MH_SWEEP_TILE = 1 };

struct ProfRec {
  int kind;
  hipEvent_t a, b;
};

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
  bool cell_tiled = false;  // 3D remap loop: LDS-staged cell kernel (MAHYCO_CELL_KERNEL=tiled; an autotuner candidate)
  bool staged = false;      // Lagrangian state written before the remap, U formed from it (mahyco_set_staged)
  int nm_len = MH_NM_L;     // target segment length of the marching dual sweeps (MAHYCO_MARCH_LEN)
  int nt_kc = MH_NT_KC;     // node planes per block of the tiled node kernel
  int ct_k = MH_CT_K;       // cell planes per block of the tiled cell kernel
  int sweep_mode = MH_SWEEP_TILE;  // MAHYCO_SWEEP=flat selects one thread per cell / node
  std::vector<const mh_output::VarSpec*> ens_vars;  // EnSight variables of this case's outputs
  std::vector<double> ens_times;                    // times of the EnSight steps written
  MhScalars* ext_sc = nullptr;  // scalars living in an ensemble's array
  bool async_alloc = false;     // ensemble case: stream-ordered device allocation and release
  size_t bytes = 0;             // device bytes of the case
  // The state arrays of a case are one device block: while `planning`, dalloc records each
  // array, and commit_arena allocates the block and hands out 256-byte aligned pieces
  // (allocs / alloc_bytes list the pieces).  `owned` holds the blocks to free.
  bool planning = false;
  std::vector<std::pair<void**, size_t>> plan;
  std::vector<void*> owned;
  hipStream_t astream = nullptr;  // stream of the case's allocation and initial uploads
};

inline int hip_check(hipError_t e, const char* what) {
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

inline void prof_begin(MhHandle* h, ProfRec& r, int kind) {
  r.kind = kind;
  (void)hipEventCreate(&r.a);
  (void)hipEventCreate(&r.b);
  (void)hipEventRecord(r.a, h->stream);
}
inline void prof_end(MhHandle* h, ProfRec& r) {
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

// blocks of the row-segment x sweep (one wavefront per MH_ROW_OUT cells of a row), a multiple of 8
inline int row_blocks(MhHandle* h) {
  const long waves = (long)h->g.ny * h->g.nz * ((h->g.nx + MH_ROW_OUT - 1) / MH_ROW_OUT);
  const long wpb = MH_BLOCK / 64;
  return (int)(((waves + wpb - 1) / wpb + 7) / 8 * 8);
}

// blocks of the row-segment dual x sweep (one wavefront per MH_ROW_OUT nodes of a node row), a multiple of 8
inline int node_row_blocks(MhHandle* h) {
  const long waves = (long)h->g.nny * h->g.nnz * ((h->g.nnx + MH_ROW_OUT - 1) / MH_ROW_OUT);
  const long wpb = MH_BLOCK / 64;
  return (int)(((waves + wpb - 1) / wpb + 7) / 8 * 8);
}

// blocks of the LDS-tiled y / z sweep (64 cells along x by MH_TILE_L along D), a multiple of 8
// k_sweep_node_march: one thread per (x node, other transverse node, segment of about nm_len nodes along D)
inline int node_march_blocks(MhHandle* h, int D) {
  const long na = D == 1 ? h->g.ny : h->g.nz;
  const long nt2 = D == 1 ? h->g.nnz : h->g.nny;
  const long threads = (long)h->g.nnx * nt2 * march_nseg((int)na + 1, h->nm_len);
  return (int)(((threads + MH_BLOCK - 1) / MH_BLOCK + 7) / 8 * 8);
}
inline int tile_blocks(MhHandle* h, int D) {
  const long na = D == 1 ? h->g.ny : h->g.nz;
  const long nt = D == 1 ? (h->g.dim == 3 ? h->g.nz : 1) : h->g.ny;
  const long tiles = (long)((h->g.nx + 63) / 64) * nt * ((na + MH_TILE_L - 1) / MH_TILE_L);
  return (int)((tiles + 7) / 8 * 8);
}

// cell and dual parts of a y or z sweep in the selected variant (the dual part is one
// thread per node in both)

// A device work list of (case, block) entries walked by the persistent batched kernels, with
// its length in device memory, so that a captured graph stays valid while cases join and
// leave; the capacity grows by doubling (a new buffer: the ensemble's graphs are re-captured).
struct EnsList {
  int2* d = nullptr;
  int* d_n = nullptr;
  int cap = 0;
  std::vector<int2> h;  // host image of d[0 .. h.size())
  bool dirty = false;
};

// Cases of one kernel set advance together.  The key fixes every template parameter of the
// group's kernels, the limiter (LIM 0..3; -1 for order 3 and the generalised ones, read at
// run time) included.
struct EnsGroup {
  int dim = 0, m = 0;
  bool o3 = false;   // order-3 remap
  bool gen = false;  // pente-borne or generalised limiters
  bool csts = false; // schema-csts
  bool adv = false;  // advection mode (sans-lagrange): prescribed node velocity, no dual sweeps
  bool lc = false;   // characteristic length other than V^(1/d): k_cell_length first
  bool rows = false; // row-segment x sweeps and LDS-tiled y / z cell sweeps (else one thread per item)
  bool ntile = false;  // 3D: z-marching node kernel
  bool cons = false;   // conservation-energie-totale: dual kinetic energy and its return to the cells
  int lim = -1;
  int bpc = 4;         // persistent blocks per CU of the group's launches (autotuned)
  std::vector<int> slots;
  // cells, nodes, x rows, y / z tiles, node tiles, y / z marching node segments
  EnsList cmap, nmap, rcmap, rnmap, tmap[2], ntmap, mmap[2];
  int* d_list = nullptr;   // slots of the group, for the per-case time step
  int* d_nlist = nullptr;
  int list_cap = 0;
  bool list_dirty = false;
};

struct MhEnsembleImpl {
  hipStream_t stream = nullptr;
  // admissions allocate and upload on their own stream, which the iterations' stream then
  // waits for: admitting a case does not wait for the iterations in flight
  hipStream_t copy_stream = nullptr;
  hipEvent_t copy_done = nullptr;
  int max_slots = 0;
  double budget = 0.;
  size_t used = 0;
  int ncu = 256;                   // compute units of the device (persistent grid sizes)
  std::vector<MhHandle*> slot;     // active cases
  std::vector<MhHandle*> pending;  // admitted, activated at the next phase boundary
  std::vector<int> pending_slot;
  std::vector<char> tab_dirty;     // slots whose state-table entries must be uploaded
  MhScalars* d_sc = nullptr;       // scalars of every slot
  MhDev* d_tab[2] = {nullptr, nullptr};  // case states by coordinate parity
  std::vector<EnsGroup> groups;    // never removed: an emptied group keeps its (empty) lists
  long launched = 0;
  long version = 0;                // bumped when the launch sequence changes (graphs re-captured)
  std::map<long, hipGraphExec_t> graphs;
  std::vector<hipGraph_t> graph_defs;
  // pinned staging of the host-to-device list and table updates, reused once the previous
  // update's copies have run
  char* stage = nullptr;
  size_t stage_cap = 0;
  hipEvent_t stage_done = nullptr;
  bool stage_pending = false;
  int tuned_bpc = 0;               // blocks per CU chosen by mahyco_ensemble_autotune (0: default)
  long graph_captures = 0;         // graphs captured so far
  // snapshot of the slot scalars for mahyco_ensemble_poll_async / _poll_result
  MhScalars* snap = nullptr;       // pinned, max_slots entries
  hipEvent_t snap_done = nullptr;
  bool snap_pending = false;
  std::vector<int> snap_state;     // slot state when the snapshot was taken: -100 empty, 2 pending, 1 active
};

// Launch sequences of one (dimension, environment count), defined in mh_lag.hip and
// mh_sweep.hip and instantiated there for dimensions 2, 3 and 1..4 environments.
template <int DIM, int M> void launch_node_t(MhHandle* h);
template <int DIM, int M> void launch_cell_t(MhHandle* h);
template <int DIM, int M> void launch_sweeps_t(MhHandle* h);
template <int DIM, int M> void launch_prepare_t(MhHandle* h);
template <int DIM, int M> void ens_group_cycle(MhEnsembleImpl* E, const EnsGroup& g);

#define MH_EXTERN_DM(fn, ARG)                                                                 \
  extern template void fn<2, 1>(ARG); extern template void fn<2, 2>(ARG);                     \
  extern template void fn<2, 3>(ARG); extern template void fn<2, 4>(ARG);                     \
  extern template void fn<3, 1>(ARG); extern template void fn<3, 2>(ARG);                     \
  extern template void fn<3, 3>(ARG); extern template void fn<3, 4>(ARG);
MH_EXTERN_DM(launch_node_t, MhHandle*)
MH_EXTERN_DM(launch_cell_t, MhHandle*)
MH_EXTERN_DM(launch_sweeps_t, MhHandle*)
MH_EXTERN_DM(launch_prepare_t, MhHandle*)
#define MH_ENS_ARGS MhEnsembleImpl*, const EnsGroup&
MH_EXTERN_DM(ens_group_cycle, MH_ENS_ARGS)

#endif
