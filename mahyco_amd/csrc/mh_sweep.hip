// mh_sweep.hip -- launch sequences of the remap (sweeps in the rotated direction order,
// finalize) for single meshes, and of a whole iteration for a batched ensemble group.
// Compiled once per environment count: -DMH_TU_M=1..4.
#include "mh_kernels.h"
#include "mh_host.h"

#ifndef MH_TU_M
#error "MH_TU_M (environment count of this unit) must be defined"
#endif

// Cell part of one sweep direction (computeGradPhiFace ... computeUremap).
template <int DIM, int M>
void launch_cell_part(MhHandle* h, int D, int sel, int first, bool gen, bool o3) {
  const MhDev& S = h->dev;
  const bool flat = h->sweep_mode == MH_SWEEP_FLAT;
  if (gen) {
    if (D == 0) LAUNCH(h, KN_SWEEP_CELL, (k_sweep_cell_gen<DIM, 0, M>), h->nblk_c, MH_BLOCK, S, sel, first);
    else if (D == 1) LAUNCH(h, KN_SWEEP_CELL + 1, (k_sweep_cell_gen<DIM, 1, M>), h->nblk_c, MH_BLOCK, S, sel, first);
    else if constexpr (DIM == 3) LAUNCH(h, KN_SWEEP_CELL + 2, (k_sweep_cell_gen<3, 2, M>), h->nblk_c, MH_BLOCK, S, sel, first);
  } else if (o3) {
    if (D == 0) LAUNCH(h, KN_SWEEP_CELL, (k_sweep_cell3<DIM, 0, M>), h->nblk_c, MH_BLOCK, S, sel, first);
    else if (D == 1) LAUNCH(h, KN_SWEEP_CELL + 1, (k_sweep_cell3<DIM, 1, M>), h->nblk_c, MH_BLOCK, S, sel, first);
    else if constexpr (DIM == 3) LAUNCH(h, KN_SWEEP_CELL + 2, (k_sweep_cell3<3, 2, M>), h->nblk_c, MH_BLOCK, S, sel, first);
  } else if (D == 0) {
    if (flat) MH_LIMSW(h, LAUNCH(h, KN_SWEEP_CELL, (k_sweep_cell<DIM, 0, M, LIM>), h->nblk_c, MH_BLOCK, S, sel, first));
    else MH_LIMSW(h, LAUNCH(h, KN_SWEEP_CELL, (k_sweep_cell_row<DIM, M, LIM>), row_blocks(h), MH_BLOCK, S, sel, first));
  } else if (D == 1) {
    if (flat) MH_LIMSW(h, LAUNCH(h, KN_SWEEP_CELL + 1, (k_sweep_cell<DIM, 1, M, LIM>), h->nblk_c, MH_BLOCK, S, sel, first));
    else MH_LIMSW(h, LAUNCH(h, KN_SWEEP_CELL + 1, (k_sweep_cell_tile<DIM, 1, M, LIM>), tile_blocks(h, 1), MH_BLOCK, S, sel, first));
  } else if constexpr (DIM == 3) {
    if (flat) MH_LIMSW(h, LAUNCH(h, KN_SWEEP_CELL + 2, (k_sweep_cell<3, 2, M, LIM>), h->nblk_c, MH_BLOCK, S, sel, first));
    else MH_LIMSW(h, LAUNCH(h, KN_SWEEP_CELL + 2, (k_sweep_cell_tile<3, 2, M, LIM>), tile_blocks(h, 2), MH_BLOCK, S, sel, first));
  }
}

// Dual part of one sweep direction (computeDualUremap), then the dual kinetic energy when
// the total energy is conserved.  None in advection mode (withDualProjection, PrepareRemap.cc:211).
template <int DIM>
void launch_dual_part(MhHandle* h, int D, int sel, int first, int last, bool gen, bool o3) {
  const MhDev& S = h->dev;
  if (S.sans_lagrange) return;
  const bool flat = h->sweep_mode == MH_SWEEP_FLAT;
  if (gen) {
    if (D == 0) LAUNCH(h, KN_SWEEP_NODE, (k_sweep_node<DIM, 0, MH_LIM_ANY>), h->nblk_n, MH_BLOCK, S, sel, first, last);
    else if (D == 1) LAUNCH(h, KN_SWEEP_NODE + 1, (k_sweep_node<DIM, 1, MH_LIM_ANY>), h->nblk_n, MH_BLOCK, S, sel, first, last);
    else if constexpr (DIM == 3) LAUNCH(h, KN_SWEEP_NODE + 2, (k_sweep_node<3, 2, MH_LIM_ANY>), h->nblk_n, MH_BLOCK, S, sel, first, last);
  } else if (D == 0) {
    if (flat) MH_LIMSW(h, LAUNCH(h, KN_SWEEP_NODE, (k_sweep_node<DIM, 0, LIM>), h->nblk_n, MH_BLOCK, S, sel, first, last));
    else MH_LIMSW(h, LAUNCH(h, KN_SWEEP_NODE, (k_sweep_node_row<DIM, LIM>), node_row_blocks(h), MH_BLOCK, S, sel, first, last));
  } else if (D == 1) {
    if (!flat && !o3)  // threads march along y over segments of nodes
      MH_LIMSW(h, LAUNCH(h, KN_SWEEP_NODE + 1, (k_sweep_node_march<DIM, 1, LIM>), node_march_blocks(h, 1), MH_BLOCK, S, sel, first, last));
    else
      MH_LIMSW(h, LAUNCH(h, KN_SWEEP_NODE + 1, (k_sweep_node<DIM, 1, LIM>), h->nblk_n, MH_BLOCK, S, sel, first, last));
  } else if constexpr (DIM == 3) {
    if (!flat && !o3)  // threads march along z over segments of nodes
      MH_LIMSW(h, LAUNCH(h, KN_SWEEP_NODE + 2, (k_sweep_node_march<3, 2, LIM>), node_march_blocks(h, 2), MH_BLOCK, S, sel, first, last));
    else
      MH_LIMSW(h, LAUNCH(h, KN_SWEEP_NODE + 2, (k_sweep_node<3, 2, LIM>), h->nblk_n, MH_BLOCK, S, sel, first, last));
  }
  if (S.dec) {  // conservation-energie-totale: the dual kinetic energy of this sweep
    if (D == 0) LAUNCH(h, KN_SWEEP_KE, (k_sweep_ke<DIM, 0>), h->nblk_n, MH_BLOCK, S, sel, first);
    else if (D == 1) LAUNCH(h, KN_SWEEP_KE, (k_sweep_ke<DIM, 1>), h->nblk_n, MH_BLOCK, S, sel, first);
    else if constexpr (DIM == 3) LAUNCH(h, KN_SWEEP_KE, (k_sweep_ke<3, 2>), h->nblk_n, MH_BLOCK, S, sel, first);
  }
}

// The remap of one iteration (RemapADIService::appliRemap, RemapADIService.cc:15-90, and
// remapVariables).
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
    launch_cell_part<DIM, M>(h, D, sel, first, gen, o3);
    launch_dual_part<DIM>(h, D, sel, first, last, gen, o3);
  }
  if (S.dec && nd == 0) LAUNCH(h, KN_SWEEP_KE, (k_dual_init<DIM>), h->nblk_n, MH_BLOCK, S);
  if (nd <= 1 && !S.sans_lagrange)
    LAUNCH(h, KN_NODE_VEL, (k_node_velocity<DIM>), h->nblk_n, MH_BLOCK, S, nd == 0 && !S.dec ? 1 : 0);
  if (S.dec) LAUNCH(h, KN_CONS_EC, (k_cons_ec<DIM>), h->nblk_c, MH_BLOCK, S);
  LAUNCH(h, KN_FINAL, (k_finalize<DIM, M>), h->nblk_c, MH_BLOCK, S, nd & 1);
  h->sens = (h->sens + 1) % dim;
}

// ---- batched ensembles: one persistent launch per phase and batch group --------------------
#define MH_LIST(l) (l).d, (l).d_n

// One sweep direction of a batch group (cell part, then dual part).
template <int DIM, int D, int M, int LIM>
void ens_sweep(const EnsGroup& g, const MhDev* tab, hipStream_t st, dim3 G, int sel, int first, int last) {
  const dim3 B(MH_BLOCK);
  if constexpr (LIM < 0) {  // order 3 or pente-borne / generalised limiters: one thread per item
    if (g.gen) {
      hipLaunchKernelGGL((k_sweep_cell_gen_batch<DIM, D, M>), G, B, 0, st, tab, MH_LIST(g.cmap), sel, first);
      if (!g.adv)
        hipLaunchKernelGGL((k_sweep_node_batch<DIM, D, MH_LIM_ANY>), G, B, 0, st, tab, MH_LIST(g.nmap), sel, first, last);
    } else {
      hipLaunchKernelGGL((k_sweep_cell3_batch<DIM, D, M>), G, B, 0, st, tab, MH_LIST(g.cmap), sel, first);
      if (!g.adv)
        hipLaunchKernelGGL((k_sweep_node_batch<DIM, D, -1>), G, B, 0, st, tab, MH_LIST(g.nmap), sel, first, last);
    }
  } else if (g.rows && D == 0) {
    hipLaunchKernelGGL((k_sweep_cell_row_batch<DIM, M, LIM>), G, B, 0, st, tab, MH_LIST(g.rcmap), sel, first);
    if (!g.adv)
      hipLaunchKernelGGL((k_sweep_node_row_batch<DIM, LIM>), G, B, 0, st, tab, MH_LIST(g.rnmap), sel, first, last);
  } else if (g.rows) {
    hipLaunchKernelGGL((k_sweep_cell_tile_batch<DIM, D, M, LIM>), G, B, 0, st, tab, MH_LIST(g.tmap[D == 0 ? 0 : D - 1]),
                       sel, first);
    if (!g.adv && !g.o3)  // threads march along D over segments of nodes
      hipLaunchKernelGGL((k_sweep_node_march_batch<DIM, (D == 0 ? 1 : D), LIM>), G, B, 0, st, tab,
                         MH_LIST(g.mmap[D == 0 ? 0 : D - 1]), sel, first, last);
    else if (!g.adv)
      hipLaunchKernelGGL((k_sweep_node_batch<DIM, D, LIM>), G, B, 0, st, tab, MH_LIST(g.nmap), sel, first, last);
  } else {
    hipLaunchKernelGGL((k_sweep_cell_batch<DIM, D, M, LIM>), G, B, 0, st, tab, MH_LIST(g.cmap), sel, first);
    if (!g.adv)
      hipLaunchKernelGGL((k_sweep_node_batch<DIM, D, LIM>), G, B, 0, st, tab, MH_LIST(g.nmap), sel, first, last);
  }
}

// One iteration of every case of a batch group: a fixed grid of E->ncu * g.bpc blocks per
// launch walks the group's work lists.
template <int DIM, int M, int LIM>
void ens_group_cycle_l(MhEnsembleImpl* E, const EnsGroup& g) {
  const int par = (int)(E->launched & 1);
  const MhDev* tab = E->d_tab[par];
  hipStream_t st = E->stream;
  const dim3 G((unsigned)(E->ncu * g.bpc)), B(MH_BLOCK);
  if (g.adv) hipLaunchKernelGGL((k_node_adv_batch<DIM>), G, B, 0, st, tab, MH_LIST(g.nmap));
  else if (DIM == 3 && g.ntile) hipLaunchKernelGGL((k_node_lag3_tiled_batch<>), G, B, 0, st, tab, MH_LIST(g.ntmap));
  else hipLaunchKernelGGL((k_node_lag_batch<DIM>), G, B, 0, st, tab, MH_LIST(g.nmap));
  if (g.lc) hipLaunchKernelGGL((k_cell_length_batch<DIM>), G, B, 0, st, tab, MH_LIST(g.cmap));
  if (g.csts) hipLaunchKernelGGL((k_cell_lag_csts_batch<DIM, M>), G, B, 0, st, tab, MH_LIST(g.cmap));
  else hipLaunchKernelGGL((k_cell_lag_batch<DIM, M>), G, B, 0, st, tab, MH_LIST(g.cmap));
  if (g.csts) hipLaunchKernelGGL((k_node_fwd_batch<DIM>), G, B, 0, st, tab, MH_LIST(g.nmap));
  const int sens = (int)(E->launched % DIM);
  if (g.o3 || g.gen) hipLaunchKernelGGL((k_hcell_batch<DIM>), G, B, 0, st, tab, MH_LIST(g.cmap));
  for (int s = 0; s < DIM; ++s) {
    const int D = (s + sens) % DIM, first = s == 0, last = s == DIM - 1, sel = s & 1;
    if (D == 0) ens_sweep<DIM, 0, M, LIM>(g, tab, st, G, sel, first, last);
    else if (D == 1) ens_sweep<DIM, 1, M, LIM>(g, tab, st, G, sel, first, last);
    else if constexpr (DIM == 3) ens_sweep<3, 2, M, LIM>(g, tab, st, G, sel, first, last);
    if (g.cons) {  // the dual kinetic energy of this sweep
      if (D == 0) hipLaunchKernelGGL((k_sweep_ke_batch<DIM, 0>), G, B, 0, st, tab, MH_LIST(g.nmap), sel, first);
      else if (D == 1) hipLaunchKernelGGL((k_sweep_ke_batch<DIM, 1>), G, B, 0, st, tab, MH_LIST(g.nmap), sel, first);
      else if constexpr (DIM == 3)
        hipLaunchKernelGGL((k_sweep_ke_batch<3, 2>), G, B, 0, st, tab, MH_LIST(g.nmap), sel, first);
    }
  }
  if (g.cons) hipLaunchKernelGGL((k_cons_ec_batch<DIM>), G, B, 0, st, tab, MH_LIST(g.cmap));
  hipLaunchKernelGGL((k_finalize_batch<DIM, M>), G, B, 0, st, tab, MH_LIST(g.cmap), DIM & 1);
  hipLaunchKernelGGL((k_dt_batch<>), dim3((unsigned)E->ncu), dim3(MH_DT_BLOCK), 0, st, tab, g.d_list, g.d_nlist);
}

template <int DIM, int M>
void ens_group_cycle(MhEnsembleImpl* E, const EnsGroup& g) {
  switch (g.lim) {
    case 0: ens_group_cycle_l<DIM, M, 0>(E, g); break;
    case 1: ens_group_cycle_l<DIM, M, 1>(E, g); break;
    case 2: ens_group_cycle_l<DIM, M, 2>(E, g); break;
    case 3: ens_group_cycle_l<DIM, M, 3>(E, g); break;
    default: ens_group_cycle_l<DIM, M, -1>(E, g); break;
  }
}

template void launch_sweeps_t<2, MH_TU_M>(MhHandle*);
template void launch_sweeps_t<3, MH_TU_M>(MhHandle*);
template void ens_group_cycle<2, MH_TU_M>(MhEnsembleImpl*, const EnsGroup&);
template void ens_group_cycle<3, MH_TU_M>(MhEnsembleImpl*, const EnsGroup&);
