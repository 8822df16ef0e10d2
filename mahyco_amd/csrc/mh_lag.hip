// mh_lag.hip -- launch sequences of the Lagrange phase (node and cell kernels).
#include "mh_kernels.h"
#include "mh_host.h"

template <int DIM, int M>
void launch_node_t(MhHandle* h) {
  if (h->dev.sans_lagrange) {
    LAUNCH(h, KN_NODE, (k_node_adv<DIM>), h->nblk_n, MH_BLOCK, h->dev);
  } else if (DIM == 3 && h->node_tiled) {
    const dim3 grid((h->g.nnx + MH_NTX - 1) / MH_NTX, (h->g.nny + MH_NTY - 1) / MH_NTY,
                    (h->g.nnz + h->dev.nt_kc - 1) / h->dev.nt_kc);
    LAUNCH(h, KN_NODE, k_node_lag3_tiled<>, grid, MH_BLOCK, h->dev);
  } else {
    LAUNCH(h, KN_NODE, (k_node_lag<DIM>), h->nblk_n, MH_BLOCK, h->dev);
  }
}
template <int DIM, int M>
void launch_cell_t(MhHandle* h) {
  if (h->dev.lc != MH_L_RACINE_CUBIQUE_VOLUME) LAUNCH(h, KN_CELL, (k_cell_length<DIM>), h->nblk_c, MH_BLOCK, h->dev);
  if (h->dev.csts) {
    LAUNCH(h, KN_CELL, (k_cell_lag_csts<DIM, M>), h->nblk_c, MH_BLOCK, h->dev);
  } else if (DIM == 3 && h->dev.with_proj && !h->dev.sans_lagrange && !h->staged && h->cell_tiled) {
    // the Lagrange-only loop keeps k_cell_lag: its CFL candidates are folded per block in cell order
    const dim3 grid((h->g.nx + MH_CT_X - 1) / MH_CT_X, (h->g.ny + MH_CT_Y - 1) / MH_CT_Y,
                    (h->g.nz + h->dev.ct_k - 1) / h->dev.ct_k);
    LAUNCH(h, KN_CELL, (k_cell_lag3_tiled<M>), grid, MH_BLOCK, h->dev);
  } else {
    LAUNCH(h, KN_CELL, (k_cell_lag<DIM, M>), h->nblk_c, MH_BLOCK, h->dev);
  }
}

// staged mode: the remap vector from the state the cell kernel wrote (and the caller may have changed)
template <int DIM, int M>
void launch_prepare_t(MhHandle* h) {
  LAUNCH(h, KN_PREPARE, (k_prepare_remap<DIM, M>), h->nblk_c, MH_BLOCK, h->dev);
}

template void launch_node_t<2, 1>(MhHandle*);
template void launch_node_t<2, 2>(MhHandle*);
template void launch_node_t<2, 3>(MhHandle*);
template void launch_node_t<2, 4>(MhHandle*);
template void launch_node_t<3, 1>(MhHandle*);
template void launch_node_t<3, 2>(MhHandle*);
template void launch_node_t<3, 3>(MhHandle*);
template void launch_node_t<3, 4>(MhHandle*);
template void launch_cell_t<2, 1>(MhHandle*);
template void launch_cell_t<2, 2>(MhHandle*);
template void launch_cell_t<2, 3>(MhHandle*);
template void launch_cell_t<2, 4>(MhHandle*);
template void launch_cell_t<3, 1>(MhHandle*);
template void launch_cell_t<3, 2>(MhHandle*);
template void launch_cell_t<3, 3>(MhHandle*);
template void launch_cell_t<3, 4>(MhHandle*);
template void launch_prepare_t<2, 1>(MhHandle*);
template void launch_prepare_t<2, 2>(MhHandle*);
template void launch_prepare_t<2, 3>(MhHandle*);
template void launch_prepare_t<2, 4>(MhHandle*);
template void launch_prepare_t<3, 1>(MhHandle*);
template void launch_prepare_t<3, 2>(MhHandle*);
template void launch_prepare_t<3, 3>(MhHandle*);
template void launch_prepare_t<3, 4>(MhHandle*);
