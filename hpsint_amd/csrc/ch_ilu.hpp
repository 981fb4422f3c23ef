// ch_ilu.hpp -- ILU(0) of the Cahn-Hilliard block in parity-colour order,
// with the inverse diagonal of the Allen-Cahn block (the GPU form of the
// reference's BlockPreconditioner2); kernels in ch_ilu.hip.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "physics.hpp"

namespace hps
{
  struct ChParams
  {
    int        dim;
    int        nc[3];      // cells per physical direction (nc[2] = 1 in 2D)
    int        nn[3];      // nodes per physical direction (nn[2] = 1 in 2D)
    int64_t    n_nodes;
    int        n_comp, G;
    // the Q1 linearisation cache: tiles of tx x ty cells and tz layers along
    // the last direction, nt tiles per direction (x, y (1 in 2D), march),
    // nt_lanes lanes per tile layer, nf fields per q-point
    int        tx, ty, tz, nt[3], nt_lanes, nf;
    double     inv_h[3], jxw;
    PhysParams pp;
    // storage of the block, colour by colour and entry by entry: entry
    // (s, a, b) of the node with in-colour index i of colour k at
    // off[k] + (4 s + 2 a + b) cnt[k] + i (consecutive nodes of a colour
    // read consecutive addresses); set by ch_layout
    int64_t    off[9], cnt[8];
  };

  // fills cp.off and cp.cnt from the node counts
  void ch_layout(ChParams &cp);

  // per colour: the lower-colour neighbours as stencil indices, ordered by
  // (colour, index); n_lower of them
  struct ChOrder
  {
    signed char order[8][26];
    int         n_lower[8];
  };

  ChOrder ch_order(int dim);
  // Cahn-Hilliard block of the tangent (FP64) from the FP64
  // linearisation cache (scalar mobility)
  void ch_assemble(const ChParams &cp, const double *cache, double *A, hipStream_t s);
  // in-place ILU(0) (FP64) and its FP32 copy
  void ch_factor(const ChParams &cp, double *A, float *A32, hipStream_t s);
  // dst = (LU)^-1 src on (c, mu), inv_diag * src on the order parameters
  template <typename T>
  void ch_apply(const ChParams &cp, const T *F, const T *inv_diag, const T *src, T *dst,
                hipStream_t s);
} // namespace hps
