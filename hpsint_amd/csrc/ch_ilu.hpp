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
    int        tens;       // tensorial mobility
    // the Q1 linearisation cache: tiles of tx x ty cells and tz layers along
    // the last direction, nt tiles per direction (x, y (1 in 2D), march),
    // nt_lanes lanes per tile layer, nf fields per q-point
    int        tx, ty, tz, nt[3], nt_lanes, nf;
    double     inv_h[3], jxw;
    PhysParams pp;
    // storage of the block, colour by colour and stencil entry by stencil
    // entry: the 2x2 block (a, b) of stencil entry s of the node with
    // in-colour index i of colour k at off[k] + 4 (s cnt[k] + i) + 2 a + b
    // (consecutive nodes of a colour read consecutive 4-vectors).  The
    // colour's nodes form an mx[k] x my[k] x (cnt[k] / (mx[k] my[k])) grid,
    // x fastest.  After the 8 colours (at off[8]) the solves keep a work
    // vector of (c, mu) pairs in the same colour-major node order: node i of
    // colour k at pair noff[k] + i.  Set by ch_layout.
    int64_t    off[9], cnt[8], noff[9];
    int        mx[8], my[8];
  };

  // fills cp.off, cp.cnt, cp.noff, cp.mx and cp.my from the node counts
  void ch_layout(ChParams &cp);
  // values (of the factor's type) to allocate: the factor and the work vector
  inline int64_t ch_storage(const ChParams &cp)
  {
    return cp.off[8] + 2 * cp.noff[8];
  }

  // Cahn-Hilliard block of the tangent (FP64) from the FP64
  // linearisation cache (scalar or tensorial mobility, 1-12 order parameters)
  void ch_assemble(const ChParams &cp, const double *cache, double *A, hipStream_t s);
  // identity rows and columns at the constrained nodes (n node indices)
  void ch_constrain(const ChParams &cp, const int64_t *nodes, int64_t n, double *A,
                    hipStream_t s);
  // in-place ILU(0) (FP64) and its FP32 copy
  void ch_factor(const ChParams &cp, double *A, float *A32, hipStream_t s);
  // dst = (LU)^-1 src on (c, mu), inv_diag * src on the order parameters
  template <typename T>
  void ch_apply(const ChParams &cp, const T *F, const T *inv_diag, const T *src, T *dst,
                hipStream_t s);
} // namespace hps
