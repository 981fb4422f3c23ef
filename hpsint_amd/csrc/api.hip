// api.hip -- C ABI of the MI355X sintering-operator library (see
// include/hpsint_amd.h for the contract of every entry point).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <string>
#include <vector>

#include "../../include/hpsint_amd.h"
#include "blas.hpp"
#include "ch_ilu.hpp"
#include "mlp.hpp"
#include "ops.hpp"

namespace
{
  thread_local std::string g_last_error;

  int fail(int code, const std::string &msg)
  {
    g_last_error = msg;
    return code;
  }

#define HPS_HIP(call)                                                              \
  do                                                                               \
    {                                                                              \
      const hipError_t e_ = (call);                                                \
      if (e_ != hipSuccess)                                                        \
        return fail(e_ == hipErrorOutOfMemory ? HPS_ERR_OUT_OF_MEMORY : HPS_ERR_HIP, \
                    std::string(#call) + ": " + hipGetErrorString(e_));            \
    }                                                                              \
  while (0)

#define HPS_LAUNCHED()                                                             \
  do                                                                               \
    {                                                                              \
      const hipError_t e_ = hipGetLastError();                                     \
      if (e_ != hipSuccess)                                                        \
        return fail(HPS_ERR_HIP, std::string("kernel launch: ") + hipGetErrorString(e_)); \
    }                                                                              \
  while (0)

  // pinned host buffer of the reductions (and of the per-block loss partials)
  constexpr int RED_HOST_DOUBLES = 4096;

  int env_int(const char *name, int dflt)
  {
    const char *v = std::getenv(name);
    return v ? std::atoi(v) : dflt;
  }
} // namespace

struct hps_operator_s
{
  hps_params           p;
  int                  dim, n_grains, nc;
  bool                 tens;
  int                  degree;
  const hps::OpTable  *ops; // Q1 kernels
  const hps::Q2Table  *q2;  // Q2 kernels
  hps::Q2Params        qp;
  hps::KParams         kp;
  hps::FixParams       fp;
  int                  n_tiles, tz;
  // 3D scalar tangent with z-column lanes (8 x 8 x tz tiles, ntz1 along y)
  bool                 zcol, zcol32, zcol_residual;
  int                  ntz1;
  // layers per z-column march of the FP64 tangent and the residual (a
  // multiple of the cache tile depth tz; the FP32 tangent marches tz)
  int                  zcol_mz;
  int64_t              n_nodes, n_dofs, cache_elems, part_elems;
  hipStream_t          stream;
  double              *cache64;
  float               *cache32;
  bool                 lin_set;
  double              *part;
  double              *red_partial, *red_out, *red_host;
  double             **ptrs_dev;
  double             **ptrs_host;
  int                  precond_kind;
  double              *diag, *inv_diag64, *blocks, *inv_blocks64;
  float               *inv_diag32, *inv_blocks32;
  double              *ch64;     // ILU(0) factors of the Cahn-Hilliard block
  float               *ch32;
  int                 *d_singular;
  // kernel timing: a ring of `tcap` event triples (before the main kernel,
  // between it and the fix-up, after the fix-up); tcount calls recorded
  std::vector<hipEvent_t> *tev;
  int                  tcap, tcount;
  bool                 timing;
  const double        *old[3];
  double               cut_tol;          // grain cut-off tolerance (<= 0: off)
  bool                 cut_active;       // a component table is set (cell_masks)
  unsigned            *cell_masks;       // per cell: bit g = eta_g active
  bool                 cut_ready;        // tile groups for the fast cut vmult
  unsigned            *cut_masks;        // per tile: union of its cell masks
  int                 *cut_tiles;        // tiles grouped by active count
  int                  cut_off[hps::CUT_MAX_GA + 2];
  int64_t              n_constrained;    // constrained nodes (none by default)
  int64_t             *constrained;      // their node indices (device)
  double              *constrained_save; // n_constrained * nc saved src entries
};

static int check_op(hps_operator op)
{
  if (!op)
    return fail(HPS_ERR_INVALID_ARGUMENT, "null operator handle");
  return HPS_OK;
}

// Q2 launch parameters with the current time weights
static hps::Q2Params q2_params(hps_operator op)
{
  hps::Q2Params q = op->qp;
  q.pp            = op->kp.pp;
  q.order         = op->kp.order;
  for (int i = 0; i < 3; ++i)
    q.w_hist[i] = op->kp.w_hist[i];
  return q;
}

// launch parameters of the z-column kernels: marches of zcol_mz layers
// (wide) or of one cache tile, over the cache tiles of kp.tz layers
static hps::KParams zcol_params(hps_operator op, bool wide)
{
  hps::KParams k = op->kp;
  k.ctz          = k.tz;
  if (wide)
    {
      k.tz    = op->zcol_mz;
      k.nt[2] = (k.nc[2] + k.tz - 1) / k.tz;
    }
  return k;
}

// workgroups of a z-column launch with parameters zk
static int zcol_tiles(hps_operator op, const hps::KParams &zk)
{
  return zk.nt[0] * op->ntz1 * zk.nt[2];
}

// event i (0: before the main kernel, 1: before the fix-up, 2: after it) of
// the current call's triple; i == 2 completes the call
static hipEvent_t timing_event(hps_operator op, int i)
{
  const int slot = op->tcount % op->tcap;
  hipEvent_t ev  = (*op->tev)[3 * slot + i];
  if (i == 2)
    ++op->tcount;
  return ev;
}

int hps_internal_fail(int code, const char *msg)
{
  return fail(code, msg);
}

extern "C" const char *hps_last_error(void)
{
  return g_last_error.c_str();
}

extern "C" int hps_create(const hps_params *prm, hps_operator *out)
{
  if (!prm || !out)
    return fail(HPS_ERR_INVALID_ARGUMENT, "null argument");
  *out = nullptr;
  if (prm->dim != 2 && prm->dim != 3)
    return fail(HPS_ERR_UNSUPPORTED, "dim must be 2 or 3");
  if (prm->fe_degree != 1 && prm->fe_degree != 2)
    return fail(HPS_ERR_UNSUPPORTED,
                "only FE_Q(1) with QGauss(2) and FE_Q(2) with QGauss(3) are built");
  if (prm->mobility != HPS_MOBILITY_SCALAR && prm->mobility != HPS_MOBILITY_TENSORIAL)
    return fail(HPS_ERR_INVALID_ARGUMENT, "unknown mobility");
  for (int d = 0; d < prm->dim; ++d)
    if (prm->n_cells[d] < 1 || !(prm->cell_size[d] > 0.0))
      return fail(HPS_ERR_INVALID_ARGUMENT, "cells and cell sizes must be positive");
  const int  G    = prm->n_components - 2;
  const bool tens = prm->mobility == HPS_MOBILITY_TENSORIAL;
  if (G < 1)
    return fail(HPS_ERR_INVALID_N_COMP, "n_components must be >= 3");
  // tile geometry: cells along y per workgroup (3D) and cells per tile along
  // the march direction; HPSINT_AMD_TILE_Y / HPSINT_AMD_TILE_M override the
  // defaults for measurements
  const int tile_y = env_int("HPSINT_AMD_TILE_Y", 4);
  const int TZ     = env_int("HPSINT_AMD_TILE_M", 16);
  if (TZ < 1 || TZ > 4096)
    return fail(HPS_ERR_INVALID_ARGUMENT, "HPSINT_AMD_TILE_M out of range");
  // Q1 3D z-column FP64 tangent and residual: most cache tiles per march
  // and the workgroups per CU a deeper march must leave
  const int zm    = env_int("HPSINT_AMD_ZCOL_MARCH", 2);
  const int zm_wg = env_int("HPSINT_AMD_ZCOL_MARCH_WG", 4);
  if (zm < 1 || zm > 64)
    return fail(HPS_ERR_INVALID_ARGUMENT, "HPSINT_AMD_ZCOL_MARCH out of range (1-64)");
  if (zm_wg < 0 || zm_wg > 1 << 20)
    return fail(HPS_ERR_INVALID_ARGUMENT, "HPSINT_AMD_ZCOL_MARCH_WG out of range");
  // Q2 3D tangent vmult: cell layers per tile march (0: colour passes)
  const int q2_tile = env_int("HPSINT_AMD_Q2_TILE", 2);
  if (q2_tile < 0 || q2_tile > 4096)
    return fail(HPS_ERR_INVALID_ARGUMENT, "HPSINT_AMD_Q2_TILE out of range");
  const bool          q2  = prm->fe_degree == 2;
  const hps::OpTable *ops = nullptr;
  const hps::Q2Table *q2t = nullptr;
  if (q2)
    q2t = hps::q2_table(prm->dim, G, tens);
  else
    ops = prm->dim == 3 ? hps::q1_table_d3(G, tens, tile_y) : hps::q1_table_d2(G, tens);
  if (!ops && !q2t)
    return fail(HPS_ERR_INVALID_N_COMP,
                "n_components = " + std::to_string(prm->n_components) +
                  " is not instantiated for this dim/degree/mobility");

  hps_operator op = new hps_operator_s();
  std::memset((void *)op, 0, sizeof(*op));
  op->p            = *prm;
  op->cut_tol      = -1.0;
  op->dim          = prm->dim;
  op->n_grains     = G;
  op->nc           = G + 2;
  op->tens         = tens;
  op->degree       = prm->fe_degree;
  op->ops          = ops;
  op->q2           = q2t;
  op->precond_kind = -1;

  hps::KParams &kp = op->kp;
  const int     ncx = prm->n_cells[0];
  const int     ncy = prm->dim == 3 ? prm->n_cells[1] : 1;
  const int     ncm = prm->dim == 3 ? prm->n_cells[2] : prm->n_cells[1];
  const int     p   = prm->fe_degree;
  kp.nc[0]         = ncx;
  kp.nc[1]         = ncy;
  kp.nc[2]         = ncm;
  kp.nn[0]         = p * ncx + 1;
  kp.nn[1]         = prm->dim == 3 ? p * ncy + 1 : 1;
  kp.nn[2]         = p * ncm + 1;
  kp.tz            = TZ;
  if (ops)
    {
      kp.nt[0] = (ncx + ops->tx - 1) / ops->tx;
      kp.nt[1] = (ncy + ops->ty - 1) / ops->ty;
      kp.nt[2] = (ncm + TZ - 1) / TZ;
    }
  double vol       = 1.0;
  for (int d = 0; d < 3; ++d)
    kp.inv_h[d] = d < prm->dim ? 1.0 / prm->cell_size[d] : 1.0;
  for (int d = 0; d < prm->dim; ++d)
    vol *= prm->cell_size[d] * 0.5;
  kp.jxw        = vol;
  kp.pp.A       = prm->A;
  kp.pp.B       = prm->B;
  kp.pp.kappa_c = prm->kappa_c;
  kp.pp.kappa_p = prm->kappa_p;
  kp.pp.Mvol    = prm->Mvol;
  kp.pp.Mvap    = prm->Mvap;
  kp.pp.Msurf   = prm->Msurf;
  kp.pp.Mgb     = prm->Mgb;
  kp.pp.L       = prm->L;
  kp.pp.w0      = 0.0;
  kp.pp.w_alg   = 1.0;
  kp.order      = 0;
  hps::set_fold_scales(kp, prm->dim);

  op->n_nodes = (int64_t)kp.nn[0] * kp.nn[1] * kp.nn[2];
  op->n_dofs  = op->n_nodes * op->nc;

  auto cleanup_fail = [&](hipError_t e) {
    hps_destroy(op);
    return fail(e == hipErrorOutOfMemory ? HPS_ERR_OUT_OF_MEMORY : HPS_ERR_HIP,
                std::string("allocation: ") + hipGetErrorString(e));
  };
  hipError_t e;
  if (q2)
    {
      // Q2 geometry: physical x, y, z (z = 1 cell in 2D)
      hps::Q2Params &qp = op->qp;
      qp.nc[0]          = prm->n_cells[0];
      qp.nc[1]          = prm->n_cells[1];
      qp.nc[2]          = prm->dim == 3 ? prm->n_cells[2] : 1;
      qp.h_prod         = 1.0;
      qp.tile_zc        = prm->dim == 3 ? q2_tile : 0;
      for (int d = 0; d < 3; ++d)
        {
          qp.nn[d]    = d < prm->dim ? 2 * qp.nc[d] + 1 : 1;
          qp.inv_h[d] = d < prm->dim ? 1.0 / prm->cell_size[d] : 1.0;
          if (d < prm->dim)
            qp.h_prod *= prm->cell_size[d];
        }
      op->cache_elems =
        hps::q2_workgroups(qp, prm->dim, q2t->cpw) * (int64_t)q2t->n_fields *
        (q2t->cpw * (prm->dim == 3 ? 27 : 9));
      op->part_elems = 0;
    }
  else
    {
      op->n_tiles = kp.nt[0] * kp.nt[1] * kp.nt[2];
      const int px = ops->tx + 1;
      const int pn = px * (prm->dim == 3 ? ops->ty + 1 : 1);
      op->cache_elems = (int64_t)op->n_tiles * TZ * ops->n_fields * ops->nt;
      op->tz          = TZ;
      op->part_elems  = (int64_t)op->n_tiles * (TZ + 1) * pn * op->nc;

      hps::FixParams &fp = op->fp;
      for (int d = 0; d < 3; ++d)
        {
          fp.nc[d] = kp.nc[d];
          fp.nn[d] = kp.nn[d];
          fp.nt[d] = kp.nt[d];
          fp.nb[d] = kp.nt[d] - 1;
        }
      fp.tsz[0] = ops->tx;
      fp.tsz[1] = prm->dim == 3 ? ops->ty : 1;
      fp.tsz[2] = TZ;
      fp.n_comp = op->nc;
      fp.px     = px;
      fp.pn     = pn;
      fp.n_m    = (int64_t)fp.nb[2] * fp.nn[0] * fp.nn[1];
      fp.n_y    = (int64_t)fp.nb[1] * fp.nn[0] * fp.nn[2];
      fp.n_x    = (int64_t)fp.nb[0] * fp.nn[1] * fp.nn[2];

      // the z-column tangent (3D scalar mobility, up to ZCOL_MAX_G order
      // parameters, the cache of 8 x 4 tiles); HPSINT_AMD_ZCOL=0 keeps the
      // one-lane-per-corner kernel
      const bool zc = prm->dim == 3 && !tens && ops->tx == hps::TX3 && ops->ty == hps::TY3 &&
                      env_int("HPSINT_AMD_ZCOL", 1) != 0;
      op->zcol   = zc && G <= hps::ZCOL_MAX_G;
      op->zcol32 = zc && G <= hps::ZCOL32_MAX_G;
      // the residual with z-column lanes as well (HPSINT_AMD_ZCOL_RES=0: not)
      op->zcol_residual = op->zcol && env_int("HPSINT_AMD_ZCOL_RES", 1) != 0;
      // FP64 tangent and residual: marches of up to HPSINT_AMD_ZCOL_MARCH
      // cache tiles (default 2: 32 layers, half the tile faces along z),
      // the deepest that still gives HPSINT_AMD_ZCOL_MARCH_WG (default 4)
      // workgroups per CU (2 run at once at 2 waves/SIMD; 128^3: two tiles,
      // 64^3: one); FP32: one (3 workgroups per CU; two tiles per march
      // left a tail)
      int dev = 0, cus = 0;
      if (hipGetDevice(&dev) != hipSuccess ||
          hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
          cus < 1)
        {
          (void)hipGetLastError();
          cus = 256; // MI355X
        }
      op->ntz1    = (ncy + 7) / 8;
      op->zcol_mz = TZ;
      for (int m = zm; m > 1; --m)
        if ((int64_t)kp.nt[0] * op->ntz1 * ((ncm + m * TZ - 1) / (m * TZ)) >=
            (int64_t)zm_wg * cus)
          {
            op->zcol_mz = m * TZ;
            break;
          }
      if (op->zcol || op->zcol32)
        {
          for (const bool wide : {false, true})
            op->part_elems =
              std::max(op->part_elems, hps::zcol_part_elems(zcol_params(op, wide), op->ntz1,
                                                            op->nc));
        }
      if ((e = hipMalloc(&op->part, op->part_elems * sizeof(double))) != hipSuccess)
        return cleanup_fail(e);
    }
  const int64_t red = (int64_t)hps::RED_BLOCKS * (hps::MAXK > op->nc ? hps::MAXK : op->nc);
  if ((e = hipMalloc(&op->red_partial, red * sizeof(double))) != hipSuccess)
    return cleanup_fail(e);
  if ((e = hipMalloc(&op->red_out, 64 * sizeof(double))) != hipSuccess)
    return cleanup_fail(e);
  if ((e = hipHostMalloc(&op->red_host, RED_HOST_DOUBLES * sizeof(double))) != hipSuccess)
    return cleanup_fail(e);
  if ((e = hipMalloc(&op->ptrs_dev, 32 * sizeof(double *))) != hipSuccess)
    return cleanup_fail(e);
  if ((e = hipHostMalloc(&op->ptrs_host, 32 * sizeof(double *))) != hipSuccess)
    return cleanup_fail(e);
  if ((e = hipMalloc(&op->d_singular, sizeof(int))) != hipSuccess)
    return cleanup_fail(e);
  op->tev = new std::vector<hipEvent_t>();
  *out = op;
  return HPS_OK;
}

extern "C" int hps_destroy(hps_operator op)
{
  if (!op)
    return HPS_OK;
  (void)hipStreamSynchronize(op->stream);
  (void)hipFree(op->cache64);
  (void)hipFree(op->cache32);
  (void)hipFree(op->part);
  (void)hipFree(op->red_partial);
  (void)hipFree(op->red_out);
  if (op->red_host)
    (void)hipHostFree(op->red_host);
  (void)hipFree(op->ptrs_dev);
  if (op->ptrs_host)
    (void)hipHostFree(op->ptrs_host);
  (void)hipFree(op->diag);
  (void)hipFree(op->inv_diag64);
  (void)hipFree(op->inv_diag32);
  (void)hipFree(op->blocks);
  (void)hipFree(op->inv_blocks64);
  (void)hipFree(op->inv_blocks32);
  (void)hipFree(op->ch64);
  (void)hipFree(op->ch32);
  (void)hipFree(op->d_singular);
  (void)hipFree(op->constrained);
  (void)hipFree(op->constrained_save);
  (void)hipFree(op->cut_masks);
  (void)hipFree(op->cut_tiles);
  (void)hipFree(op->cell_masks);
  if (op->tev)
    {
      for (hipEvent_t ev : *op->tev)
        (void)hipEventDestroy(ev);
      delete op->tev;
    }
  delete op;
  return HPS_OK;
}

extern "C" int hps_set_stream(hps_operator op, void *stream)
{
  if (int r = check_op(op))
    return r;
  op->stream = (hipStream_t)stream;
  return HPS_OK;
}

extern "C" int hps_sync(hps_operator op)
{
  if (int r = check_op(op))
    return r;
  HPS_HIP(hipStreamSynchronize(op->stream));
  return HPS_OK;
}

extern "C" int hps_sizes(hps_operator op, int64_t *n_nodes, int64_t *n_dofs)
{
  if (int r = check_op(op))
    return r;
  if (n_nodes)
    *n_nodes = op->n_nodes;
  if (n_dofs)
    *n_dofs = op->n_dofs;
  return HPS_OK;
}

extern "C" int hps_set_time_weights(hps_operator op, double w0, double w_alg,
                                    const double *w_hist, int order)
{
  if (int r = check_op(op))
    return r;
  if (order < 0 || order > 3 || (order > 0 && !w_hist))
    return fail(HPS_ERR_INVALID_ARGUMENT, "order must be 0..3 with weights");
  op->kp.pp.w0    = w0;
  op->kp.pp.w_alg = w_alg;
  op->kp.order    = order;
  for (int i = 0; i < 3; ++i)
    op->kp.w_hist[i] = i < order ? w_hist[i] : 0.0;
  return HPS_OK;
}

// the fast cut vmult (3D, scalar mobility, default tiles): tile masks as
// the union of their cells' masks, tiles grouped by their number of active
// order parameters; a tile with more than CUT_MAX_GA leaves the masked full
// kernels in use
static int group_cut_tiles(hps_operator op)
{
  op->cut_ready = false;
  if (op->dim != 3 || op->tens || op->ops->ty != hps::TY3 || op->n_grains > 32)
    return HPS_OK;
  const int nt = op->n_tiles;
  if (!op->cut_masks)
    {
      HPS_HIP(hipMalloc(&op->cut_masks, nt * sizeof(unsigned)));
      HPS_HIP(hipMalloc(&op->cut_tiles, nt * sizeof(int)));
    }
  hps::tile_masks_d3(op->kp, nt, op->cut_masks, op->stream);
  HPS_LAUNCHED();
  std::vector<unsigned> m(nt);
  HPS_HIP(hipMemcpyAsync(m.data(), op->cut_masks, nt * sizeof(unsigned), hipMemcpyDeviceToHost,
                         op->stream));
  HPS_HIP(hipStreamSynchronize(op->stream));
  int count[hps::CUT_MAX_GA + 1] = {};
  for (int i = 0; i < nt; ++i)
    {
      const int ga = __builtin_popcount(m[i]);
      if (ga > hps::CUT_MAX_GA)
        return HPS_OK;
      ++count[ga];
    }
  op->cut_off[0] = 0;
  for (int g = 0; g <= hps::CUT_MAX_GA; ++g)
    op->cut_off[g + 1] = op->cut_off[g] + count[g];
  std::vector<int> list(nt);
  int              pos[hps::CUT_MAX_GA + 1];
  for (int g = 0; g <= hps::CUT_MAX_GA; ++g)
    pos[g] = op->cut_off[g];
  for (int i = 0; i < nt; ++i)
    list[pos[__builtin_popcount(m[i])]++] = i;
  HPS_HIP(hipMemcpy(op->cut_tiles, list.data(), nt * sizeof(int), hipMemcpyHostToDevice));
  op->cut_ready = true;
  return HPS_OK;
}

// drops the component table; a cache filled with it is no longer valid
static void clear_cut(hps_operator op)
{
  if (op->cut_active)
    op->lin_set = false;
  op->cut_active = false;
  op->cut_ready  = false;
  op->kp.cmask   = nullptr;
}

extern "C" int hps_set_grain_cut_off(hps_operator op, double tol)
{
  if (int r = check_op(op))
    return r;
  clear_cut(op);
  if (!(tol > 0.0))
    {
      op->cut_tol = -1.0;
      return HPS_OK;
    }
  if (op->q2)
    return fail(HPS_ERR_UNSUPPORTED, "grain cut-off is built for Q1 operators");
  if (op->n_grains > 32)
    return fail(HPS_ERR_INVALID_ARGUMENT, "grain cut-off: at most 32 order parameters");
  op->cut_tol = tol;
  return HPS_OK;
}

extern "C" int hps_set_component_mask(hps_operator op, const double *u)
{
  if (int r = check_op(op))
    return r;
  if (!u)
    return fail(HPS_ERR_INVALID_ARGUMENT, "null vector");
  if (!(op->cut_tol > 0.0))
    return fail(HPS_ERR_NOT_INITIALIZED, "component mask without a grain cut-off tolerance");
  const int64_t n_cells = (int64_t)op->kp.nc[0] * op->kp.nc[1] * op->kp.nc[2];
  if (!op->cell_masks)
    HPS_HIP(hipMalloc(&op->cell_masks, n_cells * sizeof(unsigned)));
  clear_cut(op);
  op->lin_set = false; // the cache is filled with the new table
  hps::cell_masks(op->kp, op->dim, op->n_grains, op->cut_tol, u, op->cell_masks, op->stream);
  HPS_LAUNCHED();
  op->kp.cmask    = op->cell_masks;
  op->cut_active  = true;
  return group_cut_tiles(op);
}

extern "C" int hps_get_component_mask(hps_operator op, uint32_t *masks)
{
  if (int r = check_op(op))
    return r;
  if (!op->cut_active)
    return fail(HPS_ERR_NOT_INITIALIZED, "no component table set");
  const int64_t n_cells = (int64_t)op->kp.nc[0] * op->kp.nc[1] * op->kp.nc[2];
  HPS_HIP(hipMemcpyAsync(masks, op->cell_masks, n_cells * sizeof(unsigned),
                         hipMemcpyDeviceToDevice, op->stream));
  return HPS_OK;
}

extern "C" int hps_set_linearization(hps_operator op, const double *u)
{
  if (int r = check_op(op))
    return r;
  if (!u)
    return fail(HPS_ERR_INVALID_ARGUMENT, "null vector");
  if (!op->cache64)
    {
      HPS_HIP(hipMalloc(&op->cache64, op->cache_elems * sizeof(double)));
      HPS_HIP(hipMalloc(&op->cache32, op->cache_elems * sizeof(float)));
    }
  if (op->q2)
    op->q2->fill(q2_params(op), u, op->cache64, op->cache32, op->stream);
  else if (op->cut_active)
    op->ops->fill_cut(op->kp, op->n_tiles, u, op->cache64, op->cache32, op->stream);
  else
    op->ops->fill(op->kp, op->n_tiles, u, op->cache64, op->cache32, op->stream);
  HPS_LAUNCHED();
  op->lin_set = true;
  return HPS_OK;
}

// sums the tile-face partials of the Q1 kernels (the Q2 colour passes
// complete dst themselves); zcol_tiles: of the z-column kernel, whose
// marches are zcol_mz layers deep if zcol_wide
template <typename T>
static int finish_fixup(hps_operator op, T *dst, int ostride, int ocstride, int ooffset,
                        bool zcol_tiles = false, bool zcol_wide = false)
{
  if (op->q2)
    return HPS_OK;
  if (zcol_tiles)
    {
      hps::fixup_zcol_d3(zcol_params(op, zcol_wide), op->ntz1, op->nc, (const T *)op->part, dst,
                         op->stream);
      HPS_LAUNCHED();
      return HPS_OK;
    }
  hps::FixParams fp = op->fp;
  fp.ostride        = ostride;
  fp.ocstride       = ocstride;
  fp.ooffset        = ooffset;
  hps::launch_fixup<T>(fp, (const T *)op->part, dst, op->stream);
  HPS_LAUNCHED();
  return HPS_OK;
}

extern "C" int hps_residual(hps_operator op, int wtd, double *dst, const double *src,
                            const double *const *old)
{
  if (int r = check_op(op))
    return r;
  if (!dst || !src || wtd < 0 || wtd > 2)
    return fail(HPS_ERR_INVALID_ARGUMENT, "residual: bad arguments");
  const double *o[3] = {nullptr, nullptr, nullptr};
  if (wtd == 2)
    {
      if (op->kp.order > 0 && !old)
        return fail(HPS_ERR_INVALID_ARGUMENT, "residual: history vectors missing");
      for (int i = 0; i < op->kp.order; ++i)
        {
          if (!old[i])
            return fail(HPS_ERR_INVALID_ARGUMENT, "residual: null history vector");
          o[i] = old[i];
        }
    }
  if (op->timing)
    HPS_HIP(hipEventRecord(timing_event(op, 0), op->stream));
  // with order 0 there is no history: evaluate without the history term
  const int w = (wtd == 2 && op->kp.order == 0) ? 1 : wtd;
  const bool zres = op->zcol && !op->cut_active && op->zcol_residual;
  if (op->q2)
    op->q2->residual(q2_params(op), w, src, o[0], o[1], o[2], dst, op->stream);
  else if (zres)
    {
      const hps::KParams zk = zcol_params(op, true);
      hps::residual_zcol_d3(op->n_grains, zk, op->ntz1, zcol_tiles(op, zk), w, src, o[0], o[1],
                            o[2], dst, op->part, op->stream);
    }
  else
    (op->cut_active ? op->ops->residual_cut : op->ops->residual)(
      op->kp, op->n_tiles, w, src, o[0], o[1], o[2], dst, op->part, op->stream);
  HPS_LAUNCHED();
  if (op->timing)
    HPS_HIP(hipEventRecord(timing_event(op, 1), op->stream));
  if (int r = finish_fixup<double>(op, dst, op->nc, 1, 0, zres, true))
    return r;
  if (op->timing)
    HPS_HIP(hipEventRecord(timing_event(op, 2), op->stream));
  // FEEvaluation::distribute_local_to_global leaves constrained entries
  // untouched: their residual rows are zero
  if (op->n_constrained)
    {
      hps::constraint_set(op->stream, op->n_constrained, op->nc, op->constrained, dst, 0.0);
      HPS_LAUNCHED();
    }
  return HPS_OK;
}

extern "C" int hps_set_constrained_nodes(hps_operator op, const int64_t *nodes, int64_t n)
{
  if (int r = check_op(op))
    return r;
  if (n < 0 || (n > 0 && !nodes))
    return fail(HPS_ERR_INVALID_ARGUMENT, "constrained nodes: bad arguments");
  std::vector<int64_t> v(nodes, nodes + n);
  std::sort(v.begin(), v.end());
  for (int64_t i = 0; i < n; ++i)
    {
      if (v[i] < 0 || v[i] >= op->n_nodes)
        return fail(HPS_ERR_INVALID_ARGUMENT, "constrained node index out of range");
      if (i > 0 && v[i] == v[i - 1])
        return fail(HPS_ERR_INVALID_ARGUMENT, "constrained node listed twice");
    }
  HPS_HIP(hipStreamSynchronize(op->stream));
  // the cut-off masks and tile groups depend on the linearisation point
  // only, not on the constraints: they stay
  (void)hipFree(op->constrained);
  (void)hipFree(op->constrained_save);
  op->constrained      = nullptr;
  op->constrained_save = nullptr;
  op->n_constrained    = 0;
  if (n == 0)
    return HPS_OK;
  HPS_HIP(hipMalloc(&op->constrained, n * sizeof(int64_t)));
  HPS_HIP(hipMalloc(&op->constrained_save, n * op->nc * sizeof(double)));
  HPS_HIP(hipMemcpy(op->constrained, v.data(), n * sizeof(int64_t), hipMemcpyHostToDevice));
  op->n_constrained = n;
  return HPS_OK;
}

extern "C" int hps_vmult_f64(hps_operator op, double *dst, const double *src)
{
  if (int r = check_op(op))
    return r;
  if (!dst || !src)
    return fail(HPS_ERR_INVALID_ARGUMENT, "vmult: null vector");
  if (!op->lin_set)
    return fail(HPS_ERR_NOT_INITIALIZED, "vmult before hps_set_linearization");
  // constrained entries of src: saved and zeroed, restored after the
  // operator (src is modified in between, in stream order)
  if (op->n_constrained)
    {
      hps::constraint_zero64(op->stream, op->n_constrained, op->nc, op->constrained,
                             const_cast<double *>(src), op->constrained_save);
      HPS_LAUNCHED();
    }
  if (op->timing)
    HPS_HIP(hipEventRecord(timing_event(op, 0), op->stream));
  if (op->q2)
    op->q2->vmult64(q2_params(op), src, dst, op->cache64, op->stream);
  else if (op->cut_ready)
    for (int g = 0; g <= hps::CUT_MAX_GA; ++g)
      {
        const int count = op->cut_off[g + 1] - op->cut_off[g];
        if (count > 0)
          hps::vmult_cut_d3(g, op->kp, op->n_grains, op->cut_tiles + op->cut_off[g], count,
                            op->cut_masks, src, op->cache64, dst, op->part, op->stream);
      }
  else if (op->zcol && !op->cut_active)
    {
      const hps::KParams zk = zcol_params(op, true);
      hps::vmult_zcol_d3(op->n_grains, zk, op->ntz1, zcol_tiles(op, zk), src, op->cache64, dst,
                         op->part, op->stream);
    }
  else
    (op->cut_active ? op->ops->vmult64_cut : op->ops->vmult64)(op->kp, op->n_tiles, src, dst,
                                                               op->part, op->cache64,
                                                               op->stream);
  HPS_LAUNCHED();
  if (op->timing)
    HPS_HIP(hipEventRecord(timing_event(op, 1), op->stream));
  if (int r = finish_fixup<double>(op, dst, op->nc, 1, 0,
                                   !op->cut_ready && op->zcol && !op->cut_active, true))
    return r;
  if (op->timing)
    HPS_HIP(hipEventRecord(timing_event(op, 2), op->stream));
  if (op->n_constrained)
    {
      hps::constraint_restore64(op->stream, op->n_constrained, op->nc, op->constrained,
                                const_cast<double *>(src), dst, op->constrained_save);
      HPS_LAUNCHED();
    }
  return HPS_OK;
}

extern "C" int hps_vmult_f32(hps_operator op, float *dst, const float *src)
{
  if (int r = check_op(op))
    return r;
  if (!dst || !src)
    return fail(HPS_ERR_INVALID_ARGUMENT, "vmult: null vector");
  if (!op->lin_set)
    return fail(HPS_ERR_NOT_INITIALIZED, "vmult before hps_set_linearization");
  float *save32 = reinterpret_cast<float *>(op->constrained_save);
  if (op->n_constrained)
    {
      hps::constraint_zero32(op->stream, op->n_constrained, op->nc, op->constrained,
                             const_cast<float *>(src), save32);
      HPS_LAUNCHED();
    }
  if (op->timing)
    HPS_HIP(hipEventRecord(timing_event(op, 0), op->stream));
  if (op->q2)
    op->q2->vmult32(q2_params(op), src, dst, op->cache32, op->stream);
  else if (op->cut_ready)
    for (int g = 0; g <= hps::CUT_MAX_GA; ++g)
      {
        const int count = op->cut_off[g + 1] - op->cut_off[g];
        if (count > 0)
          hps::vmult_cut_d3(g, op->kp, op->n_grains, op->cut_tiles + op->cut_off[g], count,
                            op->cut_masks, src, op->cache32, dst, (float *)op->part,
                            op->stream);
      }
  else if (op->zcol32 && !op->cut_active)
    {
      const hps::KParams zk = zcol_params(op, false);
      hps::vmult_zcol_d3(op->n_grains, zk, op->ntz1, zcol_tiles(op, zk), src, op->cache32, dst,
                         (float *)op->part, op->stream);
    }
  else
    (op->cut_active ? op->ops->vmult32_cut : op->ops->vmult32)(
      op->kp, op->n_tiles, src, dst, (float *)op->part, op->cache32, op->stream);
  HPS_LAUNCHED();
  if (op->timing)
    HPS_HIP(hipEventRecord(timing_event(op, 1), op->stream));
  if (int r = finish_fixup<float>(op, dst, op->nc, 1, 0,
                                  !op->cut_ready && op->zcol32 && !op->cut_active))
    return r;
  if (op->timing)
    HPS_HIP(hipEventRecord(timing_event(op, 2), op->stream));
  if (op->n_constrained)
    {
      hps::constraint_restore32(op->stream, op->n_constrained, op->nc, op->constrained,
                                const_cast<float *>(src), dst, save32);
      HPS_LAUNCHED();
    }
  return HPS_OK;
}

// geometry, cache layout and material of the Cahn-Hilliard ILU kernels
static hps::ChParams ch_params(hps_operator op)
{
  hps::ChParams cp{};
  cp.dim  = op->dim;
  cp.nc[0] = op->p.n_cells[0];
  cp.nc[1] = op->p.n_cells[1];
  cp.nc[2] = op->dim == 3 ? op->p.n_cells[2] : 1;
  for (int d = 0; d < 3; ++d)
    {
      cp.nn[d]    = d < op->dim ? cp.nc[d] + 1 : 1;
      cp.nt[d]    = op->kp.nt[d];
      cp.inv_h[d] = op->kp.inv_h[d];
    }
  cp.n_nodes  = op->n_nodes;
  cp.n_comp   = op->nc;
  cp.G        = op->n_grains;
  cp.tens     = op->tens ? 1 : 0;
  cp.tx       = op->ops->tx;
  cp.ty       = op->ops->ty;
  cp.tz       = op->kp.tz;
  cp.nt_lanes = op->ops->nt;
  cp.nf       = op->ops->n_fields;
  cp.jxw      = op->kp.jxw;
  cp.pp       = op->kp.pp;
  hps::ch_layout(cp);
  return cp;
}

// diagonal / node blocks assembled by probing (FP64 cache and arithmetic,
// or FP32 for the FP32 inner solve), then inverted in FP64
static int precond_update(hps_operator op, int kind, bool fp32)
{
  if (int r = check_op(op))
    return r;
  if (!op->lin_set)
    return fail(HPS_ERR_NOT_INITIALIZED, "preconditioner update before hps_set_linearization");
  const int64_t n = op->n_dofs, nn = op->n_nodes;
  const int     nc = op->nc;
  if (kind == HPS_PRECON_BLOCK_CH_ILU)
    {
      if (op->q2)
        return fail(HPS_ERR_UNSUPPORTED, "the Cahn-Hilliard ILU preconditioner is built for Q1");
      // the Allen-Cahn block: the point-Jacobi inverse diagonal
      if (int r = precond_update(op, HPS_PRECON_POINT_JACOBI, fp32))
        return r;
      // the factors and the solves' work vector
      const int64_t nA = hps::ch_storage(ch_params(op));
      if (!op->ch64)
        {
          HPS_HIP(hipMalloc(&op->ch64, nA * sizeof(double)));
          HPS_HIP(hipMalloc(&op->ch32, nA * sizeof(float)));
        }
      hps::ch_assemble(ch_params(op), op->cache64, op->ch64, op->stream);
      HPS_LAUNCHED();
      if (op->n_constrained)
        {
          hps::ch_constrain(ch_params(op), op->constrained, op->n_constrained, op->ch64,
                            op->stream);
          HPS_LAUNCHED();
        }
      hps::ch_factor(ch_params(op), op->ch64, op->ch32, op->stream);
      HPS_LAUNCHED();
      op->precond_kind = kind;
      return HPS_OK;
    }
  if (kind == HPS_PRECON_POINT_JACOBI)
    {
      if (!op->diag)
        {
          HPS_HIP(hipMalloc(&op->diag, n * sizeof(double)));
          HPS_HIP(hipMalloc(&op->inv_diag64, n * sizeof(double)));
          HPS_HIP(hipMalloc(&op->inv_diag32, n * sizeof(float)));
        }
      if (op->q2 && fp32)
        op->q2->probe32(q2_params(op), op->cache32, op->inv_diag32, -1, nc, 1, 0, op->stream);
      else if (op->q2)
        op->q2->probe(q2_params(op), op->cache64, op->diag, -1, nc, 1, 0, op->stream);
      else if (fp32)
        (op->cut_active ? op->ops->probe32_cut : op->ops->probe32)(
          op->kp, op->n_tiles, op->cache32, op->inv_diag32, (float *)op->part, -1, nc, 1, 0,
          op->stream);
      else
        (op->cut_active ? op->ops->probe_cut : op->ops->probe)(
          op->kp, op->n_tiles, op->cache64, op->diag, op->part, -1, nc, 1, 0, op->stream);
      HPS_LAUNCHED();
      if (fp32)
        {
          if (int r = finish_fixup<float>(op, op->inv_diag32, nc, 1, 0))
            return r;
          hps::blas_f32_to_f64(op->stream, n, op->diag, op->inv_diag32);
        }
      else if (int r = finish_fixup<double>(op, op->diag, nc, 1, 0))
        return r;
      // identity rows of the constrained nodes: diagonal 1
      if (op->n_constrained)
        hps::constraint_set(op->stream, op->n_constrained, nc, op->constrained, op->diag, 1.0);
      hps::invert_diag(op->stream, n, op->diag, op->inv_diag64, op->inv_diag32);
      HPS_LAUNCHED();
    }
  else if (kind == HPS_PRECON_NODE_BLOCK_JACOBI)
    {
      if (!op->blocks)
        {
          HPS_HIP(hipMalloc(&op->blocks, nn * nc * nc * sizeof(double)));
          const int64_t ni = hps::inv_stride(nn) * nc * nc;
          HPS_HIP(hipMalloc(&op->inv_blocks64, ni * sizeof(double)));
          HPS_HIP(hipMalloc(&op->inv_blocks32, ni * sizeof(float)));
          // the padding nodes past nn are read by the FP32 apply (two nodes
          // per lane) and their results dropped, so their values do not
          // matter (the FP32 assembly stages its blocks in inv_blocks32);
          // they start as zeros
          HPS_HIP(hipMemsetAsync(op->inv_blocks64, 0, ni * sizeof(double), op->stream));
          HPS_HIP(hipMemsetAsync(op->inv_blocks32, 0, ni * sizeof(float), op->stream));
        }
      // the assembled blocks are stored per node column by column
      // (blocks[node][column][row]): a block column is a contiguous run
      if (op->q2) // all block columns in one pass over the cache
        {
          if (fp32)
            op->q2->probe32(q2_params(op), op->cache32, op->inv_blocks32, -2, nc * nc, 1, 0,
                            op->stream);
          else
            op->q2->probe(q2_params(op), op->cache64, op->blocks, -2, nc * nc, 1, 0,
                          op->stream);
          HPS_LAUNCHED();
        }
      for (int j = 0; j < nc && !op->q2; ++j)
        {
          if (fp32)
            (op->cut_active ? op->ops->probe32_cut : op->ops->probe32)(
              op->kp, op->n_tiles, op->cache32, op->inv_blocks32, (float *)op->part, j,
              nc * nc, 1, j * nc, op->stream);
          else
            (op->cut_active ? op->ops->probe_cut : op->ops->probe)(
              op->kp, op->n_tiles, op->cache64, op->blocks, op->part, j, nc * nc, 1, j * nc,
              op->stream);
          HPS_LAUNCHED();
          if (fp32)
            {
              if (int r = finish_fixup<float>(op, op->inv_blocks32, nc * nc, 1, j * nc))
                return r;
            }
          else if (int r = finish_fixup<double>(op, op->blocks, nc * nc, 1, j * nc))
            return r;
        }
      if (fp32)
        hps::blas_f32_to_f64(op->stream, nn * nc * nc, op->blocks, op->inv_blocks32);
      // cut-off: unit diagonal for the components without any entry
      if (op->cut_active && !hps::unit_empty_rows(op->stream, nn, nc, op->blocks))
        return fail(HPS_ERR_INVALID_N_COMP, "node-block Jacobi not built for this n_components");
      // identity rows and columns of the constrained nodes: identity blocks
      if (op->n_constrained)
        hps::constraint_identity_blocks(op->stream, op->n_constrained, nc, op->constrained,
                                        op->blocks);
      HPS_HIP(hipMemsetAsync(op->d_singular, 0, sizeof(int), op->stream));
      if (!hps::invert_blocks(op->stream, nn, nc, op->blocks, op->inv_blocks64,
                              op->inv_blocks32, op->d_singular))
        return fail(HPS_ERR_INVALID_N_COMP, "node-block Jacobi not built for this n_components");
      HPS_LAUNCHED();
      int singular = 0;
      HPS_HIP(hipMemcpyAsync(&singular, op->d_singular, sizeof(int), hipMemcpyDeviceToHost,
                             op->stream));
      HPS_HIP(hipStreamSynchronize(op->stream));
      if (singular)
        return fail(HPS_ERR_INVALID_ARGUMENT,
                    std::to_string(singular) + " singular node blocks");
    }
  else
    return fail(HPS_ERR_INVALID_ARGUMENT, "unknown preconditioner kind");
  op->precond_kind = kind;
  return HPS_OK;
}

extern "C" int hps_precond_update(hps_operator op, int kind)
{
  return precond_update(op, kind, false);
}

extern "C" int hps_precond_update_f32(hps_operator op, int kind)
{
  return precond_update(op, kind, true);
}

extern "C" int hps_precond_apply_f64(hps_operator op, double *dst, const double *src)
{
  if (int r = check_op(op))
    return r;
  if (op->precond_kind == HPS_PRECON_BLOCK_CH_ILU)
    hps::ch_apply<double>(ch_params(op), op->ch64, op->inv_diag64, src, dst, op->stream);
  else if (op->precond_kind == HPS_PRECON_POINT_JACOBI)
    hps::diag_apply64(op->stream, op->n_dofs, op->inv_diag64, src, dst);
  else if (op->precond_kind == HPS_PRECON_NODE_BLOCK_JACOBI)
    hps::block_apply64(op->stream, op->n_nodes, op->nc, op->inv_blocks64, src, dst);
  else
    return fail(HPS_ERR_NOT_INITIALIZED, "preconditioner not set up");
  HPS_LAUNCHED();
  return HPS_OK;
}

extern "C" int hps_precond_apply_f32(hps_operator op, float *dst, const float *src)
{
  if (int r = check_op(op))
    return r;
  if (op->precond_kind == HPS_PRECON_BLOCK_CH_ILU)
    hps::ch_apply<float>(ch_params(op), op->ch32, op->inv_diag32, src, dst, op->stream);
  else if (op->precond_kind == HPS_PRECON_POINT_JACOBI)
    hps::diag_apply32(op->stream, op->n_dofs, op->inv_diag32, src, dst);
  else if (op->precond_kind == HPS_PRECON_NODE_BLOCK_JACOBI)
    hps::block_apply32(op->stream, op->n_nodes, op->nc, op->inv_blocks32, src, dst);
  else
    return fail(HPS_ERR_NOT_INITIALIZED, "preconditioner not set up");
  HPS_LAUNCHED();
  return HPS_OK;
}

extern "C" int hps_precond_get_diagonal(hps_operator op, double *inv_diag)
{
  if (int r = check_op(op))
    return r;
  if (!op->inv_diag64)
    return fail(HPS_ERR_NOT_INITIALIZED, "no diagonal computed");
  HPS_HIP(hipMemcpyAsync(inv_diag, op->inv_diag64, op->n_dofs * sizeof(double),
                         hipMemcpyDeviceToDevice, op->stream));
  return HPS_OK;
}

extern "C" int hps_precond_get_blocks(hps_operator op, double *blocks)
{
  if (int r = check_op(op))
    return r;
  if (!op->blocks)
    return fail(HPS_ERR_NOT_INITIALIZED, "no node blocks computed");
  // stored column by column per node; returned row by row
  hps::transpose_blocks(op->stream, op->n_nodes, op->nc, op->blocks, blocks);
  HPS_LAUNCHED();
  return HPS_OK;
}

// ------------------------------------------------------------------ BLAS-1
static int fetch(hps_operator op, int k, double *result)
{
  HPS_LAUNCHED();
  HPS_HIP(hipMemcpyAsync(op->red_host, op->red_out, k * sizeof(double), hipMemcpyDeviceToHost,
                         op->stream));
  HPS_HIP(hipStreamSynchronize(op->stream));
  for (int i = 0; i < k; ++i)
    result[i] = op->red_host[i];
  return HPS_OK;
}

extern "C" int hps_dot(hps_operator op, int64_t n, const double *x, const double *y,
                       double *result)
{
  if (int r = check_op(op))
    return r;
  hps::blas_dot(op->stream, n, x, y, op->red_partial, op->red_out);
  return fetch(op, 1, result);
}

extern "C" int hps_norm2(hps_operator op, int64_t n, const double *x, double *result)
{
  if (int r = check_op(op))
    return r;
  hps::blas_norm2(op->stream, n, x, op->red_partial, op->red_out);
  return fetch(op, 1, result);
}

extern "C" int hps_block_norms(hps_operator op, const double *x, double *result)
{
  if (int r = check_op(op))
    return r;
  hps::blas_block_norms(op->stream, op->n_nodes, op->nc, x, op->red_partial, op->red_out);
  return fetch(op, op->nc, result);
}

extern "C" int hps_multi_dot(hps_operator op, int64_t n, int k, const double *V,
                             const double *w, double *h)
{
  if (int r = check_op(op))
    return r;
  if (k < 1 || k > hps::MAXK)
    return fail(HPS_ERR_INVALID_ARGUMENT, "multi_dot: k must be 1..32");
  hps::blas_multi_dot(op->stream, n, k, V, w, op->red_partial, op->red_out);
  return fetch(op, k, h);
}

extern "C" int hps_multi_axpy_norm(hps_operator op, int64_t n, int k, const double *V,
                                   const double *h, double *w, double *norm)
{
  if (int r = check_op(op))
    return r;
  if (k < 1 || k > hps::MAXK)
    return fail(HPS_ERR_INVALID_ARGUMENT, "multi_axpy_norm: k must be 1..32");
  // coefficients go through the pinned buffer into device memory
  for (int j = 0; j < k; ++j)
    op->red_host[32 + j] = h[j];
  HPS_HIP(hipMemcpyAsync(op->red_out + 32, op->red_host + 32, k * sizeof(double),
                         hipMemcpyHostToDevice, op->stream));
  hps::blas_multi_axpy_norm(op->stream, n, k, V, op->red_out + 32, w, op->red_partial,
                            op->red_out);
  return fetch(op, 1, norm);
}

extern "C" int hps_axpy(hps_operator op, int64_t n, double a, const double *x, double *y)
{
  if (int r = check_op(op))
    return r;
  hps::blas_axpy(op->stream, n, a, x, y);
  HPS_LAUNCHED();
  return HPS_OK;
}

extern "C" int hps_sadd(hps_operator op, int64_t n, double s, double a, const double *x,
                        double *y)
{
  if (int r = check_op(op))
    return r;
  hps::blas_sadd(op->stream, n, s, a, x, y);
  HPS_LAUNCHED();
  return HPS_OK;
}

extern "C" int hps_scale(hps_operator op, int64_t n, double a, double *x)
{
  if (int r = check_op(op))
    return r;
  hps::blas_scale(op->stream, n, a, x);
  HPS_LAUNCHED();
  return HPS_OK;
}

extern "C" int hps_copy(hps_operator op, int64_t n, double *dst, const double *src)
{
  if (int r = check_op(op))
    return r;
  HPS_HIP(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDeviceToDevice, op->stream));
  return HPS_OK;
}

extern "C" int hps_fill(hps_operator op, int64_t n, double value, double *x)
{
  if (int r = check_op(op))
    return r;
  hps::blas_fill(op->stream, n, value, x);
  HPS_LAUNCHED();
  return HPS_OK;
}

extern "C" int hps_convert_f64_to_f32(hps_operator op, int64_t n, float *dst, const double *src)
{
  if (int r = check_op(op))
    return r;
  hps::blas_f64_to_f32(op->stream, n, dst, src);
  HPS_LAUNCHED();
  return HPS_OK;
}

extern "C" int hps_convert_f32_to_f64(hps_operator op, int64_t n, double *dst, const float *src)
{
  if (int r = check_op(op))
    return r;
  hps::blas_f32_to_f64(op->stream, n, dst, src);
  HPS_LAUNCHED();
  return HPS_OK;
}

static int push_ptrs(hps_operator op, const void *const *ptrs)
{
  for (int c = 0; c < op->nc; ++c)
    {
      if (!ptrs[c])
        return fail(HPS_ERR_INVALID_ARGUMENT, "null block pointer");
      op->ptrs_host[c] = (double *)ptrs[c];
    }
  HPS_HIP(hipMemcpyAsync(op->ptrs_dev, op->ptrs_host, op->nc * sizeof(double *),
                         hipMemcpyHostToDevice, op->stream));
  return HPS_OK;
}

extern "C" int hps_from_blocks(hps_operator op, double *x, const double *const *blocks)
{
  if (int r = check_op(op))
    return r;
  if (!x || !blocks)
    return fail(HPS_ERR_INVALID_ARGUMENT, "null argument");
  if (int r = push_ptrs(op, (const void *const *)blocks))
    return r;
  hps::from_blocks(op->stream, op->n_nodes, op->nc, x, (const double *const *)op->ptrs_dev);
  HPS_LAUNCHED();
  HPS_HIP(hipStreamSynchronize(op->stream));
  return HPS_OK;
}

extern "C" int hps_to_blocks(hps_operator op, double *const *blocks, const double *x)
{
  if (int r = check_op(op))
    return r;
  if (!x || !blocks)
    return fail(HPS_ERR_INVALID_ARGUMENT, "null argument");
  if (int r = push_ptrs(op, (const void *const *)blocks))
    return r;
  hps::to_blocks(op->stream, op->n_nodes, op->nc, op->ptrs_dev, x);
  HPS_LAUNCHED();
  HPS_HIP(hipStreamSynchronize(op->stream));
  return HPS_OK;
}

extern "C" int hps_enable_timing(hps_operator op, int enable)
{
  if (int r = check_op(op))
    return r;
  if (enable < 0 || enable > 4096)
    return fail(HPS_ERR_INVALID_ARGUMENT, "timing ring: 0..4096 calls");
  HPS_HIP(hipStreamSynchronize(op->stream));
  op->timing = enable != 0;
  op->tcount = 0;
  if (!op->timing)
    return HPS_OK;
  while ((int)op->tev->size() < 3 * enable)
    {
      hipEvent_t ev;
      HPS_HIP(hipEventCreate(&ev));
      op->tev->push_back(ev);
    }
  op->tcap = enable;
  return HPS_OK;
}

extern "C" int hps_last_kernel_ms(hps_operator op, float *ms_main, float *ms_fixup)
{
  if (int r = check_op(op))
    return r;
  if (!op->timing || op->tcount == 0)
    return fail(HPS_ERR_NOT_INITIALIZED, "timing not enabled or nothing timed");
  const int   slot = (op->tcount - 1) % op->tcap;
  hipEvent_t *e    = op->tev->data() + 3 * slot;
  HPS_HIP(hipEventSynchronize(e[2]));
  float a = 0.f, b = 0.f;
  HPS_HIP(hipEventElapsedTime(&a, e[0], e[1]));
  HPS_HIP(hipEventElapsedTime(&b, e[1], e[2]));
  if (ms_main)
    *ms_main = a;
  if (ms_fixup)
    *ms_fixup = b;
  return HPS_OK;
}

extern "C" int hps_kernel_ms_mean(hps_operator op, float *ms_main, float *ms_fixup, int *count)
{
  if (int r = check_op(op))
    return r;
  if (!op->timing || op->tcount == 0)
    return fail(HPS_ERR_NOT_INITIALIZED, "timing not enabled or nothing timed");
  const int n = op->tcount < op->tcap ? op->tcount : op->tcap;
  HPS_HIP(hipStreamSynchronize(op->stream));
  double sa = 0.0, sb = 0.0;
  for (int k = 0; k < n; ++k)
    {
      const int   slot = (op->tcount - 1 - k) % op->tcap;
      hipEvent_t *e    = op->tev->data() + 3 * slot;
      float       a = 0.f, b = 0.f;
      HPS_HIP(hipEventElapsedTime(&a, e[0], e[1]));
      HPS_HIP(hipEventElapsedTime(&b, e[1], e[2]));
      sa += a;
      sb += b;
    }
  if (ms_main)
    *ms_main = (float)(sa / n);
  if (ms_fixup)
    *ms_fixup = (float)(sb / n);
  if (count)
    *count = n;
  op->tcount = 0;
  return HPS_OK;
}

// ------------------------------------------------------------------
// learned Newton predictor kernels (no counterpart in the reference; the
// predictor slot is S/driver.h:2002-2050)
// ------------------------------------------------------------------
extern "C" int hps_mlp_gemm(hps_operator op, int trans_a, int trans_b, int epilogue, int M,
                            int N, int K, const void *A, int lda, const void *B, int ldb,
                            void *C, int ldc, const float *bias, const void *H, int n_split)
{
  if (int r = check_op(op))
    return r;
  if (M < 0 || N < 0 || K < 0 || n_split < 1 || !A || !B || !C)
    return fail(HPS_ERR_INVALID_ARGUMENT, "mlp_gemm: bad arguments");
  if ((epilogue == hps::EPI_BIAS_TANH || epilogue == hps::EPI_BIAS_F32) && !bias)
    return fail(HPS_ERR_INVALID_ARGUMENT, "mlp_gemm: bias missing");
  if (epilogue == hps::EPI_DTANH && !H)
    return fail(HPS_ERR_INVALID_ARGUMENT, "mlp_gemm: activation missing");
  if (n_split > 1 && epilogue != hps::EPI_F32)
    return fail(HPS_ERR_INVALID_ARGUMENT, "mlp_gemm: split-K needs the FP32 epilogue");
  const bool ok = (trans_a == 0 && trans_b == 1) || (trans_a == 0 && trans_b == 0 &&
                   (epilogue == hps::EPI_DTANH || epilogue == hps::EPI_F32)) ||
                  (trans_a == 1 && epilogue == hps::EPI_F32);
  if (!ok)
    return fail(HPS_ERR_UNSUPPORTED, "mlp_gemm: variant not built");
  hps::mlp_gemm(op->stream, trans_a, trans_b, epilogue, M, N, K, (const uint16_t *)A, lda,
                (const uint16_t *)B, ldb, C, ldc, bias, (const uint16_t *)H, n_split,
                (int64_t)M * ldc);
  HPS_LAUNCHED();
  return HPS_OK;
}

extern "C" int hps_mlp_sum_splits(hps_operator op, int64_t n, int n_split, const float *part,
                                  float *out)
{
  if (int r = check_op(op))
    return r;
  hps::mlp_sum_splits(op->stream, n, n_split, part, out);
  HPS_LAUNCHED();
  return HPS_OK;
}

extern "C" int hps_mlp_colsum(hps_operator op, int M, int N, const void *X, float *out)
{
  if (int r = check_op(op))
    return r;
  hps::mlp_colsum(op->stream, M, N, (const uint16_t *)X, out);
  HPS_LAUNCHED();
  return HPS_OK;
}

extern "C" int hps_mlp_loss(hps_operator op, int64_t n, const float *Y, const float *T,
                            void *dY, float *partial, int n_blocks, float *loss)
{
  if (int r = check_op(op))
    return r;
  if (n_blocks < 1 || n_blocks > 2 * RED_HOST_DOUBLES)
    return fail(HPS_ERR_INVALID_ARGUMENT, "mlp_loss: n_blocks must be 1..8192");
  hps::mlp_loss(op->stream, n, Y, T, (uint16_t *)dY, partial, n_blocks);
  HPS_LAUNCHED();
  if (loss)
    {
      float *h = (float *)op->red_host;
      HPS_HIP(hipMemcpyAsync(h, partial, n_blocks * sizeof(float), hipMemcpyDeviceToHost,
                             op->stream));
      HPS_HIP(hipStreamSynchronize(op->stream));
      double s = 0.0;
      for (int i = 0; i < n_blocks; ++i)
        s += h[i];
      *loss = (float)s;
    }
  return HPS_OK;
}

extern "C" int hps_mlp_adam(hps_operator op, int n_out, int n_in, float *W, float *m1,
                            float *m2, const float *grad, float lr, float beta1, float beta2,
                            float eps, int step, void *W_bf16)
{
  if (int r = check_op(op))
    return r;
  if (step < 1)
    return fail(HPS_ERR_INVALID_ARGUMENT, "mlp_adam: step counts from 1");
  const float c1 = 1.f - std::pow(beta1, (float)step);
  const float c2 = 1.f - std::pow(beta2, (float)step);
  hps::mlp_adam(op->stream, n_out, n_in, W, m1, m2, grad, lr, beta1, beta2, eps, c1, c2,
                (uint16_t *)W_bf16, nullptr);
  HPS_LAUNCHED();
  return HPS_OK;
}

extern "C" int hps_mlp_features(hps_operator op, const double *u_prev, const double *u_cur,
                                double dt, double dt_scale, double inc_scale, void *X, int *top2,
                                float *lscale)
{
  if (int r = check_op(op))
    return r;
  hps::FeatParams fp;
  fp.n_nodes = op->n_nodes;
  for (int d = 0; d < 3; ++d)
    fp.nn[d] = 1;
  fp.nn[0] = op->kp.nn[0];
  if (op->dim == 3)
    {
      fp.nn[1] = op->kp.nn[1];
      fp.nn[2] = op->kp.nn[2];
    }
  else
    fp.nn[1] = op->kp.nn[2];
  fp.n_comp    = op->nc;
  fp.dt        = dt;
  fp.dt_scale  = dt_scale;
  fp.inc_scale = inc_scale;
  fp.inc_floor = inc_scale > 0.0 ? 1e-3 / inc_scale : 0.0;
  hps::mlp_features(op->stream, fp, u_prev, u_cur, (uint16_t *)X, top2, lscale);
  HPS_LAUNCHED();
  return HPS_OK;
}

extern "C" int hps_mlp_predict(hps_operator op, const double *u_prev, const double *u_cur,
                               const float *Y, const int *top2, const float *out_scale,
                               const float *lscale, double *u_pred)
{
  if (int r = check_op(op))
    return r;
  hps::mlp_predict(op->stream, op->n_nodes, op->nc, u_prev, u_cur, Y, top2, out_scale, lscale,
                   u_pred);
  HPS_LAUNCHED();
  return HPS_OK;
}

extern "C" int hps_mlp_targets(hps_operator op, const double *u_prev, const double *u_cur,
                               const double *u_next, const int *top2, const float *out_scale,
                               const float *lscale, float *T)
{
  if (int r = check_op(op))
    return r;
  hps::mlp_targets(op->stream, op->n_nodes, op->nc, u_prev, u_cur, u_next, top2, out_scale,
                   lscale, T);
  HPS_LAUNCHED();
  return HPS_OK;
}

// ---- FP32-storage vectors with FP64 accumulation (mixed-precision Krylov)
extern "C" int hps_norm2_f32(hps_operator op, int64_t n, const float *x, double *result)
{
  if (int r = check_op(op))
    return r;
  hps::blas_norm2_f32(op->stream, n, x, op->red_partial, op->red_out);
  return fetch(op, 1, result);
}

extern "C" int hps_multi_dot_f32(hps_operator op, int64_t n, int k, const float *V,
                                 const float *w, double *h)
{
  if (int r = check_op(op))
    return r;
  if (k < 1 || k > hps::MAXK)
    return fail(HPS_ERR_INVALID_ARGUMENT, "multi_dot_f32: k must be 1..32");
  hps::blas_multi_dot_f32(op->stream, n, k, V, w, op->red_partial, op->red_out);
  return fetch(op, k, h);
}

extern "C" int hps_multi_axpy_norm_f32(hps_operator op, int64_t n, int k, const float *V,
                                       const double *h, float *w, double *norm)
{
  if (int r = check_op(op))
    return r;
  if (k < 1 || k > hps::MAXK)
    return fail(HPS_ERR_INVALID_ARGUMENT, "multi_axpy_norm_f32: k must be 1..32");
  for (int j = 0; j < k; ++j)
    op->red_host[32 + j] = h[j];
  HPS_HIP(hipMemcpyAsync(op->red_out + 32, op->red_host + 32, k * sizeof(double),
                         hipMemcpyHostToDevice, op->stream));
  hps::blas_multi_axpy_norm_f32(op->stream, n, k, V, op->red_out + 32, w, op->red_partial,
                                op->red_out);
  return fetch(op, 1, norm);
}

extern "C" int hps_scale_f32(hps_operator op, int64_t n, double a, float *x)
{
  if (int r = check_op(op))
    return r;
  hps::blas_scale_f32(op->stream, n, a, x);
  HPS_LAUNCHED();
  return HPS_OK;
}

extern "C" int hps_axpy_multi(hps_operator op, int64_t n, int k, const double *a,
                              const double *X, double *y)
{
  if (int r = check_op(op))
    return r;
  if (k < 1 || k > hps::MAXK)
    return fail(HPS_ERR_INVALID_ARGUMENT, "axpy_multi: k must be 1..32");
  for (int j = 0; j < k; ++j)
    op->red_host[32 + j] = a[j];
  HPS_HIP(hipMemcpyAsync(op->red_out + 32, op->red_host + 32, k * sizeof(double),
                         hipMemcpyHostToDevice, op->stream));
  hps::blas_axpy_multi_f64(op->stream, n, k, op->red_out + 32, X, y);
  HPS_LAUNCHED();
  // the coefficients travel through the pinned buffer: wait before reuse
  HPS_HIP(hipStreamSynchronize(op->stream));
  return HPS_OK;
}

extern "C" int hps_axpy_multi_f32(hps_operator op, int64_t n, int k, const double *a,
                                  const float *X, double *y)
{
  if (int r = check_op(op))
    return r;
  if (k < 1 || k > hps::MAXK)
    return fail(HPS_ERR_INVALID_ARGUMENT, "axpy_multi_f32: k must be 1..32");
  for (int j = 0; j < k; ++j)
    op->red_host[32 + j] = a[j];
  HPS_HIP(hipMemcpyAsync(op->red_out + 32, op->red_host + 32, k * sizeof(double),
                         hipMemcpyHostToDevice, op->stream));
  hps::blas_axpy_multi_f32_f64(op->stream, n, k, op->red_out + 32, X, y);
  HPS_LAUNCHED();
  // the coefficients travel through the pinned buffer: wait before reuse
  HPS_HIP(hipStreamSynchronize(op->stream));
  return HPS_OK;
}
