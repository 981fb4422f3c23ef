/*
 * oracle.h -- CPU restatement of hpsint's sintering-operator hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this code: it is the checker that the HIP path is compared against
 * (pytest, __graft_entry__.smoke() and the cpu_baseline leg of bench.py).
 *
 * Plain C99 + optional OpenMP, FP64 throughout, written as a direct (not
 * sum-factorised) per-cell quadrature loop so that a mistake in the GPU's
 * tensor-product butterflies cannot be mirrored here.
 *
 * Conventions shared with the HIP library:
 *   - uniform Cartesian mesh of n_cells[d] cells of size h[d] (d < dim),
 *     FE_Q(p) on Gauss-Lobatto support points (p = 1: {0,1}; p = 2:
 *     {0,1/2,1}), QGauss(p+1);
 *   - nodes numbered lexicographically, x fastest: node = ix + NX*(iy + NY*iz)
 *     with NX = p*n_cells[0]+1 ...;
 *   - vectors are component-interleaved: entry = node*n_comp + comp, with
 *     comp 0 = c, 1 = mu, 2.. = eta_i (the layout of hpsint's
 *     VectorTools::merge_components_fast, include/pf-applications/numerics/
 *     vector_tools.h:40-100).
 */
#ifndef HPSINT_ORACLE_H
#define HPSINT_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAX_COMP 16
#define ORC_MAX_GRAINS (ORC_MAX_COMP - 2)

typedef struct
{
  int    dim;           /* 2 or 3 */
  int    degree;        /* 1 or 2 */
  int    n_cells[3];    /* cells per direction */
  double h[3];          /* cell size per direction */
  int    n_comp;        /* n_grains + 2 */
  int    tensorial;     /* 0: MobilityScalar, 1: MobilityTensorial */
  double A, B, kappa_c, kappa_p;
  double Mvol, Mvap, Msurf, Mgb, L;
  /* time integration: primary weight w0 (BDF weights[0]), algebraic weight
   * (1 for implicit schemes), and history weights weights[1..order] */
  double w0, w_alg;
  int    order;
  double w_hist[3];
} orc_params;

/* ---- pointwise physics (free_energy.h / mobility.h / tensor.h) ---- */

/* FreeEnergy::Evaluation, S/free_energy.h:37-176.  out[] gets
 * f, df_dc, d2f_dc2 and per grain df_detai, d2f_dcdetai, d2f_detai2
 * laid out as: out[0]=f out[1]=df_dc out[2]=d2f_dc2,
 * out[3+i]=df_detai, out[3+G+i]=d2f_dcdetai, out[3+2G+i]=d2f_detai2. */
void orc_free_energy_eval(double A, double B, double c, const double *etas,
                          int n_grains, double *out);
double orc_d2f_detaidetaj(double B, double etai, double etaj);

/* PowerHelper<n,p>::power_sum, I/numerics/power_helper.h:26-60 */
double orc_power_sum(const double *v, int n, int p);

/* unit_vector / unit_vector_filter_2, I/base/tensor.h:137-174 */
void   orc_unit_vector(const double *v, int dim, double zero_tol, double *out);
double orc_unit_vector_filter_2(const double *v, int dim, double zero_tol);

/* residual QP operator (S/operator_sintering_generic.h:254-308).
 * value[n_comp], grad[n_comp*dim]; hist[n_comp] = interpolated
 * sum_k w_k old_k (only read for wtd == 2). */
void orc_residual_qp(const orc_params *p, int wtd, const double *value,
                     const double *grad, const double *hist, double *vr,
                     double *gr);

/* tangent QP operator (S/operator_sintering_generic.h:104-190), applied at
 * the linearisation point (lin_value[n_comp], lin_grad[n_comp*dim]) to the
 * increment (value, grad). */
void orc_tangent_qp(const orc_params *p, const double *lin_value,
                    const double *lin_grad, const double *value,
                    const double *grad, double *vr, double *gr);

/* ---- mesh sizes ---- */
int64_t orc_n_nodes(const orc_params *p);
int64_t orc_n_cells(const orc_params *p);
int     orc_n_q_points(const orc_params *p);

/* ---- cell loops ---- */

/* Nonlinear residual, evaluate_nonlinear_residual<wtd>
 * (S/operator_sintering_generic.h:393-448, 877-1006).  old[k] (k < order)
 * are the old solutions; dst is overwritten. */
void orc_residual(const orc_params *p, int wtd, const double *src,
                  const double *const *old, double *dst);

/* SinteringOperatorData::fill_quadrature_point_values
 * (S/operator_sintering_data.h:301-409).  Cache layout per cell and q:
 * n_comp values followed by n_comp*dim gradients (all components; the
 * reference keeps fewer gradients for scalar mobility, which the tangent
 * never reads).  Size: n_cells * n_q * n_comp * (1 + dim). */
void orc_fill_qp(const orc_params *p, const double *lin, double *cache);

/* tangent vmult (S/operator_sintering_generic.h:539-606) on a filled cache */
void orc_vmult(const orc_params *p, const double *cache, const double *src,
               double *dst);

/* node blocks of the tangent: blocks[node][i][j] = d R_(node,i) / d u_(node,j)
 * assembled matrix-free by probing every cell with unit vectors (the recipe
 * of MatrixFreeTools::compute_diagonal / I/matrix_free/tools.h:106-127). */
void orc_node_blocks(const orc_params *p, const double *cache,
                     double *blocks);

/* dense tangent (n_dofs x n_dofs, row-major) by unit-vector vmults; for
 * tiny meshes only */
void orc_dense_tangent(const orc_params *p, const double *cache,
                       double *mat);

/* ---- grain cut-off (S/operator_sintering_data.h:150-298,
 * S/operator_sintering_generic.h:456-537, 1008-1081) ---- */

/* SinteringOperatorData::set_component_mask with one cell per batch: bit g
 * of masks[cell] (cells x fastest) is set when the l-infinity norm of
 * eta_g over the cell's nodes exceeds tol. */
void orc_cell_masks(const orc_params *p, const double *u, double tol, uint32_t *masks);
/* tangent vmult, residual and node blocks with each cell evaluated as the
 * system of c, mu and its active order parameters (cache from
 * orc_fill_qp of the full state) */
void orc_vmult_cut(const orc_params *p, const double *cache, const uint32_t *masks,
                   const double *src, double *dst);
void orc_residual_cut(const orc_params *p, int wtd, const uint32_t *masks, const double *src,
                      const double *const *old, double *dst);
void orc_node_blocks_cut(const orc_params *p, const double *cache, const uint32_t *masks,
                         double *blocks);

/* ILU(0) of a CSR matrix (sorted columns, diagonal stored), in place;
 * returns 0, or < 0 on a missing or zero pivot; and its solve */
int  orc_ilu0(int64_t n, const int64_t *rowptr, const int64_t *col, double *val);
void orc_ilu0_solve(int64_t n, const int64_t *rowptr, const int64_t *col, const double *val,
                    const double *b, double *x);

/* OpenMP threads used by the cell loops (1 when built without OpenMP) */
int  orc_max_threads(void);
void orc_set_threads(int n);

#ifdef __cplusplus
}
#endif

#endif
