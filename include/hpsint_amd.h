/*
 * hpsint_amd.h -- C ABI of the MI355X sintering-operator library.
 *
 * Drop-in for the hot path of hpsint's SinteringOperatorGeneric: residual,
 * tangent vmult, linearisation cache, block-vector BLAS-1 and Jacobi
 * preconditioners.  All vector arguments are DEVICE pointers (HBM-resident)
 * in the component-interleaved layout  x[node * n_components + component]
 * (component 0 = c, 1 = mu, 2.. = eta_i; nodes lexicographic, x fastest),
 * i.e. the layout of hpsint's VectorTools::merge_components_fast
 * (include/pf-applications/numerics/vector_tools.h:40-100).
 * hps_to_blocks / hps_from_blocks convert to and from the one-vector-per-
 * component layout of DynamicBlockVector.
 *
 * Every call is enqueued on the operator's stream (hps_set_stream) and
 * returns without waiting, except the reductions that return a value to the
 * host (hps_dot, hps_norm2, hps_multi_dot, hps_sync).  Calls on one handle
 * must not overlap from several host threads.
 *
 * Return value: HPS_OK or one of the error codes below; hps_last_error()
 * gives the message of the last failure on the calling thread.
 *
 * Reference interfaces are cited as
 *   S/ = applications/sintering/include/pf-applications/sintering/
 *   I/ = include/pf-applications/
 */
#ifndef HPSINT_AMD_H
#define HPSINT_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum
{
  HPS_OK                    = 0,
  HPS_ERR_INVALID_ARGUMENT  = 1,
  HPS_ERR_INVALID_N_COMP    = 2, /* ExcInvalidNumberOfComponents, S/instantiation.h:61-68 */
  HPS_ERR_UNSUPPORTED       = 3, /* dim / degree / mobility not built */
  HPS_ERR_OUT_OF_MEMORY     = 4,
  HPS_ERR_HIP               = 5,
  HPS_ERR_NOT_INITIALIZED   = 6, /* vmult before set_linearization etc. */
  HPS_ERR_NEWTON_NO_CONVERGENCE = 7, /* ExcNewtonDidNotConverge, I/lac/solvers_nonlinear.h:183-207 */
  HPS_ERR_LINEAR_NO_CONVERGENCE = 8  /* SolverControl::NoConvergence (GMRES) */
};

enum
{
  HPS_MOBILITY_SCALAR    = 0, /* MobilityScalar, S/mobility.h:247-568 */
  HPS_MOBILITY_TENSORIAL = 1  /* MobilityTensorial, S/mobility.h:573-961 */
};

enum
{
  HPS_PRECON_POINT_JACOBI      = 0, /* InverseDiagonalMatrix, I/lac/preconditioners.h:160-215 */
  HPS_PRECON_NODE_BLOCK_JACOBI = 1, /* InverseComponentBlockDiagonalMatrix, :218-318 */
  HPS_PRECON_BLOCK_CH_ILU      = 2  /* BlockPreconditioner2 with its default sub-preconditioners
                                       (ILU on the Cahn-Hilliard block, InverseDiagonalMatrix on
                                       the Allen-Cahn block), S/preconditioners.h:983-1282 */
};

typedef struct hps_operator_s *hps_operator;

/* Mesh, discretisation and material of one operator.  Mirrors the
 * parameters the reference reads from Material.EnergyAbstract.* and
 * Material.MobilityAbstract.* (S/parameters.h:113-134) and the uniform
 * Cartesian mesh of the "Divisions" mode (S/driver.h:575-590). */
typedef struct
{
  int    dim;           /* 2 or 3 */
  int    fe_degree;     /* 1 (FE_Q(1), QGauss(2)) or 2 (FE_Q(2), QGauss(3));
                           nodes per direction: fe_degree * n_cells + 1 */
  int    n_cells[3];    /* cells per direction */
  double cell_size[3];  /* h per direction */
  int    n_components;  /* n_grains + 2 */
  int    mobility;      /* HPS_MOBILITY_* */
  double A, B, kappa_c, kappa_p;
  double Mvol, Mvap, Msurf, Mgb, L;
} hps_params;

/* Replaces the construction of SinteringOperatorGeneric
 * (S/operator_sintering_generic.h:341-364) together with its MatrixFree
 * and SinteringOperatorData (S/operator_sintering_data.h:33-60). */
int hps_create(const hps_params *params, hps_operator *op);
int hps_destroy(hps_operator op);
const char *hps_last_error(void);

/* Stream the operator enqueues on (a hipStream_t; NULL = default stream). */
int hps_set_stream(hps_operator op, void *hip_stream);
int hps_sync(hps_operator op);

/* n_components() / initialize_dof_vector() sizes,
 * S/operator_sintering_generic.h:616-620, S/operator_base.h:143-149 */
int hps_sizes(hps_operator op, int64_t *n_nodes, int64_t *n_dofs);

/* TimeIntegratorData weights: get_primary_weight() = w0,
 * get_algebraic_weight() = w_alg, get_weights()[1..order] = w_hist
 * (I/time_integration/time_integrator_data.h:152-170). */
int hps_set_time_weights(hps_operator op, double w0, double w_alg,
                         const double *w_hist, int order);

/* SinteringOperatorData::fill_quadrature_point_values
 * (S/operator_sintering_data.h:301-409): builds the linearisation cache from
 * the Newton iterate u (FP64 and FP32 copies). */
int hps_set_linearization(hps_operator op, const double *u);

/* SinteringOperatorGeneric::evaluate_nonlinear_residual<wtd>
 * (S/operator_sintering_generic.h:393-448).  old[k] (k < order) are device
 * pointers of the old solutions (only read for wtd == 2); dst is
 * overwritten. */
int hps_residual(hps_operator op, int wtd, double *dst, const double *src,
                 const double *const *old);

/* Constrained nodes: OperatorBase::vmult's constrained_indices, from
 * MatrixFree::get_constrained_dofs (S/operator_base.h:300-314).  nodes is a
 * host array of n distinct node indices (n = 0: none, the default).  With
 * constraints, hps_vmult_* zero the constrained entries of src for the cell
 * operator (src is modified during the call and restored before the call's
 * work on the stream ends) and set dst = src on them (identity rows,
 * S/operator_base.h:316-337); hps_residual leaves their rows zero
 * (FEEvaluation::distribute_local_to_global skips constrained entries); the
 * preconditioners use diagonal 1 and identity node blocks for them. */
int hps_set_constrained_nodes(hps_operator op, const int64_t *nodes, int64_t n);

/* Grain cut-off (Parameters::grain_cut_off_tolerance, S/parameters.h:314):
 * tol > 0 enables it (the reference enables it for a nonzero tolerance),
 * tol <= 0 disables it and drops the component table.  Q1 operators (2D and
 * 3D, scalar and tensorial mobility); Q2 returns HPS_ERR_UNSUPPORTED. */
int hps_set_grain_cut_off(hps_operator op, double tol);
/* SinteringOperatorData::set_component_mask (S/operator_sintering_data.h:
 * 150-298), called by the driver at the start of every Newton solve
 * (S/driver.h:2212-2218): the component table of the state u, one cell per
 * batch -- bit g of a cell's mask is set when max |eta_g| over the cell's
 * nodes exceeds the tolerance.  From then on the residual
 * (do_evaluate_nonlinear_residual_nt, S/operator_sintering_generic.h:
 * 1008-1081), the linearisation cache (zeros for the dropped order
 * parameters, :360-395 of the data file), the tangent vmult
 * (do_vmult_range_no_template, :456-537) and the preconditioner assembly
 * evaluate every cell with c, mu and its active order parameters only; a
 * component with no entry left in a node block gets a unit diagonal.  The
 * linearisation must be set again after this call. */
int hps_set_component_mask(hps_operator op, const double *u);
/* copy of the component table: one uint32 per cell (cells x fastest) into
 * the DEVICE array masks */
int hps_get_component_mask(hps_operator op, uint32_t *masks);

/* OperatorBase::vmult -> SinteringOperatorGeneric::vmult_internal
 * (S/operator_base.h:294-353, S/operator_sintering_generic.h:579-606):
 * tangent at the linearisation point; dst is overwritten.  The f32 variant
 * uses the FP32 copy of the cache and FP32 arithmetic throughout. */
int hps_vmult_f64(hps_operator op, double *dst, const double *src);
int hps_vmult_f32(hps_operator op, float *dst, const float *src);

/* Preconditioner setup (PreconditionerBase::do_update,
 * I/lac/preconditioners.h:77-112): kind = HPS_PRECON_POINT_JACOBI computes
 * the inverse diagonal as OperatorBase::compute_inverse_diagonal does
 * (S/operator_base.h:364-389: 1/d if |d| > 1e-10 else 1);
 * HPS_PRECON_NODE_BLOCK_JACOBI inverts every node's n_c x n_c block
 * (InverseComponentBlockDiagonalMatrix::do_update, Gauss-Jordan);
 * HPS_PRECON_BLOCK_CH_ILU assembles the Cahn-Hilliard block (the tangent's
 * c and mu rows and columns) from the linearisation cache, factors it by
 * ILU(0) with the nodes in 2^dim parity colours (each colour a parallel
 * stage; c before mu within a node), and takes the inverse diagonal on the
 * order parameters (Q1, scalar or tensorial mobility; constrained nodes
 * get identity rows and columns in the block, as in the constrained
 * operator; the assembly and factors are FP64, the FP32 apply uses an FP32
 * copy). */
int hps_precond_update(hps_operator op, int kind);
/* The same setup for the FP32 inner Krylov solve: the diagonal / blocks are
 * assembled from the FP32 cache in FP32 arithmetic, then inverted in FP64
 * (Q2 operators assemble in FP64).  Both applies are valid afterwards. */
int hps_precond_update_f32(hps_operator op, int kind);
/* PreconditionerBase::vmult */
int hps_precond_apply_f64(hps_operator op, double *dst, const double *src);
int hps_precond_apply_f32(hps_operator op, float *dst, const float *src);
/* copies of the preconditioner data for inspection: the inverse diagonal
 * (n_dofs entries) or the node blocks before inversion
 * (n_nodes * n_c * n_c, row-major per node) */
int hps_precond_get_diagonal(hps_operator op, double *inv_diag);
int hps_precond_get_blocks(hps_operator op, double *blocks);

/* DynamicBlockVector BLAS-1 (I/lac/dynamic_block_vector.h:346-537) on whole
 * interleaved vectors of n entries.  Reductions write one double to host
 * memory and wait for it. */
int hps_dot(hps_operator op, int64_t n, const double *x, const double *y,
            double *result);
int hps_norm2(hps_operator op, int64_t n, const double *x, double *result);
/* per-component l2 norms (the block norms DynamicBlockVector::block(b).
 * l2_norm() would give): n_components results */
int hps_block_norms(hps_operator op, const double *x, double *result);
int hps_axpy(hps_operator op, int64_t n, double a, const double *x, double *y);
/* y = s*y + a*x  (sadd) */
int hps_sadd(hps_operator op, int64_t n, double s, double a, const double *x,
             double *y);
int hps_scale(hps_operator op, int64_t n, double a, double *x);
int hps_copy(hps_operator op, int64_t n, double *dst, const double *src);
int hps_fill(hps_operator op, int64_t n, double value, double *x);
/* classical Gram-Schmidt step of GMRES: h[j] = <V_j, w> for j < k, with the
 * basis V stored as k contiguous vectors of n entries (V + j*n) */
int hps_multi_dot(hps_operator op, int64_t n, int k, const double *V,
                  const double *w, double *h);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): hps_multi_dot without an operator handle.  h[j] =
 * <V_j, w> for j < k, k <= 32, written to DEVICE memory h (k doubles) on
 * `stream` (a hipStream_t; NULL = default stream) without waiting for it.
 * The partial sums live in a fixed 512 KiB device array of the library (one
 * per GPU, held until the process ends); calls that overlap on one GPU race
 * on it.  Allocates nothing; k > 32 returns HPS_ERR_INVALID_ARGUMENT. */
int hps_multi_dot_noalloc(void *stream, int64_t n, int k, const double *V, const double *w,
                          double *h);
/* w -= sum_j h[j] V_j, then returns ||w|| (fused CGS update + norm) */
int hps_multi_axpy_norm(hps_operator op, int64_t n, int k, const double *V,
                        const double *h, double *w, double *norm);
/* y += sum_{j<k} a[j] V_j (the GMRES solution update x += V y; a on the
 * host, V as in hps_multi_dot) */
int hps_axpy_multi(hps_operator op, int64_t n, int k, const double *a, const double *V,
                   double *y);
/* conversions FP64 <-> FP32 */
int hps_convert_f64_to_f32(hps_operator op, int64_t n, float *dst,
                           const double *src);
int hps_convert_f32_to_f64(hps_operator op, int64_t n, double *dst,
                           const float *src);

/* FP32-storage vectors with FP64 accumulation (the Krylov basis of the
 * mixed-precision GMRES): norms and Gram-Schmidt coefficients accumulate
 * in FP64; updated entries are stored in FP32 */
int hps_norm2_f32(hps_operator op, int64_t n, const float *x, double *result);
int hps_multi_dot_f32(hps_operator op, int64_t n, int k, const float *V, const float *w,
                      double *h);
int hps_multi_axpy_norm_f32(hps_operator op, int64_t n, int k, const float *V,
                            const double *h, float *w, double *norm);
int hps_scale_f32(hps_operator op, int64_t n, double a, float *x);
/* y (FP64) += sum_{j<k} a[j] X_j (FP32 vectors X + j*n), FP64 arithmetic */
int hps_axpy_multi_f32(hps_operator op, int64_t n, int k, const double *a, const float *X,
                       double *y);

/* layout conversion: merge_components_fast / split_up_components_fast
 * (I/numerics/vector_tools.h:40-100); blocks is a host array of n_c device
 * pointers of n_nodes entries each */
int hps_from_blocks(hps_operator op, double *x, const double *const *blocks);
int hps_to_blocks(hps_operator op, double *const *blocks, const double *x);

/* Kernel timing for measurement: with hps_enable_timing(op, n), n >= 1,
 * every hps_vmult_* / hps_residual call records HIP events on the
 * operator's stream around its main kernel and its shared-node fix-up, into
 * a ring of the last n calls, without waiting (n = 0: off).
 * hps_last_kernel_ms waits for the last call and returns its durations in
 * milliseconds; hps_kernel_ms_mean waits for the stream and returns the
 * mean durations over the recorded calls (at most n) and their number, then
 * starts a new count. */
int hps_enable_timing(hps_operator op, int enable);
int hps_last_kernel_ms(hps_operator op, float *ms_main, float *ms_fixup);
int hps_kernel_ms_mean(hps_operator op, float *ms_main, float *ms_fixup, int *count);

/* ------------------------------------------------------------------
 * Newton-Krylov driver (host logic over the device operator).
 * hps_newton_solve restates DampedNewtonSolver::solve
 * (I/lac/solvers_nonlinear.h:248-357) with the TimeMarchingImplicit wiring
 * (I/time_integration/time_marching.h:129-241): residual, setup_jacobian =
 * hps_set_linearization, setup_preconditioner = hps_precond_update (when
 * history_newton_iterations % threshold_newton_iter == 0 or the last linear
 * solve took more than threshold_linear_iter iterations), solve_with_jacobian
 * = restarted GMRES with left preconditioning and classical Gram-Schmidt
 * (SolverGMRESWrapper, I/lac/solvers_linear.h:58-117, ReductionControl
 * semantics), damped update (tau = 0.1, at most 100 halvings).
 * ------------------------------------------------------------------ */
typedef struct hps_newton_s *hps_newton;

typedef struct
{
  double nl_abs_tol, nl_rel_tol; /* NonLinearData, solvers_nonlinear_parameters.h:39-67 */
  int    nl_max_iter;
  double l_abs_tol, l_rel_tol;
  int    l_max_iter;
  int    l_restart;              /* GMRES basis size (<= 31) */
  int    precond;                /* HPS_PRECON_* */
  int    use_damping;
  int    threshold_newton_iter;
  int    threshold_linear_iter;
  int    inner_fp32;             /* 0: FP64 throughout; 1: FP32 vmult, preconditioner
                                    and Krylov basis storage (Gram-Schmidt
                                    coefficients and norms accumulated in FP64);
                                    2: FP32 vmult and preconditioner, FP64 basis,
                                    orthogonalisation and norms */
} hps_newton_params;

typedef struct
{
  int    newton_iterations, linear_iterations, residual_evaluations, precond_updates;
  int    converged;
  double res0, res;              /* l2 norms of the initial and final residual */
} hps_newton_stats;

void hps_newton_default_params(hps_newton_params *params);
int  hps_newton_create(hps_operator op, const hps_newton_params *params, hps_newton *out);
int  hps_newton_destroy(hps_newton solver);
/* DampedNewtonSolver::clear: the next solve sets the preconditioner up again */
int  hps_newton_clear(hps_newton solver);
/* solves R(x) = 0 in place (x: device vector, initial guess on entry) with
 * the operator's current time weights and the history vectors old[k] */
int  hps_newton_solve(hps_newton solver, double *x, const double *const *old,
                      hps_newton_stats *stats);
/* as hps_newton_solve, but the relative criterion |R(x_k)| <= nl_rel_tol *
 * r0_ref uses the given reference norm (e.g. |R(u_n)|, the initial residual
 * of the plain method) instead of the initial residual of this solve, so
 * that solves from different initial guesses stop at the same accuracy
 * (r0_ref <= 0: identical to hps_newton_solve).  No counterpart in the
 * reference; used with the learned predictor. */
int  hps_newton_solve_ref(hps_newton solver, double *x, const double *const *old,
                          double r0_ref, hps_newton_stats *stats);

/* ------------------------------------------------------------------
 * Learned Newton predictor (no counterpart in the reference: it fills the
 * driver's predictor slot, S/driver.h:2002-2050, where the reference
 * offers Euler/Midpoint/Linear extrapolation).  BF16 storage, FP32
 * accumulation on MFMA.  Matrices are row-major; bf16 arrays are passed as
 * void pointers to 16-bit storage.  All calls use the operator's stream.
 * ------------------------------------------------------------------ */
enum
{
  HPS_MLP_EPI_F32       = 0, /* C = op(A) op(B), FP32 (split-K partials at C + z*M*ldc) */
  HPS_MLP_EPI_BIAS_TANH = 1, /* C = bf16(tanh(op(A) op(B) + bias)) */
  HPS_MLP_EPI_BIAS_F32  = 2, /* C = op(A) op(B) + bias, FP32 */
  HPS_MLP_EPI_DTANH     = 3  /* C = bf16(op(A) op(B) * (1 - H^2)) */
};
/* op(A)(m,k) = trans_a ? A[k*lda+m] : A[m*lda+k];
 * op(B)(k,n) = trans_b ? B[n*ldb+k] : B[k*ldb+n];  M x N result */
int hps_mlp_gemm(hps_operator op, int trans_a, int trans_b, int epilogue, int M, int N, int K,
                 const void *A, int lda, const void *B, int ldb, void *C, int ldc,
                 const float *bias, const void *H, int n_split);
int hps_mlp_sum_splits(hps_operator op, int64_t n, int n_split, const float *part,
                       float *out);
int hps_mlp_colsum(hps_operator op, int M, int N, const void *X, float *out);
/* MSE over n entries; dY = 2 (Y - T) / n (bf16); loss summed on the host if
 * loss != NULL */
int hps_mlp_loss(hps_operator op, int64_t n, const float *Y, const float *T, void *dY,
                 float *partial, int n_blocks, float *loss);
/* Adam step (step counts from 1) on FP32 master weights, writes bf16 copy */
int hps_mlp_adam(hps_operator op, int n_out, int n_in, float *W, float *m1, float *m2,
                 const float *grad, float lr, float beta1, float beta2, float eps, int step,
                 void *W_bf16);
/* per node: 272 bf16 stencil features, the node's top-2 order parameters
 * at t_n (top2: 2 ints per node) and its local increment scale ls (lscale:
 * 1 float per node; increments are divided by it).  inc_scale = 1 / (global
 * increment scale); the floor of ls is 1e-3 / inc_scale. */
int hps_mlp_features(hps_operator op, const double *u_prev, const double *u_cur, double dt,
                     double dt_scale, double inc_scale, void *X, int *top2, float *lscale);
/* u_pred = 2 u_cur - u_prev, plus ls * out_scale[k] * Y[node*4+k] on
 * (c, mu, eta_top1, eta_top2) when Y != NULL */
int hps_mlp_predict(hps_operator op, const double *u_prev, const double *u_cur,
                    const float *Y, const int *top2, const float *out_scale,
                    const float *lscale, double *u_pred);
/* training targets: (u_next - (2 u_cur - u_prev)) / (ls * out_scale) on the
 * 4 outputs */
int hps_mlp_targets(hps_operator op, const double *u_prev, const double *u_cur,
                    const double *u_next, const int *top2, const float *out_scale,
                    const float *lscale, float *T);

#ifdef __cplusplus
}
#endif

#endif
