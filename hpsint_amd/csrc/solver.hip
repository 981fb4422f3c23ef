// solver.cpp -- host-side Newton-Krylov driver over the device operator.
//
// Restates the reference's solver logic, which stays on the host:
//   DampedNewtonSolver::solve   I/lac/solvers_nonlinear.h:248-357
//   NewtonSolverSolverControl::check                      :94-117
//   SolverGMRESWrapper (deal.II SolverGMRES, classical Gram-Schmidt,
//     ReductionControl)          I/lac/solvers_linear.h:58-117
//   TimeMarchingImplicit lambdas  I/time_integration/time_marching.h:129-241
// with I/ = include/pf-applications/.  Every vector operation runs on the
// GPU through the C ABI; the host keeps only the Hessenberg matrix, the
// Givens rotations and scalars.  With inner_fp32 the tangent vmult and the
// preconditioner inside GMRES run in FP32 (FP32 cache, FP32 arithmetic);
// the Krylov basis, orthogonalisation, norms, the Newton residual and the
// Newton update stay FP64.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/hpsint_amd.h"

struct hps_newton_s
{
  hps_operator      op;
  hps_newton_params prm;
  int64_t           n;
  double           *r, *inc, *tmp, *w, *V, *z64;
  float            *x32, *y32, *z32, *V32;
  // DampedNewtonSolver history (persists across solves)
  unsigned          history_newton_iterations;
  unsigned          history_linear_iterations_last;
  bool              precond_ready;
};

// error reporting shared with the operator entry points (api.hip)
int hps_internal_fail(int code, const char *msg);

namespace
{
  int fail(int code, const std::string &m)
  {
    return hps_internal_fail(code, m.c_str());
  }

#define HPS_TRY(call)                                                         \
  do                                                                          \
    {                                                                         \
      const int rc_ = (call);                                                 \
      if (rc_ != HPS_OK)                                                      \
        return rc_;                                                           \
    }                                                                         \
  while (0)

  // y = P^-1 A x   (FP64, or FP32 inside)
  int apply_prec_op(hps_newton nt, const double *x, double *y)
  {
    hps_operator op = nt->op;
    const int64_t n = nt->n;
    if (nt->prm.inner_fp32)
      {
        HPS_TRY(hps_convert_f64_to_f32(op, n, nt->x32, x));
        HPS_TRY(hps_vmult_f32(op, nt->y32, nt->x32));
        HPS_TRY(hps_precond_apply_f32(op, nt->z32, nt->y32));
        HPS_TRY(hps_convert_f32_to_f64(op, n, y, nt->z32));
      }
    else
      {
        HPS_TRY(hps_vmult_f64(op, nt->z64, x));
        HPS_TRY(hps_precond_apply_f64(op, y, nt->z64));
      }
    return HPS_OK;
  }

  int apply_prec(hps_newton nt, const double *x, double *y)
  {
    hps_operator op = nt->op;
    if (nt->prm.inner_fp32)
      {
        HPS_TRY(hps_convert_f64_to_f32(nt->op, nt->n, nt->x32, x));
        HPS_TRY(hps_precond_apply_f32(op, nt->z32, nt->x32));
        HPS_TRY(hps_convert_f32_to_f64(op, nt->n, y, nt->z32));
        return HPS_OK;
      }
    return hps_precond_apply_f64(op, y, x);
  }

  // Givens-rotation bookkeeping shared by both precisions
  struct Arnoldi
  {
    int                 m;
    std::vector<double> H, cs, sn, g, y, h;
    explicit Arnoldi(int m_)
      : m(m_), H((m_ + 1) * m_), cs(m_), sn(m_), g(m_ + 1), y(m_), h(m_ + 1)
    {}
    // rotates column j (h[0..j+1]) and returns the new residual estimate
    double add_column(int j)
    {
      for (int i = 0; i < j; ++i)
        {
          const double t = cs[i] * h[i] + sn[i] * h[i + 1];
          h[i + 1]       = -sn[i] * h[i] + cs[i] * h[i + 1];
          h[i]           = t;
        }
      const double den = std::hypot(h[j], h[j + 1]);
      cs[j]            = den > 0.0 ? h[j] / den : 1.0;
      sn[j]            = den > 0.0 ? h[j + 1] / den : 0.0;
      h[j]             = den;
      h[j + 1]         = 0.0;
      g[j + 1]         = -sn[j] * g[j];
      g[j]             = cs[j] * g[j];
      for (int i = 0; i <= j; ++i)
        H[i * m + j] = h[i];
      return std::fabs(g[j + 1]);
    }
    void solve(int j)
    {
      for (int i = j - 1; i >= 0; --i)
        {
          double s = g[i];
          for (int k = i + 1; k < j; ++k)
            s -= H[i * m + k] * y[k];
          y[i] = s / H[i * m + i];
        }
    }
  };

  // Restarted GMRES(m) with left preconditioning and classical Gram-Schmidt
  // (one re-orthogonalisation pass when the new vector loses more than
  // half of its norm); convergence as deal.II's ReductionControl on the
  // preconditioned residual: |r_k| <= max(abs_tol, rel_tol |r_0|).
  // Solves A x = b with x = 0 on entry.  Returns the iteration count in *its.
  int gmres(hps_newton nt, const double *b, double *x, int *its)
  {
    hps_operator  op = nt->op;
    const int64_t n  = nt->n;
    const int     m  = nt->prm.l_restart;
    double       *V  = nt->V;
    Arnoldi       a(m);
    HPS_TRY(hps_fill(op, n, 0.0, x));
    // r0 = P^-1 b
    HPS_TRY(apply_prec(nt, b, V));
    double beta = 0.0;
    HPS_TRY(hps_norm2(op, n, V, &beta));
    const double target = std::max(nt->prm.l_abs_tol, nt->prm.l_rel_tol * beta);
    int          it     = 0;
    if (beta <= target)
      {
        *its = 0;
        return HPS_OK;
      }
    bool first = true;
    while (true)
      {
        if (!first)
          {
            // restart: r = P^-1 (b - A x)
            HPS_TRY(hps_vmult_f64(op, nt->z64, x));
            HPS_TRY(hps_sadd(op, n, -1.0, 1.0, b, nt->z64));
            HPS_TRY(apply_prec(nt, nt->z64, V));
            HPS_TRY(hps_norm2(op, n, V, &beta));
            if (beta <= target)
              break;
          }
        first = false;
        HPS_TRY(hps_scale(op, n, 1.0 / beta, V));
        std::fill(a.g.begin(), a.g.end(), 0.0);
        a.g[0]         = beta;
        int  j         = 0;
        bool converged = false;
        for (; j < m; ++j)
          {
            double *vj1 = V + (int64_t)(j + 1) * n;
            HPS_TRY(apply_prec_op(nt, V + (int64_t)j * n, vj1));
            HPS_TRY(hps_multi_dot(op, n, j + 1, V, vj1, a.h.data()));
            double wnorm = 0.0;
            HPS_TRY(hps_multi_axpy_norm(op, n, j + 1, V, a.h.data(), vj1, &wnorm));
            // norm before the projection, sqrt(|w'|^2 + |h|^2) (V orthonormal)
            double hh = 0.0;
            for (int i = 0; i <= j; ++i)
              hh += a.h[i] * a.h[i];
            if (wnorm < 0.5 * std::sqrt(wnorm * wnorm + hh))
              {
                std::vector<double> h2(j + 1);
                HPS_TRY(hps_multi_dot(op, n, j + 1, V, vj1, h2.data()));
                HPS_TRY(hps_multi_axpy_norm(op, n, j + 1, V, h2.data(), vj1, &wnorm));
                for (int i = 0; i <= j; ++i)
                  a.h[i] += h2[i];
              }
            a.h[j + 1]       = wnorm;
            const double res = a.add_column(j);
            ++it;
            if (res <= target || wnorm == 0.0)
              {
                converged = true;
                ++j;
                break;
              }
            if (it >= nt->prm.l_max_iter)
              {
                ++j;
                break;
              }
            HPS_TRY(hps_scale(op, n, 1.0 / wnorm, vj1));
          }
        // back substitution H y = g, x += V y
        a.solve(j);
        HPS_TRY(hps_axpy_multi(op, n, j, a.y.data(), V, x));
        if (converged)
          break;
        if (it >= nt->prm.l_max_iter)
          {
            *its = it;
            return fail(HPS_ERR_LINEAR_NO_CONVERGENCE,
                        "GMRES did not converge in " + std::to_string(it) + " iterations");
          }
      }
    *its = it;
    return HPS_OK;
  }

  // GMRES with the Krylov basis stored in FP32: FP32 tangent and
  // preconditioner, Gram-Schmidt coefficients and norms accumulated in FP64,
  // solution accumulated in FP64.  Same iteration and stopping rules as gmres.
  int gmres_mixed(hps_newton nt, const double *b, double *x, int *its)
  {
    hps_operator op = nt->op;
    const int64_t n = nt->n;
    const int     m = nt->prm.l_restart;
    float        *V = nt->V32;
    Arnoldi       a(m);
    HPS_TRY(hps_fill(op, n, 0.0, x));
    HPS_TRY(hps_convert_f64_to_f32(op, n, nt->x32, b));
    HPS_TRY(hps_precond_apply_f32(op, V, nt->x32));
    double beta = 0.0;
    HPS_TRY(hps_norm2_f32(op, n, V, &beta));
    const double target = std::max(nt->prm.l_abs_tol, nt->prm.l_rel_tol * beta);
    int          it     = 0;
    if (beta <= target)
      {
        *its = 0;
        return HPS_OK;
      }
    bool first = true;
    while (true)
      {
        if (!first)
          {
            HPS_TRY(hps_convert_f64_to_f32(op, n, nt->x32, x));
            HPS_TRY(hps_vmult_f32(op, nt->y32, nt->x32));
            HPS_TRY(hps_convert_f32_to_f64(op, n, nt->z64, nt->y32));
            HPS_TRY(hps_sadd(op, n, -1.0, 1.0, b, nt->z64));
            HPS_TRY(hps_convert_f64_to_f32(op, n, nt->x32, nt->z64));
            HPS_TRY(hps_precond_apply_f32(op, V, nt->x32));
            HPS_TRY(hps_norm2_f32(op, n, V, &beta));
            if (beta <= target)
              break;
          }
        first = false;
        HPS_TRY(hps_scale_f32(op, n, 1.0 / beta, V));
        std::fill(a.g.begin(), a.g.end(), 0.0);
        a.g[0] = beta;
        int  j = 0;
        bool converged = false;
        for (; j < m; ++j)
          {
            float *vj1 = V + (int64_t)(j + 1) * n;
            HPS_TRY(hps_vmult_f32(op, nt->y32, V + (int64_t)j * n));
            HPS_TRY(hps_precond_apply_f32(op, vj1, nt->y32));
            HPS_TRY(hps_multi_dot_f32(op, n, j + 1, V, vj1, a.h.data()));
            double wnorm = 0.0;
            HPS_TRY(hps_multi_axpy_norm_f32(op, n, j + 1, V, a.h.data(), vj1, &wnorm));
            double hh = 0.0;
            for (int i = 0; i <= j; ++i)
              hh += a.h[i] * a.h[i];
            if (wnorm < 0.5 * std::sqrt(wnorm * wnorm + hh))
              {
                std::vector<double> h2(j + 1);
                HPS_TRY(hps_multi_dot_f32(op, n, j + 1, V, vj1, h2.data()));
                HPS_TRY(hps_multi_axpy_norm_f32(op, n, j + 1, V, h2.data(), vj1, &wnorm));
                for (int i = 0; i <= j; ++i)
                  a.h[i] += h2[i];
              }
            a.h[j + 1]       = wnorm;
            const double res = a.add_column(j);
            ++it;
            if (res <= target || wnorm == 0.0)
              {
                converged = true;
                ++j;
                break;
              }
            if (it >= nt->prm.l_max_iter)
              {
                ++j;
                break;
              }
            HPS_TRY(hps_scale_f32(op, n, 1.0 / wnorm, vj1));
          }
        a.solve(j);
        HPS_TRY(hps_axpy_multi_f32(op, n, j, a.y.data(), V, x));
        if (converged)
          break;
        if (it >= nt->prm.l_max_iter)
          {
            *its = it;
            return fail(HPS_ERR_LINEAR_NO_CONVERGENCE,
                        "GMRES did not converge in " + std::to_string(it) + " iterations");
          }
      }
    *its = it;
    return HPS_OK;
  }
} // namespace

extern "C" void hps_newton_default_params(hps_newton_params *p)
{
  // NonLinearData defaults, I/lac/solvers_nonlinear_parameters.h:39-67
  std::memset(p, 0, sizeof(*p));
  p->nl_abs_tol              = 1e-20;
  p->nl_rel_tol              = 1e-5;
  p->nl_max_iter             = 10;
  p->l_abs_tol               = 1e-10;
  p->l_rel_tol               = 1e-2;
  p->l_max_iter              = 1000;
  p->l_restart               = 30;
  p->precond                 = HPS_PRECON_NODE_BLOCK_JACOBI;
  p->use_damping             = 1;
  p->threshold_newton_iter   = 100;
  p->threshold_linear_iter   = 20;
  p->inner_fp32              = 0;
}

extern "C" int hps_newton_create(hps_operator op, const hps_newton_params *prm,
                                 hps_newton *out)
{
  if (!op || !prm || !out)
    return fail(HPS_ERR_INVALID_ARGUMENT, "null argument");
  if (prm->l_restart < 1 || prm->l_restart > 31)
    return fail(HPS_ERR_INVALID_ARGUMENT, "GMRES restart length must be 1..31");
  if (prm->inner_fp32 < 0 || prm->inner_fp32 > 2)
    return fail(HPS_ERR_INVALID_ARGUMENT, "inner_fp32 must be 0, 1 or 2");
  int64_t nn = 0, nd = 0;
  HPS_TRY(hps_sizes(op, &nn, &nd));
  hps_newton nt = new hps_newton_s();
  std::memset((void *)nt, 0, sizeof(*nt));
  nt->op  = op;
  nt->prm = *prm;
  nt->n   = nd;
  const size_t b64 = (size_t)nd * sizeof(double), b32 = (size_t)nd * sizeof(float);
  hipError_t   e   = hipSuccess;
  e = e ? e : hipMalloc(&nt->r, b64);
  e = e ? e : hipMalloc(&nt->inc, b64);
  e = e ? e : hipMalloc(&nt->tmp, b64);
  e = e ? e : hipMalloc(&nt->z64, b64);
  if (prm->inner_fp32)
    {
      e = e ? e : hipMalloc(&nt->x32, b32);
      e = e ? e : hipMalloc(&nt->y32, b32);
      e = e ? e : hipMalloc(&nt->z32, b32);
    }
  if (prm->inner_fp32 == 1)
    e = e ? e : hipMalloc(&nt->V32, b32 * (prm->l_restart + 1));
  else
    e = e ? e : hipMalloc(&nt->V, b64 * (prm->l_restart + 1));
  if (e != hipSuccess)
    {
      hps_newton_destroy(nt);
      return fail(e == hipErrorOutOfMemory ? HPS_ERR_OUT_OF_MEMORY : HPS_ERR_HIP,
                  std::string("newton allocation: ") + hipGetErrorString(e));
    }
  *out = nt;
  return HPS_OK;
}

extern "C" int hps_newton_destroy(hps_newton nt)
{
  if (!nt)
    return HPS_OK;
  (void)hps_sync(nt->op);
  for (void *p : {(void *)nt->r, (void *)nt->inc, (void *)nt->tmp, (void *)nt->z64,
                  (void *)nt->V, (void *)nt->x32, (void *)nt->y32, (void *)nt->z32,
                  (void *)nt->V32})
    if (p)
      (void)hipFree(p);
  delete nt;
  return HPS_OK;
}

extern "C" int hps_newton_clear(hps_newton nt)
{
  if (!nt)
    return fail(HPS_ERR_INVALID_ARGUMENT, "null solver");
  nt->precond_ready = false;
  return HPS_OK;
}

extern "C" int hps_newton_solve(hps_newton nt, double *x, const double *const *old,
                                hps_newton_stats *st)
{
  return hps_newton_solve_ref(nt, x, old, -1.0, st);
}

extern "C" int hps_newton_solve_ref(hps_newton nt, double *x, const double *const *old,
                                    double r0_ref, hps_newton_stats *st)
{
  if (!nt || !x)
    return fail(HPS_ERR_INVALID_ARGUMENT, "null argument");
  hps_operator        op = nt->op;
  const int64_t       n  = nt->n;
  hps_newton_stats    s{};
  auto residual = [&](const double *src, double *dst, double *norm) -> int {
    HPS_TRY(hps_residual(op, 2, dst, src, old));
    ++s.residual_evaluations;
    return hps_norm2(op, n, dst, norm);
  };
  const auto &p = nt->prm;
  double norm_r = 0.0;
  HPS_TRY(residual(x, nt->r, &norm_r));
  s.res0 = norm_r;
  if (!std::isfinite(norm_r))
    return fail(HPS_ERR_NEWTON_NO_CONVERGENCE, "non-finite initial residual");
  // NewtonSolverSolverControl::check; r0_ref > 0 replaces the initial
  // residual as the reference of the relative criterion
  const double r0   = r0_ref > 0.0 ? r0_ref : norm_r;
  auto         keep = [&](int it, double v) {
    return v > p.nl_abs_tol && v / r0 > p.nl_rel_tol && it < p.nl_max_iter;
  };
  auto ok = [&](double v) { return v <= p.nl_abs_tol || v / r0 <= p.nl_rel_tol; };
  int  it = 0;
  while (keep(it, norm_r))
    {
      HPS_TRY(hps_scale(op, n, -1.0, nt->r));
      const bool threshold_exceeded =
        (nt->history_newton_iterations % p.threshold_newton_iter == 0) ||
        (nt->history_linear_iterations_last > (unsigned)p.threshold_linear_iter);
      HPS_TRY(hps_set_linearization(op, x)); // setup_jacobian
      if (threshold_exceeded || !nt->precond_ready)
        {
          HPS_TRY(p.inner_fp32 ? hps_precond_update_f32(op, p.precond)
                               : hps_precond_update(op, p.precond));
          nt->precond_ready = true;
          ++s.precond_updates;
        }
      int lin = 0;
      // inner_fp32 == 1: FP32 Krylov basis; 2: FP64 basis around the FP32
      // tangent and preconditioner (apply_prec_op, apply_prec)
      int rc  = p.inner_fp32 == 1 ? gmres_mixed(nt, nt->r, nt->inc, &lin)
                                  : gmres(nt, nt->r, nt->inc, &lin);
      s.linear_iterations += lin;
      nt->history_linear_iterations_last = lin;
      if (rc != HPS_OK)
        {
          *st = s;
          return rc;
        }
      if (p.use_damping)
        {
          const double tau = 0.1;
          double       omega = 1.0, norm_tmp = 1.0;
          int          n_tmp = 0;
          do
            {
              HPS_TRY(hps_copy(op, n, nt->tmp, x));
              HPS_TRY(hps_axpy(op, n, omega, nt->inc, nt->tmp));
              HPS_TRY(residual(nt->tmp, nt->r, &norm_tmp));
              omega /= 2.0;
              ++n_tmp;
            }
          while ((norm_tmp >= (1.0 - tau * omega) * norm_r || !std::isfinite(norm_tmp)) &&
                 n_tmp < 100);
          if (!(norm_tmp < (1.0 - tau * omega) * norm_r))
            {
              *st = s;
              return fail(HPS_ERR_NEWTON_NO_CONVERGENCE, "damping");
            }
          HPS_TRY(hps_copy(op, n, x, nt->tmp));
          norm_r = norm_tmp;
        }
      else
        {
          HPS_TRY(hps_axpy(op, n, 1.0, nt->inc, x));
          HPS_TRY(residual(x, nt->r, &norm_r));
        }
      ++it;
      ++nt->history_newton_iterations;
    }
  s.newton_iterations = it;
  s.res               = norm_r;
  s.converged         = ok(norm_r) ? 1 : 0;
  *st                 = s;
  if (!s.converged)
    return fail(HPS_ERR_NEWTON_NO_CONVERGENCE, "Newton");
  return HPS_OK;
}
