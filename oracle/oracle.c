/*
 * oracle.c -- CPU restatement of hpsint's sintering-operator hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP library, never linked
 * into it.  Every function cites the reference file:line it restates; the
 * path abbreviations are
 *   S/ = applications/sintering/include/pf-applications/sintering/
 *   I/ = include/pf-applications/
 *
 * The arithmetic of the pointwise physics keeps the reference's operation
 * order (it decides which side of the thresholds in unit_vector, the fsurf
 * mask and the phi clamp a value falls).  The FE machinery (deal.II
 * FEEvaluation, a third-party dependency) is restated from its textbook
 * definition: Lagrange bases on Gauss-Lobatto support points, Gauss-Legendre
 * quadrature, affine Cartesian mapping (J = diag(h), JxW = prod(h) w_q).
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#define MAXN1 3            /* max 1D nodes / q-points (Q2, 3-point Gauss) */
#define MAXNL 27           /* max nodes per cell */
#define MAXQ 27            /* max q-points per cell */

/* ------------------------------------------------------------------ */
/* threads                                                              */
/* ------------------------------------------------------------------ */
int orc_max_threads(void)
{
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

void orc_set_threads(int n)
{
#ifdef _OPENMP
  if (n > 0)
    omp_set_num_threads(n);
#else
  (void)n;
#endif
}

/* ------------------------------------------------------------------ */
/* power sums: PowerHelper<n,p>::power_sum, I/numerics/power_helper.h:26-60,
 * with deal.II Utilities::fixed_power (exponentiation by squaring).     */
/* ------------------------------------------------------------------ */
static double fixed_power(double x, int p)
{
  if (p == 0)
    return 1.0;
  return (p % 2 == 1) ? x * fixed_power(x * x, p / 2) : fixed_power(x * x, p / 2);
}

double orc_power_sum(const double *v, int n, int p)
{
  double s = 0.0;
  for (int i = 0; i < n; ++i)
    s += fixed_power(v[i], p);
  return s;
}

/* ------------------------------------------------------------------ */
/* free energy: S/free_energy.h:37-176                                  */
/* ------------------------------------------------------------------ */
typedef struct
{
  double A, B, c, S2, S3;
} fe_eval;

static void fe_init(fe_eval *e, double A, double B, double c,
                    const double *etas, int G, int with_power_3)
{
  e->A  = A;
  e->B  = B;
  e->c  = c;
  e->S2 = orc_power_sum(etas, G, 2);
  e->S3 = with_power_3 ? orc_power_sum(etas, G, 3) : 0.0;
}

/* :112 */
static double fe_f(const fe_eval *e)
{
  const double A = e->A, B = e->B, c = e->c;
  return A * (c * c) * ((-c + 1.0) * (-c + 1.0)) +
         B * ((c * c) + (-6.0 * c + 6.0) * e->S2 - (-4.0 * c + 8.0) * e->S3 +
              3.0 * (e->S2 * e->S2));
}

/* :125 */
static double fe_df_dc(const fe_eval *e)
{
  const double A = e->A, B = e->B, c = e->c;
  return A * (c * c) * (2.0 * c - 2.0) + 2.0 * A * c * ((-c + 1.0) * (-c + 1.0)) +
         B * (2.0 * c - 6.0 * e->S2 + 4.0 * e->S3);
}

/* :137 */
static double fe_df_detai(const fe_eval *e, double etai)
{
  const double B = e->B, c = e->c;
  return etai * (B * 12.0) * (etai * (1.0 * c - 2.0) + (-c + 1.0) + e->S2);
}

/* :144 */
static double fe_d2f_dc2(const fe_eval *e)
{
  const double A = e->A, B = e->B, c = e->c;
  return 2.0 * A * (c * c) + 4.0 * A * c * (2.0 * c - 2.0) +
         2.0 * A * ((-c + 1.0) * (-c + 1.0)) + 2.0 * B;
}

/* :151 */
static double fe_d2f_dcdetai(const fe_eval *e, double etai)
{
  return (e->B * 12.0) * etai * (etai - 1.0);
}

/* :157 */
static double fe_d2f_detai2(const fe_eval *e, double etai)
{
  const double B = e->B, c = e->c;
  return B * (12.0 - 12.0 * c + 2.0 * etai * (12.0 * c - 24.0) +
              24.0 * (etai * etai) + 12.0 * e->S2);
}

/* :164 */
double orc_d2f_detaidetaj(double B, double etai, double etaj)
{
  return 24.0 * B * etai * etaj;
}

void orc_free_energy_eval(double A, double B, double c, const double *etas,
                          int G, double *out)
{
  fe_eval e;
  fe_init(&e, A, B, c, etas, G, 1);
  out[0] = fe_f(&e);
  out[1] = fe_df_dc(&e);
  out[2] = fe_d2f_dc2(&e);
  for (int i = 0; i < G; ++i)
    {
      out[3 + i]         = fe_df_detai(&e, etas[i]);
      out[3 + G + i]     = fe_d2f_dcdetai(&e, etas[i]);
      out[3 + 2 * G + i] = fe_d2f_detai2(&e, etas[i]);
    }
}

/* apply_d2f_detaidetaj, S/free_energy.h:234-252 */
static void apply_d2f_detaidetaj(double L, double B, const double *etas, int G,
                                 const double *value, double *value_result)
{
  if (G <= 1)
    return;
  double temp = etas[0] * value[0];
  for (int ig = 1; ig < G; ++ig)
    temp += etas[ig] * value[ig];
  for (int ig = 0; ig < G; ++ig)
    value_result[ig] += (L * 24.0 * B) * etas[ig] * (temp - etas[ig] * value[ig]);
}

/* ------------------------------------------------------------------ */
/* tensor helpers: I/base/tensor.h:126-186                              */
/* ------------------------------------------------------------------ */
static double dot(const double *a, const double *b, int dim)
{
  double s = a[0] * b[0];
  for (int d = 1; d < dim; ++d)
    s += a[d] * b[d];
  return s;
}

/* :137-155 */
void orc_unit_vector(const double *v, int dim, double zero_tol, double *out)
{
  const double nrm           = sqrt(dot(v, v, dim));
  const double filter        = nrm > zero_tol ? 1.0 : 0.0;
  const double filtered_norm = nrm < zero_tol ? 1.0 : nrm;
  for (int d = 0; d < dim; ++d)
    out[d] = (v[d] / filtered_norm) * filter;
}

/* :157-174 */
double orc_unit_vector_filter_2(const double *v, int dim, double zero_tol)
{
  const double nrm           = dot(v, v, dim);
  const double filter        = nrm > zero_tol ? 1.0 : 0.0;
  const double filtered_norm = nrm < zero_tol ? 1.0 : nrm;
  return filter / filtered_norm;
}

/* ------------------------------------------------------------------ */
/* mobility: S/mobility.h                                               */
/* ------------------------------------------------------------------ */
static double phi_clamped(double c)
{
  double phi = c * c * c * (10.0 - 15.0 * c + 6.0 * c * c);
  phi        = fmax(0.0, phi);
  phi        = fmin(1.0, phi);
  return phi;
}

/* MobilityScalar::apply_M, S/mobility.h:441-472 */
static void scalar_apply_M(const orc_params *p, double c, const double *etas,
                           int G, const double *mu_grad, int dim, double *out)
{
  double etaijSum = 0.0;
  for (int i = 0; i < G; ++i)
    for (int j = 0; j < i; ++j)
      etaijSum += etas[i] * etas[j];
  const double phi = phi_clamped(c);
  const double M   = p->Mvol * phi + p->Mvap * (1.0 - phi) +
                   p->Msurf * 4.0 * c * c * (1.0 - c) * (1.0 - c) +
                   (2.0 * p->Mgb) * etaijSum;
  for (int d = 0; d < dim; ++d)
    out[d] = M * mu_grad[d];
}

/* MobilityScalar::apply_M_derivative, S/mobility.h:477-534 */
static void scalar_apply_M_derivative(const orc_params *p, const double *lin_value,
                                      const double *lin_mu_grad, int G,
                                      const double *value, const double *mu_grad,
                                      int dim, double *out)
{
  const double  c        = lin_value[0];
  const double *lin_etas = lin_value + 2;
  double        eta_sum = 0.0, eta_ij_sum = 0.0;
  for (int i = 0; i < G; ++i)
    {
      eta_sum += lin_etas[i];
      for (int j = 0; j < i; ++j)
        eta_ij_sum += lin_etas[i] * lin_etas[j];
    }
  const double phi = phi_clamped(c);
  const double M   = p->Mvol * phi + p->Mvap * (1.0 - phi) +
                   p->Msurf * 4.0 * c * c * (1.0 - c) * (1.0 - c) +
                   (2.0 * p->Mgb) * eta_ij_sum;
  const double dphidc = 30.0 * c * c * (1.0 - c) * (1.0 - c);
  const double dMdc   = p->Mvol * dphidc - p->Mvap * dphidc +
                      p->Msurf * 8.0 * c * (1.0 - 3.0 * c + 2.0 * c * c);
  double factor = 0.0;
  for (int i = 0; i < G; ++i)
    factor += (eta_sum - lin_etas[i]) * value[2 + i];
  const double s = dMdc * value[0] + (2.0 * p->Mgb) * factor;
  for (int d = 0; d < dim; ++d)
    out[d] = M * mu_grad[d] + s * lin_mu_grad[d];
}

static const double projector_scale = 1.001; /* S/mobility.h:960 */

/* MobilityTensorial::apply_M, S/mobility.h:758-804.
 * grads: [comp][dim] of the state (c, mu, etas). */
static void tensorial_apply_M(const orc_params *p, const double *value,
                              const double *grad, int G, int dim, double *out)
{
  const double  c       = value[0];
  const double *etas    = value + 2;
  const double *c_grad  = grad;
  const double *mu_grad = grad + dim;
  const double  phi     = phi_clamped(c);
  const double  f_vol_vap = p->Mvol * phi + p->Mvap * (1.0 - phi);
  const double  fsurf = p->Msurf * 4.0 * (c * c) * ((1. - c) * (1. - c));
  double        nc[3];
  orc_unit_vector(c_grad, dim, 1e-10, nc);
  const double ncmu = dot(nc, mu_grad, dim);
  for (int d = 0; d < dim; ++d)
    out[d] = (f_vol_vap + fsurf * projector_scale) * mu_grad[d] -
             nc[d] * (fsurf * ncmu);
  if (G <= 1)
    return;
  double out_gb[3] = {0.0, 0.0, 0.0};
  for (int i = 0; i < G; ++i)
    for (int j = 0; j < i; ++j)
      {
        double egd[3];
        for (int d = 0; d < dim; ++d)
          egd[d] = grad[(2 + i) * dim + d] - grad[(2 + j) * dim + d];
        const double filter = orc_unit_vector_filter_2(egd, dim, 1e-20);
        const double emu    = dot(egd, mu_grad, dim);
        const double ee     = etas[i] * etas[j];
        for (int d = 0; d < dim; ++d)
          out_gb[d] += (mu_grad[d] - egd[d] * (filter * emu)) * ee;
      }
  for (int d = 0; d < dim; ++d)
    out[d] = out[d] + out_gb[d] * (2.0 * p->Mgb);
}

/* MobilityTensorial::apply_M_derivative, S/mobility.h:809-916.
 * lin_grad: [comp][dim] of the linearisation point; value/grad: increment. */
static void tensorial_apply_M_derivative(const orc_params *p,
                                         const double *lin_value,
                                         const double *lin_grad, int G,
                                         const double *value, const double *grad,
                                         int dim, double *out)
{
  const double  c           = lin_value[0];
  const double *lin_etas    = lin_value + 2;
  const double *lin_c_grad  = lin_grad;
  const double *lin_mu_grad = lin_grad + dim;
  const double *c_gradient  = grad;
  const double *mu_gradient = grad + dim;
  const double *etas_value  = value + 2;
  double        nc[3];
  orc_unit_vector(lin_c_grad, dim, 1e-10, nc);

  /* 1) M */
  {
    const double phi       = phi_clamped(c);
    const double f_vol_vap = p->Mvol * phi + p->Mvap * (1.0 - phi);
    const double fsurf     = p->Msurf * 4.0 * (c * c) * ((1. - c) * (1. - c));
    const double ncmu      = dot(nc, mu_gradient, dim);
    for (int d = 0; d < dim; ++d)
      out[d] = (f_vol_vap + fsurf) * mu_gradient[d] - nc[d] * (fsurf * ncmu);
  }
  /* 2) dM_dc */
  {
    const double c2_1minusc2 = 4.0 * c * c * (1. - c) * (1. - c);
    const double dphidc      = 30.0 * c2_1minusc2;
    const double f_vol_vap   = (p->Mvol - p->Mvap) * dphidc;
    const double fsurf       = p->Msurf * c2_1minusc2;
    double       dfsurf      = p->Msurf * 8.0 * c * (1. - c) * (1. - 2. * c);
    if (fsurf < 1e-6)
      dfsurf = 0.0;
    const double nclmu = dot(nc, lin_mu_grad, dim);
    for (int d = 0; d < dim; ++d)
      out[d] += ((f_vol_vap + dfsurf * projector_scale) * lin_mu_grad[d] -
                 nc[d] * (dfsurf * nclmu)) *
                value[0];
  }
  /* 3) dM_dgrad_c */
  {
    double fsurf = p->Msurf * 4.0 * (c * c) * ((1. - c) * (1. - c));
    double nrm   = sqrt(dot(lin_c_grad, lin_c_grad, dim));
    if (fsurf < 1e-6)
      fsurf = 0.0;
    if (nrm < 1e-4)
      nrm = 1.0;
    const double ncg = dot(nc, c_gradient, dim);
    double       temp[3] = {0.0, 0.0, 0.0};
    for (int d = 0; d < dim; ++d)
      temp[d] = c_gradient[d] - nc[d] * ncg;
    const double nclmu  = dot(nc, lin_mu_grad, dim);
    const double lmutmp = dot(lin_mu_grad, temp, dim);
    const double fac    = -fsurf / nrm;
    for (int d = 0; d < dim; ++d)
      out[d] += fac * (nclmu * temp[d] + nc[d] * lmutmp);
  }
  if (G <= 1)
    return;
  /* 4) GB part of M and dM_detai */
  double out_gb[3] = {0.0, 0.0, 0.0};
  for (int i = 0; i < G; ++i)
    for (int j = 0; j < i; ++j)
      {
        double egd[3];
        for (int d = 0; d < dim; ++d)
          egd[d] = lin_grad[(2 + i) * dim + d] - lin_grad[(2 + j) * dim + d];
        const double filter = orc_unit_vector_filter_2(egd, dim, 1e-20);
        const double emu    = dot(egd, mu_gradient, dim);
        const double ee     = lin_etas[i] * lin_etas[j];
        for (int d = 0; d < dim; ++d)
          out_gb[d] += (mu_gradient[d] - egd[d] * (filter * emu)) * ee;
        const double elmu = dot(egd, lin_mu_grad, dim);
        const double s2 = lin_etas[j] * etas_value[i] + lin_etas[i] * etas_value[j];
        for (int d = 0; d < dim; ++d)
          out_gb[d] += (lin_mu_grad[d] - (filter * egd[d]) * elmu) * s2;
      }
  for (int d = 0; d < dim; ++d)
    out[d] = out[d] + out_gb[d] * (2.0 * p->Mgb);
}

/* ------------------------------------------------------------------ */
/* quadrature-point operators                                           */
/* ------------------------------------------------------------------ */

/* SinteringOperatorGenericResidualQuad::operator(),
 * S/operator_sintering_generic.h:254-308 */
void orc_residual_qp(const orc_params *p, int wtd, const double *value,
                     const double *grad, const double *hist, double *vr,
                     double *gr)
{
  const int     nc  = p->n_comp;
  const int     G   = nc - 2;
  const int     dim = p->dim;
  const double *etas = value + 2;
  fe_eval       fe;
  fe_init(&fe, p->A, p->B, value[0], etas, G, 1);

  /* 1) c row */
  vr[0] = 0.0;
  if (wtd >= 1)
    vr[0] = value[0] * p->w0;
  if (wtd == 2)
    vr[0] += hist[0];
  if (p->tensorial)
    tensorial_apply_M(p, value, grad, G, dim, gr);
  else
    scalar_apply_M(p, value[0], etas, G, grad + dim, dim, gr);

  /* 2) mu row */
  vr[1] = p->w_alg * value[1] - fe_df_dc(&fe);
  for (int d = 0; d < dim; ++d)
    gr[dim + d] = -p->kappa_c * grad[d];

  /* 3) eta rows */
  for (int ig = 0; ig < G; ++ig)
    {
      vr[2 + ig] = p->L * fe_df_detai(&fe, etas[ig]);
      if (wtd >= 1)
        vr[2 + ig] += value[2 + ig] * p->w0;
      if (wtd == 2)
        vr[2 + ig] += hist[2 + ig];
      for (int d = 0; d < dim; ++d)
        gr[(2 + ig) * dim + d] = p->L * p->kappa_p * grad[(2 + ig) * dim + d];
    }
}

/* SinteringOperatorGenericQuad::operator(),
 * S/operator_sintering_generic.h:104-190 (advection disabled) */
void orc_tangent_qp(const orc_params *p, const double *lin_value,
                    const double *lin_grad, const double *value,
                    const double *grad, double *vr, double *gr)
{
  const int     nc       = p->n_comp;
  const int     G        = nc - 2;
  const int     dim      = p->dim;
  const double *lin_etas = lin_value + 2;
  fe_eval       fe;
  fe_init(&fe, p->A, p->B, lin_value[0], lin_etas, G, 0); /* EnergySecond */

  /* 1) c row */
  vr[0] = value[0] * p->w0;
  if (p->tensorial)
    tensorial_apply_M_derivative(p, lin_value, lin_grad, G, value, grad, dim, gr);
  else
    scalar_apply_M_derivative(p, lin_value, lin_grad + dim, G, value, grad + dim,
                              dim, gr);

  /* 2) mu row */
  vr[1] = p->w_alg * value[1] - fe_d2f_dc2(&fe) * value[0];
  for (int ig = 0; ig < G; ++ig)
    vr[1] -= fe_d2f_dcdetai(&fe, lin_etas[ig]) * value[ig + 2];
  for (int d = 0; d < dim; ++d)
    gr[dim + d] = -p->kappa_c * grad[d];

  /* 3) eta rows */
  for (int ig = 0; ig < G; ++ig)
    {
      vr[ig + 2] = 0.0;
      vr[ig + 2] += value[ig + 2] * p->w0 +
                    p->L * (fe_d2f_dcdetai(&fe, lin_etas[ig]) * value[0] +
                            fe_d2f_detai2(&fe, lin_etas[ig]) * value[ig + 2]);
      for (int d = 0; d < dim; ++d)
        gr[(ig + 2) * dim + d] = p->L * p->kappa_p * grad[(ig + 2) * dim + d];
    }
  apply_d2f_detaidetaj(p->L, p->B, lin_etas, G, value + 2, vr + 2);
}

/* ------------------------------------------------------------------ */
/* FE machinery (deal.II FE_Q / QGauss on an affine Cartesian cell)     */
/* ------------------------------------------------------------------ */

/* Gauss-Legendre points/weights on [0,1] by Newton iteration on P_n */
static void gauss_legendre(int n, double *x, double *w)
{
  for (int i = 0; i < n; ++i)
    {
      double z = cos(M_PI * (i + 0.75) / (n + 0.5));
      for (int it = 0; it < 100; ++it)
        {
          double p0 = 1.0, p1 = 0.0;
          for (int k = 1; k <= n; ++k)
            {
              const double p2 = p1;
              p1              = p0;
              p0              = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p2) / k;
            }
          const double dp = n * (z * p0 - p1) / (z * z - 1.0);
          const double dz = p0 / dp;
          z -= dz;
          if (fabs(dz) < 1e-17)
            break;
        }
      double p0 = 1.0, p1 = 0.0;
      for (int k = 1; k <= n; ++k)
        {
          const double p2 = p1;
          p1              = p0;
          p0              = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p2) / k;
        }
      const double dp = n * (z * p0 - p1) / (z * z - 1.0);
      /* z descends with i, so x = (1-z)/2 ascends; [-1,1] weight 2/((1-z^2)P'^2)
       * halves on [0,1] */
      x[i] = 0.5 * (1.0 - z);
      w[i] = 1.0 / ((1.0 - z * z) * dp * dp);
    }
}

/* Lagrange basis on {0,1} (p=1) or {0,1/2,1} (p=2) */
static double lagrange(int p, int i, double x)
{
  if (p == 1)
    return i == 0 ? 1.0 - x : x;
  /* p == 2, nodes 0, 0.5, 1 */
  if (i == 0)
    return (x - 0.5) * (x - 1.0) / ((0.0 - 0.5) * (0.0 - 1.0));
  if (i == 1)
    return (x - 0.0) * (x - 1.0) / ((0.5 - 0.0) * (0.5 - 1.0));
  return (x - 0.0) * (x - 0.5) / ((1.0 - 0.0) * (1.0 - 0.5));
}

static double lagrange_d(int p, int i, double x)
{
  if (p == 1)
    return i == 0 ? -1.0 : 1.0;
  if (i == 0)
    return ((x - 1.0) + (x - 0.5)) / ((0.0 - 0.5) * (0.0 - 1.0));
  if (i == 1)
    return ((x - 1.0) + (x - 0.0)) / ((0.5 - 0.0) * (0.5 - 1.0));
  return ((x - 0.5) + (x - 0.0)) / ((1.0 - 0.0) * (1.0 - 0.5));
}

typedef struct
{
  int    dim, p, n1, nq1, nl, nq;
  double S[MAXQ][MAXNL];        /* shape values */
  double G[MAXQ][MAXNL][3];     /* physical shape gradients */
  double JxW[MAXQ];
  int    nn[3];                 /* nodes per direction */
  int    nc[3];                 /* cells per direction */
} fe_tables;

static void make_tables(const orc_params *pr, fe_tables *t)
{
  memset(t, 0, sizeof(*t));
  t->dim = pr->dim;
  t->p   = pr->degree;
  t->n1  = pr->degree + 1;
  t->nq1 = pr->degree + 1;
  t->nl  = 1;
  t->nq  = 1;
  for (int d = 0; d < pr->dim; ++d)
    {
      t->nl *= t->n1;
      t->nq *= t->nq1;
    }
  for (int d = 0; d < 3; ++d)
    {
      t->nc[d] = d < pr->dim ? pr->n_cells[d] : 1;
      t->nn[d] = d < pr->dim ? pr->degree * pr->n_cells[d] + 1 : 1;
    }
  double xq[MAXN1], wq[MAXN1];
  gauss_legendre(t->nq1, xq, wq);
  double vol = 1.0;
  for (int d = 0; d < pr->dim; ++d)
    vol *= pr->h[d];
  for (int q = 0; q < t->nq; ++q)
    {
      int qi[3] = {0, 0, 0}, r = q;
      for (int d = 0; d < pr->dim; ++d)
        {
          qi[d] = r % t->nq1;
          r /= t->nq1;
        }
      double w = 1.0;
      for (int d = 0; d < pr->dim; ++d)
        w *= wq[qi[d]];
      t->JxW[q] = vol * w;
      for (int k = 0; k < t->nl; ++k)
        {
          int ki[3] = {0, 0, 0}, s = k;
          for (int d = 0; d < pr->dim; ++d)
            {
              ki[d] = s % t->n1;
              s /= t->n1;
            }
          double val = 1.0;
          for (int d = 0; d < pr->dim; ++d)
            val *= lagrange(t->p, ki[d], xq[qi[d]]);
          t->S[q][k] = val;
          for (int dd = 0; dd < pr->dim; ++dd)
            {
              double g = 1.0;
              for (int d = 0; d < pr->dim; ++d)
                g *= (d == dd) ? lagrange_d(t->p, ki[d], xq[qi[d]]) / pr->h[d]
                               : lagrange(t->p, ki[d], xq[qi[d]]);
              t->G[q][k][dd] = g;
            }
        }
    }
}

int64_t orc_n_nodes(const orc_params *p)
{
  int64_t n = 1;
  for (int d = 0; d < p->dim; ++d)
    n *= (int64_t)(p->degree * p->n_cells[d] + 1);
  return n;
}

int64_t orc_n_cells(const orc_params *p)
{
  int64_t n = 1;
  for (int d = 0; d < p->dim; ++d)
    n *= p->n_cells[d];
  return n;
}

int orc_n_q_points(const orc_params *p)
{
  int n = 1;
  for (int d = 0; d < p->dim; ++d)
    n *= p->degree + 1;
  return n;
}

static void cell_dofs(const fe_tables *t, int cx, int cy, int cz, int64_t *idx)
{
  for (int k = 0; k < t->nl; ++k)
    {
      int kx = k % t->n1, ky = (k / t->n1) % t->n1, kz = k / (t->n1 * t->n1);
      if (t->dim == 2)
        kz = 0;
      const int64_t gx = (int64_t)t->p * cx + kx;
      const int64_t gy = (int64_t)t->p * cy + ky;
      const int64_t gz = (int64_t)t->p * cz + kz;
      idx[k] = gx + (int64_t)t->nn[0] * (gy + (int64_t)t->nn[1] * gz);
    }
}

static int64_t cell_index(const fe_tables *t, int cx, int cy, int cz)
{
  return cx + (int64_t)t->nc[0] * (cy + (int64_t)t->nc[1] * cz);
}

/* interpolate local nodal values u_loc[k*nc + c] to value[c], grad[c*dim+d] */
static void interpolate(const fe_tables *t, int ncomp, int q, const double *u_loc,
                        double *value, double *grad)
{
  for (int c = 0; c < ncomp; ++c)
    {
      double v = 0.0, g[3] = {0.0, 0.0, 0.0};
      for (int k = 0; k < t->nl; ++k)
        {
          const double u = u_loc[k * ncomp + c];
          v += t->S[q][k] * u;
          for (int d = 0; d < t->dim; ++d)
            g[d] += t->G[q][k][d] * u;
        }
      value[c] = v;
      if (grad)
        for (int d = 0; d < t->dim; ++d)
          grad[c * t->dim + d] = g[d];
    }
}

/* r_loc[k*nc+c] += JxW (S vr + G . gr) */
static void integrate_add(const fe_tables *t, int ncomp, int q, const double *vr,
                          const double *gr, double *r_loc)
{
  const double jxw = t->JxW[q];
  for (int k = 0; k < t->nl; ++k)
    for (int c = 0; c < ncomp; ++c)
      {
        double s = t->S[q][k] * vr[c];
        for (int d = 0; d < t->dim; ++d)
          s += t->G[q][k][d] * gr[c * t->dim + d];
        r_loc[k * ncomp + c] += jxw * s;
      }
}

enum
{
  OP_RESIDUAL,
  OP_VMULT,
  OP_FILL,
  OP_BLOCKS
};

typedef struct
{
  const orc_params   *p;
  const fe_tables    *t;
  int                 op;
  int                 wtd;
  const double       *src;
  const double *const *old;
  double             *dst;
  const double       *cache_in;
  double             *cache_out;
  const uint32_t     *masks; /* grain cut-off: active order parameters per cell */
} loop_args;

static void do_cell(const loop_args *a, int cx, int cy, int cz)
{
  const orc_params *p   = a->p;
  const fe_tables  *t   = a->t;
  const int         nc  = p->n_comp;
  const int         dim = p->dim;
  int64_t           idx[MAXNL];
  cell_dofs(t, cx, cy, cz, idx);
  const int64_t cell = cell_index(t, cx, cy, cz);
  const int     fsz  = nc * (1 + dim); /* cache fields per q */

  double u_loc[MAXNL * ORC_MAX_COMP], r_loc[MAXNL * ORC_MAX_COMP];
  double value[ORC_MAX_COMP], grad[ORC_MAX_COMP * 3];
  double vr[ORC_MAX_COMP], gr[ORC_MAX_COMP * 3];

  if (a->op == OP_BLOCKS)
    {
      /* probe every local dof (k, j) with a unit vector */
      for (int k = 0; k < t->nl; ++k)
        for (int j = 0; j < nc; ++j)
          {
            double col[ORC_MAX_COMP];
            for (int i = 0; i < nc; ++i)
              col[i] = 0.0;
            for (int q = 0; q < t->nq; ++q)
              {
                const double *lin = a->cache_in + (cell * t->nq + q) * fsz;
                for (int i = 0; i < nc; ++i)
                  {
                    value[i] = 0.0;
                    for (int d = 0; d < dim; ++d)
                      grad[i * dim + d] = 0.0;
                  }
                value[j] = t->S[q][k];
                for (int d = 0; d < dim; ++d)
                  grad[j * dim + d] = t->G[q][k][d];
                orc_tangent_qp(p, lin, lin + nc, value, grad, vr, gr);
                for (int i = 0; i < nc; ++i)
                  {
                    double s = t->S[q][k] * vr[i];
                    for (int d = 0; d < dim; ++d)
                      s += t->G[q][k][d] * gr[i * dim + d];
                    col[i] += t->JxW[q] * s;
                  }
              }
            for (int i = 0; i < nc; ++i)
              a->dst[(idx[k] * nc + i) * nc + j] += col[i];
          }
      return;
    }

  for (int k = 0; k < t->nl; ++k)
    for (int c = 0; c < nc; ++c)
      u_loc[k * nc + c] = a->src[idx[k] * nc + c];

  if (a->op == OP_FILL)
    {
      for (int q = 0; q < t->nq; ++q)
        {
          double *out = a->cache_out + (cell * t->nq + q) * fsz;
          interpolate(t, nc, q, u_loc, out, out + nc);
        }
      return;
    }

  double h_loc[MAXNL * ORC_MAX_COMP];
  if (a->op == OP_RESIDUAL && a->wtd == 2)
    {
      /* do_evalute_history, S/operator_sintering_generic.h:927-975 */
      for (int i = 0; i < p->order; ++i)
        for (int k = 0; k < t->nl; ++k)
          for (int c = 0; c < nc; ++c)
            {
              const double v = a->old[i][idx[k] * nc + c] * p->w_hist[i];
              if (i == 0)
                h_loc[k * nc + c] = v;
              else
                h_loc[k * nc + c] += v;
            }
    }

  for (int k = 0; k < t->nl * nc; ++k)
    r_loc[k] = 0.0;

  for (int q = 0; q < t->nq; ++q)
    {
      interpolate(t, nc, q, u_loc, value, grad);
      if (a->op == OP_RESIDUAL)
        {
          double hist[ORC_MAX_COMP];
          if (a->wtd == 2)
            interpolate(t, nc, q, h_loc, hist, NULL);
          orc_residual_qp(p, a->wtd, value, grad, hist, vr, gr);
        }
      else
        {
          const double *lin = a->cache_in + (cell * t->nq + q) * fsz;
          orc_tangent_qp(p, lin, lin + nc, value, grad, vr, gr);
        }
      integrate_add(t, nc, q, vr, gr, r_loc);
    }

  for (int k = 0; k < t->nl; ++k)
    for (int c = 0; c < nc; ++c)
      a->dst[idx[k] * nc + c] += r_loc[k * nc + c];
}

/* One cell of the grain cut-off (do_vmult_range_no_template and
 * do_evaluate_nonlinear_residual_nt, S/operator_sintering_generic.h:456-537,
 * 1008-1081): the cell is evaluated as a system of c, mu and its active order
 * parameters only (bit g of masks[cell]); the linearisation values and
 * gradients of the others are left out as the compacted cache of
 * SinteringOperatorData::fill_quadrature_point_values leaves them out
 * (S/operator_sintering_data.h:360-395), and their rows get nothing from
 * the cell. */
static void do_cell_cut(const loop_args *a, int cx, int cy, int cz)
{
  const orc_params *p   = a->p;
  const fe_tables  *t   = a->t;
  const int         nc  = p->n_comp;
  const int         dim = p->dim;
  int64_t           idx[MAXNL];
  cell_dofs(t, cx, cy, cz, idx);
  const int64_t  cell = cell_index(t, cx, cy, cz);
  const uint32_t m    = a->masks[cell];
  int            L[ORC_MAX_COMP], nl = 0;
  L[nl++] = 0;
  L[nl++] = 1;
  for (int g = 0; g < nc - 2; ++g)
    if ((m >> g) & 1u)
      L[nl++] = 2 + g;
  orc_params pr = *p;
  pr.n_comp     = nl;
  const int fsz = nc * (1 + dim); /* fields per q of the full cache */

  double lv[ORC_MAX_COMP], lg[ORC_MAX_COMP * 3];
  double value[ORC_MAX_COMP], grad[ORC_MAX_COMP * 3];
  double vr[ORC_MAX_COMP], gr[ORC_MAX_COMP * 3];
  /* the linearisation point at q, reduced to the cell's components */
#define ORC_REDUCED_LIN(q)                                              \
  do                                                                    \
    {                                                                   \
      const double *lin_ = a->cache_in + (cell * t->nq + (q)) * fsz;    \
      for (int i_ = 0; i_ < nl; ++i_)                                   \
        {                                                               \
          lv[i_] = lin_[L[i_]];                                         \
          for (int d_ = 0; d_ < dim; ++d_)                              \
            lg[i_ * dim + d_] = lin_[nc + L[i_] * dim + d_];            \
        }                                                               \
    }                                                                   \
  while (0)

  if (a->op == OP_BLOCKS)
    {
      for (int k = 0; k < t->nl; ++k)
        for (int jj = 0; jj < nl; ++jj)
          {
            double col[ORC_MAX_COMP];
            for (int i = 0; i < nl; ++i)
              col[i] = 0.0;
            for (int q = 0; q < t->nq; ++q)
              {
                ORC_REDUCED_LIN(q);
                for (int i = 0; i < nl; ++i)
                  {
                    value[i] = 0.0;
                    for (int d = 0; d < dim; ++d)
                      grad[i * dim + d] = 0.0;
                  }
                value[jj] = t->S[q][k];
                for (int d = 0; d < dim; ++d)
                  grad[jj * dim + d] = t->G[q][k][d];
                orc_tangent_qp(&pr, lv, lg, value, grad, vr, gr);
                for (int i = 0; i < nl; ++i)
                  {
                    double s = t->S[q][k] * vr[i];
                    for (int d = 0; d < dim; ++d)
                      s += t->G[q][k][d] * gr[i * dim + d];
                    col[i] += t->JxW[q] * s;
                  }
              }
            for (int i = 0; i < nl; ++i)
              a->dst[(idx[k] * nc + L[i]) * nc + L[jj]] += col[i];
          }
      return;
    }

  double u_loc[MAXNL * ORC_MAX_COMP], h_loc[MAXNL * ORC_MAX_COMP];
  double r_loc[MAXNL * ORC_MAX_COMP];
  for (int k = 0; k < t->nl; ++k)
    for (int i = 0; i < nl; ++i)
      u_loc[k * nl + i] = a->src[idx[k] * nc + L[i]];
  if (a->op == OP_RESIDUAL && a->wtd == 2)
    for (int k = 0; k < t->nl; ++k)
      for (int i = 0; i < nl; ++i)
        {
          double v = 0.0;
          for (int o = 0; o < p->order; ++o)
            {
              const double w = a->old[o][idx[k] * nc + L[i]] * p->w_hist[o];
              v              = o == 0 ? w : v + w;
            }
          h_loc[k * nl + i] = v;
        }
  for (int k = 0; k < t->nl * nl; ++k)
    r_loc[k] = 0.0;
  for (int q = 0; q < t->nq; ++q)
    {
      interpolate(t, nl, q, u_loc, value, grad);
      if (a->op == OP_RESIDUAL)
        {
          double hist[ORC_MAX_COMP];
          if (a->wtd == 2)
            interpolate(t, nl, q, h_loc, hist, NULL);
          orc_residual_qp(&pr, a->wtd, value, grad, hist, vr, gr);
        }
      else
        {
          ORC_REDUCED_LIN(q);
          orc_tangent_qp(&pr, lv, lg, value, grad, vr, gr);
        }
      integrate_add(t, nl, q, vr, gr, r_loc);
    }
#undef ORC_REDUCED_LIN
  for (int k = 0; k < t->nl; ++k)
    for (int i = 0; i < nl; ++i)
      a->dst[idx[k] * nc + L[i]] += r_loc[k * nl + i];
}

/* cell loop in 2^dim parity colours: cells of one colour share no node, so
 * the scatter-add is race-free and the result independent of the number of
 * threads */
static void cell_loop(const loop_args *a)
{
  const fe_tables *t       = a->t;
  const int        ncolors = 1 << t->dim;
  for (int color = 0; color < ncolors; ++color)
    {
      const int ox = color & 1, oy = (color >> 1) & 1, oz = (color >> 2) & 1;
      const int mx = (t->nc[0] - ox + 1) / 2;
      const int my = (t->nc[1] - oy + 1) / 2;
      const int mz = t->dim == 3 ? (t->nc[2] - oz + 1) / 2 : 1;
      const int64_t total = (int64_t)mx * my * mz;
#pragma omp parallel for schedule(static)
      for (int64_t i = 0; i < total; ++i)
        {
          const int ix = (int)(i % mx);
          const int iy = (int)((i / mx) % my);
          const int iz = (int)(i / ((int64_t)mx * my));
          const int cz = t->dim == 3 ? 2 * iz + oz : 0;
          if (a->masks)
            do_cell_cut(a, 2 * ix + ox, 2 * iy + oy, cz);
          else
            do_cell(a, 2 * ix + ox, 2 * iy + oy, cz);
        }
    }
}

static void zero(double *v, int64_t n)
{
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i)
    v[i] = 0.0;
}

void orc_residual(const orc_params *p, int wtd, const double *src,
                  const double *const *old, double *dst)
{
  fe_tables t;
  make_tables(p, &t);
  zero(dst, orc_n_nodes(p) * p->n_comp);
  loop_args a = {p, &t, OP_RESIDUAL, wtd, src, old, dst, NULL, NULL, NULL};
  cell_loop(&a);
}

void orc_fill_qp(const orc_params *p, const double *lin, double *cache)
{
  fe_tables t;
  make_tables(p, &t);
  loop_args a = {p, &t, OP_FILL, 0, lin, NULL, NULL, NULL, cache, NULL};
  cell_loop(&a);
}

void orc_vmult(const orc_params *p, const double *cache, const double *src,
               double *dst)
{
  fe_tables t;
  make_tables(p, &t);
  zero(dst, orc_n_nodes(p) * p->n_comp);
  loop_args a = {p, &t, OP_VMULT, 0, src, NULL, dst, cache, NULL, NULL};
  cell_loop(&a);
}

void orc_node_blocks(const orc_params *p, const double *cache, double *blocks)
{
  fe_tables t;
  make_tables(p, &t);
  zero(blocks, orc_n_nodes(p) * p->n_comp * p->n_comp);
  loop_args a = {p, &t, OP_BLOCKS, 0, NULL, NULL, blocks, cache, NULL, NULL};
  cell_loop(&a);
}

void orc_dense_tangent(const orc_params *p, const double *cache, double *mat)
{
  const int64_t n   = orc_n_nodes(p) * p->n_comp;
  double       *e   = (double *)calloc((size_t)n, sizeof(double));
  double       *col = (double *)malloc((size_t)n * sizeof(double));
  for (int64_t j = 0; j < n; ++j)
    {
      e[j] = 1.0;
      orc_vmult(p, cache, e, col);
      e[j] = 0.0;
      for (int64_t i = 0; i < n; ++i)
        mat[i * n + j] = col[i];
    }
  free(e);
  free(col);
}

/* ---- grain cut-off ---- */

void orc_cell_masks(const orc_params *p, const double *u, double tol, uint32_t *masks)
{
  fe_tables t;
  make_tables(p, &t);
  const int     nc = p->n_comp;
  const int64_t n  = orc_n_cells(p);
#pragma omp parallel for schedule(static)
  for (int64_t c = 0; c < n; ++c)
    {
      const int cx = (int)(c % t.nc[0]);
      const int cy = (int)((c / t.nc[0]) % t.nc[1]);
      const int cz = p->dim == 3 ? (int)(c / ((int64_t)t.nc[0] * t.nc[1])) : 0;
      int64_t   idx[MAXNL];
      cell_dofs(&t, cx, cy, cz, idx);
      uint32_t m = 0u;
      for (int g = 0; g < nc - 2; ++g)
        {
          double linf = 0.0;
          for (int k = 0; k < t.nl; ++k)
            {
              const double v = fabs(u[idx[k] * nc + 2 + g]);
              linf           = v > linf ? v : linf;
            }
          if (linf > tol)
            m |= 1u << g;
        }
      masks[c] = m;
    }
}

void orc_vmult_cut(const orc_params *p, const double *cache, const uint32_t *masks,
                   const double *src, double *dst)
{
  fe_tables t;
  make_tables(p, &t);
  zero(dst, orc_n_nodes(p) * p->n_comp);
  loop_args a = {p, &t, OP_VMULT, 0, src, NULL, dst, cache, NULL, masks};
  cell_loop(&a);
}

void orc_residual_cut(const orc_params *p, int wtd, const uint32_t *masks, const double *src,
                      const double *const *old, double *dst)
{
  fe_tables t;
  make_tables(p, &t);
  zero(dst, orc_n_nodes(p) * p->n_comp);
  loop_args a = {p, &t, OP_RESIDUAL, wtd, src, old, dst, NULL, NULL, masks};
  cell_loop(&a);
}

void orc_node_blocks_cut(const orc_params *p, const double *cache, const uint32_t *masks,
                         double *blocks)
{
  fe_tables t;
  make_tables(p, &t);
  zero(blocks, orc_n_nodes(p) * p->n_comp * p->n_comp);
  loop_args a = {p, &t, OP_BLOCKS, 0, NULL, NULL, blocks, cache, NULL, masks};
  cell_loop(&a);
}

/* ---- incomplete LU without fill (ILU(0)), for the preconditioner study:
 * the reference's default preconditioner of the Cahn-Hilliard block is
 * Trilinos' ILU (BlockPreconditioner2, S/preconditioners.h:983-1282,
 * PreconditionerData default "ILU", S/preconditioner_parameters.h:57-64).
 * CSR with sorted column indices and a stored diagonal; factors in place
 * (unit-lower L below the diagonal, U on and above it). */
int orc_ilu0(int64_t n, const int64_t *rowptr, const int64_t *col, double *val)
{
  int64_t *diag = (int64_t *)malloc((size_t)n * sizeof(int64_t));
  int64_t *pos  = (int64_t *)malloc((size_t)n * sizeof(int64_t));
  if (!diag || !pos)
    {
      free(diag);
      free(pos);
      return -1;
    }
  for (int64_t i = 0; i < n; ++i)
    {
      diag[i] = -1;
      pos[i]  = -1;
    }
  for (int64_t i = 0; i < n; ++i)
    for (int64_t k = rowptr[i]; k < rowptr[i + 1]; ++k)
      if (col[k] == i)
        diag[i] = k;
  int status = 0;
  for (int64_t i = 0; i < n && status == 0; ++i)
    {
      for (int64_t k = rowptr[i]; k < rowptr[i + 1]; ++k)
        pos[col[k]] = k;
      for (int64_t k = rowptr[i]; k < rowptr[i + 1] && col[k] < i; ++k)
        {
          const int64_t j = col[k];
          if (diag[j] < 0 || val[diag[j]] == 0.0)
            {
              status = -2;
              break;
            }
          const double f = val[k] / val[diag[j]];
          val[k]         = f;
          for (int64_t m = diag[j] + 1; m < rowptr[j + 1]; ++m)
            if (pos[col[m]] >= 0)
              val[pos[col[m]]] -= f * val[m];
        }
      for (int64_t k = rowptr[i]; k < rowptr[i + 1]; ++k)
        pos[col[k]] = -1;
      if (diag[i] < 0)
        status = -2;
    }
  free(diag);
  free(pos);
  return status;
}

/* x = (LU)^-1 b with the factors of orc_ilu0 */
void orc_ilu0_solve(int64_t n, const int64_t *rowptr, const int64_t *col, const double *val,
                    const double *b, double *x)
{
  for (int64_t i = 0; i < n; ++i)
    {
      double s = b[i];
      for (int64_t k = rowptr[i]; k < rowptr[i + 1] && col[k] < i; ++k)
        s -= val[k] * x[col[k]];
      x[i] = s;
    }
  for (int64_t i = n - 1; i >= 0; --i)
    {
      double  s = x[i];
      int64_t d = -1;
      for (int64_t k = rowptr[i]; k < rowptr[i + 1]; ++k)
        if (col[k] > i)
          s -= val[k] * x[col[k]];
        else if (col[k] == i)
          d = k;
      x[i] = s / val[d];
    }
}
