/*
 * cpu_fast.c -- CPU restatement of the tangent vmult for the CPU baseline
 * of the benchmark.
 *
 * NOT THE PRODUCT and not the checker: the timed CPU baseline beside the
 * GPU numbers (bench.py's cpu_baseline), checked against the oracle in
 * tests/test_cpu_fast.py.  It restates the reference's matrix-free
 * algorithm the way deal.II runs it on a CPU
 * (S/operator_sintering_generic.h:539-606 with FEEvaluation, S/ =
 * applications/sintering/include/pf-applications/sintering/):
 *   - quadrature-point linearisation cache filled once (the QP-cache
 *     variant, SinteringOperatorData::fill_quadrature_point_values), holding
 *     c, eta_i and grad mu per point;
 *   - cells processed in batches of VW cells, one SIMD lane per cell
 *     (deal.II's VectorizedArray), with GCC vector types;
 *   - sum-factorised 2-point Gauss interpolation and integration (Q1);
 *   - OpenMP over batches of one colour (cells of one colour share no node,
 *     so the scatter-add is race-free);
 *   - the scalar-mobility quadrature-point tangent of
 *     S/operator_sintering_generic.h:104-190 with the closed forms of
 *     S/free_energy.h and S/mobility.h:441-534.
 * 3D, FE_Q(1), scalar mobility, up to 14 order parameters.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef CFB_VW
#define CFB_VW 4 /* lanes (cells) per batch */
#endif
#ifndef CFB_NAME
#define CFB_NAME(x) cfb_##x
#endif

typedef double vd __attribute__((vector_size(CFB_VW * 8)));

#define MAXC ORC_MAX_COMP
#define MAXG (ORC_MAX_COMP - 2)

typedef struct
{
  orc_params p;
  int        nc, G, nf;       /* components, order parameters, cache fields */
  int        ncell[3], nn[3];
  int        nb[8];           /* batches per x-row of each colour */
  int64_t    boff[9];         /* first batch of each colour */
  vd        *cache;           /* [batch][q][field] */
  double     sa, sb, ih[3], jxw;
} cfb_op;

static inline vd bc(double x)
{
  vd v;
  for (int l = 0; l < CFB_VW; ++l)
    v[l] = x;
  return v;
}

/* batch b of colour col: its cells (cx of lane 0, cy, cz) and lane count */
static void batch_cells(const cfb_op *o, int col, int64_t b, int *cx0, int *cy, int *cz,
                        int *lanes)
{
  const int ox = col & 1, oy = (col >> 1) & 1, oz = (col >> 2) & 1;
  const int mx = (o->ncell[0] - ox + 1) / 2;
  const int my = (o->ncell[1] - oy + 1) / 2;
  const int nb = o->nb[col];
  const int ib = (int)(b % nb);
  const int64_t r = b / nb;
  *cy  = oy + 2 * (int)(r % my);
  *cz  = oz + 2 * (int)(r / my);
  *cx0 = ox + 2 * ib * CFB_VW;
  const int first = ib * CFB_VW;
  *lanes = mx - first < CFB_VW ? mx - first : CFB_VW;
}

/* gathers the 8 nodal values of each component for the batch's cells */
static void gather(const cfb_op *o, const double *u, int cx0, int cy, int cz, int lanes,
                   vd loc[][8])
{
  const int nc = o->nc;
  for (int c = 0; c < nc; ++c)
    for (int k = 0; k < 8; ++k)
      loc[c][k] = bc(0.0);
  for (int l = 0; l < lanes; ++l)
    {
      const int cx = cx0 + 2 * l;
      for (int k = 0; k < 8; ++k)
        {
          const int64_t node = ((int64_t)(cz + (k >> 2)) * o->nn[1] + (cy + ((k >> 1) & 1))) *
                                 o->nn[0] + (cx + (k & 1));
          const double *x = u + node * nc;
          for (int c = 0; c < nc; ++c)
            loc[c][k][l] = x[c];
        }
    }
}

/* 1D value (a, b) or derivative ((in1 - in0) * ih) along axis ax of a
 * 2x2x2 array (index x + 2y + 4z) */
static inline void lvl_v(const vd *in, vd *out, int ax, double a, double b)
{
  const int s = 1 << ax;
  for (int i = 0; i < 8; ++i)
    if (!(i & s))
      {
        const vd u0 = in[i], u1 = in[i + s];
        out[i]     = a * u0 + b * u1;
        out[i + s] = b * u0 + a * u1;
      }
}
static inline void lvl_d(const vd *in, vd *out, int ax, double ih)
{
  const int s = 1 << ax;
  for (int i = 0; i < 8; ++i)
    if (!(i & s))
      {
        const vd d = (in[i + s] - in[i]) * ih;
        out[i]     = d;
        out[i + s] = d;
      }
}
/* transposes: the value matrix is symmetric; D^T (g0, g1) = (-, +)(g0+g1) ih */
static inline void lvl_dt(const vd *in, vd *out, int ax, double ih)
{
  const int s = 1 << ax;
  for (int i = 0; i < 8; ++i)
    if (!(i & s))
      {
        const vd t = (in[i] + in[i + s]) * ih;
        out[i]     = -t;
        out[i + s] = t;
      }
}

static void interp(const cfb_op *o, const vd *u, vd *val, vd (*grad)[8])
{
  vd X[8], Y[8], t[8];
  lvl_v(u, X, 0, o->sa, o->sb);     /* V_x u */
  lvl_v(X, Y, 1, o->sa, o->sb);     /* V_y V_x u */
  lvl_v(Y, val, 2, o->sa, o->sb);
  lvl_d(Y, grad[2], 2, o->ih[2]);
  lvl_d(X, t, 1, o->ih[1]);
  lvl_v(t, grad[1], 2, o->sa, o->sb);
  lvl_d(u, t, 0, o->ih[0]);
  lvl_v(t, X, 1, o->sa, o->sb);
  lvl_v(X, grad[0], 2, o->sa, o->sb);
}

/* nodal contributions of value rows vr[q] and gradient rows gr[d][q] */
static void integrate(const cfb_op *o, const vd *vr, vd (*gr)[8], vd *out)
{
  vd Tv[8], Tx[8], Ty[8], a[8], Uv[8], Ux[8];
  lvl_v(vr, Tv, 2, o->sa, o->sb);
  lvl_dt(gr[2], a, 2, o->ih[2]);
  for (int i = 0; i < 8; ++i)
    Tv[i] += a[i];
  lvl_v(gr[0], Tx, 2, o->sa, o->sb);
  lvl_v(gr[1], Ty, 2, o->sa, o->sb);
  lvl_v(Tv, Uv, 1, o->sa, o->sb);
  lvl_dt(Ty, a, 1, o->ih[1]);
  for (int i = 0; i < 8; ++i)
    Uv[i] += a[i];
  lvl_v(Tx, Ux, 1, o->sa, o->sb);
  lvl_v(Uv, out, 0, o->sa, o->sb);
  lvl_dt(Ux, a, 0, o->ih[0]);
  for (int i = 0; i < 8; ++i)
    out[i] += a[i];
}

void *CFB_NAME(create)(const orc_params *p)
{
  if (p->dim != 3 || p->degree != 1 || p->tensorial || p->n_comp < 3 || p->n_comp > MAXC)
    return NULL;
  cfb_op *o = (cfb_op *)calloc(1, sizeof(cfb_op));
  o->p  = *p;
  o->nc = p->n_comp;
  o->G  = p->n_comp - 2;
  o->nf = 1 + o->G + 3;
  for (int d = 0; d < 3; ++d)
    {
      o->ncell[d] = p->n_cells[d];
      o->nn[d]    = p->n_cells[d] + 1;
      o->ih[d]    = 1.0 / p->h[d];
    }
  o->sa  = 0.5 + 0.5 / sqrt(3.0);
  o->sb  = 0.5 - 0.5 / sqrt(3.0);
  o->jxw = p->h[0] * p->h[1] * p->h[2] / 8.0;
  o->boff[0] = 0;
  for (int col = 0; col < 8; ++col)
    {
      const int ox = col & 1, oy = (col >> 1) & 1, oz = (col >> 2) & 1;
      const int mx = (o->ncell[0] - ox + 1) / 2;
      const int my = (o->ncell[1] - oy + 1) / 2;
      const int mz = (o->ncell[2] - oz + 1) / 2;
      o->nb[col]       = (mx + CFB_VW - 1) / CFB_VW;
      o->boff[col + 1] = o->boff[col] + (int64_t)o->nb[col] * my * mz;
    }
  o->cache = (vd *)aligned_alloc(64, (size_t)o->boff[8] * 8 * o->nf * sizeof(vd));
  if (!o->cache)
    {
      free(o);
      return NULL;
    }
  return o;
}

void CFB_NAME(destroy)(void *h)
{
  cfb_op *o = (cfb_op *)h;
  if (o)
    free(o->cache);
  free(o);
}

int CFB_NAME(simd_width)(void)
{
  return CFB_VW;
}

/* linearisation cache: c, eta_i, grad mu at every quadrature point */
void CFB_NAME(fill)(void *h, const double *lin)
{
  cfb_op *o = (cfb_op *)h;
  for (int col = 0; col < 8; ++col)
    {
      const int64_t n = o->boff[col + 1] - o->boff[col];
#pragma omp parallel for schedule(static)
      for (int64_t b = 0; b < n; ++b)
        {
          int cx0, cy, cz, lanes;
          batch_cells(o, col, b, &cx0, &cy, &cz, &lanes);
          vd loc[MAXC][8], val[8], grad[3][8];
          gather(o, lin, cx0, cy, cz, lanes, loc);
          vd *cb = o->cache + (o->boff[col] + b) * 8 * o->nf;
          for (int c = 0; c < o->nc; ++c)
            {
              interp(o, loc[c], val, grad);
              for (int q = 0; q < 8; ++q)
                {
                  vd *f = cb + q * o->nf;
                  if (c == 0)
                    f[0] = val[q];
                  else if (c == 1)
                    for (int d = 0; d < 3; ++d)
                      f[1 + o->G + d] = grad[d][q];
                  else
                    f[1 + (c - 2)] = val[q];
                }
            }
        }
    }
}

void CFB_NAME(vmult)(void *h, const double *src, double *dst)
{
  cfb_op       *o   = (cfb_op *)h;
  const int     nc  = o->nc, G = o->G, nf = o->nf;
  const int64_t ndof = (int64_t)o->nn[0] * o->nn[1] * o->nn[2] * nc;
  const orc_params *p = &o->p;
  const double A = p->A, B = p->B, L = p->L, kc = p->kappa_c, kp = p->kappa_p;
  const double Mvol = p->Mvol, Mvap = p->Mvap, Msurf = p->Msurf, Mgb = p->Mgb;
  const double w0 = p->w0, walg = p->w_alg, jxw = o->jxw;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < ndof; ++i)
    dst[i] = 0.0;
  for (int col = 0; col < 8; ++col)
    {
      const int64_t n = o->boff[col + 1] - o->boff[col];
#pragma omp parallel for schedule(static)
      for (int64_t b = 0; b < n; ++b)
        {
          int cx0, cy, cz, lanes;
          batch_cells(o, col, b, &cx0, &cy, &cz, &lanes);
          vd loc[MAXC][8];
          gather(o, src, cx0, cy, cz, lanes, loc);
          vd val[MAXC][8], grad[MAXC][3][8];
          for (int c = 0; c < nc; ++c)
            interp(o, loc[c], val[c], grad[c]);
          const vd *cb = o->cache + (o->boff[col] + b) * 8 * nf;
          vd vr[MAXC][8], gr[MAXC][3][8];
          for (int q = 0; q < 8; ++q)
            {
              const vd *f    = cb + q * nf;
              const vd  c    = f[0];
              const vd *eta  = f + 1;
              const vd *lmu  = f + 1 + G;
              vd        v[MAXC];
              for (int k = 0; k < nc; ++k)
                v[k] = val[k][q];
              /* free-energy second derivatives, S/free_energy.h */
              vd S2 = bc(0.0), esum = bc(0.0), eij = bc(0.0), temp = bc(0.0);
              for (int i = 0; i < G; ++i)
                {
                  eij += eta[i] * esum;
                  S2 += eta[i] * eta[i];
                  esum += eta[i];
                  temp += eta[i] * v[2 + i];
                }
              /* MobilityScalar::apply_M and its derivative, S/mobility.h:441-534 */
              vd phi = c * c * c * (10.0 - 15.0 * c + 6.0 * c * c);
              for (int l = 0; l < CFB_VW; ++l)
                phi[l] = phi[l] < 0.0 ? 0.0 : (phi[l] > 1.0 ? 1.0 : phi[l]);
              const vd M = Mvol * phi + Mvap * (1.0 - phi) +
                           Msurf * 4.0 * c * c * (1.0 - c) * (1.0 - c) + (2.0 * Mgb) * eij;
              const vd dphidc = 30.0 * c * c * (1.0 - c) * (1.0 - c);
              const vd dMdc   = (Mvol - Mvap) * dphidc + Msurf * 8.0 * c * (1.0 - 3.0 * c + 2.0 * c * c);
              vd factor = bc(0.0);
              for (int i = 0; i < G; ++i)
                factor += (esum - eta[i]) * v[2 + i];
              const vd s = dMdc * v[0] + (2.0 * Mgb) * factor;
              /* rows, S/operator_sintering_generic.h:104-190 */
              vr[0][q] = jxw * w0 * v[0];
              for (int d = 0; d < 3; ++d)
                {
                  gr[0][d][q] = jxw * (M * grad[1][d][q] + s * lmu[d]);
                  gr[1][d][q] = (-kc * jxw) * grad[0][d][q];
                }
              const vd d2f_dc2 = 2.0 * A * c * c + 4.0 * A * c * (2.0 * c - 2.0) +
                                 2.0 * A * (1.0 - c) * (1.0 - c) + 2.0 * B;
              vd vmu = walg * v[1] - d2f_dc2 * v[0];
              for (int i = 0; i < G; ++i)
                {
                  const vd e       = eta[i];
                  const vd d2f_dce = (12.0 * B) * e * (e - 1.0);
                  const vd d2f_de2 = B * (12.0 - 12.0 * c + 2.0 * e * (12.0 * c - 24.0) +
                                          24.0 * e * e + 12.0 * S2);
                  vmu -= d2f_dce * v[2 + i];
                  vr[2 + i][q] = jxw * (w0 * v[2 + i] + L * (d2f_dce * v[0] + d2f_de2 * v[2 + i]) +
                                        (L * 24.0 * B) * e * (temp - e * v[2 + i]));
                  for (int d = 0; d < 3; ++d)
                    gr[2 + i][d][q] = (L * kp * jxw) * grad[2 + i][d][q];
                }
              vr[1][q] = jxw * vmu;
            }
          vd out[8];
          for (int k = 0; k < nc; ++k)
            {
              integrate(o, vr[k], gr[k], out);
              for (int l = 0; l < lanes; ++l)
                {
                  const int cx = cx0 + 2 * l;
                  for (int m = 0; m < 8; ++m)
                    {
                      const int64_t node =
                        ((int64_t)(cz + (m >> 2)) * o->nn[1] + (cy + ((m >> 1) & 1))) *
                          o->nn[0] + (cx + (m & 1));
                      dst[node * nc + k] += out[m][l];
                    }
                }
            }
        }
    }
}

int CFB_NAME(threads)(void)
{
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
