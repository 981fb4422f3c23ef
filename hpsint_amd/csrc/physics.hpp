// physics.hpp -- pointwise sintering physics on the device.
//
// Restates, per quadrature point, the arithmetic of
//   S/free_energy.h:37-176, 234-252       (FreeEnergy::Evaluation)
//   S/mobility.h:441-534                  (MobilityScalar)
//   S/mobility.h:758-916                  (MobilityTensorial)
//   I/base/tensor.h:137-174               (unit_vector, unit_vector_filter_2)
//   S/operator_sintering_generic.h:104-190 (tangent QP operator)
//   S/operator_sintering_generic.h:254-308 (residual QP operator)
// with S/ = applications/sintering/include/pf-applications/sintering/ and
// I/ = include/pf-applications/.  Operation order follows the reference
// expressions so that the thresholds (phi clamp, unit_vector tolerances,
// fsurf mask) see the same values.
#pragma once

#include <hip/hip_runtime.h>

namespace hps
{
  struct PhysParams
  {
    double A, B, kappa_c, kappa_p;
    double Mvol, Mvap, Msurf, Mgb, L;
    double w0, w_alg;
  };

  template <typename T>
  __device__ __forceinline__ T dmax(T a, T b)
  {
    return a > b ? a : b;
  }
  template <typename T>
  __device__ __forceinline__ T dmin(T a, T b)
  {
    return a < b ? a : b;
  }

  template <int DIM, typename T>
  __device__ __forceinline__ T vdot(const T *a, const T *b)
  {
    T s = a[0] * b[0];
#pragma unroll
    for (int d = 1; d < DIM; ++d)
      s += a[d] * b[d];
    return s;
  }

  // I/base/tensor.h:137-155
  template <int DIM, typename T>
  __device__ __forceinline__ void unit_vector(const T *v, T *out)
  {
    const T zero_tol      = T(1e-10);
    const T nrm           = sqrt(vdot<DIM>(v, v));
    const T filter        = nrm > zero_tol ? T(1) : T(0);
    const T filtered_norm = nrm < zero_tol ? T(1) : nrm;
#pragma unroll
    for (int d = 0; d < DIM; ++d)
      out[d] = (v[d] / filtered_norm) * filter;
  }

  // I/base/tensor.h:157-174
  template <int DIM, typename T>
  __device__ __forceinline__ T unit_vector_filter_2(const T *v)
  {
    const T zero_tol      = T(1e-20);
    const T nrm           = vdot<DIM>(v, v);
    const T filter        = nrm > zero_tol ? T(1) : T(0);
    const T filtered_norm = nrm < zero_tol ? T(1) : nrm;
    return filter / filtered_norm;
  }

  template <typename T>
  __device__ __forceinline__ T phi_clamped(T c)
  {
    T phi = c * c * c * (T(10.0) - T(15.0) * c + T(6.0) * c * c);
    phi   = dmax(T(0.0), phi);
    phi   = dmin(T(1.0), phi);
    return phi;
  }

  // number of cached linearisation fields per quadrature point:
  //   scalar mobility:    c, eta_0..eta_{G-1}, grad mu                (1+G+DIM)
  //   tensorial mobility: c, eta_i, grad c, grad mu and the grain-boundary
  //                       sums S, P, Q_i of gb_fields (1+G+2DIM+1+NP+G DIM)
  // (mu itself and grad c never enter the scalar tangent,
  //  S/mobility.h:477-534, S/operator_sintering_generic.h:104-190)
  template <int DIM, int G, bool TENS>
  struct CacheLayout
  {
    static constexpr int NP       = DIM * (DIM + 1) / 2; // symmetric DIM x DIM
    static constexpr int n_fields = TENS ? 1 + G + 2 * DIM + 1 + NP + G * DIM : 1 + G + DIM;
    static constexpr int c        = 0;
    static constexpr int eta      = 1;
    static constexpr int grad_c   = 1 + G;                    // TENS only
    static constexpr int grad_mu  = TENS ? 1 + G + DIM : 1 + G;
    static constexpr int gb_s     = 1 + G + 2 * DIM;          // TENS only
    static constexpr int gb_p     = gb_s + 1;                 // TENS only
    static constexpr int gb_q     = gb_p + NP;                // TENS only
  };

  // index of (a, b), a <= b, in the packed upper triangle of a DIM x DIM
  // symmetric matrix
  template <int DIM>
  __device__ __forceinline__ constexpr int sym_index(int a, int b)
  {
    return a * DIM - a * (a - 1) / 2 + (b - a);
  }

  // Grain-boundary part of MobilityTensorial::apply_M_derivative
  // (S/mobility.h:809-916): with, over the pairs i > j,
  //   e_ij = grad eta_i - grad eta_j,  f_ij = unit_vector_filter_2(e_ij),
  //   S = sum eta_i eta_j,
  //   P = sum eta_i eta_j f_ij e_ij e_ij^T,
  //   Q_i = sum_{j != i} eta_j f_ij e_ij (e_ij . grad mu),
  // the pair sum of the tangent's c row is
  //   S grad(dmu) - P grad(dmu) + (sum_i deta_i (sum_j eta_j - eta_i)) grad mu
  //   - sum_i deta_i Q_i,
  // so the linearisation cache holds S, P, Q_i and the tangent is O(G)
  // instead of O(G^2) per quadrature point.
  template <int DIM, int G, typename T>
  __device__ __forceinline__ void gb_fields(const T *etas, const T (*leg)[DIM], const T *lmu,
                                            T &S, T *P, T *Q)
  {
    constexpr int NP = DIM * (DIM + 1) / 2;
    S                = T(0);
#pragma unroll
    for (int k = 0; k < NP; ++k)
      P[k] = T(0);
#pragma unroll
    for (int k = 0; k < G * DIM; ++k)
      Q[k] = T(0);
#pragma unroll
    for (int i = 0; i < G; ++i)
#pragma unroll
      for (int j = 0; j < i; ++j)
        {
          T egd[DIM];
#pragma unroll
          for (int d = 0; d < DIM; ++d)
            egd[d] = leg[i][d] - leg[j][d];
          const T filter = unit_vector_filter_2<DIM>(egd);
          const T ee     = etas[i] * etas[j];
          S += ee;
#pragma unroll
          for (int a = 0; a < DIM; ++a)
#pragma unroll
            for (int b = a; b < DIM; ++b)
              P[sym_index<DIM>(a, b)] += ee * filter * egd[a] * egd[b];
          const T elmu = vdot<DIM>(egd, lmu);
#pragma unroll
          for (int d = 0; d < DIM; ++d)
            {
              const T r = (filter * egd[d]) * elmu;
              Q[i * DIM + d] += etas[j] * r;
              Q[j * DIM + d] += etas[i] * r;
            }
        }
  }

  // ------------------------------------------------------------------
  // tangent QP operator: SinteringOperatorGenericQuad::operator()
  // S/operator_sintering_generic.h:104-190 (advection disabled)
  // lin: cached fields; v[NC], g[NC][DIM]: increment; vr, gr: result
  // ------------------------------------------------------------------
  // VALS / GRADC select the value rows and the gradient row of c; the other
  // gradient rows are always written
  // Scales of the folded scalar tangent (FOLD = true): g holds gradients
  // without their per-lane signed 1/h, and the gradient rows come out as
  // the integration inputs (times the signed JxW/h).  The two signs cancel,
  // so a constant gradient row needs one product with a uniform scale:
  //   kj[d]  = JxW / h_d^2          (grad mu in the c row)
  //   kmu[d] = -kappa_c JxW / h_d^2 (mu row), keta[d] = L kappa_p JxW / h_d^2
  //   gsj[d] = signed JxW / h_d     (the cached grad mu of the c row)
  template <typename T>
  struct FoldScales
  {
    T kj[3], kmu[3], keta[3], gsj[3];
  };

  // c-row coefficients of the scalar-mobility tangent
  // (MobilityScalar::apply_M_derivative, S/mobility.h:477-534):
  //   grad row of c = M grad(dmu) + s grad(mu_lin),
  //   s = dM/dc dc + 2 Mgb sum_i (sum_j eta_j - eta_i) deta_i
  template <int G, typename T>
  __device__ __forceinline__ void scalar_mobility_terms(const PhysParams &pp, T c,
                                                        const T *etas, const T *v, T &M,
                                                        T &s)
  {
    const T Mvol  = T(pp.Mvol);
    const T Mvap  = T(pp.Mvap);
    const T Msurf = T(pp.Msurf);
    const T Mgb   = T(pp.Mgb);
    T       eta_sum = T(0.0), eta_ij_sum = T(0.0);
#pragma unroll
    for (int i = 0; i < G; ++i)
      {
        eta_sum += etas[i];
#pragma unroll
        for (int j = 0; j < i; ++j)
          eta_ij_sum += etas[i] * etas[j];
      }
    const T phi = phi_clamped(c);
    M           = Mvol * phi + Mvap * (T(1.0) - phi) +
        Msurf * T(4.0) * c * c * (T(1.0) - c) * (T(1.0) - c) + (T(2.0) * Mgb) * eta_ij_sum;
    const T dphidc = T(30.0) * c * c * (T(1.0) - c) * (T(1.0) - c);
    const T dMdc   = Mvol * dphidc - Mvap * dphidc +
                   Msurf * T(8.0) * c * (T(1.0) - T(3.0) * c + T(2.0) * c * c);
    T factor = T(0.0);
#pragma unroll
    for (int i = 0; i < G; ++i)
      factor += (eta_sum - etas[i]) * v[2 + i];
    s = dMdc * v[0] + (T(2.0) * Mgb) * factor;
  }

  // qsum (optional, tensorial only): sum_i deta_i Q_i formed by the caller,
  // which then need not hold the Q_i fields of lin
  template <int DIM, int G, bool TENS, bool VALS, bool GRADC, typename T>
  __device__ __forceinline__ void tangent_qp_rows(const PhysParams &pp,
                                                  const T          *lin,
                                                  const T          *v,
                                                  const T (*g)[DIM],
                                                  T                *vr,
                                                  T (*gr)[DIM],
                                                  const T          *qsum = nullptr)
  {
    using CL        = CacheLayout<DIM, G, TENS>;
    const T  c      = lin[CL::c];
    const T *etas   = lin + CL::eta;
    const T *lmu    = lin + CL::grad_mu;
    const T  A      = T(pp.A);
    const T  B      = T(pp.B);
    const T  L      = T(pp.L);
    const T  w0     = T(pp.w0);
    const T  Mvol   = T(pp.Mvol);
    const T  Mvap   = T(pp.Mvap);
    const T  Msurf  = T(pp.Msurf);
    const T  Mgb    = T(pp.Mgb);

    // FreeEnergy EnergySecond evaluation: only the power-2 sum
    T S2 = T(0.0);
#pragma unroll
    for (int i = 0; i < G; ++i)
      S2 += etas[i] * etas[i];

    // 1) c row
    vr[0] = v[0] * w0;
    if constexpr (!GRADC)
      {
      }
    else if constexpr (!TENS)
      {
        T M, s;
        scalar_mobility_terms<G>(pp, c, etas, v, M, s);
#pragma unroll
        for (int d = 0; d < DIM; ++d)
          gr[0][d] = M * g[1][d] + s * lmu[d];
      }
    else
      {
        // MobilityTensorial::apply_M_derivative, S/mobility.h:809-916
        const T *lc  = lin + CL::grad_c;
        const T  projector_scale = T(1.001);
        T        nc[DIM];
        unit_vector<DIM>(lc, nc);
        T out[DIM];
        {
          const T phi       = phi_clamped(c);
          const T f_vol_vap = Mvol * phi + Mvap * (T(1.0) - phi);
          const T fsurf = Msurf * T(4.0) * (c * c) * ((T(1.) - c) * (T(1.) - c));
          const T ncmu  = vdot<DIM>(nc, g[1]);
#pragma unroll
          for (int d = 0; d < DIM; ++d)
            out[d] = (f_vol_vap + fsurf) * g[1][d] - nc[d] * (fsurf * ncmu);
        }
        {
          const T c2_1minusc2 = T(4.0) * c * c * (T(1.) - c) * (T(1.) - c);
          const T dphidc      = T(30.0) * c2_1minusc2;
          const T f_vol_vap   = (Mvol - Mvap) * dphidc;
          const T fsurf       = Msurf * c2_1minusc2;
          T dfsurf = Msurf * T(8.0) * c * (T(1.) - c) * (T(1.) - T(2.) * c);
          if (fsurf < T(1e-6))
            dfsurf = T(0.0);
          const T nclmu = vdot<DIM>(nc, lmu);
#pragma unroll
          for (int d = 0; d < DIM; ++d)
            out[d] += ((f_vol_vap + dfsurf * projector_scale) * lmu[d] -
                       nc[d] * (dfsurf * nclmu)) *
                      v[0];
        }
        {
          T fsurf = Msurf * T(4.0) * (c * c) * ((T(1.) - c) * (T(1.) - c));
          T nrm   = sqrt(vdot<DIM>(lc, lc));
          if (fsurf < T(1e-6))
            fsurf = T(0.0);
          if (nrm < T(1e-4))
            nrm = T(1.0);
          const T ncg = vdot<DIM>(nc, g[0]);
          T       temp[DIM];
#pragma unroll
          for (int d = 0; d < DIM; ++d)
            temp[d] = g[0][d] - nc[d] * ncg;
          const T nclmu  = vdot<DIM>(nc, lmu);
          const T lmutmp = vdot<DIM>(lmu, temp);
          const T fac    = -fsurf / nrm;
#pragma unroll
          for (int d = 0; d < DIM; ++d)
            out[d] += fac * (nclmu * temp[d] + nc[d] * lmutmp);
        }
        if constexpr (G > 1)
          {
            // the pair sum through the cached S, P, Q_i (gb_fields)
            const T  S  = lin[CL::gb_s];
            const T *P  = lin + CL::gb_p;
            const T *Q  = lin + CL::gb_q;
            T        eta_sum = T(0.0);
#pragma unroll
            for (int i = 0; i < G; ++i)
              eta_sum += etas[i];
            T ls2 = T(0.0);
#pragma unroll
            for (int i = 0; i < G; ++i)
              ls2 += v[2 + i] * (eta_sum - etas[i]);
#pragma unroll
            for (int d = 0; d < DIM; ++d)
              {
                T o = S * g[1][d] + ls2 * lmu[d];
#pragma unroll
                for (int e = 0; e < DIM; ++e)
                  o -= P[d <= e ? sym_index<DIM>(d, e) : sym_index<DIM>(e, d)] * g[1][e];
                if (qsum)
                  o -= qsum[d];
                else
#pragma unroll
                  for (int i = 0; i < G; ++i)
                    o -= v[2 + i] * Q[i * DIM + d];
                out[d] = out[d] + o * (T(2.0) * Mgb);
              }
          }
#pragma unroll
        for (int d = 0; d < DIM; ++d)
          gr[0][d] = out[d];
      }

    // 2) mu row
    const T kc = -T(pp.kappa_c);
#pragma unroll
    for (int d = 0; d < DIM; ++d)
      gr[1][d] = kc * g[0][d];
    const T Lkp = L * T(pp.kappa_p);
#pragma unroll
    for (int i = 0; i < G; ++i)
#pragma unroll
      for (int d = 0; d < DIM; ++d)
        gr[i + 2][d] = Lkp * g[i + 2][d];
    if constexpr (!VALS)
      return;
    const T d2f_dc2 = T(2.0) * A * (c * c) + T(4.0) * A * c * (T(2.0) * c - T(2.0)) +
                      T(2.0) * A * ((-c + T(1.0)) * (-c + T(1.0))) + T(2.0) * B;
    T vmu = T(pp.w_alg) * v[1] - d2f_dc2 * v[0];
    T d2f_dcdeta[G > 0 ? G : 1];
#pragma unroll
    for (int i = 0; i < G; ++i)
      {
        d2f_dcdeta[i] = (B * T(12.0)) * etas[i] * (etas[i] - T(1.0));
        vmu -= d2f_dcdeta[i] * v[i + 2];
      }
    vr[1] = vmu;

    // 3) eta rows
#pragma unroll
    for (int i = 0; i < G; ++i)
      {
        const T e        = etas[i];
        const T d2f_de2  = B * (T(12.0) - T(12.0) * c + T(2.0) * e * (T(12.0) * c - T(24.0)) +
                               T(24.0) * (e * e) + T(12.0) * S2);
        vr[i + 2] = v[i + 2] * w0 + L * (d2f_dcdeta[i] * v[0] + d2f_de2 * v[i + 2]);
      }
    // apply_d2f_detaidetaj, S/free_energy.h:234-252
    if constexpr (G > 1)
      {
        T temp = etas[0] * v[2];
#pragma unroll
        for (int i = 1; i < G; ++i)
          temp += etas[i] * v[2 + i];
        const T f = L * T(24.0) * B;
#pragma unroll
        for (int i = 0; i < G; ++i)
          vr[2 + i] += f * etas[i] * (temp - etas[i] * v[2 + i]);
      }
  }

  template <int DIM, int G, bool TENS, typename T>
  __device__ __forceinline__ void tangent_qp(const PhysParams &pp, const T *lin, const T *v,
                                             const T (*g)[DIM], T *vr, T (*gr)[DIM])
  {
    tangent_qp_rows<DIM, G, TENS, true, true>(pp, lin, v, g, vr, gr);
  }

  // ------------------------------------------------------------------
  // The tangent split by output: the value rows need the increment's values
  // only; the gradient row of c needs the gradients of c and mu; the other
  // gradient rows are multiples of one input gradient each.  Kernels that
  // stream one component at a time use these pieces; tangent_qp above is
  // their composition.
  // ------------------------------------------------------------------
  template <int DIM, int G, bool TENS, typename T>
  __device__ __forceinline__ void tangent_values(const PhysParams &pp, const T *lin,
                                                 const T *v, T *vr)
  {
    T g[G + 2][DIM], gr[G + 2][DIM];
#pragma unroll
    for (int c = 0; c < G + 2; ++c)
#pragma unroll
      for (int d = 0; d < DIM; ++d)
        g[c][d] = T(0);
    tangent_qp_rows<DIM, G, TENS, true, false>(pp, lin, v, g, vr, gr);
  }

  template <int DIM, int G, bool TENS, typename T>
  __device__ __forceinline__ void tangent_grad_c(const PhysParams &pp, const T *lin,
                                                 const T *v, const T *g0, const T *g1,
                                                 T *out, const T *qsum = nullptr)
  {
    T g[G + 2][DIM], vr[G + 2], gr[G + 2][DIM];
#pragma unroll
    for (int d = 0; d < DIM; ++d)
      {
        g[0][d] = g0[d];
        g[1][d] = g1[d];
      }
#pragma unroll
    for (int c = 2; c < G + 2; ++c)
#pragma unroll
      for (int d = 0; d < DIM; ++d)
        g[c][d] = T(0);
    tangent_qp_rows<DIM, G, TENS, false, true>(pp, lin, v, g, vr, gr, qsum);
#pragma unroll
    for (int d = 0; d < DIM; ++d)
      out[d] = gr[0][d];
  }

  // ------------------------------------------------------------------
  // residual QP operator: SinteringOperatorGenericResidualQuad::operator()
  // S/operator_sintering_generic.h:254-308; hist[] = sum_k w_k old_k at q
  // ------------------------------------------------------------------
  template <int DIM, int G, bool TENS, int WTD, typename T>
  __device__ __forceinline__ void residual_qp(const PhysParams &pp,
                                              const T          *v,
                                              const T (*g)[DIM],
                                              const T          *hist,
                                              T                *vr,
                                              T (*gr)[DIM])
  {
    const T  c     = v[0];
    const T *etas  = v + 2;
    const T  A     = T(pp.A);
    const T  B     = T(pp.B);
    const T  L     = T(pp.L);
    const T  Mvol  = T(pp.Mvol);
    const T  Mvap  = T(pp.Mvap);
    const T  Msurf = T(pp.Msurf);
    const T  Mgb   = T(pp.Mgb);

    T S2 = T(0.0), S3 = T(0.0);
#pragma unroll
    for (int i = 0; i < G; ++i)
      {
        S2 += etas[i] * etas[i];
        S3 += etas[i] * (etas[i] * etas[i]);
      }

    // 1) c row
    T v0 = T(0.0);
    if constexpr (WTD >= 1)
      v0 = c * T(pp.w0);
    if constexpr (WTD == 2)
      v0 += hist[0];
    vr[0] = v0;
    if constexpr (!TENS)
      {
        // MobilityScalar::apply_M, S/mobility.h:441-472
        T etaijSum = T(0.0);
#pragma unroll
        for (int i = 0; i < G; ++i)
#pragma unroll
          for (int j = 0; j < i; ++j)
            etaijSum += etas[i] * etas[j];
        const T phi = phi_clamped(c);
        const T M   = Mvol * phi + Mvap * (T(1.0) - phi) +
                    Msurf * T(4.0) * c * c * (T(1.0) - c) * (T(1.0) - c) +
                    (T(2.0) * Mgb) * etaijSum;
#pragma unroll
        for (int d = 0; d < DIM; ++d)
          gr[0][d] = M * g[1][d];
      }
    else
      {
        // MobilityTensorial::apply_M, S/mobility.h:758-804
        const T projector_scale = T(1.001);
        const T phi             = phi_clamped(c);
        const T f_vol_vap       = Mvol * phi + Mvap * (T(1.0) - phi);
        const T fsurf = Msurf * T(4.0) * (c * c) * ((T(1.) - c) * (T(1.) - c));
        T       nc[DIM];
        unit_vector<DIM>(g[0], nc);
        const T ncmu = vdot<DIM>(nc, g[1]);
        T       out[DIM];
#pragma unroll
        for (int d = 0; d < DIM; ++d)
          out[d] = (f_vol_vap + fsurf * projector_scale) * g[1][d] - nc[d] * (fsurf * ncmu);
        if constexpr (G > 1)
          {
            T out_gb[DIM];
#pragma unroll
            for (int d = 0; d < DIM; ++d)
              out_gb[d] = T(0.0);
#pragma unroll
            for (int i = 0; i < G; ++i)
#pragma unroll
              for (int j = 0; j < i; ++j)
                {
                  T egd[DIM];
#pragma unroll
                  for (int d = 0; d < DIM; ++d)
                    egd[d] = g[2 + i][d] - g[2 + j][d];
                  const T filter = unit_vector_filter_2<DIM>(egd);
                  const T emu    = vdot<DIM>(egd, g[1]);
                  const T ee     = etas[i] * etas[j];
#pragma unroll
                  for (int d = 0; d < DIM; ++d)
                    out_gb[d] += (g[1][d] - egd[d] * (filter * emu)) * ee;
                }
#pragma unroll
            for (int d = 0; d < DIM; ++d)
              out[d] = out[d] + out_gb[d] * (T(2.0) * Mgb);
          }
#pragma unroll
        for (int d = 0; d < DIM; ++d)
          gr[0][d] = out[d];
      }

    // 2) mu row
    const T df_dc = A * (c * c) * (T(2.0) * c - T(2.0)) +
                    T(2.0) * A * c * ((-c + T(1.0)) * (-c + T(1.0))) +
                    B * (T(2.0) * c - T(6.0) * S2 + T(4.0) * S3);
    vr[1] = T(pp.w_alg) * v[1] - df_dc;
    const T kc = -T(pp.kappa_c);
#pragma unroll
    for (int d = 0; d < DIM; ++d)
      gr[1][d] = kc * g[0][d];

    // 3) eta rows
    const T Lkp = L * T(pp.kappa_p);
#pragma unroll
    for (int i = 0; i < G; ++i)
      {
        const T e  = etas[i];
        T       r  = L * (e * (B * T(12.0)) * (e * (T(1.0) * c - T(2.0)) + (-c + T(1.0)) + S2));
        if constexpr (WTD >= 1)
          r += v[2 + i] * T(pp.w0);
        if constexpr (WTD == 2)
          r += hist[2 + i];
        vr[2 + i] = r;
#pragma unroll
        for (int d = 0; d < DIM; ++d)
          gr[2 + i][d] = Lkp * g[2 + i][d];
      }
  }
  // residual rows that need the state's values only (all rows' values)
  template <int DIM, int G, bool TENS, int WTD, typename T>
  __device__ __forceinline__ void residual_values(const PhysParams &pp, const T *v,
                                                  const T *hist, T *vr)
  {
    T g[G + 2][DIM], gr[G + 2][DIM];
#pragma unroll
    for (int c = 0; c < G + 2; ++c)
#pragma unroll
      for (int d = 0; d < DIM; ++d)
        g[c][d] = T(0);
    residual_qp<DIM, G, false, WTD>(pp, v, g, hist, vr, gr);
  }

  // gradient row of c of the residual for scalar mobility: M(c, eta) grad mu
  // (MobilityScalar::apply_M, S/mobility.h:441-472)
  template <int DIM, int G, typename T>
  __device__ __forceinline__ void residual_grad_c_scalar(const PhysParams &pp, const T *v,
                                                         const T *g1, T *out)
  {
    const T  c    = v[0];
    const T *etas = v + 2;
    T        etaijSum = T(0.0);
#pragma unroll
    for (int i = 0; i < G; ++i)
#pragma unroll
      for (int j = 0; j < i; ++j)
        etaijSum += etas[i] * etas[j];
    const T phi = phi_clamped(c);
    const T M   = T(pp.Mvol) * phi + T(pp.Mvap) * (T(1.0) - phi) +
                T(pp.Msurf) * T(4.0) * c * c * (T(1.0) - c) * (T(1.0) - c) +
                (T(2.0) * T(pp.Mgb)) * etaijSum;
#pragma unroll
    for (int d = 0; d < DIM; ++d)
      out[d] = M * g1[d];
  }
} // namespace hps
