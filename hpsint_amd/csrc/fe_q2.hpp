// fe_q2.hpp -- Q2 cell kernels (3-point Gauss rule per direction) of the
// sintering operator for gfx950.
//
// Same operations as the Q1 kernels (tangent vmult, residual, linearisation
// cache, diagonal / node-block probing; S/operator_sintering_generic.h:
// 539-577, 877-1006, S/operator_sintering_data.h:301-409), for FE_Q(2) with
// QGauss(3) on a uniform Cartesian mesh.
//
// Lane mapping.  A workgroup is one wave; its lanes are split into cells of
// NP = 3^DIM lanes (2 cells per wave in 3D, 7 in 2D).  Lane p of a cell,
// p = t0 + 3 t1 + 9 t2, stands for node (t0, t1, t2) of the cell on gather
// and scatter and for Gauss point (t0, t1, t2) in between.  Sum
// factorisation runs through LDS, one pass per direction: a lane reads the
// three entries of its line along d and combines them with one row (or, to
// integrate, one column) of a 3x3 matrix:
//   S[q][n]  = L_n(x_q), Q2 Lagrange basis (nodes 0, 1/2, 1) at the Gauss
//              points x_q = 1/2 - sqrt(15)/10, 1/2, 1/2 + sqrt(15)/10;
//   Dq[q][r] = l_r'(x_q), derivative of the Lagrange basis through the
//              Gauss points (collocation derivative: gradients at the
//              Gauss points from the values at the Gauss points, exact for
//              the Q2 field because Dq S = L').
// Interpolation: values = S_x S_y S_z u, gradient_d = Dq_d (values) / h_d.
// Integration:   y = JxW v + sum_d Dq_d^T (JxW g_d / h_d),  r = S^T y.
//
// Assembly.  Cells are processed in 2^DIM colour passes (parity of the cell
// index along every direction); cells of one colour share no node, so a
// pass updates dst without atomics.  (The 3D tangent vmult instead marches
// tiles of cells, q2_vmult_tile below, with the same cache and the same
// per-wave cell body.)  The first colour that reaches a node
// stores, later colours add, always in colour order: every entry is summed
// in a fixed order (deterministic).  A cell is the first to reach its local
// node (t0, t1, t2) iff along every direction its index is even, the node is
// the cell's middle node, or the neighbour cell sharing the node does not
// exist.
//
// The linearisation cache is stored per workgroup and field in lane order,
// cache[(colour base + workgroup) * NF + field][64], one coalesced 512-byte
// read per field.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "physics.hpp"

namespace hps
{
  struct Q2Params
  {
    int        nc[3];   // cells along x, y, z (1 along z in 2D)
    int        nn[3];   // nodes along x, y, z: 2 nc + 1 (1 along z in 2D)
    int        par[3];  // colour of this launch: parity of the cell index
    int        ncol[3]; // cells of this colour along x, y, z
    int64_t    n_col;   // cells of this colour
    int64_t    wg_base; // first workgroup of this colour in the cache
    int        order;   // number of history vectors (BDF order)
    double     inv_h[3];
    double     h_prod;    // cell volume
    double     w_hist[3]; // BDF weights[1..order]
    PhysParams pp;
    int        tile_zc; // 3D vmult: cell layers per tile march (0: colour passes)
  };

  // one colour of the 3D tile march (q2_vmult_tile): tiles of 4 x NY x zc
  // cells, colour = parity of the tile index along x, y, z
  struct Q2TileParams
  {
    int     tpar[3]; // parity of this colour's tiles
    int     ntc[3];  // tiles of this colour along x, y, z
    int64_t wgb[8];  // first cache workgroup of each cell colour
  };

  template <typename T>
  struct Q2Coef
  {
    // S rows: (s_a, s_b, s_c), (0, 1, 0), (s_c, s_b, s_a)
    static constexpr T s_a = T(0.68729833462074168851792653997823996108);
    static constexpr T s_b = T(0.4);
    static constexpr T s_c = T(-0.087298334620741688517926539978239961);
    // Dq rows: (d_a, d_b, d_c), (d_c, 0, -d_c), (-d_c, -d_b, -d_a)
    static constexpr T d_a = T(-3.8729833462074168851792653997823996108);
    static constexpr T d_b = T(5.1639777949432225135723538663765328144);
    static constexpr T d_c = T(-1.2909944487358056283930884665941332036);
    // L' rows (derivative of the Q2 basis at the Gauss points):
    // (e_a, e_b, e_c), (-1, 0, 1), (-e_c, -e_b, -e_a)
    static constexpr T e_a = T(-2.5491933384829667540717061599129598443);
    static constexpr T e_b = T(3.0983866769659335081434123198259196887);
    static constexpr T e_c = T(-0.5491933384829667540717061599129598443);
    // Gauss weights on [0, 1]
    static constexpr T w_a = T(0.27777777777777777777777777777777777778);
    static constexpr T w_b = T(0.44444444444444444444444444444444444444);
  };

  template <typename T>
  __device__ __forceinline__ T sel3(int t, T a, T b, T c)
  {
    return t == 0 ? a : (t == 1 ? b : c);
  }

  // row t of S: out[n] = S[t][n]
  template <typename T>
  __device__ __forceinline__ void q2_s_row(int t, T *r)
  {
    using C = Q2Coef<T>;
    r[0]    = sel3(t, C::s_a, T(0), C::s_c);
    r[1]    = t == 1 ? T(1) : C::s_b;
    r[2]    = sel3(t, C::s_c, T(0), C::s_a);
  }
  // column t of S: out[q] = S[q][t]
  template <typename T>
  __device__ __forceinline__ void q2_s_col(int t, T *r)
  {
    using C = Q2Coef<T>;
    r[0]    = sel3(t, C::s_a, C::s_b, C::s_c);
    r[1]    = t == 1 ? T(1) : T(0);
    r[2]    = sel3(t, C::s_c, C::s_b, C::s_a);
  }
  // row t of Dq
  template <typename T>
  __device__ __forceinline__ void q2_dq_row(int t, T *r)
  {
    using C = Q2Coef<T>;
    r[0]    = sel3(t, C::d_a, C::d_c, -C::d_c);
    r[1]    = sel3(t, C::d_b, T(0), -C::d_b);
    r[2]    = sel3(t, C::d_c, -C::d_c, -C::d_a);
  }
  // column t of Dq
  template <typename T>
  __device__ __forceinline__ void q2_dq_col(int t, T *r)
  {
    using C = Q2Coef<T>;
    r[0]    = sel3(t, C::d_a, C::d_b, C::d_c);
    r[1]    = sel3(t, C::d_c, T(0), -C::d_c);
    r[2]    = sel3(t, -C::d_c, -C::d_b, -C::d_a);
  }
  // row t of L'
  template <typename T>
  __device__ __forceinline__ void q2_e_row(int t, T *r)
  {
    using C = Q2Coef<T>;
    r[0]    = sel3(t, C::e_a, T(-1), -C::e_c);
    r[1]    = sel3(t, C::e_b, T(0), -C::e_b);
    r[2]    = sel3(t, C::e_c, T(1), -C::e_a);
  }
  // column t of L': out[q] = L'[q][t] (derivative of basis t at point q)
  template <typename T>
  __device__ __forceinline__ void q2_e_col(int t, T *r)
  {
    using C = Q2Coef<T>;
    r[0]    = sel3(t, C::e_a, C::e_b, C::e_c);
    r[1]    = sel3(t, T(-1), T(0), T(1));
    r[2]    = sel3(t, -C::e_c, -C::e_b, -C::e_a);
  }
  template <typename T>
  __device__ __forceinline__ T q2_w(int t)
  {
    return t == 1 ? Q2Coef<T>::w_b : Q2Coef<T>::w_a;
  }

  // out = sum_a coef[a] * buf[base + a * str]
  template <typename T>
  __device__ __forceinline__ T q2_pass(const T *buf, int base, int str, const T *coef)
  {
    return coef[0] * buf[base] + coef[1] * buf[base + str] + coef[2] * buf[base + 2 * str];
  }

  template <int DIM>
  struct Q2Lane
  {
    static constexpr int NP  = DIM == 3 ? 27 : 9;
    static constexpr int CPW = 64 / NP; // cells per wave
    static constexpr int LP  = CPW * NP; // lanes that hold a cell's node / point
    static constexpr int str(int d)
    {
      return d == 0 ? 1 : (d == 1 ? 3 : 9);
    }

    int  l, s, p, t[3];
    bool valid; // lane belongs to an existing cell of this colour
    int  cc[3]; // its cell
    int  b0[3]; // first lane of this lane's line along x, y, z

    __device__ Q2Lane(const Q2Params &kp)
    {
      l     = threadIdx.x;
      s     = l / NP;
      p     = l - s * NP;
      t[0]  = p % 3;
      t[1]  = (p / 3) % 3;
      t[2]  = DIM == 3 ? p / 9 : 0;
      valid = s < CPW && cell(kp, (int64_t)blockIdx.x * CPW + s, cc);
      // lanes past the last cell slot read the first cell's lines, so every
      // LDS read stays inside the 64 entries of a row
#pragma unroll
      for (int d = 0; d < 3; ++d)
        b0[d] = s < CPW ? l - t[d] * str(d) : 0;
    }
    // lane `lane` of a wave whose slot (lane / NP) holds the cell c, which
    // exists in the mesh if `exists`
    __device__ Q2Lane(int lane, const int *c, bool exists)
    {
      l     = lane;
      s     = l / NP;
      p     = l - s * NP;
      t[0]  = p % 3;
      t[1]  = (p / 3) % 3;
      t[2]  = DIM == 3 ? p / 9 : 0;
      valid = s < CPW && exists;
#pragma unroll
      for (int d = 0; d < 3; ++d)
        {
          cc[d] = c[d];
          b0[d] = s < CPW ? l - t[d] * str(d) : 0;
        }
    }
    __device__ int base(int d) const
    {
      return b0[d];
    }

    // cell coordinates of the ci-th cell of the launch's colour
    __device__ static bool cell(const Q2Params &kp, int64_t ci, int *c)
    {
      if (ci >= kp.n_col)
        return false;
      const int64_t q  = ci / kp.ncol[0];
      const int     ix = (int)(ci - q * kp.ncol[0]);
      const int     iy = (int)(q % kp.ncol[1]);
      const int     iz = (int)(q / kp.ncol[1]);
      c[0]             = 2 * ix + kp.par[0];
      c[1]             = 2 * iy + kp.par[1];
      c[2]             = 2 * iz + kp.par[2];
      return true;
    }
    __device__ static int64_t node(const Q2Params &kp, const int *c, int p)
    {
      const int gx = 2 * c[0] + p % 3;
      const int gy = 2 * c[1] + (p / 3) % 3;
      const int gz = DIM == 3 ? 2 * c[2] + p / 9 : 0;
      return gx + (int64_t)kp.nn[0] * (gy + (int64_t)kp.nn[1] * gz);
    }
    // this cell is the first of the colour sequence to reach local node p
    __device__ static bool first(const Q2Params &kp, const int *c, int p)
    {
      const int tt[3] = {p % 3, (p / 3) % 3, DIM == 3 ? p / 9 : 1};
      bool      f     = true;
#pragma unroll
      for (int d = 0; d < DIM; ++d)
        f = f && ((c[d] & 1) == 0 || tt[d] == 1 || (tt[d] == 0 && c[d] == 0) ||
                  (tt[d] == 2 && c[d] == kp.nc[d] - 1));
      return f;
    }
  };

  // values and gradients at this lane's Gauss point of the field whose
  // nodal values (lane order) are in u; w0..w2: scratch rows of 64
  template <int DIM, typename T>
  __device__ __forceinline__ void q2_interp_vg(const Q2Lane<DIM> &L, const T *u, T *w0, T *w1,
                                               T *w2, const T (*srow)[3], const T (*drow)[3],
                                               const T *inv_h, T &val, T *grad)
  {
    const int l = L.l;
    if constexpr (DIM == 3)
      {
        w0[l] = q2_pass(u, L.base(0), 1, srow[0]);
        __syncthreads();
        w1[l] = q2_pass(w0, L.base(1), 3, srow[1]);
        __syncthreads();
        val   = q2_pass(w1, L.base(2), 9, srow[2]);
        w2[l] = val;
        __syncthreads();
      }
    else
      {
        w0[l] = q2_pass(u, L.base(0), 1, srow[0]);
        __syncthreads();
        val   = q2_pass(w0, L.base(1), 3, srow[1]);
        w2[l] = val;
        __syncthreads();
      }
#pragma unroll
    for (int d = 0; d < DIM; ++d)
      grad[d] = q2_pass(w2, L.base(d), Q2Lane<DIM>::str(d), drow[d]) * inv_h[d];
  }

  // values only
  template <int DIM, typename T>
  __device__ __forceinline__ T q2_interp_v(const Q2Lane<DIM> &L, const T *u, T *w0, T *w1,
                                           const T (*srow)[3])
  {
    const int l = L.l;
    w0[l]       = q2_pass(u, L.base(0), 1, srow[0]);
    __syncthreads();
    if constexpr (DIM == 3)
      {
        w1[l] = q2_pass(w0, L.base(1), 3, srow[1]);
        __syncthreads();
        const T v = q2_pass(w1, L.base(2), 9, srow[2]);
        __syncthreads();
        return v;
      }
    else
      {
        const T v = q2_pass(w0, L.base(1), 3, srow[1]);
        __syncthreads();
        return v;
      }
  }

  // nodal contribution of this lane's node from the Gauss-point value
  // coefficient vq and gradient coefficients gq[d] (JxW and 1/h applied)
  template <int DIM, typename T>
  __device__ __forceinline__ T q2_integrate(const Q2Lane<DIM> &L, T vq, const T *gq, T *w0,
                                            T *w1, T *w2, T *w3, const T (*scol)[3],
                                            const T (*dcol)[3])
  {
    const int l = L.l;
    w0[l]       = gq[0];
    w1[l]       = gq[1];
    if constexpr (DIM == 3)
      w2[l] = gq[2];
    __syncthreads();
    T y = vq;
    y += q2_pass(w0, L.base(0), 1, dcol[0]);
    y += q2_pass(w1, L.base(1), 3, dcol[1]);
    if constexpr (DIM == 3)
      y += q2_pass(w2, L.base(2), 9, dcol[2]);
    w3[l] = y;
    __syncthreads();
    T r;
    if constexpr (DIM == 3)
      {
        w0[l] = q2_pass(w3, L.base(2), 9, scol[2]);
        __syncthreads();
        w1[l] = q2_pass(w0, L.base(1), 3, scol[1]);
        __syncthreads();
        r = q2_pass(w1, L.base(0), 1, scol[0]);
      }
    else
      {
        w0[l] = q2_pass(w3, L.base(1), 3, scol[1]);
        __syncthreads();
        r = q2_pass(w0, L.base(0), 1, scol[0]);
      }
    __syncthreads();
    return r;
  }

  // 16-byte vector of T
  template <typename T>
  struct Vec16;
  template <>
  struct Vec16<double>
  {
    using type = double2;
  };
  template <>
  struct Vec16<float>
  {
    using type = float4;
  };

  // the NC consecutive values at p (16-byte loads where the row allows them)
  template <int NC, typename T>
  __device__ __forceinline__ void q2_load_row(const T *p, T *x)
  {
    constexpr int W = 16 / (int)sizeof(T);
    if constexpr (NC % W == 0)
      if ((reinterpret_cast<uintptr_t>(p) & 15) == 0)
        {
#pragma unroll
          for (int k = 0; k < NC / W; ++k)
            {
              const typename Vec16<T>::type v =
                reinterpret_cast<const typename Vec16<T>::type *>(p)[k];
              __builtin_memcpy(x + k * W, &v, 16);
            }
          return;
        }
#pragma unroll
    for (int c = 0; c < NC; ++c)
      x[c] = p[c];
  }

  template <int NC, typename T>
  __device__ __forceinline__ void q2_store_row(T *p, const T *x)
  {
    constexpr int W = 16 / (int)sizeof(T);
    if constexpr (NC % W == 0)
      if ((reinterpret_cast<uintptr_t>(p) & 15) == 0)
        {
#pragma unroll
          for (int k = 0; k < NC / W; ++k)
            {
              typename Vec16<T>::type v;
              __builtin_memcpy(&v, x + k * W, 16);
              reinterpret_cast<typename Vec16<T>::type *>(p)[k] = v;
            }
          return;
        }
#pragma unroll
    for (int c = 0; c < NC; ++c)
      p[c] = x[c];
  }

  // gather: each lane of a cell slot reads the NC components of its own
  // node (row(node, x) fills x[NC]) into buf[c * 64 + lane]; the other
  // lanes write zeros.  Two cells of one colour share no node.
  template <int DIM, int NC, typename T, typename F>
  __device__ __forceinline__ void q2_gather(const Q2Params &kp, const Q2Lane<DIM> &L, T *buf,
                                            F &&row)
  {
    T x[NC];
    if (L.valid)
      row(Q2Lane<DIM>::node(kp, L.cc, L.p), x);
    else
#pragma unroll
      for (int c = 0; c < NC; ++c)
        x[c] = T(0);
#pragma unroll
    for (int c = 0; c < NC; ++c)
      buf[c * 64 + L.l] = x[c];
  }

  // scatter: each lane of an existing cell stores (first colour to reach the
  // node) or adds buf[c * 64 + lane] to dst[node * os + c * ocs + oo]
  template <int DIM, int NC, typename T>
  __device__ __forceinline__ void q2_scatter(const Q2Params &kp, const Q2Lane<DIM> &L,
                                             const T *buf, T *dst, int os, int ocs, int oo)
  {
    using LN = Q2Lane<DIM>;
    if (!L.valid)
      return;
    const int64_t n     = LN::node(kp, L.cc, L.p);
    const bool    first = LN::first(kp, L.cc, L.p);
    T             r[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c)
      r[c] = buf[c * 64 + L.l];
    if (ocs == 1)
      {
        T *d = dst + n * os + oo;
        if (!first)
          {
            T o[NC];
            q2_load_row<NC>(d, o);
#pragma unroll
            for (int c = 0; c < NC; ++c)
              r[c] = o[c] + r[c];
          }
        q2_store_row<NC>(d, r);
      }
    else
#pragma unroll
      for (int c = 0; c < NC; ++c)
        {
          T *d = dst + n * os + (int64_t)c * ocs + oo;
          *d   = first ? r[c] : *d + r[c];
        }
  }

  // barrier of the vmult body: the whole workgroup, or (WAVE) only the
  // wave, when every wave of the workgroup works on LDS rows of its own
  template <bool WAVE>
  __device__ __forceinline__ void q2_sync()
  {
    if constexpr (WAVE)
      {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      }
    else
      __syncthreads();
  }

  // tangent vmult of the workgroup's cells.  U holds the gathered increment
  // (NC rows of 64) on entry and the integrated rows on exit; A (NC rows) is
  // the second work buffer; VR (NC rows) takes the value rows.  Every
  // sum-factorisation pass runs over all components between two barriers,
  // so a workgroup waits for LDS once per pass instead of once per
  // component and pass.  Only the gradients of c and mu enter the cached
  // tangent; the gradient of eta_i enters its own row only (L kappa_p
  // grad eta_i) and is formed from the Gauss-point values before the
  // integration.
  // cl: this lane's entry of the first cached field (fields LP entries apart)
  // EL: request all cache fields at the top (the tile march, FP64)
  template <int DIM, int G, bool TENS, typename T, bool WAVE = false, bool EL = false>
  __device__ __forceinline__ void q2_vmult_body(const Q2Params &kp, const Q2Lane<DIM> &L, T *U,
                                                T *VR, T *A, const T *__restrict__ cl,
                                                const T *inv_h)
  {
    using CL         = CacheLayout<DIM, G, TENS>;
    using LN         = Q2Lane<DIM>;
    constexpr int NC = G + 2;
    constexpr int NF = CL::n_fields;
    const int     l  = L.l;

    // the tensorial Q_i fields are requested before the interpolation
    // passes, so that they arrive while the passes run
    // EL: the other cache fields as well, or all of them with scalar
    // mobility (their latency is otherwise waited for after the
    // interpolation passes: the FP64 tile march at 96^3 with 10 order
    // parameters 4 % faster tensorial, 3 % scalar; FP32 1.4 % slower, and
    // the colour-pass kernels lose occupancy)
    constexpr bool EARLYQ = TENS && G > 1;
    constexpr bool EARLYL = EL;
    constexpr int  NL     = EARLYQ ? CL::gb_q : NF; // the other fields
    T              qf[EARLYQ ? G * DIM : 1];
    T              lin_e[EARLYL ? NL : 1];
    if constexpr (EARLYQ)
      {
#pragma unroll
        for (int f = 0; f < G * DIM; ++f)
          qf[f] = L.valid ? cl[(CacheLayout<DIM, G, TENS>::gb_q + f) * Q2Lane<DIM>::LP] : T(0);
      }
    if constexpr (EARLYL)
      {
#pragma unroll
        for (int f = 0; f < NL; ++f)
          lin_e[f] = L.valid ? cl[f * Q2Lane<DIM>::LP] : T(0);
      }
    // 1) values at the Gauss points, one pass per direction for all
    //    components (U -> A -> U -> registers in 3D, U -> A -> registers in 2D)
    T v[NC];
    {
      T srow[DIM][3];
#pragma unroll
      for (int d = 0; d < DIM; ++d)
        q2_s_row(L.t[d], srow[d]);
#pragma unroll
      for (int c = 0; c < NC; ++c)
        A[c * 64 + l] = q2_pass(U + c * 64, L.base(0), 1, srow[0]);
      q2_sync<WAVE>();
      if constexpr (DIM == 3)
        {
#pragma unroll
          for (int c = 0; c < NC; ++c)
            U[c * 64 + l] = q2_pass(A + c * 64, L.base(1), 3, srow[1]);
          q2_sync<WAVE>();
#pragma unroll
          for (int c = 0; c < NC; ++c)
            v[c] = q2_pass(U + c * 64, L.base(2), 9, srow[2]);
        }
      else
#pragma unroll
        for (int c = 0; c < NC; ++c)
          v[c] = q2_pass(A + c * 64, L.base(1), 3, srow[1]);
    }
    // QV: the buffer the last pass did not read takes the Gauss-point
    // values; WS is the other one
    T *QV = DIM == 3 ? A : U;
    T *WS = DIM == 3 ? U : A;
#pragma unroll
    for (int c = 0; c < NC; ++c)
      QV[c * 64 + l] = v[c];
    q2_sync<WAVE>();

    T drow[DIM][3];
#pragma unroll
    for (int d = 0; d < DIM; ++d)
      q2_dq_row(L.t[d], drow[d]);
    T g0[DIM], g1[DIM];
#pragma unroll
    for (int d = 0; d < DIM; ++d)
      {
        g0[d] = q2_pass(QV, L.base(d), LN::str(d), drow[d]) * inv_h[d];
        g1[d] = q2_pass(QV + 64, L.base(d), LN::str(d), drow[d]) * inv_h[d];
      }

    // 2) cache and tangent: value rows (times JxW) to VR, the c gradient row
    T jxw = T(kp.h_prod);
#pragma unroll
    for (int d = 0; d < DIM; ++d)
      jxw *= q2_w<T>(L.t[d]);
    T             gq[NC][DIM];
    constexpr int LP = LN::LP; // cache: [workgroup][field][LP lanes]
    {
      T vr[NC], gc0[DIM];
      if constexpr (TENS && G > 1)
        {
          // sum_i deta_i Q_i from the Q_i fields requested at the top
          T qsum[DIM];
#pragma unroll
          for (int d = 0; d < DIM; ++d)
            qsum[d] = T(0);
#pragma unroll
          for (int i = 0; i < G; ++i)
#pragma unroll
            for (int d = 0; d < DIM; ++d)
              qsum[d] += v[2 + i] * qf[i * DIM + d];
          T lin[CL::gb_q];
#pragma unroll
          for (int f = 0; f < CL::gb_q; ++f)
            if constexpr (EARLYL)
              lin[f] = lin_e[f];
            else
              lin[f] = L.valid ? cl[f * LP] : T(0);
          tangent_values<DIM, G, TENS>(kp.pp, lin, v, vr);
          tangent_grad_c<DIM, G, TENS>(kp.pp, lin, v, g0, g1, gc0, qsum);
        }
      else
        {
          T lin[NF];
#pragma unroll
          for (int f = 0; f < NF; ++f)
            if constexpr (EARLYL)
              lin[f] = lin_e[f];
            else
              lin[f] = L.valid ? cl[f * LP] : T(0);
          tangent_values<DIM, G, TENS>(kp.pp, lin, v, vr);
          tangent_grad_c<DIM, G, TENS>(kp.pp, lin, v, g0, g1, gc0);
        }
#pragma unroll
      for (int c = 0; c < NC; ++c)
        VR[c * 64 + l] = vr[c] * jxw;
      // 3) gradient rows with JxW / h_d: c, mu = -kappa_c grad c,
      //    eta_i = L kappa_p grad eta_i (from the Gauss-point values in QV)
      const T kc  = -T(kp.pp.kappa_c);
      const T lkp = T(kp.pp.L) * T(kp.pp.kappa_p);
#pragma unroll
      for (int d = 0; d < DIM; ++d)
        {
          const T sd = jxw * inv_h[d];
          gq[0][d]   = gc0[d] * sd;
          gq[1][d]   = kc * g0[d] * sd;
#pragma unroll
          for (int c = 2; c < NC; ++c)
            gq[c][d] =
              q2_pass(QV + c * 64, L.base(d), LN::str(d), drow[d]) * (lkp * inv_h[d] * sd);
        }
    }
    q2_sync<WAVE>();

    // 4) integration: collocation-derivative transposes of the gradient rows
    //    added to the value rows, then the transposed interpolation passes
    T scol[DIM][3], dcol[DIM][3];
#pragma unroll
    for (int d = 0; d < DIM; ++d)
      {
        q2_s_col(L.t[d], scol[d]);
        q2_dq_col(L.t[d], dcol[d]);
      }
    T y[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c)
      {
        WS[c * 64 + l] = gq[c][0];
        QV[c * 64 + l] = gq[c][1];
      }
    q2_sync<WAVE>();
#pragma unroll
    for (int c = 0; c < NC; ++c)
      y[c] = VR[c * 64 + l] + q2_pass(WS + c * 64, L.base(0), 1, dcol[0]) +
             q2_pass(QV + c * 64, L.base(1), 3, dcol[1]);
    q2_sync<WAVE>();
    if constexpr (DIM == 3)
      {
#pragma unroll
        for (int c = 0; c < NC; ++c)
          WS[c * 64 + l] = gq[c][2];
        q2_sync<WAVE>();
#pragma unroll
        for (int c = 0; c < NC; ++c)
          y[c] += q2_pass(WS + c * 64, L.base(2), 9, dcol[2]);
        q2_sync<WAVE>();
        // z, y, x: QV -> WS -> QV -> U (= WS)
#pragma unroll
        for (int c = 0; c < NC; ++c)
          QV[c * 64 + l] = y[c];
        q2_sync<WAVE>();
#pragma unroll
        for (int c = 0; c < NC; ++c)
          WS[c * 64 + l] = q2_pass(QV + c * 64, L.base(2), 9, scol[2]);
        q2_sync<WAVE>();
#pragma unroll
        for (int c = 0; c < NC; ++c)
          QV[c * 64 + l] = q2_pass(WS + c * 64, L.base(1), 3, scol[1]);
        q2_sync<WAVE>();
#pragma unroll
        for (int c = 0; c < NC; ++c)
          y[c] = q2_pass(QV + c * 64, L.base(0), 1, scol[0]);
        q2_sync<WAVE>();
#pragma unroll
        for (int c = 0; c < NC; ++c)
          U[c * 64 + l] = y[c];
      }
    else
      {
        // y, x: QV -> WS -> U (= QV)
#pragma unroll
        for (int c = 0; c < NC; ++c)
          QV[c * 64 + l] = y[c];
        q2_sync<WAVE>();
#pragma unroll
        for (int c = 0; c < NC; ++c)
          WS[c * 64 + l] = q2_pass(QV + c * 64, L.base(1), 3, scol[1]);
        q2_sync<WAVE>();
#pragma unroll
        for (int c = 0; c < NC; ++c)
          y[c] = q2_pass(WS + c * 64, L.base(0), 1, scol[0]);
        q2_sync<WAVE>();
#pragma unroll
        for (int c = 0; c < NC; ++c)
          U[c * 64 + l] = y[c];
      }
  }

  // residual of the workgroup's cells, with the passes of all components
  // between two barriers as in q2_vmult_body.  U holds the gathered
  // iterate (in: nodal, out: integrated rows), H the history combination
  // sum_k w_k old_k (OP_RES2), A is the second work buffer.
  template <int DIM, int G, bool TENS, int OP, typename T>
  __device__ __forceinline__ void q2_residual_body(const Q2Params &kp, const Q2Lane<DIM> &L,
                                                   T *U, T *H, T *A, const T *inv_h)
  {
    using LN           = Q2Lane<DIM>;
    constexpr int  NC  = G + 2;
    constexpr bool HIST = OP == OP_RES2;
    const int      l   = L.l;
    T              srow[DIM][3];
#pragma unroll
    for (int d = 0; d < DIM; ++d)
      q2_s_row(L.t[d], srow[d]);
    // values of the iterate (U -> A -> U in 3D, U -> A in 2D)
    T v[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c)
      A[c * 64 + l] = q2_pass(U + c * 64, L.base(0), 1, srow[0]);
    __syncthreads();
    if constexpr (DIM == 3)
      {
#pragma unroll
        for (int c = 0; c < NC; ++c)
          U[c * 64 + l] = q2_pass(A + c * 64, L.base(1), 3, srow[1]);
        __syncthreads();
#pragma unroll
        for (int c = 0; c < NC; ++c)
          v[c] = q2_pass(U + c * 64, L.base(2), 9, srow[2]);
      }
    else
#pragma unroll
      for (int c = 0; c < NC; ++c)
        v[c] = q2_pass(A + c * 64, L.base(1), 3, srow[1]);
    // values of the history combination (H -> A -> H in 3D, H -> U in 2D;
    // every buffer written here was last read before the previous barrier)
    T hist[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c)
      hist[c] = T(0);
    if constexpr (HIST)
      {
        if constexpr (DIM == 3)
          {
#pragma unroll
            for (int c = 0; c < NC; ++c)
              A[c * 64 + l] = q2_pass(H + c * 64, L.base(0), 1, srow[0]);
            __syncthreads();
#pragma unroll
            for (int c = 0; c < NC; ++c)
              H[c * 64 + l] = q2_pass(A + c * 64, L.base(1), 3, srow[1]);
            __syncthreads();
#pragma unroll
            for (int c = 0; c < NC; ++c)
              if (c != 1)
                hist[c] = q2_pass(H + c * 64, L.base(2), 9, srow[2]);
          }
        else
          {
#pragma unroll
            for (int c = 0; c < NC; ++c)
              U[c * 64 + l] = q2_pass(H + c * 64, L.base(0), 1, srow[0]);
            __syncthreads();
#pragma unroll
            for (int c = 0; c < NC; ++c)
              if (c != 1)
                hist[c] = q2_pass(U + c * 64, L.base(1), 3, srow[1]);
          }
      }
    __syncthreads();
    // gradients from the Gauss-point values in A
#pragma unroll
    for (int c = 0; c < NC; ++c)
      A[c * 64 + l] = v[c];
    __syncthreads();
    T g[NC][DIM];
    {
      T drow[DIM][3];
#pragma unroll
      for (int d = 0; d < DIM; ++d)
        q2_dq_row(L.t[d], drow[d]);
#pragma unroll
      for (int c = 0; c < NC; ++c)
#pragma unroll
        for (int d = 0; d < DIM; ++d)
          g[c][d] = q2_pass(A + c * 64, L.base(d), LN::str(d), drow[d]) * inv_h[d];
    }
    T vr[NC], gr[NC][DIM];
    residual_qp<DIM, G, TENS, OP - OP_RES0>(kp.pp, v, g, hist, vr, gr);
    T jxw = T(kp.h_prod);
#pragma unroll
    for (int d = 0; d < DIM; ++d)
      jxw *= q2_w<T>(L.t[d]);
    __syncthreads();
    // integration: H takes the value rows, U and A the derivative passes
    T scol[DIM][3], dcol[DIM][3];
#pragma unroll
    for (int d = 0; d < DIM; ++d)
      {
        q2_s_col(L.t[d], scol[d]);
        q2_dq_col(L.t[d], dcol[d]);
      }
    T y[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c)
      {
        y[c]           = vr[c] * jxw;
        U[c * 64 + l]  = gr[c][0] * (jxw * inv_h[0]);
        A[c * 64 + l]  = gr[c][1] * (jxw * inv_h[1]);
      }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < NC; ++c)
      y[c] += q2_pass(U + c * 64, L.base(0), 1, dcol[0]) +
              q2_pass(A + c * 64, L.base(1), 3, dcol[1]);
    __syncthreads();
    if constexpr (DIM == 3)
      {
#pragma unroll
        for (int c = 0; c < NC; ++c)
          U[c * 64 + l] = gr[c][2] * (jxw * inv_h[2]);
        __syncthreads();
#pragma unroll
        for (int c = 0; c < NC; ++c)
          y[c] += q2_pass(U + c * 64, L.base(2), 9, dcol[2]);
        __syncthreads();
        // z, y, x: A -> U -> A -> U
#pragma unroll
        for (int c = 0; c < NC; ++c)
          A[c * 64 + l] = y[c];
        __syncthreads();
#pragma unroll
        for (int c = 0; c < NC; ++c)
          U[c * 64 + l] = q2_pass(A + c * 64, L.base(2), 9, scol[2]);
        __syncthreads();
#pragma unroll
        for (int c = 0; c < NC; ++c)
          A[c * 64 + l] = q2_pass(U + c * 64, L.base(1), 3, scol[1]);
        __syncthreads();
#pragma unroll
        for (int c = 0; c < NC; ++c)
          y[c] = q2_pass(A + c * 64, L.base(0), 1, scol[0]);
      }
    else
      {
        // y, x: A -> U -> registers
#pragma unroll
        for (int c = 0; c < NC; ++c)
          A[c * 64 + l] = y[c];
        __syncthreads();
#pragma unroll
        for (int c = 0; c < NC; ++c)
          U[c * 64 + l] = q2_pass(A + c * 64, L.base(1), 3, scol[1]);
        __syncthreads();
#pragma unroll
        for (int c = 0; c < NC; ++c)
          y[c] = q2_pass(U + c * 64, L.base(0), 1, scol[0]);
      }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < NC; ++c)
      U[c * 64 + l] = y[c];
  }

  // ----------------------------------------- tangent vmult and residual
  template <int DIM, int G, bool TENS, int OP, typename T>
  __global__ void __launch_bounds__(64)
    q2_apply(const Q2Params kp, const T *__restrict__ src, const double *__restrict__ old0,
             const double *__restrict__ old1, const double *__restrict__ old2,
             const T *__restrict__ cache, T *__restrict__ dst)
  {
    using CL          = CacheLayout<DIM, G, TENS>;
    constexpr int NC  = G + 2;
    constexpr int NF  = CL::n_fields;
    constexpr bool HIST = OP == OP_RES2;

    __shared__ T U[NC * 64];
    __shared__ T H[(HIST || OP == OP_VMULT) ? NC * 64 : 1];
    __shared__ T A[NC * 64];

    const Q2Lane<DIM> L(kp);
    q2_gather<DIM, NC, T>(kp, L, U, [&](int64_t n, T *x) { q2_load_row<NC>(src + n * NC, x); });
    if constexpr (HIST)
      q2_gather<DIM, NC, T>(kp, L, H, [&](int64_t n, T *x) {
        // do_evalute_history, S/operator_sintering_generic.h:950-960
        T o[NC];
        q2_load_row<NC>(old0 + n * NC, o);
#pragma unroll
        for (int c = 0; c < NC; ++c)
          x[c] = o[c] * T(kp.w_hist[0]);
        if (kp.order > 1)
          {
            q2_load_row<NC>(old1 + n * NC, o);
#pragma unroll
            for (int c = 0; c < NC; ++c)
              x[c] += o[c] * T(kp.w_hist[1]);
          }
        if (kp.order > 2)
          {
            q2_load_row<NC>(old2 + n * NC, o);
#pragma unroll
            for (int c = 0; c < NC; ++c)
              x[c] += o[c] * T(kp.w_hist[2]);
          }
      });
    __syncthreads();

    const T inv_h[3] = {T(kp.inv_h[0]), T(kp.inv_h[1]), T(kp.inv_h[2])};
    if constexpr (OP == OP_VMULT)
      q2_vmult_body<DIM, G, TENS>(kp, L, U, H, A,
                                  cache + ((kp.wg_base + blockIdx.x) * NF) * Q2Lane<DIM>::LP +
                                    L.l,
                                  inv_h);
    else
      q2_residual_body<DIM, G, TENS, OP>(kp, L, U, H, A, inv_h);
    __syncthreads();
    q2_scatter<DIM, NC, T>(kp, L, U, dst, NC, 1, 0);
  }

  // ------------------------------------------ 3D tangent vmult, tile march
  // A workgroup owns a tile of 4 x NY cells in (x, y) and marches through up
  // to tile_zc cell layers along z.  Wave w = xp + 2 r holds the cells
  // x = 4 tx + xp + {0, 2} of row y = NY ty + r: two cells of one colour, a
  // cell pair of the linearisation cache, so a wave reads its cache entries
  // as in the colour passes.  Per layer the waves run the cell bodies on LDS
  // rows of their own (wave-level barriers); then each node of the tile's
  // three node planes is summed by one owner lane, the lane of the lowest
  // cell touching it, from the rows of the up to four cells touching it in
  // (x, y), in a fixed order, and in the bottom plane the top plane of the
  // previous layer, carried in LDS.  The two lower planes are written, the
  // top plane is carried (and written after the last layer).  Tiles of one
  // colour (parity of the tile index along x, y, z) share no node; a node
  // on a tile boundary is stored by the first colour that reaches it and
  // added by the later ones, as in the colour passes.  src and the cache of
  // the tile are read once per layer; every node of a tile is written once.
  // WPE: waves per SIMD the register allocation targets
  template <int G, bool TENS, typename T, int NY, int WPE>
  __global__ void __launch_bounds__(128 * NY) __attribute__((amdgpu_waves_per_eu(WPE)))
    q2_vmult_tile(const Q2Params kp, const Q2TileParams tp, const T *__restrict__ src,
                  const T *__restrict__ cache, T *__restrict__ dst)
  {
    using CL          = CacheLayout<3, G, TENS>;
    using LN          = Q2Lane<3>;
    constexpr int NC  = G + 2;
    constexpr int NF  = CL::n_fields;
    constexpr int NW  = 2 * NY;
    constexpr int PXY = 9 * (2 * NY + 1); // nodes of a tile plane
    __shared__ T U[NW][NC * 64];
    __shared__ T H[NW][NC * 64];
    __shared__ T A[NW][NC * 64];
    __shared__ T carry[NC * PXY];

    const int b  = (int)blockIdx.x;
    const int q  = b / tp.ntc[0];
    const int tx = 2 * (b - q * tp.ntc[0]) + tp.tpar[0];
    const int ty = 2 * (q % tp.ntc[1]) + tp.tpar[1];
    const int tz = 2 * (q / tp.ntc[1]) + tp.tpar[2];
    // wave index as a uniform value: the cell colour and the cache block
    // offsets below stay scalar
    const int w   = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
    const int l   = (int)threadIdx.x & 63;
    const int xp  = w & 1, cyl = w >> 1;
    const int sl  = l < 2 * LN::NP ? l / LN::NP : 0;
    const int cxl = xp + 2 * sl; // cell of the lane in the tile
    const int cx = 4 * tx + cxl, cy = NY * ty + cyl;
    const int z0 = kp.tile_zc * tz;
    const int nl = kp.nc[2] - z0 < kp.tile_zc ? kp.nc[2] - z0 : kp.tile_zc;
    const bool in_xy = l < 2 * LN::NP && cx < kp.nc[0] && cy < kp.nc[1];
    // cells of the wave's colour along x and y (cache pair blocks)
    const int py  = cy & 1;
    const int ncx = (kp.nc[0] - xp + 1) >> 1, ncy = (kp.nc[1] - py + 1) >> 1;
    const T   inv_h[3] = {T(kp.inv_h[0]), T(kp.inv_h[1]), T(kp.inv_h[2])};
    // owner of the lane's node in the tile and the cells adding to it
    const int  wx = w ^ 1, sx = (cxl + 1) >> 1; // cell (cxl + 1, cyl)

    // the src row of the lane's node in layer k (zeros outside the mesh);
    // the rows of layer k + 1 are requested after the body of layer k
    T    xs[NC];
    auto load_src = [&](int k) {
      // the lane index through an empty asm statement, as in the march
      // below: the point index and address are formed per call, not held
      // in registers across the cell body
      int lq = l;
      asm volatile("" : "+v"(lq));
      const int cc[3] = {cx, cy, z0 + k};
      const int p     = lq - (lq / LN::NP) * LN::NP;
      if (in_xy)
        q2_load_row<NC>(src + LN::node(kp, cc, p) * NC, xs);
      else
#pragma unroll
        for (int c = 0; c < NC; ++c)
          xs[c] = T(0);
    };
    load_src(0);

    for (int k = 0; k < nl; ++k)
      {
        // the lane index passes through an empty asm statement each layer:
        // everything derived from it (node and point indices, LDS offsets,
        // coefficient rows) is recomputed per layer instead of being held in
        // registers across the march
        int lk = l;
        asm volatile("" : "+v"(lk));
        Q2Params kq = kp;
        asm volatile("" : "+s"(kq.pp.A), "+s"(kq.pp.B), "+s"(kq.pp.kappa_c), "+s"(kq.pp.kappa_p));
        asm volatile("" : "+s"(kq.pp.Mvol), "+s"(kq.pp.Mvap), "+s"(kq.pp.Msurf), "+s"(kq.pp.Mgb));
        asm volatile("" : "+s"(kq.pp.L), "+s"(kq.pp.w0), "+s"(kq.pp.w_alg), "+s"(kq.h_prod));
        asm volatile("" : "+s"(kq.inv_h[0]), "+s"(kq.inv_h[1]), "+s"(kq.inv_h[2]));
        const T inv_hk[3] = {T(kq.inv_h[0]), T(kq.inv_h[1]), T(kq.inv_h[2])};
        const int cz    = z0 + k;
        const int cc[3] = {cx, cy, cz};
        const LN  L(lk, cc, in_xy);
#pragma unroll
        for (int c = 0; c < NC; ++c)
          U[w][c * 64 + l] = xs[c];
        q2_sync<true>();
        const int     col = xp + 2 * py + 4 * (cz & 1);
        const int64_t ci  = ((int64_t)(cz >> 1) * ncy + (cy >> 1)) * ncx + (cx >> 1);
        const T      *cl  = cache + ((tp.wgb[col] + (ci >> 1)) * NF) * LN::LP +
                      (int)(ci & 1) * LN::NP + L.p;
        q2_vmult_body<3, G, TENS, T, true, sizeof(T) == 8>(kq, L, U[w], H[w], A[w], cl, inv_hk);
        __syncthreads();
        // the next layer's src rows, in flight during the node sums (set on
        // both paths, so that this layer's rows are dead after the gather;
        // 2 % faster FP32 at 96^3 with 10 order parameters, FP64 unchanged;
        // requesting the next layer's Q_i fields here as well was 3 %
        // slower)
        if (k + 1 < nl)
          load_src(k + 1);
        else
#pragma unroll
          for (int c = 0; c < NC; ++c)
            xs[c] = T(0);

        const int  t0 = L.t[0], t1 = L.t[1], t2 = L.t[2];
        const int  pn = (2 * cxl + t0) + 9 * (2 * cyl + t1);
        const bool own = L.valid && (t0 != 0 || cxl == 0) && (t1 != 0 || cyl == 0);
        T          v[NC];
        bool       keep = false;
        if (own)
          {
            const bool xr = t0 == 2 && cxl < 3 && cx + 1 < kp.nc[0];
            const bool yu = t1 == 2 && cyl < NY - 1 && cy + 1 < kp.nc[1];
#pragma unroll
            for (int c = 0; c < NC; ++c)
              v[c] = U[w][c * 64 + l];
            if (xr)
#pragma unroll
              for (int c = 0; c < NC; ++c)
                v[c] += U[wx][c * 64 + sx * LN::NP + L.p - 2];
            if (yu)
#pragma unroll
              for (int c = 0; c < NC; ++c)
                v[c] += U[w + 2][c * 64 + l - 6];
            if (xr && yu)
#pragma unroll
              for (int c = 0; c < NC; ++c)
                v[c] += U[wx + 2][c * 64 + sx * LN::NP + L.p - 8];
            if (t2 == 0 && k > 0)
#pragma unroll
              for (int c = 0; c < NC; ++c)
                v[c] += carry[c * PXY + pn];
            if (t2 < 2 || k == nl - 1)
              {
                const int  gx = 2 * cx + t0, gy = 2 * cy + t1, gz = 2 * cz + t2;
                const bool shx = (gx == 8 * tx && tx > 0) ||
                                 (gx == 8 * tx + 8 && gx < kp.nn[0] - 1);
                const bool shy = (gy == 2 * NY * ty && ty > 0) ||
                                 (gy == 2 * NY * (ty + 1) && gy < kp.nn[1] - 1);
                const bool shz = (t2 == 0 && k == 0 && tz > 0) || (t2 == 2 && gz < kp.nn[2] - 1);
                const bool first = (!shx || (tx & 1) == 0) && (!shy || (ty & 1) == 0) &&
                                   (!shz || (tz & 1) == 0);
                T *d = dst + (gx + (int64_t)kp.nn[0] * (gy + (int64_t)kp.nn[1] * gz)) * NC;
                if (!first)
                  {
                    T o[NC];
                    q2_load_row<NC>(d, o);
#pragma unroll
                    for (int c = 0; c < NC; ++c)
                      v[c] = o[c] + v[c];
                  }
                q2_store_row<NC>(d, v);
              }
            else
              keep = true;
          }
        __syncthreads();
        if (keep)
#pragma unroll
          for (int c = 0; c < NC; ++c)
            carry[c * PXY + pn] = v[c];
      }
  }

  // ------------------------------------------------ linearisation cache
  template <int DIM, int G, bool TENS>
  __global__ void __launch_bounds__(64)
    q2_fill(const Q2Params kp, const double *__restrict__ lin, double *__restrict__ cache64,
            float *__restrict__ cache32)
  {
    using CL         = CacheLayout<DIM, G, TENS>;
    using LN         = Q2Lane<DIM>;
    constexpr int NC = G + 2;
    constexpr int NF = CL::n_fields;
    constexpr int NP = DIM * (DIM + 1) / 2;
    constexpr int LP = LN::LP;
    using T          = double;

    __shared__ T U[NC * 64];
    __shared__ T A[NC * 64];
    __shared__ T LG[TENS ? G * DIM * 64 : 1]; // grad eta_i, this lane's entries

    const LN  L(kp);
    const int l = L.l;
    q2_gather<DIM, NC, T>(kp, L, U, [&](int64_t n, T *x) { q2_load_row<NC>(lin + n * NC, x); });
    __syncthreads();
    const T inv_h[3] = {kp.inv_h[0], kp.inv_h[1], kp.inv_h[2]};

    // values at the Gauss points, one pass per direction for all components
    T v[NC];
    {
      T srow[DIM][3];
#pragma unroll
      for (int d = 0; d < DIM; ++d)
        q2_s_row(L.t[d], srow[d]);
#pragma unroll
      for (int c = 0; c < NC; ++c)
        A[c * 64 + l] = q2_pass(U + c * 64, L.base(0), 1, srow[0]);
      __syncthreads();
      if constexpr (DIM == 3)
        {
#pragma unroll
          for (int c = 0; c < NC; ++c)
            U[c * 64 + l] = q2_pass(A + c * 64, L.base(1), 3, srow[1]);
          __syncthreads();
#pragma unroll
          for (int c = 0; c < NC; ++c)
            v[c] = q2_pass(U + c * 64, L.base(2), 9, srow[2]);
        }
      else
#pragma unroll
        for (int c = 0; c < NC; ++c)
          v[c] = q2_pass(A + c * 64, L.base(1), 3, srow[1]);
    }
    T *QV = DIM == 3 ? A : U; // the buffer the last pass did not read
#pragma unroll
    for (int c = 0; c < NC; ++c)
      QV[c * 64 + l] = v[c];
    __syncthreads();

    // cache: [workgroup][field][LP lanes]; lanes past the cell slots hold
    // nothing
    const bool    wr = l < LP;
    const int64_t o  = ((kp.wg_base + blockIdx.x) * NF) * LP + l;
    auto          put = [&](int f, T x) {
      if (wr)
        {
          cache64[o + f * LP] = x;
          if (cache32)
            cache32[o + f * LP] = (float)x;
        }
    };
    put(CL::c, v[0]);
#pragma unroll
    for (int i = 0; i < G; ++i)
      put(CL::eta + i, v[2 + i]);
    T drow[DIM][3];
#pragma unroll
    for (int d = 0; d < DIM; ++d)
      q2_dq_row(L.t[d], drow[d]);
    T lmu[DIM];
#pragma unroll
    for (int d = 0; d < DIM; ++d)
      {
        lmu[d] = q2_pass(QV + 64, L.base(d), LN::str(d), drow[d]) * inv_h[d];
        put(CL::grad_mu + d, lmu[d]);
      }
    if constexpr (TENS)
      {
#pragma unroll
        for (int d = 0; d < DIM; ++d)
          put(CL::grad_c + d, q2_pass(QV, L.base(d), LN::str(d), drow[d]) * inv_h[d]);
#pragma unroll
        for (int i = 0; i < G; ++i)
#pragma unroll
          for (int d = 0; d < DIM; ++d)
            LG[(i * DIM + d) * 64 + l] =
              q2_pass(QV + (2 + i) * 64, L.base(d), LN::str(d), drow[d]) * inv_h[d];
        // the grain-boundary sums of gb_fields (physics.hpp), same pair
        // order and arithmetic; grad eta_i and eta_i are this lane's LDS
        // entries, the pair loop runs over i with j unrolled (Q in registers)
        T S = T(0), P[NP], Q[G * DIM];
#pragma unroll
        for (int k = 0; k < NP; ++k)
          P[k] = T(0);
#pragma unroll
        for (int k = 0; k < G * DIM; ++k)
          Q[k] = T(0);
#pragma unroll 1
        for (int i = 1; i < G; ++i)
          {
            T li[DIM], qi[DIM];
#pragma unroll
            for (int d = 0; d < DIM; ++d)
              {
                li[d] = LG[(i * DIM + d) * 64 + l];
                qi[d] = T(0);
              }
            const T ei = QV[(2 + i) * 64 + l];
#pragma unroll
            for (int j = 0; j < G - 1; ++j)
              if (j < i)
                {
                  T egd[DIM];
#pragma unroll
                  for (int d = 0; d < DIM; ++d)
                    egd[d] = li[d] - LG[(j * DIM + d) * 64 + l];
                  const T filter = unit_vector_filter_2<DIM>(egd);
                  const T ej     = QV[(2 + j) * 64 + l];
                  const T ee     = ei * ej;
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
                      qi[d] += ej * r;
                      Q[j * DIM + d] += ei * r;
                    }
                }
            // row i's own pairs come before its pairs as the second grain
#pragma unroll
            for (int k = 0; k < G; ++k)
              if (k == i)
#pragma unroll
                for (int d = 0; d < DIM; ++d)
                  Q[k * DIM + d] += qi[d];
          }
        put(CL::gb_s, S);
#pragma unroll
        for (int k = 0; k < NP; ++k)
          put(CL::gb_p + k, P[k]);
#pragma unroll
        for (int k = 0; k < G * DIM; ++k)
          put(CL::gb_q + k, Q[k]);
      }
  }

  // --------------------------------- diagonal / block column by probing
  // For every local node j of the cell and input component b (all b for the
  // diagonal, b = column for a block column) the cell's contribution to row
  // (j, a) of the tangent, a = b for the diagonal and every a for a block
  // column: sum_q JxW [phi_j (V_a) + grad phi_j . (G_a)] with (V, G) the
  // quadrature operator applied to the increment e_b phi_j.  The operator is
  // linear in (value, gradient) of the increment, so it is evaluated once
  // per direction m of (value, d/dx, ...) with a unit input, and the node
  // sums over the Gauss points run through LDS.  The result goes to
  // dst[node * os + a * ocs + oo].
  // ------------------------------------------------ diagonal / node blocks
  // Node block B[j][a][b] = sum_q JxW psi_j,m'(q) M_{a m', b m}(q) psi_j,m(q)
  // (psi_j,0 = phi_j, psi_j,d+1 = d phi_j / dx_d; M: the quadrature-point
  // tangent from input (component b, value or gradient d) to output
  // (component a, value or gradient d')).  M is sparse: value rows read
  // values only; the c gradient row reads the values and the gradients of
  // c and mu; the mu and eta_i gradient rows are -kappa_c grad c and
  // L kappa_p grad eta_i.  Each nonzero term is a tensor product of 1D
  // weights w_q f(x_q) g(x_q) with f, g in {s_j, s_j'/h} per direction, so
  // the q-sum is three sum-factorisation passes over the field M(q); terms
  // with the same 1D weights share a field.  Rows (NR = n_c + 12):
  //   r < n_c        phi phi           M_{r 0, b 0}          -> B[.][r][b]
  //   n_c + d        phi' phi (dir d)  M_{0 d+1, b 0}        -> B[.][0][b]
  //   n_c + 3 + d    phi' phi' (d, d)  M_{0 d+1, b d+1}      -> B[.][0][b]
  //   n_c + 6 + p    phi' phi' (d, e)  M_{0 d+1, b e+1} + M_{0 e+1, b d+1}
  //   n_c + 9 + d    phi' phi' (d, d)  -kappa_c (b = 0) -> B[.][1][0],
  //                                    L kappa_p (b >= 2) -> B[.][b][b]
  // The diagonal (column < 0) uses the same rows with the mass field of
  // row b taken from input b, the c-row fields from input c, and the
  // stiffness rows with field 1 scaled by L kappa_p for every b >= 2.
  template <int DIM>
  struct Q2ProbeRows
  {
    // 1D weight type of row r along direction d: 0 phi phi, 1 phi' phi,
    // 2 phi' phi'
    __device__ static constexpr int type(int nc, int r, int d)
    {
      if (r < nc)
        return 0;
      const int k = r - nc;
      if (k < 3)
        return k == d ? 1 : 0;
      if (k < 6 || k >= 9)
        return (k % 3) == d ? 2 : 0;
      // pairs (0,1), (0,2), (1,2)
      const int p = k - 6;
      const int e0 = p == 2 ? 1 : 0, e1 = p == 0 ? 1 : 2;
      return (d == e0 || d == e1) ? 1 : 0;
    }
  };

  template <int DIM, int G, bool TENS, typename T>
  __global__ void __launch_bounds__(64)
    q2_probe(const Q2Params kp, const T *__restrict__ cache, T *__restrict__ dst, int column,
             int os, int ocs, int oo)
  {
    using CL         = CacheLayout<DIM, G, TENS>;
    constexpr int NC = G + 2;
    constexpr int NF = CL::n_fields;
    constexpr int NR = NC + 12;
    using LN         = Q2Lane<DIM>;
    using RW         = Q2ProbeRows<DIM>;

    __shared__ T A[NR * 64];
    __shared__ T B[NR * 64];

    const LN  L(kp);
    const int l    = L.l;
    const T   kc   = -T(kp.pp.kappa_c);
    const T   lkp  = T(kp.pp.L) * T(kp.pp.kappa_p);

    // entries of block column b (b < 0: the diagonal) of this lane's node,
    // left in A[a * 64 + lane]
    auto column_entries = [&](const int b) {
      // 1) fields of the rows at this lane's Gauss point, into A
      {
        // (read again for every column, from the caches: the opaque offset
        // keeps the loads inside the column loop, so the fields are not
        // held in registers across it)
        int off = 0;
        asm volatile("" : "+v"(off));
        T        lin[NF];
        const T *cl = cache + ((kp.wg_base + blockIdx.x) * NF) * LN::LP + l + off;
#pragma unroll
        for (int f = 0; f < NF; ++f)
          lin[f] = L.valid ? cl[f * LN::LP] : T(0);
        const T zero[3] = {T(0), T(0), T(0)};
        // c-row gradient outputs for the inputs of component cb: value (m = 0)
        // and gradient d (m = d + 1)
        auto c_rows = [&](int cb) {
          T v[NC];
#pragma unroll
          for (int c = 0; c < NC; ++c)
            v[c] = c == cb ? T(1) : T(0);
          T gm[DIM + 1][DIM];
          tangent_grad_c<DIM, G, TENS>(kp.pp, lin, v, zero, zero, gm[0]);
#pragma unroll
          for (int c = 0; c < NC; ++c)
            v[c] = T(0);
#pragma unroll
          for (int d = 0; d < DIM; ++d)
            {
              T u[3] = {T(0), T(0), T(0)};
              u[d]   = T(1);
              if (cb <= 1)
                tangent_grad_c<DIM, G, TENS>(kp.pp, lin, v, cb == 0 ? u : zero,
                                             cb == 1 ? u : zero, gm[d + 1]);
              else
#pragma unroll
                for (int e = 0; e < DIM; ++e)
                  gm[d + 1][e] = T(0);
            }
#pragma unroll
          for (int d = 0; d < 3; ++d)
            {
              A[(NC + d) * 64 + l]     = d < DIM ? gm[0][d] : T(0);
              A[(NC + 3 + d) * 64 + l] = d < DIM ? gm[d + 1][d] : T(0);
            }
          // pairs (0,1), (0,2), (1,2)
          A[(NC + 6) * 64 + l] = gm[2][0] + gm[1][1];
          A[(NC + 7) * 64 + l] = DIM == 3 ? gm[3 % (DIM + 1)][0] + gm[1][2 % DIM] : T(0);
          A[(NC + 8) * 64 + l] = DIM == 3 ? gm[3 % (DIM + 1)][1] + gm[2][2 % DIM] : T(0);
        };
        if (b >= 0)
          {
            T v[NC], vr[NC];
#pragma unroll
            for (int c = 0; c < NC; ++c)
              v[c] = c == b ? T(1) : T(0);
            tangent_values<DIM, G, TENS>(kp.pp, lin, v, vr);
#pragma unroll
            for (int c = 0; c < NC; ++c)
              A[c * 64 + l] = vr[c];
            c_rows(b);
            const T st = b == 0 ? kc : (b >= 2 ? lkp : T(0));
#pragma unroll
            for (int d = 0; d < 3; ++d)
              A[(NC + 9 + d) * 64 + l] = d < DIM ? st : T(0);
          }
        else
          {
#pragma unroll 1
            for (int cb = 0; cb < NC; ++cb)
              {
                T v[NC], vr[NC];
#pragma unroll
                for (int c = 0; c < NC; ++c)
                  v[c] = c == cb ? T(1) : T(0);
                tangent_values<DIM, G, TENS>(kp.pp, lin, v, vr);
                T x = T(0);
#pragma unroll
                for (int c = 0; c < NC; ++c)
                  if (c == cb)
                    x = vr[c];
                A[cb * 64 + l] = x;
              }
            c_rows(0);
#pragma unroll
            for (int d = 0; d < 3; ++d)
              A[(NC + 9 + d) * 64 + l] = d < DIM ? T(1) : T(0);
          }
      }
      __syncthreads();

      // 2) q -> node sums, one pass per direction over all rows
      //    (A -> B -> A in 3D, the last pass into registers)
      T res[NR];
      {
        T K[3][3]; // [type][q] for the current direction
        auto coefs = [&](int d) {
          T sc[3], ec[3];
          q2_s_col(L.t[d], sc);
          q2_e_col(L.t[d], ec);
#pragma unroll
          for (int q = 0; q < 3; ++q)
            {
              const T w = q2_w<T>(q);
              const T e = ec[q] * T(kp.inv_h[d]);
              K[0][q]   = w * sc[q] * sc[q];
              K[1][q]   = w * sc[q] * e;
              K[2][q]   = w * e * e;
            }
        };
        coefs(0);
#pragma unroll
        for (int r = 0; r < NR; ++r)
          B[r * 64 + l] = q2_pass(A + r * 64, L.base(0), 1, K[RW::type(NC, r, 0)]);
        __syncthreads();
        coefs(1);
        if constexpr (DIM == 3)
          {
#pragma unroll
            for (int r = 0; r < NR; ++r)
              A[r * 64 + l] = q2_pass(B + r * 64, L.base(1), 3, K[RW::type(NC, r, 1)]);
            __syncthreads();
            coefs(2);
#pragma unroll
            for (int r = 0; r < NR; ++r)
              res[r] = q2_pass(A + r * 64, L.base(2), 9, K[RW::type(NC, r, 2)]);
          }
        else
#pragma unroll
          for (int r = 0; r < NR; ++r)
            res[r] = q2_pass(B + r * 64, L.base(1), 3, K[RW::type(NC, r, 1)]);
      }

      // 3) rows to outputs
      T acc[NC];
#pragma unroll
      for (int a = 0; a < NC; ++a)
        acc[a] = res[a];
      T c_sum = T(0), st_sum = T(0);
#pragma unroll
      for (int r = NC; r < NC + 9; ++r)
        c_sum += res[r];
#pragma unroll
      for (int r = NC + 9; r < NR; ++r)
        st_sum += res[r];
      const T hp = T(kp.h_prod);
      if (b >= 0)
        {
          acc[0] += c_sum;
#pragma unroll
          for (int a = 1; a < NC; ++a)
            if (a == (b == 0 ? 1 : b) && b != 1)
              acc[a] += st_sum;
        }
      else
        {
          acc[0] += c_sum;
#pragma unroll
          for (int a = 2; a < NC; ++a)
            acc[a] += lkp * st_sum;
        }
      __syncthreads();
#pragma unroll
      for (int a = 0; a < NC; ++a)
        A[a * 64 + l] = hp * acc[a];
      __syncthreads();
    };
    // column -2: all block columns, column b of a node's block at
    // dst + node * os + b * NC (the workgroup's cache stays in the caches
    // between columns); otherwise the one column (or the diagonal) asked for
    const bool all = column < -1;
    const int  b0 = all ? 0 : column, b1 = all ? NC : column + 1;
#pragma unroll 1
    for (int b = b0; b < b1; ++b)
      {
        column_entries(b);
        q2_scatter<DIM, NC, T>(kp, L, A, dst, os, all ? 1 : ocs, all ? b * NC : oo);
        __syncthreads();
      }
  }
} // namespace hps
