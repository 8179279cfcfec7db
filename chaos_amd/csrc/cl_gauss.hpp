// cl_gauss.hpp - the Henon-Heiles flow under the s = 3 Gauss-Legendre method as a map of the tangent
// stream: IntegratorMap(Gauss(Float64, 3; tol), HenonHeilesFlow(), h) (src/Integrators/integrator_map.jl:12-23).
//
// One step, per lane:
//  - the stage equations by the fixed-point iteration of src/Integrators/gauss.jl:122-141: start
//    Y_i = z + (h sum_j A_ij) f(z); iterate F_i = f(Y_i), Y_i = z + h sum_j (Ahi_ij + Alo_ij) F_j with the
//    compensated dot product of gauss.jl:90-101 (TwoProd by fma, TwoSum); stop when the largest stage
//    increment is 0 or no longer decreases; accept when it is at most `tol` (or 8 eps max(1, |Y|));
//  - the increment h sum_i (bhi_i + blo_i) F_i, compensated the same way;
//  - the one-step Jacobian from the converged stages. The reference solves the 12 x 12 system
//    (I - h A (x) J) dY = 1 (x) I with a dense LU (gauss.jl:176-192). With J_j = [[0, I], [-H_j, 0]] the
//    q-blocks eliminate exactly: Q_i + h^2 sum_jl a_ij a_jl H_j Q_l = E_p - h sum_j a_ij H_j E_q, a 6 x 6 system
//    with 4 right-hand sides that stays in registers; then P_i = E_q + h sum_j a_ij Q_j and
//    J_step = I + h sum_i b_i [Q_i; -H_i P_i]. Same linear algebra, a different rounding than LAPACK's.
// The reference evaluates the stage iteration twice per step (once in `tangent`, once in `forward`,
// with identical inputs); here it runs once and serves both.
#pragma once
#include "cl_core.hpp"

namespace cl {

// s = 3 Gauss coefficients, hi + lo (chaos_amd/csrc/gen_gauss_coefficients.py)
#define CL_GA_HI {{0x1.1c71c71c71c72p-3, -0x1.26b88a4e09a62p-5, 0x1.40c7cef225974p-7}, \
                  {0x1.337831ea8a881p-2, 0x1.c71c71c71c71cp-3, -0x1.7066ace18c0fbp-6}, \
                  {0x1.126b88a4e09a6p-2, 0x1.ebf38310dda69p-2, 0x1.1c71c71c71c72p-3}}
#define CL_GA_LO {{-0x1.c71c71c71c71cp-58, -0x1.f9a1c81ea4ba2p-60, 0x1.94971b6f94297p-61}, \
                  {0x1.d03218eda4366p-56, 0x1.c71c71c71c71cp-57, 0x1.87f5c5d9b2176p-61}, \
                  {0x1.1f9a1c81ea4bap-57, -0x1.194971b6f9429p-56, -0x1.c71c71c71c71cp-58}}
#define CL_GB_HI {0x1.1c71c71c71c72p-2, 0x1.c71c71c71c71cp-2, 0x1.1c71c71c71c72p-2}
#define CL_GB_LO {-0x1.c71c71c71c71cp-57, 0x1.c71c71c71c71cp-56, -0x1.c71c71c71c71cp-57}

// Compensated sum_j (wh_j + wl_j) F_j over the three stages (gauss.jl:91-101).
CL_HD double dot_comp3(const double* wh, const double* wl, const double* F) {
  double s = 0.0, c = 0.0;
#pragma unroll
  for (int j = 0; j < 3; ++j) {
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      double w = t == 0 ? wh[j] : wl[j];
      double p, pe, se;
      cl_two_prod(w, F[j], &p, &pe);
      cl_two_sum(s, p, &s, &se);
      c = c + (se + pe);
    }
  }
  return s + c;
}

struct HHGauss {
  static constexpr int D = 4;
  static constexpr int S = 4;
  double h, tol;  // tol NaN: 8 eps max(1, |Y|)
  int maxiter;
  struct Aux {
    bool ok;
    double inc[4];
    double J[16];  // the one-step Jacobian at the converged stages, column-major (all NaN when the solve failed)
  };

  // J grad H for H = |p|^2/2 + |q|^2/2 + q1^2 q2 - q2^3/3 (src/Systems/flow_models.jl:101-103, flow_core.jl:75)
  static CL_HD void vf(const double* z, double* f) {
    f[0] = z[2];
    f[1] = z[3];
    f[2] = -(z[0] + 2.0 * z[0] * z[1]);
    f[3] = -((z[1] + z[0] * z[0]) - z[1] * z[1]);
  }

  CL_HD void aux(const double* z, Aux& a) const { aux_e(z, nullptr, a); }

  // The stage solve at z with the correction e of a compensated state (null: none). The stage values are
  // z + (e + h sum_j a_ij F_j) (gauss.jl:117); the start and the first field use z alone (gauss.jl:123-124).
  CL_HD void aux_e(const double* z, const double* e, Aux& a) const {
    const double Ahi[3][3] = CL_GA_HI, Alo[3][3] = CL_GA_LO;
    const double bhi[3] = CL_GB_HI, blo[3] = CL_GB_LO;
    double f0[4];
    vf(z, f0);
    double Y[3][4], F[3][4];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      double rs = h * ((Ahi[i][0] + Ahi[i][1]) + Ahi[i][2]);
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        Y[i][c] = z[c] + rs * f0[c];
        F[i][c] = f0[c];
      }
    }
    double delta = 0.0, prev = __builtin_inf();
    bool stopped = false;
    int k = 0;
    while (k < maxiter && !stopped) {
      k += 1;
#pragma unroll
      for (int i = 0; i < 3; ++i) vf(Y[i], F[i]);
      double dmax = 0.0;
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        double di = 0.0;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          double Fc[3] = {F[0][c], F[1][c], F[2][c]};
          double ss = h * dot_comp3(Ahi[i], Alo[i], Fc);
          double yn = z[c] + ((e ? e[c] : 0.0) + ss);
          double e = cl_fabs(yn - Y[i][c]);
          di = c == 0 ? e : cl_jmax(di, e);
          Y[i][c] = yn;
        }
        dmax = i == 0 ? di : cl_jmax(dmax, di);
      }
      delta = dmax;
      if (!cl_isfinite(delta)) break;
      stopped = delta == 0.0 || (k >= 2 && delta >= prev);
      prev = delta;
    }
    double bound;
    if (cl_isnan(tol)) {
      double mx = 0.0;
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int c = 0; c < 4; ++c) mx = (i == 0 && c == 0) ? cl_fabs(Y[0][0]) : cl_jmax(mx, cl_fabs(Y[i][c]));
      bound = 8.0 * CL_EPS64 * cl_jmax(1.0, mx);
    } else {
      bound = tol;
    }
    a.ok = stopped && cl_isfinite(delta) && delta <= bound;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      double Fc[3] = {F[0][c], F[1][c], F[2][c]};
      a.inc[c] = h * dot_comp3(bhi, blo, Fc);
    }
    double q[3][2];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      q[i][0] = Y[i][0];
      q[i][1] = Y[i][1];
    }
    step_jacobian(q, a.J);
    if (!a.ok) {
#pragma unroll
      for (int e = 0; e < 16; ++e) a.J[e] = __builtin_nan("");
    }
  }

  // The one-step Jacobian (4 x 4, column-major J[c * 4 + r]) from the converged stage positions q. B is factored
  // once (partial pivoting, first largest |pivot|; the multipliers kept below the diagonal and rows swapped whole),
  // then the four right-hand sides are solved one column at a time: all the row swaps first, then the forward
  // elimination and back substitution. Every entry of a right-hand side sees the same multipliers and pivot rows in
  // the same order as when the columns are eliminated alongside B, with a quarter of the live registers.
  CL_HD void step_jacobian(const double (&q)[3][2], double* J) const {
    const double Ahi[3][3] = CL_GA_HI, Alo[3][3] = CL_GA_LO;
    const double bhi[3] = CL_GB_HI, blo[3] = CL_GB_LO;
    double ha[3][3], hb[3], H[3][2][2];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
#pragma unroll
      for (int j = 0; j < 3; ++j) ha[i][j] = h * (Ahi[i][j] + Alo[i][j]);
      hb[i] = h * (bhi[i] + blo[i]);
      H[i][0][0] = 1.0 + 2.0 * q[i][1];
      H[i][1][0] = 2.0 * q[i][0];
      H[i][0][1] = 2.0 * q[i][0];
      H[i][1][1] = 1.0 - 2.0 * q[i][1];
    }
    // B (6 x 6), row-major, block (i, l) rows 2i..2i+1, columns 2l..2l+1
    double B[6][6];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int r = 0; r < 2; ++r)
#pragma unroll
        for (int l = 0; l < 3; ++l)
#pragma unroll
          for (int c = 0; c < 2; ++c) {
            double v = (i == l && r == c) ? 1.0 : 0.0;
#pragma unroll
            for (int j = 0; j < 3; ++j) v = v + (ha[i][j] * ha[j][l]) * H[j][r][c];
            B[2 * i + r][2 * l + c] = v;
          }
    int pv[6];
#pragma unroll
    for (int p = 0; p < 6; ++p) {
      int piv = p;
      double best = cl_fabs(B[p][p]);
#pragma unroll
      for (int r = p + 1; r < 6; ++r) {
        double v = cl_fabs(B[r][p]);
        if (v > best) {
          best = v;
          piv = r;
        }
      }
      pv[p] = piv;
#pragma unroll
      for (int r = p + 1; r < 6; ++r) {  // swap rows p and piv without a dynamic register index
        if (r == piv) {
#pragma unroll
          for (int c = 0; c < 6; ++c) {
            double t = B[p][c];
            B[p][c] = B[r][c];
            B[r][c] = t;
          }
        }
      }
#pragma unroll
      for (int r = p + 1; r < 6; ++r) {
        const double f = B[r][p] / B[p][p];
        B[r][p] = f;
#pragma unroll
        for (int c = p + 1; c < 6; ++c) B[r][c] = B[r][c] - f * B[p][c];
      }
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      // right-hand side column c: -h sum_j a_ij H_j E_q (c < 2), E_p (c >= 2)
      double v[6];
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int r = 0; r < 2; ++r) {
          if (c < 2) {
            double w = 0.0;
#pragma unroll
            for (int j = 0; j < 3; ++j) w = w + ha[i][j] * H[j][r][c];
            v[2 * i + r] = -w;
          } else {
            v[2 * i + r] = (r == c - 2) ? 1.0 : 0.0;
          }
        }
#pragma unroll
      for (int p = 0; p < 6; ++p)
#pragma unroll
        for (int r = p + 1; r < 6; ++r)
          if (r == pv[p]) {
            double t = v[p];
            v[p] = v[r];
            v[r] = t;
          }
#pragma unroll
      for (int p = 0; p < 6; ++p)
#pragma unroll
        for (int r = p + 1; r < 6; ++r) v[r] = v[r] - B[r][p] * v[p];
      double Qv[6];
#pragma unroll
      for (int r = 5; r >= 0; --r) {
        double w = v[r];
#pragma unroll
        for (int t = r + 1; t < 6; ++t) w = w - B[r][t] * Qv[t];
        Qv[r] = w / B[r][r];
      }
      // P_i = E_q + sum_j ha_ij Q_j; J = I + sum_i hb_i [Q_i; -H_i P_i]
      double Jc[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) Jc[r] = (r == c) ? 1.0 : 0.0;
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        double P[2];
#pragma unroll
        for (int r = 0; r < 2; ++r) {
          double w = (r == c) ? 1.0 : 0.0;
#pragma unroll
          for (int j = 0; j < 3; ++j) w = w + ha[i][j] * Qv[2 * j + r];
          P[r] = w;
        }
#pragma unroll
        for (int r = 0; r < 2; ++r) {
          Jc[r] = Jc[r] + hb[i] * Qv[2 * i + r];
          double hp = H[i][r][0] * P[0] + H[i][r][1] * P[1];
          Jc[2 + r] = Jc[2 + r] + hb[i] * (-hp);
        }
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) J[c * 4 + r] = Jc[r];
    }
  }

  // W <- J_step W with the fused_product order (src/Foundation/product.jl:9-20); a failed stage solve
  // makes the tangent all NaN, as the reference's `W * NaN` (gauss.jl:179).
  template <int K2>
  CL_HD void tangent_apply(const Aux& a, Frame<4, K2>& W) const {
    Frame<4, K2> O;
#pragma unroll
    for (int j = 0; j < K2; ++j)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        double acc = a.J[i] * W.at(0, j);
#pragma unroll
        for (int q = 1; q < 4; ++q) acc = cl_fma(a.J[q * 4 + i], W.at(q, j), acc);
        O.at(i, j) = acc;
      }
    W = O;
  }
  // the one-step tangent at the current point, column-major (NaN when the stage solve failed)
  CL_HD void jacobian(const Aux& a, double* J) const {
#pragma unroll
    for (int e = 0; e < 16; ++e) J[e] = a.J[e];
  }
  CL_HD void forward(const Aux& a, double* z) const {
#pragma unroll
    for (int c = 0; c < 4; ++c) z[c] = a.ok ? z[c] + a.inc[c] : __builtin_nan("");
  }
  CL_HD void field(const double* z, const Aux&, double* f) const { vf(z, f); }
  // energy H(z) (flow_models.jl:101): (p1^2 + p2^2)/2 + (q1^2 + q2^2)/2 + q1^2 q2 - q2^3/3
  static CL_HD double energy(const double* z) {
    return (((z[2] * z[2] + z[3] * z[3]) / 2.0 + (z[0] * z[0] + z[1] * z[1]) / 2.0) + (z[0] * z[0]) * z[1]) -
           (z[1] * z[1]) * z[1] / 3.0;
  }
};

// The Henon-Heiles Gauss map on a compensated state (value z, correction e): the reference's compensated summation
// (advance_n(...; compensated = true), gauss.jl:149-158, stepper.jl:50-58) carried through the tangent stream. The
// correction enters the stage values; the update is z, e = TwoSum(z, inc + e) (the _kahan rule), so the state is a
// double-double pair that drifts by no systematic rounding bias. The Jacobian is that of the compensated stages.
struct HHGaussComp : HHGauss {
  static constexpr int S = 8;  // z[0..3], e[4..7]
  CL_HD void aux(const double* x, Aux& a) const { aux_e(x, x + 4, a); }
  CL_HD void forward(const Aux& a, double* x) const {
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      if (a.ok) {
        double s, err;
        cl_two_sum(x[c], a.inc[c] + x[4 + c], &s, &err);
        x[c] = s;
        x[4 + c] = err;
      } else {
        x[c] = __builtin_nan("");
        x[4 + c] = __builtin_nan("");
      }
    }
  }
};

}  // namespace cl
