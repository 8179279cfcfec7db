// cl_core.hpp - per-orbit numerics of the ensemble tangent stream, compiled for gfx950 kernels.
//
// One orbit lives in one lane: the state x, the deviation frame(s) W (column-major d x k, held in
// VGPRs through compile-time sizes) and the reducer states. The arithmetic follows the reference
// operation for operation (reference paths relative to the Chaoslab repository root):
//   per-orbit body           src/Tangent/orbit.jl:44-60
//   fused frame product      src/Foundation/product.jl:9-20 (explicit fma, ascending q)
//   Neumaier sums, norms     src/Foundation/Foundation.jl:48-58, src/Foundation/norm.jl:8-12
//   standard map             src/Systems/standard_map.jl:18-21,30-34
//   wrapped standard map     src/Systems/wrapped_standard_map.jl:16-57
//   QR spectrum              src/Lyapunov/qr_spectrum.jl:54-73, orthonormalize.jl:38-151
//   Householder/Givens       src/Lyapunov/householder.jl:17-123
//   Benettin                 src/Lyapunov/benettin.jl:27-33
//   MEGNO / SALI / FLI       src/Indicators/megno.jl:63-80, sali.jl:166-179, fli.jl:320-341
//   OFLI                     src/Indicators/ofli.jl:59-102
//   GALI                     src/Indicators/gali.jl:47-169
// so that, with cl_math.h providing the same sin/cos/log bits on host and device, a kernel
// reproduces the host restatement of the same run bit for bit.
#pragma once
#include <stdint.h>
#include "cl_math.h"
#include "../../include/chaoslab_hip.h"

#if !defined(__HIPCC__)
#define __device__
#define __host__
#endif

#define CL_EPS64 0x1p-52

namespace cl {

enum : int8_t { STATUS_OK = 0, STATUS_STATE = 1, STATUS_W = 2, STATUS_REDUCER = 3 };

// ------------------------------------------------------------------------------ foundations
struct CSum {
  double s, c;
};

CL_HD CSum csum_zero() { return CSum{0.0, 0.0}; }

// Neumaier update (Foundation.jl:48-55). The reference picks (s - t) + y or (y - t) + s by the larger
// magnitude; either is the exact rounding error of s + y, as is Knuth's branch-free TwoSum error used
// here, so the compensation term is bit-identical while s + y does not overflow.
CL_HD CSum csum_add(CSum a, double y) {
  double t, e;
  cl_two_sum(a.s, y, &t, &e);
  return CSum{t, a.c + e};
}
CL_HD double csum_total(CSum a) { return a.s + a.c; }

// Scaled Euclidean norm of v[0..n) with stride `st` (norm.jl:8-12): maximum of |v|, then
// m * sqrt(sum((v/m)^2)) summed left to right. The maximum uses the hardware max: every caller passes
// entries the handoff has already checked to be finite (or vectors formed from them), and on finite and
// infinite absolute values the hardware max equals Julia's NaN-propagating `max` bit for bit.
template <int N>
CL_HD double safe_norm(const double* v, int st = 1) {
  double m = cl_fabs(v[0]);
#pragma unroll
  for (int i = 1; i < N; ++i) m = __builtin_fmax(m, cl_fabs(v[i * st]));
  if (N == 2) {
    // the larger entry over m is exactly +-1, its square exactly 1: one division instead of two, and
    // 1 + b^2 equals b^2 + 1, so the sum is the same in either order
    const double other = cl_fabs(v[0]) >= cl_fabs(v[st]) ? v[st] : v[0];
    if (m >= 0x1p-400 && m <= 0x1p400) {
      // |b| <= 1: the plain division is x / m wherever |b| >= 2^-400, and a smaller |b| (or a zero of either sign)
      // leaves 1 + b^2 = 1 either way
      const double b = cl_div_plain(other, m);
      return m * cl_sqrt12(1.0 + b * b);
    }
    if (!(cl_isfinite(m) && m > 0.0)) return m;
    const double b = other / m;
    return m * cl_sqrt12(1.0 + b * b);
  }
  if (!(cl_isfinite(m) && m > 0.0)) return m;
  double a = v[0] / m;
  double s = a * a;
#pragma unroll
  for (int i = 1; i < N; ++i) {
    double b = v[i * st] / m;
    s = s + b * b;
  }
  return m * cl_sqrt(s);
}

// Plain dot product in StaticArrays order: a1*b1 + a2*b2 + ... (no fusion).
template <int N>
CL_HD double dot(const double* a, const double* b) {
  double s = a[0] * b[0];
#pragma unroll
  for (int i = 1; i < N; ++i) s = s + a[i] * b[i];
  return s;
}

// ------------------------------------------------------------------------------ frames
// A d x k deviation frame. E is double, or float for the FP32-tangent element type (the frame is carried in Float32
// between handoffs and widened to double for the reducers, which orthonormalise and sum logs in FP64).
template <int D, int K, class E = double>
struct Frame {
  E w[D * K];  // column-major: w[col * D + row]
  CL_HD E& at(int r, int c) { return w[c * D + r]; }
  CL_HD E at(int r, int c) const { return w[c * D + r]; }
  CL_HD E* col(int c) { return w + c * D; }
  CL_HD const E* col(int c) const { return w + c * D; }
};

template <int D, int K, class E>
CL_HD bool frame_finite(const Frame<D, K, E>& W) {
  bool ok = true;
#pragma unroll
  for (int i = 0; i < D * K; ++i) ok = ok && cl_isfinite((double)W.w[i]);
  return ok;
}

// ------------------------------------------------------------------------------ systems
// W <- M W for M = [[1 + c, 1], [c, 1]], the one-step tangent of the standard-map family, in the
// fused_product order (product.jl:9-20): C[i,j] = M[i,1] W[1,j], then fma(M[i,2], W[2,j], acc).
template <int K2>
CL_HD void map_tangent_apply(double c, Frame<2, K2>& W) {
  double m11 = 1.0 + c;
#pragma unroll
  for (int j = 0; j < K2; ++j) {
    double w0 = W.at(0, j), w1 = W.at(1, j);
    W.at(0, j) = cl_fma(1.0, w1, m11 * w0);
    W.at(1, j) = cl_fma(1.0, w1, c * w0);
  }
}

// The same product on a Float32 frame: the tangent's entries are formed in FP64 from the FP64 orbit (1 + c and c,
// exactly as above) and rounded once to Float32; the product is the explicit-fma one in Float32 (product.jl:9-20 is
// defined for Float32 as for Float64).
template <int K2>
CL_HD void map_tangent_apply(double c, Frame<2, K2, float>& W) {
  const float m11 = (float)(1.0 + c), m21 = (float)c;
#pragma unroll
  for (int j = 0; j < K2; ++j) {
    float w0 = W.at(0, j), w1 = W.at(1, j);
    W.at(0, j) = __builtin_fmaf(1.0f, w1, m11 * w0);
    W.at(1, j) = __builtin_fmaf(1.0f, w1, m21 * w0);
  }
}

// The one-step tangent [[1 + c, 1], [c, 1]] of the standard-map family, column-major.
CL_HD void map_jacobian(double c, double* J) {
  J[0] = 1.0 + c;
  J[1] = c;
  J[2] = 1.0;
  J[3] = 1.0;
}

// Standard map p' = p + K sin x, x' = x + p', Jacobian [[1 + c, 1], [c, 1]], c = K cos x.
// The kernel carries sin and cos of the current point, so one reduction serves the Jacobian,
// the step and a map displacement field read at the handoff.
struct StdMap {
  static constexpr int D = 2;
  static constexpr int S = 2;
  double K;
  struct Aux {
    double sn, cs;
  };
  CL_HD void aux(const double* x, Aux& a) const { cl_sincos(x[0], &a.sn, &a.cs); }
  // the one parameter c = K cos x of the tangent at the current point
  CL_HD double tan_c(const Aux& a) const { return K * a.cs; }
  // W <- M W with fused_product order: C[i,j] = M[i,1] W[1,j], then fma(M[i,2], W[2,j], acc).
  template <int K2, class E>
  CL_HD void tangent_apply(const Aux& a, Frame<2, K2, E>& W) const { map_tangent_apply(tan_c(a), W); }
  // the one-step tangent at the current point, column-major (the ResidualMonitor reads it)
  CL_HD void jacobian(const Aux& a, double* J) const { map_jacobian(tan_c(a), J); }
  CL_HD void forward(const Aux& a, double* x) const {
    double p = x[1] + K * a.sn;
    x[0] = x[0] + p;
    x[1] = p;
  }
  // Displacement field F(x) - x = (p + K sin x, K sin x) at the current point (the direction
  // the OFLI reducer projects out on a map).
  CL_HD void field(const double* x, const Aux& a, double* f) const {
    double ks = K * a.sn;
    f[0] = x[1] + ks;
    f[1] = ks;
  }
};

// Standard map whose kick strength rides in the state (x, p, K): an ensemble over K in one launch.
// The reference holds one K per system (src/Systems/standard_map.jl:8-10); each orbit here steps exactly
// as StandardMap(K) with its own K, and K passes through the step unchanged.
struct StdMapK {
  static constexpr int D = 2;
  static constexpr int S = 3;
  struct Aux {
    double sn, cs, K;
  };
  CL_HD void aux(const double* x, Aux& a) const {
    cl_sincos(x[0], &a.sn, &a.cs);
    a.K = x[2];
  }
  CL_HD double tan_c(const Aux& a) const { return a.K * a.cs; }
  template <int K2, class E>
  CL_HD void tangent_apply(const Aux& a, Frame<2, K2, E>& W) const { map_tangent_apply(tan_c(a), W); }
  // the one-step tangent at the current point, column-major (the ResidualMonitor reads it)
  CL_HD void jacobian(const Aux& a, double* J) const { map_jacobian(tan_c(a), J); }
  CL_HD void forward(const Aux& a, double* x) const {
    double p = x[1] + a.K * a.sn;
    x[0] = x[0] + p;
    x[1] = p;
  }
  CL_HD void field(const double* x, const Aux& a, double* f) const {
    double ks = a.K * a.sn;
    f[0] = x[1] + ks;
    f[1] = ks;
  }
};

// Wrapped standard map on the torus with exact winding counts (state x, p, m, n).
struct WrappedStdMap {
  static constexpr int D = 2;
  static constexpr int S = 4;
  double K;
  struct Aux {
    double sn, cs;
  };
  CL_HD void aux(const double* x, Aux& a) const { cl_sincos(x[0], &a.sn, &a.cs); }
  CL_HD double tan_c(const Aux& a) const { return K * a.cs; }
  template <int K2, class E>
  CL_HD void tangent_apply(const Aux& a, Frame<2, K2, E>& W) const { map_tangent_apply(tan_c(a), W); }
  // the one-step tangent at the current point, column-major (the ResidualMonitor reads it)
  CL_HD void jacobian(const Aux& a, double* J) const { map_jacobian(tan_c(a), J); }
  static CL_HD void reduce(double u, double tp, double* r, double* a) {
    double q = cl_floor(u / tp);
    double rr = u - tp * q;
    if (rr < 0.0) {
      rr += tp;
      q -= 1.0;
    }
    if (rr >= tp) {
      rr -= tp;
      q += 1.0;
    }
    *r = rr;
    *a = q;
  }
  CL_HD void forward(const Aux& au, double* z) const {
    // 2 * (pi_hi + pi_lo) in Float64 is 2 pi rounded (wrapped_standard_map.jl:40, standard_map.jl:48)
    const double tp = 2.0 * (3.141592653589793 + 1.2246467991473532e-16);
    double p, a;
    reduce(z[1] + K * au.sn, tp, &p, &a);
    double n = z[3] + a;
    double x, b;
    reduce(z[0] + p, tp, &x, &b);
    double m = z[2] + n + b;
    const double lim = 9007199254740992.0;  // maxintfloat(Float64)
    if (cl_fabs(m) < lim && cl_fabs(n) < lim) {
      z[0] = x; z[1] = p; z[2] = m; z[3] = n;
    } else {
      double nan = __builtin_nan("");
      z[0] = nan; z[1] = nan; z[2] = nan; z[3] = nan;
    }
  }
  CL_HD void field(const double* x, const Aux& a, double* f) const {
    double ks = K * a.sn;
    f[0] = x[1] + ks;
    f[1] = ks;
  }
};

// The product of a map with itself (src/Systems/pair.jl:1-68): the state is (state A; state B), each half steps by the
// base map alone, and the tangent acts block by block on the frame's rows 1..d and d+1..2d (the zero blocks are never
// multiplied, so each block's product is the single-orbit one). The monitor reads orbit A's one-step tangent
// (run_indicator_record.jl:6-10).
template <class B>
struct PairMap {
  static constexpr int D = 2 * B::D;
  static constexpr int S = 2 * B::S;
  B base;
  struct Aux {
    typename B::Aux a, b;
  };
  CL_HD void aux(const double* x, Aux& au) const {
    base.aux(x, au.a);
    base.aux(x + B::S, au.b);
  }
  template <int K2, class E>
  CL_HD void tangent_apply(const Aux& au, Frame<D, K2, E>& W) const {
    Frame<B::D, K2, E> wa, wb;
#pragma unroll
    for (int j = 0; j < K2; ++j)
#pragma unroll
      for (int r = 0; r < B::D; ++r) {
        wa.at(r, j) = W.at(r, j);
        wb.at(r, j) = W.at(B::D + r, j);
      }
    base.tangent_apply(au.a, wa);
    base.tangent_apply(au.b, wb);
#pragma unroll
    for (int j = 0; j < K2; ++j)
#pragma unroll
      for (int r = 0; r < B::D; ++r) {
        W.at(r, j) = wa.at(r, j);
        W.at(B::D + r, j) = wb.at(r, j);
      }
  }
  CL_HD void jacobian(const Aux& au, double* J) const { base.jacobian(au.a, J); }
  CL_HD void forward(const Aux& au, double* z) const {
    base.forward(au.a, z);
    base.forward(au.b, z + B::S);
  }
  CL_HD void field(const double* x, const Aux& au, double* f) const { base.field(x, au.a, f); }
};

// ------------------------------------------------------------------------------ reducer states
// Byte layouts of the reference's isbits states (Julia lays struct fields out with C alignment),
// so a host can rebuild them and call the reference's reducer_final unchanged.
template <int K>
struct QRState {
  CSum sums[K];
  int64_t n;
  double G_max, orth_max;
  int64_t count_max;
  int8_t code;
  int8_t pad_[7];
  int64_t column;
  double value, threshold;
  int64_t step;
};

struct BenettinState {
  CSum sum;
  int64_t n;
  int8_t code;
  int8_t pad_[7];
  double value;
  int64_t step;
};

struct MegnoSums {
  CSum s, R1, R2, Q, Q2;
};
struct MEGNOState {
  MegnoSums c1, c2;
  double value;
  int64_t n, K, step;
  int8_t code, column;
  int8_t pad_[6];
};

struct RLIState {
  CSum sA, sB, Q;
  double value;
  int64_t n, K, step;
  int8_t code, column;
  int8_t pad_[6];
};

struct SALIState {
  CSum s1, s2;
  double sali, value;
  int64_t n, step;
  int8_t code, column;
  int8_t pad_[6];
};

struct FLIState {
  CSum s1, s2;
  double f1, fB, value;
  int64_t n, step, step1, stepB;
  int8_t code, column, colB;
  int8_t pad_[5];
};

template <int K>
struct OFLIState {
  CSum acc[K];
  double g1, gB, inst, f1, fB, sin1, sinB, sinmin, value;
  int64_t nminf, n, step, step1, stepB;
  int8_t colB, code, column;
  int8_t pad_[5];
};

template <int K>
struct GALIState {
  double s[K];
  double c[K];
  double gali[K - 1];
  double value;
  int64_t n, step;
  int8_t code;
  int8_t pad_[7];
  int64_t column;
};

struct MonitorState {
  double det_max, symp_max, orth_max, sum_max, pair_max, G_max, M_max;
  int8_t code;
  int8_t pad_[7];
  double value, threshold;
  int64_t step;
};

// ------------------------------------------------------------------------------ QR family
enum QRAlg : int { QR_CGS = 1, QR_MGS = 2, QR_RGS = 3, QR_HOUSEHOLDER = 4, QR_AUGMENTED = 5, QR_GIVENS = 6 };
enum : int8_t { QR_OK = 0, QR_PIVOT = 1, QR_PASSES = 2 };
constexpr int MAX_PASSES = 4;

template <int D, int K>
struct QRFact {
  Frame<D, K> Q;
  double R[K * K];  // column-major
  int count;
  int8_t status;
  int column;  // 1-based failing column, 0 if none
};

// Gram-Schmidt family (orthonormalize.jl:38-105).
template <int ALG, int D, int K>
CL_HD void orthonormalize_gs(const Frame<D, K>& W, QRFact<D, K>& f) {
#pragma unroll
  for (int i = 0; i < D * K; ++i) f.Q.w[i] = 0.0;
#pragma unroll
  for (int i = 0; i < K * K; ++i) f.R[i] = 0.0;
  f.count = 1;
  f.status = QR_OK;
  f.column = 0;
  bool done = false;
#pragma unroll
  for (int m = 0; m < K; ++m) {
    if (!done) {
      double u[D], r[K];
#pragma unroll
      for (int i = 0; i < D; ++i) u[i] = W.at(i, m);
#pragma unroll
      for (int i = 0; i < K; ++i) r[i] = 0.0;
      int passes = 1;
      if (ALG == QR_CGS || ALG == QR_MGS) {
        double w0[D];
#pragma unroll
        for (int i = 0; i < D; ++i) w0[i] = u[i];
#pragma unroll
        for (int i = 0; i < m; ++i) {
          const double* qi = f.Q.col(i);
          double c = (ALG == QR_CGS) ? dot<D>(qi, w0) : dot<D>(qi, u);
          r[i] = c;
#pragma unroll
          for (int l = 0; l < D; ++l) u[l] = u[l] - c * qi[l];
        }
      } else {  // RGS, eta = 1/sqrt(2), at most MAX_PASSES passes
        const double eta = 1.0 / cl_sqrt(2.0);
        double nold = safe_norm<D>(u);
        passes = 0;
        for (;;) {
          passes += 1;
          double vold[D];
#pragma unroll
          for (int i = 0; i < D; ++i) vold[i] = u[i];
#pragma unroll
          for (int i = 0; i < m; ++i) {
            const double* qi = f.Q.col(i);
            double c = dot<D>(qi, vold);
            r[i] = r[i] + c;
#pragma unroll
            for (int l = 0; l < D; ++l) u[l] = u[l] - c * qi[l];
          }
          double nnew = safe_norm<D>(u);
          if (nnew >= nold * eta || !(cl_isfinite(nnew) && nnew > 0.0)) break;
          if (passes == MAX_PASSES) {
            passes = MAX_PASSES + 1;
            break;
          }
          nold = nnew;
        }
      }
      double gamma = safe_norm<D>(u);
#pragma unroll
      for (int i = 0; i < K; ++i) f.R[m * K + i] = (i == m) ? gamma : r[i];
      if (passes > MAX_PASSES) {
        f.status = QR_PASSES;
        f.column = m + 1;
      } else if (!(cl_isfinite(gamma) && gamma > 0.0)) {
        f.status = QR_PIVOT;
        f.column = m + 1;
      }
      if (f.status != QR_OK) {
        done = true;
      } else {
#pragma unroll
        for (int l = 0; l < D; ++l) f.Q.at(l, m) = u[l] / gamma;
        f.count = f.count > passes ? f.count : passes;
      }
    }
  }
}

// (I - 2 u u') A for A with NC columns of D rows (householder.jl:30-33).
template <int D, int NC>
CL_HD void hh_reflect(double* A, const double* u) {
#pragma unroll
  for (int l = 0; l < NC; ++l) {
    double w = dot<D>(u, A + l * D);
#pragma unroll
    for (int i = 0; i < D; ++i) A[l * D + i] = A[l * D + i] - (2.0 * u[i]) * w;
  }
}

// Reflector sweep (householder.jl:20-48) over the first NREF columns of A (D x NC).
template <int D, int NC, int NREF>
CL_HD int hh_sweep(double* A, double* U) {
  int count = 0;
#pragma unroll
  for (int j = 0; j < NREF; ++j) {
    double a[D];
#pragma unroll
    for (int i = 0; i < D; ++i) a[i] = (i >= j) ? A[j * D + i] : 0.0;
    double c = safe_norm<D>(a);
    double rho;
    if (cl_isfinite(c) && c > 0.0) {
      double s = a[j] < 0.0 ? -1.0 : 1.0;
      a[j] = a[j] + s * c;
      double nv = safe_norm<D>(a);
#pragma unroll
      for (int i = 0; i < D; ++i) a[i] = a[i] / nv;
      rho = -s * c;
      hh_reflect<D, NC>(A, a);
#pragma unroll
      for (int i = 0; i < D; ++i) U[j * D + i] = a[i];
      count += 1;
    } else {
      rho = c;
    }
    A[j * D + j] = rho;
  }
  return count;
}

// Sign fix and status (householder.jl:66-75); A has leading dimension LD.
template <int D, int K, int LD>
CL_HD void hh_finish(QRFact<D, K>& f, const double* A) {
  double sg[K];
#pragma unroll
  for (int j = 0; j < K; ++j) sg[j] = A[j * LD + j] < 0.0 ? -1.0 : 1.0;
#pragma unroll
  for (int j = 0; j < K; ++j)
#pragma unroll
    for (int i = 0; i < K; ++i) f.R[j * K + i] = (i > j) ? 0.0 : sg[i] * A[j * LD + i];
#pragma unroll
  for (int j = 0; j < K; ++j)
#pragma unroll
    for (int i = 0; i < D; ++i) f.Q.at(i, j) = f.Q.at(i, j) * sg[j];
  f.status = QR_OK;
  f.column = 0;
#pragma unroll
  for (int j = 0; j < K; ++j) {
    double p = f.R[j * K + j];
    if (f.status == QR_OK && !(cl_isfinite(p) && p > 0.0)) {
      f.status = QR_PIVOT;
      f.column = j + 1;
    }
  }
}

template <int D, int K>
CL_HD void orthonormalize_householder(const Frame<D, K>& W, QRFact<D, K>& f) {
  constexpr int NREF = K < D ? K : K - 1;
  double A[D * K], U[D * K];
#pragma unroll
  for (int i = 0; i < D * K; ++i) {
    A[i] = W.w[i];
    U[i] = 0.0;
  }
  f.count = hh_sweep<D, K, NREF>(A, U);
  // Q = H_1 ... H_r applied to I[:, 1:K], reflectors in reverse order
#pragma unroll
  for (int j = 0; j < K; ++j)
#pragma unroll
    for (int i = 0; i < D; ++i) f.Q.at(i, j) = (i == j) ? 1.0 : 0.0;
#pragma unroll
  for (int j = NREF - 1; j >= 0; --j) hh_reflect<D, K>(f.Q.w, U + j * D);
  hh_finish<D, K, D>(f, A);
}

template <int D, int K>
CL_HD void orthonormalize_augmented(const Frame<D, K>& W, QRFact<D, K>& f) {
  constexpr int NREF = K < D ? K : K - 1;
  constexpr int NC = K + D;
  double A[D * NC], U[D * K];
#pragma unroll
  for (int l = 0; l < D * NC; ++l) {
    int c = l / D, r = l % D;
    A[l] = (l < D * K) ? W.w[l] : ((r == c - K) ? 1.0 : 0.0);
  }
#pragma unroll
  for (int i = 0; i < D * K; ++i) U[i] = 0.0;
  f.count = hh_sweep<D, NC, NREF>(A, U);
#pragma unroll
  for (int b = 0; b < K; ++b)
#pragma unroll
    for (int a = 0; a < D; ++a) f.Q.at(a, b) = A[(K + a) * D + b];
  hh_finish<D, K, D>(f, A);
}

// Givens rotations on the rows of [W, I_d] (householder.jl:94-123).
template <int D, int K>
CL_HD void orthonormalize_givens(const Frame<D, K>& W, QRFact<D, K>& f) {
  constexpr int WD = K + D;
  constexpr int J = K < D - 1 ? K : D - 1;
  double rows[D][WD];
#pragma unroll
  for (int i = 0; i < D; ++i)
#pragma unroll
    for (int l = 0; l < WD; ++l) rows[i][l] = (l < K) ? W.at(i, l) : ((l - K == i) ? 1.0 : 0.0);
  int count = 0;
#pragma unroll
  for (int j = 0; j < J; ++j) {
#pragma unroll
    for (int i = D - 1; i >= j + 1; --i) {
      double ab[2] = {rows[i - 1][j], rows[i][j]};
      double rho = safe_norm<2>(ab);
      if (cl_isfinite(rho) && rho > 0.0) {
        double c = ab[0] / rho, s = ab[1] / rho;
#pragma unroll
        for (int l = 0; l < WD; ++l) {
          double ra = rows[i - 1][l], rb = rows[i][l];
          rows[i - 1][l] = c * ra + s * rb;
          rows[i][l] = c * rb - s * ra;
        }
        rows[i - 1][j] = rho;
        rows[i][j] = 0.0;
        count += 1;
      }
    }
  }
  double R0[K * K];
#pragma unroll
  for (int j = 0; j < K; ++j)
#pragma unroll
    for (int i = 0; i < K; ++i) R0[j * K + i] = (i > j) ? 0.0 : rows[i][j];
#pragma unroll
  for (int b = 0; b < K; ++b)
#pragma unroll
    for (int a = 0; a < D; ++a) f.Q.at(a, b) = rows[b][K + a];
  f.count = count;
  hh_finish<D, K, K>(f, R0);
}

template <int ALG, int D, int K>
CL_HD void orthonormalize(const Frame<D, K>& W, QRFact<D, K>& f) {
  if (ALG == QR_HOUSEHOLDER) orthonormalize_householder<D, K>(W, f);
  else if (ALG == QR_AUGMENTED) orthonormalize_augmented<D, K>(W, f);
  else if (ALG == QR_GIVENS) orthonormalize_givens<D, K>(W, f);
  else orthonormalize_gs<ALG, D, K>(W, f);
}

// G_w = ||R||_F ||R^-1||_F by back substitution, rescaled when far from unit (orthonormalize.jl:112-127).
template <int K>
CL_HD double window_Gw(const double* R0) {
  double R[K * K];
  double m = cl_fabs(R0[0]);
#pragma unroll
  for (int i = 1; i < K * K; ++i) m = cl_jmax(m, cl_fabs(R0[i]));
  bool scale = (m > 1e100 || m < 1e-100) && cl_isfinite(m) && m > 0.0;
#pragma unroll
  for (int i = 0; i < K * K; ++i) R[i] = R0[i];
  if (scale) {
    // rare: the empty volatile asm keeps the compiler from computing the K^2 divisions for every lane and
    // selecting afterwards
    __asm__ volatile("");
#pragma unroll
    for (int i = 0; i < K * K; ++i) R[i] = R0[i] / m;
  }
  double Ri[K * K];
#pragma unroll
  for (int i = 0; i < K * K; ++i) Ri[i] = 0.0;
#pragma unroll
  for (int j = 0; j < K; ++j) {
    Ri[j * K + j] = 1.0 / R[j * K + j];
#pragma unroll
    for (int i = j - 1; i >= 0; --i) {
      double s = 0.0;
#pragma unroll
      for (int l = i + 1; l <= j; ++l) s += R[l * K + i] * Ri[j * K + l];
      Ri[j * K + i] = -s / R[i * K + i];
    }
  }
  double a = R[0] * R[0], b = Ri[0] * Ri[0];
#pragma unroll
  for (int i = 1; i < K * K; ++i) {
    a = a + R[i] * R[i];
    b = b + Ri[i] * Ri[i];
  }
  return cl_sqrt(a) * cl_sqrt(b);
}

// A one-division estimate of window_Gw<2> for a 2 x 2 upper-triangular R: ||R^-1||_F = ||R||_F / |R11 R22|, so
// G_w = ||R||_F^2 / |R11 R22| in exact arithmetic. With the entries in [1e-100, 1e100] (no rescaling, no
// overflow and no underflow that moves either result by more than a unit in the last place) window_Gw's
// rounding errors stay below 10 units in the last place and this one's below 5, so the two differ by less than
// GW_QUICK_REL times the estimate. Returns -1 outside that range (and for non-finite entries).
constexpr double GW_QUICK_REL = 1e-13;
// an estimate below this settles the conditioning check (window_Gw * eps < 0.5, i.e. G_w < 2^51 = 2.25e15)
constexpr double GW_QUICK_GATE = 2.0e15;
CL_HD double gw_quick2(const double* R) {
  const double r0 = cl_fabs(R[0]), r2 = cl_fabs(R[2]), r3 = cl_fabs(R[3]);
  const bool ok = r0 >= 1e-100 && r0 <= 1e100 && r3 >= 1e-100 && r3 <= 1e100 && r2 <= 1e100 && R[1] == 0.0;
  const double a = (R[0] * R[0] + R[2] * R[2]) + R[3] * R[3];
  return ok ? a / (r0 * r3) : -1.0;
}

// G_from_Gw is non-decreasing in G_w >= 0 (every operation is, and rounding keeps order), so the largest G over
// a run of handoffs is G of the largest G_w, bit for bit.
template <int K>
CL_HD double G_from_Gw(double Gw) {
  if (K == 1) return 1.0;
  if (K == 2) return (Gw + cl_sqrt(cl_jmax(Gw * Gw - 4.0, 0.0))) / 2.0;
  return Gw;
}

// max |Q'Q - I| with each entry formed as Q[1,i] Q[1,j] then fma down the rows.
template <int D, int K>
CL_HD double orth_loss(const Frame<D, K>& Q) {
  double mx = 0.0;
  bool first = true;
#pragma unroll
  for (int j = 0; j < K; ++j)
#pragma unroll
    for (int i = 0; i < K; ++i) {
      double acc = Q.at(0, i) * Q.at(0, j);
#pragma unroll
      for (int l = 1; l < D; ++l) acc = cl_fma(Q.at(l, i), Q.at(l, j), acc);
      double e = cl_fabs(i == j ? acc - 1.0 : acc - 0.0);
      mx = first ? e : cl_max_nonneg(mx, e);  // absolute values: cl_jmax's result
      first = false;
    }
  return mx;
}

template <int ALG, int D, int K>
struct QRSpectrum {
  double tol;  // orthogonality_tol
  typedef QRState<K> State;
  static constexpr bool checks_conditioning = (ALG == QR_CGS || ALG == QR_MGS || ALG == QR_RGS);
  CL_HD void init(State& st) const {
#pragma unroll
    for (int i = 0; i < K; ++i) st.sums[i] = csum_zero();
    st.n = 0;
    st.G_max = 0.0;
    st.orth_max = 0.0;
    st.count_max = 0;
    st.code = 0;
#pragma unroll
    for (int i = 0; i < 7; ++i) st.pad_[i] = 0;
    st.column = 0;
    st.value = 0.0;
    st.threshold = 0.0;
    st.step = 0;
  }
  CL_HD static void fail(State& st, int code, int64_t column, double value, double threshold, int64_t n) {
    st.code = (int8_t)code;
    st.column = column;
    st.value = value;
    st.threshold = threshold;
    st.step = n;
  }
  // At k = 2, true when the one-division estimate of G_w (gw_quick2, within GW_QUICK_REL of window_Gw) settles
  // the handoff without window_Gw: the conditioning check passes (estimate below GW_QUICK_GATE), orth_loss is
  // below its G_w-scaled threshold by more than the estimate's error (CGS, MGS; the other algorithms' threshold
  // does not use G_w), and G of this handoff cannot exceed G_max: in exact arithmetic G(G_w) <= G_max exactly
  // when G_w <= G_max + 1/G_max, tested as G_w G_max < G_max^2 + 1 with a 2e-12 margin (G_max in [2, 1e150], so
  // G_max^2 stays finite). G is concave with slope G / sqrt(G_w^2 - 4), so that margin keeps G(G_w) at least
  // 2e-12 G_max below G_max; G's rounding is below 5 units in the last place from G_w = 2.5 on, and below G_w
  // it cannot reach G_max >= 2. The running maximum then keeps its value, as jmax would.
  CL_HD bool settled_by_estimate(const State& st, const QRFact<D, K>& f, double orth) const {
    if constexpr (K != 2) {
      return false;
    } else {
      const double q = gw_quick2(f.R);
      const double Gm = st.G_max;
      bool ok = q >= 0.0 && Gm >= 2.0 && Gm < 1e150 && q * Gm * (1.0 + 2e-12) < Gm * Gm + 1.0;
      if (checks_conditioning) ok = ok && q < GW_QUICK_GATE;
      if (ALG == QR_CGS || ALG == QR_MGS) ok = ok && orth <= tol * CL_EPS64 * 2.0 * q * (1.0 - 2.0 * GW_QUICK_REL);
      return ok;
    }
  }
  CL_HD void step(State& st, int64_t n, Frame<D, K>& W) const {
    if (st.code != 0) return;
    QRFact<D, K> f;
    orthonormalize<ALG, D, K>(W, f);
    if (f.status != QR_OK) {
      if (f.status == QR_PASSES) {
        fail(st, 2, f.column, (double)(MAX_PASSES + 1), (double)MAX_PASSES, n);
      } else {
        double pivot = 0.0;  // R[column, column], selected without a dynamic register index
#pragma unroll
        for (int j = 0; j < K; ++j)
          if (j + 1 == f.column) pivot = f.R[j * K + j];
        fail(st, 1, f.column, pivot, cl_isfinite(pivot) ? 0.0 : __builtin_inf(), n);
      }
      return;
    }
    constexpr bool GS_THR = ALG == QR_CGS || ALG == QR_MGS;  // orth_loss's threshold scales with G_w
    const double orth = K >= 2 ? orth_loss<D, K>(f.Q) : 0.0;
    if (!settled_by_estimate(st, f, orth)) {
      double Gw = window_Gw<K>(f.R);
      if (checks_conditioning && !(Gw * CL_EPS64 < 0.5)) {
        fail(st, 3, 0, Gw * CL_EPS64, 0.5, n);
        return;
      }
      double g = GS_THR ? Gw : 1.0;
      double thr = tol * CL_EPS64 * (double)K * g;
      if (K >= 2 && !(orth <= thr)) {
        fail(st, 4, 0, orth, thr, n);
        return;
      }
      st.G_max = cl_jmax(st.G_max, G_from_Gw<K>(Gw));
    } else if (!GS_THR) {
      const double thr = tol * CL_EPS64 * (double)K * 1.0;  // the threshold without G_w
      if (!(orth <= thr)) {
        fail(st, 4, 0, orth, thr, n);
        return;
      }
    }
#pragma unroll
    for (int i = 0; i < K; ++i) st.sums[i] = csum_add(st.sums[i], cl_log(f.R[i * K + i]));
    st.n = n;
    st.orth_max = cl_max_nonneg(st.orth_max, orth);  // both >= +0 or NaN: cl_jmax's result
    st.count_max = st.count_max > f.count ? st.count_max : f.count;
    st.column = 0;
    W = f.Q;
  }
};

// ------------------------------------------------------------------------------ Benettin (k = 1)
template <int D>
struct Benettin {
  typedef BenettinState State;
  CL_HD void init(State& st) const {
    st.sum = csum_zero();
    st.n = 0;
    st.code = 0;
#pragma unroll
    for (int i = 0; i < 7; ++i) st.pad_[i] = 0;
    st.value = 0.0;
    st.step = 0;
  }
  CL_HD void step(State& st, int64_t n, Frame<D, 1>& W) const {
    if (st.code != 0) return;
    double alpha = safe_norm<D>(W.w);
    if (!(cl_isfinite(alpha) && alpha > 0.0)) {
      st.code = 1;
      st.value = alpha;
      st.step = n;
      return;
    }
    st.sum = csum_add(st.sum, cl_log(alpha));
    st.n = n;
#pragma unroll
    for (int i = 0; i < D; ++i) W.w[i] = W.w[i] / alpha;
  }
};

// Column norms and logs shared by MEGNO, SALI and FLI at one handoff: the three reducers perform
// the same operations on the same arriving columns, so their log sums are bitwise equal.
struct ColLogs {
  double a1, a2, l1, l2;
  bool ok1, ok2;
};
template <int D, int K>
CL_HD ColLogs col_logs(const Frame<D, K>& W) {
  ColLogs c;
  c.a1 = safe_norm<D>(W.col(0));
  c.a2 = safe_norm<D>(W.col(1));
  c.ok1 = cl_isfinite(c.a1) && c.a1 > 0.0;
  c.ok2 = cl_isfinite(c.a2) && c.a2 > 0.0;
  c.l1 = 0.0;
  c.l2 = 0.0;
  return c;
}

// ------------------------------------------------------------------------------ MEGNO (k = 2)
CL_HD MegnoSums megno_zero() {
  MegnoSums m;
  m.s = csum_zero(); m.R1 = csum_zero(); m.R2 = csum_zero(); m.Q = csum_zero(); m.Q2 = csum_zero();
  return m;
}
CL_HD MegnoSums megno_add(MegnoSums m, double a, int64_t n) {
  double tn = (double)n;
  CSum R1 = csum_add(m.R1, tn * a);
  CSum R2 = csum_add(m.R2, tn * tn * a);
  MegnoSums o;
  o.s = csum_add(m.s, a);
  o.R1 = R1;
  o.R2 = R2;
  // 2 R1 / n by the plain division: every term n a added to R1 is 0 or a multiple of 2^-105 of magnitude at least
  // 2^-53 (a is a log: never -0 and never below 2^-53 in magnitude unless 0), so R1 is +0 or at least 2^-105 in
  // magnitude and the quotient lies in the division's exact range
  o.Q = csum_add(m.Q, cl_div_plain(2.0 * csum_total(R1), tn));
  o.Q2 = csum_add(m.Q2, 3.0 * csum_total(R2));
  return o;
}

template <int D>
struct MEGNO {
  bool replace;
  typedef MEGNOState State;
  CL_HD void init(State& st) const {
    st.c1 = megno_zero();
    st.c2 = megno_zero();
    st.value = 0.0;
    st.n = 0; st.K = 0; st.step = 0;
    st.code = 0; st.column = 0;
#pragma unroll
    for (int i = 0; i < 6; ++i) st.pad_[i] = 0;
  }
  // Reads columns 1 and 2 of the frame (k = 2 in the reference; with k > 2 it only reads them).
  template <int K>
  CL_HD void step_logs(State& st, int64_t n, Frame<D, K>& W, ColLogs& c) const {
    if (st.code != 0) return;
    if (!(c.ok1 && c.ok2)) {
      st.column = c.ok1 ? 2 : 1;
      st.value = c.ok1 ? c.a2 : c.a1;
      st.step = n;
      st.code = 1;
      return;
    }
    st.c1 = megno_add(st.c1, c.l1, n);
    st.c2 = megno_add(st.c2, c.l2, n);
    st.n = n;
    st.K = st.K + 1;
    if (replace) {
#pragma unroll
      for (int i = 0; i < D; ++i) {
        W.at(i, 0) = W.at(i, 0) / c.a1;
        W.at(i, 1) = W.at(i, 1) / c.a2;
      }
    }
  }
};

// ------------------------------------------------------------------------------ SALI (k = 2)
template <int D>
struct SALI {
  typedef SALIState State;
  CL_HD void init(State& st) const {
    st.s1 = csum_zero(); st.s2 = csum_zero();
    st.sali = 0.0; st.value = 0.0;
    st.n = 0; st.step = 0;
    st.code = 0; st.column = 0;
#pragma unroll
    for (int i = 0; i < 6; ++i) st.pad_[i] = 0;
  }
  CL_HD void step_logs(State& st, int64_t n, Frame<D, 2>& W, ColLogs& c) const {
    if (st.code != 0) return;
    if (!(c.ok1 && c.ok2)) {
      st.column = c.ok1 ? 2 : 1;
      st.value = c.ok1 ? c.a2 : c.a1;
      st.step = n;
      st.code = 1;
      return;
    }
    double u1[D], u2[D], dm[D], dp[D];
#pragma unroll
    for (int i = 0; i < D; ++i) {
      u1[i] = W.at(i, 0) / c.a1;
      u2[i] = W.at(i, 1) / c.a2;
      dm[i] = u1[i] - u2[i];
      dp[i] = u1[i] + u2[i];
    }
    double d_minus = safe_norm<D>(dm), d_plus = safe_norm<D>(dp);
    st.s1 = csum_add(st.s1, c.l1);
    st.s2 = csum_add(st.s2, c.l2);
    st.sali = cl_min_num(d_minus, d_plus);  // two norms: non-negative, never NaN or -0
    st.n = n;
#pragma unroll
    for (int i = 0; i < D; ++i) {
      W.at(i, 0) = u1[i];
      W.at(i, 1) = u2[i];
    }
  }
};

// ------------------------------------------------------------------------------ RLI (k = 2, pair systems)
// rli.jl:77-110: the two arriving columns' norms and logs (the shared ColLogs, the same operations as Benettin's),
// the off-block check (both blocks' cross entries exactly 0), DL = |(sB.s - sA.s) + (sB.c - sA.c)| / n added to Q.
template <int D>
struct RLI {
  bool replace;
  typedef RLIState State;
  CL_HD void init(State& st) const {
    st.sA = csum_zero(); st.sB = csum_zero(); st.Q = csum_zero();
    st.value = 0.0;
    st.n = 0; st.K = 0; st.step = 0;
    st.code = 0; st.column = 0;
#pragma unroll
    for (int i = 0; i < 6; ++i) st.pad_[i] = 0;
  }
  CL_HD void step_logs(State& st, int64_t n, Frame<D, 2>& W, ColLogs& c) const {
    if (st.code != 0) return;
    if (!(c.ok1 && c.ok2)) {
      st.column = c.ok1 ? 2 : 1;
      st.value = c.ok1 ? c.a2 : c.a1;
      st.step = n;
      st.code = 1;
      return;
    }
    constexpr int h = D / 2;
    double m1 = 0.0, m2 = 0.0;  // the frame is finite here, so the hardware max is Julia's
#pragma unroll
    for (int i = 0; i < h; ++i) {
      m1 = __builtin_fmax(m1, cl_fabs(W.at(h + i, 0)));
      m2 = __builtin_fmax(m2, cl_fabs(W.at(i, 1)));
    }
    if (!(m1 == 0.0 && m2 == 0.0)) {
      st.column = m1 == 0.0 ? 2 : 1;
      st.value = m1 == 0.0 ? m2 : m1;
      st.step = n;
      st.code = 2;
      return;
    }
    st.sA = csum_add(st.sA, c.l1);
    st.sB = csum_add(st.sB, c.l2);
    const double dl = cl_fabs((st.sB.s - st.sA.s) + (st.sB.c - st.sA.c)) / (double)n;
    st.Q = csum_add(st.Q, dl);
    st.n = n;
    st.K = st.K + 1;
    if (replace) {
#pragma unroll
      for (int i = 0; i < D; ++i) {
        W.at(i, 0) = W.at(i, 0) / c.a1;
        W.at(i, 1) = W.at(i, 1) / c.a2;
      }
    }
  }
};

// ------------------------------------------------------------------------------ FLI (k = 2)
template <int D>
struct FLI {
  bool replace;
  typedef FLIState State;
  CL_HD void init(State& st) const {
    st.s1 = csum_zero(); st.s2 = csum_zero();
    st.f1 = 0.0; st.fB = 0.0; st.value = 0.0;
    st.n = 0; st.step = 0; st.step1 = 0; st.stepB = 0;
    st.code = 0; st.column = 0; st.colB = 0;
#pragma unroll
    for (int i = 0; i < 5; ++i) st.pad_[i] = 0;
  }
  CL_HD void step_logs(State& st, int64_t n, Frame<D, 2>& W, ColLogs& c) const {
    if (st.code != 0) return;
    if (!(c.ok1 && c.ok2)) {
      st.column = c.ok1 ? 2 : 1;
      st.value = c.ok1 ? c.a2 : c.a1;
      st.step = n;
      st.code = 1;
      return;
    }
    CSum s1 = csum_add(st.s1, c.l1), s2 = csum_add(st.s2, c.l2);
    double S1 = csum_total(s1), S2 = csum_total(s2);
    if (S1 > st.f1) { st.f1 = S1; st.step1 = n; }
    if (S1 > st.fB) { st.fB = S1; st.stepB = n; st.colB = 1; }
    if (S2 > st.fB) { st.fB = S2; st.stepB = n; st.colB = 2; }
    st.s1 = s1;
    st.s2 = s2;
    st.n = n;
    if (replace) {
#pragma unroll
      for (int i = 0; i < D; ++i) {
        W.at(i, 0) = W.at(i, 0) / c.a1;
        W.at(i, 1) = W.at(i, 1) / c.a2;
      }
    }
  }
};

// ------------------------------------------------------------------------------ OFLI
// The direction is a field f(x) supplied by the system: the vector field of a flow, or on a map
// the displacement F(x) - x (a definition this build adds; the reference defines OFLI for flows).
template <int D, int K>
struct OFLI {
  bool replace;
  typedef OFLIState<K> State;
  CL_HD void init(State& st) const {
#pragma unroll
    for (int i = 0; i < K; ++i) st.acc[i] = csum_zero();
    double ninf = -__builtin_inf(), inf = __builtin_inf();
    st.g1 = ninf; st.gB = ninf; st.inst = ninf; st.f1 = 0.0; st.fB = 0.0;
    st.sin1 = inf; st.sinB = inf; st.sinmin = inf; st.value = 0.0;
    st.nminf = 0; st.n = 0; st.step = 0; st.step1 = 0; st.stepB = 0;
    st.colB = 0; st.code = 0; st.column = 0;
#pragma unroll
    for (int i = 0; i < 5; ++i) st.pad_[i] = 0;
  }
  // `lc` (k = 2 only): column norms and logs of this frame already formed at this handoff.
  CL_HD void step(State& st, int64_t n, const double* fld, Frame<D, K>& W, const ColLogs* lc = nullptr) const {
    if (st.code != 0) return;
    double fn = safe_norm<D>(fld);
    if (!(cl_isfinite(fn) && fn > 0x1p-1022)) {
      st.value = fn; st.step = n; st.code = 2; st.column = 0;
      return;
    }
    double u[D];
#pragma unroll
    for (int i = 0; i < D; ++i) u[i] = fld[i] / fn;
    double al[K], be[K];
#pragma unroll
    for (int j = 0; j < K; ++j) {
      const double* w = W.col(j);
      al[j] = (K == 2 && lc) ? (j == 0 ? lc->a1 : lc->a2) : safe_norm<D>(w);
      double dw = dot<D>(u, w);
      double v[D];
#pragma unroll
      for (int i = 0; i < D; ++i) v[i] = w[i] - dw * u[i];
      be[j] = safe_norm<D>(v);
    }
#pragma unroll
    for (int j = 0; j < K; ++j) {
      // with the shared column norms their finiteness flags are this test's outcome
      const bool okj = (K == 2 && lc) ? (j == 0 ? lc->ok1 : lc->ok2) : (cl_isfinite(al[j]) && al[j] > 0.0);
      if (st.code == 0 && !okj) {
        st.value = al[j]; st.step = n; st.code = 1; st.column = (int8_t)(j + 1);
      }
    }
    if (st.code != 0) return;
    double g1 = st.g1, gB = st.gB, inst = -__builtin_inf(), f1 = st.f1, fB = st.fB;
    double sin1 = st.sin1, sinB = st.sinB, sinmin = st.sinmin;
    int64_t step1 = st.step1, stepB = st.stepB, nminf = st.nminf;
    int8_t colB = st.colB;
#pragma unroll
    for (int i = 0; i < K; ++i) {
      double la = (K == 2 && lc) ? (i == 0 ? lc->l1 : lc->l2) : cl_log(al[i]);
      CSum na = csum_add(st.acc[i], la);
      // a positive normal be (the usual case) goes straight to the log's arithmetic
      const double tot = csum_total(st.acc[i]);
      double g;
      if (cl_is_pos_normal(be[i])) g = tot + cl_log_normal(be[i], 0);
      else g = be[i] > 0.0 ? tot + cl_log(be[i]) : -__builtin_inf();
      double sn = be[i] / al[i], S = csum_total(na);
      // sn >= +0 and g (finite or -inf, never -0: no log here returns -0 and the sums start at +0) are never NaN,
      // nor are the running values they meet, so Julia's min and max are the hardware's
      sinmin = cl_min_num(sinmin, sn);
      nminf += be[i] > 0.0 ? 0 : 1;
      inst = cl_max_num(inst, g);
      if (g > gB) { gB = g; stepB = n; colB = (int8_t)(i + 1); sinB = sn; }
      if (i == 0 && g > g1) { g1 = g; step1 = n; sin1 = sn; }
      if (i == 0 && S > f1) f1 = S;
      if (S > fB) fB = S;
      st.acc[i] = na;
    }
    st.g1 = g1; st.gB = gB; st.inst = inst; st.f1 = f1; st.fB = fB;
    st.sin1 = sin1; st.sinB = sinB; st.sinmin = sinmin;
    st.nminf = nminf; st.n = n; st.step1 = step1; st.stepB = stepB; st.colB = colB;
    if (replace) {
#pragma unroll
      for (int j = 0; j < K; ++j)
#pragma unroll
        for (int i = 0; i < D; ++i) W.at(i, j) = W.at(i, j) / al[j];
    }
  }
};

// ------------------------------------------------------------------------------ GALI (2 <= K <= D)
template <int D, int K>
struct GALI {
  typedef GALIState<K> State;
  CL_HD void init(State& st) const {
#pragma unroll
    for (int i = 0; i < K; ++i) { st.s[i] = 0.0; st.c[i] = 0.0; }
#pragma unroll
    for (int i = 0; i < K - 1; ++i) st.gali[i] = 0.0;
    st.value = 0.0; st.n = 0; st.step = 0; st.code = 0;
#pragma unroll
    for (int i = 0; i < 7; ++i) st.pad_[i] = 0;
    st.column = 0;
  }
  CL_HD void step(State& st, int64_t n, Frame<D, K>& W) const {
    if (st.code != 0) return;
    double a[K];
#pragma unroll
    for (int j = 0; j < K; ++j) a[j] = safe_norm<D>(W.col(j));
    int bad = 0;
#pragma unroll
    for (int j = 0; j < K; ++j)
      if (bad == 0 && !(cl_isfinite(a[j]) && a[j] >= 0x1p-1022)) bad = j + 1;
    if (bad != 0) {
      double v = a[0];
#pragma unroll
      for (int j = 0; j < K; ++j)
        if (j + 1 == bad) v = a[j];
      st.value = v; st.step = n; st.code = 1; st.column = bad;
      return;
    }
#pragma unroll
    for (int i = 0; i < K; ++i) {
      CSum acc = csum_add(CSum{st.s[i], st.c[i]}, cl_log(a[i]));
      st.s[i] = acc.s;
      st.c[i] = acc.c;
    }
    Frame<D, K> U;
#pragma unroll
    for (int j = 0; j < K; ++j)
#pragma unroll
      for (int i = 0; i < D; ++i) U.at(i, j) = W.at(i, j) / a[j];
    W = U;
    // nested MGS, each column orthogonalised twice (gali.jl:254-274)
    double g = 1.0;
#pragma unroll
    for (int j = 0; j < K; ++j) {
#pragma unroll
      for (int pass = 0; pass < 2; ++pass)
#pragma unroll
        for (int p = 0; p < j; ++p) {
          double c = 0.0;
#pragma unroll
          for (int i = 0; i < D; ++i) c += U.at(i, p) * U.at(i, j);
#pragma unroll
          for (int i = 0; i < D; ++i) U.at(i, j) -= c * U.at(i, p);
        }
      double r = safe_norm<D>(U.col(j));
      double s = r == 0.0 ? 1.0 : r;
#pragma unroll
      for (int i = 0; i < D; ++i) U.at(i, j) /= s;
      if (j > 0) {
        g *= r;
        st.gali[j - 1] = g;
      }
    }
    st.n = n;
  }
};

// ------------------------------------------------------------------------------ ResidualMonitor
// ResidualMonitor(sys, det_tol, orth_tol) (src/Diagnostics/monitor.jl:44-54): the det and symplecticity
// residuals of the one-step tangent at the current point, |det M - 1| and max |M' Omega M - Omega|
// (src/Systems/residuals.jl:19-33), against det_tol eps (1 + max|M|) and det_tol eps N^2 (1 + max|M|)^2, and for
// k >= 2 the frame's orthogonality residual against orth_tol (src/Diagnostics/residuals.jl:10,25-33); running
// maxima. det is the StaticArrays 2 x 2 rule A[1]A[4] - A[3]A[2] for d = 2 and an expansion along the first row
// for d = 4; the matrix products sum left to right.
CL_HD double det3(double a00, double a01, double a02, double a10, double a11, double a12, double a20, double a21,
                  double a22) {
  return (a00 * (a11 * a22 - a12 * a21) - a01 * (a10 * a22 - a12 * a20)) + a02 * (a10 * a21 - a11 * a20);
}
template <int D>
CL_HD double det_of(const double* J) {
  if constexpr (D == 2) {
    return J[0] * J[3] - J[2] * J[1];
  } else {
#define M_(r, c) J[(c) * 4 + (r)]
    double t0 = M_(0, 0) * det3(M_(1, 1), M_(1, 2), M_(1, 3), M_(2, 1), M_(2, 2), M_(2, 3), M_(3, 1), M_(3, 2), M_(3, 3));
    double t1 = M_(0, 1) * det3(M_(1, 0), M_(1, 2), M_(1, 3), M_(2, 0), M_(2, 2), M_(2, 3), M_(3, 0), M_(3, 2), M_(3, 3));
    double t2 = M_(0, 2) * det3(M_(1, 0), M_(1, 1), M_(1, 3), M_(2, 0), M_(2, 1), M_(2, 3), M_(3, 0), M_(3, 1), M_(3, 3));
    double t3 = M_(0, 3) * det3(M_(1, 0), M_(1, 1), M_(1, 2), M_(2, 0), M_(2, 1), M_(2, 2), M_(3, 0), M_(3, 1), M_(3, 2));
#undef M_
    return ((t0 - t1) + t2) - t3;
  }
}
template <int D>
CL_HD double omega_entry(int r, int c) {
  constexpr int H = D / 2;
  return (r < H && c == r + H) ? 1.0 : ((r >= H && c == r - H) ? -1.0 : 0.0);
}
template <int D>
CL_HD double symp_residual(const double* J) {
  double A[D * D];
#pragma unroll
  for (int i = 0; i < D; ++i)
#pragma unroll
    for (int j = 0; j < D; ++j) {
      double v = J[i * D + 0] * omega_entry<D>(0, j);
#pragma unroll
      for (int l = 1; l < D; ++l) v = v + J[i * D + l] * omega_entry<D>(l, j);
      A[i * D + j] = v;
    }
  double mx = 0.0;
#pragma unroll
  for (int i = 0; i < D; ++i)
#pragma unroll
    for (int j = 0; j < D; ++j) {
      double v = A[i * D + 0] * J[j * D + 0];
#pragma unroll
      for (int l = 1; l < D; ++l) v = v + A[i * D + l] * J[j * D + l];
      double e = cl_fabs(v - omega_entry<D>(i, j));
      mx = (i == 0 && j == 0) ? e : cl_jmax(mx, e);
    }
  return mx;
}

template <int D, int K>
struct Monitor {
  typedef MonitorState State;
  double det_tol, orth_tol;
  CL_HD void init(State& st) const {
    st.det_max = st.symp_max = st.orth_max = st.sum_max = st.pair_max = st.G_max = st.M_max = 0.0;
    st.code = 0;
#pragma unroll
    for (int i = 0; i < 7; ++i) st.pad_[i] = 0;
    st.value = st.threshold = 0.0;
    st.step = 0;
  }
  CL_HD void step(State& st, int64_t n, const double* J, const Frame<D, K>& W) const {
    if (st.code != 0) return;
    double mx = cl_fabs(J[0]);
#pragma unroll
    for (int e = 1; e < D * D; ++e) mx = cl_jmax(mx, cl_fabs(J[e]));
    const double dv = cl_fabs(det_of<D>(J) - 1.0);
    const double sv = symp_residual<D>(J);
    const double ov = K >= 2 ? orth_loss<D, K>(W) : 0.0;
    const double om = 1.0 + mx;
    const double thr_d = (det_tol * CL_EPS64) * om;
    const double thr_s = ((det_tol * CL_EPS64) * (double)((D / 2) * (D / 2))) * (om * om);
    const double thr_o = orth_tol;
    const bool bd = !(cl_isfinite(dv) && dv <= thr_d), bs = !(cl_isfinite(sv) && sv <= thr_s),
               bo = !(cl_isfinite(ov) && ov <= thr_o);
    const int code = bd ? 1 : (bs ? 2 : (bo ? 3 : 0));
    st.det_max = cl_jmax(st.det_max, dv);
    st.symp_max = cl_jmax(st.symp_max, sv);
    st.orth_max = cl_jmax(st.orth_max, ov);
    st.M_max = cl_jmax(st.M_max, mx);
    st.code = (int8_t)code;
    st.value = code == 1 ? dv : (code == 2 ? sv : (code == 3 ? ov : 0.0));
    st.threshold = code == 1 ? thr_d : (code == 2 ? thr_s : (code == 3 ? thr_o : 0.0));
    st.step = n;
  }
};

// RunMonitor(sys) (src/Experiments/run_record.jl:25-43): the recorder's non-raising monitor, running maxima of
// |det M - 1|, max |M' Omega M - Omega| and max |M| of the one-step tangent at each handoff; it never fails.
struct RunMonState {
  double det_max, symp_max, M_max;
};
template <int D>
struct RunMon {
  typedef RunMonState State;
  CL_HD void init(State& st) const { st.det_max = st.symp_max = st.M_max = 0.0; }
  CL_HD void step(State& st, int64_t, const double* J) const {
    double mx = cl_fabs(J[0]);
#pragma unroll
    for (int e = 1; e < D * D; ++e) mx = cl_jmax(mx, cl_fabs(J[e]));
    st.det_max = cl_jmax(st.det_max, cl_fabs(det_of<D>(J) - 1.0));
    st.symp_max = cl_jmax(st.symp_max, symp_residual<D>(J));
    st.M_max = cl_jmax(st.M_max, mx);
  }
};

// ------------------------------------------------------------------------------ StretchingNumbers (k = 2)
// StretchingNumbers (src/Indicators/stretching.jl:34-147): at each handoff n = K tau the two arriving columns give
// a_j = log(alpha_j) (the column logs MEGNO, FLI and SALI share, so the sums are bitwise theirs); a_j and a_j^2 go into
// compensated sums and the stretching number z_j = a_j / tau is counted in a cell of a cumulative histogram in the
// orbit's UInt32 workspace: bins cells of width h = (hi - lo) / bins over [lo, hi], cell 0 below, bins + 1 above
// (spectrum_cell, :104-108). Column j's cells are the workspace entries (j - 1)(bins + 2) + cell.
struct StretchState {
  CSum s1, s2, q1, q2;
  double value;
  int64_t n, K, step;
  int8_t code, column;
  int8_t pad_[6];
};
constexpr int64_t SN_MAX_HANDOFFS = 1000000000;  // 4-byte counts (stretching.jl:91)

template <int D>
struct Stretch {
  bool replace;
  double lo, hi;
  int bins;
  typedef StretchState State;
  CL_HD void init(State& st) const {
    st.s1 = csum_zero(); st.s2 = csum_zero(); st.q1 = csum_zero(); st.q2 = csum_zero();
    st.value = 0.0;
    st.n = 0; st.K = 0; st.step = 0;
    st.code = 0; st.column = 0;
#pragma unroll
    for (int i = 0; i < 6; ++i) st.pad_[i] = 0;
  }
  CL_HD int64_t cell(double z) const {
    if (z < lo) return 0;
    if (z > hi) return (int64_t)bins + 1;
    const int64_t c = (int64_t)cl_floor((z - lo) / ((hi - lo) / (double)bins)) + 1;
    return c < (int64_t)bins ? c : (int64_t)bins;
  }
  // The counts of the launch so far, staged in 8-bit counters (entry e at cnt[e * cs]; null: counted in ws directly),
  // added to the workspace and cleared. Staged counts stay below 256: a flush follows every 255th handoff.
  static constexpr int64_t STAGE_HANDOFFS = 255;
  CL_HD void flush(uint8_t* cnt, int cs, uint32_t* ws, int64_t stride) const {
    if (!cnt) return;
    const int64_t e = 2 * ((int64_t)bins + 2);
    for (int64_t j = 0; j < e; ++j) {
      const uint32_t v = cnt[j * cs];
      if (v) {
        ws[j * stride] += v;
        cnt[j * cs] = 0;
      }
    }
  }
  // ws: the orbit's first workspace entry, entries `stride` apart
  template <int K>
  CL_HD void step_logs(State& st, int64_t n, Frame<D, K>& W, ColLogs& c, uint32_t* ws, int64_t stride,
                       uint8_t* cnt = nullptr, int cs = 0) const {
    if (st.code != 0) return;
    if (!(c.ok1 && c.ok2)) {
      st.column = c.ok1 ? 2 : 1;
      st.value = c.ok1 ? c.a2 : c.a1;
      st.step = n;
      st.code = 1;
      return;
    }
    const int64_t K1 = st.K + 1;
    if (K1 > SN_MAX_HANDOFFS) {
      st.value = (double)K1;
      st.step = n;
      st.code = 2;
      st.column = 0;
      return;
    }
    const double tau = (double)(n / K1);
    const int64_t e1 = cell(c.l1 / tau), e2 = (int64_t)bins + 2 + cell(c.l2 / tau);
    if (cnt) {
      cnt[e1 * cs] += 1;
      cnt[e2 * cs] += 1;
      if (K1 % STAGE_HANDOFFS == 0) flush(cnt, cs, ws, stride);
    } else {
      ws[e1 * stride] += 1u;
      ws[e2 * stride] += 1u;
    }
    st.s1 = csum_add(st.s1, c.l1);
    st.s2 = csum_add(st.s2, c.l2);
    st.q1 = csum_add(st.q1, c.l1 * c.l1);
    st.q2 = csum_add(st.q2, c.l2 * c.l2);
    st.n = n;
    st.K = K1;
    if (replace) {
#pragma unroll
      for (int i = 0; i < D; ++i) {
        W.at(i, 0) = W.at(i, 0) / c.a1;
        W.at(i, 1) = W.at(i, 1) / c.a2;
      }
    }
  }
};

// ------------------------------------------------------------------------------ WeightedBirkhoff
// WeightedBirkhoff (src/Indicators/wbirkhoff_reducer.jl:4-116, wbirkhoff.jl): K observables of the orbit point over H
// horizons in one pass. Handoff j (0-based) is sample j; horizon N takes samples 0..2N-1, window 1 the first N with
// the weight w(j), window 2 the next N with w(j - N) (_wb_advance, :42-48): B += w (window 1 only, a zero weight adds
// nothing), A_i += w f (Neumaier), S = max(S, |f|). The horizons' sums (H (2 + 5K) words: B, then per observable A1,
// A2, S) are too many for registers: they stay in HBM, word-major like the other states, and each handoff reads and
// writes only the horizons that have not yet taken their 2N samples. The weights w(jj) = exp(-1 / (t^p (1-t)^p)),
// t = min(jj, N - jj) / N (birkhoff_weight, naff_reducer.jl:13-17), are the same for every orbit: a table per
// horizon, formed once per launch. The register state is the tail of WBState: n, handoffs, code.
struct WBHead {
  int64_t n;
  int32_t handoffs;
  int16_t code;
  int16_t pad_;
};
// The per-launch table (device memory): words 0..15 the windows N_h, 16..31 the offsets of their weights, then the
// weights (so the horizon loop reads its bounds with uniform loads instead of indexing a kernel argument).
constexpr int WB_TABLE_HEADER = 2 * CL_WB_MAX_WINDOWS;
struct WBParams {
  int H, K, order;
  int kind[CL_WB_MAX_OBSERVABLES], a[CL_WB_MAX_OBSERVABLES], b[CL_WB_MAX_OBSERVABLES];
  int64_t nmax2;  // 2 max(window): handoffs of a complete run
  int hw;         // words of the horizons' sums
};
// Makes v computed at this point: the compiler then cannot sink its computation into the conditional block that
// uses it, where independent computations would run one after another instead of interleaving.
CL_HD void cl_pin(double v) {
#if defined(__HIP_DEVICE_COMPILE__)
  __asm__ volatile("" ::"v"(v));
#else
  (void)v;
#endif
}

// x[a] of a point of S coordinates without indexing the register array at run time
template <int S>
CL_HD double wb_pick(const double* x, int a) {
  double v = x[0];
#pragma unroll
  for (int j = 1; j < S; ++j) {
    double xj = x[j];
#if defined(__HIP_DEVICE_COMPILE__)
    __asm__ volatile("" : "+v"(xj));  // keeps the selects on registers (else they fold into a scratch load of x[a])
#endif
    v = (a == j) ? xj : v;
  }
  return v;
}
CL_HD double wb_pow(double u, int p) {
  const double u2 = u * u;
  return p == 1 ? u : (p == 2 ? u2 : (p == 3 ? u2 * u : u2 * u2));
}
// birkhoff_weight(T(min(jj, N - jj)) / T(N), p)
CL_HD double wb_weight(int64_t jj, int64_t N, int p) {
  const int64_t mn = jj < N - jj ? jj : N - jj;
  const double t = (double)mn / (double)N;
  if (t <= 0.0 || t >= 1.0) return 0.0;
  return cl_exp(-1.0 / (wb_pow(t, p) * wb_pow(1.0 - t, p)));
}

// ------------------------------------------------------------------------------ WeightedPhaseAdvance
// naff_reducer.jl:163-230: per horizon N the two windows' weighted sums A = sum w_j Delta_j, B = sum w_j of the sample
// increments (window 1: increments 0..N-1, window 2: N..2N-1, weights birkhoff_weight at the mirrored index), the
// previous sample, the count of increments near the branch cut (planar), a non-finite sample recorded as code 1.
struct WPAHorizon {
  CSum A1, B1, A2, B2;
};
template <int H>
struct WPAState {  // the reference's WPAState{Float64, H} layout: 64 H + 32 bytes
  WPAHorizon hor[H];
  double prev1, prev2;
  int64_t n;
  int32_t handoffs;
  int16_t nwrap;
  int8_t code, pad_;
};

// phase_increments' increment (naff_reducer.jl:47-67): the principal argument of z conj(zprev), from power-of-two
// normalised samples when the product leaves [1e-280, 1e280]; NaN when a sample is exactly 0.
CL_HD void wpa_unit_scale(double re, double im, double* ore, double* oim) {
  const double m = __builtin_fmax(cl_fabs(re), cl_fabs(im));
  int e = 0;
  if (cl_isfinite(m) && m > 0.0) {
    e = (int)((cl_bits(m) >> 52) & 0x7ff) - 1023;
    e = e < -1000 ? -1000 : (e > 1000 ? 1000 : e);
  }
  *ore = __builtin_ldexp(re, -e);
  *oim = __builtin_ldexp(im, -e);
}
CL_HD double wpa_phase_inc(double pre, double pim, double re, double im) {
  if ((pre == 0.0 && pim == 0.0) || (re == 0.0 && im == 0.0)) return __builtin_nan("");
  double rr = re * pre - im * (-pim);
  double ri = re * (-pim) + im * pre;
  const double s = __builtin_fmax(cl_fabs(rr), cl_fabs(ri));
  if (!(s > 1e-280 && s < 1e280)) {
    double a, b, c, d;
    wpa_unit_scale(re, im, &a, &b);
    wpa_unit_scale(pre, pim, &c, &d);
    rr = a * c - b * (-d);
    ri = a * (-d) + b * c;
  }
  return cl_atan2(ri, rr);
}

template <int H>
struct WPA {
  int64_t window[H];
  int order, obs, ia, ib;
  int64_t wmax;
  typedef WPAState<H> State;
  CL_HD void init(State& st) const {
#pragma unroll
    for (int h = 0; h < H; ++h) st.hor[h].A1 = st.hor[h].B1 = st.hor[h].A2 = st.hor[h].B2 = csum_zero();
    st.prev1 = st.prev2 = 0.0;
    st.n = 0;
    st.handoffs = 0;
    st.nwrap = 0;
    st.code = 0;
    st.pad_ = 0;
  }
  template <class Sys>
  CL_HD void step(State& st, int64_t n, const double* x) const {
    if (st.code != 0) return;
    double a0 = wb_pick<Sys::S>(x, ia);
    double a1 = obs == CL_OBS_LIFTED_ANGLE && ib < 0 ? 0.0 : wb_pick<Sys::S>(x, ib);
    if (!(cl_isfinite(a0) && cl_isfinite(a1))) {
      st.n = n;
      st.code = 1;
      return;
    }
    const int32_t s = st.handoffs;
    const int32_t h1 = s < 2147483647 ? s + 1 : s;
    if (s == 0) {
      st.prev1 = a0;
      st.prev2 = a1;
      st.n = n;
      st.handoffs = h1;
      return;
    }
    const double tp = 2.0 * 3.141592653589793;
    const double d = obs == CL_OBS_PLANAR ? wpa_phase_inc(st.prev1, st.prev2, a0, a1)
                                          : (a0 - st.prev1) + tp * (a1 - st.prev2);
    const int64_t j = (int64_t)s - 1;
#pragma unroll
    for (int h = 0; h < H; ++h) {
      const int64_t N = window[h];
      if (j < 2 * N) {
        const double w = wb_weight(j < N ? j : j - N, N, order);
        if (w != 0.0) {
          if (j < N) {
            st.hor[h].A1 = csum_add(st.hor[h].A1, w * d);
            st.hor[h].B1 = csum_add(st.hor[h].B1, w);
          } else {
            st.hor[h].A2 = csum_add(st.hor[h].A2, w * d);
            st.hor[h].B2 = csum_add(st.hor[h].B2, w);
          }
        }
      }
    }
    const bool flagged = obs == CL_OBS_PLANAR && cl_fabs(d) > 9.0 * 3.141592653589793 / 10.0 &&
                         (int64_t)s <= 2 * wmax && st.nwrap < 32767;
    st.prev1 = a0;
    st.prev2 = a1;
    st.n = n;
    st.handoffs = h1;
    if (flagged) st.nwrap = (int16_t)(st.nwrap + 1);
  }
};

struct WBirk {
  WBParams p;
  typedef WBHead State;
  CL_HD void init(State& st) const {
    st.n = 0;
    st.handoffs = 0;
    st.code = 0;
    st.pad_ = 0;
  }
  template <class Sys>
  CL_HD double observable(int k, const double* x) const {
    const int a = p.a[k], b = p.b[k];
    const double xa = wb_pick<Sys::S>(x, a);
    if (p.kind[k] == CL_OBS_COS) return cl_cos((double)b * xa);
    if (p.kind[k] == CL_OBS_SIN) return cl_sin((double)b * xa);
    return xa + 2.0 * (3.141592653589793 + 1.2246467991473532e-16) * wb_pick<Sys::S>(x, b);
  }
  // mem: the orbit's first horizon word (words `stride` apart); wtab: the weight tables
  template <class Sys>
  CL_HD void step(State& st, int64_t n, const double* x, double* mem, int64_t stride, const double* wtab) const {
    if (st.code != 0) return;
    if ((int64_t)st.handoffs >= p.nmax2) {  // every horizon is complete: only the counters advance
      st.n = n;
      if (st.handoffs < 2147483647) st.handoffs += 1;
      return;
    }
    double f[CL_WB_MAX_OBSERVABLES];
#pragma unroll
    for (int k = 0; k < CL_WB_MAX_OBSERVABLES; ++k) f[k] = k < p.K ? observable<Sys>(k, x) : 0.0;
#pragma unroll
    for (int k = 0; k < CL_WB_MAX_OBSERVABLES; ++k)
      if (k < p.K && !cl_isfinite(f[k])) {
        st.code = (int16_t)(k + 1);
        st.n = n;
        return;
      }
    const int64_t j = st.handoffs;
    const int cw = 2 + 5 * p.K;
    const int64_t* hdr = reinterpret_cast<const int64_t*>(wtab);
    for (int h = 0; h < p.H; ++h) {
      const int64_t N = hdr[h];
      if (j >= 2 * N) continue;
      const int i = j >= N ? 1 : 0;
      const double w = wtab[WB_TABLE_HEADER + hdr[CL_WB_MAX_WINDOWS + h] + (j - i * N)];
      double* hb = mem + (int64_t)h * cw * stride;
      if (i == 0 && w != 0.0) {
        CSum B{hb[0], hb[stride]};
        B = csum_add(B, w);
        hb[0] = B.s;
        hb[stride] = B.c;
      }
#pragma unroll
      for (int k = 0; k < CL_WB_MAX_OBSERVABLES; ++k) {
        if (k >= p.K) continue;
        double* c = hb + (int64_t)(2 + 5 * k) * stride;
        c[4 * stride] = cl_jmax(c[4 * stride], cl_fabs(f[k]));
        if (w != 0.0) {
          double* A = c + 2 * i * stride;
          CSum a{A[0], A[stride]};
          a = csum_add(a, w * f[k]);
          A[0] = a.s;
          A[stride] = a.c;
        }
      }
    }
    st.n = n;
    if (j < 2147483647) st.handoffs = (int32_t)(j + 1);
  }
};

}  // namespace cl
