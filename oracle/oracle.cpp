// ORACLE - test infrastructure only. A CPU restatement of the reference's ensemble tangent stream
// (Chaoslab `propagate`), used by tests/, by __graft_entry__.smoke() and by bench.py's CPU-baseline
// leg, always as the checker or the timed CPU reference, never as the product path.
//
// What it restates (reference file:line, paths relative to the Chaoslab repository root):
//   or_propagate          _orbit_body, src/Tangent/orbit.jl:44-60; the chunk/thread split of the
//                         KernelAbstractions CPU backend, src/Tangent/propagate.jl:59-60 (16 orbits per group)
//   fused product         src/Foundation/product.jl:9-20
//   CompensatedSum/add    src/Foundation/Foundation.jl:48-58;   safe_norm  src/Foundation/norm.jl:8-12
//   StandardMap           src/Systems/standard_map.jl:18-43;   WrappedStandardMap  wrapped_standard_map.jl:16-57
//   QRSpectrum            src/Lyapunov/qr_spectrum.jl:22-93; orthonormalize.jl:38-151; householder.jl:17-123
//   Benettin              src/Lyapunov/benettin.jl:14-48
//   MEGNO                 src/Indicators/megno.jl:35-111;       SALI sali.jl:150-199;   FLI fli.jl:299-360
//   OFLI                  src/Indicators/ofli.jl:30-124 (direction field: the flow's, or the map displacement)
//   GALI                  src/Indicators/gali.jl:20-199
//   LogNorm (test reducer) test/tangent-core/tc_helpers.jl:56-70
//   ResidualMonitor       src/Diagnostics/monitor.jl:44-54; residuals src/Diagnostics/residuals.jl:4-36,
//                         src/Systems/residuals.jl:19-33
//   StandardMapChain, ChainTangent and the Householder QR spectrum on d = 2N rows: oracle_chain.hpp
//
// Independence from the product: the reducer chain is interpreted at run time over byte buffers
// (the product compiles one kernel per chain), and every reducer is written out again here. The
// elementary functions are selectable: MATH_LIBM (glibc sin/cos/log, fully independent),
// MATH_CL (the project's portable sin/cos/log from chaos_amd/csrc/cl_math.h, which the kernels use;
// selected when bit-for-bit agreement with a kernel is the test) and, in the __float128
// instantiation, libquadmath (113-bit significand, finer than Double64's 106 bits).
#include <math.h>
#include <quadmath.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <thread>
#include <vector>

#include "../include/chaoslab_hip.h"
#include "../chaos_amd/csrc/cl_math.h"

typedef __float128 quad;

enum { MATH_LIBM = 0, MATH_CL = 1 };

// Planted defects for the negative controls of the closed-form tests (tests/test_oracle_pins.py): a test switches one
// on, shows that the closed form it checks then fails, and switches it off. 0 (the default) is the oracle as written.
enum { DEFECT_NONE = 0, DEFECT_GALI_RAW = 1, DEFECT_OFLI_NOPROJ = 2, DEFECT_MEGNO_Q2 = 3, DEFECT_GALI_LASTFACTOR = 4,
       DEFECT_BLK_TNOTRANS = 5, DEFECT_BLK_TSIGN = 6 };
static int g_defect = DEFECT_NONE;

// ------------------------------------------------------------------ arithmetic policies
struct F64Libm {
  typedef double T;
  static double sin(double x) { return cl_isfinite(x) ? ::sin(x) : x - x; }
  static double cos(double x) { return cl_isfinite(x) ? ::cos(x) : x - x; }
  static double log(double x) { return ::log(x); }
  static double exp(double x) { return ::exp(x); }
  static double sqrt(double x) { return ::sqrt(x); }
  static double fma(double a, double b, double c) { return ::fma(a, b, c); }
  static double abs(double x) { return fabs(x); }
  static bool finite(double x) { return cl_isfinite(x); }
  static bool nan(double x) { return cl_isnan(x); }
  static bool sign(double x) { return cl_signbit(x); }
  static double eps() { return 0x1p-52; }
  static double floatmin() { return 0x1p-1022; }
  static double floor(double x) { return ::floor(x); }
  static double pi() { return 3.141592653589793 + 1.2246467991473532e-16; }
};
struct F64Cl : F64Libm {
  static double sin(double x) { return cl_sin(x); }
  static double cos(double x) { return cl_cos(x); }
  static double log(double x) { return cl_log(x); }
  static double exp(double x) { return cl_exp(x); }
};
struct Quad {
  typedef quad T;
  static quad sin(quad x) { return finite(x) ? sinq(x) : x - x; }
  static quad cos(quad x) { return finite(x) ? cosq(x) : x - x; }
  static quad log(quad x) { return logq(x); }
  static quad exp(quad x) { return expq(x); }
  static quad sqrt(quad x) { return sqrtq(x); }
  static quad fma(quad a, quad b, quad c) { return fmaq(a, b, c); }
  static quad abs(quad x) { return fabsq(x); }
  static bool finite(quad x) { return finiteq(x); }
  static bool nan(quad x) { return isnanq(x); }
  static bool sign(quad x) { return signbitq(x) != 0; }
  static quad eps() { return FLT128_EPSILON; }
  static quad floatmin() { return FLT128_MIN; }
  static quad floor(quad x) { return floorq(x); }
  static quad pi() { return M_PIq; }
};

// Julia float max/min: NaN propagates, -0 < +0.
template <class P, class T>
static T jmax(T x, T y) {
  bool takey = (y > x) || (P::sign(y) < P::sign(x));
  return takey ? (P::nan(x) ? x : y) : (P::nan(y) ? y : x);
}
template <class P, class T>
static T jmin(T x, T y) {
  bool takey = (y < x) || (P::sign(y) > P::sign(x));
  return takey ? (P::nan(x) ? x : y) : (P::nan(y) ? y : x);
}

template <class T>
struct Acc {
  T s, c;
};
template <class P, class T>
static Acc<T> acc_add(Acc<T> a, T y) {
  T t = a.s + y;
  T c = P::abs(a.s) >= P::abs(y) ? (a.s - t) + y : (y - t) + a.s;
  Acc<T> o;
  o.s = t;
  o.c = a.c + c;
  return o;
}
template <class T>
static T acc_total(Acc<T> a) {
  return a.s + a.c;
}

// scaled norm of v[0], v[st], ..., n entries
template <class P, class T>
static T snorm(const T* v, int n, int st = 1) {
  T m = P::abs(v[0]);
  for (int i = 1; i < n; ++i) m = jmax<P>(m, P::abs(v[i * st]));
  if (!(P::finite(m) && m > 0)) return m;
  T s = (v[0] / m) * (v[0] / m);
  for (int i = 1; i < n; ++i) {
    T b = v[i * st] / m;
    s = s + b * b;
  }
  return m * P::sqrt(s);
}
template <class T>
static T vdot(const T* a, const T* b, int n) {
  T s = a[0] * b[0];
  for (int i = 1; i < n; ++i) s = s + a[i] * b[i];
  return s;
}

// ------------------------------------------------------------------ Julia state layouts (Float64)
// Byte writer/reader for the chain's state tuple: fields at their Julia offsets.
struct Buf {
  uint8_t* p;
  int off;
  template <class V>
  void put(V v) {
    memcpy(p + off, &v, sizeof(V));
    off += sizeof(V);
  }
  template <class V>
  V get() {
    V v;
    memcpy(&v, p + off, sizeof(V));
    off += sizeof(V);
    return v;
  }
  void pad(int n) {
    memset(p + off, 0, n);
    off += n;
  }
  void skip(int n) { off += n; }
};

static int red_bytes(const cl_reducer& r, int d, int k) {
  switch (r.kind) {
    case CL_RED_QR: return 16 * k + 72;
    case CL_RED_BENETTIN: return 48;
    case CL_RED_MEGNO: return 200;
    case CL_RED_SALI: return 72;
    case CL_RED_RLI: return 88;
    case CL_RED_FLI: return 96;
    case CL_RED_OFLI: return 16 * k + 120;
    case CL_RED_GALI: return 24 * k + 32;
    case 8: return 16 * k;  // LogNorm
    case CL_RED_MONITOR: return 88;
    case CL_RED_RUNMONITOR: return 24;
    case CL_RED_STRETCHING: return 104;
    case CL_RED_BIRKHOFF: return r.birkhoff ? 8 * r.birkhoff->H * (2 + 5 * r.birkhoff->K) + 16 : -1;
    case CL_RED_PHASE_ADVANCE: return r.birkhoff ? 64 * r.birkhoff->H + 32 : -1;
  }
  return -1;
}

// UInt32 workspace entries of a chain: 2 (bins + 2) for StretchingNumbers (stretching.jl:93-96), else 0
static int64_t chain_ws(const cl_chain& ch) {
  for (int r = 0; r < ch.n; ++r)
    if (ch.r[r].kind == CL_RED_STRETCHING) return 2 * ((int64_t)ch.r[r].bins + 2);
  return 0;
}

// ------------------------------------------------------------------ reducer states in T
template <class T>
struct QRSt {
  Acc<T> sums[8];
  int64_t n, count_max, column, step;
  T G_max, orth_max, value, threshold;
  int8_t code;
};
template <class T>
struct BenSt {
  Acc<T> sum;
  int64_t n, step;
  int8_t code;
  T value;
};
template <class T>
struct MSums {
  Acc<T> s, R1, R2, Q, Q2;
};
template <class T>
struct MegSt {
  MSums<T> c[2];
  T value;
  int64_t n, K, step;
  int8_t code, column;
};
template <class T>
struct RliSt {  // RLIState (rli.jl:53-63)
  Acc<T> sA, sB, Q;
  T value;
  int64_t n, K, step;
  int8_t code, column;
};
template <class T>
struct SaliSt {
  Acc<T> s1, s2;
  T sali, value;
  int64_t n, step;
  int8_t code, column;
};
template <class T>
struct FliSt {
  Acc<T> s1, s2;
  T f1, fB, value;
  int64_t n, step, step1, stepB;
  int8_t code, column, colB;
};
template <class T>
struct OfliSt {
  Acc<T> acc[8];
  T g1, gB, inst, f1, fB, sin1, sinB, sinmin, value;
  int64_t nminf, n, step, step1, stepB;
  int8_t colB, code, column;
};
template <class T>
struct GaliSt {
  T s[8], c[8], gali[8], value;
  int64_t n, step, column;
  int8_t code;
};
template <class T>
struct LogNormSt {
  Acc<T> a[8];
};

template <class T>
struct MonSt {
  T det_max, symp_max, orth_max, sum_max, pair_max, G_max, M_max;
  int8_t code;
  T value, threshold;
  int64_t step;
};
template <class T>
struct StretchSt {  // StretchState{T} (stretching.jl:77-88)
  Acc<T> s1, s2, q1, q2;
  T value;
  int64_t n, K, step;
  int8_t code, column;
};
template <class T>
struct WBSt {  // WBState{T, H, K} (wbirkhoff_reducer.jl:36-49), up to 16 horizons and 8 observables
  Acc<T> B[CL_WB_MAX_WINDOWS];
  Acc<T> A1[CL_WB_MAX_WINDOWS][CL_WB_MAX_OBSERVABLES], A2[CL_WB_MAX_WINDOWS][CL_WB_MAX_OBSERVABLES];
  T S[CL_WB_MAX_WINDOWS][CL_WB_MAX_OBSERVABLES];
  int64_t n;
  int32_t handoffs;
  int16_t code;
};
template <class T>
struct PaSt {  // WPAState{T, H} (naff_reducer.jl:166-174), 1 or 2 horizons
  Acc<T> A1[2], B1[2], A2[2], B2[2];
  T prev1, prev2;
  int64_t n;
  int32_t handoffs;
  int16_t nwrap;
  int8_t code;
};
template <class T>
struct RState {  // one of the above, by kind
  WBSt<T> wb;
  PaSt<T> pa;
  StretchSt<T> sn;
  MonSt<T> mon;
  QRSt<T> qr;
  BenSt<T> ben;
  MegSt<T> meg;
  SaliSt<T> sali;
  RliSt<T> rli;
  FliSt<T> fli;
  OfliSt<T> ofli;
  GaliSt<T> gali;
  LogNormSt<T> ln;
};

static double dbl(double v) { return v; }
static double dbl(quad v) { return (double)v; }

// serialise a state to the Julia layout (Float64 values)
template <class T>
static void write_state(const cl_reducer& r, const RState<T>& s, int d, int k, uint8_t* out) {
  Buf b{out, 0};
  switch (r.kind) {
    case CL_RED_MONITOR: {
      const MonSt<T>& q = s.mon;
      b.put(dbl(q.det_max));
      b.put(dbl(q.symp_max));
      b.put(dbl(q.orth_max));
      b.put(dbl(q.sum_max));
      b.put(dbl(q.pair_max));
      b.put(dbl(q.G_max));
      b.put(dbl(q.M_max));
      b.put(q.code);
      b.pad(7);
      b.put(dbl(q.value));
      b.put(dbl(q.threshold));
      b.put(q.step);
      break;
    }
    case CL_RED_RUNMONITOR: {
      b.put(dbl(s.mon.det_max));
      b.put(dbl(s.mon.symp_max));
      b.put(dbl(s.mon.M_max));
      break;
    }
    case CL_RED_BIRKHOFF: {
      const WBSt<T>& q = s.wb;
      const cl_birkhoff& pb = *r.birkhoff;
      for (int h = 0; h < pb.H; ++h) {
        b.put(dbl(q.B[h].s));
        b.put(dbl(q.B[h].c));
        for (int c = 0; c < pb.K; ++c) {
          b.put(dbl(q.A1[h][c].s));
          b.put(dbl(q.A1[h][c].c));
          b.put(dbl(q.A2[h][c].s));
          b.put(dbl(q.A2[h][c].c));
          b.put(dbl(q.S[h][c]));
        }
      }
      b.put(q.n);
      b.put(q.handoffs);
      b.put(q.code);
      b.pad(2);
      break;
    }
    case CL_RED_STRETCHING: {
      const StretchSt<T>& q = s.sn;
      const Acc<T>* a4[4] = {&q.s1, &q.s2, &q.q1, &q.q2};
      for (int j = 0; j < 4; ++j) {
        b.put(dbl(a4[j]->s));
        b.put(dbl(a4[j]->c));
      }
      b.put(dbl(q.value));
      b.put(q.n);
      b.put(q.K);
      b.put(q.step);
      b.put(q.code);
      b.put(q.column);
      b.pad(6);
      break;
    }
    case CL_RED_QR: {
      const QRSt<T>& q = s.qr;
      for (int i = 0; i < k; ++i) {
        b.put(dbl(q.sums[i].s));
        b.put(dbl(q.sums[i].c));
      }
      b.put(q.n);
      b.put(dbl(q.G_max));
      b.put(dbl(q.orth_max));
      b.put(q.count_max);
      b.put(q.code);
      b.pad(7);
      b.put(q.column);
      b.put(dbl(q.value));
      b.put(dbl(q.threshold));
      b.put(q.step);
      break;
    }
    case CL_RED_BENETTIN: {
      const BenSt<T>& q = s.ben;
      b.put(dbl(q.sum.s));
      b.put(dbl(q.sum.c));
      b.put(q.n);
      b.put(q.code);
      b.pad(7);
      b.put(dbl(q.value));
      b.put(q.step);
      break;
    }
    case CL_RED_MEGNO: {
      const MegSt<T>& q = s.meg;
      for (int c = 0; c < 2; ++c) {
        const MSums<T>& m = q.c[c];
        const Acc<T>* a[5] = {&m.s, &m.R1, &m.R2, &m.Q, &m.Q2};
        for (int j = 0; j < 5; ++j) {
          b.put(dbl(a[j]->s));
          b.put(dbl(a[j]->c));
        }
      }
      b.put(dbl(q.value));
      b.put(q.n);
      b.put(q.K);
      b.put(q.step);
      b.put(q.code);
      b.put(q.column);
      b.pad(6);
      break;
    }
    case CL_RED_PHASE_ADVANCE: {
      const PaSt<T>& q = s.pa;
      for (int h = 0; h < r.birkhoff->H; ++h) {
        b.put(dbl(q.A1[h].s)); b.put(dbl(q.A1[h].c));
        b.put(dbl(q.B1[h].s)); b.put(dbl(q.B1[h].c));
        b.put(dbl(q.A2[h].s)); b.put(dbl(q.A2[h].c));
        b.put(dbl(q.B2[h].s)); b.put(dbl(q.B2[h].c));
      }
      b.put(dbl(q.prev1));
      b.put(dbl(q.prev2));
      b.put(q.n);
      b.put(q.handoffs);
      b.put(q.nwrap);
      b.put(q.code);
      b.pad(1);
      break;
    }
    case CL_RED_RLI: {
      const RliSt<T>& q = s.rli;
      b.put(dbl(q.sA.s));
      b.put(dbl(q.sA.c));
      b.put(dbl(q.sB.s));
      b.put(dbl(q.sB.c));
      b.put(dbl(q.Q.s));
      b.put(dbl(q.Q.c));
      b.put(dbl(q.value));
      b.put(q.n);
      b.put(q.K);
      b.put(q.step);
      b.put(q.code);
      b.put(q.column);
      b.pad(6);
      break;
    }
    case CL_RED_SALI: {
      const SaliSt<T>& q = s.sali;
      b.put(dbl(q.s1.s));
      b.put(dbl(q.s1.c));
      b.put(dbl(q.s2.s));
      b.put(dbl(q.s2.c));
      b.put(dbl(q.sali));
      b.put(dbl(q.value));
      b.put(q.n);
      b.put(q.step);
      b.put(q.code);
      b.put(q.column);
      b.pad(6);
      break;
    }
    case CL_RED_FLI: {
      const FliSt<T>& q = s.fli;
      b.put(dbl(q.s1.s));
      b.put(dbl(q.s1.c));
      b.put(dbl(q.s2.s));
      b.put(dbl(q.s2.c));
      b.put(dbl(q.f1));
      b.put(dbl(q.fB));
      b.put(dbl(q.value));
      b.put(q.n);
      b.put(q.step);
      b.put(q.step1);
      b.put(q.stepB);
      b.put(q.code);
      b.put(q.column);
      b.put(q.colB);
      b.pad(5);
      break;
    }
    case CL_RED_OFLI: {
      const OfliSt<T>& q = s.ofli;
      for (int i = 0; i < k; ++i) {
        b.put(dbl(q.acc[i].s));
        b.put(dbl(q.acc[i].c));
      }
      T v[9] = {q.g1, q.gB, q.inst, q.f1, q.fB, q.sin1, q.sinB, q.sinmin, q.value};
      for (int i = 0; i < 9; ++i) b.put(dbl(v[i]));
      b.put(q.nminf);
      b.put(q.n);
      b.put(q.step);
      b.put(q.step1);
      b.put(q.stepB);
      b.put(q.colB);
      b.put(q.code);
      b.put(q.column);
      b.pad(5);
      break;
    }
    case CL_RED_GALI: {
      const GaliSt<T>& q = s.gali;
      for (int i = 0; i < k; ++i) b.put(dbl(q.s[i]));
      for (int i = 0; i < k; ++i) b.put(dbl(q.c[i]));
      for (int i = 0; i < k - 1; ++i) b.put(dbl(q.gali[i]));
      b.put(dbl(q.value));
      b.put(q.n);
      b.put(q.step);
      b.put(q.code);
      b.pad(7);
      b.put(q.column);
      break;
    }
    case 8: {
      for (int i = 0; i < k; ++i) {
        b.put(dbl(s.ln.a[i].s));
        b.put(dbl(s.ln.a[i].c));
      }
      break;
    }
  }
}

template <class T>
static void init_state(const cl_reducer& r, RState<T>& s, int k) {
  memset(&s, 0, sizeof(s));
  if (r.kind == CL_RED_OFLI) {
    OfliSt<T>& q = s.ofli;
    T inf = (T)INFINITY;
    q.g1 = q.gB = q.inst = -inf;
    q.sin1 = q.sinB = q.sinmin = inf;
  }
  (void)k;
}

// ------------------------------------------------------------------ QR family
template <class P, class T>
struct Fact {
  T Q[64], R[64];
  int count, column;
  int status;  // 0 ok, 1 pivot, 2 passes
};

// W, Q column-major d x k; R k x k column-major
template <class P, class T>
static void qr_gs(int alg, const T* W, int d, int k, Fact<P, T>& f) {
  for (int i = 0; i < d * k; ++i) f.Q[i] = 0;
  for (int i = 0; i < k * k; ++i) f.R[i] = 0;
  f.count = 1;
  f.status = 0;
  f.column = 0;
  for (int m = 0; m < k; ++m) {
    T u[8], r[8], w[8];
    for (int i = 0; i < d; ++i) u[i] = w[i] = W[m * d + i];
    for (int i = 0; i < k; ++i) r[i] = 0;
    int passes = 1;
    if (alg == CL_QR_CGS || alg == CL_QR_MGS) {
      for (int i = 0; i < m; ++i) {
        const T* q = f.Q + i * d;
        T c = alg == CL_QR_CGS ? vdot(q, w, d) : vdot(q, u, d);
        r[i] = c;
        for (int l = 0; l < d; ++l) u[l] = u[l] - c * q[l];
      }
    } else {
      T eta = (T)1 / P::sqrt((T)2);
      T nold = snorm<P>(u, d);
      passes = 0;
      while (true) {
        passes++;
        T vold[8];
        for (int i = 0; i < d; ++i) vold[i] = u[i];
        for (int i = 0; i < m; ++i) {
          const T* q = f.Q + i * d;
          T c = vdot(q, vold, d);
          r[i] = r[i] + c;
          for (int l = 0; l < d; ++l) u[l] = u[l] - c * q[l];
        }
        T nnew = snorm<P>(u, d);
        if (nnew >= nold * eta || !(P::finite(nnew) && nnew > 0)) break;
        if (passes == 4) {
          passes = 5;
          break;
        }
        nold = nnew;
      }
    }
    T g = snorm<P>(u, d);
    for (int i = 0; i < k; ++i) f.R[m * k + i] = i == m ? g : r[i];
    if (passes > 4) {
      f.status = 2;
      f.column = m + 1;
    } else if (!(P::finite(g) && g > 0)) {
      f.status = 1;
      f.column = m + 1;
    }
    if (f.status) return;
    for (int l = 0; l < d; ++l) f.Q[m * d + l] = u[l] / g;
    if (passes > f.count) f.count = passes;
  }
}

// A (d rows, nc columns, column-major) <- (I - 2 u u') A
template <class P, class T>
static void reflect(T* A, const T* u, int d, int nc) {
  for (int l = 0; l < nc; ++l) {
    T w = vdot(u, A + l * d, d);
    for (int i = 0; i < d; ++i) A[l * d + i] = A[l * d + i] - ((T)2 * u[i]) * w;
  }
}

template <class P, class T>
static int sweep(T* A, T* U, int d, int nc, int nref) {
  int count = 0;
  for (int j = 0; j < nref; ++j) {
    T a[8];
    for (int i = 0; i < d; ++i) a[i] = i >= j ? A[j * d + i] : (T)0;
    T c = snorm<P>(a, d), rho;
    if (P::finite(c) && c > 0) {
      T s = a[j] < 0 ? (T)-1 : (T)1;
      a[j] = a[j] + s * c;
      T nv = snorm<P>(a, d);
      for (int i = 0; i < d; ++i) a[i] = a[i] / nv;
      rho = -s * c;
      reflect<P>(A, a, d, nc);
      for (int i = 0; i < d; ++i) U[j * d + i] = a[i];
      count++;
    } else {
      rho = c;
    }
    A[j * d + j] = rho;
  }
  return count;
}

template <class P, class T>
static void finish(Fact<P, T>& f, const T* A, int ld, int d, int k) {
  T sg[8];
  for (int j = 0; j < k; ++j) sg[j] = A[j * ld + j] < 0 ? (T)-1 : (T)1;
  for (int j = 0; j < k; ++j)
    for (int i = 0; i < k; ++i) f.R[j * k + i] = i > j ? (T)0 : sg[i] * A[j * ld + i];
  for (int j = 0; j < k; ++j)
    for (int i = 0; i < d; ++i) f.Q[j * d + i] = f.Q[j * d + i] * sg[j];
  f.status = 0;
  f.column = 0;
  for (int j = 0; j < k; ++j) {
    T p = f.R[j * k + j];
    if (!(P::finite(p) && p > 0)) {
      f.status = 1;
      f.column = j + 1;
      return;
    }
  }
}

template <class P, class T>
static void qr_hh(int alg, const T* W, int d, int k, Fact<P, T>& f) {
  int nref = k < d ? k : k - 1;
  T U[64];
  for (int i = 0; i < d * k; ++i) U[i] = 0;
  if (alg == CL_QR_HOUSEHOLDER) {
    T A[64];
    for (int i = 0; i < d * k; ++i) A[i] = W[i];
    f.count = sweep<P>(A, U, d, k, nref);
    for (int j = 0; j < k; ++j)
      for (int i = 0; i < d; ++i) f.Q[j * d + i] = i == j ? (T)1 : (T)0;
    for (int j = nref - 1; j >= 0; --j) reflect<P>(f.Q, U + j * d, d, k);
    finish<P>(f, A, d, d, k);
  } else {  // augmented [W, I]
    int nc = k + d;
    T A[128];
    for (int l = 0; l < d * nc; ++l) {
      int c = l / d, r = l % d;
      A[l] = l < d * k ? W[l] : (r == c - k ? (T)1 : (T)0);
    }
    f.count = sweep<P>(A, U, d, nc, nref);
    for (int b = 0; b < k; ++b)
      for (int a = 0; a < d; ++a) f.Q[b * d + a] = A[(k + a) * d + b];
    finish<P>(f, A, d, d, k);
  }
}

template <class P, class T>
static void qr_givens(const T* W, int d, int k, Fact<P, T>& f) {
  int wd = k + d, J = k < d - 1 ? k : d - 1;
  T rows[8][16];
  for (int i = 0; i < d; ++i)
    for (int l = 0; l < wd; ++l) rows[i][l] = l < k ? W[l * d + i] : (l - k == i ? (T)1 : (T)0);
  int count = 0;
  for (int j = 0; j < J; ++j)
    for (int i = d - 1; i >= j + 1; --i) {
      T ab[2] = {rows[i - 1][j], rows[i][j]};
      T rho = snorm<P>(ab, 2);
      if (!(P::finite(rho) && rho > 0)) continue;
      T c = ab[0] / rho, s = ab[1] / rho;
      for (int l = 0; l < wd; ++l) {
        T ra = rows[i - 1][l], rb = rows[i][l];
        rows[i - 1][l] = c * ra + s * rb;
        rows[i][l] = c * rb - s * ra;
      }
      rows[i - 1][j] = rho;
      rows[i][j] = 0;
      count++;
    }
  T R0[64];
  for (int j = 0; j < k; ++j)
    for (int i = 0; i < k; ++i) R0[j * k + i] = i > j ? (T)0 : rows[i][j];
  for (int b = 0; b < k; ++b)
    for (int a = 0; a < d; ++a) f.Q[b * d + a] = rows[b][k + a];
  f.count = count;
  finish<P>(f, R0, k, d, k);
}

template <class P, class T>
static T window_gw(const T* R0, int k) {
  T R[256], Ri[256];
  T m = P::abs(R0[0]);
  for (int i = 1; i < k * k; ++i) m = jmax<P>(m, P::abs(R0[i]));
  bool sc = (m > (T)1e100 || m < (T)1e-100) && P::finite(m) && m > 0;
  for (int i = 0; i < k * k; ++i) R[i] = sc ? R0[i] / m : R0[i];
  for (int i = 0; i < k * k; ++i) Ri[i] = 0;
  for (int j = 0; j < k; ++j) {
    Ri[j * k + j] = (T)1 / R[j * k + j];
    for (int i = j - 1; i >= 0; --i) {
      T s = 0;
      for (int l = i + 1; l <= j; ++l) s += R[l * k + i] * Ri[j * k + l];
      Ri[j * k + i] = -s / R[i * k + i];
    }
  }
  T a = R[0] * R[0], b = Ri[0] * Ri[0];
  for (int i = 1; i < k * k; ++i) {
    a = a + R[i] * R[i];
    b = b + Ri[i] * Ri[i];
  }
  return P::sqrt(a) * P::sqrt(b);
}

template <class P, class T>
static T orth_loss(const T* Q, int d, int k) {
  T mx = 0;
  bool first = true;
  for (int j = 0; j < k; ++j)
    for (int i = 0; i < k; ++i) {
      T acc = Q[i * d] * Q[j * d];
      for (int l = 1; l < d; ++l) acc = P::fma(Q[i * d + l], Q[j * d + l], acc);
      T e = P::abs(i == j ? acc - (T)1 : acc - (T)0);
      mx = first ? e : jmax<P>(mx, e);
      first = false;
    }
  return mx;
}

// ------------------------------------------------------------------ reducers
template <class P, class T>
static void red_qr(const cl_reducer& r, QRSt<T>& st, int64_t n, T* W, int d, int k) {
  if (st.code) return;
  Fact<P, T> f;
  if (r.alg == CL_QR_HOUSEHOLDER || r.alg == CL_QR_AUGMENTED) qr_hh<P>(r.alg, W, d, k, f);
  else if (r.alg == CL_QR_GIVENS) qr_givens<P>(W, d, k, f);
  else qr_gs<P>(r.alg, W, d, k, f);
  auto fail = [&](int code, int64_t col, T v, T thr) {
    st.code = (int8_t)code;
    st.column = col;
    st.value = v;
    st.threshold = thr;
    st.step = n;
  };
  if (f.status) {
    if (f.status == 2) fail(2, f.column, (T)5, (T)4);
    else {
      T piv = f.R[(f.column - 1) * k + f.column - 1];
      fail(1, f.column, piv, P::finite(piv) ? (T)0 : (T)INFINITY);
    }
    return;
  }
  T gw = window_gw<P>(f.R, k);
  bool cond = r.alg == CL_QR_CGS || r.alg == CL_QR_MGS || r.alg == CL_QR_RGS;
  if (cond && !(gw * P::eps() < (T)0.5)) {
    fail(3, 0, gw * P::eps(), (T)0.5);
    return;
  }
  T orth = k >= 2 ? orth_loss<P>(f.Q, d, k) : (T)0;
  T g = (r.alg == CL_QR_CGS || r.alg == CL_QR_MGS) ? gw : (T)1;
  T thr = (T)r.tol * P::eps() * (T)k * g;
  if (k >= 2 && !(orth <= thr)) {
    fail(4, 0, orth, thr);
    return;
  }
  for (int i = 0; i < k; ++i) st.sums[i] = acc_add<P>(st.sums[i], P::log(f.R[i * k + i]));
  T G = k == 1 ? (T)1 : (k == 2 ? (gw + P::sqrt(jmax<P>(gw * gw - (T)4, (T)0))) / (T)2 : gw);
  st.n = n;
  st.G_max = jmax<P>(st.G_max, G);
  st.orth_max = jmax<P>(st.orth_max, orth);
  if (f.count > st.count_max) st.count_max = f.count;
  st.column = 0;
  for (int i = 0; i < d * k; ++i) W[i] = f.Q[i];
}

template <class P, class T>
static MSums<T> megno_add(MSums<T> m, T a, int64_t n) {
  T tn = (T)n;
  Acc<T> R1 = acc_add<P>(m.R1, tn * a), R2 = acc_add<P>(m.R2, tn * tn * a);
  MSums<T> o;
  o.s = acc_add<P>(m.s, a);
  o.R1 = R1;
  o.R2 = R2;
  o.Q = acc_add<P>(m.Q, (T)2 * acc_total(R1) / tn);
  o.Q2 = acc_add<P>(m.Q2, (T)(g_defect == DEFECT_MEGNO_Q2 ? 2 : 3) * acc_total(R2));
  return o;
}

// Column norms of a d x 2 frame, with the zero/non-finite rule shared by MEGNO, SALI, FLI.
template <class P, class T>
static bool two_norms(const T* W, int d, T& a1, T& a2, int8_t& col, T& val) {
  a1 = snorm<P>(W, d);
  a2 = snorm<P>(W + d, d);
  bool ok1 = P::finite(a1) && a1 > 0, ok2 = P::finite(a2) && a2 > 0;
  if (ok1 && ok2) return true;
  col = ok1 ? 2 : 1;
  val = ok1 ? a2 : a1;
  return false;
}

template <class P, class T>
static void red_megno(const cl_reducer& r, MegSt<T>& st, int64_t n, T* W, int d) {
  if (st.code) return;
  T a1, a2;
  if (!two_norms<P>(W, d, a1, a2, st.column, st.value)) {
    st.step = n;
    st.code = 1;
    return;
  }
  st.c[0] = megno_add<P>(st.c[0], P::log(a1), n);
  st.c[1] = megno_add<P>(st.c[1], P::log(a2), n);
  st.n = n;
  st.K += 1;
  if (r.replace)
    for (int i = 0; i < d; ++i) {
      W[i] = W[i] / a1;
      W[d + i] = W[d + i] / a2;
    }
}

// RLI on a pair (rli.jl:77-110): both columns' scaled norms and logs (Benettin's operations), the off-block check,
// DL = |(sB.s - sA.s) + (sB.c - sA.c)| / n from the new sums into the compensated Q, the replacement.
template <class P, class T>
static void red_rli(const cl_reducer& r, RliSt<T>& st, int64_t n, T* W, int d) {
  if (st.code) return;
  T a1 = snorm<P>(W, d), a2 = snorm<P>(W + d, d);
  const bool ok1 = P::finite(a1) && a1 > 0, ok2 = P::finite(a2) && a2 > 0;
  if (!(ok1 && ok2)) {
    st.column = ok1 ? 2 : 1;
    st.value = ok1 ? a2 : a1;
    st.step = n;
    st.code = 1;
    return;
  }
  const int h = d / 2;
  T m1 = 0, m2 = 0;
  for (int i = 0; i < h; ++i) {
    m1 = jmax<P>(m1, P::abs(W[h + i]));
    m2 = jmax<P>(m2, P::abs(W[d + i]));
  }
  if (!(m1 == 0 && m2 == 0)) {
    st.column = m1 == 0 ? 2 : 1;
    st.value = m1 == 0 ? m2 : m1;
    st.step = n;
    st.code = 2;
    return;
  }
  st.sA = acc_add<P>(st.sA, P::log(a1));
  st.sB = acc_add<P>(st.sB, P::log(a2));
  st.Q = acc_add<P>(st.Q, P::abs((st.sB.s - st.sA.s) + (st.sB.c - st.sA.c)) / (T)n);
  st.n = n;
  st.K += 1;
  if (r.replace)
    for (int i = 0; i < d; ++i) {
      W[i] = W[i] / a1;
      W[d + i] = W[d + i] / a2;
    }
}

// StretchingNumbers (stretching.jl:119-138): the column logs, the cell of z = a / tau in the orbit's workspace
// (spectrum_cell, :104-108; column j at offset (j - 1)(bins + 2)), the sums of a and a^2, the replacement.
template <class P, class T>
static int64_t sn_cell(T z, T lo, T hi, int64_t B) {
  if (z < lo) return 0;
  if (z > hi) return B + 1;
  const int64_t c = (int64_t)P::floor((z - lo) / ((hi - lo) / (T)B)) + 1;
  return c < B ? c : B;
}
template <class P, class T>
static void red_stretch(const cl_reducer& r, StretchSt<T>& st, int64_t n, T* W, int d, uint32_t* ws) {
  if (st.code) return;
  T a1, a2;
  if (!two_norms<P>(W, d, a1, a2, st.column, st.value)) {
    st.step = n;
    st.code = 1;
    return;
  }
  const int64_t K = st.K + 1;
  if (K > 1000000000) {
    st.value = (T)K;
    st.step = n;
    st.code = 2;
    st.column = 0;
    return;
  }
  const T tau = (T)(n / K), lo = (T)r.lo, hi = (T)r.hi;
  const int64_t B = r.bins;
  const T l1 = P::log(a1), l2 = P::log(a2);
  ws[sn_cell<P>(l1 / tau, lo, hi, B)] += 1u;
  ws[B + 2 + sn_cell<P>(l2 / tau, lo, hi, B)] += 1u;
  st.s1 = acc_add<P>(st.s1, l1);
  st.s2 = acc_add<P>(st.s2, l2);
  st.q1 = acc_add<P>(st.q1, l1 * l1);
  st.q2 = acc_add<P>(st.q2, l2 * l2);
  st.n = n;
  st.K = K;
  if (r.replace)
    for (int i = 0; i < d; ++i) {
      W[i] = W[i] / a1;
      W[d + i] = W[d + i] / a2;
    }
}


// WeightedBirkhoff (wbirkhoff_reducer.jl:51-66, the update _wb_sample of wbirkhoff.jl:118-123): the observables of
// the point, the first non-finite one fails the orbit; sample j = handoffs goes into window 1 or 2 of every horizon
// that has not yet taken its 2N samples, with the weight exp(-1 / (t^p (1 - t)^p)), t = min(jj, N - jj) / N
// (birkhoff_weight, naff_reducer.jl:13-17; the powers by plain multiplication).
template <class P, class T>
static T wb_obs(const cl_birkhoff& b, int k, const T* x) {
  if (b.kind[k] == CL_OBS_COS) return P::cos((T)b.b[k] * x[b.a[k]]);
  if (b.kind[k] == CL_OBS_SIN) return P::sin((T)b.b[k] * x[b.a[k]]);
  return x[b.a[k]] + (T)2 * P::pi() * x[b.b[k]];
}
template <class P, class T>
static T wb_w(int64_t jj, int64_t N, int p) {
  const int64_t mn = jj < N - jj ? jj : N - jj;
  const T t = (T)mn / (T)N;
  if (t <= 0 || t >= 1) return (T)0;
  auto pw = [&](T u) {
    const T u2 = u * u;
    return p == 1 ? u : (p == 2 ? u2 : (p == 3 ? u2 * u : u2 * u2));
  };
  return P::exp((T)-1 / (pw(t) * pw((T)1 - t)));
}

// WeightedPhaseAdvance (naff_reducer.jl:199-218): the sample (planar (x[a], x[b]); lifted-angle (x[a], x[b] or 0)),
// a non-finite sample fails the orbit (code 1), the first handoff stores it, each later one adds its increment
// (phase_increments' principal argument, or (q' - q) + 2 pi (m' - m)) with weight birkhoff_weight at the mirrored
// index to the window of each horizon that holds increment j = handoffs - 1; increments with |Delta| > 0.9 pi counted.
template <class P>
static void pa_unit_scale(double re, double im, double& ore, double& oim) {  // naff_reducer.jl:40-45
  const double m = std::max(std::fabs(re), std::fabs(im));
  int e = 0;
  if (std::isfinite(m) && m > 0) {
    uint64_t u;
    memcpy(&u, &m, 8);
    e = (int)((u >> 52) & 0x7ff) - 1023;
    e = e < -1000 ? -1000 : (e > 1000 ? 1000 : e);
  }
  ore = std::ldexp(re, -e);
  oim = std::ldexp(im, -e);
}
template <class P>
static double pa_phase_inc(double pre, double pim, double re, double im) {  // naff_reducer.jl:47-56
  if ((pre == 0 && pim == 0) || (re == 0 && im == 0)) return NAN;
  // r = z conj(zprev), Julia's complex product
  double rr = re * pre - im * (-pim), ri = re * (-pim) + im * pre;
  const double s = std::max(std::fabs(rr), std::fabs(ri));
  if (!(s > 1e-280 && s < 1e280)) {
    double a, b, c, d;
    pa_unit_scale<P>(re, im, a, b);
    pa_unit_scale<P>(pre, pim, c, d);
    rr = a * c - b * (-d);
    ri = a * (-d) + b * c;
  }
  return cl_atan2(ri, rr);
}
template <class P, class T>
static void red_pa(const cl_reducer& r, PaSt<T>& st, int64_t n, const T* x) {
  if (st.code) return;
  const cl_birkhoff& b = *r.birkhoff;
  const bool planar = b.kind[0] == CL_OBS_PLANAR;
  const T a0 = x[b.a[0]], a1 = (!planar && b.b[0] < 0) ? (T)0 : x[b.b[0]];
  if (!(P::finite(a0) && P::finite(a1))) {
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
  const T tp = (T)2 * (T)3.141592653589793;
  const T d = planar ? (T)pa_phase_inc<P>((double)st.prev1, (double)st.prev2, (double)a0, (double)a1)
                     : (a0 - st.prev1) + tp * (a1 - st.prev2);
  const int64_t j = (int64_t)s - 1;
  int64_t wmax = 0;
  for (int h = 0; h < b.H; ++h) {
    const int64_t N = b.window[h];
    wmax = N > wmax ? N : wmax;
    if (j < 2 * N) {
      const T w = wb_w<P, T>(j < N ? j : j - N, N, b.order);
      if (w != 0) {
        if (j < N) {
          st.A1[h] = acc_add<P>(st.A1[h], w * d);
          st.B1[h] = acc_add<P>(st.B1[h], w);
        } else {
          st.A2[h] = acc_add<P>(st.A2[h], w * d);
          st.B2[h] = acc_add<P>(st.B2[h], w);
        }
      }
    }
  }
  const bool flagged = planar && P::abs(d) > (T)9 * (T)3.141592653589793 / (T)10 && (int64_t)s <= 2 * wmax &&
                       st.nwrap < 32767;
  st.prev1 = a0;
  st.prev2 = a1;
  st.n = n;
  st.handoffs = h1;
  if (flagged) st.nwrap += 1;
}
template <class P, class T>
static void red_wbirk(const cl_reducer& r, WBSt<T>& st, int64_t n, const T* x) {
  if (st.code) return;
  const cl_birkhoff& b = *r.birkhoff;
  int64_t mx = 0;
  for (int h = 0; h < b.H; ++h) mx = b.window[h] > mx ? b.window[h] : mx;
  if ((int64_t)st.handoffs >= 2 * mx) {
    st.n = n;
    if (st.handoffs < 2147483647) st.handoffs += 1;
    return;
  }
  T f[CL_WB_MAX_OBSERVABLES];
  for (int k = 0; k < b.K; ++k) f[k] = wb_obs<P>(b, k, x);
  for (int k = 0; k < b.K; ++k)
    if (!P::finite(f[k])) {
      st.code = (int16_t)(k + 1);
      st.n = n;
      return;
    }
  const int64_t j = st.handoffs;
  for (int h = 0; h < b.H; ++h) {
    const int64_t N = b.window[h];
    if (j >= 2 * N) continue;
    const int i = j >= N ? 1 : 0;
    const T w = wb_w<P, T>(j - i * N, N, b.order);
    if (i == 0 && w != 0) st.B[h] = acc_add<P>(st.B[h], w);
    for (int k = 0; k < b.K; ++k) {
      st.S[h][k] = jmax<P>(st.S[h][k], P::abs(f[k]));
      if (w != 0) {
        if (i == 0) st.A1[h][k] = acc_add<P>(st.A1[h][k], w * f[k]);
        else st.A2[h][k] = acc_add<P>(st.A2[h][k], w * f[k]);
      }
    }
  }
  st.n = n;
  if (j < 2147483647) st.handoffs = (int32_t)(j + 1);
}

template <class P, class T>
static void red_sali(SaliSt<T>& st, int64_t n, T* W, int d) {
  if (st.code) return;
  T a1, a2;
  if (!two_norms<P>(W, d, a1, a2, st.column, st.value)) {
    st.step = n;
    st.code = 1;
    return;
  }
  T u1[8], u2[8], dm[8], dp[8];
  for (int i = 0; i < d; ++i) {
    u1[i] = W[i] / a1;
    u2[i] = W[d + i] / a2;
    dm[i] = u1[i] - u2[i];
    dp[i] = u1[i] + u2[i];
  }
  T dmin = snorm<P>(dm, d), dplus = snorm<P>(dp, d);
  st.s1 = acc_add<P>(st.s1, P::log(a1));
  st.s2 = acc_add<P>(st.s2, P::log(a2));
  st.sali = jmin<P>(dmin, dplus);
  st.n = n;
  for (int i = 0; i < d; ++i) {
    W[i] = u1[i];
    W[d + i] = u2[i];
  }
}

template <class P, class T>
static void red_fli(const cl_reducer& r, FliSt<T>& st, int64_t n, T* W, int d) {
  if (st.code) return;
  T a1, a2;
  if (!two_norms<P>(W, d, a1, a2, st.column, st.value)) {
    st.step = n;
    st.code = 1;
    return;
  }
  Acc<T> s1 = acc_add<P>(st.s1, P::log(a1)), s2 = acc_add<P>(st.s2, P::log(a2));
  T S1 = acc_total(s1), S2 = acc_total(s2);
  if (S1 > st.f1) {
    st.f1 = S1;
    st.step1 = n;
  }
  if (S1 > st.fB) {
    st.fB = S1;
    st.stepB = n;
    st.colB = 1;
  }
  if (S2 > st.fB) {
    st.fB = S2;
    st.stepB = n;
    st.colB = 2;
  }
  st.s1 = s1;
  st.s2 = s2;
  st.n = n;
  if (r.replace)
    for (int i = 0; i < d; ++i) {
      W[i] = W[i] / a1;
      W[d + i] = W[d + i] / a2;
    }
}

template <class P, class T>
static void red_ofli(const cl_reducer& r, OfliSt<T>& st, int64_t n, const T* fld, T* W, int d, int k) {
  if (st.code) return;
  T fn = snorm<P>(fld, d);
  if (!(P::finite(fn) && fn > P::floatmin())) {
    st.value = fn;
    st.step = n;
    st.code = 2;
    st.column = 0;
    return;
  }
  T u[8], al[8], be[8];
  for (int i = 0; i < d; ++i) u[i] = fld[i] / fn;
  for (int j = 0; j < k; ++j) {
    const T* w = W + j * d;
    al[j] = snorm<P>(w, d);
    T dw = vdot(u, w, d), v[8];
    for (int i = 0; i < d; ++i) v[i] = w[i] - dw * u[i];
    be[j] = g_defect == DEFECT_OFLI_NOPROJ ? al[j] : snorm<P>(v, d);
  }
  for (int j = 0; j < k; ++j)
    if (!(P::finite(al[j]) && al[j] > 0)) {
      st.value = al[j];
      st.step = n;
      st.code = 1;
      st.column = (int8_t)(j + 1);
      return;
    }
  Acc<T> na[8];
  for (int i = 0; i < k; ++i) na[i] = acc_add<P>(st.acc[i], P::log(al[i]));
  T inst = (T)-INFINITY;
  for (int i = 0; i < k; ++i) {
    T g = be[i] > 0 ? acc_total(st.acc[i]) + P::log(be[i]) : (T)-INFINITY;
    T sn = be[i] / al[i], S = acc_total(na[i]);
    st.sinmin = jmin<P>(st.sinmin, sn);
    st.nminf += be[i] > 0 ? 0 : 1;
    inst = jmax<P>(inst, g);
    if (g > st.gB) {
      st.gB = g;
      st.stepB = n;
      st.colB = (int8_t)(i + 1);
      st.sinB = sn;
    }
    if (i == 0 && g > st.g1) {
      st.g1 = g;
      st.step1 = n;
      st.sin1 = sn;
    }
    if (i == 0 && S > st.f1) st.f1 = S;
    if (S > st.fB) st.fB = S;
  }
  for (int i = 0; i < k; ++i) st.acc[i] = na[i];
  st.inst = inst;
  st.n = n;
  if (r.replace)
    for (int j = 0; j < k; ++j)
      for (int i = 0; i < d; ++i) W[j * d + i] = W[j * d + i] / al[j];
}

template <class P, class T>
static void red_gali(GaliSt<T>& st, int64_t n, T* W, int d, int k) {
  if (st.code) return;
  T a[8];
  for (int j = 0; j < k; ++j) a[j] = snorm<P>(W + j * d, d);
  for (int j = 0; j < k; ++j)
    if (!(P::finite(a[j]) && a[j] >= P::floatmin())) {
      st.value = a[j];
      st.step = n;
      st.code = 1;
      st.column = j + 1;
      return;
    }
  for (int i = 0; i < k; ++i) {
    Acc<T> q;
    q.s = st.s[i];
    q.c = st.c[i];
    q = acc_add<P>(q, P::log(a[i]));
    st.s[i] = q.s;
    st.c[i] = q.c;
  }
  T U[64];
  for (int j = 0; j < k; ++j)
    for (int i = 0; i < d; ++i) {
      U[j * d + i] = g_defect == DEFECT_GALI_RAW ? W[j * d + i] : W[j * d + i] / a[j];
      W[j * d + i] = U[j * d + i];
    }
  T g = 1;
  for (int j = 0; j < k; ++j) {
    for (int pass = 0; pass < 2; ++pass)
      for (int p = 0; p < j; ++p) {
        T c = 0;
        for (int i = 0; i < d; ++i) c += U[p * d + i] * U[j * d + i];
        for (int i = 0; i < d; ++i) U[j * d + i] -= c * U[p * d + i];
      }
    T rr = snorm<P>(U + j * d, d);
    T s = rr == 0 ? (T)1 : rr;
    for (int i = 0; i < d; ++i) U[j * d + i] /= s;
    if (j > 0) {
      if (!(g_defect == DEFECT_GALI_LASTFACTOR && j == k - 1)) g *= rr;
      st.gali[j - 1] = g;
    }
  }
  st.n = n;
}

template <class P, class T>
static void red_lognorm(LogNormSt<T>& st, T* W, int d, int k) {
  for (int j = 0; j < k; ++j) {
    T s = W[j * d] * W[j * d];
    for (int i = 1; i < d; ++i) s = s + W[j * d + i] * W[j * d + i];
    s = P::sqrt(s);
    st.a[j] = acc_add<P>(st.a[j], P::log(s));
    for (int i = 0; i < d; ++i) W[j * d + i] = W[j * d + i] / s;
  }
}

template <class T>
static bool red_failed(const cl_reducer& r, const RState<T>& s) {
  switch (r.kind) {
    case CL_RED_MONITOR: return s.mon.code != 0;
    case CL_RED_RUNMONITOR: return false;
    case CL_RED_STRETCHING: return s.sn.code != 0;
    case CL_RED_BIRKHOFF: return s.wb.code != 0;
    case CL_RED_QR: return s.qr.code != 0;
    case CL_RED_BENETTIN: return s.ben.code != 0;
    case CL_RED_MEGNO: return s.meg.code != 0;
    case CL_RED_SALI: return s.sali.code != 0;
    case CL_RED_RLI: return s.rli.code != 0;
    case CL_RED_PHASE_ADVANCE: return s.pa.code != 0;
    case CL_RED_FLI: return s.fli.code != 0;
    case CL_RED_OFLI: return s.ofli.code != 0;
    case CL_RED_GALI: return s.gali.code != 0;
  }
  return false;
}

// ------------------------------------------------------------------ systems (d = 2 maps)
#define OR_SYS_LINEAR 100  // test-only linear map x' = A x, tangent A (test/tangent-core/tc_helpers.jl:8-10,26-28)

// Fixed-point iterations of the Gauss stage equations, counted per worker thread and summed at the end of
// a run (the work census of the Henon-Heiles workload is proportional to them).
static thread_local int64_t tl_gauss_iters = 0, tl_gauss_steps = 0;
static std::atomic<int64_t> g_gauss_iters(0), g_gauss_steps(0);

template <class P, class T>
struct MapSys {
  int kind;
  T K;
  int s;     // state length
  T A[64];   // linear map, column-major dl x dl (kind OR_SYS_LINEAR)
  int dl = 2;  // dimension of the linear map (2..8)
  bool f32 = false;  // FP32-tangent element type: frames carried in Float32 (standard maps)
  // tangent at x applied to W (d x k), explicit-fma product; then the step
  void step(T* x, T* const* W, int nf, int k) const {
    if (kind == OR_SYS_LINEAR) {
      // fused_product (product.jl:9-20): C[i,j] = A[i,1] W[1,j], then fma(A[i,q], W[q,j], C[i,j]) for q = 2..d
      for (int f = 0; f < nf; ++f)
        for (int j = 0; j < k; ++j) {
          T w[8], c[8];
          for (int q = 0; q < dl; ++q) w[q] = W[f][dl * j + q];
          for (int i = 0; i < dl; ++i) {
            c[i] = A[i] * w[0];
            for (int q = 1; q < dl; ++q) c[i] = P::fma(A[q * dl + i], w[q], c[i]);
          }
          for (int i = 0; i < dl; ++i) W[f][dl * j + i] = c[i];
        }
      // x' = A x, SMatrix * SVector order: A[i,1] x1 + A[i,2] x2 + ...
      T xo[8], xn[8];
      for (int q = 0; q < dl; ++q) xo[q] = x[q];
      for (int i = 0; i < dl; ++i) {
        xn[i] = A[i] * xo[0];
        for (int q = 1; q < dl; ++q) xn[i] = xn[i] + A[q * dl + i] * xo[q];
      }
      for (int i = 0; i < dl; ++i) x[i] = xn[i];
      return;
    }
    T sn = P::sin(x[0]), cs = P::cos(x[0]);
    const T Kx = kind == CL_SYS_STANDARD_MAP_K ? x[2] : K;  // the per-orbit K rides in the state
    T c = Kx * cs, m11 = (T)1 + c;
    if (f32) {
      // FP32 tangents: the entries formed in FP64 and rounded once to Float32, the explicit-fma product in Float32
      const float m11f = (float)m11, cf = (float)c;
      for (int f = 0; f < nf; ++f)
        for (int j = 0; j < k; ++j) {
          const float w0 = (float)W[f][2 * j], w1 = (float)W[f][2 * j + 1];
          W[f][2 * j] = (T)fmaf(1.0f, w1, m11f * w0);
          W[f][2 * j + 1] = (T)fmaf(1.0f, w1, cf * w0);
        }
    } else {
      for (int f = 0; f < nf; ++f)
        for (int j = 0; j < k; ++j) {
          T w0 = W[f][2 * j], w1 = W[f][2 * j + 1];
          W[f][2 * j] = P::fma((T)1, w1, m11 * w0);
          W[f][2 * j + 1] = P::fma((T)1, w1, c * w0);
        }
    }
    if (kind == CL_SYS_STANDARD_MAP || kind == CL_SYS_STANDARD_MAP_K) {
      T p = x[1] + Kx * sn;
      x[0] = x[0] + p;
      x[1] = p;
    } else {
      T tp = (T)2 * P::pi();
      auto red = [&](T u, T& rr, T& a) {
        a = P::floor(u / tp);
        rr = u - tp * a;
        if (rr < 0) {
          rr += tp;
          a -= 1;
        }
        if (rr >= tp) {
          rr -= tp;
          a += 1;
        }
      };
      T p, a, xx, b;
      red(x[1] + K * sn, p, a);
      T nn = x[3] + a;
      red(x[0] + p, xx, b);
      T mm = x[2] + nn + b;
      T lim = (T)9007199254740992.0;
      if (P::abs(mm) < lim && P::abs(nn) < lim) {
        x[0] = xx;
        x[1] = p;
        x[2] = mm;
        x[3] = nn;
      } else {
        T nan = (T)NAN;
        x[0] = x[1] = x[2] = x[3] = nan;
      }
    }
  }
  int jd() const { return kind == OR_SYS_LINEAR ? dl : 2; }  // the size of the one-step tangent the monitors read
  // the one-step tangent at x, column-major: [[1 + K cos x, 1], [K cos x, 1]] (standard_map.jl:30-34)
  void jac(const T* x, T* J) const {
    if (kind == OR_SYS_LINEAR) {
      for (int i = 0; i < dl * dl; ++i) J[i] = A[i];
      return;
    }
    const T Kx = kind == CL_SYS_STANDARD_MAP_K ? x[2] : K;
    T c = Kx * P::cos(x[0]);
    J[0] = (T)1 + c;
    J[1] = c;
    J[2] = (T)1;
    J[3] = (T)1;
  }
  // map displacement field F(x) - x at x (OFLI direction on a map); (A - I) x for the linear map, formed with the
  // diagonal reduced first so that a shear [[I, hI], [0, I]] gives exactly (h p, 0), the free particle's direction
  void field(const T* x, T* f) const {
    if (kind == OR_SYS_LINEAR) {
      for (int i = 0; i < dl; ++i) {
        T ax = (A[i] - (i == 0 ? (T)1 : (T)0)) * x[0];
        for (int q = 1; q < dl; ++q) ax = ax + (A[q * dl + i] - (q == i ? (T)1 : (T)0)) * x[q];
        f[i] = ax;
      }
      return;
    }
    T ks = (kind == CL_SYS_STANDARD_MAP_K ? x[2] : K) * P::sin(x[0]);
    f[0] = x[1] + ks;
    f[1] = ks;
  }
};


// The product of a map with itself (src/Systems/pair.jl:1-68): state (A; B), each half stepped by the base map with
// its own block of the frame's rows (1..2 for A, 3..4 for B), the monitor reading orbit A (run_indicator_record.jl:6-10).
template <class P, class T>
struct PairSys {
  MapSys<P, T> base;
  int s;  // the pair's state length, twice the base's
  bool f32 = false;
  int jd() const { return 2; }
  void step(T* x, T* const* W, int nf, int k) const {
    T wa[2][32], wb[2][32];
    T* pa[2] = {wa[0], wa[1]};
    T* pb[2] = {wb[0], wb[1]};
    for (int f = 0; f < nf; ++f)
      for (int j = 0; j < k; ++j)
        for (int r = 0; r < 2; ++r) {
          wa[f][2 * j + r] = W[f][4 * j + r];
          wb[f][2 * j + r] = W[f][4 * j + 2 + r];
        }
    base.step(x, pa, nf, k);
    base.step(x + base.s, pb, nf, k);
    for (int f = 0; f < nf; ++f)
      for (int j = 0; j < k; ++j)
        for (int r = 0; r < 2; ++r) {
          W[f][4 * j + r] = wa[f][2 * j + r];
          W[f][4 * j + 2 + r] = wb[f][2 * j + r];
        }
  }
  void jac(const T* x, T* J) const { base.jac(x, J); }
  void field(const T* x, T* f) const { base.field(x, f); }
};

// ------------------------------------------------------------------ Henon-Heiles under Gauss s = 3
// IntegratorMap(Gauss(Float64, 3; tol), HenonHeilesFlow(), h) (src/Integrators/integrator_map.jl:12-23):
// fixed-point stage iteration with compensated sums (src/Integrators/gauss.jl:90-141), increment
// h sum b_i F_i, and the one-step Jacobian of the stage equations. The reference solves the 12 x 12
// system with LAPACK (gauss.jl:176-192); this restatement eliminates the position blocks first
// (J_j = [[0, I], [-H_j, 0]]) and solves the remaining 6 x 6 system with partial pivoting, the order
// the kernels use, so the two agree bit for bit.
static const double OR_GA_HI[3][3] = {{0x1.1c71c71c71c72p-3, -0x1.26b88a4e09a62p-5, 0x1.40c7cef225974p-7},
                                      {0x1.337831ea8a881p-2, 0x1.c71c71c71c71cp-3, -0x1.7066ace18c0fbp-6},
                                      {0x1.126b88a4e09a6p-2, 0x1.ebf38310dda69p-2, 0x1.1c71c71c71c72p-3}};
static const double OR_GA_LO[3][3] = {{-0x1.c71c71c71c71cp-58, -0x1.f9a1c81ea4ba2p-60, 0x1.94971b6f94297p-61},
                                      {0x1.d03218eda4366p-56, 0x1.c71c71c71c71cp-57, 0x1.87f5c5d9b2176p-61},
                                      {0x1.1f9a1c81ea4bap-57, -0x1.194971b6f9429p-56, -0x1.c71c71c71c71cp-58}};
static const double OR_GB_HI[3] = {0x1.1c71c71c71c72p-2, 0x1.c71c71c71c71cp-2, 0x1.1c71c71c71c72p-2};
static const double OR_GB_LO[3] = {-0x1.c71c71c71c71cp-57, 0x1.c71c71c71c71cp-56, -0x1.c71c71c71c71cp-57};

template <class P, class T>
struct HHSys {
  int s = 4, d = 4;
  bool f32 = false;
  bool comp = false;  // compensated state (z, e), s = 8: the element type CL_ELT_DD
  int jd() const { return 4; }
  T h, tol;  // tol NaN: 8 eps max(1, |Y|)
  int maxiter;
  static void vfield(const T* z, T* f) {
    f[0] = z[2];
    f[1] = z[3];
    f[2] = -(z[0] + (T)2 * z[0] * z[1]);
    f[3] = -((z[1] + z[0] * z[0]) - z[1] * z[1]);
  }
  // sum_j (wh_j + wl_j) F_j, TwoProd by fma and TwoSum per term
  static T cdot(const double* wh, const double* wl, const T* F) {
    T s = 0, c = 0;
    for (int j = 0; j < 3; ++j)
      for (int t = 0; t < 2; ++t) {
        T w = (T)(t ? wl[j] : wh[j]);
        T p = w * F[j];
        T pe = P::fma(w, F[j], -p);
        T sn = s + p, bb = sn - s;
        T se = (s - (sn - bb)) + (p - bb);
        s = sn;
        c = c + (se + pe);
      }
    return s + c;
  }
  // The stage solve of one step at x: the increment and the one-step Jacobian J (row-major), NaN when the
  // solve failed.
  void solve(const T* x, T (&J)[4][4], T (&inc)[4], bool& ok_out) const {
    T f0[4], Y[3][4], F[3][4];
    vfield(x, f0);
    for (int i = 0; i < 3; ++i) {
      T rs = h * (((T)OR_GA_HI[i][0] + (T)OR_GA_HI[i][1]) + (T)OR_GA_HI[i][2]);
      for (int c = 0; c < 4; ++c) {
        Y[i][c] = x[c] + rs * f0[c];
        F[i][c] = f0[c];
      }
    }
    T delta = 0, prev = (T)INFINITY;
    bool stopped = false;
    int it = 0;
    while (it < maxiter && !stopped) {
      ++it;
      for (int i = 0; i < 3; ++i) vfield(Y[i], F[i]);
      T dm = 0;
      for (int i = 0; i < 3; ++i) {
        T di = 0;
        for (int c = 0; c < 4; ++c) {
          T Fc[3] = {F[0][c], F[1][c], F[2][c]};
          T yn = x[c] + ((comp ? x[4 + c] : (T)0) + h * cdot(OR_GA_HI[i], OR_GA_LO[i], Fc));
          T e = P::abs(yn - Y[i][c]);
          di = c == 0 ? e : jmax<P>(di, e);
          Y[i][c] = yn;
        }
        dm = i == 0 ? di : jmax<P>(dm, di);
      }
      delta = dm;
      if (!P::finite(delta)) break;
      stopped = delta == 0 || (it >= 2 && delta >= prev);
      prev = delta;
    }
    tl_gauss_iters += it;
    tl_gauss_steps += 1;
    T bound = tol;
    if (P::nan(tol)) {
      T mx = P::abs(Y[0][0]);
      for (int i = 0; i < 3; ++i)
        for (int c = 0; c < 4; ++c) mx = jmax<P>(mx, P::abs(Y[i][c]));
      bound = (T)8 * P::eps() * jmax<P>((T)1, mx);
    }
    bool ok = stopped && P::finite(delta) && delta <= bound;
    ok_out = ok;
    for (int c = 0; c < 4; ++c) {
      T Fc[3] = {F[0][c], F[1][c], F[2][c]};
      inc[c] = h * cdot(OR_GB_HI, OR_GB_LO, Fc);
    }
    // Jacobian of the step from the converged stages
    T ha[3][3], hb[3], H[3][2][2];
    for (int i = 0; i < 3; ++i) {
      for (int j = 0; j < 3; ++j) ha[i][j] = h * ((T)OR_GA_HI[i][j] + (T)OR_GA_LO[i][j]);
      hb[i] = h * ((T)OR_GB_HI[i] + (T)OR_GB_LO[i]);
      H[i][0][0] = (T)1 + (T)2 * Y[i][1];
      H[i][1][0] = (T)2 * Y[i][0];
      H[i][0][1] = (T)2 * Y[i][0];
      H[i][1][1] = (T)1 - (T)2 * Y[i][1];
    }
    T B[6][6], R[6][4];
    for (int i = 0; i < 3; ++i)
      for (int r = 0; r < 2; ++r) {
        for (int l = 0; l < 3; ++l)
          for (int c = 0; c < 2; ++c) {
            T v = (i == l && r == c) ? (T)1 : (T)0;
            for (int j = 0; j < 3; ++j) v = v + (ha[i][j] * ha[j][l]) * H[j][r][c];
            B[2 * i + r][2 * l + c] = v;
          }
        for (int c = 0; c < 2; ++c) {
          T v = 0;
          for (int j = 0; j < 3; ++j) v = v + ha[i][j] * H[j][r][c];
          R[2 * i + r][c] = -v;
          R[2 * i + r][2 + c] = r == c ? (T)1 : (T)0;
        }
      }
    for (int p = 0; p < 6; ++p) {
      int piv = p;
      T best = P::abs(B[p][p]);
      for (int r = p + 1; r < 6; ++r)
        if (P::abs(B[r][p]) > best) {
          best = P::abs(B[r][p]);
          piv = r;
        }
      if (piv != p) {
        for (int c = 0; c < 6; ++c) std::swap(B[p][c], B[piv][c]);
        for (int c = 0; c < 4; ++c) std::swap(R[p][c], R[piv][c]);
      }
      for (int r = p + 1; r < 6; ++r) {
        T f = B[r][p] / B[p][p];
        for (int c = p + 1; c < 6; ++c) B[r][c] = B[r][c] - f * B[p][c];
        for (int c = 0; c < 4; ++c) R[r][c] = R[r][c] - f * R[p][c];
      }
    }
    T Q[6][4];
    for (int r = 5; r >= 0; --r)
      for (int c = 0; c < 4; ++c) {
        T v = R[r][c];
        for (int t = r + 1; t < 6; ++t) v = v - B[r][t] * Q[t][c];
        Q[r][c] = v / B[r][r];
      }
    for (int r = 0; r < 4; ++r)
      for (int c = 0; c < 4; ++c) J[r][c] = r == c ? (T)1 : (T)0;
    for (int i = 0; i < 3; ++i) {
      T Pm[2][4];
      for (int r = 0; r < 2; ++r)
        for (int c = 0; c < 4; ++c) {
          T v = r == c ? (T)1 : (T)0;
          for (int j = 0; j < 3; ++j) v = v + ha[i][j] * Q[2 * j + r][c];
          Pm[r][c] = v;
        }
      for (int c = 0; c < 4; ++c)
        for (int r = 0; r < 2; ++r) {
          J[r][c] = J[r][c] + hb[i] * Q[2 * i + r][c];
          T hp = H[i][r][0] * Pm[0][c] + H[i][r][1] * Pm[1][c];
          J[2 + r][c] = J[2 + r][c] + hb[i] * (-hp);
        }
    }
    if (!ok)
      for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) J[r][c] = (T)NAN;
  }
  // The one-step tangent at x, column-major (the ResidualMonitor reads it at each handoff).
  void jac(const T* x, T* Jc) const {
    T J[4][4], inc[4];
    bool ok;
    solve(x, J, inc, ok);
    for (int c = 0; c < 4; ++c)
      for (int r = 0; r < 4; ++r) Jc[c * 4 + r] = J[r][c];
  }
  void step(T* x, T* const* W, int nf, int k) const {
    T J[4][4], inc[4];
    bool ok;
    solve(x, J, inc, ok);
    for (int f = 0; f < nf; ++f) {
      T O[16];
      for (int j = 0; j < k; ++j)
        for (int i = 0; i < 4; ++i) {
          T acc = J[i][0] * W[f][j * 4];
          for (int q = 1; q < 4; ++q) acc = P::fma(J[i][q], W[f][j * 4 + q], acc);
          O[j * 4 + i] = acc;
        }
      for (int e = 0; e < 4 * k; ++e) W[f][e] = O[e];
    }
    if (comp) {  // z, e = TwoSum(z, inc + e): the compensated update (stepper.jl:50-58)
      for (int c = 0; c < 4; ++c) {
        if (!ok) {
          x[c] = x[4 + c] = (T)NAN;
          continue;
        }
        T b = inc[c] + x[4 + c];
        T sn = x[c] + b, bb = sn - x[c];
        T err = (x[c] - (sn - bb)) + (b - bb);
        x[c] = sn;
        x[4 + c] = err;
      }
      return;
    }
    for (int c = 0; c < 4; ++c) x[c] = ok ? x[c] + inc[c] : (T)NAN;
  }
  void field(const T* x, T* f) const { vfield(x, f); }
};

// ------------------------------------------------------------------ ResidualMonitor (src/Diagnostics/monitor.jl:44-54)
// det of a column-major d x d tangent (d = 2: A[1]A[4] - A[3]A[2], the StaticArrays 2 x 2 rule; d = 4: expansion
// along the first row with 3 x 3 minors, left to right)
template <class T>
static T det3(T a00, T a01, T a02, T a10, T a11, T a12, T a20, T a21, T a22) {
  return (a00 * (a11 * a22 - a12 * a21) - a01 * (a10 * a22 - a12 * a20)) + a02 * (a10 * a21 - a11 * a20);
}
template <class T>
static T detn(const T* J, int d) {
  if (d == 2) return J[0] * J[3] - J[2] * J[1];
  auto M = [&](int r, int c) { return J[c * 4 + r]; };
  T t0 = M(0, 0) * det3(M(1, 1), M(1, 2), M(1, 3), M(2, 1), M(2, 2), M(2, 3), M(3, 1), M(3, 2), M(3, 3));
  T t1 = M(0, 1) * det3(M(1, 0), M(1, 2), M(1, 3), M(2, 0), M(2, 2), M(2, 3), M(3, 0), M(3, 2), M(3, 3));
  T t2 = M(0, 2) * det3(M(1, 0), M(1, 1), M(1, 3), M(2, 0), M(2, 1), M(2, 3), M(3, 0), M(3, 1), M(3, 3));
  T t3 = M(0, 3) * det3(M(1, 0), M(1, 1), M(1, 2), M(2, 0), M(2, 1), M(2, 2), M(3, 0), M(3, 1), M(3, 2));
  return ((t0 - t1) + t2) - t3;
}
// max |M' Omega M - Omega|, Omega = [0 I; -I 0]; products with sums left to right (symplectic_residual)
template <class P, class T>
static T symp_res(const T* J, int d) {
  const int h = d / 2;
  auto Om = [&](int r, int c) -> T { return (r < h && c == r + h) ? (T)1 : ((r >= h && c == r - h) ? (T)-1 : (T)0); };
  T A[16], mx = 0;
  for (int i = 0; i < d; ++i)
    for (int j = 0; j < d; ++j) {
      T v = J[i * d + 0] * Om(0, j);
      for (int l = 1; l < d; ++l) v = v + J[i * d + l] * Om(l, j);
      A[i * d + j] = v;  // (M' Omega)[i][j], with M'[i][l] = M[l][i] = J[i * d + l]
    }
  for (int i = 0; i < d; ++i)
    for (int j = 0; j < d; ++j) {
      T v = A[i * d + 0] * J[j * d + 0];
      for (int l = 1; l < d; ++l) v = v + A[i * d + l] * J[j * d + l];
      T e = P::abs(v - Om(i, j));
      mx = (i == 0 && j == 0) ? e : jmax<P>(mx, e);
    }
  return mx;
}
template <class P, class T, class SYS>
static void red_monitor(const cl_reducer& r, MonSt<T>& st, int64_t n, const SYS& sys, const T* x, const T* W, int d,
                        int k) {
  if (st.code) return;
  T J[64] = {};
  sys.jac(x, J);
  T mx = P::abs(J[0]);
  for (int e = 1; e < d * d; ++e) mx = jmax<P>(mx, P::abs(J[e]));
  const T dv = P::abs(detn(J, d) - (T)1);
  const T sv = symp_res<P>(J, d);
  const T ov = k >= 2 ? orth_loss<P>(W, d, k) : (T)0;
  const T one_m = (T)1 + mx;
  const T thr_d = ((T)r.tol * P::eps()) * one_m;
  const T thr_s = (((T)r.tol * P::eps()) * (T)((d / 2) * (d / 2))) * (one_m * one_m);
  const T thr_o = (T)r.tol2;
  auto bad = [](T v, T t) { return !(P::finite(v) && v <= t); };
  int code = bad(dv, thr_d) ? 1 : (bad(sv, thr_s) ? 2 : (bad(ov, thr_o) ? 3 : 0));
  st.det_max = jmax<P>(st.det_max, dv);
  st.symp_max = jmax<P>(st.symp_max, sv);
  st.orth_max = jmax<P>(st.orth_max, ov);
  st.M_max = jmax<P>(st.M_max, mx);
  st.code = (int8_t)code;
  st.value = code == 1 ? dv : (code == 2 ? sv : (code == 3 ? ov : (T)0));
  st.threshold = code == 1 ? thr_d : (code == 2 ? thr_s : (code == 3 ? thr_o : (T)0));
  st.step = n;
}

// RunMonitor (src/Experiments/run_record.jl:37-40): running maxima of |det M - 1|, the symplectic residual and
// max |M| of the one-step tangent; never fails
template <class P, class T, class SYS>
static void red_runmonitor(MonSt<T>& st, const SYS& sys, const T* x, int d) {
  T J[64] = {};
  sys.jac(x, J);
  T mx = P::abs(J[0]);
  for (int e = 1; e < d * d; ++e) mx = jmax<P>(mx, P::abs(J[e]));
  st.det_max = jmax<P>(st.det_max, P::abs(detn(J, d) - (T)1));
  st.symp_max = jmax<P>(st.symp_max, symp_res<P>(J, d));
  st.M_max = jmax<P>(st.M_max, mx);
}

// ------------------------------------------------------------------ the per-orbit body
template <class P, class T, class SYS>
static int8_t run_orbit_seg(const SYS& sys, const cl_chain& ch, int d, int k, int64_t n0, int64_t steps, int64_t N,
                            int tau, T* x, T* W0, T* W1, RState<T>* st, int64_t* hstep, uint32_t* ws = nullptr);

// N steps from fresh reducer states
template <class P, class T, class SYS>
static int8_t run_orbit(const SYS& sys, const cl_chain& ch, int d, int k, int64_t N, int tau, T* x,
                        T* W0, T* W1, RState<T>* st, int64_t* hstep, uint32_t* ws = nullptr) {
  for (int r = 0; r < ch.n; ++r) init_state(ch.r[r], st[r], k);
  if (ws) memset(ws, 0, 4 * (size_t)chain_ws(ch));
  return run_orbit_seg<P>(sys, ch, d, k, 0, N, N, tau, x, W0, W1, st, hstep, ws);
}

// One launch (orbit.jl:44-60): `steps` steps from global step n0 with the carried point, frames and reducer
// states; the final status check runs only at step N.
template <class P, class T, class SYS>
static int8_t run_orbit_seg(const SYS& sys, const cl_chain& ch, int d, int k, int64_t n0, int64_t steps, int64_t N,
                            int tau, T* x, T* W0, T* W1, RState<T>* st, int64_t* hstep, uint32_t* ws) {
  T* W[2] = {W0, W1};
  int nf = 1;
  for (int r = 0; r < ch.n; ++r)
    if (ch.r[r].frame == 1) nf = 2;
  auto status = [&]() -> int8_t {
    for (int i = 0; i < sys.s; ++i)
      if (!P::finite(x[i])) return 1;
    for (int f = 0; f < nf; ++f)
      for (int i = 0; i < d * k; ++i)
        if (!P::finite(W[f][i])) return 2;
    return 0;
  };
  int64_t n = n0;
  for (int64_t h = 0; h < steps / tau; ++h) {
    for (int t = 0; t < tau; ++t) sys.step(x, W, nf, k);
    n += tau;
    int8_t code = status();
    if (code) {
      *hstep = n;
      return code;
    }
    for (int r = 0; r < ch.n; ++r) {
      const cl_reducer& rd = ch.r[r];
      T* w = W[rd.frame];
      switch (rd.kind) {
        case CL_RED_QR: red_qr<P>(rd, st[r].qr, n, w, d, k); break;
        case CL_RED_BENETTIN: {
          BenSt<T>& b = st[r].ben;
          if (b.code) break;
          T al = snorm<P>(w, d);
          if (!(P::finite(al) && al > 0)) {
            b.code = 1;
            b.value = al;
            b.step = n;
            break;
          }
          b.sum = acc_add<P>(b.sum, P::log(al));
          b.n = n;
          for (int i = 0; i < d; ++i) w[i] = w[i] / al;
          break;
        }
        case CL_RED_MEGNO: red_megno<P>(rd, st[r].meg, n, w, d); break;
        case CL_RED_SALI: red_sali<P>(st[r].sali, n, w, d); break;
        case CL_RED_RLI: red_rli<P>(rd, st[r].rli, n, w, d); break;
        case CL_RED_FLI: red_fli<P>(rd, st[r].fli, n, w, d); break;
        case CL_RED_OFLI: {
          T f[8];
          sys.field(x, f);
          red_ofli<P>(rd, st[r].ofli, n, f, w, d, k);
          break;
        }
        case CL_RED_GALI: red_gali<P>(st[r].gali, n, w, d, k); break;
        case 8: red_lognorm<P>(st[r].ln, w, d, k); break;
        case CL_RED_MONITOR: red_monitor<P>(rd, st[r].mon, n, sys, x, w, d, k); break;
        case CL_RED_RUNMONITOR: red_runmonitor<P>(st[r].mon, sys, x, sys.jd()); break;
        case CL_RED_STRETCHING: red_stretch<P>(rd, st[r].sn, n, w, d, ws); break;
        case CL_RED_BIRKHOFF: red_wbirk<P>(rd, st[r].wb, n, x); break;
        case CL_RED_PHASE_ADVANCE: red_pa<P>(rd, st[r].pa, n, x); break;
      }
    }
    if (sys.f32)  // the frames the chain leaves are rounded back to Float32
      for (int f = 0; f < nf; ++f)
        for (int e = 0; e < d * k; ++e) W[f][e] = (T)(float)W[f][e];
    for (int r = 0; r < ch.n; ++r)
      if (red_failed(ch.r[r], st[r])) {
        *hstep = n;
        return 3;
      }
    if (ch.flags & CL_CHAIN_RESET_FRAME)  // the window kinds' constant deviation column (run_indicator_device.jl:66-80)
      for (int f = 0; f < nf; ++f)
        for (int e = 0; e < d * k; ++e) W[f][e] = (e % (d + 1) == 0) ? (T)1 : (T)0;
  }
  *hstep = n;
  return n == N ? status() : (int8_t)0;
}

template <class P>
static int propagate_all(const cl_system* sys, const cl_chain* ch, int64_t m, int s, int d, int k, int64_t N,
                         int tau, double* x, double* W, double* W2, uint8_t* states, int8_t* code, int64_t* hstep,
                         int nthreads, const double* A = nullptr, uint32_t* ws = nullptr, int elt = CL_ELT_F64) {
  typedef typename P::T T;
  const int64_t nws = chain_ws(*ch);
  const bool f32 = elt == CL_ELT_F64_F32TAN;
  MapSys<P, T> ms;
  ms.f32 = f32;
  ms.kind = sys->kind;
  ms.K = (T)sys->K;
  ms.s = s;
  ms.dl = (A && sys->kind == OR_SYS_LINEAR) ? d : 2;
  for (int i = 0; i < ms.dl * ms.dl; ++i) ms.A[i] = A ? (T)A[i] : (T)0;
  const bool pair = sys->kind == CL_SYS_PAIR_STANDARD_MAP || sys->kind == CL_SYS_PAIR_WRAPPED_STANDARD_MAP;
  PairSys<P, T> ps;
  ps.base.kind = sys->kind == CL_SYS_PAIR_STANDARD_MAP ? CL_SYS_STANDARD_MAP : CL_SYS_WRAPPED_STANDARD_MAP;
  ps.base.K = (T)sys->K;
  ps.base.s = s / 2;
  for (int i = 0; i < 4; ++i) ps.base.A[i] = (T)0;
  ps.s = s;
  HHSys<P, T> hs;
  hs.h = (T)sys->h;
  hs.tol = (T)sys->tol;
  hs.maxiter = sys->maxiter > 0 ? sys->maxiter : 50;
  hs.comp = elt == CL_ELT_DD;
  hs.s = hs.comp ? 8 : 4;
  int sb = 0;
  int off[CL_MAX_REDUCERS];
  for (int r = 0; r < ch->n; ++r) {
    off[r] = sb;
    sb += red_bytes(ch->r[r], d, k);
  }
  std::atomic<int64_t> next(0);
  auto worker = [&]() {
    std::vector<RState<T>> st(ch->n > 0 ? ch->n : 1);
    while (true) {
      int64_t lo = next.fetch_add(16);
      if (lo >= m) break;
      int64_t hi = lo + 16 < m ? lo + 16 : m;
      for (int64_t i = lo; i < hi; ++i) {
        T xx[8], w0[64], w1[64];
        for (int c = 0; c < s; ++c) xx[c] = (T)x[i * s + c];
        for (int e = 0; e < d * k; ++e) {  // FP32 tangents: the input frames rounded to Float32
          w0[e] = f32 ? (T)(float)W[i * d * k + e] : (T)W[i * d * k + e];
          w1[e] = W2 ? (f32 ? (T)(float)W2[i * d * k + e] : (T)W2[i * d * k + e]) : (T)0;
        }
        uint32_t* wsi = (ws && nws > 0) ? ws + i * nws : nullptr;
        code[i] = sys->kind == CL_SYS_HENON_HEILES_GAUSS
                      ? run_orbit<P, T>(hs, *ch, d, k, N, tau, xx, w0, w1, st.data(), &hstep[i], wsi)
                  : pair ? run_orbit<P, T>(ps, *ch, d, k, N, tau, xx, w0, w1, st.data(), &hstep[i], wsi)
                         : run_orbit<P, T>(ms, *ch, d, k, N, tau, xx, w0, w1, st.data(), &hstep[i], wsi);
        for (int c = 0; c < s; ++c) x[i * s + c] = dbl(xx[c]);
        for (int e = 0; e < d * k; ++e) {
          W[i * d * k + e] = dbl(w0[e]);
          if (W2) W2[i * d * k + e] = dbl(w1[e]);
        }
        if (states)
          for (int r = 0; r < ch->n; ++r) write_state(ch->r[r], st[r], d, k, states + i * sb + off[r]);
      }
    }
    g_gauss_iters += tl_gauss_iters;
    g_gauss_steps += tl_gauss_steps;
    tl_gauss_iters = tl_gauss_steps = 0;
  };
  if (nthreads <= 1) {
    worker();
  } else {
    std::vector<std::thread> th;
    for (int t = 0; t < nthreads; ++t) th.emplace_back(worker);
    for (auto& t : th) t.join();
  }
  return 0;
}

// ------------------------------------------------------------------ the recorder (src/Experiments/run_record.jl)
// Each orbit runs in segments of the per-orbit body to the grid steps (_record_loop!, run_record.jl:160-172); after
// a successful segment, row g of R[orbit][g][col] is the row of _dev_row! (run_device.jl:9-37): lambda (k), the
// RunMonitor's det and symp maxima, orth (the QR state's orth_max, 0 for Benettin), for k = 2 the running maxima of
// |lambda_1 + lambda_2| (eps_sum and eps_pair coincide at k = 2), G_max (1 for Benettin), M_max, then the point. A
// failing segment sets fail = g + 1 (1-based grid row) and ends the orbit's rows. The chain is (estimator,
// RunMonitor) with the estimator a QR spectrum (k = 1, 2) or Benettin (k = 1).
template <class P>
static void record_row(const cl_chain* ch, int k, int s, int g, const RState<double>* st, const double* x, double* row,
                       const double* prev) {
  double lam[2] = {0.0, 0.0}, orth = 0.0, gmax = 1.0;
  if (ch->r[0].kind == CL_RED_QR) {
    const QRSt<double>& q = st[0].qr;
    for (int j = 0; j < k; ++j) lam[j] = q.n == 0 ? 0.0 : (q.sums[j].s + q.sums[j].c) / (double)q.n;
    orth = q.orth_max;
    gmax = q.G_max;
  } else {
    const BenSt<double>& b = st[0].ben;
    lam[0] = b.n == 0 ? 0.0 : (b.sum.s + b.sum.c) / (double)b.n;
  }
  const MonSt<double>& mon = st[1].mon;
  for (int j = 0; j < k; ++j) row[j] = lam[j];
  row[k] = mon.det_max;
  row[k + 1] = mon.symp_max;
  row[k + 2] = orth;
  if (k == 2) {
    const double e = P::abs(lam[0] + lam[1]);
    row[k + 3] = g > 0 ? jmax<P>(prev[k + 3], e) : e;
    row[k + 4] = g > 0 ? jmax<P>(prev[k + 4], e) : e;
  }
  row[k + 5] = gmax;
  row[k + 6] = mon.M_max;
  for (int j = 0; j < s; ++j) row[k + 7 + j] = x[j];
}

// The indicator recorder's row (src/Experiments/run_indicator_record.jl:39-70 `_put!`, run_indicator_device.jl:13-29):
// a series kind's chain is its members then RunMonitor, and the row holds each member's final columns (FLI: fli,
// fli1, fli.jl:91-94; MEGNO: megno_series of column 1 then 2, megno.jl:97-102, with tn^3 the literal power
// (tn tn) tn; SALI: sali and L, sali.jl:62-66), then det_max, symp_max, M_max, then the point. A window kind's chain
// is RunMonitor then its reader, and the row holds the monitor and the point.
static bool pair_kind(int kind) { return kind == CL_SYS_PAIR_STANDARD_MAP || kind == CL_SYS_PAIR_WRAPPED_STANDARD_MAP; }

static void ind_row(const cl_chain* ch, int s, const RState<double>* st, const double* x, double* row) {
  const bool window = ch->r[0].kind == CL_RED_RUNMONITOR;
  const int nmem = window ? 0 : ch->n - 1;
  int col = 0;
  for (int r = 0; r < nmem; ++r) {
    if (ch->r[r].kind == CL_RED_FLI) {
      row[col++] = st[r].fli.fB;
      row[col++] = st[r].fli.f1;
    } else if (ch->r[r].kind == CL_RED_MEGNO) {
      const MegSt<double>& q = st[r].meg;
      for (int c = 0; c < 2; ++c) {
        double y = 0.0, yb = 0.0, y20 = 0.0;
        if (q.K != 0) {
          const double tn = (double)q.n;
          y = 2.0 * (q.c[c].R1.s + q.c[c].R1.c) / tn;
          yb = (q.c[c].Q.s + q.c[c].Q.c) / (double)q.K;
          y20 = (double)(q.n / q.K) * (q.c[c].Q2.s + q.c[c].Q2.c) / (tn * tn * tn);
        }
        row[col++] = y;
        row[col++] = yb;
        row[col++] = y20;
      }
    } else if (ch->r[r].kind == CL_RED_RLI) {  // rli.jl:116-119
      const RliSt<double>& q = st[r].rli;
      double dl = 0.0, rli = 0.0;
      if (q.K != 0) {
        dl = F64Libm::abs((q.sB.s - q.sA.s) + (q.sB.c - q.sA.c)) / (double)q.n;
        rli = (q.Q.s + q.Q.c) / (double)q.K;
      }
      row[col++] = dl;
      row[col++] = rli;
    } else {
      const SaliSt<double>& q = st[r].sali;
      row[col++] = q.sali;
      row[col++] = q.n == 0 ? 0.0 : -((q.s1.s + q.s2.s) + (q.s1.c + q.s2.c));
    }
  }
  const MonSt<double>& mon = st[window ? 0 : ch->n - 1].mon;
  row[col] = mon.det_max;
  row[col + 1] = mon.symp_max;
  row[col + 2] = mon.M_max;
  for (int j = 0; j < s; ++j) row[col + 3 + j] = x[j];
}

static int ind_row_width(const cl_chain* ch, int s) {
  if (ch->r[0].kind == CL_RED_RUNMONITOR) return 3 + s;
  int c = 0;
  for (int r = 0; r + 1 < ch->n; ++r) c += ch->r[r].kind == CL_RED_MEGNO ? 6 : 2;
  return c + 3 + s;
}

// The recorder loops (_record_batch, run_record.jl; _ind_series_loop! and _ind_window_loop!,
// run_indicator_record.jl:99-132): segments of the shared body end at the grid steps, the row of each grid step is
// written from the carried states; a failing segment records its 1-based grid row and ends the orbit's rows.
// ind = false: the exponent methods' row (record_row); true: the indicator row (ind_row).
template <class P>
static int record_all(const cl_system* sys, const cl_chain* ch, int64_t m, int s, int k, int64_t N, int tau,
                      const int64_t* grid, int G, double* x, double* W, double* R, int64_t* fail, int8_t* code,
                      int64_t* hstep, uint8_t* states, int nthreads, bool ind = false) {
  const int ncol = ind ? ind_row_width(ch, pair_kind(sys->kind) ? s / 2 : s) : k + 7 + s;
  const bool pair = sys->kind == CL_SYS_PAIR_STANDARD_MAP || sys->kind == CL_SYS_PAIR_WRAPPED_STANDARD_MAP;
  MapSys<P, double> ms;
  ms.kind = pair ? (sys->kind == CL_SYS_PAIR_STANDARD_MAP ? CL_SYS_STANDARD_MAP : CL_SYS_WRAPPED_STANDARD_MAP) : sys->kind;
  ms.K = sys->K;
  ms.s = pair ? s / 2 : s;
  for (int i = 0; i < 4; ++i) ms.A[i] = 0.0;
  PairSys<P, double> ps;
  ps.base = ms;
  ps.s = s;
  const int d = pair ? 4 : 2;
  int sb = 0, off[CL_MAX_REDUCERS];
  for (int r = 0; r < ch->n; ++r) {
    off[r] = sb;
    sb += red_bytes(ch->r[r], d, k);
  }
  std::atomic<int64_t> next(0);
  auto worker = [&]() {
    RState<double> st[CL_MAX_REDUCERS];
    while (true) {
      int64_t lo = next.fetch_add(16);
      if (lo >= m) break;
      int64_t hi = lo + 16 < m ? lo + 16 : m;
      for (int64_t i = lo; i < hi; ++i) {
        double xx[8], w0[8], w1[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        for (int c = 0; c < s; ++c) xx[c] = x[i * s + c];
        for (int e = 0; e < d * k; ++e) w0[e] = W[i * d * k + e];
        for (int r = 0; r < ch->n; ++r) init_state(ch->r[r], st[r], k);
        int64_t n = 0, h = 0;
        int8_t c = 0;
        fail[i] = 0;
        for (int g = 0; g < G; ++g) {
          c = pair ? run_orbit_seg<P>(ps, *ch, d, k, n, grid[g] - n, N, tau, xx, w0, w1, st, &h, nullptr)
                   : run_orbit_seg<P>(ms, *ch, d, k, n, grid[g] - n, N, tau, xx, w0, w1, st, &h, nullptr);
          if (c != 0) {
            fail[i] = g + 1;
            break;
          }
          n = grid[g];
          double* row = R + ((size_t)i * G + g) * ncol;
          if (ind) ind_row(ch, pair ? s / 2 : s, st, xx, row);
          else record_row<P>(ch, k, s, g, st, xx, row, g > 0 ? row - ncol : nullptr);
        }
        code[i] = c;
        hstep[i] = h;
        for (int cc = 0; cc < s; ++cc) x[i * s + cc] = xx[cc];
        for (int e = 0; e < d * k; ++e) W[i * d * k + e] = w0[e];
        if (states)
          for (int r = 0; r < ch->n; ++r) write_state(ch->r[r], st[r], d, k, states + i * sb + off[r]);
      }
    }
  };
  if (nthreads <= 1) {
    worker();
  } else {
    std::vector<std::thread> th;
    for (int t = 0; t < nthreads; ++t) th.emplace_back(worker);
    for (auto& t : th) t.join();
  }
  return 0;
}

// ------------------------------------------------------------------ two-trajectory estimator
// The reference orbit and a shadow at distance d0, stepped together and placed back at d0 along the separation
// every tau steps (_shadow_body, src/Lyapunov/two_trajectory.jl:45-82). Separation and placement
// (src/Lyapunov/separation.jl:9-38): the plain difference and ref + delta on the lifted map and the linear map; on
// the wrapped map the minimal image of each coordinate in (-pi, pi] and the wrapped placement whose reduction counts
// are added to the reference's.
struct ShadowRec {  // ShadowEnd{s, Float64} (two_trajectory.jl:11-19)
  double x[4], y[4];
  Acc<double> sum;
  double dprev;
  int8_t code;
  int64_t step;
  double value;
};

template <class P>
static void wrap_reduce(double u, double tp, double& r, double& a) {  // _reduce, wrapped_standard_map.jl:15-27
  a = P::floor(u / tp);
  r = u - tp * a;
  if (r < 0) {
    r += tp;
    a -= 1;
  }
  if (r >= tp) {
    r -= tp;
    a += 1;
  }
}

template <class P>
static void sh_sep(const MapSys<P, double>& ms, const double* x, const double* y, double* d) {
  if (ms.kind == CL_SYS_WRAPPED_STANDARD_MAP) {
    const double tp = 2.0 * P::pi(), h = tp / 2;
    for (int j = 0; j < 2; ++j) {
      const double u = y[j] - x[j];
      d[j] = u > h ? u - tp : (u <= -h ? u + tp : u);
    }
  } else {
    d[0] = y[0] - x[0];
    d[1] = y[1] - x[1];
  }
}

template <class P>
static void sh_place(const MapSys<P, double>& ms, const double* x, const double* d, double* y) {
  if (ms.kind == CL_SYS_WRAPPED_STANDARD_MAP) {
    const double tp = 2.0 * P::pi();
    double a, b;
    wrap_reduce<P>(x[0] + d[0], tp, y[0], a);
    wrap_reduce<P>(x[1] + d[1], tp, y[1], b);
    y[2] = x[2] + a;
    y[3] = x[3] + b;
  } else {
    y[0] = x[0] + d[0];
    y[1] = x[1] + d[1];
  }
}

template <class P>
static ShadowRec sh_end(const double* x, const double* y, int s, Acc<double> acc, double dprev, int8_t code,
                        int64_t step, double value) {
  ShadowRec e;
  memset(&e, 0, sizeof(e));
  for (int j = 0; j < s; ++j) {
    e.x[j] = x[j];
    e.y[j] = y[j];
  }
  e.sum = acc;
  e.dprev = dprev;
  e.code = code;
  e.step = step;
  e.value = value;
  return e;
}

// One launch of `steps` steps from global step n0; at n0 = 0 the shadow starts at x0 + d0 u0 (_shadow_start,
// two_trajectory.jl:31-36), later launches continue from e.
template <class P>
static ShadowRec shadow_body(const MapSys<P, double>& ms, const double* x0, const double* u0, const ShadowRec& e,
                             int64_t n0, int64_t steps, int tau, double d0) {
  const int s = ms.s;
  double x[4], y[4], dprev;
  Acc<double> acc;
  if (n0 == 0) {
    for (int j = 0; j < s; ++j) x[j] = y[j] = x0[j];
    const double dl[2] = {d0 * u0[0], d0 * u0[1]};
    double sp[2];
    sh_place(ms, x, dl, y);
    sh_sep(ms, x, y, sp);
    dprev = snorm<P>(sp, 2);
    acc.s = acc.c = 0.0;
    const int8_t code = !(P::finite(dprev) && dprev > 0) ? 2 : ((dprev < d0 / 2 || dprev > 2 * d0) ? 3 : 0);
    if (code) return sh_end<P>(x, y, s, acc, dprev, code, 0, dprev);
  } else {
    for (int j = 0; j < s; ++j) {
      x[j] = e.x[j];
      y[j] = e.y[j];
    }
    acc = e.sum;
    dprev = e.dprev;
  }
  int64_t n = n0;
  for (int64_t h = 0; h < steps / tau; ++h) {
    for (int t = 0; t < tau; ++t) {
      ms.step(x, nullptr, 0, 0);
      ms.step(y, nullptr, 0, 0);
    }
    n += tau;
    bool bad = false;
    double v = 0.0;
    for (int j = 0; j < s && !bad; ++j)
      if (!P::finite(x[j])) bad = true, v = x[j];
    for (int j = 0; j < s && !bad; ++j)
      if (!P::finite(y[j])) bad = true, v = y[j];
    if (bad) return sh_end<P>(x, y, s, acc, dprev, 1, n, v);
    double dl[2];
    sh_sep(ms, x, y, dl);
    const double d = snorm<P>(dl, 2);
    if (!(P::finite(d) && d > 0)) return sh_end<P>(x, y, s, acc, dprev, 2, n, d);
    acc = acc_add<P>(acc, P::log(d / dprev));
    const double r = d0 / d;
    const double nd[2] = {dl[0] * r, dl[1] * r};
    double sp[2];
    sh_place(ms, x, nd, y);
    sh_sep(ms, x, y, sp);
    dprev = snorm<P>(sp, 2);
    if (!(dprev >= d0 / 2 && dprev <= 2 * d0))
      return sh_end<P>(x, y, s, acc, dprev, (P::finite(dprev) && dprev > 0) ? 3 : 2, n, dprev);
  }
  return sh_end<P>(x, y, s, acc, dprev, 0, n, dprev);
}

// end record <-> the Julia byte layout of ShadowEnd{s, Float64}: 2s + 6 words
static void sh_write(const ShadowRec& e, int s, uint64_t* w) {
  memcpy(w, e.x, 8 * s);
  memcpy(w + s, e.y, 8 * s);
  memcpy(w + 2 * s, &e.sum.s, 8);
  memcpy(w + 2 * s + 1, &e.sum.c, 8);
  memcpy(w + 2 * s + 2, &e.dprev, 8);
  w[2 * s + 3] = (uint8_t)e.code;
  memcpy(w + 2 * s + 4, &e.step, 8);
  memcpy(w + 2 * s + 5, &e.value, 8);
}
static ShadowRec sh_read(const uint64_t* w, int s) {
  ShadowRec e;
  memset(&e, 0, sizeof(e));
  memcpy(e.x, w, 8 * s);
  memcpy(e.y, w + s, 8 * s);
  memcpy(&e.sum.s, w + 2 * s, 8);
  memcpy(&e.sum.c, w + 2 * s + 1, 8);
  memcpy(&e.dprev, w + 2 * s + 2, 8);
  e.code = (int8_t)(w[2 * s + 3] & 0xff);
  memcpy(&e.step, w + 2 * s + 4, 8);
  memcpy(&e.value, w + 2 * s + 5, 8);
  return e;
}

template <class F>
static void for_orbits(int64_t m, int nthreads, F f) {
  std::atomic<int64_t> next(0);
  auto worker = [&]() {
    while (true) {
      int64_t lo = next.fetch_add(64);
      if (lo >= m) break;
      for (int64_t i = lo; i < (lo + 64 < m ? lo + 64 : m); ++i) f(i);
    }
  };
  if (nthreads <= 1) {
    worker();
    return;
  }
  std::vector<std::thread> th;
  for (int t = 0; t < nthreads; ++t) th.emplace_back(worker);
  for (auto& t : th) t.join();
}

template <class P>
static MapSys<P, double> shadow_sys(const cl_system* sys, int s, const double* A) {
  MapSys<P, double> ms;
  ms.kind = A ? OR_SYS_LINEAR : sys->kind;
  ms.K = A ? 0.0 : sys->K;
  ms.s = s;
  for (int i = 0; i < 4; ++i) ms.A[i] = A ? A[i] : 0.0;
  return ms;
}

template <class P>
static void shadow_all(const cl_system* sys, const double* A, int64_t m, int s, int64_t n0, int64_t steps, int tau,
                       double d0, const double* x0, const double* u0, uint64_t* ends, int nthreads) {
  const MapSys<P, double> ms = shadow_sys<P>(sys, s, A);
  const int ew = 2 * s + 6;
  for_orbits(m, nthreads, [&](int64_t i) {
    ShadowRec e = sh_read(ends + i * ew, s);
    if (n0 > 0 && e.code != 0) return;  // failed in an earlier launch: kept, not advanced (two_trajectory.jl:113-121)
    e = shadow_body(ms, x0 + i * s, u0 + 2 * i, e, n0, steps, tau, d0);
    sh_write(e, s, ends + i * ew);
  });
}

// The recorder's L2 path (_record_shadow!, src/Experiments/run_record.jl:152-172): segments to each grid step;
// row g gets total(sum) / grid[g] in column 0 and the point in columns 8..8+s; fail = 1-based failing row.
template <class P>
static void shadow_record_all(const cl_system* sys, int64_t m, int s, int tau, double d0, const int64_t* grid, int G,
                              const double* x0, const double* u0, double* R, int64_t* fail, uint64_t* ends,
                              int nthreads) {
  const MapSys<P, double> ms = shadow_sys<P>(sys, s, nullptr);
  const int ew = 2 * s + 6, ncol = 1 + 7 + s;
  for_orbits(m, nthreads, [&](int64_t i) {
    ShadowRec e;
    memset(&e, 0, sizeof(e));
    int64_t n = 0;
    fail[i] = 0;
    for (int g = 0; g < G; ++g) {
      e = shadow_body(ms, x0 + i * s, u0 + 2 * i, e, n, grid[g] - n, tau, d0);
      if (e.code != 0) {
        fail[i] = g + 1;
        break;
      }
      n = grid[g];
      double* row = R + ((size_t)i * G + g) * ncol;
      row[0] = acc_total(e.sum) / (double)grid[g];
      for (int j = 0; j < s; ++j) row[8 + j] = e.x[j];
    }
    sh_write(e, s, ends + i * ew);
  });
}

static bool or_dims_ok(const cl_system* sys, int d, int k) {
  int want = sys->kind == CL_SYS_HENON_HEILES_GAUSS ? 4
           : (sys->kind == CL_SYS_STANDARD_MAP || sys->kind == CL_SYS_WRAPPED_STANDARD_MAP ||
              sys->kind == CL_SYS_STANDARD_MAP_K) ? 2
           : (sys->kind == CL_SYS_PAIR_STANDARD_MAP || sys->kind == CL_SYS_PAIR_WRAPPED_STANDARD_MAP) ? 4 : -1;
  return d == want && k >= 1 && k <= d;
}

#include "oracle_chain.hpp"

extern "C" {

// The coupled standard-map chain (kind CL_SYS_STANDARD_MAP_CHAIN) with no reducer or QRSpectrum(HouseholderQR()),
// k <= 16, full runs from fresh states. Buffers are orbit-major: x[i][2N], W[i][k][2N] (column-major frames),
// states[i][16k + 72]. Ns (optional) gives each orbit its own step count (a multiple of tau), else N for all.
// mode RED_ROWS (0) is the reference's row-order reductions, RED_SITES (1) and RED_BLOCKED (2) the device kernels'.
// elt CL_ELT_F64 (1) or CL_ELT_F64_F32TAN (2): Float32 tangents (chain_orbit_f32).
int or_propagate_chain(const cl_system* sys, const cl_chain* ch, int64_t m, int k, const int64_t* Ns, int64_t N,
                       int tau, double* x, double* W, void* states, int8_t* code, int64_t* hstep, int math, int mode,
                       int nthreads, int elt) {
  if (sys->kind != CL_SYS_STANDARD_MAP_CHAIN || sys->N < 2 || sys->N > 128 || (sys->ring && sys->N < 3)) return 10;
  if (k < 1 || k > 16 || k > 2 * sys->N || tau < 1) return 10;
  if (mode != RED_ROWS && (sys->N & (sys->N - 1))) return 10;  // the butterfly needs a power-of-two site count
  if (elt != CL_ELT_F64 && elt != CL_ELT_F64_F32TAN) return 3;
  const bool f32 = elt == CL_ELT_F64_F32TAN;
  if (ch->n > 1 || (ch->n == 1 && !(ch->r[0].kind == CL_RED_QR && ch->r[0].alg == CL_QR_HOUSEHOLDER))) return 3;
  for (int64_t i = 0; i < m; ++i)
    if ((Ns ? Ns[i] : N) % tau) return 10;
  if (math == MATH_CL)
    return chain_all<F64Cl>(sys, ch, m, k, Ns, N, tau, x, W, (uint8_t*)states, code, hstep, mode, nthreads, f32);
  return chain_all<F64Libm>(sys, ch, m, k, Ns, N, tau, x, W, (uint8_t*)states, code, hstep, mode, nthreads, f32);
}

// Same-orbit check of the standard map's QR spectrum (the AC4 method, test/tangent-core/tc_ac01_04.jl:66-88):
// the orbit steps in FP64 exactly as the kernels step it (portable sin), while the Jacobian at each FP64
// point, the frame products and QRSpectrum (chain ch: one QR reducer) run in __float128. Writes
// lambda[i][j] = total(sums_j) / n. x0 [m][2], W0 [m][k][2] (column-major frames).
int or_same_orbit_quad(double K, const cl_chain* ch, int64_t m, int k, int64_t N, int tau, const double* x0,
                       const double* W0, double* lambda, int nthreads) {
  if (ch->n != 1 || ch->r[0].kind != CL_RED_QR || k < 1 || k > 2 || N % tau) return 10;
  std::atomic<int64_t> next(0);
  auto worker = [&]() {
    while (true) {
      int64_t i = next.fetch_add(1);
      if (i >= m) break;
      double xd = x0[2 * i], pd = x0[2 * i + 1];
      quad W[4], O[4];
      for (int e = 0; e < 2 * k; ++e) W[e] = (quad)W0[i * 2 * k + e];
      QRSt<quad> st;
      memset(&st, 0, sizeof(st));
      for (int s2 = 0; s2 < 8; ++s2) st.sums[s2].s = st.sums[s2].c = 0;
      st.G_max = st.orth_max = st.value = st.threshold = 0;
      for (int64_t n = 1; n <= N; ++n) {
        quad c = (quad)K * cosq((quad)xd);
        quad a11 = (quad)1 + c;
        for (int j = 0; j < k; ++j) {
          quad w0 = W[2 * j], w1 = W[2 * j + 1];
          O[2 * j] = fmaq((quad)1, w1, a11 * w0);
          O[2 * j + 1] = fmaq((quad)1, w1, c * w0);
        }
        for (int e = 0; e < 2 * k; ++e) W[e] = O[e];
        double p2 = pd + K * cl_sin(xd);
        xd = xd + p2;
        pd = p2;
        if (n % tau == 0) red_qr<Quad>(ch->r[0], st, n, W, 2, k);
      }
      for (int j = 0; j < k; ++j)
        lambda[i * k + j] = st.n ? dbl((st.sums[j].s + st.sums[j].c) / (quad)st.n) : 0.0;
    }
  };
  std::vector<std::thread> th;
  for (int t = 0; t < (nthreads > 1 ? nthreads : 1); ++t) th.emplace_back(worker);
  for (auto& t : th) t.join();
  return 0;
}

// Gauss fixed-point iterations and steps summed over the runs since the last call; resets both.
void or_gauss_iterations(int64_t* iters, int64_t* steps) {
  *iters = g_gauss_iters.exchange(0);
  *steps = g_gauss_steps.exchange(0);
}

// Bytes per orbit of the chain's state tuple (Julia layout), -1 for an unknown reducer.
int64_t or_state_bytes(const cl_chain* ch, int d, int k) {
  int64_t b = 0;
  for (int r = 0; r < ch->n; ++r) {
    int x = red_bytes(ch->r[r], d, k);
    if (x < 0) return -1;
    b += x;
  }
  return b;
}

// Full runs of N steps from fresh states (n0 = 0) on host AoS buffers; `math` selects the
// elementary functions (MATH_LIBM or MATH_CL); nthreads worker threads over groups of 16 orbits.
int or_propagate(const cl_system* sys, const cl_chain* ch, int64_t m, int s, int d, int k, int64_t N, int tau,
                 double* x, double* W, double* W2, void* states, int8_t* code, int64_t* hstep, int math,
                 int nthreads) {
  if (!or_dims_ok(sys, d, k) || N % tau) return 10;
  if (math == MATH_CL)
    return propagate_all<F64Cl>(sys, ch, m, s, d, k, N, tau, x, W, W2, (uint8_t*)states, code, hstep, nthreads);
  return propagate_all<F64Libm>(sys, ch, m, s, d, k, N, tau, x, W, W2, (uint8_t*)states, code, hstep, nthreads);
}

// or_propagate with an element type: CL_ELT_F64, or CL_ELT_F64_F32TAN on the standard maps (FP64 orbit and Jacobian,
// Float32 frames between handoffs, FP64 reducers; the frames rounded to Float32 on entry and after each handoff's
// chain).
int or_propagate_elt(const cl_system* sys, const cl_chain* ch, int64_t m, int s, int d, int k, int64_t N, int tau,
                     double* x, double* W, double* W2, void* states, int8_t* code, int64_t* hstep, int math,
                     int nthreads, int elt) {
  if (!or_dims_ok(sys, d, k) || N % tau) return 10;
  const bool map = sys->kind == CL_SYS_STANDARD_MAP || sys->kind == CL_SYS_WRAPPED_STANDARD_MAP ||
                   sys->kind == CL_SYS_STANDARD_MAP_K;
  if (elt == CL_ELT_F64) return or_propagate(sys, ch, m, s, d, k, N, tau, x, W, W2, states, code, hstep, math, nthreads);
  const bool hh = sys->kind == CL_SYS_HENON_HEILES_GAUSS;
  if (!((elt == CL_ELT_F64_F32TAN && map) || (elt == CL_ELT_DD && hh && s == 8))) return 3;
  if (math == MATH_CL)
    return propagate_all<F64Cl>(sys, ch, m, s, d, k, N, tau, x, W, W2, (uint8_t*)states, code, hstep, nthreads,
                                nullptr, nullptr, elt);
  return propagate_all<F64Libm>(sys, ch, m, s, d, k, N, tau, x, W, W2, (uint8_t*)states, code, hstep, nthreads,
                                nullptr, nullptr, elt);
}

// The compensated Henon-Heiles orbit in __float128 (the state's correction words carried as well), for the energy
// and exponent comparisons of the double-double element type.
int or_propagate_quad_elt(const cl_system* sys, const cl_chain* ch, int64_t m, int s, int d, int k, int64_t N, int tau,
                          double* x, double* W, double* W2, void* states, int8_t* code, int64_t* hstep, int nthreads,
                          int elt) {
  if (!or_dims_ok(sys, d, k) || N % tau) return 10;
  return propagate_all<Quad>(sys, ch, m, s, d, k, N, tau, x, W, W2, (uint8_t*)states, code, hstep, nthreads, nullptr,
                             nullptr, elt);
}

// The recorder on a standard map (lifted, s = 2) or a wrapped standard map (s = 4) with the chain (estimator,
// RunMonitor): grid[0..G) increasing multiples of tau ending at N; R (m x G x (k + 7 + s), zeroed by the caller) gets
// the rows, fail[i] the 1-based failing grid row or 0; x, W, code, hstep and states as in or_propagate.
int or_record(const cl_system* sys, const cl_chain* ch, int64_t m, int s, int k, int64_t N, int tau,
              const int64_t* grid, int G, double* x, double* W, double* R, int64_t* fail, int8_t* code,
              int64_t* hstep, void* states, int math, int nthreads) {
  const bool est_ok = ch->n == 2 && ch->r[1].kind == CL_RED_RUNMONITOR &&
                      ((ch->r[0].kind == CL_RED_QR && (k == 1 || k == 2)) || (ch->r[0].kind == CL_RED_BENETTIN && k == 1));
  const bool sys_ok = (sys->kind == CL_SYS_STANDARD_MAP && s == 2) || (sys->kind == CL_SYS_WRAPPED_STANDARD_MAP && s == 4);
  if (!est_ok || !sys_ok || N % tau || G < 1 || grid[G - 1] != N) return 10;
  for (int g = 0; g < G; ++g)
    if (grid[g] % tau || grid[g] <= (g ? grid[g - 1] : 0)) return 10;
  if (math == MATH_CL)
    return record_all<F64Cl>(sys, ch, m, s, k, N, tau, grid, G, x, W, R, fail, code, hstep, (uint8_t*)states, nthreads);
  return record_all<F64Libm>(sys, ch, m, s, k, N, tau, grid, G, x, W, R, fail, code, hstep, (uint8_t*)states, nthreads);
}

// The indicator recorder on a standard map (lifted s = 2, wrapped s = 4): a series kind's chain (FLI, MEGNO, SALI
// members, then RunMonitor; k = 2) or a window kind's (RunMonitor then WeightedBirkhoff, k = 1, with the constant
// deviation column flag). R[m][G][row width], fail, x, W (m x k x 2), code, hstep and the states as or_record.
int or_ind_record(const cl_system* sys, const cl_chain* ch, int64_t m, int s, int k, int64_t N, int tau,
                  const int64_t* grid, int G, double* x, double* W, double* R, int64_t* fail, int8_t* code,
                  int64_t* hstep, void* states, int math, int nthreads) {
  const bool sys_ok = (sys->kind == CL_SYS_STANDARD_MAP && s == 2) || (sys->kind == CL_SYS_WRAPPED_STANDARD_MAP && s == 4) ||
                      (sys->kind == CL_SYS_PAIR_STANDARD_MAP && s == 4) ||
                      (sys->kind == CL_SYS_PAIR_WRAPPED_STANDARD_MAP && s == 8);
  bool ch_ok = ch->n >= 2;
  if (ch_ok && ch->r[0].kind == CL_RED_RUNMONITOR) {
    ch_ok = ch->n == 2 && (ch->r[1].kind == CL_RED_BIRKHOFF || ch->r[1].kind == CL_RED_PHASE_ADVANCE) && k == 1 &&
            (ch->flags & CL_CHAIN_RESET_FRAME);
  } else if (ch_ok) {
    ch_ok = k == 2 && ch->n <= 5 && ch->r[ch->n - 1].kind == CL_RED_RUNMONITOR;
    for (int r = 0; ch_ok && r + 1 < ch->n; ++r)
      ch_ok = ch->r[r].kind == CL_RED_FLI || ch->r[r].kind == CL_RED_MEGNO || ch->r[r].kind == CL_RED_SALI ||
              (ch->r[r].kind == CL_RED_RLI && pair_kind(sys->kind));
  }
  if (!ch_ok || !sys_ok || N % tau || G < 1 || grid[G - 1] != N) return 10;
  for (int g = 0; g < G; ++g)
    if (grid[g] % tau || grid[g] <= (g ? grid[g - 1] : 0)) return 10;
  if (math == MATH_CL)
    return record_all<F64Cl>(sys, ch, m, s, k, N, tau, grid, G, x, W, R, fail, code, hstep, (uint8_t*)states,
                             nthreads, true);
  return record_all<F64Libm>(sys, ch, m, s, k, N, tau, grid, G, x, W, R, fail, code, hstep, (uint8_t*)states,
                             nthreads, true);
}

// The two-trajectory estimator on a standard map (lifted s = 2, wrapped s = 4), or on the linear map x' = A x
// (A column-major 2 x 2, s = 2) when A is given: one launch of `steps` steps from global step n0 over host AoS
// buffers x0[m][s], u0[m][2] (unit directions) and the end records ends[m][2s + 6] (Julia byte layout), in place.
int or_shadow(const cl_system* sys, const double* A, int64_t m, int s, int64_t n0, int64_t steps, int64_t N, int tau,
              double d0, const double* x0, const double* u0, void* ends, int math, int nthreads) {
  const bool sys_ok = A ? s == 2
                        : ((sys->kind == CL_SYS_STANDARD_MAP && s == 2) || (sys->kind == CL_SYS_WRAPPED_STANDARD_MAP && s == 4));
  if (!sys_ok || tau < 1 || steps < 0 || steps % tau || n0 < 0 || n0 % tau || n0 + steps > N || N % tau) return 10;
  if (!(d0 > 0 && cl_isfinite(d0))) return 10;
  if (math == MATH_CL)
    shadow_all<F64Cl>(sys, A, m, s, n0, steps, tau, d0, x0, u0, (uint64_t*)ends, nthreads);
  else
    shadow_all<F64Libm>(sys, A, m, s, n0, steps, tau, d0, x0, u0, (uint64_t*)ends, nthreads);
  return 0;
}

// The recorder's two-trajectory rows: R (m x G x (8 + s), zeroed by the caller), fail[i] the 1-based failing grid row
// or 0, ends the end records after the last segment run.
int or_shadow_record(const cl_system* sys, int64_t m, int s, int64_t N, int tau, double d0, const int64_t* grid,
                     int G, const double* x0, const double* u0, double* R, int64_t* fail, void* ends, int math,
                     int nthreads) {
  const bool sys_ok = (sys->kind == CL_SYS_STANDARD_MAP && s == 2) || (sys->kind == CL_SYS_WRAPPED_STANDARD_MAP && s == 4);
  if (!sys_ok || tau < 1 || N % tau || G < 1 || grid[G - 1] != N || !(d0 > 0 && cl_isfinite(d0))) return 10;
  for (int g = 0; g < G; ++g)
    if (grid[g] % tau || grid[g] <= (g ? grid[g - 1] : 0)) return 10;
  if (math == MATH_CL)
    shadow_record_all<F64Cl>(sys, m, s, tau, d0, grid, G, x0, u0, R, fail, (uint64_t*)ends, nthreads);
  else
    shadow_record_all<F64Libm>(sys, m, s, tau, d0, grid, G, x0, u0, R, fail, (uint64_t*)ends, nthreads);
  return 0;
}

// or_propagate for a chain with a workspace reducer (StretchingNumbers): ws[m][chain_ws] UInt32 counts per orbit
// (the reference's ws[:, i]), zeroed at the start of each orbit. With A (column-major 2 x 2) the linear map
// x' = A x runs instead of sys (then s = d = 2).
int or_propagate_ws(const cl_system* sys, const double* A, const cl_chain* ch, int64_t m, int s, int d, int k,
                    int64_t N, int tau, double* x, double* W, void* states, int8_t* code, int64_t* hstep,
                    uint32_t* ws, int math, int nthreads) {
  cl_system lin;
  if (A) {
    memset(&lin, 0, sizeof(lin));
    lin.kind = OR_SYS_LINEAR;
    sys = &lin;
    if (s != 2 || d != 2) return 10;
  } else if (!or_dims_ok(sys, d, k)) {
    return 10;
  }
  if (N % tau || k < 1 || k > 2 || !ws) return 10;
  for (int r = 0; r < ch->n; ++r)
    if (ch->r[r].kind == CL_RED_STRETCHING && (k != 2 || ch->r[r].bins < 1 || !(ch->r[r].lo < ch->r[r].hi))) return 10;
  if (math == MATH_CL)
    return propagate_all<F64Cl>(sys, ch, m, s, d, k, N, tau, x, W, nullptr, (uint8_t*)states, code, hstep, nthreads,
                                A, ws);
  return propagate_all<F64Libm>(sys, ch, m, s, d, k, N, tau, x, W, nullptr, (uint8_t*)states, code, hstep, nthreads,
                                A, ws);
}

// Switch a planted defect on (id > 0) or off (0); see g_defect.
void or_plant_defect(int id) { g_defect = id; }

// A linear map x' = A x of dimension d (2..8, A column-major d x d), k <= d columns: the closed-form fixtures of the
// reference's suites in more than two dimensions (the free particle of test/ofli/of_ac01_05.jl:1-40 with A = [[I, hI],
// [0, I]], prescribed frames for GALI with A = I). math as in or_propagate; quad != 0 runs in __float128.
int or_propagate_linear_d(const double* A, int d, const cl_chain* ch, int64_t m, int k, int64_t N, int tau, double* x,
                          double* W, void* states, int8_t* code, int64_t* hstep, int quad_) {
  cl_system sys;
  memset(&sys, 0, sizeof(sys));
  sys.kind = OR_SYS_LINEAR;
  if (d < 2 || d > 8 || k < 1 || k > d || N % tau) return 10;
  if (quad_)
    return propagate_all<Quad>(&sys, ch, m, d, d, k, N, tau, x, W, nullptr, (uint8_t*)states, code, hstep, 1, A);
  return propagate_all<F64Cl>(&sys, ch, m, d, d, k, N, tau, x, W, nullptr, (uint8_t*)states, code, hstep, 1, A);
}

// A linear map x' = A x (A column-major 2 x 2): the closed-form fixtures of the reference's suites
// (Y5 / cat map [[2,1],[1,1]]). math as in or_propagate; quad != 0 runs in __float128.
int or_propagate_linear(const double* A, const cl_chain* ch, int64_t m, int k, int64_t N, int tau, double* x,
                        double* W, void* states, int8_t* code, int64_t* hstep, int quad_) {
  cl_system sys;
  memset(&sys, 0, sizeof(sys));
  sys.kind = OR_SYS_LINEAR;
  if (k < 1 || k > 2 || N % tau) return 10;
  if (quad_)
    return propagate_all<Quad>(&sys, ch, m, 2, 2, k, N, tau, x, W, nullptr, (uint8_t*)states, code, hstep, 1, A);
  return propagate_all<F64Cl>(&sys, ch, m, 2, 2, k, N, tau, x, W, nullptr, (uint8_t*)states, code, hstep, 1, A);
}

// The weights w(jj), jj < N, of WeightedBirkhoff's horizon N (math as in or_propagate).
void or_wb_weights(int64_t N, int order, double* w, int math) {
  for (int64_t jj = 0; jj < N; ++jj)
    w[jj] = math == MATH_CL ? wb_w<F64Cl, double>(jj, N, order) : wb_w<F64Libm, double>(jj, N, order);
}

// window_gw of n 2 x 2 upper-triangular factors (column-major, 4 entries each), for the test of the kernels' G_w
// estimate.
void or_window_gw2(int64_t n, const double* R, double* out) {
  for (int64_t i = 0; i < n; ++i) out[i] = window_gw<F64Cl, double>(R + 4 * i, 2);
}

// The portable atan2 of the kernels (cl_math.h), for its accuracy test against libm.
void or_atan2(int64_t n, const double* y, const double* x, double* out) {
  for (int64_t i = 0; i < n; ++i) out[i] = cl_atan2(y[i], x[i]);
}

// The portable exp of the kernels (cl_math.h), for its accuracy test against libm.
void or_exp(int64_t n, const double* x, double* out) {
  for (int64_t i = 0; i < n; ++i) out[i] = cl_exp(x[i]);
}

// n additions of x into a compensated sum (Foundation.jl:48-58), and the plain sum, for the AC9 fixture.
void or_csum(int64_t n, double x, double* comp, double* plain) {
  Acc<double> a;
  a.s = a.c = 0.0;
  double p = 0.0;
  for (int64_t i = 0; i < n; ++i) {
    a = acc_add<F64Libm>(a, x);
    p += x;
  }
  *comp = acc_total(a);
  *plain = p;
}
void or_csum_list(int64_t n, const double* v, double* comp) {
  Acc<double> a;
  a.s = a.c = 0.0;
  for (int64_t i = 0; i < n; ++i) a = acc_add<F64Libm>(a, v[i]);
  *comp = acc_total(a);
}

// The same run in __float128 (libquadmath): inputs/outputs are rounded to double, states are the
// Julia layout with each value rounded to double.
int or_propagate_quad(const cl_system* sys, const cl_chain* ch, int64_t m, int s, int d, int k, int64_t N, int tau,
                      double* x, double* W, double* W2, void* states, int8_t* code, int64_t* hstep, int nthreads) {
  if (!or_dims_ok(sys, d, k) || N % tau) return 10;
  return propagate_all<Quad>(sys, ch, m, s, d, k, N, tau, x, W, W2, (uint8_t*)states, code, hstep, nthreads);
}

// Seeded frames of the documented counter rule (the same rule the library implements):
// key = mix(mix(seed) + orbit), uniforms u(t), Box-Muller normals, two MGS passes per column, unit columns.
static uint64_t mixr(uint64_t z) {
  z += 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}
int or_initial_deviations(uint64_t seed, int64_t orbit0, int64_t m, int d, int k, double* W) {
  if (d > 256 || k > d || k < 1 || k > 16) return 10;
  std::vector<double> Gv((size_t)d * k);
  for (int64_t o = 0; o < m; ++o) {
    uint64_t key = mixr(mixr(seed) + (uint64_t)(orbit0 + o));
    double* G = Gv.data();
    for (int j = 0; j < k; ++j)
      for (int i = 0; i < d; ++i) {
        uint64_t t = (uint64_t)(j * d + i), p = t >> 1;
        double u1 = ((double)(mixr(key + (2 * p + 1) * 0x9e3779b97f4a7c15ull) >> 12) + 0.5) * 0x1p-52;
        double u2 = ((double)(mixr(key + (2 * p + 2) * 0x9e3779b97f4a7c15ull) >> 12) + 0.5) * 0x1p-52;
        double r = sqrt(-2.0 * cl_log(u1)), th = 6.283185307179586 * u2;
        G[j * d + i] = (t & 1) ? r * cl_sin(th) : r * cl_cos(th);
      }
    for (int j = 0; j < k; ++j) {
      double* g = G + j * d;
      for (int pass = 0; pass < 2; ++pass)
        for (int i = 0; i < j; ++i) {
          double c = vdot(G + i * d, g, d);
          for (int l = 0; l < d; ++l) g[l] = g[l] - c * G[i * d + l];
        }
      double nr = snorm<F64Cl>(g, d);
      for (int l = 0; l < d; ++l) g[l] = g[l] / nr;
    }
    for (int e = 0; e < d * k; ++e) W[o * d * k + e] = G[e];
  }
  return 0;
}

}  // extern "C"
