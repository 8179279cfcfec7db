// shadow.hip - the two-trajectory estimator (src/Lyapunov/two_trajectory.jl:45-82,113-121): a shadow orbit at
// distance d0 from the reference orbit, stepped with it and placed back at distance d0 along the current
// separation every tau steps; X2 = (sum of log(d / d_prev)) / N. One orbit per lane; the per-orbit end record
// (ShadowEnd{s, Float64}: x[s], y[s], sum (s, c), dprev, code and 7 pad bytes, step, value; 2s + 6 words) lives in
// HBM as 8-byte words, word-major (e[w m + i]), and is carried across launches.
//
// Separation and placement (src/Lyapunov/separation.jl): on the lifted standard map the plain difference and
// ref + delta; on the wrapped map the minimal image of each coordinate in (-pi, pi] and the wrapped placement whose
// reduction counts are added to the reference's winding counts.
#include <hip/hip_runtime.h>
#include "cl_core.hpp"
#include "../../include/chaoslab_hip.h"

namespace cl {

struct ShadowArgs {
  const double* x0;  // [s][m] starts (read at n0 = 0)
  const double* u0;  // [2][m] unit directions (read at n0 = 0)
  uint64_t* e;       // [2s + 6][m] end records
  int64_t m, n0, steps, N;
  int tau;
  double d0;
};

template <class Sys>
struct ShadowGeom;
template <>
struct ShadowGeom<StdMap> {
  CL_HD static void sep(const double* x, const double* y, double* d) {
    d[0] = y[0] - x[0];
    d[1] = y[1] - x[1];
  }
  CL_HD static void place(const double* x, const double* d, double* y) {
    y[0] = x[0] + d[0];
    y[1] = x[1] + d[1];
  }
};
// 2 pi as the wrapped map's step forms it: 2 * (pi_hi + pi_lo) in Float64 (wrapped_standard_map.jl:40)
CL_HD double shadow_two_pi() { return 2.0 * (3.141592653589793 + 1.2246467991473532e-16); }
CL_HD double minimal_image(double u, double tp) {
  const double h = tp / 2.0;
  return u > h ? u - tp : (u <= -h ? u + tp : u);
}
template <>
struct ShadowGeom<WrappedStdMap> {
  CL_HD static void sep(const double* x, const double* y, double* d) {
    const double tp = shadow_two_pi();
    d[0] = minimal_image(y[0] - x[0], tp);
    d[1] = minimal_image(y[1] - x[1], tp);
  }
  CL_HD static void place(const double* x, const double* d, double* y) {
    const double tp = shadow_two_pi();
    double a, b;
    WrappedStdMap::reduce(x[0] + d[0], tp, &y[0], &a);
    WrappedStdMap::reduce(x[1] + d[1], tp, &y[1], &b);
    y[2] = x[2] + a;
    y[3] = x[3] + b;
  }
};

// safe_norm of a separation with the reference's NaN rule (maximum(abs, v) is NaN when an entry is NaN, and then
// the norm is NaN; Foundation/norm.jl:8-12)
CL_HD double sep_norm(const double* v) {
  if (cl_isnan(v[0]) || cl_isnan(v[1])) return __builtin_nan("");
  return safe_norm<2>(v);
}

template <class Sys>
__global__ void __launch_bounds__(256) shadow_kernel(ShadowArgs a, Sys sys) {
  constexpr int S = Sys::S;
  typedef ShadowGeom<Sys> G;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.m) return;
  auto W = [&](int w) -> uint64_t& { return a.e[(int64_t)w * a.m + i]; };
  auto dw = [&](int w) { return __longlong_as_double((long long)W(w)); };
  auto put = [&](int w, double v) { W(w) = (uint64_t)__double_as_longlong(v); };
  double x[S], y[S], dprev;
  CSum acc;
  if (a.n0 == 0) {
#pragma unroll
    for (int j = 0; j < S; ++j) x[j] = a.x0[(int64_t)j * a.m + i];
    acc = csum_zero();
  } else {
    if ((int8_t)(W(2 * S + 3) & 0xff) != 0) return;  // failed in an earlier launch: not advanced
#pragma unroll
    for (int j = 0; j < S; ++j) {
      x[j] = dw(j);
      y[j] = dw(S + j);
    }
    acc = CSum{dw(2 * S), dw(2 * S + 1)};
    dprev = dw(2 * S + 2);
  }
  int8_t code = 0;
  int64_t step = 0;
  double value = 0.0;
  const double lo = a.d0 / 2.0, hi = 2.0 * a.d0;
  if (a.n0 == 0) {  // _shadow_start: the shadow at ref + d0 u0 and its realized separation
    double dl[2] = {a.d0 * a.u0[i], a.d0 * a.u0[a.m + i]}, sp[2];
#pragma unroll
    for (int j = 0; j < S; ++j) y[j] = x[j];
    G::place(x, dl, y);
    G::sep(x, y, sp);
    dprev = sep_norm(sp);
    code = !(cl_isfinite(dprev) && dprev > 0.0) ? 2 : ((dprev < lo || dprev > hi) ? 3 : 0);
    value = dprev;
  }
  int64_t n = a.n0;
  if (code == 0) {
    const int64_t nh = a.steps / a.tau;
    typename Sys::Aux ax, ay;
    for (int64_t h = 0; h < nh; ++h) {
      for (int t = 0; t < a.tau; ++t) {
        sys.aux(x, ax);
        sys.forward(ax, x);
        sys.aux(y, ay);
        sys.forward(ay, y);
      }
      n += a.tau;
      bool fx = true, fy = true;
      double bad = 0.0;
#pragma unroll
      for (int j = S - 1; j >= 0; --j) {  // the first non-finite entry, x before y
        if (!cl_isfinite(y[j])) {
          fy = false;
          bad = y[j];
        }
      }
#pragma unroll
      for (int j = S - 1; j >= 0; --j) {
        if (!cl_isfinite(x[j])) {
          fx = false;
          bad = x[j];
        }
      }
      if (!(fx && fy)) {
        code = 1;
        step = n;
        value = bad;
        break;
      }
      double dl[2];
      G::sep(x, y, dl);
      const double d = sep_norm(dl);
      if (!(cl_isfinite(d) && d > 0.0)) {
        code = 2;
        step = n;
        value = d;
        break;
      }
      acc = csum_add(acc, cl_log(d / dprev));
      const double r = a.d0 / d;
      double nd[2] = {dl[0] * r, dl[1] * r}, sp[2];
      G::place(x, nd, y);
      G::sep(x, y, sp);
      dprev = sep_norm(sp);
      if (!(dprev >= lo && dprev <= hi)) {
        code = (cl_isfinite(dprev) && dprev > 0.0) ? 3 : 2;
        step = n;
        value = dprev;
        break;
      }
    }
    if (code == 0) {
      step = n;
      value = dprev;
    }
  }
#pragma unroll
  for (int j = 0; j < S; ++j) {
    put(j, x[j]);
    put(S + j, y[j]);
  }
  put(2 * S, acc.s);
  put(2 * S + 1, acc.c);
  put(2 * S + 2, dprev);
  W(2 * S + 3) = (uint64_t)(uint8_t)code;
  W(2 * S + 4) = (uint64_t)step;
  put(2 * S + 5, value);
}

template <class Sys>
int shadow_launch_t(const cl_system* s, const ShadowArgs& a, hipStream_t st) {
  Sys sys;
  sys.K = s->K;
  hipLaunchKernelGGL((shadow_kernel<Sys>), dim3((unsigned)((a.m + 255) / 256)), dim3(256), 0, st, a, sys);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : (e == hipErrorOutOfMemory ? CL_ERR_OOM : CL_ERR_DEVICE_LOST);
}

int shadow_launch(const cl_system* s, const double* x0, const double* u0, uint64_t* e, int64_t m, int64_t n0,
                  int64_t steps, int64_t N, int tau, double d0, hipStream_t st) {
  if (m == 0) return 0;
  ShadowArgs a{x0, u0, e, m, n0, steps, N, tau, d0};
  if (s->kind == CL_SYS_STANDARD_MAP) return shadow_launch_t<StdMap>(s, a, st);
  if (s->kind == CL_SYS_WRAPPED_STANDARD_MAP) return shadow_launch_t<WrappedStdMap>(s, a, st);
  return CL_ERR_UNSUPPORTED;
}

// Row g of the recorder's array for the two-trajectory method (_shadow_record_kernel!, run_device.jl:74-98):
// lambda = total(sum) / target and the point; a failed launch sets fail = g + 1.
__global__ void k_shadow_rows(int64_t m, int s, int g, int G, int64_t target, const uint64_t* __restrict__ e,
                              int64_t* fail, double* R) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m || fail[i] != 0) return;
  if ((int8_t)(e[(int64_t)(2 * s + 3) * m + i] & 0xff) != 0) {
    fail[i] = g + 1;
    return;
  }
  const int ncol = 1 + 7 + s;
  double* row = R + ((size_t)i * G + g) * ncol;
  auto dw = [&](int w) { return __longlong_as_double((long long)e[(int64_t)w * m + i]); };
  row[0] = (dw(2 * s) + dw(2 * s + 1)) / (double)target;
  for (int j = 0; j < s; ++j) row[8 + j] = dw(j);
}

int shadow_rows_launch(int64_t m, int s, int g, int G, int64_t target, const uint64_t* e, int64_t* fail, double* R,
                       hipStream_t st) {
  if (m == 0) return 0;
  hipLaunchKernelGGL(k_shadow_rows, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, st, m, s, g, G, target, e, fail,
                     R);
  hipError_t err = hipGetLastError();
  return err == hipSuccess ? 0 : (err == hipErrorOutOfMemory ? CL_ERR_OOM : CL_ERR_DEVICE_LOST);
}

}  // namespace cl
