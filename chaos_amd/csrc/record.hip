// record.hip - the recorder's grid rows on the device (src/Experiments/run_device.jl:9-37). After the launch that
// ends at grid step g, every orbit still running gets row g of R from its chain states (estimator, RunMonitor) and
// its point; an orbit whose launch failed gets fail = g + 1 (1-based grid row, as the reference's kernel writes) and
// no further rows. The launches themselves are the propagate kernels, so a row equals the final values of the same
// estimator run directly to that step.
#include <hip/hip_runtime.h>
#include "cl_core.hpp"
#include "../../include/chaoslab_hip.h"
#include "record.hpp"

namespace cl {

// R is orbit-major, R[(i G + g) ncol + col], ncol = k + 7 + s; states are the chain's words, SoA (st[w m + i]).
__global__ void k_record_rows(int64_t m, int s, int k, int benettin, int g, int G, const double* __restrict__ x,
                              const uint64_t* __restrict__ st, const int8_t* __restrict__ code, int64_t* fail,
                              double* R) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m || fail[i] != 0) return;
  if (code[i] != STATUS_OK) {
    fail[i] = g + 1;
    return;
  }
  const int ncol = k + 7 + s;
  double* row = R + ((size_t)i * G + g) * ncol;
  auto dword = [&](int w) { return __longlong_as_double((long long)st[(int64_t)w * m + i]); };
  double lam[2] = {0.0, 0.0}, orth = 0.0, gmax = 1.0;
  int wmon;
  if (!benettin) {  // QRState<k>: sums[k] (s, c), n, G_max, orth_max, count_max, code, column, value, threshold, step
    const int64_t n = (int64_t)st[(int64_t)(2 * k) * m + i];
    for (int j = 0; j < k; ++j) lam[j] = n == 0 ? 0.0 : (dword(2 * j) + dword(2 * j + 1)) / (double)n;
    gmax = dword(2 * k + 1);
    orth = dword(2 * k + 2);
    wmon = 2 * k + 9;
  } else {  // BenettinState: sum (s, c), n, code, value, step
    const int64_t n = (int64_t)st[2 * m + i];
    lam[0] = n == 0 ? 0.0 : (dword(0) + dword(1)) / (double)n;
    wmon = 6;
  }
  for (int j = 0; j < k; ++j) row[j] = lam[j];
  row[k] = dword(wmon);          // det_max
  row[k + 1] = dword(wmon + 1);  // symp_max
  row[k + 2] = orth;
  if (k == 2) {  // eps_sum and eps_pair of a 2-spectrum are both |lambda_1 + lambda_2|; running maxima over the rows
    const double e = cl_fabs(lam[0] + lam[1]);
    const double* prev = row - ncol;
    row[k + 3] = g > 0 ? cl_jmax(prev[k + 3], e) : e;
    row[k + 4] = g > 0 ? cl_jmax(prev[k + 4], e) : e;
  }
  row[k + 5] = gmax;
  row[k + 6] = dword(wmon + 2);  // M_max
  for (int j = 0; j < s; ++j) row[k + 7 + j] = x[(int64_t)j * m + i];
}

int record_rows_launch(int64_t m, int s, int k, int benettin, int g, int G, const double* x, const uint64_t* st,
                       const int8_t* code, int64_t* fail, double* R, hipStream_t stream) {
  if (m == 0) return 0;
  hipLaunchKernelGGL(k_record_rows, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, stream, m, s, k, benettin, g,
                     G, x, st, code, fail, R);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : (e == hipErrorOutOfMemory ? CL_ERR_OOM : CL_ERR_DEVICE_LOST);
}

// The indicator recorder's rows (src/Experiments/run_indicator_device.jl:13-29 and _put!, run_indicator_record.jl:39-70):
// the columns of each member's final (FLI: fli, fli1; MEGNO: Y, Ybar, Ybar20 of column 1 then 2; SALI: sali, L; RLI:
// DL, RLI), then
// det_max, symp_max and M_max of RunMonitor, then the point (orbit A's state on a pair). A window kind has no members
// (monitor and point only).
// R is orbit-major, R[(i G + g) ncol + col]; states SoA by word. `off` is each member's first state word.
__global__ void k_ind_rows(int64_t m, int s, IndRowDesc d, int g, int G, const double* __restrict__ x,
                           const uint64_t* __restrict__ st, const int8_t* __restrict__ code, int64_t* fail,
                           double* R) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m || fail[i] != 0) return;
  if (code[i] != STATUS_OK) {
    fail[i] = g + 1;
    return;
  }
  double* row = R + ((size_t)i * G + g) * d.ncol;
  auto dword = [&](int w) { return __longlong_as_double((long long)st[(int64_t)w * m + i]); };
  auto iword = [&](int w) { return (int64_t)st[(int64_t)w * m + i]; };
  int col = 0;
#pragma unroll
  for (int j = 0; j < IND_MAX_MEMBERS; ++j) {
    if (j >= d.n) break;
    const int o = d.off[j];
    if (d.kind[j] == CL_RED_FLI) {  // FLIState: s1, s2, f1, fB, value, n, ...  (fli.jl:91-94)
      row[col++] = dword(o + 5);
      row[col++] = dword(o + 4);
    } else if (d.kind[j] == CL_RED_MEGNO) {  // MEGNOState: c1, c2 (s, R1, R2, Q, Q2), value, n, K (megno.jl:97-111)
      const int64_t n = iword(o + 21), K = iword(o + 22);
      for (int c = 0; c < 2; ++c) {
        const int b = o + 10 * c;
        double Y = 0.0, Yb = 0.0, Y20 = 0.0;
        if (K != 0) {
          const double tn = (double)n;
          Y = 2.0 * (dword(b + 2) + dword(b + 3)) / tn;
          Yb = (dword(b + 6) + dword(b + 7)) / (double)K;
          Y20 = (double)(n / K) * (dword(b + 8) + dword(b + 9)) / (tn * tn * tn);
        }
        row[col++] = Y;
        row[col++] = Yb;
        row[col++] = Y20;
      }
    } else if (d.kind[j] == CL_RED_RLI) {  // RLIState: sA, sB, Q, value, n, K  (rli.jl:116-119)
      const int64_t n = iword(o + 7), K = iword(o + 8);
      double dl = 0.0, rli = 0.0;
      if (K != 0) {
        dl = cl_fabs((dword(o + 2) - dword(o + 0)) + (dword(o + 3) - dword(o + 1))) / (double)n;
        rli = (dword(o + 4) + dword(o + 5)) / (double)K;
      }
      row[col++] = dl;
      row[col++] = rli;
    } else {  // SALIState: s1, s2, sali, value, n  (sali.jl:62-66)
      row[col++] = dword(o + 4);
      row[col++] = iword(o + 6) == 0 ? 0.0 : -((dword(o + 0) + dword(o + 2)) + (dword(o + 1) + dword(o + 3)));
    }
  }
  row[col] = dword(d.wmon);          // det_max
  row[col + 1] = dword(d.wmon + 1);  // symp_max
  row[col + 2] = dword(d.wmon + 2);  // M_max
  for (int j = 0; j < s; ++j) row[col + 3 + j] = x[(int64_t)j * m + i];
}

int ind_rows_launch(int64_t m, int s, const IndRowDesc& d, int g, int G, const double* x, const uint64_t* st,
                    const int8_t* code, int64_t* fail, double* R, hipStream_t stream) {
  if (m == 0) return 0;
  hipLaunchKernelGGL(k_ind_rows, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, stream, m, s, d, g, G, x, st,
                     code, fail, R);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : (e == hipErrorOutOfMemory ? CL_ERR_OOM : CL_ERR_DEVICE_LOST);
}

}  // namespace cl
