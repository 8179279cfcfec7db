// record.hip - the recorder's grid rows on the device (src/Experiments/run_device.jl:9-37). After the launch that
// ends at grid step g, every orbit still running gets row g of R from its chain states (estimator, RunMonitor) and
// its point; an orbit whose launch failed gets fail = g + 1 (1-based grid row, as the reference's kernel writes) and
// no further rows. The launches themselves are the propagate kernels, so a row equals the final values of the same
// estimator run directly to that step.
#include <hip/hip_runtime.h>
#include "cl_core.hpp"
#include "../../include/chaoslab_hip.h"

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

}  // namespace cl
