// chain_kernels.hpp - launch interface of the coupled standard-map chain kernels (chain_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cl {

struct ChainArgs {
  double* x;             // [m][2N], orbit-major
  double* W;             // [m][k][2N]
  uint64_t* st;          // [m][2k + 9] words of QRState{Float64, k}
  int8_t* code;          // [m]
  int64_t* hstep;        // [m]
  const int64_t* Ns;     // per-orbit step counts (ragged full runs from n0 = 0) or null
  const int64_t* order;  // launch order of the orbits (a permutation) or null
  int64_t m, n0, steps, N;
  int tau;
  double K, eps, tol;
};

// 1 when a kernel is built for N sites and k columns.
int chain_supported(int N, int k);
// Launches the kernel for (N, ring, with or without QRSpectrum(HouseholderQR())); 0 or an error code.
int chain_launch(const ChainArgs& a, int N, bool ring, bool qr, hipStream_t st);

}  // namespace cl
