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

// The persistent launch over several lattices (chain_mixed.hip): one class per site count, work items of up to
// `slots` orbits of one class taken from a device counter in list order.
struct MixedClass {
  ChainArgs a;  // the class's buffers, per-orbit step counts and launch order (n0, steps are the chunk's)
  int N;        // sites
  int slots;    // orbits per work item
};
struct MixedItem {
  int32_t cls;     // class index
  int32_t nslots;  // orbits in the item
  int64_t g0;      // first orbit slot (index into the class's launch order)
};
struct MixedArgs {
  const MixedClass* cls;   // device
  const MixedItem* items;  // device, in the order they are taken
  int64_t nitems;
  unsigned* counter;       // device work counter of this launch, zero at its start
  int64_t n0, steps;       // the chunk: global steps n0 .. n0 + steps (each orbit stops at its own count)
};
// Orbits per work item of the persistent launch for N sites (its LDS holds this many orbits' matrices), for the
// blocked order or the lane-butterfly sweep.
int mixed_slots(int N, bool blocked);
// Zeroes n work counters (a fill kernel, so the launch can be captured in a graph).
int mixed_zero(unsigned* counters, int n, hipStream_t st);
// Compute units of the current device (queried once; call before a graph capture).
int mixed_cus();
// One persistent launch over the items (ring or open chains, QRSpectrum(HouseholderQR()), k = 16, either order,
// Float64 frames or Float32 tangents).
int mixed_launch(const MixedArgs& a, bool ring, bool blocked, bool f32, hipStream_t st);

// 1 when a kernel is built for N sites and k columns.
int chain_supported(int N, int k);
// Launches the kernel for (N, ring, with or without QRSpectrum(HouseholderQR()), the blocked order on the matrix cores
// or the lane-butterfly sweep, Float64 frames or Float32 tangents); 0 or an error code.
int chain_launch(const ChainArgs& a, int N, bool ring, bool qr, bool blocked, bool f32, hipStream_t st);

}  // namespace cl
