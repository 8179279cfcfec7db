// chain_mixed.hip - the persistent mixed-lattice launch (chain_mixed_impl.hpp) with Float64 frames, its host-side
// helpers, and the dispatch to the Float32-tangent instances (chain_mixed_f32.hip).
#include "chain_mixed_impl.hpp"

namespace cl {

namespace chain_detail {

int mixed_launch_f32(const MixedArgs& a, bool ring, bool blocked, int64_t grid, hipStream_t st);

__global__ void k_zero_counters(unsigned* c, int n) {
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i < n) c[i] = 0u;
}

}  // namespace chain_detail

int mixed_slots(int N, bool blocked) {
  using namespace chain_detail;
  if (N > 64) return 1;
  const int s = MIX_BLOCK / N, cap = blocked ? MIX_BLK_SLOTS : MIX_SLOTS_MAX;
  return s < cap ? s : cap;
}

int mixed_zero(unsigned* counters, int n, hipStream_t st) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(chain_detail::k_zero_counters, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, counters, n);
  return hipGetLastError() == hipSuccess ? 0 : CL_ERR_DEVICE_LOST;
}

int mixed_cus() {
  static int cus = 0;
  if (cus == 0) {
    int dev = 0, n = 0;
    if (hipGetDevice(&dev) == hipSuccess && hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess)
      cus = n;
    if (cus <= 0) cus = 256;
  }
  return cus;
}

int mixed_launch(const MixedArgs& a, bool ring, bool blocked, bool f32, hipStream_t st) {
  using namespace chain_detail;
  if (a.nitems <= 0) return 0;
  // three groups per compute unit (the LDS pool allows three; two for the blocked order at one wave per SIMD), never
  // more groups than items
  int64_t grid = (int64_t)mixed_cus() * (blocked && CL_MIX_BLK_WAVES == 1 ? 2 : 3);
  if (grid > a.nitems) grid = a.nitems;
  return f32 ? mixed_launch_f32(a, ring, blocked, grid, st) : mixed_launch_ft<double>(a, ring, blocked, grid, st);
}

}  // namespace cl
