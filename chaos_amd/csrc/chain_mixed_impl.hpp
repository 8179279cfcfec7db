// chain_mixed_impl.hpp - one persistent launch over coupled-chain lattices of several site counts (N in {8, 16, 32, 64,
// 128}) with uneven per-orbit step counts. Each 128-thread group takes work items from a device counter until the
// list is exhausted; an item is up to `slots` orbits of one lattice, and the group runs them with that lattice's body
// (chain_orbits, the per-N kernels' own code, so an orbit's bits do not depend on the route). The host orders the
// items longest first, so the uneven lengths of every lattice pack onto the compute units together, and runs the
// launch in step chunks: an orbit whose count is reached, or that failed, retires and its items exit at once in
// later chunks.
#pragma once
#include "chain_impl.hpp"

namespace cl {

namespace chain_detail {

constexpr int MIX_BLOCK = 128;
constexpr int MIX_KC = 16;
// LDS pool (doubles): Us for 128 threads and 16 reflectors, then N = 128's exchange buffer and its two packed
// triangles; smaller N put their triangles where the exchange buffer would be (at most MIX_SLOTS_MAX orbits).
constexpr int MIX_US = 2 * MIX_KC * MIX_BLOCK;
constexpr int MIX_XB = 128 * MIX_KC;
constexpr int MIX_TRI = Tri<MIX_KC>::SIZE;
constexpr int MIX_POOL = MIX_US + MIX_XB + 2 * MIX_TRI;
constexpr int MIX_SLOTS_MAX = (MIX_POOL - MIX_US) / (2 * MIX_TRI);
// The blocked order (qr_handoff_blk): the frames of the group (32 doubles per thread), then N = 128's exchange buffer and
// its two 16 x 16 matrices and triangle, or for smaller N those of at most MIX_BLK_SLOTS orbits (three groups per
// compute unit fit the LDS either way).
constexpr int MIX_XS = 32 * MIX_BLOCK;
constexpr int MIX_BLK_SLOT = 2 * MSZ + MIX_TRI;
constexpr int MIX_BLK_SLOTS = 4;
constexpr int MIX_POOL_BLK = MIX_XS + (MIX_XB + MIX_BLK_SLOT > MIX_BLK_SLOTS * MIX_BLK_SLOT ? MIX_XB + MIX_BLK_SLOT
                                                                                           : MIX_BLK_SLOTS * MIX_BLK_SLOT);

// Each lattice's body is its own function: inlined into the one kernel, the five bodies spilled 1.7 KB per lane even
// with the whole register file (C3, blocked order: 741 ms, against 512 ms as separate functions; the butterfly order
// 640 -> 617 ms).
#define CL_MIX_ITEM_ATTR __attribute__((noinline))
template <int NS, bool RING, bool BLK, class FT>
__device__ CL_MIX_ITEM_ATTR void mixed_item(const ChainArgs& a, const MixedItem& w, double* pool) {
  if constexpr (BLK) {
    constexpr int SL = NS > 64 ? 1 : (MIX_BLOCK / NS < MIX_BLK_SLOTS ? MIX_BLOCK / NS : MIX_BLK_SLOTS);
    double* xb = NS > 64 ? pool + MIX_XS : nullptr;
    double* Gs = pool + MIX_XS + (NS > 64 ? MIX_XB : 0);
    double* Ts = Gs + SL * MSZ;
    double* Rs = Ts + SL * MSZ;
    chain_orbits<NS, MIX_KC, RING, true, MIX_BLOCK, true, SL, FT>(a, w.g0, w.nslots, xb, pool, Rs, nullptr, Gs, Ts);
  } else {
    double* Us = pool;
    if constexpr (NS > 64) {
      chain_orbits<NS, MIX_KC, RING, true, MIX_BLOCK, false, Geo<NS>::OPB, FT>(a, w.g0, w.nslots, pool + MIX_US, Us,
                                                                                pool + MIX_US + MIX_XB,
                                                                                pool + MIX_US + MIX_XB + MIX_TRI);
    } else {
      double* Rs = pool + MIX_US;
      chain_orbits<NS, MIX_KC, RING, true, MIX_BLOCK, false, Geo<NS>::OPB, FT>(a, w.g0, w.nslots, nullptr, Us, Rs,
                                                                                Rs + MIX_SLOTS_MAX * MIX_TRI);
    }
  }
}

// The lane-butterfly order at two waves per SIMD (256 registers): on the C3 workload (the five lattices, 2048 orbits
// each, uneven lengths) 617 ms. The blocked order at one wave per SIMD (the whole register file): 512 ms, against
// 683 ms at two.
#ifndef CL_MIX_BLK_WAVES
#define CL_MIX_BLK_WAVES 1
#endif
template <bool RING, bool BLK, class FT>
__global__ void __launch_bounds__(MIX_BLOCK) __attribute__((amdgpu_waves_per_eu(BLK ? CL_MIX_BLK_WAVES : 2)))
chain_mixed_kernel(MixedArgs m) {
  __shared__ double pool[BLK ? MIX_POOL_BLK : MIX_POOL];
  __shared__ unsigned item_s;
  for (;;) {
    __syncthreads();  // the previous item's LDS use is over
    if (threadIdx.x == 0) item_s = atomicAdd(m.counter, 1u);
    __syncthreads();
    const int64_t it = (int64_t)item_s;
    if (it >= m.nitems) return;  // every thread of the group reads the same value: the exit is uniform
    const MixedItem w = m.items[it];
    const MixedClass& c = m.cls[w.cls];
    ChainArgs a = c.a;
    a.n0 = m.n0;
    a.steps = m.steps;
    switch (c.N) {
      case 8: mixed_item<8, RING, BLK, FT>(a, w, pool); break;
      case 16: mixed_item<16, RING, BLK, FT>(a, w, pool); break;
      case 32: mixed_item<32, RING, BLK, FT>(a, w, pool); break;
      case 64: mixed_item<64, RING, BLK, FT>(a, w, pool); break;
      case 128: mixed_item<128, RING, BLK, FT>(a, w, pool); break;
      default: break;
    }
  }
}

// One persistent launch of the kernel for (ring, blocked order, frame type) on `grid` groups.
template <class FT>
int mixed_launch_ft(const MixedArgs& a, bool ring, bool blocked, int64_t grid, hipStream_t st) {
  if (ring)
    hipLaunchKernelGGL((blocked ? chain_mixed_kernel<true, true, FT> : chain_mixed_kernel<true, false, FT>),
                       dim3((unsigned)grid), dim3(MIX_BLOCK), 0, st, a);
  else
    hipLaunchKernelGGL((blocked ? chain_mixed_kernel<false, true, FT> : chain_mixed_kernel<false, false, FT>),
                       dim3((unsigned)grid), dim3(MIX_BLOCK), 0, st, a);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : (e == hipErrorOutOfMemory ? CL_ERR_OOM : CL_ERR_DEVICE_LOST);
}

}  // namespace chain_detail

}  // namespace cl
