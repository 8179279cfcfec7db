// chain_kernels.hip - dispatch of the coupled standard-map chain kernels (chain_impl.hpp); the instances
// for each site count are compiled in their own translation units (chain_n*.hip).
#include "chain_kernels.hpp"
#include "../../include/chaoslab_hip.h"

namespace cl {

int launch_chain_n8(const ChainArgs& a, bool ring, bool qr, bool blocked, hipStream_t st);
int launch_chain_n16(const ChainArgs& a, bool ring, bool qr, bool blocked, hipStream_t st);
int launch_chain_n32(const ChainArgs& a, bool ring, bool qr, bool blocked, hipStream_t st);
int launch_chain_n64(const ChainArgs& a, bool ring, bool qr, bool blocked, hipStream_t st);
int launch_chain_n128(const ChainArgs& a, bool ring, bool qr, bool blocked, hipStream_t st);
int launch_chain_n8_f32(const ChainArgs& a, bool ring, bool qr, bool blocked, hipStream_t st);
int launch_chain_n16_f32(const ChainArgs& a, bool ring, bool qr, bool blocked, hipStream_t st);
int launch_chain_n32_f32(const ChainArgs& a, bool ring, bool qr, bool blocked, hipStream_t st);
int launch_chain_n64_f32(const ChainArgs& a, bool ring, bool qr, bool blocked, hipStream_t st);
int launch_chain_n128_f32(const ChainArgs& a, bool ring, bool qr, bool blocked, hipStream_t st);

int chain_supported(int N, int k) {
  return (N == 8 || N == 16 || N == 32 || N == 64 || N == 128) && k == 16;
}

int chain_launch(const ChainArgs& a, int N, bool ring, bool qr, bool blocked, bool f32, hipStream_t st) {
  if (a.m == 0) return 0;
  if (f32) {
    switch (N) {
      case 8: return launch_chain_n8_f32(a, ring, qr, blocked, st);
      case 16: return launch_chain_n16_f32(a, ring, qr, blocked, st);
      case 32: return launch_chain_n32_f32(a, ring, qr, blocked, st);
      case 64: return launch_chain_n64_f32(a, ring, qr, blocked, st);
      case 128: return launch_chain_n128_f32(a, ring, qr, blocked, st);
    }
    return CL_ERR_UNSUPPORTED;
  }
  switch (N) {
    case 8: return launch_chain_n8(a, ring, qr, blocked, st);
    case 16: return launch_chain_n16(a, ring, qr, blocked, st);
    case 32: return launch_chain_n32(a, ring, qr, blocked, st);
    case 64: return launch_chain_n64(a, ring, qr, blocked, st);
    case 128: return launch_chain_n128(a, ring, qr, blocked, st);
  }
  return CL_ERR_UNSUPPORTED;
}

}  // namespace cl
