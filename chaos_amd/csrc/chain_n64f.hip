// chain_n64f.hip - instances of the chain kernel for N = 64 sites with Float32 tangents (CL_ELT_F64_F32TAN).
#include "chain_impl.hpp"

namespace cl {

int launch_chain_n64_f32(const ChainArgs& a, bool ring, bool qr, bool blocked, hipStream_t st) {
  return chain_detail::launch_n<64, float>(a, ring, qr, blocked, st);
}

}  // namespace cl
