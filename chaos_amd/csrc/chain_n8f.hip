// chain_n8f.hip - instances of the chain kernel for N = 8 sites with Float32 tangents (CL_ELT_F64_F32TAN).
#include "chain_impl.hpp"

namespace cl {

int launch_chain_n8_f32(const ChainArgs& a, bool ring, bool qr, bool blocked, hipStream_t st) {
  return chain_detail::launch_n<8, float>(a, ring, qr, blocked, st);
}

}  // namespace cl
