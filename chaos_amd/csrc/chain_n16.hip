// chain_n16.hip - instances of the chain kernel for N = 16 sites (ring and open, with and without the QR spectrum).
#include "chain_impl.hpp"

namespace cl {

int launch_chain_n16(const ChainArgs& a, bool ring, bool qr, bool blocked, hipStream_t st) {
  return chain_detail::launch_n<16, double>(a, ring, qr, blocked, st);
}

}  // namespace cl
