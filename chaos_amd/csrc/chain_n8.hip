// chain_n8.hip - instances of the chain kernel for N = 8 sites (ring and open, with and without the QR spectrum).
#include "chain_impl.hpp"

namespace cl {

int launch_chain_n8(const ChainArgs& a, bool ring, bool qr, bool blocked, hipStream_t st) {
  return chain_detail::launch_n<8, double>(a, ring, qr, blocked, st);
}

}  // namespace cl
