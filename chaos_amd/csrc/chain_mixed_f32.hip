// chain_mixed_f32.hip - the persistent mixed-lattice launch with Float32 tangents (CL_ELT_F64_F32TAN), its own
// translation unit so the instances compile in parallel with the Float64 ones.
#include "chain_mixed_impl.hpp"

namespace cl {

namespace chain_detail {

int mixed_launch_f32(const MixedArgs& a, bool ring, bool blocked, int64_t grid, hipStream_t st) {
  return mixed_launch_ft<float>(a, ring, blocked, grid, st);
}

}  // namespace chain_detail

}  // namespace cl
