// record.hpp - the recorders' row kernels (record.hip), launched by the C ABI in propagate.hip.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cl {

constexpr int IND_MAX_MEMBERS = 4;
// The members of an indicator chain (CL_RED_FLI, CL_RED_MEGNO, CL_RED_SALI) with their first state words, the
// RunMonitor's first word and the row width.
struct IndRowDesc {
  int n, kind[IND_MAX_MEMBERS], off[IND_MAX_MEMBERS], wmon, ncol;
};

int record_rows_launch(int64_t m, int s, int k, int benettin, int g, int G, const double* x, const uint64_t* st,
                       const int8_t* code, int64_t* fail, double* R, hipStream_t stream);
int ind_rows_launch(int64_t m, int s, const IndRowDesc& d, int g, int G, const double* x, const uint64_t* st,
                    const int8_t* code, int64_t* fail, double* R, hipStream_t stream);

}  // namespace cl
