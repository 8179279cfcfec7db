// inst_pipe.hip - instances of the wave-specialised pipeline (pipe_kernel.hpp) for the two-dimensional maps:
// a single QR spectrum (k = 1 and 2, every factorisation), and the recorder's QR spectrum + RunMonitor chains.
#include "registry.hpp"
#include "pipe_kernel.hpp"

namespace cl {

template <class Sys, int ALG, int K, bool MON>
int pipe_launch(const cl_system* s, const cl_chain* d, const KArgs& a, uint8_t* rerun, hipStream_t st) {
  if (a.tau < 1 || a.tau > PIPE_TAU_MAX) return CL_ERR_UNSUPPORTED;
  Sys sys = make_sys<Sys>(*s);
  QRSpectrum<ALG, 2, K> qr;
  slot_setup(qr, d->r[0]);
  const int H = pipe_handoffs(a.tau);
  const int64_t grid = (a.m + PIPE_LANES - 1) / PIPE_LANES;
  hipLaunchKernelGGL((pipe_kernel<Sys, ALG, K, MON>), dim3((unsigned)grid), dim3(pipe_waves<ALG, K, MON>() * PIPE_LANES),
                     0, st, a, sys, qr, H, rerun);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : (e == hipErrorOutOfMemory ? CL_ERR_OOM : CL_ERR_DEVICE_LOST);
}

#define PADD(SK, SYS, ALG, K) \
  if (*n < cap) out[(*n)++] = PipeInstance{SK, K, ALG, 0, &pipe_launch<SYS, ALG, K, false>};
#define PMON(SK, SYS, ALG) \
  if (*n < cap) out[(*n)++] = PipeInstance{SK, 2, ALG, 1, &pipe_launch<SYS, ALG, 2, true>};
#define PRECORD(SK, SYS)                                                                          \
  PMON(SK, SYS, QR_CGS) PMON(SK, SYS, QR_MGS) PMON(SK, SYS, QR_RGS) PMON(SK, SYS, QR_HOUSEHOLDER)   \
  PMON(SK, SYS, QR_AUGMENTED)
#define PALL(SK, SYS, K)                                                                          \
  PADD(SK, SYS, QR_CGS, K) PADD(SK, SYS, QR_MGS, K) PADD(SK, SYS, QR_RGS, K)                      \
  PADD(SK, SYS, QR_HOUSEHOLDER, K) PADD(SK, SYS, QR_AUGMENTED, K) PADD(SK, SYS, QR_GIVENS, K)

void register_pipes(PipeInstance* out, int* n, int cap) {
  PALL(CL_SYS_STANDARD_MAP, StdMap, 2)
  PALL(CL_SYS_STANDARD_MAP, StdMap, 1)
  PALL(CL_SYS_WRAPPED_STANDARD_MAP, WrappedStdMap, 2)
  PALL(CL_SYS_WRAPPED_STANDARD_MAP, WrappedStdMap, 1)
  PADD(CL_SYS_STANDARD_MAP_K, StdMapK, QR_MGS, 2) PADD(CL_SYS_STANDARD_MAP_K, StdMapK, QR_HOUSEHOLDER, 2)
  PRECORD(CL_SYS_STANDARD_MAP, StdMap)
  PRECORD(CL_SYS_WRAPPED_STANDARD_MAP, WrappedStdMap)
}

}  // namespace cl
