// registry.hpp - table of the built kernel instances, keyed by system, frame shape and chain.
//
// The reference compiles one kernel per (system type, reducer tuple type) on first use. Here the
// kernels are compiled ahead of time for gfx950; a chain that has no instance is reported as
// "kernel does not compile for the element type" (FALLBACK_REASONS[3]), as the reference's probe
// reports a chain its device cannot compile.
#pragma once
#include <hip/hip_runtime.h>
#include "orbit_kernels.hpp"

namespace cl {

struct Instance {
  int sys_kind, d, k, nf, words, nred;
  int elt;  // CL_ELT_F64 or CL_ELT_F64_F32TAN
  int red[CL_MAX_REDUCERS][3];  // kind, alg, frame
  int (*launch)(const cl_system*, const cl_chain*, const KArgs&, hipStream_t);
};

template <class Sys>
inline Sys make_sys(const cl_system& s);
template <>
inline StdMap make_sys<StdMap>(const cl_system& s) {
  StdMap m;
  m.K = s.K;
  return m;
}
template <>
inline WrappedStdMap make_sys<WrappedStdMap>(const cl_system& s) {
  WrappedStdMap m;
  m.K = s.K;
  return m;
}

template <>
inline PairMap<StdMap> make_sys<PairMap<StdMap>>(const cl_system& s) {
  PairMap<StdMap> m;
  m.base.K = s.K;
  return m;
}
template <>
inline PairMap<WrappedStdMap> make_sys<PairMap<WrappedStdMap>>(const cl_system& s) {
  PairMap<WrappedStdMap> m;
  m.base.K = s.K;
  return m;
}
template <>
inline StdMapK make_sys<StdMapK>(const cl_system&) {
  return StdMapK();
}
template <>
inline HHGauss make_sys<HHGauss>(const cl_system& s) {
  HHGauss m;
  m.h = s.h;
  m.tol = s.tol;
  m.maxiter = s.maxiter > 0 ? s.maxiter : 50;
  return m;
}

template <>
inline HHGaussComp make_sys<HHGaussComp>(const cl_system& s) {
  HHGaussComp m;
  m.h = s.h;
  m.tol = s.tol;
  m.maxiter = s.maxiter > 0 ? s.maxiter : 50;
  return m;
}

// Launch geometry: a big ensemble fills the chip with 256-lane groups; a small one (below 64K
// orbits) uses 64-lane groups so its waves spread over as many CUs as possible.
inline int block_for(int64_t m) { return m >= 65536 ? 256 : 64; }

template <class Sys, int K, int NF, class Chain, bool F32T = false>
int launch_impl(const cl_system* s, const cl_chain* d, const KArgs& a, hipStream_t st) {
  Sys sys = make_sys<Sys>(*s);
  Chain ch;
  chain_setup(ch, *d, 0);
  int block = a.block ? a.block : block_for(a.m);
  int64_t grid = (a.m + block - 1) / block;
  // the frames' LDS slots of systems that park them during the step (OrbitPlan), sized by the group
  size_t lds = OrbitPlan<Sys>::park_frames && !F32T ? (size_t)NF * Sys::D * K * block * sizeof(double) : 0;
  KArgs b = a;
  b.ws_lds = 0;
  if constexpr (HasStretch<Chain>::value) {
    // StretchingNumbers' counts staged in LDS, 2 (bins + 2) bytes per lane, when they fit in 64 KiB per group
    const size_t need = (size_t)2 * ((size_t)stretch_bins(ch) + 2) * (size_t)block;
    if (a.ws && need <= 65536 && lds == 0) {
      lds = need;
      b.ws_lds = 1;
    }
  }
  hipLaunchKernelGGL((orbit_kernel<Sys, K, NF, Chain, F32T>), dim3((unsigned)grid), dim3(block), lds, st, b, sys, ch);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : (e == hipErrorOutOfMemory ? CL_ERR_OOM : CL_ERR_DEVICE_LOST);
}

// Reducer-kind codes of the slot types, for the table keys.
template <class R>
struct Kind;
template <int ALG, int D, int K>
struct Kind<QRSpectrum<ALG, D, K>> {
  static constexpr int kind = CL_RED_QR, alg = ALG;
};
template <int D>
struct Kind<Benettin<D>> {
  static constexpr int kind = CL_RED_BENETTIN, alg = 0;
};
template <int D>
struct Kind<MEGNO<D>> {
  static constexpr int kind = CL_RED_MEGNO, alg = 0;
};
template <int D>
struct Kind<SALI<D>> {
  static constexpr int kind = CL_RED_SALI, alg = 0;
};
template <int D>
struct Kind<RLI<D>> {
  static constexpr int kind = CL_RED_RLI, alg = 0;
};
template <int D>
struct Kind<FLI<D>> {
  static constexpr int kind = CL_RED_FLI, alg = 0;
};
template <int D, int K>
struct Kind<OFLI<D, K>> {
  static constexpr int kind = CL_RED_OFLI, alg = 0;
};
template <int D, int K>
struct Kind<GALI<D, K>> {
  static constexpr int kind = CL_RED_GALI, alg = 0;
};
template <int D, int K>
struct Kind<Monitor<D, K>> {
  static constexpr int kind = CL_RED_MONITOR, alg = 0;
};
template <int D>
struct Kind<RunMon<D>> {
  static constexpr int kind = CL_RED_RUNMONITOR, alg = 0;
};
template <int D>
struct Kind<Stretch<D>> {
  static constexpr int kind = CL_RED_STRETCHING, alg = 0;
};
template <int H>
struct Kind<WPA<H>> {  // the window count is the key's alg
  static constexpr int kind = CL_RED_PHASE_ADVANCE, alg = H;
};
template <>
struct Kind<WBirk> {
  static constexpr int kind = CL_RED_BIRKHOFF, alg = 0;
};

inline void fill_key(Instance&, const Nil&, int) {}
template <class H, class T>
void fill_key(Instance& in, const Cons<H, T>& c, int i) {
  in.red[i][0] = Kind<decltype(c.head.r)>::kind;
  in.red[i][1] = Kind<decltype(c.head.r)>::alg;
  in.red[i][2] = H::frame;
  in.nred = i + 1;
  fill_key(in, c.tail, i + 1);
}

template <int SYSK, class Sys, int K, int NF, class Chain, bool F32T = false>
Instance make_instance() {
  Instance in;
  in.elt = F32T ? CL_ELT_F64_F32TAN : CL_ELT_F64;
  in.sys_kind = SYSK;
  in.d = Sys::D;
  in.k = K;
  in.nf = NF;
  in.words = ChainWords<Chain>::value;
  in.nred = 0;
  for (int i = 0; i < CL_MAX_REDUCERS; ++i) in.red[i][0] = in.red[i][1] = in.red[i][2] = 0;
  Chain c;
  fill_key(in, c, 0);
  in.launch = &launch_impl<Sys, K, NF, Chain, F32T>;
  return in;
}

// Each instance translation unit appends its entries.
void register_instances_map(Instance* out, int* n, int cap);
void register_instances_flow(Instance* out, int* n, int cap);
void register_instances_f32(Instance* out, int* n, int cap);

// Wave-specialised pipeline kernels for small map ensembles whose chain is one QRSpectrum on frame 1, alone or
// followed by RunMonitor (pipe_kernel.hpp). `launch` marks in `rerun` the orbits the caller must rerun with the fused kernel.
constexpr int PIPE_TAU_MAX = 16;  // largest tau the pipeline's LDS rings hold
struct PipeInstance {
  int sys_kind, k, alg, mon;  // mon: the chain is the QR spectrum followed by RunMonitor
  int (*launch)(const cl_system*, const cl_chain*, const KArgs&, uint8_t* rerun, hipStream_t);
};
void register_pipes(PipeInstance* out, int* n, int cap);

}  // namespace cl
