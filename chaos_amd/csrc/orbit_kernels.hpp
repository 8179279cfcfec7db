// orbit_kernels.hpp - the fused ensemble kernel: one orbit per lane, structure-of-arrays HBM
// layout, state/frames/reducer states in VGPRs for the whole launch.
//
// Per step (src/Tangent/orbit.jl:48-58): the tangent at the OLD point is applied to every frame
// with the explicit-fma product, then the map advances; sin and cos of the new point are formed
// once and serve the next Jacobian, the next step and any reducer that reads the field. Every
// `tau` steps the handoff checks finiteness (orbit.jl:53-55) and runs the reducer chain in order
// (orbit.jl:23-32,56-57), each reducer seeing the frame as replaced by the reducers before it.
#pragma once
#include <stdint.h>
#include <type_traits>
#include "cl_core.hpp"
#include "cl_gauss.hpp"
#include "../../include/chaoslab_hip.h"

namespace cl {

struct KArgs {
  double* x;     // [S][m]
  double* W0;    // [D*K][m]
  double* W1;    // second frame or null
  uint64_t* st;  // [words][m]
  int8_t* code;  // [m]
  int64_t* hstep;
  int64_t m, n0, steps, N;
  int tau;
  // optional per-handoff feature trace (float [3 * trace_len][m]); handoff t (1-based, counted from
  // trace_n0) writes rows 3(t-1) .. 3(t-1)+2 while t <= trace_len
  float* trace;
  int64_t trace_n0;
  int trace_len;
  // optional orbit mask: orbits with mask[i] == 0 are left untouched
  const uint8_t* mask;
  // reducer workspace [entries][m] (UInt32), or null when the chain requests none
  uint32_t* ws;
  // WeightedBirkhoff weight tables (one per horizon), or null
  const double* wtab;
  // restart the frame from the unit columns after every handoff (CL_CHAIN_RESET_FRAME)
  int reset_frame;
  // lanes per group of the launch (64, 128 or 256), 0 for the size rule
  int block;
  // StretchingNumbers: stage the counts in LDS (dynamic shared memory, 2 (bins + 2) bytes per lane)
  int ws_lds;
};

// Per-orbit context the reducers read: the point, the carried sin/cos, the frames and a cache of
// the column log-norms of each frame (valid until a reducer replaces that frame).
template <class Sys, int K, int NF>
struct OrbitCtx {
  Sys sys;
  double x[Sys::S];
  typename Sys::Aux aux;
  Frame<Sys::D, K> W[NF];
  bool logs_valid[NF];
  ColLogs logs[NF];
  uint32_t* ws;  // the orbit's first workspace entry (entries m apart)
  int64_t wstride;
  uint8_t* cnt;  // this lane's staged StretchingNumbers counts in LDS (entries cstride apart), or null
  int cstride;
  double* mem;   // the orbit's first WeightedBirkhoff horizon word in the state array (words m apart)
  const double* wtab;
};

// ------------------------------------------------------------------------------ chain slots
template <class R, int FR>
struct Slot {
  R r;
  typename R::State st;
  static constexpr int frame = FR;
};

struct Nil {};
template <class H, class T>
struct Cons {
  H head;
  T tail;
};

template <int D, int K, int NF, class Ctx>
CL_HD void ensure_logs(Ctx& c, int fr) {
  if (!c.logs_valid[fr]) {
    ColLogs L = col_logs(c.W[fr]);
    if (L.ok1 && L.ok2) {
      L.l1 = cl_log(L.a1);
      L.l2 = cl_log(L.a2);
    }
    c.logs[fr] = L;
    c.logs_valid[fr] = true;
  }
}

// One reducer at a handoff, by reducer type.
template <int ALG, int D, int K, int FR, class Ctx>
CL_HD void slot_step(Slot<QRSpectrum<ALG, D, K>, FR>& s, int64_t n, Ctx& c) {
  s.r.step(s.st, n, c.W[FR]);
  c.logs_valid[FR] = false;
}
template <int D, int FR, class Ctx>
CL_HD void slot_step(Slot<Benettin<D>, FR>& s, int64_t n, Ctx& c) {
  s.r.step(s.st, n, c.W[FR]);
  c.logs_valid[FR] = false;
}
template <int D, int FR, class Ctx>
CL_HD void slot_step(Slot<MEGNO<D>, FR>& s, int64_t n, Ctx& c) {
  ensure_logs<D, 2, 0>(c, FR);
  s.r.step_logs(s.st, n, c.W[FR], c.logs[FR]);
  if (s.r.replace) c.logs_valid[FR] = false;
}
template <int D, int FR, class Ctx>
CL_HD void slot_step(Slot<FLI<D>, FR>& s, int64_t n, Ctx& c) {
  ensure_logs<D, 2, 0>(c, FR);
  s.r.step_logs(s.st, n, c.W[FR], c.logs[FR]);
  if (s.r.replace) c.logs_valid[FR] = false;
}
template <int D, int FR, class Ctx>
CL_HD void slot_step(Slot<RLI<D>, FR>& s, int64_t n, Ctx& c) {
  ensure_logs<D, 2, 0>(c, FR);
  s.r.step_logs(s.st, n, c.W[FR], c.logs[FR]);
  if (s.r.replace) c.logs_valid[FR] = false;
}
template <int D, int FR, class Ctx>
CL_HD void slot_step(Slot<SALI<D>, FR>& s, int64_t n, Ctx& c) {
  ensure_logs<D, 2, 0>(c, FR);
  s.r.step_logs(s.st, n, c.W[FR], c.logs[FR]);
  c.logs_valid[FR] = false;
}
template <int D, int K, int FR, class Ctx>
CL_HD void slot_step(Slot<OFLI<D, K>, FR>& s, int64_t n, Ctx& c) {
  double f[D];
  c.sys.field(c.x, c.aux, f);
  if constexpr (K == 2) {
    ensure_logs<D, 2, 0>(c, FR);
    s.r.step(s.st, n, f, c.W[FR], &c.logs[FR]);
  } else {
    s.r.step(s.st, n, f, c.W[FR]);
  }
  if (s.r.replace) c.logs_valid[FR] = false;
}
template <int D, int K, int FR, class Ctx>
CL_HD void slot_step(Slot<GALI<D, K>, FR>& s, int64_t n, Ctx& c) {
  s.r.step(s.st, n, c.W[FR]);
  c.logs_valid[FR] = false;
}

template <int D, int K, int FR, class Ctx>
CL_HD void slot_step(Slot<Monitor<D, K>, FR>& s, int64_t n, Ctx& c) {
  double J[D * D];
  c.sys.jacobian(c.aux, J);
  s.r.step(s.st, n, J, c.W[FR]);
}

template <int D, int FR, class Ctx>
CL_HD void slot_step(Slot<Stretch<D>, FR>& s, int64_t n, Ctx& c) {
  ensure_logs<D, 2, 0>(c, FR);
  s.r.step_logs(s.st, n, c.W[FR], c.logs[FR], c.ws, c.wstride, c.cnt, c.cstride);
  if (s.r.replace) c.logs_valid[FR] = false;
}

template <int H, int FR, class Ctx>
CL_HD void slot_step(Slot<WPA<H>, FR>& s, int64_t n, Ctx& c) {
  s.r.template step<decltype(c.sys)>(s.st, n, c.x);
}
template <int FR, class Ctx>
CL_HD void slot_step(Slot<WBirk, FR>& s, int64_t n, Ctx& c) {
  s.r.template step<decltype(c.sys)>(s.st, n, c.x, c.mem, c.wstride, c.wtab);
}

// The workspace of a fresh run is zeroed by the first launch (src/Tangent/propagate.jl:29-31).
template <class R, int FR, class Ctx>
CL_HD void slot_ws_init(const Slot<R, FR>&, Ctx&) {}
template <int D, int FR, class Ctx>
CL_HD void slot_ws_init(const Slot<Stretch<D>, FR>& s, Ctx& c) {
  const int64_t e = 2 * ((int64_t)s.r.bins + 2);
  for (int64_t j = 0; j < e; ++j) c.ws[j * c.wstride] = 0u;
}
// WeightedBirkhoff's horizon sums start at zero (CompensatedSum() and S = 0 are all-zero words).
template <int FR, class Ctx>
CL_HD void slot_ws_init(const Slot<WBirk, FR>& s, Ctx& c) {
  for (int j = 0; j < s.r.p.hw; ++j) c.mem[(int64_t)j * c.wstride] = 0.0;
}
template <class Ctx>
CL_HD void chain_ws_init(const Nil&, Ctx&) {}
template <class H, class T, class Ctx>
CL_HD void chain_ws_init(const Cons<H, T>& ch, Ctx& c) {
  slot_ws_init(ch.head, c);
  chain_ws_init(ch.tail, c);
}

// The staged StretchingNumbers counts of a launch, added to the workspace at its end.
template <class R, int FR, class Ctx>
CL_HD void slot_ws_flush(const Slot<R, FR>&, Ctx&) {}
template <int D, int FR, class Ctx>
CL_HD void slot_ws_flush(const Slot<Stretch<D>, FR>& s, Ctx& c) {
  s.r.flush(c.cnt, c.cstride, c.ws, c.wstride);
}
template <class Ctx>
CL_HD void chain_ws_flush(const Nil&, Ctx&) {}
template <class H, class T, class Ctx>
CL_HD void chain_ws_flush(const Cons<H, T>& ch, Ctx& c) {
  slot_ws_flush(ch.head, c);
  chain_ws_flush(ch.tail, c);
}
// Whether a chain holds a StretchingNumbers slot, and its number of bins (the staging size).
template <class Chain>
struct HasStretch : std::false_type {};
template <int D, int FR, class T>
struct HasStretch<Cons<Slot<Stretch<D>, FR>, T>> : std::true_type {};
template <class H, class T>
struct HasStretch<Cons<H, T>> : HasStretch<T> {};
template <int D, int FR, class T>
CL_HD int stretch_bins(const Cons<Slot<Stretch<D>, FR>, T>& ch) { return ch.head.r.bins; }
template <class H, class T>
CL_HD int stretch_bins(const Cons<H, T>& ch) { return stretch_bins(ch.tail); }
CL_HD inline int stretch_bins(const Nil&) { return 0; }

template <int D, int FR, class Ctx>
CL_HD void slot_step(Slot<RunMon<D>, FR>& s, int64_t n, Ctx& c) {
  double J[D * D];
  c.sys.jacobian(c.aux, J);
  s.r.step(s.st, n, J);
}

// Reducer parameters from the descriptor.
template <int ALG, int D, int K>
CL_HD void slot_setup(QRSpectrum<ALG, D, K>& r, const cl_reducer& d) { r.tol = d.tol; }
template <int D>
CL_HD void slot_setup(Benettin<D>&, const cl_reducer&) {}
template <int D>
CL_HD void slot_setup(MEGNO<D>& r, const cl_reducer& d) { r.replace = d.replace != 0; }
template <int D>
CL_HD void slot_setup(FLI<D>& r, const cl_reducer& d) { r.replace = d.replace != 0; }
template <int D>
CL_HD void slot_setup(SALI<D>&, const cl_reducer&) {}
template <int D>
CL_HD void slot_setup(RLI<D>& r, const cl_reducer& d) { r.replace = d.replace != 0; }
template <int D, int K>
CL_HD void slot_setup(OFLI<D, K>& r, const cl_reducer& d) { r.replace = d.replace != 0; }
template <int D, int K>
CL_HD void slot_setup(GALI<D, K>&, const cl_reducer&) {}
template <int D>
CL_HD void slot_setup(RunMon<D>&, const cl_reducer&) {}
template <int H>
inline void slot_setup(WPA<H>& r, const cl_reducer& d) {
  const cl_birkhoff& b = *d.birkhoff;
  r.wmax = 0;
  for (int h = 0; h < H; ++h) {
    r.window[h] = b.window[h];
    r.wmax = r.wmax > b.window[h] ? r.wmax : b.window[h];
  }
  r.order = b.order;
  r.obs = b.kind[0];
  r.ia = b.a[0];
  r.ib = b.b[0];
}
inline void slot_setup(WBirk& r, const cl_reducer& d) {
  const cl_birkhoff& b = *d.birkhoff;
  r.p.H = b.H;
  r.p.K = b.K;
  r.p.order = b.order;
  int64_t mx = 0;
  for (int h = 0; h < b.H; ++h) mx = b.window[h] > mx ? b.window[h] : mx;
  for (int k = 0; k < CL_WB_MAX_OBSERVABLES; ++k) {
    r.p.kind[k] = k < b.K ? b.kind[k] : 0;
    r.p.a[k] = k < b.K ? b.a[k] : 0;
    r.p.b[k] = k < b.K ? b.b[k] : 0;
  }
  r.p.nmax2 = 2 * mx;
  r.p.hw = b.H * (2 + 5 * b.K);
}
template <int D>
CL_HD void slot_setup(Stretch<D>& r, const cl_reducer& d) {
  r.replace = d.replace != 0;
  r.lo = d.lo;
  r.hi = d.hi;
  r.bins = d.bins;
}
template <int D, int K>
CL_HD void slot_setup(Monitor<D, K>& r, const cl_reducer& d) {
  r.det_tol = d.tol;
  r.orth_tol = d.tol2;
}

// ------------------------------------------------------------------------------ chain recursion
CL_HD void chain_setup(Nil&, const cl_chain&, int) {}
template <class H, class T>
CL_HD void chain_setup(Cons<H, T>& c, const cl_chain& d, int i) {
  slot_setup(c.head.r, d.r[i]);
  chain_setup(c.tail, d, i + 1);
}
CL_HD void chain_init(Nil&) {}
template <class H, class T>
CL_HD void chain_init(Cons<H, T>& c) {
  c.head.r.init(c.head.st);
  chain_init(c.tail);
}
template <class Ctx>
CL_HD void chain_step(Nil&, int64_t, Ctx&) {}
template <class H, class T, class Ctx>
CL_HD void chain_step(Cons<H, T>& c, int64_t n, Ctx& ctx) {
  slot_step(c.head, n, ctx);
  chain_step(c.tail, n, ctx);
}
template <class S>
CL_HD bool state_failed(const S& s) {
  return s.code != 0;
}
CL_HD bool state_failed(const RunMonState&) { return false; }
CL_HD bool chain_failed(const Nil&) { return false; }
template <class H, class T>
CL_HD bool chain_failed(const Cons<H, T>& c) {
  return state_failed(c.head.st) || chain_failed(c.tail);
}
template <class C>
struct ChainWords;
template <>
struct ChainWords<Nil> {
  static constexpr int value = 0;
};
template <class H, class T>
struct ChainWords<Cons<H, T>> {
  static constexpr int value = (int)(sizeof(decltype(H::st)) / 8) + ChainWords<T>::value;
};

template <class S>
CL_HD void state_load(S& s, const uint64_t* st, int64_t m, int64_t i, int w0) {
  constexpr int NW = sizeof(S) / 8;
  uint64_t tmp[NW];
#pragma unroll
  for (int j = 0; j < NW; ++j) tmp[j] = st[(int64_t)(w0 + j) * m + i];
  __builtin_memcpy(&s, tmp, sizeof(S));
}
template <class S>
CL_HD void state_store(const S& s, uint64_t* st, int64_t m, int64_t i, int w0) {
  constexpr int NW = sizeof(S) / 8;
  uint64_t tmp[NW];
  __builtin_memcpy(tmp, &s, sizeof(S));
#pragma unroll
  for (int j = 0; j < NW; ++j) st[(int64_t)(w0 + j) * m + i] = tmp[j];
}
CL_HD int chain_load(Nil&, const uint64_t*, int64_t, int64_t, int w) { return w; }
template <class H, class T>
CL_HD int chain_load(Cons<H, T>& c, const uint64_t* st, int64_t m, int64_t i, int w) {
  state_load(c.head.st, st, m, i, w);
  return chain_load(c.tail, st, m, i, w + (int)(sizeof(c.head.st) / 8));
}
// WeightedBirkhoff, the last reducer of its chain: the register state (n, handoffs, code) follows the horizon sums
// in the Julia layout of WBState.
template <int FR>
CL_HD int chain_load(Cons<Slot<WBirk, FR>, Nil>& c, const uint64_t* st, int64_t m, int64_t i, int w) {
  state_load(c.head.st, st, m, i, w + c.head.r.p.hw);
  return w + c.head.r.p.hw + 2;
}
template <class C>
struct WBOffset {  // state words before a trailing WeightedBirkhoff
  static constexpr int value = 0;
};
template <class H, class T>
struct WBOffset<Cons<H, T>> {
  static constexpr int value = (int)(sizeof(decltype(H::st)) / 8) + WBOffset<T>::value;
};
template <int FR>
struct WBOffset<Cons<Slot<WBirk, FR>, Nil>> {
  static constexpr int value = 0;
};
CL_HD int chain_store(const Nil&, uint64_t*, int64_t, int64_t, int w) { return w; }
template <int FR>
CL_HD int chain_store(const Cons<Slot<WBirk, FR>, Nil>& c, uint64_t* st, int64_t m, int64_t i, int w) {
  state_store(c.head.st, st, m, i, w + c.head.r.p.hw);
  return w + c.head.r.p.hw + 2;
}
template <class H, class T>
CL_HD int chain_store(const Cons<H, T>& c, uint64_t* st, int64_t m, int64_t i, int w) {
  state_store(c.head.st, st, m, i, w);
  return chain_store(c.tail, st, m, i, w + (int)(sizeof(c.head.st) / 8));
}
// Feature trace of the prefix pass. A chain that starts MEGNO (frame 0) then FLI (frame 0) records at
// each handoff: MEGNO's Y of column 1 (2 R1 / n), FLI_B, and the log stretch log(alpha_1) of the handoff.
template <class Chain, class Ctx>
CL_HD void trace_hook(const Chain&, const Ctx&, int64_t, const KArgs&, int64_t) {}
template <int D, class Rest, class Ctx>
CL_HD void trace_hook(const Cons<Slot<MEGNO<D>, 0>, Cons<Slot<FLI<D>, 0>, Rest>>& ch, const Ctx& c, int64_t n,
                      const KArgs& a, int64_t i) {
  if (a.trace == nullptr) return;
  int64_t t = (n - a.trace_n0) / a.tau;
  if (t < 1 || t > a.trace_len) return;
  double y = 2.0 * csum_total(ch.head.st.c1.R1) / (double)n;
  double f = ch.tail.head.st.fB;
  double l = c.logs[0].l1;
  float* o = a.trace + (int64_t)(3 * (t - 1)) * a.m + i;
  o[0] = (float)y;
  o[a.m] = (float)f;
  o[2 * a.m] = (float)l;
}

template <class Sys, int K, int NF>
CL_HD int8_t frame_status(const OrbitCtx<Sys, K, NF>& c) {
  bool xf = true;
#pragma unroll
  for (int i = 0; i < Sys::S; ++i) xf = xf && cl_isfinite(c.x[i]);
  if (!xf) return STATUS_STATE;
  bool wf = true;
#pragma unroll
  for (int f = 0; f < NF; ++f) wf = wf && frame_finite(c.W[f]);
  return wf ? STATUS_OK : STATUS_W;
}

// Status of the point and of carried Float32 frames (the FP32-tangent element type).
template <class Sys, int K, int NF, class Ctx>
CL_HD int8_t frame_status_f(const Ctx& c, const Frame<Sys::D, K, float> (&Wf)[NF]) {
  bool xf = true;
#pragma unroll
  for (int i = 0; i < Sys::S; ++i) xf = xf && cl_isfinite(c.x[i]);
  if (!xf) return STATUS_STATE;
  bool wf = true;
#pragma unroll
  for (int f = 0; f < NF; ++f) wf = wf && frame_finite(Wf[f]);
  return wf ? STATUS_OK : STATUS_W;
}

// ------------------------------------------------------------------------------ the kernel
// Systems whose step needs most of the register file (the Gauss stage solve and its Jacobian) park the frames in
// LDS while the step solves its stages (a slot per thread, written after the tangent product and read back before
// the next one), so that they are not live across the solve, and ask for two waves per SIMD. On the Henon-Heiles
// workload (196608 orbits, two 4 x 4 frames, MEGNO + GALI and the Givens QR spectrum) the launch takes 5.31 s this
// way against 6.05 s with the frames in registers at one wave per SIMD, 6.12 s parked at one wave, 7.20 s parked at
// three, and 5.55 s (6.00 s unparked) when the reducer states also go to HBM between handoffs. Other systems keep
// their frames in registers for the whole launch.
template <class Sys>
struct OrbitPlan {
  static constexpr bool park_frames = false;
  static constexpr int waves = 1;
};
template <>
struct OrbitPlan<HHGauss> {
  static constexpr bool park_frames = true;
  static constexpr int waves = 2;
};
template <>
struct OrbitPlan<HHGaussComp> : OrbitPlan<HHGauss> {};

// The frames' LDS slots (dynamic shared memory, NF d k doubles per thread): entry e of thread t at
// park[e * blockDim + t], conflict-free 8-byte accesses. The empty asm with a memory clobber keeps the compiler from
// forwarding the stored values in registers past the solve.
template <int D, int K, int NF>
__device__ __attribute__((always_inline)) inline void park_store(double* park, const Frame<D, K> (&W)[NF]) {
#pragma unroll
  for (int f = 0; f < NF; ++f)
#pragma unroll
    for (int e = 0; e < D * K; ++e) park[(f * D * K + e) * blockDim.x + threadIdx.x] = W[f].w[e];
  __asm__ volatile("" ::: "memory");
}
template <int D, int K, int NF>
__device__ __attribute__((always_inline)) inline void park_load(const double* park, Frame<D, K> (&W)[NF]) {
  __asm__ volatile("" ::: "memory");
#pragma unroll
  for (int f = 0; f < NF; ++f)
#pragma unroll
    for (int e = 0; e < D * K; ++e) W[f].w[e] = park[(f * D * K + e) * blockDim.x + threadIdx.x];
}

// F32T: the FP32-tangent element type. The frames are carried in Float32 between handoffs (the input frames are
// rounded to Float32 at the start of the launch); at each handoff they are widened exactly to double, the chain runs
// in FP64 as for the Float64 type, and the frames it leaves are rounded back to Float32. The orbit stays FP64.
template <class Sys, int K, int NF, class Chain, bool F32T = false>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(OrbitPlan<Sys>::waves)))
orbit_kernel(KArgs a, Sys sys, Chain chain0) {
  constexpr bool PF = OrbitPlan<Sys>::park_frames && !F32T;
  extern __shared__ double park[];
  constexpr int D = Sys::D;
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.m) return;
  if (a.mask && !a.mask[i]) return;
  if (!(a.n0 == 0 || a.code[i] == STATUS_OK)) return;
  OrbitCtx<Sys, K, NF> c;
  c.sys = sys;
  Frame<D, K, float> Wf[NF];
#pragma unroll
  for (int j = 0; j < Sys::S; ++j) c.x[j] = a.x[(int64_t)j * a.m + i];
#pragma unroll
  for (int e = 0; e < D * K; ++e) c.W[0].w[e] = a.W0[(int64_t)e * a.m + i];
  if (NF > 1) {
#pragma unroll
    for (int e = 0; e < D * K; ++e) c.W[NF - 1].w[e] = a.W1[(int64_t)e * a.m + i];
  }
  if constexpr (F32T) {
#pragma unroll
    for (int f = 0; f < NF; ++f)
#pragma unroll
      for (int e = 0; e < D * K; ++e) Wf[f].w[e] = (float)c.W[f].w[e];
  }
  Chain ch = chain0;
  c.ws = a.ws ? a.ws + i : nullptr;
  c.wstride = a.m;
  c.cnt = nullptr;
  c.cstride = (int)blockDim.x;
  if constexpr (HasStretch<Chain>::value) {
    if (a.ws_lds && c.ws) {
      c.cnt = reinterpret_cast<uint8_t*>(park) + threadIdx.x;
      const int e = 2 * (stretch_bins(chain0) + 2);
      for (int j = 0; j < e; ++j) c.cnt[j * c.cstride] = 0;
    }
  }
  c.mem = (double*)(a.st + (int64_t)WBOffset<Chain>::value * a.m + i);
  c.wtab = a.wtab;
  if (a.n0 == 0) {
    chain_init(ch);
    chain_ws_init(ch, c);
  } else {
    chain_load(ch, a.st, a.m, i, 0);
  }
  int64_t n = a.n0;
  int8_t code = STATUS_OK;
  bool early = false;
  sys.aux(c.x, c.aux);
  const int64_t nh = a.steps / a.tau;
  for (int64_t h = 0; h < nh; ++h) {
    for (int t = 0; t < a.tau; ++t) {
#pragma unroll
      for (int f = 0; f < NF; ++f) {
        if constexpr (F32T) sys.tangent_apply(c.aux, Wf[f]);
        else sys.tangent_apply(c.aux, c.W[f]);
      }
      sys.forward(c.aux, c.x);
      if constexpr (PF) park_store<D, K, NF>(park, c.W);
      sys.aux(c.x, c.aux);
      if constexpr (PF) park_load<D, K, NF>(park, c.W);
    }
    n += a.tau;
    if constexpr (F32T) code = frame_status_f<Sys, K, NF>(c, Wf);
    else code = frame_status(c);
    if (code != STATUS_OK) {
      early = true;
      break;
    }
    if constexpr (F32T) {
#pragma unroll
      for (int f = 0; f < NF; ++f)
#pragma unroll
        for (int e = 0; e < D * K; ++e) c.W[f].w[e] = (double)Wf[f].w[e];
    }
#pragma unroll
    for (int f = 0; f < NF; ++f) c.logs_valid[f] = false;
    chain_step(ch, n, c);
    trace_hook(ch, c, n, a, i);
    if constexpr (F32T) {
#pragma unroll
      for (int f = 0; f < NF; ++f)
#pragma unroll
        for (int e = 0; e < D * K; ++e) Wf[f].w[e] = (float)c.W[f].w[e];
    }
    if (chain_failed(ch)) {
      code = STATUS_REDUCER;
      early = true;
      break;
    }
    if (a.reset_frame) {
#pragma unroll
      for (int f = 0; f < NF; ++f)
#pragma unroll
        for (int e = 0; e < D * K; ++e) {
          c.W[f].w[e] = (e % (D + 1) == 0) ? 1.0 : 0.0;
          Wf[f].w[e] = (e % (D + 1) == 0) ? 1.0f : 0.0f;
        }
    }
  }
  if constexpr (F32T) {
    if (!early) code = (n == a.N) ? frame_status_f<Sys, K, NF>(c, Wf) : STATUS_OK;
#pragma unroll
    for (int f = 0; f < NF; ++f)
#pragma unroll
      for (int e = 0; e < D * K; ++e) c.W[f].w[e] = (double)Wf[f].w[e];
  } else {
    if (!early) code = (n == a.N) ? frame_status(c) : STATUS_OK;
  }
  if constexpr (HasStretch<Chain>::value) chain_ws_flush(ch, c);
#pragma unroll
  for (int j = 0; j < Sys::S; ++j) a.x[(int64_t)j * a.m + i] = c.x[j];
#pragma unroll
  for (int e = 0; e < D * K; ++e) a.W0[(int64_t)e * a.m + i] = c.W[0].w[e];
  if (NF > 1) {
#pragma unroll
    for (int e = 0; e < D * K; ++e) a.W1[(int64_t)e * a.m + i] = c.W[NF - 1].w[e];
  }
  chain_store(ch, a.st, a.m, i, 0);
  a.code[i] = code;
  a.hstep[i] = n;
}

}  // namespace cl
