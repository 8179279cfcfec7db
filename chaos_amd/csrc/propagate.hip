// propagate.hip - the C ABI of libchaoslab_hip.so: device probe, contexts, the chunk entry points
// that replace `_orbit_kernel!` / `_run_chunk!` (src/Tangent/propagate.jl:32-50,92-119), seeded frames
// and the elementary-function test entry points.
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <algorithm>
#include <mutex>
#include <vector>
#include "registry.hpp"
#include "frames.hpp"
#include "chain_kernels.hpp"

using namespace cl;

struct cl_ctx {
  int device;
  hipStream_t stream;
  char err[256];
  cl_tuning tune;  // launch settings chosen by the caller's autotuner (zeros: the built-in rules)
  // grown-on-demand device buffers of the host entry point
  void* buf[14];
  size_t cap[14];
};

// The launch settings of the call in progress on this thread (the context's cl_tuning).
static thread_local cl_tuning g_tune = {0, 0, 0, 0};

static char g_create_err[256] = "";

// Scratch of cl_propagate_mixed_noalloc: the device staging of the work list, step counts, launch orders and work
// counters, a file-scope device array of a fixed size (each GPU the library runs on holds one copy until the
// process ends). Host code reaches it only through the address hipGetSymbolAddress returns for the device, taken
// once per device by cl_create with that device current (g_scratch_ptr); the array's name is never used as a
// device address on the host.
static constexpr size_t CL_SCRATCH_BYTES = (size_t)4 << 20;
static __device__ __align__(256) char scratch[CL_SCRATCH_BYTES];
static constexpr int CL_MAX_DEVICES = 64;
static void* g_scratch_ptr[CL_MAX_DEVICES];
static std::mutex g_scratch_mu;

// ------------------------------------------------------------------------------ instance table
static std::once_flag g_once;
static Instance g_inst[384];
static int g_ninst = 0;

static PipeInstance g_pipe[64];
static int g_npipe = 0;

static void load_table() {
  std::call_once(g_once, [] {
    register_instances_map(g_inst, &g_ninst, 384);
    register_instances_flow(g_inst, &g_ninst, 384);
    register_instances_f32(g_inst, &g_ninst, 384);
    register_pipes(g_pipe, &g_npipe, 64);
  });
}

// The pipeline kernel for a one-frame chain on a two-dimensional map that is a single QR spectrum, or a QR
// spectrum followed by RunMonitor.
static const PipeInstance* find_pipe(const cl_system* sys, const cl_chain* ch, int d, int k) {
  load_table();
  if (d != 2 || ch->n < 1 || ch->n > 2 || ch->r[0].kind != CL_RED_QR || ch->r[0].frame != 0) return nullptr;
  const int mon = ch->n == 2;
  if (mon && (ch->r[1].kind != CL_RED_RUNMONITOR || ch->r[1].frame != 0)) return nullptr;
  for (int i = 0; i < g_npipe; ++i)
    if (g_pipe[i].sys_kind == sys->kind && g_pipe[i].k == k && g_pipe[i].alg == ch->r[0].alg && g_pipe[i].mon == mon)
      return &g_pipe[i];
  return nullptr;
}

static const Instance* find_instance(const cl_system* sys, const cl_chain* ch, int d, int k, int elt = CL_ELT_F64) {
  load_table();
  for (int i = 0; i < g_ninst; ++i) {
    const Instance& in = g_inst[i];
    if (in.elt != elt || in.sys_kind != sys->kind || in.d != d || in.k != k || in.nred != ch->n) continue;
    bool same = true;
    for (int r = 0; r < ch->n && same; ++r) {
      const cl_reducer& q = ch->r[r];
      int alg = (q.kind == CL_RED_QR || q.kind == CL_RED_PHASE_ADVANCE) ? q.alg : 0;  // the QR algorithm; the windows
      same = in.red[r][0] == q.kind && in.red[r][1] == alg && in.red[r][2] == q.frame;
    }
    if (same) return &in;
  }
  return nullptr;
}

static int state_dim(int kind, int elt = CL_ELT_F64) {
  if (elt == CL_ELT_DD && kind == CL_SYS_HENON_HEILES_GAUSS) return 8;  // value and correction
  switch (kind) {
    case CL_SYS_STANDARD_MAP: return 2;
    case CL_SYS_WRAPPED_STANDARD_MAP: return 4;
    case CL_SYS_HENON_HEILES_GAUSS: return 4;
    case CL_SYS_STANDARD_MAP_K: return 3;
    case CL_SYS_PAIR_STANDARD_MAP: return 4;
    case CL_SYS_PAIR_WRAPPED_STANDARD_MAP: return 8;
    default: return -1;
  }
}
static int sys_dim(int kind) {
  switch (kind) {
    case CL_SYS_STANDARD_MAP:
    case CL_SYS_WRAPPED_STANDARD_MAP:
    case CL_SYS_STANDARD_MAP_K: return 2;
    case CL_SYS_HENON_HEILES_GAUSS:
    case CL_SYS_PAIR_STANDARD_MAP:
    case CL_SYS_PAIR_WRAPPED_STANDARD_MAP: return 4;
    default: return -1;
  }
}

static int map_err(hipError_t e) {
  if (e == hipSuccess) return CL_OK;
  if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) return CL_ERR_OOM;
  return CL_ERR_DEVICE_LOST;
}

static int set_err(cl_ctx* ctx, int code, const char* msg) {
  if (ctx) snprintf(ctx->err, sizeof(ctx->err), "%s", msg);
  return code;
}

// Validation shared by both launch entry points (src/Tangent/propagate.jl:279-289, orbit.jl:70-74).
// WeightedBirkhoff (the last reducer of its chain): the horizon sums held in the state array, H (2 + 5K) words.
static int wb_index(const cl_chain* ch) {
  for (int r = 0; r < ch->n && r < CL_MAX_REDUCERS; ++r)
    if (ch->r[r].kind == CL_RED_BIRKHOFF) return r;
  return -1;
}
static int64_t wb_mem_words(const cl_chain* ch) {
  const int r = wb_index(ch);
  if (r < 0 || !ch->r[r].birkhoff) return 0;
  const cl_birkhoff& b = *ch->r[r].birkhoff;
  return (int64_t)b.H * (2 + 5 * b.K);
}
// The descriptor's own checks (WeightedBirkhoff's constructor, wbirkhoff_reducer.jl:28-34): 1 to 16 horizons of at
// least 16 samples, 1 to 8 observables (the reference allows 31: more than 8 has no kernel), order 1 to 4, the
// observables' coordinates inside the state; the reducer must be the last of its chain.
static const char* wb_check(const cl_chain* ch, int s) {
  const int r = wb_index(ch);
  if (r < 0) return nullptr;
  if (r != ch->n - 1) return "WeightedBirkhoff must be the last reducer of its chain";
  const cl_birkhoff* b = ch->r[r].birkhoff;
  if (!b) return "WeightedBirkhoff needs its parameters";
  if (b->H < 1 || b->H > CL_WB_MAX_WINDOWS) return "WeightedBirkhoff: 1 to 16 windows";
  if (b->K < 1 || b->K > CL_WB_MAX_OBSERVABLES) return "WeightedBirkhoff: 1 to 8 observables";
  if (b->order < 1 || b->order > 4) return "WeightedBirkhoff: birkhoff_order must be 1 to 4";
  for (int h = 0; h < b->H; ++h)
    if (b->window[h] < 16 || b->window[h] > ((int64_t)1 << 40)) return "WeightedBirkhoff: each window at least 16";
  for (int k = 0; k < b->K; ++k) {
    const bool trig = b->kind[k] == CL_OBS_COS || b->kind[k] == CL_OBS_SIN;
    if (!(trig || b->kind[k] == CL_OBS_LIFTED)) return "WeightedBirkhoff: unknown observable kind";
    if (b->a[k] < 0 || b->a[k] >= s || (!trig && (b->b[k] < 0 || b->b[k] >= s)))
      return "WeightedBirkhoff: observable coordinate outside the state";
  }
  return nullptr;
}

// WeightedPhaseAdvance's descriptor (naff_reducer.jl:148-152, check_run :236-242): 1 or 2 windows of at least 16
// samples (the key's alg is the window count), order 1 to 4, a planar or lifted-angle observable inside the state.
static const char* pa_check(const cl_chain* ch, int s) {
  for (int r = 0; r < ch->n && r < CL_MAX_REDUCERS; ++r) {
    if (ch->r[r].kind != CL_RED_PHASE_ADVANCE) continue;
    const cl_birkhoff* b = ch->r[r].birkhoff;
    if (!b) return "WeightedPhaseAdvance needs its parameters";
    if (b->H < 1 || b->H > 2 || ch->r[r].alg != b->H) return "WeightedPhaseAdvance: 1 or 2 windows (alg = the count)";
    if (b->order < 1 || b->order > 4) return "WeightedPhaseAdvance: birkhoff_order must be 1 to 4";
    for (int h = 0; h < b->H; ++h)
      if (b->window[h] < 16 || b->window[h] > ((int64_t)1 << 40)) return "WeightedPhaseAdvance: each window at least 16";
    const bool planar = b->kind[0] == CL_OBS_PLANAR, lifted = b->kind[0] == CL_OBS_LIFTED_ANGLE;
    if (!(planar || lifted)) return "WeightedPhaseAdvance: the observable is planar or lifted-angle";
    if (b->a[0] < 0 || b->a[0] >= s || b->b[0] >= s || (planar && b->b[0] < 0) || (lifted && b->b[0] < -1))
      return "WeightedPhaseAdvance: observable index outside the state";
  }
  return nullptr;
}

// The per-launch table: the header (windows, offsets) and horizon h's w(jj), jj < N_h, at offset
// sum_{h' < h} N_h' after it.
__global__ void k_wb_weights(int H, cl_birkhoff b, double* wtab) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t* hdr = reinterpret_cast<int64_t*>(wtab);
  if (t < CL_WB_MAX_WINDOWS) {
    int64_t o = 0;
    for (int h = 0; h < t && h < H; ++h) o += b.window[h];
    hdr[t] = t < H ? b.window[t] : 0;
    hdr[CL_WB_MAX_WINDOWS + t] = o;
  }
  int64_t off = 0;
  for (int h = 0; h < H; ++h) {
    const int64_t N = b.window[h];
    if (t >= off && t < off + N) {
      wtab[WB_TABLE_HEADER + t] = wb_weight(t - off, N, b.order);
      return;
    }
    off += N;
  }
}

static int validate(cl_ctx* ctx, const cl_system* sys, const cl_chain* ch, int elt, int64_t m, int s, int d,
                    int k, int64_t n0, int64_t steps, int64_t N, int tau, const Instance** out) {
  if (!sys || !ch) return set_err(ctx, CL_ERR_ARGUMENT, "null system or chain descriptor");
  if (ch->n < 0 || ch->n > CL_MAX_REDUCERS) return set_err(ctx, CL_ERR_ARGUMENT, "chain length out of range");
  if (!(tau >= 1 && N >= 1 && N % tau == 0))
    return set_err(ctx, CL_ERR_ARGUMENT, "N must be a positive multiple of tau");
  if (!(steps >= 0 && steps % tau == 0 && n0 >= 0 && n0 % tau == 0 && n0 + steps <= N))
    return set_err(ctx, CL_ERR_ARGUMENT, "launch steps must be a multiple of tau within N");
  if (m < 0) return set_err(ctx, CL_ERR_ARGUMENT, "negative orbit count");
  int sd = sys_dim(sys->kind);
  if (sd < 0) return set_err(ctx, CL_ERR_UNSUPPORTED, "system kind has no kernel");
  if (d != sd) return set_err(ctx, CL_ERR_ARGUMENT, "d is not the system dimension");
  if (sys->kind == CL_SYS_HENON_HEILES_GAUSS && sys->stages != 3)
    return set_err(ctx, CL_ERR_UNSUPPORTED, "Gauss kernels are built for s = 3 stages");
  if (elt != CL_ELT_F64 && elt != CL_ELT_F64_F32TAN && elt != CL_ELT_DD)
    return set_err(ctx, CL_ERR_UNSUPPORTED, "element type has no kernel for this system");
  if (s != state_dim(sys->kind, elt)) return set_err(ctx, CL_ERR_ARGUMENT, "state length does not match the system");
  if (!(k >= 1 && k <= d)) return set_err(ctx, CL_ERR_ARGUMENT, "k must be in 1:d");
  int nws = 0;
  for (int r = 0; r < ch->n; ++r) {
    const cl_reducer& q = ch->r[r];
    if (q.kind != CL_RED_STRETCHING) continue;
    ++nws;
    if (k != 2) return set_err(ctx, CL_ERR_ARGUMENT, "StretchingNumbers needs k = 2");
    if (!(q.bins >= 1 && q.lo < q.hi && q.hi - q.lo < __builtin_inf() && q.lo > -__builtin_inf()))
      return set_err(ctx, CL_ERR_ARGUMENT, "StretchingNumbers needs bins >= 1 and a finite range lo < hi");
    if ((n0 + steps) / tau > 1000000000) return set_err(ctx, CL_ERR_ARGUMENT, "more than 10^9 handoffs");
  }
  if (nws > 1) return set_err(ctx, CL_ERR_ARGUMENT, "at most one reducer of a chain may request a workspace");
  if (const char* why = wb_check(ch, s)) return set_err(ctx, CL_ERR_ARGUMENT, why);
  if (const char* why = pa_check(ch, s)) return set_err(ctx, CL_ERR_ARGUMENT, why);
  const Instance* in = find_instance(sys, ch, d, k, elt);
  if (!in) return set_err(ctx, CL_ERR_UNSUPPORTED, "no kernel instance for this system/chain/k/element type");
  *out = in;
  return CL_OK;
}

// UInt32 workspace entries per orbit (_chain_workspace, src/Tangent/orbit.jl:8-18): 2 (bins + 2) for
// StretchingNumbers (stretching.jl:93-96), 0 when no reducer requests one, -1 when two do.
static int64_t ws_entries(const cl_chain* ch) {
  int64_t e = 0;
  int c = 0;
  for (int r = 0; r < ch->n && r < CL_MAX_REDUCERS; ++r)
    if (ch->r[r].kind == CL_RED_STRETCHING) {
      e = 2 * ((int64_t)ch->r[r].bins + 2);
      ++c;
    }
  return c > 1 ? -1 : e;
}

extern "C" int64_t cl_workspace_entries(const cl_chain* ch, int d, int k) {
  (void)d;
  (void)k;
  if (!ch || ch->n < 0 || ch->n > CL_MAX_REDUCERS) return -CL_ERR_ARGUMENT;
  const int64_t e = ws_entries(ch);
  return e < 0 ? -CL_ERR_ARGUMENT : e;
}

// The coupled chain (kind CL_SYS_STANDARD_MAP_CHAIN): state 2N, frames 2N x k, no reducer or
// QRSpectrum(HouseholderQR()); kernels for N in {8, 16, 32, 64, 128} and k = 16 (chain_kernels.hip).
// Returns the state words per orbit through *words, -1 when the chain has no kernel.
static int validate_chain(cl_ctx* ctx, const cl_system* sys, const cl_chain* ch, int elt, int64_t m, int s, int d,
                          int k, int tau, int* words, bool* qr) {
  const int N = sys->N;
  if (N < 2 || N > 128 || (sys->ring && N < 3)) return set_err(ctx, CL_ERR_ARGUMENT, "chain sites out of range");
  if (d != 2 * N || s != 2 * N) return set_err(ctx, CL_ERR_ARGUMENT, "state length and d must be 2N for the chain");
  if (!(k >= 1 && k <= d)) return set_err(ctx, CL_ERR_ARGUMENT, "k must be in 1:d");
  if (tau < 1) return set_err(ctx, CL_ERR_ARGUMENT, "tau must be positive");
  if (m < 0) return set_err(ctx, CL_ERR_ARGUMENT, "negative orbit count");
  if (elt != CL_ELT_F64 && elt != CL_ELT_F64_F32TAN)
    return set_err(ctx, CL_ERR_UNSUPPORTED, "element type has no kernel for this system");
  if (!chain_supported(N, k)) return set_err(ctx, CL_ERR_UNSUPPORTED, "no chain kernel for this N and k");
  if (ch->n == 0) {
    *qr = false;
    *words = 0;
  } else if (ch->n == 1 && ch->r[0].kind == CL_RED_QR && ch->r[0].alg == CL_QR_HOUSEHOLDER && ch->r[0].frame == 0) {
    *qr = true;
    *words = 2 * k + 9;
  } else {
    return set_err(ctx, CL_ERR_UNSUPPORTED, "chain kernels run no reducer or QRSpectrum(HouseholderQR())");
  }
  return CL_OK;
}

static ChainArgs chain_args(const cl_system* sys, const cl_chain* ch, int64_t m, double* dx, double* dW,
                            uint64_t* dst, int8_t* dcode, int64_t* dhstep, int64_t n0, int64_t steps, int64_t N,
                            int tau) {
  ChainArgs a;
  a.x = dx;
  a.W = dW;
  a.st = dst;
  a.code = dcode;
  a.hstep = dhstep;
  a.Ns = nullptr;
  a.order = nullptr;
  a.m = m;
  a.n0 = n0;
  a.steps = steps;
  a.N = N;
  a.tau = tau;
  a.K = sys->K;
  a.eps = sys->eps;
  a.tol = ch->n > 0 ? ch->r[0].tol : 8.0;
  return a;
}

// ------------------------------------------------------------------------------ helper kernels
// AoS[i][w] <-> SoA[w][i] for 8-byte words.
__global__ void k_aos_to_soa(const uint64_t* __restrict__ a, uint64_t* __restrict__ s, int64_t m, int w) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= m * w) return;
  int64_t j = t / m, i = t - j * m;
  s[t] = a[i * w + j];
}
__global__ void k_soa_to_aos(const uint64_t* __restrict__ s, uint64_t* __restrict__ a, int64_t m, int w) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= m * w) return;
  int64_t j = t / m, i = t - j * m;
  a[i * w + j] = s[t];
}

// the same for 4-byte words (the reducer workspace)
__global__ void k_aos_to_soa32(const uint32_t* __restrict__ a, uint32_t* __restrict__ s, int64_t m, int64_t w) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= m * w) return;
  int64_t j = t / m, i = t - j * m;
  s[t] = a[i * w + j];
}
__global__ void k_soa_to_aos32(const uint32_t* __restrict__ s, uint32_t* __restrict__ a, int64_t m, int64_t w) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= m * w) return;
  int64_t j = t / m, i = t - j * m;
  a[i * w + j] = s[t];
}

template <int D, int K>
__global__ void k_frames(uint64_t seed, int64_t orbit0, int64_t m, double* __restrict__ W) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  Frame<D, K> f;
  seeded_frame<D, K>(seed, orbit0 + i, f);
#pragma unroll
  for (int e = 0; e < D * K; ++e) W[(int64_t)e * m + i] = f.w[e];
}

__global__ void k_math(int64_t n, const double* __restrict__ x, double* sn, double* cs, double* lg) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s, c;
  cl_sincos(x[i], &s, &c);
  sn[i] = s;
  cs[i] = c;
  lg[i] = cl_log(x[i]);
}

__global__ void k_sqrt12(int64_t n, const double* __restrict__ x, double* __restrict__ y) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  y[i] = cl_sqrt12(x[i]);
}

static unsigned grid_for(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

// ------------------------------------------------------------------------------ probe / context
extern "C" int cl_probe(int elt, cl_probe_t* out) {
  if (!out) return CL_ERR_ARGUMENT;
  memset(out, 0, sizeof(*out));
  out->loadable = 1;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n < 1) return CL_ERR_NO_DEVICE;
  hipDeviceProp_t p;
  if (hipGetDeviceProperties(&p, 0) != hipSuccess) return CL_ERR_NO_DEVICE;
  out->functional = strncmp(p.gcnArchName, "gfx950", 6) == 0;
  snprintf(out->device, sizeof(out->device), "%s (%s)", p.name, p.gcnArchName);
  out->compiles = out->functional && (elt == CL_ELT_F64 || elt == CL_ELT_F64_F32TAN || elt == CL_ELT_DD);
  out->gpu_faster = out->compiles;
  return out->functional ? CL_OK : CL_ERR_NO_DEVICE;
}

extern "C" int64_t cl_free_memory(void) {
  size_t fr = 0, tot = 0;
  if (hipMemGetInfo(&fr, &tot) != hipSuccess) return -1;
  return (int64_t)fr;
}

extern "C" cl_ctx* cl_create(int device) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) {
    snprintf(g_create_err, sizeof(g_create_err), "no HIP device %d", device);
    return nullptr;
  }
  if (hipSetDevice(device) != hipSuccess) {
    snprintf(g_create_err, sizeof(g_create_err), "hipSetDevice(%d) failed", device);
    return nullptr;
  }
  if (device < CL_MAX_DEVICES) {  // the scratch array's address on this device, once per device
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
      void* p = nullptr;
      if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) == hipSuccess) g_scratch_ptr[device] = p;
    }
  }
  cl_ctx* c = (cl_ctx*)calloc(1, sizeof(cl_ctx));
  c->device = device;
  if (hipStreamCreate(&c->stream) != hipSuccess) {
    free(c);
    snprintf(g_create_err, sizeof(g_create_err), "hipStreamCreate failed");
    return nullptr;
  }
  load_table();
  return c;
}

extern "C" void cl_destroy(cl_ctx* c) {
  if (!c) return;
  hipSetDevice(c->device);
  for (int i = 0; i < 14; ++i)
    if (c->buf[i]) hipFree(c->buf[i]);
  hipStreamDestroy(c->stream);
  free(c);
}

extern "C" const char* cl_last_error(cl_ctx* c) { return c ? c->err : g_create_err; }

extern "C" int cl_set_tuning(cl_ctx* c, const cl_tuning* t) {
  if (!c || !t) return CL_ERR_ARGUMENT;
  if (!(t->block == 0 || t->block == 64 || t->block == 128 || t->block == 256) || t->split < 0 || t->split > 2 ||
      t->pipe < 0 || t->pipe > 2 || t->chain_qr < 0 || t->chain_qr > 2)
    return set_err(c, CL_ERR_ARGUMENT,
                   "tuning: block 0/64/128/256, split, pipe and chain_qr 0 (automatic), 1 (on), 2 (off)");
  c->tune = *t;
  return CL_OK;
}

extern "C" int cl_get_tuning(cl_ctx* c, cl_tuning* t) {
  if (!c || !t) return CL_ERR_ARGUMENT;
  *t = c->tune;
  return CL_OK;
}

// The context's stream, for the other translation units of the library (not part of the C ABI).
extern "C" __attribute__((visibility("hidden"))) hipStream_t cl_ctx_stream(cl_ctx* c) { return c->stream; }

extern "C" int cl_birkhoff_weights(int64_t N, int order, double* w) {
  if (N < 1 || order < 1 || order > 4 || !w) return CL_ERR_ARGUMENT;
  for (int64_t jj = 0; jj < N; ++jj) w[jj] = wb_weight(jj, N, order);
  return CL_OK;
}

extern "C" int64_t cl_state_bytes(const cl_system* sys, const cl_chain* ch, int elt, int d, int k) {
  if (!sys || !ch || ch->n < 0 || ch->n > CL_MAX_REDUCERS) return -CL_ERR_ARGUMENT;
  if (elt != CL_ELT_F64 && elt != CL_ELT_F64_F32TAN && elt != CL_ELT_DD) return -CL_ERR_UNSUPPORTED;
  if (sys->kind == CL_SYS_STANDARD_MAP_CHAIN) {
    int words = 0;
    bool qr = false;
    int rc = validate_chain(nullptr, sys, ch, elt, 0, d, d, k, 1, &words, &qr);
    return rc ? -(int64_t)rc : (int64_t)words * 8;
  }
  const Instance* in = find_instance(sys, ch, d, k, elt);
  if (!in) return -CL_ERR_UNSUPPORTED;
  if (wb_check(ch, state_dim(sys->kind))) return -CL_ERR_ARGUMENT;
  return ((int64_t)in->words + wb_mem_words(ch)) * 8;
}

static int grow(cl_ctx* c, int slot, size_t bytes) {
  if (bytes == 0) bytes = 8;
  if (c->cap[slot] >= bytes) return CL_OK;
  if (c->buf[slot]) hipFree(c->buf[slot]);
  c->buf[slot] = nullptr;
  c->cap[slot] = 0;
  hipError_t e = hipMalloc(&c->buf[slot], bytes);
  if (e != hipSuccess) return map_err(e);
  c->cap[slot] = bytes;
  return CL_OK;
}

static int launch(const Instance* in, const cl_system* sys, const cl_chain* ch, int64_t m, double* dx,
                  double* dW, double* dW2, uint64_t* dst, int8_t* dcode, int64_t* dhstep, int64_t n0,
                  int64_t steps, int64_t N, int tau, hipStream_t st, float* trace = nullptr,
                  int64_t trace_n0 = 0, int trace_len = 0, const uint8_t* mask = nullptr,
                  uint32_t* ws = nullptr) {
  if (m == 0) return CL_OK;
  KArgs a;
  memset(&a, 0, sizeof(a));
  a.block = g_tune.block;
  a.ws = ws;
  a.mask = mask;
  a.reset_frame = (ch->flags & CL_CHAIN_RESET_FRAME) ? 1 : 0;
  const int wbr = wb_index(ch);
  double* wtab = nullptr;
  if (wbr >= 0) {  // the weight tables, formed on the device for this launch
    const cl_birkhoff& b = *ch->r[wbr].birkhoff;
    int64_t tot = 0;
    for (int h = 0; h < b.H; ++h) tot += b.window[h];
    hipError_t e = hipMallocAsync((void**)&wtab, (size_t)(tot + WB_TABLE_HEADER) * 8, st);
    if (e != hipSuccess) return map_err(e);
    const int64_t nt = tot > CL_WB_MAX_WINDOWS ? tot : CL_WB_MAX_WINDOWS;
    hipLaunchKernelGGL(k_wb_weights, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, st, b.H, b, wtab);
    a.wtab = wtab;
  }
  a.trace = trace;
  a.trace_n0 = trace_n0;
  a.trace_len = trace_len;
  a.x = dx;
  a.W0 = dW;
  a.W1 = dW2;
  a.st = dst;
  a.code = dcode;
  a.hstep = dhstep;
  a.m = m;
  a.n0 = n0;
  a.steps = steps;
  a.N = N;
  a.tau = tau;
  int rc = in->launch(sys, ch, a, st);
  if (wtab) hipFreeAsync(wtab, st);
  return rc;
}

// ------------------------------------------------------------------------------ split execution
// A chain with reducers on both frames can run as two single-frame passes (frame-1 reducers relabelled
// to frame 0): each pass holds fewer registers, so more waves run per SIMD, and the frames never
// interact (the Jacobian applied to each is the same). The passes reproduce the fused kernel exactly
// for every orbit that completes in both; an orbit that fails in either pass (a non-finite state or
// frame, or a failed reducer) is restored from the launch's inputs and run again by the fused kernel,
// whose failure rule (status of x and both frames first, then every reducer in chain order) is the
// reference's.
static bool split_chains(const cl_chain* ch, cl_chain* ca, cl_chain* cb) {
  memset(ca, 0, sizeof(*ca));
  memset(cb, 0, sizeof(*cb));
  bool seen1 = false;
  for (int r = 0; r < ch->n; ++r) {
    cl_reducer q = ch->r[r];
    if (q.frame == 0) {
      if (seen1) return false;  // frame-0 reducers must come first (the chain order is kept)
      ca->r[ca->n++] = q;
    } else {
      seen1 = true;
      q.frame = 0;
      cb->r[cb->n++] = q;
    }
  }
  return ca->n > 0 && cb->n > 0;
}

// Measured on MI355X (bench workloads c2 and c4): for the maps the step is a few dozen instructions and
// the split runs 1.45x faster than the fused pass (C2: 637 ms against 927 ms per launch); for the Gauss
// flow the step (stage iteration and solve) dominates and running it twice makes the split 1.67x slower
// (C4: 11.46 s against 6.85 s). CL_SPLIT=0 forces the fused pass; CL_SPLIT=1 forces the split.
static bool split_enabled(int kind) {
  const char* e = getenv("CL_SPLIT");
  if (e && e[0] == '0') return false;
  if (e && e[0] == '1') return true;
  if (g_tune.split) return g_tune.split == 1;
  return kind == CL_SYS_STANDARD_MAP || kind == CL_SYS_WRAPPED_STANDARD_MAP || kind == CL_SYS_STANDARD_MAP_K;
}

// Orbits that failed in either pass (and were live at n0) are restored from the inputs and marked for
// the fused rerun.
__global__ void k_reconcile(int64_t m, int64_t n0, int S, int DK, int words, const int8_t* __restrict__ code_in,
                            const int8_t* __restrict__ codeB, const int64_t* __restrict__ h_in,
                            const double* __restrict__ x_in, const double* __restrict__ W0_in,
                            const double* __restrict__ W1_in, const uint64_t* __restrict__ st_in, double* x,
                            double* W0, double* W1, uint64_t* st, int8_t* code, int64_t* hstep, uint8_t* mask) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const bool live = n0 == 0 || code_in[i] == STATUS_OK;
  const bool rerun = live && (code[i] != STATUS_OK || codeB[i] != STATUS_OK);
  mask[i] = rerun ? 1 : 0;
  if (!rerun) return;
  for (int c = 0; c < S; ++c) x[(int64_t)c * m + i] = x_in[(int64_t)c * m + i];
  for (int e = 0; e < DK; ++e) {
    W0[(int64_t)e * m + i] = W0_in[(int64_t)e * m + i];
    W1[(int64_t)e * m + i] = W1_in[(int64_t)e * m + i];
  }
  for (int w = 0; w < words; ++w) st[(int64_t)w * m + i] = st_in[(int64_t)w * m + i];
  code[i] = code_in[i];
  hstep[i] = h_in[i];
}

// Runs a two-frame launch as two passes plus the fused rerun of failed orbits. Returns CL_ERR_UNSUPPORTED
// (without side effects) when the chain does not split or a sub-chain has no kernel instance.
static int launch_split(const Instance* in, const cl_system* sys, const cl_chain* ch, int64_t m, int s, int d,
                        int k, double* dx, double* dW, double* dW2, uint64_t* dst, int8_t* dcode, int64_t* dhstep,
                        int64_t n0, int64_t steps, int64_t N, int tau, hipStream_t st) {
  cl_chain ca, cb;
  if (!split_chains(ch, &ca, &cb)) return CL_ERR_UNSUPPORTED;
  const Instance* ia = find_instance(sys, &ca, d, k, in->elt);
  const Instance* ib = find_instance(sys, &cb, d, k, in->elt);
  if (!ia || !ib || ia->nf != 1 || ib->nf != 1 || ia->words + ib->words != in->words) return CL_ERR_UNSUPPORTED;
  if (m == 0) return CL_OK;
  const int DK = d * k, words = in->words;
  const size_t bx = (size_t)m * s * 8, bw = (size_t)m * DK * 8, bs = (size_t)m * words * 8;
  // inputs (for the rerun), and pass B's private copies of x, code and step
  size_t off[12], tot = 0;
  const size_t sz[12] = {bx, bw, bw, bs, (size_t)m, (size_t)m * 8, bx, (size_t)m, (size_t)m * 8, (size_t)m, 0, 0};
  for (int j = 0; j < 10; ++j) {
    off[j] = tot;
    tot += (sz[j] + 255) & ~(size_t)255;
  }
  char* base = nullptr;
  hipError_t e = hipMallocAsync((void**)&base, tot, st);
  if (e != hipSuccess) return map_err(e);
  double *x_in = (double*)(base + off[0]), *W0_in = (double*)(base + off[1]), *W1_in = (double*)(base + off[2]);
  uint64_t* st_in = (uint64_t*)(base + off[3]);
  int8_t* code_in = (int8_t*)(base + off[4]);
  int64_t* h_in = (int64_t*)(base + off[5]);
  double* xB = (double*)(base + off[6]);
  int8_t* codeB = (int8_t*)(base + off[7]);
  int64_t* hB = (int64_t*)(base + off[8]);
  uint8_t* mask = (uint8_t*)(base + off[9]);
  int rc = CL_OK;
  if ((e = hipMemcpyAsync(x_in, dx, bx, hipMemcpyDeviceToDevice, st)) != hipSuccess ||
      (e = hipMemcpyAsync(W0_in, dW, bw, hipMemcpyDeviceToDevice, st)) != hipSuccess ||
      (e = hipMemcpyAsync(W1_in, dW2, bw, hipMemcpyDeviceToDevice, st)) != hipSuccess ||
      (n0 > 0 && (e = hipMemcpyAsync(st_in, dst, bs, hipMemcpyDeviceToDevice, st)) != hipSuccess) ||
      (e = hipMemcpyAsync(code_in, dcode, (size_t)m, hipMemcpyDeviceToDevice, st)) != hipSuccess ||
      (e = hipMemcpyAsync(h_in, dhstep, (size_t)m * 8, hipMemcpyDeviceToDevice, st)) != hipSuccess ||
      (e = hipMemcpyAsync(xB, dx, bx, hipMemcpyDeviceToDevice, st)) != hipSuccess ||
      (e = hipMemcpyAsync(codeB, dcode, (size_t)m, hipMemcpyDeviceToDevice, st)) != hipSuccess ||
      (e = hipMemcpyAsync(hB, dhstep, (size_t)m * 8, hipMemcpyDeviceToDevice, st)) != hipSuccess) {
    rc = map_err(e);
  }
  if (!rc) rc = launch(ia, sys, &ca, m, dx, dW, nullptr, dst, dcode, dhstep, n0, steps, N, tau, st);
  if (!rc)
    rc = launch(ib, sys, &cb, m, xB, dW2, nullptr, dst + (size_t)ia->words * m, codeB, hB, n0, steps, N, tau, st);
  if (!rc) {
    hipLaunchKernelGGL(k_reconcile, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, st, m, n0, s, DK,
                       n0 > 0 ? words : 0, code_in, codeB, h_in, x_in, W0_in, W1_in, st_in, dx, dW, dW2, dst, dcode,
                       dhstep, mask);
    e = hipGetLastError();
    rc = map_err(e);
  }
  if (!rc) rc = launch(in, sys, ch, m, dx, dW, dW2, dst, dcode, dhstep, n0, steps, N, tau, st, nullptr, 0, 0, mask);
  hipFreeAsync(base, st);
  return rc;
}

// ------------------------------------------------------------------------------ pipelined small ensembles
// Below a few tens of thousands of orbits the fused kernel leaves most SIMDs idle and runs at the issue rate of
// one wave; the pipeline kernel spreads each orbit-step over four waves (pipe_kernel.hpp). Orbits it marks
// (any failure) are rerun by the fused kernel from their unchanged inputs. CL_PIPE=0 disables it, CL_PIPE=1
// uses it for every eligible launch; by default it serves launches of at most PIPE_MAX_ORBITS orbits.
static const int64_t PIPE_MAX_ORBITS = 16384;

static bool pipe_enabled(int64_t m) {
  const char* e = getenv("CL_PIPE");
  if (e && e[0] == '0') return false;
  if (e && e[0] == '1') return true;
  if (g_tune.pipe) return g_tune.pipe == 1;
  return m <= PIPE_MAX_ORBITS;
}

static int launch_pipe(const Instance* in, const cl_system* sys, const cl_chain* ch, int64_t m, int d, int k,
                       double* dx, double* dW, uint64_t* dst, int8_t* dcode, int64_t* dhstep, int64_t n0,
                       int64_t steps, int64_t N, int tau, hipStream_t st) {
  const PipeInstance* pi = find_pipe(sys, ch, d, k);
  if (!pi || in->elt != CL_ELT_F64 || in->nf != 1 || tau > PIPE_TAU_MAX || !pipe_enabled(m)) return CL_ERR_UNSUPPORTED;
  if (m == 0) return CL_OK;
  uint8_t* rerun = nullptr;
  hipError_t e = hipMallocAsync((void**)&rerun, (size_t)m, st);
  if (e != hipSuccess) return map_err(e);
  KArgs a;
  memset(&a, 0, sizeof(a));
  a.x = dx;
  a.W0 = dW;
  a.st = dst;
  a.code = dcode;
  a.hstep = dhstep;
  a.m = m;
  a.n0 = n0;
  a.steps = steps;
  a.N = N;
  a.tau = tau;
  int rc = pi->launch(sys, ch, a, rerun, st);
  if (!rc) rc = launch(in, sys, ch, m, dx, dW, nullptr, dst, dcode, dhstep, n0, steps, N, tau, st, nullptr, 0, 0, rerun);
  hipFreeAsync(rerun, st);
  return rc;
}

// One launch by the fastest applicable route: two single-frame passes for a two-frame map chain, the pipeline
// for a small single-QR ensemble, else the fused kernel. Every route gives the fused kernel's bits.
static int launch_best(const Instance* in, const cl_system* sys, const cl_chain* ch, int64_t m, int s, int d, int k,
                       double* dx, double* dW, double* dW2, uint64_t* dst, int8_t* dcode, int64_t* dhstep,
                       int64_t n0, int64_t steps, int64_t N, int tau, hipStream_t st, uint32_t* ws = nullptr) {
  if (ws)  // a workspace chain runs the fused kernel
    return launch(in, sys, ch, m, dx, dW, dW2, dst, dcode, dhstep, n0, steps, N, tau, st, nullptr, 0, 0, nullptr, ws);
  int rc;
  if (in->nf > 1 && split_enabled(sys->kind)) {
    rc = launch_split(in, sys, ch, m, s, d, k, dx, dW, dW2, dst, dcode, dhstep, n0, steps, N, tau, st);
    if (rc != CL_ERR_UNSUPPORTED) return rc;
  }
  rc = launch_pipe(in, sys, ch, m, d, k, dx, dW, dst, dcode, dhstep, n0, steps, N, tau, st);
  if (rc != CL_ERR_UNSUPPORTED) return rc;
  return launch(in, sys, ch, m, dx, dW, dW2, dst, dcode, dhstep, n0, steps, N, tau, st);
}

#include "record.hpp"

// Row g of the recorder's array after the launch that ends at grid step g (src/Experiments/run_device.jl:9-37).
extern "C" int cl_record_rows(cl_ctx* ctx, const cl_system* sys, const cl_chain* ch, int64_t m, int s, int k, int g,
                              int G, const double* dx, const uint64_t* dst, const int8_t* dcode, int64_t* dfail,
                              double* dR, void* stream) {
  if (!ctx || !sys || !ch) return CL_ERR_ARGUMENT;
  const bool est_ok = ch->n == 2 && ch->r[1].kind == CL_RED_RUNMONITOR && ch->r[0].frame == 0 &&
                      ch->r[1].frame == 0 &&
                      ((ch->r[0].kind == CL_RED_QR && (k == 1 || k == 2)) || (ch->r[0].kind == CL_RED_BENETTIN && k == 1));
  if (!est_ok) return set_err(ctx, CL_ERR_UNSUPPORTED, "the recorder's chain is (QRSpectrum or Benettin, RunMonitor)");
  const bool sys_ok = (sys->kind == CL_SYS_STANDARD_MAP && s == 2) || (sys->kind == CL_SYS_WRAPPED_STANDARD_MAP && s == 4);
  if (!sys_ok) return set_err(ctx, CL_ERR_UNSUPPORTED, "the recorder runs the standard map, lifted or wrapped");
  if (m < 0 || G < 1 || g < 0 || g >= G) return set_err(ctx, CL_ERR_ARGUMENT, "grid row out of range");
  if (m > 0 && (!dx || !dst || !dcode || !dfail || !dR)) return set_err(ctx, CL_ERR_ARGUMENT, "null buffer");
  hipSetDevice(ctx->device);
  int rc = record_rows_launch(m, s, k, ch->r[0].kind == CL_RED_BENETTIN, g, G, dx, dst, dcode, dfail, dR,
                              stream ? (hipStream_t)stream : ctx->stream);
  return rc ? set_err(ctx, rc, "record kernel launch failed") : CL_OK;
}

// Row g of the indicator recorder's array (src/Experiments/run_indicator_device.jl:13-29,31-60,62-91): a series
// kind's chain is its members (FLI, MEGNO, SALI on frame 1) then RunMonitor; a window kind's is RunMonitor then
// one reader, whose row holds the monitor and the point only.
extern "C" int cl_ind_record_rows(cl_ctx* ctx, const cl_system* sys, const cl_chain* ch, int64_t m, int s, int g,
                                  int G, const double* dx, const uint64_t* dst, const int8_t* dcode, int64_t* dfail,
                                  double* dR, void* stream) {
  if (!ctx || !sys || !ch) return CL_ERR_ARGUMENT;
  const bool pair = sys->kind == CL_SYS_PAIR_STANDARD_MAP || sys->kind == CL_SYS_PAIR_WRAPPED_STANDARD_MAP;
  const bool sys_ok = ((sys->kind == CL_SYS_STANDARD_MAP || sys->kind == CL_SYS_PAIR_STANDARD_MAP) ||
                       (sys->kind == CL_SYS_WRAPPED_STANDARD_MAP || sys->kind == CL_SYS_PAIR_WRAPPED_STANDARD_MAP)) &&
                      s == state_dim(sys->kind);
  if (!sys_ok)
    return set_err(ctx, CL_ERR_UNSUPPORTED, "the indicator recorder runs the standard map or its pair, lifted or wrapped");
  const int ns = pair ? s / 2 : s;  // the point columns hold orbit A's state (the first ns components)
  IndRowDesc d;
  memset(&d, 0, sizeof(d));
  const bool window = ch->n >= 2 && ch->r[0].kind == CL_RED_RUNMONITOR;
  if (window) {
    d.n = 0;
    d.wmon = 0;
  } else {
    if (ch->n < 2 || ch->n - 1 > IND_MAX_MEMBERS || ch->r[ch->n - 1].kind != CL_RED_RUNMONITOR)
      return set_err(ctx, CL_ERR_UNSUPPORTED, "a series kind's chain is 1 to 4 members then RunMonitor");
    int w = 0, cols = 0;
    for (int j = 0; j + 1 < ch->n; ++j) {
      const int kd = ch->r[j].kind;
      if (ch->r[j].frame != 0) return set_err(ctx, CL_ERR_UNSUPPORTED, "members read frame 1");
      d.kind[j] = kd;
      d.off[j] = w;
      if (kd == CL_RED_FLI) { w += (int)(sizeof(FLIState) / 8); cols += 2; }
      else if (kd == CL_RED_MEGNO) { w += (int)(sizeof(MEGNOState) / 8); cols += 6; }
      else if (kd == CL_RED_SALI) { w += (int)(sizeof(SALIState) / 8); cols += 2; }
      else if (kd == CL_RED_RLI && pair) { w += (int)(sizeof(RLIState) / 8); cols += 2; }
      else return set_err(ctx, CL_ERR_UNSUPPORTED, "series members are FLI, MEGNO and SALI (and RLI on a pair)");
    }
    d.n = ch->n - 1;
    d.wmon = w;
    d.ncol = cols;
  }
  d.ncol += 3 + ns;
  if (m < 0 || G < 1 || g < 0 || g >= G) return set_err(ctx, CL_ERR_ARGUMENT, "grid row out of range");
  if (m > 0 && (!dx || !dst || !dcode || !dfail || !dR)) return set_err(ctx, CL_ERR_ARGUMENT, "null buffer");
  hipSetDevice(ctx->device);
  int rc = ind_rows_launch(m, ns, d, g, G, dx, dst, dcode, dfail, dR, stream ? (hipStream_t)stream : ctx->stream);
  return rc ? set_err(ctx, rc, "indicator row kernel launch failed") : CL_OK;
}

extern "C" int cl_ind_row_width(const cl_chain* ch, int s) {
  if (!ch || ch->n < 1) return CL_ERR_ARGUMENT;
  if (ch->r[0].kind == CL_RED_RUNMONITOR) return 3 + s;
  int cols = 0;
  for (int j = 0; j + 1 < ch->n; ++j) {
    const int kd = ch->r[j].kind;
    cols += kd == CL_RED_MEGNO ? 6 : (kd == CL_RED_FLI || kd == CL_RED_SALI || kd == CL_RED_RLI) ? 2 : 0;
  }
  return cols + 3 + s;
}

// The order of the chain's HouseholderQR for this context's calls: the blocked factorisation on the matrix cores
// unless the tuning asks for the lane-butterfly sweep.
static bool chain_blocked(const cl_ctx* c) { return c->tune.chain_qr != 2; }

static int chain_device(cl_ctx* ctx, const cl_system* sys, const cl_chain* ch, int elt, int64_t m, int s, int d,
                        int k, int64_t n0, int64_t steps, int64_t N, int tau, double* dx, double* dW, uint64_t* dst,
                        int8_t* dcode, int64_t* dhstep, hipStream_t st) {
  int words = 0;
  bool qr = false;
  int rc = validate_chain(ctx, sys, ch, elt, m, s, d, k, tau, &words, &qr);
  if (rc) return rc;
  if (!(N >= 1 && N % tau == 0 && steps >= 0 && steps % tau == 0 && n0 >= 0 && n0 % tau == 0 && n0 + steps <= N))
    return set_err(ctx, CL_ERR_ARGUMENT, "launch steps must be a multiple of tau within N");
  rc = chain_launch(chain_args(sys, ch, m, dx, dW, dst, dcode, dhstep, n0, steps, N, tau), sys->N, sys->ring != 0,
                    qr, chain_blocked(ctx), elt == CL_ELT_F64_F32TAN, st);
  return rc ? set_err(ctx, rc, "chain kernel launch failed") : CL_OK;
}

extern "C" int cl_propagate_device(cl_ctx* ctx, const cl_system* sys, const cl_chain* ch, int elt, int64_t m,
                                   int s, int d, int k, int64_t n0, int64_t steps, int64_t N, int tau,
                                   double* dx, double* dW, double* dW2, uint64_t* dst, int8_t* dcode,
                                   int64_t* dhstep, void* stream) {
  if (!ctx) return CL_ERR_ARGUMENT;
  if (sys && ch && sys->kind == CL_SYS_STANDARD_MAP_CHAIN) {
    if (dW2) return set_err(ctx, CL_ERR_UNSUPPORTED, "chain kernels carry one frame");
    hipSetDevice(ctx->device);
    return chain_device(ctx, sys, ch, elt, m, s, d, k, n0, steps, N, tau, dx, dW, dst, dcode, dhstep,
                        stream ? (hipStream_t)stream : ctx->stream);
  }
  const Instance* in = nullptr;
  int rc = validate(ctx, sys, ch, elt, m, s, d, k, n0, steps, N, tau, &in);
  if (rc) return rc;
  if (in->nf > 1 && !dW2) return set_err(ctx, CL_ERR_ARGUMENT, "the chain uses a second frame: W2 is required");
  if (ws_entries(ch) != 0)
    return set_err(ctx, CL_ERR_ARGUMENT, "the chain requests a workspace: call cl_propagate_device_ws");
  hipSetDevice(ctx->device);
  hipStream_t st = stream ? (hipStream_t)stream : ctx->stream;
  g_tune = ctx->tune;
  rc = launch_best(in, sys, ch, m, s, d, k, dx, dW, dW2, dst, dcode, dhstep, n0, steps, N, tau, st);
  if (rc) return set_err(ctx, rc, "kernel launch failed");
  return CL_OK;
}

extern "C" int cl_propagate_device_ws(cl_ctx* ctx, const cl_system* sys, const cl_chain* ch, int elt, int64_t m,
                                      int s, int d, int k, int64_t n0, int64_t steps, int64_t N, int tau,
                                      double* dx, double* dW, double* dW2, uint64_t* dst, int8_t* dcode,
                                      int64_t* dhstep, uint32_t* dws, void* stream) {
  if (!ctx) return CL_ERR_ARGUMENT;
  if (sys && sys->kind == CL_SYS_STANDARD_MAP_CHAIN)
    return set_err(ctx, CL_ERR_UNSUPPORTED, "chain kernels run no workspace reducer");
  const Instance* in = nullptr;
  int rc = validate(ctx, sys, ch, elt, m, s, d, k, n0, steps, N, tau, &in);
  if (rc) return rc;
  if (in->nf > 1 && !dW2) return set_err(ctx, CL_ERR_ARGUMENT, "the chain uses a second frame: W2 is required");
  const int64_t e = ws_entries(ch);
  if (e == 0) return set_err(ctx, CL_ERR_ARGUMENT, "the chain requests no workspace: call cl_propagate_device");
  if (m > 0 && !dws) return set_err(ctx, CL_ERR_ARGUMENT, "null workspace");
  hipSetDevice(ctx->device);
  hipStream_t st = stream ? (hipStream_t)stream : ctx->stream;
  g_tune = ctx->tune;
  rc = launch_best(in, sys, ch, m, s, d, k, dx, dW, dW2, dst, dcode, dhstep, n0, steps, N, tau, st, dws);
  if (rc) return set_err(ctx, rc, "kernel launch failed");
  return CL_OK;
}

extern "C" int cl_propagate_device_traced(cl_ctx* ctx, const cl_system* sys, const cl_chain* ch, int elt,
                                          int64_t m, int s, int d, int k, int64_t n0, int64_t steps, int64_t N,
                                          int tau, double* dx, double* dW, double* dW2, uint64_t* dst, int8_t* dcode,
                                          int64_t* dhstep, float* dtrace, int64_t trace_n0, int trace_len,
                                          void* stream) {
  if (!ctx) return CL_ERR_ARGUMENT;
  const Instance* in = nullptr;
  int rc = validate(ctx, sys, ch, elt, m, s, d, k, n0, steps, N, tau, &in);
  if (rc) return rc;
  if (in->nf > 1 && !dW2) return set_err(ctx, CL_ERR_ARGUMENT, "the chain uses a second frame: W2 is required");
  if (trace_len < 0 || (trace_len > 0 && !dtrace)) return set_err(ctx, CL_ERR_ARGUMENT, "bad trace buffer");
  hipSetDevice(ctx->device);
  hipStream_t st = stream ? (hipStream_t)stream : ctx->stream;
  rc = launch(in, sys, ch, m, dx, dW, dW2, dst, dcode, dhstep, n0, steps, N, tau, st, dtrace, trace_n0, trace_len);
  if (rc) return set_err(ctx, rc, "kernel launch failed");
  return CL_OK;
}

enum { B_X, B_W, B_W2, B_ST, B_CODE, B_H, B_XA, B_WA, B_W2A, B_STA, B_WS, B_WSA, B_MLP, B_MLPS };

// The surrogate GEMM's split-K partial sums (mlp.hip): one buffer per context, grown on demand, so GEMMs of one
// context that split K must not overlap (the surrogate issues them in order on one stream). Null on failure.
extern "C" __attribute__((visibility("hidden"))) void* cl_ctx_mlp_workspace(cl_ctx* c, size_t bytes) {
  return grow(c, B_MLP, bytes) == CL_OK ? c->buf[B_MLP] : nullptr;
}

// The surrogate training step's activations, gradients and loss partials (cl_mlp_train_step): one buffer per
// context, grown on demand; steps of one context run in order on one stream. Null on failure.
extern "C" __attribute__((visibility("hidden"))) void* cl_ctx_mlp_step_buffer(cl_ctx* c, size_t bytes) {
  return grow(c, B_MLPS, bytes) == CL_OK ? c->buf[B_MLPS] : nullptr;
}

extern "C" int cl_propagate_ragged(cl_ctx* ctx, const cl_system* sys, const cl_chain* ch, int elt, int64_t m, int s,
                                   int d, int k, const int64_t* Ns, const int64_t* order, int tau, double* dx,
                                   double* dW, uint64_t* dst, int8_t* dcode, int64_t* dhstep, void* stream) {
  if (!ctx || !sys || !ch) return CL_ERR_ARGUMENT;
  if (sys->kind != CL_SYS_STANDARD_MAP_CHAIN)
    return set_err(ctx, CL_ERR_UNSUPPORTED, "per-orbit step counts are built for the chain kernels");
  int words = 0;
  bool qr = false;
  int rc = validate_chain(ctx, sys, ch, elt, m, s, d, k, tau, &words, &qr);
  if (rc) return rc;
  if (m == 0) return CL_OK;
  if (!Ns) return set_err(ctx, CL_ERR_ARGUMENT, "null step counts");
  int64_t Nmax = 0;
  for (int64_t i = 0; i < m; ++i) {
    if (!(Ns[i] >= 1 && Ns[i] % tau == 0)) return set_err(ctx, CL_ERR_ARGUMENT, "each N_i must be a positive multiple of tau");
    Nmax = Ns[i] > Nmax ? Ns[i] : Nmax;
  }
  if (order) {
    std::vector<char> seen((size_t)m, 0);
    for (int64_t i = 0; i < m; ++i) {
      if (order[i] < 0 || order[i] >= m || seen[(size_t)order[i]]) return set_err(ctx, CL_ERR_ARGUMENT, "order is not a permutation");
      seen[(size_t)order[i]] = 1;
    }
  }
  hipSetDevice(ctx->device);
  hipStream_t st = stream ? (hipStream_t)stream : ctx->stream;
  // step counts and order in stream-ordered device memory, freed after the kernel on the same stream, so
  // calls on different streams can run concurrently
  int64_t* dbuf = nullptr;
  const size_t nb = (size_t)m * 8 * (order ? 2 : 1);
  hipError_t e = hipMallocAsync((void**)&dbuf, nb, st);
  if (e != hipSuccess) return set_err(ctx, map_err(e), "device allocation failed");
  e = hipMemcpyAsync(dbuf, Ns, (size_t)m * 8, hipMemcpyHostToDevice, st);
  if (e == hipSuccess && order) e = hipMemcpyAsync(dbuf + m, order, (size_t)m * 8, hipMemcpyHostToDevice, st);
  if (e != hipSuccess) {
    hipFreeAsync(dbuf, st);
    return set_err(ctx, map_err(e), "copy of step counts failed");
  }
  ChainArgs a = chain_args(sys, ch, m, dx, dW, dst, dcode, dhstep, 0, Nmax, Nmax, tau);
  a.Ns = dbuf;
  a.order = order ? dbuf + m : nullptr;
  rc = chain_launch(a, sys->N, sys->ring != 0, qr, chain_blocked(ctx), elt == CL_ELT_F64_F32TAN, st);
  e = hipFreeAsync(dbuf, st);
  if (rc) return set_err(ctx, rc, "chain kernel launch failed");
  return e == hipSuccess ? CL_OK : set_err(ctx, map_err(e), "stream-ordered free failed");
}

// ------------------------------------------------------------------------------ mixed lattices, one persistent launch
// The host side of cl_propagate_mixed: validation, the per-lattice launch orders (longest first), the work list
// (items of up to mixed_slots(N) orbits, longest item first over all lattices, weighted by N: a chain step costs in
// proportion to its sites) and the device staging image, copied in one transfer.
struct MixedPlan {
  std::vector<char> image;  // staging image: classes | items | Ns per lattice | orders per lattice | counters
  size_t off_items = 0, off_counters = 0;
  int64_t nitems = 0, Nmax = 0;
  int nchunks = 0;
  bool f32 = false;  // Float32 tangents
};

static int mixed_plan(cl_ctx* ctx, const cl_system* sys, const cl_chain* ch, int elt, int nlat, const cl_lattice* lat,
                      int k, int tau, int64_t chunk, MixedPlan& P, std::vector<MixedClass>& cls) {
  if (!sys || !ch || (nlat > 0 && !lat)) return set_err(ctx, CL_ERR_ARGUMENT, "null descriptor");
  if (sys->kind != CL_SYS_STANDARD_MAP_CHAIN) return set_err(ctx, CL_ERR_UNSUPPORTED, "the mixed launch runs chains");
  if (elt != CL_ELT_F64 && elt != CL_ELT_F64_F32TAN)
    return set_err(ctx, CL_ERR_UNSUPPORTED, "element type has no kernel for this system");
  P.f32 = elt == CL_ELT_F64_F32TAN;
  if (!(ch->n == 1 && ch->r[0].kind == CL_RED_QR && ch->r[0].alg == CL_QR_HOUSEHOLDER && ch->r[0].frame == 0))
    return set_err(ctx, CL_ERR_UNSUPPORTED, "the mixed launch runs QRSpectrum(HouseholderQR())");
  if (k != 16) return set_err(ctx, CL_ERR_UNSUPPORTED, "the mixed launch is built for k = 16");
  if (nlat < 1 || nlat > 64) return set_err(ctx, CL_ERR_ARGUMENT, "1 to 64 lattices");
  if (tau < 1 || chunk < 0 || chunk % tau) return set_err(ctx, CL_ERR_ARGUMENT, "chunk must be a multiple of tau");
  const int words = 2 * k + 9;
  size_t tot_orbits = 0;
  for (int c = 0; c < nlat; ++c) {
    const cl_lattice& L = lat[c];
    if (!chain_supported(L.N, k)) return set_err(ctx, CL_ERR_UNSUPPORTED, "no chain kernel for this N");
    if (sys->ring && L.N < 3) return set_err(ctx, CL_ERR_ARGUMENT, "a ring needs N >= 3");
    if (L.m < 0) return set_err(ctx, CL_ERR_ARGUMENT, "negative orbit count");
    if (L.m > 0 && (!L.Ns || !L.dx || !L.dW || !L.dstates || !L.dcode || !L.dhstep))
      return set_err(ctx, CL_ERR_ARGUMENT, "null buffer");
    for (int64_t i = 0; i < L.m; ++i) {
      if (!(L.Ns[i] >= 1 && L.Ns[i] % tau == 0))
        return set_err(ctx, CL_ERR_ARGUMENT, "each N_i must be a positive multiple of tau");
      P.Nmax = L.Ns[i] > P.Nmax ? L.Ns[i] : P.Nmax;
    }
    tot_orbits += (size_t)L.m;
  }
  const int64_t len = chunk == 0 ? P.Nmax : chunk;
  P.nchunks = P.Nmax == 0 ? 0 : (int)((P.Nmax + len - 1) / len);
  if (P.nchunks > 100000) return set_err(ctx, CL_ERR_ARGUMENT, "more than 10^5 chunks");
  // launch orders (longest first, ties by orbit index) and the work list
  struct Item { MixedItem w; int64_t cost; };
  std::vector<Item> items;
  std::vector<std::vector<int64_t>> order((size_t)nlat);
  for (int c = 0; c < nlat; ++c) {
    const cl_lattice& L = lat[c];
    std::vector<int64_t>& o = order[(size_t)c];
    o.resize((size_t)L.m);
    for (int64_t i = 0; i < L.m; ++i) o[(size_t)i] = i;
    std::stable_sort(o.begin(), o.end(), [&](int64_t a, int64_t b) { return L.Ns[a] > L.Ns[b]; });
    const int sl = mixed_slots(L.N, chain_blocked(ctx));
    for (int64_t g = 0; g < L.m; g += sl) {
      Item it;
      it.w.cls = c;
      it.w.nslots = (int32_t)(L.m - g < sl ? L.m - g : sl);
      it.w.g0 = g;
      it.cost = L.Ns[o[(size_t)g]] * (int64_t)L.N;
      items.push_back(it);
    }
  }
  std::stable_sort(items.begin(), items.end(), [](const Item& a, const Item& b) { return a.cost > b.cost; });
  P.nitems = (int64_t)items.size();
  auto al = [](size_t v) { return (v + 255) & ~(size_t)255; };
  size_t off = al(sizeof(MixedClass) * (size_t)nlat);
  P.off_items = off;
  off = al(off + sizeof(MixedItem) * items.size());
  const size_t off_ns = off;
  off = al(off + 8 * tot_orbits);
  const size_t off_ord = off;
  off = al(off + 8 * tot_orbits);
  P.off_counters = off;
  off = al(off + 4 * (size_t)(P.nchunks > 0 ? P.nchunks : 1));
  P.image.assign(off, 0);
  for (size_t j = 0; j < items.size(); ++j)
    memcpy(P.image.data() + P.off_items + j * sizeof(MixedItem), &items[j].w, sizeof(MixedItem));
  cls.resize((size_t)nlat);
  size_t o_ns = off_ns, o_ord = off_ord;
  for (int c = 0; c < nlat; ++c) {
    const cl_lattice& L = lat[c];
    if (L.m > 0) {
      memcpy(P.image.data() + o_ns, L.Ns, 8 * (size_t)L.m);
      memcpy(P.image.data() + o_ord, order[(size_t)c].data(), 8 * (size_t)L.m);
    }
    MixedClass& mc = cls[(size_t)c];
    memset(&mc, 0, sizeof(mc));
    mc.N = L.N;
    mc.slots = mixed_slots(L.N, chain_blocked(ctx));
    mc.a.x = L.dx;
    mc.a.W = L.dW;
    mc.a.st = L.dstates;
    mc.a.code = L.dcode;
    mc.a.hstep = L.dhstep;
    mc.a.Ns = (const int64_t*)(uintptr_t)o_ns;   // offsets; rebased onto the device address by mixed_run
    mc.a.order = (const int64_t*)(uintptr_t)o_ord;
    mc.a.m = L.m;
    mc.a.N = P.Nmax;
    mc.a.tau = tau;
    mc.a.K = sys->K;
    mc.a.eps = sys->eps;
    mc.a.tol = ch->r[0].tol;
    o_ns += 8 * (size_t)L.m;
    o_ord += 8 * (size_t)L.m;
  }
  (void)words;
  return CL_OK;
}

// Copies the staging image to `base` (device memory of at least P.image.size() bytes) and runs the chunks on st:
// one graph holding the counter fill and every chunk's launch (CL_GRAPH=0: the same launches directly).
static int mixed_run(cl_ctx* ctx, MixedPlan& P, std::vector<MixedClass>& cls, char* base, bool ring, int64_t chunk,
                     hipStream_t st, bool allow_graph = true) {
  for (MixedClass& mc : cls) {
    mc.a.Ns = (const int64_t*)(base + (uintptr_t)mc.a.Ns);
    mc.a.order = (const int64_t*)(base + (uintptr_t)mc.a.order);
  }
  memcpy(P.image.data(), cls.data(), sizeof(MixedClass) * cls.size());
  hipError_t e = hipMemcpyAsync(base, P.image.data(), P.image.size(), hipMemcpyHostToDevice, st);
  if (e != hipSuccess) return set_err(ctx, map_err(e), "copy of the work list failed");
  if (P.nchunks == 0) return CL_OK;
  const int64_t len = chunk == 0 ? P.Nmax : chunk;
  unsigned* counters = (unsigned*)(base + P.off_counters);
  auto enqueue = [&](hipStream_t s) -> int {
    int rc = mixed_zero(counters, P.nchunks, s);
    for (int c = 0; c < P.nchunks && !rc; ++c) {
      MixedArgs a;
      a.cls = (const MixedClass*)base;
      a.items = (const MixedItem*)(base + P.off_items);
      a.nitems = P.nitems;
      a.counter = counters + c;
      a.n0 = (int64_t)c * len;
      a.steps = P.Nmax - a.n0 < len ? P.Nmax - a.n0 : len;
      rc = mixed_launch(a, ring, chain_blocked(ctx), P.f32, s);
    }
    return rc;
  };
  const char* ge = getenv("CL_GRAPH");
  const bool graph = allow_graph && P.nchunks > 1 && !(ge && ge[0] == '0');
  if (!graph) {
    int rc = enqueue(st);
    return rc ? set_err(ctx, rc, "mixed chain launch failed") : CL_OK;
  }
  // capture on the context's own stream (never the legacy default stream), replay on st
  hipGraph_t g = nullptr;
  hipGraphExec_t ge2 = nullptr;
  e = hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) return set_err(ctx, map_err(e), "graph capture failed to start");
  int rc = enqueue(ctx->stream);
  hipError_t e2 = hipStreamEndCapture(ctx->stream, &g);
  if (rc || e2 != hipSuccess) {
    if (g) hipGraphDestroy(g);
    return set_err(ctx, rc ? rc : map_err(e2), "graph capture of the chunks failed");
  }
  e = hipGraphInstantiate(&ge2, g, nullptr, nullptr, 0);
  if (e == hipSuccess) e = hipGraphLaunch(ge2, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (ge2) hipGraphExecDestroy(ge2);
  hipGraphDestroy(g);
  return e == hipSuccess ? CL_OK : set_err(ctx, map_err(e), "graph launch of the chunks failed");
}

extern "C" int cl_propagate_mixed(cl_ctx* ctx, const cl_system* sys, const cl_chain* ch, int elt, int nlat,
                                  const cl_lattice* lat, int k, int tau, int64_t chunk, void* stream) {
  if (!ctx) return CL_ERR_ARGUMENT;
  MixedPlan P;
  std::vector<MixedClass> cls;
  int rc = mixed_plan(ctx, sys, ch, elt, nlat, lat, k, tau, chunk, P, cls);
  if (rc) return rc;
  hipSetDevice(ctx->device);
  mixed_cus();
  hipStream_t st = stream ? (hipStream_t)stream : ctx->stream;
  char* base = nullptr;
  hipError_t e = hipMallocAsync((void**)&base, P.image.size(), st);
  if (e != hipSuccess) return set_err(ctx, map_err(e), "device allocation failed");
  rc = mixed_run(ctx, P, cls, base, sys->ring != 0, chunk, st);
  hipFreeAsync(base, st);
  e = hipStreamSynchronize(st);
  if (rc) return rc;
  return e == hipSuccess ? CL_OK : set_err(ctx, map_err(e), "mixed chain launch failed");
}

// cl_propagate_mixed without any device allocation: the staging image goes to this device's scratch array (at most
// CL_SCRATCH_BYTES) and the chunks are launched directly on the stream (no graph is instantiated). Calls that overlap
// on one device share the array and race.
extern "C" int cl_propagate_mixed_noalloc(cl_ctx* ctx, const cl_system* sys, const cl_chain* ch, int elt, int nlat,
                                          const cl_lattice* lat, int k, int tau, int64_t chunk, void* stream) {
  if (!ctx) return CL_ERR_ARGUMENT;
  MixedPlan P;
  std::vector<MixedClass> cls;
  int rc = mixed_plan(ctx, sys, ch, elt, nlat, lat, k, tau, chunk, P, cls);
  if (rc) return rc;
  if (P.image.size() > CL_SCRATCH_BYTES)
    return set_err(ctx, CL_ERR_ARGUMENT, "work list, step counts and orders exceed the 4 MiB scratch (about 2e5 orbits)");
  char* base = ctx->device < CL_MAX_DEVICES ? (char*)g_scratch_ptr[ctx->device] : nullptr;
  if (!base) return set_err(ctx, CL_ERR_DEVICE_LOST, "the scratch array has no address on this device");
  hipSetDevice(ctx->device);
  mixed_cus();
  hipStream_t st = stream ? (hipStream_t)stream : ctx->stream;
  rc = mixed_run(ctx, P, cls, base, sys->ring != 0, chunk, st, false);
  hipError_t e = hipStreamSynchronize(st);
  if (rc) return rc;
  return e == hipSuccess ? CL_OK : set_err(ctx, map_err(e), "mixed chain launch failed");
}

// The chunk entry point for the chain: host and device buffers share the orbit-major layout.
static int chain_chunk(cl_ctx* ctx, const cl_system* sys, const cl_chain* ch, int elt, int64_t m, int s, int d, int k,
                       int64_t n0, int64_t steps, int64_t N, int tau, double* x, double* W, void* states,
                       int8_t* code, int64_t* hstep) {
  int words = 0;
  bool qr = false;
  int rc = validate_chain(ctx, sys, ch, elt, m, s, d, k, tau, &words, &qr);
  if (rc) return rc;
  if (m == 0) return CL_OK;
  if (!x || !W || !code || !hstep || (words > 0 && !states)) return set_err(ctx, CL_ERR_ARGUMENT, "null buffer");
  hipSetDevice(ctx->device);
  hipStream_t st = ctx->stream;
  size_t bx = (size_t)m * s * 8, bw = (size_t)m * d * k * 8, bs = (size_t)m * words * 8;
  if ((rc = grow(ctx, B_X, bx)) || (rc = grow(ctx, B_W, bw)) || (rc = grow(ctx, B_ST, bs)) ||
      (rc = grow(ctx, B_CODE, (size_t)m)) || (rc = grow(ctx, B_H, (size_t)m * 8)))
    return set_err(ctx, rc, "device allocation failed");
  double *dx = (double*)ctx->buf[B_X], *dW = (double*)ctx->buf[B_W];
  uint64_t* dst = (uint64_t*)ctx->buf[B_ST];
  int8_t* dcode = (int8_t*)ctx->buf[B_CODE];
  int64_t* dh = (int64_t*)ctx->buf[B_H];
  hipError_t e = hipSuccess;
#define CK(call)                                      \
  if ((e = (call)) != hipSuccess) {                   \
    return set_err(ctx, map_err(e), #call " failed"); \
  }
  CK(hipMemcpyAsync(dx, x, bx, hipMemcpyHostToDevice, st));
  CK(hipMemcpyAsync(dW, W, bw, hipMemcpyHostToDevice, st));
  if (n0 > 0) {
    if (words > 0) CK(hipMemcpyAsync(dst, states, bs, hipMemcpyHostToDevice, st));
    CK(hipMemcpyAsync(dcode, code, (size_t)m, hipMemcpyHostToDevice, st));
    CK(hipMemcpyAsync(dh, hstep, (size_t)m * 8, hipMemcpyHostToDevice, st));
  }
  rc = chain_device(ctx, sys, ch, elt, m, s, d, k, n0, steps, N, tau, dx, dW, dst, dcode, dh, st);
  if (rc) return rc;
  CK(hipMemcpyAsync(x, dx, bx, hipMemcpyDeviceToHost, st));
  CK(hipMemcpyAsync(W, dW, bw, hipMemcpyDeviceToHost, st));
  if (words > 0) CK(hipMemcpyAsync(states, dst, bs, hipMemcpyDeviceToHost, st));
  CK(hipMemcpyAsync(code, dcode, (size_t)m, hipMemcpyDeviceToHost, st));
  CK(hipMemcpyAsync(hstep, dh, (size_t)m * 8, hipMemcpyDeviceToHost, st));
  CK(hipStreamSynchronize(st));
#undef CK
  return CL_OK;
}

// ------------------------------------------------------------------------------ two-trajectory estimator
namespace cl {
int shadow_launch(const cl_system* s, const double* x0, const double* u0, uint64_t* e, int64_t m, int64_t n0,
                  int64_t steps, int64_t N, int tau, double d0, hipStream_t st);
int shadow_rows_launch(int64_t m, int s, int g, int G, int64_t target, const uint64_t* e, int64_t* fail, double* R,
                       hipStream_t st);
}

static int validate_shadow(cl_ctx* ctx, const cl_system* sys, int64_t m, int s, int64_t n0, int64_t steps, int64_t N,
                           int tau, double d0) {
  if (!ctx || !sys) return CL_ERR_ARGUMENT;
  const bool sys_ok = (sys->kind == CL_SYS_STANDARD_MAP && s == 2) || (sys->kind == CL_SYS_WRAPPED_STANDARD_MAP && s == 4);
  if (!sys_ok) return set_err(ctx, CL_ERR_UNSUPPORTED, "the two-trajectory kernel runs the standard map, lifted or wrapped");
  if (m < 0 || !(tau >= 1 && N >= 1 && N % tau == 0 && steps >= 0 && steps % tau == 0 && n0 >= 0 && n0 % tau == 0 &&
                 n0 + steps <= N))
    return set_err(ctx, CL_ERR_ARGUMENT, "launch steps must be a multiple of tau within N");
  if (!(d0 > 0.0 && d0 < __builtin_inf())) return set_err(ctx, CL_ERR_ARGUMENT, "d0 must be finite and positive");
  return CL_OK;
}

extern "C" int cl_shadow_device(cl_ctx* ctx, const cl_system* sys, int64_t m, int s, int64_t n0, int64_t steps,
                                int64_t N, int tau, double d0, const double* dx0, const double* du0, uint64_t* dends,
                                void* stream) {
  int rc = validate_shadow(ctx, sys, m, s, n0, steps, N, tau, d0);
  if (rc) return rc;
  if (m > 0 && (!dends || (n0 == 0 && (!dx0 || !du0)))) return set_err(ctx, CL_ERR_ARGUMENT, "null buffer");
  hipSetDevice(ctx->device);
  rc = shadow_launch(sys, dx0, du0, dends, m, n0, steps, N, tau, d0, stream ? (hipStream_t)stream : ctx->stream);
  return rc ? set_err(ctx, rc, "two-trajectory kernel launch failed") : CL_OK;
}

extern "C" int cl_propagate_shadow_chunk(cl_ctx* ctx, const cl_system* sys, int64_t m, int s, int64_t n0,
                                         int64_t steps, int64_t N, int tau, double d0, const double* x0,
                                         const double* u0, void* ends) {
  int rc = validate_shadow(ctx, sys, m, s, n0, steps, N, tau, d0);
  if (rc) return rc;
  if (m == 0) return CL_OK;
  if (!x0 || !u0 || !ends) return set_err(ctx, CL_ERR_ARGUMENT, "null buffer");
  hipSetDevice(ctx->device);
  hipStream_t st = ctx->stream;
  const int ew = 2 * s + 6;
  const size_t bx = (size_t)m * s * 8, bu = (size_t)m * 16, be = (size_t)m * ew * 8;
  char* base = nullptr;
  const size_t tot = 2 * (bx + bu + be) + 1024;
  hipError_t e = hipMallocAsync((void**)&base, tot, st);
  if (e != hipSuccess) return set_err(ctx, map_err(e), "device allocation failed");
  uint64_t *xa = (uint64_t*)base, *ua = (uint64_t*)(base + bx), *ea = (uint64_t*)(base + bx + bu);
  uint64_t *xs = (uint64_t*)(base + bx + bu + be), *us = (uint64_t*)((char*)xs + bx), *es = (uint64_t*)((char*)us + bu);
  const int B = 256;
  if ((e = hipMemcpyAsync(xa, x0, bx, hipMemcpyHostToDevice, st)) != hipSuccess ||
      (e = hipMemcpyAsync(ua, u0, bu, hipMemcpyHostToDevice, st)) != hipSuccess ||
      (e = hipMemcpyAsync(ea, ends, be, hipMemcpyHostToDevice, st)) != hipSuccess) {
    hipFreeAsync(base, st);
    return set_err(ctx, map_err(e), "copy to the device failed");
  }
  hipLaunchKernelGGL(k_aos_to_soa, dim3(grid_for(m * s, B)), dim3(B), 0, st, xa, xs, m, s);
  hipLaunchKernelGGL(k_aos_to_soa, dim3(grid_for(m * 2, B)), dim3(B), 0, st, ua, us, m, 2);
  hipLaunchKernelGGL(k_aos_to_soa, dim3(grid_for(m * ew, B)), dim3(B), 0, st, ea, es, m, ew);
  rc = shadow_launch(sys, (double*)xs, (double*)us, es, m, n0, steps, N, tau, d0, st);
  if (!rc) {
    hipLaunchKernelGGL(k_soa_to_aos, dim3(grid_for(m * ew, B)), dim3(B), 0, st, es, ea, m, ew);
    e = hipMemcpyAsync(ends, ea, be, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) rc = map_err(e);
  }
  hipFreeAsync(base, st);
  hipStreamSynchronize(st);
  return rc ? set_err(ctx, rc, "two-trajectory chunk failed") : CL_OK;
}

extern "C" int cl_shadow_record_rows(cl_ctx* ctx, int64_t m, int s, int g, int G, int64_t target,
                                     const uint64_t* dends, int64_t* dfail, double* dR, void* stream) {
  if (!ctx) return CL_ERR_ARGUMENT;
  if (m < 0 || (s != 2 && s != 4) || G < 1 || g < 0 || g >= G || target < 1)
    return set_err(ctx, CL_ERR_ARGUMENT, "grid row out of range");
  if (m > 0 && (!dends || !dfail || !dR)) return set_err(ctx, CL_ERR_ARGUMENT, "null buffer");
  hipSetDevice(ctx->device);
  int rc = shadow_rows_launch(m, s, g, G, target, dends, dfail, dR, stream ? (hipStream_t)stream : ctx->stream);
  return rc ? set_err(ctx, rc, "record kernel launch failed") : CL_OK;
}

extern "C" int cl_propagate_chunk(cl_ctx* ctx, const cl_system* sys, const cl_chain* ch, int elt, int64_t m,
                                  int s, int d, int k, int64_t n0, int64_t steps, int64_t N, int tau, double* x,
                                  double* W, double* W2, void* states, int8_t* code, int64_t* hstep,
                                  uint32_t* ws) {
  if (!ctx) return CL_ERR_ARGUMENT;
  if (sys && ch && sys->kind == CL_SYS_STANDARD_MAP_CHAIN) {
    if (W2) return set_err(ctx, CL_ERR_UNSUPPORTED, "chain kernels carry one frame");
    return chain_chunk(ctx, sys, ch, elt, m, s, d, k, n0, steps, N, tau, x, W, states, code, hstep);
  }
  const Instance* in = nullptr;
  int rc = validate(ctx, sys, ch, elt, m, s, d, k, n0, steps, N, tau, &in);
  if (rc) return rc;
  if (m == 0) return CL_OK;
  const int64_t nws = ws_entries(ch);
  if (!x || !W || !code || !hstep || (in->words + wb_mem_words(ch) > 0 && !states) || (in->nf > 1 && !W2) ||
      (nws > 0 && !ws))
    return set_err(ctx, CL_ERR_ARGUMENT, "null buffer");
  hipSetDevice(ctx->device);
  hipStream_t st = ctx->stream;
  const int dk = d * k, nw = in->words + (int)wb_mem_words(ch);
  uint32_t *dws = nullptr, *wsa = nullptr;
  const size_t bws = (size_t)m * (size_t)(nws > 0 ? nws : 0) * 4;
  if (nws > 0) {
    if ((rc = grow(ctx, B_WS, bws)) || (rc = grow(ctx, B_WSA, bws))) return set_err(ctx, rc, "device allocation failed");
    dws = (uint32_t*)ctx->buf[B_WS];
    wsa = (uint32_t*)ctx->buf[B_WSA];
  }
  size_t bx = (size_t)m * s * 8, bw = (size_t)m * dk * 8, bs = (size_t)m * nw * 8;
  if ((rc = grow(ctx, B_X, bx)) || (rc = grow(ctx, B_W, bw)) || (rc = grow(ctx, B_ST, bs)) ||
      (rc = grow(ctx, B_CODE, (size_t)m)) || (rc = grow(ctx, B_H, (size_t)m * 8)) || (rc = grow(ctx, B_XA, bx)) ||
      (rc = grow(ctx, B_WA, bw)) || (rc = grow(ctx, B_STA, bs)))
    return set_err(ctx, rc, "device allocation failed");
  if (in->nf > 1 && ((rc = grow(ctx, B_W2, bw)) || (rc = grow(ctx, B_W2A, bw))))
    return set_err(ctx, rc, "device allocation failed");
  uint64_t *dx = (uint64_t*)ctx->buf[B_X], *dW = (uint64_t*)ctx->buf[B_W], *dW2 = (uint64_t*)ctx->buf[B_W2];
  uint64_t *dst = (uint64_t*)ctx->buf[B_ST], *xa = (uint64_t*)ctx->buf[B_XA], *wa = (uint64_t*)ctx->buf[B_WA];
  uint64_t *w2a = (uint64_t*)ctx->buf[B_W2A], *sta = (uint64_t*)ctx->buf[B_STA];
  int8_t* dcode = (int8_t*)ctx->buf[B_CODE];
  int64_t* dh = (int64_t*)ctx->buf[B_H];
  hipError_t e = hipSuccess;
#define CK(call)                                   \
  if ((e = (call)) != hipSuccess) {                \
    return set_err(ctx, map_err(e), #call " failed"); \
  }
  const int B = 256;
  CK(hipMemcpyAsync(xa, x, bx, hipMemcpyHostToDevice, st));
  CK(hipMemcpyAsync(wa, W, bw, hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_aos_to_soa, dim3(grid_for(m * s, B)), dim3(B), 0, st, xa, dx, m, s);
  hipLaunchKernelGGL(k_aos_to_soa, dim3(grid_for(m * dk, B)), dim3(B), 0, st, wa, dW, m, dk);
  if (in->nf > 1) {
    CK(hipMemcpyAsync(w2a, W2, bw, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_aos_to_soa, dim3(grid_for(m * dk, B)), dim3(B), 0, st, w2a, dW2, m, dk);
  }
  if (n0 > 0) {
    if (nw > 0) {
      CK(hipMemcpyAsync(sta, states, bs, hipMemcpyHostToDevice, st));
      hipLaunchKernelGGL(k_aos_to_soa, dim3(grid_for(m * nw, B)), dim3(B), 0, st, sta, dst, m, nw);
    }
    CK(hipMemcpyAsync(dcode, code, (size_t)m, hipMemcpyHostToDevice, st));
    CK(hipMemcpyAsync(dh, hstep, (size_t)m * 8, hipMemcpyHostToDevice, st));
    if (nws > 0) {
      CK(hipMemcpyAsync(wsa, ws, bws, hipMemcpyHostToDevice, st));
      hipLaunchKernelGGL(k_aos_to_soa32, dim3(grid_for(m * nws, B)), dim3(B), 0, st, wsa, dws, m, nws);
    }
  }
  g_tune = ctx->tune;
  rc = launch_best(in, sys, ch, m, s, d, k, (double*)dx, (double*)dW, (double*)dW2, dst, dcode, dh, n0, steps, N,
                   tau, st, dws);
  if (rc) return set_err(ctx, rc, "kernel launch failed");
  if (nws > 0) {
    hipLaunchKernelGGL(k_soa_to_aos32, dim3(grid_for(m * nws, B)), dim3(B), 0, st, dws, wsa, m, nws);
    CK(hipMemcpyAsync(ws, wsa, bws, hipMemcpyDeviceToHost, st));
  }
  hipLaunchKernelGGL(k_soa_to_aos, dim3(grid_for(m * s, B)), dim3(B), 0, st, dx, xa, m, s);
  hipLaunchKernelGGL(k_soa_to_aos, dim3(grid_for(m * dk, B)), dim3(B), 0, st, dW, wa, m, dk);
  CK(hipMemcpyAsync(x, xa, bx, hipMemcpyDeviceToHost, st));
  CK(hipMemcpyAsync(W, wa, bw, hipMemcpyDeviceToHost, st));
  if (in->nf > 1) {
    hipLaunchKernelGGL(k_soa_to_aos, dim3(grid_for(m * dk, B)), dim3(B), 0, st, dW2, w2a, m, dk);
    CK(hipMemcpyAsync(W2, w2a, bw, hipMemcpyDeviceToHost, st));
  }
  if (nw > 0) {
    hipLaunchKernelGGL(k_soa_to_aos, dim3(grid_for(m * nw, B)), dim3(B), 0, st, dst, sta, m, nw);
    CK(hipMemcpyAsync(states, sta, bs, hipMemcpyDeviceToHost, st));
  }
  CK(hipMemcpyAsync(code, dcode, (size_t)m, hipMemcpyDeviceToHost, st));
  CK(hipMemcpyAsync(hstep, dh, (size_t)m * 8, hipMemcpyDeviceToHost, st));
  CK(hipStreamSynchronize(st));
#undef CK
  return CL_OK;
}

// ------------------------------------------------------------------------------ frames
template <int D, int K>
static void frames_host(uint64_t seed, int64_t orbit0, int64_t m, double* W) {
  for (int64_t i = 0; i < m; ++i) {
    Frame<D, K> f;
    seeded_frame<D, K>(seed, orbit0 + i, f);
    for (int e = 0; e < D * K; ++e) W[i * D * K + e] = f.w[e];
  }
}

#define FRAME_SIZES(X) X(2, 1) X(2, 2) X(4, 1) X(4, 2) X(4, 3) X(4, 4)

// Frames of the chain sizes (d > 4): one 64-lane group per orbit, the d x k frame in LDS (d k doubles). The lanes
// draw the entries and apply each projection and scaling over the rows; the dot products and the norm's sum run in
// row order in every lane (the same serial sums and maximum as seeded_frame_dyn, so the bits are the host's).
__global__ void __launch_bounds__(64) k_frames_dyn(uint64_t seed, int64_t orbit0, int64_t m, int d, int k,
                                                   double* __restrict__ W) {
  extern __shared__ double G[];
  const int64_t i = blockIdx.x;
  if (i >= m) return;
  const int lane = (int)threadIdx.x;
  const uint64_t key = orbit_key(seed, orbit0 + i);
  for (int t = lane; t < d * k; t += 64) G[t] = normal_draw(key, (uint64_t)t);
  __syncthreads();
  for (int j = 0; j < k; ++j) {
    double* g = G + j * d;
    for (int pass = 0; pass < 2; ++pass)
      for (int q = 0; q < j; ++q) {
        const double* gq = G + q * d;
        double c = gq[0] * g[0];
        for (int l = 1; l < d; ++l) c = c + gq[l] * g[l];
        __syncthreads();  // every lane has read g before it changes
        for (int l = lane; l < d; l += 64) g[l] = g[l] - c * gq[l];
        __syncthreads();
      }
    double mx = cl_fabs(g[0]);
    for (int l = 1; l < d; ++l) mx = cl_jmax(mx, cl_fabs(g[l]));
    double nr = mx;
    if (cl_isfinite(mx) && mx > 0.0) {
      double s = (g[0] / mx) * (g[0] / mx);
      for (int l = 1; l < d; ++l) {
        const double b = g[l] / mx;
        s = s + b * b;
      }
      nr = mx * cl_sqrt(s);
    }
    __syncthreads();
    for (int l = lane; l < d; l += 64) g[l] = g[l] / nr;
    __syncthreads();
  }
  double* out = W + i * (int64_t)d * k;
  for (int t = lane; t < d * k; t += 64) out[t] = G[t];
}

static bool dyn_frame_ok(int d, int k) { return d > 4 && d <= 256 && k >= 1 && k <= d && k <= 16; }

extern "C" int cl_initial_deviations(uint64_t seed, int64_t orbit0, int64_t m, int d, int k, double* W) {
  if (m < 0 || !W) return CL_ERR_ARGUMENT;
#define X(D, K) \
  if (d == D && k == K) { frames_host<D, K>(seed, orbit0, m, W); return CL_OK; }
  FRAME_SIZES(X)
#undef X
  if (dyn_frame_ok(d, k)) {
    for (int64_t i = 0; i < m; ++i) seeded_frame_dyn(seed, orbit0 + i, d, k, W + i * (int64_t)d * k);
    return CL_OK;
  }
  return CL_ERR_UNSUPPORTED;
}

extern "C" int cl_initial_deviations_device(cl_ctx* ctx, uint64_t seed, int64_t orbit0, int64_t m, int d, int k,
                                            double* dW, void* stream) {
  if (!ctx || m < 0 || !dW) return CL_ERR_ARGUMENT;
  if (m == 0) return CL_OK;
  hipSetDevice(ctx->device);
  hipStream_t st = stream ? (hipStream_t)stream : ctx->stream;
#define X(D, K)                                                                                    \
  if (d == D && k == K) {                                                                          \
    hipLaunchKernelGGL((k_frames<D, K>), dim3(grid_for(m, 256)), dim3(256), 0, st, seed, orbit0, m, dW); \
    return map_err(hipGetLastError());                                                             \
  }
  FRAME_SIZES(X)
#undef X
  if (dyn_frame_ok(d, k)) {
    hipLaunchKernelGGL(k_frames_dyn, dim3((unsigned)m), dim3(64), (size_t)d * k * sizeof(double), st, seed, orbit0, m, d,
                       k, dW);
    return map_err(hipGetLastError());
  }
  return set_err(ctx, CL_ERR_UNSUPPORTED, "no frame kernel for this d, k");
}

// ------------------------------------------------------------------------------ math test entry points
extern "C" void cl_math_host(int64_t n, const double* x, double* sn, double* cs, double* lg) {
  for (int64_t i = 0; i < n; ++i) {
    cl_sincos(x[i], &sn[i], &cs[i]);
    lg[i] = cl_log(x[i]);
  }
}

extern "C" void cl_gw_host(int64_t n, const double* R, const double* G_max, double* quick, double* exact,
                           uint8_t* settled) {
  QRSpectrum<QR_HOUSEHOLDER, 2, 2> qr;
  qr.tol = 8.0;
  for (int64_t i = 0; i < n; ++i) {
    quick[i] = gw_quick2(R + 4 * i);
    exact[i] = window_Gw<2>(R + 4 * i);
    QRState<2> st;
    qr.init(st);
    st.G_max = G_max[i];
    QRFact<2, 2> f;
    for (int e = 0; e < 4; ++e) f.R[e] = R[4 * i + e];
    settled[i] = qr.settled_by_estimate(st, f, 0.0) ? 1 : 0;
  }
}

extern "C" int cl_math_device(cl_ctx* ctx, int64_t n, const double* dx, double* dsn, double* dcs, double* dlg,
                              void* stream) {
  if (!ctx || n < 0) return CL_ERR_ARGUMENT;
  if (n == 0) return CL_OK;
  hipSetDevice(ctx->device);
  hipStream_t st = stream ? (hipStream_t)stream : ctx->stream;
  hipLaunchKernelGGL(k_math, dim3(grid_for(n, 256)), dim3(256), 0, st, n, dx, dsn, dcs, dlg);
  return map_err(hipGetLastError());
}

extern "C" int cl_sqrt12_device(cl_ctx* ctx, int64_t n, const double* dx, double* dy, void* stream) {
  if (!ctx || n < 0) return CL_ERR_ARGUMENT;
  if (n == 0) return CL_OK;
  hipSetDevice(ctx->device);
  hipStream_t st = stream ? (hipStream_t)stream : ctx->stream;
  hipLaunchKernelGGL(k_sqrt12, dim3(grid_for(n, 256)), dim3(256), 0, st, n, dx, dy);
  return map_err(hipGetLastError());
}
