/* chaoslab_hip.h - C ABI of the MI355X (gfx950) ensemble tangent-stream library.
 *
 * The library replaces the device path of Chaoslab's `propagate`: the KernelAbstractions kernel
 * `_orbit_kernel!` (src/Tangent/propagate.jl:32-50) and the chunk driver around it
 * (`_run_chunk!`, src/Tangent/propagate.jl:92-119), plus the device hooks that the CUDA package
 * extension installs (ext/ChaoslabCUDAExt.jl:22-91). A Julia `ChaoslabHIPExt` binds these symbols
 * with `ccall`; Python binds them with ctypes. No torch or HIP types appear in the signatures:
 * buffers are plain pointers and sizes, streams are opaque `void*`.
 *
 * Layouts
 *   Host buffers (cl_propagate_chunk) are the reference's own array-of-structs layouts:
 *     x      double[m][s]        (Vector{SVector{s,Float64}})
 *     W      double[m][k][d]     (Vector{SMatrix{d,k,Float64}}, column-major per orbit)
 *     states bytes[m][state_bytes] (the chain's tuple of isbits reducer states, Julia layout)
 *     code   int8[m], hstep int64[m]
 *   Device buffers (cl_propagate_device) are structure-of-arrays, one orbit per lane:
 *     x[c*m + i], W[(col*d + row)*m + i], states as 8-byte words st[w*m + i].
 *   The coupled chain (CL_SYS_STANDARD_MAP_CHAIN) spreads one orbit over 2N-row frames held by N
 *   threads; its device buffers keep the orbit-major host layout above (x[i][2N], W[i][k][2N],
 *   states[i][words]), and so do seeded frames of d > 4.
 *
 * Return codes are the indices of the reference's FALLBACK_REASONS (src/Foundation/Foundation.jl:71-75):
 *   0 ok, 2 no functional device, 3 kernel not built for this system/chain/element type,
 *   4 device out of memory, 9 device lost; 10 is an invalid argument (the reference throws
 *   ArgumentError before any work). Per-orbit numerical failures are not errors: they come back
 *   in code[] / hstep[] and in the reducer states exactly as the reference kernel writes them.
 */
#ifndef CHAOSLAB_HIP_H
#define CHAOSLAB_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ status codes */
#define CL_OK 0
#define CL_ERR_NO_DEVICE 2
#define CL_ERR_UNSUPPORTED 3
#define CL_ERR_OOM 4
#define CL_ERR_DEVICE_LOST 9
#define CL_ERR_ARGUMENT 10

/* element types: the orbit and frame arithmetic of a run */
#define CL_ELT_F64 1         /* Float64 orbit and tangents */
#define CL_ELT_F64_F32TAN 2  /* Float64 orbit, Float32 tangent frames, FP64 reorthonormalisation and log sums
                                (the standard-map family and the coupled chain) */
#define CL_ELT_DD 3          /* Double64-equivalent orbit (double-double on fma), Float64 tangents */

/* ------------------------------------------------------------------ system descriptors */
#define CL_SYS_STANDARD_MAP 1          /* StandardMap(K), src/Systems/standard_map.jl:8-43 */
#define CL_SYS_WRAPPED_STANDARD_MAP 2  /* WrappedStandardMap(K), src/Systems/wrapped_standard_map.jl:9-81 */
#define CL_SYS_STANDARD_MAP_CHAIN 3    /* StandardMapChain(N, K, eps; boundary), src/Systems/standard_map_chain.jl:11-53 */
#define CL_SYS_HENON_HEILES_GAUSS 4    /* IntegratorMap(Gauss(Float64, s; tol), HenonHeilesFlow(), h), src/Integrators/integrator_map.jl:12-23 */
#define CL_SYS_STANDARD_MAP_K 5        /* StandardMap with its K carried per orbit: state (x, p, K); no reference
                                          counterpart (the reference holds one K per system, standard_map.jl:8-10) */
#define CL_SYS_PAIR_STANDARD_MAP 6     /* PairMap(StandardMap(K)), src/Systems/pair.jl:1-68: state (A; B), s = 4, d = 4 */
#define CL_SYS_PAIR_WRAPPED_STANDARD_MAP 7 /* PairMap(WrappedStandardMap(K)): s = 8, d = 4 */

typedef struct {
  int32_t kind;     /* CL_SYS_* */
  int32_t N;        /* chain sites (kind 3) */
  int32_t ring;     /* chain boundary: 1 ring, 0 open (kind 3) */
  int32_t stages;   /* Gauss stages s (kind 4) */
  int32_t maxiter;  /* Gauss fixed-point iteration cap (kind 4; 0 means the reference default 50) */
  int32_t pad_;
  double K;         /* kick strength (kinds 1-3) */
  double eps;       /* chain coupling (kind 3) */
  double h;         /* step size (kind 4) */
  double tol;       /* Gauss acceptance bound; NaN means 8 eps max(1, |Y|) (kind 4) */
} cl_system;

/* ------------------------------------------------------------------ reducer chain */
#define CL_RED_QR 1        /* QRSpectrum(alg; orthogonality_tol), src/Lyapunov/qr_spectrum.jl:14-96 */
#define CL_RED_BENETTIN 2  /* Benettin(), src/Lyapunov/benettin.jl:11-53 */
#define CL_RED_MEGNO 3     /* MEGNO(; replace), src/Indicators/megno.jl:29-113 */
#define CL_RED_SALI 4      /* SALI(), src/Indicators/sali.jl:14-68 */
#define CL_RED_FLI 5       /* FLI(; replace), src/Indicators/fli.jl:26-96 */
#define CL_RED_OFLI 6      /* OFLI(flow; replace), src/Indicators/ofli.jl:23-126 */
#define CL_RED_GALI 7      /* GALI(), src/Indicators/gali.jl:17-201 */
#define CL_RED_MONITOR 9   /* ResidualMonitor(sys, det_tol, orth_tol), src/Diagnostics/monitor.jl:9-54 */
#define CL_RED_RUNMONITOR 10 /* RunMonitor(sys), the recorder's non-raising monitor, src/Experiments/run_record.jl:25-43 */
#define CL_RED_STRETCHING 11 /* StretchingNumbers(; range, bins, interval, min_per_cell, replace), k = 2, with a
                                UInt32 workspace of 2 (bins + 2) counts per orbit, src/Indicators/stretching.jl:34-147 */
#define CL_RED_BIRKHOFF 12   /* WeightedBirkhoff(; window, observables, birkhoff_order), any k, the last reducer of its
                                chain, src/Indicators/wbirkhoff_reducer.jl:4-116; parameters in cl_reducer.birkhoff */
#define CL_RED_RLI 13        /* RLI(; replace), k = 2 on a pair system, src/Indicators/rli.jl:1-140 */
#define CL_RED_PHASE_ADVANCE 14 /* WeightedPhaseAdvance(; window, observable, birkhoff_order), any k, a reader of the point,
                                   src/Indicators/naff_reducer.jl:142-254; 1 or 2 windows, parameters in cl_reducer.birkhoff
                                   (window, order, and the observable in kind[0], a[0], b[0]) */

/* Observables of WeightedBirkhoff (src/Indicators/wbirkhoff.jl:17-77), coordinates 0-based:
 * CL_OBS_COS cos(b x[a]), CL_OBS_SIN sin(b x[a]) (b the integer frequency), CL_OBS_LIFTED x[a] + 2 pi x[b]. */
#define CL_OBS_COS 1
#define CL_OBS_SIN 2
#define CL_OBS_LIFTED 3
/* Observables of WeightedPhaseAdvance (naff_reducer.jl:88-120): CL_OBS_PLANAR the sample x[a] + i x[b], its increment
 * the principal argument of z' conj(z); CL_OBS_LIFTED_ANGLE the pair (x[a], x[b]) of a coordinate and its winding count
 * (b = -1: none), its increment (q' - q) + 2 pi (m' - m). */
#define CL_OBS_PLANAR 4
#define CL_OBS_LIFTED_ANGLE 5
#define CL_WB_MAX_WINDOWS 16
#define CL_WB_MAX_OBSERVABLES 8
typedef struct {
  int32_t H;        /* horizons, 1..CL_WB_MAX_WINDOWS */
  int32_t K;        /* observables, 1..CL_WB_MAX_OBSERVABLES */
  int32_t order;    /* birkhoff_order p, 1..4 */
  int32_t pad_;
  int64_t window[CL_WB_MAX_WINDOWS];   /* N_h >= 16; a run has 2 max(N_h) handoffs */
  int32_t kind[CL_WB_MAX_OBSERVABLES]; /* CL_OBS_* */
  int32_t a[CL_WB_MAX_OBSERVABLES];
  int32_t b[CL_WB_MAX_OBSERVABLES];
} cl_birkhoff;

#define CL_QR_CGS 1
#define CL_QR_MGS 2
#define CL_QR_RGS 3
#define CL_QR_HOUSEHOLDER 4  /* HouseholderQR(), V-H */
#define CL_QR_AUGMENTED 5    /* AugmentedQR(), V-A */
#define CL_QR_GIVENS 6       /* GivensQR(), V-G */

typedef struct {
  int32_t kind;     /* CL_RED_* */
  int32_t alg;      /* CL_QR_* for CL_RED_QR */
  int32_t replace;  /* MEGNO / FLI / OFLI chain-position flag */
  int32_t frame;    /* 0: the frame W; 1: the second frame W2 (a second tangent frame on the same
                       orbit, so a replacing indicator chain and a QR spectrum can share one pass) */
  double tol;       /* QRSpectrum orthogonality_tol; ResidualMonitor det_tol */
  double tol2;      /* ResidualMonitor orth_tol */
  double lo, hi;    /* StretchingNumbers range (s_lo, s_hi) */
  int32_t bins;     /* StretchingNumbers cells B (the workspace holds 2 (B + 2) counts per orbit) */
  int32_t pad2_;
  const cl_birkhoff *birkhoff; /* WeightedBirkhoff parameters (host memory, read at each call) */
} cl_reducer;

#define CL_MAX_REDUCERS 8
/* cl_chain.flags: CL_CHAIN_RESET_FRAME restarts the deviation frame from the unit columns e_1..e_k after every
 * handoff, as the indicator recorder's window kinds run their readers ("one handoff at a time with the constant
 * deviation column", _ind_window_kernel!, src/Experiments/run_indicator_device.jl:66-80); the caller passes the
 * unit columns as W. */
#define CL_CHAIN_RESET_FRAME 1
typedef struct {
  int32_t n;
  int32_t flags;
  cl_reducer r[CL_MAX_REDUCERS];
} cl_chain;

/* ------------------------------------------------------------------ device probe and context */
typedef struct {
  int32_t loadable;    /* the library loaded (always 1 when this call returns) */
  int32_t functional;  /* a gfx950 device is visible */
  int32_t compiles;    /* kernels are built for the element type */
  int32_t gpu_faster;  /* always 1 for Float64 on MI355X */
  char device[64];
} cl_probe_t;

typedef struct cl_ctx cl_ctx;

/* Launch settings an autotuner picks per context (zeros: the built-in rules): `block` lanes per group of the orbit
 * kernel (64, 128 or 256; the rule is 256 from 65536 orbits on, else 64), `split` the two single-frame passes of a
 * two-frame map chain and `pipe` the wave-specialised pipeline of small QR ensembles (1 on, 2 off). These give the
 * same bits under every setting. `chain_qr` picks the order of the coupled chain's HouseholderQR (k = 16): 1 the
 * blocked factorisation on the FP64 matrix cores (two panels of 8 reflectors, compact WY trailing update), 2 the
 * unblocked sweep with the sums over sites as lane butterflies, 0 the built-in rule (blocked); the two orders agree to
 * rounding, each bit for bit with its restatement in the oracle. No reference counterpart: the reference's launch plan
 * is its measured one-wave throughput record (ext/ChaoslabCUDAExt.jl:42-62, src/Tangent/propagate.jl:144-193). */
typedef struct {
  int32_t block;
  int32_t split;
  int32_t pipe;
  int32_t chain_qr;
} cl_tuning;
int cl_set_tuning(cl_ctx *ctx, const cl_tuning *t);
int cl_get_tuning(cl_ctx *ctx, cl_tuning *t);

/* Foundation.probe(::HIPDevice, T) (src/Foundation/Foundation.jl:92-93; ext/ChaoslabCUDAExt.jl:87-91). */
int cl_probe(int elt, cl_probe_t *out);
/* Foundation.free_memory(::HIPDevice) (ext/ChaoslabCUDAExt.jl:31); -1 without a device. */
int64_t cl_free_memory(void);
/* A context owns one HIP stream (default priority) and grown-on-demand staging buffers. */
cl_ctx *cl_create(int device);
void cl_destroy(cl_ctx *ctx);
/* Text of the last error on this context (or of the last cl_create failure when ctx is NULL). */
const char *cl_last_error(cl_ctx *ctx);

/* Bytes per orbit of the chain's state tuple (reducer_bytes, src/Tangent/Tangent.jl:108-111), or a
 * negative code when the system/chain/d/k combination has no kernel (then -3) or is invalid (-10). */
int64_t cl_state_bytes(const cl_system *sys, const cl_chain *chain, int elt, int d, int k);

/* One launch of `steps` steps from global step n0 over m orbits, in place on host buffers in the
 * reference layouts: the work of one `_orbit_kernel!` launch (src/Tangent/propagate.jl:32-50)
 * including the host<->device copies of `_run_chunk!` (src/Tangent/propagate.jl:95-110).
 * n0 == 0 initialises the reducer states (the input states are ignored); an orbit whose code is not
 * 0 at n0 > 0 is left untouched. W2 is the second frame (NULL when no reducer uses frame 1).
 * ws is the reducer workspace, ws[i][cl_workspace_entries(...)] UInt32 counts per orbit (the reference's
 * ws[:, i]), zeroed by the launch at n0 == 0; NULL when the chain requests none. Synchronous. */
int cl_propagate_chunk(cl_ctx *ctx, const cl_system *sys, const cl_chain *chain, int elt,
                       int64_t m, int s, int d, int k, int64_t n0, int64_t steps, int64_t N, int tau,
                       double *x, double *W, double *W2, void *states, int8_t *code, int64_t *hstep,
                       uint32_t *ws);

/* The same launch on device-resident structure-of-arrays buffers, enqueued on `stream` (a
 * hipStream_t, NULL for the context's stream). Asynchronous. */
int cl_propagate_device(cl_ctx *ctx, const cl_system *sys, const cl_chain *chain, int elt,
                        int64_t m, int s, int d, int k, int64_t n0, int64_t steps, int64_t N, int tau,
                        double *dx, double *dW, double *dW2, uint64_t *dstates, int8_t *dcode,
                        int64_t *dhstep, void *stream);

/* cl_propagate_device for a chain with a workspace reducer (StretchingNumbers): dws[e][m], e < cl_workspace_entries,
 * UInt32 (the reference's ws[:, i], word-major), zeroed by the launch at n0 == 0 and carried across launches like the
 * states (src/Tangent/propagate.jl:29-50). cl_propagate_device refuses such a chain. */
int cl_propagate_device_ws(cl_ctx *ctx, const cl_system *sys, const cl_chain *chain, int elt,
                           int64_t m, int s, int d, int k, int64_t n0, int64_t steps, int64_t N, int tau,
                           double *dx, double *dW, double *dW2, uint64_t *dstates, int8_t *dcode,
                           int64_t *dhstep, uint32_t *dws, void *stream);
/* The unnormalised weights w(jj) = exp(-1 / (t^p (1 - t)^p)), t = min(jj, N - jj) / N, jj = 0..N-1, of a horizon N
 * (_wb and birkhoff_weight, src/Indicators/naff_reducer.jl:13-17), computed on the host with the kernels' exp, so the
 * pure function on a stored series (weighted_birkhoff, src/Indicators/wbirkhoff.jl:139-162) uses the device's weights. */
int cl_birkhoff_weights(int64_t N, int order, double *w);

/* UInt32 workspace entries per orbit of the chain (_chain_workspace, src/Tangent/orbit.jl:8-18): 2 (bins + 2) for
 * StretchingNumbers, 0 when no reducer requests one; -10 when two reducers request one. */
int64_t cl_workspace_entries(const cl_chain *chain, int d, int k);

/* cl_propagate_device plus a feature trace of the first `trace_len` handoffs after global step trace_n0:
 * float trace[3 * trace_len][m], rows 3(t-1)..3(t-1)+2 of handoff t = MEGNO's Y of column 1, FLI_B and
 * log(alpha_1). Only chains that start MEGNO then FLI on the first frame record (others write nothing).
 * No reference counterpart: this is the prefix pass of the surrogate triage. */
int cl_propagate_device_traced(cl_ctx *ctx, const cl_system *sys, const cl_chain *chain, int elt,
                               int64_t m, int s, int d, int k, int64_t n0, int64_t steps, int64_t N, int tau,
                               double *dx, double *dW, double *dW2, uint64_t *dstates, int8_t *dcode,
                               int64_t *dhstep, float *dtrace, int64_t trace_n0, int trace_len, void *stream);

/* Full runs of the coupled chain with one step count per orbit (Ns[i], host array, each a positive
 * multiple of tau) on device buffers in the orbit-major layout, fresh states. `order` (host, optional)
 * is a permutation giving the launch order of the orbits: longest first balances uneven step counts,
 * as orbit groups are dispatched to the compute units in order. Asynchronous on `stream` (NULL: the
 * context's stream); the host arrays are copied before the call returns, so calls on different streams
 * run concurrently.
 * No reference counterpart (the reference runs one N per call): the ensemble over
 * StandardMapChain(N, K, eps; boundary) (src/Systems/standard_map_chain.jl:11-53) with uneven lengths. */
int cl_propagate_ragged(cl_ctx *ctx, const cl_system *sys, const cl_chain *chain, int elt,
                        int64_t m, int s, int d, int k, const int64_t *Ns, const int64_t *order, int tau,
                        double *dx, double *dW, uint64_t *dstates, int8_t *dcode, int64_t *dhstep, void *stream);

/* One lattice of the mixed launch: N sites (8, 16, 32, 64 or 128), m orbits with their step counts (host array Ns[m],
 * positive multiples of tau) and their device buffers in the chain's orbit-major layout (x[i][2N], W[i][k][2N],
 * states[i][2k + 9] words, code[i], hstep[i]); the states start fresh. */
typedef struct {
  int32_t N;
  int32_t pad_;
  int64_t m;
  const int64_t *Ns;
  double *dx;
  double *dW;
  uint64_t *dstates;
  int8_t *dcode;
  int64_t *dhstep;
} cl_lattice;

/* Full runs of several coupled-chain lattices in ONE persistent launch (per chunk): every lattice shares the system's
 * K, eps and boundary (sys->N is ignored: each lattice carries its own) and the chain QRSpectrum(HouseholderQR())
 * with k = 16. Work items of a few orbits of one lattice are taken from a device counter, longest first over all
 * lattices. The run is split into chunks of `chunk` steps (a positive multiple of tau; 0: one chunk), captured in one
 * HIP graph when there is more than one; an orbit whose count is reached, or that failed, retires and is skipped by
 * the later chunks. Every orbit's result is bit for bit that of cl_propagate_ragged on its lattice alone.
 * Synchronous (returns after the work on `stream` (NULL: the context's stream) has finished); the device staging of
 * the step counts, the launch orders and the work list is allocated and freed by the call.
 * No reference counterpart (the reference runs one StandardMapChain per call, src/Systems/standard_map_chain.jl:11-53,
 * and relaunches in step chunks, src/Tangent/propagate.jl:144-193). */
int cl_propagate_mixed(cl_ctx *ctx, const cl_system *sys, const cl_chain *chain, int elt, int nlat,
                       const cl_lattice *lat, int k, int tau, int64_t chunk, void *stream);

/* cl_propagate_mixed without any device allocation, for callers that cannot allocate device memory (tight serving
 * loops, harnesses that forbid hipMalloc). No counterpart in the reference. The staging of the work list, step counts, launch orders and work counters lives in a fixed 4 MiB array
 * of the library on each device (about 2e5 orbits; larger requests return 10 before any device work); the chunks are
 * launched directly on the stream, no graph is instantiated. The array keeps the last call's contents and every
 * field a call reads is rewritten by that call (counters cleared by a fill kernel). Calls that overlap on one device
 * share the array and race: serialise them. Host-side planning uses ordinary heap memory. Synchronous. */
int cl_propagate_mixed_noalloc(cl_ctx *ctx, const cl_system *sys, const cl_chain *chain, int elt, int nlat,
                               const cl_lattice *lat, int k, int tau, int64_t chunk, void *stream);

/* The recorder's rows (replaces the row writes of _record_kernel!, src/Experiments/run_device.jl:9-37,41-72):
 * after the launch (cl_propagate_device) that ends at grid step g of G, row g of the orbit-major record array
 * dR[m][G][k + 7 + s] is written for every orbit still running: lambda (k), det_max, symp_max, orth, the running
 * maxima of the sum and pairing residuals (k = 2; zero columns for k = 1), G_max, M_max, the point (s = 2 lifted,
 * 4 wrapped). An orbit whose launch failed gets dfail[i] = g + 1 and no further rows; dfail and dR start zeroed.
 * The chain is (QRSpectrum or Benettin, RunMonitor) on a standard map; device buffers; asynchronous. */
int cl_record_rows(cl_ctx *ctx, const cl_system *sys, const cl_chain *chain, int64_t m, int s, int k, int g,
                   int G, const double *dx, const uint64_t *dstates, const int8_t *dcode, int64_t *dfail,
                   double *dR, void *stream);

/* The indicator recorder's rows (replaces the row writes of _ind_series_kernel! and _ind_window_kernel!,
 * src/Experiments/run_indicator_device.jl:13-29,31-60,62-91): after the launch that ends at grid step g of G, row g
 * of dR[m][G][cl_ind_row_width] is written for every orbit still running. A series kind's chain is its members
 * (FLI, MEGNO, SALI in any order, at most 4, on frame 1) then RunMonitor; the row is each member's final columns
 * (run_indicator_record.jl:39-70: FLI fli, fli1; MEGNO Y, Ybar, Ybar20 of column 1 then column 2; SALI sali, L),
 * then det_max, symp_max, M_max, then the point (s = 2 lifted, 4 wrapped). A window kind's chain is RunMonitor
 * then its reader; its row is the monitor's three maxima and the point. An orbit whose launch failed gets
 * dfail[i] = g + 1 (its 1-based grid row) and no further rows; dfail and dR start zeroed. Device buffers;
 * asynchronous. */
int cl_ind_record_rows(cl_ctx *ctx, const cl_system *sys, const cl_chain *chain, int64_t m, int s, int g, int G,
                       const double *dx, const uint64_t *dstates, const int8_t *dcode, int64_t *dfail, double *dR,
                       void *stream);
/* Columns of a row of cl_ind_record_rows for this chain and state length s (negative: bad argument). */
int cl_ind_row_width(const cl_chain *chain, int s);

/* The two-trajectory estimator (replaces _shadow_kernel!, src/Lyapunov/two_trajectory.jl:113-121, and its body
 * _shadow_body, :45-82): the reference orbit and a shadow at distance d0, placed back at d0 along the separation
 * every tau steps; the end record of each orbit (ShadowEnd{s, Float64}: x[s], y[s], sum (s, c), dprev, code + 7 pad
 * bytes, step, value; 2s + 6 words) is carried across launches. At n0 = 0 the shadow starts at x0 + d0 u0 (u0 unit,
 * 2 per orbit); later launches continue orbits whose code is 0. Standard map, lifted (s = 2, plain separation) or
 * wrapped (s = 4, minimal image, src/Lyapunov/separation.jl).
 * Device version: SoA buffers (x0[c*m + i], u0[c*m + i], ends[w*m + i]), asynchronous.
 * Host chunk version (replaces one launch of _shadow_chunk!, :124-143): AoS host buffers (x0[i][s], u0[i][2],
 * ends[i][2s + 6] words in the Julia byte layout), synchronous. */
int cl_shadow_device(cl_ctx *ctx, const cl_system *sys, int64_t m, int s, int64_t n0, int64_t steps, int64_t N,
                     int tau, double d0, const double *dx0, const double *du0, uint64_t *dends, void *stream);
int cl_propagate_shadow_chunk(cl_ctx *ctx, const cl_system *sys, int64_t m, int s, int64_t n0, int64_t steps,
                              int64_t N, int tau, double d0, const double *x0, const double *u0, void *ends);
/* The recorder's rows for the two-trajectory method (replaces _shadow_record_kernel!'s row writes,
 * src/Experiments/run_device.jl:74-98): after the launch that ends at grid step g (global step `target`), row g of
 * dR[m][G][1 + 7 + s] gets lambda = total(sum) / target in column 0 and the point in columns 8..; an orbit whose
 * record has a nonzero code gets dfail[i] = g + 1. Device buffers, asynchronous. */
int cl_shadow_record_rows(cl_ctx *ctx, int64_t m, int s, int g, int G, int64_t target, const uint64_t *dends,
                          int64_t *dfail, double *dR, void *stream);

/* Seeded orthonormal frames, replacing initial_deviations(seed, orbit, Val(d), Val(k), T)
 * (src/Tangent/initial.jl:27-34): orbit ids orbit0 .. orbit0+m-1 (1-based ids as the reference),
 * key mix(mix(seed) + id) with the reference's splitmix64 finaliser, Gaussian d x k matrix from a
 * counter stream, two modified Gram-Schmidt passes per column and unit columns (R_jj > 0).
 * Column j does not depend on k.
 * Host version: AoS output W[m][k][d]. Device version: SoA output (orbit-major for d > 4), asynchronous.
 * Sizes: d in {2, 4} with k <= d, or 4 < d <= 256 with k <= 16. */
int cl_initial_deviations(uint64_t seed, int64_t orbit0, int64_t m, int d, int k, double *W);
int cl_initial_deviations_device(cl_ctx *ctx, uint64_t seed, int64_t orbit0, int64_t m, int d, int k,
                                 double *dW, void *stream);

/* Elementary functions of the kernels, evaluated on the host or on the device (for the parity
 * tests of sin/cos/log; the kernels and these calls share one implementation). */
void cl_math_host(int64_t n, const double *x, double *sn, double *cs, double *lg);
int cl_math_device(cl_ctx *ctx, int64_t n, const double *dx, double *dsn, double *dcs, double *dlg,
                   void *stream);
/* The square root the kernels' 2-entry norms take of 1 + b^2 in [1, 2], on the device (for its parity test against
 * the host's IEEE square root; no counterpart in the reference, whose norm calls sqrt). */
int cl_sqrt12_device(cl_ctx *ctx, int64_t n, const double *dx, double *dy, void *stream);
/* G_w = ||R||_F ||R^-1||_F of n 2 x 2 upper-triangular factors R (column-major, 4 entries each) as the kernels
 * form it on the host: window_Gw (orthonormalize.jl:112-127) and the one-division estimate the kernels' checks
 * use to skip it (-1 outside its range); settled[i] is the fused QR step's decision (Householder, orth_loss 0)
 * that the estimate settles the handoff with G_max[i] as the running maximum. For the tests of the estimate's
 * error bound and of that decision. */
void cl_gw_host(int64_t n, const double *R, const double *G_max, double *quick, double *exact, uint8_t *settled);

/* ------------------------------------------------------------------ surrogate network kernels
 * No reference counterpart: the learned triage of the ensemble (a dense network on the prefix trace).
 * Matrices are row-major; bf16 values are uint16_t bit patterns; activations are feature-major
 * (features x orbits). All calls are asynchronous on `stream` (NULL: the context's stream). */

/* C = epi(op(A) op(B)): op(A) is M x K (A stored K x M when transA), op(B) is K x N (B stored N x K when
 * transB); BF16 inputs, FP32 accumulation on the matrix cores. epi: 0 none, 1 + bias[m] then ReLU,
 * 2 + bias[m], 3 zero where H[m][n] <= 0 (ReLU backward). C is FP32 when c_f32, else BF16 (round to nearest even).
 * A shape with few 64 x 64 output tiles and a long K is split over K: the FP32 partials go to a device buffer the
 * context owns (grown on demand; CL_ERR_OOM if that fails) and are added in split order, so results depend on the
 * shape only. Calls of one context that split K must not run concurrently on different streams. */
int cl_mlp_gemm(cl_ctx *ctx, int transA, int transB, int M, int N, int K, const uint16_t *A, int lda,
                const uint16_t *B, int ldb, void *C, int ldc, int c_f32, const float *bias, int epi,
                const uint16_t *H, int ldh, void *stream);
/* out[m] = sum_n X[m][n] (FP32 sums of BF16 values): the bias gradient. */
int cl_mlp_rowsum(cl_ctx *ctx, int M, int N, const uint16_t *X, int ld, float *out, void *stream);
/* Batch-mean loss: BCE-with-logits of out[0][:] against y, plus wreg * w * (out[1][:] - t)^2 / 2; writes
 * d(loss)/d(out) as BF16 (2 x B) and per-group partial sums partial[2 g + {0, 1}] (ceil(B/256) groups). */
int cl_mlp_loss(cl_ctx *ctx, int B, const float *out, const float *y, const float *t, const float *w, float wreg,
                uint16_t *dout, float *partial, void *stream);
/* One Adam step (step >= 1) on FP32 parameters p with moments m, v; writes the BF16 copy pb when non-NULL. */
int cl_mlp_adam(cl_ctx *ctx, int64_t n, float *p, uint16_t *pb, const float *g, float *m, float *v, int step,
                float lr, float b1, float b2, float eps, float wd, void *stream);
/* X = bf16((trace - mean[f]) * istd[f]) for an F x B FP32 trace (non-finite results become 0). */
int cl_mlp_normalize(cl_ctx *ctx, int F, int64_t B, const float *trace, const float *mean, const float *istd,
                     uint16_t *X, void *stream);
/* Per-feature mean and 1/std of an F x B FP32 trace. */
int cl_mlp_feature_stats(cl_ctx *ctx, int F, int64_t B, const float *trace, float *mean, float *istd, void *stream);
/* dst = bf16(src), n values. */
int cl_f32_to_bf16(cl_ctx *ctx, int64_t n, const float *src, uint16_t *dst, void *stream);

/* A dense ReLU network for cl_mlp_train_step: layer l maps sizes[l] inputs to sizes[l + 1] outputs with FP32 master
 * weights W[l] (sizes[l+1] x sizes[l], row-major), their BF16 copy Wb[l], bias b[l] and the Adam moments of both. */
#define CL_MLP_MAX_LAYERS 8
typedef struct cl_mlp_net {
  int32_t nlayers;
  int32_t sizes[CL_MLP_MAX_LAYERS + 1];
  float *W[CL_MLP_MAX_LAYERS];
  uint16_t *Wb[CL_MLP_MAX_LAYERS];
  float *b[CL_MLP_MAX_LAYERS];
  float *mW[CL_MLP_MAX_LAYERS];
  float *vW[CL_MLP_MAX_LAYERS];
  float *mb[CL_MLP_MAX_LAYERS];
  float *vb[CL_MLP_MAX_LAYERS];
} cl_mlp_net;

/* One training step on the B columns of X (sizes[0] x B, row stride ldx): forward (bias + ReLU, the last layer FP32
 * with sizes[nlayers] = 2), the loss of cl_mlp_loss, backward, and Adam (step >= 1, no weight decay) on every weight
 * matrix and bias; loss_sum[0..1] (device, FP64) gains the batch-mean BCE and regression terms. The same kernels in
 * the same order as the separate calls, so the parameters equal theirs bit for bit. The intermediates use a device
 * buffer the context owns (grown on demand; CL_ERR_OOM if that fails): steps of one context must not overlap. */
int cl_mlp_train_step(cl_ctx *ctx, const cl_mlp_net *net, const uint16_t *X, int ldx, int B, const float *y,
                      const float *t, const float *w, float wreg, int step, float lr, float b1, float b2, float eps,
                      double *loss_sum, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* CHAOSLAB_HIP_H */
