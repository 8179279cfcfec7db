// mlp.hip - the surrogate network's kernels: BF16 MFMA GEMM with fused epilogues (forward and both
// backward products), the loss, Adam, and feature normalisation. Activations are stored feature-major
// (a layer's input is a K x B row-major matrix, one column per orbit), so a dense layer is
// Z = W A + b with W (out x in, row-major), and its gradients are dA = W^T dZ and dW = dZ A^T.
//
// GEMM: 256-thread groups own a 64 x 64 tile of C; four wave64s each own 32 x 32 as 2 x 2 tiles of
// v_mfma_f32_16x16x32_bf16 (gfx950): lane l supplies A row (l mod 16) and B column (l mod 16), the eight
// K values (l div 16) * 8 .. +7 of a 32-wide K step, and receives C rows 4 (l div 16) .. +3 of column
// (l mod 16). Tiles of A and B (64 x 64 K values each) are staged through LDS K-contiguous, read from HBM as
// 16-byte vectors of eight values, and the next K tile's vectors are loaded into registers while the matrix
// cores work on the current one; edges and unaligned operands fall back to guarded 2-byte loads and zeros.
// A product with few output tiles and a long K (the weight gradient dW = dZ A^T, K = the batch) is split over K:
// each split writes its FP32 partial tile to the context's workspace and a second kernel adds the partials in
// split order, so the result does not depend on scheduling.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "../../include/chaoslab_hip.h"

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

struct cl_ctx;
extern "C" hipStream_t cl_ctx_stream(cl_ctx* ctx);
extern "C" void* cl_ctx_mlp_workspace(cl_ctx* ctx, size_t bytes);
extern "C" void* cl_ctx_mlp_step_buffer(cl_ctx* ctx, size_t bytes);

namespace {

__device__ inline float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
// round to nearest even (NaN kept quiet)
__device__ inline uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

constexpr int BM = 64, BN = 64, BK = 64, LDK = BK + 8;  // LDS rows padded by 8 bf16 (16 B)

// Eight consecutive bf16 values p[0..7] as one 16-byte vector: a vector load when all eight are in range and
// `vec` (16-byte aligned rows), else guarded scalar loads with zeros past `n` valid values.
__device__ inline uint4 ld8(const uint16_t* __restrict__ p, int n, bool vec) {
  if (vec && n >= 8) return *reinterpret_cast<const uint4*>(p);
  uint32_t w[4] = {0u, 0u, 0u, 0u};
#pragma unroll
  for (int e = 0; e < 8; ++e)
    if (e < n) w[e >> 1] |= (uint32_t)p[e] << ((e & 1) * 16);
  return uint4{w[0], w[1], w[2], w[3]};
}

// One operand's share of a K tile: two vectors per thread. K-contiguous (A untransposed, B transposed): vector
// v = tid + 256 i covers row v / 8, K values (v mod 8) * 8 .. +7. Otherwise (stored K x rows): K value v / 8,
// rows (v mod 8) * 8 .. +7.
template <bool KCONTIG>
__device__ inline void tile_load(uint4 (&r)[2], const uint16_t* __restrict__ X, int ld, int rows, int r0, int k0,
                                 int kend, bool vec, int tid) {
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const int v = tid + 256 * i;
    if (KCONTIG) {
      const int row = r0 + (v >> 3), gk = k0 + (v & 7) * 8;
      const int n = row < rows ? kend - gk : 0;
      r[i] = ld8(X + (int64_t)(row < rows ? row : 0) * ld + gk, n, vec);
    } else {
      const int gk = k0 + (v >> 3), row = r0 + (v & 7) * 8;
      const int n = gk < kend ? rows - row : 0;
      r[i] = ld8(X + (int64_t)(gk < kend ? gk : k0) * ld + row, n, vec);
    }
  }
}

template <bool KCONTIG>
__device__ inline void tile_store(uint16_t* __restrict__ S, const uint4 (&r)[2], int tid) {
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const int v = tid + 256 * i;
    if (KCONTIG) {
      *reinterpret_cast<uint4*>(S + (v >> 3) * LDK + (v & 7) * 8) = r[i];
    } else {
      const int kk = v >> 3, row = (v & 7) * 8;
      const uint32_t w[4] = {r[i].x, r[i].y, r[i].z, r[i].w};
#pragma unroll
      for (int e = 0; e < 8; ++e) S[(row + e) * LDK + kk] = (uint16_t)(w[e >> 1] >> ((e & 1) * 16));
    }
  }
}

// EPI: 0 plain, 1 bias[m] then ReLU, 2 bias[m], 3 multiply by (H(m, n) > 0) (ReLU backward).
template <int EPI, bool OUT_F32>
__device__ inline void epi_store(float x, int gm, int gn, void* __restrict__ C, int ldc, const float* __restrict__ bias,
                                 const uint16_t* __restrict__ H, int ldh) {
  if (EPI == 1 || EPI == 2) x += bias[gm];
  if (EPI == 1) x = x > 0.f ? x : 0.f;
  if (EPI == 3) x = bf2f(H[(int64_t)gm * ldh + gn]) > 0.f ? x : 0.f;
  if (OUT_F32)
    reinterpret_cast<float*>(C)[(int64_t)gm * ldc + gn] = x;
  else
    reinterpret_cast<uint16_t*>(C)[(int64_t)gm * ldc + gn] = f2bf(x);
}

// C(m, n) = sum_k A(m, k) B(k, n); A(m, k) = TA ? A[k lda + m] : A[m lda + k]; B(k, n) = TB ? B[n ldb + k] : B[k ldb + n].
// Split z = blockIdx.z sums k in [z kc, min(K, (z + 1) kc)); with SPLIT its raw FP32 tile goes to
// ws[z M N + m N + n], else the epilogue writes C.
template <bool TA, bool TB, int EPI, bool OUT_F32, bool SPLIT>
__global__ void __launch_bounds__(256) k_gemm(int M, int N, int K, int kc, const uint16_t* __restrict__ A, int lda,
                                              int vec_a, const uint16_t* __restrict__ B, int ldb, int vec_b,
                                              void* __restrict__ C, int ldc, const float* __restrict__ bias,
                                              const uint16_t* __restrict__ H, int ldh, float* __restrict__ ws) {
  __shared__ __attribute__((aligned(16))) uint16_t As[BM * LDK];
  __shared__ __attribute__((aligned(16))) uint16_t Bs[BN * LDK];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int bm = blockIdx.y * BM, bn = blockIdx.x * BN;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  const int kbeg = blockIdx.z * kc, kend = min(K, kbeg + kc);
  floatx4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};

  uint4 ra[2], rb[2];
  tile_load<!TA>(ra, A, lda, M, bm, kbeg, kend, vec_a != 0, tid);
  tile_load<TB>(rb, B, ldb, N, bn, kbeg, kend, vec_b != 0, tid);
  for (int k0 = kbeg; k0 < kend; k0 += BK) {
    tile_store<!TA>(As, ra, tid);
    tile_store<TB>(Bs, rb, tid);
    __syncthreads();
    if (k0 + BK < kend) {  // the next tile's loads stay in flight over this tile's products
      tile_load<!TA>(ra, A, lda, M, bm, k0 + BK, kend, vec_a != 0, tid);
      tile_load<TB>(rb, B, ldb, N, bn, k0 + BK, kend, vec_b != 0, tid);
    }
#pragma unroll
    for (int ks = 0; ks < BK; ks += 32) {
      const int kq = ks + (lane >> 4) * 8;
      bf16x8 af[2], bfr[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        af[i] = *reinterpret_cast<const bf16x8*>(As + (wm + i * 16 + (lane & 15)) * LDK + kq);
        bfr[i] = *reinterpret_cast<const bf16x8*>(Bs + (wn + i * 16 + (lane & 15)) * LDK + kq);
      }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i], bfr[j], acc[i][j], 0, 0, 0);
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int gm = bm + wm + i * 16 + (lane >> 4) * 4 + v;
        const int gn = bn + wn + j * 16 + (lane & 15);
        if (gm >= M || gn >= N) continue;
        if (SPLIT)
          ws[((int64_t)blockIdx.z * M + gm) * N + gn] = acc[i][j][v];
        else
          epi_store<EPI, OUT_F32>(acc[i][j][v], gm, gn, C, ldc, bias, H, ldh);
      }
}

// C(m, n) = epilogue(sum over splits z = 0, 1, ... of ws[z M N + m N + n]), added in split order.
template <int EPI, bool OUT_F32>
__global__ void __launch_bounds__(256) k_split_sum(int M, int N, int S, const float* __restrict__ ws,
                                                   void* __restrict__ C, int ldc, const float* __restrict__ bias,
                                                   const uint16_t* __restrict__ H, int ldh) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x, mn = (int64_t)M * N;
  if (i >= mn) return;
  float x = ws[i];
  for (int z = 1; z < S; ++z) x += ws[z * mn + i];
  epi_store<EPI, OUT_F32>(x, (int)(i / N), (int)(i % N), C, ldc, bias, H, ldh);
}

// Row sums of an M x N bf16 matrix in FP32 (the bias gradient); one 256-thread group per row.
__global__ void __launch_bounds__(256) k_rowsum(int M, int N, const uint16_t* __restrict__ X, int ld,
                                                float* __restrict__ out) {
  __shared__ float part[256];
  int r = blockIdx.x;
  float s = 0.f;
  for (int n = threadIdx.x; n < N; n += 256) s += bf2f(X[(int64_t)r * ld + n]);
  part[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[r] = part[0];
}

// Loss per orbit b: BCE with logits on out[0][b] against y[b] in {0, 1}, plus wreg * w[b] * (out[1][b] - t[b])^2 / 2;
// mean over the batch. dout = d(loss)/d(out) in bf16 (the next GEMM's input); partial sums of the two terms per
// group go to partial[2 * group + {0, 1}].
__global__ void __launch_bounds__(256) k_loss(int Bn, const float* __restrict__ out, const float* __restrict__ y,
                                              const float* __restrict__ t, const float* __restrict__ w, float wreg,
                                              uint16_t* __restrict__ dout, float* __restrict__ partial) {
  __shared__ float pb[256], pr[256];
  int b = blockIdx.x * 256 + threadIdx.x;
  float lb = 0.f, lr = 0.f;
  if (b < Bn) {
    float z = out[b], yy = y[b];
    float sp = z > 0.f ? z + log1pf(expf(-z)) : log1pf(expf(z));  // softplus(z)
    lb = sp - yy * z;
    float sig = 1.f / (1.f + expf(-z));
    float dz = (sig - yy) / (float)Bn;
    float e = out[Bn + b] - t[b];
    lr = 0.5f * wreg * w[b] * e * e;
    float de = wreg * w[b] * e / (float)Bn;
    dout[b] = f2bf(dz);
    dout[Bn + b] = f2bf(de);
  }
  pb[threadIdx.x] = lb;
  pr[threadIdx.x] = lr;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      pb[threadIdx.x] += pb[threadIdx.x + s];
      pr[threadIdx.x] += pr[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    partial[2 * blockIdx.x] = pb[0];
    partial[2 * blockIdx.x + 1] = pr[0];
  }
}

// Adam (Kingma and Ba 2015) on FP32 master weights; writes the BF16 copy the GEMMs read.
__global__ void k_adam(int64_t n, float* __restrict__ p, uint16_t* __restrict__ pb, const float* __restrict__ g,
                       float* __restrict__ m, float* __restrict__ v, float lr, float b1, float b2, float eps, float c1,
                       float c2, float wd) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float gi = g[i] + wd * p[i];
  float mi = b1 * m[i] + (1.f - b1) * gi;
  float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  float mh = mi / c1, vh = vi / c2;
  float pi = p[i] - lr * mh / (sqrtf(vh) + eps);
  p[i] = pi;
  if (pb) pb[i] = f2bf(pi);
}

// X[f][b] = bf16((trace[f][b] - mean[f]) * istd[f]), trace F x B FP32.
__global__ void k_normalize(int F, int64_t Bn, const float* __restrict__ tr, const float* __restrict__ mean,
                            const float* __restrict__ istd, uint16_t* __restrict__ X) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)F * Bn) return;
  int f = (int)(i / Bn);
  float x = (tr[i] - mean[f]) * istd[f];
  X[i] = f2bf(isfinite(x) ? x : 0.f);
}

// Per-feature mean and 1/std over B columns (FP64 sums), one group per feature.
__global__ void __launch_bounds__(256) k_feature_stats(int F, int64_t Bn, const float* __restrict__ tr,
                                                       float* __restrict__ mean, float* __restrict__ istd) {
  __shared__ double s1[256], s2[256];
  int f = blockIdx.x;
  double a = 0.0, b = 0.0;
  for (int64_t n = threadIdx.x; n < Bn; n += 256) {
    double x = tr[(int64_t)f * Bn + n];
    a += x;
    b += x * x;
  }
  s1[threadIdx.x] = a;
  s2[threadIdx.x] = b;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      s1[threadIdx.x] += s1[threadIdx.x + w];
      s2[threadIdx.x] += s2[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    double mu = s1[0] / (double)Bn, var = s2[0] / (double)Bn - mu * mu;
    mean[f] = (float)mu;
    istd[f] = (float)(1.0 / sqrt(var > 1e-12 ? var : 1e-12));
  }
}

// loss_sum[c] += (sum over the groups of partial[2 g + c], in double) / B: the step's batch-mean loss terms.
__global__ void k_loss_accum(int ng, int Bn, const float* __restrict__ partial, double* __restrict__ loss_sum) {
  if (threadIdx.x >= 2) return;
  const int c = threadIdx.x;
  double s = 0.0;
  for (int g = 0; g < ng; ++g) s += (double)partial[2 * g + c];
  loss_sum[c] += s / (double)Bn;
}

__global__ void k_f32_to_bf16(int64_t n, const float* __restrict__ s, uint16_t* __restrict__ d) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) d[i] = f2bf(s[i]);
}

// Splits over K for an M x N x K product with fewer output tiles than half the compute units of a 256-CU device:
// enough groups for two per compute unit, each split at least 256 K values, split lengths a multiple of BK.
// Depends on the shape only.
void split_plan(int M, int N, int K, int* S, int* kc) {
  const int64_t tiles = (int64_t)((N + BN - 1) / BN) * ((M + BM - 1) / BM);
  int64_t s = tiles >= 128 ? 1 : (512 + tiles - 1) / tiles;
  if (s > K / 256) s = K / 256;
  if (s < 1) s = 1;
  int64_t c = (K + s - 1) / s;
  c = (c + BK - 1) / BK * BK;
  *kc = (int)c;
  *S = (int)((K + c - 1) / c);
}

bool vec_ok(const uint16_t* p, int ld) { return ld % 8 == 0 && ((uintptr_t)p & 15) == 0; }

template <bool TA, bool TB, int EPI, bool F32>
void gemm_launch(int M, int N, int K, int S, int kc, float* ws, const uint16_t* A, int lda, const uint16_t* B, int ldb,
                 void* C, int ldc, const float* bias, const uint16_t* H, int ldh, hipStream_t st) {
  dim3 grid((N + BN - 1) / BN, (M + BM - 1) / BM, S);
  const int va = vec_ok(A, lda), vb = vec_ok(B, ldb);
  if (S == 1) {
    hipLaunchKernelGGL((k_gemm<TA, TB, EPI, F32, false>), grid, dim3(256), 0, st, M, N, K, kc, A, lda, va, B, ldb, vb,
                       C, ldc, bias, H, ldh, (float*)nullptr);
    return;
  }
  hipLaunchKernelGGL((k_gemm<TA, TB, 0, true, true>), grid, dim3(256), 0, st, M, N, K, kc, A, lda, va, B, ldb, vb,
                     (void*)nullptr, 0, (const float*)nullptr, (const uint16_t*)nullptr, 0, ws);
  const int64_t mn = (int64_t)M * N;
  hipLaunchKernelGGL((k_split_sum<EPI, F32>), dim3((unsigned)((mn + 255) / 256)), dim3(256), 0, st, M, N, S, ws, C,
                     ldc, bias, H, ldh);
}

int rc_of(hipError_t e) { return e == hipSuccess ? CL_OK : (e == hipErrorOutOfMemory ? CL_ERR_OOM : CL_ERR_DEVICE_LOST); }

}  // namespace

extern "C" int cl_mlp_gemm(cl_ctx* ctx, int transA, int transB, int M, int N, int K, const uint16_t* A, int lda,
                           const uint16_t* B, int ldb, void* C, int ldc, int c_f32, const float* bias, int epi,
                           const uint16_t* H, int ldh, void* stream) {
  if (!ctx || M <= 0 || N <= 0 || K <= 0 || !A || !B || !C) return CL_ERR_ARGUMENT;
  if ((epi == 1 || epi == 2) && !bias) return CL_ERR_ARGUMENT;
  if (epi == 3 && !H) return CL_ERR_ARGUMENT;
  hipStream_t st = stream ? (hipStream_t)stream : cl_ctx_stream(ctx);
  int S = 1, kc = K;
  split_plan(M, N, K, &S, &kc);
  float* ws = nullptr;
  if (S > 1) {
    ws = (float*)cl_ctx_mlp_workspace(ctx, (size_t)S * M * N * sizeof(float));
    if (!ws) return CL_ERR_OOM;
  }
#define DISPATCH(TA, TB)                                                                                              \
  if (transA == TA && transB == TB) {                                                                                 \
    if (c_f32) {                                                                                                      \
      if (epi == 0) gemm_launch<TA, TB, 0, true>(M, N, K, S, kc, ws, A, lda, B, ldb, C, ldc, bias, H, ldh, st);       \
      else if (epi == 1) gemm_launch<TA, TB, 1, true>(M, N, K, S, kc, ws, A, lda, B, ldb, C, ldc, bias, H, ldh, st);  \
      else if (epi == 2) gemm_launch<TA, TB, 2, true>(M, N, K, S, kc, ws, A, lda, B, ldb, C, ldc, bias, H, ldh, st);  \
      else gemm_launch<TA, TB, 3, true>(M, N, K, S, kc, ws, A, lda, B, ldb, C, ldc, bias, H, ldh, st);                \
    } else {                                                                                                          \
      if (epi == 0) gemm_launch<TA, TB, 0, false>(M, N, K, S, kc, ws, A, lda, B, ldb, C, ldc, bias, H, ldh, st);      \
      else if (epi == 1) gemm_launch<TA, TB, 1, false>(M, N, K, S, kc, ws, A, lda, B, ldb, C, ldc, bias, H, ldh, st); \
      else if (epi == 2) gemm_launch<TA, TB, 2, false>(M, N, K, S, kc, ws, A, lda, B, ldb, C, ldc, bias, H, ldh, st); \
      else gemm_launch<TA, TB, 3, false>(M, N, K, S, kc, ws, A, lda, B, ldb, C, ldc, bias, H, ldh, st);               \
    }                                                                                                                 \
    return rc_of(hipGetLastError());                                                                                  \
  }
  DISPATCH(0, 0)
  DISPATCH(1, 0)
  DISPATCH(0, 1)
  DISPATCH(1, 1)
#undef DISPATCH
  return CL_ERR_ARGUMENT;
}

extern "C" int cl_mlp_rowsum(cl_ctx* ctx, int M, int N, const uint16_t* X, int ld, float* out, void* stream) {
  if (!ctx || M <= 0 || N <= 0 || !X || !out) return CL_ERR_ARGUMENT;
  hipStream_t st = stream ? (hipStream_t)stream : cl_ctx_stream(ctx);
  hipLaunchKernelGGL(k_rowsum, dim3(M), dim3(256), 0, st, M, N, X, ld, out);
  return rc_of(hipGetLastError());
}

extern "C" int cl_mlp_loss(cl_ctx* ctx, int Bn, const float* out, const float* y, const float* t, const float* w,
                           float wreg, uint16_t* dout, float* partial, void* stream) {
  if (!ctx || Bn <= 0 || !out || !y || !t || !w || !dout || !partial) return CL_ERR_ARGUMENT;
  hipStream_t st = stream ? (hipStream_t)stream : cl_ctx_stream(ctx);
  hipLaunchKernelGGL(k_loss, dim3((Bn + 255) / 256), dim3(256), 0, st, Bn, out, y, t, w, wreg, dout, partial);
  return rc_of(hipGetLastError());
}

extern "C" int cl_mlp_adam(cl_ctx* ctx, int64_t n, float* p, uint16_t* pb, const float* g, float* m, float* v,
                           int step, float lr, float b1, float b2, float eps, float wd, void* stream) {
  if (!ctx || n <= 0 || !p || !g || !m || !v || step < 1) return CL_ERR_ARGUMENT;
  hipStream_t st = stream ? (hipStream_t)stream : cl_ctx_stream(ctx);
  float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
  hipLaunchKernelGGL(k_adam, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, p, pb, g, m, v, lr, b1, b2, eps,
                     c1, c2, wd);
  return rc_of(hipGetLastError());
}

extern "C" int cl_mlp_normalize(cl_ctx* ctx, int F, int64_t Bn, const float* trace, const float* mean,
                                const float* istd, uint16_t* X, void* stream) {
  if (!ctx || F <= 0 || Bn <= 0 || !trace || !mean || !istd || !X) return CL_ERR_ARGUMENT;
  hipStream_t st = stream ? (hipStream_t)stream : cl_ctx_stream(ctx);
  int64_t n = (int64_t)F * Bn;
  hipLaunchKernelGGL(k_normalize, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, F, Bn, trace, mean, istd, X);
  return rc_of(hipGetLastError());
}

extern "C" int cl_mlp_feature_stats(cl_ctx* ctx, int F, int64_t Bn, const float* trace, float* mean, float* istd,
                                    void* stream) {
  if (!ctx || F <= 0 || Bn <= 0 || !trace || !mean || !istd) return CL_ERR_ARGUMENT;
  hipStream_t st = stream ? (hipStream_t)stream : cl_ctx_stream(ctx);
  hipLaunchKernelGGL(k_feature_stats, dim3(F), dim3(256), 0, st, F, Bn, trace, mean, istd);
  return rc_of(hipGetLastError());
}

extern "C" int cl_f32_to_bf16(cl_ctx* ctx, int64_t n, const float* src, uint16_t* dst, void* stream) {
  if (!ctx || n <= 0 || !src || !dst) return CL_ERR_ARGUMENT;
  hipStream_t st = stream ? (hipStream_t)stream : cl_ctx_stream(ctx);
  hipLaunchKernelGGL(k_f32_to_bf16, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, src, dst);
  return rc_of(hipGetLastError());
}

// One training step in the order of the host-driven calls (cl_mlp_gemm forward with bias and ReLU, the last layer in
// FP32; cl_mlp_loss; per layer from the last: the weight gradient, the bias gradient, the masked input gradient;
// cl_mlp_adam on every weight matrix then its bias), so the parameters after it are those calls' bit for bit. The
// intermediates live in a buffer the context owns (grown on demand).
extern "C" int cl_mlp_train_step(cl_ctx* ctx, const cl_mlp_net* net, const uint16_t* X, int ldx, int Bn,
                                 const float* y, const float* t, const float* w, float wreg, int step, float lr,
                                 float b1, float b2, float eps, double* loss_sum, void* stream) {
  if (!ctx || !net || !X || !y || !t || !w || !loss_sum || Bn <= 0 || ldx < Bn || step < 1) return CL_ERR_ARGUMENT;
  const int L = net->nlayers;
  if (L < 1 || L > CL_MLP_MAX_LAYERS) return CL_ERR_ARGUMENT;
  int maxin = 0;
  for (int l = 0; l <= L; ++l)
    if (net->sizes[l] <= 0) return CL_ERR_ARGUMENT;
  for (int l = 0; l < L; ++l) {
    if (!net->W[l] || !net->Wb[l] || !net->b[l] || !net->mW[l] || !net->vW[l] || !net->mb[l] || !net->vb[l])
      return CL_ERR_ARGUMENT;
    if (l > 0 && net->sizes[l] > maxin) maxin = net->sizes[l];
  }
  if (net->sizes[L] != 2) return CL_ERR_ARGUMENT;  // the loss reads the logit and the predicted exponent
  hipStream_t st = stream ? (hipStream_t)stream : cl_ctx_stream(ctx);
  // buffer plan: activations, output, output gradient, loss partials, weight and bias gradients, two input gradients
  const int ng = (Bn + 255) / 256;
  size_t off = 0;
  auto take = [&off](size_t bytes) {
    const size_t o = off;
    off += (bytes + 255) & ~(size_t)255;
    return o;
  };
  size_t o_act[CL_MLP_MAX_LAYERS], o_gw[CL_MLP_MAX_LAYERS], o_gb[CL_MLP_MAX_LAYERS];
  for (int l = 0; l + 1 < L; ++l) o_act[l] = take((size_t)net->sizes[l + 1] * Bn * 2);
  const size_t o_out = take((size_t)2 * Bn * 4), o_dout = take((size_t)2 * Bn * 2), o_part = take((size_t)2 * ng * 4);
  for (int l = 0; l < L; ++l) {
    o_gw[l] = take((size_t)net->sizes[l + 1] * net->sizes[l] * 4);
    o_gb[l] = take((size_t)net->sizes[l + 1] * 4);
  }
  const size_t o_da0 = take((size_t)maxin * Bn * 2), o_da1 = take((size_t)maxin * Bn * 2);
  char* base = (char*)cl_ctx_mlp_step_buffer(ctx, off);
  if (!base) return CL_ERR_OOM;
  auto bf = [base](size_t o) { return (uint16_t*)(base + o); };
  auto f32 = [base](size_t o) { return (float*)(base + o); };
  int rc = CL_OK;
  // forward
  const uint16_t* A = X;
  int lda = ldx;
  for (int l = 0; l < L && !rc; ++l) {
    const int M = net->sizes[l + 1], K = net->sizes[l];
    if (l == L - 1) {
      rc = cl_mlp_gemm(ctx, 0, 0, M, Bn, K, net->Wb[l], K, A, lda, f32(o_out), Bn, 1, net->b[l], 2, nullptr, 0, st);
    } else {
      rc = cl_mlp_gemm(ctx, 0, 0, M, Bn, K, net->Wb[l], K, A, lda, bf(o_act[l]), Bn, 0, net->b[l], 1, nullptr, 0, st);
      A = bf(o_act[l]);
      lda = Bn;
    }
  }
  if (!rc) rc = cl_mlp_loss(ctx, Bn, f32(o_out), y, t, w, wreg, bf(o_dout), f32(o_part), st);
  if (!rc) {
    hipLaunchKernelGGL(k_loss_accum, dim3(1), dim3(64), 0, st, ng, Bn, f32(o_part), loss_sum);
    rc = rc_of(hipGetLastError());
  }
  // backward
  const uint16_t* dZ = bf(o_dout);
  for (int l = L - 1; l >= 0 && !rc; --l) {
    const int M = net->sizes[l + 1], K = net->sizes[l];
    const uint16_t* Ap = l == 0 ? X : bf(o_act[l - 1]);
    const int ldp = l == 0 ? ldx : Bn;
    rc = cl_mlp_gemm(ctx, 0, 1, M, K, Bn, dZ, Bn, Ap, ldp, f32(o_gw[l]), K, 1, nullptr, 0, nullptr, 0, st);
    if (!rc) rc = cl_mlp_rowsum(ctx, M, Bn, dZ, Bn, f32(o_gb[l]), st);
    if (!rc && l > 0) {
      uint16_t* dA = bf(((L - 1 - l) & 1) ? o_da1 : o_da0);  // never the buffer dZ is read from
      rc = cl_mlp_gemm(ctx, 1, 0, K, Bn, M, net->Wb[l], K, dZ, Bn, dA, Bn, 0, nullptr, 3, bf(o_act[l - 1]), Bn, st);
      dZ = dA;
    }
  }
  // Adam, weights then bias of each layer
  for (int l = 0; l < L && !rc; ++l) {
    const int64_t nw = (int64_t)net->sizes[l + 1] * net->sizes[l];
    rc = cl_mlp_adam(ctx, nw, net->W[l], net->Wb[l], f32(o_gw[l]), net->mW[l], net->vW[l], step, lr, b1, b2, eps, 0.f,
                     st);
    if (!rc)
      rc = cl_mlp_adam(ctx, net->sizes[l + 1], net->b[l], nullptr, f32(o_gb[l]), net->mb[l], net->vb[l], step, lr, b1,
                       b2, eps, 0.f, st);
  }
  return rc;
}
