// mlp.hip - the surrogate network's kernels: BF16 MFMA GEMM with fused epilogues (forward and both
// backward products), the loss, Adam, and feature normalisation. Activations are stored feature-major
// (a layer's input is a K x B row-major matrix, one column per orbit), so a dense layer is
// Z = W A + b with W (out x in, row-major), and its gradients are dA = W^T dZ and dW = dZ A^T.
//
// GEMM: 256-thread groups own a 64 x 64 tile of C; four wave64s each own 32 x 32 as 2 x 2 tiles of
// v_mfma_f32_16x16x32_bf16 (gfx950): lane l supplies A row (l mod 16) and B column (l mod 16), the eight
// K values (l div 16) * 8 .. +7 of a 32-wide K step, and receives C rows 4 (l div 16) .. +3 of column
// (l mod 16). Tiles of A and B are staged through LDS K-contiguous; edges are zero-filled.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "../../include/chaoslab_hip.h"

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

struct cl_ctx;
extern "C" hipStream_t cl_ctx_stream(cl_ctx* ctx);

namespace {

__device__ inline float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
// round to nearest even (NaN kept quiet)
__device__ inline uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

constexpr int BM = 64, BN = 64, BK = 32, LDK = BK + 8;  // LDS rows padded by 8 bf16 (16 B)

// C(m, n) = sum_k A(m, k) B(k, n); A(m, k) = TA ? A[k lda + m] : A[m lda + k]; B(k, n) = TB ? B[n ldb + k] : B[k ldb + n].
// EPI: 0 plain, 1 bias[m] then ReLU, 2 bias[m], 3 multiply by (H(m, n) > 0) (ReLU backward).
template <bool TA, bool TB, int EPI, bool OUT_F32>
__global__ void __launch_bounds__(256) k_gemm(int M, int N, int K, const uint16_t* __restrict__ A, int lda,
                                              const uint16_t* __restrict__ B, int ldb, void* __restrict__ C, int ldc,
                                              const float* __restrict__ bias, const uint16_t* __restrict__ H, int ldh) {
  __shared__ __attribute__((aligned(16))) uint16_t As[BM * LDK];
  __shared__ __attribute__((aligned(16))) uint16_t Bs[BN * LDK];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int bm = blockIdx.y * BM, bn = blockIdx.x * BN;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  floatx4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};

  for (int k0 = 0; k0 < K; k0 += BK) {
    // stage A: 64 x 32 values, 8 per thread
    {
      if (!TA) {  // row-major M x K: a thread reads 8 consecutive k of one row
        int r = tid >> 2, kk = (tid & 3) * 8;
        int gm = bm + r;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          int gk = k0 + kk + e;
          As[r * LDK + kk + e] = (gm < M && gk < K) ? A[(int64_t)gm * lda + gk] : (uint16_t)0;
        }
      } else {  // stored K x M: a thread reads 8 consecutive m of one k
        int kk = tid >> 3, r = (tid & 7) * 8;
        int gk = k0 + kk;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          int gm = bm + r + e;
          As[(r + e) * LDK + kk] = (gm < M && gk < K) ? A[(int64_t)gk * lda + gm] : (uint16_t)0;
        }
      }
      if (!TB) {  // row-major K x N: 8 consecutive n of one k
        int kk = tid >> 3, c = (tid & 7) * 8;
        int gk = k0 + kk;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          int gn = bn + c + e;
          Bs[(c + e) * LDK + kk] = (gn < N && gk < K) ? B[(int64_t)gk * ldb + gn] : (uint16_t)0;
        }
      } else {  // stored N x K: 8 consecutive k of one n
        int c = tid >> 2, kk = (tid & 3) * 8;
        int gn = bn + c;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          int gk = k0 + kk + e;
          Bs[c * LDK + kk + e] = (gn < N && gk < K) ? B[(int64_t)gn * ldb + gk] : (uint16_t)0;
        }
      }
    }
    __syncthreads();
    const int kq = (lane >> 4) * 8;
    bf16x8 af[2], bfr[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const uint16_t* pa = As + (wm + i * 16 + (lane & 15)) * LDK + kq;
      const uint16_t* pb = Bs + (wn + i * 16 + (lane & 15)) * LDK + kq;
      af[i] = *reinterpret_cast<const bf16x8*>(pa);
      bfr[i] = *reinterpret_cast<const bf16x8*>(pb);
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i], bfr[j], acc[i][j], 0, 0, 0);
    __syncthreads();
  }
  // epilogue
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        int gm = bm + wm + i * 16 + (lane >> 4) * 4 + v;
        int gn = bn + wn + j * 16 + (lane & 15);
        if (gm >= M || gn >= N) continue;
        float x = acc[i][j][v];
        if (EPI == 1 || EPI == 2) x += bias[gm];
        if (EPI == 1) x = x > 0.f ? x : 0.f;
        if (EPI == 3) x = bf2f(H[(int64_t)gm * ldh + gn]) > 0.f ? x : 0.f;
        if (OUT_F32)
          reinterpret_cast<float*>(C)[(int64_t)gm * ldc + gn] = x;
        else
          reinterpret_cast<uint16_t*>(C)[(int64_t)gm * ldc + gn] = f2bf(x);
      }
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

__global__ void k_f32_to_bf16(int64_t n, const float* __restrict__ s, uint16_t* __restrict__ d) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) d[i] = f2bf(s[i]);
}

template <bool TA, bool TB, int EPI, bool F32>
void gemm_launch(int M, int N, int K, const uint16_t* A, int lda, const uint16_t* B, int ldb, void* C, int ldc,
                 const float* bias, const uint16_t* H, int ldh, hipStream_t st) {
  dim3 grid((N + BN - 1) / BN, (M + BM - 1) / BM);
  hipLaunchKernelGGL((k_gemm<TA, TB, EPI, F32>), grid, dim3(256), 0, st, M, N, K, A, lda, B, ldb, C, ldc, bias, H, ldh);
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
#define DISPATCH(TA, TB)                                                                                         \
  if (transA == TA && transB == TB) {                                                                            \
    if (c_f32) {                                                                                                 \
      if (epi == 0) gemm_launch<TA, TB, 0, true>(M, N, K, A, lda, B, ldb, C, ldc, bias, H, ldh, st);             \
      else if (epi == 1) gemm_launch<TA, TB, 1, true>(M, N, K, A, lda, B, ldb, C, ldc, bias, H, ldh, st);        \
      else if (epi == 2) gemm_launch<TA, TB, 2, true>(M, N, K, A, lda, B, ldb, C, ldc, bias, H, ldh, st);        \
      else gemm_launch<TA, TB, 3, true>(M, N, K, A, lda, B, ldb, C, ldc, bias, H, ldh, st);                      \
    } else {                                                                                                     \
      if (epi == 0) gemm_launch<TA, TB, 0, false>(M, N, K, A, lda, B, ldb, C, ldc, bias, H, ldh, st);            \
      else if (epi == 1) gemm_launch<TA, TB, 1, false>(M, N, K, A, lda, B, ldb, C, ldc, bias, H, ldh, st);       \
      else if (epi == 2) gemm_launch<TA, TB, 2, false>(M, N, K, A, lda, B, ldb, C, ldc, bias, H, ldh, st);       \
      else gemm_launch<TA, TB, 3, false>(M, N, K, A, lda, B, ldb, C, ldc, bias, H, ldh, st);                     \
    }                                                                                                            \
    return rc_of(hipGetLastError());                                                                             \
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
