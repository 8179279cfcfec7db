// pipe_kernel.hpp - a wave-specialised pipeline for small map ensembles whose chain is one QR spectrum.
//
// A small ensemble (a few thousand orbits) fills a few dozen waves on a 1024-SIMD device, so the fused
// kernel runs at the issue rate of one wave per SIMD: its time is the length of the per-step instruction
// stream (sin/cos of the point, the frame product, the factorisation, its checks and two logs), step after
// step. The work of one orbit-step splits into four parts with one-way data flow:
//
//   A  the orbit: sin/cos of x_n, the step, the parameter c_n = K cos x_n of the tangent (standard_map.jl:30-34)
//   B  the frame: W <- M(c) W in the fused_product order, then the factorisation W = QR (carries W = Q)
//   C  the checks of the factorisation: window_Gw conditioning and orth_loss (qr_spectrum.jl:54-73),
//      and the running maxima G_max, orth_max
//   D  the exponent sums: add(sums[i], log R_ii)
//   E  (chains QRSpectrum + RunMonitor, the recorder's) the monitor's running maxima of the one-step tangent's
//      det and symplecticity residuals and of max |M| at each handoff point, from c at that point
//
// Each part runs on its own waves of a 384-lane group (64 orbits, one per lane of every wave): A and B one
// wave each; C two waves, for the even and the odd handoffs (a maximum does not depend on the order of its
// terms, and both start from the same value); D one wave per column (each column's compensated sum stays in
// step order). One orbit-step occupies six SIMDs, and B, whose factorisation carries the frame from step to
// step, sets the pace. The parts hand data forward through double-buffered LDS rings: A writes c for a chunk
// of steps while B consumes the previous chunk's c and writes R and Q, which C and D read one chunk later.
// One barrier per chunk. The arithmetic of every part is the fused kernel's, in the same order, so the
// results are bitwise the same.
//
// A failure (non-finite point or frame at a handoff, a failed factorisation or check) is not handled here:
// the group marks the orbit, writes nothing back for it, and the caller reruns the marked orbits with the
// fused kernel from the unchanged inputs, which reproduces the reference's stop rule exactly
// (orbit.jl:48-58; qr_spectrum.jl:54-96).
#pragma once
#include "registry.hpp"

namespace cl {

constexpr int PIPE_LANES = 64;   // orbits per group
constexpr int PIPE_WAVES = 6;    // A, B, C even, C odd, D column 1, D column 2 (and E for a monitored chain)
constexpr int PIPE_STEPS = 16;   // map steps per chunk at most (tau <= 16)
constexpr int PIPE_HAND = 8;     // handoffs per chunk at most

// handoffs per chunk for a given tau: chunks of about 8 steps, at least one handoff
CL_HD int pipe_handoffs(int tau) { return tau <= PIPE_HAND ? PIPE_HAND / tau : 1; }

// MON: the chain is QRSpectrum then RunMonitor (its state words follow the QR state's)
template <class Sys, int ALG, int K, bool MON>
__global__ void __launch_bounds__((MON ? 7 : 6) * PIPE_LANES)
pipe_kernel(KArgs a, Sys sys, QRSpectrum<ALG, 2, K> qr, int H, uint8_t* rerun) {
  constexpr int D = 2;
  constexpr int NW = MON ? 7 : 6;
  constexpr int WMON = (int)(sizeof(QRState<K>) / 8);  // the monitor's first state word
  typedef QRState<K> St;
  __shared__ double ringC[2][PIPE_STEPS][PIPE_LANES];
  __shared__ double ringR[2][PIPE_HAND][K * K][PIPE_LANES];
  __shared__ double ringQ[2][PIPE_HAND][D * K][PIPE_LANES];
  __shared__ double ringM[MON ? 2 : 1][MON ? PIPE_HAND : 1][PIPE_LANES];  // c at each handoff point
  __shared__ int bad_by[NW][PIPE_LANES];
  __shared__ double outG[2][PIPE_LANES], outO[2][PIPE_LANES];
  __shared__ CSum outS[2][PIPE_LANES];  // [1]: column 2's sum, from the wave that formed it
  __shared__ int64_t outCount[PIPE_LANES];

  // Waves w and w + 4 of a group share a SIMD (round-robin placement over the CU's four SIMDs). Measured:
  // the step time is set by B whatever it shares with (placements differ by under 3%); the two C waves get a
  // SIMD each, A shares with D of column 2 and B with D of column 1 (E, when present, with C even).
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int role = wave < 4 ? wave : (wave == 4 ? 5 : (wave == 5 ? 4 : 6));
  const int64_t i = (int64_t)blockIdx.x * PIPE_LANES + lane;
  const bool in_range = i < a.m;
  const bool live = in_range && (a.n0 == 0 || a.code[i] == STATUS_OK);
  const int tau = a.tau;

  St st;
  qr.init(st);
  if (live && a.n0 != 0) state_load(st, a.st, a.m, i, 0);
  RunMonState mst;
  const RunMon<2> mon = {};
  mon.init(mst);
  if (MON && role == 6 && live && a.n0 != 0) state_load(mst, a.st, a.m, i, WMON);

  double x[Sys::S];
  typename Sys::Aux aux;
  Frame<D, K> W;
#pragma unroll
  for (int j = 0; j < Sys::S; ++j) x[j] = 0.0;
#pragma unroll
  for (int e = 0; e < D * K; ++e) W.w[e] = 0.0;
  if (role == 0 && in_range) {
#pragma unroll
    for (int j = 0; j < Sys::S; ++j) x[j] = a.x[(int64_t)j * a.m + i];
  }
  if (role == 1 && in_range) {
#pragma unroll
    for (int e = 0; e < D * K; ++e) W.w[e] = a.W0[(int64_t)e * a.m + i];
  }
  if (role == 0) sys.aux(x, aux);

  const int64_t nh = a.steps / tau;
  const int64_t nc = (nh + H - 1) / H;
  bool bad = false;
  for (int64_t p = 0; p < nc + 2; ++p) {
    if (role == 0) {
      if (p < nc) {
        const int slot = (int)(p & 1);
        const int64_t h0 = p * H, h1 = (h0 + H < nh) ? h0 + H : nh;
        for (int64_t h = h0; h < h1; ++h) {
          for (int t = 0; t < tau; ++t) {
            ringC[slot][(int)(h - h0) * tau + t][lane] = sys.tan_c(aux);
            sys.forward(aux, x);
            sys.aux(x, aux);
          }
          bool xf = true;
#pragma unroll
          for (int j = 0; j < Sys::S; ++j) xf = xf && cl_isfinite(x[j]);
          bad = bad || !xf;
          if constexpr (MON) ringM[slot][(int)(h - h0)][lane] = sys.tan_c(aux);
        }
      }
    } else if (role == 1) {
      const int64_t q = p - 1;
      if (q >= 0 && q < nc) {
        const int slot = (int)(q & 1);
        const int64_t h0 = q * H, h1 = (h0 + H < nh) ? h0 + H : nh;
        // the next step's c is read one step ahead, so the LDS latency overlaps this step's work instead of
        // sitting on the frame's step-to-step chain (the index is clamped to the chunk's last step)
        const int ns = (int)(h1 - h0) * tau;
        int si = 0;
        double cn = ringC[slot][0][lane];
        for (int64_t h = h0; h < h1; ++h) {
          const int r = (int)(h - h0);
          for (int t = 0; t < tau; ++t) {
            const double c = cn;
            ++si;
            cn = ringC[slot][si < ns ? si : ns - 1][lane];
            map_tangent_apply(c, W);
          }
          bad = bad || !frame_finite(W);
          QRFact<D, K> f;
          orthonormalize<ALG, D, K>(W, f);
          bad = bad || f.status != QR_OK;
#pragma unroll
          for (int e = 0; e < K * K; ++e) ringR[slot][r][e][lane] = f.R[e];
#pragma unroll
          for (int e = 0; e < D * K; ++e) ringQ[slot][r][e][lane] = f.Q.w[e];
          W = f.Q;
          st.count_max = st.count_max > f.count ? st.count_max : f.count;
        }
      }
    } else if (role <= 3) {
      const int64_t q = p - 2;
      if (q >= 0 && q < nc) {
        const int slot = (int)(q & 1);
        const int64_t h0 = q * H, h1 = (h0 + H < nh) ? h0 + H : nh;
        for (int64_t h = h0 + ((h0 + role) & 1); h < h1; h += 2) {  // role 2: even h, role 3: odd h
          const int r = (int)(h - h0);
          double R[K * K];
#pragma unroll
          for (int e = 0; e < K * K; ++e) R[e] = ringR[slot][r][e][lane];
          {
            Frame<D, K> Q;
#pragma unroll
            for (int e = 0; e < D * K; ++e) Q.w[e] = ringQ[slot][r][e][lane];
            const double Gw = window_Gw<K>(R);
            if (QRSpectrum<ALG, D, K>::checks_conditioning && !(Gw * CL_EPS64 < 0.5)) bad = true;
            const double orth = K >= 2 ? orth_loss<D, K>(Q) : 0.0;
            const double g = (ALG == QR_CGS || ALG == QR_MGS) ? Gw : 1.0;
            const double thr = qr.tol * CL_EPS64 * (double)K * g;
            if (K >= 2 && !(orth <= thr)) bad = true;
            st.G_max = cl_jmax(st.G_max, G_from_Gw<K>(Gw));
            st.orth_max = cl_jmax(st.orth_max, orth);
          }
        }
      }
    } else if (role <= 5) {
      const int col = role - 4;
      const int64_t q = p - 2;
      if (col < K && q >= 0 && q < nc) {
        const int slot = (int)(q & 1);
        const int64_t h0 = q * H, h1 = (h0 + H < nh) ? h0 + H : nh;
        // the column is uniform over the wave; selects keep the state out of indexed (scratch) memory
        CSum sm = col == 0 ? st.sums[0] : st.sums[K - 1];
        for (int64_t h = h0; h < h1; ++h) {
          const int r = (int)(h - h0);
          sm = csum_add(sm, cl_log(ringR[slot][r][col * K + col][lane]));
        }
        if (col == 0) st.sums[0] = sm;
        else st.sums[K - 1] = sm;
        st.n = a.n0 + h1 * tau;
      }
    } else if constexpr (MON) {
      const int64_t q = p - 1;
      if (q >= 0 && q < nc) {
        const int slot = (int)(q & 1);
        const int64_t h0 = q * H, h1 = (h0 + H < nh) ? h0 + H : nh;
        for (int64_t h = h0; h < h1; ++h) {
          double J[4];
          map_jacobian(ringM[slot][(int)(h - h0)][lane], J);
          mon.step(mst, a.n0 + (h + 1) * tau, J);
        }
      }
    }
    __syncthreads();
  }
  bad_by[role][lane] = bad ? 1 : 0;
  if (role == 1) outCount[lane] = st.count_max;
  if (role == 2 || role == 3) {
    outG[role - 2][lane] = st.G_max;
    outO[role - 2][lane] = st.orth_max;
  }
  if (role == 5 && K > 1) outS[1][lane] = st.sums[K - 1];
  __syncthreads();
  bool failed = false;
#pragma unroll
  for (int w = 0; w < NW; ++w) failed = failed || bad_by[w][lane] != 0;
  if (role == 0 && in_range) rerun[i] = (live && failed) ? 1 : 0;
  if (!live) return;
  const bool keep = !failed && nh > 0;
  if (role == 0) {
    if (keep) {
#pragma unroll
      for (int j = 0; j < Sys::S; ++j) a.x[(int64_t)j * a.m + i] = x[j];
    }
  } else if (role == 1) {
    if (keep) {
#pragma unroll
      for (int e = 0; e < D * K; ++e) a.W0[(int64_t)e * a.m + i] = W.w[e];
    }
  } else if (role == 4) {
    if (keep) {
      // the maxima over the even and the odd handoffs, each started from the launch's input value
      st.G_max = cl_jmax(outG[0][lane], outG[1][lane]);
      st.orth_max = cl_jmax(outO[0][lane], outO[1][lane]);
      if (K > 1) st.sums[K - 1] = outS[1][lane];
      st.count_max = outCount[lane];
      st.column = 0;
      state_store(st, a.st, a.m, i, 0);
      a.code[i] = STATUS_OK;
      a.hstep[i] = a.n0 + nh * tau;
    }
  } else if (MON && role == 6) {
    if (keep) state_store(mst, a.st, a.m, i, WMON);
  }
}

}  // namespace cl
