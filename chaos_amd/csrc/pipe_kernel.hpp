// pipe_kernel.hpp - a wave-specialised pipeline for small map ensembles whose chain is one QR spectrum.
//
// A small ensemble (a few thousand orbits) fills a few dozen waves on a 1024-SIMD device, so the fused
// kernel runs at the issue rate of one wave per SIMD: its time is the length of the per-step instruction
// stream (sin/cos of the point, the frame product, the factorisation, its checks and two logs), step after
// step. The work of one orbit-step splits into parts with one-way data flow:
//
//   A  the orbit: sin/cos of x_n, the step, the parameter c_n = K cos x_n of the tangent (standard_map.jl:30-34)
//   B  the frame: W <- M(c) W in the fused_product order, then the factorisation W = QR (carries W = Q). For the
//      Gram-Schmidt family at k = 2 the first column's recurrence (normalise M q_1) does not read the second
//      column, so B splits into B1, column 1, and B2, column 2, which projects on the q_1 that B1 formed one
//      chunk earlier: two step-to-step chains of half the length instead of one.
//   C  the checks of the factorisation: window_Gw conditioning and orth_loss (qr_spectrum.jl:54-73),
//      and the running maxima G_max, orth_max. At k = 2 a one-division estimate of G_w (gw_quick2) settles the
//      checks and the maximum for almost every handoff; window_Gw runs only where the estimate's error bound
//      leaves the outcome open, so every decision and every stored value is window_Gw's
//   D  the exponent sums: add(sums[i], log R_ii), for diagonals that are positive normal numbers
//   E  (chains QRSpectrum + RunMonitor, the recorder's) the monitor's running maxima of the one-step tangent's
//      det and symplecticity residuals and of max |M| at each handoff point, from c at that point
//
// Each part runs on its own waves of a group of 64 orbits (one per lane of every wave): A, B (or B1 and B2) and
// E one wave each; C two waves, for the even and the odd handoffs (a maximum does not depend on the order of its
// terms, and both start from the same value); D one wave per column (each column's compensated sum stays in
// step order). The parts hand data forward through triple-buffered LDS rings, one chunk of steps per stage:
// at iteration p, A writes c for chunk p, B (B1) factorises chunk p - 1, B2 chunk p - 2, C and D read the
// factors of the chunk the last B stage finished one iteration earlier. One barrier per iteration. C and D
// form their per-handoff values for the whole chunk before folding them in order, so the handoffs' logs and
// checks overlap instead of queuing on the running sum. The arithmetic of every part is the fused kernel's,
// in the same order, so the results are bitwise the same.
//
// A failure (non-finite point or frame at a handoff, a failed factorisation or check) is not handled here:
// the group marks the orbit, writes nothing back for it, and the caller reruns the marked orbits with the
// fused kernel from the unchanged inputs, which reproduces the reference's stop rule exactly
// (orbit.jl:48-58; qr_spectrum.jl:54-96).
#pragma once
#include "registry.hpp"

namespace cl {

constexpr int PIPE_LANES = 64;   // orbits per group
constexpr int PIPE_STEPS = 16;   // map steps per chunk at most (tau <= 16)
constexpr int PIPE_HAND = 8;     // handoffs per chunk at most
constexpr int PIPE_SLOTS = 3;    // ring buffers: the chunk an iteration writes, the one it reads, the one in flight

// handoffs per chunk for a given tau: chunks of about 8 steps, at least one handoff
CL_HD int pipe_handoffs(int tau) { return tau <= PIPE_HAND ? PIPE_HAND / tau : 1; }

// B splits into two column waves for the Gram-Schmidt family at k = 2
template <int ALG, int K>
constexpr bool pipe_split() {
  return K == 2 && (ALG == QR_CGS || ALG == QR_MGS || ALG == QR_RGS);
}
template <int ALG, int K, bool MON>
constexpr int pipe_waves() {
  return 6 + (pipe_split<ALG, K>() ? 1 : 0) + (MON ? 1 : 0);
}

// roles, and their waves: wave w runs on SIMD w mod 4. Split: A has a SIMD of its own, the two C waves share
// one, B1 and B2 each share theirs with a D wave (and A with E in the recorder's chain).
enum : int { PR_A = 0, PR_B = 1, PR_C0 = 2, PR_C1 = 3, PR_D0 = 4, PR_D1 = 5, PR_E = 6, PR_B2 = 7 };
template <bool SPLIT>
CL_HD int pipe_role(int wave) {
  if (SPLIT)  // waves 0..7: C even, B1, B2, A, C odd, D col 1, D col 2, E
    return wave == 0 ? PR_C0 : wave == 1 ? PR_B : wave == 2 ? PR_B2 : wave == 3 ? PR_A : wave == 4 ? PR_C1
         : wave == 5 ? PR_D0 : wave == 6 ? PR_D1 : PR_E;
  // waves 0..6: A, B, C even, C odd, D col 2, D col 1, E
  return wave == 0 ? PR_A : wave == 1 ? PR_B : wave == 2 ? PR_C0 : wave == 3 ? PR_C1 : wave == 4 ? PR_D1
       : wave == 5 ? PR_D0 : PR_E;
}

// Column m of the Gram-Schmidt factorisation of a 2 x 2 frame (orthonormalize_gs, m = 0 and m = 1, the same
// operations in the same order): u is the column after the tangent, q1 the first column of Q (m = 1).
// Returns the norm gamma; r01 gets the projection coefficient (m = 1), passes the reorthogonalisation passes.
template <int ALG>
CL_HD double pipe_gs_column(int m, double* u, const double* q1, double* r01, int* passes, int8_t* status) {
  *passes = 1;
  *r01 = 0.0;
  if (m == 1) {
    if (ALG == QR_CGS || ALG == QR_MGS) {
      const double c = dot<2>(q1, u);
      *r01 = c;
#pragma unroll
      for (int l = 0; l < 2; ++l) u[l] = u[l] - c * q1[l];
    } else {  // RGS
      const double eta = 1.0 / cl_sqrt(2.0);
      double nold = safe_norm<2>(u);
      int ps = 0;
      double r = 0.0;
      for (;;) {
        ps += 1;
        double vold[2] = {u[0], u[1]};
        const double c = dot<2>(q1, vold);
        r = r + c;
#pragma unroll
        for (int l = 0; l < 2; ++l) u[l] = u[l] - c * q1[l];
        double nnew = safe_norm<2>(u);
        if (nnew >= nold * eta || !(cl_isfinite(nnew) && nnew > 0.0)) break;
        if (ps == MAX_PASSES) {
          ps = MAX_PASSES + 1;
          break;
        }
        nold = nnew;
      }
      *r01 = r;
      *passes = ps;
    }
  }
  const double gamma = safe_norm<2>(u);
  *status = *passes > MAX_PASSES ? QR_PASSES : (!(cl_isfinite(gamma) && gamma > 0.0) ? QR_PIVOT : QR_OK);
  return gamma;
}

// MON: the chain is QRSpectrum then RunMonitor (its state words follow the QR state's)
template <class Sys, int ALG, int K, bool MON>
__global__ void __launch_bounds__((pipe_waves<ALG, K, MON>() * PIPE_LANES))
pipe_kernel(KArgs a, Sys sys, QRSpectrum<ALG, 2, K> qr, int H, uint8_t* rerun) {
  constexpr int D = 2;
  constexpr bool SPLIT = pipe_split<ALG, K>();
  constexpr int NW = pipe_waves<ALG, K, MON>();
  constexpr int WMON = (int)(sizeof(QRState<K>) / 8);  // the monitor's first state word
  constexpr int LAG = SPLIT ? 3 : 2;                    // iterations from A's chunk to C's and D's
  typedef QRState<K> St;
  __shared__ double ringC[PIPE_SLOTS][PIPE_STEPS][PIPE_LANES];
  __shared__ double ringR[PIPE_SLOTS][PIPE_HAND][K * K][PIPE_LANES];
  __shared__ double ringQ[PIPE_SLOTS][PIPE_HAND][D * K][PIPE_LANES];
  __shared__ double ringM[MON ? 2 : 1][MON ? PIPE_HAND : 1][PIPE_LANES];  // c at each handoff point
  __shared__ int bad_by[8][PIPE_LANES];
  __shared__ double outG[2][PIPE_LANES], outO[2][PIPE_LANES];
  __shared__ CSum outS[2][PIPE_LANES];  // [1]: column 2's sum, from the wave that formed it
  __shared__ int64_t outCount[2][PIPE_LANES];  // [1]: B2's (the reorthogonalisation passes of column 2)

  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int role = pipe_role<SPLIT>(wave);
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
  if (MON && role == PR_E && live && a.n0 != 0) state_load(mst, a.st, a.m, i, WMON);

  double x[Sys::S];
  typename Sys::Aux aux;
  Frame<D, K> W;
#pragma unroll
  for (int j = 0; j < Sys::S; ++j) x[j] = 0.0;
#pragma unroll
  for (int e = 0; e < D * K; ++e) W.w[e] = 0.0;
  if (role == PR_A && in_range) {
#pragma unroll
    for (int j = 0; j < Sys::S; ++j) x[j] = a.x[(int64_t)j * a.m + i];
  }
  if ((role == PR_B || role == PR_B2) && in_range) {
#pragma unroll
    for (int e = 0; e < D * K; ++e) W.w[e] = a.W0[(int64_t)e * a.m + i];
  }
  if (role == PR_A) sys.aux(x, aux);

  const int64_t nh = a.steps / tau;
  const int64_t nc = (nh + H - 1) / H;
  bool bad = false;
  int64_t cmax = st.count_max;
  constexpr bool GS_THR = ALG == QR_CGS || ALG == QR_MGS;  // orth_loss's threshold scales with G_w
  const double c0 = qr.tol * CL_EPS64 * (double)K;
  double gw_max = 0.0;  // C: the largest G_w of its handoffs (G_w >= 0)
  for (int64_t p = 0; p < nc + LAG; ++p) {
    if (role == PR_A) {
      if (p < nc) {
        const int slot = (int)(p % PIPE_SLOTS);
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
          if constexpr (MON) ringM[(int)(p & 1)][(int)(h - h0)][lane] = sys.tan_c(aux);
        }
      }
    } else if (role == PR_B || role == PR_B2) {
      // B (or B1, column 1) works on chunk p - 1; B2 (column 2) on chunk p - 2, with B1's q_1 of that chunk
      const int64_t q = p - (role == PR_B ? 1 : 2);
      if (q >= 0 && q < nc) {
        const int slot = (int)(q % PIPE_SLOTS);
        const int64_t h0 = q * H, h1 = (h0 + H < nh) ? h0 + H : nh;
        // the next step's c is read one step ahead, so the LDS latency overlaps this step's work instead of
        // sitting on the frame's step-to-step chain (the index is clamped to the chunk's last step)
        const int ns = (int)(h1 - h0) * tau;
        int si = 0;
        double cn = ringC[slot][0][lane];
        for (int64_t h = h0; h < h1; ++h) {
          const int r = (int)(h - h0);
          if constexpr (SPLIT) {
            const int col = role == PR_B ? 0 : 1;
            Frame<D, 1> w1;
            w1.w[0] = W.w[2 * col];
            w1.w[1] = W.w[2 * col + 1];
            for (int t = 0; t < tau; ++t) {
              const double c = cn;
              ++si;
              cn = ringC[slot][si < ns ? si : ns - 1][lane];
              map_tangent_apply(c, w1);
            }
            bad = bad || !frame_finite(w1);
            double u[2] = {w1.w[0], w1.w[1]};
            double q1[2] = {0.0, 0.0};
            if (role == PR_B2) {
              q1[0] = ringQ[slot][r][0][lane];
              q1[1] = ringQ[slot][r][1][lane];
            }
            double r01;
            int passes;
            int8_t status;
            const double gamma = pipe_gs_column<ALG>(col, u, q1, &r01, &passes, &status);
            bad = bad || status != QR_OK;
            const double qa = u[0] / gamma, qb = u[1] / gamma;
            ringQ[slot][r][2 * col][lane] = qa;
            ringQ[slot][r][2 * col + 1][lane] = qb;
            // R column-major: column 1 is (gamma, 0), column 2 (r01, gamma)
            ringR[slot][r][2 * col][lane] = col == 0 ? gamma : r01;
            ringR[slot][r][2 * col + 1][lane] = col == 0 ? 0.0 : gamma;
            W.w[2 * col] = qa;
            W.w[2 * col + 1] = qb;
            cmax = cmax > passes ? cmax : (int64_t)passes;
          } else {
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
            cmax = cmax > f.count ? cmax : (int64_t)f.count;
          }
        }
      }
    } else if (role == PR_C0 || role == PR_C1) {
      const int64_t q = p - LAG;
      if (q >= 0 && q < nc) {
        const int slot = (int)(q % PIPE_SLOTS);
        const int64_t h0 = q * H, h1 = (h0 + H < nh) ? h0 + H : nh;
        const int par = (int)((h0 + (role == PR_C0 ? 0 : 1)) & 1);  // C0: even h, C1: odd h
        // the checks as QRSpectrum::step makes them, from window_Gw; gw_max collects G_w (the largest G is
        // G_from_Gw of the largest G_w)
        auto exact = [&](double Gw, double orth) {
          if (QRSpectrum<ALG, D, K>::checks_conditioning && !(Gw * CL_EPS64 < 0.5)) bad = true;
          const double g = GS_THR ? Gw : 1.0;
          const double thr = qr.tol * CL_EPS64 * (double)K * g;
          if (K >= 2 && !(orth <= thr)) bad = true;
          gw_max = cl_jmax(gw_max, Gw);
        };
        double gw[PIPE_HAND / 2], ol[PIPE_HAND / 2];
#pragma unroll
        // for every entry of the slot (entries past the chunk's end hold an earlier chunk's factors and go unused),
        // as straight-line code
        for (int u = 0; u < PIPE_HAND / 2; ++u) {
          const int r = par + 2 * u;
          double R[K * K];
#pragma unroll
          for (int e = 0; e < K * K; ++e) R[e] = ringR[slot][r][e][lane];
          Frame<D, K> Q;
#pragma unroll
          for (int e = 0; e < D * K; ++e) Q.w[e] = ringQ[slot][r][e][lane];
          gw[u] = K == 2 ? gw_quick2(R) : window_Gw<K>(R);
          ol[u] = K >= 2 ? orth_loss<D, K>(Q) : 0.0;
        }
#pragma unroll
        for (int u = 0; u < PIPE_HAND / 2; ++u) {
          cl_pin(gw[u]);
          cl_pin(ol[u]);
        }
#pragma unroll
        for (int u = 0; u < PIPE_HAND / 2; ++u) {
          const int r = par + 2 * u;
          if (h0 + r < h1) {
            if constexpr (K == 2) {
              // window_Gw only where the estimate (within GW_QUICK_REL of it) cannot settle a check or might
              // raise the running maximum: a new maximum, G_w near 2^52, orth_loss near its threshold
              const double g = gw[u];
              bool sure = g >= 0.0 && g < GW_QUICK_GATE && g * (1.0 + GW_QUICK_REL) < gw_max;
              sure = sure && ol[u] <= (GS_THR ? c0 * g * (1.0 - 2.0 * GW_QUICK_REL) : c0 * 1.0);
              if (!sure) {
                __asm__ volatile("");  // keeps the rare branch a branch
                double R[K * K];
#pragma unroll
                for (int e = 0; e < K * K; ++e) R[e] = ringR[slot][r][e][lane];
                exact(window_Gw<K>(R), ol[u]);
              }
            } else {
              exact(gw[u], ol[u]);
            }
            st.orth_max = cl_jmax(st.orth_max, ol[u]);
          }
        }
      }
    } else if (role == PR_D0 || role == PR_D1) {
      const int col = role - PR_D0;
      const int64_t q = p - LAG;
      if (col < K && q >= 0 && q < nc) {
        const int slot = (int)(q % PIPE_SLOTS);
        const int64_t h0 = q * H, h1 = (h0 + H < nh) ? h0 + H : nh;
        const int nr = (int)(h1 - h0);
        // the logs of the chunk's handoffs do not depend on each other: form them all, then fold in step order.
        // cl_log_normal is cl_log's arithmetic without its branches, so the logs interleave; a diagonal outside
        // its domain (zero, negative, subnormal, not finite) sends the orbit to the rerun
        double lg[PIPE_HAND];
#pragma unroll
        for (int r = 0; r < PIPE_HAND; ++r) {
          const double v = ringR[slot][r][col * K + col][lane];
          const bool ok = r >= nr || cl_is_pos_normal(v);
          bad = bad || !ok;
          lg[r] = cl_log_normal(r < nr && ok ? v : 1.0, 0);
        }
        // computed here, not sunk into the conditional folds below (where they would run one after another)
#pragma unroll
        for (int r = 0; r < PIPE_HAND; ++r) cl_pin(lg[r]);
        // the column is uniform over the wave; selects keep the state out of indexed (scratch) memory
        CSum sm = col == 0 ? st.sums[0] : st.sums[K - 1];
#pragma unroll
        for (int r = 0; r < PIPE_HAND; ++r)
          if (r < nr) sm = csum_add(sm, lg[r]);
        if (col == 0) st.sums[0] = sm;
        else st.sums[K - 1] = sm;
        st.n = a.n0 + h1 * tau;
      }
    } else if constexpr (MON) {
      const int64_t q = p - 1;
      if (q >= 0 && q < nc) {
        const int64_t h0 = q * H, h1 = (h0 + H < nh) ? h0 + H : nh;
        for (int64_t h = h0; h < h1; ++h) {
          double J[4];
          map_jacobian(ringM[(int)(q & 1)][(int)(h - h0)][lane], J);
          mon.step(mst, a.n0 + (h + 1) * tau, J);
        }
      }
    }
    __syncthreads();
  }
  bad_by[wave][lane] = bad ? 1 : 0;
  if (role == PR_B) outCount[0][lane] = cmax;
  if (role == PR_B2) outCount[1][lane] = cmax;
  if (role == PR_C0 || role == PR_C1) {
    outG[role - PR_C0][lane] = gw_max;
    outO[role - PR_C0][lane] = st.orth_max;
  }
  if (role == PR_D1 && K > 1) outS[1][lane] = st.sums[K - 1];
  __syncthreads();
  bool failed = false;
#pragma unroll
  for (int w = 0; w < NW; ++w) failed = failed || bad_by[w][lane] != 0;
  // the count is the maximum over both columns (B2 holds column 2's reorthogonalisation passes)
  int64_t cnt = outCount[0][lane];
  if (SPLIT) cnt = cnt > outCount[1][lane] ? cnt : outCount[1][lane];
  if (role == PR_A && in_range) rerun[i] = (live && failed) ? 1 : 0;
  if (!live) return;
  const bool keep = !failed && nh > 0;
  if (role == PR_A) {
    if (keep) {
#pragma unroll
      for (int j = 0; j < Sys::S; ++j) a.x[(int64_t)j * a.m + i] = x[j];
    }
  } else if (role == PR_B2) {
    if (keep) {
      a.W0[(int64_t)2 * a.m + i] = W.w[2];
      a.W0[(int64_t)3 * a.m + i] = W.w[3];
    }
  } else if (role == PR_B) {
    if (keep) {
      const int ne = SPLIT ? 2 : D * K;
#pragma unroll
      for (int e = 0; e < D * K; ++e)
        if (e < ne) a.W0[(int64_t)e * a.m + i] = W.w[e];
    }
  } else if (role == PR_D0) {
    if (keep) {
      // the maxima over the even and the odd handoffs, each started from the launch's input value; G_max from
      // the largest G_w (nh > 0 handoffs, each with G_w > 0)
      st.G_max = cl_jmax(st.G_max, G_from_Gw<K>(cl_jmax(outG[0][lane], outG[1][lane])));
      st.orth_max = cl_jmax(outO[0][lane], outO[1][lane]);
      if (K > 1) st.sums[K - 1] = outS[1][lane];
      st.count_max = cnt;
      st.column = 0;
      state_store(st, a.st, a.m, i, 0);
      a.code[i] = STATUS_OK;
      a.hstep[i] = a.n0 + nh * tau;
    }
  } else if (MON && role == PR_E) {
    if (keep) state_store(mst, a.st, a.m, i, WMON);
  }
}

}  // namespace cl
