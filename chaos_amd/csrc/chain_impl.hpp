// chain_impl.hpp - the coupled standard-map chain (StandardMapChain, src/Systems/standard_map_chain.jl:1-53)
// with its banded tangent (ChainTangent, src/Systems/chain_tangent.jl:21-97) and the QR spectrum with
// Householder reflectors (src/Lyapunov/householder.jl:17-65, src/Lyapunov/qr_spectrum.jl:56-78) on frames of
// d = 2N rows and k = 16 columns.
//
// One thread per chain site: thread t of an orbit holds q_t, p_t and rows t (q part) and N+t (p part) of
// the frame, 2k doubles. An orbit spans N lanes of a wave (N <= 64; 64/N orbits per 64-lane group) or a
// 128-thread group (N = 128). Per step each thread evaluates sin/cos of its position and of the bond to its
// right neighbour; the left bond's values come from the left neighbour, exact because q_j - q_l is exactly
// -(q_l - q_j) and the portable sin is odd and cos even bit for bit. The banded product needs the
// neighbours' q rows of the frame (cross-lane moves, LDS for N = 128).
//
// Column reductions (norms and dot products over the 2N rows) take a partial per site over its q row (when
// in range) then its p row, combined by a butterfly over the N sites, smallest level first (pairs of
// sites, then pairs of pairs, ...); every thread ends with the same value because IEEE addition commutes.
// The in-wave levels use DPP lane moves and the gfx950 permlane swaps (no LDS); at N = 128 the two waves'
// sums are added last through LDS. The CPU oracle restates this order (oracle/oracle_chain.hpp, RED_SITES)
// and is checked against the reference's row order.
//
// HBM layout (orbit-major, the reference's host layout): x[i][2N], W[i][k][2N], states[i][2k+9 words].
// Before each QR the thread stores its frame rows to W (a failed factorisation leaves the frame it was
// given, qr_spectrum.jl:64-66). Per-orbit step counts and an orbit order (longest first) come from the
// ragged entry point; groups are dispatched in order, which balances uneven orbits over the CUs.
#pragma once
#include <hip/hip_runtime.h>
#include <type_traits>
#include "cl_core.hpp"
#include "chain_kernels.hpp"
#include "../../include/chaoslab_hip.h"

namespace cl {

namespace chain_detail {

// Register-resident frame rows: a vector type, so that a uniform runtime column index (the reflector loops)
// is an indexed register move rather than a private-memory array.
template <int K>
using dvec = double __attribute__((ext_vector_type(K)));
template <int K>
using fvec = float __attribute__((ext_vector_type(K)));
// Frame registers of the element type: Float64, or Float32 tangents (CL_ELT_F64_F32TAN).
template <class FT, int K>
using frame_vec = typename std::conditional<std::is_same<FT, float>::value, fvec<K>, dvec<K>>::type;

template <int NS>
struct Geo {
  static constexpr int BLOCK = NS > 64 ? NS : 64;
  static constexpr int OPB = BLOCK / NS;
};

// ---------------------------------------------------------------- cross-lane primitives within one orbit
// In-wave butterfly level L (smallest first). L = 0, 1: the lane ^ 1, ^ 2 (DPP quad_perm); L = 2, 3: the lane
// mirrored in its 8-lane half-row, 16-lane row (DPP row_half_mirror, row_mirror): after the lower levels every
// lane of a 4-lane (8-lane) group holds the group's value, so the mirrored lane pairs the same two groups as
// lane ^ 4 (^ 8) would; L = 4, 5: v_permlane16_swap, v_permlane32_swap exchange the 16-lane rows and the
// 32-lane halves and leave both values of a pair in every lane. `pair2` returns the two values whose sum
// (maximum, or) the level forms, in an order that may differ between the lanes of a pair (the operations
// used commute).
__device__ __attribute__((always_inline)) inline double dbl_of(unsigned lo, unsigned hi) {
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
template <int L>
__device__ __attribute__((always_inline)) inline void pair2(double v, double& a, double& b) {
  const unsigned long long bits = (unsigned long long)__double_as_longlong(v);
  const unsigned lo = (unsigned)bits, hi = (unsigned)(bits >> 32);
  if constexpr (L < 4) {
    constexpr int ctrl = L == 0 ? 0xB1 : (L == 1 ? 0x4E : (L == 2 ? 0x141 : 0x140));
    a = v;
    b = dbl_of((unsigned)__builtin_amdgcn_mov_dpp((int)lo, ctrl, 0xf, 0xf, false),
               (unsigned)__builtin_amdgcn_mov_dpp((int)hi, ctrl, 0xf, 0xf, false));
  } else if constexpr (L == 4) {
    const auto l = __builtin_amdgcn_permlane16_swap(lo, lo, false, false);
    const auto h = __builtin_amdgcn_permlane16_swap(hi, hi, false, false);
    a = dbl_of(l[0], h[0]);
    b = dbl_of(l[1], h[1]);
  } else {
    const auto l = __builtin_amdgcn_permlane32_swap(lo, lo, false, false);
    const auto h = __builtin_amdgcn_permlane32_swap(hi, hi, false, false);
    a = dbl_of(l[0], h[0]);
    b = dbl_of(l[1], h[1]);
  }
}
template <int L>
__device__ __attribute__((always_inline)) inline int or_level(int v) {
  if constexpr (L < 4) {
    constexpr int ctrl = L == 0 ? 0xB1 : (L == 1 ? 0x4E : (L == 2 ? 0x141 : 0x140));
    return v | __builtin_amdgcn_mov_dpp(v, ctrl, 0xf, 0xf, false);
  } else if constexpr (L == 4) {
    const auto x = __builtin_amdgcn_permlane16_swap((unsigned)v, (unsigned)v, false, false);
    return (int)(x[0] | x[1]);
  } else {
    const auto x = __builtin_amdgcn_permlane32_swap((unsigned)v, (unsigned)v, false, false);
    return (int)(x[0] | x[1]);
  }
}

// Barrier of the orbit's two waves (N = 128) for their LDS exchanges: the fences cover LDS only, so the barrier does not
// wait for the group's outstanding stores to HBM (a full __syncthreads would drain them at every exchange).
__device__ __attribute__((always_inline)) inline void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

template <int NS>
struct Orb {
  // in-wave butterfly levels of an orbit: log2 of its lanes in one wave
  static constexpr int LV = NS >= 64 ? 6 : (NS == 32 ? 5 : (NS == 16 ? 4 : (NS == 8 ? 3 : (NS == 4 ? 2 : 1))));

  double* xb;  // LDS exchange buffer (N = 128 only): 128 * 16 doubles
  int t;       // site
  int base;    // first lane of this orbit inside its wave (N <= 64)

  __device__ __attribute__((always_inline)) double bcast(double v, int owner) const {
    if constexpr (NS == 64) {  // the orbit is the wave: a scalar read of the owner lane
      const long long b = __double_as_longlong(v);
      const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), owner);
      const int hi = __builtin_amdgcn_readlane((int)(b >> 32), owner);
      return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
    } else if constexpr (NS > 64) {
      lds_barrier();
      if (t == owner) xb[0] = v;
      lds_barrier();
      return xb[0];
    } else {
      return __shfl(v, base + owner, 64);
    }
  }
  __device__ __attribute__((always_inline)) int any(int v) const {
    if constexpr (NS > 64) {
      lds_barrier();
      if (t == 0) reinterpret_cast<int*>(xb)[0] = 0;
      lds_barrier();
      if (v) atomicOr(reinterpret_cast<int*>(xb), v);
      lds_barrier();
      return reinterpret_cast<int*>(xb)[0];
    } else {
      v = or_level<0>(v);
      if constexpr (LV > 1) v = or_level<1>(v);
      if constexpr (LV > 2) v = or_level<2>(v);
      if constexpr (LV > 3) v = or_level<3>(v);
      if constexpr (LV > 4) v = or_level<4>(v);
      if constexpr (LV > 5) v = or_level<5>(v);
      return v;
    }
  }
  template <int L, int M>
  __device__ __attribute__((always_inline)) static void sum_level(double (&v)[M], int lo, int hi) {
#pragma unroll
    for (int j = 0; j < M; ++j)
      if (j >= lo && j <= hi) {
        double a, b;
        pair2<L>(v[j], a, b);
        v[j] = a + b;
      }
  }
  // butterfly sums of v[lo..hi] (the other entries are left untouched)
  template <int M>
  __device__ __attribute__((always_inline)) void sum(double (&v)[M], int lo = 0, int hi = M - 1) const {
    sum_level<0>(v, lo, hi);
    if constexpr (LV > 1) sum_level<1>(v, lo, hi);
    if constexpr (LV > 2) sum_level<2>(v, lo, hi);
    if constexpr (LV > 3) sum_level<3>(v, lo, hi);
    if constexpr (LV > 4) sum_level<4>(v, lo, hi);
    if constexpr (LV > 5) sum_level<5>(v, lo, hi);
    if constexpr (NS > 64) {
      lds_barrier();
#pragma unroll
      for (int j = 0; j < M; ++j)
        if (j >= lo && j <= hi) xb[j * 128 + t] = v[j];
      lds_barrier();
#pragma unroll
      for (int j = 0; j < M; ++j)
        if (j >= lo && j <= hi) v[j] = v[j] + xb[j * 128 + (t ^ 64)];
    }
  }
  __device__ __attribute__((always_inline)) double sum1(double v) const {
    double a[1] = {v};
    sum<1>(a);
    return a[0];
  }
  template <int L>
  __device__ __attribute__((always_inline)) static double max_level(double v) {
    double a, b;
    pair2<L>(v, a, b);
    return fmax(a, b);
  }
  __device__ __attribute__((always_inline)) static double max_wave(double v) {
    v = max_level<0>(v);
    if constexpr (LV > 1) v = max_level<1>(v);
    if constexpr (LV > 2) v = max_level<2>(v);
    if constexpr (LV > 3) v = max_level<3>(v);
    if constexpr (LV > 4) v = max_level<4>(v);
    if constexpr (LV > 5) v = max_level<5>(v);
    return v;
  }
  __device__ __attribute__((always_inline)) double max1(double v) const {
    v = max_wave(v);
    if constexpr (NS > 64) {
      lds_barrier();
      xb[t] = v;
      lds_barrier();
      v = fmax(v, xb[t ^ 64]);
    }
    return v;
  }
  __device__ __attribute__((always_inline)) void max2(double& u, double& v) const {
    u = max_wave(u);
    v = max_wave(v);
    if constexpr (NS > 64) {
      lds_barrier();
      xb[t] = u;
      xb[128 + t] = v;
      lds_barrier();
      u = fmax(u, xb[t ^ 64]);
      v = fmax(v, xb[128 + (t ^ 64)]);
    }
  }
  // One reduction of three values over the orbit's sites: the largest of mx and the butterfly sums of a and b (the
  // order of sum(), so each equals its own butterfly sum bit for bit).
  __device__ __attribute__((always_inline)) void msum2(double& mx, double& a, double& b) const {
    auto level = [&](auto L) {
      double x, y;
      pair2<decltype(L)::value>(mx, x, y);
      mx = fmax(x, y);
      pair2<decltype(L)::value>(a, x, y);
      a = x + y;
      pair2<decltype(L)::value>(b, x, y);
      b = x + y;
    };
    level(std::integral_constant<int, 0>());
    if constexpr (LV > 1) level(std::integral_constant<int, 1>());
    if constexpr (LV > 2) level(std::integral_constant<int, 2>());
    if constexpr (LV > 3) level(std::integral_constant<int, 3>());
    if constexpr (LV > 4) level(std::integral_constant<int, 4>());
    if constexpr (LV > 5) level(std::integral_constant<int, 5>());
    if constexpr (NS > 64) {
      lds_barrier();
      xb[t] = mx;
      xb[128 + t] = a;
      xb[256 + t] = b;
      lds_barrier();
      mx = fmax(mx, xb[t ^ 64]);
      a = a + xb[128 + (t ^ 64)];
      b = b + xb[256 + (t ^ 64)];
    }
  }
  __device__ __attribute__((always_inline)) double right1(double v) const {
    if constexpr (NS > 64) {
      lds_barrier();
      xb[t] = v;
      lds_barrier();
      return xb[(t + 1) & (NS - 1)];
    } else {
      return __shfl(v, base + ((t + 1) & (NS - 1)), 64);
    }
  }
  __device__ __attribute__((always_inline)) void left2(double u, double v, double& ul, double& vl) const {
    if constexpr (NS > 64) {
      lds_barrier();
      xb[t] = u;
      xb[128 + t] = v;
      lds_barrier();
      ul = xb[(t + NS - 1) & (NS - 1)];
      vl = xb[128 + ((t + NS - 1) & (NS - 1))];
    } else {
      const int sl = base + ((t + NS - 1) & (NS - 1));
      ul = __shfl(u, sl, 64);
      vl = __shfl(v, sl, 64);
    }
  }
  // values of v at the left and right neighbour sites (ring order inside the orbit)
  template <int M, class V, class E = double>
  __device__ __attribute__((always_inline)) void nbr(const V& v, V& l, V& r) const {
    if constexpr (NS > 64) {
      E* xe = reinterpret_cast<E*>(xb);  // the exchange buffer holds elements of the frame's type
      lds_barrier();
#pragma unroll
      for (int j = 0; j < M; ++j) xe[j * 128 + t] = v[j];
      lds_barrier();
#pragma unroll
      for (int j = 0; j < M; ++j) {
        l[j] = xe[j * 128 + ((t + NS - 1) & (NS - 1))];
        r[j] = xe[j * 128 + ((t + 1) & (NS - 1))];
      }
    } else {
      const int sl = base + ((t + NS - 1) & (NS - 1)), sr = base + ((t + 1) & (NS - 1));
#pragma unroll
      for (int j = 0; j < M; ++j) {
        l[j] = __shfl(v[j], sl, 64);
        r[j] = __shfl(v[j], sr, 64);
      }
    }
  }
};

// Safe norm (src/Foundation/norm.jl:8-12) of one column over the rows >= j: the largest |entry|, then
// m sqrt(sum (a/m)^2), the sum per site (q row then p row) and by the butterfly. The factorisation only
// runs on finite frames (the handoff's status check comes first) and orthogonal reflections keep the
// entries finite, so the hardware max of the absolute values equals Julia's NaN-propagating max.
template <int NS>
__device__ __attribute__((always_inline)) double col_norm_m(const Orb<NS>& o, double mx, double aq, double ap, bool qin, bool pin) {
  // branch-free: a zero or non-finite maximum is the result, selected after the arithmetic has run on 1
  const bool ok = cl_isfinite(mx) && mx > 0.0;
  const double m = ok ? mx : 1.0;
  const double b = aq / m, c = ap / m;
  const double part = qin ? b * b + c * c : (pin ? c * c : 0.0);
  const double r = m * cl_sqrt(o.sum1(part));
  return ok ? r : mx;
}

struct QRChainState {
  // sums of the columns this thread owns (c = t, t + N, ...), the rest replicated
  int64_t n, count_max, column, step;
  double G_max, orth_max, value, threshold;
  int8_t code;
};

__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// LDS visibility between the threads of one orbit: the whole group for N = 128, the wave otherwise.
template <int NS>
__device__ inline void osync() {
  if constexpr (NS > 64) lds_barrier();
  else wave_sync();
}


// One handoff of QRSpectrum(HouseholderQR()) (qr_spectrum.jl:56-78) on the orbit's frame (Wq, Wp rows of this
// thread): reflector sweep (householder.jl:17-44), Q formation (:47-53), sign fix (:55-65), window_Gw
// (orthonormalize.jl:112-127), orth_loss (qr_spectrum.jl:4) and the log sums. The reflector loops are
// unrolled (column indices are compile-time register indices); each thread keeps its rows of the reflectors
// in LDS (Us, one slot per thread and reflector), the triangle R and its inverse too, per orbit (Ro, Rio).
// Upper-triangular k x k factors in LDS, packed by columns: entry (r, l), r <= l, at l (l + 1) / 2 + r.
template <int KC>
struct Tri {
  static constexpr int SIZE = KC * (KC + 1) / 2;
  static __device__ __attribute__((always_inline)) constexpr int at(int r, int l) { return l * (l + 1) / 2 + r; }
};

// Sign fix and pivot check (householder.jl:55-65) on the packed triangle Ro of the swept matrix, R(r, l) = sgn_r A(r, l)
// (neg: bit r set when sgn_r = -1; piv: the first pivot that is zero or not finite, 1-based, or 0), and with all pivots
// valid window_Gw (orthonormalize.jl:112-127): R^-1 column by column into Rio, then |R|_F |R^-1|_F.
// R^-1 column by column into Rio (orthonormalize.jl:116-124) and |R|_F |R^-1|_F, from R = sgn_r Ro (divided by mR
// when SC: the window's rescaling of a triangle with entries beyond 1e+-100).
template <int NS, int KC, bool SC>
__device__ __attribute__((always_inline)) inline double window_gw_chain(const Orb<NS>& o, const double* Ro, double* Rio,
                                                                        uint32_t neg, double mR, bool wr) {
  using TR = Tri<KC>;
  constexpr int SPT = (KC + NS - 1) / NS;
  const int t = o.t;
  auto sgn = [&](int r) -> double { return (neg >> r) & 1u ? -1.0 : 1.0; };
  auto getR = [&](int r, int l) -> double {
    const double v = sgn(r) * Ro[TR::at(r, l)];
    return SC ? v / mR : v;
  };
  // R^-1 column cc by its owner thread(s), rows from the bottom up
#pragma unroll
  for (int s2 = 0; s2 < SPT; ++s2) {
    const int cc = t + s2 * NS;
    if (cc < KC) {
      double Ri[KC];
#pragma unroll
      for (int r = 0; r < KC; ++r) Ri[r] = 0.0;
      // one division per row for every lane (the owners of different columns do not diverge): row r of column cc is
      // 1 / R(r, r) at r = cc, -(sum over l in (r, cc] of R(r, l) Ri[l], in l order from 0.0) / R(r, r) above it, 0 below
#pragma unroll
      for (int r = KC - 1; r >= 0; --r) {
        double acc = 0.0;
#pragma unroll
        for (int l = r + 1; l < KC; ++l) acc = l <= cc ? acc + getR(r, l) * Ri[l] : acc;
        const double q = (r == cc ? 1.0 : -acc) / getR(r, r);
        Ri[r] = r <= cc ? q : 0.0;
      }
#pragma unroll
      for (int r = 0; r < KC; ++r)
        if (wr && r <= cc) Rio[TR::at(r, cc)] = Ri[r];
    }
  }
  osync<NS>();
  // Frobenius norms in column-major order (the zeros below the diagonal add nothing); the squares do not depend on the
  // row signs
  double sa, sb;
  {
    const double v = SC ? Ro[0] / mR : Ro[0];
    sa = v * v;
    const double vi = Rio[0];
    sb = vi * vi;
  }
#pragma unroll
  for (int e = 1; e < TR::SIZE; ++e) {
    const double u = SC ? Ro[e] / mR : Ro[e], ui = Rio[e];
    sa = sa + u * u;
    sb = sb + ui * ui;
  }
  return cl_sqrt(sa) * cl_sqrt(sb);
}

// Sign fix and pivot check (householder.jl:55-65) on the packed triangle Ro of the swept matrix, R(r, l) = sgn_r A(r, l)
// (neg: bit r set when sgn_r = -1; piv: the first pivot that is zero or not finite, 1-based, or 0), and with all pivots
// valid window_Gw (orthonormalize.jl:112-127): R^-1 column by column into Rio, then |R|_F |R^-1|_F.
template <int NS, int KC>
__device__ __attribute__((always_inline)) inline void r_checks(const Orb<NS>& o, const double* Ro, double* Rio,
                                                               uint32_t& neg, int& piv, double& gw_val,
                                                               bool wr = true) {
  using TR = Tri<KC>;
  const int t = o.t;
  neg = 0;
  piv = 0;
  gw_val = 0.0;
#pragma unroll
  for (int j = 0; j < KC; ++j) {
    const double dj = Ro[TR::at(j, j)];
    if (dj < 0.0) neg |= 1u << j;
    const double pv = (dj < 0.0 ? -1.0 : 1.0) * dj;
    if (piv == 0 && !(cl_isfinite(pv) && pv > 0.0)) piv = j + 1;
  }
  if (piv == 0) {
    // the largest |entry| (order-free: every entry is finite here), the entries spread over the orbit's lanes
    double mR = 0.0;
#pragma unroll
    for (int i = 0; i < (TR::SIZE + NS - 1) / NS; ++i) {
      const int e = t + i * NS;
      if (e < TR::SIZE) mR = fmax(mR, cl_fabs(Ro[e]));
    }
    mR = o.max1(mR);
    const bool sc = (mR > 1e100 || mR < 1e-100) && cl_isfinite(mR) && mR > 0.0;
    // the rescaled form is rare: its own instance, so the usual one carries no divisions by mR
    if (sc)
      gw_val = window_gw_chain<NS, KC, true>(o, Ro, Rio, neg, mR, wr);
    else
      gw_val = window_gw_chain<NS, KC, false>(o, Ro, Rio, neg, mR, wr);
  }
}

template <int NS, int KC, int BLOCK>
__device__ __attribute__((always_inline)) inline void qr_handoff(const Orb<NS>& o, const ChainArgs& a, int64_t n,
                                                                 dvec<KC>& Wq, dvec<KC>& Wp, double* Wo,
                                                                 CSum (&sums)[(KC + NS - 1) / NS], QRChainState& qs,
                                                                 double* Us, double* Ro, double* Rio) {
  constexpr int D = 2 * NS;
  constexpr int NREF = KC < D ? KC : KC - 1;
  constexpr int SPT = (KC + NS - 1) / NS;
  const int t = o.t;
  // the frame this factorisation starts from: what a failed handoff leaves in place
#pragma unroll
  for (int c = 0; c < KC; ++c) {
    Wo[c * D + t] = Wq[c];
    Wo[c * D + NS + t] = Wp[c];
  }
  int count = 0;
  using TR = Tri<KC>;
  const int tid = (int)threadIdx.x;
#pragma unroll
  for (int j = 0; j < NREF; ++j) {
    const bool qin = t >= j, pin = NS + t >= j;
    const double aq = Wq[j], ap = Wp[j];
    // largest |entry| over the rows >= j and over the rows > j (the reflector vector differs from the column
    // only in row j, so its largest entry follows from the second without another reduction)
    double mge = qin ? cl_fabs(aq) : 0.0, mgt = t > j ? cl_fabs(aq) : 0.0;
    if (pin) mge = fmax(mge, cl_fabs(ap));
    if (NS + t > j) mgt = fmax(mgt, cl_fabs(ap));
    o.max2(mge, mgt);
    const double c = col_norm_m<NS>(o, mge, aq, ap, qin, pin);
    // A zero or non-finite column norm leaves the column as it is (no reflector, rho = c). The reflector's work runs
    // in any case on a stand-in norm and its results are selected, so that no branch splits the unrolled sweep
    // (measured on the C3 workload: N = 128 alone 460 -> 453 ms, the five streams and the persistent launch about
    // 1% faster).
    const bool ok = cl_isfinite(c) && c > 0.0;
    const double cc = ok ? c : 1.0;
    const double ajv = o.bcast(j < NS ? aq : ap, j % NS);
    const double sg = ajv < 0.0 ? -1.0 : 1.0;
    double vq = aq, vp = ap;
    if (j < NS) {
      if (t == j) vq = ajv + sg * cc;
    } else {
      if (t == j - NS) vp = ajv + sg * cc;
    }
    const double nv = col_norm_m<NS>(o, fmax(mgt, cl_fabs(ajv + sg * cc)), vq, vp, qin, pin);
    const double uq = ok && qin ? vq / nv : 0.0;
    const double up = ok && pin ? vp / nv : 0.0;
    {
      double w[KC];
#pragma unroll
      for (int l = 0; l < KC; ++l) w[l] = qin ? uq * Wq[l] + up * Wp[l] : (pin ? up * Wp[l] : 0.0);
      o.template sum<KC>(w, j + 1, KC - 1);
#pragma unroll
      for (int l = 0; l < KC; ++l)
        if (l > j) {
          const double nq = Wq[l] - (2.0 * uq) * w[l], np = Wp[l] - (2.0 * up) * w[l];
          Wq[l] = ok && qin ? nq : Wq[l];
          Wp[l] = ok && pin ? np : Wp[l];
        }
    }
    count += ok ? 1 : 0;
    const double rho = ok ? -sg * c : c;
    Us[(2 * j) * BLOCK + tid] = uq;
    Us[(2 * j + 1) * BLOCK + tid] = up;
    if (j < NS) {
      if (t == j) Wq[j] = rho;
    } else {
      if (t == j - NS) Wp[j] = rho;
    }
  }
  // the upper triangle of the swept matrix (rows < k), unsigned, packed by columns
  if (t < KC) {
#pragma unroll
    for (int l = 0; l < KC; ++l)
      if (l >= t) Ro[TR::at(t, l)] = Wq[l];
  }
  if (NS < KC && NS + t < KC) {
#pragma unroll
    for (int l = 0; l < KC; ++l)
      if (l >= NS + t) Ro[TR::at(NS + t, l)] = Wp[l];
  }
  osync<NS>();
  uint32_t neg = 0;
  int piv = 0;
  double gw_val = 0.0;
  r_checks<NS, KC>(o, Ro, Rio, neg, piv, gw_val);
  auto sgn = [&](int r) -> double { return (neg >> r) & 1u ? -1.0 : 1.0; };
  // Q = H_1 ... H_nref I[:, 1:k], reflectors in reverse order, into the frame registers
#pragma unroll
  for (int l = 0; l < KC; ++l) {
    Wq[l] = t == l ? 1.0 : 0.0;
    Wp[l] = NS + t == l ? 1.0 : 0.0;
  }
#pragma unroll
  for (int j = NREF - 1; j >= 0; --j) {
    const bool qin = t >= j, pin = NS + t >= j;
    const double uq = Us[(2 * j) * BLOCK + tid], up = Us[(2 * j + 1) * BLOCK + tid];
    double w[KC];
#pragma unroll
    for (int l = 0; l < KC; ++l) w[l] = qin ? uq * Wq[l] + up * Wp[l] : (pin ? up * Wp[l] : 0.0);
    o.template sum<KC>(w, j, KC - 1);
#pragma unroll
    for (int l = 0; l < KC; ++l)
      if (l >= j) {
        if (qin) Wq[l] = Wq[l] - (2.0 * uq) * w[l];
        if (pin) Wp[l] = Wp[l] - (2.0 * up) * w[l];
      }
  }
#pragma unroll
  for (int l = 0; l < KC; ++l) {
    Wq[l] = Wq[l] * sgn(l);
    Wp[l] = Wp[l] * sgn(l);
  }
  if (piv != 0) {
    const double pv = sgn(piv - 1) * Ro[TR::at(piv - 1, piv - 1)];
    qs.code = 1;
    qs.column = piv;
    qs.value = pv;
    qs.threshold = cl_isfinite(pv) ? 0.0 : __builtin_inf();
    qs.step = n;
  } else {
    // orth_loss: largest |Q'Q - I| over the entries l <= b (the matrix is symmetric entry for entry). Each entry
    // is its own butterfly sum, so only the upper triangle is formed, two columns to a call (b and KC - 2 - b; the
    // last column and, for even KC, the middle one alone): KC entries per call, k(k+1)/2 sums in all.
    double orth = 0.0, onan = 0.0;
    int nan_at = KC * KC;  // b KC + l of the first NaN in column order, the one Julia's max returns
    constexpr int NG = KC / 2 + 1;
#pragma unroll
    for (int g = 0; g < NG; ++g) {
      const int b1 = g < NG - 1 ? g : KC - 1;
      const int b2 = (g < NG - 1 && KC - 2 - g > g) ? KC - 2 - g : -1;
      const int ne = b1 + 1 + (b2 >= 0 ? b2 + 1 : 0);
      double v[KC];
      {
        const double qb = Wq[b1], pb = Wp[b1];
#pragma unroll
        for (int l = 0; l < KC; ++l)
          if (l <= b1) v[l] = cl_fma(Wp[l], pb, Wq[l] * qb);
      }
      if (b2 >= 0) {
        const double qb = Wq[b2], pb = Wp[b2];
#pragma unroll
        for (int l = 0; l < KC; ++l)
          if (l <= b2) v[b1 + 1 + l] = cl_fma(Wp[l], pb, Wq[l] * qb);
      }
      o.template sum<KC>(v, 0, ne - 1);
#pragma unroll
      for (int e = 0; e < KC; ++e)
        if (e < ne) {
          const int b = e <= b1 ? b1 : b2, l = e <= b1 ? e : e - b1 - 1;
          const double er = cl_fabs(l == b ? v[e] - 1.0 : v[e] - 0.0);
          if (cl_isnan(er) && b * KC + l < nan_at) {
            nan_at = b * KC + l;
            onan = er;
          }
          orth = fmax(orth, er);  // er >= +0, so starting from +0 equals starting from the first entry
        }
    }
    if (nan_at < KC * KC) orth = onan;  // Julia's max: NaN wins
    const double thr = a.tol * CL_EPS64 * (double)KC * 1.0;
    if (!(orth <= thr)) {
      qs.code = 4;
      qs.column = 0;
      qs.value = orth;
      qs.threshold = thr;
      qs.step = n;
    } else {
#pragma unroll
      for (int s2 = 0; s2 < SPT; ++s2) {
        const int cc = t + s2 * NS;
        if (cc < KC) sums[s2] = csum_add(sums[s2], cl_log(sgn(cc) * Ro[TR::at(cc, cc)]));
      }
      qs.n = n;
      qs.G_max = cl_jmax(qs.G_max, gw_val);
      qs.orth_max = cl_jmax(qs.orth_max, orth);
      if ((int64_t)count > qs.count_max) qs.count_max = count;
      qs.column = 0;
    }
  }
  osync<NS>();  // Ro / Rio / Us are rewritten at the next handoff
}

// ---------------------------------------------------------------- blocked Householder on the matrix cores
// The blocked order of the same factorisation (oracle_chain.hpp, RED_BLOCKED): two panels of 8 columns, each swept like
// qr_handoff over its own columns; after the first panel the trailing columns take its 8 reflectors at once in compact
// WY form, A <- A - U (T' (U' A)) (Schreiber and Van Loan, SIAM J. Sci. Stat. Comput. 10 (1989) 53-57); then
// Q = E - U (T U_top') over all reflectors and orth_loss from Q'Q. The sums over the 2N rows (the Gram matrices U'A,
// U'U, Q'Q) and the rank-8 and rank-16 updates run on the FP64 matrix instruction; the frame goes through LDS for them.
typedef double v4d __attribute__((ext_vector_type(4)));

// Row r, column c of an orbit's frame in LDS at r * 16 + (c ^ (r & 15)): the XOR keeps a thread's row, a 16-lane
// column operand and a tile's rows free of bank conflicts.
__device__ __attribute__((always_inline)) inline int xs_at(int r, int c) { return r * 16 + (c ^ (r & 15)); }

// D = A B + C on a 16 x 16 x 4 FP64 tile: lane l holds A[l % 16][l / 16], B[l / 16][l % 16] and D[l / 16 + 4 i][l % 16],
// i = 0..3. Each entry of D is the fma chain over k = 0..3 in order from C (measured on gfx950, bit for bit, including
// subnormal operands and signed zeros).
__device__ __attribute__((always_inline)) inline v4d mma4(double a, double b, v4d c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// 16 x 16 matrices in LDS (Gram, T, Y, M), column major with 17 doubles per column: a 16-lane access along a row
// then meets 16 different banks.
constexpr int MLD = 17;
constexpr int MSZ = 16 * MLD;

template <int NS>
struct Blk {
  static constexpr int D = 2 * NS;
  static constexpr int OPW = NS >= 64 ? 1 : 64 / NS;  // orbit slots per wave
  static constexpr int WROWS = NS > 64 ? 128 : D;      // rows of one orbit that one wave covers
  static constexpr int SLICES = WROWS / 4;
  static constexpr int TILES = WROWS / 16;
};

// The largest of v over the 64 lanes (every lane ends with it).
__device__ __attribute__((always_inline)) inline double wave_max64(double v) {
  double a, b;
  pair2<0>(v, a, b); v = fmax(a, b);
  pair2<1>(v, a, b); v = fmax(a, b);
  pair2<2>(v, a, b); v = fmax(a, b);
  pair2<3>(v, a, b); v = fmax(a, b);
  pair2<4>(v, a, b); v = fmax(a, b);
  pair2<5>(v, a, b); v = fmax(a, b);
  return v;
}

// LDS of the blocked handoff: Xs the frames (per slot 2N x 16, xs_at), Gs and Ts one 16 x 16 matrix per slot (column
// major), Rs the packed triangles, xb N = 128's exchange buffer (also the two waves' partial Gram matrices).
struct BlkLds {
  double* Xs;
  double* Gs;
  double* Ts;
  double* Rs;
  double* xb;
};

// Gram matrix X'X of each orbit slot of this wave (its rows: all 2N, or for N = 128 the wave's 128), summed over the
// two waves for N = 128 (a + b, the same bits in either wave).
template <int NS, int SLOTS>
__device__ __attribute__((always_inline)) inline void blk_gram(const BlkLds& L, v4d (&acc)[Blk<NS>::OPW]) {
  using B = Blk<NS>;
  const int lane = (int)threadIdx.x & 63, w = (int)threadIdx.x >> 6, g = lane >> 4, j = lane & 15;
#pragma unroll
  for (int o = 0; o < B::OPW; ++o) {
    const int so = NS > 64 ? 0 : w * B::OPW + o;
    const double* Xo = L.Xs + so * B::D * 16;
    const int r0 = NS > 64 ? 128 * w : 0;
    v4d c = {0.0, 0.0, 0.0, 0.0};
    if (so >= SLOTS) {  // a slot the group holds no LDS for (uniform over the wave)
      acc[o] = c;
      continue;
    }
#pragma unroll
    for (int s = 0; s < B::SLICES; ++s) {
      const int r = r0 + 4 * s + g;
      const double v = Xo[xs_at(r, j)];
      c = mma4(v, v, c);
    }
    acc[o] = c;
  }
  if constexpr (NS > 64) {
    lds_barrier();  // the exchange buffer's previous use is over
#pragma unroll
    for (int i = 0; i < 4; ++i) L.xb[(w * 4 + i) * 64 + lane] = acc[0][i];
    lds_barrier();
#pragma unroll
    for (int i = 0; i < 4; ++i) acc[0][i] = acc[0][i] + L.xb[((1 - w) * 4 + i) * 64 + lane];
  }
}

// Stores each slot's 16 x 16 tile (C layout) into its Gs matrix (one wave of N = 128 stores).
template <int NS, int SLOTS>
__device__ __attribute__((always_inline)) inline void blk_store(double* M0, const v4d (&acc)[Blk<NS>::OPW]) {
  using B = Blk<NS>;
  const int lane = (int)threadIdx.x & 63, w = (int)threadIdx.x >> 6, g = lane >> 4, j = lane & 15;
  if (NS > 64 && w != 0) return;
#pragma unroll
  for (int o = 0; o < B::OPW; ++o) {
    const int so = NS > 64 ? 0 : w * B::OPW + o;
    if (so >= SLOTS) continue;
    double* Mo = M0 + so * MSZ;
#pragma unroll
    for (int i = 0; i < 4; ++i) Mo[(g + 4 * i) + MLD * j] = acc[o][i];
  }
}

// Rows m = t, t + N (< 16) of the compact WY factor of the first NB reflectors from their Gram matrix (Gso):
// T[m][m] = 2, T[m][j] = -2 (fma chain over p < j of T[m][p] G[p][j] from +0) for j > m, zero below the diagonal.
template <int NS, int NB>
__device__ __attribute__((always_inline)) inline void blk_tmat(int t, const double* Gso, double* Tso) {
#pragma unroll
  for (int mm = 0; mm < (NS >= 16 ? 1 : 2); ++mm) {
    const int m = t + mm * NS;
    if (m < 16) {
      double Tr[NB];
#pragma unroll
      for (int p = 0; p < NB; ++p) Tr[p] = p == m ? 2.0 : 0.0;
#pragma unroll
      for (int jj = 1; jj < NB; ++jj) {
        double z = 0.0;
#pragma unroll
        for (int p = 0; p < jj; ++p) z = cl_fma(Tr[p], Gso[p + MLD * jj], z);
        Tr[jj] = jj > m ? -2.0 * z : Tr[jj];
      }
#pragma unroll
      for (int p = 0; p < NB; ++p) Tso[m + MLD * p] = Tr[p];
    }
  }
}

// One panel (columns J0 .. J1 - 1) of the reflector sweep, as in qr_handoff with the dot products and updates over the
// panel's columns only; u_j goes to the frame rows in LDS. `act` selects every register change (the orbits of a wave
// run the handoff together; an orbit without a factorisation keeps its registers).
template <int NS, int KC, int J0, int J1>
__device__ __attribute__((always_inline)) inline void blk_panel(const Orb<NS>& o, dvec<KC>& Wq, dvec<KC>& Wp, double* Xo,
                                                                bool act, int& count) {
  constexpr int D = 2 * NS;
  constexpr int NREF = KC < D ? KC : KC - 1;
  const int t = o.t;
#pragma unroll
  for (int j = J0; j < J1; ++j) {
    if (j >= NREF) {  // k = 2N: the last column has no reflector
      Xo[xs_at(t, j)] = 0.0;
      Xo[xs_at(NS + t, j)] = 0.0;
      continue;
    }
    const bool qin = t >= j, pin = NS + t >= j;
    const double aq = Wq[j], ap = Wp[j];
    // one reduction: the largest |entry| over the rows >= j, their plain sum of squares and the pivot entry a_j (its
    // owner adds it, the other sites +0)
    double mge = qin ? cl_fabs(aq) : 0.0;
    if (pin) mge = fmax(mge, cl_fabs(ap));
    double ss = qin ? aq * aq + ap * ap : (pin ? ap * ap : 0.0);
    double ajv = j < NS ? (t == j ? aq : 0.0) : (t == j - NS ? ap : 0.0);
    o.msum2(mge, ss, ajv);
    double c = mge;
    if (cl_isfinite(mge) && mge > 0.0) {
      if (mge >= 0x1p-400 && mge <= 0x1p400)
        c = cl_sqrt(ss);
      else
        c = col_norm_m<NS>(o, mge, aq, ap, qin, pin);  // safe_norm's scaled sum, for columns of extreme size
    }
    const bool ok = cl_isfinite(c) && c > 0.0;
    const double cc = ok ? c : 1.0;
    const double sg = ajv < 0.0 ? -1.0 : 1.0;
    double vq = aq, vp = ap;
    if (j < NS) {
      if (t == j) vq = ajv + sg * cc;
    } else {
      if (t == j - NS) vp = ajv + sg * cc;
    }
    const double nv = cl_sqrt((2.0 * cc) * (cc + cl_fabs(ajv)));  // |v| = sqrt(2 c (c + |a_j|))
    const double uq = ok && qin ? vq / nv : 0.0;
    const double up = ok && pin ? vp / nv : 0.0;
    if (j + 1 < J1) {
      double w[KC];
#pragma unroll
      for (int l = 0; l < KC; ++l) w[l] = qin ? uq * Wq[l] + up * Wp[l] : (pin ? up * Wp[l] : 0.0);
      o.template sum<KC>(w, j + 1, J1 - 1);
#pragma unroll
      for (int l = 0; l < KC; ++l)
        if (l > j && l < J1) {
          const double nq = Wq[l] - (2.0 * uq) * w[l], np = Wp[l] - (2.0 * up) * w[l];
          Wq[l] = act && ok && qin ? nq : Wq[l];
          Wp[l] = act && ok && pin ? np : Wp[l];
        }
    }
    count += ok ? 1 : 0;
    const double rho = ok ? -sg * c : c;
    Xo[xs_at(t, j)] = uq;
    Xo[xs_at(NS + t, j)] = up;
    if (j < NS) {
      if (act && t == j) Wq[j] = rho;
    } else {
      if (act && t == j - NS) Wp[j] = rho;
    }
  }
}

// One handoff of QRSpectrum(HouseholderQR()) in the blocked order, run by every orbit of the wave (N = 128: of the
// group) together; `act` marks the orbits that factorise, the others keep registers and state. The same steps and
// checks as qr_handoff: the frame is stored first (a failed handoff leaves it), then the sweep, R, its checks and
// window_Gw, Q and orth_loss, the log sums.
template <int NS, int KC, int BLOCK, int SLOTS>
__device__ __attribute__((always_inline)) inline void qr_handoff_blk(const Orb<NS>& o, const ChainArgs& a, int64_t n,
                                                                     dvec<KC>& Wq, dvec<KC>& Wp, double* Wo,
                                                                     CSum (&sums)[(KC + NS - 1) / NS], QRChainState& qs,
                                                                     const BlkLds& L, bool act_in) {
  static_assert(KC == 16, "the blocked handoff is built for k = 16");
  // one orbit per wave (N >= 64): the handoff only runs when that orbit factorises, and its registers need no guard
  const bool act = NS >= 64 ? true : act_in;
  using B = Blk<NS>;
  using TR = Tri<KC>;
  constexpr int D = 2 * NS;
  constexpr int SPT = (KC + NS - 1) / NS;
  const int t = o.t;
  const int slot = (int)threadIdx.x / NS;
  const int lane = (int)threadIdx.x & 63, w = (int)threadIdx.x >> 6, g = lane >> 4, jl = lane & 15;
  // Slots at or past SLOTS (the persistent launch's groups hold LDS for fewer orbits than they have lanes) hold no
  // orbit: their lanes read slot 0's matrices and write none. Every slot has its frame rows in Xs.
  const bool live = slot < SLOTS;
  const int sl = live ? slot : 0;
  double* Xo = L.Xs + slot * D * 16;  // this thread's orbit
  double* Gso = L.Gs + sl * MSZ;
  double* Tso = L.Ts + sl * MSZ;
  double* Ro = L.Rs + sl * TR::SIZE;
  if (act) {
#pragma unroll
    for (int c = 0; c < KC; ++c) {
      Wo[c * D + t] = Wq[c];
      Wo[c * D + NS + t] = Wp[c];
    }
  }
  int count = 0;
  // ---- panel 1 and the trailing columns in LDS; R's columns 0..7 (rows < 8) are final
  blk_panel<NS, KC, 0, 8>(o, Wq, Wp, Xo, act, count);
  if (live && t < 8) {
#pragma unroll
    for (int l = 0; l < 8; ++l)
      if (l >= t) Ro[TR::at(t, l)] = Wq[l];
  }
#pragma unroll
  for (int c = 8; c < KC; ++c) {
    Xo[xs_at(t, c)] = Wq[c];
    Xo[xs_at(NS + t, c)] = Wp[c];
  }
  osync<NS>();
  v4d acc[B::OPW];
  blk_gram<NS, SLOTS>(L, acc);  // [U_1 | A_trail]' [U_1 | A_trail]
  blk_store<NS, SLOTS>(L.Gs, acc);
  osync<NS>();
  if (live && t < 8) blk_tmat<NS, 8>(t, Gso, Tso);
  osync<NS>();
  // Y = T_1' W (W = U_1' A_trail, the Gram's upper right block) into Ts rows 0..7, columns 8..15
  if (NS <= 64 || w == 0) {
#pragma unroll
    for (int oo = 0; oo < B::OPW; ++oo) {
      const int so = NS > 64 ? 0 : w * B::OPW + oo;
      if (so >= SLOTS) continue;
      const double* Gq = L.Gs + so * MSZ;
      double* Tq = L.Ts + so * MSZ;
      v4d c = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int kk = 0; kk < 2; ++kk) c = mma4(Tq[(4 * kk + g) + MLD * jl], Gq[(4 * kk + g) + MLD * jl], c);
#pragma unroll
      for (int i = 0; i < 4; ++i)
        if (g + 4 * i < 8 && jl >= 8) Tq[(g + 4 * i) + MLD * jl] = c[i];
    }
  }
  osync<NS>();
  // A_trail <- A_trail - U_1 Y on 16-row tiles (columns 8..15 written back)
#pragma unroll
  for (int oo = 0; oo < B::OPW; ++oo) {
    const int so = NS > 64 ? 0 : w * B::OPW + oo;
    if (so >= SLOTS) continue;
    double* Xq = L.Xs + so * D * 16;
    const double* Tq = L.Ts + so * MSZ;
    const int r0 = NS > 64 ? 128 * w : 0;
    const double b0 = Tq[g + MLD * jl], b1 = Tq[(4 + g) + MLD * jl];
#pragma unroll
    for (int tt = 0; tt < B::TILES; ++tt) {
      const int rb = r0 + 16 * tt;
      const double a0 = -Xq[xs_at(rb + jl, g)], a1 = -Xq[xs_at(rb + jl, 4 + g)];
      v4d c;
#pragma unroll
      for (int i = 0; i < 4; ++i) c[i] = Xq[xs_at(rb + g + 4 * i, jl)];
      c = mma4(a0, b0, c);
      c = mma4(a1, b1, c);
      if (jl >= 8) {
#pragma unroll
        for (int i = 0; i < 4; ++i) Xq[xs_at(rb + g + 4 * i, jl)] = c[i];
      }
    }
  }
  osync<NS>();
#pragma unroll
  for (int c = 8; c < KC; ++c) {
    const double q = Xo[xs_at(t, c)], p = Xo[xs_at(NS + t, c)];
    Wq[c] = act ? q : Wq[c];
    Wp[c] = act ? p : Wp[c];
  }
  // ---- panel 2
  blk_panel<NS, KC, 8, KC>(o, Wq, Wp, Xo, act, count);
  // the rest of the upper triangle of the swept matrix (rows < k, columns 8..15), unsigned, packed by columns
  if (live && t < KC) {
#pragma unroll
    for (int l = 8; l < KC; ++l)
      if (l >= t) Ro[TR::at(t, l)] = Wq[l];
  }
  if (live && NS < KC && NS + t < KC) {
#pragma unroll
    for (int l = 8; l < KC; ++l)
      if (l >= NS + t) Ro[TR::at(NS + t, l)] = Wp[l];
  }
  osync<NS>();
  uint32_t neg = 0;
  int piv = 0;
  double gw_val = 0.0;
  r_checks<NS, KC>(o, Ro, Gso, neg, piv, gw_val, live);  // R^-1 in Gs, free until the next Gram
  auto sgn = [&](int r) -> double { return (neg >> r) & 1u ? -1.0 : 1.0; };
  // ---- Q = E - U M, M = T U_top'
  blk_gram<NS, SLOTS>(L, acc);  // U'U
  osync<NS>();                  // window_Gw's reads of Gs are over
  blk_store<NS, SLOTS>(L.Gs, acc);
  osync<NS>();
  if (live) blk_tmat<NS, KC>(t, Gso, Tso);
  osync<NS>();
  if (NS <= 64 || w == 0) {
#pragma unroll
    for (int oo = 0; oo < B::OPW; ++oo) {
      const int so = NS > 64 ? 0 : w * B::OPW + oo;
      if (so >= SLOTS) continue;
      const double* Tq = L.Ts + so * MSZ;
      const double* Xq = L.Xs + so * D * 16;
      v4d c = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int kk = 0; kk < 4; ++kk) c = mma4(Tq[jl + MLD * (4 * kk + g)], Xq[xs_at(jl, 4 * kk + g)], c);
      double* Gq = L.Gs + so * MSZ;
#pragma unroll
      for (int i = 0; i < 4; ++i) Gq[(g + 4 * i) + MLD * jl] = c[i];
    }
  }
  osync<NS>();
#pragma unroll
  for (int oo = 0; oo < B::OPW; ++oo) {
    const int so = NS > 64 ? 0 : w * B::OPW + oo;
    if (so >= SLOTS) continue;
    double* Xq = L.Xs + so * D * 16;
    const double* Gq = L.Gs + so * MSZ;
    const double* Rq = L.Rs + so * TR::SIZE;
    const double sj = Rq[TR::at(jl, jl)] < 0.0 ? -1.0 : 1.0;
    const int r0 = NS > 64 ? 128 * w : 0;
    double b[4];
#pragma unroll
    for (int kk = 0; kk < 4; ++kk) b[kk] = Gq[(4 * kk + g) + MLD * jl];
#pragma unroll
    for (int tt = 0; tt < B::TILES; ++tt) {
      const int rb = r0 + 16 * tt;
      double av[4];
#pragma unroll
      for (int kk = 0; kk < 4; ++kk) av[kk] = -Xq[xs_at(rb + jl, 4 * kk + g)];
      v4d c;
#pragma unroll
      for (int i = 0; i < 4; ++i) c[i] = rb + g + 4 * i == jl ? 1.0 : 0.0;
#pragma unroll
      for (int kk = 0; kk < 4; ++kk) c = mma4(av[kk], b[kk], c);
#pragma unroll
      for (int i = 0; i < 4; ++i) Xq[xs_at(rb + g + 4 * i, jl)] = c[i] * sj;
    }
  }
  osync<NS>();
#pragma unroll
  for (int c = 0; c < KC; ++c) {
    const double q = Xo[xs_at(t, c)], p = Xo[xs_at(NS + t, c)];
    Wq[c] = act ? q : Wq[c];
    Wp[c] = act ? p : Wp[c];
  }
  // ---- orth_loss: largest |Q'Q - I|, NaN first in column order (Julia's max)
  blk_gram<NS, SLOTS>(L, acc);
  double orth = 0.0;
#pragma unroll
  for (int oo = 0; oo < B::OPW; ++oo) {
    double e = 0.0;
    bool nan = false;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const double v = acc[oo][i];
      const double er = cl_fabs(g + 4 * i == jl ? v - 1.0 : v - 0.0);
      nan = nan || cl_isnan(er);
      e = fmax(e, er);  // er >= +0
    }
    double mo = wave_max64(e);
    if (__builtin_amdgcn_ballot_w64(nan) != 0) {
      int idx = 256;  // b * 16 + a of this lane's first NaN entry (a row, b column)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const double v = acc[oo][i];
        const double er = cl_fabs(g + 4 * i == jl ? v - 1.0 : v - 0.0);
        const int ix = jl * 16 + g + 4 * i;
        if (cl_isnan(er) && ix < idx) idx = ix;
      }
#pragma unroll
      for (int s2 = 1; s2 < 64; s2 <<= 1) {
        const int u = __shfl_xor(idx, s2, 64);
        idx = u < idx ? u : idx;
      }
      const int ai = idx & 15, bj = idx >> 4;
      const int it = ai >> 2;
      double mine = 0.0;
#pragma unroll
      for (int i = 0; i < 4; ++i)
        if (i == it) mine = cl_fabs(g + 4 * i == jl ? acc[oo][i] - 1.0 : acc[oo][i] - 0.0);
      mo = __shfl(mine, bj + 16 * (ai & 3), 64);
    }
    if (NS > 64 || slot - w * B::OPW == oo) orth = mo;
  }
  // ---- checks and sums (qr_spectrum.jl:56-78)
  if (act && live) {
    if (piv != 0) {
      const double pv = sgn(piv - 1) * Ro[TR::at(piv - 1, piv - 1)];
      qs.code = 1;
      qs.column = piv;
      qs.value = pv;
      qs.threshold = cl_isfinite(pv) ? 0.0 : __builtin_inf();
      qs.step = n;
    } else {
      const double thr = a.tol * CL_EPS64 * (double)KC * 1.0;
      if (!(orth <= thr)) {
        qs.code = 4;
        qs.column = 0;
        qs.value = orth;
        qs.threshold = thr;
        qs.step = n;
      } else {
#pragma unroll
        for (int s2 = 0; s2 < SPT; ++s2) {
          const int cc = t + s2 * NS;
          if (cc < KC) sums[s2] = csum_add(sums[s2], cl_log(sgn(cc) * Ro[TR::at(cc, cc)]));
        }
        qs.n = n;
        qs.G_max = cl_jmax(qs.G_max, gw_val);
        qs.orth_max = cl_jmax(qs.orth_max, orth);
        if ((int64_t)count > qs.count_max) qs.count_max = count;
        qs.column = 0;
      }
    }
  }
  osync<NS>();  // the LDS is rewritten at the next handoff
}

// The orbits of one thread block: orbit slots g0 .. g0 + nslots - 1 of the launch order (slot = thread / N), each
// run from the launch's step n0 for `steps` steps or, with per-orbit step counts, to min(N_i, n0 + steps); an orbit
// whose count is reached (N_i <= n0) or that failed before n0 is retired and left untouched. LDS: xb (N = 128 only,
// 128 k doubles), Us (2 NREF BLOCK doubles), Rs and Ris (one packed triangle per orbit slot each).
template <int NS, int KC, bool RING, bool HASQR, int BLOCK, bool BLK = false, int SLOTS = Geo<NS>::OPB,
          class FT = double>
__device__ __attribute__((always_inline)) inline void chain_orbits(const ChainArgs& a, int64_t g0, int nslots,
                                                                   double* xb, double* Us, double* Rs, double* Ris,
                                                                   double* Gs = nullptr, double* Ts = nullptr) {
  constexpr int D = 2 * NS;
  constexpr int SPT = (KC + NS - 1) / NS;  // QR sums owned per thread
  constexpr int WORDS = 2 * KC + 9;
  using TR = Tri<KC>;

  Orb<NS> o;
  o.xb = xb;
  o.t = (int)threadIdx.x % NS;
  o.base = ((int)threadIdx.x & 63) - ((int)threadIdx.x & 63) % NS;
  const int t = o.t;
  const int slot = (int)threadIdx.x / NS;
  double* Ro = Rs + (HASQR ? slot * TR::SIZE : 0);
  double* Rio = Ris + (HASQR ? slot * TR::SIZE : 0);
  const int64_t g = g0 + slot;
  const bool exists = slot < nslots && g < a.m;
  const int64_t i = exists ? (a.order ? a.order[g] : g) : 0;
  const int64_t Ni = exists ? (a.Ns ? a.Ns[i] : a.N) : 0;
  bool active = exists && (a.n0 == 0 || a.code[i] == STATUS_OK) && (!a.Ns || Ni > a.n0);
  const int64_t nend = a.Ns ? (Ni < a.n0 + a.steps ? Ni : a.n0 + a.steps) : a.n0 + a.steps;
  const int64_t Hi = active ? (nend - a.n0) / a.tau : 0;
  // handoffs of the group: every lane of a 64-lane group (or the 128-thread group) loops the same count
  int64_t Hmax = Hi;
  if constexpr (NS <= 64) {
#pragma unroll
    for (int s = 32; s >= 1; s >>= 1) {
      int64_t v = __shfl_xor(Hmax, s, 64);
      Hmax = v > Hmax ? v : Hmax;
    }
  }

  double* xo = a.x + i * D;
  double* Wo = a.W + i * (int64_t)D * KC;
  uint64_t* so = a.st + i * (int64_t)WORDS;
  double q = 0.0, p = 0.0;
  // Float32 tangents: the frame is rounded once at the launch start and carried in Float32 between handoffs
  using FV = frame_vec<FT, KC>;
  FV Wq, Wp;
#pragma unroll
  for (int c = 0; c < KC; ++c) Wq[c] = Wp[c] = (FT)0;
  CSum sums[SPT];
  QRChainState qs;
  qs.n = qs.count_max = qs.column = qs.step = 0;
  qs.G_max = qs.orth_max = qs.value = qs.threshold = 0.0;
  qs.code = 0;
#pragma unroll
  for (int s = 0; s < SPT; ++s) sums[s] = csum_zero();
  if (active) {
    q = xo[t];
    p = xo[NS + t];
#pragma unroll
    for (int c = 0; c < KC; ++c) {
      Wq[c] = (FT)Wo[c * D + t];
      Wp[c] = (FT)Wo[c * D + NS + t];
    }
    if (HASQR && a.n0 > 0) {
#pragma unroll
      for (int s = 0; s < SPT; ++s) {
        int c = t + s * NS;
        if (c < KC) sums[s] = CSum{__longlong_as_double(so[2 * c]), __longlong_as_double(so[2 * c + 1])};
      }
      qs.n = (int64_t)so[2 * KC];
      qs.G_max = __longlong_as_double(so[2 * KC + 1]);
      qs.orth_max = __longlong_as_double(so[2 * KC + 2]);
      qs.count_max = (int64_t)so[2 * KC + 3];
      qs.code = (int8_t)(so[2 * KC + 4] & 0xff);
      qs.column = (int64_t)so[2 * KC + 5];
      qs.value = __longlong_as_double(so[2 * KC + 6]);
      qs.threshold = __longlong_as_double(so[2 * KC + 7]);
      qs.step = (int64_t)so[2 * KC + 8];
    }
  }
  const bool hasL = RING || t > 0, hasR = RING || t < NS - 1;
  int64_t n = a.n0;
  int8_t code = STATUS_OK;
  bool done = !active, qr_failed = false;

  for (int64_t h = 0; h < Hmax; ++h) {
    const bool run = !done && h < Hi;
    if (NS > 64 && !run) break;  // the 128-thread group is one orbit: uniform
    bool want = false;           // blocked order: this orbit factorises at the handoff below
    if (run) {
      for (int s = 0; s < a.tau; ++s) {
        // ---- tangent at the old point, then the step (orbit.jl:52-53)
        const double qr0 = o.right1(q);
        double sq, cq, sb, cb;
        cl_sincos(q, &sq, &cq);
        cl_sincos(q - qr0, &sb, &cb);  // bond to the right neighbour
        double bl[2];  // the left neighbour's right bond: sin(q - q_l) = -bl[0], cos(q - q_l) = bl[1]
        o.left2(sb, cb, bl[0], bl[1]);
        double gq = a.K * cq;
        double cL = hasL ? a.eps * bl[1] : 0.0;
        double cR = hasR ? a.eps * cb : 0.0;
        double dg = (gq + cL) + cR, gl = -cL, gr = -cR;
        // the band's entries in the frame's type (formed in FP64, rounded once), the product in that type
        const FT dgf = (FT)dg, glf = (FT)gl, grf = (FT)gr;
        FV wl, wr;
        o.template nbr<KC, FV, FT>(Wq, wl, wr);
#pragma unroll
        for (int c = 0; c < KC; ++c) {
          FT gw = dgf * Wq[c];
          if (hasL) gw = gw + glf * wl[c];
          if (hasR) gw = gw + grf * wr[c];
          FT np = gw + Wp[c];
          Wq[c] = (Wq[c] + gw) + Wp[c];
          Wp[c] = np;
        }
        double fs = 0.0;
        if (hasL) fs = fs + (-bl[0]);
        if (hasR) fs = fs + sb;
        double pn = p + (a.K * sq + a.eps * fs);
        q = q + pn;
        p = pn;
      }
      n += a.tau;
      // ---- status (orbit.jl:1-3,55-56)
      int badx = !(cl_isfinite(q) && cl_isfinite(p)), badw = 0;
#pragma unroll
      for (int c = 0; c < KC; ++c) badw |= !(cl_isfinite(Wq[c]) && cl_isfinite(Wp[c]));
      {
        const int f = o.any(badx | (badw << 1));
        badx = f & 1;
        badw = f >> 1;
      }
      if (badx || badw) {
        code = badx ? STATUS_STATE : STATUS_W;
        done = true;
      } else if (HASQR && qs.code == 0) {
        if constexpr (BLK) {
          want = true;
        } else {
          if constexpr (std::is_same<FT, double>::value) {
            qr_handoff<NS, KC, BLOCK>(o, a, n, Wq, Wp, Wo, sums, qs, Us, Ro, Rio);
          } else {  // the factorisation in FP64 on the exactly widened frame, its Q rounded back
            dvec<KC> Wqd, Wpd;
#pragma unroll
            for (int c = 0; c < KC; ++c) {
              Wqd[c] = (double)Wq[c];
              Wpd[c] = (double)Wp[c];
            }
            qr_handoff<NS, KC, BLOCK>(o, a, n, Wqd, Wpd, Wo, sums, qs, Us, Ro, Rio);
#pragma unroll
            for (int c = 0; c < KC; ++c) {
              Wq[c] = (FT)Wqd[c];
              Wp[c] = (FT)Wpd[c];
            }
          }
          if (qs.code != 0) {
            code = STATUS_REDUCER;
            done = true;
            qr_failed = true;  // the frame before this handoff's factorisation is already in W
          }
        }
      }
    }
    if constexpr (BLK && HASQR) {
      // the matrix instructions need the whole wave: the orbits of a wave factorise together (N = 128: the group's
      // one orbit, whose two waves agree on `want`)
      if (NS > 64 ? want : __builtin_amdgcn_ballot_w64(want) != 0) {
        BlkLds L{Us, Gs, Ts, Rs, xb};
        if constexpr (std::is_same<FT, double>::value) {
          qr_handoff_blk<NS, KC, BLOCK, SLOTS>(o, a, n, Wq, Wp, Wo, sums, qs, L, want);
        } else {  // the factorisation in FP64 on the exactly widened frame, its Q rounded back
          dvec<KC> Wqd, Wpd;
#pragma unroll
          for (int c = 0; c < KC; ++c) {
            Wqd[c] = (double)Wq[c];
            Wpd[c] = (double)Wp[c];
          }
          qr_handoff_blk<NS, KC, BLOCK, SLOTS>(o, a, n, Wqd, Wpd, Wo, sums, qs, L, want);
#pragma unroll
          for (int c = 0; c < KC; ++c) {
            Wq[c] = (FT)Wqd[c];
            Wp[c] = (FT)Wpd[c];
          }
        }
        if (want && qs.code != 0) {
          code = STATUS_REDUCER;
          done = true;
          qr_failed = true;
        }
      }
    }
  }
  if (!active) return;
  if (!done) {
    int badx = !(cl_isfinite(q) && cl_isfinite(p)), badw = 0;
#pragma unroll
    for (int c = 0; c < KC; ++c) badw |= !(cl_isfinite(Wq[c]) && cl_isfinite(Wp[c]));
    {
      const int f = o.any(badx | (badw << 1));
      badx = f & 1;
      badw = f >> 1;
    }
    code = n == Ni ? (badx ? STATUS_STATE : (badw ? STATUS_W : STATUS_OK)) : STATUS_OK;
  }
  xo[t] = q;
  xo[NS + t] = p;
  if (!qr_failed) {
#pragma unroll
    for (int c = 0; c < KC; ++c) {
      Wo[c * D + t] = (double)Wq[c];
      Wo[c * D + NS + t] = (double)Wp[c];
    }
  }
  if (HASQR) {
#pragma unroll
    for (int s = 0; s < SPT; ++s) {
      int c = t + s * NS;
      if (c < KC) {
        so[2 * c] = (uint64_t)__double_as_longlong(sums[s].s);
        so[2 * c + 1] = (uint64_t)__double_as_longlong(sums[s].c);
      }
    }
    if (t == 0) {
      so[2 * KC] = (uint64_t)qs.n;
      so[2 * KC + 1] = (uint64_t)__double_as_longlong(qs.G_max);
      so[2 * KC + 2] = (uint64_t)__double_as_longlong(qs.orth_max);
      so[2 * KC + 3] = (uint64_t)qs.count_max;
      so[2 * KC + 4] = (uint64_t)(uint8_t)qs.code;
      so[2 * KC + 5] = (uint64_t)qs.column;
      so[2 * KC + 6] = (uint64_t)__double_as_longlong(qs.value);
      so[2 * KC + 7] = (uint64_t)__double_as_longlong(qs.threshold);
      so[2 * KC + 8] = (uint64_t)qs.step;
    }
  }
  if (t == 0) {
    a.code[i] = code;
    a.hstep[i] = n;
  }
}

// N >= 64: ask the allocator for two waves per SIMD (256 registers, a few spilled): measured on the C3 workload,
// N = 128 925 ms against 1001 ms and N = 64 552 against 573; at N = 32 the spills cost more than the second
// wave (484 against 453), so smaller N keep the whole register file.
template <int NS, int KC, bool RING, bool HASQR, class FT>
__global__ void __launch_bounds__(Geo<NS>::BLOCK) __attribute__((amdgpu_waves_per_eu(NS >= 64 ? 2 : 1)))
chain_kernel(ChainArgs a) {
  constexpr int D = 2 * NS;
  constexpr int NREF = KC < D ? KC : KC - 1;
  constexpr int OPB = Geo<NS>::OPB;
  constexpr int BLOCK = Geo<NS>::BLOCK;
  __shared__ double xb[NS > 64 ? 128 * KC : 1];
  __shared__ double Us[HASQR ? 2 * NREF * BLOCK : 1];
  __shared__ double Rs[HASQR ? OPB * Tri<KC>::SIZE : 1];
  __shared__ double Ris[HASQR ? OPB * Tri<KC>::SIZE : 1];
  chain_orbits<NS, KC, RING, HASQR, BLOCK, false, OPB, FT>(a, (int64_t)blockIdx.x * OPB, OPB, xb, Us, Rs, Ris);
}

// The blocked order on the matrix cores (qr_handoff_blk). LDS per group: the frames (2N x 16 per orbit, 32 BLOCK
// doubles), two 16 x 16 matrices and one packed triangle per orbit, and N = 128's exchange buffer.
// N >= 64 in the blocked order: one wave per SIMD (the whole register file, no spills). Measured on the C3 workload
// (N = 128 alone 356 -> 235 ms, the five lattices on five streams 532 -> 485 ms) against two waves per SIMD at 256
// registers with 0.6 KB spilled per lane.
#ifndef CL_BLK_WAVES
#define CL_BLK_WAVES 1
#endif
template <int NS, int KC, bool RING, class FT>
__global__ void __launch_bounds__(Geo<NS>::BLOCK) __attribute__((amdgpu_waves_per_eu(NS >= 64 ? CL_BLK_WAVES : 1)))
chain_kernel_blk(ChainArgs a) {
  constexpr int OPB = Geo<NS>::OPB;
  constexpr int BLOCK = Geo<NS>::BLOCK;
  __shared__ double xb[NS > 64 ? 128 * KC : 1];
  __shared__ double Xs[32 * BLOCK];
  __shared__ double Gs[OPB * MSZ];
  __shared__ double Ts[OPB * MSZ];
  __shared__ double Rs[OPB * Tri<KC>::SIZE];
  chain_orbits<NS, KC, RING, true, BLOCK, true, OPB, FT>(a, (int64_t)blockIdx.x * OPB, OPB, xb, Xs, Rs, nullptr, Gs,
                                                         Ts);
}

template <int NS, int KC, bool RING, bool HASQR, class FT>
int launch_chain(const ChainArgs& a, hipStream_t st, bool blocked) {
  constexpr int B = Geo<NS>::BLOCK, OPB = Geo<NS>::OPB;
  int64_t grid = (a.m + OPB - 1) / OPB;
  if (HASQR && blocked)
    hipLaunchKernelGGL((chain_kernel_blk<NS, KC, RING, FT>), dim3((unsigned)grid), dim3(B), 0, st, a);
  else
    hipLaunchKernelGGL((chain_kernel<NS, KC, RING, HASQR, FT>), dim3((unsigned)grid), dim3(B), 0, st, a);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : (e == hipErrorOutOfMemory ? CL_ERR_OOM : CL_ERR_DEVICE_LOST);
}

template <int NS, class FT>
int launch_n(const ChainArgs& a, bool ring, bool qr, bool blocked, hipStream_t st) {
  if (ring)
    return qr ? launch_chain<NS, 16, true, true, FT>(a, st, blocked) : launch_chain<NS, 16, true, false, FT>(a, st, false);
  return qr ? launch_chain<NS, 16, false, true, FT>(a, st, blocked) : launch_chain<NS, 16, false, false, FT>(a, st, false);
}

}  // namespace chain_detail

}  // namespace cl
