// ORACLE - test infrastructure only (included by oracle.cpp). The coupled standard-map chain and its
// QR spectrum with Householder reflectors, for frames of d = 2N rows and up to 16 columns.
//
// What it restates (paths relative to the Chaoslab repository root):
//   StandardMapChain forward   src/Systems/standard_map_chain.jl:34-41 (_force), :68-76 / chain_tangent.jl:79-86
//   ChainTangent * W           src/Systems/chain_tangent.jl:21-64 (_band, the banded product), standard_map_chain.jl:44-51
//                              (_grow) and chain_tangent.jl:90-97 (_chain_tangent)
//   HouseholderQR              src/Lyapunov/householder.jl:17-65 (_reflector, _reflect, _hh_sweep, _form_q, _finish)
//   QRSpectrum step            src/Lyapunov/qr_spectrum.jl:56-78; orth_loss :4; window_Gw orthonormalize.jl:112-127
//   orbit body                 src/Tangent/orbit.jl:44-60
//
// Reductions over the 2N rows of a frame column come in two orders:
//   RED_ROWS  (0): the reference's left fold over the rows in row order (q rows 1..N, then p rows), the dot and sum
//                  of a StaticArrays vector;
//   RED_SITES (1): the order of the device kernel, which gives each chain site one thread: a partial per site over its
//                  q row (when that row is in range) then its p row, then a butterfly over the N sites,
//                  v[t] <- v[t] + v[t ^ o] for o = 1, 2, 4, ..., N/2 (pairs of sites first), whose result every site
//                  holds (IEEE addition commutes). In this order rows that hold structural zeros (rows above the reflector's pivot, and
//                  columns a reflector leaves unchanged) are skipped; the results differ from RED_ROWS by rounding only.
//   RED_BLOCKED (2): the blocked Householder factorisation of the device's matrix-core path (two panels of 8 columns):
//                  each panel is swept in the RED_SITES order over its own columns only, each column's norm from one
//                  reduction (chain_panel); after the first panel the
//                  trailing 8 columns take the panel's reflectors at once in compact WY form, A <- A - U (T' (U'A))
//                  (Schreiber and Van Loan, "A storage-efficient WY representation for products of Householder
//                  transformations", SIAM J. Sci. Stat. Comput. 10 (1989) 53-57); Q = E - U (T U_top') over all
//                  reflectors, and orth_loss from the Gram matrix Q'Q. Every product that sums over the 2N rows is a
//                  fused multiply-add chain over the rows in order from +0 (the FP64 matrix instruction's own order:
//                  four fma per instruction, k = 0..3), one chain per 128 rows, the chains' results added in order.
//                  Mathematically the same factorisation; the results differ from RED_ROWS by rounding only.
// The reference's left fold is the restatement; RED_SITES and RED_BLOCKED are what the device kernels compute bit for
// bit.

enum { RED_ROWS = 0, RED_SITES = 1, RED_BLOCKED = 2 };

template <class T>
static T butterfly(std::vector<T>& v) {
  int N = (int)v.size();
  std::vector<T> w(N);
  for (int o = 1; o < N; o *= 2) {
    for (int t = 0; t < N; ++t) w[t] = v[t] + v[t ^ o];
    v.swap(w);
  }
  return v[0];
}

// sum over rows r >= r0 of terms[r] (terms already formed per row)
template <class T>
static T chain_sum(const T* terms, int N, int r0, int mode) {
  int d = 2 * N;
  if (mode == RED_ROWS) {
    T s = terms[r0];
    for (int r = r0 + 1; r < d; ++r) s = s + terms[r];
    return s;
  }
  std::vector<T> v(N);
  for (int t = 0; t < N; ++t) v[t] = t >= r0 ? terms[t] + terms[N + t] : terms[N + t];
  return butterfly(v);
}

// safe_norm of rows r0..d-1 of a column (norm.jl:8-12): largest |entry|, then m sqrt(sum (a/m)^2)
template <class P, class T>
static T chain_snorm(const T* a, int N, int r0, int mode) {
  int d = 2 * N;
  T m = P::abs(a[r0]);
  for (int r = r0 + 1; r < d; ++r) m = jmax<P>(m, P::abs(a[r]));
  if (!(P::finite(m) && m > 0)) return m;
  std::vector<T> t(d);
  for (int r = r0; r < d; ++r) {
    T b = a[r] / m;
    t[r] = b * b;
  }
  return m * P::sqrt(chain_sum(t.data(), N, r0, mode));
}

template <class T>
static T chain_dot(const T* a, const T* b, int N, int r0, int mode) {
  int d = 2 * N;
  std::vector<T> t(d);
  for (int r = r0; r < d; ++r) t[r] = a[r] * b[r];
  return chain_sum(t.data(), N, r0, mode);
}

template <class T>
struct ChainFact {
  std::vector<T> Q, R;  // d x k and k x k, column-major
  int count, status, column;
};

// HouseholderQR of a d x k frame A (column-major; overwritten).
template <class P, class T>
static void chain_householder(T* A, int N, int k, int mode, ChainFact<T>& f) {
  const int d = 2 * N;
  const int nref = k < d ? k : k - 1;
  std::vector<T> U((size_t)d * k, (T)0), a(d), v(d), u(d);
  int count = 0;
  for (int j = 0; j < nref; ++j) {
    T* col = A + (size_t)j * d;
    T c, rho;
    bool ok;
    if (mode == RED_ROWS) {
      for (int r = 0; r < d; ++r) a[r] = r >= j ? col[r] : (T)0;
      c = chain_snorm<P>(a.data(), N, 0, mode);
    } else {
      for (int r = 0; r < d; ++r) a[r] = col[r];
      c = chain_snorm<P>(a.data(), N, j, mode);
    }
    ok = P::finite(c) && c > 0;
    if (ok) {
      T s = a[j] < 0 ? (T)-1 : (T)1;
      for (int r = 0; r < d; ++r) v[r] = a[r];
      v[j] = a[j] + s * c;
      int r0 = mode == RED_ROWS ? 0 : j;
      T nv = chain_snorm<P>(v.data(), N, r0, mode);
      for (int r = 0; r < d; ++r) u[r] = r >= r0 ? v[r] / nv : (T)0;
      int l0 = mode == RED_ROWS ? 0 : j + 1;
      for (int l = l0; l < k; ++l) {
        T* Al = A + (size_t)l * d;
        T w = chain_dot(u.data(), Al, N, r0, mode);
        for (int r = r0; r < d; ++r) Al[r] = Al[r] - ((T)2 * u[r]) * w;
      }
      for (int r = 0; r < d; ++r) U[(size_t)j * d + r] = u[r];
      ++count;
      rho = -s * c;
    } else {
      rho = c;
    }
    col[j] = rho;
  }
  // Q = H_1 ... H_nref applied to I[:, 1:k], reflectors in reverse order
  std::vector<T> E((size_t)d * k, (T)0);
  for (int l = 0; l < k; ++l) E[(size_t)l * d + l] = (T)1;
  for (int j = nref - 1; j >= 0; --j) {
    const T* uj = U.data() + (size_t)j * d;
    int r0 = mode == RED_ROWS ? 0 : j, l0 = mode == RED_ROWS ? 0 : j;
    for (int l = l0; l < k; ++l) {
      T* El = E.data() + (size_t)l * d;
      T w = chain_dot(uj, El, N, r0, mode);
      for (int r = r0; r < d; ++r) El[r] = El[r] - ((T)2 * uj[r]) * w;
    }
  }
  // sign fix and status (householder.jl:55-65)
  f.Q.assign((size_t)d * k, (T)0);
  f.R.assign((size_t)k * k, (T)0);
  std::vector<T> sg(k);
  for (int j = 0; j < k; ++j) sg[j] = A[(size_t)j * d + j] < 0 ? (T)-1 : (T)1;
  for (int j = 0; j < k; ++j)
    for (int i = 0; i <= j; ++i) f.R[(size_t)j * k + i] = sg[i] * A[(size_t)j * d + i];
  for (int j = 0; j < k; ++j)
    for (int r = 0; r < d; ++r) f.Q[(size_t)j * d + r] = E[(size_t)j * d + r] * sg[j];
  f.count = count;
  f.status = 0;
  f.column = 0;
  for (int j = 0; j < k; ++j) {
    T p = f.R[(size_t)j * k + j];
    if (!(P::finite(p) && p > 0)) {
      f.status = 1;
      f.column = j + 1;
      break;
    }
  }
}

// ---------------------------------------------------------------- RED_BLOCKED
// Gram matrix G = X'X of the first k columns of a d x k column-major X: G[i][j] (i + 16 j) an fma chain over the rows
// in order from +0, one chain per 128 rows, the chains' results added in order. G[i][j] equals G[j][i] bit for bit.
template <class P, class T>
static void chain_gram(const T* X, int d, int k, T* G) {
  for (int j = 0; j < 16; ++j)
    for (int i = 0; i < 16; ++i) {
      T tot = (T)0;
      for (int rb = 0; rb < d; rb += 128) {
        T acc = (T)0;
        if (i < k && j < k)
          for (int r = rb; r < rb + 128 && r < d; ++r) acc = P::fma(X[(size_t)i * d + r], X[(size_t)j * d + r], acc);
        tot = rb == 0 ? acc : tot + acc;
      }
      G[i + 16 * j] = tot;
    }
}

// Upper-triangular factor of the compact WY form of H_0 H_1 ... H_{nb-1}, H_j = I - 2 u_j u_j' (a reflector that was
// not applied has u_j = 0): T[j][j] = 2, T[m][j] = -2 (sum over p < j of T[m][p] G[p][j]), the sum an fma chain in
// p order from +0 that includes the zeros T[m][p], p < m (they leave the chain at +0: G is finite); zero below the
// diagonal. T (i + 16 j) for i, j < nb.
template <class P, class T>
static void chain_tmat(const T* G, int nb, T* Tm) {
  for (int e = 0; e < 256; ++e) Tm[e] = (T)0;
  for (int j = 0; j < nb; ++j) {
    Tm[j + 16 * j] = (T)2;
    for (int m = 0; m < j; ++m) {
      T z = (T)0;
      for (int p = 0; p < j; ++p) z = P::fma(Tm[m + 16 * p], G[p + 16 * j], z);
      Tm[m + 16 * j] = (T)(g_defect == DEFECT_BLK_TSIGN ? 2 : -2) * z;
    }
  }
}

// One panel of the sweep (columns j0 .. j1 - 1), the reflectors applied to the panel's columns only; writes u_j into U
// and rho into the diagonal. The column norm c: m the largest |entry| over rows >= j; for m in [2^-400, 2^400] the
// square root of the plain sum of squares (no square overflows or falls below the normal range), in the RED_SITES
// order, else safe_norm (norm.jl:8-12) in the RED_SITES order; ||v|| of v = a + sign(a_j) c e_j from its closed form
// sqrt(2c (c + |a_j|)), so one reduction gives m, the sum and a_j together. The same reflector as the reference's to
// rounding.
template <class P, class T>
static void chain_panel(T* A, int N, int j0, int j1, int nref, T* U, int& count) {
  const int d = 2 * N;
  std::vector<T> sq(d), u(d);
  for (int j = j0; j < j1 && j < nref; ++j) {
    T* col = A + (size_t)j * d;
    T m = P::abs(col[j]);
    for (int r = j + 1; r < d; ++r) m = jmax<P>(m, P::abs(col[r]));
    T c;
    if (!(P::finite(m) && m > 0)) {
      c = m;
    } else if (m >= (T)0x1p-400 && m <= (T)0x1p400) {
      for (int r = 0; r < d; ++r) sq[r] = r >= j ? col[r] * col[r] : (T)0;
      c = P::sqrt(chain_sum(sq.data(), N, j, RED_SITES));
    } else {
      c = chain_snorm<P>(col, N, j, RED_SITES);
    }
    T rho;
    if (P::finite(c) && c > 0) {
      T s = col[j] < 0 ? (T)-1 : (T)1;
      const T vj = col[j] + s * c;
      const T nv = P::sqrt(((T)2 * c) * (c + P::abs(col[j])));
      for (int r = 0; r < d; ++r) u[r] = r > j ? col[r] / nv : (r == j ? vj / nv : (T)0);
      for (int l = j + 1; l < j1; ++l) {
        T* Al = A + (size_t)l * d;
        T w = chain_dot(u.data(), Al, N, j, RED_SITES);
        for (int r = j; r < d; ++r) Al[r] = Al[r] - ((T)2 * u[r]) * w;
      }
      for (int r = 0; r < d; ++r) U[(size_t)j * d + r] = u[r];
      ++count;
      rho = -s * c;
    } else {
      rho = c;
    }
    col[j] = rho;
  }
}

// HouseholderQR of a d x 16 frame A (column-major; overwritten) in the RED_BLOCKED order.
template <class P, class T>
static void chain_householder_blocked(T* A, int N, int k, ChainFact<T>& f) {
  const int d = 2 * N;
  const int nref = k < d ? k : k - 1;
  std::vector<T> U((size_t)d * k, (T)0), X((size_t)d * k);
  T G[256], Tm[256], Y[256], M[256];
  int count = 0;
  // panel 1, then the trailing columns 8..15 take H_7 ... H_0 = I - U T' U' at once: A <- A - U (T' W), W = U'A
  chain_panel<P>(A, N, 0, 8, nref, U.data(), count);
  for (int c = 0; c < k; ++c)
    for (int r = 0; r < d; ++r) X[(size_t)c * d + r] = c < 8 ? U[(size_t)c * d + r] : A[(size_t)c * d + r];
  chain_gram<P>(X.data(), d, k, G);
  chain_tmat<P>(G, 8, Tm);
  for (int m = 0; m < 8; ++m)
    for (int c = 8; c < k; ++c) {
      T y = (T)0;
      for (int p = 0; p < 8; ++p)  // T' W, the zeros included
        y = P::fma(g_defect == DEFECT_BLK_TNOTRANS ? Tm[m + 16 * p] : Tm[p + 16 * m], G[p + 16 * c], y);
      Y[m + 16 * c] = y;
    }
  for (int c = 8; c < k; ++c)
    for (int r = 0; r < d; ++r) {
      T acc = A[(size_t)c * d + r];
      for (int m = 0; m < 8; ++m) acc = P::fma(-U[(size_t)m * d + r], Y[m + 16 * c], acc);
      A[(size_t)c * d + r] = acc;
    }
  // panel 2
  chain_panel<P>(A, N, 8, k, nref, U.data(), count);
  // Q = H_0 ... H_15 E = E - U (T U_top'), U_top the first 16 rows of U
  chain_gram<P>(U.data(), d, k, G);
  chain_tmat<P>(G, k, Tm);
  for (int m = 0; m < k; ++m)
    for (int c = 0; c < k; ++c) {
      T acc = (T)0;
      for (int p = 0; p < k; ++p) acc = P::fma(Tm[m + 16 * p], U[(size_t)p * d + c], acc);  // the zeros included
      M[m + 16 * c] = acc;
    }
  // sign fix and status (householder.jl:55-65)
  f.Q.assign((size_t)d * k, (T)0);
  f.R.assign((size_t)k * k, (T)0);
  std::vector<T> sg(k);
  for (int j = 0; j < k; ++j) sg[j] = A[(size_t)j * d + j] < 0 ? (T)-1 : (T)1;
  for (int j = 0; j < k; ++j)
    for (int i = 0; i <= j; ++i) f.R[(size_t)j * k + i] = sg[i] * A[(size_t)j * d + i];
  for (int c = 0; c < k; ++c)
    for (int r = 0; r < d; ++r) {
      T acc = r == c ? (T)1 : (T)0;
      for (int m = 0; m < k; ++m) acc = P::fma(-U[(size_t)m * d + r], M[m + 16 * c], acc);
      f.Q[(size_t)c * d + r] = acc * sg[c];
    }
  f.count = count;
  f.status = 0;
  f.column = 0;
  for (int j = 0; j < k; ++j) {
    T p = f.R[(size_t)j * k + j];
    if (!(P::finite(p) && p > 0)) {
      f.status = 1;
      f.column = j + 1;
      break;
    }
  }
}

// largest |Q'Q - I| (qr_spectrum.jl:4): entries accumulated with fma in row order, per site then the butterfly, or
// (RED_BLOCKED) the Gram matrix's chains
template <class P, class T>
static T chain_orth_loss(const T* Q, int N, int k, int mode) {
  const int d = 2 * N;
  T mx = 0;
  bool first = true;
  for (int b = 0; b < k; ++b)
    for (int a = 0; a < k; ++a) {
      const T *qa = Q + (size_t)a * d, *qb = Q + (size_t)b * d;
      T acc;
      if (mode == RED_BLOCKED) {
        acc = (T)0;
        for (int rb = 0; rb < d; rb += 128) {
          T part = (T)0;
          for (int r = rb; r < rb + 128 && r < d; ++r) part = P::fma(qa[r], qb[r], part);
          acc = rb == 0 ? part : acc + part;
        }
      } else if (mode == RED_ROWS) {
        acc = qa[0] * qb[0];
        for (int r = 1; r < d; ++r) acc = P::fma(qa[r], qb[r], acc);
      } else {
        std::vector<T> v(N);
        for (int t = 0; t < N; ++t) v[t] = P::fma(qa[N + t], qb[N + t], qa[t] * qb[t]);
        acc = butterfly(v);
      }
      T e = P::abs(a == b ? acc - (T)1 : acc - (T)0);
      mx = first ? e : jmax<P>(mx, e);
      first = false;
    }
  return mx;
}

template <class T>
struct ChainQRSt {
  Acc<T> sums[16];
  int64_t n, count_max, column, step;
  T G_max, orth_max, value, threshold;
  int8_t code;
};

template <class P, class T>
static void chain_qr_step(const cl_reducer& r, ChainQRSt<T>& st, int64_t n, T* W, int N, int k, int mode) {
  if (st.code) return;
  ChainFact<T> f;
  std::vector<T> A(W, W + (size_t)2 * N * k);
  if (mode == RED_BLOCKED)
    chain_householder_blocked<P>(A.data(), N, k, f);
  else
    chain_householder<P>(A.data(), N, k, mode, f);
  auto fail = [&](int code, int64_t col, T v, T thr) {
    st.code = (int8_t)code;
    st.column = col;
    st.value = v;
    st.threshold = thr;
    st.step = n;
  };
  if (f.status) {
    T piv = f.R[(size_t)(f.column - 1) * k + f.column - 1];
    fail(1, f.column, piv, P::finite(piv) ? (T)0 : (T)INFINITY);
    return;
  }
  T gw = window_gw<P>(f.R.data(), k);
  T orth = k >= 2 ? chain_orth_loss<P>(f.Q.data(), N, k, mode) : (T)0;
  T thr = (T)r.tol * P::eps() * (T)k * (T)1;
  if (k >= 2 && !(orth <= thr)) {
    fail(4, 0, orth, thr);
    return;
  }
  for (int i = 0; i < k; ++i) st.sums[i] = acc_add<P>(st.sums[i], P::log(f.R[(size_t)i * k + i]));
  T G = k == 1 ? (T)1 : (k == 2 ? (gw + P::sqrt(jmax<P>(gw * gw - (T)4, (T)0))) / (T)2 : gw);
  st.n = n;
  st.G_max = jmax<P>(st.G_max, G);
  st.orth_max = jmax<P>(st.orth_max, orth);
  if (f.count > st.count_max) st.count_max = f.count;
  st.column = 0;
  for (size_t e = 0; e < (size_t)2 * N * k; ++e) W[e] = f.Q[e];
}

template <class T>
static void chain_write_qr(const ChainQRSt<T>& q, int k, uint8_t* out) {
  Buf b{out, 0};
  for (int i = 0; i < k; ++i) {
    b.put(dbl(q.sums[i].s));
    b.put(dbl(q.sums[i].c));
  }
  b.put(q.n);
  b.put(dbl(q.G_max));
  b.put(dbl(q.orth_max));
  b.put(q.count_max);
  b.put(q.code);
  b.pad(7);
  b.put(q.column);
  b.put(dbl(q.value));
  b.put(dbl(q.threshold));
  b.put(q.step);
}

template <class P, class T>
struct ChainSys {
  int N;
  bool ring;
  T K, eps;
  int left(int j) const { return j > 0 ? j - 1 : (ring ? N - 1 : -1); }
  int right(int j) const { return j < N - 1 ? j + 1 : (ring ? 0 : -1); }
  // W <- ChainTangent(x) * W, then x <- forward(x); x = (q_1..q_N, p_1..p_N), W d x k column-major. F is the frame's
  // type: T, or float for Float32 tangents (the band's entries formed in T and rounded to float once, the product in
  // float in the same operation order).
  template <class F>
  void step(T* x, F* W, int k, std::vector<F>& tmp) const {
    const int d = 2 * N;
    tmp.resize((size_t)d * k);
    const T* q = x;
    for (int j = 0; j < N; ++j) {
      int kl = left(j), kr = right(j);
      T g = K * P::cos(q[j]);
      T cl = kl >= 0 ? eps * P::cos(q[j] - q[kl]) : (T)0;
      T cr = kr >= 0 ? eps * P::cos(q[j] - q[kr]) : (T)0;
      T dg = (g + cl) + cr, gl = -cl, gr = -cr;
      const F dgf = (F)dg, glf = (F)gl, grf = (F)gr;
      for (int c = 0; c < k; ++c) {
        const F* Wc = W + (size_t)c * d;
        F gw = dgf * Wc[j];
        if (kl >= 0) gw = gw + glf * Wc[kl];
        if (kr >= 0) gw = gw + grf * Wc[kr];
        tmp[(size_t)c * d + N + j] = gw + Wc[N + j];
        tmp[(size_t)c * d + j] = (Wc[j] + gw) + Wc[N + j];
      }
    }
    for (size_t e = 0; e < (size_t)d * k; ++e) W[e] = tmp[e];
    std::vector<T> p(N);
    for (int j = 0; j < N; ++j) {
      int kl = left(j), kr = right(j);
      T s = 0;
      if (kl >= 0) s = s + P::sin(q[j] - q[kl]);
      if (kr >= 0) s = s + P::sin(q[j] - q[kr]);
      p[j] = x[N + j] + (K * P::sin(q[j]) + eps * s);
    }
    for (int j = 0; j < N; ++j) {
      x[N + j] = p[j];
      x[j] = x[j] + p[j];
    }
  }
};

// One orbit of the chain with the QR-spectrum chain (or no reducer), full run of Nsteps from fresh states. Float32
// tangents (f32): the frame is rounded to float once at the start and carried in float between handoffs; each
// factorisation runs in T on the exactly widened frame and its Q is rounded back (the frame a failed factorisation
// leaves is the widened one it was given).
template <class P, class T>
static int8_t chain_orbit_f32(const ChainSys<P, T>& sys, const cl_chain& ch, int k, int64_t Nsteps, int tau, T* x, T* W,
                              ChainQRSt<T>& st, int64_t* hstep, int mode) {
  const int d = 2 * sys.N;
  const size_t dk = (size_t)d * k;
  std::vector<float> Wf(dk), tmp;
  for (size_t e = 0; e < dk; ++e) Wf[e] = (float)W[e];
  auto status = [&]() -> int8_t {
    for (int c = 0; c < d; ++c)
      if (!P::finite(x[c])) return 1;
    for (size_t e = 0; e < dk; ++e)
      if (!std::isfinite(Wf[e])) return 2;
    return 0;
  };
  auto widen = [&]() {
    for (size_t e = 0; e < dk; ++e) W[e] = (T)Wf[e];
  };
  int64_t n = 0;
  for (int64_t h = 0; h < Nsteps / tau; ++h) {
    for (int t = 0; t < tau; ++t) sys.step(x, Wf.data(), k, tmp);
    n += tau;
    int8_t code = status();
    if (code) {
      widen();
      *hstep = n;
      return code;
    }
    if (ch.n > 0) {
      widen();
      chain_qr_step<P>(ch.r[0], st, n, W, sys.N, k, mode);
      if (st.code) {
        *hstep = n;
        return 3;
      }
      for (size_t e = 0; e < dk; ++e) Wf[e] = (float)W[e];
    }
  }
  widen();
  *hstep = n;
  return n == Nsteps ? status() : (int8_t)0;
}

template <class P, class T>
static int8_t chain_orbit(const ChainSys<P, T>& sys, const cl_chain& ch, int k, int64_t Nsteps, int tau, T* x, T* W,
                          ChainQRSt<T>& st, int64_t* hstep, int mode) {
  const int d = 2 * sys.N;
  std::vector<T> tmp;
  auto status = [&]() -> int8_t {
    for (int c = 0; c < d; ++c)
      if (!P::finite(x[c])) return 1;
    for (size_t e = 0; e < (size_t)d * k; ++e)
      if (!P::finite(W[e])) return 2;
    return 0;
  };
  int64_t n = 0;
  for (int64_t h = 0; h < Nsteps / tau; ++h) {
    for (int t = 0; t < tau; ++t) sys.step(x, W, k, tmp);
    n += tau;
    int8_t code = status();
    if (code) {
      *hstep = n;
      return code;
    }
    if (ch.n > 0) {
      chain_qr_step<P>(ch.r[0], st, n, W, sys.N, k, mode);
      if (st.code) {
        *hstep = n;
        return 3;
      }
    }
  }
  *hstep = n;
  return n == Nsteps ? status() : (int8_t)0;
}

template <class P>
static int chain_all(const cl_system* sys, const cl_chain* ch, int64_t m, int k, const int64_t* Ns, int64_t N,
                     int tau, double* x, double* W, uint8_t* states, int8_t* code, int64_t* hstep, int mode,
                     int nthreads, bool f32 = false) {
  typedef typename P::T T;
  ChainSys<P, T> cs;
  cs.N = sys->N;
  cs.ring = sys->ring != 0;
  cs.K = (T)sys->K;
  cs.eps = (T)sys->eps;
  const int d = 2 * cs.N;
  const int sb = ch->n > 0 ? 16 * k + 72 : 0;
  std::atomic<int64_t> next(0);
  auto worker = [&]() {
    std::vector<T> xx(d), ww((size_t)d * k);
    while (true) {
      int64_t lo = next.fetch_add(1);
      if (lo >= m) break;
      int64_t i = lo;
      for (int c = 0; c < d; ++c) xx[c] = (T)x[i * d + c];
      for (size_t e = 0; e < (size_t)d * k; ++e) ww[e] = (T)W[i * d * k + e];
      ChainQRSt<T> st;
      memset(&st, 0, sizeof(st));
      for (int s = 0; s < 16; ++s) st.sums[s].s = st.sums[s].c = (T)0;
      st.G_max = st.orth_max = st.value = st.threshold = (T)0;
      int64_t Ni = Ns ? Ns[i] : N;
      code[i] = f32 ? chain_orbit_f32<P, T>(cs, *ch, k, Ni, tau, xx.data(), ww.data(), st, &hstep[i], mode)
                    : chain_orbit<P, T>(cs, *ch, k, Ni, tau, xx.data(), ww.data(), st, &hstep[i], mode);
      for (int c = 0; c < d; ++c) x[i * d + c] = dbl(xx[c]);
      for (size_t e = 0; e < (size_t)d * k; ++e) W[i * d * k + e] = dbl(ww[e]);
      if (states && sb) chain_write_qr(st, k, states + i * sb);
    }
  };
  if (nthreads <= 1) {
    worker();
  } else {
    std::vector<std::thread> th;
    for (int t = 0; t < nthreads; ++t) th.emplace_back(worker);
    for (auto& t : th) t.join();
  }
  return 0;
}
