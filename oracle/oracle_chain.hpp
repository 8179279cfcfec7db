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
// The reference's left fold is the restatement; RED_SITES is what the device kernel computes bit for bit.

enum { RED_ROWS = 0, RED_SITES = 1 };

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

// largest |Q'Q - I| (qr_spectrum.jl:4): entries accumulated with fma in row order, or per site then the butterfly
template <class P, class T>
static T chain_orth_loss(const T* Q, int N, int k, int mode) {
  const int d = 2 * N;
  T mx = 0;
  bool first = true;
  for (int b = 0; b < k; ++b)
    for (int a = 0; a < k; ++a) {
      const T *qa = Q + (size_t)a * d, *qb = Q + (size_t)b * d;
      T acc;
      if (mode == RED_ROWS) {
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
  // W <- ChainTangent(x) * W, then x <- forward(x); x = (q_1..q_N, p_1..p_N), W d x k column-major
  void step(T* x, T* W, int k, std::vector<T>& tmp) const {
    const int d = 2 * N;
    tmp.resize((size_t)d * k);
    const T* q = x;
    for (int j = 0; j < N; ++j) {
      int kl = left(j), kr = right(j);
      T g = K * P::cos(q[j]);
      T cl = kl >= 0 ? eps * P::cos(q[j] - q[kl]) : (T)0;
      T cr = kr >= 0 ? eps * P::cos(q[j] - q[kr]) : (T)0;
      T dg = (g + cl) + cr, gl = -cl, gr = -cr;
      for (int c = 0; c < k; ++c) {
        const T* Wc = W + (size_t)c * d;
        T gw = dg * Wc[j];
        if (kl >= 0) gw = gw + gl * Wc[kl];
        if (kr >= 0) gw = gw + gr * Wc[kr];
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

// One orbit of the chain with the QR-spectrum chain (or no reducer), full run of Nsteps from fresh states.
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
                     int nthreads) {
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
      code[i] = chain_orbit<P, T>(cs, *ch, k, Ni, tau, xx.data(), ww.data(), st, &hstep[i], mode);
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
