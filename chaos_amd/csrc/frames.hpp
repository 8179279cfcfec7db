// frames.hpp - seeded initial deviation frames and counter-based uniform draws.
//
// The reference keys each orbit's frame only by (seed, orbit) through the splitmix64 finaliser,
// `StableRNG(_mix(_mix(seed) + orbit))` (src/Tangent/initial.jl:1-7,27-34), draws a standard normal
// d x d matrix and keeps the first k columns of its QR factor with a positive R diagonal. This build
// keeps the key rule and the Haar construction, with its own counter stream so that a frame can be
// formed on the device, in any order, with the same bits as on the host:
//   key      = mix(mix(seed) + orbit)                         (orbit ids are 1-based)
//   u(t)     = ((mix(key + (t+1) * 0x9e3779b97f4a7c15) >> 12) + 1/2) 2^-52   in (0, 1)
//   z(2p)    = r cos(2 pi u(2p+1)), z(2p+1) = r sin(2 pi u(2p+1)), r = sqrt(-2 log u(2p))   (Box-Muller)
//   G[i, j]  = z(j d + i)      (column j depends only on draws of column j and below)
//   W        = G[:, 1:k] orthonormalised column by column with two modified Gram-Schmidt passes
//              against the earlier columns, then divided by its scaled norm (the two-pass rule of
//              the reference's `_orthonormalize`, src/Tangent/initial.jl:10-25), so R_jj > 0.
// Column j depends only on G's first j columns, so it does not depend on k, the property the
// reference tests for its frames (test/tangent-core/tc_ac05_08.jl:37-79).
#pragma once
#include "cl_core.hpp"

namespace cl {

CL_HD uint64_t mix64(uint64_t z) {
  z += 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

CL_HD uint64_t orbit_key(uint64_t seed, int64_t orbit) { return mix64(mix64(seed) + (uint64_t)orbit); }

CL_HD double uniform01(uint64_t key, uint64_t t) {
  uint64_t v = mix64(key + (t + 1) * 0x9e3779b97f4a7c15ull);
  return ((double)(v >> 12) + 0.5) * 0x1p-52;
}

CL_HD double normal_draw(uint64_t key, uint64_t t) {
  uint64_t p = t >> 1;
  double u1 = uniform01(key, 2 * p), u2 = uniform01(key, 2 * p + 1);
  double r = cl_sqrt(-2.0 * cl_log(u1));
  double sn, cs;
  cl_sincos(6.283185307179586 * u2, &sn, &cs);
  return (t & 1) ? r * sn : r * cs;
}

template <int D, int K>
CL_HD void seeded_frame(uint64_t seed, int64_t orbit, Frame<D, K>& W) {
  uint64_t key = orbit_key(seed, orbit);
  Frame<D, K> G;
#pragma unroll
  for (int j = 0; j < K; ++j)
#pragma unroll
    for (int i = 0; i < D; ++i) G.at(i, j) = normal_draw(key, (uint64_t)(j * D + i));
#pragma unroll
  for (int j = 0; j < K; ++j) {
#pragma unroll
    for (int pass = 0; pass < 2; ++pass)
#pragma unroll
      for (int i = 0; i < j; ++i) {
        double c = dot<D>(G.col(i), G.col(j));
#pragma unroll
        for (int l = 0; l < D; ++l) G.at(l, j) = G.at(l, j) - c * G.at(l, i);
      }
    double nr = safe_norm<D>(G.col(j));
#pragma unroll
    for (int l = 0; l < D; ++l) G.at(l, j) = G.at(l, j) / nr;
  }
  W = G;
}

// The same rule for the chain sizes (d up to 256, k up to 16), with run-time sizes, in place in G
// (column-major d x k): draws, two modified Gram-Schmidt passes per column, division by the scaled norm.
CL_HD void seeded_frame_dyn(uint64_t seed, int64_t orbit, int d, int k, double* G) {
  uint64_t key = orbit_key(seed, orbit);
  for (int j = 0; j < k; ++j)
    for (int i = 0; i < d; ++i) G[j * d + i] = normal_draw(key, (uint64_t)(j * d + i));
  for (int j = 0; j < k; ++j) {
    double* g = G + j * d;
    for (int pass = 0; pass < 2; ++pass)
      for (int i = 0; i < j; ++i) {
        const double* gi = G + i * d;
        double c = gi[0] * g[0];
        for (int l = 1; l < d; ++l) c = c + gi[l] * g[l];
        for (int l = 0; l < d; ++l) g[l] = g[l] - c * gi[l];
      }
    double mx = cl_fabs(g[0]);
    for (int l = 1; l < d; ++l) mx = cl_jmax(mx, cl_fabs(g[l]));
    double nr = mx;
    if (cl_isfinite(mx) && mx > 0.0) {
      double s = (g[0] / mx) * (g[0] / mx);
      for (int l = 1; l < d; ++l) {
        double b = g[l] / mx;
        s = s + b * b;
      }
      nr = mx * cl_sqrt(s);
    }
    for (int l = 0; l < d; ++l) g[l] = g[l] / nr;
  }
}

}  // namespace cl
