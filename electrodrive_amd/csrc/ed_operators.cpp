// Host construction of the FMM translation operators (FP64).
//
// Solid harmonics with Condon-Shortley associated Legendre functions:
//   R_n^m(x) = r^n P_n^m(cos t) e^{i m phi} / (n+m)!
//   S_n^m(x) = (n-m)! P_n^m(cos t) e^{i m phi} / r^(n+1)
//   X_n^{-m} = (-1)^m conj(X_n^m),  1/|x-y| = sum conj(R_n^m(y)) S_n^m(x).
// Coefficients are scaled by the box side s, so the operators only depend on
// the lattice offset (M2L) or the child octant (M2M, L2L):
//   M2L  L_n^m = (-1)^(n+m) sum_{j,k} M_j^k S_{j+n}^{k-m}(o)        (times 1/s)
//   M2M  M'_n^m = sum_{j<=n,k} M_j^k 2^-j conj(R_{n-j}^{m-k}(d))
//   L2L  L'_j^k = sum_{n>=j,m} L_n^m 2^-j R_{n-j}^{m-k}(d)
// with d = (c_child - c_parent)/s_parent.  The operators act on the real
// packing of the m >= 0 half: index n^2 for Re X_n^0 and n^2+2m-1, n^2+2m for
// Re, Im X_n^m.  The recurrences used here are mirrored by the device P2M/L2P.
#include <cmath>
#include <complex>
#include <cstring>
#include <vector>

#include "ed_internal.h"

namespace ed {

using cplx = std::complex<double>;

namespace {

struct HTable {
  int p;
  std::vector<cplx> v;  // (p+1) x (p+1), entry [n][m] for m <= n
  explicit HTable(int p_) : p(p_), v((size_t)(p_ + 1) * (p_ + 1)) {}
  cplx& at(int n, int m) { return v[(size_t)n * (p + 1) + m]; }
  // X_n^m for any m (zero outside |m| <= n)
  cplx full(int n, int m) const {
    if (n < 0 || n > p || m > n || m < -n) return cplx(0.0, 0.0);
    if (m >= 0) return v[(size_t)n * (p + 1) + m];
    const cplx c = std::conj(v[(size_t)n * (p + 1) + (-m)]);
    return ((-m) & 1) ? -c : c;
  }
};

HTable regular(double x, double y, double z, int p) {
  HTable R(p);
  const double r2 = x * x + y * y + z * z;
  const cplx w(x, y);
  R.at(0, 0) = 1.0;
  for (int m = 1; m <= p; ++m) R.at(m, m) = -w / (2.0 * m) * R.at(m - 1, m - 1);
  for (int m = 0; m < p; ++m) {
    R.at(m + 1, m) = z * R.at(m, m);
    for (int n = m + 1; n < p; ++n)
      R.at(n + 1, m) = ((2.0 * n + 1.0) * z * R.at(n, m) - r2 * R.at(n - 1, m)) /
                       (double)((n + 1) * (n + 1) - m * m);
  }
  return R;
}

HTable irregular(double x, double y, double z, int p) {
  HTable S(p);
  const double r2 = x * x + y * y + z * z;
  const double ir2 = 1.0 / r2;
  const cplx w(x, y);
  S.at(0, 0) = 1.0 / std::sqrt(r2);
  for (int m = 1; m <= p; ++m) S.at(m, m) = -(2.0 * m - 1.0) * w * ir2 * S.at(m - 1, m - 1);
  for (int m = 0; m < p; ++m) {
    S.at(m + 1, m) = (2.0 * m + 1.0) * z * ir2 * S.at(m, m);
    for (int n = m + 1; n < p; ++n)
      S.at(n + 1, m) = ((2.0 * n + 1.0) * z * S.at(n, m) -
                        (double)(n * n - m * m) * S.at(n - 1, m)) * ir2;
  }
  return S;
}

// Accumulate "out_(n,m) += c * in_(j,k)" for m >= 0 outputs and any input k,
// into the real matrix T (rows: packed out, cols: packed in).
inline void accum(double* T, int ncol, int n, int m, int j, int k, cplx c) {
  const int ak = k < 0 ? -k : k;
  const int col_re = j * j + (ak == 0 ? 0 : 2 * ak - 1);
  const int col_im = ak == 0 ? -1 : j * j + 2 * ak;
  const int row_re = n * n + (m == 0 ? 0 : 2 * m - 1);
  const int row_im = m == 0 ? -1 : n * n + 2 * m;
  double cr = c.real(), ci = c.imag();
  if (k >= 0) {
    // c * (a + i b)
    T[(size_t)row_re * ncol + col_re] += cr;
    if (col_im >= 0) T[(size_t)row_re * ncol + col_im] += -ci;
    if (row_im >= 0) {
      T[(size_t)row_im * ncol + col_re] += ci;
      if (col_im >= 0) T[(size_t)row_im * ncol + col_im] += cr;
    }
  } else {
    // X_j^k = (-1)^|k| conj(X_j^|k|): c * sgn * (a - i b)
    const double sg = (ak & 1) ? -1.0 : 1.0;
    cr *= sg;
    ci *= sg;
    T[(size_t)row_re * ncol + col_re] += cr;
    if (col_im >= 0) T[(size_t)row_re * ncol + col_im] += ci;
    if (row_im >= 0) {
      T[(size_t)row_im * ncol + col_re] += ci;
      if (col_im >= 0) T[(size_t)row_im * ncol + col_im] += -cr;
    }
  }
}

inline void child_offset(int octant, double* d) {
  d[0] = (octant & 1) ? 0.25 : -0.25;
  d[1] = (octant & 2) ? 0.25 : -0.25;
  d[2] = (octant & 4) ? 0.25 : -0.25;
}

}  // namespace

void m2l_matrix(int p, double ox, double oy, double oz, double* T) {
  const int nc = n_real(p);
  std::memset(T, 0, sizeof(double) * (size_t)nc * nc);
  const HTable S = irregular(ox, oy, oz, 2 * p);
  for (int n = 0; n <= p; ++n)
    for (int m = 0; m <= n; ++m) {
      const double sg = ((n + m) & 1) ? -1.0 : 1.0;
      for (int j = 0; j <= p; ++j)
        for (int k = -j; k <= j; ++k) {
          const cplx c = sg * S.full(j + n, k - m);
          if (c == cplx(0.0, 0.0)) continue;
          accum(T, nc, n, m, j, k, c);
        }
    }
}

void m2m_matrix(int p, int octant, double* T) {
  const int nc = n_real(p);
  std::memset(T, 0, sizeof(double) * (size_t)nc * nc);
  double d[3];
  child_offset(octant, d);
  const HTable R = regular(d[0], d[1], d[2], p);
  for (int n = 0; n <= p; ++n)
    for (int m = 0; m <= n; ++m)
      for (int j = 0; j <= n; ++j) {
        const double f = std::ldexp(1.0, -j);
        for (int k = -j; k <= j; ++k) {
          const cplx c = f * std::conj(R.full(n - j, m - k));
          if (c == cplx(0.0, 0.0)) continue;
          accum(T, nc, n, m, j, k, c);
        }
      }
}

void l2l_matrix(int p, int octant, double* T) {
  const int nc = n_real(p);
  std::memset(T, 0, sizeof(double) * (size_t)nc * nc);
  double d[3];
  child_offset(octant, d);
  const HTable R = regular(d[0], d[1], d[2], p);
  for (int j = 0; j <= p; ++j)
    for (int k = 0; k <= j; ++k) {
      const double f = std::ldexp(1.0, -j);
      for (int n = j; n <= p; ++n)
        for (int m = -n; m <= n; ++m) {
          const cplx c = f * R.full(n - j, m - k);
          if (c == cplx(0.0, 0.0)) continue;
          accum(T, nc, j, k, n, m, c);
        }
    }
}

}  // namespace ed
