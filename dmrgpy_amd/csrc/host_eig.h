// Host-side small eigenproblems of the Krylov solvers: the projected
// tridiagonal matrices of Lanczos and of the Lanczos exponential, and the
// dense ground state of operators of dimension <= 3.  Plain C++ (no HIP), so
// a host compiler can build and check these routines on their own.
#pragma once

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <complex>
#include <vector>

#include "dmrg_error.h"

namespace dmrg {

// Symmetric tridiagonal eigen-decomposition by the implicit QL method with
// Wilkinson shifts.  d: diagonal (n), e: off-diagonal (n-1).  On return d
// holds the eigenvalues ascending; if z is given it receives the
// eigenvectors as columns of an n x n row-major matrix.
inline void tridiag_eig(std::vector<double>& d, std::vector<double> e, std::vector<double>* z) {
  const int n = (int)d.size();
  e.resize(n, 0.0);
  if (n > 0) e[n - 1] = 0.0;
  if (z) {
    z->assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) (*z)[(size_t)i * n + i] = 1.0;
  }
  for (int l = 0; l < n; ++l) {
    int iter = 0;
    for (;;) {
      int m = l;
      for (; m < n - 1; ++m) {
        const double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
        if (std::fabs(e[m]) <= DBL_EPSILON * dd) break;
      }
      if (m == l) break;
      if (++iter > 100) throw Error("tridiag_eig: no convergence");
      double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
      double r = std::hypot(g, 1.0);
      g = d[m] - d[l] + e[l] / (g + std::copysign(r, g));
      double s = 1.0, cth = 1.0, p = 0.0;
      bool deflated = false;
      for (int i = m - 1; i >= l; --i) {
        const double f = s * e[i], b = cth * e[i];
        r = std::hypot(f, g);
        e[i + 1] = r;
        if (r == 0.0) {
          d[i + 1] -= p;
          e[m] = 0.0;
          deflated = true;
          break;
        }
        s = f / r;
        cth = g / r;
        g = d[i + 1] - p;
        r = (d[i] - g) * s + 2.0 * cth * b;
        p = s * r;
        d[i + 1] = g + p;
        g = cth * r - b;
        if (z) {
          for (int k = 0; k < n; ++k) {
            double* zr = z->data() + (size_t)k * n;
            const double t = zr[i + 1];
            zr[i + 1] = s * zr[i] + cth * t;
            zr[i] = cth * zr[i] - s * t;
          }
        }
      }
      if (deflated) continue;
      d[l] -= p;
      e[l] = g;
      e[m] = 0.0;
    }
  }
  // sort ascending (selection sort keeps it simple; n <= a few hundred)
  for (int i = 0; i < n - 1; ++i) {
    int k = i;
    for (int j = i + 1; j < n; ++j)
      if (d[j] < d[k]) k = j;
    if (k != i) {
      std::swap(d[i], d[k]);
      if (z)
        for (int r2 = 0; r2 < n; ++r2) std::swap((*z)[(size_t)r2 * n + i], (*z)[(size_t)r2 * n + k]);
    }
  }
}

// Dense real symmetric eigen-decomposition by cyclic Jacobi rotations
// (a: n x n row-major, destroyed).  Used only for operators of dimension
// <= 3 (complex ones through their real 2n x 2n embedding).
inline void jacobi_eig(std::vector<double>& a, int n, std::vector<double>& w, std::vector<double>& v) {
  v.assign((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) v[(size_t)i * n + i] = 1.0;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0;
    for (int p = 0; p < n; ++p)
      for (int q = p + 1; q < n; ++q) off += a[(size_t)p * n + q] * a[(size_t)p * n + q];
    if (off < 1e-300) break;
    for (int p = 0; p < n; ++p)
      for (int q = p + 1; q < n; ++q) {
        const double apq = a[(size_t)p * n + q];
        if (apq == 0.0) continue;
        const double theta = (a[(size_t)q * n + q] - a[(size_t)p * n + p]) / (2.0 * apq);
        const double t = std::copysign(1.0, theta) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        const double cs = 1.0 / std::sqrt(t * t + 1.0), sn = t * cs;
        for (int k = 0; k < n; ++k) {
          const double akp = a[(size_t)k * n + p], akq = a[(size_t)k * n + q];
          a[(size_t)k * n + p] = cs * akp - sn * akq;
          a[(size_t)k * n + q] = sn * akp + cs * akq;
        }
        for (int k = 0; k < n; ++k) {
          const double apk = a[(size_t)p * n + k], aqk = a[(size_t)q * n + k];
          a[(size_t)p * n + k] = cs * apk - sn * aqk;
          a[(size_t)q * n + k] = sn * apk + cs * aqk;
        }
        for (int k = 0; k < n; ++k) {
          const double vkp = v[(size_t)k * n + p], vkq = v[(size_t)k * n + q];
          v[(size_t)k * n + p] = cs * vkp - sn * vkq;
          v[(size_t)k * n + q] = sn * vkp + cs * vkq;
        }
      }
  }
  w.resize(n);
  for (int i = 0; i < n; ++i) w[i] = a[(size_t)i * n + i];
}

// First column of exp(coeff T) for the symmetric tridiagonal T with diagonal
// al (k) and off-diagonal be (k - 1): Z exp(coeff Lambda) Z^T e_1 from the
// eigen-decomposition T = Z Lambda Z^T.
inline std::vector<std::complex<double>> tridiag_expm_col0(const std::vector<double>& al,
                                                           const std::vector<double>& be,
                                                           std::complex<double> coeff) {
  const int k = (int)al.size();
  std::vector<std::complex<double>> col(k);
  if (k == 1) {
    col[0] = std::exp(coeff * al[0]);
    return col;
  }
  std::vector<double> d(al), z;
  tridiag_eig(d, be, &z);
  for (int r = 0; r < k; ++r) {
    std::complex<double> s = 0;
    for (int j = 0; j < k; ++j) s += z[(size_t)r * k + j] * std::exp(coeff * d[j]) * z[(size_t)0 * k + j];
    col[r] = s;
  }
  return col;
}

}  // namespace dmrg
