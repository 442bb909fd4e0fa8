// Host side of the H_eff middle step: the MPO site tensors of a two-site
// (or one-site) operator fused into one small dense operator, and its
// compressed rows for the sparse pass.  Plain C++ (no HIP), so a host
// compiler can build and check it on its own.
//
// MPO site tensor layout: W[wl][wr][s_out][s_in] (row-major, FP64; a complex
// tensor holds its imaginary plane right after the real one).
#pragma once

#include <cstddef>
#include <vector>

namespace dmrg {

// Two-site operator (src/dmrgpy/pyitensor/kernels.py:353-410, steps 2-3 of
// the L, W_i, W_{i+1}, R chain done as one matrix):
//   M[o][i] = sum_m W1[a][m][s1o][s1] W2[m][b][s2o][s2],
//   o = (s1o d2 + s2o) wr + b,  i = a d1 d2 + s1 d2 + s2,
// a Kout x Kin matrix with Kout = d1 d2 wr, Kin = wl d1 d2 (im empty for a
// real operator).
inline void fuse_two_site(const double* W1, const double* W2, bool wc, int wl, int wm, int wr, int d1, int d2,
                          std::vector<double>& re_out, std::vector<double>& im_out) {
  const int P = d1 * d2, Kin = wl * P, Kout = P * wr;
  re_out.assign((size_t)Kout * Kin, 0.0);
  im_out.assign(wc ? (size_t)Kout * Kin : 0, 0.0);
  const size_t n1 = (size_t)wl * wm * d1 * d1, n2 = (size_t)wm * wr * d2 * d2;
  auto w1 = [&](size_t plane, int a, int m, int so, int s) {
    return W1[plane + (((size_t)a * wm + m) * d1 + so) * d1 + s];
  };
  auto w2 = [&](size_t plane, int m, int b, int so, int s) {
    return W2[plane + (((size_t)m * wr + b) * d2 + so) * d2 + s];
  };
  for (int a = 0; a < wl; ++a)
    for (int b = 0; b < wr; ++b)
      for (int s1o = 0; s1o < d1; ++s1o)
        for (int s2o = 0; s2o < d2; ++s2o)
          for (int s1 = 0; s1 < d1; ++s1)
            for (int s2 = 0; s2 < d2; ++s2) {
              double re = 0, im = 0;
              for (int m = 0; m < wm; ++m) {
                const double ar = w1(0, a, m, s1o, s1), br = w2(0, m, b, s2o, s2);
                const double ai = wc ? w1(n1, a, m, s1o, s1) : 0, bi = wc ? w2(n2, m, b, s2o, s2) : 0;
                re += ar * br - ai * bi;
                im += ar * bi + ai * br;
              }
              const int o = (s1o * d2 + s2o) * wr + b, i = a * P + s1 * d2 + s2;
              re_out[(size_t)o * Kin + i] = re;
              if (wc) im_out[(size_t)o * Kin + i] = im;
            }
}

// One-site operator (the middle step of one_site_heff, dmrg.py:251-275):
//   M[so wr + b][a d + s] = W[a][b][so][s].
inline void one_site_op(const double* W, bool wc, int wl, int wr, int dd, std::vector<double>& re_out,
                        std::vector<double>& im_out) {
  const int P = dd, Kin = wl * P, Kout = P * wr;
  re_out.assign((size_t)Kout * Kin, 0.0);
  im_out.assign(wc ? (size_t)Kout * Kin : 0, 0.0);
  const size_t n1 = (size_t)wl * wr * dd * dd;
  for (int a = 0; a < wl; ++a)
    for (int b = 0; b < wr; ++b)
      for (int so = 0; so < dd; ++so)
        for (int s = 0; s < dd; ++s) {
          const int o = so * wr + b, i = a * P + s;
          const size_t at = (((size_t)a * wr + b) * dd + so) * dd + s;
          re_out[(size_t)o * Kin + i] = W[at];
          if (wc) im_out[(size_t)o * Kin + i] = W[n1 + at];
        }
}

// Compressed rows of a real Kout x Kin operator: its nonzeros row by row.
inline void compress_rows(const std::vector<double>& w, int Kout, int Kin, std::vector<int>& rowptr,
                          std::vector<int>& col, std::vector<double>& val) {
  rowptr.assign(Kout + 1, 0);
  col.clear();
  val.clear();
  for (int o = 0; o < Kout; ++o) {
    for (int i = 0; i < Kin; ++i) {
      const double v = w[(size_t)o * Kin + i];
      if (v != 0.0) {
        col.push_back(i);
        val.push_back(v);
      }
    }
    rowptr[o + 1] = (int)col.size();
  }
}

// Row ranges of the sparse pass's four waves over rows with the given
// compressed-row pointers: contiguous, each about a quarter of (nonzeros +
// one store per row); grp[0] = 0, grp[4] = Kout, non-decreasing.
inline void wave_groups(const std::vector<int>& rowptr, int Kout, int grp[5]) {
  grp[0] = 0;
  for (int g = 1; g < 5; ++g) grp[g] = Kout;
  const double total = (double)rowptr[Kout] + Kout;
  int g = 1;
  double run = 0;
  for (int o = 0; o < Kout && g < 4; ++o) {
    run += (rowptr[o + 1] - rowptr[o]) + 1;
    if (run >= total * g / 4.0) grp[g++] = o + 1;
  }
}

}  // namespace dmrg
