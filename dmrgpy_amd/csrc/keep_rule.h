// The truncated SVD's keep rule (src/dmrgpy/pyitensor/svd.py:32-68): how
// many singular values a split keeps and the weight it discards.  Plain C++
// (no HIP), so a host compiler can build and check it on its own.
#pragma once

#include <algorithm>
#include <vector>

namespace dmrg {

// The reference's keep rule on a descending spectrum.  `extra` is weight
// (in the units of s^2) known to lie outside the listed values -- the part
// of a matrix a subspace split does not capture, or the trailing block an
// early-stopped tridiagonal reduction leaves out; it counts towards the
// total and is always discarded (0 for an exact spectrum).  maxdim <= 0:
// no cap.
inline int truncate_rule(const std::vector<double>& s, double cutoff, int maxdim, int mindim,
                         double* discarded, double extra) {
  const int n = (int)s.size();
  std::vector<double> p(n);
  double total = extra;
  for (int i = 0; i < n; ++i) {
    p[i] = s[i] * s[i];
    total += p[i];
  }
  if (total <= 0) {
    if (discarded) *discarded = 0.0;
    return std::max(1, mindim);
  }
  for (int i = 0; i < n; ++i) p[i] /= total;
  const double pextra = extra / total;
  const int hi = maxdim > 0 ? std::min(maxdim, n) : n;
  const int j_upper = std::max(0, n - mindim);
  const int j_forced = std::min(std::max(0, n - hi), j_upper);
  int j;
  if (j_upper <= j_forced) {
    j = j_forced;
  } else {
    // tail(t) = sum of the j_forced + t smallest, t = 0..j_upper-j_forced;
    // keep the largest t with tail <= cutoff (ascending prefix scan)
    double forced = pextra;
    for (int t = 0; t < j_forced; ++t) forced += p[n - 1 - t];
    int best = 0;
    double tail = forced;
    if (tail <= cutoff) best = 0;
    bool any = tail <= cutoff;
    for (int t = 1; t <= j_upper - j_forced; ++t) {
      tail += p[n - j_forced - t];
      if (tail <= cutoff) {
        best = t;
        any = true;
      }
    }
    j = j_forced + (any ? best : 0);
  }
  const int keep = n - j;
  if (discarded) {
    double dsc = pextra;
    for (int i = keep; i < n; ++i) dsc += p[i];
    *discarded = dsc;
  }
  return keep;
}

}  // namespace dmrg
