// Algorithmic work of one H_eff apply: the FLOP count and HBM bytes the
// library's H_eff profiling (and the roofline figures built on it) divides by
// the measured time.  Plain C++ (no HIP), so a host compiler can build and
// check them.
#pragma once

#include "../../include/dmrg_hip.h"

namespace dmrg {

// MACs of the contraction chain in the order the library runs it -- two-site:
// L, W_i, W_{i+1}, R (src/dmrgpy/pyitensor/kernels.py:298-410); one-site:
// L, W, R (src/dmrgpy/pyitensor/dmrg.py:251-275) -- as FLOPs: 2 per MAC for
// a real field, 8 for a complex one (field 1).
inline double heff_flops(const dmrg_heff_desc& d) {
  const double Dl = d.Dl, Dr = d.Dr, wl = d.wl, wr = d.wr;
  double mac;
  if (d.nsite == 2) {
    const double d1 = d.d1, d2 = d.d2, wm = d.wm;
    mac = Dl * Dl * wl * d1 * d2 * Dr + Dl * Dr * d2 * wl * wm * d1 * d1 + Dl * Dr * d1 * wm * wr * d2 * d2 +
          Dl * d1 * d2 * wr * Dr * Dr;
  } else {
    const double dd = d.d1;
    mac = Dl * Dl * wl * dd * Dr + Dl * Dr * wl * wr * dd * dd + Dl * dd * wr * Dr * Dr;
  }
  return (d.field == 1 ? 8.0 : 2.0) * mac;
}

// Algorithmic HBM bytes of one apply: every operand of the chain read once
// and every result written once in the contraction precision (es = 8 for
// prec 0, 4 otherwise; two planes for a complex field) -- L-GEMM: L, x in,
// T1 out; MPO pass: T1 in, T2 out; R-GEMM: T2, R in, y out -- plus, below
// FP64, x converted to the contraction precision (8 + es per element) and y
// back to FP64 (es + 8), except that the FP32 real R-GEMM writes y in FP64
// itself.  The MPO tensors (a few KB) are not counted.
inline double heff_bytes(const dmrg_heff_desc& d) {
  const double P = d.nsite == 2 ? (double)d.d1 * d.d2 : (double)d.d1;
  const double Dl = d.Dl, Dr = d.Dr, wl = d.wl, wr = d.wr;
  const double np_ = d.field == 1 ? 2.0 : 1.0, es = d.prec == 0 ? 8.0 : 4.0;
  const double n = Dl * P * Dr, t1 = wl * Dl * P * Dr, t2 = Dl * P * wr * Dr;
  const bool direct64 = d.prec == 1 && d.field == 0;
  double b = np_ * es * (wl * Dl * Dl + n + t1     // L-GEMM: L, x in, T1 out
                         + t1 + t2                 // MPO pass: T1 in, T2 out
                         + t2 + wr * Dr * Dr);     // R-GEMM: T2, R in
  b += np_ * n * (direct64 ? 8.0 : es);            // R-GEMM: y out
  if (d.prec != 0) b += np_ * n * (8.0 + es);      // x to the contraction precision
  if (d.prec != 0 && !direct64) b += np_ * n * (es + 8.0);  // y back to FP64
  return b;
}

}  // namespace dmrg
