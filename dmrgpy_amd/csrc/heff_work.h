// Algorithmic work of one H_eff apply: the FLOP count the library's H_eff
// profiling (and the roofline figures built on it) divides by the measured
// time.  Plain C++ (no HIP), so a host compiler can build and check it.
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

}  // namespace dmrg
