// Workgroup -> output tile order of the tiled GEMM kernels (gemm.hip).  The
// map is plain integer arithmetic shared by the kernels and the host, so a
// host program can check it for every grid shape without a GPU.
#pragma once

#ifdef __HIPCC__
#define DMRG_TILE_FN __host__ __device__ __forceinline__
#else
#define DMRG_TILE_FN inline
#endif

namespace dmrg {

// Tile (row tr, column tc) of workgroup (bx, by) of an nbx x nby grid of
// tiles; its linear id is lin = by * nbx + bx.  Workgroups are dispatched
// round-robin over the 8 XCDs (linear id lin on XCD lin % 8), so XCD x runs
// lin = x, x + 8, ... in that order; that sequence is mapped onto a
// contiguous run of the tiles in grouped order -- XCD_GM row tiles at a
// time, column by column inside a group -- so the workgroups resident on
// one XCD at the same time cover a near-square block of tiles and share both
// their A row panels and their B column panels in that XCD's L2.  A run of
// row-major tiles shares only the A panel: the wide L-step GEMM of H_eff
// then re-fetched its whole B operand once per 128-row band (3.3 GB per
// D=1024 apply, 1.1 GB in this order; group heights 4 and 8 measured
// equal, 16 worse; kernel times unchanged within 1%).  Any tile count
// works (the first nblk % 8 XCDs take one tile more); every tile is taken
// by exactly one workgroup.
constexpr int XCD_GM = 8;
DMRG_TILE_FN void xcd_tile_of(int bx, int by, int nbx, int nby, int& tr, int& tc) {
  const int nblk = nbx * nby;
  const int lin = by * nbx + bx;
  const int x = lin & 7, t = lin >> 3;
  const int per = nblk >> 3, rem = nblk & 7;
  const int idx = x * per + (x < rem ? x : rem) + t;
  const int gsz = XCD_GM * nbx;
  const int g = idx / gsz, first = g * XCD_GM;
  const int rows = nby - first < XCD_GM ? nby - first : XCD_GM;
  const int w = idx - g * gsz;
  tr = first + w % rows;
  tc = w / rows;
}

// A Hermitian product (a Gram matrix) computes only the tiles on or below
// the diagonal: the BM x BN tile (tr, tc) lies strictly above it when its
// first column tc * BN is past its last row (tr + 1) * BM - 1.
DMRG_TILE_FN bool tile_above_diagonal(int tr, int tc, int BM, int BN) { return tc * BN >= (tr + 1) * BM; }
// the same test from the tile's first row m0 = tr * BM and column n0 = tc * BN
DMRG_TILE_FN bool block_above_diagonal(int m0, int n0, int BM) { return n0 >= m0 + BM; }

}  // namespace dmrg
