"""The GEMM kernels' workgroup -> tile maps (dmrgpy_amd/csrc/tile_order.h),
compiled for the host with g++ and checked exhaustively (CPU).

* every grid shape up to 72 x 72 tiles (and long thin ones): the XCD-grouped
  order hands every tile to exactly one workgroup, none out of range;
* the order's purpose: the 32 workgroups an XCD holds at once (one per CU)
  cover a near-square block of tiles -- on average at most 13.5 distinct A
  row panels + B column panels per window, against 12 for the best 32-tile
  block -- and fewer than a contiguous row-major run of the same tiles;
* the lower-triangle exit of Hermitian products, element by element: no
  tile holding an element on or below the diagonal exits, every tile wholly
  above it does, and the tile-index and first-row/column forms agree.
"""

import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, os.pardir, "dmrgpy_amd", "csrc")

HARNESS = r"""
#include "tile_order.h"
#include <cstdio>
#include <set>
#include <vector>
using namespace dmrg;

static long long bijection_failures(int nbx, int nby) {
  std::vector<int> hit((size_t)nbx * nby, 0);
  long long bad = 0;
  for (int by = 0; by < nby; ++by)
    for (int bx = 0; bx < nbx; ++bx) {
      int tr = -1, tc = -1;
      xcd_tile_of(bx, by, nbx, nby, tr, tc);
      if (tr < 0 || tr >= nby || tc < 0 || tc >= nbx) { ++bad; continue; }
      ++hit[(size_t)tr * nbx + tc];
    }
  for (int h : hit) bad += (h != 1);
  return bad;
}

// distinct row tiles + column tiles over windows of `win` consecutive
// workgroups of each XCD (XCD x runs lin = x, x + 8, ...), summed
static long long panels(int nbx, int nby, int win, bool grouped, long long* windows) {
  const int nblk = nbx * nby;
  long long total = 0;
  *windows = 0;
  for (int x = 0; x < 8; ++x) {
    std::vector<int> lins;
    for (int lin = x; lin < nblk; lin += 8) lins.push_back(lin);
    const int per = nblk / 8, rem = nblk % 8;
    for (size_t w0 = 0; w0 < lins.size(); w0 += win) {
      std::set<int> rows, cols;
      for (size_t k = w0; k < lins.size() && k < w0 + win; ++k) {
        int tr, tc;
        if (grouped) {
          xcd_tile_of(lins[k] % nbx, lins[k] / nbx, nbx, nby, tr, tc);
        } else {   // the same contiguous run of tiles, row-major
          const int idx = x * per + (x < rem ? x : rem) + (int)k;
          tr = idx / nbx;
          tc = idx % nbx;
        }
        rows.insert(tr);
        cols.insert(tc);
      }
      total += (long long)rows.size() + (long long)cols.size();
      ++*windows;
    }
  }
  return total;
}

static long long diagonal_failures(int n, int BM, int BN) {
  const int nty = (n + BM - 1) / BM, ntx = (n + BN - 1) / BN;
  long long bad = 0;
  for (int tr = 0; tr < nty; ++tr)
    for (int tc = 0; tc < ntx; ++tc) {
      const bool skip = tile_above_diagonal(tr, tc, BM, BN);
      if (skip != block_above_diagonal(tr * BM, tc * BN, BM)) ++bad;
      bool holds_lower = false;   // an element (i, j), j <= i, inside the matrix
      for (int i = tr * BM; i < (tr + 1) * BM && i < n; ++i)
        for (int j = tc * BN; j < (tc + 1) * BN && j < n; ++j)
          if (j <= i) holds_lower = true;
      bool wholly_above = true;   // every element of the full tile has j > i
      for (int i = tr * BM; i < (tr + 1) * BM; ++i)
        for (int j = tc * BN; j < (tc + 1) * BN; ++j)
          if (j <= i) wholly_above = false;
      if (skip && holds_lower) ++bad;      // would leave part of the lower triangle unwritten
      if (!skip && wholly_above) ++bad;    // would compute a tile the mirror overwrites
    }
  return bad;
}

int main() {
  long long shapes = 0, bad = 0;
  for (int nby = 1; nby <= 72; ++nby)
    for (int nbx = 1; nbx <= 72; ++nbx) { bad += bijection_failures(nbx, nby); ++shapes; }
  const int thin[][2] = {{1, 4096}, {4096, 1}, {700, 9}, {9, 700}, {129, 33}, {33, 129}, {256, 256}, {8, 1000}};
  for (auto& s : thin) { bad += bijection_failures(s[0], s[1]); ++shapes; }
  std::printf("bijection %lld %lld\n", shapes, bad);

  const int grids[][2] = {{32, 8}, {32, 32}, {64, 16}, {48, 24}, {100, 30}, {16, 64}};
  for (auto& g : grids) {
    long long wg = 0, wr = 0;
    const long long pg = panels(g[0], g[1], 32, true, &wg), pr = panels(g[0], g[1], 32, false, &wr);
    std::printf("panels %d %d %lld %lld %lld\n", g[0], g[1], wg, pg, pr);
  }

  long long dshapes = 0, dbad = 0;
  const int bms[] = {32, 64, 128}, bns[] = {32, 64, 128};
  const int ns[] = {1, 2, 3, 7, 31, 32, 33, 63, 64, 65, 96, 127, 128, 129, 200, 255, 256, 257, 300};
  for (int BM : bms)
    for (int BN : bns)
      for (int n : ns) { dbad += diagonal_failures(n, BM, BN); ++dshapes; }
  std::printf("diagonal %lld %lld\n", dshapes, dbad);
  return 0;
}
"""


@pytest.fixture(scope="module")
def harness_output(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    d = tmp_path_factory.mktemp("tile_order")
    src, exe = d / "harness.cpp", d / "harness"
    src.write_text(HARNESS)
    subprocess.run([gxx, "-O2", "-std=c++17", "-Wall", "-Werror", "-I", CSRC, str(src), "-o", str(exe)],
                   check=True, capture_output=True, text=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True, timeout=120).stdout
    return [line.split() for line in out.splitlines()]


def test_xcd_order_is_a_bijection(harness_output):
    (_, shapes, bad), = [r for r in harness_output if r[0] == "bijection"]
    assert int(shapes) == 72 * 72 + 8
    assert int(bad) == 0


def test_xcd_order_shares_panels(harness_output):
    rows = [r for r in harness_output if r[0] == "panels"]
    assert len(rows) == 6
    for _, nbx, nby, windows, grouped, rowmajor in rows:
        g, r = int(grouped) / int(windows), int(rowmajor) / int(windows)
        print("grid", nbx, "x", nby, "panels per XCD window: grouped", g, "row-major", r)
        assert g <= 13.5 and g < 0.7 * r, (nbx, nby, g, r)


def test_lower_triangle_exit(harness_output):
    (_, shapes, bad), = [r for r in harness_output if r[0] == "diagonal"]
    assert int(shapes) == 3 * 3 * 19
    assert int(bad) == 0
