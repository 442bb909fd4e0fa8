"""The H_eff middle step's host side (dmrgpy_amd/csrc/small_op.h), compiled
for the host with g++ and checked against NumPy (CPU):

* ``fuse_two_site``: the two MPO site tensors fused into the Kout x Kin
  operator, equal to ``einsum('amxs,mbyt->xybast', W1, W2)`` reshaped in C
  order to (d1 d2 wr, wl d1 d2) -- the row index (s1', s2', b), the column
  index (a, s1, s2) of T1 -- for real and complex tensors of many shapes
  (d = 1 ... 4, w = 1 ... 6) and the Hubbard / Heisenberg automaton MPOs;
* ``one_site_op``: W[a][b][s'][s] as the (d wr, wl d) matrix;
* ``compress_rows`` + ``wave_groups``: the compressed rows reproduce the
  dense operator (explicit zeros dropped, column order kept), and the four
  waves' row ranges cover every row once, each within one row of a quarter
  of the work.
"""

import ctypes
import os
import shutil
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, os.pardir, "dmrgpy_amd", "csrc")

WRAPPER = r"""
#include "small_op.h"
#include <cstring>
using namespace dmrg;

extern "C" void t_fuse2(const double* W1, const double* W2, int wc, int wl, int wm, int wr, int d1, int d2,
                        double* re, double* im) {
  std::vector<double> r, i;
  fuse_two_site(W1, W2, wc != 0, wl, wm, wr, d1, d2, r, i);
  std::memcpy(re, r.data(), sizeof(double) * r.size());
  if (wc) std::memcpy(im, i.data(), sizeof(double) * i.size());
}

extern "C" void t_one(const double* W, int wc, int wl, int wr, int dd, double* re, double* im) {
  std::vector<double> r, i;
  one_site_op(W, wc != 0, wl, wr, dd, r, i);
  std::memcpy(re, r.data(), sizeof(double) * r.size());
  if (wc) std::memcpy(im, i.data(), sizeof(double) * i.size());
}

// returns nnz; rowptr (Kout + 1), col / val (Kout * Kin capacity), grp (5)
extern "C" int t_csr(const double* w, int Kout, int Kin, int* rowptr, int* col, double* val, int* grp) {
  std::vector<double> wv(w, w + (size_t)Kout * Kin);
  std::vector<int> rp, c;
  std::vector<double> v;
  compress_rows(wv, Kout, Kin, rp, c, v);
  wave_groups(rp, Kout, grp);
  std::memcpy(rowptr, rp.data(), sizeof(int) * rp.size());
  if (!c.empty()) {
    std::memcpy(col, c.data(), sizeof(int) * c.size());
    std::memcpy(val, v.data(), sizeof(double) * v.size());
  }
  return (int)c.size();
}
"""

DP = ctypes.POINTER(ctypes.c_double)
IP = ctypes.POINTER(ctypes.c_int)


@pytest.fixture(scope="module")
def lib(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    d = tmp_path_factory.mktemp("small_op")
    src, so = d / "wrap.cpp", d / "libsmallop.so"
    src.write_text(WRAPPER)
    subprocess.run([gxx, "-O2", "-std=c++17", "-Wall", "-Werror", "-fPIC", "-shared", "-I", CSRC, str(src),
                    "-o", str(so)], check=True, capture_output=True, text=True)
    lib = ctypes.CDLL(str(so))
    I = ctypes.c_int
    lib.t_fuse2.argtypes = [DP, DP, I, I, I, I, I, I, DP, DP]
    lib.t_fuse2.restype = None
    lib.t_one.argtypes = [DP, I, I, I, I, DP, DP]
    lib.t_one.restype = None
    lib.t_csr.argtypes = [DP, I, I, IP, IP, DP, IP]
    lib.t_csr.restype = ctypes.c_int
    return lib


def planes(W):
    """[re, im] stacked as the library reads a complex site tensor."""
    return np.ascontiguousarray(np.stack([W.real, W.imag]) if np.iscomplexobj(W) else W, dtype=np.float64)


def fuse2(lib, W1, W2):
    wl, wm, d1, _ = W1.shape
    _, wr, d2, _ = W2.shape
    wc = np.iscomplexobj(W1) or np.iscomplexobj(W2)
    A, B = (planes(W1.astype(complex)), planes(W2.astype(complex))) if wc else (planes(W1), planes(W2))
    kout, kin = d1 * d2 * wr, wl * d1 * d2
    re, im = np.empty((kout, kin)), np.empty((kout, kin))
    lib.t_fuse2(A.ctypes.data_as(DP), B.ctypes.data_as(DP), int(wc), wl, wm, wr, d1, d2,
                re.ctypes.data_as(DP), im.ctypes.data_as(DP))
    return re + 1j * im if wc else re


def fused_reference(W1, W2):
    # W1[a][m][s1'][s1], W2[m][b][s2'][s2] -> M[s1', s2', b, a, s1, s2]
    M = np.einsum("amxs,mbyt->xybast", W1, W2)
    d1, d2, wr, wl = M.shape[0], M.shape[1], M.shape[2], M.shape[3]
    return M.reshape(d1 * d2 * wr, wl * d1 * d2)


SHAPES = [(1, 1, 1, 1, 1), (5, 5, 5, 2, 2), (6, 6, 6, 4, 4), (5, 5, 5, 3, 3), (3, 4, 2, 2, 3), (1, 5, 5, 2, 2),
          (5, 5, 1, 2, 2), (2, 3, 4, 4, 1), (4, 1, 3, 3, 2)]


@pytest.mark.parametrize("shape", SHAPES, ids=["x".join(map(str, s)) for s in SHAPES])
@pytest.mark.parametrize("cplx", [False, True])
def test_fuse_two_site_matches_einsum(lib, shape, cplx):
    wl, wm, wr, d1, d2 = shape
    rng = np.random.default_rng(hash(shape) % 2 ** 32)
    W1 = rng.standard_normal((wl, wm, d1, d1))
    W2 = rng.standard_normal((wm, wr, d2, d2))
    if cplx:
        W1 = W1 + 1j * rng.standard_normal(W1.shape)
        W2 = W2 + 1j * rng.standard_normal(W2.shape)
    got = fuse2(lib, W1, W2)
    ref = fused_reference(W1, W2)
    assert got.shape == ref.shape
    assert np.abs(got - ref).max() <= 1e-14 * max(1.0, np.abs(ref).max())


@pytest.mark.parametrize("codes,terms", [
    ([2] * 4, "heisenberg"),
    ([1] * 4, "hubbard"),
])
def test_fuse_two_site_on_model_mpos(lib, codes, terms):
    """The automaton MPOs of the Heisenberg and Hubbard chains: the fused
    operator of the middle pair equals the einsum of its two tensors."""
    from dmrgpy_amd.mpo import build_mpo
    from oracle import ed
    if terms == "heisenberg":
        tl = ed.heisenberg_terms(len(codes))
    else:
        n = len(codes)
        tl = []
        for i in range(1, n):
            for s in ("up", "dn"):
                tl += [(-1.0, [("Cdag" + s, i), ("C" + s, i + 1)]), (-1.0, [("Cdag" + s, i + 1), ("C" + s, i)])]
        tl += [(4.0, [("Nup", i), ("Ndn", i)]) for i in range(1, n + 1)]
    m = build_mpo(codes, tl)
    # the library's layout W[wl][wr][s'][s] (the automaton stores [np][wl][wr][s'][s])
    W = [w[0] if w.shape[0] == 1 else w[0] + 1j * w[1] for w in m.W]
    i = len(codes) // 2 - 1
    got = fuse2(lib, W[i], W[i + 1])
    ref = fused_reference(W[i], W[i + 1])
    assert np.abs(got - ref).max() <= 1e-14 * max(1.0, np.abs(ref).max())
    assert np.count_nonzero(ref) > 0


@pytest.mark.parametrize("cplx", [False, True])
def test_one_site_op(lib, cplx):
    rng = np.random.default_rng(3)
    for wl, wr, dd in [(1, 1, 1), (5, 5, 2), (6, 6, 4), (1, 5, 3), (5, 1, 3)]:
        W = rng.standard_normal((wl, wr, dd, dd))
        if cplx:
            W = W + 1j * rng.standard_normal(W.shape)
        A = planes(W)
        re, im = np.empty((dd * wr, wl * dd)), np.empty((dd * wr, wl * dd))
        lib.t_one(A.ctypes.data_as(DP), int(cplx), wl, wr, dd, re.ctypes.data_as(DP), im.ctypes.data_as(DP))
        got = re + 1j * im if cplx else re
        ref = np.einsum("abxs->xbas", W).reshape(dd * wr, wl * dd)
        assert np.array_equal(got, ref)


def test_compressed_rows_and_wave_groups(lib):
    rng = np.random.default_rng(9)
    cases = [np.zeros((7, 5)), np.eye(12), rng.standard_normal((1, 1)), np.ones((96, 96))]
    for _ in range(40):
        kout, kin = int(rng.integers(1, 130)), int(rng.integers(1, 130))
        w = rng.standard_normal((kout, kin)) * (rng.random((kout, kin)) < rng.uniform(0.02, 0.6))
        if kout > 4:
            w[rng.integers(0, kout, 3)] = 0.0                  # some empty rows
        cases.append(w)
    for w in cases:
        w = np.ascontiguousarray(w, dtype=np.float64)
        kout, kin = w.shape
        rowptr = np.empty(kout + 1, dtype=np.int32)
        col = np.empty(max(1, kout * kin), dtype=np.int32)
        val = np.empty(max(1, kout * kin))
        grp = np.empty(5, dtype=np.int32)
        nnz = lib.t_csr(w.ctypes.data_as(DP), kout, kin, rowptr.ctypes.data_as(IP), col.ctypes.data_as(IP),
                        val.ctypes.data_as(DP), grp.ctypes.data_as(IP))
        assert nnz == np.count_nonzero(w) and rowptr[0] == 0 and rowptr[-1] == nnz
        assert np.all(np.diff(rowptr) >= 0)
        dense = np.zeros_like(w)
        for o in range(kout):
            cols = col[rowptr[o]:rowptr[o + 1]]
            assert np.all(np.diff(cols) > 0)                  # ascending, no duplicates
            dense[o, cols] = val[rowptr[o]:rowptr[o + 1]]
        assert np.array_equal(dense, w)
        # the four waves: contiguous row ranges covering [0, Kout)
        assert grp[0] == 0 and grp[4] == kout and np.all(np.diff(grp) >= 0)
        work = np.diff(rowptr) + 1
        total = work.sum()
        heaviest = work.max()
        for g in range(4):
            share = work[grp[g]:grp[g + 1]].sum()
            assert share <= total / 4.0 + heaviest, (kout, kin, grp, share, total)
