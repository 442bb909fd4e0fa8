"""The H_eff FLOP count and algorithmic bytes behind the roofline figures
(dmrgpy_amd/csrc/heff_work.h), compiled for the host with g++ and checked
against counts derived from the contraction chain itself (CPU):

* FLOPs: each pairwise contraction of the chain costs the product of the
  extents of every index its two operands carry (the kept ones and the
  summed ones); a MAC is 2 FLOPs real, 8 complex;
* bytes: every operand of every step read once and every result written
  once -- sizes from the same index sets -- in the contraction precision,
  plus the FP64 <-> contraction-precision conversions of x and y below FP64
  (none for y when the FP32 real R-GEMM writes FP64 itself).

A dropped step, a wrong extent or a swapped site dimension changes a count
at one of the shapes below.
"""

import ctypes
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, os.pardir, "dmrgpy_amd", "csrc")

WRAPPER = r"""
#include "heff_work.h"
extern "C" double t_flops(int nsite, int field, int Dl, int Dr, int d1, int d2, int wl, int wm, int wr) {
  dmrg_heff_desc d = {};
  d.nsite = nsite; d.field = field; d.Dl = Dl; d.Dr = Dr; d.d1 = d1; d.d2 = d2;
  d.wl = wl; d.wm = wm; d.wr = wr;
  return dmrg::heff_flops(d);
}
extern "C" double t_bytes(int nsite, int field, int prec, int Dl, int Dr, int d1, int d2, int wl, int wm, int wr) {
  dmrg_heff_desc d = {};
  d.nsite = nsite; d.field = field; d.prec = prec; d.Dl = Dl; d.Dr = Dr; d.d1 = d1; d.d2 = d2;
  d.wl = wl; d.wm = wm; d.wr = wr;
  return dmrg::heff_bytes(d);
}
"""

# The chains as einsum-style steps (operand index letters; the extents come
# from `dims`).  Two-site: x = wl, m = wm, r = wr, p = a' (bra row of L),
# a = a (ket), s / t = s1 / s2, u / v = s1' / s2', b = b, q = b'.
TWO_SITE = [("xpa", "astb"), ("xpstb", "xmus"), ("mputb", "mrvt"), ("rpuvb", "rqb")]
ONE_SITE = [("xpa", "asb"), ("xpsb", "xrus"), ("rpub", "rqb")]

# The kernels the library runs for the same chains: (inputs, output) read /
# written through HBM.  The two MPO steps are one pass over T1; the MPO
# tensors themselves (a few KB) are not counted.
TWO_SITE_PASSES = [(["xpa", "astb"], "xpstb"),        # L-GEMM: L, x -> T1
                   (["xpstb"], "rpuvb"),              # MPO pass: T1 -> T2
                   (["rpuvb", "rqb"], "puvq")]        # R-GEMM: T2, R -> y
ONE_SITE_PASSES = [(["xpa", "asb"], "xpsb"), (["xpsb"], "rpub"), (["rpub", "rqb"], "puq")]


def chain_macs(steps, dims):
    total = 0
    for a, b in steps:
        n = 1
        for ch in set(a) | set(b):
            n *= dims[ch]
        total += n
    return total


def size(idx, dims):
    n = 1
    for ch in idx:
        n *= dims[ch]
    return n


def chain_bytes(passes, dims, field, prec):
    npl, es = (2 if field else 1), (8 if prec == 0 else 4)
    x, y = passes[0][0][1], passes[-1][1]
    total = 0
    for ins, out in passes:
        for t in ins:
            total += npl * es * size(t, dims)
        if out != y:
            total += npl * es * size(out, dims)
    direct64 = prec == 1 and field == 0
    total += npl * size(y, dims) * (8 if direct64 else es)
    if prec != 0:
        total += npl * size(x, dims) * (8 + es)
        if not direct64:
            total += npl * size(y, dims) * (es + 8)
    return total


@pytest.fixture(scope="module")
def lib(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    d = tmp_path_factory.mktemp("heff_work")
    src, so = d / "wrap.cpp", d / "libheffwork.so"
    src.write_text(WRAPPER)
    subprocess.run([gxx, "-O2", "-std=c++17", "-Wall", "-Werror", "-fPIC", "-shared", "-I", CSRC, str(src),
                    "-o", str(so)], check=True, capture_output=True, text=True)
    lib = ctypes.CDLL(str(so))
    lib.t_flops.argtypes = [ctypes.c_int] * 9
    lib.t_flops.restype = ctypes.c_double
    lib.t_bytes.argtypes = [ctypes.c_int] * 10
    lib.t_bytes.restype = ctypes.c_double
    return lib


@pytest.fixture(scope="module")
def flops(lib):
    return lib.t_flops


SHAPES = [  # Dl, Dr, d1, d2, wl, wm, wr
    (1024, 1024, 4, 4, 6, 6, 6),      # the D=1024 Hubbard bulk bond
    (1024, 1024, 2, 2, 5, 5, 5),      # Heisenberg
    (1024, 1024, 3, 3, 5, 5, 5),      # S=1 Heisenberg
    (730, 512, 4, 4, 6, 6, 6),        # unequal bonds
    (1, 4, 4, 4, 1, 6, 6),            # left edge
    (64, 1, 2, 3, 5, 4, 1),           # right edge, different site dimensions
    (37, 91, 3, 2, 2, 7, 3),
]


@pytest.mark.parametrize("shape", SHAPES, ids=["x".join(map(str, s)) for s in SHAPES])
@pytest.mark.parametrize("field", [0, 1])
def test_two_site_flops(flops, shape, field):
    Dl, Dr, d1, d2, wl, wm, wr = shape
    dims = dict(p=Dl, a=Dl, b=Dr, q=Dr, s=d1, u=d1, t=d2, v=d2, x=wl, m=wm, r=wr)
    expect = (8 if field else 2) * chain_macs(TWO_SITE, dims)
    assert flops(2, field, Dl, Dr, d1, d2, wl, wm, wr) == pytest.approx(expect, rel=1e-15)


@pytest.mark.parametrize("shape", SHAPES, ids=["x".join(map(str, s)) for s in SHAPES])
@pytest.mark.parametrize("field", [0, 1])
def test_one_site_flops(flops, shape, field):
    Dl, Dr, d1, _, wl, _, wr = shape
    dims = dict(p=Dl, a=Dl, b=Dr, q=Dr, s=d1, u=d1, x=wl, r=wr)
    expect = (8 if field else 2) * chain_macs(ONE_SITE, dims)
    assert flops(1, field, Dl, Dr, d1, 0, wl, 0, wr) == pytest.approx(expect, rel=1e-15)


def test_bulk_hubbard_figure(flops):
    """The bulk D=1024 Hubbard bond (d=4, w=6): M_H = 2 w d^2 D^3 + 2 w^2 d^3
    D^2 = 2.11e11 MAC, i.e. 4.22e11 FLOP per apply."""
    f = flops(2, 0, 1024, 1024, 4, 4, 6, 6, 6)
    D, d, w = 1024.0, 4.0, 6.0
    assert f == 2 * (2 * w * d ** 2 * D ** 3 + 2 * w ** 2 * d ** 3 * D ** 2)
    assert 4.21e11 < f < 4.23e11


@pytest.mark.parametrize("shape", SHAPES, ids=["x".join(map(str, s)) for s in SHAPES])
@pytest.mark.parametrize("field", [0, 1])
@pytest.mark.parametrize("prec", [0, 1, 2])
def test_bytes(lib, shape, field, prec):
    Dl, Dr, d1, d2, wl, wm, wr = shape
    dims = dict(p=Dl, a=Dl, b=Dr, q=Dr, s=d1, u=d1, t=d2, v=d2, x=wl, m=wm, r=wr)
    assert lib.t_bytes(2, field, prec, Dl, Dr, d1, d2, wl, wm, wr) == pytest.approx(
        chain_bytes(TWO_SITE_PASSES, dims, field, prec), rel=1e-15)
    assert lib.t_bytes(1, field, prec, Dl, Dr, d1, 0, wl, 0, wr) == pytest.approx(
        chain_bytes(ONE_SITE_PASSES, dims, field, prec), rel=1e-15)
