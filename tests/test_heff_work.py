"""The H_eff FLOP count behind the roofline figures
(dmrgpy_amd/csrc/heff_work.h), compiled for the host with g++ and checked
against a count derived from the contraction chain itself (CPU): each
pairwise contraction of the chain costs the product of the extents of every
index its two operands carry (the kept ones and the summed ones), and a MAC
is 2 FLOPs real, 8 complex.  A dropped step, a wrong extent or a swapped
site dimension changes the count at one of the shapes below.
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
"""

# The chains as einsum-style steps (operand index letters; the extents come
# from `dims`).  Two-site: x = wl, m = wm, r = wr, p = a' (bra row of L),
# a = a (ket), s / t = s1 / s2, u / v = s1' / s2', b = b, q = b'.
TWO_SITE = [("xpa", "astb"), ("xpstb", "xmus"), ("mputb", "mrvt"), ("rpuvb", "rqb")]
ONE_SITE = [("xpa", "asb"), ("xpsb", "xrus"), ("rpub", "rqb")]


def chain_macs(steps, dims):
    total = 0
    for a, b in steps:
        n = 1
        for ch in set(a) | set(b):
            n *= dims[ch]
        total += n
    return total


@pytest.fixture(scope="module")
def flops(tmp_path_factory):
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
