"""The Krylov solvers' host-side small eigenproblems
(dmrgpy_amd/csrc/host_eig.h), compiled for the host with g++ and checked
against LAPACK and SciPy (CPU).

* ``tridiag_eig`` (implicit QL, Wilkinson shifts): every Lanczos iteration's
  Ritz value, the Ritz vector and the Krylov exponential -- eigenvalues
  against ``numpy.linalg.eigvalsh`` of the dense matrix, eigenvectors by
  residual and orthonormality, on random, diagonal, degenerate, Wilkinson
  (near-degenerate pairs), graded, badly scaled and genuine Lanczos
  tridiagonals, n = 1 ... 120;
* ``tridiag_expm_col0``: the first column of exp(c T) against a 32-digit
  mpmath eigen-decomposition, for real (imaginary-time), imaginary
  (real-time) and complex coefficients, to rounding relative to
  |exp(c T)|_2 (``scipy.linalg.expm`` itself misses by 4e-13 of that norm at
  c = -3, so it is not the reference);
* ``jacobi_eig``: the dense ground state of operators of dimension <= 3
  (complex ones through their real 2n x 2n embedding).
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
#include "host_eig.h"
#include <cstring>
using namespace dmrg;

extern "C" int t_tridiag_eig(int n, const double* d, const double* e, double* w, double* z) {
  try {
    std::vector<double> dd(d, d + n), ee(e, e + (n > 0 ? n - 1 : 0)), zz;
    tridiag_eig(dd, ee, z ? &zz : nullptr);
    std::memcpy(w, dd.data(), sizeof(double) * n);
    if (z) std::memcpy(z, zz.data(), sizeof(double) * n * n);
    return 0;
  } catch (const Error&) {
    return 1;
  }
}

extern "C" int t_expm_col0(int k, const double* al, const double* be, double cre, double cim, double* out) {
  try {
    std::vector<double> a(al, al + k), b(be, be + (k > 0 ? k - 1 : 0));
    const std::vector<std::complex<double>> col = tridiag_expm_col0(a, b, {cre, cim});
    for (int r = 0; r < k; ++r) {
      out[2 * r] = col[r].real();
      out[2 * r + 1] = col[r].imag();
    }
    return 0;
  } catch (const Error&) {
    return 1;
  }
}

extern "C" void t_jacobi_eig(int n, const double* a, double* w, double* v) {
  std::vector<double> aa(a, a + n * n), ww, vv;
  jacobi_eig(aa, n, ww, vv);
  std::memcpy(w, ww.data(), sizeof(double) * n);
  std::memcpy(v, vv.data(), sizeof(double) * n * n);
}
"""

DP = ctypes.POINTER(ctypes.c_double)


@pytest.fixture(scope="module")
def lib(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    d = tmp_path_factory.mktemp("host_eig")
    src, so = d / "wrap.cpp", d / "libhosteig.so"
    src.write_text(WRAPPER)
    subprocess.run([gxx, "-O2", "-std=c++17", "-Wall", "-Werror", "-fPIC", "-shared", "-I", CSRC, str(src),
                    "-o", str(so)], check=True, capture_output=True, text=True)
    lib = ctypes.CDLL(str(so))
    lib.t_tridiag_eig.argtypes = [ctypes.c_int, DP, DP, DP, DP]
    lib.t_tridiag_eig.restype = ctypes.c_int
    lib.t_expm_col0.argtypes = [ctypes.c_int, DP, DP, ctypes.c_double, ctypes.c_double, DP]
    lib.t_expm_col0.restype = ctypes.c_int
    lib.t_jacobi_eig.argtypes = [ctypes.c_int, DP, DP, DP]
    lib.t_jacobi_eig.restype = None
    return lib


def _p(a):
    return a.ctypes.data_as(DP)


def tridiag_eig(lib, d, e, vectors=True):
    d = np.ascontiguousarray(d, dtype=np.float64)
    e = np.ascontiguousarray(e, dtype=np.float64)
    n = d.size
    w = np.empty(n)
    z = np.empty((n, n)) if vectors else None
    rc = lib.t_tridiag_eig(n, _p(d), _p(e) if e.size else None, _p(w), _p(z) if vectors else None)
    assert rc == 0, "tridiag_eig raised (no convergence)"
    return w, z


def dense(d, e):
    return np.diag(d) + np.diag(e, 1) + np.diag(e, -1)


def lanczos_tridiagonal(A, k, rng):
    """(alpha, beta) of k Lanczos steps with full re-orthogonalisation."""
    n = A.shape[0]
    Q = np.zeros((n, k))
    q = rng.standard_normal(n)
    Q[:, 0] = q / np.linalg.norm(q)
    al, be = [], []
    for j in range(k):
        w = A @ Q[:, j]
        al.append(Q[:, j] @ w)
        w -= Q[:, :j + 1] @ (Q[:, :j + 1].T @ w)
        w -= Q[:, :j + 1] @ (Q[:, :j + 1].T @ w)
        if j + 1 < k:
            be.append(np.linalg.norm(w))
            Q[:, j + 1] = w / be[-1]
    return np.array(al), np.array(be)


def tridiagonal_cases():
    rng = np.random.default_rng(20261203)
    cases = [("n1", np.array([-0.7]), np.array([]))]
    for n in (2, 3, 5, 17, 64, 120):
        cases.append(("random%d" % n, rng.standard_normal(n), rng.standard_normal(n - 1)))
    cases.append(("diagonal", rng.standard_normal(12), np.zeros(11)))
    cases.append(("degenerate", np.full(9, 2.5), np.zeros(8)))
    cases.append(("constant", np.zeros(40), np.ones(39)))                       # -2 cos(pi j / 41)
    m = 10
    cases.append(("wilkinson21", np.abs(np.arange(-m, m + 1)).astype(float), np.ones(2 * m)))
    cases.append(("graded", 10.0 ** -np.arange(14.0), 10.0 ** -(np.arange(13.0) + 0.5)))
    cases.append(("tiny_offdiag", rng.standard_normal(20), 1e-200 * rng.standard_normal(19)))
    cases.append(("split", rng.standard_normal(15), np.where(np.arange(14) % 5 == 4, 0.0,
                                                             rng.standard_normal(14))))
    cases.append(("huge", 1e150 * rng.standard_normal(25), 1e150 * rng.standard_normal(24)))
    cases.append(("small", 1e-150 * rng.standard_normal(25), 1e-150 * rng.standard_normal(24)))
    # genuine Lanczos tridiagonals: a spectrum with a cluster at the bottom
    # (the ground state of a gapless chain) and a random symmetric matrix
    ev = np.concatenate([np.full(5, -3.0) + 1e-9 * np.arange(5), rng.uniform(-2.0, 4.0, 195)])
    U, _ = np.linalg.qr(rng.standard_normal((200, 200)))
    al, be = lanczos_tridiagonal((U * ev) @ U.T, 60, rng)
    cases.append(("lanczos_cluster", al, be))
    S = rng.standard_normal((150, 150))
    al, be = lanczos_tridiagonal(S + S.T, 80, rng)
    cases.append(("lanczos_random", al, be))
    return cases


CASES = tridiagonal_cases()


@pytest.mark.parametrize("name,d,e", CASES, ids=[c[0] for c in CASES])
def test_tridiag_eig(lib, name, d, e):
    T = dense(d, e)
    scale = max(np.abs(T).max(), 1e-300)
    w, Z = tridiag_eig(lib, d, e)
    assert np.all(np.diff(w) >= 0), "eigenvalues not ascending"
    ref = np.linalg.eigvalsh(T)
    assert np.abs(w - ref).max() <= 1e-13 * scale * max(1, np.sqrt(d.size)), (name, np.abs(w - ref).max())
    n = d.size
    assert np.abs(Z.T @ Z - np.eye(n)).max() < 1e-13 * max(1, n), name
    assert np.abs(T @ Z - Z * w).max() <= 1e-13 * scale * max(1, n), name
    w2, _ = tridiag_eig(lib, d, e, vectors=False)
    assert np.array_equal(w, w2), "eigenvalues depend on whether vectors are accumulated"


_EXACT = {}


def exact_decomposition(k):
    """The k x k test tridiagonal and its eigen-decomposition at 32 digits
    (mpmath), cached per k."""
    if k not in _EXACT:
        mpmath = pytest.importorskip("mpmath")
        rng = np.random.default_rng(1000 + k)
        al = rng.uniform(-4.0, 4.0, k)
        be = rng.uniform(0.05, 2.0, k - 1)
        with mpmath.workdps(32):
            lam, Q = mpmath.eigsy(mpmath.matrix(dense(al, be).tolist()))
        _EXACT[k] = (al, be, lam, Q)
    return _EXACT[k]


@pytest.mark.parametrize("k", [1, 2, 3, 8, 20, 50])
@pytest.mark.parametrize("coeff", [-0.05, -1.0j * 0.05, -0.025 - 0.5j, -3.0])
def test_tridiag_expm_col0(lib, k, coeff):
    mpmath = pytest.importorskip("mpmath")
    al, be, lam, Q = exact_decomposition(k)
    c = complex(coeff)
    out = np.empty(2 * k)
    assert lib.t_expm_col0(k, _p(al), _p(be) if k > 1 else None, c.real, c.imag, _p(out)) == 0
    got = out[0::2] + 1j * out[1::2]
    with mpmath.workdps(32):
        cm = mpmath.mpc(c.real, c.imag)
        ex = [sum(Q[r, j] * mpmath.exp(cm * lam[j]) * Q[0, j] for j in range(k)) for r in range(k)]
        ref = np.array([complex(v) for v in ex])
        norm = max(abs(mpmath.exp(cm * lam[j])) for j in range(k))
    # to rounding relative to |exp(c T)|_2 = max_j |exp(c lambda_j)| (1 for the
    # real-time steps' c = -i dt); the first column alone can be far smaller
    # when e_1 barely overlaps the dominant eigenvector (c = -3 here)
    assert np.abs(got - ref).max() <= 1e-13 * max(1.0, np.sqrt(k)) * float(norm), (k, coeff)


@pytest.mark.parametrize("n", [1, 2, 3, 4, 6])
def test_jacobi_eig(lib, n):
    rng = np.random.default_rng(n)
    for trial in range(20):
        if n % 2 == 0 and trial % 2:
            # the real 2m x 2m embedding [[Re, -Im], [Im, Re]] of a Hermitian m x m
            m = n // 2
            H = rng.standard_normal((m, m)) + 1j * rng.standard_normal((m, m))
            H = H + H.conj().T
            A = np.block([[H.real, -H.imag], [H.imag, H.real]])
        else:
            A = rng.standard_normal((n, n))
            A = A + A.T
        w, V = np.empty(n), np.empty((n, n))
        Ac = np.ascontiguousarray(A)
        lib.t_jacobi_eig(n, _p(Ac), _p(w), _p(V))
        assert np.abs(np.sort(w) - np.linalg.eigvalsh(A)).max() < 1e-13 * max(1.0, np.abs(A).max())
        assert np.abs(V.T @ V - np.eye(n)).max() < 1e-13
        assert np.abs(A @ V - V * w).max() < 1e-12 * max(1.0, np.abs(A).max())
