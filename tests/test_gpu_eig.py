"""Dense symmetric eigensolver of the Gram route (dmrg_sym_eig /
dmrg_set_eigensolver): the library's own two-launch-per-column Householder
tridiagonalisation followed by rocSOLVER's stedc/ormtr (mode 1, the
default), against numpy.linalg.eigh (LAPACK) and rocSOLVER's syevd (mode 0).

Orders cover a single column block (n < 32), panel boundaries (32, 33, 64,
65), ragged sizes and the Gram shapes of the sweeps; matrices include
positive semi-definite Gram matrices with spectra decaying to 1e-14 of the
largest value (what the truncation sees), matrices with exactly zero
sub-columns (Householder tau = 0) and a diagonal matrix.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def engine():
    from dmrgpy_amd.engine import get_engine
    return get_engine(0)


def _dev(engine, a):
    import torch
    return torch.tensor(np.ascontiguousarray(a), dtype=torch.float64, device=engine.device)


def _check(engine, A, mode, tol_rel=1e-13):
    import torch
    n = A.shape[0]
    w, Z = engine.sym_eig(_dev(engine, A), mode)
    torch.cuda.synchronize()
    w, Z = w.cpu().numpy(), Z.cpu().numpy()
    ref = np.linalg.eigvalsh(A)
    scale = max(np.abs(ref).max(), 1e-300)
    assert np.all(np.diff(w) >= -1e-12 * scale)
    assert np.abs(w - ref).max() <= tol_rel * scale * max(1.0, np.sqrt(n) / 4)
    # rows of Z are eigenvectors: residual and orthonormality
    R = A @ Z.T - Z.T * w
    assert np.abs(R).max() <= 1e-12 * scale * np.sqrt(n)
    assert np.abs(Z @ Z.T - np.eye(n)).max() <= 1e-12 * np.sqrt(n)
    return w


@pytest.mark.parametrize("n", [2, 3, 17, 32, 33, 64, 65, 100, 257])
@pytest.mark.parametrize("mode", [1, 0])
def test_random_symmetric(engine, n, mode):
    rng = np.random.default_rng(n + 100 * mode)
    X = rng.standard_normal((n, n))
    _check(engine, X + X.T, mode)


@pytest.mark.parametrize("n", [96, 384, 700])
def test_gram_decaying_spectrum(engine, n):
    """PSD Gram matrix with eigenvalues 1 .. 1e-14 (the regime of the keep
    rule): eigenvalues to ~1e-14 of the largest, modes agree to 1e-13."""
    rng = np.random.default_rng(7 * n)
    Q = np.linalg.qr(rng.standard_normal((n, n)))[0]
    lam = np.logspace(0, -14, n)
    A = (Q * lam) @ Q.T
    A = (A + A.T) / 2
    w1 = _check(engine, A, 1, tol_rel=2e-14)
    w0 = _check(engine, A, 0, tol_rel=2e-14)
    assert np.abs(w1 - w0).max() <= 1e-13


def test_zero_subcolumns(engine):
    """Block-diagonal and diagonal matrices: Householder vectors of zero
    norm (tau = 0) in many columns."""
    rng = np.random.default_rng(3)
    n = 90
    A = np.zeros((n, n))
    for b0, b1 in [(0, 10), (10, 45), (45, 46), (46, 90)]:
        X = rng.standard_normal((b1 - b0, b1 - b0))
        A[b0:b1, b0:b1] = X + X.T
    _check(engine, A, 1)
    _check(engine, np.diag(rng.standard_normal(70)), 1)


def test_svd_split_modes_agree(engine):
    """dmrg_svd_split with the two eigensolvers on a DMRG-like two-site
    matrix (spectrum decaying to 1e-8): same keep count and discarded
    weight, singular values and truncated product within the Gram-route
    rounding bound (eps * s_max^2 / s)."""
    import torch
    rng = np.random.default_rng(11)
    m, n, k = 320, 288, 288
    U0 = np.linalg.qr(rng.standard_normal((m, k)))[0]
    V0 = np.linalg.qr(rng.standard_normal((n, k)))[0]
    s0 = np.logspace(0, -8, k)
    M = (U0 * s0) @ V0.T
    Md = _dev(engine, M[None])
    out = {}
    try:
        for mode in (0, 1):
            engine.set_eigensolver(mode)
            U, V, keep, disc, sv = engine.svd_split(0, Md, m, n, 1e-12, 0, 1, 0)
            torch.cuda.synchronize()
            out[mode] = (U.cpu().numpy()[0], V.cpu().numpy()[0], keep, disc, sv)
    finally:
        engine.set_eigensolver(1)
    U0_, V0_, k0_, d0_, s0_ = out[0]
    U1_, V1_, k1_, d1_, s1_ = out[1]
    assert k0_ == k1_
    assert abs(d0_ - d1_) < 1e-15
    # Gram route: s comes from an eigenvalue known to ~eps*s_max^2, i.e. to
    # ~eps*s_max^2/s absolute, whichever eigensolver runs
    eps = 2.2e-16
    assert np.all(np.abs(s0_ - s1_) <= 1e-13 + 8 * eps / np.maximum(s0_, 1e-300))
    P0, P1 = U0_ @ V0_, U1_ @ V1_
    skeep = s0_[k0_ - 1]
    assert np.linalg.norm(P0 - P1) < (1e-12 + 100 * eps / skeep) * np.linalg.norm(P0)
    assert np.abs(U1_.T @ U1_ - np.eye(k1_)).max() < 1e-12
