"""Dense eigensolvers of the Gram route (dmrg_sym_eig, dmrg_herm_eig,
dmrg_set_eigensolver): the library's own two-launch-per-column Householder
tridiagonalisation (real symmetric, and complex Hermitian to a real
tridiagonal) followed by rocSOLVER's stedc/ormtr/unmtr (mode 1, the
default), against numpy.linalg.eigh (LAPACK) and rocSOLVER's syevd/heevd
(mode 0).

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


@pytest.mark.parametrize("n", [2, 3, 17, 32, 33, 64, 65, 100, 127, 128, 129, 130, 257, 300])
@pytest.mark.parametrize("mode", [1, 2, 0])
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


@pytest.mark.parametrize("n", [300, 1100, 2500])
def test_back_transformation_matches_ormtr(engine, n):
    """The blocked compact-WY back-transformation (mode 1) against rocSOLVER
    dormtr (mode 2) after the same reduction and divide and conquer: equal
    eigenvectors to rounding (several 128-reflector blocks, split inner
    products), on a Gram-like spectrum."""
    import torch
    rng = np.random.default_rng(n + 3)
    Q = np.linalg.qr(rng.standard_normal((n, n)))[0]
    lam = np.maximum(1e-9 ** (np.arange(n) / 256.0), 1e-17)
    A = (Q * lam) @ Q.T
    A = (A + A.T) / 2
    out = {}
    for mode in (1, 2):
        w, Z = engine.sym_eig(_dev(engine, A), mode)
        torch.cuda.synchronize()
        out[mode] = (w.cpu().numpy(), Z.cpu().numpy())
    assert np.array_equal(out[1][0], out[2][0])   # same reduction and stedc
    assert np.abs(out[1][1] - out[2][1]).max() <= 1e-12 * np.sqrt(n)
    Z = out[1][1]
    assert np.abs(Z @ Z.T - np.eye(n)).max() <= 1e-12 * np.sqrt(n)


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


@pytest.mark.parametrize("other", [1, 2])
def test_svd_split_modes_agree(engine, other):
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
        for mode in (0, other):
            engine.set_eigensolver(mode)
            U, V, keep, disc, sv = engine.svd_split(0, Md, m, n, 1e-12, 0, 1, 0)
            torch.cuda.synchronize()
            out[mode] = (U.cpu().numpy()[0], V.cpu().numpy()[0], keep, disc, sv)
    finally:
        engine.set_eigensolver(1)   # the default
    U0_, V0_, k0_, d0_, s0_ = out[0]
    U1_, V1_, k1_, d1_, s1_ = out[other]
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


# ---- complex Hermitian (dmrg_herm_eig): the own reduction to a real
# tridiagonal (zhetrd conventions) + stedc + zunmtr (mode 1), zheevd (mode 0)

def _check_herm(engine, A, mode, tol_rel=1e-13):
    import torch
    n = A.shape[0]
    w, Z = engine.herm_eig(_dev(engine, np.stack([A.real, A.imag])), mode)
    torch.cuda.synchronize()
    w, Z = w.cpu().numpy(), Z.cpu().numpy()
    Z = Z[0] + 1j * Z[1]
    ref = np.linalg.eigvalsh(A)
    scale = max(np.abs(ref).max(), 1e-300)
    assert np.all(np.diff(w) >= -1e-12 * scale)
    assert np.abs(w - ref).max() <= tol_rel * scale * max(1.0, np.sqrt(n) / 4)
    # rows of Z are eigenvectors of A (not of conj(A)): residual, orthonormality
    R = A @ Z.T - Z.T * w
    assert np.abs(R).max() <= 1e-12 * scale * np.sqrt(n)
    assert np.abs(Z.conj() @ Z.T - np.eye(n)).max() <= 1e-12 * np.sqrt(n)
    return w


@pytest.mark.parametrize("n", [2, 3, 17, 32, 33, 64, 65, 66, 100, 129, 257])
@pytest.mark.parametrize("mode", [1, 2, 0])
def test_random_hermitian(engine, n, mode):
    rng = np.random.default_rng(n + 100 * mode + 7)
    X = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    _check_herm(engine, X + X.conj().T, mode)


@pytest.mark.parametrize("n", [96, 390])
def test_hermitian_gram_decaying_spectrum(engine, n):
    """Complex PSD Gram matrix with eigenvalues 1 .. 1e-14: the own
    reduction and zheevd agree to 1e-13 of the largest value."""
    rng = np.random.default_rng(5 * n)
    Q = np.linalg.qr(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))[0]
    lam = np.logspace(0, -14, n)
    A = (Q * lam) @ Q.conj().T
    A = (A + A.conj().T) / 2
    w1 = _check_herm(engine, A, 1, tol_rel=2e-14)
    w0 = _check_herm(engine, A, 0, tol_rel=2e-14)
    assert np.abs(w1 - w0).max() <= 1e-13


@pytest.mark.parametrize("mode", [1, 2])
def test_hermitian_special(engine, mode):
    """Block-diagonal matrices (exactly zero sub-columns: the reduction takes
    the reflection tau = 2 there, since rocSOLVER's zunmtr -- mode 2 --
    returns NaN vectors when some reflectors have tau = 0, and zheevd itself
    does on these matrices), a real symmetric matrix given as complex, and a
    diagonal; with the library's back-transformation (mode 1) and zunmtr."""
    rng = np.random.default_rng(9)
    n = 70
    A = np.zeros((n, n), complex)
    for b0, b1 in [(0, 10), (10, 45), (45, 46), (46, 70)]:
        X = rng.standard_normal((b1 - b0, b1 - b0)) + 1j * rng.standard_normal((b1 - b0, b1 - b0))
        A[b0:b1, b0:b1] = X + X.conj().T
    A[11, 10] = 0.5j   # a lone complex entry right below the diagonal
    A[10, 11] = -0.5j
    _check_herm(engine, A, mode)
    X = rng.standard_normal((40, 40))
    _check_herm(engine, (X + X.T).astype(complex), mode)
    _check_herm(engine, np.diag(rng.standard_normal(50)).astype(complex), mode)


@pytest.mark.parametrize("other", [1, 2])
def test_svd_split_complex_modes_agree(engine, other):
    """Complex dmrg_svd_split with the own Hermitian reduction and with
    zheevd: same keep and discarded weight, values and truncated product
    within the Gram-route rounding bound."""
    import torch
    rng = np.random.default_rng(13)
    m, n, k = 192, 160, 160
    U0 = np.linalg.qr(rng.standard_normal((m, k)) + 1j * rng.standard_normal((m, k)))[0]
    V0 = np.linalg.qr(rng.standard_normal((n, k)) + 1j * rng.standard_normal((n, k)))[0]
    s0 = np.logspace(0, -8, k)
    M = (U0 * s0) @ V0.conj().T
    Md = _dev(engine, np.stack([M.real, M.imag]))
    out = {}
    try:
        for mode in (0, other):
            engine.set_eigensolver(mode)
            U, V, keep, disc, sv = engine.svd_split(1, Md, m, n, 1e-12, 0, 1, 0)
            torch.cuda.synchronize()
            Uc = U.cpu().numpy()
            Vc = V.cpu().numpy()
            out[mode] = (Uc[0] + 1j * Uc[1], Vc[0] + 1j * Vc[1], keep, disc, sv)
    finally:
        engine.set_eigensolver(1)   # the default
    U0_, V0_, k0_, d0_, s0_ = out[0]
    U1_, V1_, k1_, d1_, s1_ = out[other]
    assert k0_ == k1_
    assert abs(d0_ - d1_) < 1e-15
    eps = 2.2e-16
    assert np.all(np.abs(s0_ - s1_) <= 1e-13 + 8 * eps / np.maximum(s0_, 1e-300))
    P0, P1 = U0_ @ V0_, U1_ @ V1_
    skeep = s0_[k0_ - 1]
    assert np.linalg.norm(P0 - P1) < (1e-12 + 100 * eps / skeep) * np.linalg.norm(P0)
    assert np.abs(U1_.conj().T @ U1_ - np.eye(k1_)).max() < 1e-12
    assert np.linalg.norm(P1 - M) <= 1e-6 * np.linalg.norm(M)


def test_svd_split_complex_block_structure(engine):
    """A complex two-site matrix with conserved-charge block structure (its
    Gram matrix is block-diagonal: exactly zero sub-columns): the default
    split returns finite factors that reproduce M to the discarded weight."""
    import torch
    rng = np.random.default_rng(17)
    m, n = 96, 80
    M = np.zeros((m, n), complex)
    for (r0, r1), (c0, c1) in [((0, 30), (0, 20)), ((30, 70), (20, 60)), ((70, 96), (60, 80))]:
        k = min(r1 - r0, c1 - c0)
        Ub = np.linalg.qr(rng.standard_normal((r1 - r0, k)) + 1j * rng.standard_normal((r1 - r0, k)))[0]
        Vb = np.linalg.qr(rng.standard_normal((c1 - c0, k)) + 1j * rng.standard_normal((c1 - c0, k)))[0]
        M[r0:r1, c0:c1] = (Ub * np.logspace(0, -5, k)) @ Vb.conj().T
    U, V, keep, disc, sv = engine.svd_split(1, _dev(engine, np.stack([M.real, M.imag])), m, n, 1e-12, 0, 1, 0)
    torch.cuda.synchronize()
    Uc, Vc = U.cpu().numpy(), V.cpu().numpy()
    assert np.all(np.isfinite(Uc)) and np.all(np.isfinite(Vc))
    P = (Uc[0] + 1j * Uc[1]) @ (Vc[0] + 1j * Vc[1])
    ref = np.linalg.svd(M, compute_uv=False)
    assert np.abs(sv[:keep] - ref[:keep]).max() <= 1e-13 + 8 * 2.2e-16 / ref[keep - 1]
    assert np.linalg.norm(P - M) <= (1.01 * np.sqrt(disc) + 1e-11) * np.linalg.norm(M)


@pytest.mark.parametrize("n", [150, 400])
def test_hermitian_back_transformation_matches_unmtr(engine, n):
    """The blocked complex back-transformation (mode 1) against rocSOLVER
    zunmtr (mode 2) after the same reduction: equal eigenvectors to rounding
    (several 64-reflector blocks, split inner products)."""
    import torch
    rng = np.random.default_rng(n + 5)
    Q = np.linalg.qr(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))[0]
    lam = np.maximum(1e-9 ** (np.arange(n) / 100.0), 1e-17)
    A = (Q * lam) @ Q.conj().T
    A = (A + A.conj().T) / 2
    out = {}
    for mode in (1, 2):
        w, Z = engine.herm_eig(_dev(engine, np.stack([A.real, A.imag])), mode)
        torch.cuda.synchronize()
        Zn = Z.cpu().numpy()
        out[mode] = (w.cpu().numpy(), Zn[0] + 1j * Zn[1])
    assert np.array_equal(out[1][0], out[2][0])
    assert np.abs(out[1][1] - out[2][1]).max() <= 1e-12 * np.sqrt(n)
    Z = out[1][1]
    assert np.abs(Z.conj() @ Z.T - np.eye(n)).max() <= 1e-12 * np.sqrt(n)


@pytest.mark.parametrize("field", ["real", "complex"])
@pytest.mark.parametrize("n", [40, 64])
def test_reduction_is_deterministic(engine, field, n):
    """Regression: the tridiagonal reduction once read, in its second launch
    per column, a value that workgroup 0 of the same launch overwrites, so
    about 1 call in 100-400 at orders 40-64 returned different eigenpairs
    (and DMRG energies then drifted from run to run).  Identical calls must
    now return bit-identical eigenvalues and eigenvectors, with the own
    back-transformation and with rocSOLVER's."""
    import torch
    rng = np.random.default_rng(n)
    X = rng.standard_normal((n, n + 3)) * np.logspace(0, -5, n)[:, None]
    if field == "complex":
        X = X + 1j * rng.standard_normal((n, n + 3)) * np.logspace(0, -5, n)[:, None]
    A = X @ X.conj().T
    reps = 800
    for mode in (1, 2):
        if field == "real":
            Ad = _dev(engine, A)
            run = lambda: engine.sym_eig(Ad, mode)
        else:
            Ad = _dev(engine, np.stack([A.real, A.imag]))
            run = lambda: engine.herm_eig(Ad, mode)
        w0, Z0 = run()
        bad = 0
        for _ in range(reps):
            w, Z = run()
            if not (torch.equal(w, w0) and torch.equal(Z, Z0)):
                bad += 1
        torch.cuda.synchronize()
        assert bad == 0, f"mode {mode}: {bad} of {reps} identical calls differed"


# ---- the Gram orders of the D=1024 Hubbard sweeps (min(Dl d, d Dr) up to
# 4096): eigenvalues against the constructed spectrum, the eigenvectors the
# truncation keeps (top 1024) by residual and orthonormality, and the default
# back-transformation (mode 1) against rocSOLVER's dormtr (mode 2)

def _dmrg_like_spectrum(rng, n, keep=900):
    """Eigenvalues decaying to 1e-15 of the largest at index `keep` (the
    weight cut of the 1e-12 truncation), then a noise floor of 1e-18..2e-18."""
    k = np.arange(n)
    return np.maximum(10.0 ** (-15.0 * k / keep), 1e-18 * (1.0 + rng.random(n)))


@pytest.mark.parametrize("n", [3500, 4096])
def test_large_gram_orders(engine, n):
    import time
    import torch
    rng = np.random.default_rng(n)
    Q = np.linalg.qr(rng.standard_normal((n, n)))[0]
    lam = _dmrg_like_spectrum(rng, n)
    A = (Q * lam) @ Q.T
    A = (A + A.T) / 2
    ref = np.sort(lam)
    out = {}
    for mode in (1, 2):
        t0 = time.perf_counter()
        w, Z = engine.sym_eig(_dev(engine, A), mode)
        torch.cuda.synchronize()
        out[mode] = (w.cpu().numpy(), Z.cpu().numpy())
        print("order", n, "mode", mode, "sym_eig {:.1f} ms".format(1e3 * (time.perf_counter() - t0)), flush=True)
    w, Z = out[1]
    assert np.all(np.diff(w) >= -1e-12)
    # |A| = 1 and A carries rounding of ~n eps from its construction
    err_w = np.abs(w - ref).max()
    top = Z[-1024:]
    R = A @ top.T - top.T * w[-1024:]
    err_r = np.abs(R).max()
    err_o = np.abs(Z @ Z.T - np.eye(n)).max()
    err_bt = np.abs(out[1][1][-1024:] - out[2][1][-1024:]).max()
    # leading eigenvectors (well separated) against the constructed ones
    align = np.abs(np.abs(np.einsum("kn,nk->k", Z[-64:][::-1], Q[:, :64])) - 1.0).max()
    print("order", n, "max|w - lambda|", err_w, "residual", err_r, "orthonormality", err_o,
          "mode 1 vs 2 (kept rows)", err_bt, "top-64 alignment", align, flush=True)
    assert err_w <= 1e-14 * np.sqrt(n)
    assert err_r <= 1e-12 * np.sqrt(n)
    assert err_o <= 1e-12 * np.sqrt(n)
    assert np.array_equal(out[1][0], out[2][0])
    assert err_bt <= 1e-12 * np.sqrt(n)
    assert align <= 1e-10


def test_large_hermitian_order(engine):
    """Complex Hermitian Gram matrix of order 1024 (TDVP bonds): the own
    reduction + back-transformation against the constructed spectrum and
    zheevd (mode 0)."""
    import torch
    n = 1024
    rng = np.random.default_rng(77)
    Q = np.linalg.qr(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))[0]
    lam = _dmrg_like_spectrum(rng, n, keep=500)
    A = (Q * lam) @ Q.conj().T
    A = (A + A.conj().T) / 2
    res = {}
    for mode in (1, 0):
        w, Z = engine.herm_eig(_dev(engine, np.stack([A.real, A.imag])), mode)
        torch.cuda.synchronize()
        w, Z = w.cpu().numpy(), Z.cpu().numpy()
        res[mode] = (w, Z[0] + 1j * Z[1])
    w, Z = res[1]
    err_w = np.abs(w - np.sort(lam)).max()
    R = A @ Z.T - Z.T * w
    err_r = np.abs(R).max()
    err_o = np.abs(Z.conj() @ Z.T - np.eye(n)).max()
    err_0 = np.abs(res[1][0] - res[0][0]).max()
    print("hermitian order", n, "max|w - lambda|", err_w, "residual", err_r, "orthonormality", err_o,
          "vs zheevd", err_0, flush=True)
    assert err_w <= 1e-14 * np.sqrt(n)
    assert err_r <= 1e-12 * np.sqrt(n)
    assert err_o <= 1e-12 * np.sqrt(n)
    assert err_0 <= 1e-14 * np.sqrt(n)


# ---- early stop of the reduction (mode 1) against the complete reduction
# (mode 3): the split must keep the same count and discard the same weight

@pytest.mark.parametrize("field,shape", [(0, (2400, 2560)), (0, (1800, 1700)), (0, (330, 340)),
                                         (1, (1600, 1700)), (1, (1400, 1300))])
def test_split_early_stop_matches_full_reduction(engine, field, shape):
    """A graded two-site-like matrix M = diag(r) Z diag(c) (Z Gaussian; r, c
    falling by 1e-10 across the rows / columns, as Schmidt weights of the
    neighbouring bonds grade a DMRG theta), so the Gram matrix's rounding
    noise is graded too and its trailing blocks fall far below eps |G|_F:
    the reduction of mode 1 stops once the trailing block is below 2e-15 of
    |G|_F (at about half the order here, CPU Householder); the split must
    then agree with the complete reduction (mode 3) -- keep count,
    discarded weight to 1e-14, kept singular values and the truncated
    product within the Gram-route rounding bound, isometry -- and must have
    reduced fewer columns (algorithmic bytes of the tridiagonalisation).
    Field 1: the same with a complex Gaussian Z (Hermitian reduction)."""
    import time
    import torch
    rng = np.random.default_rng(shape[0])
    m, n = shape
    r = 10.0 ** (-10.0 * np.arange(m) / m)
    c = 10.0 ** (-10.0 * np.arange(n) / n)
    Z = rng.standard_normal((m, n))
    if field:
        Z = Z + 1j * rng.standard_normal((m, n))
    M = (r[:, None] * Z) * c[None, :]
    Md = _dev(engine, np.stack([M.real, M.imag]) if field else M[None])
    out = {}
    try:
        for mode in (3, 1):
            engine.set_eigensolver(mode)
            engine.svd_split(field, Md, m, n, 1e-12, 0, 1, 0)     # warm-up (workspace)
            engine.prof_enable(True)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            U, V, keep, disc, sv = engine.svd_split(field, Md, m, n, 1e-12, 0, 1, 0)
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
            trd_bytes = engine.prof_read_bytes()[4]
            engine.prof_enable(False)
            Uh, Vh = U.cpu().numpy(), V.cpu().numpy()
            if field:
                Uh, Vh = Uh[0] + 1j * Uh[1], Vh[0] + 1j * Vh[1]
            else:
                Uh, Vh = Uh[0], Vh[0]
            out[mode] = (Uh, Vh, keep, disc, sv, dt, trd_bytes)
    finally:
        engine.set_eigensolver(1)   # the default
    Uf, Vf, kf, df, sf, tf, bf = out[3]
    Ue, Ve, ke, de, se, te, be = out[1]
    print("shape", shape, "keep", kf, ke, "discarded", df, de, "split ms full {:.1f} early {:.1f}".format(
        1e3 * tf, 1e3 * te), "reduction bytes ratio", be / bf, flush=True)
    assert ke == kf
    assert abs(de - df) < 1e-14
    eps = 2.2e-16
    assert np.all(np.abs(se[:ke] - sf[:kf]) <= 1e-13 + 8 * eps / np.maximum(sf[:kf], 1e-300))
    P0, P1 = Uf @ Vf, Ue @ Ve
    skeep = sf[kf - 1]
    assert np.linalg.norm(P0 - P1) < (1e-12 + 100 * eps / skeep) * np.linalg.norm(P0)
    assert np.abs(Ue.conj().T @ Ue - np.eye(ke)).max() < 1e-12
    # the reduction stopped early: at ~45% of the order on the large cases,
    # which saves 1 - (1 - 0.45)^3 ~ 16% of its bytes (the first columns
    # stream the largest trailing blocks) and ~40% of its time (two
    # launches per column, latency bound)
    assert be < 0.95 * bf


@pytest.mark.parametrize("field", ["real", "auto"])
def test_sweeps_early_stop_match_full_reduction(engine, field):
    """Two D=256 Hubbard sweeps (N=16, Gram orders up to 1024) from the same
    state with the early-stopping reduction (mode 1) and the complete one
    (mode 3): identical keep counts at every bond and energies to 1e-12 per
    site; real tensors, and complex ones (the reference's random complex
    start, field "auto") through the Hermitian reduction."""
    from dmrgpy_amd import algorithms as alg
    from dmrgpy_amd.chain import Chain
    from dmrgpy_amd.multioperator import hubbard_terms
    n = 16
    ch = Chain([1] * n, field=field)
    ch.set_hamiltonian(hubbard_terms(n, 1.0, 4.0))
    engine = ch.eng
    np.random.seed(5)
    psi = ch._default_mps(maxm=64)
    for m in (64, 128):
        alg.dmrg_sweep(engine, psi, ch.H, m, 1e-12, 2, 0)
    res = {}
    try:
        for mode in (3, 1):
            engine.set_eigensolver(mode)
            p = psi.copy()
            st = alg.Stats()
            engine.prof_enable(True)
            es = [alg.dmrg_sweep(engine, p, ch.H, 256, 1e-12, 2, 0, st) for _ in range(2)]
            nbytes = engine.prof_read_bytes()[4]
            engine.prof_enable(False)
            res[mode] = (es, [t[1] for t in st.truncation], [t[2] for t in st.truncation], nbytes)
    finally:
        engine.set_eigensolver(1)
    (ef, kf, df, bf), (ee, ke, de, be) = res[3], res[1]
    print("energies full", ef, "early", ee, "max keep", max(kf), "reduction bytes ratio", be / bf, flush=True)
    assert be <= bf
    assert kf == ke
    assert max(abs(a - b) for a, b in zip(ef, ee)) / n < 1e-12
    assert max(abs(a - b) for a, b in zip(df, de)) < 1e-14
