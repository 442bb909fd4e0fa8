"""CPU pin of the early-stopping tridiagonalisation the Gram-route split
uses (the library's sytrd_lower / svd_gram in mode 1): a NumPy Householder
reduction (dsytrd, lower) of graded Gram matrices, stopped after the first
32-column panel from column 64 on whose trailing block has Frobenius norm
at most 2e-15 |G|_F, that block taken as zero and its trace counted as
discarded weight.  Against numpy's eigvalsh of the same G: the kept
eigenvalues move by at most the stop tolerance (Weyl), the reference keep
rule (oracle.linalg.truncate, svd.py:32-68) gives the same keep count and
discarded weight, and the back-transformed kept eigenvectors match.  The
GPU tests compare the library's mode 1 with its complete reduction; this
file checks the mathematics those share."""

import numpy as np
import pytest

from oracle.linalg import truncate

STOP_TOL, STOP_MIN, NB = 2e-15, 64, 32


def graded(m, n, decay, seed, cx=False):
    """M = diag(r) Z diag(c), r and c falling by 10^-decay across the rows
    and columns (Schmidt weights of the neighbouring bonds grade theta)."""
    rng = np.random.default_rng(seed)
    Z = rng.standard_normal((m, n))
    if cx:
        Z = Z + 1j * rng.standard_normal((m, n))
    r = 10.0 ** (-decay * np.arange(m) / m)
    c = 10.0 ** (-decay * np.arange(n) / n)
    M = (r[:, None] * Z) * c[None, :]
    return M / np.abs(M).max()


def tridiagonalise(G, tol):
    """Householder reduction of Hermitian G with the early stop; returns
    (d, e, reflectors [(v, tau)], s, dropped trace); s = columns reduced."""
    A = G.copy()
    n = A.shape[0]
    stop2 = tol * tol * np.vdot(A, A).real
    refl, d, e = [], [], []
    s_done, dropped = n - 1, 0.0
    for j in range(n - 1):
        x = A[j + 1:, j].copy()
        alpha = x[0]
        xn = np.linalg.norm(x)
        if xn == 0.0:
            v, tau, beta = np.zeros_like(x), 0.0, alpha
            v[0] = 1.0
        else:
            beta = -np.copysign(xn, alpha.real)   # real (zlarfg), so T is real
            v = x / (alpha - beta)
            v[0] = 1.0
            tau = (beta - alpha) / beta
        B = A[j + 1:, j + 1:]
        # H = I - tau v v^H applied on both sides: B <- H^H B H
        # B <- H^H B H, H = I - tau v v^H (zhetd2's rank-2 form)
        w = tau * (B @ v)
        w = w - 0.5 * np.conj(tau) * np.vdot(v, w) * v
        B -= np.outer(v, np.conj(w)) + np.outer(w, np.conj(v))
        d.append(A[j, j].real)
        e.append(float(np.real(beta)))
        refl.append((v, tau))
        A[j + 1:, j] = 0.0
        s = j + 1
        if (s % NB == 0 and s >= STOP_MIN and s < n - 1 and
                np.vdot(A[s:, s:], A[s:, s:]).real <= stop2):
            s_done = s
            dropped = max(0.0, float(np.trace(A[s:, s:]).real))
            break
    if s_done == n - 1:
        d.append(A[n - 1, n - 1].real)
    else:
        d.append(0.0)   # d[s]: the first diagonal entry of the dropped block
    return np.array(d), np.array(e), refl, s_done, dropped


@pytest.mark.parametrize("m,n,decay,cx", [(420, 460, 10.0, False), (380, 360, 12.0, False),
                                          (300, 320, 10.0, True)])
def test_early_stop_keeps_the_kept_spectrum(m, n, decay, cx):
    M = graded(m, n, decay, m, cx)
    G = M @ M.conj().T
    g = G.shape[0]
    lam = np.linalg.eigvalsh(G)[::-1]
    gf = np.linalg.norm(G)
    d, e, refl, s, dropped = tridiagonalise(G, STOP_TOL)
    assert s < 0.7 * g, "the reduction should stop well before the end on a graded matrix"
    # the (s+1)-order tridiagonal (real: beta is real for complex G too)
    T = np.diag(d) + np.diag(e[:s], 1) + np.diag(e[:s], -1)
    lt, Zt = np.linalg.eigh(T)
    lt, Zt = lt[::-1], Zt[:, ::-1]
    # Weyl: every eigenvalue within the stop tolerance (plus FP64 rounding)
    k = s + 1
    assert np.abs(lt - lam[:k]).max() <= STOP_TOL * gf + 50 * 2.2e-16 * lam[0]
    # the keep rule on the stopped spectrum plus the dropped trace equals the
    # keep rule on the exact spectrum
    sv = np.sqrt(np.maximum(lam, 0.0))
    keep, disc = truncate(sv, 1e-12, None, 1)
    st = np.concatenate([np.sqrt(np.maximum(lt, 0.0)), np.zeros(g - k), [np.sqrt(dropped)]])
    keep_s, disc_s = truncate(st, 1e-12, None, 1)
    assert keep_s == keep and keep < k
    assert abs(disc_s - disc) < 1e-14
    # the trace correction matters: the corrected discarded weight is much
    # closer than the dropped weight itself (without it the error would be
    # about dropped / total)
    total = lam[lam > 0].sum()
    assert abs(disc_s - disc) <= 0.25 * dropped / total + 1e-17
    # the kept eigenvectors: Q (reflectors) times the embedded tridiagonal vectors
    U = np.zeros((g, keep), dtype=G.dtype)
    U[:k] = Zt[:, :keep]
    for j in range(s - 1, -1, -1):
        v, tau = refl[j]
        U[j + 1:] -= tau * np.outer(v, np.conj(v) @ U[j + 1:])
    # residuals of the leading pairs (well-separated eigenvalues), and the
    # truncated projector's action on G against eigh's
    res = np.linalg.norm(G @ U[:, :keep // 2] - U[:, :keep // 2] * lam[:keep // 2], axis=0)
    assert res.max() <= 100 * 2.2e-16 * gf
    W = np.linalg.eigh(G)[1][:, ::-1][:, :keep]
    PG_stop = U @ (U.conj().T @ G)
    PG_ref = W @ (W.conj().T @ G)
    assert np.linalg.norm(PG_stop - PG_ref) <= 1e-12 * gf
