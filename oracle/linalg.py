"""Dense linear-algebra pieces of the reference hot path, restated in NumPy
(oracle -- test infrastructure only).

* ``truncate``         -- ``src/dmrgpy/pyitensor/svd.py:32-68`` (_truncate)
* ``svd_truncated``    -- ``src/dmrgpy/pyitensor/svd.py:94-151`` (Gram route + guard)
* ``lanczos_ground``   -- ``src/dmrgpy/pyitensor/dmrg.py:61-142``
* ``lanczos_expm``     -- ``src/dmrgpy/pyitensor/tdvp.py:79-187``
"""

import numpy as np
from scipy.linalg import eigh_tridiagonal

GRAM_MIN_RELATIVE_SV = 1e-7   # svd.py:84
GRAM_MIN_DIM = 16             # svd.py:88
KRYLOV_ERRGOAL = 1e-10        # tdvp.py:79


def truncate(s, cutoff, maxdim, mindim):
    """(keep, discarded_weight) for descending singular values ``s``
    (svd.py:32-68).  Drops the largest tail whose normalised weight stays
    <= cutoff, never below mindim, and always enforces maxdim."""
    n = len(s)
    p = np.asarray(s, dtype=float) ** 2
    total = p.sum()
    if total <= 0:
        return max(1, mindim), 0.0
    p = p / total
    hi = min(maxdim, n) if maxdim is not None else n
    j_upper = max(0, n - mindim)
    j_forced = min(max(0, n - hi), j_upper)
    if j_upper <= j_forced:
        j = j_forced
    else:
        p_asc = p[::-1]
        forced = float(p_asc[:j_forced].sum())
        tail = forced + np.concatenate(([0.0], np.cumsum(p_asc[j_forced:j_upper])))
        ok = np.nonzero(tail <= cutoff)[0]
        j = j_forced + (int(ok.max()) if ok.size else 0)
    keep = n - j
    return keep, float(p[keep:].sum())


def svd_truncated(mat, cutoff, maxdim, mindim=1):
    """(U_keep, S_all, Vh_keep, keep, discarded) -- svd.py:108-151.  Gram
    route when min(m,n) >= 16 and the smallest kept value is >= 1e-7 of the
    largest, else LAPACK SVD."""
    m, n = mat.shape
    if min(m, n) >= GRAM_MIN_DIM:
        left = m <= n
        gram = mat @ mat.conj().T if left else mat.conj().T @ mat
        evals, evecs = np.linalg.eigh(gram)
        evals = evals[::-1]
        evecs = evecs[:, ::-1]
        S = np.sqrt(np.clip(evals, 0.0, None))
        keep, disc = truncate(S, cutoff, maxdim if maxdim else None, max(1, mindim))
        if S[0] > 0.0 and S[keep - 1] >= GRAM_MIN_RELATIVE_SV * S[0]:
            W = evecs[:, :keep]
            if left:
                U = W
                Vh = (W.conj().T @ mat) / S[:keep, None]
            else:
                Vh = W.conj().T
                U = (mat @ W) / S[None, :keep]
            return U, S, Vh, keep, disc
    U, S, Vh = np.linalg.svd(mat, full_matrices=False)
    keep, disc = truncate(S, cutoff, maxdim if maxdim else None, max(1, mindim))
    return U[:, :keep], S, Vh[:keep, :], keep, disc


def _tridiag(alphas, betas):
    a = np.asarray(alphas, dtype=float)
    if len(a) == 1:
        return a.reshape(1, 1)
    b = np.asarray(betas, dtype=float)
    return np.diag(a) + np.diag(b, 1) + np.diag(b, -1)


def lanczos_ground(matvec, v0, niter=30, tol=1e-12, stats=None):
    """Lowest eigenpair by Lanczos with full re-orthogonalisation
    (dmrg.py:95-142): warm start v0, stop once the lowest Ritz value moves
    by < tol*max(1,|E|), at most min(niter, dim) iterations."""
    beta0 = np.linalg.norm(v0)
    if beta0 == 0:
        v0 = np.random.default_rng(0).standard_normal(v0.shape) + 0j
        beta0 = np.linalg.norm(v0)
    q = v0 / beta0
    qs = [q]
    alphas, betas = [], []
    w = matvec(q)
    alpha = np.vdot(q, w).real
    alphas.append(alpha)
    w = w - alpha * q
    prev = float(np.linalg.eigvalsh(_tridiag(alphas, betas))[0])
    m = min(niter, v0.size)
    nmv = 1
    for _ in range(1, m):
        beta = np.linalg.norm(w)
        if beta < tol:
            break
        betas.append(beta)
        qn = w / beta
        qs.append(qn)
        w = matvec(qn)
        nmv += 1
        alpha = np.vdot(qn, w).real
        alphas.append(alpha)
        w = w - alpha * qn - beta * qs[-2]
        for qk in qs[:-1]:
            w = w - np.vdot(qk, w) * qk
        cur = float(np.linalg.eigvalsh(_tridiag(alphas, betas))[0])
        if abs(cur - prev) < tol * max(1.0, abs(cur)):
            break
        prev = cur
    ev, evec = np.linalg.eigh(_tridiag(alphas, betas))
    if stats is not None:
        stats.append(nmv)
    return ev[0], np.column_stack(qs) @ evec[:, 0]


def eigh_tridiagonal_robust(alphas, betas):
    """tdvp.py:175-187: stemr, falling back to stebz."""
    try:
        return eigh_tridiagonal(alphas, betas)
    except np.linalg.LinAlgError:
        return eigh_tridiagonal(alphas, betas, lapack_driver="stebz")


def lanczos_expm(matvec, v0, coeff, niter=30, tol=1e-12, errgoal=KRYLOV_ERRGOAL, stats=None):
    """expm(coeff*A) v0 via a Lanczos basis with full re-orthogonalisation
    and Saad's a-posteriori stop beta0*beta_k*|e_k^T exp(coeff T_k) e_1| <
    errgoal (tdvp.py:82-172)."""
    beta0 = np.linalg.norm(v0)
    if beta0 == 0:
        return v0.copy()
    m = min(niter, v0.size)
    Q = np.empty((v0.size, m), dtype=complex)
    Q[:, 0] = v0 / beta0
    alphas, betas = [], []

    def first_col(k):
        if k == 1:
            return np.array([np.exp(coeff * alphas[0])], dtype=complex)
        ev, U = eigh_tridiagonal_robust(np.array(alphas[:k]), np.array(betas[:k - 1]))
        return U @ (np.exp(coeff * ev) * U[0, :])

    beta = 0.0
    for it in range(m):
        w = matvec(Q[:, it])
        alpha = np.vdot(Q[:, it], w).real
        alphas.append(alpha)
        w = w - alpha * Q[:, it]
        if it > 0:
            w = w - beta * Q[:, it - 1]
        Qk = Q[:, :it + 1]
        w = w - Qk @ (Qk.conj().T @ w)
        beta = np.linalg.norm(w)
        c = first_col(it + 1)
        if beta < tol or it == m - 1 or beta0 * beta * abs(c[-1]) < errgoal:
            if stats is not None:
                stats.append(it + 1)
            return beta0 * (Qk @ c)
        betas.append(beta)
        Q[:, it + 1] = w / beta
    raise AssertionError("unreachable")
