"""Two-site TDVP (oracle -- test infrastructure only).

Restates ``src/dmrgpy/pyitensor/tdvp.py``: ``_half_sweep_lr`` (:220-246),
``_half_sweep_rl`` (:249-273), ``_evolve_two_site``/``_evolve_one_site``
(:190-199) and ``tdvp_step`` (:350-368, two-site branch), and the session
drivers ``Chain.quench_tdvp`` / ``Chain.evolve_and_measure_tdvp``
(``src/dmrgpy/pyitensor/chain.py:520-564``).
"""

import numpy as np

from .dmrg import ONE_ENV, heff1_matvec, heff2_matvec, left_envs, right_envs, theta2
from .linalg import lanczos_expm, svd_truncated
from .mps import apply_mpo, extend_left, extend_right, inner, inner3, normalize


def _evolve2(L, W1, W2, R, theta, coeff, niter, stats):
    mv = heff2_matvec(L, W1, W2, R, theta.shape)
    return lanczos_expm(mv, theta.reshape(-1), coeff, niter=niter, stats=stats).reshape(theta.shape)


def _evolve1(L, W, R, C, coeff, niter, stats):
    mv = heff1_matvec(L, W, R, C.shape)
    return lanczos_expm(mv, C.reshape(-1), coeff, niter=niter, stats=stats).reshape(C.shape)


def half_sweep_lr(psi, Ws, tau, cutoff, maxdim, niter, stats=None):
    n = len(psi)
    R = right_envs(Ws, psi)
    L = {0: ONE_ENV}
    for i in range(1, n):
        th = _evolve2(L[i - 1], Ws[i - 1], Ws[i], R[i + 2], theta2(psi, i), -1j * tau, niter, stats)
        Dl, d1, d2, Dr = th.shape
        U, S, Vh, keep, _ = svd_truncated(th.reshape(Dl * d1, d2 * Dr), cutoff, maxdim, 1)
        psi.A[i - 1] = U.reshape(Dl, d1, keep)
        C = (S[:keep, None] * Vh).reshape(keep, d2, Dr)
        psi.A[i] = C
        L[i] = extend_left(L[i - 1], psi.A[i - 1], Ws[i - 1], psi.A[i - 1])
        if i < n - 1:
            psi.A[i] = _evolve1(L[i], Ws[i], R[i + 2], C, 1j * tau, niter, stats)
    psi.center = n


def half_sweep_rl(psi, Ws, tau, cutoff, maxdim, niter, stats=None):
    n = len(psi)
    L = left_envs(Ws, psi)
    R = {n + 1: ONE_ENV}
    for i in range(n - 1, 0, -1):
        th = _evolve2(L[i - 1], Ws[i - 1], Ws[i], R[i + 2], theta2(psi, i), -1j * tau, niter, stats)
        Dl, d1, d2, Dr = th.shape
        U, S, Vh, keep, _ = svd_truncated(th.reshape(Dl * d1, d2 * Dr), cutoff, maxdim, 1)
        psi.A[i] = Vh.reshape(keep, d2, Dr)
        C = (U * S[None, :keep]).reshape(Dl, d1, keep)
        psi.A[i - 1] = C
        R[i + 1] = extend_right(R[i + 2], psi.A[i], Ws[i], psi.A[i])
        if i > 1:
            psi.A[i - 1] = _evolve1(L[i - 1], Ws[i - 1], R[i + 1], C, 1j * tau, niter, stats)
    psi.center = 1


def tdvp_step(psi, Ws, dt, cutoff, maxdim, niter=50, stats=None):
    tau = dt / 2.0
    half_sweep_lr(psi, Ws, tau, cutoff, maxdim, niter, stats)
    half_sweep_rl(psi, Ws, tau, cutoff, maxdim, niter, stats)
    return psi


def quench(wf0, H, Hshift, A1, A2, nt, dt, cutoff, maxdim, niter=50):
    """chain.py:520-542: C(k dt) = <A2 psi0 | psi1(k dt)>, psi1 = A1 psi0
    evolved under H - E_GS, renormalised to its initial norm every step."""
    psi1 = apply_mpo(A1, wf0, cutoff, maxdim)
    psi2 = apply_mpo(A2, wf0, cutoff, maxdim)
    norm0 = np.sqrt(inner(psi1, psi1))
    corr = []
    for _ in range(nt):
        corr.append(inner(psi2, psi1))
        tdvp_step(psi1, Hshift, dt, cutoff, maxdim, niter)
        normalize(psi1)
        c = psi1.center or 1
        psi1.A[c - 1] = psi1.A[c - 1] * norm0
    return corr, psi1


def evolve_and_measure(wf, H, A, nt, dt, cutoff, maxdim, niter=50):
    """chain.py:544-564: measure <psi|A|psi> before each step."""
    psi = wf.copy()
    corr = []
    for _ in range(nt):
        corr.append(inner3(psi, A, psi))
        tdvp_step(psi, H, dt, cutoff, maxdim, niter)
    return corr, psi
