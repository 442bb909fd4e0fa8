"""Two-site DMRG sweeps (oracle -- test infrastructure only).

Restates ``src/dmrgpy/pyitensor/dmrg.py``:

* ``heff2_matvec``   -- ``two_site_heff`` (:224-248) with the planned operand
  order ``[L, W_i, W_{i+1}, R]`` of ``kernels._make_matvec_planned``
  (``kernels.py:298-410``); output in bra order ``[a', s1', s2', b']``
* ``heff1_matvec``   -- ``one_site_heff`` (:251-275)
* ``local_ground``   -- ``_local_ground_state`` (:302-333), dense when dim <= 3,
  Lanczos cap ``max(niter, 200)``
* ``update``         -- ``_apply_local_update`` (:336-352)
* ``sweep``/``dmrg`` -- ``_dmrg_one_sweep`` (:355-382) / ``dmrg`` (:385-398),
  energy = last local eigenvalue (bond 1 of the final right-to-left half)
"""

import numpy as np

from .linalg import lanczos_ground, svd_truncated
from .mps import extend_left, extend_right

ONE_ENV = np.ones((1, 1, 1), complex)


def heff2_matvec(L, W1, W2, R, shape):
    def mv(v):
        t = v.reshape(shape)
        t = np.tensordot(t, L, axes=([0], [2]))            # [s1, s2, b, a', wl]
        t = np.tensordot(t, W1, axes=([0, 4], [1, 0]))     # [s2, b, a', s1', wm]
        t = np.tensordot(t, W2, axes=([0, 4], [1, 0]))     # [b, a', s1', s2', wr]
        t = np.tensordot(t, R, axes=([0, 4], [2, 1]))      # [a', s1', s2', b']
        return t.reshape(-1)
    return mv


def heff1_matvec(L, W, R, shape):
    def mv(v):
        t = v.reshape(shape)
        t = np.tensordot(t, L, axes=([0], [2]))            # [s, b, a', wl]
        t = np.tensordot(t, W, axes=([0, 3], [1, 0]))      # [b, a', s', wr]
        t = np.tensordot(t, R, axes=([0, 3], [2, 1]))      # [a', s', b']
        return t.reshape(-1)
    return mv


def right_envs(Ws, psi):
    """{i: R_i} for i = N+1..2 (1-based), R_i covering sites i..N (dmrg.py:204-211)."""
    n = len(psi)
    env = {n + 1: ONE_ENV}
    for i in range(n, 1, -1):
        env[i] = extend_right(env[i + 1], psi.A[i - 1], Ws[i - 1], psi.A[i - 1])
    return env


def left_envs(Ws, psi):
    """{i: L_i} for i = 0..N-1, L_i covering sites 1..i (dmrg.py:214-221)."""
    n = len(psi)
    env = {0: ONE_ENV}
    for i in range(1, n):
        env[i] = extend_left(env[i - 1], psi.A[i - 1], Ws[i - 1], psi.A[i - 1])
    return env


def theta2(psi, i):
    """Two-site tensor at 1-based bond (i, i+1), layout [Dl, d1, d2, Dr]."""
    return np.tensordot(psi.A[i - 1], psi.A[i], axes=([2], [0]))


def local_ground(L, W1, W2, R, theta, niter, stats=None):
    shape = theta.shape
    mv = heff2_matvec(L, W1, W2, R, shape)
    x0 = theta.reshape(-1)
    dim = x0.size
    if dim <= 3:
        basis = np.eye(dim, dtype=complex)
        H = np.column_stack([mv(basis[:, k]) for k in range(dim)])
        w, v = np.linalg.eigh((H + H.conj().T) / 2)
        e, vec = w[0], v[:, 0]
    else:
        e, vec = lanczos_ground(mv, x0, niter=max(niter, 200), stats=stats)
    return float(np.real(e)), vec.reshape(shape)


def update(psi, i, theta, cutoff, maxdim, direction, spectra=None):
    Dl, d1, d2, Dr = theta.shape
    U, S, Vh, keep, disc = svd_truncated(theta.reshape(Dl * d1, d2 * Dr), cutoff, maxdim, 1)
    if spectra is not None:
        spectra.append((keep, disc))
    if direction == "right":
        psi.A[i - 1] = U.reshape(Dl, d1, keep)
        psi.A[i] = (S[:keep, None] * Vh).reshape(keep, d2, Dr)
        psi.center = i + 1
    else:
        psi.A[i - 1] = (U * S[None, :keep]).reshape(Dl, d1, keep)
        psi.A[i] = Vh.reshape(keep, d2, Dr)
        psi.center = i


def sweep(psi, Ws, maxdim, cutoff, niter, stats=None, spectra=None):
    n = len(psi)
    energy = None
    R = right_envs(Ws, psi)
    L = {0: ONE_ENV}
    for i in range(1, n):
        energy, th = local_ground(L[i - 1], Ws[i - 1], Ws[i], R[i + 2], theta2(psi, i), niter, stats)
        update(psi, i, th, cutoff, maxdim, "right", spectra)
        L[i] = extend_left(L[i - 1], psi.A[i - 1], Ws[i - 1], psi.A[i - 1])
    R = {n + 1: ONE_ENV}
    for i in range(n - 1, 0, -1):
        energy, th = local_ground(L[i - 1], Ws[i - 1], Ws[i], R[i + 2], theta2(psi, i), niter, stats)
        update(psi, i, th, cutoff, maxdim, "left", spectra)
        R[i + 1] = extend_right(R[i + 2], psi.A[i], Ws[i], psi.A[i])
    return energy


def dmrg(psi, Ws, maxdims, cutoff=1e-12, niter=2, callback=None):
    """Run one sweep per entry of ``maxdims`` (``Sweeps`` with per-sweep
    maxdim, ``sweeps.py:14-48``); returns the last sweep's energy."""
    energy = None
    for k, m in enumerate(maxdims):
        energy = sweep(psi, Ws, m, cutoff, niter)
        if callback is not None:
            callback(k, energy, psi)
    return energy
