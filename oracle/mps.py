"""MPS containers and overlaps (oracle -- test infrastructure only).

* ``random_mps``  -- ``src/dmrgpy/pyitensor/mpsalgebra.py:315-353`` (same draw
  order from the legacy global ``np.random``: per site ``randn(*shape) +
  1j*randn(*shape)`` with the reference's boundary shapes), then the lossless
  right-moving canonicalisation ``position(N)`` and ``normalize``
  (``mpscontainer.py:81-138``).
* ``inner`` / ``inner3`` -- ``mpsalgebra.py:66-106``.
* ``apply_mpo``  -- zip-up ``mpsalgebra.py:226-302``.

An MPS is a list of arrays ``A[Dl, d, Dr]`` plus an integer orthogonality
centre (1-based, as in the reference).
"""

import numpy as np

from .linalg import svd_truncated


class MPS:
    def __init__(self, tensors, center=None):
        self.A = [np.asarray(t, dtype=complex) for t in tensors]
        self.center = center

    def copy(self):
        return MPS([a.copy() for a in self.A], self.center)

    def __len__(self):
        return len(self.A)

    def to_dense(self):
        v = self.A[0][0]                       # [d, Dr]
        for a in self.A[1:]:
            v = np.tensordot(v, a, axes=([v.ndim - 1], [0]))
        return v.reshape(-1)


def shift_right(psi, i, cutoff=0.0, maxdim=None):
    """mpscontainer.py:104-119 (0-based site i)."""
    T = psi.A[i]
    Dl, d, Dr = T.shape
    U, S, Vh, keep, _ = svd_truncated(T.reshape(Dl * d, Dr), cutoff, maxdim, 1)
    psi.A[i] = U.reshape(Dl, d, keep)
    SV = S[:keep, None] * Vh
    psi.A[i + 1] = np.tensordot(SV, psi.A[i + 1], axes=([1], [0]))


def shift_left(psi, i, cutoff=0.0, maxdim=None):
    """mpscontainer.py:121-130 (0-based site i)."""
    T = psi.A[i]
    Dl, d, Dr = T.shape
    U, S, Vh, keep, _ = svd_truncated(T.reshape(Dl, d * Dr), cutoff, maxdim, 1)
    psi.A[i] = Vh.reshape(keep, d, Dr)
    US = U * S[None, :keep]
    psi.A[i - 1] = np.tensordot(psi.A[i - 1], US, axes=([2], [0]))


def position(psi, b, cutoff=0.0, maxdim=None):
    """Move the 1-based centre to b (mpscontainer.py:81-102)."""
    while psi.center < b:
        shift_right(psi, psi.center - 1, cutoff, maxdim)
        psi.center += 1
    while psi.center > b:
        shift_left(psi, psi.center - 1, cutoff, maxdim)
        psi.center -= 1
    return psi


def random_bond_dims(dims, m):
    n = len(dims)
    left, cum = [], 1
    for d in dims[:-1]:
        cum = min(cum * d, m)
        left.append(cum)
    bonds = list(left)
    cum = 1
    for k in range(n - 2, -1, -1):
        cum = min(cum * dims[k + 1], m)
        bonds[k] = min(bonds[k], cum)
    return bonds


def random_draws(dims, m):
    """The raw Gaussian site arrays, in the reference's draw order."""
    n = len(dims)
    bonds = random_bond_dims(dims, m)
    out = []
    for i in range(n):
        shape = ([bonds[i - 1]] if i > 0 else []) + [dims[i]] + ([bonds[i]] if i < n - 1 else [])
        arr = np.random.randn(*shape) + 1j * np.random.randn(*shape)
        Dl = bonds[i - 1] if i > 0 else 1
        Dr = bonds[i] if i < n - 1 else 1
        out.append(arr.reshape(Dl, dims[i], Dr))
    return out


def random_mps(dims, m):
    psi = MPS(random_draws(dims, m), center=1)
    position(psi, len(dims))
    normalize(psi)
    return psi


def inner(a, b):
    """<a|b> (mpsalgebra.py:89-98)."""
    env = np.ones((1, 1), complex)
    for x, y in zip(a.A, b.A):
        t = np.tensordot(env, y, axes=([1], [0]))             # [a', s, b]
        env = np.tensordot(x.conj(), t, axes=([0, 1], [0, 1]))  # [b', b]
    return complex(env[0, 0])


def inner3(bra, Ws, ket):
    """<bra|M|ket> (mpsalgebra.py:99-114)."""
    env = np.ones((1, 1, 1), complex)
    for x, W, y in zip(bra.A, Ws, ket.A):
        env = extend_left(env, x, W, y)
    return complex(env[0, 0, 0])


def normalize(psi):
    nrm = inner(psi, psi).real ** 0.5
    c = psi.center or 1
    psi.A[c - 1] = psi.A[c - 1] * (1.0 / nrm)
    return psi


def extend_left(L, bra, W, ket):
    """new[b', wr, b] = L[a', wl, a] conj(bra[a', s', b']) W[wl, s, s', wr] ket[a, s, b]."""
    t = np.tensordot(L, ket, axes=([2], [0]))               # [a', wl, s, b]
    t = np.tensordot(t, W, axes=([1, 2], [0, 1]))           # [a', b, s', wr]
    t = np.tensordot(bra.conj(), t, axes=([0, 1], [0, 2]))  # [b', b, wr]
    return t.transpose(0, 2, 1)


def extend_right(R, bra, W, ket):
    """new[a', wl, a] = R[b', wr, b] conj(bra[a', s', b']) W[wl, s, s', wr] ket[a, s, b]."""
    t = np.tensordot(ket, R, axes=([2], [2]))               # [a, s, b', wr]
    t = np.tensordot(t, W, axes=([1, 3], [1, 3]))           # [a, b', wl, s']
    t = np.tensordot(t, bra.conj(), axes=([1, 3], [2, 1]))  # [a, wl, a']
    return t.transpose(2, 1, 0)


def apply_mpo(Ws, psi, cutoff=0.0, maxdim=None):
    """Zip-up K|x> (mpsalgebra.py:226-274): contract site by site, split off
    a left-orthonormal U at every cut, carry S*V forward; centre N."""
    n = len(psi)
    out = []
    left = None  # [k, wr, Dr]
    for i in range(n):
        W, X = Ws[i], psi.A[i]
        # piece[k, s', wr, Dr] = left[k, wl, Dl] W[wl, s, s', wr] X[Dl, s, Dr]
        if left is None:
            left = np.ones((1, 1, 1), complex)
        t = np.tensordot(left, X, axes=([2], [0]))             # [k, wl, s, Dr]
        t = np.tensordot(t, W, axes=([1, 2], [0, 1]))         # [k, Dr, s', wr]
        k, Dr, d, wr = t.shape
        if i == n - 1:
            out.append(t.transpose(0, 2, 3, 1).reshape(k, d, wr * Dr))
            break
        mat = t.transpose(0, 2, 3, 1).reshape(k * d, wr * Dr)
        U, S, Vh, keep, _ = svd_truncated(mat, cutoff, maxdim, 1)
        out.append(U.reshape(k, d, keep))
        left = (S[:keep, None] * Vh).reshape(keep, wr, Dr)
    return MPS(out, center=n)
