"""Terms -> MPO and terms -> dense matrix (oracle -- test infrastructure only).

* ``resolve_term``  -- ``src/dmrgpy/pyitensor/autompo.py:36-103`` (HTerm.add's
  fermion-sign insertion sort and resolve()'s F-string threading)
* ``dense_matrix``  -- ``autompo.py:144-164``
* ``to_mpo``        -- ``src/dmrgpy/pyitensor/mpobuilder.py:40-103`` with
  ``mpsalgebra.sum_many`` (``mpsalgebra.py:143-223``): one bond-1 MPO per term,
  block-diagonal concatenation, then truncating right- and left-moving SVD
  sweeps (``mpscontainer.py:81-138``).

Terms use the session wire format ``[(coef, [(name, site_1based), ...]), ...]``.
"""

import numpy as np

from . import sites as _sites
from .linalg import svd_truncated


def resolve_term(codes, coef, factors):
    """(coef', [standard-convention matrix per site]) for one term."""
    ops = []
    c = complex(coef)
    for name, site in factors:
        idx = 0
        while idx < len(ops) and ops[idx][1] <= site:
            idx += 1
        if idx < len(ops) and _sites.is_fermionic(name):
            if sum(1 for nm, _ in ops[idx:] if _sites.is_fermionic(nm)) % 2 == 1:
                c = -c
        ops.insert(idx, (name, site))
    by_site = {}
    for name, site in ops:
        by_site.setdefault(site, []).append(name)
    n = len(codes)
    mats = []
    carry = False
    for site in range(1, n + 1):
        code = codes[site - 1]
        names = by_site.get(site)
        if names is None:
            mats.append(_sites.op(code, "F") if carry else _sites.op(code, "Id"))
            continue
        odd = sum(1 for nm in names if _sites.is_fermionic(nm)) % 2 == 1
        need_f = carry != odd
        ms = [_sites.op(code, nm) for nm in names]
        if need_f:
            ms.append(_sites.op(code, "F"))
        comb = ms[0]
        for mm in ms[1:]:
            comb = comb @ mm
        mats.append(comb)
        carry = carry != odd
    return c, mats


def dense_matrix(codes, terms):
    dims = [_sites.dim(c) for c in codes]
    tot = int(np.prod(dims))
    out = np.zeros((tot, tot), complex)
    for coef, factors in terms:
        c, mats = resolve_term(codes, coef, factors)
        full = mats[0]
        for m in mats[1:]:
            full = np.kron(full, m)
        out += c * full
    return out


def _svd_split(mat, cutoff, maxdim):
    U, S, Vh, keep, _ = svd_truncated(mat, cutoff, maxdim, 1)
    return U, S[:keep], Vh


def to_mpo(codes, terms, cutoff=1e-14, maxdim=5000):
    """MPO tensors W_i[wl, s_in, s_out, wr] (dimension-1 links at the ends)."""
    n = len(codes)
    dims = [_sites.dim(c) for c in codes]
    resolved = [resolve_term(codes, coef, f) for coef, f in terms]
    if not resolved:
        return [np.zeros((1, d, d, 1), complex) for d in dims]
    K = len(resolved)
    Ws = []
    for i in range(n):
        wl = 1 if i == 0 else K
        wr = 1 if i == n - 1 else K
        W = np.zeros((wl, dims[i], dims[i], wr), complex)
        for t, (c, mats) in enumerate(resolved):
            st = mats[i].T * (c if i == 0 else 1.0)   # (in,out) storage
            W[0 if i == 0 else t, :, :, 0 if i == n - 1 else t] = st
        Ws.append(W)
    if n == 1:
        return Ws
    # sum_many's own truncating left-to-right sweep, then to_mpo's
    # there-and-back pair (the first of which is a no-op at center n)
    Ws = _right_sweep(Ws, cutoff, maxdim)
    return compress_left(Ws, cutoff, maxdim)


def _right_sweep(Ws, cutoff, maxdim):
    Ws = [w.copy() for w in Ws]
    for i in range(len(Ws) - 1):
        wl, d1, d2, wr = Ws[i].shape
        U, S, Vh = _svd_split(Ws[i].reshape(wl * d1 * d2, wr), cutoff, maxdim)
        k = len(S)
        Ws[i] = U.reshape(wl, d1, d2, k)
        Ws[i + 1] = np.tensordot(S[:, None] * Vh, Ws[i + 1], axes=([1], [0]))
    return Ws


def compress_left(Ws, cutoff, maxdim):
    Ws = [w.copy() for w in Ws]
    for i in range(len(Ws) - 1, 0, -1):
        wl, d1, d2, wr = Ws[i].shape
        U, S, Vh = _svd_split(Ws[i].reshape(wl, d1 * d2 * wr), cutoff, maxdim)
        k = len(S)
        Ws[i] = Vh.reshape(k, d1, d2, wr)
        Ws[i - 1] = np.tensordot(Ws[i - 1], U * S[None, :], axes=([3], [0]))
    return Ws


def mpo_to_dense(Ws):
    """Dense standard-convention matrix of an MPO (small N only)."""
    T = Ws[0][0]                                   # [s_in, s_out, wr]
    din, dout = T.shape[0], T.shape[1]
    T = T.transpose(1, 0, 2)                       # [out, in, wr]
    for W in Ws[1:]:
        # T[O, I, w] x W[w, s, s', w'] -> [O s', I s, w']
        T = np.tensordot(T, W, axes=([2], [0]))    # [O, I, s, s', w']
        O, I, s, sp, w2 = T.shape
        T = T.transpose(0, 3, 1, 2, 4).reshape(O * sp, I * s, w2)
    return T[:, :, 0]
