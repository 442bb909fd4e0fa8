"""Exact diagonalisation and closed forms (oracle -- test infrastructure only).

The reference's ED backend (``src/dmrgpy/edtk/one2many.py:4-20``,
``src/dmrgpy/algebra/algebra.py:150-172``) stops ARPACK at tol=1e-7, too loose
for 1e-10-per-site parity, so this builds the same sparse Kronecker
Hamiltonian from the terms (``oracle.mpo.resolve_term``) and solves it with
dense ``eigh`` (dim <= 4096) or ARPACK at tol=0 (machine precision).
"""

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from . import sites as _sites
from .mpo import resolve_term


def sparse_hamiltonian(codes, terms):
    dims = [_sites.dim(c) for c in codes]
    tot = int(np.prod(dims))
    H = sp.csr_matrix((tot, tot), dtype=complex)
    for coef, factors in terms:
        c, mats = resolve_term(codes, coef, factors)
        full = sp.csr_matrix(mats[0])
        for m in mats[1:]:
            full = sp.kron(full, sp.csr_matrix(m), format="csr")
        H = H + c * full
    return H


def ground_state(codes, terms):
    """(E0, psi0) of the Hamiltonian given as session terms."""
    H = sparse_hamiltonian(codes, terms)
    if H.shape[0] <= 4096:
        w, v = np.linalg.eigh(H.toarray())
        return float(w[0]), v[:, 0]
    w, v = spla.eigsh(H, k=1, which="SA", tol=0, ncv=40)
    return float(w[0]), v[:, 0]


def heisenberg_terms(n, J=1.0, Jz=None, site_offset=1):
    """Sum_i J_i (Sx Sx + Sy Sy) + Jz_i Sz Sz for i = 1..n-1 (1-based)."""
    J = np.broadcast_to(np.asarray(J, dtype=float), (n - 1,))
    Jz = J if Jz is None else np.broadcast_to(np.asarray(Jz, dtype=float), (n - 1,))
    terms = []
    for i in range(n - 1):
        a, b = i + site_offset, i + 1 + site_offset
        terms.append((J[i], [("Sx", a), ("Sx", b)]))
        terms.append((J[i], [("Sy", a), ("Sy", b)]))
        terms.append((Jz[i], [("Sz", a), ("Sz", b)]))
    return terms


def free_fermion_xx(n, J=1.0):
    """Open XX chain H = J sum (Sx Sx + Sy Sy): by Jordan-Wigner a free
    hopping chain with t = J/2, ground energy = sum of negative single-particle
    energies (the XX-chain closed form also used by the reference's
    tests/_free_fermion_reference.py:79-141)."""
    h = np.zeros((n, n))
    for i in range(n - 1):
        h[i, i + 1] = h[i + 1, i] = J / 2.0
    e = np.linalg.eigvalsh(h)
    return float(e[e < 0].sum()), h


def free_fermion_xx_szsz(n, J=1.0):
    """<Sz_i Sz_j> in the XX ground state by Wick's theorem (n even)."""
    _, h = free_fermion_xx(n, J)
    e, U = np.linalg.eigh(h)
    occ = U[:, e < 0]
    G = occ @ occ.T                                # <c_i^dag c_j>
    szsz = np.zeros((n, n))
    for i in range(n):
        for j in range(n):
            if i == j:
                szsz[i, j] = 0.25
            else:
                # (n_i - 1/2)(n_j - 1/2) with <n_i n_j> = G_ii G_jj - G_ij^2
                szsz[i, j] = (G[i, i] * G[j, j] - G[i, j] ** 2) - 0.5 * G[i, i] - 0.5 * G[j, j] + 0.25
    return szsz
