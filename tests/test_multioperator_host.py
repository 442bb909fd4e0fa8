"""The front end's operator algebra (dmrgpy_amd/multioperator.py) against
dense matrices (CPU): for spins, spinless fermions and native spinful
fermions, with complex coefficients, the wire format of a sum, a product, a
scalar multiple and a dagger is the sum, the matrix product, the multiple
and the conjugate transpose of the dense operators (the oracle's
``ed.sparse_hamiltonian`` builds them, Jordan-Wigner strings included).
The product and the dagger of fermion operators pin the Jordan-Wigner
transform's consistency: it must commute with both.
"""

import numpy as np
import pytest

from dmrgpy_amd.multioperator import obj2MO
from oracle import ed, sites

CASES = {
    "spin_half": ([2] * 3, ["Sx", "Sy", "Sz", "Sp", "Sm"]),
    "spin_one": ([3] * 3, ["Sx", "Sy", "Sz", "Sp", "Sm"]),
    "spinless": ([0] * 4, ["C", "Cdag", "N"]),
    "spinful": ([1] * 3, ["Cup", "Cdagup", "Cdn", "Cdagdn", "Nup", "Ndn"]),
}


def dense(codes, op):
    return ed.sparse_hamiltonian(codes, op.to_terms()).toarray()


def random_operator(rng, n, names, nterms=3, maxlen=3):
    op = 0
    for _ in range(nterms):
        k = int(rng.integers(1, maxlen + 1))
        fac = [[names[int(rng.integers(len(names)))], int(rng.integers(n))] for _ in range(k)]
        c = complex(rng.standard_normal(), rng.standard_normal())
        op = op + c * obj2MO(fac)
    return op


@pytest.mark.parametrize("case", sorted(CASES))
def test_algebra_matches_dense(case):
    codes, names = CASES[case]
    n = len(codes)
    rng = np.random.default_rng(sum(map(ord, case)))
    for _ in range(12):
        A = random_operator(rng, n, names)
        B = random_operator(rng, n, names)
        dA, dB = dense(codes, A), dense(codes, B)
        z = complex(rng.standard_normal(), rng.standard_normal())
        assert np.allclose(dense(codes, A + B), dA + dB, atol=1e-12)
        assert np.allclose(dense(codes, A - B), dA - dB, atol=1e-12)
        assert np.allclose(dense(codes, z * A), z * dA, atol=1e-12)
        assert np.allclose(dense(codes, A * B), dA @ dB, atol=1e-12)
        assert np.allclose(dense(codes, A.get_dagger()), dA.conj().T, atol=1e-12)
        assert np.allclose(dense(codes, (A * B).get_dagger()), dB.conj().T @ dA.conj().T, atol=1e-12)
        # a scalar added is that multiple of the identity
        assert np.allclose(dense(codes, A + 2.5), dA + 2.5 * np.eye(dA.shape[0]), atol=1e-12)


def test_fermion_anticommutation():
    """{c_i, c_j^dag} = delta_ij and {c_i, c_j} = 0 through the wire format,
    for the spinless chain and both flavours of the native spinful site."""
    for codes, pairs in (([0] * 4, [("C", "Cdag")]), ([1] * 3, [("Cup", "Cdagup"), ("Cdn", "Cdagdn")])):
        n = len(codes)
        dim = int(np.prod([sites.dim(c) for c in codes]))
        for i in range(n):
            for j in range(n):
                for a, adag in pairs:
                    for b, bdag in pairs:
                        ci, cjd = obj2MO([[a, i]]), obj2MO([[bdag, j]])
                        acomm = dense(codes, ci * cjd + cjd * ci)
                        expect = np.eye(dim) if (i == j and a == b) else np.zeros((dim, dim))
                        assert np.allclose(acomm, expect, atol=1e-12), (codes, a, i, bdag, j)
                        cj = obj2MO([[b, j]])
                        assert np.allclose(dense(codes, ci * cj + cj * ci), 0, atol=1e-12)
