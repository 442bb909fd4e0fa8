"""Host-side MPO builder of the backend (dmrgpy_amd/mpo.py) against the
oracle's dense Kronecker Hamiltonian (the reference's AutoMPO.dense_matrix
semantics, src/dmrgpy/pyitensor/autompo.py:144-164)."""

import numpy as np
import pytest

from dmrgpy_amd import mpo as pmpo
from dmrgpy_amd import sites as psites
from oracle import ed
from oracle import mpo as ompo
from oracle import sites as osites


def hubbard_terms(n, t=1.0, U=4.0):
    """Native spinful chain after Jordan-Wigner (front-end form)."""
    from dmrgpy_amd.multioperator import hubbard_terms as ht
    return ht(n, t, U)


CASES = {
    "heis_s12": ([2] * 6, ed.heisenberg_terms(6)),
    "heis_s1": ([3] * 4, ed.heisenberg_terms(4)),
    "heis_s32": ([4] * 3, ed.heisenberg_terms(3)),
    "xxz_random": ([2] * 5, ed.heisenberg_terms(5, np.linspace(0.6, 1.4, 4), np.linspace(1.2, 0.3, 4))),
    "fields": ([2] * 4, [(0.3, [("Sx", 1)]), (-0.7, [("Sz", 4)]), (0.2, [("Sy", 2)]),
                         (1.0, [("Sz", 1), ("Sz", 4)]), (0.5, [("Sp", 2), ("Sm", 3)]),
                         (0.5, [("Sm", 2), ("Sp", 3)])]),
    "spinless": ([0] * 5, [(-1.0, [("Adag", i), ("A", i + 1)]) for i in range(1, 5)] +
                 [(-1.0, [("Adag", i + 1), ("A", i)]) for i in range(1, 5)] +
                 [(0.5, [("N", i), ("N", i + 1)]) for i in range(1, 5)]),
    "cnames": ([0] * 4, [(-1.0, [("Cdag", 1), ("C", 3)]), (-1.0, [("Cdag", 3), ("C", 1)]),
                         (0.3, [("C", 2), ("Cdag", 4)]), (0.3, [("C", 4), ("Cdag", 2)])]),
    "identity_shift": ([2] * 3, ed.heisenberg_terms(3) + [(-2.5, [("Id", 1)])]),
    "complex_coupling": ([2] * 3, [(1j, [("Sp", 1), ("Sm", 2)]), (-1j, [("Sm", 1), ("Sp", 2)])]),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_build_mpo_dense(name):
    codes, terms = CASES[name]
    m = pmpo.build_mpo(codes, terms)
    H = ompo.dense_matrix(codes, terms)
    assert np.abs(m.to_dense() - H).max() < 1e-13


def test_heisenberg_is_real_w5():
    m = pmpo.build_mpo([2] * 10, ed.heisenberg_terms(10))
    assert not m.complex
    assert max(m.bond_dims()) == 5


def test_hubbard_is_real_w6():
    codes = [1] * 4
    terms = hubbard_terms(4)
    m = pmpo.build_mpo(codes, terms)
    assert not m.complex
    assert max(m.bond_dims()) == 6
    assert np.abs(m.to_dense() - ompo.dense_matrix(codes, terms)).max() < 1e-13


def test_complex_mpo_kept_complex():
    codes, terms = CASES["complex_coupling"]
    assert pmpo.build_mpo(codes, terms).complex


@pytest.mark.parametrize("code", [0, 1, 2, 3, 4, 5, 6])
def test_site_tables_match_oracle(code):
    # two independent restatements of the reference site tables agree
    d = psites.dim(code)
    assert d == osites.dim(code)
    names = ["Sz", "Sx", "Sy", "Sp", "Sm"] if code >= 2 else (
        ["A", "Adag", "N", "F"] if code == 0 else
        ["Aup", "Adn", "Adagup", "Adagdn", "Cdn", "Cdagdn", "Nup", "Ndn", "F", "Fup", "Sz", "Sp", "Sm"])
    for nm in names:
        assert np.array_equal(psites.matrix(code, nm), osites.op(code, nm)), nm


@pytest.mark.parametrize("code", [2, 3, 4, 5, 6])
def test_spin_algebra(code):
    S = (code - 1) / 2.0
    sx, sy, sz = (psites.matrix(code, k) for k in ("Sx", "Sy", "Sz"))
    assert np.abs(sx @ sy - sy @ sx - 1j * sz).max() < 1e-12
    assert np.abs(sx @ sx + sy @ sy + sz @ sz - S * (S + 1) * np.eye(psites.dim(code))).max() < 1e-12


def _golden(name):
    import json
    import os
    cases = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "dmrg_golden.json")))
    c = next(c for c in cases if c["name"] == name)
    return c["codes"], [(complex(*t[0]), [tuple(f) for f in t[1]]) for t in c["terms"]]


def test_front_end_jordan_wigner_matches_reference_terms():
    """The front end's JW transform gives the same operator as the terms the
    reference's own MultiOperator.to_terms produced (golden fixture)."""
    from dmrgpy_amd.multioperator import hubbard_terms as ht, obj2MO
    codes, ref_terms = _golden("hubbard_n6_m64_s6")
    diff = ed.sparse_hamiltonian(codes, ht(6, 1.0, 4.0)) - ed.sparse_hamiltonian(codes, ref_terms)
    assert abs(diff).max() < 1e-13
    codes, ref_terms = _golden("spinless_n10_m32_s5")
    H = 0
    for i in range(9):
        H = H + (-1.0) * (obj2MO([["Cdag", i], ["C", i + 1]]) + obj2MO([["Cdag", i + 1], ["C", i]]))
        H = H + 0.5 * obj2MO([["N", i], ["N", i + 1]])
    diff = ed.sparse_hamiltonian(codes, H.to_terms()) - ed.sparse_hamiltonian(codes, ref_terms)
    assert abs(diff).max() < 1e-13


def test_pruned_automaton_bond_dims_and_exact_zipup():
    """Unused start/done channels are dropped: Heisenberg gets 4,5,...,5,4,
    a single-site operator w=1, and the zip-up application with the
    product MPO (oracle zip-up on the same tensors) is exact at the same
    maxdim as the reference's compressed MPO."""
    from oracle import mps as om
    from dmrgpy_amd.multioperator import heisenberg_terms
    assert pmpo.build_mpo([2] * 8, heisenberg_terms(8)).bond_dims() == [4, 5, 5, 5, 5, 5, 4]
    assert pmpo.build_mpo([2] * 4, [(1.0, [("Sz", 2)])]).bond_dims() == [1, 1, 1]
    n = 6
    np.random.seed(6)
    psi = om.random_mps([2] * n, 8)
    om.position(psi, 1)
    terms = [(1.0, [("Sz", 3)]), (0.5, [("Sx", 2), ("Sx", 5)])]
    m = pmpo.build_mpo([2] * n, terms)
    assert m.bond_dims() == [1, 2, 2, 2, 1]
    W = [(w[0] if w.shape[0] == 1 else w[0] + 1j * w[1]).transpose(0, 3, 2, 1) for w in m.W]
    out = om.apply_mpo(W, psi, 1e-12, 8).to_dense()
    exact = ompo.dense_matrix([2] * n, terms) @ psi.to_dense()
    assert np.linalg.norm(out - exact) < 1e-13 * np.linalg.norm(exact)


from dmrgpy_amd.mpo import build_mpo  # noqa: E402


def _long_range_terms(n, lam=0.5, J=1.0):
    """sum_{i<j} J lam^(j-i-1) S_i.S_j (exponentially decaying couplings)."""
    t = []
    for i in range(1, n + 1):
        for j in range(i + 1, n + 1):
            c = J * lam ** (j - i - 1)
            for op in ("Sx", "Sy", "Sz"):
                t.append((complex(c), [(op, i), (op, j)]))
    return t


def test_compressed_mpo_long_range_is_minimal_and_exact():
    """Exponentially decaying long-range couplings: the automaton keeps one
    state per open S_i (w = O(N)); the SVD compression of
    mpobuilder.py:92-103 (bidirectional truncating sweeps at the 1e-14 build
    cutoff) reaches the minimal w = 5 with the same dense H."""
    from dmrgpy_amd.mpo import compress_mpo, hamiltonian_mpo
    n = 9
    codes = [2] * n
    terms = _long_range_terms(n)
    A = build_mpo(codes, terms)
    C = compress_mpo(A)
    H = hamiltonian_mpo(codes, terms)
    assert max(A.bond_dims()) > 5
    assert max(C.bond_dims()) == 5
    assert max(H.bond_dims()) == 5 and not H.complex
    Hd = A.to_dense()
    assert np.abs(C.to_dense() - Hd).max() < 1e-12 * np.abs(Hd).max()
    from oracle import mpo as omp
    ref = omp.dense_matrix(codes, terms)       # Kronecker sum of the terms (oracle)
    assert np.abs(C.to_dense() - ref).max() < 1e-12 * np.abs(ref).max()


def test_compressed_mpo_keeps_nearest_neighbour_automaton():
    """Nearest-neighbour chains keep the (sparse, minimal) automaton MPO;
    its compression would not lower w."""
    from dmrgpy_amd.mpo import compress_mpo, hamiltonian_mpo
    from dmrgpy_amd.multioperator import heisenberg_terms, hubbard_terms
    for codes, terms, w in (([2] * 10, heisenberg_terms(10), 5), ([1] * 6, hubbard_terms(6, 1.0, 4.0), 6)):
        H = hamiltonian_mpo(codes, terms)
        assert H.coupling_sites and max(H.bond_dims()) == w
        C = compress_mpo(build_mpo(codes, terms))
        assert max(C.bond_dims()) == w
        D = build_mpo(codes, terms).to_dense()
        assert np.abs(C.to_dense() - D).max() < 1e-12 * np.abs(D).max()


def test_term_gradients_refuse_compressed_mpo():
    import pytest
    from dmrgpy_amd.gradients import contract_term_gradients
    from dmrgpy_amd.mpo import compress_mpo
    C = compress_mpo(build_mpo([2] * 6, _long_range_terms(6)))
    with pytest.raises(ValueError):
        contract_term_gradients([w for w in C.W], C, 3)
