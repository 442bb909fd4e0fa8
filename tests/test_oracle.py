"""Pins the CPU oracle (oracle/) before anything is checked against it.

* golden vectors written from the reference's own pure-Python backend
  (tests/golden/make_golden.py): seeded per-sweep DMRG energies, initial
  random-MPS energies (pins the draw order), vev values, TDVP correlators;
* closed forms and ED values from the reference's tests
  (tests/reference_data.py:15-17, tests/test_spin_chain.py:21-99);
* mathematics: brute-force truncation rule, Lanczos against eigh, Krylov
  exponential against scipy.linalg.expm, TDVP against dense expm, the XX
  chain against free fermions, MPO against the dense Kronecker sum.
"""

import json
import os

import numpy as np
import pytest
from scipy.linalg import expm

from oracle import dmrg, ed, linalg, mpo, mps, sites, tdvp

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "dmrg_golden.json")))
TGOLD = json.load(open(os.path.join(HERE, "golden", "tdvp_golden.json")))


def terms_of(case):
    return [(complex(*t[0]), [tuple(f) for f in t[1]]) for t in case["terms"]]


def run_case(case, nsweeps=None):
    codes = case["codes"]
    Ws = mpo.to_mpo(codes, terms_of(case))
    np.random.seed(case["seed"])
    psi = mps.random_mps([sites.dim(c) for c in codes], case["maxm"])
    e0 = mps.inner3(psi, Ws, psi).real
    es = [dmrg.sweep(psi, Ws, case["maxm"], case["cutoff"], 2)
          for _ in range(nsweeps or case["nsweeps"])]
    return e0, es, psi, Ws


FAST = [c for c in GOLD if len(c["codes"]) <= 12]


@pytest.mark.parametrize("case", FAST, ids=[c["name"] for c in FAST])
def test_golden_trajectory(case):
    e0, es, psi, Ws = run_case(case)
    n = len(case["codes"])
    assert abs(e0 - case["initial_energy"]) < 1e-12
    # first sweep starts from a non-orthonormal right basis and amplifies
    # rounding; later sweeps agree to rounding
    assert abs(es[0] - case["energies"][0]) / n < 1e-10
    assert max(abs(a - b) for a, b in zip(es[1:], case["energies"][1:])) / n < 1e-12
    for name, val in case.get("vev", []):
        fs = {"SzSz_5_6": [(1.0, [("Sz", 5), ("Sz", 6)])], "SxSx_1_2": [(1.0, [("Sx", 1), ("Sx", 2)])],
              "N_3": [(1.0, [("N", 3)])], "NupNdn_3": [(1.0, [("Nup", 3), ("Ndn", 3)])]}[name]
        A = mpo.to_mpo(case["codes"], fs)
        v = mps.inner3(psi, A, psi) / mps.inner(psi, psi)
        assert abs(v - complex(*val)) < 1e-10


@pytest.mark.slow
def test_golden_trajectory_n32_seed0():
    case = next(c for c in GOLD if c["name"] == "heis_n32_m64_s0")
    e0, es, psi, Ws = run_case(case)
    assert abs(e0 - case["initial_energy"]) < 1e-12
    assert max(abs(a - b) for a, b in zip(es, case["energies"])) / 32 < 1e-10


def test_ed_pins():
    # tests/reference_data.py:15-17 (N=10 Heisenberg ED)
    assert abs(ed.ground_state([2] * 10, ed.heisenberg_terms(10))[0] - (-4.258035207282873)) < 1e-11
    # dimer and N=3 closed forms (tests/test_spin_chain.py:21-71)
    assert abs(ed.ground_state([2] * 2, ed.heisenberg_terms(2))[0] + 0.75) < 1e-14
    assert abs(ed.ground_state([2] * 3, ed.heisenberg_terms(3))[0] + 1.0) < 1e-14
    # transverse-field Ising N=4 (tests/test_spin_chain.py:43-63): -sum Sz Sz + sum Sx
    t = [(-1.0, [("Sz", i), ("Sz", i + 1)]) for i in range(1, 4)]
    t += [(1.0, [("Sx", i)]) for i in range(1, 5)]
    assert abs(ed.ground_state([2] * 4, t)[0] - (-2.0941996592125895)) < 1e-10


def test_xx_free_fermions():
    n = 10
    e_ff, _ = ed.free_fermion_xx(n)
    e_ed = ed.ground_state([2] * n, ed.heisenberg_terms(n, 1.0, 0.0))[0]
    assert abs(e_ff - e_ed) < 1e-12
    Ws = mpo.to_mpo([2] * n, ed.heisenberg_terms(n, 1.0, 0.0))
    np.random.seed(3)
    psi = mps.random_mps([2] * n, 32)
    e = dmrg.dmrg(psi, Ws, [32] * 4)
    assert abs(e - e_ff) < 1e-10
    # Sz Sz correlator by Wick's theorem
    zz = ed.free_fermion_xx_szsz(n)
    A = mpo.to_mpo([2] * n, [(1.0, [("Sz", 3), ("Sz", 6)])])
    assert abs(mps.inner3(psi, A, psi).real - zz[2, 5]) < 1e-8


@pytest.mark.parametrize("codes,terms", [
    ([2] * 5, ed.heisenberg_terms(5)),
    ([3] * 4, ed.heisenberg_terms(4)),
    ([0] * 5, [(-1.0, [("Adag", i), ("A", i + 1)]) for i in range(1, 5)] +
     [(-1.0, [("Adag", i + 1), ("A", i)]) for i in range(1, 5)] + [(0.5, [("N", 2), ("N", 3)])]),
    ([1] * 3, [(-1.0, [("Cdagup", 1), ("Cup", 2)]), (-1.0, [("Cdagup", 2), ("Cup", 1)]),
               (4.0, [("Nup", 2), ("Ndn", 2)])]),
])
def test_to_mpo_dense(codes, terms):
    Ws = mpo.to_mpo(codes, terms)
    H = mpo.dense_matrix(codes, terms)
    assert np.abs(mpo.mpo_to_dense(Ws) - H).max() < 1e-12
    assert np.abs(H - H.conj().T).max() < 1e-14


def test_heisenberg_mpo_bond_dim_is_five():
    Ws = mpo.to_mpo([2] * 8, ed.heisenberg_terms(8))
    assert max(w.shape[3] for w in Ws[:-1]) == 5


def brute_keep(s, cutoff, maxdim, mindim):
    """Smallest keep k (>= mindim) whose dropped normalised tail stays <=
    cutoff, then capped by maxdim (svd.py:32-68 semantics)."""
    n = len(s)
    p = np.asarray(s, float) ** 2
    p = p / p.sum()
    hi = min(maxdim, n) if maxdim else n
    best = n
    for k in range(n, 0, -1):
        if k < mindim:
            break
        if p[k:].sum() <= cutoff:
            best = k
    return max(min(best, hi), min(mindim, n))


@pytest.mark.parametrize("seed", range(20))
def test_truncate_rule_brute_force(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(1, 40))
    s = np.sort(np.abs(rng.standard_normal(n)) ** rng.uniform(1, 12))[::-1]
    if seed % 4 == 0:
        s[n // 2:] = s[n // 2]          # ties at the cut
    cutoff = float(10.0 ** rng.uniform(-14, -1))
    maxdim = int(rng.integers(1, n + 3)) if seed % 3 else None
    mindim = int(rng.integers(1, 4))
    keep, disc = linalg.truncate(s, cutoff, maxdim, mindim)
    assert keep == brute_keep(s, cutoff, maxdim, mindim)
    p = s ** 2 / (s ** 2).sum()
    assert abs(disc - p[keep:].sum()) < 1e-14


def test_svd_truncated_reconstructs():
    rng = np.random.default_rng(1)
    for m, n in [(20, 30), (40, 17), (5, 9)]:
        M = rng.standard_normal((m, n)) + 1j * rng.standard_normal((m, n))
        U, S, Vh, keep, disc = linalg.svd_truncated(M, 0.0, None, 1)
        assert keep == min(m, n)
        assert np.abs((U * S[:keep]) @ Vh - M).max() < 1e-12
        np.testing.assert_allclose(S, np.linalg.svd(M, compute_uv=False), rtol=1e-12)


def test_lanczos_ground_vs_eigh():
    rng = np.random.default_rng(2)
    A = rng.standard_normal((60, 60)) + 1j * rng.standard_normal((60, 60))
    A = A + A.conj().T
    e, v = linalg.lanczos_ground(lambda x: A @ x, rng.standard_normal(60) + 0j, niter=200)
    w, U = np.linalg.eigh(A)
    assert abs(e - w[0]) < 1e-10
    assert abs(abs(np.vdot(v, U[:, 0])) - 1) < 1e-8


def test_lanczos_expm_vs_expm():
    rng = np.random.default_rng(3)
    A = rng.standard_normal((50, 50))
    A = A + A.T
    v = rng.standard_normal(50) + 1j * rng.standard_normal(50)
    out = linalg.lanczos_expm(lambda x: A @ x, v, -0.03j, niter=50)
    assert np.linalg.norm(out - expm(-0.03j * A) @ v) < 1e-9 * np.linalg.norm(v)


def test_tdvp_matches_dense_expm():
    n = 6
    terms = ed.heisenberg_terms(n)
    Ws = mpo.to_mpo([2] * n, terms)
    H = mpo.dense_matrix([2] * n, terms)
    np.random.seed(4)
    psi = mps.random_mps([2] * n, 8)
    mps.position(psi, 1)            # right-canonical start
    v0 = psi.to_dense()
    dt = 0.05
    for _ in range(4):
        tdvp.tdvp_step(psi, Ws, dt, 1e-14, 64)
    exact = expm(-1j * 4 * dt * H) @ v0
    assert abs(abs(np.vdot(psi.to_dense(), exact)) - 1) < 1e-10


def test_tdvp_golden_evolve_and_measure():
    """The oracle's evolve_and_measure (chain.py:544-564) against the
    reference's correlators: from the seeded ground state, S^x_2 applied by
    zip-up at the session's cutoff / maxm, then <S^z_4> before each step."""
    t = TGOLD[0]
    codes = t["codes"]
    terms = [(complex(*x[0]), [tuple(f) for f in x[1]]) for x in t["terms"]]
    Ws = mpo.to_mpo(codes, terms)
    np.random.seed(t["seed"])
    psi = mps.random_mps([2] * len(codes), t["maxm"])
    for _ in range(t["nsweeps"]):
        dmrg.sweep(psi, Ws, t["maxm"], t["cutoff"], 2)
    assert t["evolve"]["apply"] == "Sx_2" and t["evolve"]["measure"] == "Sz_4"
    phi = mps.apply_mpo(mpo.to_mpo(codes, [(1.0, [("Sx", 2)])]), psi, t["cutoff"], t["maxm"])
    A = mpo.to_mpo(codes, [(1.0, [("Sz", 4)])])
    corr, _ = tdvp.evolve_and_measure(phi, Ws, A, t["evolve"]["nt"], t["evolve"]["dt"], t["cutoff"], t["maxm"])
    ref = np.array([complex(*z) for z in t["evolve"]["corr"]])
    assert np.abs(np.array(corr) - ref).max() < 1e-12


def test_tdvp_golden_quench():
    t = TGOLD[0]
    codes = t["codes"]
    terms = [(complex(*x[0]), [tuple(f) for f in x[1]]) for x in t["terms"]]
    Ws = mpo.to_mpo(codes, terms)
    np.random.seed(t["seed"])
    psi = mps.random_mps([2] * len(codes), t["maxm"])
    for _ in range(t["nsweeps"]):
        dmrg.sweep(psi, Ws, t["maxm"], t["cutoff"], 2)
    EGS = (mps.inner3(psi, Ws, psi) / mps.inner(psi, psi)).real
    Hs = mpo.to_mpo(codes, terms + [(-EGS, [("Id", 1)])])
    A = mpo.to_mpo(codes, [(1.0, [("Sz", 4)])])
    corr, _ = tdvp.quench(psi, Ws, Hs, A, A, t["quench"]["nt"], t["quench"]["dt"], 1e-12, t["maxm"])
    ref = np.array([complex(*z) for z in t["quench"]["corr"]])
    assert np.abs(np.array(corr) - ref).max() < 1e-12
    # C(0) = <Sz Sz> (tests/test_time_evolution.py:497-541)
    assert abs(corr[0] - 0.25) < 1e-12
