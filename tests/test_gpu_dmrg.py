"""End-to-end parity of the MI355X session (dmrgpy_amd.chain.Chain) with the
reference's pure-Python backend (GPU).

The golden fixtures in tests/golden/ were written by the reference's own
``pyitensor.chain.Chain`` from seeded random MPS; the session here draws the
same MPS from the same seed (legacy ``np.random``) and must reproduce every
sweep's energy to 1e-10 per site in FP64.  Further pins: ED at N=16 (the
N=32 configuration's ED check), vev values, TDVP correlators, and the
reduced-precision paths against the FP64 path (|dE/N| < 1e-9).
"""

import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "dmrg_golden.json")))
TGOLD = json.load(open(os.path.join(HERE, "golden", "tdvp_golden.json")))


def terms_of(case):
    return [(complex(*t[0]), [tuple(f) for f in t[1]]) for t in case["terms"]]


def make_session(codes, terms, maxm, cutoff=1e-12, precision="fp64", field="auto"):
    from dmrgpy_amd.chain import Chain
    ch = Chain(codes, precision=precision, field=field)
    ch.set_sweep_params(maxm, 1, cutoff, 0.0)
    ch.set_verbose(False)
    ch.set_mpomaxm(max(maxm, 5000))
    ch.set_hamiltonian(terms)
    return ch


@pytest.mark.parametrize("case", GOLD, ids=[c["name"] for c in GOLD])
def test_golden_sweeps_fp64(case):
    from dmrgpy_amd import algorithms as alg
    print("case", case["name"], flush=True)
    ch = make_session(case["codes"], terms_of(case), case["maxm"], case["cutoff"])
    np.random.seed(case["seed"])
    ch.wf0 = ch._default_mps()
    e0 = alg.expectation(ch.eng, ch.wf0, ch.H, ch.wf0).real
    assert abs(e0 - case["initial_energy"]) < 1e-11
    n = len(case["codes"])
    es = [ch.gs_energy() for _ in range(case["nsweeps"])]
    err = np.abs(np.array(es) - np.array(case["energies"])) / n
    # The first sweep starts from the canonicalised random MPS (centre N):
    # its left-to-right half runs on a non-orthonormal right basis, exactly
    # as the reference does, which amplifies rounding differences (the
    # oracle itself differs from the reference there by up to 4e-11/site).
    # Every later sweep must agree to rounding.
    print("case", case["name"], "first-sweep |dE|/N", err[0], "later sweeps max", err[1:].max(), flush=True)
    assert err[0] < 1e-10, err
    assert err[1:].max() < 1e-11, err
    for name, val in case.get("vev", []):
        fs = {"SzSz_5_6": [(1.0, [("Sz", 5), ("Sz", 6)])], "SxSx_1_2": [(1.0, [("Sx", 1), ("Sx", 2)])],
              "SzSz_16_17": [(1.0, [("Sz", 16), ("Sz", 17)])], "N_3": [(1.0, [("N", 3)])],
              "NupNdn_3": [(1.0, [("Nup", 3), ("Ndn", 3)])]}[name]
        v = ch.vev(fs, ch.gs_wavefunction())
        assert abs(v - complex(*val)) < 1e-9, (name, v, val)


def test_ed_n16():
    from oracle import ed
    n = 16
    terms = ed.heisenberg_terms(n)
    e_ed = ed.ground_state([2] * n, terms)[0]
    ch = make_session([2] * n, terms, 128)
    ch.set_sweep_params(128, 6, 1e-12, 0.0)
    np.random.seed(0)
    e = ch.gs_energy()
    assert abs(e - e_ed) / n < 1e-10


def test_tdvp_golden():
    t = TGOLD[0]
    terms = terms_of(t)
    ch = make_session(t["codes"], terms, t["maxm"], t["cutoff"])
    np.random.seed(t["seed"])
    ch.wf0 = ch._default_mps()
    for _ in range(t["nsweeps"]):
        ch.gs_energy()
    gs = ch.gs_wavefunction().copy()
    sz4 = [(1.0, [("Sz", 4)])]
    corr, _ = ch.quench_tdvp(terms, sz4, sz4, t["quench"]["nt"], t["quench"]["dt"])
    ref = np.array([complex(*z) for z in t["quench"]["corr"]])
    assert np.abs(np.array(corr) - ref).max() < 1e-10
    ch.wf0 = gs.copy()
    psi = ch.apply_operator([(1.0, [("Sx", 2)])], gs.copy())
    corr2, _ = ch.evolve_and_measure_tdvp(terms, sz4, psi, t["evolve"]["nt"], t["evolve"]["dt"])
    ref2 = np.array([complex(*z) for z in t["evolve"]["corr"]])
    assert np.abs(np.array(corr2) - ref2).max() < 1e-10


def _heisenberg32_energies(schedules, seed=1):
    from dmrgpy_amd.multioperator import heisenberg_terms
    n, D = 32, 64
    terms = heisenberg_terms(n)
    out = {}
    for name, sched in schedules.items():
        ch = make_session([2] * n, terms, D, precision="fp64", field="real")
        ch.set_sweep_params(D, len(sched), 1e-12, 0.0)
        ch.precision_schedule = sched
        np.random.seed(seed)
        out[name] = ch.gs_energy()
    return n, out


def test_fp32_vs_fp64():
    """Six FP32-contraction sweeps (FP64 Lanczos/SVD), final FP64 Rayleigh
    quotient, against six FP64 sweeps from the same seeded MPS: the
    reduced-precision target |dE/N| < 1e-9."""
    n, e = _heisenberg32_energies({"fp64": ["fp64"] * 6, "fp32": ["fp32"] * 6})
    err = abs(e["fp32"] - e["fp64"]) / n
    print("fp32 vs fp64 |dE|/N", err, flush=True)
    assert err < 1e-9, e


# Measured accuracy of sweeps that stay in split-BF16x3 to the end (FP64
# Lanczos / SVD and the final FP64 Rayleigh quotient): the operator's ~1e-5
# relative rounding stops the local solves at the 1e-4 Ritz-residual floor, so
# this path does NOT reach the 1e-9 per-site target on its own; the bound
# below is what it reaches on this chain (N=32, D=64), with a margin.
BF16X3_MEASURED_BOUND = 1e-8


def test_bf16x3_sweeps_accuracy():
    n, e = _heisenberg32_energies({"fp64": ["fp64"] * 6, "bf16x3": ["bf16x3"] * 6})
    err = abs(e["bf16x3"] - e["fp64"]) / n
    print("bf16x3 (all sweeps) vs fp64 |dE|/N", err, flush=True)
    assert e["bf16x3"] > e["fp64"] - 1e-9 * n      # variational: never below the FP64 energy
    assert err < BF16X3_MEASURED_BOUND, e


def test_bf16x3_then_fp32_schedule():
    """BF16x3 for the early sweeps, FP32 for the last two (the mixed
    schedule): reaches the 1e-9 per-site target without an FP64 sweep."""
    n, e = _heisenberg32_energies({"fp64": ["fp64"] * 6, "mixed": ["bf16x3"] * 4 + ["fp32"] * 2})
    err = abs(e["mixed"] - e["fp64"]) / n
    print("bf16x3x4 + fp32x2 vs fp64 |dE|/N", err, flush=True)
    assert err < 1e-9, e


def test_bf16x3_then_fp64_schedule():
    """BF16x3 sweeps finished by one FP64 sweep (masks BF16x3 rounding; kept
    as a separately named check of the schedule mechanics)."""
    n, e = _heisenberg32_energies({"fp64": ["fp64"] * 6, "mixed": ["bf16x3"] * 5 + ["fp64"]})
    err = abs(e["mixed"] - e["fp64"]) / n
    print("bf16x3x5 + fp64 vs fp64 |dE|/N", err, flush=True)
    assert err < 1e-9, e


def test_front_end_spin_chain():
    from dmrgpy_amd.frontend import Spin_Chain
    n = 10
    sc = Spin_Chain(["S=1/2"] * n, itensor_version="mi355x")
    h = 0
    for i in range(n - 1):
        h = h + sc.SS(i, i + 1)
    sc.set_hamiltonian(h)
    sc.maxm = 40
    sc.nsweeps = 6
    e = sc.gs_energy()
    assert abs(e - (-4.258035207282873)) < 1e-9    # tests/reference_data.py:15-17
    assert abs(sc.vev(sc.Sz[4] * sc.Sz[5]).imag) < 1e-12
    # SU(2): <S_i.S_{i+1}> summed = E
    ss = sum(sc.vev(sc.SS(i, i + 1)).real for i in range(n - 1))
    assert abs(ss - e) < 1e-9


def test_front_end_hubbard_u0_free_fermions():
    from dmrgpy_amd.frontend import Spinful_Fermionic_Chain_Native
    n = 6
    fc = Spinful_Fermionic_Chain_Native(n, itensor_version="mi355x")
    h = 0
    for i in range(n - 1):
        h = h - (fc.Cdagup[i] * fc.Cup[i + 1] + fc.Cdagup[i + 1] * fc.Cup[i])
        h = h - (fc.Cdagdn[i] * fc.Cdn[i + 1] + fc.Cdagdn[i + 1] * fc.Cdn[i])
    fc.set_hamiltonian(h)
    fc.maxm = 64
    fc.nsweeps = 8
    e = fc.gs_energy()
    eps = np.linalg.eigvalsh(-(np.eye(n, k=1) + np.eye(n, k=-1)))
    assert abs(e - 2 * eps[eps < 0].sum()) < 1e-9


def test_errors_are_python_exceptions():
    from dmrgpy_amd.chain import Chain
    ch = Chain([2] * 4)
    with pytest.raises(RuntimeError):
        ch.gs_energy()
    with pytest.raises(RuntimeError):
        ch.gs_wavefunction()


@pytest.mark.parametrize("center,D", [(6, 8), (1, 8), (6, 4)])
def test_tdvp_step_vs_oracle(center, D):
    """One TDVP step on the device against the oracle's from the same MPS
    (dense states compared, so gauge does not matter)."""
    from dmrgpy_amd import algorithms as alg
    from dmrgpy_amd.engine import get_engine
    from dmrgpy_amd.mpo import build_mpo
    from dmrgpy_amd.mps import DeviceMPS
    from dmrgpy_amd.multioperator import heisenberg_terms
    from oracle import mpo as omp
    from oracle import mps as om
    from oracle import tdvp as ot
    n = 6
    terms = heisenberg_terms(n)
    np.random.seed(5)
    psi = om.random_mps([2] * n, 8)
    om.position(psi, center)
    ref = psi.copy()
    ot.tdvp_step(ref, omp.to_mpo([2] * n, terms), 0.05, 1e-12, D, 50)
    eng = get_engine(0)
    dpsi = DeviceMPS.from_numpy(psi.A, eng.device, field=1, center=psi.center)
    alg.tdvp_step(eng, dpsi, build_mpo([2] * n, terms), 0.05, 1e-12, D, 50)
    v, r = dpsi.to_dense(), ref.to_dense()
    assert np.linalg.norm(v - r) < 1e-10 * np.linalg.norm(r)


def test_apply_mpo_vs_oracle():
    from dmrgpy_amd import algorithms as alg
    from dmrgpy_amd.engine import get_engine
    from dmrgpy_amd.mpo import build_mpo
    from dmrgpy_amd.mps import DeviceMPS
    from oracle import mpo as omp
    from oracle import mps as om
    n = 6
    np.random.seed(6)
    psi = om.random_mps([2] * n, 8)
    om.position(psi, 1)
    terms = [(1.0, [("Sz", 3)]), (0.5, [("Sx", 2), ("Sx", 5)])]
    ref = om.apply_mpo(omp.to_mpo([2] * n, terms), psi, 1e-12, 8)
    eng = get_engine(0)
    dpsi = DeviceMPS.from_numpy(psi.A, eng.device, field=1, center=1)
    out = alg.apply_mpo(eng, build_mpo([2] * n, terms), dpsi, 1e-12, 8)
    v, r = out.to_dense(), ref.to_dense()
    assert np.linalg.norm(v - r) < 1e-11 * np.linalg.norm(r)
    assert out.center == n


def test_reduced_precision_sweeps_keep_isometries():
    """FP32 contractions leave a noise floor in theta, so cutoff 1e-12 keeps
    singular values near the Gram-route guard.  The tensors a sweep leaves
    behind must still be isometries (B B^dag = I), otherwise H_eff of later
    bonds is no projection of H and FP64 Lanczos finds Ritz values below the
    ground state.  Pinned by ED: every energy stays above E_0."""
    from dmrgpy_amd import algorithms as alg
    from dmrgpy_amd.chain import RES_TOL
    from dmrgpy_amd.multioperator import hubbard_terms
    from oracle import ed
    n = 8
    terms = hubbard_terms(n, 1.0, 4.0)
    e_ed = ed.ground_state([1] * n, terms)[0]
    ch = make_session([1] * n, terms, 256, precision="fp32", field="real")
    np.random.seed(3)
    psi = ch._default_mps(maxm=64)
    for _ in range(3):
        alg.dmrg_sweep(ch.eng, psi, ch.H, 256, 1e-12, 2, 1, None, res_tol=RES_TOL[1])
    assert psi.center == 1
    for i in range(2, n + 1):
        B = psi.A(i).double().cpu().numpy()
        B = B[0] + 1j * B[1] if B.shape[0] == 2 else B[0]
        Bm = B.reshape(B.shape[0], -1)
        assert np.abs(Bm @ Bm.conj().T - np.eye(Bm.shape[0])).max() < 1e-12, i
    stats = alg.Stats()
    e64 = alg.dmrg_sweep(ch.eng, psi, ch.H, 256, 1e-12, 2, 0, stats)
    e_ray = (alg.expectation(ch.eng, psi, ch.H, psi) / alg.overlap(ch.eng, psi, psi)).real
    assert e64 > e_ed - 1e-9 and e_ray > e_ed - 1e-9
    assert abs(e_ray - e_ed) / n < 1e-8


def test_excited_states_vs_ed():
    """Overlap-penalty excited states (chain.py:253-275, dmrg.py:520-681)
    against exact diagonalisation: random-bond XXZ chain with random
    longitudinal fields (non-degenerate spectrum), N=8 at full bond
    dimension; energies, vanishing energy fluctuations, mutual orthogonality."""
    from dmrgpy_amd import algorithms as alg
    from dmrgpy_amd.chain import Chain
    from dmrgpy_amd.gradients import xxz_terms
    from oracle import ed
    rng = np.random.default_rng(23)
    n = 8
    terms = xxz_terms(rng.uniform(0.5, 1.5, n - 1), rng.uniform(0.5, 1.5, n - 1))
    terms += [(complex(h), [("Sz", i + 1)]) for i, h in enumerate(rng.uniform(-0.4, 0.4, n))]
    w = np.linalg.eigvalsh(ed.sparse_hamiltonian([2] * n, terms).toarray())
    ch = Chain([2] * n, field="real")
    ch.set_sweep_params(16, 10, 1e-12, 0.0)
    ch.set_hamiltonian(terms)
    np.random.seed(1)
    energies, fluct, wfs = ch.excited_states(3, scale_lagrange=2.0)
    assert np.abs(np.array(energies) - w[:3]).max() < 1e-8, (energies, w[:3])
    # the session's fluctuation uses H|psi> by zip-up truncated at maxm, as
    # the reference does (chain.py:1225-1231); the exact variance is dense
    assert max(fluct) < 1e-6, fluct
    Hd = ed.sparse_hamiltonian([2] * n, terms).toarray()
    for wf in wfs:
        v = wf.to_dense().reshape(-1)
        v = v / np.linalg.norm(v)
        hv = Hd @ v
        assert abs(np.vdot(hv, hv) - np.vdot(v, hv) ** 2) < 1e-9
    for a in range(3):
        for b in range(a):
            assert abs(alg.overlap(ch.eng, wfs[a], wfs[b])) < 1e-6


def test_excited_states_gram_schmidt_vs_ed():
    """excited_states(do_gram_schmidt=True) (chain.py:253-275,1233-1241):
    the penalty states orthogonalised by MPS sums; energies against ED and
    mutual overlaps at the rounding of the sums (N=8, full bond dimension)."""
    from dmrgpy_amd import algorithms as alg
    from dmrgpy_amd.chain import Chain
    from dmrgpy_amd.gradients import xxz_terms
    from oracle import ed
    rng = np.random.default_rng(23)
    n = 8
    terms = xxz_terms(rng.uniform(0.5, 1.5, n - 1), rng.uniform(0.5, 1.5, n - 1))
    terms += [(complex(h), [("Sz", i + 1)]) for i, h in enumerate(rng.uniform(-0.4, 0.4, n))]
    w = np.linalg.eigvalsh(ed.sparse_hamiltonian([2] * n, terms).toarray())
    ch = Chain([2] * n, field="real")
    ch.set_sweep_params(16, 10, 1e-12, 0.0)
    ch.set_hamiltonian(terms)
    np.random.seed(1)
    energies, fluct, wfs = ch.excited_states(3, scale_lagrange=2.0, do_gram_schmidt=True)
    print("Gram-Schmidt excited states", energies, "ED", w[:3], flush=True)
    assert np.abs(np.array(energies) - w[:3]).max() < 1e-8, (energies, w[:3])
    ov = max(abs(alg.overlap(ch.eng, wfs[a], wfs[b])) for a in range(3) for b in range(a))
    print("max overlap after Gram-Schmidt", ov, flush=True)
    assert ov < 1e-10
    for wf in wfs:
        assert abs(alg.overlap(ch.eng, wf, wf) - 1.0) < 1e-12


@pytest.mark.parametrize("field", [0, 1])
def test_mps_sum_vs_dense(field):
    """Direct sums of MPS (mpsalgebra.py:143-232) against the dense sum of
    the states: two and three operands, without truncation exactly (1e-12),
    and with maxdim the best truncation a left-to-right SVD sweep gives."""
    from dmrgpy_amd import algorithms as alg
    from dmrgpy_amd.engine import get_engine
    eng = get_engine(0)
    np.random.seed(4)
    dims = [2, 3, 2, 2, 3, 2]
    a = alg.random_mps(eng, dims, 5, field)
    b = alg.random_mps(eng, dims, 4, field)
    c = alg.scaled(alg.random_mps(eng, dims, 3, field), 0.7 - (0.2j if field else 0.0))
    s2 = alg.mps_sum(eng, a, b)
    s3 = alg.mps_sum_many(eng, [a, b, c])
    va, vb, vc = a.to_dense(), b.to_dense(), c.to_dense()
    e2 = np.linalg.norm(s2.to_dense() - (va + vb)) / np.linalg.norm(va + vb)
    e3 = np.linalg.norm(s3.to_dense() - (va + vb + vc)) / np.linalg.norm(va + vb + vc)
    print("mps sum errors", e2, e3, "bond dims", s2.bond_dims(), s3.bond_dims(), flush=True)
    assert e2 < 1e-12 and e3 < 1e-12
    assert s2.center == len(dims)
    # lossless sums stay within the exact bond dimensions of the sum
    assert max(s3.bond_dims()) <= 5 + 4 + 3
    t = alg.mps_sum(eng, a, b, 0.0, 3)
    assert max(t.bond_dims()) <= 3
    assert np.linalg.norm(t.to_dense() - (va + vb)) < np.linalg.norm(va + vb)


def test_generalized_eigenvalue_vs_dense():
    """gs_energy_generalized (chain.py:230-251): smallest lambda of
    H psi = lambda A psi with A = 1 + sum_i c_i Sz_i^2-like positive terms,
    against scipy.linalg.eigh(H, A) on the dense matrices (N=8)."""
    from scipy.linalg import eigh
    from dmrgpy_amd.chain import Chain
    from dmrgpy_amd.multioperator import heisenberg_terms
    from oracle import ed
    n = 8
    rng = np.random.default_rng(29)
    terms_h = heisenberg_terms(n)
    terms_a = [(1.0, [("Id", 1)])] + [(float(c), [("Sz", i + 1), ("Sz", i + 1)])
                                     for i, c in enumerate(rng.uniform(0.2, 0.8, n))]
    H = ed.sparse_hamiltonian([2] * n, terms_h).toarray()
    A = ed.sparse_hamiltonian([2] * n, terms_a).toarray()
    lam_ref = eigh(H, A, eigvals_only=True)[0]
    ch = Chain([2] * n, field="real")
    ch.set_sweep_params(16, 12, 1e-12, 0.0)
    ch.set_hamiltonian(terms_h)
    np.random.seed(5)
    lam = ch.gs_energy_generalized(terms_a)
    assert abs(lam - lam_ref) < 1e-9, (lam, lam_ref)


def test_long_range_hamiltonian_uses_compressed_mpo():
    """Exponentially decaying long-range Heisenberg couplings: the session
    sweeps with the SVD-compressed MPO (w = 5 instead of the automaton's
    O(N)) and reaches the ED ground-state energy (N=10)."""
    from dmrgpy_amd.chain import Chain
    from oracle import ed
    n = 10
    terms = []
    for i in range(1, n + 1):
        for j in range(i + 1, n + 1):
            for op in ("Sx", "Sy", "Sz"):
                terms.append((complex(0.6 ** (j - i - 1)), [(op, i), (op, j)]))
    e_ed = ed.ground_state([2] * n, terms)[0]
    ch = Chain([2] * n, field="real")
    ch.set_sweep_params(64, 8, 1e-12, 0.0)
    ch.set_hamiltonian(terms)
    assert max(ch.H.bond_dims()) == 5
    np.random.seed(3)
    e = ch.gs_energy()
    print("long-range: E", e, "ED", e_ed, flush=True)
    assert abs(e - e_ed) / n < 1e-10


def test_front_end_interleaved_spinful_chain_matches_native():
    """Spinful_Fermionic_Chain (fermionchain.py:210-233: 2n interleaved
    spinless sites) and Spinful_Fermionic_Chain_Native (n d=4 sites) give
    the same Hubbard ground-state energy, equal to ED, and the same
    <Sz_i Sz_j> and double occupancy (N=4 orbitals, t=1, U=4)."""
    from dmrgpy_amd.frontend import Spinful_Fermionic_Chain, Spinful_Fermionic_Chain_Native
    from dmrgpy_amd.multioperator import hubbard_terms
    from oracle import ed
    n = 4
    out = {}
    for cls in (Spinful_Fermionic_Chain, Spinful_Fermionic_Chain_Native):
        fc = cls(n, itensor_version="mi355x")
        h = 0
        for i in range(n - 1):
            h = h - (fc.Cdagup[i] * fc.Cup[i + 1] + fc.Cdagup[i + 1] * fc.Cup[i])
            h = h - (fc.Cdagdn[i] * fc.Cdn[i + 1] + fc.Cdagdn[i + 1] * fc.Cdn[i])
        for i in range(n):
            h = h + 4.0 * (fc.Nup[i] - 0.5) * (fc.Ndn[i] - 0.5)
        fc.set_hamiltonian(h)
        fc.maxm = 64
        fc.nsweeps = 8
        e = fc.gs_energy()
        zz = fc.vev(fc.Sz[0] * fc.Sz[1]).real
        docc = sum(fc.vev(fc.Nup[i] * fc.Ndn[i]).real for i in range(n)) / n
        out[cls.__name__] = (e, zz, docc)
    e_ed = ed.ground_state([1] * n, hubbard_terms(n, 1.0, 4.0))[0]
    print("interleaved vs native", out, "ED", e_ed, flush=True)
    (e1, z1, d1), (e2, z2, d2) = out["Spinful_Fermionic_Chain"], out["Spinful_Fermionic_Chain_Native"]
    assert abs(e1 - e_ed) < 1e-9 and abs(e2 - e_ed) < 1e-9
    assert abs(z1 - z2) < 1e-8 and abs(d1 - d2) < 1e-8
