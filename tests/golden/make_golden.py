"""Write the golden fixtures in this folder from the reference's pure-Python
backend (dmrgpy ``pyitensor``), driven through its session class
``pyitensor.chain.Chain`` exactly as dmrgpy's front end drives it
(``src/dmrgpy/groundstate.py:27-58``): set_sweep_params -> set_hamiltonian ->
gs_energy, one sweep per call so every sweep's energy is recorded.

Only quantities that nothing else pins are written here: the seeded
finite-bond-dimension DMRG energy trajectories, the seeded initial-MPS energy
(pins the random-MPS draw order), TDVP correlators on a seeded ground state,
and ``vev`` values.  ED / closed forms pin the rest inside the tests.

Run (in a container that has the reference checked out at /root/reference):

    python tests/golden/make_golden.py            # writes tests/golden/*.json

The reference never travels with the repository: only the JSON vectors do.
"""

import json
import os
import sys
import time

import numpy as np

REF = "/root/reference/src"
HERE = os.path.dirname(os.path.abspath(__file__))


def _c(z):
    z = complex(z)
    return [z.real, z.imag]


def heis_terms(n, J=None, Jz=None):
    J = [1.0] * (n - 1) if J is None else list(J)
    Jz = J if Jz is None else list(Jz)
    t = []
    for i in range(1, n):
        t.append((J[i - 1], [("Sx", i), ("Sx", i + 1)]))
        t.append((J[i - 1], [("Sy", i), ("Sy", i + 1)]))
        t.append((Jz[i - 1], [("Sz", i), ("Sz", i + 1)]))
    return t


def spinless_terms(n, t=1.0, V=0.5):
    """-t sum (c^dag_i c_{i+1} + h.c.) + V sum n_i n_{i+1}, Jordan-Wigner
    transformed by the reference's own MultiOperator machinery."""
    from dmrgpy import multioperator as mo
    H = 0
    for i in range(n - 1):
        H = H + (-t) * (mo.obj2MO([["Cdag", i], ["C", i + 1]]) + mo.obj2MO([["Cdag", i + 1], ["C", i]]))
        H = H + V * mo.obj2MO([["N", i], ["N", i + 1]])
    return H.to_terms()


def hubbard_terms(n, t=1.0, U=4.0):
    """Native spinful chain: -t sum_sigma (c^dag c + h.c.) + U sum (n_up-1/2)(n_dn-1/2)."""
    from dmrgpy import multioperator as mo
    H = 0
    for i in range(n - 1):
        for up, dn in (("Cdagup", "Cup"), ("Cdagdn", "Cdn")):
            H = H + (-t) * (mo.obj2MO([[up, i], [dn, i + 1]]) + mo.obj2MO([[up, i + 1], [dn, i]]))
    for i in range(n):
        nu = mo.obj2MO([["Nup", i]]) - 0.5 * mo.obj2MO([["Id", i]])
        nd = mo.obj2MO([["Ndn", i]]) - 0.5 * mo.obj2MO([["Id", i]])
        H = H + U * nu * nd
    return H.to_terms()


def dmrg_case(name, codes, terms, maxm, nsweeps, seed, vev_terms=None, cutoff=1e-12):
    from dmrgpy.pyitensor.chain import Chain
    from dmrgpy.pyitensor.mpsalgebra import inner
    ch = Chain(codes)
    ch.set_sweep_params(maxm, 1, cutoff, 0.0)
    ch.set_verbose(False)
    ch.set_mpomaxm(max(maxm, 5000))
    ch.set_hamiltonian(terms)
    np.random.seed(seed)
    ch.wf0 = ch._default_mps()
    e_init = inner(ch.wf0, ch.H, ch.wf0).real
    t0 = time.time()
    energies = [float(ch.gs_energy()) for _ in range(nsweeps)]
    wall = time.time() - t0
    out = {"name": name, "codes": list(codes), "terms": [[_c(c), [list(f) for f in fs]] for c, fs in terms],
           "maxm": maxm, "nsweeps": nsweeps, "seed": seed, "cutoff": cutoff,
           "initial_energy": float(e_init), "energies": energies, "wall_s": wall,
           "host": "reference pyitensor on this container's CPU"}
    if vev_terms:
        wf = ch.gs_wavefunction()
        out["vev"] = [[vt_name, _c(ch.vev(vt, wf))] for vt_name, vt in vev_terms]
    print(name, energies[-1], "%.1fs" % wall, flush=True)
    return out, ch


def main():
    sys.path.insert(0, REF)
    cases = []

    # N=10 Heisenberg at the session default maxm=30 (ED pin: tests/reference_data.py:15-17)
    out, _ = dmrg_case("heis_n10_m30_s0", [2] * 10, heis_terms(10), 30, 6, 0,
                       vev_terms=[("SzSz_5_6", [(1.0, [("Sz", 5), ("Sz", 6)])]),
                                  ("SxSx_1_2", [(1.0, [("Sx", 1), ("Sx", 2)])])])
    cases.append(out)

    # configuration 1: N=32, D=64, 6 sweeps, seeds 0,1,2
    for s in (0, 1, 2):
        out, _ = dmrg_case("heis_n32_m64_s%d" % s, [2] * 32, heis_terms(32), 64, 6, s,
                           vev_terms=[("SzSz_16_17", [(1.0, [("Sz", 16), ("Sz", 17)])])] if s == 0 else None)
        cases.append(out)

    # S=1 Heisenberg, N=12, D=32
    out, _ = dmrg_case("s1_n12_m32_s3", [3] * 12, heis_terms(12), 32, 4, 3)
    cases.append(out)

    # random-bond XXZ, N=12, D=32 (couplings from default_rng(7): J then Delta)
    rng = np.random.default_rng(7)
    J = rng.uniform(0.5, 1.5, 11)
    Dz = rng.uniform(0.5, 1.5, 11)
    out, _ = dmrg_case("xxz_n12_m32_s4", [2] * 12, heis_terms(12, J, J * Dz), 32, 4, 4)
    out["J"] = list(J)
    out["Delta"] = list(Dz)
    cases.append(out)

    # spinless fermions, N=10, JW terms from the reference's MultiOperator
    out, _ = dmrg_case("spinless_n10_m32_s5", [0] * 10, spinless_terms(10), 32, 4, 5,
                       vev_terms=[("N_3", [(1.0, [("N", 3)])])])
    cases.append(out)

    # native Hubbard (Electron sites), N=6, t=1, U=4
    out, _ = dmrg_case("hubbard_n6_m64_s6", [1] * 6, hubbard_terms(6), 64, 4, 6,
                       vev_terms=[("NupNdn_3", [(1.0, [("Nup", 3), ("Ndn", 3)])])])
    cases.append(out)

    with open(os.path.join(HERE, "dmrg_golden.json"), "w") as f:
        json.dump(cases, f, indent=1)

    # TDVP: quench correlator and evolve-and-measure on a seeded N=8 ground state
    tcases = []
    from dmrgpy.pyitensor.mpsalgebra import inner
    n = 8
    out, ch = dmrg_case("tdvp_gs_n8_m16_s8", [2] * n, heis_terms(n), 16, 6, 8)
    wf = ch.gs_wavefunction().copy()
    sz4 = [(1.0, [("Sz", 4)])]
    corr, _ = ch.quench_tdvp(heis_terms(n), sz4, sz4, 12, 0.05)
    out["quench"] = {"site": 4, "nt": 12, "dt": 0.05, "corr": [_c(c) for c in corr]}
    ch.wf0 = wf.copy()
    corr2, _ = ch.evolve_and_measure_tdvp(heis_terms(n), sz4, ch.apply_operator(
        [(1.0, [("Sx", 2)])], wf.copy()), 10, 0.05)
    out["evolve"] = {"apply": "Sx_2", "measure": "Sz_4", "nt": 10, "dt": 0.05,
                     "corr": [_c(c) for c in corr2]}
    tcases.append(out)
    with open(os.path.join(HERE, "tdvp_golden.json"), "w") as f:
        json.dump(tcases, f, indent=1)


if __name__ == "__main__":
    main()
