"""Coupling gradients on the GPU (dmrg_mpo_grad + dmrgpy_amd.gradients).

* kernel: G = d<bra|H|ket>/dW for one site against a NumPy einsum of the
  same contraction (real/complex, FP64/FP32, bra != ket);
* dE/dJ_i, dE/dDelta_i of a random-bond XXZ chain (couplings from
  default_rng(7), as in the golden XXZ case) against Hellmann-Feynman values
  from the session's own vev of the bond operators, and against central
  finite differences of the DMRG energy (relative 1e-6);
* the torch.autograd.Function returns the same gradient.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rnd(rng, shape, cplx):
    a = rng.standard_normal(shape)
    return a + 1j * rng.standard_normal(shape) if cplx else a


def planar(a):
    return np.stack([a.real, a.imag]) if np.iscomplexobj(a) else a[None]


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("pitch", [0, 3])
def test_mpo_grad_kernel(engine, prec, cplx, pitch):
    """G = d<bra|H|ket>/dW against einsum; with pitch > 0 the environments'
    last index is padded with NaN columns that must not be read."""
    import torch
    rng = np.random.default_rng(3 + prec + 2 * cplx)
    Dl, Dr, Dlb, Drb, d, wl, wr = 9, 13, 7, 11, 3, 4, 5
    L = rnd(rng, (wl, Dlb, Dl), cplx)       # [w][bra][ket]
    R = rnd(rng, (wr, Dr, Drb), cplx)       # [w][ket][bra]
    ket = rnd(rng, (Dl, d, Dr), cplx)
    bra = rnd(rng, (Dlb, d, Drb), cplx)
    ref = np.einsum("lpa,pqB,asb,rbB->lrqs", L, bra.conj(), ket, R)
    et = engine.env_dtype(prec)
    dv = lambda a, dt=torch.float64: torch.from_numpy(np.ascontiguousarray(planar(a))).to(dt).to(engine.device)

    def pad(t):
        if not pitch:
            return t
        out = torch.full(t.shape[:-1] + (t.shape[-1] + pitch,), float("nan"), dtype=t.dtype, device=t.device)
        out[..., :t.shape[-1]] = t
        return out

    G = engine.mpo_grad(int(cplx), prec, pad(dv(L, et)), dv(bra), dv(ket), pad(dv(R, et)), wl, wr)
    Gc = G[0] + 1j * G[1] if cplx else G[0]
    err = np.linalg.norm(Gc - ref) / np.linalg.norm(ref)
    assert err < (1e-12 if prec == 0 else 1e-5), err


def _xxz_case():
    rng = np.random.default_rng(7)
    J = rng.uniform(0.5, 1.5, 11)
    Dz = rng.uniform(0.5, 1.5, 11)
    return J, Dz


def _solve(J, Dz, wf=None, nsweeps=6, D=32):
    from dmrgpy_amd.chain import Chain
    from dmrgpy_amd.gradients import xxz_terms
    ch = Chain([2] * (len(J) + 1), field="real")
    ch.set_sweep_params(D, nsweeps, 1e-12, 0.0)
    ch.set_hamiltonian(xxz_terms(J, Dz))
    if wf is not None:
        ch.set_wavefunction(wf.copy())
    else:
        np.random.seed(4)
    return ch, ch.gs_energy()


def test_xxz_gradients_hellmann_feynman_and_fd():
    from dmrgpy_amd.gradients import xxz_gradients
    J, Dz = _xxz_case()
    ch, e = _solve(J, Dz)
    psi = ch.gs_wavefunction()
    gJ, gD = xxz_gradients(ch.eng, psi, J, Dz)
    # Hellmann-Feynman through the session's vev
    for i in range(len(J)):
        hJ = [(1.0, [("Sx", i + 1), ("Sx", i + 2)]), (1.0, [("Sy", i + 1), ("Sy", i + 2)]),
              (Dz[i], [("Sz", i + 1), ("Sz", i + 2)])]
        assert abs(gJ[i] - ch.vev(hJ, psi).real) < 1e-10
        assert abs(gD[i] - J[i] * ch.vev([(1.0, [("Sz", i + 1), ("Sz", i + 2)])], psi).real) < 1e-10
    # central finite differences (warm-started DMRG), three couplings each
    h = 1e-4
    for i in (0, 5, 10):
        Jp, Jm = J.copy(), J.copy()
        Jp[i] += h
        Jm[i] -= h
        fd = (_solve(Jp, Dz, psi, 4)[1] - _solve(Jm, Dz, psi, 4)[1]) / (2 * h)
        assert abs(fd - gJ[i]) < 1e-6 * abs(gJ[i]), (i, fd, gJ[i])
        Dp, Dm = Dz.copy(), Dz.copy()
        Dp[i] += h
        Dm[i] -= h
        fd = (_solve(J, Dp, psi, 4)[1] - _solve(J, Dm, psi, 4)[1]) / (2 * h)
        assert abs(fd - gD[i]) < 1e-6 * abs(gD[i]), (i, fd, gD[i])


def test_autograd_function():
    import torch
    from dmrgpy_amd.chain import Chain
    from dmrgpy_amd.gradients import ground_state_energy, xxz_gradients
    J, Dz = _xxz_case()
    ch = Chain([2] * (len(J) + 1), field="real")
    ch.set_sweep_params(32, 6, 1e-12, 0.0)
    np.random.seed(4)
    Jt = torch.tensor(J, requires_grad=True)
    Dt = torch.tensor(Dz, requires_grad=True)
    E = ground_state_energy(Jt, Dt, ch)
    E.backward()
    gJ, gD = xxz_gradients(ch.eng, ch.gs_wavefunction(), J, Dz)
    assert np.abs(Jt.grad.numpy() - gJ).max() < 1e-12
    assert np.abs(Dt.grad.numpy() - gD).max() < 1e-12


def test_tdvp_correlator_gradient_vs_exact():
    """Adjoint TDVP: C(T) = <psi2|exp(-i(H - E0)T)|psi1>, psi1 = psi2 =
    Sz_4 |psi0>, for a random-bond XXZ chain of 8 sites at full bond
    dimension, where two-site TDVP from the canonical gauge is exact, against
    dense exact evolution (C to 1e-10) and central finite differences of the
    dense C(T) in every coupling.  What remains is the Simpson quadrature of
    the adjoint integral, fourth order in dt: 1e-6 of max|dC/dJ| at
    dt = 0.025, and the error falls >= 8x from dt = 0.05 (a wrong sign,
    factor, time alignment or missing term would not converge)."""
    from scipy.linalg import expm
    from dmrgpy_amd.gradients import xxz_correlator_gradients, xxz_terms
    from oracle import ed
    rng = np.random.default_rng(11)
    n = 8
    J = rng.uniform(0.5, 1.5, n - 1)
    Dz = rng.uniform(0.5, 1.5, n - 1)
    ch, e0 = _solve(J, Dz, D=16, nsweeps=8)
    gs = ch.gs_wavefunction()
    psi1 = ch.apply_operator([(1.0, [("Sz", 4)])], gs.copy())
    v = psi1.to_dense().reshape(-1)
    T = 1.0

    def dense_C(Jx, Dx):
        H = ed.sparse_hamiltonian([2] * n, xxz_terms(Jx, Dx)).toarray()
        return np.vdot(v, expm(-1j * (H - e0 * np.eye(len(v))) * T) @ v)

    H0 = ed.sparse_hamiltonian([2] * n, xxz_terms(J, Dz)).toarray()
    g = gs.to_dense().reshape(-1)
    assert abs(np.vdot(g, H0 @ g) / np.vdot(g, g) - e0) < 1e-9      # same basis order
    h = 1e-6
    exJ = np.zeros(n - 1, complex)
    exD = np.zeros(n - 1, complex)
    for i in range(n - 1):
        Jp, Jm = J.copy(), J.copy()
        Jp[i] += h
        Jm[i] -= h
        exJ[i] = (dense_C(Jp, Dz) - dense_C(Jm, Dz)) / (2 * h)
        Dp, Dm = Dz.copy(), Dz.copy()
        Dp[i] += h
        Dm[i] -= h
        exD[i] = (dense_C(J, Dp) - dense_C(J, Dm)) / (2 * h)
    scale = max(np.abs(exJ).max(), np.abs(exD).max())
    C0 = dense_C(J, Dz)
    errs, cerrs = [], []
    for nt in (20, 40):
        C, gJ, gD = xxz_correlator_gradients(ch.eng, psi1, psi1, e0, J, Dz, T / nt, nt, 16, 1e-14,
                                             checkpoint=10)
        errs.append(max(np.abs(gJ - exJ).max(), np.abs(gD - exD).max()) / scale)
        cerrs.append(abs(C - C0))
    assert errs[1] < errs[0] / 8 and errs[1] < 1e-6, errs
    assert max(cerrs) < 1e-10, cerrs


def test_fit_bond_couplings():
    """Couplings fitted to the bond correlations <S_i.S_{i+1}> of a seeded
    hidden Heisenberg chain (mean of J fixed to the hidden one's): the
    fitted ground state reproduces the profile to 1e-6."""
    from dmrgpy_amd.chain import Chain
    from dmrgpy_amd.gradients import fit_bond_couplings, heisenberg_bond_terms, term_gradients
    from dmrgpy_amd.mpo import build_mpo
    rng = np.random.default_rng(19)
    n = 10
    Jh = rng.uniform(0.5, 1.5, n - 1)
    ch = Chain([2] * n, field="real")
    ch.set_sweep_params(32, 8, 1e-12, 0.0)
    ch.set_hamiltonian(heisenberg_bond_terms(Jh))
    np.random.seed(2)
    ch.gs_energy()
    g = term_gradients(ch.eng, ch.gs_wavefunction(), build_mpo([2] * n, heisenberg_bond_terms(Jh)),
                       3 * (n - 1)).real
    target = g[0::3] + g[1::3] + g[2::3]
    for i in (0, 4):   # Hellmann-Feynman = <S_i.S_{i+1}> through the session's vev
        ss = [(1.0, [(o, i + 1), (o, i + 2)]) for o in ("Sx", "Sy", "Sz")]
        assert abs(ch.vev(ss, ch.gs_wavefunction()).real - target[i]) < 1e-10
    fit = Chain([2] * n, field="real")
    fit.set_sweep_params(32, 4, 1e-12, 0.0)
    np.random.seed(3)
    J, err, hist = fit_bond_couplings(fit, target, np.full(n - 1, Jh.mean()), steps=80, tol=1e-8)
    assert err < 1e-6, hist
    # the profile pins J only loosely (bond correlations respond weakly to
    # smooth modulations of J), so the couplings agree to about 1e-2
    assert np.abs(J - Jh).max() < 3e-2, (J, Jh)
