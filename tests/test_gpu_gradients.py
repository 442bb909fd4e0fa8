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
def test_mpo_grad_kernel(engine, prec, cplx):
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
    G = engine.mpo_grad(int(cplx), prec, dv(L, et), dv(bra), dv(ket), dv(R, et), wl, wr)
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
