"""CPU check of the coupling-gradient bookkeeping (dmrgpy_amd/mpo.py
coupling_sites + gradients.term_gradients' formula): with the backward
contraction G_s = d<psi|H|psi>/dW_s evaluated in NumPy (the same
contraction the dmrg_mpo_grad kernel computes on the GPU), dE/dc_t must
equal <psi|T_t|psi> (Hellmann-Feynman) for every term and
sum_t c_t dE/dc_t = E."""

import numpy as np

from dmrgpy_amd.gradients import contract_term_gradients, xxz_terms
from dmrgpy_amd.mpo import build_mpo
from oracle import dmrg as od
from oracle import mpo as omp
from oracle import mps as om


def test_term_gradients_are_hellmann_feynman():
    n = 8
    rng = np.random.default_rng(1)
    J, Dz = rng.uniform(0.5, 1.5, n - 1), rng.uniform(0.5, 1.5, n - 1)
    terms = xxz_terms(J, Dz) + [(0.3, [("Sx", 3)]), (-0.2, [("Sy", 5)])]
    m = build_mpo([2] * n, terms)
    Wor = [(w[0] if w.shape[0] == 1 else w[0] + 1j * w[1]).transpose(0, 3, 2, 1) for w in m.W]
    np.random.seed(0)
    psi = om.random_mps([2] * n, 16)
    for _ in range(4):
        E = od.sweep(psi, Wor, 16, 1e-12, 2)
    om.normalize(psi)
    L, R = od.left_envs(Wor, psi), od.right_envs(Wor, psi)
    G = [np.einsum("pla,pqB,asb,Brb->lrqs", L[s - 1], psi.A[s - 1].conj(), psi.A[s - 1], R[s + 1])
         for s in range(1, n + 1)]
    g = contract_term_gradients([np.stack([x.real, x.imag]) for x in G], m, len(terms))
    for ti, (c, f) in enumerate(terms):
        A = omp.to_mpo([2] * n, [(1.0, f)])
        assert abs(g[ti] - om.inner3(psi, A, psi)) < 1e-12
    assert abs(sum(complex(c) * g[i] for i, (c, f) in enumerate(terms)) - E) < 1e-10


def test_pair_grads_chain_rule_vs_torch_autograd():
    """dmrgpy_amd.autodiff.pair_grads (host chain rule of the two-site H_eff
    backward, from the gradient of the pre-contracted pair W12 to W1 and W2)
    against torch autograd of the same pre-contraction on the CPU."""
    import torch
    from dmrgpy_amd.autodiff import pair_grads
    rng = np.random.default_rng(9)
    wl, wm, wr, d1, d2 = 3, 4, 5, 2, 3
    W1 = rng.standard_normal((wl, wm, d1, d1))
    W2 = rng.standard_normal((wm, wr, d2, d2))
    G = rng.standard_normal((wl, wr, d1 * d2, d1 * d2))
    t1, t2 = torch.tensor(W1, requires_grad=True), torch.tensor(W2, requires_grad=True)
    W12 = torch.einsum("lmac,mrbd->lrabcd", t1, t2).reshape(wl, wr, d1 * d2, d1 * d2)
    (W12 * torch.from_numpy(G)).sum().backward()
    g1, g2 = pair_grads(G, W1, W2)
    assert np.abs(g1 - t1.grad.numpy()).max() < 1e-12
    assert np.abs(g2 - t2.grad.numpy()).max() < 1e-12
