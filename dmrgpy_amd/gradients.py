"""Coupling gradients of the DMRG ground-state energy.

At a converged, normalised ground state E(J) = <psi|H(J)|psi>, so
dE/dJ = <psi|dH/dJ|psi> (Hellmann-Feynman).  With the automaton MPO every
input term's coefficient sits in exactly one MPO element, so all
derivatives come from one backward contraction per site,
G_s = d<psi|H|psi>/dW_s (libdmrg_hip ``dmrg_mpo_grad``), evaluated against
the left and right environments of the final state:
dE/dc_t = source_t * sum(G_s[wl_t, wr_t] * M_t).

``GroundStateEnergy`` wraps a DMRG solve as a ``torch.autograd.Function``
whose backward applies these kernels, so couplings can be optimised with
torch optimisers.  The reference has no gradient machinery; these values
are checked against central finite differences and against ``vev`` of the
bond operators (tests/test_gpu_gradients.py).
"""

import numpy as np

from . import algorithms as alg


def mpo_gradients(eng, psi, mpo, prec=0):
    """[G_s] for s = 1..N: d<psi|H|psi>/dW_s for the normalised psi, host
    arrays [np, wl, wr, d, d]."""
    field = max(psi.field, 1 if mpo.complex else 0)
    p = psi.as_complex() if field and not psi.field else psi.copy()
    alg.normalize(eng, p)
    L = alg.left_envs(eng, p, mpo, prec)
    R = alg.right_envs(eng, p, mpo, prec)
    n = p.length()
    out = []
    for s in range(1, n + 1):
        W = mpo.W[s - 1]
        A = p.A(s)
        out.append(eng.mpo_grad(field, prec, L[s - 1], A, A, R[s + 1], W.shape[1], W.shape[2]))
    return out


def term_gradients(eng, psi, mpo, nterms, prec=0):
    """dE/dc_t for every input term t (complex array of length nterms)."""
    G = mpo_gradients(eng, psi, mpo, prec)
    return contract_term_gradients(G, mpo, nterms)


def contract_term_gradients(G, mpo, nterms):
    """dE/dc_t = source_t * sum(G_s[wl_t, wr_t] * M_t) over the MPO elements
    that carry term t's coefficient."""
    g = np.zeros(nterms, dtype=complex)
    for ti, s, fr, to, m, source in mpo.coupling_sites:
        Gs = G[s]
        Gc = Gs[0] + 1j * Gs[1] if Gs.shape[0] == 2 else Gs[0].astype(complex)
        g[ti] += source * np.sum(Gc[fr, to] * m)
    return g


def xxz_terms(J, Delta):
    J = np.asarray(J, float)
    Delta = np.asarray(Delta, float)
    t = []
    for i in range(len(J)):
        t.append((complex(J[i]), [("Sx", i + 1), ("Sx", i + 2)]))
        t.append((complex(J[i]), [("Sy", i + 1), ("Sy", i + 2)]))
        t.append((complex(J[i] * Delta[i]), [("Sz", i + 1), ("Sz", i + 2)]))
    return t


def xxz_gradients(eng, psi, J, Delta, prec=0):
    """(dE/dJ_i, dE/dDelta_i) for H = sum_i J_i (SxSx + SySy + Delta_i SzSz)."""
    from .mpo import build_mpo
    J = np.asarray(J, float)
    Delta = np.asarray(Delta, float)
    mpo = build_mpo([2] * (len(J) + 1), xxz_terms(J, Delta))
    g = term_gradients(eng, psi, mpo, 3 * len(J), prec).real
    gx, gy, gz = g[0::3], g[1::3], g[2::3]
    return gx + gy + Delta * gz, J * gz


def _autograd():
    import torch

    class GroundStateEnergy(torch.autograd.Function):
        """E0(J, Delta) of the XXZ chain by DMRG on the device; backward by
        the Hellmann-Feynman MPO-gradient kernels."""

        @staticmethod
        def forward(ctx, J, Delta, chain):
            from .mpo import build_mpo
            Jn = J.detach().cpu().double().numpy()
            Dn = Delta.detach().cpu().double().numpy()
            chain.set_hamiltonian(xxz_terms(Jn, Dn))
            e = chain.gs_energy()
            gJ, gD = xxz_gradients(chain.eng, chain.gs_wavefunction(), Jn, Dn)
            ctx.save_for_backward(torch.from_numpy(gJ), torch.from_numpy(gD))
            return torch.tensor(e, dtype=torch.float64)

        @staticmethod
        def backward(ctx, grad):
            gJ, gD = ctx.saved_tensors
            return grad * gJ, grad * gD, None

    return GroundStateEnergy


def ground_state_energy(J, Delta, chain):
    """Differentiable E0(J, Delta) (torch tensors in, scalar tensor out)."""
    return _autograd().apply(J, Delta, chain)
