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
torch optimisers; ``fit_bond_couplings`` uses it to recover couplings from
target bond correlations.  ``tdvp_correlator_gradient`` differentiates a
TDVP-evolved correlator with respect to every coupling by the adjoint
method with checkpoints.  The reference has no gradient machinery; these
values are checked against central finite differences, against ``vev`` of
the bond operators and against exact dense time evolution
(tests/test_gpu_gradients.py).
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
    if not mpo.coupling_sites and nterms > 0:
        raise ValueError("this MPO has no per-term bookkeeping (SVD-compressed): build it with mpo.build_mpo")
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


# ---------------------------------------------------------------------------
# transition gradients <bra| dH/dc_t |ket>
# ---------------------------------------------------------------------------
def transition_mpo_gradients(eng, bra, ket, mpo, prec=0):
    """[G_s]: d<bra|H|ket>/dW_s for every site (bra and ket may differ in
    bond dimension), from mixed environments and the ``dmrg_mpo_grad``
    kernel; host arrays [2, wl, wr, d, d]."""
    bra = bra if bra.field else bra.as_complex()
    ket = ket if ket.field else ket.as_complex()
    n = ket.length()
    L = {0: alg.ones_env(eng, 1, prec)}
    for i in range(1, n):
        L[i] = alg.env_left(eng, L[i - 1], ket.A(i), mpo.W[i - 1], 1, prec, bra=bra.A(i))
    R = {n + 1: alg.ones_env(eng, 1, prec)}
    for i in range(n, 1, -1):
        R[i] = alg.env_right(eng, R[i + 1], ket.A(i), mpo.W[i - 1], 1, prec, bra=bra.A(i))
    out = []
    for s in range(1, n + 1):
        W = mpo.W[s - 1]
        out.append(eng.mpo_grad(1, prec, L[s - 1], bra.A(s), ket.A(s), R[s + 1], W.shape[1], W.shape[2]))
    return out


def transition_term_gradients(eng, bra, ket, mpo, nterms, prec=0):
    """<bra| dH/dc_t |ket> for every input term t (complex array)."""
    return contract_term_gradients(transition_mpo_gradients(eng, bra, ket, mpo, prec), mpo, nterms)


# ---------------------------------------------------------------------------
# adjoint TDVP
# ---------------------------------------------------------------------------
def simpson_weights(nt, dt):
    """Composite Simpson weights on the grid s_k = k dt, k = 0..nt (nt even)."""
    if nt % 2:
        raise ValueError("Simpson's rule needs an even number of steps")
    w = np.full(nt + 1, 2.0)
    w[1::2] = 4.0
    w[0] = w[-1] = 1.0
    return w * dt / 3.0


def tdvp_correlator_gradient(eng, codes, terms, psi1, psi2, e0, dt, nt, maxdim, cutoff=1e-12, niter=50,
                             checkpoint=10, prec=0, value_only=False):
    """C(T) = <psi2| psi(T)> and dC(T)/dc_t for every input term t of H.

    psi(t) = exp(-i (H - e0) t) psi1 is evolved by the two-site TDVP steps of
    quench_tdvp (src/dmrgpy/pyitensor/chain.py:520-542: shifted Hamiltonian,
    renormalisation to the initial norm after each step; here from the
    canonical gauge with the centre on site 1), T = nt dt.  With
    psi1, psi2 and e0 held fixed,
        dC/dc_t = -i int_0^T <phi(s)| dH/dc_t |psi(s)> ds,
        phi(s)  = exp(+i (H - e0)(T - s)) psi2,
    evaluated with Simpson's rule on the step grid.  phi runs backward from
    T by TDVP steps of -dt; psi(s) is recomputed segment by segment from the
    checkpoints stored every ``checkpoint`` steps of the forward pass, so at
    most nt / checkpoint + checkpoint + 2 states are held at once.  The
    integrand is the transition gradient <phi|dH/dc_t|psi> of the MPO
    gradient kernel.  Returns (C(T), complex array over the terms), or
    (C(T), None) with value_only (the forward evolution alone)."""
    from .mpo import build_mpo
    nterm = len(terms)
    Hs = build_mpo(codes, list(terms) + [(-e0, [("Id", 1)])], force_complex=False)
    w = simpson_weights(nt, dt)
    p1 = psi1.as_complex() if not psi1.field else psi1.copy()
    p2 = psi2.as_complex() if not psi2.field else psi2.copy()
    # orthogonality centre on site 1 before the first step: a first
    # left-to-right half-sweep from another gauge (as quench_tdvp does right
    # after the zip-up operator application, which leaves the centre on the
    # last site) evolves with an H_eff that is no projection and makes an
    # O(dt) error; from the canonical gauge the step is exact at full bond
    # dimension and second order otherwise
    alg.position(eng, p1, 1)
    alg.position(eng, p2, 1)
    n1 = np.sqrt(abs(alg.overlap(eng, p1, p1)))
    n2 = np.sqrt(abs(alg.overlap(eng, p2, p2)))

    def step(psi, d, norm):
        alg.tdvp_step(eng, psi, Hs, d, cutoff, maxdim, niter, prec)
        alg.normalize(eng, psi)
        alg.scale_site(psi, psi.center or 1, norm)

    ckpt = {0: p1.copy()}
    psi = p1
    for k in range(1, nt + 1):
        step(psi, dt, n1)
        if not value_only and (k % checkpoint == 0 or k == nt):
            ckpt[k] = psi.copy()
    C = alg.overlap(eng, p2, psi)
    if value_only:
        return C, None
    g = np.zeros(nterm + 1, dtype=complex)
    phi = p2
    marks = sorted(ckpt)
    for seg in range(len(marks) - 2, -1, -1):
        k0, k1 = marks[seg], marks[seg + 1]
        states = {}
        p = ckpt[k0].copy()
        for k in range(k0 + 1, k1 + 1):
            step(p, dt, n1)
            states[k] = p.copy()
        for k in range(k1, k0, -1):
            g += w[k] * transition_term_gradients(eng, phi, states[k], Hs, nterm + 1, prec)
            step(phi, -dt, n2)
        del states
    g += w[0] * transition_term_gradients(eng, phi, ckpt[0], Hs, nterm + 1, prec)
    return C, -1j * g[:nterm]


def xxz_correlator_gradients(eng, psi1, psi2, e0, J, Delta, dt, nt, maxdim, cutoff=1e-12, checkpoint=10, prec=0):
    """(C(T), dC/dJ_i, dC/dDelta_i) for H = sum_i J_i (SxSx + SySy + Delta_i SzSz)."""
    J = np.asarray(J, float)
    Delta = np.asarray(Delta, float)
    C, g = tdvp_correlator_gradient(eng, [2] * (len(J) + 1), xxz_terms(J, Delta), psi1, psi2, e0, dt, nt,
                                    maxdim, cutoff, checkpoint=checkpoint, prec=prec)
    gx, gy, gz = g[0::3], g[1::3], g[2::3]
    return C, gx + gy + Delta * gz, J * gz


# ---------------------------------------------------------------------------
# E0 as a differentiable function of the couplings of any term list
# ---------------------------------------------------------------------------
def coupled_terms(coeffs, operators):
    """H = sum_k c_k O_k as a session term list: each O_k is a term list
    (its own coefficients kept as written) scaled by c_k."""
    t = []
    for c, op in zip(coeffs, operators):
        for a, f in op:
            t.append((complex(c) * complex(a), list(f)))
    return t


def operator_expectations(eng, psi, codes, operators, prec=0):
    """<O_k> of the normalised psi for every term list O_k, from one
    backward pass of the MPO-gradient kernels (the gradient of
    <psi|sum_k c_k O_k|psi> at c = 1)."""
    from .mpo import build_mpo
    ones = np.ones(len(operators))
    terms = coupled_terms(ones, operators)
    g = term_gradients(eng, psi, build_mpo(codes, terms), len(terms), prec)
    out = np.zeros(len(operators), dtype=complex)
    t = 0
    for k, op in enumerate(operators):
        for a, _ in op:
            out[k] += complex(a) * g[t]
            t += 1
    return out


def _autograd_terms():
    import torch

    class CoupledGroundStateEnergy(torch.autograd.Function):
        """E0(c) of H = sum_k c_k O_k by DMRG on the device (warm-started
        from the session's current state); backward: dE0/dc_k = <O_k>
        (Hellmann-Feynman) by the MPO-gradient kernels."""

        @staticmethod
        def forward(ctx, c, operators, chain):
            cn = c.detach().cpu().double().numpy()
            chain.set_hamiltonian(coupled_terms(cn, operators))
            e = chain.gs_energy()
            g = operator_expectations(chain.eng, chain.gs_wavefunction(), chain.site_types, operators).real
            ctx.save_for_backward(torch.from_numpy(g))
            return torch.tensor(e, dtype=torch.float64)

        @staticmethod
        def backward(ctx, grad):
            (g,) = ctx.saved_tensors
            return grad * g.to(grad.device), None, None

    return CoupledGroundStateEnergy


def coupled_ground_state_energy(c, operators, chain):
    """Differentiable E0(c) for H = sum_k c_k O_k with O_k arbitrary term
    lists (torch tensor c in, scalar tensor out); the session keeps the
    ground state, so successive calls warm-start DMRG."""
    return _autograd_terms().apply(c, operators, chain)


def xxz_bond_operators(n):
    """[O_xy_1, .., O_xy_{n-1}, O_zz_1, .., O_zz_{n-1}] with
    O_xy_i = Sx_i Sx_{i+1} + Sy_i Sy_{i+1}, O_zz_i = Sz_i Sz_{i+1}: the XXZ
    chain is H = sum_i a_i O_xy_i + b_i O_zz_i with a_i = J_i,
    b_i = J_i Delta_i."""
    xy = [[(1.0, [("Sx", i + 1), ("Sx", i + 2)]), (1.0, [("Sy", i + 1), ("Sy", i + 2)])] for i in range(n - 1)]
    zz = [[(1.0, [("Sz", i + 1), ("Sz", i + 2)])] for i in range(n - 1)]
    return xy + zz


def xxz_bond_correlations(eng, psi, n, prec=0):
    """(<XX+YY>_i, <ZZ>_i) for the n-1 bonds of the normalised psi."""
    v = operator_expectations(eng, psi, [2] * n, xxz_bond_operators(n), prec).real
    return v[:n - 1], v[n - 1:]


def fit_couplings(chain, operators, targets, c0, scale_index=None, scale=None, steps=200, tol=1e-9,
                  callback=None):
    """Couplings c of H(c) = sum_k c_k O_k whose ground state has <O_k> = t_k.

    E0(c) is concave (a minimum of functions linear in c) with gradient
    <O_k>_c (Hellmann-Feynman), so F(c) = E0(c) - c.t is concave and its
    maximiser reproduces the targets.  E0 is homogeneous of degree one, so
    the overall scale is not fixed by correlations: the penalty
    (mean(c[scale_index]) - scale)^2 pins it (defaults: all couplings, the
    mean of c0).  L-BFGS on -F; every evaluation re-solves DMRG warm-started
    from the previous ground state with the chain's sweep settings, and the
    gradient comes from one backward pass of the device MPO-gradient
    kernels.  Returns (c, max |<O> - t|, history of that maximum)."""
    from scipy.optimize import minimize
    targets = np.asarray(targets, float)
    c0 = np.asarray(c0, float)
    idx = np.arange(len(c0)) if scale_index is None else np.asarray(scale_index)
    m0 = c0[idx].mean() if scale is None else float(scale)
    hist = []

    def fun(c):
        chain.set_hamiltonian(coupled_terms(c, operators))
        e = chain.gs_energy()
        ex = operator_expectations(chain.eng, chain.gs_wavefunction(), chain.site_types, operators).real
        res = ex - targets
        hist.append(float(np.abs(res).max()))
        if callback is not None:
            callback(len(hist), c, hist[-1])
        pen = c[idx].mean() - m0
        grad = -res
        grad[idx] += 2.0 * pen / len(idx)
        return -(e - c @ targets) + pen ** 2, grad

    r = minimize(fun, c0, jac=True, method="L-BFGS-B",
                 options={"maxiter": steps, "maxfun": steps, "gtol": tol, "ftol": 1e-16, "maxcor": 30})
    fun(r.x)
    return r.x, hist[-1], hist


def fit_xxz_couplings(chain, t_xy, t_zz, J0, Delta0, scale=None, steps=200, tol=1e-9, callback=None):
    """(J, Delta, max residual, history): the XXZ couplings whose ground
    state has bond correlations <XX+YY>_i = t_xy_i and <ZZ>_i = t_zz_i, by
    fit_couplings over a_i = J_i, b_i = J_i Delta_i with the scale of the
    J_i pinned to `scale` (default mean(J0))."""
    n = len(J0) + 1
    J0 = np.asarray(J0, float)
    c0 = np.concatenate([J0, J0 * np.asarray(Delta0, float)])
    c, resid, hist = fit_couplings(chain, xxz_bond_operators(n), np.concatenate([t_xy, t_zz]), c0,
                                   scale_index=np.arange(n - 1), scale=J0.mean() if scale is None else scale,
                                   steps=steps, tol=tol, callback=callback)
    a, b = c[:n - 1], c[n - 1:]
    return a, b / a, resid, hist


# ---------------------------------------------------------------------------
# fitting couplings to target bond correlations
# ---------------------------------------------------------------------------
def heisenberg_bond_terms(J):
    t = []
    for i, j in enumerate(np.asarray(J, float)):
        for op in ("Sx", "Sy", "Sz"):
            t.append((complex(j), [(op, i + 1), (op, i + 2)]))
    return t


def fit_bond_couplings(chain, target, J0, steps=60, tol=1e-7, callback=None):
    """Couplings J of H(J) = sum_i J_i S_i.S_{i+1} whose ground state has the
    target bond correlations t_i = <S_i.S_{i+1}>.

    E0(J) is concave (a minimum of functions linear in J) and, by
    Hellmann-Feynman, dE0/dJ_i = <S_i.S_{i+1}>_J, so F(J) = E0(J) - J.t is
    concave with gradient <S_i.S_{i+1}>_J - t_i: its maximiser reproduces
    the targets.  E0 is homogeneous of degree one in J, so the overall scale
    is pinned by the penalty (mean(J) - mean(J0))^2.  L-BFGS on -F with the
    device-side MPO-gradient kernels; every evaluation warm-starts DMRG from
    the previous ground state.  Returns (J, max |<S.S> - t|, history)."""
    from scipy.optimize import minimize
    from .mpo import build_mpo
    target = np.asarray(target, float)
    J0 = np.asarray(J0, float)
    m0 = J0.mean()
    n = len(J0)
    hist = []

    def fun(J):
        chain.set_hamiltonian(heisenberg_bond_terms(J))
        e = chain.gs_energy()
        mpo = build_mpo(chain.site_types, heisenberg_bond_terms(J))
        g = term_gradients(chain.eng, chain.gs_wavefunction(), mpo, 3 * n).real
        corr = g[0::3] + g[1::3] + g[2::3]
        res = corr - target
        hist.append(float(np.abs(res).max()))
        if callback is not None:
            callback(len(hist), J, hist[-1])
        pen = J.mean() - m0
        return -(e - J @ target) + pen ** 2, -res + 2.0 * pen / n

    r = minimize(fun, J0, jac=True, method="L-BFGS-B",
                 options={"maxiter": steps, "gtol": tol, "ftol": 1e-15, "maxcor": 20})
    fun(r.x)
    return r.x, hist[-1], hist
