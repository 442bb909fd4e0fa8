"""Backward passes of the differentiable H_eff apply and environment update
(dmrgpy_amd.autodiff; GPU).

Each Function's gradients with respect to every input (theta or the MPS
sites, the environments, the MPO tensors) are compared with torch's own
autograd of the same contraction written as float64 einsums on the CPU
(the plain reference of the op), for FP64 contractions (relative 1e-11) and
FP32 contractions (relative 2e-5), with environment dimensions that are not
multiples of the row pitch.  At a converged DMRG ground state the gradient
of E = <theta|H_eff|theta> / <theta|theta> with respect to the two MPO
tensors of the bond must equal the Hellmann-Feynman site gradients of the
MPO-gradient kernel, and dE/dtheta must vanish (theta is an eigenvector).
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = {0: 1e-11, 1: 2e-5}


def _rel(a, b):
    a = np.asarray(a, dtype=float)
    b = np.asarray(b, dtype=float)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _env(engine, a, prec):
    """[1, w, D, pitch] device environment from a [w, D, D] array."""
    import torch
    w, r, c = a.shape
    t = torch.zeros((1, w, r, engine.env_pitch(c)), dtype=engine.env_dtype(prec), device=engine.device)
    t[0, :, :, :c] = torch.from_numpy(a).to(t.dtype)
    return t


def _unenv(t, r, c):
    return t[0, :, :r, :c].double().cpu().numpy()


@pytest.mark.parametrize("prec", [0, 1])
def test_heff2_backward_vs_torch_einsum(engine, prec):
    import torch
    from dmrgpy_amd.autodiff import heff2_apply
    rng = np.random.default_rng(5 + prec)
    Dl, Dr, d1, d2, wl, wm, wr = 13, 11, 2, 3, 3, 4, 5
    L = rng.standard_normal((wl, Dl, Dl))
    R = rng.standard_normal((wr, Dr, Dr))
    W1 = rng.standard_normal((wl, wm, d1, d1))
    W2 = rng.standard_normal((wm, wr, d2, d2))
    th = rng.standard_normal((Dl, d1, d2, Dr))
    g = rng.standard_normal((Dl, d1, d2, Dr))
    if prec:   # inputs representable in FP32, so the references agree to FP32 accumulation
        L, R = L.astype(np.float32).astype(float), R.astype(np.float32).astype(float)
    # reference: torch autograd of the float64 einsum on the CPU
    tL, tR, tW1, tW2, tt = (torch.tensor(a, requires_grad=True) for a in (L, R, W1, W2, th))
    W12 = torch.einsum("lmac,mrbd->lrabcd", tW1, tW2)
    y_ref = torch.einsum("lxa,lrpqcd,acdb,rby->xpqy", tL, W12, tt, tR)
    (y_ref * torch.from_numpy(g)).sum().backward()
    # device Function
    Ld, Rd = _env(engine, L, prec).requires_grad_(), _env(engine, R, prec).requires_grad_()
    W1d = torch.from_numpy(W1[None].copy()).requires_grad_()
    W2d = torch.from_numpy(W2[None].copy()).requires_grad_()
    thd = torch.from_numpy(th.reshape(1, -1)).to(engine.device).requires_grad_()
    y = heff2_apply(engine, thd, Ld, Rd, W1d, W2d, (Dl, d1, d2, Dr), prec)
    (y * torch.from_numpy(g.reshape(1, -1)).to(engine.device)).sum().backward()
    errs = {
        "y": _rel(y.detach().cpu().numpy().reshape(th.shape), y_ref.detach().numpy()),
        "theta": _rel(thd.grad.cpu().numpy().reshape(th.shape), tt.grad.numpy()),
        "L": _rel(_unenv(Ld.grad, Dl, Dl), tL.grad.numpy()),
        "R": _rel(_unenv(Rd.grad, Dr, Dr), tR.grad.numpy()),
        "W1": _rel(W1d.grad[0].numpy(), tW1.grad.numpy()),
        "W2": _rel(W2d.grad[0].numpy(), tW2.grad.numpy()),
    }
    print("heff2 prec", prec, errs, flush=True)
    assert max(errs.values()) < TOL[prec], errs
    # the pitch padding of the environment gradients stays zero
    assert float(Ld.grad[0, :, :, Dl:].abs().max()) == 0.0 if Ld.grad.shape[-1] > Dl else True


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("direction", ["left", "right"])
def test_env_update_backward_vs_torch_einsum(engine, prec, direction):
    import torch
    from dmrgpy_amd.autodiff import env_update
    rng = np.random.default_rng(17 + prec + 3 * (direction == "left"))
    Dl, Dr, d, wl, wr = 9, 14, 3, 4, 5
    W = rng.standard_normal((wl, wr, d, d))
    ket = rng.standard_normal((Dl, d, Dr))
    bra = rng.standard_normal((Dl, d, Dr))
    E = rng.standard_normal((wl, Dl, Dl) if direction == "left" else (wr, Dr, Dr))
    if prec:
        E = E.astype(np.float32).astype(float)
    tE, tk, tb, tW = (torch.tensor(a, requires_grad=True) for a in (E, ket, bra, W))
    if direction == "left":      # out[wr][b'][b]
        out_ref = torch.einsum("lxa,xsy,lrsq,aqb->ryb", tE, tb, tW, tk)
    else:                        # out[wl][a][a']
        out_ref = torch.einsum("aqb,lrsq,xsy,rby->lax", tk, tW, tb, tE)
    G = rng.standard_normal(tuple(out_ref.shape))
    (out_ref * torch.from_numpy(G)).sum().backward()
    Ed = _env(engine, E, prec).requires_grad_()
    kd = torch.from_numpy(ket[None].copy()).to(engine.device).requires_grad_()
    bd = torch.from_numpy(bra[None].copy()).to(engine.device).requires_grad_()
    Wd = torch.from_numpy(W[None].copy()).requires_grad_()
    out = env_update(engine, direction, Ed, kd, bd, Wd, prec)
    r, c = out_ref.shape[1], out_ref.shape[2]
    Gd = torch.zeros_like(out)
    Gd[0, :, :r, :c] = torch.from_numpy(G).to(out.dtype)
    (out * Gd).sum().backward()
    errs = {
        "out": _rel(_unenv(out.detach(), r, c), out_ref.detach().numpy()),
        "env": _rel(_unenv(Ed.grad, E.shape[1], E.shape[2]), tE.grad.numpy()),
        "ket": _rel(kd.grad[0].cpu().numpy(), tk.grad.numpy()),
        "bra": _rel(bd.grad[0].cpu().numpy(), tb.grad.numpy()),
        "W": _rel(Wd.grad[0].numpy(), tW.grad.numpy()),
    }
    print("env", direction, "prec", prec, errs, flush=True)
    assert max(errs.values()) < TOL[prec], errs


def test_bond_energy_gradient_is_hellmann_feynman(engine):
    """E(W1, W2, theta) = <theta|H_eff|theta>/<theta|theta> at bond 1 of a
    converged ground state (centre on site 1): autograd through the H_eff
    Function gives dE/dW_1, dE/dW_2 equal to the MPO-gradient kernel's site
    gradients of <psi|H|psi>, and dE/dtheta = 2 (H_eff theta - E theta) = 0."""
    import torch
    from dmrgpy_amd import algorithms as alg
    from dmrgpy_amd.autodiff import heff2_apply
    from dmrgpy_amd.chain import Chain
    from dmrgpy_amd.gradients import mpo_gradients, xxz_terms
    rng = np.random.default_rng(31)
    n = 10
    ch = Chain([2] * n, field="real")
    ch.set_sweep_params(64, 6, 1e-12, 0.0)
    ch.set_hamiltonian(xxz_terms(rng.uniform(0.5, 1.5, n - 1), rng.uniform(0.5, 1.5, n - 1)))
    np.random.seed(0)
    e0 = ch.gs_energy()
    psi = ch.gs_wavefunction()
    assert psi.center == 1
    G = mpo_gradients(engine, psi, ch.H)
    R = alg.right_envs(engine, psi, ch.H, 0)[3]
    L = alg.ones_env(engine, 0, 0)
    th, dims = alg.theta2(engine, psi, 1)
    thd = th.clone().requires_grad_()
    W1 = torch.from_numpy(ch.H.W[0].copy()).requires_grad_()
    W2 = torch.from_numpy(ch.H.W[1].copy()).requires_grad_()
    y = heff2_apply(engine, thd, L, R, W1, W2, dims, 0)
    E = (thd * y).sum() / (thd * thd).sum()
    E.backward()
    print("bond energy", E.item(), "DMRG", e0, flush=True)
    assert abs(E.item() - e0) < 1e-10
    e1 = _rel(W1.grad.numpy(), G[0])
    e2 = _rel(W2.grad.numpy(), G[1])
    gth = float(thd.grad.abs().max())
    print("dE/dW1 vs site-1 kernel gradient", e1, "dE/dW2 vs site 2", e2, "max |dE/dtheta|", gth, flush=True)
    assert e1 < 1e-10 and e2 < 1e-10
    assert gth < 1e-5
