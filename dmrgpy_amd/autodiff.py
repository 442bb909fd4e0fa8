"""torch.autograd.Function wrappers of the H_eff apply and the environment
update, with hand-written backward passes on the library's kernels.

The reference has no gradients; its forward operations are the two-site
matvec ``two_site_heff`` (src/dmrgpy/pyitensor/dmrg.py:224-248, built by
``kernels._make_matvec_planned``, kernels.py:353-410) and the environment
extension ``_extend_left`` / ``_extend_right`` (dmrg.py:170-201).  Here both
are differentiable with respect to every input:

``heff2_apply(eng, theta, L, R, W1, W2, dims, prec)``
    y[a' p' b'] = sum L[wl][a'][a] W12[wl wr][p' p] theta[a p b] R[wr][b][b'],
    W12 = sum_wm W1[wl][wm] (x) W2[wm][wr] (p = (s1, s2)).  Backward, for the
    output gradient g:
      * d/dtheta = H_eff^T g: one apply of the transposed operator
        (L and R transposed per channel, W1 and W2 transposed in their
        physical indices);
      * d/dW: G12 = d<g|H_eff|theta>/dW12 by the MPO-gradient kernel
        (``dmrg_mpo_grad`` with the pair treated as one site of dimension
        d1 d2), then the chain rule to W1 and W2 on the host;
      * d/dL[w] = g Y_w^T with Y_w = H_eff(L -> delta_w) theta, and
        d/dR[w] = Y'_w^T g with Y'_w = H_eff(R -> delta_w) theta: one apply
        per channel with a unit environment in that channel, then one GEMM.
``env_update(eng, direction, env_in, ket, bra, W, prec)``
    left:  out[wr][b'][b] = sum L[wl][a'][a] bra[a' s' b'] W[wl wr][s' s] ket[a s b];
    right: out[wl][a][a'] = sum ket[a s b] W[wl wr][s' s] bra[a' s' b'] R[wr][b][b'].
    Backward: d/dW by the MPO-gradient kernel (the output gradient acting as
    the missing environment), d/dket and d/dbra as one-site H_eff applies of
    the (transposed) operator with the output gradient as environment, and
    d/denv_in as one environment update in the opposite direction.

Real field (the arithmetic of every real Hamiltonian's DMRG run); FP64
(prec 0) or FP32 contractions (prec 1, environments float32).  Device
tensors use the engine layouts (a leading plane axis of size 1; environments
[1, w, D, pitch]); MPO tensors are CPU float64 tensors [1, wl, wr, s_out, s_in]
(W1 of a pair [1, wl, wm, d1, d1]).  There is no CPU fallback: every
operation is a call into libdmrg_hip.so.
"""

import numpy as np


def _torch():
    import torch
    return torch


def pair_grads(G12, W1, W2):
    """Chain rule from the gradient of the pre-contracted pair
    W12[wl][wr][(s1o s2o)][(s1 s2)] = sum_wm W1[wl][wm][s1o][s1] W2[wm][wr][s2o][s2]
    to (dW1, dW2); arrays [wl, wr, P, P], [wl, wm, d1, d1], [wm, wr, d2, d2]."""
    wl, wm, d1, _ = W1.shape
    _, wr, d2, _ = W2.shape
    G = np.asarray(G12).reshape(wl, wr, d1, d2, d1, d2)
    dW1 = np.einsum("lrabcd,mrbd->lmac", G, W2)
    dW2 = np.einsum("lrabcd,lmac->mrbd", G, W1)
    return dW1, dW2


def _pitched(eng, t2d_batches, dtype):
    """[1, w, rows, pitch] device tensor holding the given [w, rows, cols]
    data (cols padded to the engine's environment pitch)."""
    torch = _torch()
    w, rows, cols = t2d_batches.shape
    out = torch.zeros((1, w, rows, eng.env_pitch(cols)), dtype=dtype, device=eng.device)
    out[0, :, :, :cols] = t2d_batches
    return out


def _env_T(eng, env, rows, cols):
    """Per-channel transpose of a pitched environment [1, w, rows, >=cols]."""
    return _pitched(eng, env[0, :, :rows, :cols].transpose(-1, -2), env.dtype)


def _like(t, data, rows, cols):
    """Zero tensor shaped like `t` (a pitched environment) holding data[:, :rows, :cols]."""
    torch = _torch()
    out = torch.zeros_like(t)
    out[0, :, :rows, :cols] = data[0, :, :rows, :cols].to(t.dtype)
    return out


def _delta_env(eng, w, which, D, dtype):
    """Unit environment in channel `which`: E[w'][a][b] = delta(w', which) delta(a, b)."""
    torch = _torch()
    e = torch.zeros((1, w, D, eng.env_pitch(D)), dtype=dtype, device=eng.device)
    e[0, which, :, :D] = torch.eye(D, dtype=dtype, device=eng.device)
    return e


def _heff2(eng, L, R, W1, W2, dims, prec, x):
    Dl, d1, d2, Dr = dims
    desc = eng.heff_desc(L, R, W1, W2, Dl, Dr, d1, d2, 0, prec)
    y = eng.empty(x.shape)
    eng.heff_apply(desc, x.contiguous(), y)
    return y


def _heff1(eng, L, R, W, dims, prec, x):
    Dl, d, Dr = dims
    desc = eng.heff_desc(L, R, W, None, Dl, Dr, d, 0, 0, prec)
    y = eng.empty(x.shape)
    eng.heff_apply(desc, x.contiguous(), y)
    return y


def _gemm64(eng, A, B, M, N, K, transA=False, transB=False):
    C = eng.empty((M, N))
    eng.gemm(0, 0, A.contiguous(), B.contiguous(), C, M, N, K, transA=transA, transB=transB)
    return C


def _check_real(*ts):
    for t in ts:
        if t.shape[0] != 1:
            raise NotImplementedError("autodiff: the differentiable H_eff / environment update is real-field only")


class Heff2Apply:
    """Namespace of the autograd Function for the two-site H_eff apply."""

    _fn = None

    @classmethod
    def function(cls):
        if cls._fn is not None:
            return cls._fn
        torch = _torch()

        class _Heff2(torch.autograd.Function):
            @staticmethod
            def forward(ctx, theta, L, R, W1, W2, eng, dims, prec):
                _check_real(theta, L, R, W1, W2)
                W1n, W2n = W1.detach().double().numpy(), W2.detach().double().numpy()
                y = _heff2(eng, L.detach(), R.detach(), W1n, W2n, dims, prec, theta.detach())
                ctx.save_for_backward(theta.detach(), L.detach(), R.detach())
                ctx.W = (W1n, W2n)
                ctx.eng, ctx.dims, ctx.prec = eng, dims, prec
                return y

            @staticmethod
            def backward(ctx, g):
                theta, L, R = ctx.saved_tensors
                W1n, W2n = ctx.W
                eng, dims, prec = ctx.eng, ctx.dims, ctx.prec
                Dl, d1, d2, Dr = dims
                P = d1 * d2
                wl, wr = W1n.shape[1], W2n.shape[2]
                g = g.contiguous()
                gt = gL = gR = gW1 = gW2 = None
                if ctx.needs_input_grad[0]:
                    W1t = np.ascontiguousarray(W1n.swapaxes(-1, -2))
                    W2t = np.ascontiguousarray(W2n.swapaxes(-1, -2))
                    gt = _heff2(eng, _env_T(eng, L, Dl, Dl), _env_T(eng, R, Dr, Dr), W1t, W2t, dims, prec, g)
                if ctx.needs_input_grad[3] or ctx.needs_input_grad[4]:
                    G12 = eng.mpo_grad(0, prec, L, g.view(1, Dl, P, Dr), theta.view(1, Dl, P, Dr), R, wl, wr)[0]
                    dW1, dW2 = pair_grads(G12, W1n[0], W2n[0])
                    gW1 = torch.from_numpy(dW1[None].copy())
                    gW2 = torch.from_numpy(dW2[None].copy())
                if ctx.needs_input_grad[1]:
                    gL = torch.zeros_like(L)
                    for w in range(wl):
                        Y = _heff2(eng, _delta_env(eng, wl, w, Dl, L.dtype), R, W1n, W2n, dims, prec, theta)
                        gL[0, w, :, :Dl] = _gemm64(eng, g.view(Dl, P * Dr), Y.view(Dl, P * Dr), Dl, Dl, P * Dr,
                                                   transB=True).to(L.dtype)
                if ctx.needs_input_grad[2]:
                    gR = torch.zeros_like(R)
                    for w in range(wr):
                        Y = _heff2(eng, L, _delta_env(eng, wr, w, Dr, R.dtype), W1n, W2n, dims, prec, theta)
                        gR[0, w, :, :Dr] = _gemm64(eng, Y.view(Dl * P, Dr), g.view(Dl * P, Dr), Dr, Dr, Dl * P,
                                                   transA=True).to(R.dtype)
                return gt, gL, gR, gW1, gW2, None, None, None

        cls._fn = _Heff2
        return _Heff2


def heff2_apply(eng, theta, L, R, W1, W2, dims, prec=0):
    """y = H_eff theta (two sites), differentiable with respect to theta
    ([1, Dl*d1*d2*Dr] float64), L / R (environments [1, w, D, pitch] in the
    contraction precision) and W1 / W2 (CPU float64 [1, wl, wm, d1, d1] and
    [1, wm, wr, d2, d2]).  dims = (Dl, d1, d2, Dr)."""
    return Heff2Apply.function().apply(theta, L, R, W1, W2, eng, tuple(dims), int(prec))


class EnvUpdate:
    """Namespace of the autograd Function for the environment update."""

    _fn = None

    @classmethod
    def function(cls):
        if cls._fn is not None:
            return cls._fn
        torch = _torch()
        from . import algorithms as alg

        class _Env(torch.autograd.Function):
            @staticmethod
            def forward(ctx, env_in, ket, bra, W, eng, direction, prec):
                _check_real(env_in, ket, bra, W)
                if ket.shape != bra.shape:
                    raise NotImplementedError("autodiff env_update: bra and ket must have the same shape")
                Wn = W.detach().double().numpy()
                fwd = alg.env_left if direction == "left" else alg.env_right
                out = fwd(eng, env_in.detach(), ket.detach().contiguous(), Wn, 0, prec, bra=bra.detach().contiguous())
                ctx.save_for_backward(env_in.detach(), ket.detach(), bra.detach())
                ctx.W, ctx.eng, ctx.direction, ctx.prec = Wn, eng, direction, prec
                return out

            @staticmethod
            def backward(ctx, g):
                env_in, ket, bra = ctx.saved_tensors
                Wn, eng, prec = ctx.W, ctx.eng, ctx.prec
                _, Dl, d, Dr = ket.shape
                wl, wr = Wn.shape[1], Wn.shape[2]
                Wt = np.ascontiguousarray(Wn.swapaxes(-1, -2))
                g = g.to(env_in.dtype)
                g_env = g_ket = g_bra = g_W = None
                if ctx.direction == "left":
                    # env_in L[wl][a'][a], output gradient G[wr][b'][b]
                    GT = _env_T(eng, g, Dr, Dr)                      # [wr][b][b']
                    if ctx.needs_input_grad[3]:
                        g_W = torch.from_numpy(eng.mpo_grad(0, prec, env_in, bra, ket, GT, wl, wr)[:1].copy())
                    if ctx.needs_input_grad[1]:
                        g_ket = _heff1(eng, _env_T(eng, env_in, Dl, Dl), g, Wt, (Dl, d, Dr), prec,
                                       bra.reshape(1, -1)).view(ket.shape)
                    if ctx.needs_input_grad[2]:
                        g_bra = _heff1(eng, env_in, GT, Wn, (Dl, d, Dr), prec, ket.reshape(1, -1)).view(bra.shape)
                    if ctx.needs_input_grad[0]:
                        out = alg.env_right(eng, GT, ket.contiguous(), Wn, 0, prec, bra=bra.contiguous())
                        g_env = _like(env_in, _env_T(eng, out, Dl, Dl), Dl, Dl)
                else:
                    # env_in R[wr][b][b'], output gradient G[wl][a][a']
                    GT = _env_T(eng, g, Dl, Dl)                      # [wl][a'][a]
                    if ctx.needs_input_grad[3]:
                        g_W = torch.from_numpy(eng.mpo_grad(0, prec, GT, bra, ket, env_in, wl, wr)[:1].copy())
                    if ctx.needs_input_grad[1]:
                        g_ket = _heff1(eng, g, _env_T(eng, env_in, Dr, Dr), Wt, (Dl, d, Dr), prec,
                                       bra.reshape(1, -1)).view(ket.shape)
                    if ctx.needs_input_grad[2]:
                        g_bra = _heff1(eng, GT, env_in, Wn, (Dl, d, Dr), prec, ket.reshape(1, -1)).view(bra.shape)
                    if ctx.needs_input_grad[0]:
                        out = alg.env_left(eng, GT, ket.contiguous(), Wn, 0, prec, bra=bra.contiguous())
                        g_env = _like(env_in, _env_T(eng, out, Dr, Dr), Dr, Dr)
                return g_env, g_ket, g_bra, g_W, None, None, None

        cls._fn = _Env
        return _Env


def env_update(eng, direction, env_in, ket, bra, W, prec=0):
    """One environment extension (direction "left": L grows by one site to
    the right; "right": R grows to the left), differentiable with respect to
    env_in, ket, bra ([1, D, d, D] float64, same shape) and W (CPU float64
    [1, wl, wr, d, d])."""
    return EnvUpdate.function().apply(env_in, ket, bra, W, eng, direction, int(prec))
