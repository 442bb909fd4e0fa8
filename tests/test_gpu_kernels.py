"""Kernel-level parity of the HIP engine against the oracle (GPU).

Every kernel is called through the C ABI (dmrgpy_amd.engine) on seeded
inputs and compared with the oracle's NumPy FP64 restatement of the
reference contraction (oracle/dmrg.py, oracle/mps.py, oracle/linalg.py):

* MFMA GEMM (FP64 / FP32 / split-BF16x3, real / complex, all operand modes)
  against a float64 matmul of the same inputs;
* H_eff two-site and one-site applies, left / right environment updates;
* truncated SVD (keep rule, reconstruction, orthonormality);
* Lanczos ground state and Krylov exponential.

Tolerances (relative Frobenius): FP64 1e-12, FP32 2e-6, split-BF16x3 1e-4.
"""

import numpy as np
import pytest

from oracle import dmrg as odmrg
from oracle import linalg as olin
from oracle import mps as omps

pytestmark = pytest.mark.gpu

TOL = {0: 1e-12, 1: 2e-6, 2: 1e-4}


def rnd(rng, shape, cplx):
    a = rng.standard_normal(shape)
    if cplx:
        a = a + 1j * rng.standard_normal(shape)
    return a


def planar(a):
    a = np.asarray(a)
    if np.iscomplexobj(a):
        return np.stack([a.real, a.imag])
    return a[None]


def unplanar(p):
    p = np.asarray(p)
    return p[0] + 1j * p[1] if p.shape[0] == 2 else p[0]


def dev(engine, a, dtype=None):
    import torch
    t = torch.from_numpy(np.ascontiguousarray(planar(a), dtype=np.float64))
    if dtype is not None:
        t = t.to(dtype)
    return t.to(engine.device)


def host(t):
    return unplanar(t.double().cpu().numpy())


def relerr(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


# layout conversions oracle -> engine
def L_eng(L):   # oracle L[a', w, a] -> engine [w, a', a]
    return L.transpose(1, 0, 2)


def R_eng(R):   # oracle R[b', w, b] -> engine [w, b, b']
    return R.transpose(1, 2, 0)


def W_eng(W):   # oracle W[wl, s_in, s_out, wr] -> engine [wl, wr, s_out, s_in]
    return planar(W.transpose(0, 3, 2, 1))


@pytest.mark.parametrize("prec", [0, 1, 2])
@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("ta,tb", [(False, False), (True, False), (False, True), (True, True)])
@pytest.mark.parametrize("shape", [(37, 53, 29), (160, 96, 200), (2048, 2176, 48)])
def test_gemm(engine, prec, cplx, ta, tb, shape):
    import torch
    M, N, K = shape
    rng = np.random.default_rng(hash((prec, cplx, ta, tb, M)) % 2**32)
    A = rnd(rng, (K, M) if ta else (M, K), cplx)
    B = rnd(rng, (N, K) if tb else (K, N), cplx)
    C0 = rnd(rng, (M, N), cplx)
    dt = torch.float64 if prec == 0 else torch.float32
    Ad, Bd, Cd = dev(engine, A, dt), dev(engine, B, dt), dev(engine, C0, dt)
    # inputs as the kernel sees them (rounded to the storage type)
    Ah, Bh, Ch = host(Ad), host(Bd), host(Cd)
    conjA, conjB = cplx and ta, cplx and tb
    opA = (Ah.T.conj() if conjA else Ah.T) if ta else Ah
    opB = (Bh.T.conj() if conjB else Bh.T) if tb else Bh
    ref = 0.7 * opA @ opB + 0.3 * Ch
    engine.gemm(prec, int(cplx), Ad, Bd, Cd, M, N, K, transA=ta, transB=tb, conjA=conjA, conjB=conjB,
                alpha=0.7, beta=0.3, planeA=A.size, planeB=B.size, planeC=M * N)
    torch.cuda.synchronize()
    err = relerr(host(Cd), ref)
    assert err < TOL[prec], err


def test_gemm_batched_strided(engine):
    import torch
    rng = np.random.default_rng(5)
    bs, M, N, K = 3, 40, 24, 33
    A = rng.standard_normal((bs, M, K))
    B = rng.standard_normal((K, N))          # shared (stride 0)
    Ad, Bd = dev(engine, A), dev(engine, B)
    Cd = engine.zeros((1, bs, M, N))
    engine.gemm(0, 0, Ad, Bd, Cd, M, N, K, strideA=M * K, strideB=0, strideC=M * N, batch=bs)
    torch.cuda.synchronize()
    assert relerr(host(Cd), np.einsum("bmk,kn->bmn", A, B)) < 1e-13


def heff_case(rng, cplx, Dl=23, Dr=31, d1=2, d2=3, wl=4, wm=5, wr=3):
    L = rnd(rng, (Dl, wl, Dl), cplx)
    R = rnd(rng, (Dr, wr, Dr), cplx)
    W1 = rnd(rng, (wl, d1, d1, wm), cplx)
    W2 = rnd(rng, (wm, d2, d2, wr), cplx)
    th = rnd(rng, (Dl, d1, d2, Dr), cplx)
    return L, R, W1, W2, th


@pytest.mark.parametrize("prec", [0, 1, 2])
@pytest.mark.parametrize("cplx", [False, True])
def test_heff2(engine, prec, cplx):
    import torch
    rng = np.random.default_rng(11 + prec)
    L, R, W1, W2, th = heff_case(rng, cplx)
    ref = odmrg.heff2_matvec(L, W1, W2, R, th.shape)(th.reshape(-1).astype(complex)).reshape(th.shape)
    et = engine.env_dtype(prec)
    Ld, Rd = dev(engine, L_eng(L), et), dev(engine, R_eng(R), et)
    desc = engine.heff_desc(Ld, Rd, W_eng(W1), W_eng(W2), th.shape[0], th.shape[3], th.shape[1],
                            th.shape[2], int(cplx), prec)
    x = dev(engine, th)
    y = engine.zeros(x.shape)
    engine.heff_apply(desc, x, y)
    torch.cuda.synchronize()
    err = relerr(host(y), ref if cplx else ref.real)
    assert err < TOL[prec] * 10, err


@pytest.mark.parametrize("prec", [0, 1, 2])
@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("dims", [(70, 66, 4, 4, 6, 6, 6), (40, 130, 3, 3, 5, 5, 5), (9, 300, 2, 2, 5, 5, 5)])
def test_heff2_real_mpo(engine, prec, cplx, dims):
    """Real (realified) MPO with real or complex vectors at shapes where the
    small operator runs on MFMA (Hubbard-like d=4, w=6: K = 96)."""
    import torch
    Dl, Dr, d1, d2, wl, wm, wr = dims
    rng = np.random.default_rng(sum(dims) + prec + 7 * cplx)
    L = rnd(rng, (Dl, wl, Dl), cplx)
    R = rnd(rng, (Dr, wr, Dr), cplx)
    W1 = rnd(rng, (wl, d1, d1, wm), False)
    W2 = rnd(rng, (wm, d2, d2, wr), False)
    th = rnd(rng, (Dl, d1, d2, Dr), cplx)
    ref = odmrg.heff2_matvec(L, W1, W2, R, th.shape)(th.reshape(-1).astype(complex)).reshape(th.shape)
    et = engine.env_dtype(prec)
    desc = engine.heff_desc(dev(engine, L_eng(L), et), dev(engine, R_eng(R), et), W_eng(W1), W_eng(W2),
                            Dl, Dr, d1, d2, int(cplx), prec)
    x = dev(engine, th)
    y = engine.zeros(x.shape)
    engine.heff_apply(desc, x, y)
    torch.cuda.synchronize()
    err = relerr(host(y), ref if cplx else ref.real)
    assert err < TOL[prec] * 10, err


def pitched(engine, a, dtype, extra=3):
    """Device copy of a planar array whose last index has `extra` padding
    columns filled with NaN: a kernel that reads the padding poisons its
    output."""
    import torch
    t = dev(engine, a, dtype)
    out = torch.full(t.shape[:-1] + (t.shape[-1] + extra,), float("nan"), dtype=t.dtype, device=t.device)
    out[..., :t.shape[-1]] = t
    return out


def unpitch(t, n):
    return t[..., :n].contiguous()


@pytest.mark.parametrize("prec", [0, 1, 2])
@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("nsite", [1, 2])
def test_heff_pitched_environments(engine, prec, cplx, nsite):
    """Environments whose last index has a row pitch (padding = NaN): the
    H_eff apply reads only the logical columns; d = 3 makes theta's rows
    (9 Dr) need a pitched copy as well."""
    import torch
    rng = np.random.default_rng(31 + prec + 2 * cplx + 5 * nsite)
    Dl, Dr, d, w = 21, 35, 3, 4
    L = rnd(rng, (Dl, w, Dl), cplx)
    R = rnd(rng, (Dr, w, Dr), cplx)
    W1 = rnd(rng, (w, d, d, w), False)
    W2 = rnd(rng, (w, d, d, w), False)
    et = engine.env_dtype(prec)
    if nsite == 2:
        th = rnd(rng, (Dl, d, d, Dr), cplx)
        ref = odmrg.heff2_matvec(L, W1, W2, R, th.shape)(th.reshape(-1).astype(complex)).reshape(th.shape)
        desc = engine.heff_desc(pitched(engine, L_eng(L), et), pitched(engine, R_eng(R), et, 1),
                                W_eng(W1), W_eng(W2), Dl, Dr, d, d, int(cplx), prec)
    else:
        th = rnd(rng, (Dl, d, Dr), cplx)
        ref = odmrg.heff1_matvec(L, W1, R, th.shape)(th.reshape(-1).astype(complex)).reshape(th.shape)
        desc = engine.heff_desc(pitched(engine, L_eng(L), et), pitched(engine, R_eng(R), et, 1),
                                W_eng(W1), None, Dl, Dr, d, 0, int(cplx), prec)
    x = dev(engine, th)
    y = engine.zeros(x.shape)
    engine.heff_apply(desc, x, y)
    torch.cuda.synchronize()
    err = relerr(host(y), ref if cplx else ref.real)
    assert err < TOL[prec] * 10, err


@pytest.mark.parametrize("prec", [0, 1, 2])
@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("model", ["hubbard", "heisenberg"])
def test_heff2_model_mpo(engine, prec, cplx, model):
    """Bulk MPO tensors of the automaton MPO (sparse: the pre-contracted W
    pair of the Hubbard chain has 124 nonzeros out of 96 x 96), which run on
    the sparse small-operator kernel; inner sizes not multiples of 64."""
    import torch
    from dmrgpy_amd.mpo import build_mpo
    from dmrgpy_amd.multioperator import heisenberg_terms, hubbard_terms
    if model == "hubbard":
        mpo = build_mpo([1] * 8, hubbard_terms(8, 1.0, 4.0))
    else:
        mpo = build_mpo([2] * 8, heisenberg_terms(8))
    We1, We2 = mpo.W[3], mpo.W[4]                       # engine layout [1, wl, wr, s_out, s_in]
    W1 = We1[0].transpose(0, 3, 2, 1)                   # oracle [wl, s_in, s_out, wr]
    W2 = We2[0].transpose(0, 3, 2, 1)
    wl, d1, wm, d2, wr = W1.shape[0], W1.shape[1], W1.shape[3], W2.shape[1], W2.shape[3]
    Dl, Dr = 45, 77
    rng = np.random.default_rng(3 + prec + 5 * cplx)
    L = rnd(rng, (Dl, wl, Dl), cplx)
    R = rnd(rng, (Dr, wr, Dr), cplx)
    th = rnd(rng, (Dl, d1, d2, Dr), cplx)
    ref = odmrg.heff2_matvec(L, W1, W2, R, th.shape)(th.reshape(-1).astype(complex)).reshape(th.shape)
    et = engine.env_dtype(prec)
    desc = engine.heff_desc(dev(engine, L_eng(L), et), dev(engine, R_eng(R), et), We1, We2,
                            Dl, Dr, d1, d2, int(cplx), prec)
    x = dev(engine, th)
    y = engine.zeros(x.shape)
    engine.heff_apply(desc, x, y)
    torch.cuda.synchronize()
    err = relerr(host(y), ref if cplx else ref.real)
    assert err < TOL[prec] * 10, err


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("cplx", [False, True])
def test_heff1(engine, prec, cplx):
    import torch
    rng = np.random.default_rng(21 + prec)
    Dl, Dr, d, wl, wr = 17, 29, 3, 5, 4
    L = rnd(rng, (Dl, wl, Dl), cplx)
    R = rnd(rng, (Dr, wr, Dr), cplx)
    W = rnd(rng, (wl, d, d, wr), cplx)
    C = rnd(rng, (Dl, d, Dr), cplx)
    ref = odmrg.heff1_matvec(L, W, R, C.shape)(C.reshape(-1).astype(complex)).reshape(C.shape)
    et = engine.env_dtype(prec)
    desc = engine.heff_desc(dev(engine, L_eng(L), et), dev(engine, R_eng(R), et), W_eng(W), None,
                            Dl, Dr, d, 0, int(cplx), prec)
    x = dev(engine, C)
    y = engine.zeros(x.shape)
    engine.heff_apply(desc, x, y)
    torch.cuda.synchronize()
    assert relerr(host(y), ref if cplx else ref.real) < TOL[prec] * 10


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("direction", ["left", "right"])
@pytest.mark.parametrize("pitch", [0, 3])
def test_env(engine, prec, cplx, direction, pitch):
    """Environment update against the oracle; with pitch > 0 the input
    environment's padding columns are NaN and the output is written into a
    pitched buffer (padding untouched)."""
    import torch
    rng = np.random.default_rng(31 + prec)
    Dl, Dr, Dlb, Drb, d, wl, wr = 19, 27, 13, 21, 3, 4, 5
    ket = rnd(rng, (Dl, d, Dr), cplx)
    bra = rnd(rng, (Dlb, d, Drb), cplx)
    W = rnd(rng, (wl, d, d, wr), cplx)
    et = engine.env_dtype(prec)
    if direction == "left":
        Lin = rnd(rng, (Dlb, wl, Dl), cplx)
        ref = omps.extend_left(Lin, bra, W, ket)               # [Drb, wr, Dr]
        ein = pitched(engine, L_eng(Lin), et, pitch) if pitch else dev(engine, L_eng(Lin), et)
        eout = engine.zeros((ein.shape[0], wr, Drb, Dr + pitch), dtype=et)
        ref_e, n_out = L_eng(ref), Dr
    else:
        Rin = rnd(rng, (Drb, wr, Dr), cplx)
        ref = omps.extend_right(Rin, bra, W, ket)              # [Dlb, wl, Dl]
        ein = pitched(engine, R_eng(Rin), et, pitch) if pitch else dev(engine, R_eng(Rin), et)
        eout = engine.zeros((ein.shape[0], wl, Dl, Dlb + pitch), dtype=et)
        ref_e, n_out = R_eng(ref), Dlb
    engine.env_update(direction, int(cplx), prec, dev(engine, ket), dev(engine, bra), W_eng(W), ein, eout)
    torch.cuda.synchronize()
    assert relerr(host(unpitch(eout, n_out)), ref_e if cplx else ref_e.real) < TOL[prec] * 10
    if pitch:
        assert float(eout[..., n_out:].abs().max()) == 0.0


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("shape", [(24, 40), (64, 30), (8, 12), (128, 128)])
@pytest.mark.parametrize("absorb", [0, 1, 2])
def test_svd(engine, cplx, shape, absorb):
    import torch
    m, n = shape
    rng = np.random.default_rng(m * 7 + n)
    # decaying spectrum so cutoff and maxdim both matter
    k = min(m, n)
    U0 = np.linalg.qr(rnd(rng, (m, k), cplx))[0]
    V0 = np.linalg.qr(rnd(rng, (n, k), cplx))[0]
    s0 = np.exp(-0.35 * np.arange(k))
    M = (U0 * s0) @ V0.conj().T
    maxdim = max(2, k - 3)
    U, V, keep, disc, sv = engine.svd_split(int(cplx), dev(engine, M), m, n, 1e-10, maxdim, 1, absorb)
    torch.cuda.synchronize()
    Ur, Sr, Vr, keep_ref, disc_ref = olin.svd_truncated(M, 1e-10, maxdim, 1)
    assert keep == keep_ref
    assert abs(disc - disc_ref) < 1e-12
    # Gram route (both here and in the reference, svd.py:108-151): a singular
    # value s comes from an eigenvalue known to ~eps*s_max^2, i.e. to
    # ~eps*s_max^2/s absolute; the factor built by dividing by s inherits
    # ~eps*(s_max/s)^2 loss of orthonormality
    eps = 2.2e-16
    smax, skeep = Sr[0], Sr[keep - 1]
    assert np.all(np.abs(sv - Sr[:len(sv)]) <= 1e-13 + 8 * eps * smax ** 2 / np.maximum(Sr[:len(sv)], 1e-300))
    Uh, Vh = host(U), host(V)
    rec = (Uh * sv[:keep]) @ Vh if absorb == 2 else Uh @ Vh
    ref = (Ur * Sr[:keep]) @ Vr
    assert relerr(rec, ref) < 1e-10 + 100 * eps * (smax / skeep) ** 2
    orth = 1e-12 + 100 * eps * (smax / skeep) ** 2
    # the factor a sweep leaves behind (U when S goes right, Vh when S goes
    # left) is an isometry to rounding whatever the spectrum
    if absorb == 0:
        assert np.abs(Uh.conj().T @ Uh - np.eye(keep)).max() < 1e-12
    if absorb == 1:
        assert np.abs(Vh @ Vh.conj().T - np.eye(keep)).max() < 1e-12
    if absorb == 2:
        assert np.abs(Uh.conj().T @ Uh - np.eye(keep)).max() < orth
        assert np.abs(Vh @ Vh.conj().T - np.eye(keep)).max() < orth


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("shape", [(120, 200), (200, 120), (160, 160)])
@pytest.mark.parametrize("absorb", [0, 1])
def test_svd_isometry_near_guard(engine, cplx, shape, absorb):
    """Spectrum down to just above the Gram-route guard (1e-7 s_0): the
    reference's division by S would leave ~1e-2 orthonormality errors in the
    derived factor; the stay-behind factor must still be an isometry to
    1e-12 and the product must reconstruct M."""
    import torch
    m, n = shape
    rng = np.random.default_rng(m + 3 * n + absorb)
    k = min(m, n)
    U0 = np.linalg.qr(rnd(rng, (m, k), cplx))[0]
    V0 = np.linalg.qr(rnd(rng, (n, k), cplx))[0]
    s0 = np.logspace(0, -6.9, k)
    M = (U0 * s0) @ V0.conj().T
    U, V, keep, disc, sv = engine.svd_split(int(cplx), dev(engine, M), m, n, 0.0, 0, 1, absorb)
    torch.cuda.synchronize()
    assert keep == k
    Uh, Vh = host(U), host(V)
    if absorb == 0:
        assert np.abs(Uh.conj().T @ Uh - np.eye(keep)).max() < 1e-12
    else:
        assert np.abs(Vh @ Vh.conj().T - np.eye(keep)).max() < 1e-12
    assert relerr(Uh @ Vh, M) < 1e-12


def _hermitian_heff(rng, cplx, Dl=9, Dr=11, d=2, w=3):
    """Random L/R/W pieces giving a Hermitian two-site H_eff."""
    def herm_env(D, w):
        X = rnd(rng, (w, D, D), cplx)
        return (X + X.conj().transpose(0, 2, 1)) / 2
    Lw = herm_env(Dl, w)            # [w, a', a]
    Rw = herm_env(Dr, w)            # [w, b', b]
    def herm_w(wl, wr, d):
        X = rnd(rng, (wl, d, d, wr), cplx)
        return (X + X.conj().transpose(0, 2, 1, 3)) / 2   # W[wl, s, s', wr] hermitian in (s, s')
    W1 = herm_w(w, w, d)
    W2 = herm_w(w, w, d)
    L = Lw.transpose(1, 0, 2)        # oracle [a', w, a]
    R = Rw.transpose(1, 0, 2)        # oracle [b', w, b]
    return L, R, W1, W2


@pytest.mark.parametrize("cplx", [False, True])
def test_lanczos(engine, cplx):
    import torch
    rng = np.random.default_rng(41)
    L, R, W1, W2 = _hermitian_heff(rng, cplx)
    Dl, Dr, d = L.shape[0], R.shape[0], 2
    th = rnd(rng, (Dl, d, d, Dr), cplx)
    mv = odmrg.heff2_matvec(L, W1, W2, R, th.shape)
    e_ref, v_ref = olin.lanczos_ground(mv, th.reshape(-1).astype(complex), niter=200)
    H = np.column_stack([mv(col) for col in np.eye(th.size, dtype=complex)])
    e_exact = np.linalg.eigvalsh((H + H.conj().T) / 2)[0]
    desc = engine.heff_desc(dev(engine, L_eng(L)), dev(engine, R_eng(R)), W_eng(W1), W_eng(W2),
                            Dl, Dr, d, d, int(cplx), 0)
    x = dev(engine, th)
    e, it = engine.lanczos(desc, x, 200, 1e-12)
    torch.cuda.synchronize()
    assert abs(e - e_ref) < 1e-10 * max(1, abs(e_ref))
    assert abs(e - e_exact) < 1e-9 * max(1, abs(e_exact))
    v = host(x).reshape(-1)
    assert abs(np.linalg.norm(v) - 1) < 1e-12
    assert abs(abs(np.vdot(v, v_ref)) - 1) < 1e-8


def test_expm(engine):
    import torch
    from scipy.linalg import expm
    rng = np.random.default_rng(51)
    L, R, W1, W2 = _hermitian_heff(rng, True, Dl=6, Dr=7)
    Dl, Dr, d = 6, 7, 2
    th = rnd(rng, (Dl, d, d, Dr), True)
    mv = odmrg.heff2_matvec(L, W1, W2, R, th.shape)
    H = np.column_stack([mv(col) for col in np.eye(th.size, dtype=complex)])
    H = (H + H.conj().T) / 2
    coeff = -0.05j
    exact = expm(coeff * H) @ th.reshape(-1)
    ref = olin.lanczos_expm(mv, th.reshape(-1), coeff, niter=50)
    desc = engine.heff_desc(dev(engine, L_eng(L)), dev(engine, R_eng(R)), W_eng(W1), W_eng(W2),
                            Dl, Dr, d, d, 1, 0)
    x = dev(engine, th)
    it = engine.expm(desc, x, coeff, 50, 1e-12, 1e-10)
    torch.cuda.synchronize()
    out = host(x).reshape(-1)
    assert np.linalg.norm(out - ref) < 1e-11 * np.linalg.norm(ref)
    assert np.linalg.norm(out - exact) < 1e-9 * np.linalg.norm(exact)
    assert it >= 2
