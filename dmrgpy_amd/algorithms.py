"""DMRG / TDVP drivers over the device engine.

The sweep structure restates the reference pure-Python backend so results
match it bond for bond:

* ``dmrg_sweep``    -- ``_dmrg_one_sweep`` (src/dmrgpy/pyitensor/dmrg.py:355-382):
  right environments rebuilt from scratch, left-to-right then right-to-left,
  energy = last local eigenvalue;
* ``local_ground``  -- ``_local_ground_state`` (dmrg.py:302-333), Lanczos cap
  ``max(niter, 200)``, tolerance 1e-12;
* ``update``        -- ``_apply_local_update`` (dmrg.py:336-352);
* ``tdvp_step``     -- ``tdvp_step`` / ``_half_sweep_lr`` / ``_half_sweep_rl``
  (src/dmrgpy/pyitensor/tdvp.py:220-273, 350-368);
* ``apply_mpo``     -- zip-up ``applyMPO`` (src/dmrgpy/pyitensor/mpsalgebra.py:226-302);
* ``random_mps``    -- ``randomMPS`` (mpsalgebra.py:315-353), same draws from
  the legacy global ``np.random``.

Every contraction, Krylov step and SVD runs in libdmrg_hip.so; this module
only sequences calls and keeps the device buffers.
"""

import math

import numpy as np

from .mps import DeviceMPS

LANCZOS_TOL = 1e-12         # dmrg.py:95
LANCZOS_CAP = 200           # dmrg.py:330
KRYLOV_ERRGOAL = 1e-10      # tdvp.py:79
KRYLOV_TOL = 1e-12          # tdvp.py:82


class Stats:
    """Counters of one run (matvecs per local solve, truncation)."""

    def __init__(self):
        self.lanczos_iters = []
        self.krylov_iters = []
        self.truncation = []      # (bond, keep, discarded weight)
        self.svd_guided = []      # 1 where the guided subspace split produced the result
        self.svd_resid = []       # its measured weight outside the subspace (-1: not attempted)
        self.bond_shapes = []     # (bond, Dl, d1, d2, Dr) of every two-site local solve

    def reset(self):
        self.__init__()


def ones_env(eng, field, prec):
    t = eng.zeros((2 if field else 1, 1, 1, 1), dtype=eng.env_dtype(prec))
    t[0] = 1.0
    return t


def theta2(eng, psi, i):
    """A_i * A_{i+1} as a flat [np, Dl*d1*d2*Dr] FP64 vector (1-based bond i)."""
    A, B = psi.A(i), psi.A(i + 1)
    npl, Dl, d1, Dm = A.shape
    _, _, d2, Dr = B.shape
    th = eng.empty((npl, Dl * d1 * d2 * Dr))
    eng.gemm(0, psi.field, A, B, th, Dl * d1, d2 * Dr, Dm, planeA=Dl * d1 * Dm, planeB=Dm * d2 * Dr,
             planeC=Dl * d1 * d2 * Dr)
    return th, (Dl, d1, d2, Dr)


def env_left(eng, L, A, W, field, prec, bra=None):
    """L[w][D_bra][D_ket] one site further right; the last index is stored
    with a row pitch rounded up to a multiple of 4 (Engine.env_pitch)."""
    bra = A if bra is None else bra
    out = eng.empty((L.shape[0], W.shape[2], bra.shape[3], eng.env_pitch(A.shape[3])), dtype=L.dtype)
    eng.env_update("left", field, prec, A, bra, W, L, out)
    return out


def env_right(eng, R, A, W, field, prec, bra=None):
    """R[w][D_ket][D_bra] one site further left (last index pitched like env_left)."""
    bra = A if bra is None else bra
    out = eng.empty((R.shape[0], W.shape[1], A.shape[1], eng.env_pitch(bra.shape[1])), dtype=R.dtype)
    eng.env_update("right", field, prec, A, bra, W, R, out)
    return out


def right_envs(eng, psi, mpo, prec):
    """{i: R_i}, R_i covering sites i..N (dmrg.py:204-211)."""
    n = psi.length()
    env = {n + 1: ones_env(eng, psi.field, prec)}
    for i in range(n, 1, -1):
        env[i] = env_right(eng, env[i + 1], psi.A(i), mpo.W[i - 1], psi.field, prec)
    return env


def left_envs(eng, psi, mpo, prec):
    """{i: L_i}, L_i covering sites 1..i (dmrg.py:214-221)."""
    n = psi.length()
    env = {0: ones_env(eng, psi.field, prec)}
    for i in range(1, n):
        env[i] = env_left(eng, env[i - 1], psi.A(i), mpo.W[i - 1], psi.field, prec)
    return env


def heff2(eng, L, R, mpo, i, dims, field, prec):
    Dl, d1, d2, Dr = dims
    return eng.heff_desc(L, R, mpo.W[i - 1], mpo.W[i], Dl, Dr, d1, d2, field, prec)


def heff1(eng, L, R, mpo, i, dims, field, prec):
    Dl, d, Dr = dims
    return eng.heff_desc(L, R, mpo.W[i - 1], None, Dl, Dr, d, 0, field, prec)


# Ritz-residual floor of the split-BF16x3 operator (the first stage of a
# mixed-precision local solve)
BF16X3_RES_TOL = 1e-4


def local_ground(eng, L, R, mpo, psi, i, niter, prec, stats=None, tol=LANCZOS_TOL, res_tol=0.0):
    th, dims = theta2(eng, psi, i)
    if prec == 3:
        # mixed: BF16x3 Lanczos down to its residual floor, then FP32 Lanczos
        # warm-started from that vector down to res_tol
        e, it0 = eng.lanczos(heff2(eng, L, R, mpo, i, dims, psi.field, 2), th, max(niter, LANCZOS_CAP), tol,
                             max(res_tol, BF16X3_RES_TOL))
        e, it = eng.lanczos(heff2(eng, L, R, mpo, i, dims, psi.field, 1), th, max(niter, LANCZOS_CAP), tol,
                            res_tol)
        it += it0
    else:
        desc = heff2(eng, L, R, mpo, i, dims, psi.field, prec)
        e, it = eng.lanczos(desc, th, max(niter, LANCZOS_CAP), tol, res_tol)
    if stats is not None:
        stats.lanczos_iters.append(it)
        stats.bond_shapes.append((i,) + tuple(dims))
    return e, th, dims


def update(eng, psi, i, th, dims, cutoff, maxdim, direction, stats=None):
    """Split theta back into sites i, i+1 (dmrg.py:336-352).  In a sweep the
    site the split leaves behind was an isometry before the update (A_{i+1}
    right-orthonormal when moving right, A_i left-orthonormal when moving
    left) whose span holds almost all of theta; it guides the subspace split
    when eng.svd_guided is set (the library falls back to the full split
    whenever the subspace misses weight)."""
    Dl, d1, d2, Dr = dims
    npl = th.shape[0]
    guide, side = None, 0
    if getattr(eng, "svd_guided", False):
        guide, side = (psi.A(i + 1), 0) if direction == "right" else (psi.A(i), 1)
    U, V, keep, disc, _ = eng.svd_split(psi.field, th, Dl * d1, d2 * Dr, cutoff, maxdim, 1,
                                        0 if direction == "right" else 1, guide=guide, guide_side=side)
    psi.set_A(i, U.view(npl, Dl, d1, keep))
    psi.set_A(i + 1, V.view(npl, keep, d2, Dr))
    psi.center = i + 1 if direction == "right" else i
    if stats is not None:
        stats.truncation.append((i, keep, disc))
        stats.svd_guided.append(eng.last_svd_guided)
        stats.svd_resid.append(eng.last_svd_resid)
    return keep, disc


def dmrg_sweep(eng, psi, mpo, maxdim, cutoff, niter, prec=0, stats=None, tol=LANCZOS_TOL, res_tol=0.0):
    n = psi.length()
    energy = None
    R = right_envs(eng, psi, mpo, prec)
    L = {0: ones_env(eng, psi.field, prec)}
    for i in range(1, n):
        energy, th, dims = local_ground(eng, L[i - 1], R[i + 2], mpo, psi, i, niter, prec, stats, tol, res_tol)
        update(eng, psi, i, th, dims, cutoff, maxdim, "right", stats)
        L[i] = env_left(eng, L[i - 1], psi.A(i), mpo.W[i - 1], psi.field, prec)
        R.pop(i + 2, None)
    R = {n + 1: ones_env(eng, psi.field, prec)}
    for i in range(n - 1, 0, -1):
        energy, th, dims = local_ground(eng, L[i - 1], R[i + 2], mpo, psi, i, niter, prec, stats, tol, res_tol)
        update(eng, psi, i, th, dims, cutoff, maxdim, "left", stats)
        R[i + 1] = env_right(eng, R[i + 2], psi.A(i + 1), mpo.W[i], psi.field, prec)
        L.pop(i, None)
        R.pop(i + 2, None)
    return energy


# ---------------------------------------------------------------------------
# excited states: overlap-penalty DMRG (src/dmrgpy/pyitensor/dmrg.py:520-681)
# ---------------------------------------------------------------------------
def _ones_mpo_site(d):
    return np.eye(d)[None, None, None]


def overlap_left(eng, L, wk, psi, i):
    """<wk|psi> left environment one site further ([1][D_wk][D_psi])."""
    return env_left(eng, L, psi.A(i), _ones_mpo_site(psi.A(i).shape[2]), psi.field, 0, bra=wk.A(i))


def overlap_right(eng, R, wk, psi, i):
    return env_right(eng, R, psi.A(i), _ones_mpo_site(psi.A(i).shape[2]), psi.field, 0, bra=wk.A(i))


def bond_projection(eng, Lov, wk, Rov, i, dims, field):
    """wk projected onto psi's two-site basis at bond i (dmrg.py:556-561):
    p[a][s1 s2][b] = sum_{a' b'} conj(Lov[a'][a]) theta_wk[a'][s1 s2][b'] conj(Rov[b][b'])."""
    Dl, d1, d2, Dr = dims
    th, (Dla, _, _, Drb) = theta2(eng, wk, i)
    npl = 2 if field else 1
    ldL, ldR = Lov.shape[-1], Rov.shape[-1]
    T = eng.empty((npl, Dl * d1 * d2 * Drb))
    eng.gemm(0, field, Lov, th, T, Dl, d1 * d2 * Drb, Dla, transA=True, conjA=bool(field), lda=ldL,
             planeA=Dla * ldL, planeB=Dla * d1 * d2 * Drb, planeC=Dl * d1 * d2 * Drb)
    P = eng.empty((npl, Dl * d1 * d2 * Dr))
    eng.gemm(0, field, T, Rov, P, Dl * d1 * d2, Dr, Drb, transB=True, conjB=bool(field), ldb=ldR,
             planeA=Dl * d1 * d2 * Drb, planeB=Dr * ldR, planeC=Dl * d1 * d2 * Dr)
    return P


def dmrg_sweep_excited(eng, psi, mpo, penalty, weight, maxdim, cutoff, niter, prec=0, stats=None,
                       tol=LANCZOS_TOL, res_tol=0.0):
    """One penalised sweep (dmrg_excited, dmrg.py:620-681): the local
    operator is H_eff + weight * sum_k |p_k><p_k| with p_k the projection of
    the k-th penalty state onto the current two-site basis; overlap
    environments <wk|psi> are rebuilt from the right at the start of the
    sweep and extended with every update, like the H environments."""
    n = psi.length()
    field = psi.field
    energy = None
    R = right_envs(eng, psi, mpo, prec)
    one = lambda: ones_env(eng, field, 0)
    Rov = []
    for wk in penalty:
        env = {n + 1: one()}
        for i in range(n, 1, -1):
            env[i] = overlap_right(eng, env[i + 1], wk, psi, i)
        Rov.append(env)
    L = {0: ones_env(eng, field, prec)}
    Lov = [{0: one()} for _ in penalty]

    def solve(i):
        th, dims = theta2(eng, psi, i)
        desc = heff2(eng, L[i - 1], R[i + 2], mpo, i, dims, field, prec)
        if penalty:
            import torch
            P = torch.stack([bond_projection(eng, Lov[k][i - 1], wk, Rov[k][i + 2], i, dims, field)
                             for k, wk in enumerate(penalty)])
            eng.set_projectors(desc, P, weight)
        e, it = eng.lanczos(desc, th, max(niter, LANCZOS_CAP), tol, res_tol)
        if stats is not None:
            stats.lanczos_iters.append(it)
        return e, th, dims

    for i in range(1, n):
        energy, th, dims = solve(i)
        update(eng, psi, i, th, dims, cutoff, maxdim, "right", stats)
        L[i] = env_left(eng, L[i - 1], psi.A(i), mpo.W[i - 1], field, prec)
        for k, wk in enumerate(penalty):
            Lov[k][i] = overlap_left(eng, Lov[k][i - 1], wk, psi, i)
        R.pop(i + 2, None)
    R = {n + 1: ones_env(eng, field, prec)}
    Rov = [{n + 1: one()} for _ in penalty]
    for i in range(n - 1, 0, -1):
        energy, th, dims = solve(i)
        update(eng, psi, i, th, dims, cutoff, maxdim, "left", stats)
        R[i + 1] = env_right(eng, R[i + 2], psi.A(i + 1), mpo.W[i], field, prec)
        for k, wk in enumerate(penalty):
            Rov[k][i + 1] = overlap_right(eng, Rov[k][i + 2], wk, psi, i + 1)
        L.pop(i, None)
    return energy


def expectation(eng, bra, mpo, ket, prec=0):
    """<bra|M|ket> by a full left-to-right environment sweep (mpsalgebra.py:99-114)."""
    field = max(bra.field, ket.field, 1 if mpo.complex else 0)
    bra = bra.as_complex() if field and not bra.field else bra
    ket = ket.as_complex() if field and not ket.field else ket
    L = ones_env(eng, field, prec)
    for i in range(1, ket.length() + 1):
        L = env_left(eng, L, ket.A(i), mpo.W[i - 1], field, prec, bra=bra.A(i))
    v = L.double().cpu().numpy()
    return complex(v[0, 0, 0, 0], v[1, 0, 0, 0] if field else 0.0)


def identity_mpo(dims):
    from .mpo import MPO
    return MPO([np.eye(d)[None, None, None] for d in dims], codes=[None] * len(dims))


def overlap(eng, bra, ket):
    """<bra|ket> (mpsalgebra.py:89-98)."""
    return expectation(eng, bra, identity_mpo(ket.phys_dims()), ket)


def scale_site(psi, i, factor):
    """Multiply the state by a scalar through site i (host-side complex
    factor applied to the planes)."""
    t = psi.A(i)
    f = complex(factor)
    if psi.field == 0:
        if f.imag != 0:
            raise ValueError("complex factor on a real MPS")
        t.mul_(f.real)
    else:
        re, im = t[0].clone(), t[1].clone()
        t[0] = f.real * re - f.imag * im
        t[1] = f.real * im + f.imag * re


def normalize(eng, psi):
    nrm = math.sqrt(max(overlap(eng, psi, psi).real, 0.0))
    scale_site(psi, psi.center or 1, 1.0 / nrm)
    return psi


def shift_right(eng, psi, i, cutoff=0.0, maxdim=None):
    """mpscontainer.py:104-119 (1-based site i)."""
    T, nxt = psi.A(i), psi.A(i + 1)
    npl, Dl, d, Dr = T.shape
    U, V, keep, _, _ = eng.svd_split(psi.field, T, Dl * d, Dr, cutoff, maxdim, 1, 0)
    psi.set_A(i, U.view(npl, Dl, d, keep))
    _, _, d2, Dr2 = nxt.shape
    out = eng.empty((npl, keep, d2, Dr2))
    eng.gemm(0, psi.field, V, nxt, out, keep, d2 * Dr2, Dr, planeA=keep * Dr, planeB=Dr * d2 * Dr2,
             planeC=keep * d2 * Dr2)
    psi.set_A(i + 1, out)


def shift_left(eng, psi, i, cutoff=0.0, maxdim=None):
    """mpscontainer.py:121-130 (1-based site i)."""
    T, prv = psi.A(i), psi.A(i - 1)
    npl, Dl, d, Dr = T.shape
    U, V, keep, _, _ = eng.svd_split(psi.field, T, Dl, d * Dr, cutoff, maxdim, 1, 1)
    psi.set_A(i, V.view(npl, keep, d, Dr))
    _, Dl0, d0, _ = prv.shape
    out = eng.empty((npl, Dl0, d0, keep))
    eng.gemm(0, psi.field, prv, U, out, Dl0 * d0, keep, Dl, planeA=Dl0 * d0 * Dl, planeB=Dl * keep,
             planeC=Dl0 * d0 * keep)
    psi.set_A(i - 1, out)


def position(eng, psi, b, cutoff=0.0, maxdim=None):
    while psi.center < b:
        shift_right(eng, psi, psi.center, cutoff, maxdim)
        psi.center += 1
    while psi.center > b:
        shift_left(eng, psi, psi.center, cutoff, maxdim)
        psi.center -= 1
    return psi


def random_bond_dims(dims, m):
    n = len(dims)
    bonds, cum = [], 1
    for d in dims[:-1]:
        cum = min(cum * d, m)
        bonds.append(cum)
    cum = 1
    for k in range(n - 2, -1, -1):
        cum = min(cum * dims[k + 1], m)
        bonds[k] = min(bonds[k], cum)
    return bonds


def random_mps_arrays(dims, m, field=1, rng=None):
    """The host site arrays of randomMPS (mpsalgebra.py:315-353), each
    [Dl][d][Dr]: per site ``randn(*shape) + 1j*randn(*shape)`` from the
    legacy global numpy generator (or ``rng``) with the reference's boundary
    shapes; field 0 keeps the real part of the same draws."""
    rs = np.random if rng is None else rng
    n = len(dims)
    bonds = random_bond_dims(dims, m)
    arrays = []
    for i in range(n):
        shape = ([bonds[i - 1]] if i > 0 else []) + [dims[i]] + ([bonds[i]] if i < n - 1 else [])
        a = rs.randn(*shape) + 1j * rs.randn(*shape)
        Dl = bonds[i - 1] if i > 0 else 1
        Dr = bonds[i] if i < n - 1 else 1
        a = a.reshape(Dl, dims[i], Dr)
        arrays.append(a if field else a.real)
    return arrays


def random_mps(eng, dims, m, field=1, rng=None):
    """randomMPS (mpsalgebra.py:315-353): the draws of random_mps_arrays,
    canonicalised to centre N on the device and normalised."""
    arrays = random_mps_arrays(dims, m, field, rng)
    n = len(dims)
    psi = DeviceMPS.from_numpy(arrays, eng.device, field=field, center=1)
    position(eng, psi, n)
    normalize(eng, psi)
    return psi


# ---------------------------------------------------------------------------
# MPS sums (mpsalgebra.py:143-232)
# ---------------------------------------------------------------------------
def mps_sum_many(eng, states, cutoff=0.0, maxdim=None):
    """Direct sum of K states (sum_many, mpsalgebra.py:143-223): interior
    sites block-diagonal in the link spaces, the boundary sites concatenated
    along their one link; then one truncating left-to-right SVD sweep from
    site 1 to site N (position(N, cutoff, maxdim)), as the reference does."""
    states = list(states)
    if not states:
        raise ValueError("mps_sum: no states given")
    n = states[0].length()
    if any(s.length() != n for s in states):
        raise ValueError("mps_sum: states of different lengths")
    if any(s.phys_dims() != states[0].phys_dims() for s in states):
        raise ValueError("mps_sum: different physical dimensions")
    field = max(s.field for s in states)
    states = [s.as_complex() if field and not s.field else s for s in states]
    npl = 2 if field else 1
    tensors = []
    for i in range(1, n + 1):
        Ts = [s.A(i) for s in states]
        d = Ts[0].shape[2]
        Dl = 1 if i == 1 else sum(t.shape[1] for t in Ts)
        Dr = 1 if i == n else sum(t.shape[3] for t in Ts)
        out = eng.zeros((npl, Dl, d, Dr))
        lo = ro = 0
        for t in Ts:
            _, a, _, b = t.shape
            ls = slice(0, 1) if i == 1 else slice(lo, lo + a)
            rs = slice(0, 1) if i == n else slice(ro, ro + b)
            out[:, ls, :, rs] = t
            lo += a
            ro += b
        tensors.append(out)
    res = DeviceMPS(tensors, center=1, field=field)
    position(eng, res, n, cutoff, maxdim)
    return res


def mps_sum(eng, a, b, cutoff=0.0, maxdim=None):
    """a + b (sum, mpsalgebra.py:226-232)."""
    return mps_sum_many(eng, [a, b], cutoff, maxdim)


def scaled(psi, factor):
    """factor * psi as a new state (the scalar goes into the centre site,
    or site 1)."""
    f = complex(factor)
    p = psi.as_complex() if (f.imag != 0 and not psi.field) else psi.copy()
    scale_site(p, p.center or 1, f)
    return p


def gram_schmidt(eng, wfs, cutoff=0.0, maxdim=None):
    """Sequential Gram-Schmidt of MPS (Chain._gram_schmidt, chain.py:1233-1241):
    wf_i <- normalise(wf_i - sum over j < i, one at a time, of wf_j <wf_j|wf_i>),
    each subtraction an MPS sum truncated to cutoff / maxdim."""
    wfs = list(wfs)
    for i in range(1, len(wfs)):
        for j in range(i):
            c = overlap(eng, wfs[j], wfs[i])
            wf = mps_sum(eng, wfs[i], scaled(wfs[j], -c), cutoff, maxdim)
            normalize(eng, wf)
            wfs[i] = wf
    return wfs


# ---------------------------------------------------------------------------
# MPO x MPS
# ---------------------------------------------------------------------------
def apply_mpo(eng, mpo, psi, cutoff=0.0, maxdim=None):
    """Zip-up K|x> (mpsalgebra.py:226-302): per site contract the carried
    remainder with the MPS and MPO tensors, split off a left-orthonormal U
    with a truncated SVD, carry S*V on; centre N."""
    field = max(psi.field, 1 if mpo.complex else 0)
    x = psi.as_complex() if field and not psi.field else psi
    npl = 2 if field else 1
    n = x.length()
    left = eng.zeros((npl, 1, 1, 1))          # [k, wl, Dl]
    left[0] = 1.0
    out = []
    for i in range(1, n + 1):
        X = x.A(i)
        W = mpo.W[i - 1]                       # [np_w, wl, wr, s', s]
        _, k, wl, Dl = left.shape
        _, _, d, Dr = X.shape
        wr = W.shape[2]
        # T1[(k wl)][(s Dr)] = left[(k wl)][Dl] X[Dl][(s Dr)]
        T1 = eng.empty((npl, k * wl * d * Dr))
        eng.gemm(0, field, left, X, T1, k * wl, d * Dr, Dl, planeA=k * wl * Dl, planeB=Dl * d * Dr,
                 planeC=k * wl * d * Dr)
        # T2[k][(s' wr)][Dr] = Wm[(s' wr)][(wl s)] T1[k][(wl s)][Dr]
        Wc = W[0] + (1j * W[1] if W.shape[0] == 2 else 0)
        Wm = Wc.transpose(2, 1, 0, 3).reshape(d * wr, wl * d)
        Wd = eng.torch.from_numpy(np.ascontiguousarray(
            np.stack([Wm.real] + ([Wm.imag] if field else [])), dtype=np.float64)).to(eng.device)
        T2 = eng.empty((npl, k * d * wr * Dr))
        eng.gemm(0, field, Wd, T1, T2, d * wr, Dr, wl * d, strideA=0, strideB=wl * d * Dr,
                 strideC=d * wr * Dr, planeA=d * wr * wl * d, planeB=k * wl * d * Dr,
                 planeC=k * d * wr * Dr, batch=k)
        if i == n:
            out.append(T2.view(npl, k, d, wr * Dr))
            break
        U, V, keep, _, _ = eng.svd_split(field, T2, k * d, wr * Dr, cutoff, maxdim, 1, 0)
        out.append(U.view(npl, k, d, keep))
        left = V.view(npl, keep, wr, Dr)
    return DeviceMPS(out, center=n, field=field)


def apply_local(eng, psi, site, M):
    """Exact single-site operator application (standard-convention matrix
    M[out, in]) without changing the gauge."""
    field = max(psi.field, 1 if np.iscomplexobj(M) and np.abs(np.imag(M)).max() > 0 else 0)
    p = psi.as_complex() if field and not psi.field else psi.copy()
    A = p.A(site)
    npl, Dl, d, Dr = A.shape
    Mc = np.asarray(M, dtype=complex)
    Md = eng.torch.from_numpy(np.ascontiguousarray(
        np.stack([Mc.real] + ([Mc.imag] if field else [])), dtype=np.float64)).to(eng.device)
    out = eng.empty(A.shape)
    eng.gemm(0, field, Md, A, out, d, Dr, d, strideA=0, strideB=d * Dr, strideC=d * Dr,
             planeA=d * d, planeB=Dl * d * Dr, planeC=Dl * d * Dr, batch=Dl)
    p.set_A(site, out)
    return p


# ---------------------------------------------------------------------------
# TDVP
# ---------------------------------------------------------------------------
def _evolve2(eng, L, R, mpo, psi, i, coeff, niter, prec, stats):
    th, dims = theta2(eng, psi, i)
    desc = heff2(eng, L, R, mpo, i, dims, psi.field, prec)
    it = eng.expm(desc, th, coeff, niter, KRYLOV_TOL, KRYLOV_ERRGOAL)
    if stats is not None:
        stats.krylov_iters.append(it)
    return th, dims


def _evolve1(eng, L, R, mpo, C, i, coeff, niter, prec, stats):
    npl, Dl, d, Dr = C.shape
    x = C.reshape(npl, Dl * d * Dr).clone()
    desc = heff1(eng, L, R, mpo, i, (Dl, d, Dr), 1, prec)
    it = eng.expm(desc, x, coeff, niter, KRYLOV_TOL, KRYLOV_ERRGOAL)
    if stats is not None:
        stats.krylov_iters.append(it)
    return x.view(npl, Dl, d, Dr)


def half_sweep_lr(eng, psi, mpo, tau, cutoff, maxdim, niter, prec=0, stats=None):
    n = psi.length()
    R = right_envs(eng, psi, mpo, prec)
    L = {0: ones_env(eng, 1, prec)}
    for i in range(1, n):
        th, dims = _evolve2(eng, L[i - 1], R[i + 2], mpo, psi, i, -1j * tau, niter, prec, stats)
        update(eng, psi, i, th, dims, cutoff, maxdim, "right", stats)
        L[i] = env_left(eng, L[i - 1], psi.A(i), mpo.W[i - 1], 1, prec)
        if i < n - 1:
            psi.set_A(i + 1, _evolve1(eng, L[i], R[i + 2], mpo, psi.A(i + 1), i + 1, 1j * tau, niter, prec,
                                      stats))
    psi.center = n


def half_sweep_rl(eng, psi, mpo, tau, cutoff, maxdim, niter, prec=0, stats=None):
    n = psi.length()
    L = left_envs(eng, psi, mpo, prec)
    R = {n + 1: ones_env(eng, 1, prec)}
    for i in range(n - 1, 0, -1):
        th, dims = _evolve2(eng, L[i - 1], R[i + 2], mpo, psi, i, -1j * tau, niter, prec, stats)
        update(eng, psi, i, th, dims, cutoff, maxdim, "left", stats)
        R[i + 1] = env_right(eng, R[i + 2], psi.A(i + 1), mpo.W[i], 1, prec)
        if i > 1:
            psi.set_A(i, _evolve1(eng, L[i - 1], R[i + 1], mpo, psi.A(i), i, 1j * tau, niter, prec, stats))
    psi.center = 1


def tdvp_step(eng, psi, mpo, dt, cutoff, maxdim, niter=50, prec=0, stats=None):
    """One step exp(-i dt H) (tdvp.py:350-368, two-site); psi is made complex."""
    if not psi.field:
        c = psi.as_complex()
        psi.tensors, psi.field = c.tensors, 1
    tau = dt / 2.0
    half_sweep_lr(eng, psi, mpo, tau, cutoff, maxdim, niter, prec, stats)
    half_sweep_rl(eng, psi, mpo, tau, cutoff, maxdim, niter, prec, stats)
    return psi
