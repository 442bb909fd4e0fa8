"""Subspace-guided truncated SVD (dmrg_svd_split_guided) against the exact
truncated SVD of the oracle (oracle/linalg.py::svd_truncated, the
reference's svd.py:108-151 rule) and against the library's full split.

The guided split is the exact truncated SVD of M restricted to span(guide)
enlarged by a random sketch; it is accepted only when the weight of M
outside that subspace is below 0.1 * cutoff.  So (a) when M lies in a
subspace the sketch can capture, the keep count, discarded weight,
singular values and the reconstructed product equal the exact ones to
rounding; (b) when it does not, the library falls back to the full split
and the result is identical to dmrg_svd_split; (c) in DMRG sweeps the
guided and full paths give the same energies.
"""

import numpy as np
import pytest

from oracle import linalg as olin

pytestmark = pytest.mark.gpu


def rnd(rng, shape, cplx):
    a = rng.standard_normal(shape)
    if cplx:
        a = a + 1j * rng.standard_normal(shape)
    return a


def dev(engine, a):
    import torch
    a = np.asarray(a)
    p = np.stack([a.real, a.imag]) if np.iscomplexobj(a) else a[None]
    return torch.tensor(np.ascontiguousarray(p), dtype=torch.float64, device=engine.device)


def host(t):
    a = t.double().cpu().numpy()
    return a[0] + 1j * a[1] if a.shape[0] == 2 else a[0]


@pytest.fixture(scope="module")
def engine():
    from dmrgpy_amd.engine import get_engine
    return get_engine(0)


def _problem(rng, cplx, m, n, k0, side, smin, delta, delta_rank):
    """M = A Q0 (side 0, Q0 k0 x n orthonormal rows) or Q0 A (side 1, Q0
    m x k0 orthonormal columns), spectrum of the guided part from 1 down to
    smin, plus a perturbation of norm `delta` and rank delta_rank (full
    rank if None)."""
    s = np.logspace(0, np.log10(smin), k0)
    if side == 0:
        Q0 = np.linalg.qr(rnd(rng, (n, k0), cplx))[0].conj().T
        Ua = np.linalg.qr(rnd(rng, (m, k0), cplx))[0]
        M = (Ua * s) @ Q0
    else:
        Q0 = np.linalg.qr(rnd(rng, (m, k0), cplx))[0]
        Va = np.linalg.qr(rnd(rng, (n, k0), cplx))[0].conj().T
        M = (Q0 * s) @ Va
    if delta_rank is None:
        P = rnd(rng, (m, n), cplx)
    else:
        P = rnd(rng, (m, delta_rank), cplx) @ rnd(rng, (delta_rank, n), cplx)
    M = M + delta * P / np.linalg.norm(P)
    return M, Q0


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("side", [0, 1])
@pytest.mark.parametrize("smin", [1e-5, 1e-7])
def test_guided_equals_exact(engine, cplx, side, smin):
    import torch
    m, n, k0 = 480, 512, 96
    rng = np.random.default_rng(11 + 2 * side + int(cplx) + int(-np.log10(smin)))
    M, Q0 = _problem(rng, cplx, m, n, k0, side, smin, 1e-4, 20)
    absorb = 0 if side == 0 else 1
    U, V, keep, disc, sv = engine.svd_split(int(cplx), dev(engine, M), m, n, 1e-12, 0, 1, absorb,
                                            guide=dev(engine, Q0), guide_side=side, sketch=64)
    torch.cuda.synchronize()
    assert engine.last_svd_guided == 1
    Ur, Sr, Vr, keep_ref, disc_ref = olin.svd_truncated(M, 1e-12, None, 1)
    assert keep == keep_ref
    assert abs(disc - disc_ref) < 1e-14
    eps = 2.2e-16
    assert np.all(np.abs(sv[:keep] - Sr[:keep]) <= 1e-13 + 8 * eps / Sr[:keep])
    assert np.all(sv[k0 + 64:] == 0.0)
    Uh, Vh = host(U), host(V)
    ref = (Ur * Sr[:keep]) @ Vr
    skeep = Sr[keep - 1]
    assert np.linalg.norm(Uh @ Vh - ref) <= (1e-10 + 100 * eps / skeep ** 2) * np.linalg.norm(ref)
    if absorb == 0:
        assert np.abs(Uh.conj().T @ Uh - np.eye(keep)).max() < 1e-12
    else:
        assert np.abs(Vh @ Vh.conj().T - np.eye(keep)).max() < 1e-12


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("side", [0, 1])
def test_guided_falls_back(engine, cplx, side):
    """A full-rank perturbation of weight 1e-6 >> cutoff cannot be captured
    by a 64-row sketch: the result must be the full split's, bit for bit."""
    import torch
    m, n, k0 = 400, 448, 80
    rng = np.random.default_rng(5 + side + 2 * int(cplx))
    M, Q0 = _problem(rng, cplx, m, n, k0, side, 1e-4, 1e-3, None)
    absorb = 0 if side == 0 else 1
    Md = dev(engine, M)
    U, V, keep, disc, sv = engine.svd_split(int(cplx), Md, m, n, 1e-12, 200, 1, absorb,
                                            guide=dev(engine, Q0), guide_side=side, sketch=64)
    assert engine.last_svd_guided == 0
    U2, V2, keep2, disc2, sv2 = engine.svd_split(int(cplx), Md, m, n, 1e-12, 200, 1, absorb)
    torch.cuda.synchronize()
    assert keep == keep2 and disc == disc2
    assert np.array_equal(sv, sv2)
    assert torch.equal(U, U2) and torch.equal(V, V2)


def test_guided_not_worth_small(engine):
    """k0 + sketch above 0.6 min(m, n): the full split runs."""
    import torch
    rng = np.random.default_rng(2)
    M, Q0 = _problem(rng, False, 64, 64, 32, 0, 1e-3, 1e-5, 4)
    engine.svd_split(0, dev(engine, M), 64, 64, 1e-12, 0, 1, 0, guide=dev(engine, Q0), guide_side=0, sketch=16)
    torch.cuda.synchronize()
    assert engine.last_svd_guided == 0


def _sweeps(guided, maxdim, sketch, nsweeps=5):
    from dmrgpy_amd import algorithms as alg
    from dmrgpy_amd.chain import Chain
    from dmrgpy_amd.engine import get_engine
    from dmrgpy_amd.multioperator import hubbard_terms
    n = 16
    eng = get_engine(0)
    try:
        ch = Chain([1] * n, precision="fp64", field="real")
        ch.set_sweep_params(maxdim, 1, 1e-12, 0.0)
        ch.set_verbose(False)
        ch.set_hamiltonian(hubbard_terms(n, 1.0, 4.0))
        assert ch.eng is eng
        eng.svd_guided = guided
        eng.svd_sketch = sketch
        np.random.seed(7)
        psi = ch._default_mps(maxm=maxdim)
        es = []
        for _ in range(nsweeps):
            stats = alg.Stats()
            es.append(alg.dmrg_sweep(eng, psi, ch.H, maxdim, 1e-12, 2, 0, stats))
    finally:
        eng.svd_guided = False
        eng.svd_sketch = 128
    resid = [r for r in stats.svd_resid if r >= 0]
    return np.array(es), float(np.mean(stats.svd_guided)), resid


def test_guided_sweeps_maxdim_limited():
    """Hubbard N=16 (native d=4 sites) at maxdim 32, where maxdim forces every
    bulk truncation: the guided split carries the converged sweeps and the
    energies follow the full split's within a small fraction of the
    truncation error (E(D=32) - E(D=64) is ~1e-4 here)."""
    e_full, f_full, _ = _sweeps(False, 32, 32)
    e_gd, f_gd, resid = _sweeps(True, 32, 32)
    print("maxdim 32: guided fraction", f_gd, "max resid", max(resid) if resid else None)
    print("energies full", e_full, "guided", e_gd)
    assert f_full == 0.0
    assert f_gd > 0.5
    assert abs(e_gd[-1] - e_full[-1]) / 16 < 1e-7


def test_guided_sweeps_cutoff_limited():
    """Same chain at maxdim 64 with cutoff 1e-12 (cutoff-limited bonds): the
    guided split is accepted only where the missed weight is below a tenth
    of the cutoff, so the energies equal the full split's to 1e-10 per site
    whatever the acceptance rate."""
    e_full, _, _ = _sweeps(False, 64, 24)
    e_gd, f_gd, resid = _sweeps(True, 64, 24)
    print("maxdim 64: guided fraction", f_gd, "resid range", (min(resid), max(resid)) if resid else None)
    assert np.abs(e_gd - e_full).max() / 16 < 1e-10
