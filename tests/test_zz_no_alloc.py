"""dmrg_heff_apply_noalloc (GPU): the allocation-free variant of
dmrg_heff_apply, whose workspace is the library's static device scratch
array, must give the same y as dmrg_heff_apply (same kernels, so bit for
bit) on two calls in a row, the second with a smaller operator, for every
contraction precision, real and complex vectors, sparse (model) and dense
MPO pairs and penalty projectors; operators too large for the array are
rejected with an error before any kernel runs."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _operator(engine, rng, Dl, Dr, prec, cplx, model):
    from dmrgpy_amd.mpo import build_mpo
    from dmrgpy_amd.multioperator import heisenberg_terms, hubbard_terms
    if model == "hubbard":
        mpo = build_mpo([1] * 6, hubbard_terms(6, 1.0, 4.0))
        W1, W2 = mpo.W[2], mpo.W[3]                     # [1, wl, wr, s_out, s_in]
    elif model == "heisenberg":
        mpo = build_mpo([2] * 6, heisenberg_terms(6))
        W1, W2 = mpo.W[2], mpo.W[3]
    else:                                               # dense random pair
        W1 = rng.standard_normal((1, 4, 5, 3, 3))
        W2 = rng.standard_normal((1, 5, 3, 3, 3))
    npl = 2 if cplx else 1
    wl, wm, d1 = W1.shape[1], W1.shape[2], W1.shape[3]
    wr, d2 = W2.shape[2], W2.shape[3]
    et = engine.env_dtype(prec)
    import torch
    L = torch.from_numpy(rng.standard_normal((npl, wl, Dl, engine.env_pitch(Dl)))).to(et).to(engine.device)
    R = torch.from_numpy(rng.standard_normal((npl, wr, Dr, engine.env_pitch(Dr)))).to(et).to(engine.device)
    desc = engine.heff_desc(L, R, W1, W2, Dl, Dr, d1, d2, int(cplx), prec)
    x = torch.from_numpy(rng.standard_normal((npl, Dl * d1 * d2 * Dr))).to(engine.device)
    return desc, x


@pytest.mark.parametrize("model", ["hubbard", "heisenberg", "dense"])
@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("prec", [0, 1, 2])
def test_noalloc_matches_heff_apply(engine, prec, cplx, model):
    import torch
    rng = np.random.default_rng(100 + 10 * prec + 3 * cplx + len(model))
    for Dl, Dr in ((96, 83), (40, 21)):                 # second call smaller
        print("noalloc", model, "prec", prec, "cplx", cplx, "Dl", Dl, "Dr", Dr, flush=True)
        desc, x = _operator(engine, rng, Dl, Dr, prec, cplx, model)
        y_ref = torch.full_like(x, float("nan"))
        y = torch.full_like(x, float("nan"))
        engine.heff_apply(desc, x, y_ref)
        engine.heff_apply_noalloc(desc, x, y)
        torch.cuda.synchronize()
        assert torch.isfinite(y).all()
        diff = (y - y_ref).abs().max().item()
        print("   max |y_noalloc - y| =", diff, flush=True)
        assert torch.equal(y, y_ref), diff


def test_noalloc_with_projectors(engine):
    import torch
    rng = np.random.default_rng(7)
    for Dl, Dr in ((64, 48), (24, 16)):
        desc, x = _operator(engine, rng, Dl, Dr, 0, False, "heisenberg")
        P = torch.from_numpy(rng.standard_normal((3, 1, x.shape[1]))).to(engine.device)
        engine.set_projectors(desc, P, 2.5)
        y_ref = torch.full_like(x, float("nan"))
        y = torch.full_like(x, float("nan"))
        engine.heff_apply(desc, x, y_ref)
        engine.heff_apply_noalloc(desc, x, y)
        torch.cuda.synchronize()
        assert torch.equal(y, y_ref), (y - y_ref).abs().max().item()


def test_noalloc_rejects_operators_beyond_the_scratch_array(engine):
    """A real FP64 two-site operator at D=512 (d=2, w=5) needs ~84 MB of
    workspace, beyond the 64 MiB array: rejected with a message, y untouched.
    The buffers are allocated at full size all the same."""
    import torch
    rng = np.random.default_rng(3)
    desc, x = _operator(engine, rng, 512, 512, 0, False, "heisenberg")
    y = torch.full_like(x, 7.0)
    with pytest.raises(RuntimeError, match="scratch array"):
        engine.heff_apply_noalloc(desc, x, y)
    torch.cuda.synchronize()
    assert bool((y == 7.0).all())
    # the regular entry point still serves it, and the allocation-free one
    # still works afterwards on a size that fits
    engine.heff_apply(desc, x, y)
    desc2, x2 = _operator(engine, rng, 32, 32, 0, False, "heisenberg")
    y2, y2_ref = torch.empty_like(x2), torch.empty_like(x2)
    engine.heff_apply(desc2, x2, y2_ref)
    engine.heff_apply_noalloc(desc2, x2, y2)
    torch.cuda.synchronize()
    assert torch.equal(y2, y2_ref)
