"""The MPS handle (dmrgpy_amd/mps.py) as a container, on CPU tensors (CPU):
``copy()`` is independent of the original -- the reference's handles are
mutable and dmrgpy copies through the handle's own ``copy()``
(src/dmrgpy/mps.py:45-95; tests/test_time_evolution.py:453-494 checks the
same independence) -- and the host round trips keep the state: numpy ->
planes -> numpy, the dense vector against the oracle's, real -> complex
storage, and the field detection of ``from_numpy``.
"""

import numpy as np
import torch

from dmrgpy_amd.mps import DeviceMPS
from oracle import mps as om


def _state(dims=(2, 3, 2, 2), m=4, cplx=True):
    arrays = om.random_draws(list(dims), m)
    return [a if cplx else a.real for a in arrays]


def test_copy_is_independent():
    np.random.seed(1)
    psi = DeviceMPS.from_numpy(_state(), "cpu", field=1, center=2)
    before = [t.clone() for t in psi.tensors]
    cp = psi.copy()
    assert cp.center == 2 and cp.field == 1 and cp.length() == psi.length()
    # in-place work on the copy (what a TDVP step does) leaves the original alone
    for t in cp.tensors:
        t.mul_(-3.0).add_(1.0)
    cp.set_A(1, torch.zeros_like(cp.A(1)))
    cp.center = 4
    assert psi.center == 2
    assert all(torch.equal(a, b) for a, b in zip(psi.tensors, before))
    # and the other way round
    cp2 = psi.copy()
    psi.tensors[2].zero_()
    assert not torch.equal(cp2.tensors[2], psi.tensors[2])
    assert torch.equal(cp2.tensors[2], before[2])


def test_host_round_trips():
    np.random.seed(2)
    arrays = _state()
    psi = DeviceMPS.from_numpy(arrays, "cpu")
    assert psi.field == 1                                   # complex draws detected
    assert psi.bond_dims() == [a.shape[2] for a in arrays[:-1]]
    assert psi.phys_dims() == [a.shape[1] for a in arrays]
    for a, b in zip(psi.to_numpy(), arrays):
        assert np.array_equal(a, b)
    ref = om.MPS([a.copy() for a in arrays], center=1).to_dense()
    assert np.array_equal(psi.to_dense(), ref)
    # real storage: one plane, detected from real input; complex view adds a zero plane
    np.random.seed(3)
    real = _state(cplx=False)
    pr = DeviceMPS.from_numpy(real, "cpu")
    assert pr.field == 0 and all(t.shape[0] == 1 for t in pr.tensors)
    pc = pr.as_complex()
    assert pc.field == 1 and all(t.shape[0] == 2 and torch.count_nonzero(t[1]) == 0 for t in pc.tensors)
    assert np.array_equal(pc.to_dense(), pr.to_dense())
    assert pc.as_complex() is pc
