"""Device-resident MPS handle.

Site tensors live in HBM as float64 torch tensors ``[np, Dl, d, Dr]`` (np = 1
real, 2 complex planes), dimension-1 links at the open ends.  The handle is
mutable like the reference's pure-Python MPS and ``copy()`` returns an
independent device copy (``src/dmrgpy/mps.py:45-95`` copies through the
handle's own ``copy()``).
"""

import numpy as np


class DeviceMPS:
    def __init__(self, tensors, center=None, field=1):
        self.tensors = list(tensors)
        self.center = center          # 1-based orthogonality centre, or None
        self.field = field            # 0 real, 1 complex

    def length(self):
        return len(self.tensors)

    __len__ = length

    def A(self, i):
        return self.tensors[i - 1]

    def set_A(self, i, t):
        self.tensors[i - 1] = t

    def copy(self):
        return DeviceMPS([t.clone() for t in self.tensors], self.center, self.field)

    def bond_dims(self):
        return [t.shape[3] for t in self.tensors[:-1]]

    def phys_dims(self):
        return [t.shape[2] for t in self.tensors]

    def to_numpy(self):
        """Site tensors as complex host arrays [Dl, d, Dr]."""
        out = []
        for t in self.tensors:
            a = t.detach().cpu().numpy()
            out.append(a[0] + 1j * a[1] if a.shape[0] == 2 else a[0].astype(complex))
        return out

    def to_dense(self):
        arrs = self.to_numpy()
        v = arrs[0][0]
        for a in arrs[1:]:
            v = np.tensordot(v, a, axes=([v.ndim - 1], [0]))
        return v.reshape(-1)

    @staticmethod
    def from_numpy(arrays, device, field=None, center=None):
        import torch
        arrays = [np.asarray(a) for a in arrays]
        if field is None:
            field = 1 if any(np.iscomplexobj(a) and np.abs(a.imag).max() > 0 for a in arrays) else 0
        ts = []
        for a in arrays:
            a = a.astype(complex)
            planes = [a.real] + ([a.imag] if field == 1 else [])
            ts.append(torch.from_numpy(np.ascontiguousarray(np.stack(planes), dtype=np.float64)).to(device))
        return DeviceMPS(ts, center, field)

    def as_complex(self):
        """The same state with complex (two-plane) storage."""
        import torch
        if self.field == 1:
            return self
        ts = [torch.cat([t, torch.zeros_like(t)], dim=0) for t in self.tensors]
        return DeviceMPS(ts, self.center, 1)
