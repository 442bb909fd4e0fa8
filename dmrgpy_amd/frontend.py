"""dmrgpy-compatible front end for the MI355X backend.

The user API of the reference (src/dmrgpy/manybodychain.py:38-576,
spinchain.py:48-63, fermionchain.py:210-463) selects a numerical backend
with ``itensor_version=`` and then calls ``set_hamiltonian``,
``gs_energy``, ``get_gs`` and ``vev`` on MultiOperators.  This module keeps
those names and call semantics for the backend value ``"mi355x"``; the
session is ``dmrgpy_amd.chain.Chain``.  Only the hot-path calls are
provided (ground state, expectation values, TDVP time evolution).  Chains:
``Spin_Chain``, ``Fermionic_Chain``, ``Spinful_Fermionic_Chain`` (2n
interleaved spinless sites) and ``Spinful_Fermionic_Chain_Native`` (n d=4
sites).
"""

import numpy as np

from . import multioperator as mo

BACKEND = "mi355x"

LABEL2SITE = {"1/2": 2, "S=1/2": 2, 2: 2, "1": 3, "S=1": 3, 3: 3, "3/2": 4, "S=3/2": 4, 4: 4,
              "2": 5, "S=2": 5, 5: 5, "5/2": 6, "S=5/2": 6, 6: 6}   # spinchain.py:17-34


def available(version):
    """cppext.available analogue (src/dmrgpy/cppext.py:38-64)."""
    if version != BACKEND:
        return False
    try:
        from . import _lib
        _lib.load()
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


class Many_Body_Chain:
    def __init__(self, sites, itensor_version=BACKEND, precision="fp64", field="auto", device=0):
        if itensor_version != BACKEND:
            raise ValueError("this package provides only itensor_version='{}'".format(BACKEND))
        self.sites = list(sites)
        self.ns = len(self.sites)
        self.itensor_version = itensor_version
        self.maxm = 30
        self.nsweeps = 15
        self.cutoff = 1e-12
        self.noise = 1e-1
        self.mpomaxm = 5000
        self.verbose = False
        self.hamiltonian = None
        self.wf0 = None
        self.e0 = None
        self.computed_gs = False
        self.skip_dmrg_gs = False
        self.tevol_method = "TDVP"
        from .chain import Chain
        self._session = Chain(self.sites, device=device, precision=precision, field=field)
        self._ham_key = None

    # operators
    def get_operator(self, name, i):
        return mo.obj2MO([[name, i]])

    def set_hamiltonian(self, h):
        self.hamiltonian = h
        self.computed_gs = False
        self.wf0 = None
        self.e0 = None
        self._session.wf0 = None

    # ground state (src/dmrgpy/groundstate.py:21-84)
    def gs_energy(self, wf0=None, reconverge=None):
        if self.computed_gs and reconverge is None and wf0 is None:
            return self.e0
        s = self._session
        s.set_sweep_params(self.maxm, self.nsweeps, self.cutoff, self.noise)
        s.set_verbose(self.verbose)
        s.set_mpomaxm(max(self.maxm, self.mpomaxm))
        terms = self.hamiltonian.to_terms()
        key = (self.maxm, self.nsweeps, self.cutoff, self.noise, max(self.maxm, self.mpomaxm), repr(terms))
        if key != self._ham_key:
            s.set_hamiltonian(terms)
            self._ham_key = key
        if wf0 is not None:
            s.set_wavefunction(wf0)
        if reconverge is not None:
            self.skip_dmrg_gs = not reconverge
        out = s.gs_energy(skip_dmrg=self.skip_dmrg_gs)
        self.e0 = out
        self.computed_gs = True
        self.skip_dmrg_gs = True
        self.wf0 = s.gs_wavefunction().copy()
        return out

    def get_gs(self):
        if not self.computed_gs:
            self.gs_energy()
        return self.wf0

    def vev(self, MO, npow=1, wf=None):
        """<GS|MO|GS> (src/dmrgpy/vev.py:23-35)."""
        if wf is None:
            wf = self.get_gs()
        return self._session.vev(MO.to_terms(), wf, npow)

    def overlap(self, wf1, wf2):
        return self._session.overlap(wf1, wf2)

    def applyoperator(self, MO, wf):
        return self._session.apply_operator(MO.to_terms(), wf)

    # real-time evolution (src/dmrgpy/timedependent.py:52-247, TDVP branch)
    def evolution_DC(self, A, B, nt=100, dt=0.05):
        """C(t) = <GS|B^dag(t) A|GS>-type quench correlator via quench_tdvp;
        returns (times, values)."""
        self.get_gs()
        corr, _ = self._session.quench_tdvp(self.hamiltonian.to_terms(), A.to_terms(),
                                            B.get_dagger().to_terms(), nt, dt)
        return np.arange(nt) * dt, np.array(corr)

    def evolve_and_measure(self, operator, nt=100, dt=0.05, wf=None):
        wf = self.get_gs() if wf is None else wf
        corr, _ = self._session.evolve_and_measure_tdvp(self.hamiltonian.to_terms(), operator.to_terms(),
                                                        wf, nt, dt)
        return np.arange(nt) * dt, np.array(corr)


class Spin_Chain(Many_Body_Chain):
    """spinchain.Spin_Chain: sites given by labels ("S=1/2", "1", ...)."""

    def __init__(self, sites, **kwargs):
        codes = [LABEL2SITE[s] for s in sites]
        Many_Body_Chain.__init__(self, codes, **kwargs)
        self.Sx = [self.get_operator("Sx", i) for i in range(self.ns)]
        self.Sy = [self.get_operator("Sy", i) for i in range(self.ns)]
        self.Sz = [self.get_operator("Sz", i) for i in range(self.ns)]
        self.Si = [self.Sx, self.Sy, self.Sz]

    def SS(self, i, j):
        return self.Sx[i] * self.Sx[j] + self.Sy[i] * self.Sy[j] + self.Sz[i] * self.Sz[j]


class Fermionic_Chain(Many_Body_Chain):
    """Spinless fermions (site code 0) with C, Cdag, N operators."""

    def __init__(self, n, **kwargs):
        Many_Body_Chain.__init__(self, [0] * n, **kwargs)
        self.C = [self.get_operator("C", i) for i in range(n)]
        self.Cdag = [self.get_operator("Cdag", i) for i in range(n)]
        self.N = [self.get_operator("N", i) for i in range(n)]


class Spinful_Fermionic_Chain_Native(Many_Body_Chain):
    """Native spinful fermions (site code 1, d = 4), fermionchain.py:426-463."""

    def __init__(self, n, **kwargs):
        Many_Body_Chain.__init__(self, [1] * n, **kwargs)
        g = self.get_operator
        self.Cup = [g("Cup", i) for i in range(n)]
        self.Cdagup = [g("Cdagup", i) for i in range(n)]
        self.Cdn = [g("Cdn", i) for i in range(n)]
        self.Cdagdn = [g("Cdagdn", i) for i in range(n)]
        self.Nup = [g("Nup", i) for i in range(n)]
        self.Ndn = [g("Ndn", i) for i in range(n)]
        self.Ntot = [self.Nup[i] + self.Ndn[i] for i in range(n)]
        self.Sz = [0.5 * self.Nup[i] - 0.5 * self.Ndn[i] for i in range(n)]


class Spinful_Fermionic_Chain(Fermionic_Chain):
    """fermionchain.py:210-233: n orbitals as 2n interleaved spinless-fermion
    sites (up on site 2i, down on site 2i+1, 0-based), with the spin,
    pairing and number operators built from them."""

    def __init__(self, n, **kwargs):
        Fermionic_Chain.__init__(self, 2 * n, **kwargs)
        C, Cdag, N = self.C, self.Cdag, self.N
        self.Sx = [0.5 * Cdag[2 * i] * C[2 * i + 1] + 0.5 * Cdag[2 * i + 1] * C[2 * i] for i in range(n)]
        self.Sy = [-0.5j * Cdag[2 * i] * C[2 * i + 1] + 0.5j * Cdag[2 * i + 1] * C[2 * i] for i in range(n)]
        self.Sz = [0.5 * N[2 * i] + (-0.5) * N[2 * i + 1] for i in range(n)]
        self.Delta = [0.5 * C[2 * i] * C[2 * i + 1] for i in range(n)]
        self.Cup = [C[2 * i] for i in range(n)]
        self.Cdagup = [Cdag[2 * i] for i in range(n)]
        self.Cdn = [C[2 * i + 1] for i in range(n)]
        self.Cdagdn = [Cdag[2 * i + 1] for i in range(n)]
        self.Nup = [N[2 * i] for i in range(n)]
        self.Ndn = [N[2 * i + 1] for i in range(n)]
        self.Ntot = [self.Nup[i] + self.Ndn[i] for i in range(n)]
