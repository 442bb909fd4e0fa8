"""Chain: the session object of the MI355X backend.

Drop-in for the reference's session protocol, ``pyitensor.chain.Chain``
(src/dmrgpy/pyitensor/chain.py:142-564), which dmrgpy's front end creates
with ``backend.Chain(site_types)`` (src/dmrgpy/sites.py:20-24) and drives
through ``set_sweep_params`` -> ``set_verbose`` -> ``set_mpomaxm`` ->
``set_hamiltonian`` -> ``gs_energy`` (src/dmrgpy/groundstate.py:27-58).
Same method names, argument meaning, defaults (maxm 30, nsweeps 15, cutoff
1e-12) and errors (RuntimeError before set_hamiltonian).  Wave-function
handles are ``DeviceMPS`` objects (mutable, ``copy()`` gives an independent
device copy).

Backend-specific settings (all optional, defaults reproduce the reference):

* ``precision``: "fp64" (reference arithmetic), "fp32", "bf16x3" or "mixed"
  (each local solve in split-BF16x3 down to its residual floor, then FP32)
  for the H_eff / environment contractions; Krylov vectors, the tridiagonal
  problems and the SVD stay FP64, and reduced-precision runs finish with an
  FP64 Rayleigh quotient <psi|H|psi>/<psi|psi>;
* ``precision_schedule``: per-sweep precisions (e.g. reduced precision
  early, FP64 in the last sweep);
* ``maxdim_schedule``: per-sweep bond dimensions (a D ramp);
* ``field``: "auto" (complex random MPS as the reference draws it) or
  "real" (real part of the same draws; real arithmetic when H is real).
"""

import numpy as np

from . import algorithms as alg
from . import sites as _sites
from .engine import PREC_NAMES, get_engine
from .mpo import build_mpo, hamiltonian_mpo
from .mps import DeviceMPS

BUILD_CUTOFF = 1e-14   # chain.py:38 (MPO build cutoff; the automaton MPO is exact)

# reduced-precision Lanczos: relative Ritz-residual stop above the rounding floor
RES_TOL = {1: 2e-6, 2: 1e-4, 3: 2e-6}


class Chain:
    def __init__(self, site_types, device=0, precision="fp64", field="auto"):
        self.site_types = list(site_types)
        self.dims = [_sites.dim(c) for c in self.site_types]
        self.eng = get_engine(device)
        self.H = None
        self.have_H = False
        self.wf0 = None
        self._wf0_energy = None
        self.maxm = 30
        self.nsweeps = 15
        self.cutoff = 1e-12
        self.noise = 1e-1
        self.mpomaxm = 5000
        self.verbose = False
        self.precision = precision
        self.precision_schedule = None
        self.maxdim_schedule = None
        self.field = field
        self.niter = 2
        self.stats = alg.Stats()
        self.sweep_energies = []

    # -- session protocol --------------------------------------------------
    def num_sites(self):
        return len(self.site_types)

    def site_dim(self, site):
        return self.dims[site - 1]

    def random_mps(self):
        return self._default_mps()

    def set_sweep_params(self, maxm, nsweeps, cutoff, noise):
        self.maxm = int(maxm)
        self.nsweeps = int(nsweeps)
        self.cutoff = float(cutoff)
        self.noise = noise

    def set_mpomaxm(self, mpomaxm):
        self.mpomaxm = mpomaxm

    def set_verbose(self, verbose):
        self.verbose = bool(verbose)

    def set_hamiltonian(self, terms):
        self.H = hamiltonian_mpo(self.site_types, terms, BUILD_CUTOFF)
        self.H_terms = list(terms)
        self.have_H = True
        self._wf0_energy = None
        self._bandwidth_max = None

    def build_operator(self, terms):
        return build_mpo(self.site_types, terms)

    def gs_energy(self, skip_dmrg=False):
        if not self.have_H:
            raise RuntimeError("Chain.gs_energy called before set_hamiltonian")
        if skip_dmrg and self.wf0 is not None and self._wf0_energy is not None:
            return self._wf0_energy
        if self.wf0 is None:
            self.wf0 = self._default_mps()
        self._match_field(self.wf0, self.H)
        energy = self._dmrg(self.wf0, self.H)
        self._wf0_energy = energy
        return energy

    def gs_wavefunction(self):
        if self.wf0 is None:
            raise RuntimeError("Chain.gs_wavefunction called before gs_energy")
        return self.wf0

    def set_wavefunction(self, wf):
        if not isinstance(wf, DeviceMPS):
            wf = DeviceMPS.from_numpy(list(wf), self.eng.device, center=None)
        self.wf0 = wf
        self._wf0_energy = None

    def vev(self, terms, wf, npow=1):
        """<wf|A^npow|wf>/<wf|wf> (chain.py:277-287)."""
        A = build_mpo(self.site_types, terms)
        psi = wf.copy()
        nrm = max(alg.overlap(self.eng, psi, psi).real, 0.0) ** 0.5
        alg.scale_site(psi, psi.center or 1, 1.0 / nrm)
        if npow == 1:
            return alg.expectation(self.eng, psi, A, psi)
        psi1 = psi
        for _ in range(npow - 1):
            psi1 = self._apply_mpo(A, psi1)
        return alg.expectation(self.eng, psi, A, psi1)

    def overlap(self, wf1, wf2):
        return alg.overlap(self.eng, wf1, wf2)

    def overlap_aMb(self, wf1, terms, wf2):
        return alg.expectation(self.eng, wf1, build_mpo(self.site_types, terms), wf2)

    def apply_operator(self, terms, wf):
        return self._apply_mpo(build_mpo(self.site_types, terms), wf)

    def sum_mps(self, wf1, wf2):
        """wf1 + wf2 truncated to the session's cutoff / maxm (chain.py:300-301)."""
        return alg.mps_sum(self.eng, wf1, wf2, self.cutoff, self.maxm)

    def _gram_schmidt(self, wfs):
        """chain.py:1233-1241."""
        return alg.gram_schmidt(self.eng, wfs, self.cutoff, self.maxm)

    def tdvp_step(self, H, wf, dt, num_center=2):
        """(H, wf, dt) argument order as chain.py:397-420."""
        if num_center != 2:
            raise NotImplementedError("the MI355X backend implements two-site TDVP only")
        psi = wf.copy()
        return alg.tdvp_step(self.eng, psi, H, dt, self.cutoff, self.maxm, 50, self._prec(None), self.stats)

    def quench_tdvp(self, terms_h, terms_i, terms_j, nt, dt):
        """chain.py:520-542."""
        if self.wf0 is None:
            self.gs_energy()
        H = build_mpo(self.site_types, terms_h)
        wf0 = self.wf0
        EGS = (alg.expectation(self.eng, wf0, H, wf0) / alg.overlap(self.eng, wf0, wf0)).real
        Hshift = build_mpo(self.site_types, list(terms_h) + [(-EGS, [("Id", 1)])])
        A1 = build_mpo(self.site_types, terms_i)
        A2 = build_mpo(self.site_types, terms_j)
        psi1 = self._apply_mpo(A1, wf0)
        psi2 = self._apply_mpo(A2, wf0)
        norm0 = np.sqrt(complex(alg.overlap(self.eng, psi1, psi1)))
        corr = []
        for _ in range(int(nt)):
            corr.append(alg.overlap(self.eng, psi2, psi1))
            psi1 = alg.tdvp_step(self.eng, psi1, Hshift, dt, self.cutoff, self.maxm, 50, self._prec(None),
                                 self.stats)
            alg.normalize(self.eng, psi1)
            alg.scale_site(psi1, psi1.center or 1, norm0)
        return corr, psi1

    def evolve_and_measure_tdvp(self, terms_h, terms_op, wf, nt, dt):
        """chain.py:544-564: measure <psi|A|psi> before every step."""
        H = build_mpo(self.site_types, terms_h)
        A = build_mpo(self.site_types, terms_op)
        psi = wf.copy()
        corr = []
        for _ in range(int(nt)):
            corr.append(alg.expectation(self.eng, psi, A, psi))
            psi = alg.tdvp_step(self.eng, psi, H, dt, self.cutoff, self.maxm, 50, self._prec(None), self.stats)
        return corr, psi

    # -- helpers ------------------------------------------------------------
    def _prec(self, sweep):
        name = self.precision
        if sweep is not None and self.precision_schedule:
            name = self.precision_schedule[min(sweep, len(self.precision_schedule) - 1)]
        return PREC_NAMES[name] if isinstance(name, str) else int(name)

    def _default_mps(self, maxm=None):
        field = 0 if (self.field == "real" and not (self.H is not None and self.H.complex)) else 1
        m = maxm if maxm is not None else (self.maxdim_schedule[0] if self.maxdim_schedule else self.maxm)
        return alg.random_mps(self.eng, self.dims, m, field)

    def _match_field(self, psi, H):
        if H.complex and not psi.field:
            c = psi.as_complex()
            psi.tensors, psi.field = c.tensors, 1

    def _apply_mpo(self, K, x):
        return alg.apply_mpo(self.eng, K, x, self.cutoff, self.maxm)

    # -- generalized eigenproblem (chain.py:230-251, dmrg.py:425-517) -------
    def gs_energy_generalized(self, terms_a, lam0=None):
        """Smallest lambda of H|psi> = lambda A|psi> (A Hermitian positive
        definite): each sweep is an ordinary DMRG sweep on H - lambda A
        (built as one automaton MPO from both term lists), after which lambda
        becomes <psi|H|psi>/<psi|A|psi>; lambda starts from lam0 or, when
        lam0 is None or NaN, from the current state's quotient.  Mutates
        wf0 and invalidates the plain-energy cache, like the reference."""
        if not self.have_H:
            raise RuntimeError("Chain.gs_energy_generalized called before set_hamiltonian")
        A = build_mpo(self.site_types, terms_a)
        if self.wf0 is None:
            self.wf0 = self._default_mps()
        psi = self.wf0
        self._match_field(psi, self.H)
        self._match_field(psi, A)
        ratio = lambda: (alg.expectation(self.eng, psi, self.H, psi).real,
                         alg.expectation(self.eng, psi, A, psi).real)
        unset = lam0 is None or np.isnan(complex(lam0).real)
        if unset:
            h0, a0 = ratio()
            lam = h0 / a0 if abs(a0) > 1e-14 else 0.0
        else:
            lam = complex(lam0).real
        for k in range(self.nsweeps):
            maxdim = self.maxdim_schedule[min(k, len(self.maxdim_schedule) - 1)] if self.maxdim_schedule \
                else self.maxm
            Hl = build_mpo(self.site_types, list(self.H_terms) + [(-lam * complex(c), f) for c, f in terms_a])
            self._match_field(psi, Hl)
            prec = self._prec(k)
            alg.dmrg_sweep(self.eng, psi, Hl, maxdim, self.cutoff, self.niter, prec, self.stats,
                           res_tol=RES_TOL.get(prec, 0.0))
            h, a = ratio()
            if not abs(a) > 1e-14:
                raise RuntimeError("gs_energy_generalized: <psi|A|psi> collapsed to ~0 (or NaN)")
            lam = h / a
            if self.verbose:
                print("sweep {}: lambda = {}".format(k, lam))
        self._wf0_energy = None
        return lam

    # -- excited states (chain.py:253-275, 1201-1231) ---------------------
    def _maximum_energy(self):
        """Reduced-effort DMRG on -H from a fresh random MPS (<= 5 sweeps,
        maxdim <= 20): a spectral bound for the penalty weight."""
        if getattr(self, "_bandwidth_max", None) is None:
            psi = self._default_mps()
            neg = build_mpo(self.site_types, [(-complex(c), f) for c, f in self.H_terms])
            self._match_field(psi, neg)
            e = None
            for _ in range(min(self.nsweeps, 5)):
                e = alg.dmrg_sweep(self.eng, psi, neg, min(self.maxm, 20), self.cutoff, self.niter, 0)
            self._bandwidth_max = -e
        return self._bandwidth_max

    def _bandwidth(self):
        return self._maximum_energy() - self.gs_energy(skip_dmrg=True)

    def _energy_fluctuation(self, psi, H):
        """<H^2> - <H>^2 of the normalised state, with H|psi> by zip-up."""
        p = psi.copy()
        alg.normalize(self.eng, p)
        hp = self._apply_mpo(H, p)
        de = alg.overlap(self.eng, p, hp).real
        return alg.overlap(self.eng, hp, hp).real - de * de

    def excited_states(self, n, scale_lagrange=1.0, do_gram_schmidt=False):
        """The n lowest eigenstates by overlap-penalty DMRG (chain.py:253-275):
        state k minimises <psi|H + w sum_{j<k} |psi_j><psi_j| |psi> from a
        fresh random MPS, w = bandwidth * scale_lagrange.  Returns
        (energies <psi|H|psi>, fluctuations <H^2> - <H>^2, wave functions)."""
        if not self.have_H:
            raise RuntimeError("Chain.excited_states called before set_hamiltonian")
        if self.wf0 is None:
            self.gs_energy()
        psi0 = self.wf0.copy()
        alg.normalize(self.eng, psi0)
        wfs = [psi0]
        weight = self._bandwidth() * scale_lagrange
        for _ in range(1, int(n)):
            psi1 = self._default_mps()
            self._match_field(psi1, self.H)
            if psi1.field != psi0.field:
                wfs = [w.as_complex() if not w.field else w for w in wfs]
                psi1 = psi1.as_complex() if not psi1.field else psi1
            for k in range(self.nsweeps):
                maxdim = self.maxdim_schedule[min(k, len(self.maxdim_schedule) - 1)] if self.maxdim_schedule \
                    else self.maxm
                prec = self._prec(k)
                alg.dmrg_sweep_excited(self.eng, psi1, self.H, wfs, weight, maxdim, self.cutoff, self.niter,
                                       prec, self.stats, res_tol=RES_TOL.get(prec, 0.0))
            alg.normalize(self.eng, psi1)
            wfs.append(psi1)
        if do_gram_schmidt:
            wfs = self._gram_schmidt(wfs)
        energies, fluct = [], []
        for wf in wfs:
            fluct.append(self._energy_fluctuation(wf, self.H))
            energies.append((alg.expectation(self.eng, wf, self.H, wf) / alg.overlap(self.eng, wf, wf)).real)
        return energies, fluct, wfs

    def last_sweep_truncation(self):
        """[(bond, keep, discarded weight)] of every split of the last
        sweep: the discarded weight is the reference's Spectrum.truncerr
        (src/dmrgpy/pyitensor/svd.py:16-29,195), which its dmrg computes and
        drops."""
        return list(self._last_truncation)

    def truncation_weight(self):
        """Total discarded weight of the last sweep (sum over its splits)."""
        return float(sum(t[2] for t in self._last_truncation))

    _last_truncation = ()

    def _dmrg(self, psi, H, callback=None):
        energy = None
        self.sweep_energies = []
        for k in range(self.nsweeps):
            t0 = len(self.stats.truncation)
            maxdim = self.maxdim_schedule[min(k, len(self.maxdim_schedule) - 1)] if self.maxdim_schedule \
                else self.maxm
            prec = self._prec(k)
            energy = alg.dmrg_sweep(self.eng, psi, H, maxdim, self.cutoff, self.niter, prec, self.stats,
                                    res_tol=RES_TOL.get(prec, 0.0))
            self.sweep_energies.append(energy)
            self._last_truncation = tuple(self.stats.truncation[t0:])
            if self.verbose:
                print("sweep {}: energy = {}".format(k, energy))
            if callback is not None:
                callback(k, energy, psi)
        if self._prec(self.nsweeps - 1) != 0:
            # FP64 Rayleigh quotient of the final state
            energy = (alg.expectation(self.eng, psi, H, psi) / alg.overlap(self.eng, psi, psi)).real
        return energy
