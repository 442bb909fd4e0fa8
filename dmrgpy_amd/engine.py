"""Device engine: owns one libdmrg_hip context on one GPU and exposes the C
ABI entry points over torch device buffers.

Tensors handed to the engine follow the library layouts (see
include/dmrg_hip.h): a leading "plane" axis of size 1 (real) or 2 (complex,
real plane then imaginary plane), then row-major data.  MPS site tensors
and Krylov vectors are float64; environments are float64 for FP64
contractions and float32 for the FP32 / split-BF16x3 modes.  MPO tensors
stay on the host as numpy float64 arrays ``[np, wl, wr, s_out, s_in]``.
"""

import ctypes

import numpy as np

from . import _lib
from ._lib import EnvDesc, HeffDesc, check

PREC_NAMES = {"fp64": 0, "fp32": 1, "bf16x3": 2, "mixed": 3}
# "mixed" (host-level mode 3): FP32 environments; each local Lanczos solve
# runs split-BF16x3 contractions to the BF16x3 residual floor, then continues
# from that vector with FP32 contractions (algorithms.local_ground)
MIXED = 3
# rows of the random sketch that enlarges a guide subspace in svd_split
SVD_SKETCH = 128


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _hptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class Engine:
    def __init__(self, device=0):
        import torch
        if not torch.cuda.is_available():
            raise RuntimeError("dmrgpy_amd: no GPU visible (the engine runs only on a HIP device)")
        self.torch = torch
        self.device = torch.device("cuda", device)
        torch.cuda.set_device(self.device)
        self.stream = torch.cuda.current_stream(self.device)
        self.lib = _lib.load()
        ctx = ctypes.c_void_p()
        check(self.lib.dmrg_ctx_create(device, ctypes.c_void_p(self.stream.cuda_stream), ctypes.byref(ctx)),
              "dmrg_ctx_create")
        self.ctx = ctx
        self._keep = []
        self.last_svd_guided = 0
        self.last_svd_resid = -1.0
        # two-site splits in sweeps use the neighbouring isometry as a guide
        # subspace (dmrg_svd_split_guided) when True.  Off by default: the
        # full split is the reference's rule; the guided one pays where
        # maxdim limits the bonds or the cutoff is loose.
        self.svd_guided = False
        self.svd_sketch = SVD_SKETCH

    def close(self):
        if getattr(self, "ctx", None):
            self.lib.dmrg_ctx_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- allocation helpers --------------------------------------------
    def empty(self, shape, dtype=None):
        return self.torch.empty(shape, dtype=dtype or self.torch.float64, device=self.device)

    def zeros(self, shape, dtype=None):
        return self.torch.zeros(shape, dtype=dtype or self.torch.float64, device=self.device)

    @staticmethod
    def env_pitch(D):
        """Row pitch for an environment's last index: a multiple of 4
        elements, so the H_eff and environment GEMMs load 16-byte vectors."""
        return (int(D) + 3) // 4 * 4

    def env_dtype(self, prec):
        return self.torch.float64 if prec == 0 else self.torch.float32

    @staticmethod
    def contraction_prec(prec):
        """Library precision code for a host mode (mixed contracts environments in FP32)."""
        return 1 if prec == MIXED else int(prec)

    def sync(self):
        check(self.lib.dmrg_sync(self.ctx), "dmrg_sync")

    # -- descriptors ------------------------------------------------------
    def heff_desc(self, L, R, W1, W2, Dl, Dr, d1, d2, field, prec):
        """W1/W2: host float64 arrays [np, wl, wm|wr, d, d] (W2 None for one site)."""
        d = HeffDesc()
        d.nsite = 2 if W2 is not None else 1
        d.field = field
        d.prec = self.contraction_prec(prec)
        d.Dl, d.Dr, d.d1 = Dl, Dr, d1
        d.d2 = d2 if W2 is not None else 0
        W1 = np.ascontiguousarray(W1, dtype=np.float64)
        d.w_complex = 1 if W1.shape[0] == 2 else 0
        d.wl = W1.shape[1]
        if W2 is not None:
            W2 = np.ascontiguousarray(W2, dtype=np.float64)
            if W2.shape[0] != W1.shape[0]:
                raise ValueError("heff_desc: W1/W2 plane mismatch")
            d.wm = W1.shape[2]
            d.wr = W2.shape[2]
            d.W2 = _hptr(W2)
        else:
            d.wm = 0
            d.wr = W1.shape[2]
            d.W2 = None
        d.W1 = _hptr(W1)
        d.L = _ptr(L)
        d.R = _ptr(R)
        d.ldL, d.ldR = int(L.shape[-1]), int(R.shape[-1])   # row pitch of the last index
        d.nproj, d.proj, d.proj_weight = 0, None, 0.0
        d._hold = (W1, W2, L, R)
        return d

    @staticmethod
    def set_projectors(desc, P, weight):
        """Penalty projectors for an H_eff descriptor: P float64 device
        [nproj, np, n] (the layout of the vectors the operator acts on); the
        apply then adds weight * sum_k p_k <p_k|x>."""
        if P is None or P.shape[0] == 0:
            desc.nproj, desc.proj, desc.proj_weight = 0, None, 0.0
            return desc
        desc.nproj = int(P.shape[0])
        desc.proj = _ptr(P)
        desc.proj_weight = float(weight)
        desc._hold = desc._hold + (P,)
        return desc

    # -- entry points -----------------------------------------------------
    def heff_apply(self, desc, x, y):
        check(self.lib.dmrg_heff_apply(self.ctx, ctypes.byref(desc), _ptr(x), _ptr(y)), "dmrg_heff_apply")

    def heff_apply_noalloc(self, desc, x, y):
        """heff_apply through the allocation-free entry point (workspace in
        the library's static scratch array; raises for operators too large
        for it)."""
        check(self.lib.dmrg_heff_apply_noalloc(self.ctx, ctypes.byref(desc), _ptr(x), _ptr(y)),
              "dmrg_heff_apply_noalloc")

    def rayleigh(self, desc, x):
        v = ctypes.c_double()
        check(self.lib.dmrg_rayleigh(self.ctx, ctypes.byref(desc), _ptr(x), ctypes.byref(v)), "dmrg_rayleigh")
        return v.value

    def lanczos(self, desc, x, maxiter=200, tol=1e-12, res_tol=0.0):
        e = ctypes.c_double()
        it = ctypes.c_int()
        check(self.lib.dmrg_lanczos_ground(self.ctx, ctypes.byref(desc), _ptr(x), maxiter, tol, res_tol,
                                           ctypes.byref(e), ctypes.byref(it)), "dmrg_lanczos_ground")
        return e.value, it.value

    def expm(self, desc, x, coeff, maxiter=50, tol=1e-12, errgoal=1e-10):
        it = ctypes.c_int()
        c = complex(coeff)
        check(self.lib.dmrg_krylov_expm(self.ctx, ctypes.byref(desc), _ptr(x), c.real, c.imag, maxiter, tol,
                                        errgoal, ctypes.byref(it)), "dmrg_krylov_expm")
        return it.value

    def env_update(self, direction, field, prec, ket, bra, W, env_in, env_out):
        """ket/bra: [np, Dl, d, Dr] float64; W host [np, wl, wr, d, d]."""
        e = EnvDesc()
        e.dir = 0 if direction == "left" else 1
        e.field = field
        e.prec = self.contraction_prec(prec)
        _, e.Dl, e.d, e.Dr = ket.shape
        _, e.Dlb, db, e.Drb = bra.shape
        if db != e.d:
            raise ValueError("env_update: physical dims differ")
        W = np.ascontiguousarray(W, dtype=np.float64)
        e.w_complex = 1 if W.shape[0] == 2 else 0
        e.wl, e.wr = W.shape[1], W.shape[2]
        e.ket, e.bra, e.W = _ptr(ket), _ptr(bra), _hptr(W)
        e.env_in, e.env_out = _ptr(env_in), _ptr(env_out)
        e.ld_in, e.ld_out = int(env_in.shape[-1]), int(env_out.shape[-1])
        check(self.lib.dmrg_env_update(self.ctx, ctypes.byref(e)), "dmrg_env_update")

    def svd_split(self, field, M, m, n, cutoff, maxdim, mindim, absorb, guide=None, guide_side=0,
                  sketch=None):
        """M: float64 device [np, m, n].  Returns (U [np,m,keep], V [np,keep,n],
        keep, discarded, singular values (numpy, descending)).

        guide: an isometry whose span should hold most of M -- guide_side 0:
        [np, k0, n] with orthonormal rows, 1: [np, m, k0] with orthonormal
        columns -- for the subspace split of dmrg_svd_split_guided;
        self.last_svd_guided tells whether that path produced the result."""
        npl = 2 if field == 1 else 1
        kcap = min(m, n) if not maxdim else min(m, n, maxdim)
        kcap = max(kcap, max(1, mindim))
        Ub = self.empty((npl * m * kcap,))
        Vb = self.empty((npl * kcap * n,))
        keep = ctypes.c_int()
        disc = ctypes.c_double()
        sv = np.zeros(min(m, n))
        svp = sv.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        if guide is None:
            check(self.lib.dmrg_svd_split(self.ctx, field, _ptr(M), m, n, cutoff, int(maxdim or 0), mindim, absorb,
                                          _ptr(Ub), _ptr(Vb), kcap, ctypes.byref(keep), ctypes.byref(disc), svp),
                  "dmrg_svd_split")
            self.last_svd_guided = 0
            self.last_svd_resid = -1.0
        else:
            k0 = guide.numel() // (npl * (n if guide_side == 0 else m))
            if guide.dtype != self.torch.float64 or not guide.is_contiguous() or \
                    guide.numel() != npl * k0 * (n if guide_side == 0 else m):
                raise ValueError("svd_split: guide must be a contiguous float64 [np, k0, n] or [np, m, k0] tensor")
            used = ctypes.c_int()
            resid = ctypes.c_double()
            sketch = self.svd_sketch if sketch is None else sketch
            check(self.lib.dmrg_svd_split_guided(self.ctx, field, _ptr(M), m, n, cutoff, int(maxdim or 0), mindim,
                                                 absorb, _ptr(guide), k0, guide_side, sketch, _ptr(Ub), _ptr(Vb),
                                                 kcap, ctypes.byref(keep), ctypes.byref(disc), svp,
                                                 ctypes.byref(used), ctypes.byref(resid)), "dmrg_svd_split_guided")
            self.last_svd_guided = used.value
            self.last_svd_resid = resid.value
        k = keep.value
        U = Ub[:npl * m * k].view(npl, m, k).clone()
        V = Vb[:npl * k * n].view(npl, k, n).clone()
        return U, V, k, disc.value, sv

    def set_eigensolver(self, mode):
        """Eigensolver of the Gram-route splits.  1 (default): own
        tridiagonalisation (real or Hermitian) + rocSOLVER stedc + the
        library's blocked back-transformation of the kept eigenvectors, the
        reduction stopping once its trailing block is below 2e-15 of
        the Gram matrix in Frobenius norm (orders above 96); 3: as 1 with
        the complete reduction; 2: as 3 with rocSOLVER ormtr/unmtr for the
        back-transformation; 0: rocSOLVER syevd/heevd."""
        check(self.lib.dmrg_set_eigensolver(self.ctx, int(mode)), "dmrg_set_eigensolver")

    def herm_eig(self, A, mode=1):
        """Eigenpairs of a Hermitian device matrix given as planar float64
        [2, n, n] (copied): (w ascending [n], Z [2, n, n] planar with row k
        the eigenvector of w[k])."""
        n = A.shape[-1]
        Ac = A.reshape(2, n, n).clone()
        w = self.empty((n,))
        Z = self.empty((2, n, n))
        check(self.lib.dmrg_herm_eig(self.ctx, int(mode), _ptr(Ac), n, _ptr(w), _ptr(Z)), "dmrg_herm_eig")
        return w, Z

    def sym_eig(self, A, mode=1):
        """Eigenpairs of a symmetric float64 device matrix [n, n] (copied):
        (w ascending [n], Z [n, n] with row k the eigenvector of w[k])."""
        n = A.shape[-1]
        Ac = A.reshape(n, n).clone()
        w = self.empty((n,))
        Z = self.empty((n, n))
        check(self.lib.dmrg_sym_eig(self.ctx, int(mode), _ptr(Ac), n, _ptr(w), _ptr(Z)), "dmrg_sym_eig")
        return w, Z

    def mpo_grad(self, field, prec, L, bra, ket, R, wl, wr):
        """dE/dW of one site: host [np, wl, wr, d, d] float64."""
        _, Dl, d, Dr = ket.shape
        _, Dlb, _, Drb = bra.shape
        npl = 2 if field == 1 else 1
        G = np.zeros((npl, wl, wr, d, d))
        check(self.lib.dmrg_mpo_grad(self.ctx, field, prec, _ptr(L), _ptr(bra), _ptr(ket), _ptr(R), Dl, Dr, Dlb,
                                     Drb, d, wl, wr, int(L.shape[-1]), int(R.shape[-1]), _hptr(G)), "dmrg_mpo_grad")
        return G

    def gemm(self, prec, field, A, B, C, M, N, K, transA=False, transB=False, conjA=False, conjB=False,
             alpha=1.0, beta=0.0, lda=None, ldb=None, ldc=None, strideA=0, strideB=0, strideC=0,
             planeA=0, planeB=0, planeC=0, batch=1):
        lda = lda if lda is not None else (M if transA else K)
        ldb = ldb if ldb is not None else (K if transB else N)
        ldc = ldc if ldc is not None else N
        check(self.lib.dmrg_gemm(self.ctx, prec, field, int(transA), int(transB), int(conjA), int(conjB), M, N, K,
                                 alpha, _ptr(A), lda, strideA, planeA, _ptr(B), ldb, strideB, planeB, beta,
                                 _ptr(C), ldc, strideC, planeC, batch), "dmrg_gemm")

    def prof_enable(self, on=True):
        check(self.lib.dmrg_prof_enable(self.ctx, int(on)), "dmrg_prof_enable")

    def prof_read(self):
        """(time_ms, flops, calls) per slot: H_eff, env, SVD, Krylov,
        tridiagonalisation."""
        t = (ctypes.c_double * 5)()
        f = (ctypes.c_double * 5)()
        n = (ctypes.c_longlong * 5)()
        check(self.lib.dmrg_prof_read(self.ctx, t, f, n), "dmrg_prof_read")
        return list(t), list(f), list(n)

    def prof_read_bytes(self):
        b = (ctypes.c_double * 5)()
        check(self.lib.dmrg_prof_read_bytes(self.ctx, b), "dmrg_prof_read_bytes")
        return list(b)

    def matvec_count(self):
        return int(self.lib.dmrg_matvec_count(self.ctx))


_ENGINES = {}


def get_engine(device=0):
    """One engine (context) per device per process."""
    eng = _ENGINES.get(device)
    if eng is None:
        eng = Engine(device)
        _ENGINES[device] = eng
    return eng
