/* C ABI of the MI355X (gfx950) two-site DMRG / TDVP engine, libdmrg_hip.so.
 *
 * Every entry point takes plain pointers and sizes, runs on the HIP stream
 * the context was created with, and returns 0 on success or a negative
 * status; dmrg_last_error() then holds a message.  Nothing here aborts the
 * process.  Device pointers are HIP device memory; "host" pointers are
 * ordinary host memory read during the call.
 *
 * Storage conventions (row-major):
 *   field 0 = real, 1 = complex stored as two planes (real plane then
 *   imaginary plane, each of the tensor's full size);
 *   prec  0 = FP64 contractions, 1 = FP32 MFMA contractions, 2 = split-BF16x3
 *   MFMA contractions (FP32 accumulation); Krylov vectors, MPS tensors and
 *   SVD operands are always FP64, environments are stored in the
 *   contraction precision (double for prec 0, float otherwise).
 *   MPS site A[Dl][d][Dr]; MPO site W[wl][wr][s_out][s_in];
 *   left environment L[w][D_bra][D_ket]; right environment R[w][D_ket][D_bra];
 *   two-site vector theta[Dl][d1][d2][Dr]; one-site vector C[Dl][d][Dr].
 *
 * The reference computes all of this in the pure-Python backend
 * src/dmrgpy/pyitensor/ (complex128 NumPy); each entry point names the
 * reference function it replaces.
 */
#ifndef DMRG_HIP_H
#define DMRG_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct dmrg_ctx dmrg_ctx;

/* Effective Hamiltonian of one bond (nsite 2) or one site (nsite 1). */
typedef struct dmrg_heff_desc {
  int nsite;           /* 1 or 2 */
  int field;           /* 0 real, 1 complex (planar) */
  int prec;            /* 0 FP64, 1 FP32, 2 BF16x3 */
  int Dl, Dr;          /* outer link dims of the vector */
  int d1, d2;          /* physical dims (d2 unused for nsite 1) */
  int wl, wm, wr;      /* MPO link dims (wm unused for nsite 1) */
  const void* L;       /* device, [wl][Dl][Dl] (bra, ket), contraction precision */
  const void* R;       /* device, [wr][Dr][Dr] (ket, bra), contraction precision */
  const double* W1;    /* host, nsite 2: [wl][wm][d1][d1]; nsite 1: [wl][wr][d1][d1] */
  const double* W2;    /* host, [wm][wr][d2][d2] (nsite 2 only) */
  int w_complex;       /* W1/W2 carry an imaginary plane after the real one */
  int ldL, ldR;        /* row pitch (elements) of the last index of L and R; 0 = Dl / Dr.  A pitch
                          that is a multiple of 4 lets every H_eff GEMM use 16-byte loads */
  int nproj;           /* penalty projectors (excited states, src/dmrgpy/pyitensor/dmrg.py:562-599): */
  const double* proj;  /* device FP64 [nproj][np][n] vectors p_k in the layout of x; the apply adds */
  double proj_weight;  /* proj_weight * sum_k p_k <p_k|x>; nproj = 0 for the plain H_eff */
} dmrg_heff_desc;

/* One environment extension <bra|W|ket> by one site. */
typedef struct dmrg_env_desc {
  int dir;             /* 0: left env grows rightward, 1: right env grows leftward */
  int field, prec;
  int Dl, Dr;          /* ket site link dims */
  int Dlb, Drb;        /* bra site link dims */
  int d;               /* physical dim */
  int wl, wr;          /* MPO link dims of the site */
  const double* ket;   /* device FP64 [Dl][d][Dr] */
  const double* bra;   /* device FP64 [Dlb][d][Drb]; may alias ket */
  const double* W;     /* host FP64 [wl][wr][d][d] (+ imaginary plane) */
  int w_complex;
  const void* env_in;  /* device: dir 0 [wl][Dlb][Dl]; dir 1 [wr][Dr][Drb] */
  void* env_out;       /* device: dir 0 [wr][Drb][Dr]; dir 1 [wl][Dl][Dlb] */
  int ld_in, ld_out;   /* row pitch of the last index of env_in / env_out; 0 = dense */
} dmrg_env_desc;

/* Library version (major*100 + minor). */
int dmrg_version(void);
/* Message of the last failed call on this thread. */
const char* dmrg_last_error(void);

/* Create a context on `device` issuing work on `stream` (a hipStream_t, may
 * be NULL for the null stream).  The context owns its workspace. */
int dmrg_ctx_create(int device, void* stream, dmrg_ctx** out);
int dmrg_ctx_destroy(dmrg_ctx* ctx);
/* Wait for all work issued through the context. */
int dmrg_sync(dmrg_ctx* ctx);

/* y = H_eff x.  Replaces the matvec built by two_site_heff / one_site_heff
 * (src/dmrgpy/pyitensor/dmrg.py:224-275) through kernels.make_matvec
 * (src/dmrgpy/pyitensor/kernels.py:298-445).  x, y: device FP64, length
 * Dl*d1*d2*Dr (nsite 2) or Dl*d1*Dr (nsite 1) per plane. */
int dmrg_heff_apply(dmrg_ctx* ctx, const dmrg_heff_desc* op, const double* x, double* y);

/* dmrg_heff_apply that allocates nothing: no counterpart in the reference;
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops).  Same arguments and result as
 * dmrg_heff_apply, but every intermediate (the uploaded MPO pair, theta in
 * the contraction precision, T1, T2, y before conversion, penalty-projector
 * coefficients) lives in one static 64 MiB device array per GPU that the
 * library maps when the context is created; nothing is allocated or freed
 * and no profiling events are recorded.  An operator whose workspace would
 * exceed 64 MiB is rejected with a status and a message naming the size
 * (a real two-site H_eff with d=2, w=5 fits up to about D=450 in FP64 and
 * D=610 in FP32).  All
 * calls on one GPU share that array and it keeps the last call's contents
 * (nothing in it needs to start at zero): calls that overlap on one GPU,
 * from several contexts, streams or threads, race. */
int dmrg_heff_apply_noalloc(dmrg_ctx* ctx, const dmrg_heff_desc* op, const double* x, double* y);

/* <x|H_eff|x> / <x|x> in the operator's precision. */
int dmrg_rayleigh(dmrg_ctx* ctx, const dmrg_heff_desc* op, const double* x, double* value);

/* Environment extension.  Replaces _extend_left / _extend_right
 * (src/dmrgpy/pyitensor/dmrg.py:170-201) and the per-site step of inner()
 * (src/dmrgpy/pyitensor/mpsalgebra.py:66-106). */
int dmrg_env_update(dmrg_ctx* ctx, const dmrg_env_desc* env);

/* Lowest eigenpair of H_eff by Lanczos with full re-orthogonalisation,
 * warm-started from x (overwritten with the normalised eigenvector); stops
 * when the lowest Ritz value moves by < tol*max(1,|E|) or after
 * min(maxiter, dim) steps; dims <= 3 are diagonalised densely.  res_tol > 0
 * adds a second stop, Ritz residual <= res_tol*max(1,|E|), for the
 * reduced-precision operators (0 keeps the reference rule alone).  Replaces
 * _lanczos_ground_state / _local_ground_state
 * (src/dmrgpy/pyitensor/dmrg.py:95-142, 302-333). */
int dmrg_lanczos_ground(dmrg_ctx* ctx, const dmrg_heff_desc* op, double* x, int maxiter, double tol,
                        double res_tol, double* energy, int* iters);

/* x <- expm((cre + i cim) * H_eff) x by a Lanczos basis with full
 * re-orthogonalisation and Saad's a-posteriori stop
 * beta0*beta_k*|e_k^T exp(c T_k) e_1| < errgoal.  Replaces
 * _lanczos_expm_multiply (src/dmrgpy/pyitensor/tdvp.py:82-172). */
int dmrg_krylov_expm(dmrg_ctx* ctx, const dmrg_heff_desc* op, double* x, double cre, double cim,
                     int maxiter, double tol, double errgoal, int* iters);

/* Truncated SVD of the m x n FP64 matrix M (planar if field 1): keeps the
 * count given by the cutoff / maxdim / mindim rule of _truncate
 * (src/dmrgpy/pyitensor/svd.py:32-68), Gram route with the exact-SVD
 * fallback of _svd_truncated (svd.py:108-151).  absorb 0: U left-orthonormal,
 * V = S*Vh (a right-moving step); absorb 1: U*S, V = Vh (left-moving);
 * absorb 2: U, Vh with S only returned.  U is written [m][keep], V
 * [keep][n] (planar), both dense; kcap bounds keep (buffers must hold
 * m*kcap and kcap*n elements per plane).  keep <= min(m, n): an all-zero M
 * with mindim above min(m, n) keeps min(m, n), the columns the reference's
 * slicing leaves.  svals_host (length min(m,n), may
 * be NULL) receives all singular values, descending.  Replaces svd()
 * (svd.py:154-195) as called by _apply_local_update
 * (src/dmrgpy/pyitensor/dmrg.py:336-352). */
int dmrg_svd_split(dmrg_ctx* ctx, int field, const double* M, int m, int n, double cutoff, int maxdim,
                   int mindim, int absorb, double* U, double* V, int kcap, int* keep,
                   double* discarded, double* svals_host);

/* dmrg_svd_split with a guide: G holds an isometry whose span should contain
 * most of M -- gside 0: k0 x n with orthonormal rows (in a right-moving
 * sweep, the tensor right of the bond), gside 1: m x k0 with orthonormal
 * columns (left-moving: the tensor left of the bond).  The split is then
 * computed in span(G) enlarged by a `sketch`-column random sketch of the
 * rest (an eigenproblem of order k0 + sketch instead of min(m, n)), and
 * accepted only when the weight of M outside that subspace is at most
 * 0.1 * cutoff * |M|^2, or, where maxdim forces the truncation, at most a
 * tenth of the weight maxdim discards; that weight counts as discarded in
 * the keep rule of _truncate (src/dmrgpy/pyitensor/svd.py:32-68).
 * Otherwise, or when k0 + sketch > 0.6 * min(m, n), it runs dmrg_svd_split.
 * *used = 1 when the guided path produced the result; *resid (may be NULL)
 * receives the measured weight outside the subspace over |M|^2 (-1 when the
 * guided path was not attempted).  The reference always
 * runs the full split (svd.py:108-151, called from dmrg.py:336-352); this
 * entry has no counterpart there and is what makes the truncation cost
 * scale with the kept dimension.  svals_host (length min(m, n)) receives
 * the k0 + sketch leading values, zeros after. */
int dmrg_svd_split_guided(dmrg_ctx* ctx, int field, const double* M, int m, int n, double cutoff, int maxdim,
                          int mindim, int absorb, const double* G, int k0, int gside, int sketch, double* U,
                          double* V, int kcap, int* keep, double* discarded, double* svals_host, int* used,
                          double* resid);

/* Dense eigensolver of the Gram route (the np.linalg.eigh call of
 * _svd_truncated, src/dmrgpy/pyitensor/svd.py:134): mode 1 (the default)
 * reduces the real symmetric or complex Hermitian Gram matrix to a real
 * tridiagonal with this library's two-launch-per-column Householder
 * kernels, runs rocSOLVER's divide and conquer on the tridiagonal and
 * back-transforms only the eigenvectors the truncation keeps with this
 * library's blocked compact-WY GEMMs.  For a real Gram matrix of order
 * above 96, mode 1 stops the reduction after the first full panel (at or
 * past column 64) whose trailing block has Frobenius norm at most 2e-15
 * of the Gram matrix's: that block's eigenvalues are then taken as zero
 * (they lie below the FP64 rounding of forming the Gram matrix), its trace
 * is counted as discarded weight, and the divide and conquer runs on the
 * reduced order; a split that would keep more values than the reduced
 * order falls back to the full SVD.  Mode 3
 * is mode 1 with the complete reduction; mode 2 is mode 3 with rocSOLVER's
 * dormtr / zunmtr for the back-transformation; mode 0 calls rocSOLVER
 * dsyevd / zheevd. */
int dmrg_set_eigensolver(dmrg_ctx* ctx, int mode);

/* All eigenpairs of the real symmetric n x n matrix A (device, row-major;
 * overwritten): w ascending, row k of Z the eigenvector of w[k].  mode as
 * in dmrg_set_eigensolver.  The eigensolver of dmrg_svd_split exposed for
 * its parity tests; replaces np.linalg.eigh (svd.py:134). */
int dmrg_sym_eig(dmrg_ctx* ctx, int mode, double* A, int n, double* w, double* Z);

/* All eigenpairs of the Hermitian n x n matrix A (device, row-major, planar
 * complex: real plane then imaginary plane, plane stride n*n; overwritten):
 * w ascending, row k of Z (planar, same layout) the eigenvector of w[k].
 * mode 1 (the split's default): this library's reduction to a real
 * tridiagonal (zhetrd conventions), rocSOLVER dstedc and this library's
 * blocked back-transformation; mode 2: the same with rocSOLVER zunmtr for
 * the back-transformation; mode 0: rocSOLVER zheevd.  The
 * eigensolver of the complex dmrg_svd_split exposed for its parity tests;
 * replaces np.linalg.eigh on the complex Gram matrix (svd.py:134). */
int dmrg_herm_eig(dmrg_ctx* ctx, int mode, double* A, int n, double* w, double* Z);

/* Gradient of <bra|H|ket> with respect to one MPO site tensor W[wl][wr][s'][s]:
 * G = sum L[wl][a'][a] conj(bra[a'][s'][b']) ket[a][s][b] R[wr][b][b'], written
 * to host memory as G[np][wl][wr][d][d] (FP64).  L [wl][Dlb][Dl] and
 * R [wr][Dr][Drb] are environments in the contraction precision, with row
 * pitches ldL >= Dl and ldR >= Drb of their last index (0 = dense).  Has no
 * counterpart in the reference (which has no gradients); with a normalised
 * ground state it gives dE/dW (Hellmann-Feynman) for every coupling. */
int dmrg_mpo_grad(dmrg_ctx* ctx, int field, int prec, const void* L, const double* bra, const double* ket,
                  const void* R, int Dl, int Dr, int Dlb, int Drb, int d, int wl, int wr, int ldL, int ldR,
                  double* G_host);

/* General batched GEMM in a given precision (used by the gradient
 * contractions and by the tests): C = alpha op(A) op(B) + beta C with the
 * conventions of the engine's kernels; A, B, C are FP64 device buffers for
 * prec 0 and FP32 otherwise. */
int dmrg_gemm(dmrg_ctx* ctx, int prec, int field, int transA, int transB, int conjA, int conjB, int M,
              int N, int K, double alpha, const void* A, int lda, long long strideA, long long planeA,
              const void* B, int ldb, long long strideB, long long planeB, double beta, void* C, int ldc,
              long long strideC, long long planeC, int batch);

/* Profiling of H_eff applies (slot 0), environment updates (1), SVDs (2),
 * Krylov vector work (3) and the tridiagonalisations of the Gram
 * eigensolver (4): HIP events around each call on the context stream.  The
 * read functions fill arrays of 5; bytes are algorithmic (H_eff: every
 * operand read once and every result written once; tridiagonalisation: the
 * lower triangle of the trailing matrix read once per column). */
int dmrg_prof_enable(dmrg_ctx* ctx, int on);
int dmrg_prof_read(dmrg_ctx* ctx, double* time_ms, double* flops, long long* calls);
/* Algorithmic HBM bytes of the profiled calls per slot (H_eff: every operand
 * read once and every intermediate written once; the roofline denominator
 * of the measured traffic).  No reference counterpart (measurement only). */
int dmrg_prof_read_bytes(dmrg_ctx* ctx, double* bytes);
/* Total H_eff applies issued through the context. */
long long dmrg_matvec_count(dmrg_ctx* ctx);

#ifdef __cplusplus
}
#endif

#endif /* DMRG_HIP_H */
