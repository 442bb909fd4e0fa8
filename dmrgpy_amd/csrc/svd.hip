// FP64 truncated SVD of a bond tensor, Gram-matrix route:
//   m <= n:  G = M M^dag (MFMA GEMM), eigen-decompose G, S = sqrt(eig),
//            E = top-keep eigenvectors (rows, descending), U = E^dag,
//            S Vh = E M  (so a right-moving split needs no scaling at all)
//   m >  n:  G = M^dag M, Vh = E, U S = M E^dag
// The dense symmetric eigen-decomposition is rocSOLVER's syevd/heevd.  The
// keep count follows the cutoff / maxdim / mindim rule of the reference's
// _truncate (src/dmrgpy/pyitensor/svd.py:32-68) on the host, and the same
// fallback to a full SVD (rocSOLVER gesvd) applies when the matrix is small
// (min(m,n) < 16) or the smallest kept singular value is below 1e-7 of the
// largest (svd.py:108-151).
// One deliberate difference: the reference divides by S to get the factor on
// the non-Gram side (svd.py:140-145), whose orthonormality then degrades as
// eps * (s_0 / s)^2.  Where that factor is the one a sweep leaves behind as
// an isometry, it is re-orthonormalised here (lowdin below), which keeps
// H_eff of later bonds a true projection of H.
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>

#include "dmrg_internal.h"
#include "keep_rule.h"

namespace dmrg {

static void check_rb(rocblas_status s, const char* what) {
  if (s != rocblas_status_success) throw Error(std::string(what) + ": rocblas status " + std::to_string((int)s));
}

// E[r][j] = G[(top - r)][j] (row-major rows of the eigenvector buffer taken
// in descending eigenvalue order), planar output.  interleaved input when CX.
template <bool CX>
__global__ void take_rows_k(const double* __restrict__ G, int ld, int top, int keep, int cols,
                            double* __restrict__ E, long long planeE) {
  long long idx = (long long)blockIdx.x * 256 + threadIdx.x;
  const long long tot = (long long)keep * cols;
  for (; idx < tot; idx += (long long)gridDim.x * 256) {
    const int r = (int)(idx / cols), j = (int)(idx % cols);
    const long long src = (long long)(top - r) * ld + j;
    if (CX) {
      E[idx] = G[2 * src];
      E[idx + planeE] = G[2 * src + 1];
    } else {
      E[idx] = G[src];
    }
  }
}

// dst[j][r] = conj(src[r][j]) * (scale ? s[r] : 1)   (src rows x cols, planar)
template <bool CX>
__global__ void conj_transpose_k(const double* __restrict__ src, int rows, int cols, long long psrc,
                                 double* __restrict__ dst, long long pdst, const double* __restrict__ s,
                                 int use_s) {
  long long idx = (long long)blockIdx.x * 256 + threadIdx.x;
  const long long tot = (long long)rows * cols;
  for (; idx < tot; idx += (long long)gridDim.x * 256) {
    const int j = (int)(idx / rows), r = (int)(idx % rows);  // dst[j][r]
    const double f = use_s ? s[r] : 1.0;
    dst[idx] = f * src[(long long)r * cols + j];
    if (CX) dst[idx + pdst] = -f * src[(long long)r * cols + j + psrc];
  }
}

// rows (rowscale) or columns (colscale) of a planar matrix multiplied by
// s[r]^pw (pw = +1 or -1)
template <bool CX>
__global__ void scale_k(double* __restrict__ X, int rows, int cols, long long plane,
                        const double* __restrict__ s, int by_row, int inverse) {
  long long idx = (long long)blockIdx.x * 256 + threadIdx.x;
  const long long tot = (long long)rows * cols;
  for (; idx < tot; idx += (long long)gridDim.x * 256) {
    const int r = (int)(idx / cols), j = (int)(idx % cols);
    double f = s[by_row ? r : j];
    if (inverse) f = 1.0 / f;
    X[idx] *= f;
    if (CX) X[idx + plane] *= f;
  }
}

// dst (interleaved complex) = src (real), n elements
__global__ void real_to_inter_k(const double* __restrict__ src, long long n, double* __restrict__ dst) {
  long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  for (; i < n; i += (long long)gridDim.x * 256) {
    dst[2 * i] = src[i];
    dst[2 * i + 1] = 0.0;
  }
}

// dst (planar, plane stride n) = conj(src (interleaved)), n elements
__global__ void inter_to_planar_conj_k(const double* __restrict__ src, long long n, double* __restrict__ dst) {
  long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  for (; i < n; i += (long long)gridDim.x * 256) {
    dst[i] = src[2 * i];
    dst[i + n] = -src[2 * i + 1];
  }
}

__global__ void planar_to_inter_k(const double* __restrict__ src, long long n, double* __restrict__ dst) {
  long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  for (; i < n; i += (long long)gridDim.x * 256) {
    dst[2 * i] = src[i];
    dst[2 * i + 1] = src[i + n];
  }
}

// dst[r][c] (planar, ld dcols) = src[r*lds + c] (interleaved if CX), r<rows, c<cols
template <bool CX>
__global__ void copy_block_k(const double* __restrict__ src, int lds, int rows, int cols, double* __restrict__ dst,
                             long long pdst) {
  long long idx = (long long)blockIdx.x * 256 + threadIdx.x;
  const long long tot = (long long)rows * cols;
  for (; idx < tot; idx += (long long)gridDim.x * 256) {
    const int r = (int)(idx / cols), c = (int)(idx % cols);
    const long long s = (long long)r * lds + c;
    if (CX) {
      dst[idx] = src[2 * s];
      dst[idx + pdst] = src[2 * s + 1];
    } else {
      dst[idx] = src[s];
    }
  }
}

// dst[r][c] = src[r][c] for c < m, 0 for m <= c < g (keep rows): the kept
// eigenvectors of a leading order-m tridiagonal block as rows of length g
__global__ void embed_rows_k(const double* __restrict__ src, int m, int keep, int g, double* __restrict__ dst) {
  const long long tot = (long long)keep * g;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < tot; i += (long long)gridDim.x * blockDim.x) {
    const int r = (int)(i / g), cc = (int)(i % g);
    dst[i] = cc < m ? src[(size_t)r * m + cc] : 0.0;
  }
}

// Two-pass max reductions (the maximum is order independent, so the result
// is deterministic): block maxima into part[], then one block over part[].
// mode 0: max |x_i| over n elements; mode 1: max |T - I| over a k x k planar
// matrix (n = k * k, imaginary plane at +plane when cx).
__global__ __launch_bounds__(256) void max_partial_k(const double* __restrict__ x, long long n, int mode, int k,
                                                     long long plane, int cx, double* __restrict__ part) {
  __shared__ double red[256];
  double m = 0.0;
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
    if (mode == 0) {
      m = fmax(m, fabs(x[i]));
    } else {
      m = fmax(m, fabs(x[i] - ((i / k == i % k) ? 1.0 : 0.0)));
      if (cx) m = fmax(m, fabs(x[i + plane]));
    }
  }
  red[threadIdx.x] = m;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(256) void max_final_k(const double* __restrict__ part, int nb, double* __restrict__ out) {
  __shared__ double red[256];
  double m = 0.0;
  for (int i = threadIdx.x; i < nb; i += 256) m = fmax(m, part[i]);
  red[threadIdx.x] = m;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = red[0];
}

static void dev_max(Ctx* c, const double* x, long long n, int mode, int k, long long plane, int cx, double* out) {
  long long nb = (n + 4095) / 4096;
  if (nb > 1024) nb = 1024;
  if (nb < 1) nb = 1;
  double* part = (double*)ws_get(c, WS_SVD_RED, sizeof(double) * 1024);
  max_partial_k<<<(unsigned)nb, 256, 0, c->stream>>>(x, n, mode, k, plane, cx, part);
  DMRG_CHECK_HIP(hipGetLastError());
  max_final_k<<<1, 256, 0, c->stream>>>(part, (int)nb, out);
  DMRG_CHECK_HIP(hipGetLastError());
}

__global__ void scaled_copy_k(const double* __restrict__ x, long long n, const double* __restrict__ amax,
                              double* __restrict__ y) {
  const double f = amax[0] > 0.0 ? 1.0 / amax[0] : 1.0;
  long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  for (; i < n; i += (long long)gridDim.x * 256) y[i] = f * x[i];
}

static unsigned g1(long long n) {
  long long b = (n + 255) / 256;
  if (b > 4096) b = 4096;
  if (b < 1) b = 1;
  return (unsigned)b;
}

// T = s * I (k x k planar; the imaginary plane, if any, is zeroed)
__global__ void set_identity_k(double* __restrict__ T, int k, long long plane, int cx, double s) {
  long long idx = (long long)blockIdx.x * 256 + threadIdx.x;
  const long long tot = (long long)k * k;
  for (; idx < tot; idx += (long long)gridDim.x * 256) {
    T[idx] = (idx / k == idx % k) ? s : 0.0;
    if (cx) T[idx + plane] = 0.0;
  }
}

// Loewdin re-orthonormalisation of the factor a Gram-route split derives by
// dividing by the singular values.  In exact arithmetic its rows (rows:
// X is keep x len) or columns (X is len x keep) are orthonormal; in floating
// point they carry errors of order eps * (s_0 / s)^2, which for singular
// values near the 1e-7 s_0 guard reaches 1e-2 and breaks the isometry the
// environments of the next bonds rely on.  With O = X X^dag (X^dag X) = I + E,
// the coupled Newton-Schulz iteration  T = (3I - Z Y) / 2,  Y <- Y T,
// Z <- T Z  from Y = O, Z = I converges quadratically to Y = O^{1/2},
// Z = O^{-1/2}.  X becomes Z X (X Z), the closest matrix with orthonormal
// rows (columns), and the partner factor P becomes P Y (Y P), so the
// product of the two factors is unchanged (P may be null).  Returns false,
// with nothing modified, when the iteration does not converge.
static bool lowdin(Ctx* c, int field, double* X, int keep, int len, bool rows, double* P, int plen) {
  const bool cx = field == CPLX;
  const size_t es = cx ? 2 : 1;
  const long long kk = (long long)keep * keep;
  const long long pX = (long long)keep * len, pP = (long long)keep * plen;
  double* buf = (double*)ws_get(c, WS_SVD_FIX, sizeof(double) * (es * 5 * (size_t)kk + 16));
  double* Y = buf;
  double* Z = Y + es * kk;
  double* T = Z + es * kk;
  double* Yn = T + es * kk;
  double* Zn = Yn + es * kk;
  double* dev = Zn + es * kk;
  auto sq = [&](const double* A, const double* B, double* C, double alpha, double beta) {
    GemmDesc d = gdesc(keep, keep, keep, keep, keep, keep, field, 4);
    d.pA = d.pB = d.pC = kk;
    d.alpha = alpha;
    d.beta = beta;
    gemm_f64(c, d, A, B, C);
  };
  {  // Y = O
    GemmDesc d = gdesc(keep, keep, len, rows ? len : keep, rows ? len : keep, keep, field, 4);
    if (rows) { d.transB = true; d.conjB = true; } else { d.transA = true; d.conjA = true; }
    d.pA = d.pB = pX;
    d.pC = kk;
    gemm_f64(c, d, X, X, Y);
  }
  set_identity_k<<<g1(kk), 256, 0, c->stream>>>(Z, keep, kk, cx, 1.0);
  DMRG_CHECK_HIP(hipGetLastError());
  double prev = 1e300;
  bool ok = false;
  for (int it = 0; it < 16; ++it) {
    set_identity_k<<<g1(kk), 256, 0, c->stream>>>(T, keep, kk, cx, 1.5);
    DMRG_CHECK_HIP(hipGetLastError());
    sq(Z, Y, T, -0.5, 1.0);
    dev_max(c, T, kk, 1, keep, kk, cx, dev);
    double h = 0.0;
    DMRG_CHECK_HIP(hipMemcpyAsync(&h, dev, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    DMRG_CHECK_HIP(hipStreamSynchronize(c->stream));
    if (!(h < 0.25 || (it > 0 && h < prev))) break;   // |E| too large, or diverging
    if (h < 1e-15 || (h < 1e-12 && h > 0.5 * prev)) {  // converged (or at the rounding floor)
      ok = true;
      break;
    }
    prev = h;
    sq(Y, T, Yn, 1.0, 0.0);
    sq(T, Z, Zn, 1.0, 0.0);
    std::swap(Y, Yn);
    std::swap(Z, Zn);
  }
  if (!ok) return false;
  const long long big = std::max((long long)len, (long long)(P ? plen : 0)) * keep;
  double* tmp = (double*)ws_get(c, WS_SVD_GRAM, sizeof(double) * es * (size_t)big);
  if (rows) {  // X <- Z X (keep x len), P <- P Y (plen x keep)
    GemmDesc d = gdesc(keep, len, keep, keep, len, len, field, 4);
    d.pA = kk; d.pB = pX; d.pC = pX;
    gemm_f64(c, d, Z, X, tmp);
    DMRG_CHECK_HIP(hipMemcpyAsync(X, tmp, sizeof(double) * es * pX, hipMemcpyDeviceToDevice, c->stream));
    if (!P) return true;
    GemmDesc e = gdesc(plen, keep, keep, keep, keep, keep, field, 4);
    e.pA = pP; e.pB = kk; e.pC = pP;
    gemm_f64(c, e, P, Y, tmp);
    DMRG_CHECK_HIP(hipMemcpyAsync(P, tmp, sizeof(double) * es * pP, hipMemcpyDeviceToDevice, c->stream));
  } else {     // X <- X Z (len x keep), P <- Y P (keep x plen)
    GemmDesc d = gdesc(len, keep, keep, keep, keep, keep, field, 4);
    d.pA = pX; d.pB = kk; d.pC = pX;
    gemm_f64(c, d, X, Z, tmp);
    DMRG_CHECK_HIP(hipMemcpyAsync(X, tmp, sizeof(double) * es * pX, hipMemcpyDeviceToDevice, c->stream));
    if (!P) return true;
    GemmDesc e = gdesc(keep, plen, keep, keep, plen, plen, field, 4);
    e.pA = kk; e.pB = pP; e.pC = pP;
    gemm_f64(c, e, Y, P, tmp);
    DMRG_CHECK_HIP(hipMemcpyAsync(P, tmp, sizeof(double) * es * pP, hipMemcpyDeviceToDevice, c->stream));
  }
  return true;
}

// The keep rule (truncate_rule) is in keep_rule.h.

struct SvdOut {
  int keep;
  double discarded;
  double resid = -1.0;   // guided split: weight outside the subspace / |M|^2 (-1: not computed)
};

SvdOut svd_split(Ctx* c, int field, const double* M, int m, int n, double cutoff, int maxdim, int mindim,
                 int absorb, double* U, double* V, int kcap, std::vector<double>& svals);

static SvdOut svd_full(Ctx* c, int field, const double* M, int m, int n, double cutoff, int maxdim, int mindim,
                       int absorb, double* U, double* V, int kcap, std::vector<double>& svals) {
  const bool cx = field == CPLX;
  const int k = std::min(m, n);
  const long long mn = (long long)m * n;
  const size_t es = cx ? 2 : 1;
  // rocSOLVER is column-major: the row-major M is the column-major M^T (n x m).
  double* A = (double*)ws_get(c, WS_SVD_TMP, sizeof(double) * es * mn);
  double* Ux = (double*)ws_get(c, WS_SVD_GRAM, sizeof(double) * es * ((size_t)n * k + (size_t)k * m + 2 * k + 16));
  double* Vx = Ux + es * (size_t)n * k;
  double* S = Vx + es * (size_t)k * m;
  double* Ework = S + k;
  rocblas_int* info = (rocblas_int*)ws_get(c, WS_SVD_INFO, 64);
  // scale by max|M| so the factorisation never sees overflowing norms
  // (un-normalised MPS tensors reach 1e100 during canonicalisation)
  double* amax = (double*)ws_get(c, WS_SVD_EV, sizeof(double) * (2 * (size_t)k + 16));
  dev_max(c, M, (long long)es * mn, 0, 0, 0, 0, amax);
  if (cx) {
    double* Ms = (double*)ws_get(c, WS_SVD_TMP2, sizeof(double) * es * mn);
    scaled_copy_k<<<g1(es * mn), 256, 0, c->stream>>>(M, es * mn, amax, Ms);
    planar_to_inter_k<<<g1(mn), 256, 0, c->stream>>>(Ms, mn, A);
    // M^T as column-major of the interleaved row-major M is exact (no conj)
    check_rb(rocsolver_zgesvd(c->blas, rocblas_svect_singular, rocblas_svect_singular, n, m,
                              (rocblas_double_complex*)A, n, S, (rocblas_double_complex*)Ux, n,
                              (rocblas_double_complex*)Vx, k, Ework, rocblas_outofplace, info),
             "zgesvd");
  } else {
    scaled_copy_k<<<g1(mn), 256, 0, c->stream>>>(M, mn, amax, A);
    check_rb(rocsolver_dgesvd(c->blas, rocblas_svect_singular, rocblas_svect_singular, n, m, A, n, S, Ux, n, Vx, k,
                              Ework, rocblas_outofplace, info),
             "dgesvd");
  }
  svals.resize(k);
  int hinfo = 0;
  double hmax = 0.0;
  DMRG_CHECK_HIP(hipMemcpyAsync(svals.data(), S, sizeof(double) * k, hipMemcpyDeviceToHost, c->stream));
  DMRG_CHECK_HIP(hipMemcpyAsync(&hinfo, info, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  DMRG_CHECK_HIP(hipMemcpyAsync(&hmax, amax, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  DMRG_CHECK_HIP(hipStreamSynchronize(c->stream));
  DMRG_REQUIRE(hinfo == 0, "gesvd did not converge");
  const double scale = hmax > 0.0 ? hmax : 1.0;
  for (auto& s : svals) s *= scale;
  DMRG_CHECK_HIP(hipMemcpyAsync(S, svals.data(), sizeof(double) * k, hipMemcpyHostToDevice, c->stream));
  DMRG_CHECK_HIP(hipStreamSynchronize(c->stream));   // svals belongs to the caller: consume it now
  SvdOut o;
  // at most the k computed singular triplets: the rule keeps max(1, mindim)
  // of an all-zero spectrum, which the reference's slicing U[:, :keep]
  // (src/dmrgpy/pyitensor/svd.py:148-151) also cuts to k columns
  o.keep = std::min(truncate_rule(svals, cutoff, maxdim, mindim, &o.discarded, 0.0), k);
  DMRG_REQUIRE(o.keep <= kcap, "svd: keep exceeds output capacity");
  const int keep = o.keep;
  // U_M (row-major m x k) is the Vx buffer, Vh_M (k x n) is the Ux buffer.
  if (cx) {
    copy_block_k<true><<<g1((long long)m * keep), 256, 0, c->stream>>>(Vx, k, m, keep, U, (long long)m * keep);
    copy_block_k<true><<<g1((long long)keep * n), 256, 0, c->stream>>>(Ux, n, keep, n, V, (long long)keep * n);
  } else {
    copy_block_k<false><<<g1((long long)m * keep), 256, 0, c->stream>>>(Vx, k, m, keep, U, 0);
    copy_block_k<false><<<g1((long long)keep * n), 256, 0, c->stream>>>(Ux, n, keep, n, V, 0);
  }
  DMRG_CHECK_HIP(hipGetLastError());
  if (absorb == 0 || absorb == 1) {
    double* Sd = S;  // device singular values (first keep used)
    if (absorb == 0) {
      if (cx) scale_k<true><<<g1((long long)keep * n), 256, 0, c->stream>>>(V, keep, n, (long long)keep * n, Sd, 1, 0);
      else scale_k<false><<<g1((long long)keep * n), 256, 0, c->stream>>>(V, keep, n, 0, Sd, 1, 0);
    } else {
      if (cx) scale_k<true><<<g1((long long)m * keep), 256, 0, c->stream>>>(U, m, keep, (long long)m * keep, Sd, 0, 0);
      else scale_k<false><<<g1((long long)m * keep), 256, 0, c->stream>>>(U, m, keep, 0, Sd, 0, 0);
    }
    DMRG_CHECK_HIP(hipGetLastError());
  }
  return o;
}

// Gram route with the eigenproblem on the left (M M^dag, side 0) or right
// (M^dag M, side 1) index; side -1 picks the smaller one, as the reference
// does.  A forced side larger than min(m, n) keeps the min(m, n) largest
// eigenvalues (the rest are rounding noise of a rank-deficient Gram).
static SvdOut svd_gram(Ctx* c, int field, const double* M, int m, int n, double cutoff, int maxdim, int mind,
                       int absorb, double* U, double* V, int kcap, std::vector<double>& svals, int side,
                       double extra = 0.0) {
  const bool cx = field == CPLX;
  const bool left = side < 0 ? m <= n : side == 0;
  const int g = left ? m : n;
  const int rank = std::min(m, n);
  const long long mn = (long long)m * n;
  const size_t es = cx ? 2 : 1;
  double* ev = (double*)ws_get(c, WS_SVD_EV, sizeof(double) * (2 * (size_t)g + 16));
  double* ework = ev + g;
  double* amax_dev = ev + 2 * (size_t)g + 8;
  rocblas_int* info = (rocblas_int*)ws_get(c, WS_SVD_INFO, 64);
  double* Gsolve = nullptr;
  double *trd_Z = nullptr, *trd_tau = nullptr, *trd_refl = nullptr, *trd_Zs = nullptr;
  int trd_nrefl = g - 1, trd_m = g;   // reflectors computed, order of the tridiagonal solved
  double trd_dropped = 0.0;             // trace of the block a stopped reduction took as zero
  {
    ProfScope ps(c, 2, (cx ? 8.0 : 2.0) * (double)g * g * (left ? n : m));
    double* Gplanar = (double*)ws_get(c, WS_SVD_TMP2, sizeof(double) * es * (size_t)g * g);
    // Gram of M / max|M|: the eigensolver never sees overflowing norms
    // (un-normalised MPS tensors reach 1e100 during canonicalisation)
    double* Ms = (double*)ws_get(c, WS_SVD_TMP, sizeof(double) * es * mn);
    dev_max(c, M, (long long)es * mn, 0, 0, 0, 0, amax_dev);
    scaled_copy_k<<<g1(es * mn), 256, 0, c->stream>>>(M, es * mn, amax_dev, Ms);
    DMRG_CHECK_HIP(hipGetLastError());
    GemmDesc d = gdesc(g, g, left ? n : m, 0, 0, g, field, 4);
    if (left) {  // G = M M^dag
      d.lda = n; d.ldb = n; d.transB = true; d.conjB = true;
    } else {     // G = M^dag M
      d.lda = n; d.ldb = n; d.transA = true; d.conjA = true;
    }
    d.pA = mn; d.pB = mn; d.pC = (long long)g * g;
    d.lower = true;   // Hermitian: the lower triangle, then its mirror
    gemm_f64(c, d, Ms, Ms, Gplanar);
    mirror_lower(c, Gplanar, g, g, (long long)g * g, field);
    if (cx && c->eig_mode >= 1) {
      // complex Hermitian: own reduction to a real tridiagonal (stopping
      // early in mode 1 as the real one does), rocSOLVER's divide and
      // conquer on it; unmtr on the kept columns after keep
      trd_Z = (double*)ws_get(c, WS_TRD_Z, sizeof(double) * ((size_t)g * g + 2 * (size_t)g + 16));
      trd_tau = trd_Z + (size_t)g * g;
      trd_nrefl = hetrd_lower(c, Gplanar, g, ev, ework, trd_tau, c->eig_mode == 1 ? TRD_STOP_TOL : 0.0,
                              &trd_dropped);
      trd_m = trd_nrefl < g - 1 ? trd_nrefl + 1 : g;
      if (trd_m < g) {
        trd_Zs = (double*)ws_get(c, WS_TRD_ZS, sizeof(double) * (size_t)trd_m * trd_m);
        check_rb(rocsolver_dstedc(c->blas, rocblas_evect_tridiagonal, trd_m, ev, ework, trd_Zs, trd_m, info),
                 "dstedc");
      } else {
        check_rb(rocsolver_dstedc(c->blas, rocblas_evect_tridiagonal, g, ev, ework, trd_Z, g, info), "dstedc");
      }
      trd_refl = Gplanar;
    } else if (cx) {
      double* Gi = (double*)ws_get(c, WS_SVD_CPLX, sizeof(double) * 2 * (size_t)g * g);
      planar_to_inter_k<<<g1((long long)g * g), 256, 0, c->stream>>>(Gplanar, (long long)g * g, Gi);
      DMRG_CHECK_HIP(hipGetLastError());
      check_rb(rocsolver_zheevd(c->blas, rocblas_evect_original, rocblas_fill_upper, g, (rocblas_double_complex*)Gi,
                                g, ev, ework, info),
               "zheevd");
      Gsolve = Gi;
    } else if (c->eig_mode >= 1 && g >= 32) {
      // tridiagonalise here, eigenpairs of the tridiagonal by rocSOLVER's
      // divide and conquer; the back-transformation waits for keep.  In
      // mode 1 the reduction stops once the trailing block is below
      // TRD_STOP_TOL |G|_F (sytrd_lower), and divide and conquer then runs
      // on the leading order-(nrefl + 1) tridiagonal only; the rest of the
      // spectrum is exactly zero.
      trd_Z = (double*)ws_get(c, WS_TRD_Z, sizeof(double) * ((size_t)g * g + g + 16));
      trd_tau = trd_Z + (size_t)g * g;
      trd_nrefl = sytrd_lower(c, Gplanar, g, ev, ework, trd_tau, c->eig_mode == 1 ? TRD_STOP_TOL : 0.0,
                              &trd_dropped);
      trd_m = trd_nrefl < g - 1 ? trd_nrefl + 1 : g;
      if (trd_m < g) {
        trd_Zs = (double*)ws_get(c, WS_TRD_ZS, sizeof(double) * (size_t)trd_m * trd_m);
        check_rb(rocsolver_dstedc(c->blas, rocblas_evect_tridiagonal, trd_m, ev, ework, trd_Zs, trd_m, info),
                 "dstedc");
      } else {
        check_rb(rocsolver_dstedc(c->blas, rocblas_evect_tridiagonal, g, ev, ework, trd_Z, g, info), "dstedc");
      }
      Gsolve = trd_Z;
      trd_refl = Gplanar;
    } else {
      check_rb(rocsolver_dsyevd(c->blas, rocblas_evect_original, rocblas_fill_upper, g, Gplanar, g, ev, ework, info),
               "dsyevd");
      Gsolve = Gplanar;
    }
  }
  std::vector<double> evh(g, 0.0);
  int hinfo = 0;
  double hmax = 0.0;
  // ascending; a reduction that stopped early leaves g - trd_m exact zeros
  // below the trd_m computed values
  DMRG_CHECK_HIP(hipMemcpyAsync(evh.data() + (g - trd_m), ev, sizeof(double) * trd_m, hipMemcpyDeviceToHost,
                                c->stream));
  DMRG_CHECK_HIP(hipMemcpyAsync(&hinfo, info, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  DMRG_CHECK_HIP(hipMemcpyAsync(&hmax, amax_dev, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  DMRG_CHECK_HIP(hipStreamSynchronize(c->stream));
  DMRG_REQUIRE(hinfo == 0, "syevd/heevd did not converge");
  const double scale = hmax > 0.0 ? hmax : 1.0;
  svals.resize(rank);
  for (int i = 0; i < rank; ++i) svals[i] = scale * std::sqrt(std::max(evh[g - 1 - i], 0.0));
  SvdOut o;
  // a stopped reduction's dropped block: its eigenvalues (all below the
  // stop tolerance) are counted as discarded weight, in units of s^2
  o.keep = truncate_rule(svals, cutoff, maxdim, mind, &o.discarded, extra + scale * scale * trd_dropped);
  const int keep = o.keep;
  if (!(svals[0] > 0.0 && svals[keep - 1] >= 1e-7 * svals[0]) || keep > trd_m)
    return svd_full(c, field, M, m, n, cutoff, maxdim, mind, absorb, U, V, kcap, svals);
  DMRG_REQUIRE(keep <= kcap, "svd: keep exceeds output capacity");
  if (trd_refl && !cx) {
    // eigenvectors of G for the kept (largest) values only: Q times the
    // last keep eigenvector columns of the tridiagonal
    ProfScope ps(c, 2, 2.0 * (double)g * trd_nrefl * keep);
    double* Zt = trd_Z + (size_t)(g - keep) * g;
    if (trd_m < g) {   // the kept eigenvectors of the leading block, zero beyond it
      const long long tot = (long long)keep * g;
      embed_rows_k<<<g1(tot), 256, 0, c->stream>>>(trd_Zs + (size_t)(trd_m - keep) * trd_m, trd_m, keep, g, Zt);
      DMRG_CHECK_HIP(hipGetLastError());
    }
    if (c->eig_mode != 2)
      ormtr_lower_rows(c, trd_refl, g, trd_tau, Zt, keep, trd_nrefl);
    else
      check_rb(rocsolver_dormtr(c->blas, rocblas_side_left, rocblas_fill_lower, rocblas_operation_none, g, keep,
                                trd_refl, g, trd_tau, trd_Z + (size_t)(g - keep) * g, g),
               "dormtr");
  } else if (trd_refl) {
    // complex: Q (complex) times the kept tridiagonal eigenvectors, then
    // interleaved columns as the heevd route leaves them
    ProfScope ps(c, 2, 8.0 * (double)g * g * keep);
    double* Zc = (double*)ws_get(c, WS_TRD_ZC, sizeof(double) * 2 * (size_t)g * g);
    double* Zk = Zc + 2 * (size_t)(g - keep) * g;
    const long long kg = (long long)keep * g;
    if (c->eig_mode != 2) {   // own blocked product on planar rows
      double* Et = (double*)ws_get(c, WS_SVD_CPLX, sizeof(double) * 2 * (size_t)kg);
      if (trd_m < g) {   // the kept eigenvectors of the leading block, zero beyond it
        embed_rows_k<<<g1(kg), 256, 0, c->stream>>>(trd_Zs + (size_t)(trd_m - keep) * trd_m, trd_m, keep, g, Et);
        DMRG_CHECK_HIP(hipGetLastError());
      } else {
        DMRG_CHECK_HIP(hipMemcpyAsync(Et, trd_Z + (size_t)(g - keep) * g, sizeof(double) * kg,
                                      hipMemcpyDeviceToDevice, c->stream));
      }
      DMRG_CHECK_HIP(hipMemsetAsync(Et + kg, 0, sizeof(double) * kg, c->stream));
      unmtr_lower_rows(c, trd_refl, g, trd_tau, Et, keep, trd_nrefl);
      planar_to_inter_k<<<g1(kg), 256, 0, c->stream>>>(Et, kg, Zk);
      DMRG_CHECK_HIP(hipGetLastError());
    } else {   // rocSOLVER zunmtr on interleaved reflectors
      double* Gi = (double*)ws_get(c, WS_SVD_CPLX, sizeof(double) * 2 * (size_t)g * g);
      real_to_inter_k<<<g1(kg), 256, 0, c->stream>>>(trd_Z + (size_t)(g - keep) * g, kg, Zk);
      DMRG_CHECK_HIP(hipGetLastError());
      planar_to_inter_k<<<g1((long long)g * g), 256, 0, c->stream>>>(trd_refl, (long long)g * g, Gi);
      DMRG_CHECK_HIP(hipGetLastError());
      check_rb(rocsolver_zunmtr(c->blas, rocblas_side_left, rocblas_fill_lower, rocblas_operation_none, g, keep,
                                (rocblas_double_complex*)Gi, g, (rocblas_double_complex*)trd_tau,
                                (rocblas_double_complex*)Zk, g),
               "zunmtr");
    }
    Gsolve = Zc;
  }
  // device copy of the kept singular values
  double* sdev = ev;  // reuse: ev no longer needed
  {
    std::vector<double> sk(svals.begin(), svals.begin() + keep);
    DMRG_CHECK_HIP(hipMemcpyAsync(sdev, sk.data(), sizeof(double) * keep, hipMemcpyHostToDevice, c->stream));
    // E: keep x g, rows in descending eigenvalue order
    double* E = (double*)ws_get(c, WS_SVD_TMP, sizeof(double) * es * (size_t)keep * g);
    const long long pE = (long long)keep * g;
    if (cx) take_rows_k<true><<<g1(pE), 256, 0, c->stream>>>(Gsolve, g, g - 1, keep, g, E, pE);
    else take_rows_k<false><<<g1(pE), 256, 0, c->stream>>>(Gsolve, g, g - 1, keep, g, E, pE);
    DMRG_CHECK_HIP(hipGetLastError());
    if (left) {
      // U = E^dag (m x keep), S Vh = E M (keep x n)
      if (cx) conj_transpose_k<true><<<g1(pE), 256, 0, c->stream>>>(E, keep, g, pE, U, (long long)m * keep, sdev, absorb == 1);
      else conj_transpose_k<false><<<g1(pE), 256, 0, c->stream>>>(E, keep, g, pE, U, 0, sdev, absorb == 1);
      DMRG_CHECK_HIP(hipGetLastError());
      GemmDesc d = gdesc(keep, n, m, m, n, n, field, 4);
      d.pA = pE; d.pB = mn; d.pC = (long long)keep * n;
      gemm_f64(c, d, E, M, V);
      if (absorb != 0) {  // Vh = (E M) / S
        if (cx) scale_k<true><<<g1((long long)keep * n), 256, 0, c->stream>>>(V, keep, n, (long long)keep * n, sdev, 1, 1);
        else scale_k<false><<<g1((long long)keep * n), 256, 0, c->stream>>>(V, keep, n, 0, sdev, 1, 1);
      }
    } else {
      // Vh = E (keep x n), U S = M E^dag (m x keep)
      DMRG_CHECK_HIP(hipMemcpyAsync(V, E, sizeof(double) * pE, hipMemcpyDeviceToDevice, c->stream));
      if (cx) DMRG_CHECK_HIP(hipMemcpyAsync(V + (long long)keep * n, E + pE, sizeof(double) * pE,
                                             hipMemcpyDeviceToDevice, c->stream));
      if (absorb == 0) {
        if (cx) scale_k<true><<<g1(pE), 256, 0, c->stream>>>(V, keep, n, (long long)keep * n, sdev, 1, 0);
        else scale_k<false><<<g1(pE), 256, 0, c->stream>>>(V, keep, n, 0, sdev, 1, 0);
      }
      GemmDesc d = gdesc(m, keep, n, n, n, keep, field, 4);
      d.transB = true; d.conjB = true;
      d.pA = mn; d.pB = pE; d.pC = (long long)m * keep;
      gemm_f64(c, d, M, E, U);
      if (absorb != 1) {  // U = M E^dag / S
        if (cx) scale_k<true><<<g1((long long)m * keep), 256, 0, c->stream>>>(U, m, keep, (long long)m * keep, sdev, 0, 1);
        else scale_k<false><<<g1((long long)m * keep), 256, 0, c->stream>>>(U, m, keep, 0, sdev, 0, 1);
      }
    }
    DMRG_CHECK_HIP(hipGetLastError());
    // The factor that stays behind in a sweep (U when S goes right, Vh when
    // S goes left) must be an isometry.  When the Gram side made it the
    // derived factor (divided by S), re-orthonormalise it; if that fails,
    // redo the split with the Gram on its side, where it is an
    // eigenvector matrix.
    bool fixed = true;
    if (left && absorb == 1) fixed = lowdin(c, field, V, keep, n, true, U, m);
    else if (!left && absorb == 0) fixed = lowdin(c, field, U, keep, m, false, V, n);
    // sk / E must outlive the async copies and kernels above
    DMRG_CHECK_HIP(hipStreamSynchronize(c->stream));
    if (!fixed)
      return svd_gram(c, field, M, m, n, cutoff, maxdim, mind, absorb, U, V, kcap, svals, left ? 1 : 0, extra);
  }
  return o;
}

// ---------------------------------------------------------------------------
// Subspace-guided split.  In a sweep the two-site matrix of bond i is
// theta = A_i A_{i+1} with one factor an isometry from the previous sweep
// (A_{i+1} with orthonormal rows when moving right, A_i with orthonormal
// columns when moving left), and the local update changes theta only a
// little.  So the rows of theta lie almost entirely in the row space of the
// guide Q0 (k0 x n).  The split:
//   1. sketch the part outside Q0: Y = Omega M (p x n, Omega uniform random),
//      projected off Q0 twice; orthonormal rows Q1 from the eigenvectors of
//      Y Y^dag (directions below 1e-7 of the largest dropped);
//   2. Q = [Q0; Q1] re-orthonormalised (Loewdin), B = M Q^dag (m x l);
//   3. residual weight r = |M - B Q|^2 -- exactly the weight the subspace
//      misses; accepted only when r <= 0.1 * cutoff * |M|^2, or, when maxdim
//      forces the truncation, r <= 0.1 * the weight maxdim discards;
//      otherwise the caller runs the full split;
//   4. the Gram-route split of B (l x l eigenproblem instead of min(m, n)),
//      with r counted as discarded weight in the keep rule, and V = Vb Q.
// The result is the exact truncated SVD of M Q^dag Q, which differs from
// M by less than a tenth of the weight the keep rule discards anyway.  At a
// cutoff-limited bond the weight an update moves outside the old span is
// typically of the order of the cutoff itself (measured 1e-12..2e-11 at
// cutoff 1e-12 on a D=1024 Hubbard chain, spread over hundreds of
// directions), so there the full split runs; the guided split pays where
// maxdim limits the bond or the cutoff is looser.
// ---------------------------------------------------------------------------

// uniform entries in [-1, 1) from a counter hash (splitmix64 finaliser)
__global__ void sketch_k(double* __restrict__ O, long long n, unsigned long long seed) {
  long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  for (; i < n; i += (long long)gridDim.x * 256) {
    unsigned long long z = seed + 0x9E3779B97F4A7C15ull * (unsigned long long)(i + 1);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    O[i] = (double)(z >> 11) * (2.0 / 9007199254740992.0) - 1.0;
  }
}

// sum of squares, two passes in a fixed order (deterministic)
__global__ __launch_bounds__(256) void sumsq_partial_k(const double* __restrict__ x, long long n,
                                                       double* __restrict__ part) {
  __shared__ double red[256];
  double a = 0.0;
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256)
    a += x[i] * x[i];
  red[threadIdx.x] = a;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(256) void sum_final_k(const double* __restrict__ part, int nb, double* __restrict__ out) {
  __shared__ double red[256];
  double a = 0.0;
  for (int i = threadIdx.x; i < nb; i += 256) a += part[i];
  red[threadIdx.x] = a;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = red[0];
}

static void dev_sumsq(Ctx* c, const double* x, long long n, double* out) {
  long long nb = (n + 4095) / 4096;
  if (nb > 1024) nb = 1024;
  if (nb < 1) nb = 1;
  double* part = (double*)ws_get(c, WS_SVD_RED, sizeof(double) * 1024);
  sumsq_partial_k<<<(unsigned)nb, 256, 0, c->stream>>>(x, n, part);
  DMRG_CHECK_HIP(hipGetLastError());
  sum_final_k<<<1, 256, 0, c->stream>>>(part, (int)nb, out);
  DMRG_CHECK_HIP(hipGetLastError());
}

// planar block copy of `rows` rows of length `cols` (per plane)
static void copy_rows(Ctx* c, bool cx, const double* src, long long psrc, double* dst, long long pdst,
                      long long count) {
  DMRG_CHECK_HIP(hipMemcpyAsync(dst, src, sizeof(double) * count, hipMemcpyDeviceToDevice, c->stream));
  if (cx)
    DMRG_CHECK_HIP(
        hipMemcpyAsync(dst + pdst, src + psrc, sizeof(double) * count, hipMemcpyDeviceToDevice, c->stream));
}

// Row-guided split of M (m x n) with guide rows Q0 (k0 x n).  Returns false
// when the subspace misses too much weight (nothing written that the caller
// relies on).
static bool svd_rowguided(Ctx* c, int field, const double* M, int m, int n, const double* Q0, int k0, int sketch,
                          double cutoff, int maxdim, int mind, int absorb, double* U, double* V, int kcap,
                          std::vector<double>& svals, SvdOut& out, double* resid) {
  const bool cx = field == CPLX;
  const size_t es = cx ? 2 : 1;
  const long long mn = (long long)m * n;
  const int p = std::min(sketch, n - k0);
  if (p <= 0 || k0 <= 0) return false;
  // scratch: Omega (p x m), Y (p x n), C (p x k0), Gy (p x p, interleaved if cx), E (p x p), Y1 (p x n)
  const long long nO = (long long)p * m, nY = (long long)p * n, nC = (long long)p * k0, nG = (long long)p * p;
  double* S = (double*)ws_get(c, WS_GSV_S, sizeof(double) * (es * (nO + 2 * nY + nC + 2 * nG) + 2 * (size_t)p + 64));
  double* Om = S;
  double* Y = Om + es * nO;
  double* Cc = Y + es * nY;
  double* Gy = Cc + es * nC;
  double* E = Gy + es * nG;
  double* Y1 = E + es * nG;
  double* ev = Y1 + es * nY;
  double* scal = ev + p;   // device scalars: [0] |R|^2, [1] |M|^2
  rocblas_int* info = (rocblas_int*)ws_get(c, WS_SVD_INFO, 64);
  {
    ProfScope ps(c, 2, (cx ? 8.0 : 2.0) * (double)p * ((double)m * n + 4.0 * n * k0 + (double)p * n));
    sketch_k<<<g1(nO), 256, 0, c->stream>>>(Om, nO, 0x5EEDC0FFEEull);
    DMRG_CHECK_HIP(hipGetLastError());
    if (cx) DMRG_CHECK_HIP(hipMemsetAsync(Om + nO, 0, sizeof(double) * nO, c->stream));
    {  // Y = Omega M
      GemmDesc d = gdesc(p, n, m, m, n, n, field, 4);
      d.pA = nO; d.pB = mn; d.pC = nY;
      gemm_f64(c, d, Om, M, Y);
    }
    for (int pass = 0; pass < 2; ++pass) {  // Y <- Y - (Y Q0^dag) Q0
      GemmDesc d = gdesc(p, k0, n, n, n, k0, field, 4);
      d.transB = true; d.conjB = true;
      d.pA = nY; d.pB = (long long)k0 * n; d.pC = nC;
      gemm_f64(c, d, Y, Q0, Cc);
      GemmDesc e = gdesc(p, n, k0, k0, n, n, field, 4);
      e.alpha = -1.0; e.beta = 1.0;
      e.pA = nC; e.pB = (long long)k0 * n; e.pC = nY;
      gemm_f64(c, e, Cc, Q0, Y);
    }
    {  // Gy = Y Y^dag
      GemmDesc d = gdesc(p, p, n, n, n, p, field, 4);
      d.transB = true; d.conjB = true;
      d.pA = d.pB = nY; d.pC = nG;
      gemm_f64(c, d, Y, Y, E);
    }
    if (cx) {
      planar_to_inter_k<<<g1(nG), 256, 0, c->stream>>>(E, nG, Gy);
      DMRG_CHECK_HIP(hipGetLastError());
      check_rb(rocsolver_zheevd(c->blas, rocblas_evect_original, rocblas_fill_upper, p, (rocblas_double_complex*)Gy,
                                p, ev, Y1, info),
               "zheevd");
    } else {
      DMRG_CHECK_HIP(hipMemcpyAsync(Gy, E, sizeof(double) * nG, hipMemcpyDeviceToDevice, c->stream));
      check_rb(rocsolver_dsyevd(c->blas, rocblas_evect_original, rocblas_fill_upper, p, Gy, p, ev, Y1, info),
               "dsyevd");
    }
  }
  std::vector<double> evh(p);
  int hinfo = 0;
  DMRG_CHECK_HIP(hipMemcpyAsync(evh.data(), ev, sizeof(double) * p, hipMemcpyDeviceToHost, c->stream));
  DMRG_CHECK_HIP(hipMemcpyAsync(&hinfo, info, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  DMRG_CHECK_HIP(hipStreamSynchronize(c->stream));
  if (hinfo != 0) return false;
  const double evmax = evh[p - 1];
  int r = 0;
  if (evmax > 0.0)
    while (r < p && evh[p - 1 - r] > 1e-14 * evmax) ++r;
  const int l = k0 + r;
  const long long nQ = (long long)l * n;
  double* Q = (double*)ws_get(c, WS_GSV_Q, sizeof(double) * es * (size_t)nQ);
  copy_rows(c, cx, Q0, (long long)k0 * n, Q, nQ, (long long)k0 * n);
  if (r > 0) {
    // Q1 = diag(ev^-1/2) E Y with E the top-r eigenvector rows (descending)
    std::vector<double> inv(r);
    for (int i = 0; i < r; ++i) inv[i] = std::sqrt(evh[p - 1 - i]);
    double* sdev = ev;   // ev already read back
    DMRG_CHECK_HIP(hipMemcpyAsync(sdev, inv.data(), sizeof(double) * r, hipMemcpyHostToDevice, c->stream));
    const long long nE = (long long)r * p;
    if (cx) take_rows_k<true><<<g1(nE), 256, 0, c->stream>>>(Gy, p, p - 1, r, p, E, nE);
    else take_rows_k<false><<<g1(nE), 256, 0, c->stream>>>(Gy, p, p - 1, r, p, E, nE);
    DMRG_CHECK_HIP(hipGetLastError());
    GemmDesc d = gdesc(r, n, p, p, n, n, field, 4);
    d.pA = nE; d.pB = nY; d.pC = nQ;
    gemm_f64(c, d, E, Y, Q + (long long)k0 * n);
    if (cx) scale_k<true><<<g1((long long)r * n), 256, 0, c->stream>>>(Q + (long long)k0 * n, r, n, nQ, sdev, 1, 1);
    else scale_k<false><<<g1((long long)r * n), 256, 0, c->stream>>>(Q + (long long)k0 * n, r, n, 0, sdev, 1, 1);
    DMRG_CHECK_HIP(hipGetLastError());
    DMRG_CHECK_HIP(hipStreamSynchronize(c->stream));   // inv must outlive its copy
  }
  if (!lowdin(c, field, Q, l, n, true, nullptr, 0)) return false;
  // B = M Q^dag (m x l); residual R = M - B Q
  const long long nB = (long long)m * l;
  double* B = (double*)ws_get(c, WS_GSV_B, sizeof(double) * es * (size_t)nB);
  double* R = (double*)ws_get(c, WS_GSV_R, sizeof(double) * es * (size_t)mn);
  {
    ProfScope ps(c, 2, (cx ? 8.0 : 2.0) * 2.0 * (double)m * n * l);
    GemmDesc d = gdesc(m, l, n, n, n, l, field, 4);
    d.transB = true; d.conjB = true;
    d.pA = mn; d.pB = nQ; d.pC = nB;
    gemm_f64(c, d, M, Q, B);
    DMRG_CHECK_HIP(hipMemcpyAsync(R, M, sizeof(double) * es * mn, hipMemcpyDeviceToDevice, c->stream));
    GemmDesc e = gdesc(m, n, l, l, n, n, field, 4);
    e.alpha = -1.0; e.beta = 1.0;
    e.pA = nB; e.pB = nQ; e.pC = mn;
    gemm_f64(c, e, B, Q, R);
    dev_sumsq(c, R, (long long)es * mn, scal);
    dev_sumsq(c, M, (long long)es * mn, scal + 1);
  }
  double hs[2] = {0.0, 0.0};
  DMRG_CHECK_HIP(hipMemcpyAsync(hs, scal, sizeof(double) * 2, hipMemcpyDeviceToHost, c->stream));
  DMRG_CHECK_HIP(hipStreamSynchronize(c->stream));
  const double res2 = hs[0], tot2 = hs[1];
  *resid = tot2 > 0.0 ? res2 / tot2 : -1.0;
  if (!(tot2 > 0.0) || !std::isfinite(tot2)) return false;
  const double rel = res2 / tot2;
  const bool within_cutoff = rel <= 0.1 * cutoff;
  // a split that maxdim truncates discards more than the cutoff anyway; the
  // missed weight may then be up to a tenth of that forced weight
  const bool may_be_forced = maxdim > 0 && l > maxdim;
  if (!within_cutoff && !may_be_forced) return false;
  // split of B on its right Gram (l x l); V = Vb Q
  const int kc = std::min(kcap, l);
  double* Vb = (double*)ws_get(c, WS_GSV_V, sizeof(double) * es * (size_t)kc * l);
  SvdOut o = svd_gram(c, field, B, m, l, cutoff, maxdim, mind, absorb, U, Vb, kc, svals, 1, res2);
  if (!within_cutoff && !(o.keep >= maxdim && rel <= 0.1 * (o.discarded - rel))) return false;
  {
    GemmDesc d = gdesc(o.keep, n, l, l, n, n, field, 4);
    d.pA = (long long)o.keep * l; d.pB = nQ; d.pC = (long long)o.keep * n;
    gemm_f64(c, d, Vb, Q, V);
  }
  out = o;
  out.resid = *resid;
  return true;
}

// Guided split entry: gside 0 = Q0 is k0 x n with orthonormal rows (the
// factor right of the bond), gside 1 = Q0 is m x k0 with orthonormal columns
// (the factor left of the bond; handled as the row-guided split of M^dag).
// Falls back to svd_split when the guide is too large to pay off or the
// subspace misses weight; *used reports which path ran.
SvdOut svd_split_guided(Ctx* c, int field, const double* M, int m, int n, double cutoff, int maxdim, int mindim,
                        int absorb, const double* G, int k0, int gside, int sketch, double* U, double* V, int kcap,
                        std::vector<double>& svals, int* used, double* resid) {
  DMRG_REQUIRE(m > 0 && n > 0, "svd: empty matrix");
  DMRG_REQUIRE(absorb >= 0 && absorb <= 2, "svd: absorb must be 0, 1 or 2");
  DMRG_REQUIRE(gside == 0 || gside == 1, "svd: guide side must be 0 (rows) or 1 (columns)");
  const int mind = std::max(1, mindim);
  *used = 0;
  *resid = -1.0;
  const int len = gside == 0 ? n : m;   // guide vector length
  const bool worth = G && k0 > 0 && sketch > 0 && (cutoff > 0.0 || maxdim > 0) && std::min(m, n) >= 16 &&
                     (double)(k0 + sketch) <= 0.6 * std::min(m, n) && k0 < len;
  if (worth) {
    SvdOut o;
    bool ok;
    if (gside == 0) {
      ok = svd_rowguided(c, field, M, m, n, G, k0, sketch, cutoff, maxdim, mind, absorb, U, V, kcap, svals, o,
                         resid);
    } else {
      const bool cx = field == CPLX;
      const size_t es = cx ? 2 : 1;
      const long long mn = (long long)m * n, mk = (long long)m * k0;
      double* Mt = (double*)ws_get(c, WS_GSV_T, sizeof(double) * es * mn);
      double* Gt = (double*)ws_get(c, WS_GSV_T2, sizeof(double) * es * mk);
      if (cx) {
        conj_transpose_k<true><<<g1(mn), 256, 0, c->stream>>>(M, m, n, mn, Mt, mn, nullptr, 0);
        conj_transpose_k<true><<<g1(mk), 256, 0, c->stream>>>(G, m, k0, mk, Gt, mk, nullptr, 0);
      } else {
        conj_transpose_k<false><<<g1(mn), 256, 0, c->stream>>>(M, m, n, 0, Mt, 0, nullptr, 0);
        conj_transpose_k<false><<<g1(mk), 256, 0, c->stream>>>(G, m, k0, 0, Gt, 0, nullptr, 0);
      }
      DMRG_CHECK_HIP(hipGetLastError());
      const int kc = std::min(kcap, std::min(m, n));
      double* Ut = (double*)ws_get(c, WS_GSV_OU, sizeof(double) * es * (size_t)n * kc);
      double* Vt = (double*)ws_get(c, WS_GSV_OV, sizeof(double) * es * (size_t)kc * m);
      const int absorb_t = absorb == 0 ? 1 : (absorb == 1 ? 0 : 2);
      ok = svd_rowguided(c, field, Mt, n, m, Gt, k0, sketch, cutoff, maxdim, mind, absorb_t, Ut, Vt, kc, svals, o,
                         resid);
      if (ok) {
        // M^dag = Ut S Vt  =>  M = Vt^dag S Ut^dag
        const int k = o.keep;
        const long long pU = (long long)m * k, pV = (long long)k * n;
        if (cx) {
          conj_transpose_k<true><<<g1(pU), 256, 0, c->stream>>>(Vt, k, m, (long long)k * m, U, pU, nullptr, 0);
          conj_transpose_k<true><<<g1(pV), 256, 0, c->stream>>>(Ut, n, k, (long long)n * k, V, pV, nullptr, 0);
        } else {
          conj_transpose_k<false><<<g1(pU), 256, 0, c->stream>>>(Vt, k, m, 0, U, 0, nullptr, 0);
          conj_transpose_k<false><<<g1(pV), 256, 0, c->stream>>>(Ut, n, k, 0, V, 0, nullptr, 0);
        }
        DMRG_CHECK_HIP(hipGetLastError());
      }
    }
    if (ok) {
      DMRG_REQUIRE(o.keep <= kcap, "svd: keep exceeds output capacity");
      *used = 1;
      return o;
    }
  }
  return svd_split(c, field, M, m, n, cutoff, maxdim, mindim, absorb, U, V, kcap, svals);
}

SvdOut svd_split(Ctx* c, int field, const double* M, int m, int n, double cutoff, int maxdim, int mindim,
                 int absorb, double* U, double* V, int kcap, std::vector<double>& svals) {
  DMRG_REQUIRE(m > 0 && n > 0, "svd: empty matrix");
  DMRG_REQUIRE(absorb >= 0 && absorb <= 2, "svd: absorb must be 0, 1 or 2");
  const int mind = std::max(1, mindim);
  if (std::min(m, n) < 16) return svd_full(c, field, M, m, n, cutoff, maxdim, mind, absorb, U, V, kcap, svals);
  return svd_gram(c, field, M, m, n, cutoff, maxdim, mind, absorb, U, V, kcap, svals, -1);
}

int svd_split_entry(Ctx* c, int field, const double* M, int m, int n, double cutoff, int maxdim, int mindim,
                    int absorb, double* U, double* V, int kcap, int* keep, double* discarded, double* svals_host) {
  std::vector<double> sv;
  SvdOut o = svd_split(c, field, M, m, n, cutoff, maxdim, mindim, absorb, U, V, kcap, sv);
  if (keep) *keep = o.keep;
  if (discarded) *discarded = o.discarded;
  if (svals_host)
    for (size_t i = 0; i < sv.size(); ++i) svals_host[i] = sv[i];
  return 0;
}

// All eigenpairs of a real symmetric matrix (row-major = column-major):
// w ascending, Z[k][:] the eigenvector of w[k] (rows of the row-major Z).
// modes 1 and 3: own tridiagonalisation (always complete here) + stedc + own
// back-transformation, mode 2: the same with rocSOLVER's ormtr, mode 0:
// syevd.  A is overwritten.
int sym_eig_entry(Ctx* c, int mode, double* A, int n, double* w, double* Z) {
  DMRG_REQUIRE(n >= 2, "sym_eig: order must be at least 2");
  double* ework = (double*)ws_get(c, WS_SVD_EV, sizeof(double) * ((size_t)2 * n + 16));
  double* tauv = ework + n;
  rocblas_int* info = (rocblas_int*)ws_get(c, WS_SVD_INFO, 64);
  if (mode >= 1) {   // every eigenvector is wanted: always the full reduction
    sytrd_lower(c, A, n, w, ework, tauv);
    check_rb(rocsolver_dstedc(c->blas, rocblas_evect_tridiagonal, n, w, ework, Z, n, info), "dstedc");
    if (mode != 2)
      ormtr_lower_rows(c, A, n, tauv, Z, n);
    else
      check_rb(rocsolver_dormtr(c->blas, rocblas_side_left, rocblas_fill_lower, rocblas_operation_none, n, n, A, n,
                                tauv, Z, n),
               "dormtr");
  } else {
    check_rb(rocsolver_dsyevd(c->blas, rocblas_evect_original, rocblas_fill_upper, n, A, n, w, ework, info),
             "dsyevd");
    DMRG_CHECK_HIP(hipMemcpyAsync(Z, A, sizeof(double) * n * n, hipMemcpyDeviceToDevice, c->stream));
  }
  int hinfo = 0;
  DMRG_CHECK_HIP(hipMemcpyAsync(&hinfo, info, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  DMRG_CHECK_HIP(hipStreamSynchronize(c->stream));
  DMRG_REQUIRE(hinfo == 0, "sym_eig: eigensolver did not converge");
  return 0;
}

// All eigenpairs of a Hermitian matrix A (row-major, planar complex, plane
// stride n*n): w ascending, Z planar with Z[k][:] the eigenvector of w[k].
// modes 1 and 3: own reduction + stedc + own back-transformation, mode 2:
// the same with rocSOLVER's unmtr, mode 0: heevd.  A is overwritten.
// The column-major view of the buffer is conj(A); its eigenvectors come out
// as rows of the interleaved buffer and are conjugated on the way out.
int herm_eig_entry(Ctx* c, int mode, double* A, int n, double* w, double* Z) {
  DMRG_REQUIRE(n >= 2, "herm_eig: order must be at least 2");
  const long long nn = (long long)n * n;
  double* ework = (double*)ws_get(c, WS_SVD_EV, sizeof(double) * ((size_t)3 * n + 16));
  double* tauv = ework + n;   // complex, interleaved
  rocblas_int* info = (rocblas_int*)ws_get(c, WS_SVD_INFO, 64);
  double* Zc = (double*)ws_get(c, WS_TRD_ZC, sizeof(double) * 2 * (size_t)nn);
  double* Ai = (double*)ws_get(c, WS_SVD_CPLX, sizeof(double) * 2 * (size_t)nn);
  if (mode >= 1) {
    double* Zr = (double*)ws_get(c, WS_TRD_Z, sizeof(double) * (size_t)nn);
    hetrd_lower(c, A, n, w, ework, tauv);
    check_rb(rocsolver_dstedc(c->blas, rocblas_evect_tridiagonal, n, w, ework, Zr, n, info), "dstedc");
    if (mode != 2) {   // Ai holds the planar rows of Q Y
      DMRG_CHECK_HIP(hipMemcpyAsync(Ai, Zr, sizeof(double) * nn, hipMemcpyDeviceToDevice, c->stream));
      DMRG_CHECK_HIP(hipMemsetAsync(Ai + nn, 0, sizeof(double) * nn, c->stream));
      unmtr_lower_rows(c, A, n, tauv, Ai, n);
      planar_to_inter_k<<<g1(nn), 256, 0, c->stream>>>(Ai, nn, Zc);
      DMRG_CHECK_HIP(hipGetLastError());
    } else {
      real_to_inter_k<<<g1(nn), 256, 0, c->stream>>>(Zr, nn, Zc);
      DMRG_CHECK_HIP(hipGetLastError());
      planar_to_inter_k<<<g1(nn), 256, 0, c->stream>>>(A, nn, Ai);
      DMRG_CHECK_HIP(hipGetLastError());
      check_rb(rocsolver_zunmtr(c->blas, rocblas_side_left, rocblas_fill_lower, rocblas_operation_none, n, n,
                                (rocblas_double_complex*)Ai, n, (rocblas_double_complex*)tauv,
                                (rocblas_double_complex*)Zc, n),
               "zunmtr");
    }
  } else {
    planar_to_inter_k<<<g1(nn), 256, 0, c->stream>>>(A, nn, Zc);
    DMRG_CHECK_HIP(hipGetLastError());
    check_rb(rocsolver_zheevd(c->blas, rocblas_evect_original, rocblas_fill_upper, n, (rocblas_double_complex*)Zc,
                              n, w, ework, info),
             "zheevd");
  }
  inter_to_planar_conj_k<<<g1(nn), 256, 0, c->stream>>>(Zc, nn, Z);
  DMRG_CHECK_HIP(hipGetLastError());
  int hinfo = 0;
  DMRG_CHECK_HIP(hipMemcpyAsync(&hinfo, info, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  DMRG_CHECK_HIP(hipStreamSynchronize(c->stream));
  DMRG_REQUIRE(hinfo == 0, "herm_eig: eigensolver did not converge");
  return 0;
}

int svd_split_guided_entry(Ctx* c, int field, const double* M, int m, int n, double cutoff, int maxdim,
                           int mindim, int absorb, const double* G, int k0, int gside, int sketch, double* U,
                           double* V, int kcap, int* keep, double* discarded, double* svals_host, int* used,
                           double* resid) {
  std::vector<double> sv;
  int u = 0;
  double rs = -1.0;
  SvdOut o = svd_split_guided(c, field, M, m, n, cutoff, maxdim, mindim, absorb, G, k0, gside, sketch, U, V, kcap,
                              sv, &u, &rs);
  if (resid) *resid = rs;
  if (keep) *keep = o.keep;
  if (discarded) *discarded = o.discarded;
  if (used) *used = u;
  if (svals_host) {
    const size_t full = (size_t)std::min(m, n);
    for (size_t i = 0; i < full; ++i) svals_host[i] = i < sv.size() ? sv[i] : 0.0;
  }
  return 0;
}

}  // namespace dmrg
