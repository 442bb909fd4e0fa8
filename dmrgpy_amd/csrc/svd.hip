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
// product of the two factors is unchanged.  Returns false, with nothing
// modified, when the iteration does not converge.
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
  const long long big = std::max((long long)len, (long long)plen) * keep;
  double* tmp = (double*)ws_get(c, WS_SVD_GRAM, sizeof(double) * es * (size_t)big);
  if (rows) {  // X <- Z X (keep x len), P <- P Y (plen x keep)
    GemmDesc d = gdesc(keep, len, keep, keep, len, len, field, 4);
    d.pA = kk; d.pB = pX; d.pC = pX;
    gemm_f64(c, d, Z, X, tmp);
    DMRG_CHECK_HIP(hipMemcpyAsync(X, tmp, sizeof(double) * es * pX, hipMemcpyDeviceToDevice, c->stream));
    GemmDesc e = gdesc(plen, keep, keep, keep, keep, keep, field, 4);
    e.pA = pP; e.pB = kk; e.pC = pP;
    gemm_f64(c, e, P, Y, tmp);
    DMRG_CHECK_HIP(hipMemcpyAsync(P, tmp, sizeof(double) * es * pP, hipMemcpyDeviceToDevice, c->stream));
  } else {     // X <- X Z (len x keep), P <- Y P (keep x plen)
    GemmDesc d = gdesc(len, keep, keep, keep, keep, keep, field, 4);
    d.pA = pX; d.pB = kk; d.pC = pX;
    gemm_f64(c, d, X, Z, tmp);
    DMRG_CHECK_HIP(hipMemcpyAsync(X, tmp, sizeof(double) * es * pX, hipMemcpyDeviceToDevice, c->stream));
    GemmDesc e = gdesc(keep, plen, keep, keep, plen, plen, field, 4);
    e.pA = kk; e.pB = pP; e.pC = pP;
    gemm_f64(c, e, Y, P, tmp);
    DMRG_CHECK_HIP(hipMemcpyAsync(P, tmp, sizeof(double) * es * pP, hipMemcpyDeviceToDevice, c->stream));
  }
  return true;
}

// The reference's keep rule on a descending spectrum.
int truncate_rule(const std::vector<double>& s, double cutoff, int maxdim, int mindim, double* discarded) {
  const int n = (int)s.size();
  std::vector<double> p(n);
  double total = 0;
  for (int i = 0; i < n; ++i) {
    p[i] = s[i] * s[i];
    total += p[i];
  }
  if (total <= 0) {
    if (discarded) *discarded = 0.0;
    return std::max(1, mindim);
  }
  for (int i = 0; i < n; ++i) p[i] /= total;
  const int hi = maxdim > 0 ? std::min(maxdim, n) : n;
  const int j_upper = std::max(0, n - mindim);
  const int j_forced = std::min(std::max(0, n - hi), j_upper);
  int j;
  if (j_upper <= j_forced) {
    j = j_forced;
  } else {
    // tail(t) = sum of the j_forced + t smallest, t = 0..j_upper-j_forced;
    // keep the largest t with tail <= cutoff (ascending prefix scan)
    double forced = 0;
    for (int t = 0; t < j_forced; ++t) forced += p[n - 1 - t];
    int best = 0;
    double tail = forced;
    if (tail <= cutoff) best = 0;
    bool any = tail <= cutoff;
    for (int t = 1; t <= j_upper - j_forced; ++t) {
      tail += p[n - j_forced - t];
      if (tail <= cutoff) {
        best = t;
        any = true;
      }
    }
    j = j_forced + (any ? best : 0);
  }
  const int keep = n - j;
  if (discarded) {
    double dsc = 0;
    for (int i = keep; i < n; ++i) dsc += p[i];
    *discarded = dsc;
  }
  return keep;
}

struct SvdOut {
  int keep;
  double discarded;
};

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
  SvdOut o;
  o.keep = truncate_rule(svals, cutoff, maxdim, mindim, &o.discarded);
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
                       int absorb, double* U, double* V, int kcap, std::vector<double>& svals, int side) {
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
    gemm_f64(c, d, Ms, Ms, Gplanar);
    if (cx) {
      double* Gi = (double*)ws_get(c, WS_SVD_CPLX, sizeof(double) * 2 * (size_t)g * g);
      planar_to_inter_k<<<g1((long long)g * g), 256, 0, c->stream>>>(Gplanar, (long long)g * g, Gi);
      DMRG_CHECK_HIP(hipGetLastError());
      check_rb(rocsolver_zheevd(c->blas, rocblas_evect_original, rocblas_fill_upper, g, (rocblas_double_complex*)Gi,
                                g, ev, ework, info),
               "zheevd");
      Gsolve = Gi;
    } else {
      check_rb(rocsolver_dsyevd(c->blas, rocblas_evect_original, rocblas_fill_upper, g, Gplanar, g, ev, ework, info),
               "dsyevd");
      Gsolve = Gplanar;
    }
  }
  std::vector<double> evh(g);
  int hinfo = 0;
  double hmax = 0.0;
  DMRG_CHECK_HIP(hipMemcpyAsync(evh.data(), ev, sizeof(double) * g, hipMemcpyDeviceToHost, c->stream));
  DMRG_CHECK_HIP(hipMemcpyAsync(&hinfo, info, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  DMRG_CHECK_HIP(hipMemcpyAsync(&hmax, amax_dev, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  DMRG_CHECK_HIP(hipStreamSynchronize(c->stream));
  DMRG_REQUIRE(hinfo == 0, "syevd/heevd did not converge");
  const double scale = hmax > 0.0 ? hmax : 1.0;
  svals.resize(rank);
  for (int i = 0; i < rank; ++i) svals[i] = scale * std::sqrt(std::max(evh[g - 1 - i], 0.0));
  SvdOut o;
  o.keep = truncate_rule(svals, cutoff, maxdim, mind, &o.discarded);
  const int keep = o.keep;
  if (!(svals[0] > 0.0 && svals[keep - 1] >= 1e-7 * svals[0]))
    return svd_full(c, field, M, m, n, cutoff, maxdim, mind, absorb, U, V, kcap, svals);
  DMRG_REQUIRE(keep <= kcap, "svd: keep exceeds output capacity");
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
    if (!fixed) return svd_gram(c, field, M, m, n, cutoff, maxdim, mind, absorb, U, V, kcap, svals, left ? 1 : 0);
  }
  return o;
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

}  // namespace dmrg
