// Krylov solvers on the device: Lanczos ground state and Lanczos matrix
// exponential, both with full re-orthogonalisation.  Vectors stay in HBM
// (FP64, planar complex); only the projected tridiagonal matrix and a handful
// of scalars per iteration come to the host, once per iteration.
//
// Vector work per iteration k (n elements per plane, sizeof = 8):
//   dot + axpy + block re-orthogonalisation (k dots, k axpys) + norm,
//   about (2k + 6) * n * 8 bytes of HBM traffic: bandwidth bound.
#include "dmrg_internal.h"
#include "host_eig.h"

#include <cfloat>
#include <cmath>
#include <complex>
#include <random>

namespace dmrg {

// ---------------------------------------------------------------------------
// device kernels
// ---------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

// part[j*nblk + blk] (re) and part[(k+j)*nblk + blk] (im): block-partial
// sums of conj(Q_j) . w over a contiguous chunk of the vector.
template <bool CX>
__global__ __launch_bounds__(256) void dot_block_k(const double* __restrict__ Q, long long ldq, int k,
                                                   const double* __restrict__ w, long long n,
                                                   double* __restrict__ part, int nblk) {
  __shared__ double red[2][4];
  const long long chunk = (n + nblk - 1) / nblk;
  const long long beg = (long long)blockIdx.x * chunk;
  const long long end = beg + chunk < n ? beg + chunk : n;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int j = 0; j < k; ++j) {
    const double* q = Q + (long long)j * ldq;
    double sr = 0, si = 0;
    for (long long i = beg + threadIdx.x; i < end; i += 256) {
      const double qr = q[i], wr = w[i];
      if (CX) {
        const double qi = q[i + n], wi = w[i + n];
        sr += qr * wr + qi * wi;
        si += qr * wi - qi * wr;
      } else {
        sr += qr * wr;
      }
    }
    sr = wave_sum(sr);
    if (CX) si = wave_sum(si);
    if (lane == 0) {
      red[0][wv] = sr;
      red[1][wv] = si;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      part[(long long)j * nblk + blockIdx.x] = red[0][0] + red[0][1] + red[0][2] + red[0][3];
      if (CX) part[(long long)(k + j) * nblk + blockIdx.x] = red[1][0] + red[1][1] + red[1][2] + red[1][3];
    }
    __syncthreads();
  }
}

// out[row] = sum_b part[row*nblk + b]  (fixed order: deterministic)
__global__ __launch_bounds__(256) void reduce_rows_k(const double* __restrict__ part, int nblk,
                                                     double* __restrict__ out) {
  __shared__ double red[4];
  const int row = blockIdx.x;
  double s = 0;
  for (int b = threadIdx.x; b < nblk; b += 256) s += part[(long long)row * nblk + b];
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) out[row] = red[0] + red[1] + red[2] + red[3];
}

// w -= sum_j coef_j Q_j (coef complex when CX: re[0..k), im[k..2k))
template <bool CX>
__global__ __launch_bounds__(256) void sub_block_k(const double* __restrict__ Q, long long ldq, int k,
                                                   const double* __restrict__ coef, double* __restrict__ w,
                                                   long long n) {
  long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  const long long stride = (long long)gridDim.x * 256;
  for (; i < n; i += stride) {
    double ar = w[i], ai = CX ? w[i + n] : 0.0;
    for (int j = 0; j < k; ++j) {
      const double* q = Q + (long long)j * ldq;
      const double cr = coef[j];
      const double qr = q[i];
      if (CX) {
        const double ci = coef[k + j], qi = q[i + n];
        ar -= cr * qr - ci * qi;
        ai -= cr * qi + ci * qr;
      } else {
        ar -= cr * qr;
      }
    }
    w[i] = ar;
    if (CX) w[i + n] = ai;
  }
}

// x = scale * sum_j coef_j Q_j, coef real (COEF_CX false) or complex
template <bool CX, bool COEF_CX>
__global__ __launch_bounds__(256) void combine_k(const double* __restrict__ Q, long long ldq, int k,
                                                 const double* __restrict__ coef, double scale,
                                                 double* __restrict__ x, long long n) {
  long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  const long long stride = (long long)gridDim.x * 256;
  for (; i < n; i += stride) {
    double ar = 0, ai = 0;
    for (int j = 0; j < k; ++j) {
      const double* q = Q + (long long)j * ldq;
      const double cr = coef[j], ci = COEF_CX ? coef[k + j] : 0.0;
      const double qr = q[i], qi = CX ? q[i + n] : 0.0;
      ar += cr * qr - ci * qi;
      ai += cr * qi + ci * qr;
    }
    x[i] = scale * ar;
    if (CX) x[i + n] = scale * ai;
  }
}

__global__ void scale_copy_k(const double* __restrict__ src, double* __restrict__ dst, double s, long long n) {
  long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  const long long stride = (long long)gridDim.x * 256;
  for (; i < n; i += stride) dst[i] = s * src[i];
}

static unsigned vgrid(long long n) {
  long long b = (n + 255) / 256;
  if (b > 2048) b = 2048;
  if (b < 1) b = 1;
  return (unsigned)b;
}

static int dot_blocks(long long n) {
  long long b = (n + 2047) / 2048;
  if (b > 1024) b = 1024;
  if (b < 1) b = 1;
  return (int)b;
}

void vec_dot_block(Ctx* c, int field, const double* Q, long long ldq, int k, const double* w, long long n,
                   double* out) {
  if (k <= 0) return;
  // bytes: the k basis vectors and w read once
  ProfScope ps(c, 3, 0.0, 8.0 * (field == CPLX ? 2 : 1) * (double)(k + 1) * (double)n);
  const int nb = dot_blocks(n);
  const int rows = (field == CPLX ? 2 : 1) * k;
  double* part = (double*)ws_get(c, WS_KRYLOV_RED, sizeof(double) * (size_t)rows * nb);
  if (field == CPLX) dot_block_k<true><<<nb, 256, 0, c->stream>>>(Q, ldq, k, w, n, part, nb);
  else dot_block_k<false><<<nb, 256, 0, c->stream>>>(Q, ldq, k, w, n, part, nb);
  DMRG_CHECK_HIP(hipGetLastError());
  reduce_rows_k<<<rows, 256, 0, c->stream>>>(part, nb, out);
  DMRG_CHECK_HIP(hipGetLastError());
}

void vec_sub_block(Ctx* c, int field, const double* Q, long long ldq, int k, const double* coef, double* w,
                   long long n) {
  if (k <= 0) return;
  // bytes: the k basis vectors read once, w read and written
  ProfScope ps(c, 3, 0.0, 8.0 * (field == CPLX ? 2 : 1) * (double)(k + 2) * (double)n);
  if (field == CPLX) sub_block_k<true><<<vgrid(n), 256, 0, c->stream>>>(Q, ldq, k, coef, w, n);
  else sub_block_k<false><<<vgrid(n), 256, 0, c->stream>>>(Q, ldq, k, coef, w, n);
  DMRG_CHECK_HIP(hipGetLastError());
}

void vec_combine(Ctx* c, int field, const double* Q, long long ldq, int k, const double* coef, int coef_cplx,
                 double scale, double* x, long long n) {
  DMRG_REQUIRE(!(coef_cplx && field == REAL), "combine: complex coefficients need a complex vector");
  if (field == CPLX) {
    if (coef_cplx) combine_k<true, true><<<vgrid(n), 256, 0, c->stream>>>(Q, ldq, k, coef, scale, x, n);
    else combine_k<true, false><<<vgrid(n), 256, 0, c->stream>>>(Q, ldq, k, coef, scale, x, n);
  } else {
    combine_k<false, false><<<vgrid(n), 256, 0, c->stream>>>(Q, ldq, k, coef, scale, x, n);
  }
  DMRG_CHECK_HIP(hipGetLastError());
}

static void scale_copy(Ctx* c, const double* src, double* dst, double s, long long ntot) {
  scale_copy_k<<<vgrid(ntot), 256, 0, c->stream>>>(src, dst, s, ntot);
  DMRG_CHECK_HIP(hipGetLastError());
}

// Re <a|b> with a host round trip (used outside the hot loop)
double vec_dot_real(Ctx* c, int field, const double* a, const double* b, long long n) {
  double* out = (double*)ws_get(c, WS_KRYLOV_COEF, sizeof(double) * 4);
  vec_dot_block(c, field, a, 0, 1, b, n, out);
  DMRG_CHECK_HIP(hipMemcpyAsync(c->host_scalars, out, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  DMRG_CHECK_HIP(hipStreamSynchronize(c->stream));
  return c->host_scalars[0];
}

// The host-side small eigenproblems (tridiag_eig, jacobi_eig,
// tridiag_expm_col0) are in host_eig.h.

// ---------------------------------------------------------------------------
// Lanczos ground state
// ---------------------------------------------------------------------------
static double lowest_tridiag(const std::vector<double>& al, const std::vector<double>& be) {
  std::vector<double> d(al);
  tridiag_eig(d, be, nullptr);
  return d[0];
}

static void dense_ground(Ctx* c, const HeffOp& h, double* x, double* energy) {
  const int field = h.d.field;
  const long long n = h.n;
  const int np = field == CPLX ? 2 : 1;
  const int dim = (int)n;
  double* e = (double*)ws_get(c, WS_KRYLOV_W, sizeof(double) * n * np * 2);
  std::vector<double> hcol((size_t)n * np), hm((size_t)dim * np * dim * np, 0.0), host_e(n * np);
  const int N = dim * np;  // real embedding size
  for (int col = 0; col < N; ++col) {
    std::fill(host_e.begin(), host_e.end(), 0.0);
    host_e[col] = 1.0;  // unit vector (real or imaginary part)
    DMRG_CHECK_HIP(hipMemcpyAsync(e, host_e.data(), sizeof(double) * n * np, hipMemcpyHostToDevice, c->stream));
    heff_apply(c, h, e, e + n * np);
    DMRG_CHECK_HIP(hipMemcpyAsync(hcol.data(), e + n * np, sizeof(double) * n * np, hipMemcpyDeviceToHost,
                                  c->stream));
    DMRG_CHECK_HIP(hipStreamSynchronize(c->stream));
    for (int row = 0; row < N; ++row) hm[(size_t)row * N + col] = hcol[row];
  }
  // symmetrise ((H + H^dag)/2 of the reference is the symmetric part of the
  // real embedding)
  for (int i = 0; i < N; ++i)
    for (int j = i + 1; j < N; ++j) {
      const double s = 0.5 * (hm[(size_t)i * N + j] + hm[(size_t)j * N + i]);
      hm[(size_t)i * N + j] = hm[(size_t)j * N + i] = s;
    }
  std::vector<double> w, v;
  jacobi_eig(hm, N, w, v);
  int k = 0;
  for (int i = 1; i < N; ++i)
    if (w[i] < w[k]) k = i;
  std::vector<double> xv(N);
  for (int i = 0; i < N; ++i) xv[i] = v[(size_t)i * N + k];
  DMRG_CHECK_HIP(hipMemcpyAsync(x, xv.data(), sizeof(double) * N, hipMemcpyHostToDevice, c->stream));
  DMRG_CHECK_HIP(hipStreamSynchronize(c->stream));
  *energy = w[k];
}

void lanczos_ground(Ctx* c, const HeffOp& h, double* x, int maxiter, double tol, double res_tol, double* energy,
                    int* iters) {
  const int field = h.d.field;
  const int np = field == CPLX ? 2 : 1;
  const long long n = h.n, ntot = n * np;
  if (n <= 3) {
    dense_ground(c, h, x, energy);
    if (iters) *iters = 0;
    return;
  }
  const int m = (int)std::min<long long>(maxiter, n);
  DMRG_REQUIRE(m >= 1, "lanczos: maxiter must be >= 1");
  const long long ldq = ntot;
  double* Q = (double*)ws_get(c, WS_KRYLOV_Q, sizeof(double) * (size_t)ldq * m);
  double* w = (double*)ws_get(c, WS_KRYLOV_W, sizeof(double) * ntot);
  // coefficient scratch: [0..2m) reorth coefficients, then scalars
  double* coef = (double*)ws_get(c, WS_KRYLOV_COEF, sizeof(double) * (4 * (size_t)m + 16));
  double* scal = coef + 4 * (size_t)m;  // scal[0] beta_prev, scal[1] alpha(re), scal[2] alpha(im), scal[4] |w|^2
  double* hs = c->host_scalars;

  // beta0 = |x|
  vec_dot_block(c, field, x, 0, 1, x, n, scal + 4);
  DMRG_CHECK_HIP(hipMemcpyAsync(hs, scal + 4, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  DMRG_CHECK_HIP(hipStreamSynchronize(c->stream));
  double beta0 = std::sqrt(std::max(hs[0], 0.0));
  if (beta0 == 0.0) {
    // zero start vector: seeded Gaussian restart (dmrg.py:102-105)
    std::mt19937_64 gen(0);
    std::normal_distribution<double> nd;
    std::vector<double> r(ntot, 0.0);
    for (long long i = 0; i < n; ++i) r[i] = nd(gen);
    DMRG_CHECK_HIP(hipMemcpyAsync(x, r.data(), sizeof(double) * ntot, hipMemcpyHostToDevice, c->stream));
    DMRG_CHECK_HIP(hipStreamSynchronize(c->stream));
    double s2 = 0;
    for (long long i = 0; i < n; ++i) s2 += r[i] * r[i];
    beta0 = std::sqrt(s2);
  }
  scale_copy(c, x, Q, 1.0 / beta0, ntot);

  std::vector<double> al, be;
  // first step
  heff_apply(c, h, Q, w);
  vec_dot_block(c, field, Q, 0, 1, w, n, scal + 1);  // alpha (re, im)
  if (field == CPLX) {
    // only the real part of alpha enters (dmrg.py:109): zero the imaginary
    DMRG_CHECK_HIP(hipMemsetAsync(scal + 2, 0, sizeof(double), c->stream));
  }
  vec_sub_block(c, field, Q, ldq, 1, scal + 1, w, n);
  vec_dot_block(c, field, w, 0, 1, w, n, scal + 4);
  DMRG_CHECK_HIP(hipMemcpyAsync(hs, scal, sizeof(double) * 5, hipMemcpyDeviceToHost, c->stream));
  DMRG_CHECK_HIP(hipStreamSynchronize(c->stream));
  al.push_back(hs[1]);
  double beta = std::sqrt(std::max(hs[4], 0.0));
  double prev = al[0];
  int k = 1;
  for (int it = 1; it < m; ++it) {
    if (beta < tol) break;
    be.push_back(beta);
    double* q = Q + (long long)it * ldq;
    scale_copy(c, w, q, 1.0 / beta, ntot);
    heff_apply(c, h, q, w);
    // alpha; then w -= alpha q + beta q_prev  (coefficient pair [beta, alpha])
    vec_dot_block(c, field, q, 0, 1, w, n, scal + 1);
    if (field == CPLX) DMRG_CHECK_HIP(hipMemsetAsync(scal + 2, 0, sizeof(double), c->stream));
    hs[0] = beta;
    DMRG_CHECK_HIP(hipMemcpyAsync(scal, hs, sizeof(double), hipMemcpyHostToDevice, c->stream));
    {
      // coefficient layout for sub_block with k=2 over (q_prev, q): re [beta, alpha], im [0, 0]
      double* pair = coef + 4 * (size_t)m - 4;  // 4 doubles of scratch at the end of the coef block
      DMRG_CHECK_HIP(hipMemcpyAsync(pair, scal, sizeof(double), hipMemcpyDeviceToDevice, c->stream));
      DMRG_CHECK_HIP(hipMemcpyAsync(pair + 1, scal + 1, sizeof(double), hipMemcpyDeviceToDevice, c->stream));
      DMRG_CHECK_HIP(hipMemsetAsync(pair + 2, 0, 2 * sizeof(double), c->stream));
      vec_sub_block(c, field, Q + (long long)(it - 1) * ldq, ldq, 2, pair, w, n);
    }
    // full re-orthogonalisation against q_0..q_{it-1} (dmrg.py:128-129)
    vec_dot_block(c, field, Q, ldq, it, w, n, coef);
    vec_sub_block(c, field, Q, ldq, it, coef, w, n);
    vec_dot_block(c, field, w, 0, 1, w, n, scal + 4);
    DMRG_CHECK_HIP(hipMemcpyAsync(hs + 8, scal, sizeof(double) * 5, hipMemcpyDeviceToHost, c->stream));
    DMRG_CHECK_HIP(hipStreamSynchronize(c->stream));
    al.push_back(hs[8 + 1]);
    beta = std::sqrt(std::max(hs[8 + 4], 0.0));
    k = it + 1;
    const double cur = lowest_tridiag(al, be);
    if (std::fabs(cur - prev) < tol * std::max(1.0, std::fabs(cur))) break;
    if (res_tol > 0.0) {
      // Ritz residual |H y - theta y| = beta_k |last component of the
      // lowest tridiagonal eigenvector|: the stop for reduced-precision
      // operators, whose Ritz values stall at the rounding floor
      std::vector<double> dd(al), zz;
      tridiag_eig(dd, be, &zz);
      const double res = beta * std::fabs(zz[(size_t)(k - 1) * k + 0]);
      if (res <= res_tol * std::max(1.0, std::fabs(cur))) break;
    }
    prev = cur;
  }
  // Ritz vector x = Q v
  std::vector<double> d(al), z;
  std::vector<double> ee(be.begin(), be.begin() + (k - 1));
  d.resize(k);
  tridiag_eig(d, ee, &z);
  std::vector<double> v(k);
  for (int i = 0; i < k; ++i) v[i] = z[(size_t)i * k + 0];
  DMRG_CHECK_HIP(hipMemcpyAsync(coef, v.data(), sizeof(double) * k, hipMemcpyHostToDevice, c->stream));
  vec_combine(c, field, Q, ldq, k, coef, 0, 1.0, x, n);
  DMRG_CHECK_HIP(hipStreamSynchronize(c->stream));
  *energy = d[0];
  if (iters) *iters = k;
}

double rayleigh(Ctx* c, const HeffOp& h, const double* x) {
  const int np = h.d.field == CPLX ? 2 : 1;
  double* y = (double*)ws_get(c, WS_REFINE, sizeof(double) * h.n * np);
  heff_apply(c, h, x, y);
  const double xy = vec_dot_real(c, h.d.field, x, y, h.n);
  const double xx = vec_dot_real(c, h.d.field, x, x, h.n);
  return xy / xx;
}

// ---------------------------------------------------------------------------
// Krylov exponential
// ---------------------------------------------------------------------------
void krylov_expm(Ctx* c, const HeffOp& h, double* x, std::complex<double> coeff, int maxiter, double tol,
                 double errgoal, int* iters) {
  const int field = h.d.field;
  const int np = field == CPLX ? 2 : 1;
  const long long n = h.n, ntot = n * np;
  DMRG_REQUIRE(field == CPLX || coeff.imag() == 0.0, "expm: complex coefficient needs a complex vector");
  const int m = (int)std::min<long long>(maxiter, n);
  DMRG_REQUIRE(m >= 1, "expm: maxiter must be >= 1");
  const long long ldq = ntot;
  double* Q = (double*)ws_get(c, WS_KRYLOV_Q, sizeof(double) * (size_t)ldq * m);
  double* w = (double*)ws_get(c, WS_KRYLOV_W, sizeof(double) * ntot);
  double* coef = (double*)ws_get(c, WS_KRYLOV_COEF, sizeof(double) * (4 * (size_t)m + 16));
  double* scal = coef + 4 * (size_t)m;
  double* hs = c->host_scalars;

  vec_dot_block(c, field, x, 0, 1, x, n, scal + 4);
  DMRG_CHECK_HIP(hipMemcpyAsync(hs, scal + 4, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  DMRG_CHECK_HIP(hipStreamSynchronize(c->stream));
  const double beta0 = std::sqrt(std::max(hs[0], 0.0));
  if (beta0 == 0.0) {
    if (iters) *iters = 0;
    return;
  }
  scale_copy(c, x, Q, 1.0 / beta0, ntot);
  std::vector<double> al, be;
  double beta = 0.0;
  for (int it = 0; it < m; ++it) {
    double* q = Q + (long long)it * ldq;
    heff_apply(c, h, q, w);
    vec_dot_block(c, field, q, 0, 1, w, n, scal + 1);
    if (field == CPLX) DMRG_CHECK_HIP(hipMemsetAsync(scal + 2, 0, sizeof(double), c->stream));
    hs[0] = beta;
    DMRG_CHECK_HIP(hipMemcpyAsync(scal, hs, sizeof(double), hipMemcpyHostToDevice, c->stream));
    double* pair = coef + 4 * (size_t)m - 4;
    if (it > 0) {
      DMRG_CHECK_HIP(hipMemcpyAsync(pair, scal, sizeof(double), hipMemcpyDeviceToDevice, c->stream));
      DMRG_CHECK_HIP(hipMemcpyAsync(pair + 1, scal + 1, sizeof(double), hipMemcpyDeviceToDevice, c->stream));
      DMRG_CHECK_HIP(hipMemsetAsync(pair + 2, 0, 2 * sizeof(double), c->stream));
      vec_sub_block(c, field, Q + (long long)(it - 1) * ldq, ldq, 2, pair, w, n);
    } else {
      vec_sub_block(c, field, q, ldq, 1, scal + 1, w, n);
    }
    // w -= Q_k (Q_k^dag w) over every basis vector so far (tdvp.py:138-139)
    vec_dot_block(c, field, Q, ldq, it + 1, w, n, coef);
    vec_sub_block(c, field, Q, ldq, it + 1, coef, w, n);
    vec_dot_block(c, field, w, 0, 1, w, n, scal + 4);
    DMRG_CHECK_HIP(hipMemcpyAsync(hs + 8, scal, sizeof(double) * 5, hipMemcpyDeviceToHost, c->stream));
    DMRG_CHECK_HIP(hipStreamSynchronize(c->stream));
    al.push_back(hs[8 + 1]);
    beta = std::sqrt(std::max(hs[8 + 4], 0.0));
    // first column of exp(coeff T_k)
    const int k = it + 1;
    const std::vector<std::complex<double>> col = tridiag_expm_col0(al, be, coeff);
    if (beta < tol || it == m - 1 || beta0 * beta * std::abs(col[k - 1]) < errgoal) {
      std::vector<double> cc(2 * (size_t)k);
      for (int r = 0; r < k; ++r) {
        cc[r] = col[r].real();
        cc[k + r] = col[r].imag();
      }
      DMRG_CHECK_HIP(hipMemcpyAsync(coef, cc.data(), sizeof(double) * 2 * k, hipMemcpyHostToDevice, c->stream));
      vec_combine(c, field, Q, ldq, k, coef, field == CPLX ? 1 : 0, beta0, x, n);
      DMRG_CHECK_HIP(hipStreamSynchronize(c->stream));
      if (iters) *iters = k;
      return;
    }
    be.push_back(beta);
    scale_copy(c, w, Q + (long long)(it + 1) * ldq, 1.0 / beta, ntot);
  }
  throw Error("expm: unreachable");
}

}  // namespace dmrg
