// Householder tridiagonalisation of a real symmetric matrix (LAPACK dsytrd,
// lower storage conventions), two kernel launches per column.
//
// The dense symmetric eigensolver behind every Gram-route split spends most
// of its time reducing the matrix to tridiagonal form, one column at a time.
// rocSOLVER issues five small kernels per column there; this reduction needs
// two.  Cross-workgroup reductions never synchronise inside a kernel (on a
// multi-XCD device that costs a device-scope cache write-back per workgroup):
// every workgroup writes its partial sums, and every workgroup of the NEXT
// launch re-reduces them in the same fixed order (a few KB), so the stream
// order is the only synchronisation.
//
//   launch 1 (rows r >= j): alpha of column j-1 from the partials of the
//     previous launch 2; finish column j-1 of W (w + alpha v); update column
//     j of the panel-start matrix with the panel so far,
//       a = A[j][:] - V W[j,:]^T - W V[j,:]^T      (A symmetric: row j)
//     and partial sums of |a_{j+2:}|^2, W^T a, V^T a per workgroup;
//   launch 2 (rows r >= j+1): the Householder reflector (beta, tau, the
//     1/(a_{j+1} - beta) scale) and the corrections y = W^T v, z = V^T v
//     from launch 1's partials; p = A22 v with v built on the fly from a;
//     w = tau (p - V y - W z); partial sums of w^T v.
//
// After each panel of TRD_NB columns the trailing matrix is updated with two
// GEMMs, A22 -= V W^T + W V^T, over the full square (the matrix-vector
// products read whole rows).  Outputs match dsytrd(uplo = 'L') on the
// column-major view (the matrix is symmetric, so the row-major buffer is the
// same matrix): d, e, tau, and reflector j in row j, columns j+2.., ready
// for rocsolver_dstedc + rocsolver_dormtr.  All sums run in a fixed order
// (deterministic).
#include "dmrg_internal.h"

namespace dmrg {

constexpr int TRD_NB = 32;

struct TrdState {
  double alpha, scale, t, pad;
  double y[TRD_NB], z[TRD_NB];
};

constexpr int TRD_PS = 1 + 2 * TRD_NB;   // partials per workgroup in launch 1

__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// block-wide sum of one value per thread (256 threads); result valid in thread 0
__device__ __forceinline__ double block_sum(double v, double* sh) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[wave] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0) s = sh[0] + sh[1] + sh[2] + sh[3];
  return s;
}

// Sum of part[0..nb) with all 256 threads (strided partial sums, then a
// fixed butterfly and wave order); every thread gets the result.
__device__ __forceinline__ double sum_one(const double* part, int nb, double* red4) {
  double s = 0.0;
  for (int b = threadIdx.x; b < nb; b += 256) s += part[b];
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) red4[threadIdx.x >> 6] = s;
  __syncthreads();
  const double r = red4[0] + red4[1] + red4[2] + red4[3];
  __syncthreads();
  return r;
}

constexpr int TRD_ROWS1 = 64;   // rows per workgroup in launch 1
constexpr int TRD_ROWS2 = 4;    // rows per workgroup in launch 2

// Both launches issue every global load they need up front (all of them
// miss the caches: the previous launch wrote the data, possibly from another
// XCD), then reduce and compute; a launch costs about two memory round trips.
__global__ __launch_bounds__(256) void trd_col_k(const double* __restrict__ A, int n, int j, int jj,
                                                 const double* __restrict__ V, double* __restrict__ W,
                                                 const double* __restrict__ wpre, double* __restrict__ avec,
                                                 double* __restrict__ part1, const double* __restrict__ part2,
                                                 int g2prev, TrdState* __restrict__ st) {
  __shared__ double sV[TRD_ROWS1][TRD_NB + 1], sW[TRD_ROWS1][TRD_NB + 1];
  __shared__ double sWj[TRD_NB], sVj[TRD_NB], sh[8];
  __shared__ double sa[TRD_ROWS1];
  __shared__ double red4[256];
  const int tid = threadIdx.x;
  const int r0 = j + blockIdx.x * TRD_ROWS1;
  const int nr = max(0, min(TRD_ROWS1, n - r0));
  const int r = r0 + tid;
  // ---- loads
  const double tprev = st->t;
  double ps = 0.0;   // this thread's share of column j-1's w^T v partials
  if (jj > 0) {
#pragma unroll 4
    for (int b = tid; b < g2prev; b += 256) ps += part2[b];
  }
  // panel rows into registers first (all loads in flight at once), then LDS
  constexpr int SPT = TRD_ROWS1 * TRD_NB / 256;
  double lv[SPT], lw[SPT];
#pragma unroll
  for (int u = 0; u < SPT; ++u) {
    const int idx = tid + 256 * u, k = idx % TRD_NB;
    lv[u] = lw[u] = 0.0;
    if (idx < nr * TRD_NB && k < jj) {
      lv[u] = V[(size_t)r0 * TRD_NB + idx];
      lw[u] = W[(size_t)r0 * TRD_NB + idx];
    }
  }
  double vj = 0.0, wj = 0.0;
  if (tid < jj) {
    vj = V[(size_t)j * TRD_NB + tid];
    wj = tid == jj - 1 ? wpre[j] : W[(size_t)j * TRD_NB + tid];
  }
  double arow = 0.0, wpr = 0.0;
  if (tid < nr) {
    arow = A[(size_t)j * n + r];
    if (jj > 0) wpr = wpre[r];
  }
#pragma unroll
  for (int u = 0; u < SPT; ++u) {
    const int idx = tid + 256 * u;
    sV[idx / TRD_NB][idx % TRD_NB] = lv[u];
    sW[idx / TRD_NB][idx % TRD_NB] = lw[u];
  }
  // ---- alpha of column j-1 (fixed-order reduction)
  double alpha = 0.0;
  if (jj > 0) {
    ps = wave_sum(ps);
    if ((tid & 63) == 0) red4[tid >> 6] = ps;
  }
  __syncthreads();
  if (jj > 0) {
    alpha = -0.5 * tprev * (red4[0] + red4[1] + red4[2] + red4[3]);
    if (blockIdx.x == 0 && tid == 0) st->alpha = alpha;
  }
  if (tid < jj) {
    sVj[tid] = vj;
    sWj[tid] = tid == jj - 1 ? wj + alpha * vj : wj;
  }
  if (jj > 0 && tid < nr) {   // finish column jj-1 of W: w + alpha v
    const double w = wpr + alpha * sV[tid][jj - 1];
    sW[tid][jj - 1] = w;
    W[(size_t)r * TRD_NB + jj - 1] = w;
  }
  __syncthreads();
  // ---- column j of the panel-start matrix, updated with the panel so far
  double pn = 0.0;
  if (tid < TRD_ROWS1) sa[tid] = 0.0;
  if (tid < nr) {
    double a = arow;
    for (int k = 0; k < jj; ++k) a -= sV[tid][k] * sWj[k] + sW[tid][k] * sVj[k];
    avec[r] = a;
    if (r >= j + 2) {
      sa[tid] = a;
      pn = a * a;
    }
  }
  const double pns = block_sum(pn, sh);
  {  // partial W^T a (k < jj) and V^T a (jj <= k < 2 jj) over this block's rows
    const int k = tid & 63, grp = tid >> 6;
    double acc = 0.0;
    if (k < 2 * jj)
      for (int i = grp; i < nr; i += 4) acc += (k < jj ? sW[i][k] : sV[i][k - jj]) * sa[i];
    red4[grp * 64 + k] = acc;
  }
  __syncthreads();
  double* myp = part1 + (size_t)blockIdx.x * TRD_PS;
  if (tid == 0) myp[0] = pns;
  if (tid < 2 * jj) myp[1 + tid] = red4[tid] + red4[64 + tid] + red4[128 + tid] + red4[192 + tid];
}

// Reflector of column j (from launch 1's partials, in every workgroup), then
// p = A22 v for TRD_ROWS2 rows per workgroup: every thread walks columns
// c = j+1+tid, +256, ... for all rows at once, accumulating A a (v = e_{j+1}
// + scale a_{j+2:} is applied after the reduction, so the matrix stream does
// not wait for the reflector); the corrections V y + W z are one (row, k)
// product per thread of the first wave.
__global__ __launch_bounds__(256) void trd_symv_k(double* __restrict__ A, int n, int j, int jj,
                                                  double* __restrict__ V, const double* __restrict__ W,
                                                  double* __restrict__ wpre, const double* __restrict__ avec,
                                                  const double* __restrict__ part1, int g1,
                                                  double* __restrict__ part2, TrdState* __restrict__ st,
                                                  double* __restrict__ d, double* __restrict__ e,
                                                  double* __restrict__ tau) {
  __shared__ double sy[TRD_NB], sz[TRD_NB];
  __shared__ double red[4][TRD_ROWS2];
  __shared__ double red4[256];
  __shared__ double tot[TRD_PS];
  __shared__ double sst[2];   // t, scale
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int nk = 1 + 2 * jj;
  const int rbase = j + 1 + blockIdx.x * TRD_ROWS2;
  const int nr = min(TRD_ROWS2, n - rbase);   // >= 1
  // ---- loads: launch-1 partials, the reflector inputs, the correction rows
  double ps = 0.0;
  {
    const int k = tid & 63, grp = tid >> 6;
    if (k < nk) {
#pragma unroll 4
      for (int b = grp; b < g1; b += 4) ps += part1[(size_t)b * TRD_PS + k];
    }
  }
  double a0 = 0.0, aj = 0.0;
  if (tid == 0) {
    a0 = avec[j + 1];
    aj = avec[j];
  }
  double v1 = 0.0, w1 = 0.0;
  if (tid < jj) {
    v1 = V[(size_t)(j + 1) * TRD_NB + tid];
    w1 = tid == jj - 1 ? wpre[j + 1] + st->alpha * v1 : W[(size_t)(j + 1) * TRD_NB + tid];
  }
  double mrow[TRD_ROWS2];
#pragma unroll
  for (int i = 0; i < TRD_ROWS2; ++i) mrow[i] = 0.0;
  if (tid < 64 && (lane & 31) < jj) {
    const double* M = lane < 32 ? V : W;
#pragma unroll
    for (int i = 0; i < TRD_ROWS2; ++i)
      if (i < nr) mrow[i] = M[(size_t)(rbase + i) * TRD_NB + (lane & 31)];
  }
  // ---- the matrix stream: acc = sum_{c >= j+2} A[r][c] a[c], a1 = A[r][j+1]
  double acc[TRD_ROWS2], a1[TRD_ROWS2];
#pragma unroll
  for (int i = 0; i < TRD_ROWS2; ++i) acc[i] = a1[i] = 0.0;
#pragma unroll 4
  for (int c = j + 1 + tid; c < n; c += 256) {
    const bool first = c == j + 1;
    const double v = first ? 0.0 : avec[c];
#pragma unroll
    for (int i = 0; i < TRD_ROWS2; ++i) {
      const int r = rbase + min(i, nr - 1);   // rows past the end repeat the last (discarded)
      const double x = A[(size_t)r * n + c];
      acc[i] += x * v;
      if (first) a1[i] = x;
    }
  }
  // ---- reflector (every workgroup, same fixed-order sums)
  red4[tid] = ps;
  __syncthreads();
  if (tid < nk) tot[tid] = red4[tid] + red4[64 + tid] + red4[128 + tid] + red4[192 + tid];
  __syncthreads();
  if (tid == 0) {
    const double xn = sqrt(tot[0]);
    double t = 0.0, beta = a0, scale = 0.0;
    if (xn != 0.0) {
      beta = -copysign(hypot(a0, xn), a0);
      t = (beta - a0) / beta;
      scale = 1.0 / (a0 - beta);
    }
    sst[0] = t;
    sst[1] = scale;
    if (blockIdx.x == 0) {
      d[j] = aj;
      e[j] = beta;
      tau[j] = t;
      st->t = t;
    }
  }
  __syncthreads();
  const double t = sst[0], scale = sst[1];
  if (tid < jj) {
    sy[tid] = w1 + scale * tot[1 + tid];
    sz[tid] = v1 + scale * tot[1 + jj + tid];
  }
  __syncthreads();
  // ---- p - V y - W z per row
  const double f = tid < 64 && (lane & 31) < jj ? (lane < 32 ? sy[lane & 31] : sz[lane & 31]) : 0.0;
#pragma unroll
  for (int i = 0; i < TRD_ROWS2; ++i) {
    const double q = wave_sum(a1[i] + scale * acc[i] - mrow[i] * f);
    if (lane == 0) red[wave][i] = q;
  }
  __syncthreads();
  if (tid < TRD_ROWS2) {
    double pwv = 0.0;
    if (tid < nr) {
      const int r = rbase + tid;
      const double q = red[0][tid] + red[1][tid] + red[2][tid] + red[3][tid];
      const double wp = t * q;
      const double vr = r == j + 1 ? 1.0 : scale * avec[r];
      wpre[r] = wp;
      V[(size_t)r * TRD_NB + jj] = vr;
      if (r >= j + 2) A[(size_t)j * n + r] = vr;   // reflector storage (row j)
      pwv = wp * vr;
    }
    red4[tid] = pwv;
  }
  __syncthreads();
  if (tid == 0) {
    double s = 0.0;
    for (int i = 0; i < TRD_ROWS2; ++i) s += red4[i];
    part2[blockIdx.x] = s;
  }
}

// W[r][k] = wpre[r] + alpha V[r][k] for the panel's last column, rows >= s
__global__ __launch_bounds__(256) void trd_fin_k(int n, int s, int k, const double* __restrict__ V,
                                                 double* __restrict__ W, const double* __restrict__ wpre,
                                                 const double* __restrict__ part2, int g2,
                                                 const TrdState* __restrict__ st) {
  __shared__ double red4[256];
  const double alpha = -0.5 * st->t * sum_one(part2, g2, red4);
  const int r = s + blockIdx.x * 256 + threadIdx.x;
  if (r < n) W[(size_t)r * TRD_NB + k] = wpre[r] + alpha * V[(size_t)r * TRD_NB + k];
}

// A (n x n, symmetric, row-major = column-major) -> d[n], e[n-1], tau[n-1]
// and reflectors in place (dsytrd, uplo 'L').
void sytrd_lower(Ctx* c, double* A, int n, double* d, double* e, double* tau) {
  DMRG_REQUIRE(n >= 2, "sytrd: order must be at least 2");
  const size_t vw = (size_t)n * TRD_NB;
  const int gmax1 = (n + TRD_ROWS1 - 1) / TRD_ROWS1;
  const int gmax2 = (n + TRD_ROWS2 - 1) / TRD_ROWS2;
  double* buf = (double*)ws_get(c, WS_TRD, sizeof(double) * (2 * vw + 2 * (size_t)n + (size_t)gmax1 * TRD_PS +
                                                              (size_t)gmax2 + 8) +
                                               sizeof(TrdState) + 64);
  double* V = buf;
  double* W = V + vw;
  double* wpre = W + vw;
  double* avec = wpre + n;
  double* part1 = avec + n;
  double* part2 = part1 + (size_t)gmax1 * TRD_PS;
  TrdState* st = reinterpret_cast<TrdState*>(part2 + gmax2 + 8);
  DMRG_CHECK_HIP(hipMemsetAsync(st, 0, sizeof(TrdState), c->stream));
  // algorithmic bytes: the lower triangle of the trailing matrix, once per column
  double bytes = 0.0;
  for (int j = 0; j < n - 1; ++j) bytes += 8.0 * (double)(n - j - 1) * (n - j) / 2.0;
  ProfScope ps(c, 4, 0.0, bytes);
  int g2prev = 1;
  for (int j0 = 0; j0 < n - 1; j0 += TRD_NB) {
    const int nbp = std::min(TRD_NB, n - 1 - j0);
    for (int jj = 0; jj < nbp; ++jj) {
      const int j = j0 + jj;
      const int g1 = (n - j + TRD_ROWS1 - 1) / TRD_ROWS1;       // rows j .. n-1
      const int g2 = (n - j - 1 + TRD_ROWS2 - 1) / TRD_ROWS2;   // rows j+1 .. n-1
      DMRG_REQUIRE(g1 >= 1 && g1 <= gmax1 && g2 >= 1 && g2 <= gmax2, "sytrd: grid exceeds partial buffers");
      trd_col_k<<<g1, 256, 0, c->stream>>>(A, n, j, jj, V, W, wpre, avec, part1, part2, g2prev, st);
      DMRG_CHECK_HIP(hipGetLastError());
      trd_symv_k<<<g2, 256, 0, c->stream>>>(A, n, j, jj, V, W, wpre, avec, part1, g1, part2, st, d, e, tau);
      DMRG_CHECK_HIP(hipGetLastError());
      g2prev = g2;
    }
    const int s = j0 + nbp;
    if (s < n) {
      trd_fin_k<<<(n - s + 255) / 256, 256, 0, c->stream>>>(n, s, nbp - 1, V, W, wpre, part2, g2prev, st);
      DMRG_CHECK_HIP(hipGetLastError());
      const int m = n - s;
      GemmDesc g = gdesc(m, m, nbp, TRD_NB, TRD_NB, n, REAL, 4);
      g.transB = true;
      g.alpha = -1.0;
      g.beta = 1.0;
      double* C = A + (size_t)s * n + s;
      gemm_f64(c, g, V + (size_t)s * TRD_NB, W + (size_t)s * TRD_NB, C);
      gemm_f64(c, g, W + (size_t)s * TRD_NB, V + (size_t)s * TRD_NB, C);
    }
  }
  DMRG_CHECK_HIP(hipMemcpyAsync(d + (n - 1), A + (size_t)(n - 1) * n + (n - 1), sizeof(double),
                                hipMemcpyDeviceToDevice, c->stream));
}

}  // namespace dmrg
