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
// After each panel of TRD_NB columns the trailing matrix is updated with one
// GEMM, A22 -= [V W] [W V]^T = V W^T + W V^T, over the full square (the
// matrix-vector products read whole rows).  Outputs match dsytrd(uplo = 'L')
// on the column-major view (the matrix is symmetric, so the row-major buffer
// is the same matrix): d, e, tau, and reflector j in row j, columns j+2..,
// ready for rocsolver_dstedc and the back-transformation below.  All sums
// run in a fixed order (deterministic).
#include "dmrg_internal.h"

namespace dmrg {

constexpr int TRD_NB = 32;

struct TrdState {
  double alpha, scale, t, pad;
  double y[TRD_NB], z[TRD_NB];
};

constexpr int TRD_PS = 1 + 2 * TRD_NB;   // partials per workgroup in launch 1
// The panel's V and W columns are stored per row as [V | W | V] (row pitch
// TRD_P): one GEMM with inner dimension 2 TRD_NB then applies both halves of
// the trailing update, A22 -= [V W] [W V]^T, reading and writing A22 once.
constexpr int TRD_P = 3 * TRD_NB;

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

// Unroll depth of the loops that stream the trailing matrix (launch 2) and
// the previous launch's partial sums: each unrolled batch keeps that many
// independent loads in flight per thread.  Deeper unrolling was measured no
// faster (sym_eig at order 3000: 61.8 / 63.2 / 67.8 ms with 4 / 8 / 16; 4
// rows per workgroup in launch 2 instead of 2: 63.8 ms)
#ifndef TRD_UNROLL_STREAM
#define TRD_UNROLL_STREAM 4
#endif
#ifndef TRD_UNROLL_PART
#define TRD_UNROLL_PART 4
#endif
#define TRD_STR(x) #x
#define TRD_UNROLL(n) _Pragma(TRD_STR(unroll n))

constexpr int TRD_ROWS1 = 64;   // rows per workgroup in launch 1 (32 / 64 / 128 measured; 64 best at orders >= 2000)
#ifndef TRD_ROWS2_CFG
#define TRD_ROWS2_CFG 2
#endif
constexpr int TRD_ROWS2 = TRD_ROWS2_CFG;    // rows per workgroup in launch 2 (measured best of 1, 2, 4, 8, 16)

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
TRD_UNROLL(TRD_UNROLL_PART)
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
      lv[u] = V[(size_t)(r0 + idx / TRD_NB) * TRD_P + k];
      lw[u] = W[(size_t)(r0 + idx / TRD_NB) * TRD_P + k];
    }
  }
  double vj = 0.0, wj = 0.0;
  if (tid < jj) {
    vj = V[(size_t)j * TRD_P + tid];
    wj = tid == jj - 1 ? wpre[j] : W[(size_t)j * TRD_P + tid];
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
    W[(size_t)r * TRD_P + jj - 1] = w;
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
TRD_UNROLL(TRD_UNROLL_PART)
      for (int b = grp; b < g1; b += 4) ps += part1[(size_t)b * TRD_PS + k];
    }
  }
  double a0 = 0.0, aj = 0.0;
  if (tid == 0) {
    a0 = avec[j + 1];
    aj = avec[j];
  }
  // row j+1 of V and W (column jj-1 of W was finished by launch 1, which
  // covers rows >= j; wpre[j+1] must not be read here: workgroup 0 of this
  // launch overwrites it with the new column)
  double v1 = 0.0, w1 = 0.0;
  if (tid < jj) {
    v1 = V[(size_t)(j + 1) * TRD_P + tid];
    w1 = W[(size_t)(j + 1) * TRD_P + tid];
  }
  double mrow[TRD_ROWS2];
#pragma unroll
  for (int i = 0; i < TRD_ROWS2; ++i) mrow[i] = 0.0;
  if (tid < 64 && (lane & 31) < jj) {
    const double* M = lane < 32 ? V : W;
#pragma unroll
    for (int i = 0; i < TRD_ROWS2; ++i)
      if (i < nr) mrow[i] = M[(size_t)(rbase + i) * TRD_P + (lane & 31)];
  }
  // ---- the matrix stream: acc = sum_{c >= j+2} A[r][c] a[c], a1 = A[r][j+1]
  double acc[TRD_ROWS2], a1[TRD_ROWS2];
#pragma unroll
  for (int i = 0; i < TRD_ROWS2; ++i) acc[i] = a1[i] = 0.0;
TRD_UNROLL(TRD_UNROLL_STREAM)
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
      V[(size_t)r * TRD_P + jj] = vr;
      V[(size_t)r * TRD_P + 2 * TRD_NB + jj] = vr;   // the copy after W
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
  if (r < n) W[(size_t)r * TRD_P + k] = wpre[r] + alpha * V[(size_t)r * TRD_P + k];
}

// Sum of squares of the trailing block A[s:, s:] (pitch n): per-block
// partial sums over whole rows, then one fixed-order sum (deterministic).
constexpr int TRD_SQ_BLOCKS = 512;
__global__ __launch_bounds__(256) void trd_sumsq_k(const double* __restrict__ A, int n, int s, int planes,
                                                   double* __restrict__ part) {
  __shared__ double red4[4];
  double acc = 0.0;
  for (int pl = 0; pl < planes; ++pl)   // planar complex: both planes (stride n*n)
    for (int r = s + blockIdx.x; r < n; r += gridDim.x) {
      const double* row = A + (size_t)pl * n * n + (size_t)r * n;
      for (int cc = s + threadIdx.x; cc < n; cc += 256) {
        const double x = row[cc];
        acc += x * x;
      }
    }
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) red4[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = red4[0] + red4[1] + red4[2] + red4[3];
}

__global__ __launch_bounds__(256) void trd_sumpart_k(const double* __restrict__ part, int nb,
                                                     double* __restrict__ out) {
  __shared__ double red4[256];
  const double v = sum_one(part, nb, red4);
  if (threadIdx.x == 0) *out = v;
}

// d[s..n-1] = 0, e[s..n-2] = 0, tau[s..n-2] = 0 (tstride doubles per tau:
// 2 for interleaved complex): the tridiagonal of a reduction stopped after s
// columns (the trailing block taken as zero)
__global__ void trd_clear_tail_k(int n, int s, double* __restrict__ d, double* __restrict__ e,
                                 double* __restrict__ tau, int tstride) {
  for (int i = s + blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    d[i] = 0.0;
    if (i < n - 1) {
      e[i] = 0.0;
      for (int q = 0; q < tstride; ++q) tau[(size_t)i * tstride + q] = 0.0;
    }
  }
}

// trace of the trailing block, sum_{r >= s} A[r][r] (real plane), one
// workgroup, fixed order: the weight a stopped reduction drops, which the
// keep rule adds back to the discarded weight
__global__ __launch_bounds__(256) void trd_trace_k(const double* __restrict__ A, int n, int s,
                                                   double* __restrict__ out) {
  __shared__ double red4[4];
  double acc = 0.0;
  for (int r = s + threadIdx.x; r < n; r += 256) acc += A[(size_t)r * n + r];
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) red4[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) *out = red4[0] + red4[1] + red4[2] + red4[3];
}

static double trailing_trace(Ctx* c, const double* A, int n, int s, double* out) {
  trd_trace_k<<<1, 256, 0, c->stream>>>(A, n, s, out);
  DMRG_CHECK_HIP(hipGetLastError());
  DMRG_CHECK_HIP(hipMemcpyAsync(c->host_scalars + 34, out, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  DMRG_CHECK_HIP(hipStreamSynchronize(c->stream));
  return std::max(0.0, c->host_scalars[34]);
}

// |A[s:, s:]|_F^2 into host_scalars[32 + slot] (pinned), with event
// trd_ev[slot] recorded after the copy: the caller waits on the event, so a
// check can stay in flight while the next panel is queued
static void trailing_sumsq_async(Ctx* c, const double* A, int n, int s, int planes, double* part, int cap,
                                 double* out, int slot) {
  const int nb = std::max(1, std::min(std::min(TRD_SQ_BLOCKS, n - s), cap));
  trd_sumsq_k<<<nb, 256, 0, c->stream>>>(A, n, s, planes, part);
  DMRG_CHECK_HIP(hipGetLastError());
  trd_sumpart_k<<<1, 256, 0, c->stream>>>(part, nb, out + slot);
  DMRG_CHECK_HIP(hipGetLastError());
  DMRG_CHECK_HIP(hipMemcpyAsync(c->host_scalars + 32 + slot, out + slot, sizeof(double), hipMemcpyDeviceToHost,
                                c->stream));
  if (!c->trd_ev[slot]) DMRG_CHECK_HIP(hipEventCreateWithFlags(&c->trd_ev[slot], hipEventDisableTiming));
  DMRG_CHECK_HIP(hipEventRecord(c->trd_ev[slot], c->stream));
}

static double trailing_sumsq_wait(Ctx* c, int slot) {
  DMRG_CHECK_HIP(hipEventSynchronize(c->trd_ev[slot]));
  return c->host_scalars[32 + slot];
}

// A (n x n, symmetric, row-major = column-major) -> d[n], e[n-1], tau[n-1]
// and reflectors in place (dsytrd, uplo 'L').  Returns the number of
// reflectors computed: n - 1, or, with stop_tol > 0, possibly fewer -- after
// each panel (from column TRD_STOP_MIN on) the Frobenius norm of the
// trailing block is compared with stop_tol |A|_F, and once it is below, the
// block is taken as zero: the remaining d, e and tau are set to 0 and the
// tridiagonal is complete after s columns.  The host reads a panel's check
// while the next panel is already queued (no idle GPU between panels), so
// the reduction stops one panel after the first panel whose trailing block
// passed; the trailing norm never grows (each step is an orthogonal
// similarity of the trailing block, of which the next one is a principal
// submatrix), so the later block passes too.  That perturbs A by at most
// stop_tol |A|_F (every eigenvalue by at most as much); the caller's
// stop_tol sits below the rounding of forming a Gram matrix in FP64.  The
// dense eigenproblem then shrinks to order s + 1 plus zeros, and the trace
// of the dropped block (its eigenvalues' sum) is returned in *dropped.
constexpr int TRD_STOP_MIN = 64;
int sytrd_lower(Ctx* c, double* A, int n, double* d, double* e, double* tau, double stop_tol, double* dropped) {
  DMRG_REQUIRE(n >= 2, "sytrd: order must be at least 2");
  const size_t vwv = (size_t)n * TRD_P;   // the [V | W | V] panel store
  const int gmax1 = (n + TRD_ROWS1 - 1) / TRD_ROWS1;
  const int gmax2 = (n + TRD_ROWS2 - 1) / TRD_ROWS2;
  double* buf = (double*)ws_get(c, WS_TRD, sizeof(double) * (vwv + 2 * (size_t)n + (size_t)gmax1 * TRD_PS +
                                                              (size_t)gmax2 + 8) +
                                               sizeof(TrdState) + 64);
  double* V = buf;            // row r: V at [0, NB), W at [NB, 2 NB), V again at [2 NB, 3 NB)
  double* W = V + TRD_NB;
  double* wpre = V + vwv;
  double* avec = wpre + n;
  double* part1 = avec + n;
  double* part2 = part1 + (size_t)gmax1 * TRD_PS;
  TrdState* st = reinterpret_cast<TrdState*>(part2 + gmax2 + 8);
  DMRG_CHECK_HIP(hipMemsetAsync(st, 0, sizeof(TrdState), c->stream));
  ProfScope ps(c, 4, 0.0, 0.0);
  // early stop: |A|_F of the input, and scratch for the trailing sums (the
  // launch-1 partial buffer is free between panels)
  const bool early = stop_tol > 0.0 && n > TRD_STOP_MIN + TRD_NB;
  double* sq_out = part2 + gmax2;    // 8 spare doubles before the state (two used)
  double stop2 = 0.0;
  const int sq_cap = gmax1 * TRD_PS;   // partial sums the launch-1 buffer holds
  if (early) {
    trailing_sumsq_async(c, A, n, 0, 1, part1, sq_cap, sq_out, 0);
    stop2 = stop_tol * stop_tol * trailing_sumsq_wait(c, 0);
  }
  int done = n - 1;   // reflectors computed
  int pend = -1;      // slot of a trailing-norm check in flight
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
      double* C = A + (size_t)s * n + s;
      const double* Vs = V + (size_t)s * TRD_P;
      if (nbp == TRD_NB) {   // A22 -= [V W] [W V]^T in one pass over A22
        GemmDesc g = gdesc(m, m, 2 * TRD_NB, TRD_P, TRD_P, n, REAL, 4);
        g.transB = true;
        g.alpha = -1.0;
        g.beta = 1.0;
        gemm_f64(c, g, Vs, Vs + TRD_NB, C);
      } else {   // a short last panel: columns past nbp still hold the previous panel's values
        GemmDesc g = gdesc(m, m, nbp, TRD_P, TRD_P, n, REAL, 4);
        g.transB = true;
        g.alpha = -1.0;
        g.beta = 1.0;
        gemm_f64(c, g, Vs, Vs + TRD_NB, C);
        gemm_f64(c, g, Vs + TRD_NB, Vs, C);
      }
      int slot = -1;
      if (early && nbp == TRD_NB && s >= TRD_STOP_MIN && s < n - 1) {
        slot = pend == 0 ? 1 : 0;
        trailing_sumsq_async(c, A, n, s, 1, part1, sq_cap, sq_out, slot);
      }
      // the previous panel's check: its block passed => this one (already
      // queued and applied) passes too; stop here
      if (pend >= 0 && trailing_sumsq_wait(c, pend) <= stop2 && s < n - 1) {
        if (dropped) *dropped = trailing_trace(c, A, n, s, sq_out + 2);
        trd_clear_tail_k<<<(n - s + 255) / 256, 256, 0, c->stream>>>(n, s, d, e, tau, 1);
        DMRG_CHECK_HIP(hipGetLastError());
        done = s;
        break;
      }
      pend = slot;
    }
  }
  if (done == n - 1)
    DMRG_CHECK_HIP(hipMemcpyAsync(d + (n - 1), A + (size_t)(n - 1) * n + (n - 1), sizeof(double),
                                  hipMemcpyDeviceToDevice, c->stream));
  // algorithmic bytes: the lower triangle of the trailing matrix, once per reduced column
  for (int j = 0; j < done; ++j) ps.bytes += 8.0 * (double)(n - j - 1) * (n - j) / 2.0;
  return done;
}


// ---------------------------------------------------------------------------
// Complex Hermitian reduction (LAPACK zhetrd, lower, on the column-major view
// B of the row-major buffer: buf[r][c] = B(c, r), so row j of the buffer is
// column j of B and the matrix-vector product reads conj(buf)).  Same two
// launches per column as the real reduction, on planar complex data (real
// plane, then imaginary plane).  Reflector j: beta = -sign(Re a) |a_{j+1:}|,
// tau = ((beta - Re a0) / beta, -Im a0 / beta), v = a_{j+2:} / (a0 - beta)
// (zlarfg), except that an exactly zero sub-column with a real a0 takes the
// reflection tau = 2, v = e_{j+1}; d and e are real, tau complex
// (interleaved, as rocSOLVER's zunmtr takes it).  The panel update is
// A22 -= V W^H + W V^H.
// ---------------------------------------------------------------------------
constexpr int TRC_PS = 1 + 4 * TRD_NB;   // |a|^2, W^H a and V^H a (re, im) per workgroup

struct TrcState {
  double alr, ali;   // alpha of the previous column
  double tr, ti;     // tau of the previous column
};

__global__ __launch_bounds__(256) void trc_col_k(const double* __restrict__ A, int n, int j, int jj,
                                                 const double* __restrict__ V, double* __restrict__ W,
                                                 const double* __restrict__ wpre, double* __restrict__ avec,
                                                 double* __restrict__ part1, const double* __restrict__ part2,
                                                 int g2prev, TrcState* __restrict__ st) {
  __shared__ double sVr[TRD_ROWS1][TRD_NB + 1], sVi[TRD_ROWS1][TRD_NB + 1];
  __shared__ double sWr[TRD_ROWS1][TRD_NB + 1], sWi[TRD_ROWS1][TRD_NB + 1];
  __shared__ double sWjr[TRD_NB], sWji[TRD_NB], sVjr[TRD_NB], sVji[TRD_NB], sh[8];
  __shared__ double sar[TRD_ROWS1], sai[TRD_ROWS1];
  __shared__ double redr[256], redi[256];
  const int tid = threadIdx.x;
  const size_t pA = (size_t)n * n, pV = (size_t)n * TRD_P;
  const int r0 = j + blockIdx.x * TRD_ROWS1;
  const int nr = max(0, min(TRD_ROWS1, n - r0));
  const int r = r0 + tid;
  // ---- loads
  const double tr = st->tr, ti = st->ti;
  double psr = 0.0, psi = 0.0;   // this thread's share of sum conj(w) v of column j-1
  if (jj > 0)
    for (int b = tid; b < g2prev; b += 256) {
      psr += part2[2 * b];
      psi += part2[2 * b + 1];
    }
  constexpr int SPT = TRD_ROWS1 * TRD_NB / 256;
  double lvr[SPT], lvi[SPT], lwr[SPT], lwi[SPT];
#pragma unroll
  for (int u = 0; u < SPT; ++u) {
    const int idx = tid + 256 * u, k = idx % TRD_NB;
    lvr[u] = lvi[u] = lwr[u] = lwi[u] = 0.0;
    if (idx < nr * TRD_NB && k < jj) {
      const size_t o = (size_t)(r0 + idx / TRD_NB) * TRD_P + k;
      lvr[u] = V[o];
      lvi[u] = V[o + pV];
      lwr[u] = W[o];
      lwi[u] = W[o + pV];
    }
  }
  double vjr = 0.0, vji = 0.0, wjr = 0.0, wji = 0.0;
  if (tid < jj) {
    const size_t o = (size_t)j * TRD_P + tid;
    vjr = V[o];
    vji = V[o + pV];
    if (tid == jj - 1) {
      wjr = wpre[j];
      wji = wpre[j + n];
    } else {
      wjr = W[o];
      wji = W[o + pV];
    }
  }
  double arr = 0.0, ari = 0.0, wpr = 0.0, wpi = 0.0;
  if (tid < nr) {
    arr = A[(size_t)j * n + r];
    ari = A[(size_t)j * n + r + pA];
    if (jj > 0) {
      wpr = wpre[r];
      wpi = wpre[r + n];
    }
  }
#pragma unroll
  for (int u = 0; u < SPT; ++u) {
    const int idx = tid + 256 * u;
    sVr[idx / TRD_NB][idx % TRD_NB] = lvr[u];
    sVi[idx / TRD_NB][idx % TRD_NB] = lvi[u];
    sWr[idx / TRD_NB][idx % TRD_NB] = lwr[u];
    sWi[idx / TRD_NB][idx % TRD_NB] = lwi[u];
  }
  // ---- alpha of column j-1: -tau/2 sum conj(w) v
  if (jj > 0) {
    psr = wave_sum(psr);
    psi = wave_sum(psi);
    if ((tid & 63) == 0) {
      redr[tid >> 6] = psr;
      redi[tid >> 6] = psi;
    }
  }
  __syncthreads();
  double alr = 0.0, ali = 0.0;
  if (jj > 0) {
    const double sr = redr[0] + redr[1] + redr[2] + redr[3];
    const double si = redi[0] + redi[1] + redi[2] + redi[3];
    alr = -0.5 * (tr * sr - ti * si);
    ali = -0.5 * (tr * si + ti * sr);
    if (blockIdx.x == 0 && tid == 0) {
      st->alr = alr;
      st->ali = ali;
    }
  }
  if (tid < jj) {
    sVjr[tid] = vjr;
    sVji[tid] = vji;
    if (tid == jj - 1) {
      sWjr[tid] = wjr + alr * vjr - ali * vji;
      sWji[tid] = wji + alr * vji + ali * vjr;
    } else {
      sWjr[tid] = wjr;
      sWji[tid] = wji;
    }
  }
  if (jj > 0 && tid < nr) {   // finish column jj-1 of W: w + alpha v
    const double xr = sVr[tid][jj - 1], xi = sVi[tid][jj - 1];
    const double wr = wpr + alr * xr - ali * xi, wi = wpi + alr * xi + ali * xr;
    sWr[tid][jj - 1] = wr;
    sWi[tid][jj - 1] = wi;
    W[(size_t)r * TRD_P + jj - 1] = wr;
    W[(size_t)r * TRD_P + jj - 1 + pV] = wi;
    W[(size_t)r * TRD_P + 2 * TRD_NB + jj - 1] = wr;   // the copy after V
    W[(size_t)r * TRD_P + 2 * TRD_NB + jj - 1 + pV] = wi;
  }
  __syncthreads();
  // ---- column j of B, updated with the panel: a -= V conj(W_j)^T + W conj(V_j)^T
  double pn = 0.0;
  if (tid < TRD_ROWS1) sar[tid] = sai[tid] = 0.0;
  if (tid < nr) {
    double ar = arr, ai = ari;
    for (int k = 0; k < jj; ++k) {
      // V[r][k] conj(Wj[k]) + W[r][k] conj(Vj[k])
      ar -= sVr[tid][k] * sWjr[k] + sVi[tid][k] * sWji[k] + sWr[tid][k] * sVjr[k] + sWi[tid][k] * sVji[k];
      ai -= sVi[tid][k] * sWjr[k] - sVr[tid][k] * sWji[k] + sWi[tid][k] * sVjr[k] - sWr[tid][k] * sVji[k];
    }
    avec[r] = ar;
    avec[r + n] = ai;
    if (r >= j + 2) {
      sar[tid] = ar;
      sai[tid] = ai;
      pn = ar * ar + ai * ai;
    }
  }
  const double pns = block_sum(pn, sh);
  {  // partial conj(W)^T a (k < jj) and conj(V)^T a (jj <= k < 2 jj) over this block's rows
    const int k = tid & 63, grp = tid >> 6;
    double accr = 0.0, acci = 0.0;
    if (k < 2 * jj)
      for (int i = grp; i < nr; i += 4) {
        const double xr = k < jj ? sWr[i][k] : sVr[i][k - jj];
        const double xi = k < jj ? sWi[i][k] : sVi[i][k - jj];
        accr += xr * sar[i] + xi * sai[i];
        acci += xr * sai[i] - xi * sar[i];
      }
    redr[grp * 64 + k] = accr;
    redi[grp * 64 + k] = acci;
  }
  __syncthreads();
  double* myp = part1 + (size_t)blockIdx.x * TRC_PS;
  if (tid == 0) myp[0] = pns;
  if (tid < 2 * jj) {
    const double re = redr[tid] + redr[64 + tid] + redr[128 + tid] + redr[192 + tid];
    const double im = redi[tid] + redi[64 + tid] + redi[128 + tid] + redi[192 + tid];
    const int base = tid < jj ? 1 + tid : 1 + 2 * jj + (tid - jj);   // W^H a: [1, 1+2jj), V^H a: [1+2jj, 1+4jj)
    myp[base] = re;
    myp[base + jj] = im;
  }
}

__global__ __launch_bounds__(256) void trc_symv_k(double* __restrict__ A, int n, int j, int jj,
                                                  double* __restrict__ V, const double* __restrict__ W,
                                                  double* __restrict__ wpre, const double* __restrict__ avec,
                                                  const double* __restrict__ part1, int g1,
                                                  double* __restrict__ part2, TrcState* __restrict__ st,
                                                  double* __restrict__ d, double* __restrict__ e,
                                                  double* __restrict__ tau) {
  __shared__ double syr[TRD_NB], syi[TRD_NB], szr[TRD_NB], szi[TRD_NB];
  __shared__ double redr[4][TRD_ROWS2], redi[4][TRD_ROWS2];
  __shared__ double red4[256];
  __shared__ double tot[TRC_PS];
  __shared__ double sst[4];   // tau (re, im), scale (re, im)
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const size_t pA = (size_t)n * n, pV = (size_t)n * TRD_P;
  const int nk = 1 + 4 * jj;   // <= 125
  const int rbase = j + 1 + blockIdx.x * TRD_ROWS2;
  const int nr = min(TRD_ROWS2, n - rbase);   // >= 1
  // ---- loads: launch-1 partials (two groups of 128), reflector inputs, correction rows
  double ps = 0.0;
  {
    const int k = tid & 127, grp = tid >> 7;
    if (k < nk) {
#pragma unroll 4
      for (int b = grp; b < g1; b += 2) ps += part1[(size_t)b * TRC_PS + k];
    }
  }
  double a0r = 0.0, a0i = 0.0, aj = 0.0;
  if (tid == 0) {
    a0r = avec[j + 1];
    a0i = avec[j + 1 + n];
    aj = avec[j];
  }
  // row j+1 of V and W (column jj-1 of W was finished by launch 1, which
  // covers rows >= j; wpre[j+1] must not be read here: workgroup 0 of this
  // launch overwrites it with the new column)
  double v1r = 0.0, v1i = 0.0, w1r = 0.0, w1i = 0.0;
  if (tid < jj) {
    const size_t o = (size_t)(j + 1) * TRD_P + tid;
    v1r = V[o];
    v1i = V[o + pV];
    w1r = W[o];
    w1i = W[o + pV];
  }
  double mr[TRD_ROWS2], mi[TRD_ROWS2];
#pragma unroll
  for (int i = 0; i < TRD_ROWS2; ++i) mr[i] = mi[i] = 0.0;
  if (tid < 64 && (lane & 31) < jj) {
    const double* M = lane < 32 ? V : W;
#pragma unroll
    for (int i = 0; i < TRD_ROWS2; ++i)
      if (i < nr) {
        const size_t o = (size_t)(rbase + i) * TRD_P + (lane & 31);
        mr[i] = M[o];
        mi[i] = M[o + pV];
      }
  }
  // ---- the matrix stream: acc = sum_{c >= j+2} conj(buf[r][c]) a[c], a1 = conj(buf[r][j+1])
  double accr[TRD_ROWS2], acci[TRD_ROWS2], a1r[TRD_ROWS2], a1i[TRD_ROWS2];
#pragma unroll
  for (int i = 0; i < TRD_ROWS2; ++i) accr[i] = acci[i] = a1r[i] = a1i[i] = 0.0;
#pragma unroll 4
  for (int c = j + 1 + tid; c < n; c += 256) {
    const bool first = c == j + 1;
    const double ar = first ? 0.0 : avec[c], ai = first ? 0.0 : avec[c + n];
#pragma unroll
    for (int i = 0; i < TRD_ROWS2; ++i) {
      const int r = rbase + min(i, nr - 1);   // rows past the end repeat the last (discarded)
      const double xr = A[(size_t)r * n + c], xi = A[(size_t)r * n + c + pA];
      accr[i] += xr * ar + xi * ai;
      acci[i] += xr * ai - xi * ar;
      if (first) {
        a1r[i] = xr;
        a1i[i] = -xi;
      }
    }
  }
  // ---- reflector (every workgroup, same fixed-order sums)
  red4[tid] = ps;
  __syncthreads();
  if (tid < nk) tot[tid] = red4[tid] + red4[128 + tid];
  __syncthreads();
  if (tid == 0) {
    const double xn2 = tot[0];
    // an exactly zero sub-column with a real a0 takes tau = 2, v = e_{j+1}
    // (H = diag(-1, 1, ...), beta = -a0) instead of LAPACK's tau = 0:
    // rocSOLVER's complex back-transformation (unmtr) returns NaN vectors
    // when some but not all reflectors have tau = 0
    double tr = 2.0, ti = 0.0, beta = -a0r, scr = 0.0, sci = 0.0;
    if (xn2 != 0.0 || a0i != 0.0) {
      beta = -copysign(hypot(hypot(a0r, a0i), sqrt(xn2)), a0r);
      tr = (beta - a0r) / beta;
      ti = -a0i / beta;
      // scale = 1 / (a0 - beta)
      const double dr = a0r - beta, di = a0i;
      const double den = dr * dr + di * di;
      scr = dr / den;
      sci = -di / den;
    }
    sst[0] = tr;
    sst[1] = ti;
    sst[2] = scr;
    sst[3] = sci;
    if (blockIdx.x == 0) {
      d[j] = aj;
      e[j] = beta;
      tau[2 * j] = tr;
      tau[2 * j + 1] = ti;
      st->tr = tr;
      st->ti = ti;
    }
  }
  __syncthreads();
  const double tr = sst[0], ti = sst[1], scr = sst[2], sci = sst[3];
  if (tid < jj) {   // y = W^H v = conj(W[j+1]) + scale W^H a, z = V^H v likewise
    const double war = tot[1 + tid], wai = tot[1 + jj + tid];
    const double var = tot[1 + 2 * jj + tid], vai = tot[1 + 3 * jj + tid];
    syr[tid] = w1r + scr * war - sci * wai;
    syi[tid] = -w1i + scr * wai + sci * war;
    szr[tid] = v1r + scr * var - sci * vai;
    szi[tid] = -v1i + scr * vai + sci * var;
  }
  __syncthreads();
  // ---- p - V y - W z per row
  double fr = 0.0, fi = 0.0;
  if (tid < 64 && (lane & 31) < jj) {
    fr = lane < 32 ? syr[lane & 31] : szr[lane & 31];
    fi = lane < 32 ? syi[lane & 31] : szi[lane & 31];
  }
#pragma unroll
  for (int i = 0; i < TRD_ROWS2; ++i) {
    const double qr = wave_sum(a1r[i] + scr * accr[i] - sci * acci[i] - (mr[i] * fr - mi[i] * fi));
    const double qi = wave_sum(a1i[i] + scr * acci[i] + sci * accr[i] - (mr[i] * fi + mi[i] * fr));
    if (lane == 0) {
      redr[wave][i] = qr;
      redi[wave][i] = qi;
    }
  }
  __syncthreads();
  if (tid < TRD_ROWS2) {
    double pr = 0.0, pi = 0.0;
    if (tid < nr) {
      const int r = rbase + tid;
      const double qr = redr[0][tid] + redr[1][tid] + redr[2][tid] + redr[3][tid];
      const double qi = redi[0][tid] + redi[1][tid] + redi[2][tid] + redi[3][tid];
      const double wr = tr * qr - ti * qi, wi = tr * qi + ti * qr;
      double vr = 1.0, vi = 0.0;
      if (r != j + 1) {
        const double ar = avec[r], ai = avec[r + n];
        vr = scr * ar - sci * ai;
        vi = scr * ai + sci * ar;
      }
      wpre[r] = wr;
      wpre[r + n] = wi;
      V[(size_t)r * TRD_P + jj] = vr;
      V[(size_t)r * TRD_P + jj + pV] = vi;
      if (r >= j + 2) {   // reflector storage (row j)
        A[(size_t)j * n + r] = vr;
        A[(size_t)j * n + r + pA] = vi;
      }
      pr = wr * vr + wi * vi;   // conj(w) v
      pi = wr * vi - wi * vr;
    }
    red4[tid] = pr;
    red4[8 + tid] = pi;
  }
  __syncthreads();
  if (tid == 0) {
    double sr = 0.0, si = 0.0;
    for (int i = 0; i < TRD_ROWS2; ++i) {
      sr += red4[i];
      si += red4[8 + i];
    }
    part2[2 * blockIdx.x] = sr;
    part2[2 * blockIdx.x + 1] = si;
  }
}

__global__ __launch_bounds__(256) void trc_fin_k(int n, int s, int k, const double* __restrict__ V,
                                                 double* __restrict__ W, const double* __restrict__ wpre,
                                                 const double* __restrict__ part2, int g2,
                                                 const TrcState* __restrict__ st) {
  __shared__ double red4[256];
  double psr = 0.0, psi = 0.0;
  for (int b = threadIdx.x; b < g2; b += 256) {
    psr += part2[2 * b];
    psi += part2[2 * b + 1];
  }
  psr = wave_sum(psr);
  psi = wave_sum(psi);
  if ((threadIdx.x & 63) == 0) {
    red4[threadIdx.x >> 6] = psr;
    red4[4 + (threadIdx.x >> 6)] = psi;
  }
  __syncthreads();
  const double sr = red4[0] + red4[1] + red4[2] + red4[3], si = red4[4] + red4[5] + red4[6] + red4[7];
  const double tr = st->tr, ti = st->ti;
  const double alr = -0.5 * (tr * sr - ti * si), ali = -0.5 * (tr * si + ti * sr);
  const size_t pV = (size_t)n * TRD_P;
  const int r = s + blockIdx.x * 256 + threadIdx.x;
  if (r < n) {
    const size_t o = (size_t)r * TRD_P + k;
    const double vr = V[o], vi = V[o + pV];
    const double wr = wpre[r] + alr * vr - ali * vi, wi = wpre[r + n] + alr * vi + ali * vr;
    W[o] = wr;
    W[o + pV] = wi;
    W[o + 2 * TRD_NB] = wr;   // the copy after V
    W[o + 2 * TRD_NB + pV] = wi;
  }
}

// A (n x n, planar complex with plane stride n*n, Hermitian; the buffer read
// column-major) -> d[n], e[n-1] (real), tau[n-1] (complex, interleaved) and
// reflectors in place (zhetrd, uplo 'L', on the column-major view).  Returns
// the number of reflectors computed; stop_tol > 0 stops the reduction as
// sytrd_lower does (trailing block, both planes, below stop_tol |A|_F).
int hetrd_lower(Ctx* c, double* A, int n, double* d, double* e, double* tau, double stop_tol, double* dropped) {
  DMRG_REQUIRE(n >= 2, "hetrd: order must be at least 2");
  const size_t vw = (size_t)n * TRD_P;   // one plane of the [W | V | W] panel store
  const int gmax1 = (n + TRD_ROWS1 - 1) / TRD_ROWS1;
  const int gmax2 = (n + TRD_ROWS2 - 1) / TRD_ROWS2;
  double* buf = (double*)ws_get(c, WS_TRD, sizeof(double) * (2 * vw + 4 * (size_t)n + (size_t)gmax1 * TRC_PS +
                                                              2 * (size_t)gmax2 + 8) +
                                               sizeof(TrcState) + 64);
  double* W = buf;             // planar (plane stride vw); row r: W at [0, NB), V at [NB, 2 NB), W again
  double* V = W + TRD_NB;
  double* wpre = buf + 2 * vw;   // planar, stride n
  double* avec = wpre + 2 * (size_t)n;
  double* part1 = avec + 2 * (size_t)n;
  double* part2 = part1 + (size_t)gmax1 * TRC_PS;
  TrcState* st = reinterpret_cast<TrcState*>(part2 + 2 * (size_t)gmax2 + 8);
  DMRG_CHECK_HIP(hipMemsetAsync(st, 0, sizeof(TrcState), c->stream));
  ProfScope ps(c, 4, 0.0, 0.0);
  const bool early = stop_tol > 0.0 && n > TRD_STOP_MIN + TRD_NB;
  double* sq_out = part2 + 2 * (size_t)gmax2;   // 8 spare doubles before the state (two used)
  double stop2 = 0.0;
  const int sq_cap = gmax1 * TRC_PS;
  if (early) {
    trailing_sumsq_async(c, A, n, 0, 2, part1, sq_cap, sq_out, 0);
    stop2 = stop_tol * stop_tol * trailing_sumsq_wait(c, 0);
  }
  int done = n - 1;
  int pend = -1;
  int g2prev = 1;
  for (int j0 = 0; j0 < n - 1; j0 += TRD_NB) {
    const int nbp = std::min(TRD_NB, n - 1 - j0);
    for (int jj = 0; jj < nbp; ++jj) {
      const int j = j0 + jj;
      const int g1 = (n - j + TRD_ROWS1 - 1) / TRD_ROWS1;
      const int g2 = (n - j - 1 + TRD_ROWS2 - 1) / TRD_ROWS2;
      DMRG_REQUIRE(g1 >= 1 && g1 <= gmax1 && g2 >= 1 && g2 <= gmax2, "hetrd: grid exceeds partial buffers");
      trc_col_k<<<g1, 256, 0, c->stream>>>(A, n, j, jj, V, W, wpre, avec, part1, part2, g2prev, st);
      DMRG_CHECK_HIP(hipGetLastError());
      trc_symv_k<<<g2, 256, 0, c->stream>>>(A, n, j, jj, V, W, wpre, avec, part1, g1, part2, st, d, e, tau);
      DMRG_CHECK_HIP(hipGetLastError());
      g2prev = g2;
    }
    const int s = j0 + nbp;
    if (s < n) {
      trc_fin_k<<<(n - s + 255) / 256, 256, 0, c->stream>>>(n, s, nbp - 1, V, W, wpre, part2, g2prev, st);
      DMRG_CHECK_HIP(hipGetLastError());
      // buf(r, c) -= conj(W[r]) . V[c] + conj(V[r]) . W[c]   (B22 -= V W^H + W V^H):
      // one GEMM over [W V] and [V W] for a full panel
      const int m = n - s;
      const bool full = nbp == TRD_NB;
      GemmDesc g = gdesc(m, m, full ? 2 * TRD_NB : nbp, TRD_P, TRD_P, n, CPLX, 4);
      g.transB = true;
      g.conjA = true;
      g.alpha = -1.0;
      g.beta = 1.0;
      g.pA = (long long)vw;
      g.pB = (long long)vw;
      g.pC = (long long)n * n;
      double* C = A + (size_t)s * n + s;
      const double* Ws = W + (size_t)s * TRD_P;
      gemm_f64(c, g, Ws, Ws + TRD_NB, C);
      if (!full) gemm_f64(c, g, Ws + TRD_NB, Ws, C);   // short last panel: columns past nbp are stale
      int slot = -1;
      if (early && full && s >= TRD_STOP_MIN && s < n - 1) {
        slot = pend == 0 ? 1 : 0;
        trailing_sumsq_async(c, A, n, s, 2, part1, sq_cap, sq_out, slot);
      }
      if (pend >= 0 && trailing_sumsq_wait(c, pend) <= stop2 && s < n - 1) {
        if (dropped) *dropped = trailing_trace(c, A, n, s, sq_out + 2);   // real plane: Hermitian diagonal
        trd_clear_tail_k<<<(n - s + 255) / 256, 256, 0, c->stream>>>(n, s, d, e, tau, 2);
        DMRG_CHECK_HIP(hipGetLastError());
        done = s;
        break;
      }
      pend = slot;
    }
  }
  if (done == n - 1)
    DMRG_CHECK_HIP(hipMemcpyAsync(d + (n - 1), A + (size_t)(n - 1) * n + (n - 1), sizeof(double),
                                  hipMemcpyDeviceToDevice, c->stream));
  // algorithmic bytes: the lower triangle of the trailing matrix (both planes) once per reduced column
  for (int j = 0; j < done; ++j) ps.bytes += 16.0 * (double)(n - j - 1) * (n - j) / 2.0;
  return done;
}


// ---------------------------------------------------------------------------
// Back-transformation of kept eigenvectors (dormtr, side L, lower, no
// transpose, on the reflectors sytrd_lower leaves): E = Q Y with
// Q = H_0 H_1 ... H_{n-2}, H_j = I - tau_j v_j v_j^T.  The reflectors are
// grouped into blocks of BT_NB (compact WY: Q_b = I - V_b T_b V_b^T, T_b upper
// triangular, LAPACK dlarft forward/columnwise), and each block is applied
// to the eigenvectors as three GEMMs, last block first.  The eigenvectors
// are the rows of Zt (keep x n, row-major), so a block is Zt <- Zt Q_b^T:
//   Y = Zt V_b (split along its long inner dimension into BT_SPLIT partial
//   products written side by side), Y2 = [Y_0 .. Y_{S-1}] [T_b^T; ..; T_b^T],
//   Zt -= Y2 V_b^T.
// ---------------------------------------------------------------------------
constexpr int BT_NB = 128;
constexpr int BT_SPLIT = 8;

// Vall[j][r] = v_j[r] for rows j < n - 1 (0 for r <= j, 1 at r = j + 1, the
// stored entry A[j][r] beyond), zero rows up to the block multiple
__global__ void bt_vall_k(const double* __restrict__ A, int n, int nrefl, long long total,
                          double* __restrict__ Vall) {
  for (long long idx = (long long)blockIdx.x * 256 + threadIdx.x; idx < total; idx += (long long)gridDim.x * 256) {
    const int j = (int)(idx / n), r = (int)(idx % n);
    double v = 0.0;
    if (j < nrefl) {
      if (r == j + 1) v = 1.0;
      else if (r > j + 1) v = A[(size_t)j * n + r];
    }
    Vall[idx] = v;
  }
}

// T_b from S_b = V_b V_b^T and tau (one workgroup of 256 threads per block),
// then BT_SPLIT stacked copies of T_b^T.  T is built by blocked recursion:
// its 16 x 16 diagonal blocks by the column recurrence (thread k owns row
// k: T[k][k] = tau_k, T[k][i] = -tau_i sum_{l=k}^{i-1} T[k][l] S[l][i]),
// then pairs of diagonal blocks of size h = 16, 32, 64 are joined by the
// compact-WY identity T12 = -T1 (S12 T2) (LAPACK dlarft's recursive form:
// (I - V1 T1 V1^T)(I - V2 T2 V2^T) = I - V T V^T with that T12), every
// output an independent dot product.  The serial recurrence over the whole
// block made row 0 a chain of 127 dependent steps (279 us per call).  The
// strict upper triangle of S (column-packed) and the upper triangle of T
// (row-packed) sit in LDS; X = S12 T2 is kept in T12's slots until
// T12 = -T1 X replaces it (through registers, so no slot is read after it
// is overwritten).
constexpr int BT_TRI = BT_NB * (BT_NB + 1) / 2;
__device__ __forceinline__ int bt_s_idx(int l, int i) { return i * (i - 1) / 2 + l; }            // l < i
__device__ __forceinline__ int bt_t_idx(int k, int l) { return k * BT_NB - k * (k - 1) / 2 + (l - k); }  // l >= k
constexpr int BT_LEAF = 16;

__global__ __launch_bounds__(256) void bt_larft_k(const double* __restrict__ S, const double* __restrict__ tau,
                                                  int nrefl, double* __restrict__ TtS) {
  __shared__ double sS[BT_NB * (BT_NB - 1) / 2];
  __shared__ double sT[BT_TRI];
  __shared__ double stau[BT_NB];
  const int b = blockIdx.x, tid = threadIdx.x;
  const int j0 = b * BT_NB;
  const double* Sb = S + (size_t)b * BT_NB * BT_NB;
  if (tid < BT_NB) stau[tid] = j0 + tid < nrefl ? tau[j0 + tid] : 0.0;
  for (int idx = tid; idx < BT_NB * BT_NB; idx += 256) {   // S[t][k], kept for t < k
    const int t = idx / BT_NB, k = idx % BT_NB;
    if (t < k) sS[bt_s_idx(t, k)] = Sb[idx];
  }
  __syncthreads();
  if (tid < BT_NB) {   // diagonal blocks of BT_LEAF by the recurrence
    const int k = tid, end = (k / BT_LEAF + 1) * BT_LEAF;
    const int tk = bt_t_idx(k, 0);   // row k of T is contiguous
    sT[tk + k] = stau[k];
    for (int i = k + 1; i < end; ++i) {
      const int si = bt_s_idx(0, i);   // column i of S is contiguous
      double a = 0.0;
      for (int l = k; l < i; ++l) a += sT[tk + l] * sS[si + l];
      sT[tk + i] = -stau[i] * a;
    }
  }
  __syncthreads();
  constexpr int MAXPER = BT_NB * (BT_NB / 2) / 2 / 256;   // outputs per thread at the last level (16)
  for (int h = BT_LEAF; h < BT_NB; h *= 2) {
    const int per = BT_NB * h / 2 / 256;   // (BT_NB / 2h) pairs x h x h outputs over 256 threads
    double x[MAXPER];
    // X = S12 T2 into the T12 slots (T2 is read, T12 only written: disjoint)
#pragma unroll
    for (int e = 0; e < MAXPER; ++e) {
      if (e < per) {
        const int o = tid + 256 * e;
        const int p = o / (h * h), r = o % (h * h);
        const int b0 = 2 * h * p, k = b0 + r / h, i = b0 + h + r % h;
        double a = 0.0;
        for (int l = b0 + h; l <= i; ++l) a += sS[bt_s_idx(k, l)] * sT[bt_t_idx(l, i)];
        x[e] = a;
      }
    }
#pragma unroll
    for (int e = 0; e < MAXPER; ++e) {
      if (e < per) {
        const int o = tid + 256 * e;
        const int p = o / (h * h), r = o % (h * h);
        const int b0 = 2 * h * p, k = b0 + r / h, i = b0 + h + r % h;
        sT[bt_t_idx(k, i)] = x[e];
      }
    }
    __syncthreads();
    // T12 = -T1 X
#pragma unroll
    for (int e = 0; e < MAXPER; ++e) {
      if (e < per) {
        const int o = tid + 256 * e;
        const int p = o / (h * h), r = o % (h * h);
        const int b0 = 2 * h * p, k = b0 + r / h, i = b0 + h + r % h;
        const int tk = bt_t_idx(k, 0);
        double a = 0.0;
        for (int l = k; l < b0 + h; ++l) a += sT[tk + l] * sT[bt_t_idx(l, i)];
        x[e] = -a;
      }
    }
    __syncthreads();
#pragma unroll
    for (int e = 0; e < MAXPER; ++e) {
      if (e < per) {
        const int o = tid + 256 * e;
        const int p = o / (h * h), r = o % (h * h);
        const int b0 = 2 * h * p, k = b0 + r / h, i = b0 + h + r % h;
        sT[bt_t_idx(k, i)] = x[e];
      }
    }
    __syncthreads();
  }
  double* out = TtS + (size_t)b * BT_SPLIT * BT_NB * BT_NB;
  for (int idx = tid; idx < BT_NB * BT_NB; idx += 256) {   // row k of T^T = column k of T
    const int k = idx / BT_NB, i = idx % BT_NB;
    const double v = i <= k ? sT[bt_t_idx(i, k)] : 0.0;
    for (int s = 0; s < BT_SPLIT; ++s) out[((size_t)s * BT_NB + k) * BT_NB + i] = v;
  }
}

// Y[c][i] = sum_s Ycat[c][s nb + i] (the split inner products of a block, in
// order), for every plane: rows*nb outputs per plane, plane strides pin/pout
__global__ void bt_sum_k(const double* __restrict__ Ycat, int rows, int nb, int S, long long pin,
                         double* __restrict__ Y, long long pout, int planes) {
  const long long tot = (long long)rows * nb;
  for (long long idx = (long long)blockIdx.x * 256 + threadIdx.x; idx < tot * planes;
       idx += (long long)gridDim.x * 256) {
    const int pl = (int)(idx / tot);
    const long long o = idx % tot;
    const int c = (int)(o / nb), i = (int)(o % nb);
    const double* src = Ycat + pl * pin + (size_t)c * S * nb + i;
    double s = 0.0;
    for (int q = 0; q < S; ++q) s += src[(size_t)q * nb];
    Y[pl * pout + o] = s;
  }
}

// Zt (keep x n, row-major, the rows eigenvectors of the tridiagonal) <- rows
// of Q Y, Q from the reflectors in A (row-major n x n as sytrd_lower leaves
// it) and tau
void ormtr_lower_rows(Ctx* c, const double* A, int n, const double* tau, double* Zt, int keep, int nrefl) {
  if (nrefl < 0) nrefl = n - 1;
  DMRG_REQUIRE(n >= 2 && keep >= 1 && keep <= n && nrefl >= 1 && nrefl <= n - 1, "ormtr: bad sizes");
  const int nblk = (nrefl + BT_NB - 1) / BT_NB;
  const size_t vrows = (size_t)nblk * BT_NB;
  double* Vall = (double*)ws_get(c, WS_BT_V, sizeof(double) * vrows * n);
  const size_t nS = (size_t)nblk * BT_NB * BT_NB, nT = (size_t)nblk * BT_SPLIT * BT_NB * BT_NB;
  const size_t nY = (size_t)keep * (BT_SPLIT + 1) * BT_NB, nY2 = (size_t)keep * BT_NB;   // S partials + their sum
  double* Sm = (double*)ws_get(c, WS_BT_W, sizeof(double) * (nS + nT + nY + nY2));
  double* TtS = Sm + nS;
  double* Ycat = TtS + nT;
  double* Y2 = Ycat + nY;
  const long long total = (long long)vrows * n;
  bt_vall_k<<<(int)std::min<long long>((total + 255) / 256, 65536), 256, 0, c->stream>>>(A, n, nrefl, total, Vall);
  DMRG_CHECK_HIP(hipGetLastError());
  {  // S_b = V_b V_b^T for every block at once
    GemmDesc g = gdesc(BT_NB, BT_NB, n, n, n, BT_NB, REAL, 4);
    g.transB = true;
    g.batch = nblk;
    g.sA = g.sB = (long long)BT_NB * n;
    g.sC = (long long)BT_NB * BT_NB;
    gemm_f64(c, g, Vall, Vall, Sm);
  }
  bt_larft_k<<<nblk, 256, 0, c->stream>>>(Sm, tau, nrefl, TtS);
  DMRG_CHECK_HIP(hipGetLastError());
  for (int b = nblk - 1; b >= 0; --b) {
    const int b0 = b * BT_NB;
    const int len = n - b0 - 1;   // columns b0 + 1 .. n - 1 carry block b
    if (len <= 0) continue;
    int S = BT_SPLIT, kc = len, start = b0 + 1;
    for (; S >= 1; --S) {   // S equal chunks ending at column n; columns left of b0 + 1 meet zero rows of V_b
      kc = (len + S - 1) / S;
      start = n - S * kc;
      if (start >= 0) break;
    }
    DMRG_REQUIRE(S >= 1 && start >= 0 && start <= b0 + 1, "ormtr: bad split");
    const double* Vb = Vall + (size_t)b * BT_NB * n;
    GemmDesc g1 = gdesc(keep, BT_NB, kc, n, n, S * BT_NB, REAL, 4);
    g1.transB = true;
    g1.batch = S;
    g1.sA = g1.sB = kc;
    g1.sC = BT_NB;
    gemm_f64(c, g1, Zt + start, Vb + start, Ycat);
    const double* Y = Ycat;
    if (S > 1) {   // sum the split products in place of an inner dimension S*NB in the next GEMM
      double* Ys = Ycat + (size_t)keep * S * BT_NB;   // room after the S partials
      const long long tot = (long long)keep * BT_NB;
      bt_sum_k<<<(int)std::min<long long>((tot + 255) / 256, 4096), 256, 0, c->stream>>>(Ycat, keep, BT_NB, S, 0, Ys,
                                                                                           0, 1);
      DMRG_CHECK_HIP(hipGetLastError());
      Y = Ys;
    }
    GemmDesc g2 = gdesc(keep, BT_NB, BT_NB, BT_NB, BT_NB, BT_NB, REAL, 4);
    gemm_f64(c, g2, Y, TtS + (size_t)b * BT_SPLIT * BT_NB * BT_NB, Y2);
    GemmDesc g3 = gdesc(keep, len, BT_NB, BT_NB, n, n, REAL, 4);
    g3.alpha = -1.0;
    g3.beta = 1.0;
    gemm_f64(c, g3, Y2, Vb + b0 + 1, Zt + b0 + 1);
  }
}


// ---------------------------------------------------------------------------
// Complex counterpart (zunmtr L/lower/N on hetrd_lower's reflectors), planar
// complex throughout, blocks of BTC_NB reflectors: Q_b = I - V_b T_b V_b^H
// (T_b(k, i) = -tau_i sum_l T_b(k, l) v_l^H v_i), applied to the rows of
// Et = (Q Y)^T as Et <- Et - (Et conj(V_b)) T_b^T V_b^T, last block first.
// ---------------------------------------------------------------------------
constexpr int BTC_NB = 64;
constexpr int BTC_TRI = BTC_NB * (BTC_NB + 1) / 2;

__global__ void btc_vall_k(const double* __restrict__ A, int n, int nrefl, long long total,
                           double* __restrict__ Vall) {
  const size_t pA = (size_t)n * n;
  for (long long idx = (long long)blockIdx.x * 256 + threadIdx.x; idx < total; idx += (long long)gridDim.x * 256) {
    const int j = (int)(idx / n), r = (int)(idx % n);
    double vr = 0.0, vi = 0.0;
    if (j < nrefl) {
      if (r == j + 1) {
        vr = 1.0;
      } else if (r > j + 1) {
        vr = A[(size_t)j * n + r];
        vi = A[(size_t)j * n + r + pA];
      }
    }
    Vall[idx] = vr;
    Vall[idx + total] = vi;
  }
}

__device__ __forceinline__ int btc_s_idx(int l, int i) { return i * (i - 1) / 2 + l; }
__device__ __forceinline__ int btc_t_idx(int k, int l) { return k * BTC_NB - k * (k - 1) / 2 + (l - k); }

// S planar (plane stride pS), tau interleaved; TtS planar (plane stride pT).
// The same blocked recursion as bt_larft_k in complex arithmetic
// (S[l][i] = v_l^H v_i, T12 = -T1 S12 T2 from (I - V1 T1 V1^H)(I - V2 T2 V2^H)),
// 256 threads: 16 x 16 diagonal blocks by the recurrence, then levels 16, 32.
__global__ __launch_bounds__(256) void btc_larft_k(const double* __restrict__ S, long long pS,
                                                  const double* __restrict__ tau, int nrefl,
                                                  double* __restrict__ TtS, long long pT) {
  __shared__ double sSr[BTC_TRI], sSi[BTC_TRI], sTr[BTC_TRI], sTi[BTC_TRI];
  __shared__ double str[BTC_NB], sti[BTC_NB];
  const int b = blockIdx.x, tid = threadIdx.x;
  const int j0 = b * BTC_NB;
  const double* Sb = S + (size_t)b * BTC_NB * BTC_NB;
  if (tid < BTC_NB) {
    str[tid] = j0 + tid < nrefl ? tau[2 * (j0 + tid)] : 0.0;
    sti[tid] = j0 + tid < nrefl ? tau[2 * (j0 + tid) + 1] : 0.0;
  }
  for (int idx = tid; idx < BTC_NB * BTC_NB; idx += 256) {   // S[t][k], kept for t < k
    const int t = idx / BTC_NB, k = idx % BTC_NB;
    if (t < k) {
      sSr[btc_s_idx(t, k)] = Sb[idx];
      sSi[btc_s_idx(t, k)] = Sb[idx + pS];
    }
  }
  __syncthreads();
  if (tid < BTC_NB) {   // diagonal blocks of BT_LEAF by the recurrence
    const int k = tid, end = (k / BT_LEAF + 1) * BT_LEAF;
    const int tk = btc_t_idx(k, 0);
    sTr[tk + k] = str[k];
    sTi[tk + k] = sti[k];
    for (int i = k + 1; i < end; ++i) {
      const int si = btc_s_idx(0, i);
      double ar = 0.0, ai = 0.0;
      for (int l = k; l < i; ++l) {
        const double xr = sTr[tk + l], xi = sTi[tk + l], yr = sSr[si + l], yi = sSi[si + l];
        ar += xr * yr - xi * yi;
        ai += xr * yi + xi * yr;
      }
      sTr[tk + i] = -(str[i] * ar - sti[i] * ai);
      sTi[tk + i] = -(str[i] * ai + sti[i] * ar);
    }
  }
  __syncthreads();
  constexpr int MAXPER = BTC_NB * (BTC_NB / 2) / 2 / 256;   // 4
  for (int h = BT_LEAF; h < BTC_NB; h *= 2) {
    const int per = BTC_NB * h / 2 / 256;
    double xr[MAXPER], xi[MAXPER];
#pragma unroll
    for (int e = 0; e < MAXPER; ++e) {   // X = S12 T2 into the T12 slots
      if (e < per) {
        const int o = tid + 256 * e;
        const int p = o / (h * h), r = o % (h * h);
        const int b0 = 2 * h * p, k = b0 + r / h, i = b0 + h + r % h;
        double ar = 0.0, ai = 0.0;
        for (int l = b0 + h; l <= i; ++l) {
          const double sr = sSr[btc_s_idx(k, l)], si = sSi[btc_s_idx(k, l)];
          const double tr = sTr[btc_t_idx(l, i)], ti = sTi[btc_t_idx(l, i)];
          ar += sr * tr - si * ti;
          ai += sr * ti + si * tr;
        }
        xr[e] = ar;
        xi[e] = ai;
      }
    }
#pragma unroll
    for (int e = 0; e < MAXPER; ++e) {
      if (e < per) {
        const int o = tid + 256 * e;
        const int p = o / (h * h), r = o % (h * h);
        const int b0 = 2 * h * p, k = b0 + r / h, i = b0 + h + r % h;
        sTr[btc_t_idx(k, i)] = xr[e];
        sTi[btc_t_idx(k, i)] = xi[e];
      }
    }
    __syncthreads();
#pragma unroll
    for (int e = 0; e < MAXPER; ++e) {   // T12 = -T1 X
      if (e < per) {
        const int o = tid + 256 * e;
        const int p = o / (h * h), r = o % (h * h);
        const int b0 = 2 * h * p, k = b0 + r / h, i = b0 + h + r % h;
        const int tk = btc_t_idx(k, 0);
        double ar = 0.0, ai = 0.0;
        for (int l = k; l < b0 + h; ++l) {
          const double ur = sTr[tk + l], ui = sTi[tk + l];
          const double vr = sTr[btc_t_idx(l, i)], vi = sTi[btc_t_idx(l, i)];
          ar += ur * vr - ui * vi;
          ai += ur * vi + ui * vr;
        }
        xr[e] = -ar;
        xi[e] = -ai;
      }
    }
    __syncthreads();
#pragma unroll
    for (int e = 0; e < MAXPER; ++e) {
      if (e < per) {
        const int o = tid + 256 * e;
        const int p = o / (h * h), r = o % (h * h);
        const int b0 = 2 * h * p, k = b0 + r / h, i = b0 + h + r % h;
        sTr[btc_t_idx(k, i)] = xr[e];
        sTi[btc_t_idx(k, i)] = xi[e];
      }
    }
    __syncthreads();
  }
  double* out = TtS + (size_t)b * BT_SPLIT * BTC_NB * BTC_NB;
  for (int idx = tid; idx < BTC_NB * BTC_NB; idx += 256) {   // row r of T^T (column r of T)
    const int r = idx / BTC_NB, k = idx % BTC_NB;
    const double vr = k <= r ? sTr[btc_t_idx(k, r)] : 0.0, vi = k <= r ? sTi[btc_t_idx(k, r)] : 0.0;
    for (int s = 0; s < BT_SPLIT; ++s) {
      const size_t o = ((size_t)s * BTC_NB + r) * BTC_NB + k;
      out[o] = vr;
      out[o + pT] = vi;
    }
  }
}

// Et (keep x n planar, plane stride keep*n) <- rows of Q Y for hetrd_lower's
// reflectors in A (planar, plane stride n*n) and tau (interleaved); nrefl
// reflectors (n - 1 when negative)
void unmtr_lower_rows(Ctx* c, const double* A, int n, const double* tau, double* Et, int keep, int nrefl) {
  if (nrefl < 0) nrefl = n - 1;
  DMRG_REQUIRE(n >= 2 && keep >= 1 && keep <= n && nrefl >= 1 && nrefl <= n - 1, "unmtr: bad sizes");
  const int NB = BTC_NB;
  const int nblk = (nrefl + NB - 1) / NB;
  const size_t vrows = (size_t)nblk * NB;
  const long long pV = (long long)vrows * n;
  double* Vall = (double*)ws_get(c, WS_BT_V, sizeof(double) * 2 * (size_t)pV);
  const size_t nS = (size_t)nblk * NB * NB, nT = (size_t)nblk * BT_SPLIT * NB * NB;
  const size_t nY = (size_t)keep * (BT_SPLIT + 1) * NB, nY2 = (size_t)keep * NB;   // S partials + their sum
  double* Sm = (double*)ws_get(c, WS_BT_W, sizeof(double) * 2 * (nS + nT + nY + nY2));
  double* TtS = Sm + 2 * nS;
  double* Ycat = TtS + 2 * nT;
  double* Y2 = Ycat + 2 * nY;
  btc_vall_k<<<(int)std::min<long long>((pV + 255) / 256, 65536), 256, 0, c->stream>>>(A, n, nrefl, pV, Vall);
  DMRG_CHECK_HIP(hipGetLastError());
  {  // S_b = V_b^H V_b (rows of Vall are the reflectors): conj(Vall_b) Vall_b^T
    GemmDesc g = gdesc(NB, NB, n, n, n, NB, CPLX, 4);
    g.transB = true;
    g.conjA = true;
    g.batch = nblk;
    g.sA = g.sB = (long long)NB * n;
    g.sC = (long long)NB * NB;
    g.pA = g.pB = pV;
    g.pC = (long long)nS;
    gemm_f64(c, g, Vall, Vall, Sm);
  }
  btc_larft_k<<<nblk, 256, 0, c->stream>>>(Sm, (long long)nS, tau, nrefl, TtS, (long long)nT);
  DMRG_CHECK_HIP(hipGetLastError());
  const long long pE = (long long)keep * n;
  for (int b = nblk - 1; b >= 0; --b) {
    const int b0 = b * NB;
    const int len = n - b0 - 1;
    if (len <= 0) continue;
    int S = BT_SPLIT, kc = len, start = b0 + 1;
    for (; S >= 1; --S) {
      kc = (len + S - 1) / S;
      start = n - S * kc;
      if (start >= 0) break;
    }
    DMRG_REQUIRE(S >= 1 && start >= 0 && start <= b0 + 1, "unmtr: bad split");
    const double* Vb = Vall + (size_t)b * NB * n;
    GemmDesc g1 = gdesc(keep, NB, kc, n, n, S * NB, CPLX, 4);   // Y = Et conj(V_b)
    g1.transB = true;
    g1.conjB = true;
    g1.batch = S;
    g1.sA = g1.sB = kc;
    g1.sC = NB;
    g1.pA = pE;
    g1.pB = pV;
    g1.pC = (long long)nY;
    gemm_f64(c, g1, Et + start, Vb + start, Ycat);
    const double* Y = Ycat;
    if (S > 1) {   // sum the split products (both planes)
      double* Ys = Ycat + (size_t)keep * S * NB;
      const long long tot = (long long)keep * NB;
      bt_sum_k<<<(int)std::min<long long>((2 * tot + 255) / 256, 4096), 256, 0, c->stream>>>(
          Ycat, keep, NB, S, (long long)nY, Ys, (long long)nY, 2);
      DMRG_CHECK_HIP(hipGetLastError());
      Y = Ys;
    }
    GemmDesc g2 = gdesc(keep, NB, NB, NB, NB, NB, CPLX, 4);   // Y2 = Y T_b^T
    g2.pA = (long long)nY;
    g2.pB = (long long)nT;
    g2.pC = (long long)nY2;
    gemm_f64(c, g2, Y, TtS + (size_t)b * BT_SPLIT * NB * NB, Y2);
    GemmDesc g3 = gdesc(keep, len, NB, NB, n, n, CPLX, 4);   // Et -= Y2 V_b^T
    g3.alpha = -1.0;
    g3.beta = 1.0;
    g3.pA = (long long)nY2;
    g3.pB = pV;
    g3.pC = pE;
    gemm_f64(c, g3, Y2, Vb + b0 + 1, Et + b0 + 1);
  }
}

}  // namespace dmrg
