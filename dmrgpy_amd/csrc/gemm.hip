// MFMA GEMM kernels for gfx950 (CDNA4): FP64 and FP32 via the 16x16x4
// MFMA shapes, split-BF16x3 via the 16x16x32 BF16 MFMA.  256 threads = four
// 64-wide waves in a 2x2 arrangement; every wave owns a (16*TM) x (16*TN)
// patch of the block tile, so a block computes (32*TM) x (32*TN).  Operand
// tiles are staged through LDS with the next tile prefetched into registers
// while the current one is multiplied.
//
// MFMA lane layout (16x16xK shapes): lane l supplies A[i = l%16][k-slot l/16]
// and B[k-slot l/16][j = l%16]; the accumulator element r (0..3) holds
// C[4*(l/16) + r][l%16] for the FP32 and BF16 shapes and C[4*r + l/16][l%16]
// for the FP64 shape.  The k-slots of one instruction may map to any 4 (FP32/FP64) or
// 32 (BF16, 8 per slot) distinct k values as long as A and B agree, so each
// lane reads 4 (resp. 8) consecutive k from LDS in one vector access.
#include "dmrg_common.h"
#include "tile_order.h"

namespace dmrg {

typedef double dbl4 __attribute__((ext_vector_type(4)));
typedef float flt4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

template <typename T> struct Mfma;
template <> struct Mfma<double> {
  typedef dbl4 acc_t;
  static __device__ __forceinline__ acc_t op(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
};
template <> struct Mfma<float> {
  typedef flt4 acc_t;
  static __device__ __forceinline__ acc_t op(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
};

struct GemmArgs {
  int M, N, K, lda, ldb, ldc;
  long long sA, sB, sC, pA, pB, pC;
  double alpha, beta;
  int conjA, conjB;
  int vec;     // 16-byte vector loads allowed (aligned pointers and strides)
  int lower;   // skip tiles strictly above the diagonal (Hermitian C)
};

// ---------------------------------------------------------------------------
// FP64 / FP32 kernel
// ---------------------------------------------------------------------------
template <typename T, bool CX, int TM, int TN, bool TA, bool TB, int TAG>
__global__ __launch_bounds__(256) void gemm_mfma(const T* __restrict__ A, const T* __restrict__ B,
                                                 T* __restrict__ C, GemmArgs g) {
  constexpr int BM = 32 * TM, BN = 32 * TN, BK = 16;
  constexpr int PAD = (sizeof(T) == 8) ? 2 : 4;
  constexpr int NP = CX ? 2 : 1;
  constexpr int LA = BM * BK / 256, LB = BN * BK / 256;
  typedef typename Mfma<T>::acc_t acc_t;

  __shared__ T As[NP][BM][BK + PAD];
  __shared__ T Bs[NP][BN][BK + PAD];

  const long long bz = blockIdx.z;
  A += bz * g.sA;
  B += bz * g.sB;
  C += bz * g.sC;
  const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
  if (g.lower && block_above_diagonal(m0, n0, BM)) return;   // whole block above the diagonal
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int li = lane & 15, lk = lane >> 4;
  const int M = g.M, N = g.N, K = g.K;

  acc_t acc[NP][TM][TN];
#pragma unroll
  for (int p = 0; p < NP; ++p)
#pragma unroll
    for (int a = 0; a < TM; ++a)
#pragma unroll
      for (int b = 0; b < TN; ++b) acc[p][a][b] = acc_t{0, 0, 0, 0};

  T ra[NP][LA], rb[NP][LB];
  const T sgnA = (CX && g.conjA) ? T(-1) : T(1);
  const T sgnB = (CX && g.conjB) ? T(-1) : T(1);

  auto load_regs = [&](int k0) {
#pragma unroll
    for (int e = 0; e < LA; ++e) {
      int idx = tid + 256 * e;
      int r, c;
      if (!TA) { r = idx / BK; c = idx % BK; } else { c = idx / BM; r = idx % BM; }
      int gm = m0 + r, gk = k0 + c;
      bool ok = (gm < M) && (gk < K);
      long long off = TA ? (long long)gk * g.lda + gm : (long long)gm * g.lda + gk;
#pragma unroll
      for (int p = 0; p < NP; ++p) ra[p][e] = ok ? A[off + p * g.pA] : T(0);
    }
#pragma unroll
    for (int e = 0; e < LB; ++e) {
      int idx = tid + 256 * e;
      int r, c;  // r: n within tile, c: k within tile
      if (!TB) { c = idx / BN; r = idx % BN; } else { r = idx / BK; c = idx % BK; }
      int gn = n0 + r, gk = k0 + c;
      bool ok = (gn < N) && (gk < K);
      long long off = TB ? (long long)gn * g.ldb + gk : (long long)gk * g.ldb + gn;
#pragma unroll
      for (int p = 0; p < NP; ++p) rb[p][e] = ok ? B[off + p * g.pB] : T(0);
    }
  };
  auto store_regs = [&]() {
#pragma unroll
    for (int e = 0; e < LA; ++e) {
      int idx = tid + 256 * e;
      int r, c;
      if (!TA) { r = idx / BK; c = idx % BK; } else { c = idx / BM; r = idx % BM; }
      As[0][r][c] = ra[0][e];
      if (CX) As[NP - 1][r][c] = sgnA * ra[NP - 1][e];
    }
#pragma unroll
    for (int e = 0; e < LB; ++e) {
      int idx = tid + 256 * e;
      int r, c;
      if (!TB) { c = idx / BN; r = idx % BN; } else { r = idx / BK; c = idx % BK; }
      Bs[0][r][c] = rb[0][e];
      if (CX) Bs[NP - 1][r][c] = sgnB * rb[NP - 1][e];
    }
  };

  load_regs(0);
  for (int k0 = 0; k0 < K; k0 += BK) {
    store_regs();
    __syncthreads();
    if (k0 + BK < K) load_regs(k0 + BK);
    T fa[NP][TM][4], fb[NP][TN][4];
#pragma unroll
    for (int p = 0; p < NP; ++p) {
#pragma unroll
      for (int a = 0; a < TM; ++a)
#pragma unroll
        for (int t = 0; t < 4; ++t) fa[p][a][t] = As[p][wm * TM * 16 + a * 16 + li][lk * 4 + t];
#pragma unroll
      for (int b = 0; b < TN; ++b)
#pragma unroll
        for (int t = 0; t < 4; ++t) fb[p][b][t] = Bs[p][wn * TN * 16 + b * 16 + li][lk * 4 + t];
    }
#pragma unroll
    for (int t = 0; t < 4; ++t) {
#pragma unroll
      for (int a = 0; a < TM; ++a) {
#pragma unroll
        for (int b = 0; b < TN; ++b) {
          if (!CX) {
            acc[0][a][b] = Mfma<T>::op(fa[0][a][t], fb[0][b][t], acc[0][a][b]);
          } else {
            acc[0][a][b] = Mfma<T>::op(fa[0][a][t], fb[0][b][t], acc[0][a][b]);
            acc[0][a][b] = Mfma<T>::op(-fa[NP - 1][a][t], fb[NP - 1][b][t], acc[0][a][b]);
            acc[NP - 1][a][b] = Mfma<T>::op(fa[0][a][t], fb[NP - 1][b][t], acc[NP - 1][a][b]);
            acc[NP - 1][a][b] = Mfma<T>::op(fa[NP - 1][a][t], fb[0][b][t], acc[NP - 1][a][b]);
          }
        }
      }
    }
    __syncthreads();
  }

  const T alpha = T(g.alpha), beta = T(g.beta);
#pragma unroll
  for (int a = 0; a < TM; ++a)
#pragma unroll
    for (int b = 0; b < TN; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        // accumulator row of element r: FP32 16x16x4 holds rows 4*(l/16)+r,
        // FP64 16x16x4 holds rows 4*r+(l/16) (measured lane layouts)
        const int rr = (sizeof(T) == 8) ? (4 * r + lk) : (lk * 4 + r);
        int row = m0 + wm * TM * 16 + a * 16 + rr;
        int col = n0 + wn * TN * 16 + b * 16 + li;
        if (row < M && col < N) {
          long long off = (long long)row * g.ldc + col;
#pragma unroll
          for (int p = 0; p < NP; ++p) {
            T v = alpha * acc[p][a][b][r];
            if (g.beta != 0.0) v += beta * C[off + p * g.pC];
            C[off + p * g.pC] = v;
          }
        }
      }
}

// ---------------------------------------------------------------------------
// Real FP64 / FP32 kernel (the H_eff and environment GEMMs): 16-byte global
// loads when the host found base pointers, leading dimensions and batch
// strides multiples of 16 bytes (g.vec), element loads otherwise;
// double-buffered LDS (one barrier per k-tile) and an XCD-aware tile order
// (xcd_tile, tile_order.h).
// ---------------------------------------------------------------------------

// Tile (row, column) of this workgroup in the XCD-grouped order of
// tile_order.h.
__device__ __forceinline__ void xcd_tile(int nbx, int nby, int& tr, int& tc) {
  xcd_tile_of(blockIdx.x, blockIdx.y, nbx, nby, tr, tc);
}
template <typename T> struct Vec;
template <> struct Vec<float> { typedef flt4 type; static constexpr int W = 4; };
template <> struct Vec<double> { typedef double dbl2 __attribute__((ext_vector_type(2))); typedef dbl2 type; static constexpr int W = 2; };

template <typename T, int TM, int TN, bool TA, bool TB, int TAG, typename CT = T>
__global__ __launch_bounds__(256) void gemm_fast(const T* __restrict__ A, const T* __restrict__ B,
                                                 CT* __restrict__ C, GemmArgs g) {
  constexpr int BM = 32 * TM, BN = 32 * TN, BK = 16;
  constexpr int PAD = 16 / sizeof(T);     // keeps rows 16-byte aligned
  constexpr int VW = Vec<T>::W;
  typedef typename Vec<T>::type vec_t;
  constexpr int NA = BM * BK / (VW * 256), NB = BN * BK / (VW * 256);
  typedef typename Mfma<T>::acc_t acc_t;

  // Each operand tile keeps the orientation of its global rows, so every
  // 16-byte global vector lands in LDS with one 16-byte store (no transposed
  // scalar stores and their bank conflicts): [m][k] / [n][k] when k is the
  // contiguous global index, [k][m] / [k][n] otherwise.  MFMA fragments read
  // 4 consecutive k either as one vector or as 4 scalars from 4 rows.
  constexpr int AR = TA ? BK : BM, AC = (TA ? BM : BK) + PAD;
  constexpr int BR = TB ? BN : BK, BC = (TB ? BK : BN) + PAD;
  __shared__ __attribute__((aligned(16))) T As[2][AR][AC];
  __shared__ __attribute__((aligned(16))) T Bs[2][BR][BC];

  const long long bz = blockIdx.z;
  A += bz * g.sA;
  B += bz * g.sB;
  C += bz * g.sC;
  int tr, tc;
  if (g.lower) {
    // Hermitian C: row-major tiles (every XCD gets an even share of the
    // lower-triangle tiles), those strictly above the diagonal exit at once
    const int lin = blockIdx.y * gridDim.x + blockIdx.x;
    tr = lin / gridDim.x;
    tc = lin % gridDim.x;
    if (tile_above_diagonal(tr, tc, BM, BN)) return;
  } else {
    xcd_tile(gridDim.x, gridDim.y, tr, tc);
  }
  const int m0 = tr * BM, n0 = tc * BN;

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int li = lane & 15, lk = lane >> 4;
  const int M = g.M, N = g.N, K = g.K;

  acc_t acc[TM][TN];
#pragma unroll
  for (int a = 0; a < TM; ++a)
#pragma unroll
    for (int b = 0; b < TN; ++b) acc[a][b] = acc_t{0, 0, 0, 0};

  vec_t ra[NA], rb[NB];

  auto load_regs = [&](int k0) {
#pragma unroll
    for (int e = 0; e < NA; ++e) {
      const int v = tid + 256 * e;
      int r, c;  // element (m = m0 + r, k = k0 + c) is the first of the vector
      if (!TA) { r = v / (BK / VW); c = (v % (BK / VW)) * VW; } else { c = v / (BM / VW); r = (v % (BM / VW)) * VW; }
      const int gm = m0 + r, gk = k0 + c;
      if (!TA) {
        const T* p = A + (long long)gm * g.lda + gk;
        if (g.vec && gm < M && gk + VW <= K) {
          ra[e] = *reinterpret_cast<const vec_t*>(p);
        } else {
#pragma unroll
          for (int j = 0; j < VW; ++j) ra[e][j] = (gm < M && gk + j < K) ? p[j] : T(0);
        }
      } else {
        const T* p = A + (long long)gk * g.lda + gm;
        if (g.vec && gk < K && gm + VW <= M) {
          ra[e] = *reinterpret_cast<const vec_t*>(p);
        } else {
#pragma unroll
          for (int j = 0; j < VW; ++j) ra[e][j] = (gk < K && gm + j < M) ? p[j] : T(0);
        }
      }
    }
#pragma unroll
    for (int e = 0; e < NB; ++e) {
      const int v = tid + 256 * e;
      int r, c;  // element (n = n0 + r, k = k0 + c)
      if (!TB) { c = v / (BN / VW); r = (v % (BN / VW)) * VW; } else { r = v / (BK / VW); c = (v % (BK / VW)) * VW; }
      const int gn = n0 + r, gk = k0 + c;
      if (!TB) {
        const T* p = B + (long long)gk * g.ldb + gn;
        if (g.vec && gk < K && gn + VW <= N) {
          rb[e] = *reinterpret_cast<const vec_t*>(p);
        } else {
#pragma unroll
          for (int j = 0; j < VW; ++j) rb[e][j] = (gk < K && gn + j < N) ? p[j] : T(0);
        }
      } else {
        const T* p = B + (long long)gn * g.ldb + gk;
        if (g.vec && gn < N && gk + VW <= K) {
          rb[e] = *reinterpret_cast<const vec_t*>(p);
        } else {
#pragma unroll
          for (int j = 0; j < VW; ++j) rb[e][j] = (gn < N && gk + j < K) ? p[j] : T(0);
        }
      }
    }
  };
  auto store_regs = [&](int buf) {
#pragma unroll
    for (int e = 0; e < NA; ++e) {
      const int v = tid + 256 * e;
      if (!TA) {
        const int r = v / (BK / VW), c = (v % (BK / VW)) * VW;     // As[m][k]
        *reinterpret_cast<vec_t*>(&As[buf][r][c]) = ra[e];
      } else {
        const int c = v / (BM / VW), r = (v % (BM / VW)) * VW;     // As[k][m]
        *reinterpret_cast<vec_t*>(&As[buf][c][r]) = ra[e];
      }
    }
#pragma unroll
    for (int e = 0; e < NB; ++e) {
      const int v = tid + 256 * e;
      if (!TB) {
        const int c = v / (BN / VW), r = (v % (BN / VW)) * VW;     // Bs[k][n]
        *reinterpret_cast<vec_t*>(&Bs[buf][c][r]) = rb[e];
      } else {
        const int r = v / (BK / VW), c = (v % (BK / VW)) * VW;     // Bs[n][k]
        *reinterpret_cast<vec_t*>(&Bs[buf][r][c]) = rb[e];
      }
    }
  };

  // FP64 64 x 64 tiles with beta != 0 (among them the tridiagonalisation's
  // trailing updates, inner dimension 64): the C tile is loaded before the
  // k-loop, so its read does not wait behind the MFMA work in the epilogue
  // (the larger FP64 tile keeps its two waves per SIMD without it)
  constexpr bool CPRE = sizeof(T) == 8 && sizeof(CT) == 8 && TM * TN <= 4;
  CT cpre[CPRE ? TM : 1][CPRE ? TN : 1][4];
  const bool pre = CPRE && g.beta != 0.0;
  if (CPRE && pre) {
#pragma unroll
    for (int a = 0; a < (CPRE ? TM : 1); ++a)
#pragma unroll
      for (int b = 0; b < (CPRE ? TN : 1); ++b)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int rr = (sizeof(T) == 8) ? (4 * r + lk) : (lk * 4 + r);
          const int row = m0 + wm * TM * 16 + a * 16 + rr;
          const int col = n0 + wn * TN * 16 + b * 16 + li;
          cpre[a][b][r] = (row < M && col < N) ? C[(long long)row * g.ldc + col] : CT(0);
        }
  }

  load_regs(0);
  store_regs(0);
  __syncthreads();
  int cur = 0;
  for (int k0 = 0; k0 < K; k0 += BK) {
    const bool more = k0 + BK < K;
    if (more) load_regs(k0 + BK);
    T fa[TM][4], fb[TN][4];
#pragma unroll
    for (int a = 0; a < TM; ++a) {
      const int row = wm * TM * 16 + a * 16 + li;
#pragma unroll
      for (int t = 0; t < 4; ++t) fa[a][t] = TA ? As[cur][lk * 4 + t][row] : As[cur][row][lk * 4 + t];
    }
#pragma unroll
    for (int b = 0; b < TN; ++b) {
      const int col = wn * TN * 16 + b * 16 + li;
#pragma unroll
      for (int t = 0; t < 4; ++t) fb[b][t] = TB ? Bs[cur][col][lk * 4 + t] : Bs[cur][lk * 4 + t][col];
    }
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
      for (int a = 0; a < TM; ++a)
#pragma unroll
        for (int b = 0; b < TN; ++b) acc[a][b] = Mfma<T>::op(fa[a][t], fb[b][t], acc[a][b]);
    if (more) store_regs(cur ^ 1);
    __syncthreads();
    cur ^= 1;
  }

  const T alpha = T(g.alpha), beta = T(g.beta);
#pragma unroll
  for (int a = 0; a < TM; ++a)
#pragma unroll
    for (int b = 0; b < TN; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int rr = (sizeof(T) == 8) ? (4 * r + lk) : (lk * 4 + r);
        const int row = m0 + wm * TM * 16 + a * 16 + rr;
        const int col = n0 + wn * TN * 16 + b * 16 + li;
        if (row < M && col < N) {
          const long long off = (long long)row * g.ldc + col;
          T v = alpha * acc[a][b][r];
          if (g.beta != 0.0) {
            T cold;
            if constexpr (CPRE) cold = (T)cpre[CPRE ? a : 0][CPRE ? b : 0][r];
            else cold = (T)C[off];
            v += beta * cold;
          }
          C[off] = (CT)v;
        }
      }
}

// ---------------------------------------------------------------------------
// split-BF16x3 kernel: FP32 operands, x = hi + lo with hi = bf16(x) and
// lo = bf16(x - hi); C += Ahi*Bhi + Ahi*Blo + Alo*Bhi with FP32 accumulation.
// The split happens once per element when the tile is written to LDS.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void split_bf16(float x, __bf16& hi, __bf16& lo) {
  hi = (__bf16)x;
  lo = (__bf16)(x - (float)hi);
}

template <bool CX, int TM, int TN, bool TA, bool TB, int TAG>
__global__ __launch_bounds__(256) void gemm_bf16x3_k(const float* __restrict__ A, const float* __restrict__ B,
                                                     float* __restrict__ C, GemmArgs g) {
  constexpr int BM = 32 * TM, BN = 32 * TN, BK = 32;
  constexpr int PAD = 8;
  constexpr int NP = CX ? 2 : 1;
  constexpr int LA = BM * BK / 256, LB = BN * BK / 256;

  __shared__ __bf16 Ah[NP][BM][BK + PAD], Al[NP][BM][BK + PAD];
  __shared__ __bf16 Bh[NP][BN][BK + PAD], Bl[NP][BN][BK + PAD];

  const long long bz = blockIdx.z;
  A += bz * g.sA;
  B += bz * g.sB;
  C += bz * g.sC;
  const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int li = lane & 15, lk = lane >> 4;
  const int M = g.M, N = g.N, K = g.K;

  flt4 acc[NP][TM][TN];
#pragma unroll
  for (int p = 0; p < NP; ++p)
#pragma unroll
    for (int a = 0; a < TM; ++a)
#pragma unroll
      for (int b = 0; b < TN; ++b) acc[p][a][b] = flt4{0, 0, 0, 0};

  float ra[NP][LA], rb[NP][LB];
  const float sgnA = (CX && g.conjA) ? -1.f : 1.f;
  const float sgnB = (CX && g.conjB) ? -1.f : 1.f;

  auto load_regs = [&](int k0) {
#pragma unroll
    for (int e = 0; e < LA; ++e) {
      int idx = tid + 256 * e;
      int r, c;
      if (!TA) { r = idx / BK; c = idx % BK; } else { c = idx / BM; r = idx % BM; }
      int gm = m0 + r, gk = k0 + c;
      bool ok = (gm < M) && (gk < K);
      long long off = TA ? (long long)gk * g.lda + gm : (long long)gm * g.lda + gk;
#pragma unroll
      for (int p = 0; p < NP; ++p) ra[p][e] = ok ? A[off + p * g.pA] : 0.f;
    }
#pragma unroll
    for (int e = 0; e < LB; ++e) {
      int idx = tid + 256 * e;
      int r, c;
      if (!TB) { c = idx / BN; r = idx % BN; } else { r = idx / BK; c = idx % BK; }
      int gn = n0 + r, gk = k0 + c;
      bool ok = (gn < N) && (gk < K);
      long long off = TB ? (long long)gn * g.ldb + gk : (long long)gk * g.ldb + gn;
#pragma unroll
      for (int p = 0; p < NP; ++p) rb[p][e] = ok ? B[off + p * g.pB] : 0.f;
    }
  };
  auto store_regs = [&]() {
#pragma unroll
    for (int e = 0; e < LA; ++e) {
      int idx = tid + 256 * e;
      int r, c;
      if (!TA) { r = idx / BK; c = idx % BK; } else { c = idx / BM; r = idx % BM; }
#pragma unroll
      for (int p = 0; p < NP; ++p) {
        float x = (p == 1) ? sgnA * ra[p][e] : ra[p][e];
        __bf16 h, l;
        split_bf16(x, h, l);
        Ah[p][r][c] = h;
        Al[p][r][c] = l;
      }
    }
#pragma unroll
    for (int e = 0; e < LB; ++e) {
      int idx = tid + 256 * e;
      int r, c;
      if (!TB) { c = idx / BN; r = idx % BN; } else { r = idx / BK; c = idx % BK; }
#pragma unroll
      for (int p = 0; p < NP; ++p) {
        float x = (p == 1) ? sgnB * rb[p][e] : rb[p][e];
        __bf16 h, l;
        split_bf16(x, h, l);
        Bh[p][r][c] = h;
        Bl[p][r][c] = l;
      }
    }
  };

  load_regs(0);
  for (int k0 = 0; k0 < K; k0 += BK) {
    store_regs();
    __syncthreads();
    if (k0 + BK < K) load_regs(k0 + BK);
    bf16x8 ah[NP][TM], al[NP][TM], bh[NP][TN], bl[NP][TN];
#pragma unroll
    for (int p = 0; p < NP; ++p) {
#pragma unroll
      for (int a = 0; a < TM; ++a) {
        int row = wm * TM * 16 + a * 16 + li;
        ah[p][a] = *reinterpret_cast<const bf16x8*>(&Ah[p][row][lk * 8]);
        al[p][a] = *reinterpret_cast<const bf16x8*>(&Al[p][row][lk * 8]);
      }
#pragma unroll
      for (int b = 0; b < TN; ++b) {
        int row = wn * TN * 16 + b * 16 + li;
        bh[p][b] = *reinterpret_cast<const bf16x8*>(&Bh[p][row][lk * 8]);
        bl[p][b] = *reinterpret_cast<const bf16x8*>(&Bl[p][row][lk * 8]);
      }
    }
#pragma unroll
    for (int a = 0; a < TM; ++a)
#pragma unroll
      for (int b = 0; b < TN; ++b) {
        if (!CX) {
          acc[0][a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[0][a], bh[0][b], acc[0][a][b], 0, 0, 0);
          acc[0][a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[0][a], bl[0][b], acc[0][a][b], 0, 0, 0);
          acc[0][a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[0][a], bh[0][b], acc[0][a][b], 0, 0, 0);
        } else {
          bf16x8 nah = -ah[1][a], nal = -al[1][a];
          acc[0][a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[0][a], bh[0][b], acc[0][a][b], 0, 0, 0);
          acc[0][a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[0][a], bl[0][b], acc[0][a][b], 0, 0, 0);
          acc[0][a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[0][a], bh[0][b], acc[0][a][b], 0, 0, 0);
          acc[0][a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(nal, bh[1][b], acc[0][a][b], 0, 0, 0);
          acc[0][a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(nah, bl[1][b], acc[0][a][b], 0, 0, 0);
          acc[0][a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(nah, bh[1][b], acc[0][a][b], 0, 0, 0);
          acc[1][a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[0][a], bh[1][b], acc[1][a][b], 0, 0, 0);
          acc[1][a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[0][a], bl[1][b], acc[1][a][b], 0, 0, 0);
          acc[1][a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[0][a], bh[1][b], acc[1][a][b], 0, 0, 0);
          acc[1][a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[1][a], bh[0][b], acc[1][a][b], 0, 0, 0);
          acc[1][a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[1][a], bl[0][b], acc[1][a][b], 0, 0, 0);
          acc[1][a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[1][a], bh[0][b], acc[1][a][b], 0, 0, 0);
        }
      }
    __syncthreads();
  }

  const float alpha = (float)g.alpha, beta = (float)g.beta;
#pragma unroll
  for (int a = 0; a < TM; ++a)
#pragma unroll
    for (int b = 0; b < TN; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        int row = m0 + wm * TM * 16 + a * 16 + lk * 4 + r;
        int col = n0 + wn * TN * 16 + b * 16 + li;
        if (row < M && col < N) {
          long long off = (long long)row * g.ldc + col;
#pragma unroll
          for (int p = 0; p < NP; ++p) {
            float v = alpha * acc[p][a][b][r];
            if (g.beta != 0.0) v += beta * C[off + p * g.pC];
            C[off + p * g.pC] = v;
          }
        }
      }
}

// ---------------------------------------------------------------------------
// split-BF16x3 kernel on pre-split operands: C[m][n] = sum_k A[m][k] B[n][k]
// with A and B each given as a (hi, lo) pair of bf16 arrays whose rows are
// k-contiguous (row pitches multiples of 8 elements, 16-byte aligned), so
// every global load and every LDS store is a 16-byte vector and each MFMA
// operand fragment (8 consecutive k) is one 16-byte LDS read.  The split
// happens once per element outside the GEMM (environments once per bond,
// theta once per apply, the MPO pass writes T2 split).  128x128 block tile,
// BK = 32 (one v_mfma_f32_16x16x32_bf16 k-step), double-buffered LDS with
// one barrier per k-step, XCD-aware tile order; C is FP32.
// ---------------------------------------------------------------------------
struct B3Args {
  int M, N, K, lda, ldb, ldc;
};

template <int TM, int TN, int TAG>
__global__ __launch_bounds__(256) void gemm_b3nt_k(const __bf16* __restrict__ Ah, const __bf16* __restrict__ Al,
                                                   const __bf16* __restrict__ Bh, const __bf16* __restrict__ Bl,
                                                   float* __restrict__ C, B3Args g) {
  constexpr int BM = 32 * TM, BN = 32 * TN, BK = 32, LDK = BK + 8;   // row pitch 80 bytes
  constexpr int NA = BM * BK / (8 * 256), NB = BN * BK / (8 * 256); // 16-byte vectors per thread
  __shared__ __attribute__((aligned(16))) __bf16 sAh[2][BM][LDK], sAl[2][BM][LDK];
  __shared__ __attribute__((aligned(16))) __bf16 sBh[2][BN][LDK], sBl[2][BN][LDK];
  int tr, tc;
  xcd_tile(gridDim.x, gridDim.y, tr, tc);
  const int m0 = tr * BM, n0 = tc * BN;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int li = lane & 15, lk = lane >> 4;
  const int M = g.M, N = g.N, K = g.K;
  flt4 acc[TM][TN];
#pragma unroll
  for (int a = 0; a < TM; ++a)
#pragma unroll
    for (int b = 0; b < TN; ++b) acc[a][b] = flt4{0.f, 0.f, 0.f, 0.f};
  bf16x8 rah[NA], ral[NA], rbh[NB], rbl[NB];
  const bf16x8 zero8 = bf16x8{};
  auto load = [&](int k0) {
#pragma unroll
    for (int e = 0; e < NA; ++e) {
      const int v = tid + 256 * e, r = v / (BK / 8), c = (v % (BK / 8)) * 8;
      const int gm = m0 + r, gk = k0 + c;
      if (gm < M && gk < K) {   // the row pitch covers the vector; k >= K is masked
        const long long off = (long long)gm * g.lda + gk;
        rah[e] = *reinterpret_cast<const bf16x8*>(Ah + off);
        ral[e] = *reinterpret_cast<const bf16x8*>(Al + off);
        if (gk + 8 > K) {
#pragma unroll
          for (int j = 0; j < 8; ++j)
            if (gk + j >= K) { rah[e][j] = (__bf16)0.f; ral[e][j] = (__bf16)0.f; }
        }
      } else {
        rah[e] = zero8;
        ral[e] = zero8;
      }
    }
#pragma unroll
    for (int e = 0; e < NB; ++e) {
      const int v = tid + 256 * e, r = v / (BK / 8), c = (v % (BK / 8)) * 8;
      const int gn = n0 + r, gk = k0 + c;
      if (gn < N && gk < K) {
        const long long off = (long long)gn * g.ldb + gk;
        rbh[e] = *reinterpret_cast<const bf16x8*>(Bh + off);
        rbl[e] = *reinterpret_cast<const bf16x8*>(Bl + off);
        if (gk + 8 > K) {
#pragma unroll
          for (int j = 0; j < 8; ++j)
            if (gk + j >= K) { rbh[e][j] = (__bf16)0.f; rbl[e][j] = (__bf16)0.f; }
        }
      } else {
        rbh[e] = zero8;
        rbl[e] = zero8;
      }
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int e = 0; e < NA; ++e) {
      const int v = tid + 256 * e, r = v / (BK / 8), c = (v % (BK / 8)) * 8;
      *reinterpret_cast<bf16x8*>(&sAh[buf][r][c]) = rah[e];
      *reinterpret_cast<bf16x8*>(&sAl[buf][r][c]) = ral[e];
    }
#pragma unroll
    for (int e = 0; e < NB; ++e) {
      const int v = tid + 256 * e, r = v / (BK / 8), c = (v % (BK / 8)) * 8;
      *reinterpret_cast<bf16x8*>(&sBh[buf][r][c]) = rbh[e];
      *reinterpret_cast<bf16x8*>(&sBl[buf][r][c]) = rbl[e];
    }
  };
  load(0);
  store(0);
  __syncthreads();
  int cur = 0;
  for (int k0 = 0; k0 < K; k0 += BK) {
    const bool more = k0 + BK < K;
    if (more) load(k0 + BK);
    bf16x8 ah[TM], al[TM], bh[TN], bl[TN];
#pragma unroll
    for (int a = 0; a < TM; ++a) {
      const int row = wm * TM * 16 + a * 16 + li;
      ah[a] = *reinterpret_cast<const bf16x8*>(&sAh[cur][row][lk * 8]);
      al[a] = *reinterpret_cast<const bf16x8*>(&sAl[cur][row][lk * 8]);
    }
#pragma unroll
    for (int b = 0; b < TN; ++b) {
      const int col = wn * TN * 16 + b * 16 + li;
      bh[b] = *reinterpret_cast<const bf16x8*>(&sBh[cur][col][lk * 8]);
      bl[b] = *reinterpret_cast<const bf16x8*>(&sBl[cur][col][lk * 8]);
    }
#pragma unroll
    for (int a = 0; a < TM; ++a)
#pragma unroll
      for (int b = 0; b < TN; ++b) {
        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[a], bh[b], acc[a][b], 0, 0, 0);
        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[a], bl[b], acc[a][b], 0, 0, 0);
        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[a], bh[b], acc[a][b], 0, 0, 0);
      }
    if (more) store(cur ^ 1);
    __syncthreads();
    cur ^= 1;
  }
#pragma unroll
  for (int a = 0; a < TM; ++a)
#pragma unroll
    for (int b = 0; b < TN; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = m0 + wm * TM * 16 + a * 16 + lk * 4 + r;
        const int col = n0 + wn * TN * 16 + b * 16 + li;
        if (row < M && col < N) C[(long long)row * g.ldc + col] = acc[a][b][r];
      }
}

template <int TM, int TN>
static void launch_b3nt(Ctx* c, dim3 grid, int tag, const __bf16* Ah, const __bf16* Al, const __bf16* Bh,
                        const __bf16* Bl, float* C, const B3Args& g) {
  if (tag == 1) gemm_b3nt_k<TM, TN, 1><<<grid, 256, 0, c->stream>>>(Ah, Al, Bh, Bl, C, g);
  else if (tag == 2) gemm_b3nt_k<TM, TN, 2><<<grid, 256, 0, c->stream>>>(Ah, Al, Bh, Bl, C, g);
  else gemm_b3nt_k<TM, TN, 0><<<grid, 256, 0, c->stream>>>(Ah, Al, Bh, Bl, C, g);
}

void gemm_b3nt(Ctx* c, int M, int N, int K, const __bf16* Ah, const __bf16* Al, int lda, const __bf16* Bh,
               const __bf16* Bl, int ldb, float* C, int ldc, int tag) {
  DMRG_REQUIRE(M > 0 && N > 0 && K > 0, "gemm_b3nt: bad sizes");
  DMRG_REQUIRE(lda % 8 == 0 && ldb % 8 == 0 && lda >= K && ldb >= K, "gemm_b3nt: pitches must be multiples of 8");
  auto al = [](const void* p) { return ((uintptr_t)p & 15) == 0; };
  DMRG_REQUIRE(al(Ah) && al(Al) && al(Bh) && al(Bl), "gemm_b3nt: operands must be 16-byte aligned");
  B3Args g{M, N, K, lda, ldb, ldc};
  const long long tiles128 = (long long)((M + 127) / 128) * ((N + 127) / 128);
  if (tiles128 >= 256) {
    dim3 grid((N + 127) / 128, (M + 127) / 128);
    DMRG_REQUIRE(grid.y <= 65535, "gemm_b3nt: M too large");
    launch_b3nt<4, 4>(c, grid, tag, Ah, Al, Bh, Bl, C, g);
  } else {
    dim3 grid((N + 63) / 64, (M + 63) / 64);
    DMRG_REQUIRE(grid.y <= 65535, "gemm_b3nt: M too large");
    launch_b3nt<2, 2>(c, grid, tag, Ah, Al, Bh, Bl, C, g);
  }
  DMRG_CHECK_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// host dispatch
// ---------------------------------------------------------------------------
static GemmArgs make_args(const GemmDesc& d) {
  GemmArgs g;
  g.M = d.M; g.N = d.N; g.K = d.K;
  g.lda = d.lda; g.ldb = d.ldb; g.ldc = d.ldc;
  g.sA = d.sA; g.sB = d.sB; g.sC = d.sC;
  g.pA = d.pA; g.pB = d.pB; g.pC = d.pC;
  g.alpha = d.alpha; g.beta = d.beta;
  g.conjA = d.conjA; g.conjB = d.conjB;
  g.vec = 0;
  g.lower = d.lower ? 1 : 0;
  return g;
}

static void check_desc(const GemmDesc& d) {
  DMRG_REQUIRE(d.M >= 0 && d.N >= 0 && d.K >= 0 && d.batch >= 1, "gemm: negative size");
  DMRG_REQUIRE(d.batch <= 65535, "gemm: batch too large");
  DMRG_REQUIRE(d.ldc >= d.N || d.M <= 1, "gemm: ldc < N");
  DMRG_REQUIRE(d.transA ? (d.lda >= d.M || d.K <= 1) : (d.lda >= d.K || d.M <= 1), "gemm: bad lda");
  DMRG_REQUIRE(d.transB ? (d.ldb >= d.K || d.N <= 1) : (d.ldb >= d.N || d.K <= 1), "gemm: bad ldb");
}

// Launch helper: picks the instantiation for (transA, transB, tag).
#define DMRG_GEMM_TAGS(X) X(0) X(1) X(2) X(3) X(4) X(5)

template <typename T, bool CX, int TM, int TN, bool TA, bool TB>
static void launch_tagged(Ctx* c, const GemmDesc& d, const T* A, const T* B, T* C) {
  GemmArgs g = make_args(d);
  dim3 grid((d.N + 32 * TN - 1) / (32 * TN), (d.M + 32 * TM - 1) / (32 * TM), d.batch);
  DMRG_REQUIRE(grid.y <= 65535, "gemm: M too large for one launch");
  switch (d.tag) {
#define CASE(t) \
  case t: gemm_mfma<T, CX, TM, TN, TA, TB, t><<<grid, 256, 0, c->stream>>>(A, B, C, g); break;
    DMRG_GEMM_TAGS(CASE)
#undef CASE
    default: throw Error("gemm: bad tag");
  }
  DMRG_CHECK_HIP(hipGetLastError());
}

template <typename T, bool CX, int TM, int TN>
static void launch_tt(Ctx* c, const GemmDesc& d, const T* A, const T* B, T* C) {
  if (!d.transA && !d.transB) launch_tagged<T, CX, TM, TN, false, false>(c, d, A, B, C);
  else if (d.transA && !d.transB) launch_tagged<T, CX, TM, TN, true, false>(c, d, A, B, C);
  else if (!d.transA && d.transB) launch_tagged<T, CX, TM, TN, false, true>(c, d, A, B, C);
  else launch_tagged<T, CX, TM, TN, true, true>(c, d, A, B, C);
}

template <typename T, int TM, int TN, bool TA, bool TB, typename CT = T>
static void launch_fast_tagged(Ctx* c, const GemmDesc& d, const T* A, const T* B, CT* C, bool vec) {
  GemmArgs g = make_args(d);
  g.vec = vec ? 1 : 0;
  dim3 grid((d.N + 32 * TN - 1) / (32 * TN), (d.M + 32 * TM - 1) / (32 * TM), d.batch);
  DMRG_REQUIRE(grid.y <= 65535, "gemm: M too large for one launch");
  switch (d.tag) {
#define CASE(t) \
  case t: gemm_fast<T, TM, TN, TA, TB, t, CT><<<grid, 256, 0, c->stream>>>(A, B, C, g); break;
    DMRG_GEMM_TAGS(CASE)
#undef CASE
    default: throw Error("gemm: bad tag");
  }
  DMRG_CHECK_HIP(hipGetLastError());
}

template <typename T, int TM, int TN, typename CT = T>
static void launch_fast(Ctx* c, const GemmDesc& d, const T* A, const T* B, CT* C, bool vec) {
  if (!d.transA && !d.transB) launch_fast_tagged<T, TM, TN, false, false, CT>(c, d, A, B, C, vec);
  else if (d.transA && !d.transB) launch_fast_tagged<T, TM, TN, true, false, CT>(c, d, A, B, C, vec);
  else if (!d.transA && d.transB) launch_fast_tagged<T, TM, TN, false, true, CT>(c, d, A, B, C, vec);
  else launch_fast_tagged<T, TM, TN, true, true, CT>(c, d, A, B, C, vec);
}

// 16-byte vector loads are legal when every row start is 16-byte aligned
template <typename T>
static bool vec_ok(const GemmDesc& d, const T* A, const T* B) {
  const int vw = 16 / (int)sizeof(T);
  auto al = [](const void* p) { return ((uintptr_t)p & 15) == 0; };
  return al(A) && al(B) && d.lda % vw == 0 && d.ldb % vw == 0 && d.sA % vw == 0 && d.sB % vw == 0;
}

// Real FP32 GEMM writing an FP64 result (the R-step of the FP32 H_eff apply
// writes y straight into the FP64 Krylov vector, with no conversion pass).
void gemm_f32_c64(Ctx* c, const GemmDesc& d, const float* A, const float* B, double* C) {
  DMRG_REQUIRE(d.field == REAL && d.beta == 0.0, "gemm_f32_c64: real field and beta = 0 only");
  check_desc(d);
  if (d.M == 0 || d.N == 0) return;
  const long long tiles128 = (long long)((d.M + 127) / 128) * ((d.N + 127) / 128) * d.batch;
  const int vw = 4;
  auto al = [](const void* p) { return ((uintptr_t)p & 15) == 0; };
  const bool vec = al(A) && al(B) && d.lda % vw == 0 && d.ldb % vw == 0 && d.sA % vw == 0 && d.sB % vw == 0;
  if (tiles128 >= 256) launch_fast<float, 4, 4, double>(c, d, A, B, C, vec);
  else launch_fast<float, 2, 2, double>(c, d, A, B, C, vec);
}

// Tile choice: large problems use 128x128 (real) / 64x64 (complex) tiles;
// problems that would leave the 256 CUs mostly idle drop to 64x64 / 32x32.
// Real operands always take the fast kernel; complex ones the planar one.
template <typename T>
static void gemm_any(Ctx* c, const GemmDesc& d, const T* A, const T* B, T* C) {
  check_desc(d);
  if (d.M == 0 || d.N == 0) return;
  long long tiles128 = (long long)((d.M + 127) / 128) * ((d.N + 127) / 128) * d.batch;
  if (d.field == REAL) {
    const bool vec = vec_ok(d, A, B);
    // FP64 uses 128x64 tiles: a 128x128 FP64 tile needs ~380 registers per
    // lane (one wave per SIMD); 128x64 keeps two
    if (tiles128 >= 256) {
      if (sizeof(T) == 8) launch_fast<T, 4, 2>(c, d, A, B, C, vec);
      else launch_fast<T, 4, 4>(c, d, A, B, C, vec);
    } else {
      launch_fast<T, 2, 2>(c, d, A, B, C, vec);
    }
    return;
  }
  long long tiles64 = (long long)((d.M + 63) / 64) * ((d.N + 63) / 64) * d.batch;
  if (tiles64 >= 512) launch_tt<T, true, 2, 2>(c, d, A, B, C);
  else launch_tt<T, true, 1, 1>(c, d, A, B, C);
}

// C[i][j] = conj(C[j][i]) for j > i (n x n, row pitch ld, imaginary plane at
// +plane when cx): the strict upper triangle of a Hermitian product computed
// with GemmDesc::lower.  One 32 x 32 tile pair per block: the lower tile
// (J, I) is read coalesced, transposed through LDS and written as (I, J).
__global__ __launch_bounds__(256) void mirror_lower_k(double* __restrict__ C, int n, int ld, long long plane,
                                                      int cx) {
  __shared__ double t[32][33];
  const int I = blockIdx.y, J = blockIdx.x;   // destination tile (I, J), I <= J
  if (I > J) return;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;   // 32 x 8
  for (int pl = 0; pl < (cx ? 2 : 1); ++pl) {
    double* P = C + pl * plane;
    for (int r = ty; r < 32; r += 8) {   // source rows J*32 + r, columns I*32 + tx
      const int sr = J * 32 + r, sc = I * 32 + tx;
      t[r][tx] = (sr < n && sc < n) ? P[(size_t)sr * ld + sc] : 0.0;
    }
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {   // destination rows I*32 + r, columns J*32 + tx
      const int dr = I * 32 + r, dc = J * 32 + tx;
      if (dr < n && dc < n && dc > dr) P[(size_t)dr * ld + dc] = pl ? -t[tx][r] : t[tx][r];
    }
    __syncthreads();
  }
}

void mirror_lower(Ctx* c, double* C, int n, int ld, long long plane, int field) {
  if (n <= 1) return;
  const int nt = (n + 31) / 32;
  mirror_lower_k<<<dim3(nt, nt), 256, 0, c->stream>>>(C, n, ld, plane, field == CPLX ? 1 : 0);
  DMRG_CHECK_HIP(hipGetLastError());
}

void gemm_f64(Ctx* c, const GemmDesc& d, const double* A, const double* B, double* C) {
  gemm_any<double>(c, d, A, B, C);
}
void gemm_f32(Ctx* c, const GemmDesc& d, const float* A, const float* B, float* C) {
  gemm_any<float>(c, d, A, B, C);
}

template <bool CX, int TM, int TN, bool TA, bool TB>
static void launch_b3_tagged(Ctx* c, const GemmDesc& d, const float* A, const float* B, float* C) {
  GemmArgs g = make_args(d);
  dim3 grid((d.N + 32 * TN - 1) / (32 * TN), (d.M + 32 * TM - 1) / (32 * TM), d.batch);
  DMRG_REQUIRE(grid.y <= 65535, "gemm: M too large for one launch");
  switch (d.tag) {
#define CASE(t) \
  case t: gemm_bf16x3_k<CX, TM, TN, TA, TB, t><<<grid, 256, 0, c->stream>>>(A, B, C, g); break;
    DMRG_GEMM_TAGS(CASE)
#undef CASE
    default: throw Error("gemm: bad tag");
  }
  DMRG_CHECK_HIP(hipGetLastError());
}

template <bool CX, int TM, int TN>
static void launch_b3_tt(Ctx* c, const GemmDesc& d, const float* A, const float* B, float* C) {
  if (!d.transA && !d.transB) launch_b3_tagged<CX, TM, TN, false, false>(c, d, A, B, C);
  else if (d.transA && !d.transB) launch_b3_tagged<CX, TM, TN, true, false>(c, d, A, B, C);
  else if (!d.transA && d.transB) launch_b3_tagged<CX, TM, TN, false, true>(c, d, A, B, C);
  else launch_b3_tagged<CX, TM, TN, true, true>(c, d, A, B, C);
}

void gemm_bf16x3(Ctx* c, const GemmDesc& d, const float* A, const float* B, float* C) {
  check_desc(d);
  if (d.M == 0 || d.N == 0) return;
  long long tiles128 = (long long)((d.M + 127) / 128) * ((d.N + 127) / 128) * d.batch;
  if (d.field == CPLX) {
    long long tiles64 = (long long)((d.M + 63) / 64) * ((d.N + 63) / 64) * d.batch;
    if (tiles64 >= 512) launch_b3_tt<true, 2, 2>(c, d, A, B, C);
    else launch_b3_tt<true, 1, 1>(c, d, A, B, C);
  } else {
    if (tiles128 >= 256) launch_b3_tt<false, 4, 4>(c, d, A, B, C);
    else launch_b3_tt<false, 2, 2>(c, d, A, B, C);
  }
}

}  // namespace dmrg
