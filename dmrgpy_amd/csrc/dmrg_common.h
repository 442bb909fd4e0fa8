// Shared definitions for the MI355X (gfx950) DMRG/TDVP engine.
//
// Conventions used by every kernel and entry point in this library:
//  * row-major storage, 64-bit element offsets;
//  * "field": 0 = real, 1 = complex stored as two planes (re plane, then im
//    plane, separated by a plane stride); every complex operand is planar;
//  * "prec":  0 = FP64 contractions, 1 = FP32 contractions (FP32 MFMA with
//    FP32 accumulation), 2 = split-BF16x3 contractions (three BF16 MFMA
//    products hi*hi + hi*lo + lo*hi with FP32 accumulation).  Vectors handed
//    to the Krylov solvers, SVD inputs and outputs are always FP64.
#pragma once

#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include "dmrg_error.h"

#include <cstdint>
#include <string>
#include <vector>

namespace dmrg {

enum Field { REAL = 0, CPLX = 1 };
enum Prec { P64 = 0, P32 = 1, PB3 = 2 };

#define DMRG_CHECK_HIP(expr)                                                  \
  do {                                                                        \
    hipError_t e_ = (expr);                                                   \
    if (e_ != hipSuccess)                                                     \
      throw ::dmrg::Error(std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define DMRG_REQUIRE(cond, msg)                            \
  do {                                                     \
    if (!(cond)) throw ::dmrg::Error(std::string(msg));    \
  } while (0)

// A device buffer that only grows.  Growth synchronises the stream first so
// the old allocation is not freed under a running kernel.
struct DevBuf {
  void* ptr = nullptr;
  size_t bytes = 0;
  void* get(size_t need, hipStream_t s) {
    if (need > bytes) {
      if (ptr) {
        DMRG_CHECK_HIP(hipStreamSynchronize(s));
        DMRG_CHECK_HIP(hipFree(ptr));
        ptr = nullptr;
      }
      size_t nb = need + need / 8 + 256;
      DMRG_CHECK_HIP(hipMalloc(&ptr, nb));
      bytes = nb;
    }
    return ptr;
  }
  void release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
  }
};

// Workspace slots.  Each slot is owned by one stage of one algorithm, so
// nested calls never alias (a Lanczos loop's matvec uses the H_eff slots,
// the Lanczos basis uses its own).
enum Slot {
  WS_T1 = 0, WS_T2, WS_XIN, WS_YOUT, WS_WSMALL, WS_ENV1, WS_ENV2,
  WS_KRYLOV_Q, WS_KRYLOV_W, WS_KRYLOV_RED, WS_KRYLOV_COEF,
  WS_SVD_GRAM, WS_SVD_EV, WS_SVD_TMP, WS_SVD_TMP2, WS_SVD_INFO, WS_SVD_CPLX, WS_SVD_FIX, WS_SVD_RED, WS_PROJ, WS_B3_L, WS_B3_R,
  WS_BF16_A, WS_BF16_B, WS_REFINE,
  WS_GSV_Q, WS_GSV_B, WS_GSV_R, WS_GSV_S, WS_GSV_V, WS_GSV_T, WS_GSV_T2, WS_GSV_OU, WS_GSV_OV,
  WS_TRD, WS_TRD_Z, WS_TRD_ZC, WS_BT_V, WS_BT_W, WS_TRD_ZS,
  WS_NSLOTS
};

struct ProfEntry {
  hipEvent_t start, stop;
  double flops;
  double bytes;
  int kind;
};

struct Ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  rocblas_handle blas = nullptr;
  DevBuf ws[WS_NSLOTS];
  double* host_scalars = nullptr;  // pinned, for scalar read-back
  // profiling: H_eff applies (kind 0), environment updates (1), SVDs (2),
  // Krylov vector work (3), tridiagonalisations (4)
  bool prof_on = false;
  std::vector<ProfEntry> prof;
  std::vector<hipEvent_t> event_pool;
  double prof_time_ms[5] = {0, 0, 0, 0, 0};
  double prof_flops[5] = {0, 0, 0, 0, 0};
  double prof_bytes[5] = {0, 0, 0, 0, 0};   // algorithmic HBM bytes
  long long prof_calls[5] = {0, 0, 0, 0, 0};
  long long matvecs = 0;
  // dense eigensolver of the Gram route: 1 (default) = own two-launch
  // tridiagonalisation (real or complex Hermitian; a real reduction stops
  // once the trailing block is below TRD_STOP_TOL |G|_F) + rocSOLVER stedc +
  // the library's blocked back-transformation of the kept eigenvectors;
  // 3 = the same with the reduction always complete; 2 = complete reduction
  // with rocSOLVER ormtr / unmtr for the back-transformation; 0 = rocSOLVER
  // syevd / heevd
  int eig_mode = 1;
  hipEvent_t trd_ev[2] = {nullptr, nullptr};   // early-stop checks of the reduction in flight
  // allocation-free calls (dmrg_heff_apply_noalloc): while `noalloc` is set,
  // ws_get carves the workspace out of this device's static scratch array
  // (device address `scratch`, mapped at context creation) instead of the
  // growing DevBuf slots; nothing is allocated or freed then
  unsigned char* scratch = nullptr;
  bool noalloc = false;
  size_t scratch_used = 0;
  void* scratch_slot[WS_NSLOTS] = {};
  size_t scratch_slot_bytes[WS_NSLOTS] = {};
};

void* ws_get(Ctx* c, Slot s, size_t bytes);

// Early stop of the Gram matrices' tridiagonalisation (eigensolver mode 1):
// relative Frobenius norm below which the trailing block is taken as zero.
// Every eigenvalue moves by at most this times |G|_F (about 9 eps |G|_F),
// an order of magnitude below the rounding of forming G = M M^dag in FP64
// (~sqrt(k) eps tr(G) in Frobenius norm for inner dimension k ~ 10^3); the
// block's trace is added to the discarded weight, so the keep rule sees the
// whole spectrum's weight.  2e-16 was measured too: same keep counts, ~17%
// more reduced columns.
constexpr double TRD_STOP_TOL = 2e-15;

// ---- GEMM (gemm.hip) ------------------------------------------------------
// C[b] = alpha * op(A[b]) * op(B[b]) + beta * C[b], row-major, logical A is
// M x K, B is K x N.  transA: A stored K x M (element (m,k) at k*lda+m).
// transB: B stored N x K.  Complex (field==CPLX): planar operands, conjA /
// conjB negate the imaginary planes.  tag only distinguishes kernel names
// in profiles (0 generic, 1 H_eff L-step, 2 H_eff R-step, 3 env, 4 svd,
// 5 backward).
struct GemmDesc {
  int M, N, K;
  int lda, ldb, ldc;
  long long sA, sB, sC;      // batch strides (elements)
  long long pA, pB, pC;      // plane strides (elements), complex only
  int batch;
  bool transA, transB;
  bool conjA, conjB;
  double alpha, beta;
  int field;
  int tag;
  // C Hermitian by construction (A = B^dag products such as a Gram matrix):
  // only the tiles on or below the diagonal are computed; the caller
  // mirrors the strict upper triangle (mirror_lower)
  bool lower = false;
};
void gemm_f64(Ctx* c, const GemmDesc& g, const double* A, const double* B, double* C);
// the strict upper triangle of a Hermitian n x n result from its lower one
// (after a GemmDesc::lower product); plane: imaginary-plane stride
void mirror_lower(Ctx* c, double* C, int n, int ld, long long plane, int field);
void gemm_f32(Ctx* c, const GemmDesc& g, const float* A, const float* B, float* C);
// real FP32 operands, FP64 result (beta must be 0)
void gemm_f32_c64(Ctx* c, const GemmDesc& g, const float* A, const float* B, double* C);
// split-BF16x3: FP32 operands split on the fly into bf16 hi/lo pairs.
void gemm_bf16x3(Ctx* c, const GemmDesc& g, const float* A, const float* B, float* C);
// split-BF16x3 on pre-split operands: C[M][N] = A[M][K] * B[N][K]^T, A and B
// as (hi, lo) bf16 pairs with k-contiguous rows (pitches multiples of 8 and
// >= K, 16-byte aligned); the pitch padding may hold anything (masked).
// tag as in GemmDesc (1 H_eff L-step, 2 R-step).
void gemm_b3nt(Ctx* c, int M, int N, int K, const __bf16* Ah, const __bf16* Al, int lda, const __bf16* Bh,
               const __bf16* Bl, int ldb, float* C, int ldc, int tag);

}  // namespace dmrg
