// Internal declarations shared by the HIP translation units of the library.
#pragma once

#include <vector>

#include "dmrg_common.h"
#include "../../include/dmrg_hip.h"
#include "heff_work.h"

namespace dmrg {

struct SmallOp {
  const void* W = nullptr;     // [KoutPad][Kin] (+ imag plane), contraction precision
  const int* ioff = nullptr;   // [Kin]
  const int* ooff = nullptr;   // [Kout]
  int Kin = 0, Kout = 0, KoutPad = 0;
  bool wcplx = false;
  // compressed rows of a real W (MPO tensors are sparse): rowptr[Kout + 1],
  // col[nnz], val[nnz] (contraction precision), and the row ranges handled
  // by each of the four waves of a block, balanced by nonzeros: grp[5]
  const int* rowptr = nullptr;
  const int* col = nullptr;
  const void* val = nullptr;
  const int* grp = nullptr;
  int nnz = 0;
};

struct HeffOp {
  dmrg_heff_desc d;
  SmallOp sop;
  int P = 1;            // product of the physical dims acted on
  long long n = 0;      // vector length per plane
  double flops_per_apply = 0;
  double bytes_per_apply = 0;   // algorithmic HBM bytes (each operand read once, each result written once)
  // split-BF16x3 on pre-split operands (real field): L as (hi, lo) rows
  // [(wl a')][a] and R transposed as (hi, lo) rows [b'][(wr b)]
  bool b3 = false;
  const __bf16 *Lh = nullptr, *Ll = nullptr, *RTh = nullptr, *RTl = nullptr;
  int ldLb = 0, ldRb = 0;
};

// Records a pair of HIP events around a region on the context stream when
// profiling is on; kind 0 = H_eff apply, 1 = environment update, 2 = SVD,
// 3 = Krylov vector work, 4 = tridiagonalisation.
struct ProfScope {
  Ctx* c;
  int kind;
  double flops, bytes;
  hipEvent_t a = nullptr, b = nullptr;
  ProfScope(Ctx* c_, int kind_, double flops_, double bytes_ = 0.0);
  ~ProfScope();
};

size_t prec_size(int prec);
void to_prec(Ctx* c, int prec, const double* src, void* dst, long long n);
void from_prec(Ctx* c, int prec, const void* src, double* dst, long long n);
void to_prec_pitched(Ctx* c, int prec, const double* src, void* dst, long long rows, int cols, int ld, int planes);
SmallOp upload_smallop(Ctx* c, Slot slot, int prec, const std::vector<double>& wre,
                       const std::vector<double>& wim, bool wcplx, int Kin, int Kout,
                       const std::vector<int>& ioff, const std::vector<int>& ooff);
void run_smallop(Ctx* c, int prec, int field, const SmallOp& s, const void* in, void* out, int X,
                 int inner, long long sx_in, long long sx_out, long long pin, long long pout);
void gemm_prec(Ctx* c, int prec, const GemmDesc& g, const void* A, const void* B, void* C);
bool smallop_sparse_ok(const SmallOp& s);
// dsytrd (uplo 'L') of a symmetric n x n matrix in place: d[n], e[n-1],
// tau[n-1]; returns the number of reflectors computed (n - 1, or fewer when
// stop_tol > 0 and the trailing block fell below stop_tol |A|_F: the rest of
// d, e, tau is then zero, and *dropped, when given, receives the trace of
// the block taken as zero)
int sytrd_lower(Ctx* c, double* A, int n, double* d, double* e, double* tau, double stop_tol = 0.0,
                double* dropped = nullptr);
// zhetrd (uplo 'L') of the column-major view of a Hermitian n x n planar
// complex buffer (plane stride n*n) in place: d[n], e[n-1] real, tau[n-1]
// complex interleaved; returns the number of reflectors computed (as
// sytrd_lower, with the same early stop over both planes)
int hetrd_lower(Ctx* c, double* A, int n, double* d, double* e, double* tau, double stop_tol = 0.0,
                double* dropped = nullptr);
// rows of Zt (keep x n, row-major: eigenvectors of the tridiagonal) <- rows
// of Q Y for the first nrefl reflectors sytrd_lower left in A (all n - 1 by
// default; dormtr L/lower/N by blocked compact-WY GEMMs)
void ormtr_lower_rows(Ctx* c, const double* A, int n, const double* tau, double* Zt, int keep, int nrefl = -1);
// complex: rows of Et (keep x n, planar, plane stride keep*n) <- rows of Q Y
// for the first nrefl of hetrd_lower's reflectors (all n - 1 by default;
// zunmtr L/lower/N)
void unmtr_lower_rows(Ctx* c, const double* A, int n, const double* tau, double* Et, int keep, int nrefl = -1);
GemmDesc gdesc(int M, int N, int K, int lda, int ldb, int ldc, int field, int tag);

// Upper bound of the workspace bytes heff_prepare + heff_apply request for
// the descriptor (each slot rounded up to 256 bytes), for the allocation-free
// entry point's size check
size_t heff_scratch_bytes(const dmrg_heff_desc& d);
HeffOp heff_prepare(Ctx* c, const dmrg_heff_desc& d);
void heff_apply(Ctx* c, const HeffOp& h, const double* x, double* y);
void env_update(Ctx* c, const dmrg_env_desc& e);

// Krylov vector primitives (FP64, planar complex when field == CPLX); vectors
// of length n per plane, a "block" of k vectors is stored with stride ldq
// elements between consecutive vectors (each vector's planes are n apart).
void vec_dot_block(Ctx* c, int field, const double* Q, long long ldq, int k, const double* w, long long n,
                   double* out /* device, k (re) or 2k (re..., im...) */);
void vec_sub_block(Ctx* c, int field, const double* Q, long long ldq, int k, const double* coef,
                   double* w, long long n);
void vec_combine(Ctx* c, int field, const double* Q, long long ldq, int k, const double* coef,
                 int coef_cplx, double scale, double* x, long long n);
double vec_dot_real(Ctx* c, int field, const double* a, const double* b, long long n);

}  // namespace dmrg
