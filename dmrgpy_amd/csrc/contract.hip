// Effective-Hamiltonian apply and environment updates, built as ordered
// chains of MFMA GEMMs plus one HBM-bound "small operator" pass that applies
// the MPO tensors (K = w*d <= ~100) between them.
//
// Layouts (row-major; complex = two planes):
//   MPS site      A[Dl][d][Dr]                       (FP64)
//   MPO site      W[wl][wr][s_out][s_in]             (FP64, host)
//   left env      L[w][D_bra][D_ket]                 (contraction precision)
//   right env     R[w][D_ket][D_bra]                 (contraction precision)
//   two-site vec  theta[Dl][d1][d2][Dr]              (FP64)
//
// H_eff two-site (reference operand order L, W_i, W_{i+1}, R):
//   T1[(wl a')][(p b)]  = L[(wl a')][a] * theta[a][(p b)]          GEMM
//   T2[a'][(p' wr)][b]  = sum_{wl,p} W12[(p' wr)][(wl p)] T1[...]  small-op
//   y[(a' p')][b']      = T2[(a' p')][(wr b)] * R[(wr b)][b']      GEMM
// with W12 = sum_wm W_i W_{i+1} pre-contracted on the host once per bond.
#include "dmrg_common.h"
#include "dmrg_internal.h"
#include "small_op.h"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace dmrg {

// ---------------------------------------------------------------------------
// small operator: out[x][o][inner] = sum_i Wsm[o][i] * in[x][i][inner], with
// the (o, i) -> memory offset maps given by ooff / ioff.  One thread per
// (x, inner) element and a chunk of OCH outputs (32 real, 16 complex, so the
// input is re-read KoutPad/OCH times); W is uniform across the wave and is
// read through the scalar cache.
// ---------------------------------------------------------------------------
constexpr int OCH_PAD = 32;   // Kout is padded to a multiple of this

template <typename T, bool CXV, bool CXW, int OCH>
__global__ __launch_bounds__(256) void smallop_k(const T* __restrict__ in, T* __restrict__ out,
                                                 const T* __restrict__ Wsm, const int* __restrict__ ioff,
                                                 const int* __restrict__ ooff, int X, int inner,
                                                 long long sx_in, long long sx_out, int Kin, int Kout,
                                                 int KoutPad, long long pin, long long pout) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  const int x = blockIdx.y;
  const int oc = blockIdx.z * OCH;
  if (b >= inner) return;
  const T* ip = in + (long long)x * sx_in + b;
  T* op = out + (long long)x * sx_out + b;
  const long long wplane = (long long)KoutPad * Kin;
  T ar[OCH], ai[OCH];
#pragma unroll
  for (int o = 0; o < OCH; ++o) { ar[o] = T(0); ai[o] = T(0); }
  for (int i = 0; i < Kin; ++i) {
    const long long off = ioff[i];
    const T vr = ip[off];
    const T vi = CXV ? ip[off + pin] : T(0);
    const T* wrow = Wsm + (long long)oc * Kin + i;
#pragma unroll
    for (int o = 0; o < OCH; ++o) {
      const T w = wrow[(long long)o * Kin];
      ar[o] += w * vr;
      if (CXV) ai[o] += w * vi;
      if (CXW) {
        const T wi = wrow[(long long)o * Kin + wplane];
        ar[o] -= wi * vi;
        ai[o] += wi * vr;
      }
    }
  }
#pragma unroll
  for (int o = 0; o < OCH; ++o) {
    if (oc + o < Kout) {
      const long long off = ooff[oc + o];
      op[off] = ar[o];
      if (CXV) op[off + pout] = ai[o];
    }
  }
}

template <typename T>
__global__ void convert_k(const double* __restrict__ src, T* __restrict__ dst, long long n) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long stride = (long long)gridDim.x * blockDim.x;
  for (; i < n; i += stride) dst[i] = (T)src[i];
}
template <typename T>
__global__ void convert_back_k(const T* __restrict__ src, double* __restrict__ dst, long long n) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long stride = (long long)gridDim.x * blockDim.x;
  for (; i < n; i += stride) dst[i] = (double)src[i];
}

__global__ void scale_vec_k(double* __restrict__ v, int n, double s) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) v[i] *= s;
}

static unsigned grid1d(long long n) {
  long long b = (n + 255) / 256;
  if (b > 65536) b = 65536;
  if (b < 1) b = 1;
  return (unsigned)b;
}

void to_prec(Ctx* c, int prec, const double* src, void* dst, long long n) {
  if (n <= 0) return;
  if (prec == P64) {
    if (src != dst)
      DMRG_CHECK_HIP(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
  } else {
    convert_k<float><<<grid1d(n), 256, 0, c->stream>>>(src, (float*)dst, n);
    DMRG_CHECK_HIP(hipGetLastError());
  }
}
void from_prec(Ctx* c, int prec, const void* src, double* dst, long long n) {
  if (n <= 0) return;
  if (prec == P64) {
    if (src != dst)
      DMRG_CHECK_HIP(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
  } else {
    convert_back_k<float><<<grid1d(n), 256, 0, c->stream>>>((const float*)src, dst, n);
    DMRG_CHECK_HIP(hipGetLastError());
  }
}

size_t prec_size(int prec) { return prec == P64 ? sizeof(double) : sizeof(float); }

// dst[pl][r][0..cols) (row pitch ld) = src[pl][r][0..cols) (dense), converted
// to T; planes of rows*cols (src) and rows*ld (dst) elements
template <typename T>
__global__ void convert_pitched_k(const double* __restrict__ src, T* __restrict__ dst, long long rows, int cols,
                                  int ld, int planes) {
  const long long tot = (long long)planes * rows * cols;
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (; i < tot; i += stride) {
    const long long pr = i / cols;   // plane * rows + row
    const int col = (int)(i - pr * cols);
    dst[pr * ld + col] = (T)src[i];
  }
}

// hi/lo bf16 split of FP32/FP64 rows: dst[r][c] (pitch ldd) for c < cols
template <typename S>
__global__ void split_rows_k(const S* __restrict__ src, long long rows, int cols, int lds, __bf16* __restrict__ hi,
                             __bf16* __restrict__ lo, int ldd) {
  const long long tot = rows * cols;
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (; i < tot; i += stride) {
    const long long r = i / cols;
    const int col = (int)(i - r * cols);
    const float x = (float)src[r * lds + col];
    const __bf16 h = (__bf16)x;
    hi[r * ldd + col] = h;
    lo[r * ldd + col] = (__bf16)(x - (float)h);
  }
}

// hi/lo bf16 split with transposition: dst[c][r] (pitch ldd) = src[r][c]
// (pitch lds); 32x32 tiles through LDS so reads and writes are coalesced
template <typename S>
__global__ __launch_bounds__(256) void split_transpose_k(const S* __restrict__ src, int rows, int cols, int lds,
                                                         __bf16* __restrict__ hi, __bf16* __restrict__ lo, int ldd) {
  __shared__ float tile[32][33];
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  const int r0 = blockIdx.y * 32, c0 = blockIdx.x * 32;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int r = r0 + ty + 8 * j, col = c0 + tx;
    tile[ty + 8 * j][tx] = (r < rows && col < cols) ? (float)src[(long long)r * lds + col] : 0.f;
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int col = c0 + ty + 8 * j, r = r0 + tx;
    if (col < cols && r < rows) {
      const float x = tile[tx][ty + 8 * j];
      const __bf16 h = (__bf16)x;
      hi[(long long)col * ldd + r] = h;
      lo[(long long)col * ldd + r] = (__bf16)(x - (float)h);
    }
  }
}

template <typename S>
void split_rows(Ctx* c, const S* src, long long rows, int cols, int lds, __bf16* hi, __bf16* lo, int ldd) {
  split_rows_k<S><<<grid1d(rows * cols), 256, 0, c->stream>>>(src, rows, cols, lds, hi, lo, ldd);
  DMRG_CHECK_HIP(hipGetLastError());
}

template <typename S>
void split_transpose(Ctx* c, const S* src, int rows, int cols, int lds, __bf16* hi, __bf16* lo, int ldd) {
  DMRG_REQUIRE((rows + 31) / 32 <= 65535, "split_transpose: too many rows");
  dim3 g((cols + 31) / 32, (rows + 31) / 32);
  split_transpose_k<S><<<g, 256, 0, c->stream>>>(src, rows, cols, lds, hi, lo, ldd);
  DMRG_CHECK_HIP(hipGetLastError());
}

void to_prec_pitched(Ctx* c, int prec, const double* src, void* dst, long long rows, int cols, int ld, int planes) {
  if (ld == cols) {
    to_prec(c, prec, src, dst, rows * cols * planes);
    return;
  }
  const long long tot = (long long)planes * rows * cols;
  if (prec == P64)
    convert_pitched_k<double><<<grid1d(tot), 256, 0, c->stream>>>(src, (double*)dst, rows, cols, ld, planes);
  else
    convert_pitched_k<float><<<grid1d(tot), 256, 0, c->stream>>>(src, (float*)dst, rows, cols, ld, planes);
  DMRG_CHECK_HIP(hipGetLastError());
}

// Uploads a small operator (W rows padded to a multiple of OCH_PAD), its
// offset tables and, for a real W, its compressed rows into a workspace slot;
// returns device pointers.
SmallOp upload_smallop(Ctx* c, Slot slot, int prec, const std::vector<double>& wre,
                       const std::vector<double>& wim, bool wcplx, int Kin, int Kout,
                       const std::vector<int>& ioff, const std::vector<int>& ooff) {
  SmallOp s;
  s.Kin = Kin;
  s.Kout = Kout;
  s.KoutPad = ((Kout + OCH_PAD - 1) / OCH_PAD) * OCH_PAD;
  s.wcplx = wcplx;
  const size_t es = prec_size(prec);
  auto up = [](size_t b) { return ((b + 255) / 256) * 256; };
  // compressed rows (real W only) and the four waves' row ranges
  std::vector<int> rowptr(Kout + 1, 0), col;
  std::vector<double> val;
  if (!wcplx) compress_rows(wre, Kout, Kin, rowptr, col, val);
  s.nnz = (int)col.size();
  int grp[5];
  wave_groups(rowptr, Kout, grp);
  const size_t nw = (size_t)s.KoutPad * Kin * (wcplx ? 2 : 1);
  const size_t bytes_w = up(nw * es);
  const size_t bytes_o = up(sizeof(int) * (Kin + Kout + Kout + 1 + 5 + col.size()));
  const size_t bytes = bytes_w + bytes_o + up(val.size() * es) + 256;
  std::vector<unsigned char> host(bytes, 0);
  for (int pl = 0; pl < (wcplx ? 2 : 1); ++pl) {
    const std::vector<double>& src = pl ? wim : wre;
    for (int o = 0; o < Kout; ++o)
      for (int i = 0; i < Kin; ++i) {
        size_t idx = (size_t)pl * s.KoutPad * Kin + (size_t)o * Kin + i;
        double v = src[(size_t)o * Kin + i];
        if (prec == P64) reinterpret_cast<double*>(host.data())[idx] = v;
        else reinterpret_cast<float*>(host.data())[idx] = (float)v;
      }
  }
  int* ints = reinterpret_cast<int*>(host.data() + bytes_w);
  std::memcpy(ints, ioff.data(), sizeof(int) * Kin);
  std::memcpy(ints + Kin, ooff.data(), sizeof(int) * Kout);
  std::memcpy(ints + Kin + Kout, rowptr.data(), sizeof(int) * (Kout + 1));
  std::memcpy(ints + Kin + 2 * Kout + 1, grp, sizeof(int) * 5);
  if (!col.empty()) std::memcpy(ints + Kin + 2 * Kout + 6, col.data(), sizeof(int) * col.size());
  unsigned char* vb = host.data() + bytes_w + bytes_o;
  for (size_t k = 0; k < val.size(); ++k) {
    if (prec == P64) reinterpret_cast<double*>(vb)[k] = val[k];
    else reinterpret_cast<float*>(vb)[k] = (float)val[k];
  }
  unsigned char* dev = (unsigned char*)ws_get(c, slot, bytes);
  DMRG_CHECK_HIP(hipMemcpyAsync(dev, host.data(), bytes, hipMemcpyHostToDevice, c->stream));
  // the host vector dies at return: make the copy complete first
  DMRG_CHECK_HIP(hipStreamSynchronize(c->stream));
  const int* di = reinterpret_cast<const int*>(dev + bytes_w);
  s.W = dev;
  s.ioff = di;
  s.ooff = di + Kin;
  s.rowptr = di + Kin + Kout;
  s.grp = di + Kin + 2 * Kout + 1;
  s.col = di + Kin + 2 * Kout + 6;
  s.val = dev + bytes_w + bytes_o;
  return s;
}

// ---------------------------------------------------------------------------
// Sparse small operator for a real W: out[x][o][b] = sum over the nonzeros
// (i, w) of row o of  w * in[x][i][b].  One block per (x, up to
// SP_CHUNKS chunks of 64 inner elements, plane); the chunks are staged in LDS in turn
// (prefetched into registers one chunk ahead), each lane owns one column and
// each wave a contiguous range of output rows with about a quarter of the
// nonzeros.  Row bounds, column indices and values are uniform across a wave
// (scalar loads); every output row is written as one 256-byte segment per
// wave.  HBM traffic: the input once and the output once.
// ---------------------------------------------------------------------------
constexpr int SP_BN = 64, SP_LDI = SP_BN + 1, SP_CHUNKS = 2, SP_KMAX = 128, SP_PRE = SP_KMAX * SP_BN / 256;

// Blocks along the inner index and chunks of SP_BN columns per block: at
// most SP_CHUNKS chunks, spread evenly (a bond dimension just above a
// multiple of SP_CHUNKS * SP_BN would otherwise leave every second block
// with one chunk of work).  Measured on the FP32 Hubbard apply at D = 500 …
// 1024: at most 2 chunks per block (more blocks in flight) 1-4% faster per
// apply at D = 500 … 749 than the former 8 unbalanced, 4 or 8 balanced in
// between, equal at 800 … 1024.
static void sp_split(int inner, int& nbx, int& cpb) {
  const int nch = (inner + SP_BN - 1) / SP_BN;
  nbx = (nch + SP_CHUNKS - 1) / SP_CHUNKS;
  cpb = (nch + nbx - 1) / nbx;
}

template <typename T, bool SPLIT>
__global__ __launch_bounds__(256) void smallop_sparse_k(const T* __restrict__ in, T* __restrict__ out,
                                                        const int* __restrict__ rowptr,
                                                        const int* __restrict__ colidx, const T* __restrict__ val,
                                                        const int* __restrict__ grp, const int* __restrict__ ioff,
                                                        const int* __restrict__ ooff, int inner, long long sx_in,
                                                        long long sx_out, int Kin, long long pin, long long pout,
                                                        __bf16* __restrict__ out_hi, __bf16* __restrict__ out_lo,
                                                        int cpb) {
  extern __shared__ double smem_d[];
  T* Is = reinterpret_cast<T*>(smem_d);                 // [Kin][SP_LDI]
  int* io = reinterpret_cast<int*>(Is + Kin * SP_LDI);  // [Kin]
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);   // wave-uniform
  const int x = blockIdx.y, pl = blockIdx.z;
  in += (long long)x * sx_in + pl * pin;
  if (SPLIT) {
    out_hi += (long long)x * sx_out;
    out_lo += (long long)x * sx_out;
  } else {
    out += (long long)x * sx_out + pl * pout;
  }
  for (int i = tid; i < Kin; i += 256) io[i] = ioff[i];
  const int r0 = grp[wave], r1 = grp[wave + 1];
  T pre[SP_PRE];
  auto fetch = [&](int b0) {
#pragma unroll
    for (int u = 0; u < SP_PRE; ++u) {
      const int e = u * 256 + tid;
      const int i = e / SP_BN, b = b0 + e % SP_BN;
      pre[u] = (i < Kin && b < inner) ? in[(long long)io[i] + b] : T(0);
    }
  };
  const int bfirst = blockIdx.x * cpb * SP_BN;
  __syncthreads();   // offsets ready
  fetch(bfirst);
  for (int ch = 0; ch < cpb; ++ch) {
    const int b0 = bfirst + ch * SP_BN;
    if (b0 >= inner) break;
    __syncthreads();   // previous chunk's LDS reads done
#pragma unroll
    for (int u = 0; u < SP_PRE; ++u) {
      const int e = u * 256 + tid;
      if (e < Kin * SP_BN) Is[(e / SP_BN) * SP_LDI + e % SP_BN] = pre[u];
    }
    __syncthreads();
    if (ch + 1 < cpb && b0 + SP_BN < inner) fetch(b0 + SP_BN);
    const int b = b0 + lane;
    for (int o = r0; o < r1; ++o) {
      T acc = T(0);
      const int k1 = rowptr[o + 1];
      for (int k = rowptr[o]; k < k1; ++k) acc += val[k] * Is[colidx[k] * SP_LDI + lane];
      if (b < inner) {
        const long long off = (long long)ooff[o] + b;
        if (SPLIT) {   // split-BF16x3 operand for the next GEMM: hi = bf16(acc), lo = bf16(acc - hi)
          const __bf16 h = (__bf16)acc;
          out_hi[off + pl * pout] = h;
          out_lo[off + pl * pout] = (__bf16)(acc - (float)h);
        } else {
          out[off] = acc;
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// MFMA small operator for FP32 vectors and a real W (every realified MPO):
// per outer index x the operation is the GEMM  out[x] (Kout x inner) =
// W (Kout x Kin) * in[x] (Kin x inner)  with row offsets ioff / ooff.  One
// block per (x, 512 inner elements, plane): W is staged in LDS once, then
// eight chunks of 64 inner columns are staged in turn; each wave owns 16
// columns and all ceil(Kout/16) = MT row tiles (v_mfma_f32_16x16x4_f32).
// Reads the input once and writes the output once.
// ---------------------------------------------------------------------------
typedef float flt4 __attribute__((ext_vector_type(4)));
constexpr int SOP_BN = 64, SOP_LDI = SOP_BN + 4, SOP_CHUNKS = 8, SOP_U = 8;
constexpr int SOP_KMAX = 128, SOP_PRE = SOP_KMAX * SOP_BN / 256;   // Kin <= SOP_KMAX

static size_t smallop_mfma_lds(int Kin, int Kout) {
  const int KP = (Kin + 3) / 4 * 4, MT = (Kout + 15) / 16;
  return sizeof(float) * ((size_t)MT * 16 * (KP + 1) + (size_t)KP * SOP_LDI) + sizeof(int) * (Kin + Kout);
}

template <int MT>
__global__ __launch_bounds__(256) void smallop_mfma_k(const float* __restrict__ in, float* __restrict__ out,
                                                      const float* __restrict__ Wsm, const int* __restrict__ ioff,
                                                      const int* __restrict__ ooff, int inner, long long sx_in,
                                                      long long sx_out, int Kin, int Kout, long long pin,
                                                      long long pout) {
  extern __shared__ float smem[];
  const int KP = (Kin + 3) / 4 * 4, LDW = KP + 1;
  float* Ws = smem;                      // [MT * 16][LDW]
  float* Is = Ws + MT * 16 * LDW;        // [KP][SOP_LDI]
  int* io = reinterpret_cast<int*>(Is + KP * SOP_LDI);
  int* oo = io + Kin;
  const int tid = threadIdx.x;
  const int x = blockIdx.y, pl = blockIdx.z;
  in += (long long)x * sx_in + pl * pin;
  out += (long long)x * sx_out + pl * pout;
  for (int i = tid; i < Kin; i += 256) io[i] = ioff[i];
  for (int o = tid; o < Kout; o += 256) oo[o] = ooff[o];
  // staging loops issue SOP_U independent loads before the first LDS write,
  // so a block waits for memory once per SOP_U elements per thread
  for (int base = 0; base < MT * 16 * KP; base += 256 * SOP_U) {
    float v[SOP_U];
#pragma unroll
    for (int u = 0; u < SOP_U; ++u) {
      const int e = base + u * 256 + tid;
      const int o = e / KP, i = e % KP;
      v[u] = (e < MT * 16 * KP && o < Kout && i < Kin) ? Wsm[(long long)o * Kin + i] : 0.f;
    }
#pragma unroll
    for (int u = 0; u < SOP_U; ++u) {
      const int e = base + u * 256 + tid;
      if (e < MT * 16 * KP) Ws[(e / KP) * LDW + e % KP] = v[u];
    }
  }
  const int lane = tid & 63, wave = tid >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int col0 = wave * 16;
  // input chunks are prefetched into registers one chunk ahead, so their
  // loads are in flight while the previous chunk is multiplied
  float pre[SOP_PRE];
  auto fetch = [&](int b0) {
#pragma unroll
    for (int u = 0; u < SOP_PRE; ++u) {
      const int e = u * 256 + tid;
      const int i = e / SOP_BN, b = b0 + e % SOP_BN;
      pre[u] = (i < Kin && b < inner) ? in[(long long)io[i] + b] : 0.f;
    }
  };
  const int bfirst = blockIdx.x * SOP_CHUNKS * SOP_BN;
  __syncthreads();   // offsets ready
  fetch(bfirst);
  for (int ch = 0; ch < SOP_CHUNKS; ++ch) {
    const int b0 = bfirst + ch * SOP_BN;
    if (b0 >= inner) break;
    __syncthreads();   // W ready; previous chunk's LDS reads done
#pragma unroll
    for (int u = 0; u < SOP_PRE; ++u) {
      const int e = u * 256 + tid;
      if (e < KP * SOP_BN) Is[(e / SOP_BN) * SOP_LDI + e % SOP_BN] = pre[u];
    }
    __syncthreads();
    if (ch + 1 < SOP_CHUNKS && b0 + SOP_BN < inner) fetch(b0 + SOP_BN);
    flt4 acc[MT];
#pragma unroll
    for (int t = 0; t < MT; ++t) acc[t] = flt4{0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < KP; k0 += 4) {
      const float bv = Is[(k0 + lk) * SOP_LDI + col0 + li];
#pragma unroll
      for (int t = 0; t < MT; ++t)
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(Ws[(t * 16 + li) * LDW + k0 + lk], bv, acc[t], 0, 0, 0);
    }
    // accumulator element r of tile t: row t*16 + 4*(lane/16) + r, column lane%16
    const int b = b0 + col0 + li;
    if (b < inner) {
#pragma unroll
      for (int t = 0; t < MT; ++t)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int o = t * 16 + 4 * lk + r;
          if (o < Kout) out[(long long)oo[o] + b] = acc[t][r];
        }
    }
  }
}

bool smallop_sparse_ok(const SmallOp& s) {
  return !s.wcplx && s.Kin <= SP_KMAX && 8LL * s.nnz <= (long long)s.Kin * s.Kout;
}

// The sparse pass on one real FP32 plane with its output split into bf16
// (hi, lo) arrays (the A operand of the split-BF16x3 R-step GEMM).
void run_smallop_split(Ctx* c, const SmallOp& s, const float* in, __bf16* hi, __bf16* lo, int X, int inner,
                       long long sx_in, long long sx_out) {
  if (X <= 0 || inner <= 0 || s.Kout <= 0) return;
  DMRG_REQUIRE(X <= 65535, "smallop: outer extent too large");
  DMRG_REQUIRE(smallop_sparse_ok(s), "smallop: split output needs the sparse real pass");
  int nbx, cpb;
  sp_split(inner, nbx, cpb);
  dim3 g(nbx, X, 1);
  const size_t lds = sizeof(float) * (size_t)s.Kin * SP_LDI + sizeof(int) * s.Kin;
  smallop_sparse_k<float, true><<<g, 256, lds, c->stream>>>(in, nullptr, s.rowptr, s.col, (const float*)s.val, s.grp,
                                                            s.ioff, s.ooff, inner, sx_in, sx_out, s.Kin, 0, 0, hi,
                                                            lo, cpb);
  DMRG_CHECK_HIP(hipGetLastError());
}

void run_smallop(Ctx* c, int prec, int field, const SmallOp& s, const void* in, void* out, int X,
                 int inner, long long sx_in, long long sx_out, long long pin, long long pout) {
  if (X <= 0 || inner <= 0 || s.Kout <= 0) return;
  DMRG_REQUIRE(X <= 65535, "smallop: outer extent too large");
  const bool cv = field == CPLX;
  if (smallop_sparse_ok(s)) {
    int nbx, cpb;
    sp_split(inner, nbx, cpb);
    dim3 g(nbx, X, cv ? 2 : 1);
    const size_t es = prec_size(prec);
    const size_t lds = es * (size_t)s.Kin * SP_LDI + sizeof(int) * s.Kin;
    if (prec == P64)
      smallop_sparse_k<double, false><<<g, 256, lds, c->stream>>>((const double*)in, (double*)out, s.rowptr, s.col,
                                                                  (const double*)s.val, s.grp, s.ioff, s.ooff, inner,
                                                                  sx_in, sx_out, s.Kin, pin, pout, nullptr, nullptr,
                                                                  cpb);
    else
      smallop_sparse_k<float, false><<<g, 256, lds, c->stream>>>((const float*)in, (float*)out, s.rowptr, s.col,
                                                                 (const float*)s.val, s.grp, s.ioff, s.ooff, inner,
                                                                 sx_in, sx_out, s.Kin, pin, pout, nullptr, nullptr,
                                                                 cpb);
    DMRG_CHECK_HIP(hipGetLastError());
    return;
  }
  const size_t lds = smallop_mfma_lds(s.Kin, s.Kout);
  if (prec != P64 && !s.wcplx && s.Kout <= 128 && s.Kin <= SOP_KMAX && lds <= 65536) {
    dim3 g((inner + SOP_BN * SOP_CHUNKS - 1) / (SOP_BN * SOP_CHUNKS), X, cv ? 2 : 1);
    const int MT = (s.Kout + 15) / 16;
    switch (MT) {
#define CASE(m)                                                                                               \
  case m:                                                                                                     \
    smallop_mfma_k<m><<<g, 256, lds, c->stream>>>((const float*)in, (float*)out, (const float*)s.W, s.ioff, \
                                                   s.ooff, inner, sx_in, sx_out, s.Kin, s.Kout, pin, pout);    \
    break;
      CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8)
#undef CASE
      default: throw Error("smallop: bad tile count");
    }
    DMRG_CHECK_HIP(hipGetLastError());
    return;
  }
  const int och = cv ? 16 : 32;
  dim3 grid((inner + 255) / 256, X, s.KoutPad / och);
#define LAUNCH(T, CV, CW, O)                                                                         \
  smallop_k<T, CV, CW, O><<<grid, 256, 0, c->stream>>>((const T*)in, (T*)out, (const T*)s.W, s.ioff, \
                                                        s.ooff, X, inner, sx_in, sx_out, s.Kin,      \
                                                        s.Kout, s.KoutPad, pin, pout)
  if (prec == P64) {
    if (!cv) { DMRG_REQUIRE(!s.wcplx, "smallop: complex MPO needs complex vectors"); LAUNCH(double, false, false, 32); }
    else if (s.wcplx) LAUNCH(double, true, true, 16);
    else LAUNCH(double, true, false, 16);
  } else {
    if (!cv) { DMRG_REQUIRE(!s.wcplx, "smallop: complex MPO needs complex vectors"); LAUNCH(float, false, false, 32); }
    else if (s.wcplx) LAUNCH(float, true, true, 16);
    else LAUNCH(float, true, false, 16);
  }
#undef LAUNCH
  DMRG_CHECK_HIP(hipGetLastError());
}

void gemm_prec(Ctx* c, int prec, const GemmDesc& g, const void* A, const void* B, void* C) {
  if (prec == P64) gemm_f64(c, g, (const double*)A, (const double*)B, (double*)C);
  else if (prec == P32) gemm_f32(c, g, (const float*)A, (const float*)B, (float*)C);
  else gemm_bf16x3(c, g, (const float*)A, (const float*)B, (float*)C);
}

GemmDesc gdesc(int M, int N, int K, int lda, int ldb, int ldc, int field, int tag) {
  GemmDesc g;
  g.M = M; g.N = N; g.K = K;
  g.lda = lda; g.ldb = ldb; g.ldc = ldc;
  g.sA = g.sB = g.sC = 0;
  g.pA = g.pB = g.pC = 0;
  g.batch = 1;
  g.transA = g.transB = false;
  g.conjA = g.conjB = false;
  g.alpha = 1.0; g.beta = 0.0;
  g.field = field;
  g.tag = tag;
  g.lower = false;
  return g;
}

// ---------------------------------------------------------------------------
// H_eff
// ---------------------------------------------------------------------------
// Row pitches that keep every H_eff GEMM operand row 16-byte aligned: the
// T2 intermediate [(a' p')][(wr b)] and, when P*Dr is not a multiple of 4,
// the contraction-precision copy of theta.
static int pad4(long long v) { return (int)((v + 3) / 4 * 4); }
static int pad8(long long v) { return (int)((v + 7) / 8 * 8); }
static int t2_pitch(int wr, int Dr) { return pad8((long long)wr * Dr); }

static void heff_check(const dmrg_heff_desc& d) {
  DMRG_REQUIRE(d.nsite == 1 || d.nsite == 2, "heff: nsite must be 1 or 2");
  DMRG_REQUIRE(d.Dl > 0 && d.Dr > 0 && d.d1 > 0 && d.wl > 0 && d.wr > 0, "heff: bad dims");
  DMRG_REQUIRE(d.field == REAL || d.field == CPLX, "heff: bad field");
  DMRG_REQUIRE(d.prec == P64 || d.prec == P32 || d.prec == PB3, "heff: bad precision");
  DMRG_REQUIRE(!(d.w_complex && d.field == REAL), "heff: complex MPO requires complex field");
  DMRG_REQUIRE((d.ldL == 0 || d.ldL >= d.Dl) && (d.ldR == 0 || d.ldR >= d.Dr),
               "heff: environment row pitch smaller than the dimension");
  DMRG_REQUIRE(d.nproj >= 0 && d.nproj <= 64 && (d.nproj == 0 || d.proj != nullptr),
               "heff: bad penalty projectors");
  if (d.nsite == 2) DMRG_REQUIRE(d.d2 > 0 && d.wm > 0, "heff: bad two-site dims");
}

size_t heff_scratch_bytes(const dmrg_heff_desc& d) {
  heff_check(d);
  // every request is rounded up to 256 bytes, plus 256 of slack per slot
  auto up = [](double b) { return std::ceil(b / 256.0) * 256.0 + 256.0; };
  const double P = d.nsite == 2 ? (double)d.d1 * d.d2 : (double)d.d1;
  const double Dl = d.Dl, Dr = d.Dr, wl = d.wl, wr = d.wr;
  const double np = d.field == CPLX ? 2.0 : 1.0, es = (double)prec_size(d.prec);
  const double n = Dl * P * Dr, nT = wl * n, ldT2 = t2_pitch(d.wr, d.Dr), nT2 = Dl * P * ldT2;
  // small operator (upload_smallop): padded dense W, offset tables and the
  // compressed rows at their largest (every entry nonzero)
  const double Kin = wl * P, Kout = P * wr, KoutPad = std::ceil(Kout / OCH_PAD) * OCH_PAD;
  double total = up(up(KoutPad * Kin * (d.w_complex ? 2.0 : 1.0) * es) + up(4.0 * (Kin + 2 * Kout + 6 + Kout * Kin)) +
                    up(Kout * Kin * es) + 256.0);
  // generic path: theta in the contraction precision (pitched), y in the
  // contraction precision, T1, T2
  const double Pdr = P * Dr, ldx = std::fmod(Pdr, 4.0) == 0.0 ? Pdr : (double)pad4((long long)Pdr);
  double gen = up(nT * np * es) + up(nT2 * np * es);
  if (d.prec != P64 || ldx != Pdr) gen += up(Dl * ldx * np * es);
  if (d.prec != P64 && !(d.prec == P32 && d.field == REAL)) gen += up(n * np * es);
  double path = gen;
  if (d.prec == PB3 && d.field == REAL) {
    // pre-split operands: L and R^T (hi, lo), theta^T (hi, lo), T1 FP32,
    // T2 (hi, lo), y FP32
    const double ldLb = pad8(d.Dl), ldRb = pad8((long long)d.wr * d.Dr);
    const double b3 = up(4.0 * wl * Dl * ldLb) + up(4.0 * Dr * ldRb) + up(4.0 * P * Dr * pad8(d.Dl)) + up(4.0 * nT) +
                      up(4.0 * Dl * P * ldT2) + up(4.0 * n);
    path = std::max(path, b3);
  }
  total += path;
  // penalty projectors: coefficients and the partial sums of the block dot
  // products (at most 1024 partial blocks per row)
  if (d.nproj > 0) total += up(16.0 * d.nproj) + up(8.0 * np * d.nproj * 1024.0);
  return (size_t)total;
}

HeffOp heff_prepare(Ctx* c, const dmrg_heff_desc& d) {
  HeffOp h;
  h.d = d;
  heff_check(d);
  const int Dl = d.Dl, Dr = d.Dr, wl = d.wl, wr = d.wr;
  const bool wc = d.w_complex != 0;
  {
    // offsets in the small-op tables are 32-bit: check before building them
    const long long Pll = d.nsite == 2 ? (long long)d.d1 * d.d2 : (long long)d.d1;
    const long long t2p = ((long long)wr * Dr + 7) / 8 * 8;   // t2_pitch without int overflow
    DMRG_REQUIRE((long long)wl * Dl * Pll * Dr < (1LL << 31) && (long long)Dl * Pll * t2p < (1LL << 31),
                 "heff: tensor too large for 32-bit offsets");
  }
  if (d.nsite == 2) {
    const int d1 = d.d1, d2 = d.d2, wm = d.wm, P = d1 * d2;
    const int Kin = wl * P, Kout = P * wr;
    std::vector<double> wre, wim;
    fuse_two_site(d.W1, d.W2, wc, wl, wm, wr, d1, d2, wre, wim);
    std::vector<int> ioff(Kin), ooff(Kout);
    for (int a = 0; a < wl; ++a)
      for (int p = 0; p < P; ++p) ioff[a * P + p] = a * Dl * P * Dr + p * Dr;
    for (int p = 0; p < P; ++p)
      for (int b = 0; b < wr; ++b) ooff[p * wr + b] = p * t2_pitch(wr, Dr) + b * Dr;
    h.sop = upload_smallop(c, WS_WSMALL, d.prec, wre, wim, wc, Kin, Kout, ioff, ooff);
    h.P = P;
  } else {
    const int dd = d.d1, P = dd;
    const int Kin = wl * P, Kout = P * wr;
    std::vector<double> wre, wim;
    one_site_op(d.W1, wc, wl, wr, dd, wre, wim);
    std::vector<int> ioff(Kin), ooff(Kout);
    for (int a = 0; a < wl; ++a)
      for (int p = 0; p < P; ++p) ioff[a * P + p] = a * Dl * P * Dr + p * Dr;
    for (int p = 0; p < P; ++p)
      for (int b = 0; b < wr; ++b) ooff[p * wr + b] = p * t2_pitch(wr, Dr) + b * Dr;
    h.sop = upload_smallop(c, WS_WSMALL, d.prec, wre, wim, wc, Kin, Kout, ioff, ooff);
    h.P = P;
  }
  h.n = (long long)Dl * h.P * Dr;
  h.flops_per_apply = heff_flops(d);
  h.bytes_per_apply = heff_bytes(d);
  if (d.prec == PB3 && d.field == REAL && smallop_sparse_ok(h.sop)) {
    // environments split once per operator (they are fixed for the whole
    // Krylov solve): L rows [(wl a')][a], R transposed to rows [b'][(wr b)]
    const int ldL = d.ldL ? d.ldL : Dl, ldR = d.ldR ? d.ldR : Dr;
    h.ldLb = pad8(Dl);
    h.ldRb = pad8((long long)wr * Dr);
    const size_t nL = (size_t)wl * Dl * h.ldLb, nR = (size_t)Dr * h.ldRb;
    __bf16* bl = (__bf16*)ws_get(c, WS_B3_L, sizeof(__bf16) * 2 * nL);
    __bf16* br = (__bf16*)ws_get(c, WS_B3_R, sizeof(__bf16) * 2 * nR);
    split_rows<float>(c, (const float*)d.L, (long long)wl * Dl, Dl, ldL, bl, bl + nL, h.ldLb);
    split_transpose<float>(c, (const float*)d.R, wr * Dr, Dr, ldR, br, br + nR, h.ldRb);
    h.Lh = bl;
    h.Ll = bl + nL;
    h.RTh = br;
    h.RTl = br + nR;
    h.b3 = true;
  }
  return h;
}

// heff_flops: heff_work.h.

// y += w sum_k p_k <p_k|x> for the penalty projectors of the operator (FP64;
// the coefficients stay on the device)
static void apply_projectors(Ctx* c, const HeffOp& h, const double* x, double* y) {
  const dmrg_heff_desc& d = h.d;
  if (d.nproj <= 0) return;
  const int np = d.field == CPLX ? 2 : 1;
  double* coef = (double*)ws_get(c, WS_PROJ, sizeof(double) * 2 * (size_t)d.nproj);
  const long long ldp = h.n * np;
  vec_dot_block(c, d.field, d.proj, ldp, d.nproj, x, h.n, coef);
  scale_vec_k<<<1, 128, 0, c->stream>>>(coef, np * d.nproj, -d.proj_weight);
  DMRG_CHECK_HIP(hipGetLastError());
  vec_sub_block(c, d.field, d.proj, ldp, d.nproj, coef, y, h.n);
}

void run_smallop_split(Ctx* c, const SmallOp& s, const float* in, __bf16* hi, __bf16* lo, int X, int inner,
                       long long sx_in, long long sx_out);

// y = H_eff x, x and y FP64 (planar if complex), length h.n (per plane).
void heff_apply(Ctx* c, const HeffOp& h, const double* x, double* y) {
  const dmrg_heff_desc& d = h.d;
  const int Dl = d.Dl, Dr = d.Dr, wl = d.wl, wr = d.wr, P = h.P;
  const int ldL = d.ldL ? d.ldL : Dl, ldR = d.ldR ? d.ldR : Dr;
  const int ldT2 = t2_pitch(wr, Dr);
  const int np = d.field == CPLX ? 2 : 1;
  const size_t es = prec_size(d.prec);
  const long long n = h.n;
  const long long nT = (long long)wl * Dl * P * Dr;   // T1 per plane
  const long long nT2 = (long long)Dl * P * ldT2;     // T2 per plane (rows pitched)
  ProfScope ps(c, 0, h.flops_per_apply, h.bytes_per_apply);
  if (h.b3) {
    // split-BF16x3 on pre-split operands: theta^T split, L-GEMM, MPO pass
    // writing T2 split, R-GEMM, y back to FP64
    const int ldxb = pad8(Dl);
    const size_t nxt = (size_t)P * Dr * ldxb, nt2 = (size_t)Dl * P * ldT2;
    __bf16* xt = (__bf16*)ws_get(c, WS_XIN, sizeof(__bf16) * 2 * nxt);
    split_transpose<double>(c, x, Dl, P * Dr, P * Dr, xt, xt + nxt, ldxb);
    float* T1f = (float*)ws_get(c, WS_T1, sizeof(float) * nT);
    gemm_b3nt(c, wl * Dl, P * Dr, Dl, h.Lh, h.Ll, h.ldLb, xt, xt + nxt, ldxb, T1f, P * Dr, 1);
    __bf16* t2 = (__bf16*)ws_get(c, WS_T2, sizeof(__bf16) * 2 * nt2);
    run_smallop_split(c, h.sop, T1f, t2, t2 + nt2, Dl, Dr, (long long)P * Dr, (long long)P * ldT2);
    float* y32 = (float*)ws_get(c, WS_YOUT, sizeof(float) * n);
    gemm_b3nt(c, Dl * P, Dr, wr * Dr, t2, t2 + nt2, ldT2, h.RTh, h.RTl, h.ldRb, y32, Dr, 2);
    from_prec(c, P32, y32, y, n);
    apply_projectors(c, h, x, y);
    c->matvecs++;
    return;
  }
  // theta as the L-GEMM operand: rows of P*Dr elements, pitched to a multiple
  // of 4 when needed; converted to the contraction precision on the way
  const int ldx = ((long long)P * Dr) % 4 == 0 ? P * Dr : pad4((long long)P * Dr);
  const long long nx = (long long)Dl * ldx;
  const void* xin = x;
  void* yout = y;
  if (d.prec != P64 || ldx != P * Dr) {
    void* xb = ws_get(c, WS_XIN, nx * np * es);
    to_prec_pitched(c, d.prec, x, xb, Dl, P * Dr, ldx, np);
    xin = xb;
  }
  // FP32 real: the R-GEMM writes y in FP64 directly (no conversion pass)
  const bool direct64 = d.prec == P32 && d.field == REAL;
  if (d.prec != P64 && !direct64) yout = ws_get(c, WS_YOUT, n * np * es);
  void* T1 = ws_get(c, WS_T1, nT * np * es);
  void* T2 = ws_get(c, WS_T2, nT2 * np * es);
  {  // T1[(wl a')][(p b)] = L[(wl a')][a] * x[a][(p b)]
    GemmDesc g = gdesc(wl * Dl, P * Dr, Dl, ldL, ldx, P * Dr, d.field, 1);
    g.pA = (long long)wl * Dl * ldL; g.pB = nx; g.pC = nT;
    gemm_prec(c, d.prec, g, d.L, xin, T1);
  }
  run_smallop(c, d.prec, d.field, h.sop, T1, T2, Dl, Dr, (long long)P * Dr, (long long)P * ldT2, nT, nT2);
  {  // y[(a' p')][b'] = T2[(a' p')][(wr b)] * R[(wr b)][b']
    GemmDesc g = gdesc(Dl * P, Dr, wr * Dr, ldT2, ldR, Dr, d.field, 2);
    g.pA = nT2; g.pB = (long long)wr * Dr * ldR; g.pC = n;
    if (direct64) gemm_f32_c64(c, g, (const float*)T2, (const float*)d.R, y);
    else gemm_prec(c, d.prec, g, T2, d.R, yout);
  }
  if (d.prec != P64 && !direct64) from_prec(c, d.prec, yout, y, n * np);
  apply_projectors(c, h, x, y);
  c->matvecs++;
}

// ---------------------------------------------------------------------------
// environment updates (bra and ket may differ in bond dimension)
// ---------------------------------------------------------------------------
void env_update(Ctx* c, const dmrg_env_desc& e) {
  DMRG_REQUIRE(e.dir == 0 || e.dir == 1, "env: dir must be 0 (left) or 1 (right)");
  DMRG_REQUIRE(e.Dl > 0 && e.Dr > 0 && e.Dlb > 0 && e.Drb > 0 && e.d > 0 && e.wl > 0 && e.wr > 0,
               "env: bad dims");
  DMRG_REQUIRE(!(e.w_complex && e.field == REAL), "env: complex MPO requires complex field");
  const int Dl = e.Dl, Dr = e.Dr, Dlb = e.Dlb, Drb = e.Drb, dd = e.d, wl = e.wl, wr = e.wr;
  const int np = e.field == CPLX ? 2 : 1;
  // row pitches of the environments' last index (dir 0: Dl in, Dr out; dir 1: Drb in, Dlb out)
  const int ldIn = e.ld_in ? e.ld_in : (e.dir == 0 ? Dl : Drb);
  const int ldOut = e.ld_out ? e.ld_out : (e.dir == 0 ? Dr : Dlb);
  DMRG_REQUIRE(ldIn >= (e.dir == 0 ? Dl : Drb) && ldOut >= (e.dir == 0 ? Dr : Dlb),
               "env: row pitch smaller than the dimension");
  const size_t es = prec_size(e.prec);
  const bool wc = e.w_complex != 0;
  const long long nk = (long long)Dl * dd * Dr, nb = (long long)Dlb * dd * Drb;
  double mac = e.dir == 0
                   ? (double)wl * Dlb * Dl * dd * Dr + (double)wl * wr * dd * dd * Dlb * Dr + (double)wr * Drb * Dlb * dd * Dr
                   : (double)wr * Dl * dd * Dr * Drb + (double)wl * wr * dd * dd * Dl * Drb + (double)wl * Dl * dd * Drb * Dlb;
  ProfScope ps(c, 1, (e.field == CPLX ? 8.0 : 2.0) * mac);
  // MPS tensors in contraction precision
  const void* ket = e.ket;
  const void* bra = e.bra;
  if (e.prec != P64) {
    void* kb = ws_get(c, WS_BF16_A, nk * np * es);
    to_prec(c, e.prec, (const double*)e.ket, kb, nk * np);
    ket = kb;
    if (e.bra == e.ket) {
      bra = kb;
    } else {
      void* bb = ws_get(c, WS_BF16_B, nb * np * es);
      to_prec(c, e.prec, (const double*)e.bra, bb, nb * np);
      bra = bb;
    }
  }
  // small operator
  const int Kin = (e.dir == 0 ? wl : wr) * dd, Kout = (e.dir == 0 ? wr : wl) * dd;
  std::vector<double> wre((size_t)Kout * Kin, 0.0), wim(wc ? (size_t)Kout * Kin : 0, 0.0);
  const size_t nW = (size_t)wl * wr * dd * dd;
  for (int a = 0; a < wl; ++a)
    for (int b = 0; b < wr; ++b)
      for (int so = 0; so < dd; ++so)
        for (int s = 0; s < dd; ++s) {
          size_t src = (((size_t)a * wr + b) * dd + so) * dd + s;
          int o, i;
          if (e.dir == 0) { o = b * dd + so; i = a * dd + s; }
          else { o = a * dd + so; i = b * dd + s; }
          wre[(size_t)o * Kin + i] = e.W[src];
          if (wc) wim[(size_t)o * Kin + i] = e.W[src + nW];
        }
  if (e.dir == 0) {
    // E1[(wl a')][(s b)] = L[(wl a')][a] * ket[a][(s b)]
    const long long n1 = (long long)wl * Dlb * dd * Dr, n2 = (long long)wr * Dlb * dd * Dr;
    // the MPO pass's offset tables (channel strides, below n1 / n2) are 32-bit
    DMRG_REQUIRE(n1 < (1LL << 31) && n2 < (1LL << 31), "env: tensor too large for 32-bit offsets");
    void* E1 = ws_get(c, WS_ENV1, n1 * np * es);
    void* E2 = ws_get(c, WS_ENV2, n2 * np * es);
    GemmDesc g = gdesc(wl * Dlb, dd * Dr, Dl, ldIn, dd * Dr, dd * Dr, e.field, 3);
    g.pA = (long long)wl * Dlb * ldIn; g.pB = nk; g.pC = n1;
    gemm_prec(c, e.prec, g, e.env_in, ket, E1);
    std::vector<int> ioff(Kin), ooff(Kout);
    for (int a = 0; a < wl; ++a)
      for (int s = 0; s < dd; ++s) ioff[a * dd + s] = a * Dlb * dd * Dr + s * Dr;
    for (int b = 0; b < wr; ++b)
      for (int s = 0; s < dd; ++s) ooff[b * dd + s] = b * Dlb * dd * Dr + s * Dr;
    SmallOp so = upload_smallop(c, WS_WSMALL, e.prec, wre, wim, wc, Kin, Kout, ioff, ooff);
    run_smallop(c, e.prec, e.field, so, E1, E2, Dlb, Dr, (long long)dd * Dr, (long long)dd * Dr, n1, n2);
    // Lnew[wr][b'][b] = sum_{(a' s')} conj(bra[(a' s')][b']) E2[wr][(a' s')][b]
    GemmDesc h = gdesc(Drb, Dr, Dlb * dd, Drb, Dr, ldOut, e.field, 3);
    h.transA = true; h.conjA = true;
    h.batch = wr;
    h.sA = 0; h.sB = (long long)Dlb * dd * Dr; h.sC = (long long)Drb * ldOut;
    h.pA = nb; h.pB = n2; h.pC = (long long)wr * Drb * ldOut;
    gemm_prec(c, e.prec, h, bra, E2, e.env_out);
  } else {
    // F1[wr][(a s)][b'] = ket[(a s)][b] * R[wr][b][b']
    const long long n1 = (long long)wr * Dl * dd * Drb, n2 = (long long)wl * Dl * dd * Drb;
    DMRG_REQUIRE(n1 < (1LL << 31) && n2 < (1LL << 31), "env: tensor too large for 32-bit offsets");
    void* F1 = ws_get(c, WS_ENV1, n1 * np * es);
    void* F2 = ws_get(c, WS_ENV2, n2 * np * es);
    GemmDesc g = gdesc(Dl * dd, Drb, Dr, Dr, ldIn, Drb, e.field, 3);
    g.batch = wr;
    g.sA = 0; g.sB = (long long)Dr * ldIn; g.sC = (long long)Dl * dd * Drb;
    g.pA = nk; g.pB = (long long)wr * Dr * ldIn; g.pC = n1;
    gemm_prec(c, e.prec, g, ket, e.env_in, F1);
    std::vector<int> ioff(Kin), ooff(Kout);
    for (int b = 0; b < wr; ++b)
      for (int s = 0; s < dd; ++s) ioff[b * dd + s] = b * Dl * dd * Drb + s * Drb;
    for (int a = 0; a < wl; ++a)
      for (int s = 0; s < dd; ++s) ooff[a * dd + s] = a * Dl * dd * Drb + s * Drb;
    SmallOp so = upload_smallop(c, WS_WSMALL, e.prec, wre, wim, wc, Kin, Kout, ioff, ooff);
    run_smallop(c, e.prec, e.field, so, F1, F2, Dl, Drb, (long long)dd * Drb, (long long)dd * Drb, n1, n2);
    // Rnew[wl][a][a'] = sum_{(s' b')} F2[wl][a][(s' b')] conj(bra[a'][(s' b')])
    GemmDesc h = gdesc(Dl, Dlb, dd * Drb, dd * Drb, dd * Drb, ldOut, e.field, 3);
    h.transB = true; h.conjB = true;
    h.batch = wl;
    h.sA = (long long)Dl * dd * Drb; h.sB = 0; h.sC = (long long)Dl * ldOut;
    h.pA = n2; h.pB = nb; h.pC = (long long)wl * Dl * ldOut;
    gemm_prec(c, e.prec, h, F2, bra, e.env_out);
  }
}

}  // namespace dmrg
