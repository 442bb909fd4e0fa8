// Gradient of <bra|H|ket> with respect to one MPO site tensor:
//   G[wl][wr][s'][s] = sum L[wl][a'][a] conj(bra[a'][s'][b']) ket[a][s][b] R[wr][b][b']
// so that <bra|H|ket> = sum G .* W_site for any site, and the derivative of
// the energy with respect to any MPO element (hence any coupling that enters
// the MPO linearly) is an element of G.  Chain of MFMA GEMMs:
//   X1[(wl a')][(s b)]      = L[(wl a')][a] ket[a][(s b)]
//   X2[wr][(wl a' s)][b']   = X1[(wl a' s)][b] R[wr][b][b']           (batched over wr)
//   X2p[(wr wl s)][(a' b')] = X2 with the a' and s axes swapped
//   G[(wr wl s)][s']        = X2p conj(Bp)^T,  Bp[s'][(a' b')] = bra with s' first
// The last GEMM has a long reduction (K = D_bra^2) and a tiny output, so it
// is split over K in chunks (batched GEMM) and the partials are summed.
#include "dmrg_internal.h"

namespace dmrg {

// dst[p][r][q][s] = src[p][q][r][s]
__global__ void swap_mid_k(const void* __restrict__ src_, void* __restrict__ dst_, int P, int Q, int Rr, int S,
                           long long plane, int np, int es) {
  long long tot = (long long)P * Q * Rr * S;
  long long idx = (long long)blockIdx.x * 256 + threadIdx.x;
  for (; idx < tot; idx += (long long)gridDim.x * 256) {
    const int s = (int)(idx % S);
    long long t = idx / S;
    const int r = (int)(t % Rr);
    t /= Rr;
    const int q = (int)(t % Q);
    const long long p = t / Q;
    const long long di = ((p * Rr + r) * Q + q) * S + s;
    for (int pl = 0; pl < np; ++pl) {
      if (es == 8) ((double*)dst_)[di + pl * plane] = ((const double*)src_)[idx + pl * plane];
      else ((float*)dst_)[di + pl * plane] = ((const float*)src_)[idx + pl * plane];
    }
  }
}

// out[i] = sum_c part[c*n + i] (per plane)
template <typename T>
__global__ void sum_parts_k(const T* __restrict__ part, int nc, long long n, double* __restrict__ out) {
  long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  double s = 0;
  for (int c = 0; c < nc; ++c) s += (double)part[(long long)c * n + i];
  out[i] = s;
}

static unsigned gg(long long n) {
  long long b = (n + 255) / 256;
  if (b > 8192) b = 8192;
  if (b < 1) b = 1;
  return (unsigned)b;
}

void mpo_grad(Ctx* c, int field, int prec, const void* L, const double* bra64, const double* ket64, const void* R,
              int Dl, int Dr, int Dlb, int Drb, int d, int wl, int wr, int ldL, int ldR, double* G_host) {
  const int np = field == CPLX ? 2 : 1;
  const size_t es = prec_size(prec);
  const long long nk = (long long)Dl * d * Dr, nb = (long long)Dlb * d * Drb;
  const void* ket = ket64;
  const void* bra = bra64;
  if (prec != P64) {
    void* kb = ws_get(c, WS_BF16_A, nk * np * es);
    to_prec(c, prec, ket64, kb, nk * np);
    ket = kb;
    void* bb = ws_get(c, WS_BF16_B, nb * np * es);
    to_prec(c, prec, bra64, bb, nb * np);
    bra = bb;
  }
  const long long n1 = (long long)wl * Dlb * d * Dr;
  const long long n2 = (long long)wr * wl * Dlb * d * Drb;
  void* X1 = ws_get(c, WS_T1, n1 * np * es);
  void* X2 = ws_get(c, WS_T2, n2 * np * es);
  void* X2p = ws_get(c, WS_ENV1, n2 * np * es);
  void* Bp = ws_get(c, WS_ENV2, nb * np * es);
  {
    GemmDesc g = gdesc(wl * Dlb, d * Dr, Dl, ldL, d * Dr, d * Dr, field, 5);
    g.pA = (long long)wl * Dlb * ldL; g.pB = nk; g.pC = n1;
    gemm_prec(c, prec, g, L, ket, X1);
  }
  {
    GemmDesc g = gdesc(wl * Dlb * d, Drb, Dr, Dr, ldR, Drb, field, 5);
    g.batch = wr;
    g.sA = 0; g.sB = (long long)Dr * ldR; g.sC = (long long)wl * Dlb * d * Drb;
    g.pA = n1; g.pB = (long long)wr * Dr * ldR; g.pC = n2;
    gemm_prec(c, prec, g, X1, R, X2);
  }
  // X2 [(wr wl)][a'][s][b'] -> X2p [(wr wl)][s][a'][b'];  bra [a'][s'][b'] -> Bp [s'][a'][b']
  swap_mid_k<<<gg(n2), 256, 0, c->stream>>>(X2, X2p, wr * wl, Dlb, d, Drb, n2, np, (int)es);
  DMRG_CHECK_HIP(hipGetLastError());
  swap_mid_k<<<gg(nb), 256, 0, c->stream>>>(bra, Bp, 1, Dlb, d, Drb, nb, np, (int)es);
  DMRG_CHECK_HIP(hipGetLastError());
  const int M = wr * wl * d, N = d;
  const long long K = (long long)Dlb * Drb;
  DMRG_REQUIRE(K < (1LL << 31), "mpo_grad: reduction too long");
  int nc = (int)std::min<long long>(256, std::max<long long>(1, K / 1024));
  const long long chunk = (K + nc - 1) / nc;
  nc = (int)((K + chunk - 1) / chunk);
  const long long mn = (long long)M * N;
  void* part = ws_get(c, WS_SVD_TMP2, (size_t)nc * mn * np * es);
  // partial GEMMs over K chunks; the last chunk may be shorter, so run the
  // full chunks batched and the remainder separately
  const int nfull = (int)(K / chunk);
  const long long rem = K - (long long)nfull * chunk;
  if (nfull > 0) {
    GemmDesc g = gdesc(M, N, (int)chunk, (int)K, (int)K, N, field, 5);
    g.transB = true; g.conjB = true;
    g.batch = nfull;
    g.sA = chunk; g.sB = chunk; g.sC = mn;
    g.pA = n2; g.pB = nb; g.pC = (long long)nc * mn;
    gemm_prec(c, prec, g, X2p, Bp, part);
  }
  if (rem > 0) {
    GemmDesc g = gdesc(M, N, (int)rem, (int)K, (int)K, N, field, 5);
    g.transB = true; g.conjB = true;
    g.pA = n2; g.pB = nb; g.pC = (long long)nc * mn;
    const size_t off = (size_t)nfull * chunk * es;
    gemm_prec(c, prec, g, (const char*)X2p + off, (const char*)Bp + off, (char*)part + (size_t)nfull * mn * es);
  }
  double* Gd = (double*)ws_get(c, WS_REFINE, sizeof(double) * mn * np);
  for (int pl = 0; pl < np; ++pl) {
    if (prec == P64)
      sum_parts_k<double><<<gg(mn), 256, 0, c->stream>>>((const double*)part + (size_t)pl * nc * mn, nc, mn, Gd + pl * mn);
    else
      sum_parts_k<float><<<gg(mn), 256, 0, c->stream>>>((const float*)part + (size_t)pl * nc * mn, nc, mn, Gd + pl * mn);
    DMRG_CHECK_HIP(hipGetLastError());
  }
  std::vector<double> h((size_t)mn * np);
  DMRG_CHECK_HIP(hipMemcpyAsync(h.data(), Gd, sizeof(double) * mn * np, hipMemcpyDeviceToHost, c->stream));
  DMRG_CHECK_HIP(hipStreamSynchronize(c->stream));
  // h[(wr wl s)][s'] -> G[wl][wr][s'][s]
  for (int pl = 0; pl < np; ++pl)
    for (int b = 0; b < wr; ++b)
      for (int a = 0; a < wl; ++a)
        for (int s = 0; s < d; ++s)
          for (int sp = 0; sp < d; ++sp)
            G_host[(size_t)pl * wl * wr * d * d + (((size_t)a * wr + b) * d + sp) * d + s] =
                h[(size_t)pl * mn + (((size_t)b * wl + a) * d + s) * d + sp];
}

}  // namespace dmrg
