// extern "C" entry points of libdmrg_hip.so (declared in include/dmrg_hip.h).
// Every entry point catches library errors and returns a status code; the
// message is kept per thread for dmrg_last_error().
#include <complex>
#include <cstring>
#include <mutex>

#include "dmrg_internal.h"

struct dmrg_ctx {
  dmrg::Ctx c;
};

// Workspace of the allocation-free entry point (dmrg_heff_apply_noalloc):
// one static device array per GPU, held until the process ends.  Host code
// never passes `scratch` itself to a kernel (in host code the name is a
// host-side stand-in): dmrg_ctx_create maps each device's copy once with
// hipGetSymbolAddress and the library uses that device address.
static constexpr size_t kScratchBytes = size_t(64) << 20;   // 64 MiB
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch_addr[kMaxDevices] = {};

// Device address of `scratch` on `device` (which must be current).
unsigned char* scratch_address(int device) {
  DMRG_REQUIRE(device >= 0 && device < kMaxDevices, "ctx_create: device index beyond the scratch table");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_addr[device]) {
    void* ptr = nullptr;
    DMRG_CHECK_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    DMRG_REQUIRE(ptr != nullptr, "ctx_create: hipGetSymbolAddress returned a null scratch address");
    g_scratch_addr[device] = (unsigned char*)ptr;
  }
  return g_scratch_addr[device];
}

// Switches a context to carving its workspace out of the scratch array for
// one allocation-free call, and back on every exit path.
struct NoallocScope {
  dmrg::Ctx* c;
  explicit NoallocScope(dmrg::Ctx* c_) : c(c_) {
    c->noalloc = true;
    c->scratch_used = 0;
    for (int i = 0; i < dmrg::WS_NSLOTS; ++i) {
      c->scratch_slot[i] = nullptr;
      c->scratch_slot_bytes[i] = 0;
    }
  }
  ~NoallocScope() { c->noalloc = false; }
};
}  // namespace

namespace dmrg {

void lanczos_ground(Ctx* c, const HeffOp& h, double* x, int maxiter, double tol, double res_tol, double* energy,
                    int* iters);
void krylov_expm(Ctx* c, const HeffOp& h, double* x, std::complex<double> coeff, int maxiter, double tol,
                 double errgoal, int* iters);
double rayleigh(Ctx* c, const HeffOp& h, const double* x);
int svd_split_entry(Ctx* c, int field, const double* M, int m, int n, double cutoff, int maxdim, int mindim,
                    int absorb, double* U, double* V, int kcap, int* keep, double* discarded, double* svals_host);
int sym_eig_entry(Ctx* c, int mode, double* A, int n, double* w, double* Z);
int herm_eig_entry(Ctx* c, int mode, double* A, int n, double* w, double* Z);
int svd_split_guided_entry(Ctx* c, int field, const double* M, int m, int n, double cutoff, int maxdim,
                           int mindim, int absorb, const double* G, int k0, int gside, int sketch, double* U,
                           double* V, int kcap, int* keep, double* discarded, double* svals_host, int* used,
                           double* resid);
void mpo_grad(Ctx* c, int field, int prec, const void* L, const double* bra64, const double* ket64, const void* R,
              int Dl, int Dr, int Dlb, int Drb, int d, int wl, int wr, int ldL, int ldR, double* G_host);

void* ws_get(Ctx* c, Slot s, size_t bytes) {
  if (!c->noalloc) return c->ws[s].get(bytes, c->stream);
  // allocation-free call: one 256-byte-aligned region of the scratch array
  // per slot, handed out in order
  if (c->scratch_slot[s] && c->scratch_slot_bytes[s] >= bytes) return c->scratch_slot[s];
  const size_t need = (bytes + 255) / 256 * 256;
  if (c->scratch_used + need > kScratchBytes)
    throw Error("allocation-free call: workspace exceeds the " + std::to_string(kScratchBytes) +
                "-byte scratch array (needs at least " + std::to_string(c->scratch_used + need) + " bytes)");
  void* p = c->scratch + c->scratch_used;
  c->scratch_used += need;
  c->scratch_slot[s] = p;
  c->scratch_slot_bytes[s] = need;
  return p;
}

static hipEvent_t pool_event(Ctx* c) {
  if (!c->event_pool.empty()) {
    hipEvent_t e = c->event_pool.back();
    c->event_pool.pop_back();
    return e;
  }
  hipEvent_t e;
  DMRG_CHECK_HIP(hipEventCreate(&e));
  return e;
}

// Folds finished event pairs into the per-kind totals once the list grows.
static void prof_drain(Ctx* c, bool all) {
  if (c->prof.empty()) return;
  if (all) DMRG_CHECK_HIP(hipStreamSynchronize(c->stream));
  std::vector<ProfEntry> keep;
  for (auto& p : c->prof) {
    if (!all && hipEventQuery(p.stop) != hipSuccess) {
      keep.push_back(p);
      continue;
    }
    float ms = 0;
    DMRG_CHECK_HIP(hipEventElapsedTime(&ms, p.start, p.stop));
    c->prof_time_ms[p.kind] += ms;
    c->prof_flops[p.kind] += p.flops;
    c->prof_bytes[p.kind] += p.bytes;
    c->prof_calls[p.kind] += 1;
    c->event_pool.push_back(p.start);
    c->event_pool.push_back(p.stop);
  }
  c->prof.swap(keep);
}

ProfScope::ProfScope(Ctx* c_, int kind_, double flops_, double bytes_)
    : c(c_), kind(kind_), flops(flops_), bytes(bytes_) {
  if (!c->prof_on || c->noalloc) return;   // allocation-free calls create no events
  a = pool_event(c);
  b = pool_event(c);
  DMRG_CHECK_HIP(hipEventRecord(a, c->stream));
}
ProfScope::~ProfScope() {
  if (!a) return;
  if (hipEventRecord(b, c->stream) != hipSuccess) return;
  c->prof.push_back(ProfEntry{a, b, flops, bytes, kind});
  if (c->prof.size() > 4096) {
    try {
      prof_drain(c, false);
    } catch (...) {
    }
  }
}

}  // namespace dmrg

using dmrg::Error;

static thread_local std::string g_err;

#define API_BEGIN try {
#define API_END                     \
  return 0;                         \
  }                                 \
  catch (const Error& e) {          \
    g_err = e.msg;                  \
    return -1;                      \
  }                                 \
  catch (const std::exception& e) { \
    g_err = e.what();               \
    return -2;                      \
  }                                 \
  catch (...) {                     \
    g_err = "unknown error";        \
    return -3;                      \
  }

extern "C" {

int dmrg_version(void) { return 100; }

const char* dmrg_last_error(void) { return g_err.c_str(); }

int dmrg_ctx_create(int device, void* stream, dmrg_ctx** out) {
  API_BEGIN
  DMRG_REQUIRE(out != nullptr, "ctx_create: null out");
  int ndev = 0;
  DMRG_CHECK_HIP(hipGetDeviceCount(&ndev));
  DMRG_REQUIRE(device >= 0 && device < ndev, "ctx_create: no such device");
  DMRG_CHECK_HIP(hipSetDevice(device));
  dmrg_ctx* h = new dmrg_ctx();
  h->c.device = device;
  h->c.stream = (hipStream_t)stream;
  if (rocblas_create_handle(&h->c.blas) != rocblas_status_success) {
    delete h;
    throw Error("rocblas_create_handle failed");
  }
  rocblas_set_stream(h->c.blas, h->c.stream);
  DMRG_CHECK_HIP(hipHostMalloc((void**)&h->c.host_scalars, 64 * sizeof(double), hipHostMallocDefault));
  std::memset(h->c.host_scalars, 0, 64 * sizeof(double));
  try {
    h->c.scratch = scratch_address(device);   // device current (hipSetDevice above)
  } catch (...) {
    (void)hipHostFree(h->c.host_scalars);
    rocblas_destroy_handle(h->c.blas);
    delete h;
    throw;
  }
  *out = h;
  API_END
}

int dmrg_ctx_destroy(dmrg_ctx* ctx) {
  API_BEGIN
  if (!ctx) return 0;
  dmrg::Ctx& c = ctx->c;
  (void)hipStreamSynchronize(c.stream);
  for (auto& p : c.prof) {
    (void)hipEventDestroy(p.start);
    (void)hipEventDestroy(p.stop);
  }
  for (auto e : c.event_pool) (void)hipEventDestroy(e);
  for (auto e : c.trd_ev)
    if (e) (void)hipEventDestroy(e);
  for (int i = 0; i < dmrg::WS_NSLOTS; ++i) c.ws[i].release();
  if (c.host_scalars) (void)hipHostFree(c.host_scalars);
  if (c.blas) rocblas_destroy_handle(c.blas);
  delete ctx;
  API_END
}

int dmrg_sync(dmrg_ctx* ctx) {
  API_BEGIN
  DMRG_CHECK_HIP(hipStreamSynchronize(ctx->c.stream));
  API_END
}

int dmrg_heff_apply(dmrg_ctx* ctx, const dmrg_heff_desc* op, const double* x, double* y) {
  API_BEGIN
  DMRG_REQUIRE(ctx && op && x && y, "heff_apply: null argument");
  dmrg::HeffOp h = dmrg::heff_prepare(&ctx->c, *op);
  dmrg::heff_apply(&ctx->c, h, x, y);
  API_END
}

int dmrg_heff_apply_noalloc(dmrg_ctx* ctx, const dmrg_heff_desc* op, const double* x, double* y) {
  API_BEGIN
  DMRG_REQUIRE(ctx && op && x && y, "heff_apply_noalloc: null argument");
  dmrg::Ctx* c = &ctx->c;
  DMRG_REQUIRE(c->scratch != nullptr, "heff_apply_noalloc: no scratch array mapped for this context");
  const size_t need = dmrg::heff_scratch_bytes(*op);
  if (need > kScratchBytes)
    throw Error("heff_apply_noalloc: this operator needs up to " + std::to_string(need) +
                " bytes of workspace, more than the " + std::to_string(kScratchBytes) +
                "-byte scratch array; use dmrg_heff_apply");
  NoallocScope guard(c);
  dmrg::HeffOp h = dmrg::heff_prepare(c, *op);
  dmrg::heff_apply(c, h, x, y);
  API_END
}

int dmrg_rayleigh(dmrg_ctx* ctx, const dmrg_heff_desc* op, const double* x, double* value) {
  API_BEGIN
  DMRG_REQUIRE(ctx && op && x && value, "rayleigh: null argument");
  dmrg::HeffOp h = dmrg::heff_prepare(&ctx->c, *op);
  *value = dmrg::rayleigh(&ctx->c, h, x);
  API_END
}

int dmrg_env_update(dmrg_ctx* ctx, const dmrg_env_desc* env) {
  API_BEGIN
  DMRG_REQUIRE(ctx && env && env->ket && env->bra && env->W && env->env_in && env->env_out,
               "env_update: null argument");
  dmrg::env_update(&ctx->c, *env);
  API_END
}

int dmrg_lanczos_ground(dmrg_ctx* ctx, const dmrg_heff_desc* op, double* x, int maxiter, double tol,
                        double res_tol, double* energy, int* iters) {
  API_BEGIN
  DMRG_REQUIRE(ctx && op && x && energy, "lanczos: null argument");
  dmrg::HeffOp h = dmrg::heff_prepare(&ctx->c, *op);
  dmrg::lanczos_ground(&ctx->c, h, x, maxiter, tol, res_tol, energy, iters);
  API_END
}

int dmrg_krylov_expm(dmrg_ctx* ctx, const dmrg_heff_desc* op, double* x, double cre, double cim, int maxiter,
                     double tol, double errgoal, int* iters) {
  API_BEGIN
  DMRG_REQUIRE(ctx && op && x, "expm: null argument");
  dmrg::HeffOp h = dmrg::heff_prepare(&ctx->c, *op);
  dmrg::krylov_expm(&ctx->c, h, x, std::complex<double>(cre, cim), maxiter, tol, errgoal, iters);
  API_END
}

int dmrg_svd_split(dmrg_ctx* ctx, int field, const double* M, int m, int n, double cutoff, int maxdim,
                   int mindim, int absorb, double* U, double* V, int kcap, int* keep, double* discarded,
                   double* svals_host) {
  API_BEGIN
  DMRG_REQUIRE(ctx && M && U && V && keep, "svd_split: null argument");
  DMRG_REQUIRE(field == 0 || field == 1, "svd_split: bad field");
  dmrg::svd_split_entry(&ctx->c, field, M, m, n, cutoff, maxdim, mindim, absorb, U, V, kcap, keep, discarded,
                        svals_host);
  API_END
}

int dmrg_svd_split_guided(dmrg_ctx* ctx, int field, const double* M, int m, int n, double cutoff, int maxdim,
                          int mindim, int absorb, const double* G, int k0, int gside, int sketch, double* U,
                          double* V, int kcap, int* keep, double* discarded, double* svals_host, int* used,
                          double* resid) {
  API_BEGIN
  DMRG_REQUIRE(ctx && M && U && V && keep, "svd_split_guided: null argument");
  DMRG_REQUIRE(field == 0 || field == 1, "svd_split_guided: bad field");
  DMRG_REQUIRE(k0 >= 0 && sketch >= 0, "svd_split_guided: negative guide size");
  dmrg::svd_split_guided_entry(&ctx->c, field, M, m, n, cutoff, maxdim, mindim, absorb, G, k0, gside, sketch, U, V,
                               kcap, keep, discarded, svals_host, used, resid);
  API_END
}

int dmrg_set_eigensolver(dmrg_ctx* ctx, int mode) {
  API_BEGIN
  DMRG_REQUIRE(ctx, "set_eigensolver: null context");
  DMRG_REQUIRE(mode >= 0 && mode <= 3,
               "set_eigensolver: mode must be 0 (rocSOLVER syevd), 1 (own tridiagonalisation with early stop "
               "and back-transformation), 2 (own tridiagonalisation, rocSOLVER ormtr) or 3 (as 1, the "
               "reduction always complete)");
  ctx->c.eig_mode = mode;
  API_END
}

int dmrg_sym_eig(dmrg_ctx* ctx, int mode, double* A, int n, double* w, double* Z) {
  API_BEGIN
  DMRG_REQUIRE(ctx && A && w && Z, "sym_eig: null argument");
  DMRG_REQUIRE(mode >= 0 && mode <= 2, "sym_eig: mode must be 0, 1 or 2");
  dmrg::sym_eig_entry(&ctx->c, mode, A, n, w, Z);
  API_END
}

int dmrg_herm_eig(dmrg_ctx* ctx, int mode, double* A, int n, double* w, double* Z) {
  API_BEGIN
  DMRG_REQUIRE(ctx && A && w && Z, "herm_eig: null argument");
  DMRG_REQUIRE(mode >= 0 && mode <= 2, "herm_eig: mode must be 0, 1 or 2");
  dmrg::herm_eig_entry(&ctx->c, mode, A, n, w, Z);
  API_END
}

int dmrg_mpo_grad(dmrg_ctx* ctx, int field, int prec, const void* L, const double* bra, const double* ket,
                  const void* R, int Dl, int Dr, int Dlb, int Drb, int d, int wl, int wr, int ldL, int ldR,
                  double* G_host) {
  API_BEGIN
  DMRG_REQUIRE(ctx && L && bra && ket && R && G_host, "mpo_grad: null argument");
  DMRG_REQUIRE(field == 0 || field == 1, "mpo_grad: bad field");
  DMRG_REQUIRE(prec >= 0 && prec <= 2, "mpo_grad: bad precision");
  DMRG_REQUIRE(Dl > 0 && Dr > 0 && Dlb > 0 && Drb > 0 && d > 0 && wl > 0 && wr > 0, "mpo_grad: bad dims");
  if (ldL == 0) ldL = Dl;
  if (ldR == 0) ldR = Drb;
  DMRG_REQUIRE(ldL >= Dl && ldR >= Drb, "mpo_grad: row pitch smaller than the dimension");
  dmrg::mpo_grad(&ctx->c, field, prec, L, bra, ket, R, Dl, Dr, Dlb, Drb, d, wl, wr, ldL, ldR, G_host);
  API_END
}

int dmrg_gemm(dmrg_ctx* ctx, int prec, int field, int transA, int transB, int conjA, int conjB, int M, int N,
              int K, double alpha, const void* A, int lda, long long strideA, long long planeA, const void* B,
              int ldb, long long strideB, long long planeB, double beta, void* C, int ldc, long long strideC,
              long long planeC, int batch) {
  API_BEGIN
  DMRG_REQUIRE(ctx && A && B && C, "gemm: null argument");
  DMRG_REQUIRE(prec >= 0 && prec <= 2 && (field == 0 || field == 1), "gemm: bad prec/field");
  dmrg::GemmDesc g = dmrg::gdesc(M, N, K, lda, ldb, ldc, field, 0);
  g.transA = transA != 0;
  g.transB = transB != 0;
  g.conjA = conjA != 0;
  g.conjB = conjB != 0;
  g.alpha = alpha;
  g.beta = beta;
  g.sA = strideA; g.sB = strideB; g.sC = strideC;
  g.pA = planeA; g.pB = planeB; g.pC = planeC;
  g.batch = batch;
  dmrg::gemm_prec(&ctx->c, prec, g, A, B, C);
  API_END
}

int dmrg_prof_enable(dmrg_ctx* ctx, int on) {
  API_BEGIN
  dmrg::Ctx& c = ctx->c;
  dmrg::prof_drain(&c, true);
  for (int i = 0; i < 5; ++i) {
    c.prof_time_ms[i] = 0;
    c.prof_flops[i] = 0;
    c.prof_bytes[i] = 0;
    c.prof_calls[i] = 0;
  }
  c.prof_on = on != 0;
  API_END
}

int dmrg_prof_read(dmrg_ctx* ctx, double* time_ms, double* flops, long long* calls) {
  API_BEGIN
  dmrg::Ctx& c = ctx->c;
  dmrg::prof_drain(&c, true);
  for (int i = 0; i < 5; ++i) {
    if (time_ms) time_ms[i] = c.prof_time_ms[i];
    if (flops) flops[i] = c.prof_flops[i];
    if (calls) calls[i] = c.prof_calls[i];
  }
  API_END
}

int dmrg_prof_read_bytes(dmrg_ctx* ctx, double* bytes) {
  API_BEGIN
  dmrg::Ctx& c = ctx->c;
  dmrg::prof_drain(&c, true);
  for (int i = 0; i < 5; ++i)
    if (bytes) bytes[i] = c.prof_bytes[i];
  API_END
}

long long dmrg_matvec_count(dmrg_ctx* ctx) { return ctx ? ctx->c.matvecs : -1; }

}  // extern "C"
