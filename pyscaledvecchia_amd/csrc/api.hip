// C ABI of libsvh.so: argument checks, host<->device staging, error mapping.
// Each entry point is declared, with the reference function it replaces, in
// include/svh.h.

#include <cstring>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "svh_internal.h"
#include "bessel.h"

#include <cmath>

void launch_to_soa(svh_ctx* ctx, const double* dX, int64_t n, int d, const double* dlam,
                   double* dXT);
void vecchia_setup_device(svh_ctx* ctx, const double* Xo, const double* yo, const double* Zo,
                          int64_t n, int d, int q, int64_t nblk, const int64_t* rows,
                          const int64_t* nbr, const int32_t* kcnt, int mmax,
                          const int32_t* chunk_grad, const int32_t* chunk_val, int nchunk_grad,
                          int nchunk_val);
void predict_blocks_device(svh_ctx* ctx, const double* dXn, const int64_t* dnb, int nb_stride,
                           const double* dXt, const int64_t* dtg, int tg_stride,
                           const double* dmask, int noise_free, int64_t nblk, int K, int d,
                           const double* theta, double nu, const double* dresid, int mode,
                           double* dout_a, double* dmean, double* dvar, int* jitter_power_max,
                           int precision);

void batch_release(svh_ctx* ctx);
void ordering_batch_device(svh_ctx* ctx, int nd, const int64_t* n, const int32_t* d,
                           const int64_t* first, const double* X, int m, int64_t* order,
                           int64_t* nn);
void batch_setup_device(svh_ctx* ctx, int ndata, const int32_t* d, const int32_t* q,
                        const int32_t* nu_code, const int64_t* n, const double* X, const double* y,
                        const double* Z, const int64_t* nbr, const int32_t* kcnt,
                        const int32_t* chunk, const int32_t* nchunk, int mmax, const double* lvt,
                        const int32_t* fixed_nug, const double* log_nug, double lambda_max);
void batch_update_device(svh_ctx* ctx, int i, const double* X, const double* y, const double* Z,
                         const int64_t* nbr, const int32_t* kcnt);
void batch_fit_start_device(svh_ctx* ctx, const double* theta0, int max_iter, double tol,
                            double var_penalty, int use_graph);
void batch_fit_step_device(svh_ctx* ctx, int nrep, int32_t* status);
void batch_get_device(svh_ctx* ctx, double* theta, double* ll, double* beta, int32_t* status,
                      int32_t* iters, int32_t* nevals);
void batch_set_reorder_device(svh_ctx* ctx, int i, int rebuilt);
void batch_eval_device(svh_ctx* ctx, const double* theta, double* ll, double* grad,
                       double* fisher, double* beta, int32_t* jitter_rounds);

namespace {

// Fixed working array of svh_predict_noalloc: a device global, so it exists on
// every GPU the library is loaded on (each GPU its own copy) for the life of
// the process.  Host code must not pass `scratch` itself to a kernel or copy:
// on the host that name is only a stand-in; the device address comes from
// hipGetSymbolAddress (scratch_address) and is kept per device.
constexpr size_t kScratchBytes = size_t(64) << 20;
__device__ __align__(256) char scratch[kScratchBytes];

// Device address of this device's `scratch`, looked up once per device with
// that device current (called from svh_create).  nullptr if the lookup fails.
void* scratch_address(int device) {
  static std::mutex mu;
  static std::unordered_map<int, void*> addr;
  std::lock_guard<std::mutex> lock(mu);
  auto it = addr.find(device);
  if (it != addr.end()) return it->second;
  void* p = nullptr;
  if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess) {
    (void)hipGetLastError();
    p = nullptr;
  }
  addr[device] = p;
  return p;
}

template <class F>
int guarded(svh_ctx* ctx, F&& f) {
  if (!ctx) return SVH_ERR_ARG;
  try {
    SVH_HIP(hipSetDevice(ctx->device));
    f();
    ctx->last_error.clear();
    return SVH_OK;
  } catch (const SvhError& e) {
    ctx->last_error = e.msg;
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    return e.code;
  } catch (const std::exception& e) {
    ctx->last_error = e.what();
    return SVH_ERR_HIP;
  }
}

void upload(svh_ctx* ctx, void* dst, const void* src, size_t bytes) {
  if (bytes) SVH_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
}

void download(svh_ctx* ctx, void* dst, const void* src, size_t bytes) {
  if (bytes) SVH_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
}

void sync(svh_ctx* ctx) {
  SVH_HIP(hipStreamSynchronize(ctx->stream));
  ctx->timers.collect();
}

}  // namespace

extern "C" {

int svh_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

const char* svh_version(void) { return "pyscaledvecchia_amd 0.1.0 (gfx950)"; }

int svh_create(int device, svh_ctx** out) {
  if (!out) return SVH_ERR_ARG;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return SVH_ERR_HIP;
  svh_ctx* ctx = new svh_ctx();
  ctx->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&ctx->stream) != hipSuccess) {
    delete ctx;
    return SVH_ERR_HIP;
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) ctx->num_cu = prop.multiProcessorCount;
  ctx->scratch = scratch_address(device);
  *out = ctx;
  return SVH_OK;
}

void svh_destroy(svh_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  batch_release(ctx);
  mlp_release(ctx);
  ctx->timers.release();
  ctx->arena.release();
  ctx->stage.release();
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

const char* svh_last_error(const svh_ctx* ctx) { return ctx ? ctx->last_error.c_str() : ""; }

int svh_maximin(svh_ctx* ctx, const double* X, int64_t n, int32_t d, int64_t first,
                int64_t* order) {
  return guarded(ctx, [&] {
    SVH_REQUIRE(n >= 0 && d >= 1 && d <= 32, "maximin: bad n/d");
    if (n == 0) return;
    SVH_REQUIRE(first >= 0 && first < n, "maximin: first out of range");
    SVH_REQUIRE(X && order, "maximin: null pointer");
    double* dX = ctx->arena.get_t<double>("mm.X", (size_t)n * d);
    double* dXT = ctx->arena.get_t<double>("mm.XT", (size_t)n * d);
    int64_t* dord = ctx->arena.get_t<int64_t>("mm.order", n);
    upload(ctx, dX, X, (size_t)n * d * sizeof(double));
    launch_to_soa(ctx, dX, n, d, nullptr, dXT);
    launch_maximin(ctx, dXT, n, d, first, dord);
    download(ctx, order, dord, (size_t)n * sizeof(int64_t));
    sync(ctx);
  });
}

int svh_ordered_knn(svh_ctx* ctx, const double* X, int64_t n, int32_t d, int32_t m, int64_t start,
                    int64_t* nn) {
  return guarded(ctx, [&] {
    SVH_REQUIRE(n >= 0 && d >= 1 && d <= 32 && m >= 0, "ordered_knn: bad n/d/m");
    if (n == 0) return;
    SVH_REQUIRE(n < (1ll << 31), "ordered_knn: n too large");
    SVH_REQUIRE(X && nn, "ordered_knn: null pointer");
    double* dX = ctx->arena.get_t<double>("nn.X", (size_t)n * d);
    double* dXT = ctx->arena.get_t<double>("nn.XT", (size_t)n * d);
    int64_t* dnn = ctx->arena.get_t<int64_t>("nn.out", (size_t)n * (m + 1));
    upload(ctx, dX, X, (size_t)n * d * sizeof(double));
    launch_to_soa(ctx, dX, n, d, nullptr, dXT);
    launch_ordered_knn(ctx, dXT, n, d, m, start, dnn);
    SVH_HIP(hipGetLastError());
    download(ctx, nn, dnn, (size_t)n * (m + 1) * sizeof(int64_t));
    sync(ctx);
  });
}

int svh_ordered_knn_rows(svh_ctx* ctx, const double* X, int64_t n, int32_t d, int32_t m,
                         int64_t start, int64_t* nn) {
  return guarded(ctx, [&] {
    SVH_REQUIRE(n >= 0 && d >= 1 && d <= 32 && m >= 0, "ordered_knn_rows: bad n/d/m");
    SVH_REQUIRE(start >= 0 && start <= n, "ordered_knn_rows: need 0 <= start <= n");
    if (n == start) return;
    SVH_REQUIRE(n < (1ll << 31), "ordered_knn_rows: n too large");
    SVH_REQUIRE(X && nn, "ordered_knn_rows: null pointer");
    double* dX = ctx->arena.get_t<double>("nn.X", (size_t)n * d);
    double* dXT = ctx->arena.get_t<double>("nn.XT", (size_t)n * d);
    int64_t* dnn = ctx->arena.get_t<int64_t>("nn.out", (size_t)(n - start) * (m + 1));
    upload(ctx, dX, X, (size_t)n * d * sizeof(double));
    launch_to_soa(ctx, dX, n, d, nullptr, dXT);
    launch_ordered_knn(ctx, dXT, n, d, m, start, dnn, start);
    SVH_HIP(hipGetLastError());
    download(ctx, nn, dnn, (size_t)(n - start) * (m + 1) * sizeof(int64_t));
    sync(ctx);
  });
}

int svh_knn_query(svh_ctx* ctx, const double* Xref, int64_t nref, const double* Xq, int64_t nq,
                  int32_t d, int32_t k, int64_t* nb) {
  return guarded(ctx, [&] {
    SVH_REQUIRE(nref >= 0 && nq >= 0 && d >= 1 && d <= 32, "knn_query: bad sizes");
    SVH_REQUIRE(k >= 1 && k <= nref, "knn_query: need 1 <= k <= nref");
    SVH_REQUIRE(nref < (1ll << 31), "knn_query: nref too large");
    if (nq == 0) return;
    double* dR = ctx->arena.get_t<double>("q.R", (size_t)nref * d);
    double* dRT = ctx->arena.get_t<double>("q.RT", (size_t)nref * d);
    double* dQ = ctx->arena.get_t<double>("q.Q", (size_t)nq * d);
    double* dQT = ctx->arena.get_t<double>("q.QT", (size_t)nq * d);
    int64_t* dnb = ctx->arena.get_t<int64_t>("q.nb", (size_t)nq * k);
    upload(ctx, dR, Xref, (size_t)nref * d * sizeof(double));
    upload(ctx, dQ, Xq, (size_t)nq * d * sizeof(double));
    launch_to_soa(ctx, dR, nref, d, nullptr, dRT);
    launch_to_soa(ctx, dQ, nq, d, nullptr, dQT);
    launch_knn_query(ctx, dRT, nref, dQT, nq, d, k, dnb, k);
    download(ctx, nb, dnb, (size_t)nq * k * sizeof(int64_t));
    sync(ctx);
  });
}

int svh_vecchia_setup(svh_ctx* ctx, const double* Xo, const double* yo, const double* Zo,
                      int64_t n, int32_t d, int32_t q, int64_t nblk, const int64_t* rows,
                      const int64_t* nbr, const int32_t* kcnt, int32_t mmax,
                      const int32_t* chunk_grad, const int32_t* chunk_val, int32_t nchunk_grad,
                      int32_t nchunk_val) {
  return guarded(ctx, [&] {
    SVH_REQUIRE(Xo && yo && rows && kcnt && chunk_grad && chunk_val, "setup: null pointer");
    SVH_REQUIRE(q == 0 || Zo, "setup: Zo is NULL with q > 0");
    SVH_REQUIRE(mmax == 0 || nbr, "setup: nbr is NULL");
    vecchia_setup_device(ctx, Xo, yo, Zo, n, d, q, nblk, rows, nbr, kcnt, mmax, chunk_grad,
                         chunk_val, nchunk_grad, nchunk_val);
  });
}

int svh_vecchia_rebuild(svh_ctx* ctx, const double* Xe, const double* ye, const double* Ze,
                        int64_t n, int32_t d, int32_t q, const double* lam, int64_t first,
                        int32_t m, int32_t upload_data, int64_t* order) {
  return guarded(ctx, [&] {
    SVH_REQUIRE(lam && order, "rebuild: null pointer");
    for (int l = 0; l < d && l < 32; ++l)
      SVH_REQUIRE(std::isfinite(lam[l]) && lam[l] > 0.0, "rebuild: ranges must be positive");
    vecchia_rebuild_device(ctx, Xe, ye, Ze, n, d, q, lam, first, m, upload_data, order);
  });
}

int svh_vecchia_eval(svh_ctx* ctx, const double* theta, int32_t P, double nu, int32_t precision,
                     int32_t need_grad, double var_penalty, double log_var_target, double* ll,
                     double* grad, double* fisher, double* beta, int32_t* jitter_power_max) {
  return guarded(ctx, [&] {
    SVH_REQUIRE(theta && ll, "eval: null pointer");
    SVH_REQUIRE(!need_grad || (grad && fisher), "eval: grad/fisher NULL with need_grad");
    SVH_REQUIRE(ctx->vs.q == 0 || beta, "eval: beta NULL");
    int jp = 0;
    vecchia_eval_device(ctx, theta, P, nu, precision, need_grad, var_penalty, log_var_target, ll,
                        grad, fisher, beta, &jp);
    if (jitter_power_max) *jitter_power_max = jp;
  });
}

int svh_predict(svh_ctx* ctx, const double* Xtr, const double* resid, int64_t ntr,
                const double* Xq, int64_t nq, int32_t d, int32_t m, const double* theta,
                double nu, int32_t precision, int32_t noise_free, double* mean, double* var,
                int64_t* nb_out) {
  return guarded(ctx, [&] {
    SVH_REQUIRE(ntr >= 1 && nq >= 0 && d >= 1 && d <= 32, "predict: bad sizes");
    SVH_REQUIRE(m >= 1 && m <= ntr, "predict: need 1 <= m <= ntr");
    SVH_REQUIRE(precision == SVH_FP64 || precision == SVH_MIXED, "predict: bad precision mode");
    SVH_REQUIRE(Xtr && resid && theta && mean && var, "predict: null pointer");
    if (nq == 0) return;
    double* dR = ctx->arena.get_t<double>("pr.R", (size_t)ntr * d);
    double* dRT = ctx->arena.get_t<double>("pr.RT", (size_t)ntr * d);
    double* dQ = ctx->arena.get_t<double>("pr.Q", (size_t)nq * d);
    double* dQT = ctx->arena.get_t<double>("pr.QT", (size_t)nq * d);
    double* dres = ctx->arena.get_t<double>("pr.resid", ntr);
    double* dlam = ctx->arena.get_t<double>("pr.lam", 32);
    int64_t* dnb = ctx->arena.get_t<int64_t>("pr.nb", (size_t)nq * m);
    double* dmean = ctx->arena.get_t<double>("pr.mean", nq);
    double* dvar = ctx->arena.get_t<double>("pr.var", nq);
    double lam[32];
    for (int l = 0; l < d; ++l) lam[l] = std::exp(theta[1 + l]);
    upload(ctx, dR, Xtr, (size_t)ntr * d * sizeof(double));
    upload(ctx, dQ, Xq, (size_t)nq * d * sizeof(double));
    upload(ctx, dres, resid, (size_t)ntr * sizeof(double));
    upload(ctx, dlam, lam, d * sizeof(double));
    launch_to_soa(ctx, dR, ntr, d, dlam, dRT);
    launch_to_soa(ctx, dQ, nq, d, dlam, dQT);
    launch_knn_query(ctx, dRT, ntr, dQT, nq, d, m, dnb, m);
    predict_blocks_device(ctx, dR, dnb, m, dQ, nullptr, 0, nullptr, noise_free, nq, m + 1, d,
                          theta, nu, dres, 1, nullptr, dmean, dvar, nullptr, precision);
    download(ctx, mean, dmean, (size_t)nq * sizeof(double));
    download(ctx, var, dvar, (size_t)nq * sizeof(double));
    if (nb_out) download(ctx, nb_out, dnb, (size_t)nq * m * sizeof(int64_t));
    sync(ctx);
  });
}

int svh_predict_noalloc(svh_ctx* ctx, const double* Xtr, const double* resid, int64_t ntr,
                        const double* Xq, int64_t nq, int32_t d, int32_t m, const double* theta,
                        double nu, int32_t precision, int32_t noise_free, double* mean,
                        double* var) {
  return guarded(ctx, [&] {
    SVH_REQUIRE(ntr >= 1 && nq >= 0 && d >= 1 && d <= 32, "predict_noalloc: bad sizes");
    SVH_REQUIRE(m >= 1 && m <= ntr, "predict_noalloc: need 1 <= m <= ntr");
    SVH_REQUIRE(precision == SVH_FP64 || precision == SVH_MIXED,
                "predict_noalloc: bad precision mode");
    SVH_REQUIRE(Xtr && resid && theta && mean && var, "predict_noalloc: null pointer");
    SVH_REQUIRE(ctx->scratch, "predict_noalloc: the fixed device array is not available");
    // working memory: SoA copies of both point sets, the neighbour lists, and
    // at most 64 KiB of small per-call buffers (ranges, jitter levels/flags)
    const double need = 8.0 * ((double)(ntr + nq) * d + (double)nq * m) + 65536.0;
    SVH_REQUIRE(need <= (double)kScratchBytes,
                "predict_noalloc: 8 (ntr + nq) d + 8 nq m bytes exceed the 64 MiB fixed device "
                "array; split the queries");
    if (nq == 0) return;
    // every request of this call is carved from `scratch` (no hipMalloc), the
    // host staging is a small array of the context, timers and tuner are off
    struct Fixed {
      svh_ctx* c;
      bool timers;
      int tune;
      Fixed(svh_ctx* c_) : c(c_), timers(c_->timers.enabled), tune(c_->tune.mode) {
        c->arena.fix_base = static_cast<char*>(c->scratch);
        c->arena.fix_bytes = kScratchBytes;
        c->arena.fix_used = 0;
        c->stage.fixed = true;
        c->timers.enabled = false;
        c->tune.mode = 0;
      }
      ~Fixed() {
        c->arena.fix_base = nullptr;
        c->arena.fix_bytes = c->arena.fix_used = 0;
        c->stage.fixed = false;
        c->timers.enabled = timers;
        c->tune.mode = tune;
      }
    } fixed(ctx);
    double* dRT = ctx->arena.get_t<double>("pr.RT", (size_t)ntr * d);
    double* dQT = ctx->arena.get_t<double>("pr.QT", (size_t)nq * d);
    int64_t* dnb = ctx->arena.get_t<int64_t>("pr.nb", (size_t)nq * m);
    double* dlam = ctx->arena.get_t<double>("pr.lam", 32);
    double lam[32];
    for (int l = 0; l < d; ++l) lam[l] = std::exp(theta[1 + l]);
    upload(ctx, dlam, lam, d * sizeof(double));
    launch_to_soa(ctx, Xtr, ntr, d, dlam, dRT);
    launch_to_soa(ctx, Xq, nq, d, dlam, dQT);
    launch_knn_query(ctx, dRT, ntr, dQT, nq, d, m, dnb, m);
    predict_blocks_device(ctx, Xtr, dnb, m, Xq, nullptr, 0, nullptr, noise_free, nq, m + 1, d,
                          theta, nu, resid, 1, nullptr, mean, var, nullptr, precision);
    sync(ctx);
  });
}

int svh_block_columns(svh_ctx* ctx, const double* Xall, int64_t nall, int32_t d,
                      const int64_t* blocks, const double* mask, int64_t nblk, int32_t K,
                      const double* theta, double nu, int32_t precision, double* a) {
  return guarded(ctx, [&] {
    SVH_REQUIRE(nall >= 1 && d >= 1 && d <= 32 && K >= 1 && nblk >= 0, "block_columns: sizes");
    SVH_REQUIRE(Xall && blocks && theta && a, "block_columns: null pointer");
    SVH_REQUIRE(precision == SVH_FP64 || precision == SVH_MIXED, "block_columns: precision");
    if (nblk == 0) return;
    for (int64_t e = 0; e < nblk * K; ++e)
      SVH_REQUIRE(blocks[e] >= 0 && blocks[e] < nall, "block_columns: index out of range");
    double* dX = ctx->arena.get_t<double>("bc.X", (size_t)nall * d);
    int64_t* dblk = ctx->arena.get_t<int64_t>("bc.blk", (size_t)nblk * K);
    double* dmask = mask ? ctx->arena.get_t<double>("bc.mask", (size_t)nblk * K) : nullptr;
    double* da = ctx->arena.get_t<double>("bc.a", (size_t)nblk * K);
    upload(ctx, dX, Xall, (size_t)nall * d * sizeof(double));
    upload(ctx, dblk, blocks, (size_t)nblk * K * sizeof(int64_t));
    if (mask) upload(ctx, dmask, mask, (size_t)nblk * K * sizeof(double));
    predict_blocks_device(ctx, dX, dblk, K, dX, dblk + (K - 1), K, dmask, 0, nblk, K, d, theta, nu,
                          nullptr, 0, da, nullptr, nullptr, nullptr, precision);
    download(ctx, a, da, (size_t)nblk * K * sizeof(double));
    sync(ctx);
  });
}

int svh_ordering_batch(svh_ctx* ctx, int32_t ndata, const int64_t* n, const int32_t* d,
                       const int64_t* first, const double* X, int32_t m, int64_t* order,
                       int64_t* nn) {
  return guarded(ctx, [&] {
    SVH_REQUIRE(ndata >= 1 && n && d && first && X && order && nn && m >= 0,
                "ordering_batch: bad argument");
    ordering_batch_device(ctx, ndata, n, d, first, X, m, order, nn);
  });
}

int svh_batch_setup(svh_ctx* ctx, int32_t ndata, const int32_t* d, const int32_t* q,
                    const int32_t* nu_code, const int64_t* n, const double* X, const double* y,
                    const double* Z, const int64_t* nbr, const int32_t* kcnt,
                    const int32_t* chunk, const int32_t* nchunk, int32_t mmax, const double* lvt,
                    const int32_t* fixed_nug, const double* log_nug, double lambda_max) {
  return guarded(ctx, [&] {
    SVH_REQUIRE(d && q && nu_code && n && chunk && nchunk && lvt && fixed_nug && log_nug,
                "batch_setup: null pointer");
    SVH_REQUIRE((X == nullptr) == (y == nullptr), "batch_setup: X and y both given or both NULL");
    SVH_REQUIRE((nbr == nullptr) == (kcnt == nullptr),
                "batch_setup: nbr and kcnt both given or both NULL");
    SVH_REQUIRE(ndata >= 1, "batch_setup: no datasets");
    if (X && !Z)
      for (int i = 0; i < ndata; ++i)
        SVH_REQUIRE(q[i] == 0, "batch_setup: Z is NULL with q > 0");
    batch_setup_device(ctx, ndata, d, q, nu_code, n, X, y, Z, nbr, kcnt, chunk, nchunk, mmax, lvt,
                       fixed_nug, log_nug, lambda_max);
  });
}

int svh_batch_update(svh_ctx* ctx, int32_t i, const double* X, const double* y, const double* Z,
                     const int64_t* nbr, const int32_t* kcnt) {
  return guarded(ctx, [&] {
    SVH_REQUIRE(X && y && nbr && kcnt, "batch_update: null pointer");
    batch_update_device(ctx, i, X, y, Z, nbr, kcnt);
  });
}

int svh_joint_solve(svh_ctx* ctx, int64_t nobs, int64_t ns, int32_t K, const int64_t* blocks,
                    const double* a, const double* resid, int32_t nrhs, const double* rhs,
                    double* mean, double* x, int32_t exact_var, double* var) {
  return guarded(ctx, [&] {
    SVH_REQUIRE(blocks && a, "joint_solve: null pointer");
    if (ns == 0) return;
    joint_solve_device(ctx, nobs, ns, K, blocks, a, resid, nrhs, rhs, mean, x, exact_var, var);
  });
}

int svh_batch_data(svh_ctx* ctx, const double* Xe, const double* ye, const double* Ze) {
  return guarded(ctx, [&] {
    SVH_REQUIRE(Xe && ye, "batch_data: null pointer");
    SVH_REQUIRE(batch_layout(ctx, 0).nz == 0 || Ze, "batch_data: Ze is NULL with q > 0");
    batch_data_device(ctx, Xe, ye, Ze);
  });
}

int svh_batch_rebuild(svh_ctx* ctx, int32_t count, const int32_t* which, const double* lam,
                      const int64_t* first, const int32_t* m) {
  return guarded(ctx, [&] {
    SVH_REQUIRE(count >= 0 && (count == 0 || (which && lam && first && m)),
                "batch_rebuild: null pointer");
    batch_rebuild_device(ctx, count, which, lam, first, m);
  });
}

int svh_batch_fit_start(svh_ctx* ctx, const double* theta0, int32_t max_iter, double tol,
                        double var_penalty, int32_t use_graph) {
  return guarded(ctx, [&] {
    SVH_REQUIRE(theta0, "batch_fit_start: null pointer");
    batch_fit_start_device(ctx, theta0, max_iter, tol, var_penalty, use_graph);
  });
}

int svh_batch_fit_step(svh_ctx* ctx, int32_t nrep, int32_t* status) {
  return guarded(ctx, [&] {
    SVH_REQUIRE(status && nrep >= 1, "batch_fit_step: bad argument");
    batch_fit_step_device(ctx, nrep, status);
  });
}

int svh_batch_get(svh_ctx* ctx, double* theta, double* ll, double* beta, int32_t* status,
                  int32_t* iters, int32_t* nevals) {
  return guarded(ctx, [&] {
    SVH_REQUIRE(theta && ll && beta && status && iters && nevals, "batch_get: null pointer");
    batch_get_device(ctx, theta, ll, beta, status, iters, nevals);
  });
}

int svh_batch_set_reorder(svh_ctx* ctx, int32_t i, int32_t rebuilt) {
  return guarded(ctx, [&] { batch_set_reorder_device(ctx, i, rebuilt); });
}

int svh_batch_eval(svh_ctx* ctx, const double* theta, double* ll, double* grad, double* fisher,
                   double* beta, int32_t* jitter_rounds) {
  return guarded(ctx, [&] {
    SVH_REQUIRE(theta && ll && grad && fisher && beta, "batch_eval: null pointer");
    batch_eval_device(ctx, theta, ll, grad, fisher, beta, jitter_rounds);
  });
}

int svh_timing_enable(svh_ctx* ctx, int32_t enable) {
  return guarded(ctx, [&] { ctx->timers.enabled = enable != 0; });
}

int svh_timing_read(svh_ctx* ctx, const char* name, double* total_ms, int64_t* count) {
  return guarded(ctx, [&] {
    SVH_REQUIRE(name && total_ms && count, "timing_read: null pointer");
    if (std::string(name) == "maximin_rounds") {
      *total_ms = 0.0;
      *count = ctx->mm_rounds;
      return;
    }
    if (std::string(name) == "maximin_fallbacks") {
      *total_ms = 0.0;
      *count = ctx->mm_fallbacks;
      return;
    }
    auto it = ctx->timers.acc.find(name);
    *total_ms = it == ctx->timers.acc.end() ? 0.0 : it->second.ms;
    *count = it == ctx->timers.acc.end() ? 0 : it->second.count;
  });
}

int svh_timing_reset(svh_ctx* ctx) {
  return guarded(ctx, [&] { ctx->timers.acc.clear(); });
}

}  // extern "C"

int svh_matern_general(double nu, const double* r, int64_t n, double* M, double* g) {
  if (!(nu > 0.0) || n < 0 || (n > 0 && (!r || !M || !g))) return SVH_ERR_ARG;
  const svh_bessel::MaternNu m = svh_bessel::matern_nu(nu);
  for (int64_t i = 0; i < n; ++i) svh_bessel::matern_general(m, r[i], M[i], g[i]);
  return SVH_OK;
}
