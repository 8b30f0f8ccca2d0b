/*
 * svh.h -- C ABI of libsvh.so, the gfx950 (MI355X) scaled-Vecchia hot path.
 *
 * Plain pointers and sizes only.  Every host array is caller-owned,
 * C-contiguous, float64 / int64 / int32 as declared; no host pointer is kept
 * past a call.  Device buffers belong to the context and persist across calls
 * (setup once, evaluate many times).  Every call is synchronous on return.
 * A context is not thread-safe: use one per host thread.
 *
 * The reference (jdtuck/pyscaledvecchia) is a Python package with no FFI; each
 * entry point below replaces the Python function cited beside it, and the
 * Python package pyscaledvecchia_amd binds them with ctypes.
 */
#ifndef SVH_H
#define SVH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define SVH_OK 0
#define SVH_ERR_CHOLESKY 1 /* block Cholesky failed after the jitter ladder  */
#define SVH_ERR_ARG 2      /* bad argument (shape, size, value)              */
#define SVH_ERR_HIP 3      /* HIP runtime error                              */

/* precision modes of svh_vecchia_eval / svh_predict */
#define SVH_FP64 0  /* every stage in float64                                  */
#define SVH_MIXED 1 /* likelihood: float32 derivative terms dS_p a, W solves and
                       Fisher / cross products (float32 accumulation); float64
                       covariance, Cholesky, log-det, a = L^-T e_K, GLS.
                       prediction blocks: float32 matrix-core trailing updates */
#define SVH_MIXED_F32COV 2 /* svh_vecchia_eval only: SVH_MIXED plus the
                       off-diagonal covariance entries (distances, Matern)
                       assembled in float32; diagonal and factorisation float64 */

typedef struct svh_ctx svh_ctx;

/* number of visible HIP devices (0 when none) */
int svh_device_count(void);
/* library version string */
const char* svh_version(void);

/* Create a context on HIP device `device` (one stream, default priority). */
int svh_create(int device, svh_ctx** out);
void svh_destroy(svh_ctx* ctx);
/* Message of the last failed call on this context ("" if none). */
const char* svh_last_error(const svh_ctx* ctx);

/* General Matern correlation on the host: M[i] = M_nu(r[i]) and g[i] =
 * M_nu'(r[i]) / r[i] for any nu > 0 (2^(1-nu)/Gamma(nu) c^nu K_nu(c),
 * c = sqrt(2 nu) r; M = 1 and g = 0 below r = 1e-7), by the same code the
 * kernels run for smoothness outside {0.5, 1.5, 2.5}.  Needs no device or
 * context.  Replaces _matern_corr_general / _matern_g_general
 * (_matern_general.py:92-116). */
int svh_matern_general(double nu, const double* r, int64_t n, double* M, double* g);

/* ---------------------------------------------------------------- (a) --- */

/* Exact maximin order of the n rows of X (n x d, scaled coordinates),
 * starting at row `first`.  Ties take the lowest index.  Writes order[n].
 * Replaces ordering.py:27-73 (_maximin_core / maximin_order; the caller
 * computes `first` as ordering.py:72 does). */
int svh_maximin(svh_ctx* ctx, const double* X, int64_t n, int32_t d, int64_t first,
                int64_t* order);

/* Ordered nearest neighbours: row i of nn (n x (m+1)) gets i in column 0 and
 * the min(i, m) nearest j < i (by squared distance, then index; nearest first)
 * in columns 1..; -1 elsewhere.  Rows below `start` get only column 0.
 * Replaces ordering.py:76-116 (find_ordered_nn). */
int svh_ordered_knn(svh_ctx* ctx, const double* X, int64_t n, int32_t d, int32_t m,
                    int64_t start, int64_t* nn);

/* Rows start..n-1 of svh_ordered_knn only: writes nn ((n - start) x (m+1)),
 * its row i - start being row i of the full result.  The form gp.py:326-333
 * reads (find_ordered_nn with start = n, ordering.py:76-116, of which only
 * the prediction rows are used), without the n x (m+1) rows before start. */
int svh_ordered_knn_rows(svh_ctx* ctx, const double* X, int64_t n, int32_t d, int32_t m,
                         int64_t start, int64_t* nn);

/* k nearest rows of Xref (nref x d) for each row of Xq (nq x d), sorted by
 * (squared distance, index).  Writes nb (nq x k).
 * Replaces the cKDTree query of gp.py:253-254. */
int svh_knn_query(svh_ctx* ctx, const double* Xref, int64_t nref, const double* Xq,
                  int64_t nq, int32_t d, int32_t k, int64_t* nb);

/* Several datasets at once: for dataset i (rows n[i] x d[i] of X, stored back
 * to back), the maximin order from first[i] (svh_maximin) and the ordered
 * neighbours of the reordered rows (svh_ordered_knn with start 0).  Orders are
 * dataset-local, written back to back to order; nn rows (m+1 wide) likewise.
 * Replaces the rebuild of gp.py:182-191 for many fits at once (d <= 12). */
int svh_ordering_batch(svh_ctx* ctx, int32_t ndata, const int64_t* n, const int32_t* d,
                       const int64_t* first, const double* X, int32_t m, int64_t* order,
                       int64_t* nn);

/* ---------------------------------------------------------------- (b) --- */

/* Upload one Vecchia structure (inputs already in Vecchia order).
 *   Xo (n x d), yo (n), Zo (n x q, may be NULL when q == 0)
 *   nblk blocks; block b conditions row rows[b] on nbr[b*mmax .. +kcnt[b]]
 *   (its neighbours; self is placed last).  kcnt[b] <= mmax <= 127.
 *   chunk_grad[b] / chunk_val[b]: the jitter-chunk id of block b when the
 *   gradient is / is not requested (_covariance.py:255-259 slicing of each
 *   conditioning-size group, likelihood.py:111-114).
 * Replaces the per-rebuild state of gp.py:182-191 + ordering.py:119-127. */
int svh_vecchia_setup(svh_ctx* ctx, const double* Xo, const double* yo, const double* Zo,
                      int64_t n, int32_t d, int32_t q, int64_t nblk, const int64_t* rows,
                      const int64_t* nbr, const int32_t* kcnt, int32_t mmax,
                      const int32_t* chunk_grad, const int32_t* chunk_val, int32_t nchunk_grad,
                      int32_t nchunk_val);

/* The whole rebuild of a fit at new ranges, on the device: the estimation
 * data Xe (n x d), ye (n), Ze (n x q, NULL when q == 0) -- uploaded when
 * upload_data is nonzero, otherwise the copy of the previous call with the
 * same n, d, q is reused -- scaled by lam (d positive ranges, X / lam as the
 * host computes it), the maximin order from `first` (svh_maximin; first = -1:
 * the row closest to the scaled centroid, ordering.py:72, found on the device
 * with the host's NumPy arithmetic -- column means summed in row order,
 * pairwise row sums -- so the same index), the
 * ordered m nearest neighbours of the reordered scaled rows (svh_ordered_knn,
 * start 0), and the structure svh_vecchia_setup would receive for the groups
 * of ordering.py:119-127: blocks in row order, block i conditioning row i on
 * its min(i, m) neighbours, jitter chunks of _covariance.py:255-259.  Writes
 * the order (n) to `order`; the neighbour lists stay on the device.
 * Requires 0 <= m < n, m <= 126, d <= 12.  One call replaces
 * gp.py:182-191 (maximin_order, find_ordered_nn, _nn_groups and the
 * structure upload) without moving the neighbour lists through the host. */
int svh_vecchia_rebuild(svh_ctx* ctx, const double* Xe, const double* ye, const double* Ze,
                        int64_t n, int32_t d, int32_t q, const double* lam, int64_t first,
                        int32_t m, int32_t upload_data, int64_t* order);

/* One profile Vecchia log-likelihood evaluation on the uploaded structure.
 *   theta (P = d+2): log sigma^2, log lambda_1..d, log tau, for a fixed
 *   smoothness nu > 0 (closed forms at 0.5, 1.5, 2.5; any other value uses
 *   the Bessel form of _matern_general.py:92-116).  nu == 0: the smoothness
 *   is estimated, theta (P = d+3) = log sigma^2, log lambda_1..d, log nu,
 *   log tau, and the log-nu derivative is a central difference in nu with
 *   step 1e-3 (_matern_general.py:134-179); the structure's gradient jitter
 *   chunks must then follow P = d+3 (svh_vecchia_rebuild provides both).
 *   Outputs: ll, grad[P], fisher[P*P] (row-major), beta[q]; grad/fisher are
 *   untouched when need_grad == 0.  *jitter_power_max receives the largest
 *   jitter power used by any chunk (0 = none, else -12..-4).
 * Replaces likelihood.py:66-175 (vecchia_profile_loglik), fixed or estimated nu. */
int svh_vecchia_eval(svh_ctx* ctx, const double* theta, int32_t P, double nu,
                     int32_t precision, int32_t need_grad, double var_penalty,
                     double log_var_target, double* ll, double* grad, double* fisher,
                     double* beta, int32_t* jitter_power_max);

/* ---------------------------------------------------------------- (c) --- */

/* Marginal prediction (gp.py:242-277, _predict_scaled): for each of the nq
 * rows of Xq, its m nearest training rows in the scaled space x/lambda
 * (sorted), the block [neighbours..., target] with a noise-free target
 * (noise_free != 0), and
 *   mean[i] = -a[0:m].resid[nb] / a[m],   var[i] = 1 / a[m]^2
 * where a = L^-T e_K.  Xtr (ntr x d) and Xq are standardised inputs, resid
 * (ntr) the centred responses, theta = log sigma^2, log lambda_1..d, log tau
 * (d+2 entries) and nu > 0 the smoothness (a fitted one too).
 * nb_out (nq x m) receives the neighbour indices when not NULL. */
int svh_predict(svh_ctx* ctx, const double* Xtr, const double* resid, int64_t ntr,
                const double* Xq, int64_t nq, int32_t d, int32_t m, const double* theta,
                double nu, int32_t precision, int32_t noise_free, double* mean,
                double* var, int64_t* nb_out);

/* svh_predict without any device allocation, for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  No counterpart in the reference.  Every array is a
 * DEVICE pointer of this context's GPU: Xtr (ntr x d), resid (ntr), Xq
 * (nq x d) in, mean and var (nq) out; theta stays on the host.  Same results
 * as svh_predict (neighbour indices are not returned).  Working memory comes
 * from a fixed 64 MiB device array that the library holds on each GPU for the
 * life of the process: 8 (ntr + nq) d + 8 nq m bytes (+ 64 KiB) must fit,
 * otherwise SVH_ERR_ARG (split the queries).  Calls on one GPU share that
 * array: calls that overlap in time on one GPU (other threads, other
 * contexts) race.  Synchronous on return. */
int svh_predict_noalloc(svh_ctx* ctx, const double* Xtr, const double* resid, int64_t ntr,
                        const double* Xq, int64_t nq, int32_t d, int32_t m, const double* theta,
                        double nu, int32_t precision, int32_t noise_free, double* mean,
                        double* var);

/* Columns of the sparse inverse Cholesky factor for arbitrary blocks
 * (gp.py:332-348, _joint_factor): block b is rows blocks[b*K .. b*K+K-1] of
 * Xall (self last), nugget mask mask[b*K + j] (NULL = all ones); writes
 * a[b*K + j] = (L^-T e_K)_j. */
int svh_block_columns(svh_ctx* ctx, const double* Xall, int64_t nall, int32_t d,
                      const int64_t* blocks, const double* mask, int64_t nblk, int32_t K,
                      const double* theta, double nu, int32_t precision, double* a);

/* Joint prediction solves with the factor of svh_block_columns: `blocks`
 * (ns x K, indices into the observed-first ordering, -1 for none; the last
 * entry of block i is nobs + i) and its values a (ns x K).  Lt = U_pp^T has
 * row i = block i's prediction-side entries, the diagonal last.
 *   mean (ns, may be NULL): Lt^-1 b with b_i = -sum over the observed
 *       entries of a_ik resid[block[i,k]] (resid: nobs, in observed order);
 *   x (ns x nrhs, row-major): Lt^-1 rhs for nrhs right-hand sides;
 *   var (ns, when exact_var): row sums of squares of Lt^-1.
 * Solved level by level on the device (a row's level = 1 + the deepest level
 * among its prediction neighbours).  Replaces the spsolve_triangular calls
 * of gp.py:380-401 (predict_joint / sample_joint / exact_var). */
int svh_joint_solve(svh_ctx* ctx, int64_t nobs, int64_t ns, int32_t K, const int64_t* blocks,
                    const double* a, const double* resid, int32_t nrhs, const double* rhs,
                    double* mean, double* x, int32_t exact_var, double* var);

/* ------------------------------------------------------------ batched --- */
/* Many independent fits at once (no counterpart in the reference, which fits
 * one dataset per ScaledVecchiaGP.fit, gp.py:138-239).  All blocks of all
 * evaluating datasets go through one launch per kernel; a per-dataset
 * Fisher-scoring state machine on the device restates optimize.py:11-73
 * (ridge-until-ascent step, step cap 3, <= 20 strict-increase halvings, stop
 * on |alpha step . g| < tol, reorder pause at k = 2, 4, 8, ...).
 * Limits: d <= 10, q <= 11, m <= 31 (K <= 32).  Parameter arrays use a stride
 * of 12 per dataset, beta a stride of 11, Fisher 144.
 *   setup:     concatenated inputs of ndata datasets already in Vecchia order;
 *              n[i] rows each; nbr/kcnt/chunk are per row (dataset-local
 *              neighbour indices, stride mmax; chunk ids local to the dataset
 *              as _covariance.py:255-259 slices them).  lvt = log variance
 *              target of the sigma^2 penalty; fixed_nug/log_nug hold a fixed
 *              nugget (gp.py:196-208).  X, y, Z, nbr and kcnt may be NULL
 *              when every dataset is then built on the device
 *              (svh_batch_data + svh_batch_rebuild): the buffers are
 *              allocated, neighbour lists empty.
 *   update:    replace dataset i's ordered inputs/neighbours after a rebuild.
 *   fit_start: initial log-parameters, state machines reset; use_graph != 0
 *              captures one step (evaluation + update) in a hipGraph.
 *   fit_step:  replays nrep steps; status[i] = 0 evaluating, 1 searching,
 *              2 reorder due (host decides with svh_batch_set_reorder),
 *              3 done, 4 failed (Cholesky at the start / after a rebuild).
 *   eval:      one objective evaluation of every dataset at theta. */
int svh_batch_setup(svh_ctx* ctx, int32_t ndata, const int32_t* d, const int32_t* q,
                    const int32_t* nu_code, const int64_t* n, const double* X, const double* y,
                    const double* Z, const int64_t* nbr, const int32_t* kcnt,
                    const int32_t* chunk, const int32_t* nchunk, int32_t mmax, const double* lvt,
                    const int32_t* fixed_nug, const double* log_nug, double lambda_max);
int svh_batch_update(svh_ctx* ctx, int32_t i, const double* X, const double* y, const double* Z,
                     const int64_t* nbr, const int32_t* kcnt);

/* The unordered estimation data of every dataset of the batch (Xe, ye, Ze in
 * the layout of svh_batch_setup's X, y, Z), kept on the device for
 * svh_batch_rebuild.  Call after svh_batch_setup. */
int svh_batch_data(svh_ctx* ctx, const double* Xe, const double* ye, const double* Ze);

/* Rebuild of `count` datasets of the batch (indices which[k]) on the device:
 * the data of svh_batch_data scaled by lam[k * 12 + l] (d positive ranges),
 * maximin order from first[k] (first[k] = -1: the row closest to the scaled
 * centroid, ordering.py:72, found on the device with the host's NumPy
 * arithmetic, as for svh_vecchia_rebuild), ordered m[k] nearest neighbours, and the
 * ordered rows, neighbour lists and counts written where the fitting kernels
 * read them -- svh_batch_update without the host round trip.  Replaces the
 * rebuild of gp.py:182-191 for each dataset (d <= 10, m[k] <= the batch's
 * mmax). */
int svh_batch_rebuild(svh_ctx* ctx, int32_t count, const int32_t* which, const double* lam,
                      const int64_t* first, const int32_t* m);
int svh_batch_fit_start(svh_ctx* ctx, const double* theta0, int32_t max_iter, double tol,
                        double var_penalty, int32_t use_graph);
int svh_batch_fit_step(svh_ctx* ctx, int32_t nrep, int32_t* status);
int svh_batch_get(svh_ctx* ctx, double* theta, double* ll, double* beta, int32_t* status,
                  int32_t* iters, int32_t* nevals);
int svh_batch_set_reorder(svh_ctx* ctx, int32_t i, int32_t rebuilt);
int svh_batch_eval(svh_ctx* ctx, const double* theta, double* ll, double* grad, double* fisher,
                   double* beta, int32_t* jitter_rounds);

/* ------------------------------------------------------------ warm start --- */
/* A small fully connected network that maps summary features of a dataset to
 * initial log-parameters for Fisher scoring.  No counterpart in the
 * reference, whose starting point (gp.py:169-177) stays the fallback.
 * One network per context.  x (n_in) -> [Linear -> GELU] x n_layers (width
 * n_hidden) -> Linear (n_out); parameters are one float32 vector, per layer
 * W (fan_in x fan_out, row-major) then b.  Matrix products run as bf16 x bf16
 * with float32 accumulation (bf16 weights); loss = masked mean squared error
 * sum(M (out - T)^2) / sum(M); Adam(W) updates in float32.
 *   init:        allocate for batches up to max_batch rows, parameters zero
 *   set/get:     the flat parameter vector (set also resets the Adam state)
 *   forward:     out (B x n_out) for X (B x n_in)
 *   loss_grad:   loss and gradient of a batch, no update (checking)
 *   load_data:   training rows X (n x n_in), targets T and masks M (n x n_out)
 *   train_epoch: one pass in the order perm (n indices), minibatches of B,
 *                one Adam step each (bias-corrected, step count kept);
 *                *mean_loss = mean minibatch loss */
int svh_mlp_init(svh_ctx* ctx, int32_t n_in, int32_t n_hidden, int32_t n_layers, int32_t n_out,
                 int32_t max_batch);
int svh_mlp_num_params(svh_ctx* ctx, int64_t* n);
int svh_mlp_set_params(svh_ctx* ctx, const float* params);
int svh_mlp_get_params(svh_ctx* ctx, float* params);
int svh_mlp_forward(svh_ctx* ctx, const float* X, int32_t B, float* out);
int svh_mlp_loss_grad(svh_ctx* ctx, const float* X, const float* T, const float* M, int32_t B,
                      float* loss, float* grads);
int svh_mlp_load_data(svh_ctx* ctx, const float* X, const float* T, const float* M, int32_t n);
int svh_mlp_train_epoch(svh_ctx* ctx, const int32_t* perm, int32_t B, float lr, float beta1,
                        float beta2, float eps, float weight_decay, float* mean_loss);

/* Matern GP draws by random Fourier features, for the warm-start training
 * datasets (no counterpart in the reference): for dataset i (n[i] x d[i]
 * rows of X, already divided by the ranges, back to back; d[i] <= 16) with
 * frequencies W (F x d[i]) and phases b (F), stored back to back,
 *   y_j = amp[i] * sum_f cos(W_f . x_j + b_f). */
int svh_simulate_rff(svh_ctx* ctx, int32_t ndata, const int64_t* n, const int32_t* d, int32_t F,
                     const double* X, const double* W, const double* b, const double* amp,
                     double* y);

/* -------------------------------------------------------------- autotune --- */
/* Launch-schedule autotuner (no counterpart in the reference).  Tunable launch
 * sites -- the likelihood pass-1 kernel (split-row kernel at 2..32 workgroups
 * per CU, or the general LDS kernel) and the prediction blocks (packed kernel,
 * or the matrix-core kernel at 1, 2, 4 workgroups per CU) -- key their shape
 * bucket (block size, d, q, precision, log2 of the block count).
 *   mode 0: built-in schedules (default); 1: cached choices only;
 *   2: the first call of a new bucket times every candidate (one warm-up,
 *      median of three, HIP events) and keeps the fastest.
 * cache_path (may be NULL): text file "key index ms" per line, read at first
 * use and rewritten when a bucket is tuned.  Results do not depend on the
 * schedule except for the grouping of the likelihood's partial sums.
 * svh_autotune_read writes the choices so far ("key index ms" lines) into
 * buf; SVH_ERR_ARG if they do not fit in len bytes. */
int svh_autotune(svh_ctx* ctx, int32_t mode, const char* cache_path);
int svh_autotune_read(svh_ctx* ctx, char* buf, int32_t len);

/* ------------------------------------------------------- instrumentation --- */
/* No counterpart in the reference (which has no timers).  When enabled, the
 * library brackets each launch of its dominant kernels with HIP events on the
 * context stream and accumulates their durations under a kernel name
 * ("vecchia_pass1", "predict_blocks", "knn", "maximin"; "maximin_rounds"
 * reads the number of selection rounds of the last maximin call as `count`,
 * "maximin_fallbacks" how many of its rounds had more survivors than the
 * selection holds and accepted only the largest candidate).
 * svh_timing_read returns the accumulated milliseconds and launch count. */
int svh_timing_enable(svh_ctx* ctx, int32_t enable);
int svh_timing_read(svh_ctx* ctx, const char* name, double* total_ms, int64_t* count);
int svh_timing_reset(svh_ctx* ctx);

#ifdef __cplusplus
}
#endif

#endif /* SVH_H */
