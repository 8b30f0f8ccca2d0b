// Internal declarations shared by the translation units of libsvh.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/svh.h"

// ---------------------------------------------------------------- errors
struct SvhError {
  int code;
  std::string msg;
};

#define SVH_HIP(call)                                                              \
  do {                                                                             \
    hipError_t e_ = (call);                                                        \
    if (e_ != hipSuccess)                                                          \
      throw SvhError{SVH_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)}; \
  } while (0)

#define SVH_REQUIRE(cond, text)                                  \
  do {                                                           \
    if (!(cond)) throw SvhError{SVH_ERR_ARG, std::string(text)}; \
  } while (0)

// ---------------------------------------------------------------- buffers
// Named device buffers that only grow: repeated calls of one size allocate
// once.  Pointers stay valid until the same name is requested larger.
//
// Fixed-region mode (the entry point that allocates nothing): while a region
// is set, every request is carved from it in call order (256-byte aligned,
// nothing cached, no hipMalloc); a request that does not fit throws
// SVH_ERR_ARG.
struct DevArena {
  struct Buf {
    void* p = nullptr;
    size_t bytes = 0;
  };
  std::unordered_map<std::string, Buf> bufs;
  char* fix_base = nullptr;
  size_t fix_bytes = 0, fix_used = 0;
  void* get(const std::string& name, size_t bytes) {
    if (fix_base) {
      const size_t at = (fix_used + 255) & ~size_t(255);
      if (bytes > fix_bytes || at > fix_bytes - bytes)
        throw SvhError{SVH_ERR_ARG, "working memory of the call exceeds the fixed device "
                                    "array (request '" + name + "')"};
      fix_used = at + bytes;
      return fix_base + at;
    }
    Buf& b = bufs[name];
    if (bytes == 0) bytes = 1;
    if (b.bytes < bytes) {
      if (b.p) SVH_HIP(hipFree(b.p));
      b.p = nullptr;
      SVH_HIP(hipMalloc(&b.p, bytes));
      b.bytes = bytes;
    }
    return b.p;
  }
  template <class T>
  T* get_t(const std::string& name, size_t count) {
    return static_cast<T*>(get(name, count * sizeof(T)));
  }
  void release() {
    for (auto& kv : bufs)
      if (kv.second.p) (void)hipFree(kv.second.p);
    bufs.clear();
  }
};

// Pinned host staging for small results.  In fixed mode (no allocation) a
// small pageable array of the context serves instead.
struct HostStage {
  void* p = nullptr;
  size_t bytes = 0;
  bool fixed = false;
  alignas(64) char fixed_buf[256];
  void* get(size_t want) {
    if (fixed) {
      if (want > sizeof fixed_buf) throw SvhError{SVH_ERR_ARG, "host staging too small"};
      return fixed_buf;
    }
    if (bytes < want) {
      if (p) SVH_HIP(hipHostFree(p));
      p = nullptr;
      SVH_HIP(hipHostMalloc(&p, want, hipHostMallocDefault));
      bytes = want;
    }
    return p;
  }
  void release() {
    if (p) (void)hipHostFree(p);
    p = nullptr;
    bytes = 0;
  }
};

// State uploaded by svh_vecchia_setup.
struct VecchiaState {
  bool ready = false;
  int64_t n = 0, nblk = 0;
  int d = 0, q = 0, mmax = 0, kp = 0;
  int nchunk_grad = 0, nchunk_val = 0, nchunk_grad_nu = 0;
  int zmode = 0;   // trend columns formed in the kernels: 0 none, 1 [1, x], 2 constant
  // device pointers live in the arena under "v.*" names
};

// Kernel timers: HIP events on the context stream around named launches.
struct KernelTimers {
  bool enabled = false;
  struct Acc {
    double ms = 0.0;
    int64_t count = 0;
  };
  std::unordered_map<std::string, Acc> acc;
  struct Pending {
    std::string name;
    hipEvent_t a, b;
  };
  std::vector<Pending> pending;
  std::vector<hipEvent_t> pool;
  hipEvent_t take() {
    if (!pool.empty()) {
      hipEvent_t e = pool.back();
      pool.pop_back();
      return e;
    }
    hipEvent_t e;
    SVH_HIP(hipEventCreate(&e));
    return e;
  }
  // returns the start event to record before the launch (nullptr if disabled)
  void begin(hipStream_t s, const char* name) {
    if (!enabled) return;
    Pending p{name, take(), take()};
    SVH_HIP(hipEventRecord(p.a, s));
    pending.push_back(p);
  }
  // a count without a clock (which of several kernels served a call)
  void count(const char* name) {
    if (enabled) acc[name].count += 1;
  }
  void end(hipStream_t s) {
    if (!enabled || pending.empty()) return;
    SVH_HIP(hipEventRecord(pending.back().b, s));
  }
  // call after a stream synchronisation
  void collect() {
    for (auto& p : pending) {
      float ms = 0.f;
      SVH_HIP(hipEventElapsedTime(&ms, p.a, p.b));
      Acc& a = acc[p.name];
      a.ms += ms;
      a.count += 1;
      pool.push_back(p.a);
      pool.push_back(p.b);
    }
    pending.clear();
  }
  void release() {
    for (auto& p : pending) {
      (void)hipEventDestroy(p.a);
      (void)hipEventDestroy(p.b);
    }
    for (auto e : pool) (void)hipEventDestroy(e);
    pending.clear();
    pool.clear();
  }
};

// Launch-schedule autotuner (tune.hip): per shape-bucket key, the index of the
// fastest of a launch site's candidates, timed on the context stream.
struct Autotune {
  int mode = 0;        // 0: built-in defaults, 1: cached choices only, 2: also tune new keys
  std::string path;    // cache file, "key index ms" lines ("" = in memory only)
  bool loaded = false;
  std::unordered_map<std::string, int> best;
  std::unordered_map<std::string, double> best_ms;
};

struct MlpState;   // warm-start network (mlp.hip)

struct svh_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  int num_cu = 0;
  DevArena arena;
  HostStage stage;
  VecchiaState vs;
  KernelTimers timers;
  int mm_rounds = 0;  // rounds of the last maximin call
  int mm_fallbacks = 0;  // its rounds that kept only the largest candidate
  int64_t rb_n = 0;   // sizes of the estimation data resident for rebuilds
  int rb_d = 0, rb_q = 0, rb_zmode = 0;
  bool batch_data = false;   // unordered batch data uploaded for the current batch
  // the evaluation's chunk levels and failure flags ("v.levfail"): the buffer
  // and length last left all zero (an evaluation without jitter escalation)
  const void* levfail_zero = nullptr;
  int64_t levfail_zero_n = 0;
  MlpState* mlp = nullptr;
  Autotune tune;
  void* scratch = nullptr;   // this device's fixed working array (api.hip), or nullptr
  std::string last_error;
};

// ---------------------------------------------------------------- launchers
// ordering.hip
void launch_maximin(svh_ctx* ctx, const double* dXT, int64_t n, int d, int64_t first,
                    int64_t* dorder);
void launch_ordered_knn(svh_ctx* ctx, const double* dXT, int64_t n, int d, int m,
                        int64_t start, int64_t* dnn, int64_t row0 = 0);
void launch_knn_query(svh_ctx* ctx, const double* dXrefT, int64_t nref, const double* dXqT,
                      int64_t nq, int d, int k, int64_t* dnb, int nb_stride);
void launch_to_soa(svh_ctx* ctx, const double* dX, int64_t n, int d, const double* dlam,
                   double* dXT);
// XTo[l][i] = XT[l][order[i]] (SoA, n x d)
void launch_gather_soa(svh_ctx* ctx, const double* dXT, const int64_t* dorder, int64_t n, int d,
                       double* dXTo);
// rebuilds of selected datasets of the batch on the device
void batch_rebuild_device(svh_ctx* ctx, int count, const int32_t* which, const double* lam,
                          const int64_t* first, const int32_t* m);
void batch_data_device(svh_ctx* ctx, const double* Xe, const double* ye, const double* Ze);
// offsets of one dataset in the batch buffers
struct BatchLayout {
  int64_t xoff, yoff, zoff, blk0, n;
  int d, q, mmax;
  int64_t nx, ny, nz;
};
BatchLayout batch_layout(svh_ctx* ctx, int i);
// loglik.hip
void vecchia_rebuild_device(svh_ctx* ctx, const double* Xe, const double* ye, const double* Ze,
                            int64_t n, int d, int q, const double* lam, int64_t first, int m,
                            int upload_data, int64_t* order);
void vecchia_eval_device(svh_ctx* ctx, const double* theta, int P, double nu, int precision,
                         int need_grad, double var_penalty, double log_var_target,
                         double* ll, double* grad, double* fisher, double* beta,
                         int* jitter_power_max);
// joint.hip
void joint_solve_device(svh_ctx* ctx, int64_t nobs, int64_t ns, int K, const int64_t* block,
                        const double* av, const double* resid, int nrhs, const double* rhs,
                        double* mean, double* x, int exact_var, double* var);
// mlp.hip
void mlp_release(svh_ctx* ctx);
// tune.hip: the candidate to launch at a tunable site.  Returns the cached
// choice for `key`, or (mode 2) times launch(i) for every candidate (one
// warm-up, median of three, HIP events on the context stream), caches and
// returns the fastest; `dflt` when tuning is off.  launch(i) must be
// repeatable (it may run several times before the real launch).
int tune_choose(svh_ctx* ctx, const std::string& key, int ncand, int dflt,
                void (*launch)(void*, int), void* user);
// Jitter factors 10^p, p = -12..-4 (index 1..9; index 0 = no jitter).
extern const double kJitterPow10[10];
