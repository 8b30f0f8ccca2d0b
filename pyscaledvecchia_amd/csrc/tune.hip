// Launch-schedule autotuner.  A tunable launch site (likelihood pass 1,
// prediction blocks) offers a list of candidate schedules -- kernel variant,
// grid size, workgroups per CU -- and a key naming its shape bucket (block
// size, d, q, precision, log2 of the block count).  The first time a key is
// seen in tuning mode every candidate is timed on the context stream and the
// fastest is kept in memory and in the cache file; later calls with the same
// key launch the cached choice directly.  No counterpart in the reference.

#include "svh_internal.h"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>

namespace {

void tune_load(svh_ctx* ctx) {
  Autotune& T = ctx->tune;
  T.loaded = true;
  if (T.path.empty()) return;
  std::ifstream f(T.path);
  std::string key;
  int idx;
  double ms;
  while (f >> key >> idx >> ms) {
    T.best[key] = idx;
    T.best_ms[key] = ms;
  }
}

void tune_save(svh_ctx* ctx) {
  Autotune& T = ctx->tune;
  if (T.path.empty()) return;
  std::ofstream f(T.path, std::ios::trunc);
  for (const auto& kv : T.best) f << kv.first << " " << kv.second << " " << T.best_ms[kv.first] << "\n";
}

}  // namespace

int tune_choose(svh_ctx* ctx, const std::string& key, int ncand, int dflt,
                void (*launch)(void*, int), void* user) {
  Autotune& T = ctx->tune;
  if (T.mode == 0 || ncand <= 1) return dflt;
  if (!T.loaded) tune_load(ctx);
  auto it = T.best.find(key);
  if (it != T.best.end() && it->second >= 0 && it->second < ncand) return it->second;
  if (T.mode == 1) return dflt;
  hipStream_t s = ctx->stream;
  hipEvent_t e0, e1;
  SVH_HIP(hipEventCreate(&e0));
  SVH_HIP(hipEventCreate(&e1));
  int best = dflt;
  double best_ms = 1e300;
  for (int i = 0; i < ncand; ++i) {
    launch(user, i);   // warm-up (code load, caches)
    SVH_HIP(hipGetLastError());
    double t[3];
    for (int r = 0; r < 3; ++r) {
      SVH_HIP(hipEventRecord(e0, s));
      launch(user, i);
      SVH_HIP(hipEventRecord(e1, s));
      SVH_HIP(hipEventSynchronize(e1));
      float ms = 0.f;
      SVH_HIP(hipEventElapsedTime(&ms, e0, e1));
      t[r] = ms;
    }
    std::sort(t, t + 3);
    if (std::getenv("SVH_TUNE_LOG"))
      std::fprintf(stderr, "[svh tune] %s candidate %d: %.4f ms\n", key.c_str(), i, t[1]);
    if (t[1] < best_ms) {
      best_ms = t[1];
      best = i;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  T.best[key] = best;
  T.best_ms[key] = best_ms;
  tune_save(ctx);
  return best;
}

extern "C" {

int svh_autotune(svh_ctx* ctx, int32_t mode, const char* cache_path) {
  if (!ctx || mode < 0 || mode > 2) return SVH_ERR_ARG;
  ctx->tune.mode = mode;
  ctx->tune.path = cache_path ? cache_path : "";
  ctx->tune.loaded = false;
  ctx->tune.best.clear();
  ctx->tune.best_ms.clear();
  return SVH_OK;
}

int svh_autotune_read(svh_ctx* ctx, char* buf, int32_t len) {
  if (!ctx || !buf || len < 1) return SVH_ERR_ARG;
  std::ostringstream o;
  for (const auto& kv : ctx->tune.best)
    o << kv.first << " " << kv.second << " " << ctx->tune.best_ms[kv.first] << "\n";
  const std::string str = o.str();
  const int n = std::min<int>((int)str.size(), len - 1);
  std::memcpy(buf, str.data(), n);
  buf[n] = '\0';
  return (int)str.size() < len ? SVH_OK : SVH_ERR_ARG;
}

}  // extern "C"
