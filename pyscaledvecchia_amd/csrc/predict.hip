// (c) Prediction blocks on gfx950 (restates the block loops of
// gp.py:242-277 (_predict_scaled) and gp.py:332-348 (_joint_factor) with
// _covariance.py:86-108, 220-252).
//
// One wave per block [neighbours..., target].  The covariance lower triangle
// is held packed in LDS (row i at i(i+1)/2), factorised right-looking with the
// chunk jitter level, and a = L^-T e_K is formed by a column sweep.  Mode 0
// writes a (a column of the sparse inverse Cholesky factor); mode 1 writes
// mean = -a[0:K-1].resid[nb] / a[K-1] and var = 1 / a[K-1]^2.

#include "svh_internal.h"

#include <cmath>
#include <vector>

namespace {

constexpr int kWave = 64;

__constant__ double kJitDevP[10] = {0.0,  1e-12, 1e-11, 1e-10, 1e-09,
                                    1e-08, 1e-07, 1e-06, 1e-05, 1e-04};

// p -> (i, j) with i >= j in row-major lower-triangle order
__device__ __forceinline__ void tri_pair(int p, int& i, int& j) {
  int r = (int)((sqrtf(8.0f * (float)p + 1.0f) - 1.0f) * 0.5f);
  int base = r * (r + 1) / 2;
  while (base > p) {
    --r;
    base = r * (r + 1) / 2;
  }
  while (base + r + 1 <= p) {
    ++r;
    base = r * (r + 1) / 2;
  }
  i = r;
  j = p - base;
}

__device__ __forceinline__ int tri(int i) { return i * (i + 1) / 2; }

struct PredArgs {
  const double* Xn;      // neighbour coordinates (rows of length d)
  const int64_t* nb;     // neighbour indices, nb[b*nb_stride + t], t < K-1
  const double* Xt;      // target coordinates
  const int64_t* tg;     // target index tg[b*tg_stride] (NULL: b)
  const double* mask;    // nugget mask (b*K + t), NULL: ones (target: !noise_free)
  const double* resid;   // mode 1: residual at neighbour index
  const double* inv_lam;
  const int* level;
  int* fail;
  double* out_a;
  double* mean;
  double* varo;
  int64_t nblk;
  int nb_stride, tg_stride, noise_free, chunk_step, d, K, nu_code, mode;
  double var, nug;
};

__device__ __forceinline__ double matern_M(int nu_code, double r) {
  if (nu_code == 0) return exp(-r);
  if (nu_code == 1) {
    double c = 1.7320508075688772 * r;
    return (1.0 + c) * exp(-c);
  }
  double c = 2.23606797749979 * r;
  return (1.0 + c + c * c / 3.0) * exp(-c);
}

__global__ __launch_bounds__(kWave) void pred_blocks_kernel(PredArgs a) {
  extern __shared__ double lds[];
  const int K = a.K, d = a.d, lane = threadIdx.x;
  double* Lp = lds;                 // K(K+1)/2
  double* xs = Lp + tri(K);         // K*d
  double* av = xs + K * d;          // K
  double* msk = av + K;             // K
  for (int64_t b = blockIdx.x; b < a.nblk; b += gridDim.x) {
    const int ch = (int)(b / a.chunk_step);
    const int lev = a.level[ch];
    __syncthreads();
    for (int e = lane; e < K * d; e += kWave) {
      int t = e / d, l = e - t * d;
      const double* src;
      if (t < K - 1) {
        src = a.Xn + a.nb[b * a.nb_stride + t] * d;
      } else {
        int64_t ti = a.tg ? a.tg[b * a.tg_stride] : b;
        src = a.Xt + ti * d;
      }
      xs[e] = src[l] * a.inv_lam[l];
    }
    for (int t = lane; t < K; t += kWave)
      msk[t] = a.mask ? a.mask[b * K + t] : ((t == K - 1 && a.noise_free) ? 0.0 : 1.0);
    __syncthreads();
    const int T = tri(K);
    for (int p = lane; p < T; p += kWave) {
      int i, j;
      tri_pair(p, i, j);
      double s = 0.0;
      for (int l = 0; l < d; ++l) {
        double t = xs[i * d + l] - xs[j * d + l];
        s += t * t;
      }
      double M = (i == j) ? 1.0 : matern_M(a.nu_code, sqrt(s));
      double nug = (i == j) ? a.nug * msk[i] : 0.0;
      Lp[p] = a.var * (M + nug);
    }
    __syncthreads();
    if (lev > 0) {
      double tr = 0.0;
      for (int t = 0; t < K; ++t) tr += Lp[tri(t) + t];
      double jit = kJitDevP[lev] * (tr / K);
      __syncthreads();
      for (int t = lane; t < K; t += kWave) Lp[tri(t) + t] += jit;
      __syncthreads();
    }
    bool failed = false;
    for (int k = 0; k < K; ++k) {
      double dkk = Lp[tri(k) + k];
      if (!(dkk > 0.0)) {
        failed = true;
        break;
      }
      double lkk = sqrt(dkk);
      double inv = 1.0 / lkk;
      for (int t = k + 1 + lane; t < K; t += kWave) Lp[tri(t) + k] *= inv;
      __syncthreads();
      if (lane == 0) Lp[tri(k) + k] = lkk;
      const int nt = (K - 1 - k) * (K - k) / 2;
      for (int p = lane; p < nt; p += kWave) {
        int i, j;
        tri_pair(p, i, j);
        i += k + 1;
        j += k + 1;
        Lp[tri(i) + j] -= Lp[tri(i) + k] * Lp[tri(j) + k];
      }
      __syncthreads();
    }
    if (failed) {
      if (lane == 0) atomicOr(a.fail + ch, 1);
      continue;
    }
    for (int t = lane; t < K; t += kWave) av[t] = (t == K - 1) ? 1.0 : 0.0;
    __syncthreads();
    for (int j = K - 1; j >= 0; --j) {
      double aj = av[j] / Lp[tri(j) + j];
      __syncthreads();
      if (lane == 0) av[j] = aj;
      for (int t = lane; t < j; t += kWave) av[t] -= Lp[tri(j) + t] * aj;
      __syncthreads();
    }
    if (a.mode == 0) {
      for (int t = lane; t < K; t += kWave) a.out_a[b * K + t] = av[t];
    } else {
      double s = 0.0;
      for (int t = lane; t < K - 1; t += kWave) s += av[t] * a.resid[a.nb[b * a.nb_stride + t]];
      for (int off = 32; off >= 1; off >>= 1) s += __shfl_xor(s, off, kWave);
      if (lane == 0) {
        double aK = av[K - 1];
        a.mean[b] = -s / aK;
        a.varo[b] = 1.0 / (aK * aK);
      }
    }
  }
}

__global__ void p_fill_int(int* p, int v, int64_t n) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void p_escalate(int* __restrict__ level, int* __restrict__ fail, int nchunk,
                           int* __restrict__ nfail) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < nchunk) {
    if (fail[c]) {
      level[c] += 1;
      atomicAdd(nfail, 1);
    }
    fail[c] = 0;
  }
}

}  // namespace

// Runs the block kernel with the chunk jitter ladder (_covariance.py:241-252
// over _chunks(nblk, K, 1)).  Host pointers are not used here: all arrays are
// device arrays.
void predict_blocks_device(svh_ctx* ctx, const double* dXn, const int64_t* dnb, int nb_stride,
                           const double* dXt, const int64_t* dtg, int tg_stride,
                           const double* dmask, int noise_free, int64_t nblk, int K, int d,
                           const double* theta, double nu, const double* dresid, int mode,
                           double* dout_a, double* dmean, double* dvar, int* jitter_power_max) {
  SVH_REQUIRE(K >= 1, "predict: empty block");
  int nu_code;
  if (nu == 0.5)
    nu_code = 0;
  else if (nu == 1.5)
    nu_code = 1;
  else if (nu == 2.5)
    nu_code = 2;
  else
    throw SvhError{SVH_ERR_ARG, "predict: nu must be 0.5, 1.5 or 2.5"};
  if (nblk == 0) return;
  hipStream_t s = ctx->stream;
  const size_t lds = (size_t)(K * (K + 1) / 2 + K * d + 2 * K) * sizeof(double);
  SVH_REQUIRE(lds <= 160 * 1024, "predict: block too large for LDS");
  double hl[32];
  for (int l = 0; l < d; ++l) hl[l] = 1.0 / std::exp(theta[1 + l]);
  double* inv_lam = ctx->arena.get_t<double>("p.invlam", 32);
  SVH_HIP(hipMemcpyAsync(inv_lam, hl, d * sizeof(double), hipMemcpyHostToDevice, s));
  const int64_t step64 = std::max<int64_t>(1, 4000000 / ((int64_t)K * K));
  const int step = (int)std::min<int64_t>(step64, 1 << 30);
  const int nchunk = (int)((nblk + step - 1) / step);
  int* level = ctx->arena.get_t<int>("p.level", nchunk);
  int* fail = ctx->arena.get_t<int>("p.fail", nchunk);
  int* nfail = ctx->arena.get_t<int>("p.nfail", 1);
  hipLaunchKernelGGL(p_fill_int, dim3((nchunk + 255) / 256), dim3(256), 0, s, level, 0,
                     (int64_t)nchunk);
  hipLaunchKernelGGL(p_fill_int, dim3((nchunk + 255) / 256), dim3(256), 0, s, fail, 0,
                     (int64_t)nchunk);
  PredArgs a;
  a.Xn = dXn;
  a.nb = dnb;
  a.Xt = dXt;
  a.tg = dtg;
  a.mask = dmask;
  a.resid = dresid;
  a.inv_lam = inv_lam;
  a.level = level;
  a.fail = fail;
  a.out_a = dout_a;
  a.mean = dmean;
  a.varo = dvar;
  a.nblk = nblk;
  a.nb_stride = nb_stride;
  a.tg_stride = tg_stride;
  a.noise_free = noise_free;
  a.chunk_step = step;
  a.d = d;
  a.K = K;
  a.nu_code = nu_code;
  a.mode = mode;
  a.var = std::exp(theta[0]);
  a.nug = std::exp(theta[1 + d]);
  const int G = (int)std::min<int64_t>(nblk, 8192);
  int* hn = static_cast<int*>(ctx->stage.get(sizeof(int)));
  int rounds = 0;
  for (;;) {
    ctx->timers.begin(s, "predict_blocks");
    hipLaunchKernelGGL(pred_blocks_kernel, dim3(G), dim3(kWave), lds, s, a);
    ctx->timers.end(s);
    hipLaunchKernelGGL(p_fill_int, dim3(1), dim3(1), 0, s, nfail, 0, (int64_t)1);
    hipLaunchKernelGGL(p_escalate, dim3((nchunk + 255) / 256), dim3(256), 0, s, level, fail,
                       nchunk, nfail);
    SVH_HIP(hipGetLastError());
    SVH_HIP(hipMemcpyAsync(hn, nfail, sizeof(int), hipMemcpyDeviceToHost, s));
    SVH_HIP(hipStreamSynchronize(s));
    ctx->timers.collect();
    if (*hn == 0) break;
    if (rounds == 9) throw SvhError{SVH_ERR_CHOLESKY, "block Cholesky failed even with jitter"};
    ++rounds;
  }
  if (jitter_power_max) *jitter_power_max = rounds > 0 ? -13 + rounds : 0;
}
