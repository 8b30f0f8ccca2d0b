// Learned warm start on gfx950: a small fully connected network that maps
// summary features of a dataset to initial log-parameters for Fisher scoring.
// It has no counterpart in the reference, whose starting point
// (gp.py:169-177) stays the fallback.
//
// Network: x (n_in) -> [Linear -> GELU] x n_layers (width n_hidden) ->
// Linear (n_out); GELU is 0.5 z (1 + erf(z / sqrt 2)).  Parameters are one
// flat float32 vector, per layer W (fan_in x fan_out, row-major) then b.
// Every matrix product runs on the matrix cores as bf16 x bf16 with float32
// accumulation (v_mfma_f32_16x16x16_bf16): operands are rounded to bfloat16
// (nearest even) as they are loaded, so the weights act as bf16 weights; the
// biases, activations in memory, gradients and the Adam state are float32.
//
// Training step (all on the device, no host round trip per step):
//   gather the minibatch rows, forward (GEMM + bias + GELU epilogue keeping
//   the pre-activations), masked MSE loss and its gradient, backward (GEMMs
//   for the weight gradients and the activation gradients with the GELU
//   derivative in the epilogue, column sums for the bias gradients), Adam.
//
// Also here: a batched random-Fourier-feature simulator of Matern GP draws
// (training datasets for the network).

#include "svh_internal.h"

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

namespace {

constexpr int kMW = 64;
typedef short mlp_s4 __attribute__((ext_vector_type(4)));
typedef float mlp_f4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ short to_bf16(float f) {
  unsigned u = __float_as_uint(f);
  u = (u + 0x7fffu + ((u >> 16) & 1u)) >> 16;
  return (short)u;
}

__device__ __forceinline__ float gelu_f(float z) {
  return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f));
}

__device__ __forceinline__ float gelu_grad_f(float z) {
  return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) +
         z * expf(-0.5f * z * z) * 0.39894228040143268f;
}

// C (M x N) = A (M x K) B (K x N) with A(m,k) = A[m*sam + k*sak],
// B(k,n) = B[k*sbk + n*sbn]; epilogue:
//   0: C = acc (+ bias[n])
//   1: C = z = acc + bias[n] (pre-activation), H = gelu(z)
//   2: C = acc * gelu'(Zin)   (activation gradient -> pre-activation gradient)
struct GemmArgs {
  const float* A;
  const float* B;
  float* C;
  float* H;
  const float* Zin;
  const float* bias;
  long long sam, sak, sbk, sbn;
  int M, N, K, ldc, epi;
};

// One wave per 16 x 16 output tile, four tiles along N per workgroup.
// Operand layout of v_mfma_f32_16x16x16_bf16: lane L holds A[L & 15][4(L >> 4) + j]
// and B[4(L >> 4) + j][L & 15] for j = 0..3; the result D[4(L >> 4) + v][L & 15].
__global__ __launch_bounds__(256) void mlp_gemm(GemmArgs g) {
  const int lane = threadIdx.x & (kMW - 1), w = threadIdx.x / kMW;
  const int m0 = blockIdx.y * 16, n0 = (blockIdx.x * 4 + w) * 16;
  if (n0 >= g.N) return;   // whole wave; the kernel has no barrier
  const int li = lane & 15, lg = lane >> 4;
  const int am = m0 + li, bn = n0 + li;
  mlp_f4 acc = {0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < g.K; k0 += 16) {
    mlp_s4 av, bv;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int k = k0 + 4 * lg + j;
      const float a = (am < g.M && k < g.K) ? g.A[am * g.sam + k * g.sak] : 0.f;
      const float b = (bn < g.N && k < g.K) ? g.B[k * g.sbk + bn * g.sbn] : 0.f;
      av[j] = to_bf16(a);
      bv[j] = to_bf16(b);
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc, 0, 0, 0);
  }
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int m = m0 + 4 * lg + v, n = n0 + li;
    if (m < g.M && n < g.N) {
      const long long o = (long long)m * g.ldc + n;
      float z = acc[v];
      if (g.epi == 0) {
        if (g.bias) z += g.bias[n];
        g.C[o] = z;
      } else if (g.epi == 1) {
        z += g.bias[n];
        g.C[o] = z;
        g.H[o] = gelu_f(z);
      } else {
        g.C[o] = z * gelu_grad_f(g.Zin[o]);
      }
    }
  }
}

// Masked MSE over B x n rows: loss = sum(M (out - T)^2) / sum(M);
// dout = 2 M (out - T) / sum(M).  One workgroup.
__global__ __launch_bounds__(256) void mlp_loss(const float* __restrict__ out,
                                                const float* __restrict__ T,
                                                const float* __restrict__ Mk, int total,
                                                float* __restrict__ dout,
                                                float* __restrict__ loss_acc) {
  __shared__ float red[256];
  float ms = 0.f;
  for (int e = threadIdx.x; e < total; e += 256) ms += Mk[e];
  red[threadIdx.x] = ms;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  const float msum = fmaxf(red[0], 1.0f);
  __syncthreads();
  float ls = 0.f;
  for (int e = threadIdx.x; e < total; e += 256) {
    const float r = Mk[e] * (out[e] - T[e]);
    ls += r * (out[e] - T[e]);
    dout[e] = 2.0f * r / msum;
  }
  red[threadIdx.x] = ls;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) loss_acc[0] += red[0] / msum;
}

// db[n] = sum_m dZ[m][n]
__global__ void mlp_colsum(const float* __restrict__ dZ, int M, int N, float* __restrict__ db) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  float s = 0.f;
  for (int m = 0; m < M; ++m) s += dZ[(long long)m * N + n];
  db[n] = s;
}

__global__ void mlp_adam(float* __restrict__ p, const float* __restrict__ g,
                         float* __restrict__ m, float* __restrict__ v, long long n, float lr,
                         float b1, float b2, float eps, float wd, float c1, float c2) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * ((mi / c1) / (sqrtf(vi / c2) + eps) + wd * p[i]);
  }
}

// rows of the minibatch: Xb[i] = X[idx[i]] (and the targets, masks)
__global__ void mlp_gather(const float* __restrict__ X, const int* __restrict__ idx, int B,
                           int w, float* __restrict__ Xb) {
  const long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (e >= (long long)B * w) return;
  const int i = (int)(e / w), c = (int)(e - (long long)i * w);
  Xb[e] = X[(long long)idx[i] * w + c];
}

__global__ void mlp_zero(float* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.f;
}

// y_i = amp * sum_f cos(w_f . x_i + b_f) for the rows of each dataset
// (random Fourier features; coordinates are already divided by the ranges).
struct RffDesc {
  long long xoff, yoff;   // rows start at X + xoff (n x d), y + yoff
  long long woff, boff;   // frequencies W + woff (F x d), phases b + boff (F)
  int n, d, F, pad;
  double amp;
};

__global__ __launch_bounds__(256) void rff_kernel(const RffDesc* __restrict__ desc,
                                                  const double* __restrict__ X,
                                                  const double* __restrict__ W,
                                                  const double* __restrict__ bph,
                                                  double* __restrict__ y) {
  const RffDesc D = desc[blockIdx.y];
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= D.n) return;
  double x[16];
#pragma unroll
  for (int l = 0; l < 16; ++l) x[l] = (l < D.d) ? X[D.xoff + (long long)i * D.d + l] : 0.0;
  double s = 0.0;
  for (int f = 0; f < D.F; ++f) {
    const double* wf = W + D.woff + (long long)f * D.d;
    double t = bph[D.boff + f];
#pragma unroll
    for (int l = 0; l < 16; ++l)
      if (l < D.d) t = fma(wf[l], x[l], t);
    s += cos(t);
  }
  y[D.yoff + i] = D.amp * s;
}

}  // namespace

// ---------------------------------------------------------------- state
struct MlpState {
  int nin = 0, nh = 0, nl = 0, nout = 0, np = 0, bmax = 0;
  long long step = 0;
  int ndata = 0;
  std::vector<int> woff, boff, fin, fout;   // per layer offsets and fans
};

static MlpState& mlp_of(svh_ctx* ctx) {
  if (!ctx->mlp) ctx->mlp = new MlpState();
  return *ctx->mlp;
}

void mlp_release(svh_ctx* ctx) {
  delete ctx->mlp;
  ctx->mlp = nullptr;
}

namespace {

template <class F>
int mguard(svh_ctx* ctx, F&& f) {
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

void gemm(svh_ctx* ctx, const float* A, long long sam, long long sak, const float* B,
          long long sbk, long long sbn, float* C, int ldc, int M, int N, int K, int epi,
          const float* bias, float* H, const float* Zin) {
  GemmArgs g{A, B, C, H, Zin, bias, sam, sak, sbk, sbn, M, N, K, ldc, epi};
  dim3 grid((N + 63) / 64, (M + 15) / 16);
  hipLaunchKernelGGL(mlp_gemm, grid, dim3(256), 0, ctx->stream, g);
}

// activation buffers for a batch of B rows
struct Acts {
  std::vector<float*> Z, Hh;   // per layer: pre-activation, activation (hidden layers)
  float* out;
  float* dA;   // gradient buffers (two, ping-pong)
  float* dB;
};

Acts acts_of(svh_ctx* ctx, const MlpState& S, int B) {
  Acts a;
  const int L = S.nl + 1;
  for (int l = 0; l < L; ++l) {
    const size_t cnt = (size_t)B * S.fout[l];
    a.Z.push_back(ctx->arena.get_t<float>("mlp.Z" + std::to_string(l), cnt));
    a.Hh.push_back(l < L - 1 ? ctx->arena.get_t<float>("mlp.H" + std::to_string(l), cnt) : nullptr);
  }
  a.out = a.Z[L - 1];
  const int wmax = std::max(S.nh, std::max(S.nin, S.nout));
  a.dA = ctx->arena.get_t<float>("mlp.dA", (size_t)B * wmax);
  a.dB = ctx->arena.get_t<float>("mlp.dB", (size_t)B * wmax);
  return a;
}

void forward(svh_ctx* ctx, const MlpState& S, const float* params, const float* X, int B,
             const Acts& a) {
  const int L = S.nl + 1;
  const float* h = X;
  for (int l = 0; l < L; ++l) {
    const float* W = params + S.woff[l];
    const float* b = params + S.boff[l];
    const int fi = S.fin[l], fo = S.fout[l];
    if (l < L - 1) {
      gemm(ctx, h, fi, 1, W, fo, 1, a.Z[l], fo, B, fo, fi, 1, b, a.Hh[l], nullptr);
      h = a.Hh[l];
    } else {
      gemm(ctx, h, fi, 1, W, fo, 1, a.Z[l], fo, B, fo, fi, 0, b, nullptr, nullptr);
    }
  }
}

// gradient of the masked MSE into grads; loss added to *dloss (device)
void backward(svh_ctx* ctx, const MlpState& S, const float* params, const float* X,
              const float* T, const float* Mk, int B, const Acts& a, float* grads,
              float* dloss) {
  const int L = S.nl + 1;
  float* dz = a.dA;
  float* other = a.dB;
  hipLaunchKernelGGL(mlp_loss, dim3(1), dim3(256), 0, ctx->stream, a.out, T, Mk, B * S.nout, dz,
                     dloss);
  for (int l = L - 1; l >= 0; --l) {
    const int fi = S.fin[l], fo = S.fout[l];
    const float* hin = (l == 0) ? X : a.Hh[l - 1];
    // dW = hin^T dz   (fi x fo)
    gemm(ctx, hin, 1, fi, dz, fo, 1, grads + S.woff[l], fo, fi, fo, B, 0, nullptr, nullptr,
         nullptr);
    hipLaunchKernelGGL(mlp_colsum, dim3((fo + 255) / 256), dim3(256), 0, ctx->stream, dz, B, fo,
                       grads + S.boff[l]);
    if (l > 0) {
      // dz_prev = (dz W^T) * gelu'(Z_prev)   (B x fi)
      gemm(ctx, dz, fo, 1, params + S.woff[l], 1, fo, other, fi, B, fi, fo, 2, nullptr, nullptr,
           a.Z[l - 1]);
      std::swap(dz, other);
    }
  }
}

}  // namespace

extern "C" {

int svh_mlp_init(svh_ctx* ctx, int32_t n_in, int32_t n_hidden, int32_t n_layers, int32_t n_out,
                 int32_t max_batch) {
  return mguard(ctx, [&] {
    SVH_REQUIRE(n_in >= 1 && n_hidden >= 1 && n_layers >= 1 && n_out >= 1 && max_batch >= 1,
                "mlp_init: bad sizes");
    MlpState& S = mlp_of(ctx);
    S = MlpState();
    S.nin = n_in;
    S.nh = n_hidden;
    S.nl = n_layers;
    S.nout = n_out;
    S.bmax = max_batch;
    int o = 0;
    for (int l = 0; l <= n_layers; ++l) {
      const int fi = (l == 0) ? n_in : n_hidden, fo = (l == n_layers) ? n_out : n_hidden;
      S.fin.push_back(fi);
      S.fout.push_back(fo);
      S.woff.push_back(o);
      o += fi * fo;
      S.boff.push_back(o);
      o += fo;
    }
    S.np = o;
    float* z[3] = {ctx->arena.get_t<float>("mlp.params", S.np),
                   ctx->arena.get_t<float>("mlp.m", S.np), ctx->arena.get_t<float>("mlp.v", S.np)};
    for (float* p : z)
      hipLaunchKernelGGL(mlp_zero, dim3((S.np + 255) / 256), dim3(256), 0, ctx->stream, p, S.np);
    SVH_HIP(hipStreamSynchronize(ctx->stream));
  });
}

int svh_mlp_num_params(svh_ctx* ctx, int64_t* n) {
  return mguard(ctx, [&] {
    SVH_REQUIRE(ctx->mlp && n, "mlp: not initialised");
    *n = ctx->mlp->np;
  });
}

int svh_mlp_set_params(svh_ctx* ctx, const float* params) {
  return mguard(ctx, [&] {
    SVH_REQUIRE(ctx->mlp && params, "mlp: not initialised");
    MlpState& S = *ctx->mlp;
    float* p = ctx->arena.get_t<float>("mlp.params", S.np);
    SVH_HIP(hipMemcpyAsync(p, params, (size_t)S.np * 4, hipMemcpyHostToDevice, ctx->stream));
    float* m = ctx->arena.get_t<float>("mlp.m", S.np);
    float* v = ctx->arena.get_t<float>("mlp.v", S.np);
    hipLaunchKernelGGL(mlp_zero, dim3((S.np + 255) / 256), dim3(256), 0, ctx->stream, m, S.np);
    hipLaunchKernelGGL(mlp_zero, dim3((S.np + 255) / 256), dim3(256), 0, ctx->stream, v, S.np);
    S.step = 0;
    SVH_HIP(hipStreamSynchronize(ctx->stream));
  });
}

int svh_mlp_get_params(svh_ctx* ctx, float* params) {
  return mguard(ctx, [&] {
    SVH_REQUIRE(ctx->mlp && params, "mlp: not initialised");
    MlpState& S = *ctx->mlp;
    float* p = ctx->arena.get_t<float>("mlp.params", S.np);
    SVH_HIP(hipMemcpyAsync(params, p, (size_t)S.np * 4, hipMemcpyDeviceToHost, ctx->stream));
    SVH_HIP(hipStreamSynchronize(ctx->stream));
  });
}

int svh_mlp_forward(svh_ctx* ctx, const float* X, int32_t B, float* out) {
  return mguard(ctx, [&] {
    SVH_REQUIRE(ctx->mlp && X && out && B >= 1, "mlp_forward: bad argument");
    MlpState& S = *ctx->mlp;
    float* dX = ctx->arena.get_t<float>("mlp.fx", (size_t)B * S.nin);
    SVH_HIP(hipMemcpyAsync(dX, X, (size_t)B * S.nin * 4, hipMemcpyHostToDevice, ctx->stream));
    Acts a = acts_of(ctx, S, B);
    forward(ctx, S, ctx->arena.get_t<float>("mlp.params", S.np), dX, B, a);
    SVH_HIP(hipGetLastError());
    SVH_HIP(hipMemcpyAsync(out, a.out, (size_t)B * S.nout * 4, hipMemcpyDeviceToHost,
                           ctx->stream));
    SVH_HIP(hipStreamSynchronize(ctx->stream));
  });
}

int svh_mlp_loss_grad(svh_ctx* ctx, const float* X, const float* T, const float* Mk, int32_t B,
                      float* loss, float* grads) {
  return mguard(ctx, [&] {
    SVH_REQUIRE(ctx->mlp && X && T && Mk && loss && grads && B >= 1, "mlp_loss_grad: argument");
    MlpState& S = *ctx->mlp;
    float* dX = ctx->arena.get_t<float>("mlp.gx", (size_t)B * S.nin);
    float* dT = ctx->arena.get_t<float>("mlp.gt", (size_t)B * S.nout);
    float* dM = ctx->arena.get_t<float>("mlp.gm", (size_t)B * S.nout);
    float* dg = ctx->arena.get_t<float>("mlp.grads", S.np);
    float* dl = ctx->arena.get_t<float>("mlp.loss", 1);
    SVH_HIP(hipMemcpyAsync(dX, X, (size_t)B * S.nin * 4, hipMemcpyHostToDevice, ctx->stream));
    SVH_HIP(hipMemcpyAsync(dT, T, (size_t)B * S.nout * 4, hipMemcpyHostToDevice, ctx->stream));
    SVH_HIP(hipMemcpyAsync(dM, Mk, (size_t)B * S.nout * 4, hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(mlp_zero, dim3(1), dim3(64), 0, ctx->stream, dl, 1);
    Acts a = acts_of(ctx, S, B);
    const float* p = ctx->arena.get_t<float>("mlp.params", S.np);
    forward(ctx, S, p, dX, B, a);
    backward(ctx, S, p, dX, dT, dM, B, a, dg, dl);
    SVH_HIP(hipGetLastError());
    SVH_HIP(hipMemcpyAsync(grads, dg, (size_t)S.np * 4, hipMemcpyDeviceToHost, ctx->stream));
    SVH_HIP(hipMemcpyAsync(loss, dl, 4, hipMemcpyDeviceToHost, ctx->stream));
    SVH_HIP(hipStreamSynchronize(ctx->stream));
  });
}

int svh_mlp_load_data(svh_ctx* ctx, const float* X, const float* T, const float* Mk,
                      int32_t n) {
  return mguard(ctx, [&] {
    SVH_REQUIRE(ctx->mlp && X && T && Mk && n >= 1, "mlp_load_data: bad argument");
    MlpState& S = *ctx->mlp;
    float* dX = ctx->arena.get_t<float>("mlp.dataX", (size_t)n * S.nin);
    float* dT = ctx->arena.get_t<float>("mlp.dataT", (size_t)n * S.nout);
    float* dM = ctx->arena.get_t<float>("mlp.dataM", (size_t)n * S.nout);
    SVH_HIP(hipMemcpyAsync(dX, X, (size_t)n * S.nin * 4, hipMemcpyHostToDevice, ctx->stream));
    SVH_HIP(hipMemcpyAsync(dT, T, (size_t)n * S.nout * 4, hipMemcpyHostToDevice, ctx->stream));
    SVH_HIP(hipMemcpyAsync(dM, Mk, (size_t)n * S.nout * 4, hipMemcpyHostToDevice, ctx->stream));
    SVH_HIP(hipStreamSynchronize(ctx->stream));
    S.ndata = n;
  });
}

// One pass over the loaded data in the order `perm` (n indices), minibatches
// of B rows (the last one may be short), one Adam step each.  *mean_loss
// receives the mean minibatch loss.
int svh_mlp_train_epoch(svh_ctx* ctx, const int32_t* perm, int32_t B, float lr, float beta1,
                        float beta2, float eps, float weight_decay, float* mean_loss) {
  return mguard(ctx, [&] {
    SVH_REQUIRE(ctx->mlp && perm && mean_loss && B >= 1, "mlp_train_epoch: bad argument");
    MlpState& S = *ctx->mlp;
    SVH_REQUIRE(S.ndata >= 1, "mlp_train_epoch: no data loaded");
    const int n = S.ndata;
    int* dperm = ctx->arena.get_t<int>("mlp.perm", n);
    SVH_HIP(hipMemcpyAsync(dperm, perm, (size_t)n * 4, hipMemcpyHostToDevice, ctx->stream));
    const float* X = ctx->arena.get_t<float>("mlp.dataX", (size_t)n * S.nin);
    const float* T = ctx->arena.get_t<float>("mlp.dataT", (size_t)n * S.nout);
    const float* Mk = ctx->arena.get_t<float>("mlp.dataM", (size_t)n * S.nout);
    float* Xb = ctx->arena.get_t<float>("mlp.bx", (size_t)B * S.nin);
    float* Tb = ctx->arena.get_t<float>("mlp.bt", (size_t)B * S.nout);
    float* Mb = ctx->arena.get_t<float>("mlp.bm", (size_t)B * S.nout);
    float* p = ctx->arena.get_t<float>("mlp.params", S.np);
    float* m = ctx->arena.get_t<float>("mlp.m", S.np);
    float* v = ctx->arena.get_t<float>("mlp.v", S.np);
    float* g = ctx->arena.get_t<float>("mlp.grads", S.np);
    float* dl = ctx->arena.get_t<float>("mlp.loss", 1);
    hipLaunchKernelGGL(mlp_zero, dim3(1), dim3(64), 0, ctx->stream, dl, 1);
    Acts a = acts_of(ctx, S, B);
    int nb = 0;
    for (int s = 0; s < n; s += B) {
      const int b = std::min(B, n - s);
      const int* ix = dperm + s;
      hipLaunchKernelGGL(mlp_gather, dim3((b * S.nin + 255) / 256), dim3(256), 0, ctx->stream, X,
                         ix, b, S.nin, Xb);
      hipLaunchKernelGGL(mlp_gather, dim3((b * S.nout + 255) / 256), dim3(256), 0, ctx->stream, T,
                         ix, b, S.nout, Tb);
      hipLaunchKernelGGL(mlp_gather, dim3((b * S.nout + 255) / 256), dim3(256), 0, ctx->stream, Mk,
                         ix, b, S.nout, Mb);
      forward(ctx, S, p, Xb, b, a);
      backward(ctx, S, p, Xb, Tb, Mb, b, a, g, dl);
      ++S.step;
      const float c1 = 1.f - std::pow(beta1, (float)S.step);
      const float c2 = 1.f - std::pow(beta2, (float)S.step);
      hipLaunchKernelGGL(mlp_adam, dim3(std::min(1024, (S.np + 255) / 256)), dim3(256), 0,
                         ctx->stream, p, g, m, v, (long long)S.np, lr, beta1, beta2, eps,
                         weight_decay, c1, c2);
      ++nb;
    }
    SVH_HIP(hipGetLastError());
    float hl = 0.f;
    SVH_HIP(hipMemcpyAsync(&hl, dl, 4, hipMemcpyDeviceToHost, ctx->stream));
    SVH_HIP(hipStreamSynchronize(ctx->stream));
    *mean_loss = hl / nb;
  });
}

int svh_simulate_rff(svh_ctx* ctx, int32_t ndata, const int64_t* n, const int32_t* d,
                     int32_t F, const double* X, const double* W, const double* b,
                     const double* amp, double* y) {
  return mguard(ctx, [&] {
    SVH_REQUIRE(ndata >= 1 && n && d && X && W && b && amp && y && F >= 1,
                "simulate_rff: bad argument");
    std::vector<RffDesc> h(ndata);
    long long xo = 0, yo = 0, wo = 0, bo = 0, nmax = 0;
    for (int i = 0; i < ndata; ++i) {
      SVH_REQUIRE(n[i] >= 1 && n[i] < (1ll << 31) && d[i] >= 1 && d[i] <= 16,
                  "simulate_rff: n or d out of range (d <= 16)");
      h[i] = RffDesc{xo, yo, wo, bo, (int)n[i], d[i], F, 0, amp[i]};
      xo += n[i] * d[i];
      yo += n[i];
      wo += (long long)F * d[i];
      bo += F;
      nmax = std::max<long long>(nmax, n[i]);
    }
    SVH_REQUIRE(ndata <= 65535, "simulate_rff: at most 65535 datasets per call");
    hipStream_t s = ctx->stream;
    RffDesc* dd = ctx->arena.get_t<RffDesc>("rff.desc", ndata);
    double* dX = ctx->arena.get_t<double>("rff.X", xo);
    double* dW = ctx->arena.get_t<double>("rff.W", wo);
    double* db = ctx->arena.get_t<double>("rff.b", bo);
    double* dy = ctx->arena.get_t<double>("rff.y", yo);
    SVH_HIP(hipMemcpyAsync(dd, h.data(), ndata * sizeof(RffDesc), hipMemcpyHostToDevice, s));
    SVH_HIP(hipMemcpyAsync(dX, X, xo * sizeof(double), hipMemcpyHostToDevice, s));
    SVH_HIP(hipMemcpyAsync(dW, W, wo * sizeof(double), hipMemcpyHostToDevice, s));
    SVH_HIP(hipMemcpyAsync(db, b, bo * sizeof(double), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(rff_kernel, dim3((unsigned)((nmax + 255) / 256), ndata), dim3(256), 0, s,
                       dd, dX, dW, db, dy);
    SVH_HIP(hipGetLastError());
    SVH_HIP(hipMemcpyAsync(y, dy, yo * sizeof(double), hipMemcpyDeviceToHost, s));
    SVH_HIP(hipStreamSynchronize(s));
  });
}

}  // extern "C"
