// Joint prediction solves on gfx950 (gp.py:360-402): the sparse triangular
// systems with Lt = U_pp^T of the observed-first prediction factor.
//
// Row i of Lt holds block i's prediction-side entries: its neighbours among
// earlier prediction points (columns block[i,k] - nobs for block[i,k] >=
// nobs) and itself (the last entry of the block, the diagonal).  The solve
//   x_i = (b_i - sum_k a_ik x_{c_k}) / a_i,self
// runs level by level: a row's level is one more than the deepest level among
// its prediction neighbours, so the rows of a level depend only on earlier
// levels and are solved in parallel, one wave per row.  Several right-hand
// sides (sample paths, identity columns) share the row's indices and values:
// lanes run over right-hand sides, x is stored row-major (rows x nrhs) so each
// neighbour row is one coalesced read.  The observed part of the mean's
// right-hand side, b_i = -sum_{k: block[i,k] < nobs} a_ik r_{block[i,k]}, is
// one wave reduction per row.  No counterpart kernels in the reference (it
// calls scipy.sparse.linalg.spsolve_triangular on the host).

#include "svh_internal.h"

#include <algorithm>
#include <vector>

namespace {

constexpr int kJWave = 64;
constexpr int kJRows = 4;   // rows (waves) per workgroup

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, kJWave);
  return v;
}

// b_i = -sum over observed entries of block i of a_ik * resid[block[i,k]]
__global__ __launch_bounds__(kJWave* kJRows) void joint_rhs_kernel(
    const int64_t* __restrict__ block, const double* __restrict__ av,
    const double* __restrict__ resid, int64_t ns, int K, int64_t nobs, double* __restrict__ b) {
  const int lane = threadIdx.x & (kJWave - 1);
  const int64_t i = (int64_t)blockIdx.x * kJRows + threadIdx.x / kJWave;
  if (i >= ns) return;
  double s = 0.0;
  for (int k = lane; k < K - 1; k += kJWave) {
    const int64_t c = block[i * K + k];
    if (c >= 0 && c < nobs) s += av[i * K + k] * resid[c];
  }
  s = wave_sum(s);
  if (lane == 0) b[i] = -s;
}

// One level of the forward substitution, in place on x (rows x nrhs, holding
// the right-hand sides of unsolved rows).
__global__ __launch_bounds__(kJWave* kJRows) void joint_level_kernel(
    const int* __restrict__ rows, int nrows, const int64_t* __restrict__ block,
    const double* __restrict__ av, int K, int64_t nobs, int nrhs, double* __restrict__ x) {
  const int lane = threadIdx.x & (kJWave - 1);
  const int w = blockIdx.x * kJRows + threadIdx.x / kJWave;
  if (w >= nrows) return;
  const int64_t i = rows[w];
  const int64_t* bi = block + i * K;
  const double* ai = av + i * K;
  const double dinv = 1.0 / ai[K - 1];
  for (int r0 = 0; r0 < nrhs; r0 += kJWave) {
    const int r = r0 + lane;
    if (r >= nrhs) break;
    double s = x[i * nrhs + r];
#pragma unroll 4
    for (int k = 0; k < K - 1; ++k) {
      const int64_t c = bi[k];
      if (c >= nobs) s -= ai[k] * x[(c - nobs) * nrhs + r];
    }
    x[i * nrhs + r] = s * dinv;
  }
}

__global__ void joint_zero(double* __restrict__ v, int64_t n) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t < n) v[t] = 0.0;
}

// x = the identity columns c0 .. c0 + nc - 1 (rows x nc)
__global__ void joint_identity(double* __restrict__ x, int64_t ns, int nc, int64_t c0) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= ns * nc) return;
  const int64_t i = t / nc, j = t - i * nc;
  x[t] = (i == c0 + j) ? 1.0 : 0.0;
}

// var[i] += sum_j x[i, j]^2
__global__ __launch_bounds__(kJWave* kJRows) void joint_rowsq(const double* __restrict__ x,
                                                             int64_t ns, int nc,
                                                             double* __restrict__ var) {
  const int lane = threadIdx.x & (kJWave - 1);
  const int64_t i = (int64_t)blockIdx.x * kJRows + threadIdx.x / kJWave;
  if (i >= ns) return;
  double s = 0.0;
  for (int j = lane; j < nc; j += kJWave) {
    const double v = x[i * nc + j];
    s += v * v;
  }
  s = wave_sum(s);
  if (lane == 0) var[i] += s;
}

struct Levels {
  std::vector<int> rows;   // rows grouped by level
  std::vector<int> off;    // level l: rows[off[l] .. off[l+1])
};

Levels joint_levels(const int64_t* block, int64_t ns, int K, int64_t nobs) {
  std::vector<int> lev((size_t)ns, 0);
  int maxl = 0;
  for (int64_t i = 0; i < ns; ++i) {
    int l = 0;
    for (int k = 0; k < K - 1; ++k) {
      const int64_t c = block[i * K + k];
      if (c >= nobs) {
        SVH_REQUIRE(c - nobs < i, "joint: a prediction neighbour must precede its row");
        l = std::max(l, lev[(size_t)(c - nobs)] + 1);
      }
    }
    SVH_REQUIRE(block[i * K + K - 1] == nobs + i, "joint: the last block entry must be the row");
    lev[(size_t)i] = l;
    maxl = std::max(maxl, l);
  }
  Levels L;
  L.off.assign((size_t)maxl + 2, 0);
  for (int64_t i = 0; i < ns; ++i) L.off[(size_t)lev[(size_t)i] + 1] += 1;
  for (int l = 0; l <= maxl; ++l) L.off[(size_t)l + 1] += L.off[(size_t)l];
  L.rows.resize((size_t)ns);
  std::vector<int> pos(L.off.begin(), L.off.end() - 1);
  for (int64_t i = 0; i < ns; ++i) L.rows[(size_t)pos[(size_t)lev[(size_t)i]]++] = (int)i;
  return L;
}

void solve_levels(svh_ctx* ctx, const Levels& L, const int* drows, const int64_t* dblk,
                  const double* dav, int K, int64_t nobs, int nrhs, double* dx) {
  const int nlev = (int)L.off.size() - 1;
  for (int l = 0; l < nlev; ++l) {
    const int a = L.off[(size_t)l], nr = L.off[(size_t)l + 1] - a;
    if (nr == 0) continue;
    hipLaunchKernelGGL(joint_level_kernel, dim3((unsigned)((nr + kJRows - 1) / kJRows)),
                       dim3(kJWave * kJRows), 0, ctx->stream, drows + a, nr, dblk, dav, K, nobs,
                       nrhs, dx);
  }
  SVH_HIP(hipGetLastError());
}

}  // namespace

void joint_solve_device(svh_ctx* ctx, int64_t nobs, int64_t ns, int K, const int64_t* block,
                        const double* av, const double* resid, int nrhs, const double* rhs,
                        double* mean, double* x, int exact_var, double* var) {
  SVH_REQUIRE(nobs >= 0 && ns >= 1 && K >= 1 && nrhs >= 0, "joint: bad sizes");
  SVH_REQUIRE(ns < (1ll << 31), "joint: too many prediction points");
  for (int64_t e = 0; e < ns * K; ++e)
    SVH_REQUIRE(block[e] >= -1 && block[e] < nobs + ns, "joint: index out of range");
  const Levels L = joint_levels(block, ns, K, nobs);
  hipStream_t s = ctx->stream;
  int64_t* dblk = ctx->arena.get_t<int64_t>("jt.blk", (size_t)ns * K);
  double* dav = ctx->arena.get_t<double>("jt.a", (size_t)ns * K);
  int* drows = ctx->arena.get_t<int>("jt.rows", (size_t)ns);
  SVH_HIP(hipMemcpyAsync(dblk, block, (size_t)ns * K * sizeof(int64_t), hipMemcpyHostToDevice, s));
  SVH_HIP(hipMemcpyAsync(dav, av, (size_t)ns * K * sizeof(double), hipMemcpyHostToDevice, s));
  SVH_HIP(hipMemcpyAsync(drows, L.rows.data(), (size_t)ns * sizeof(int), hipMemcpyHostToDevice, s));
  const unsigned gr = (unsigned)((ns + kJRows - 1) / kJRows);
  if (mean) {
    SVH_REQUIRE(resid || nobs == 0, "joint: resid is NULL");
    double* dres = ctx->arena.get_t<double>("jt.res", (size_t)std::max<int64_t>(nobs, 1));
    double* dm = ctx->arena.get_t<double>("jt.mean", (size_t)ns);
    if (nobs > 0)
      SVH_HIP(hipMemcpyAsync(dres, resid, (size_t)nobs * sizeof(double), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(joint_rhs_kernel, dim3(gr), dim3(kJWave * kJRows), 0, s, dblk, dav, dres,
                       ns, K, nobs, dm);
    solve_levels(ctx, L, drows, dblk, dav, K, nobs, 1, dm);
    SVH_HIP(hipMemcpyAsync(mean, dm, (size_t)ns * sizeof(double), hipMemcpyDeviceToHost, s));
  }
  if (nrhs > 0) {
    SVH_REQUIRE(rhs && x, "joint: rhs / x is NULL");
    double* dx = ctx->arena.get_t<double>("jt.x", (size_t)ns * nrhs);
    SVH_HIP(hipMemcpyAsync(dx, rhs, (size_t)ns * nrhs * sizeof(double), hipMemcpyHostToDevice, s));
    solve_levels(ctx, L, drows, dblk, dav, K, nobs, nrhs, dx);
    SVH_HIP(hipMemcpyAsync(x, dx, (size_t)ns * nrhs * sizeof(double), hipMemcpyDeviceToHost, s));
  }
  if (exact_var) {
    SVH_REQUIRE(var, "joint: var is NULL");
    // row sums of squares of Lt^-1, identity columns in chunks
    const int nc = (int)std::min<int64_t>(ns, 1024);
    double* dx = ctx->arena.get_t<double>("jt.xi", (size_t)ns * nc);
    double* dv = ctx->arena.get_t<double>("jt.var", (size_t)ns);
    hipLaunchKernelGGL(joint_zero, dim3((unsigned)((ns + 255) / 256)), dim3(256), 0, s, dv, ns);
    for (int64_t c0 = 0; c0 < ns; c0 += nc) {
      const int w = (int)std::min<int64_t>(nc, ns - c0);
      hipLaunchKernelGGL(joint_identity, dim3((unsigned)((ns * w + 255) / 256)), dim3(256), 0, s,
                         dx, ns, w, c0);
      solve_levels(ctx, L, drows, dblk, dav, K, nobs, w, dx);
      hipLaunchKernelGGL(joint_rowsq, dim3(gr), dim3(kJWave * kJRows), 0, s, dx, ns, w, dv);
    }
    SVH_HIP(hipGetLastError());
    SVH_HIP(hipMemcpyAsync(var, dv, (size_t)ns * sizeof(double), hipMemcpyDeviceToHost, s));
  }
  SVH_HIP(hipStreamSynchronize(s));
}
