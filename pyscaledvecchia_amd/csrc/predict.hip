// (c) Prediction blocks on gfx950 (restates the block loops of
// gp.py:242-277 (_predict_scaled) and gp.py:332-348 (_joint_factor) with
// _covariance.py:86-108, 220-252).
//
// pred_mfma_kernel (blocks of 48 to 143 points, e.g. m_pred = 140): one
// 512-thread workgroup per block.  The covariance is held as 16 x 16 tiles of
// its lower triangle in registers, in the matrix-core accumulator layout,
// spread over the eight waves.  It is assembled one 16-column panel at a time
// through LDS.  The Cholesky factorisation is right-looking by 16-column
// panels:
//   * the panel's diagonal tile is factorised in FP64 by one wave (rows in
//     lanes, column values broadcast through LDS);
//   * the rows below are solved against it in FP64, one row per thread;
//   * the trailing update A22 -= L21 L21^T runs on the matrix cores: FP64
//     16x16x4 MFMA in FP64 mode, FP32 16x16x4 MFMA (FP32 accumulation,
//     subtracted from the FP64 tiles) in mixed mode.
// Mode 1 (marginal prediction) appends the residual vector r~ = [r_nb, 0] as
// one more row: the factor of [[Sigma, r~], [r~^T, inf]] has L^-1 r~ as its
// last row, so with a = L^-T e_K
//   a . r~ = (L^-1 r~)_K,  mean = -(a . r~) / a_K = -(L^-1 r~)_K L_KK,
//   var = 1 / a_K^2 = L_KK^2
// and no back substitution is needed.  Mode 0 (columns of the sparse inverse
// factor) keeps each factorised panel in a per-workgroup scratch area and
// forms a = L^-T e_K by a column sweep.
//
// pred_blocks_kernel (smaller and larger blocks): one wave per block, the
// covariance lower triangle packed in LDS, factorised right-looking;
// a = L^-T e_K by a column sweep.
//
// The chunk jitter level (_covariance.py:241-252) adds 10^p tr(Sigma)/K to
// the K real diagonal entries.

#include "svh_internal.h"
#include "bessel.h"

#include <cmath>
#include <cstdlib>
#include <string>
#include <vector>

namespace {

// v of lane `src` of this lane's 16-lane row (DPP row_newbcast); src is a
// compile-time constant after unrolling, the switch folds away.
__device__ __forceinline__ double row_bcast(double v, int src) {
  switch (src) {
#define SVH_RB(n) \
  case n:         \
    return __builtin_amdgcn_update_dpp(0.0, v, 0x150 + n, 0xF, 0xF, false);
    SVH_RB(0) SVH_RB(1) SVH_RB(2) SVH_RB(3) SVH_RB(4) SVH_RB(5) SVH_RB(6) SVH_RB(7)
    SVH_RB(8) SVH_RB(9) SVH_RB(10) SVH_RB(11) SVH_RB(12) SVH_RB(13) SVH_RB(14) SVH_RB(15)
#undef SVH_RB
    default:
      return 0.0;
  }
}

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
  svh_bessel::MaternNu gnu;   // nu_code 3: general smoothness (Bessel K_nu)
};

__device__ __forceinline__ double matern_M(int nu_code, double r,
                                           const svh_bessel::MaternNu& gnu) {
  if (nu_code == 3) return svh_bessel::matern_general_M(gnu, r);
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
      double M = (i == j) ? 1.0 : matern_M(a.nu_code, sqrt(s), a.gnu);
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

// ------------------------------------------------------------------ tiled
constexpr int kPT = 512;       // threads per block of pred_mfma_kernel (8 waves)
constexpr int kPTiles = 9;     // at most 9 x 16 = 144 rows (K + 1 <= 144)
constexpr int kPSlots = 6;     // ceil(45 tiles / 8 waves)
constexpr int kPStride = 17;   // panel row stride (doubles) in LDS

typedef double pdvec4 __attribute__((ext_vector_type(4)));
typedef float pfvec4 __attribute__((ext_vector_type(4)));

// Row of accumulator register v held by lane group lk = lane >> 4 (the column
// is lane & 15): FP64 16x16x4 MFMA D[lk + 4v][.], FP32 16x16x4 MFMA
// D[4 lk + v][.].  Operands: lane L carries A[L & 15][L >> 4] and
// B[L >> 4][L & 15] for both.
template <bool MIX>
__device__ __forceinline__ int acc_row(int lk, int v) {
  return MIX ? 4 * lk + v : lk + 4 * v;
}

__device__ __forceinline__ double readlane_d(double x, int l) {
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), l);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

struct PredTileArgs {
  double* scratch;   // mode 0: gridDim.x * Kp * Kp doubles (factor rows)
};

template <bool MIX>
__global__ __launch_bounds__(kPT) void pred_mfma_kernel(PredArgs a, PredTileArgs ta) {
  extern __shared__ double lds[];
  const int Kb = a.K, d = a.d, mode = a.mode;
  const int aug = (mode == 1) ? 1 : 0;
  const int Kp = ((Kb + aug + 15) / 16) * 16;
  const int KT = Kp / 16;
  const int ntile = KT * (KT + 1) / 2;
  const int tid = threadIdx.x, lane = tid & (kWave - 1);
  const int wid = __builtin_amdgcn_readfirstlane(tid / kWave);   // wave-uniform
  const int li = lane & 15, lk = lane >> 4;
  double* xs = lds;                   // Kp * d scaled coordinates
  double* rv = xs + Kp * d;           // Kp: residual of each neighbour (mode 1)
  double* msk = rv + Kp;              // Kp: nugget mask
  double* P = msk + Kp;               // Kp * kPStride: current panel
  double* misc = P + Kp * kPStride;   // [0..15] 1/L_cc, [16] L_KK, [17] (L^-1 r~)_K, [18] fail,
                                      // [32..63] column broadcast
  const double kInf = __longlong_as_double(0x7ff0000000000000ll);
  // this wave's tiles: tile t = 8 s + wid, (I, J) in row-major lower order
  int tI[kPSlots], tJ[kPSlots];
#pragma unroll
  for (int s = 0; s < kPSlots; ++s) {
    const int t = 8 * s + wid;
    int I = 0;
    while ((I + 1) * (I + 2) / 2 <= t) ++I;
    tI[s] = (t < ntile) ? I : -1;
    tJ[s] = (t < ntile) ? t - I * (I + 1) / 2 : -1;
  }
  double* scr = (mode == 0) ? ta.scratch + (int64_t)blockIdx.x * Kp * Kp : nullptr;
  for (int64_t b = blockIdx.x; b < a.nblk; b += gridDim.x) {
    const int ch = (int)(b / a.chunk_step);
    const int lev = a.level[ch];
    __syncthreads();
    for (int e = tid; e < Kp * d; e += kPT) {
      const int t = e / d, l = e - t * d;
      double v = 0.0;
      if (t < Kb) {
        const double* src;
        if (t < Kb - 1) {
          src = a.Xn + a.nb[b * a.nb_stride + t] * d;
        } else {
          const int64_t ti = a.tg ? a.tg[b * a.tg_stride] : b;
          src = a.Xt + ti * d;
        }
        v = src[l] * a.inv_lam[l];
      }
      xs[e] = v;
    }
    for (int t = tid; t < Kp; t += kPT) {
      msk[t] = (t < Kb) ? (a.mask ? a.mask[b * Kb + t] : ((t == Kb - 1 && a.noise_free) ? 0.0 : 1.0))
                        : 0.0;
      rv[t] = (aug && t < Kb - 1) ? a.resid[a.nb[b * a.nb_stride + t]] : 0.0;
    }
    if (tid < 32) misc[tid] = 0.0;
    __syncthreads();
    double jit = 0.0;
    if (lev > 0) {
      double tr = 0.0;
      for (int t = 0; t < Kb; ++t) tr += a.var * (1.0 + a.nug * msk[t]);
      jit = kJitDevP[lev] * (tr / Kb);
    }
    // ---- covariance tiles, one 16-column panel at a time through LDS
    // (padding rows: identity; appended row: r~ and an infinite diagonal)
    pdvec4 T[kPSlots];
    for (int J = 0; J < KT; ++J) {
      const int npr = Kp - 16 * J;
      __syncthreads();
      for (int e = tid; e < npr * 16; e += kPT) {
        const int row = 16 * J + (e >> 4), col = 16 * J + (e & 15);
        double val = 0.0;
        if (row < Kb && col < Kb) {
          if (row == col) {
            val = a.var * (1.0 + a.nug * msk[row]) + jit;
          } else {
            double s2 = 0.0;
            for (int l = 0; l < d; ++l) {
              const double t = xs[row * d + l] - xs[col * d + l];
              s2 = fma(t, t, s2);
            }
            val = a.var * matern_M(a.nu_code, sqrt(s2), a.gnu);
          }
        } else if (aug && row == Kb) {
          val = (col < Kb - 1) ? rv[col] : (col == Kb ? kInf : 0.0);
        } else if (row == col) {
          val = 1.0;
        }
        P[(e >> 4) * kPStride + (e & 15)] = val;
      }
      __syncthreads();
#pragma unroll
      for (int s = 0; s < kPSlots; ++s)
        if (tJ[s] == J) {
#pragma unroll
          for (int v = 0; v < 4; ++v)
            T[s][v] = P[(16 * (tI[s] - J) + acc_row<MIX>(lk, v)) * kPStride + li];
        }
    }
    // ---- right-looking factorisation by 16-column panels
    for (int J = 0; J < KT; ++J) {
      const int npr = Kp - 16 * J;   // panel rows
      __syncthreads();               // previous panel's readers are done
#pragma unroll
      for (int s = 0; s < kPSlots; ++s)
        if (tJ[s] == J) {
#pragma unroll
          for (int v = 0; v < 4; ++v)
            P[(16 * (tI[s] - J) + acc_row<MIX>(lk, v)) * kPStride + li] = T[s][v];
        }
      __syncthreads();
      // diagonal tile: FP64, lane r holds row r & 15 (four copies per wave,
      // one per 16-lane row).  Step c: the rows' current column-c values are
      // broadcast inside each 16-lane row by DPP row_newbcast (no LDS), the
      // pivot among them.
      if (wid == 0) {
        const int r = li;
        double R[16];
#pragma unroll
        for (int c = 0; c < 16; ++c) R[c] = (c <= r) ? P[r * kPStride + c] : 0.0;
        double dg = 1.0;
        bool f = false;
#pragma unroll
        for (int c = 0; c < 16; ++c) {
          double g[16];
#pragma unroll
          for (int c2 = c; c2 < 16; ++c2) g[c2] = row_bcast(R[c], c2);
          const double piv = g[c];
          // 1 / sqrt(piv) by the hardware estimate and three Newton steps
          // (faithful to an ulp), sqrt(piv) = piv / sqrt(piv)
          double inv = __builtin_amdgcn_rsq(piv);
          const double h = 0.5 * piv;
          inv = inv * fma(-h * inv, inv, 1.5);
          inv = inv * fma(-h * inv, inv, 1.5);
          inv = inv * fma(-h * inv, inv, 1.5);
          const double lpp = piv * inv;
          if (16 * J + c < Kb && !(piv > 0.0)) f = true;
          if (16 * J + c == Kb - 1) dg = lpp;   // L_KK
          R[c] = (r > c) ? R[c] * inv : ((r == c) ? lpp : R[c]);
#pragma unroll
          for (int c2 = c + 1; c2 < 16; ++c2)
            if (r >= c2) R[c2] = fma(-R[c], g[c2] * inv, R[c2]);
        }
        if (lane < 16) {
#pragma unroll
          for (int c = 0; c < 16; ++c)
            if (c <= r) P[r * kPStride + c] = R[c];
#pragma unroll
          for (int c = 0; c < 16; ++c)
            if (c == r) misc[r] = 1.0 / R[c];
        }
        if (lane == 0) {
          if (f) misc[18] = 1.0;
          if (Kb - 1 >= 16 * J && Kb - 1 < 16 * J + 16) misc[16] = dg;
        }
      }
      __syncthreads();
      // rows below the diagonal tile: x L_JJ^T = a, one row per thread
      // (npr - 16 <= 128 rows: one pass)
      if (16 + tid < npr) {
        const int rr = 16 + tid;
        double x[16];
#pragma unroll
        for (int c = 0; c < 16; ++c) x[c] = P[rr * kPStride + c];
#pragma unroll
        for (int c = 0; c < 16; ++c) {
          x[c] *= misc[c];
#pragma unroll
          for (int c2 = c + 1; c2 < 16; ++c2) x[c2] = fma(-x[c], P[c2 * kPStride + c], x[c2]);
        }
#pragma unroll
        for (int c = 0; c < 16; ++c) P[rr * kPStride + c] = x[c];
      }
      __syncthreads();
      if (aug && tid == 0 && Kb - 1 >= 16 * J && Kb - 1 < 16 * J + 16)
        misc[17] = P[(Kb - 16 * J) * kPStride + (Kb - 1 - 16 * J)];
      if (mode == 0) {
        // keep the factorised panel: rows 16J.., columns 16J..16J+15
        for (int e = tid; e < npr * 16; e += kPT) {
          const int rr = e >> 4, c = e & 15;
          scr[(int64_t)(16 * J + rr) * Kp + 16 * J + c] = P[rr * kPStride + c];
        }
      }
      // trailing update of this wave's tiles (I, K'), K' > J, on the matrix cores
#pragma unroll
      for (int s = 0; s < kPSlots; ++s)
        if (tJ[s] > J) {
          const double* pa = P + (16 * (tI[s] - J) + li) * kPStride;
          const double* pb = P + (16 * (tJ[s] - J) + li) * kPStride;
          if constexpr (!MIX) {
#pragma unroll
            for (int c4 = 0; c4 < 4; ++c4)
              T[s] = __builtin_amdgcn_mfma_f64_16x16x4f64(-pa[4 * c4 + lk], pb[4 * c4 + lk], T[s],
                                                          0, 0, 0);
          } else {
            pfvec4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int c4 = 0; c4 < 4; ++c4)
              acc = __builtin_amdgcn_mfma_f32_16x16x4f32((float)pa[4 * c4 + lk],
                                                         (float)pb[4 * c4 + lk], acc, 0, 0, 0);
#pragma unroll
            for (int v = 0; v < 4; ++v) T[s][v] -= (double)acc[v];
          }
        }
    }
    __syncthreads();
    const bool failed = misc[18] != 0.0;
    if (failed) {
      if (tid == 0) atomicOr(a.fail + ch, 1);
      continue;
    }
    if (mode == 1) {
      if (tid == 0) {
        const double LK = misc[16];
        a.mean[b] = -misc[17] * LK;
        a.varo[b] = LK * LK;
      }
    } else if (wid == 0) {
      // a = L^-T e_K by a column sweep over the kept factor rows:
      // a_i = t_i / L_ii, then t_j -= L_ij a_i for j < i
      double tv[3];
#pragma unroll
      for (int j = 0; j < 3; ++j) tv[j] = (lane + 64 * j == Kb - 1) ? 1.0 : 0.0;
      for (int i = Kb - 1; i >= 0; --i) {
        const double* Li = scr + (int64_t)i * Kp;
        const int sl = i >> 6;
        const double ti = readlane_d(sl == 0 ? tv[0] : (sl == 1 ? tv[1] : tv[2]), i & 63);
        const double ai = ti / Li[i];
        if (lane == 0) a.out_a[b * Kb + i] = ai;
#pragma unroll
        for (int j = 0; j < 3; ++j) {
          const int c = lane + 64 * j;
          if (c < i) tv[j] = fma(-Li[c], ai, tv[j]);
        }
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
                           double* dout_a, double* dmean, double* dvar, int* jitter_power_max,
                           int precision) {
  SVH_REQUIRE(K >= 1, "predict: empty block");
  int nu_code;
  if (nu == 0.5)
    nu_code = 0;
  else if (nu == 1.5)
    nu_code = 1;
  else if (nu == 2.5)
    nu_code = 2;
  else if (nu > 0.0 && std::isfinite(nu))
    nu_code = 3;
  else
    throw SvhError{SVH_ERR_ARG, "predict: nu must be positive"};
  if (nblk == 0) return;
  hipStream_t s = ctx->stream;
  // tiled matrix-core kernel while K (+1 appended row) fits 9 tiles of 16
  const int Kp = ((K + (mode == 1 ? 1 : 0) + 15) / 16) * 16;
  // SVH_PRED_KERNEL=packed selects the one-wave packed-LDS kernel (A/B timing)
  static const bool force_packed = [] {
    const char* e = std::getenv("SVH_PRED_KERNEL");
    return e && std::string(e) == "packed";
  }();
  // small blocks (fewer than 4 tiles) run faster one wave per block
  const bool tiled_ok = Kp >= 32 && Kp <= 16 * kPTiles && !force_packed;
  const bool tiled_dflt = tiled_ok && Kp >= 64;
  const size_t lds_t = (size_t)(Kp * d + 2 * Kp + Kp * kPStride + 64) * sizeof(double);
  const size_t lds_p = (size_t)(K * (K + 1) / 2 + K * d + 2 * K) * sizeof(double);
  SVH_REQUIRE((tiled_ok ? lds_t : lds_p) <= 160 * 1024, "predict: block too large for LDS");
  const bool packed_ok = lds_p <= 160 * 1024;
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
  a.gnu = svh_bessel::matern_nu(nu_code == 3 ? nu : 2.5);
  a.mode = mode;
  a.var = std::exp(theta[0]);
  a.nug = std::exp(theta[1 + d]);
  // launch-schedule candidates for the autotuner: the packed kernel, the tiled
  // kernel at 1, 2, 4 workgroups per CU; defaults as fixed before
  const int ncu = ctx->num_cu > 0 ? ctx->num_cu : 256;
  struct PCand {
    bool tiled;
    int G;
  };
  std::vector<PCand> cands;
  if (packed_ok) cands.push_back({false, (int)std::min<int64_t>(nblk, 8192)});
  int dflt = 0;
  if (tiled_ok) {
    for (int f : {1, 2, 4}) {
      if (tiled_dflt && f == (mode == 0 ? 2 : 4)) dflt = (int)cands.size();
      cands.push_back({true, (int)std::min<int64_t>(nblk, (int64_t)f * ncu)});
    }
  }
  int Gs = 1;
  for (const PCand& c : cands)
    if (c.tiled && mode == 0) Gs = std::max(Gs, c.G);
  PredTileArgs ta;
  ta.scratch = (tiled_ok && mode == 0)
                   ? ctx->arena.get_t<double>("p.scratch", (size_t)Gs * Kp * Kp)
                   : nullptr;
  struct PLaunch {
    hipStream_t s;
    const std::vector<PCand>* cands;
    size_t lds_t, lds_p;
    const PredArgs* a;
    PredTileArgs ta;
    bool mixed;
    static void run(void* u, int i) {
      const PLaunch& L = *static_cast<PLaunch*>(u);
      const PCand& c = (*L.cands)[i];
      if (!c.tiled)
        hipLaunchKernelGGL(pred_blocks_kernel, dim3(c.G), dim3(kWave), L.lds_p, L.s, *L.a);
      else if (L.mixed)
        hipLaunchKernelGGL(pred_mfma_kernel<true>, dim3(c.G), dim3(kPT), L.lds_t, L.s, *L.a, L.ta);
      else
        hipLaunchKernelGGL(pred_mfma_kernel<false>, dim3(c.G), dim3(kPT), L.lds_t, L.s, *L.a,
                           L.ta);
    }
  };
  PLaunch lau{s, &cands, lds_t, lds_p, &a, ta, precision == SVH_MIXED};
  int pick = dflt;
  if (ctx->tune.mode != 0 && cands.size() > 1) {
    char key[160];
    int lg = 0;
    while ((int64_t(1) << (lg + 1)) <= nblk) ++lg;
    std::snprintf(key, sizeof key, "pred_blocks:K%d:m%d:d%d:p%d:nb2^%d", K, mode, d, precision,
                  lg);
    pick = tune_choose(ctx, key, (int)cands.size(), dflt, &PLaunch::run, &lau);
    hipLaunchKernelGGL(p_fill_int, dim3((nchunk + 255) / 256), dim3(256), 0, s, fail, 0,
                       (int64_t)nchunk);
  }
  int* hn = static_cast<int*>(ctx->stage.get(sizeof(int)));
  int rounds = 0;
  for (;;) {
    ctx->timers.begin(s, "predict_blocks");
    PLaunch::run(&lau, pick);
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
