// (b) Profile Vecchia log-likelihood, gradient and Fisher information on
// gfx950 (restates likelihood.py:66-175 with _covariance.py:86-259).
//
// Per evaluation:
//   ll_pass1  (grid-stride, one wave per conditioning block): gather the
//             block [neighbours..., self] in scaled coordinates, assemble
//             Sigma (lower triangle) and g(r) (upper triangle) in LDS,
//             Cholesky with the chunk jitter level, a = L^-T e_K, right-hand
//             sides dS_p a, y, Z and one forward solve for all of them
//             (W, L^-1 y, L^-1 Z).  Fisher and the GLS normal equations are
//             accumulated per workgroup; the beta-dependent remainder of each
//             block is cached as a short record.
//   ll_reduce1 (1 workgroup): fixed-order sum of the partials, beta =
//             solve(XtX + 1e-12 I, Zty) by partial-pivot elimination.
//   ll_pass2  (grid-stride, 1 + P lanes per block record): z_K = u - v.beta,
//             W.z = s - T beta, loglik and gradient terms.
//   ll_reduce2 (1 workgroup): fixed-order sum, failed-chunk count.
// All reductions are deterministic.
//
// In the split-row kernel (K <= 32) two phases are small matrix products on
// the matrix cores (16x16x4 MFMA): the derivative right-hand sides dS_p a,
// written as G [a | a x~ | a x~^2] with coordinates centred on the block's
// own point, and the cross products W^T [W | L^-1 y | L^-1 Z] behind the
// Fisher information and the gradient records.  FP64 mode runs them as FP64
// MFMA; the mixed modes as FP32 MFMA with FP32 accumulation.  In the rows
// kernel the mixed modes also compute dS_p a (distances, g(r)) and the solve
// W = L^-1 (dS a) in FP32; SVH_MIXED_F32COV in addition assembles the
// off-diagonal covariance entries in FP32.  The diagonal, the Cholesky,
// log-determinant, a = L^-T e_K, L^-1 [y Z] and the GLS solve are FP64 in
// every mode.
//
// Cached record per block (need_grad):  u, v[q], L_KK^2, wK[P], s[P], T[P*q]
// (the logarithm of L_KK^2 is taken in pass 2, where it is batched over
// records instead of costing a whole wave per block in pass 1)
// where u = (L^-1 y)_K, v = (L^-1 Z)_K, W_p = L^-1 dS_p a, s_p = W_p.(L^-1 y),
// T_pc = W_p.(L^-1 Z_c).  Then z = L^-1 y - (L^-1 Z) beta gives
// z_K = u - v.beta and W_p.z = s_p - T_p.beta exactly (linearity of L^-1).

#include "svh_internal.h"
#include "bessel.h"

#include <cmath>
#include <cstddef>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <string>
#include <unordered_map>
#include <vector>

const double kJitterPow10[10] = {0.0,  1e-12, 1e-11, 1e-10, 1e-09,
                                 1e-08, 1e-07, 1e-06, 1e-05, 1e-04};

namespace {

constexpr int kWave = 64;
constexpr int kMaxKP = 128;

// lower-triangle pair table: p -> (i, j), i >= j, row-major (0,0),(1,0),(1,1),...
struct PairTable {
  unsigned short v[kMaxKP * (kMaxKP + 1) / 2];
  constexpr PairTable() : v() {
    int p = 0;
    for (int i = 0; i < kMaxKP; ++i)
      for (int j = 0; j <= i; ++j) v[p++] = (unsigned short)(i | (j << 8));
  }
};
__constant__ PairTable kPairs = PairTable();
__constant__ double kJitDev[10] = {0.0,  1e-12, 1e-11, 1e-10, 1e-09,
                                   1e-08, 1e-07, 1e-06, 1e-05, 1e-04};

__device__ __forceinline__ void pair_of(int p, int& i, int& j) {
  unsigned short w = kPairs.v[p];
  i = w & 0xff;
  j = w >> 8;
}

struct LLArgs {
  const double* X;
  const double* y;
  const double* Z;
  const int* rows;
  const int* nbr;
  const int* kcnt;
  const int* chunk;
  const int* level;
  int* fail;
  double inv_lam[16];   // 1 / range per input (kernel argument, no device copy)
  double* cache;
  double* partial;
  int64_t nblk;
  int d, q, mmax, P, need_grad, nu_code;
  double var, nug;
  int rec_stride;  // doubles between consecutive block records
  unsigned long long* prof;  // per-phase shader-clock sums (profiling runs), or NULL
  int zmode;       // trend columns: 0 gathered from Z, 1 Z = [1, x], 2 Z = 1
  int sig32;       // mixed kernels: off-diagonal covariance entries in float32
  // nu_code 3: general smoothness (Bessel K_nu); with est_nu the log-nu
  // derivative column (central difference in nu, steps gnu_p / gnu_m) sits
  // at 1 + d, before the nugget's
  svh_bessel::MaternNu gnu, gnu_p, gnu_m;
  int est_nu;
};

constexpr double kNuStep = 1e-3;   // central-difference step in nu (_matern_general.py:135)

__device__ __forceinline__ void matern(int nu_code, double r, double& M, double& g) {
  if (nu_code == 0) {  // nu = 0.5
    if (r > 0.0) {
      double e = exp(-r);
      M = e;
      g = -e / r;
    } else {
      M = 1.0;
      g = 0.0;
    }
  } else if (nu_code == 1) {  // nu = 1.5
    double c = 1.7320508075688772 * r;
    double e = exp(-c);
    M = (1.0 + c) * e;
    g = -3.0 * e;
  } else {  // nu = 2.5
    double c = 2.23606797749979 * r;
    double e = exp(-c);
    M = (1.0 + c + c * c / 3.0) * e;
    g = -(5.0 / 3.0) * (1.0 + c) * e;
  }
}

template <int KP>
__global__ __launch_bounds__(kWave) void ll_pass1(LLArgs a) {
  extern __shared__ double lds[];
  constexpr int LD = KP + 1;
  const int d = a.d, q = a.q, P = a.P, lane = threadIdx.x;
  const int NR = a.need_grad ? P + 1 + q : 1 + q;
  const int cy = a.need_grad ? P : 0;  // column of y in R
  const int NACC = a.need_grad ? P * P + q * q + q : q * q + q;
  const int F0 = 0, X0 = a.need_grad ? P * P : 0, Y0 = X0 + q * q;
  const int NC = a.need_grad ? 2 + q + 2 * P + P * q : 2 + q;
  double* xs = lds;                 // KP * d
  double* S = xs + KP * d;          // KP * LD
  double* R = S + KP * LD;          // NR * KP
  double* av = R + NR * KP;         // KP
  double* acc = av + KP;            // NACC
  double* Dn = acc + NACC;          // KP * LD when the smoothness is estimated
  int* pts = reinterpret_cast<int*>(Dn + (a.est_nu ? KP * LD : 0));  // KP

  for (int e = lane; e < NACC; e += kWave) acc[e] = 0.0;

  for (int64_t b = blockIdx.x; b < a.nblk; b += gridDim.x) {
    const int K = a.kcnt[b] + 1;
    const int row = a.rows[b];
    const int ch = a.chunk[b];
    const int lev = a.level[ch];
    __syncthreads();
    for (int t = lane; t < K; t += kWave) pts[t] = (t < K - 1) ? a.nbr[b * a.mmax + t] : row;
    __syncthreads();
    for (int e = lane; e < K * d; e += kWave) {
      int t = e / d, l = e - t * d;
      xs[e] = a.X[(int64_t)pts[t] * d + l] * a.inv_lam[l];
    }
    for (int t = lane; t < K; t += kWave) {
      int pt = pts[t];
      R[cy * KP + t] = a.y[pt];
      for (int c = 0; c < q; ++c) R[(cy + 1 + c) * KP + t] = a.Z[(int64_t)pt * q + c];
    }
    __syncthreads();
    // ---- covariance (lower) and g (strict upper)
    const int T = K * (K + 1) / 2;
    for (int p = lane; p < T; p += kWave) {
      int i, j;
      pair_of(p, i, j);
      double s = 0.0;
      for (int l = 0; l < d; ++l) {
        double t = xs[i * d + l] - xs[j * d + l];
        s += t * t;
      }
      double M, g;
      const double rr = sqrt(s);
      if (a.nu_code == 3)
        svh_bessel::matern_general(a.gnu, rr, M, g);
      else
        matern(a.nu_code, rr, M, g);
      if (i == j) {
        S[i * LD + i] = a.var * (M + a.nug);
        if (a.est_nu) Dn[i * LD + i] = 0.0;
      } else {
        S[i * LD + j] = a.var * M;
        S[j * LD + i] = g;
        if (a.est_nu) {
          const double dm = (svh_bessel::matern_general_M(a.gnu_p, rr) -
                             svh_bessel::matern_general_M(a.gnu_m, rr)) / (2.0 * kNuStep);
          Dn[i * LD + j] = dm;
          Dn[j * LD + i] = dm;
        }
      }
    }
    __syncthreads();
    double jit = 0.0;
    if (lev > 0) {
      double tr = 0.0;
      for (int t = 0; t < K; ++t) tr += S[t * LD + t];
      jit = kJitDev[lev] * (tr / K);
      __syncthreads();
      for (int t = lane; t < K; t += kWave) S[t * LD + t] += jit;
      __syncthreads();
    }
    // ---- Cholesky, right-looking, in place (lower)
    bool failed = false;
    for (int k = 0; k < K; ++k) {
      double dkk = S[k * LD + k];
      if (!(dkk > 0.0)) {
        failed = true;
        break;
      }
      double lkk = sqrt(dkk);
      double inv = 1.0 / lkk;
      for (int t = k + 1 + lane; t < K; t += kWave) S[t * LD + k] *= inv;
      __syncthreads();
      if (lane == 0) S[k * LD + k] = lkk;
      const int nt = (K - 1 - k) * (K - k) / 2;
      for (int p = lane; p < nt; p += kWave) {
        int i, j;
        pair_of(p, i, j);
        i += k + 1;
        j += k + 1;
        S[i * LD + j] -= S[i * LD + k] * S[j * LD + k];
      }
      __syncthreads();
    }
    if (failed) {
      if (lane == 0) atomicOr(a.fail + ch, 1);
      continue;
    }
    // ---- a = L^-T e_K
    for (int t = lane; t < K; t += kWave) av[t] = (t == K - 1) ? 1.0 : 0.0;
    __syncthreads();
    for (int j = K - 1; j >= 0; --j) {
      double aj = av[j] / S[j * LD + j];
      __syncthreads();
      if (lane == 0) av[j] = aj;
      for (int t = lane; t < j; t += kWave) av[t] -= S[j * LD + t] * aj;
      __syncthreads();
    }
    // ---- right-hand sides dS_p a
    if (a.need_grad) {
      const double lKK = S[(K - 1) * LD + (K - 1)];
      for (int t = lane; t < K; t += kWave) {
        // unjittered Sigma a = L (L^T a) - jit a = L_KK e_K - jit a  (L^T a = e_K)
        R[0 * KP + t] = ((t == K - 1) ? lKK : 0.0) - jit * av[t];
        R[(P - 1) * KP + t] = a.var * a.nug * av[t];
      }
      for (int e = lane; e < K * d; e += kWave) {
        int l = e / K, i = e - l * K;
        double xil = xs[i * d + l];
        double s = 0.0;
        for (int j = 0; j < K; ++j) {
          if (j == i) continue;
          double g = (j < i) ? S[j * LD + i] : S[i * LD + j];
          double t = xil - xs[j * d + l];
          s += g * (t * t) * av[j];
        }
        R[(1 + l) * KP + i] = -a.var * s;
      }
      if (a.est_nu) {
        // d Sigma / d log nu = var nu dM/dnu (zero diagonal)
        for (int i = lane; i < K; i += kWave) {
          double s = 0.0;
          for (int j = 0; j < K; ++j) s += Dn[i * LD + j] * av[j];
          R[(1 + d) * KP + i] = a.var * a.gnu.nu * s;
        }
      }
      __syncthreads();
    }
    // ---- forward solve L X = R (all NR columns)
    for (int j = 0; j < K; ++j) {
      double ljj = S[j * LD + j];
      for (int c = lane; c < NR; c += kWave) R[c * KP + j] = R[c * KP + j] / ljj;
      __syncthreads();
      const int rem = K - 1 - j;
      for (int p = lane; p < rem * NR; p += kWave) {
        int c = p / rem, i = j + 1 + (p - c * rem);
        R[c * KP + i] -= S[i * LD + j] * R[c * KP + j];
      }
      __syncthreads();
    }
    // ---- block record and accumulators
    double* rec = a.cache + b * NC;
    const int km = K - 1;
    for (int e = lane; e < 2 + q; e += kWave) {
      double v;
      if (e == 0)
        v = R[cy * KP + km];
      else if (e <= q)
        v = R[(cy + e) * KP + km];
      else
        v = S[km * LD + km] * S[km * LD + km];   // L_KK^2
      rec[e] = v;
    }
    // GLS normal equations
    for (int e = lane; e < q * q + q; e += kWave) {
      if (e < q * q) {
        int c1 = e / q, c2 = e - c1 * q;
        acc[X0 + e] += R[(cy + 1 + c1) * KP + km] * R[(cy + 1 + c2) * KP + km];
      } else {
        int c = e - q * q;
        acc[Y0 + c] += R[(cy + 1 + c) * KP + km] * R[cy * KP + km];
      }
    }
    if (a.need_grad) {
      const int nF = P * (P + 1) / 2;
      const int ntask = nF + P + P * q;
      for (int e = lane; e < P; e += kWave) rec[2 + q + e] = R[e * KP + km];
      for (int e = lane; e < ntask; e += kWave) {
        int c1, c2;
        if (e < nF) {
          pair_of(e, c1, c2);  // c1 >= c2
        } else if (e < nF + P) {
          c1 = e - nF;
          c2 = cy;
        } else {
          int r = e - nF - P;
          c1 = r / q;
          c2 = cy + 1 + (r - c1 * q);
        }
        const double* w1 = R + c1 * KP;
        const double* w2 = R + c2 * KP;
        double s = 0.0;
        for (int i = 0; i < K; ++i) s += w1[i] * w2[i];
        if (e < nF) {
          double f = s - 0.5 * w1[km] * w2[km];
          acc[F0 + c1 * P + c2] += f;
          if (c1 != c2) acc[F0 + c2 * P + c1] += f;
        } else if (e < nF + P) {
          rec[2 + q + P + c1] = s;
        } else {
          int r = e - nF - P;
          rec[2 + q + 2 * P + r] = s;
        }
      }
    }
  }
  __syncthreads();
  for (int e = lane; e < NACC; e += kWave) a.partial[(int64_t)blockIdx.x * NACC + e] = acc[e];
}

// ---------------------------------------------------------------- pass 1, split rows
// One conditioning block per wave for K <= 32, d <= 10, q <= 11.  Lane (r, h)
// (r = lane & 31, h = lane >> 5) owns row r, columns 16h..16h+15 of the block
// matrix in registers; rows K..31 are identity padding.
//  * Pairs are split cyclically (row r takes (r, r+t mod 32), t = 0..16; half h
//    takes t in [9h, 9h+9)), so each exp/sqrt runs once; Sigma goes to the lower
//    triangle of the LDS scratch Mt, g(r) to the upper triangle.
//  * Cholesky in column-scaled form: L'_ij = L_ij / L_jj (unit lower) and pivots
//    p_j = L_jj^2; a step costs one reciprocal, no square root.  L' is stored
//    column-major to LDS (Lc) as it is produced.
//  * Right-hand sides live split by half: half 0 holds y and Z (slots 0..11),
//    half 1 the derivative columns dS_p a (slot 0 sigma^2, 1..10 ranges,
//    11 nugget).  The forward solve runs on the unit factor and scales rows by
//    1/L_rr = rsqrt(p_r) at the end.
constexpr int kSK = 32, kSH = 16, kSD = 10, kSQ = 11, kSC = 12;

__device__ __forceinline__ double rcp_nr(double x) {
  double y = __builtin_amdgcn_rcp(x);
  double e = fma(-x, y, 1.0);
  y = fma(y, e, y);
  e = fma(-x, y, 1.0);
  return fma(y, e, y);
}

__device__ __forceinline__ double rsqrt_nr(double x) {
  double y = __builtin_amdgcn_rsq(x);
  double h = 0.5 * x;
  y = y * fma(-h * y, y, 1.5);
  y = y * fma(-h * y, y, 1.5);
  return y * fma(-h * y, y, 1.5);
}

template <int NU>
__device__ __forceinline__ void matern_t(double r, double& M, double& g, int nu_rt = 2) {
  if constexpr (NU < 0) {
    if (nu_rt == 0)
      matern_t<0>(r, M, g);
    else if (nu_rt == 1)
      matern_t<1>(r, M, g);
    else
      matern_t<2>(r, M, g);
  } else if constexpr (NU == 0) {
    if (r > 0.0) {
      double e = exp(-r);
      M = e;
      g = -e / r;
    } else {
      M = 1.0;
      g = 0.0;
    }
  } else if constexpr (NU == 1) {
    double c = 1.7320508075688772 * r;
    double e = exp(-c);
    M = (1.0 + c) * e;
    g = -3.0 * e;
  } else {
    double c = 2.23606797749979 * r;
    double e = exp(-c);
    M = (1.0 + c + c * c * (1.0 / 3.0)) * e;
    g = -(5.0 / 3.0) * (1.0 + c) * e;
  }
}

// A double constant built in scalar registers where it is used (two scalar
// moves in a volatile asm, BITS = its IEEE bit pattern): neither hoisted into
// vector registers nor kept live across the kernel.
template <unsigned long long BITS>
__device__ __forceinline__ double sconst() {
  unsigned lo, hi;
  asm volatile("s_mov_b32 %0, %2\n\ts_mov_b32 %1, %3"
               : "=s"(lo), "=s"(hi)
               : "i"((unsigned)(BITS & 0xffffffffull)), "i"((unsigned)(BITS >> 32)));
  return __hiloint2double((int)hi, (int)lo);
}

// a * b + c with c in scalar registers as one v_fma_f64 (left to itself the
// compiler copies c into vector registers for a two-operand v_fmac_f64).
__device__ __forceinline__ double fma_s(double a, double b, double c) {
  double r;
  asm("v_fma_f64 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "s"(c));
  return r;
}

// e^-c for c >= 0: n = rint(-c log2 e), f = -c - n ln 2 with ln 2 in two
// parts, e^f by its Taylor polynomial to degree 13 (|f| <= ln 2 / 2, so the
// truncation is below 5e-18 relative), scaled by 2^n.  Past c = 746 the result
// is below the smallest subnormal: the clamp lets ldexp return 0.
__device__ __forceinline__ double exp_neg(double c) {
  const double x = -fmin(c, 746.0);
  const double n = rint(x * sconst<0x3FF71547652B82FEull>());        // log2 e
  double f = fma(-n, sconst<0x3FE62E42FEFA39EFull>(), x);            // ln 2 (high part)
  f = fma(-n, sconst<0x3C7ABC9E3B39803Full>(), f);                   // ln 2 (low part)
  double p = fma_s(sconst<0x3DE6124613A86D09ull>(), f, sconst<0x3E21EED8EFF8D898ull>());  // 1/13!, 1/12!
  p = fma_s(p, f, sconst<0x3E5AE64567F544E4ull>());                  // 1/11!
  p = fma_s(p, f, sconst<0x3E927E4FB7789F5Cull>());                  // 1/10!
  p = fma_s(p, f, sconst<0x3EC71DE3A556C734ull>());                  // 1/9!
  p = fma_s(p, f, sconst<0x3EFA01A01A01A01Aull>());                  // 1/8!
  p = fma_s(p, f, sconst<0x3F2A01A01A01A01Aull>());                  // 1/7!
  p = fma_s(p, f, sconst<0x3F56C16C16C16C17ull>());                  // 1/6!
  p = fma_s(p, f, sconst<0x3F81111111111111ull>());                  // 1/5!
  p = fma_s(p, f, sconst<0x3FA5555555555555ull>());                  // 1/4!
  p = fma_s(p, f, sconst<0x3FC5555555555555ull>());                  // 1/3!
  p = fma(p, f, 0.5);
  p = fma(p, f, 1.0);
  p = fma(p, f, 1.0);
  return ldexp(p, (int)n);
}

// sqrt(a) for a >= 0 from the hardware reciprocal square root and two
// correction steps (Goldschmidt iteration: g -> sqrt(a), h -> 1/(2 sqrt(a)));
// inputs below 2^-900 (squared distances of coincident points) give 0.
__device__ __forceinline__ double sqrt_pos(double a) {
  const double y = __builtin_amdgcn_rsq(a);
  double g = a * y, h = 0.5 * y;
  const double r = fma(-g, h, 0.5);
  g = fma(g, r, g);
  h = fma(h, r, h);
  double e = fma(-g, g, a);
  g = fma(e, h, g);
  e = fma(-g, g, a);
  g = fma(e, h, g);
  return (a > 0x1p-900) ? g : 0.0;
}

// Matern M(r) and g = M'(r)/r at r = sqrt(s2) for the closed-form
// smoothness values, with exp_neg and sqrt_pos (the rows kernel's covariance
// pairs).
template <int NU>
__device__ __forceinline__ void matern_s2(double s2, double& M, double& g, int nu_rt = 2) {
  if constexpr (NU < 0) {
    if (nu_rt == 0)
      matern_s2<0>(s2, M, g);
    else if (nu_rt == 1)
      matern_s2<1>(s2, M, g);
    else
      matern_s2<2>(s2, M, g);
  } else {
    const double r = sqrt_pos(s2);
    if constexpr (NU == 0) {
      const double e = exp_neg(r);
      M = e;
      g = (r > 0.0) ? -e / r : 0.0;
    } else if constexpr (NU == 1) {
      const double c = 1.7320508075688772 * r;
      const double e = exp_neg(c);
      M = (1.0 + c) * e;
      g = -3.0 * e;
    } else {
      const double c = 2.23606797749979 * r;
      const double e = exp_neg(c);
      M = (1.0 + c + c * c * (1.0 / 3.0)) * e;
      g = -(5.0 / 3.0) * (1.0 + c) * e;
    }
  }
}

// Matern M and g = M'/r in float32 (the mixed mode's covariance assembly and
// derivative terms): the same closed forms, hardware exp2 and sqrt.
template <int NU>
__device__ __forceinline__ void matern_f(float r, float& M, float& g, int nu_rt = 2) {
  if constexpr (NU < 0) {
    if (nu_rt == 0)
      matern_f<0>(r, M, g);
    else if (nu_rt == 1)
      matern_f<1>(r, M, g);
    else
      matern_f<2>(r, M, g);
  } else if constexpr (NU == 0) {
    if (r > 0.0f) {
      const float e = __expf(-r);
      M = e;
      g = -e / r;
    } else {
      M = 1.0f;
      g = 0.0f;
    }
  } else if constexpr (NU == 1) {
    const float c = 1.7320508f * r;
    const float e = __expf(-c);
    M = (1.0f + c) * e;
    g = -3.0f * e;
  } else {
    const float c = 2.2360680f * r;
    const float e = __expf(-c);
    M = (1.0f + c + c * c * (1.0f / 3.0f)) * e;
    g = -(5.0f / 3.0f) * (1.0f + c) * e;
  }
}

// 16x16x4 matrix-core products.  Operands: lane L carries A[L & 15][L >> 4]
// and B[L >> 4][L & 15].  FP64 (v_mfma_f64_16x16x4f64): result register v of
// lane L is D[(L >> 4) + 4v][L & 15].  FP32 (v_mfma_f32_16x16x4f32, the mixed
// mode): D[4 (L >> 4) + v][L & 15].
typedef double dvec4 __attribute__((ext_vector_type(4)));
typedef float fvec4 __attribute__((ext_vector_type(4)));
template <bool MIX>
struct Mfma;
template <>
struct Mfma<false> {
  typedef dvec4 acc_t;
  static __device__ __forceinline__ acc_t zero() { return acc_t{0.0, 0.0, 0.0, 0.0}; }
  static __device__ __forceinline__ acc_t mma(double x, double y, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(x, y, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lk, int v) { return lk + 4 * v; }
};
template <>
struct Mfma<true> {
  typedef fvec4 acc_t;
  static __device__ __forceinline__ acc_t zero() { return acc_t{0.f, 0.f, 0.f, 0.f}; }
  static __device__ __forceinline__ acc_t mma(double x, double y, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32((float)x, (float)y, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lk, int v) { return 4 * lk + v; }
};

// LDS doubles per wave: Cb 32 | Bc 24 | Misc 8 | Lc 32*32 | Xs 32*d | Mt 32*33
__host__ __device__ inline int split_lds_doubles(int d) {
  return 32 + 24 + 8 + 32 * 32 + 32 * d + 32 * 33;
}

// One block of the split-row kernel (body shared by the single-structure and
// the batched kernels).  Accumulates this block's Fisher / GLS contributions
// into the lane's task accumulators and writes the block record.
// PROF adds per-phase shader-clock sums (s_memtime) for tuning runs.
#define SVH_PHASE(k)                                            \
  if constexpr (PROF) {                                         \
    const unsigned long long t_ = __builtin_amdgcn_s_memtime(); \
    ph[k] += t_ - ph_t;                                         \
    ph_t = t_;                                                  \
  }

template <int NU, bool PROF, bool MIX>
__device__ __forceinline__ void split_block(const LLArgs& a, int64_t b, double* lds,
                                            double (&accF)[4], double (&accX)[3],
                                            unsigned long long (&ph)[8]) {
  unsigned long long ph_t = 0ull;
  if constexpr (PROF) ph_t = __builtin_amdgcn_s_memtime();
  const int lane = threadIdx.x, h = lane >> 5, r = lane & 31;
  const int d = a.d, q = a.q, P = a.P;
  const bool grad = a.need_grad != 0;
  double* Cb = lds;          // one column of the block matrix / the vector a
  double* Bc = Cb + 32;      // one row of right-hand sides (24 slots)
  double* Misc = Bc + 24;    // scalars of row K-1
  double* Lc = Misc + 8;     // L' column-major: Lc[j*32 + i] = L'_ij (i > j)
  double* Xs = Lc + 32 * 32; // 32 x d scaled coordinates
  double* Mt = Xs + 32 * d;  // 32 x 33 scratch matrix
  const int nX = q * q + q;
  const int dlo = h ? (d + 1) / 2 : 0, dhi = h ? d : (d + 1) / 2;   // this half's dims
  const int K = a.kcnt[b] + 1;
  const int ch = a.chunk[b];
  const int lev = a.level[ch];
  const bool act = r < K;
  int pt = 0;
  if (act) pt = (r < K - 1) ? a.nbr[b * a.mmax + r] : a.rows[b];
  __syncthreads();
  for (int l = dlo; l < dhi; ++l)
    Xs[r * d + l] = act ? a.X[(int64_t)pt * d + l] * a.inv_lam[l] : 0.0;
  __syncthreads();
  double xr[kSD];
#pragma unroll
  for (int l = 0; l < kSD; ++l) xr[l] = (l < d) ? Xs[r * d + l] : 0.0;
  SVH_PHASE(0)
  // ---- pairs: Sigma lower, g upper
#pragma unroll 1
  for (int t = 9 * h; t < 9 * h + 9 && t <= 16; ++t) {
    const int c = (r + t) & 31;
    double Mv = 0.0, gg = 0.0;
    if (t == 0) {
      Mv = act ? a.var * (1.0 + a.nug) : 1.0;
    } else if (act && c < K) {
      double s2 = 0.0;
#pragma unroll
      for (int l = 0; l < kSD; ++l)
        if (l < d) {
          const double dd = xr[l] - Xs[c * d + l];
          s2 = fma(dd, dd, s2);
        }
      matern_t<NU>(sqrt(s2), Mv, gg, a.nu_code);
      Mv *= a.var;
    }
    const int hi = r > c ? r : c, lo = r > c ? c : r;
    Mt[hi * 33 + lo] = Mv;
    if (t != 0) Mt[lo * 33 + hi] = gg;
  }
  __syncthreads();
  double S[kSH];
#pragma unroll
  for (int cc = 0; cc < kSH; ++cc) {
    const int c = kSH * h + cc;
    S[cc] = (c <= r) ? Mt[r * 33 + c] : Mt[c * 33 + r];
  }
  SVH_PHASE(1)
  // ---- chunk jitter (rare): Sigma += 10^p tr(Sigma)/K I on the real rows
  double jit = 0.0;
  if (lev > 0) {
    __syncthreads();
    if (h == 0) Cb[r] = act ? Mt[r * 33 + r] : 0.0;
    __syncthreads();
    double tr = 0.0;
    for (int t = 0; t < K; ++t) tr += Cb[t];
    jit = kJitDev[lev] * (tr / K);
#pragma unroll
    for (int cc = 0; cc < kSH; ++cc)
      if (kSH * h + cc == r && act) S[cc] += jit;
  }
  // ---- Cholesky, column-scaled
  bool fail = false;
  double piv = 1.0;
#pragma unroll
  for (int k = 0; k < kSK; ++k) {
    const int hk = k / kSH, kc = k % kSH;
    __syncthreads();
    if (h == hk) Cb[r] = S[kc];
    __syncthreads();
    const double pk = Cb[k];
    fail |= (k < K) && !(pk > 0.0);
    const double lp = Cb[r] * rcp_nr(pk);
    if (r == k) piv = pk;
    Lc[k * 32 + r] = lp;   // column k of L' (rows <= k hold junk, never read)
    // trailing update of this lane's columns c > k (columns <= k hold junk)
#pragma unroll
    for (int cc = 0; cc < kSH; ++cc) {
      if (k >= kSH && cc <= k - kSH) continue;   // no lane owns a live column
      S[cc] = fma(-lp, Cb[kSH * h + cc], S[cc]);
    }
    // keep each step's update in place: without this the scheduler defers
    // the FMAs across steps and holds every step's operands in registers
#pragma unroll
    for (int cc = 0; cc < kSH; ++cc) asm volatile("" : "+v"(S[cc]));
  }
  const double dinv = rsqrt_nr(piv);   // 1 / L_rr
  if (r == K - 1 && h == 0) {
    Misc[0] = dinv;
    Misc[1] = piv;
    Misc[2] = piv;   // L_KK^2
  }
  __syncthreads();
  SVH_PHASE(2)
  // ---- a = L^-T e_K:  L'^T at = e_K, a = at / L_KK  (both halves, same values)
  double at = (r == K - 1) ? 1.0 : 0.0;
#pragma unroll 4
  for (int i = kSK - 1; i >= 0; --i) {
    const double ai = __shfl(at, i, 32);
    if (r < i) at = fma(-Lc[r * 32 + i], ai, at);
  }
  const double dK = Misc[0];
  const double pK = Misc[1];
  const double av = at * dK;
  if (h == 0) Cb[r] = av;
  __syncthreads();
  SVH_PHASE(3)
  // ---- right-hand sides
  double R[kSC];
#pragma unroll
  for (int c = 0; c < kSC; ++c) R[c] = 0.0;
  if (grad) {
    // dS_{1+l} a = -var sum_j g_rj a_j (x_rl - x_jl)^2.  With coordinates
    // centred on the block's own point (x~ = x - x_{K-1}) this is
    //   -var [x~_rl^2 E_r0 - 2 x~_rl E_r(1+l) + E_r(1+d+l)],  E = G B,
    // B = [a | a x~_l | a x~_l^2]: a 32 x 32 x (1 + 2d) product on the
    // matrix cores (2 x 2 output tiles, 8 steps of 4).
    typedef Mfma<MIX> MF;
    const int li = lane & 15, lk = lane >> 4;
    const double* xK = Xs + (K - 1) * d;
    const int NB = 1 + 2 * d;
    // row tiles one at a time (16 accumulator registers live, not 32); the
    // products go to registers first and to Mt only after every g is read
    typename MF::acc_t Et[2][2];
#pragma unroll
    for (int rt = 0; rt < 2; ++rt) {
      typename MF::acc_t A0 = MF::zero(), A1 = MF::zero();
#pragma unroll 1
      for (int st = 0; st < 8; ++st) {
        const int j = 4 * st + lk;
        const int rw = 16 * rt + li;
        const double gv = (j == rw) ? 0.0 : (j < rw ? Mt[j * 33 + rw] : Mt[rw * 33 + j]);
        const double aj = Cb[j];
        double b0, b1;
        {
          const int c = li;   // tile 0: a, a x~_l, a x~_l^2 ...
          const int l = (c <= d) ? c - 1 : c - 1 - d;
          const double xt = (c >= 1 && c <= 2 * d) ? Xs[j * d + l] - xK[l] : 0.0;
          b0 = (c == 0) ? aj : (c <= d ? aj * xt : (c <= 2 * d ? aj * xt * xt : 0.0));
        }
        {
          const int c = 16 + li;
          const int l = (c <= d) ? c - 1 : c - 1 - d;
          const double xt = (c <= 2 * d) ? Xs[j * d + l] - xK[l] : 0.0;
          b1 = (c <= d) ? aj * xt : (c <= 2 * d ? aj * xt * xt : 0.0);
        }
        A0 = MF::mma(gv, b0, A0);
        A1 = MF::mma(gv, b1, A1);
      }
      Et[rt][0] = A0;
      Et[rt][1] = A1;
    }
    __syncthreads();   // every read of g is done: E goes to Mt, column-major
#pragma unroll
    for (int rt = 0; rt < 2; ++rt)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int rr = 16 * rt + MF::row(lk, v);
        if (li < NB) Mt[li * 33 + rr] = (double)Et[rt][0][v];
        if (16 + li < NB) Mt[(16 + li) * 33 + rr] = (double)Et[rt][1][v];
      }
    __syncthreads();
    if (h == 1) {
      const double e0 = Mt[r];
#pragma unroll
      for (int l = 0; l < kSD; ++l)
        if (l < d) {
          const double xt = Xs[r * d + l] - xK[l];
          const double v = xt * (xt * e0 - 2.0 * Mt[(1 + l) * 33 + r]) + Mt[(1 + d + l) * 33 + r];
          R[1 + l] = -a.var * v;
        }
      R[0] = ((r == K - 1) ? pK * dK : 0.0) - jit * av;
      R[kSC - 1] = a.var * a.nug * av;
    }
  }
  if (h == 0 && act) {
    R[0] = a.y[pt];
#pragma unroll
    for (int c = 0; c < kSQ; ++c)
      if (c < q) R[1 + c] = a.Z[(int64_t)pt * q + c];
  }
  SVH_PHASE(4)
  // ---- forward solve with the unit factor L', then scale by 1/L_rr
  const int ncol = grad ? 2 : 1;   // halves that carry columns
#pragma unroll 1
  for (int j = 0; j < kSK - 1; ++j) {
    __syncthreads();
    if (r == j && h < ncol) {
#pragma unroll
      for (int c = 0; c < kSC; ++c) Bc[kSC * h + c] = R[c];
    }
    __syncthreads();
    if (h < ncol) {
      const double wj = (j < r) ? Lc[j * 32 + r] : 0.0;   // L'_rj
#pragma unroll
      for (int c = 0; c < kSC; ++c) R[c] = fma(-wj, Bc[kSC * h + c], R[c]);
    }
  }
#pragma unroll
  for (int c = 0; c < kSC; ++c) R[c] *= dinv;
  const bool ok = !fail;
  if (fail && lane == 0) atomicOr(a.fail + ch, 1);
  __syncthreads();
  SVH_PHASE(5)
  // ---- columns to LDS by parameter index: W_p (p < P), then yt, Zt_c
  const int cyl = grad ? P : 0;
  if (h == 1 && grad) {
    Mt[0 * 33 + r] = R[0];
#pragma unroll
    for (int l = 0; l < kSD; ++l)
      if (l < d) Mt[(1 + l) * 33 + r] = R[1 + l];
    Mt[(d + 1) * 33 + r] = R[kSC - 1];
  }
  if (h == 0) {
    Mt[cyl * 33 + r] = R[0];
#pragma unroll
    for (int c = 0; c < kSQ; ++c)
      if (c < q) Mt[(cyl + 1 + c) * 33 + r] = R[1 + c];
  }
  __syncthreads();
  const int km = K - 1;
  double* rec = a.cache + b * a.rec_stride;
  for (int e = lane; e < 2 + q + (grad ? P : 0); e += kWave) {
    double v;
    if (e == 0)
      v = Mt[cyl * 33 + km];
    else if (e <= q)
      v = Mt[(cyl + e) * 33 + km];
    else if (e == q + 1)
      v = Misc[2];
    else
      v = Mt[(e - q - 2) * 33 + km];
    if (ok) rec[e] = v;
  }
  if (grad) {
    // D = W^T [W | yt | Zt] on the matrix cores: one 16-row tile (p < P),
    // two 16-column tiles, 8 steps of 4 rows.  Column c of the scratch is
    // both A row c and B column c.
    typedef Mfma<MIX> MF;
    const int li = lane & 15, lk = lane >> 4;
    typename MF::acc_t D0 = MF::zero(), D1 = MF::zero();
#pragma unroll 2
    for (int st = 0; st < 8; ++st) {
      const int i = 4 * st + lk;
      const double w0 = Mt[li * 33 + i];
      const double w1 = Mt[(16 + li) * 33 + i];
      D0 = MF::mma(w0, w0, D0);
      D1 = MF::mma(w0, w1, D1);
    }
    const int N = cyl + 1 + q;
    const double wKc0 = Mt[li * 33 + km];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int p = MF::row(lk, v);
      if (p < P) {
        const double d0 = (double)D0[v], d1 = (double)D1[v];
        if (li < P) {
          const double f = d0 - 0.5 * Mt[p * 33 + km] * wKc0;
          if (ok) accF[v] += f;
        } else if (li == cyl) {
          if (ok) rec[2 + q + P + p] = d0;
        } else if (li < N) {
          if (ok) rec[2 + q + 2 * P + p * q + (li - cyl - 1)] = d0;
        }
        if (16 + li < N && ok) rec[2 + q + 2 * P + p * q + (16 + li - cyl - 1)] = d1;
      }
    }
  }
  for (int e = lane; e < nX; e += kWave) {
    double v;
    if (e < q * q) {
      const int c1 = e / q, c2 = e - c1 * q;
      v = Mt[(cyl + 1 + c1) * 33 + km] * Mt[(cyl + 1 + c2) * 33 + km];
    } else {
      const int c = e - q * q;
      v = Mt[(cyl + 1 + c) * 33 + km] * Mt[cyl * 33 + km];
    }
    if (ok) {
      if (e < kWave)
        accX[0] += v;
      else if (e < 2 * kWave)
        accX[1] += v;
      else
        accX[2] += v;
    }
  }
  SVH_PHASE(6)
}

// Fisher partial sums of one workgroup: lane L accumulated entries
// (Mfma row(L >> 4, v), L & 15) of the P x P matrix.
template <bool MIX>
__device__ __forceinline__ void write_fisher(double* part, const double (&accF)[4], int P,
                                             int lane) {
  const int li = lane & 15, lk = lane >> 4;
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int p = Mfma<MIX>::row(lk, v);
    if (p < P && li < P) part[p * P + li] = accF[v];
  }
}

template <int NU, bool PROF, bool MIX>
__global__ __launch_bounds__(kWave, 2) void ll_pass1_split(LLArgs a) {
  extern __shared__ double lds[];
  const int lane = threadIdx.x;
  const int P = a.P, q = a.q;
  const bool grad = a.need_grad != 0;
  const int nX = q * q + q;
  double accF[4] = {0.0, 0.0, 0.0, 0.0};   // Fisher entries of this lane
  double accX[3] = {0.0, 0.0, 0.0};  // GLS tasks lane, +64, +128 (nX <= 132)
  unsigned long long ph[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int64_t b = blockIdx.x; b < a.nblk; b += gridDim.x)
    split_block<NU, PROF, MIX>(a, b, lds, accF, accX, ph);
  if (PROF && lane == 0)
    for (int k = 0; k < 7; ++k) atomicAdd(a.prof + k, ph[k]);
  // this workgroup's partials (layout of ll_pass1: F (P*P), XtX (q*q), Zty (q))
  const int NACC = grad ? P * P + q * q + q : q * q + q;
  const int X0 = grad ? P * P : 0;
  double* part = a.partial + (int64_t)blockIdx.x * NACC;
  if (grad) write_fisher<MIX>(part, accF, P, lane);
  for (int e = lane; e < nX; e += kWave)
    part[X0 + e] = (e < kWave) ? accX[0] : (e < 2 * kWave ? accX[1] : accX[2]);
}

template <int NU>
void launch_split_nu(hipStream_t st, int G, size_t lds, const LLArgs& a, bool mixed) {
  if (a.prof)
    hipLaunchKernelGGL((ll_pass1_split<NU, true, false>), dim3(G), dim3(kWave), lds, st, a);
  else if (mixed)
    hipLaunchKernelGGL((ll_pass1_split<NU, false, true>), dim3(G), dim3(kWave), lds, st, a);
  else
    hipLaunchKernelGGL((ll_pass1_split<NU, false, false>), dim3(G), dim3(kWave), lds, st, a);
}

// ---------------------------------------------------------------- pass 1, rows in lanes
// Two conditioning blocks per wave (K <= 32, d <= 10, q <= 11): half
// h = lane >> 5 takes block 2w + h, lane r = lane & 31 owns row r of it (rows
// K..31 are identity padding).  All per-row state lives in registers:
//   * coordinates (DM >= d, zero padded) and the Sigma row, then the factor row;
//   * the factorisation is right-looking in column-scaled form (unit lower L',
//     pivots p_r = L_rr^2): step k broadcasts the current column k and the
//     finished right-hand-side row z_k of [y Z] through LDS (one write, broadcast
//     reads; a one-wave workgroup needs no barrier), so L^-1 [y Z] comes out
//     of the factorisation itself;
//   * t = L'^-T e_{K-1} by 31 half-wave sum reductions (rows hold their own
//     column entries l'_rc), a = t / L_KK;
//   * dS_p a row by row: each lane recomputes its row's distances and g(r)
//     (no matrix of g is kept) -- sigma^2: L e_K - jitter a; ranges:
//     -var sum_c g_rc (x_rl - x_cl)^2 a_c; nugget: var tau a;
//   * W = L^-1 (dS a) by a column sweep, then the products W^T [W | L^-1 y |
//     L^-1 Z] on the matrix cores, block by block (as in the split kernel).

// The rows kernels run one wave per workgroup, and a wave's LDS operations
// complete in issue order: exchanging values between its lanes through LDS
// needs only a compiler fence, not a workgroup barrier with its wait for every
// outstanding LDS operation.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// A double moved between lanes by a DPP control (both dwords), all lanes
// active and every source lane valid (no old value to keep).
template <int CTRL>
__device__ __forceinline__ double dpp_double(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffll), CTRL, 0xF, 0xF, true);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, true);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// Sum over each 16-lane row, every lane gets its row's sum: quad permutes
// (lane ^ 1, lane ^ 2) and rotations by 4 and 8 inside the row.
__device__ __forceinline__ double row_sum(double v) {
  v += dpp_double<0xB1>(v);    // quad_perm [1,0,3,2]
  v += dpp_double<0x4E>(v);    // quad_perm [2,3,0,1]
  v += dpp_double<0x124>(v);   // row_ror:4
  v += dpp_double<0x128>(v);   // row_ror:8
  return v;
}

// Sum over each half-wave (32 lanes), every lane gets its half's sum: the row
// sums, then the row pair swap of v_permlane16_swap (rows 0/1 and 2/3),
// whose two outputs summed give row 0 + row 1 (row 2 + row 3) in every lane
// of the pair.
__device__ __forceinline__ double half_sum(double v) {
  v = row_sum(v);
  const long long b = __double_as_longlong(v);
  const unsigned lo = (unsigned)(b & 0xffffffffll), hi = (unsigned)(b >> 32);
  const auto sl = __builtin_amdgcn_permlane16_swap(lo, lo, false, false);
  const auto sh = __builtin_amdgcn_permlane16_swap(hi, hi, false, false);
  const double a0 = __longlong_as_double(((long long)sh[0] << 32) | sl[0]);
  const double a1 = __longlong_as_double(((long long)sh[1] << 32) | sl[1]);
  return a0 + a1;
}

// LDS doubles per wave of the rows kernel: per half [ab 32 | xs 32*dm |
// squared centred coordinates 32*dm], then
// the packed lower triangles (diagonal included) of Sigma, later g(r), of
// both halves (2 x 528; the product operands Mt, 24 x 33, reuse that space
// once g is consumed), then misc (16).
constexpr int kGTri = 528;   // 32 * 33 / 2 lower entries
// first coordinate of padding rows (scaled inputs are at most 1e8 in size)
constexpr double kPadFar = 1e17;
// Coordinate registers of the rows kernel for d inputs and q trend columns:
// the kernel carries DM + 2 right-hand-side columns, which must hold both
// [y Z] (1 + q) and the d + 2 derivative columns.
__host__ __device__ inline int rows_dm(int d, int q) {
  const int k = d > q - 1 ? d : q - 1;
  return k <= 4 ? 4 : (k <= 8 ? 8 : 10);
}
__host__ __device__ inline int rows_half_doubles(int dm) { return 32 + 64 * dm; }
__host__ __device__ inline int rows_lds_doubles(int dm) {
  return 2 * rows_half_doubles(dm) + 2 * kGTri + 16;
}
// Entry (i, j), i >= j, of a lower 32 x 32 triangle packed by columns:
// column j holds rows j..31 contiguously, so g_base(j) + i.  For i < j the
// same expression stays inside the array (0 .. 526): another entry.
__host__ __device__ constexpr int g_base(int j) { return 31 * j - (j * (j - 1)) / 2; }
// The same for run-time indices 0 <= j <= 31 (24-bit multiply, full rate).
__device__ __forceinline__ int g_base_rt(int j) { return (int)(__umul24((unsigned)j, (unsigned)(63 - j)) >> 1); }
// Coordinate l of point p of a block in its 32 x dm slot: pairs of
// coordinates interleaved by point, so a lane's two coordinates are one
// 16-byte read and 32 lanes reading 32 points touch distinct banks.
__host__ __device__ constexpr int xs_at(int l, int p) { return (l >> 1) * 64 + 2 * p + (l & 1); }

// ---- fused lane broadcasts
// acc -= src[lane N of this lane's 16-lane row] * m as one v_fmac_f64 with a
// DPP row_newbcast source operand (the only DPP control that 64-bit VALU
// operations take).  The compiler neither forms this instruction nor tracks
// the DPP read-after-VALU-write hazard through inline assembly, so each
// statement starts with the two wait states that hazard needs.  Callers run
// them with EXEC unchanged (no divergent branches around them).
#define SVH_BF(acc, src, m, n) \
  "v_fmac_f64_dpp %" #acc ", %" #src ", -%" #m " row_newbcast:%" #n " row_mask:0xf bank_mask:0xf bound_ctrl:1"
template <int N>
__device__ __forceinline__ void bf1(double& a0, double s0, double m) {
  asm volatile("s_nop 1\n\t" SVH_BF(0, 1, 2, 3) : "+v"(a0) : "v"(s0), "v"(m), "i"(N));
}
// four accumulators, one source, four source lanes
template <int N0, int N1, int N2, int N3>
__device__ __forceinline__ void bf4_lanes(double& a0, double& a1, double& a2, double& a3,
                                          double s, double m) {
  asm volatile("s_nop 1\n\t" SVH_BF(0, 4, 5, 6) "\n\t" SVH_BF(1, 4, 5, 7) "\n\t" SVH_BF(
                   2, 4, 5, 8) "\n\t" SVH_BF(3, 4, 5, 9)
               : "+v"(a0), "+v"(a1), "+v"(a2), "+v"(a3)
               : "v"(s), "v"(m), "i"(N0), "i"(N1), "i"(N2), "i"(N3));
}
// four accumulators, each its own source, one source lane
template <int N>
__device__ __forceinline__ void bf4_self(double& a0, double& a1, double& a2, double& a3,
                                         double m) {
  asm volatile("s_nop 1\n\t" SVH_BF(0, 0, 4, 5) "\n\t" SVH_BF(1, 1, 4, 5) "\n\t" SVH_BF(
                   2, 2, 4, 5) "\n\t" SVH_BF(3, 3, 4, 5)
               : "+v"(a0), "+v"(a1), "+v"(a2), "+v"(a3)
               : "v"(m), "i"(N));
}
// four accumulators, four separate sources, one source lane
template <int N>
__device__ __forceinline__ void bf4_src(double& a0, double& a1, double& a2, double& a3,
                                        double s0, double s1, double s2, double s3, double m) {
  asm volatile("s_nop 1\n\t" SVH_BF(0, 4, 8, 9) "\n\t" SVH_BF(1, 5, 8, 9) "\n\t" SVH_BF(
                   2, 6, 8, 9) "\n\t" SVH_BF(3, 7, 8, 9)
               : "+v"(a0), "+v"(a1), "+v"(a2), "+v"(a3)
               : "v"(s0), "v"(s1), "v"(s2), "v"(s3), "v"(m), "i"(N));
}
#define SVH_BFF(acc, src, m, n) \
  "v_fmac_f32_dpp %" #acc ", %" #src ", -%" #m " row_newbcast:%" #n " row_mask:0xf bank_mask:0xf bound_ctrl:1"
// two accumulators, each its own source / two separate sources, one lane
template <int N>
__device__ __forceinline__ void bf2_self(double& a0, double& a1, double m) {
  asm volatile("s_nop 1\n\t" SVH_BF(0, 0, 2, 3) "\n\t" SVH_BF(1, 1, 2, 3)
               : "+v"(a0), "+v"(a1)
               : "v"(m), "i"(N));
}
template <int N>
__device__ __forceinline__ void bf2_src(double& a0, double& a1, double s0, double s1, double m) {
  asm volatile("s_nop 1\n\t" SVH_BF(0, 2, 4, 5) "\n\t" SVH_BF(1, 3, 4, 5)
               : "+v"(a0), "+v"(a1)
               : "v"(s0), "v"(s1), "v"(m), "i"(N));
}
// float32 forms (the mixed mode's W sweep)
template <int N>
__device__ __forceinline__ void bf4f_self(float& a0, float& a1, float& a2, float& a3, float m) {
  asm volatile("s_nop 1\n\t" SVH_BFF(0, 0, 4, 5) "\n\t" SVH_BFF(1, 1, 4, 5) "\n\t" SVH_BFF(
                   2, 2, 4, 5) "\n\t" SVH_BFF(3, 3, 4, 5)
               : "+v"(a0), "+v"(a1), "+v"(a2), "+v"(a3)
               : "v"(m), "i"(N));
}
template <int N>
__device__ __forceinline__ void bf4f_src(float& a0, float& a1, float& a2, float& a3, float s0,
                                         float s1, float s2, float s3, float m) {
  asm volatile("s_nop 1\n\t" SVH_BFF(0, 4, 8, 9) "\n\t" SVH_BFF(1, 5, 8, 9) "\n\t" SVH_BFF(
                   2, 6, 8, 9) "\n\t" SVH_BFF(3, 7, 8, 9)
               : "+v"(a0), "+v"(a1), "+v"(a2), "+v"(a3)
               : "v"(s0), "v"(s1), "v"(s2), "v"(s3), "v"(m), "i"(N));
}
template <int N>
__device__ __forceinline__ void bf2f_self(float& a0, float& a1, float m) {
  asm volatile("s_nop 1\n\t" SVH_BFF(0, 0, 2, 3) "\n\t" SVH_BFF(1, 1, 2, 3)
               : "+v"(a0), "+v"(a1)
               : "v"(m), "i"(N));
}
template <int N>
__device__ __forceinline__ void bf2f_src(float& a0, float& a1, float s0, float s1, float m) {
  asm volatile("s_nop 1\n\t" SVH_BFF(0, 2, 4, 5) "\n\t" SVH_BFF(1, 3, 4, 5)
               : "+v"(a0), "+v"(a1)
               : "v"(s0), "v"(s1), "v"(m), "i"(N));
}
#undef SVH_BFF
#undef SVH_BF

// A copy of a lane index that the compiler cannot see through: each phase
// compares its own copy, so lane masks are formed where they are used rather
// than shared (and kept live) across phases.
__device__ __forceinline__ int opaque(int v) {
  asm volatile("" : "+v"(v));
  return v;
}

// The kernel's argument block (its only argument, at offset 0 of the kernarg
// segment) through an address the compiler cannot see to be loop-invariant:
// fields are loaded from the segment by scalar loads where they are used,
// instead of held in scalar registers across the loop over block pairs
// (which spilled them to VGPR lanes and read every 16-register tuple back
// with readlanes at each use).
template <class T>
__device__ __forceinline__ const T& kernarg_fresh() {
  typedef const __attribute__((address_space(4))) char* kptr;
  kptr p = (kptr)__builtin_amdgcn_kernarg_segment_ptr();
#ifdef __OPTIMIZE__
  // (unoptimised builds keep locals in memory, where no scalar register
  // operand exists; the hint matters only to the optimised register plan)
  asm volatile("" : "+s"(p));
#endif
  return *(const T*)(const __attribute__((address_space(4))) T*)p;
}

// v of lane N of this lane's 16-lane row (compiler-managed v_mov_b64 DPP)
template <int N>
__device__ __forceinline__ double row_bc(double v) {
  return __builtin_amdgcn_update_dpp(0.0, v, 0x150 + N, 0xF, 0xF, true);
}

// The first 16-lane row's v in both 16-lane rows of each half
// (v_permlane16_swap: output 0 carries the even rows to both rows of a pair).
__device__ __forceinline__ double from_lo_row(double v) {
  const long long b = __double_as_longlong(v);
  const unsigned lo = (unsigned)(b & 0xffffffffll), hi = (unsigned)(b >> 32);
  const auto sl = __builtin_amdgcn_permlane16_swap(lo, lo, false, false);
  const auto sh = __builtin_amdgcn_permlane16_swap(hi, hi, false, false);
  return __longlong_as_double(((long long)sh[0] << 32) | sl[0]);
}

// Trailing update of step k for columns C..E-1: S[c] -= src(lane c - LOFF) * m
template <int C, int E, int LOFF>
__device__ __forceinline__ void fac_trail(double (&S)[32], double src, double m) {
  if constexpr (E - C >= 4) {
    bf4_lanes<C - LOFF, C + 1 - LOFF, C + 2 - LOFF, C + 3 - LOFF>(S[C], S[C + 1], S[C + 2],
                                                                  S[C + 3], src, m);
    fac_trail<C + 4, E, LOFF>(S, src, m);
  } else if constexpr (E - C >= 1) {
    bf1<C - LOFF>(S[C], src, m);
    fac_trail<C + 1, E, LOFF>(S, src, m);
  }
}

// NC right-hand-side columns (NC even): V[c] -= V[c](lane N) * m, in groups
// of four and a last pair.
template <int N, int C, int NC, class T>
__device__ __forceinline__ void rhs_self_from(T (&V)[NC], T m) {
  if constexpr (NC - C >= 4) {
    if constexpr (sizeof(T) == 8)
      bf4_self<N>(V[C], V[C + 1], V[C + 2], V[C + 3], m);
    else
      bf4f_self<N>(V[C], V[C + 1], V[C + 2], V[C + 3], m);
    rhs_self_from<N, C + 4, NC>(V, m);
  } else if constexpr (NC - C >= 2) {
    if constexpr (sizeof(T) == 8)
      bf2_self<N>(V[C], V[C + 1], m);
    else
      bf2f_self<N>(V[C], V[C + 1], m);
    rhs_self_from<N, C + 2, NC>(V, m);
  }
}
template <int N, int NC, class T>
__device__ __forceinline__ void rhs_self(T (&V)[NC], T m) {
  static_assert(NC % 2 == 0, "even column count");
  rhs_self_from<N, 0, NC>(V, m);
}

// Rows 16..31 of each half take their updates from the finished rows 0..15:
// V[c] -= sum_{j<16} l_rj V_j[c] in order j = 0..15 (l_rj = S[j] there).
// Four (or two) columns C.. at a time: W holds the first rows' values of
// those columns copied to both 16-lane rows.
template <int C, int G, int J, int NC, class T>
__device__ __forceinline__ void rhs_from_lo(T (&V)[NC], const T (&W)[4], double (&S)[32],
                                            double hm) {
  if constexpr (J < 16) {
    asm volatile("" : "+v"(S[J]));
    const T m = (T)(S[J] * hm);   // hm: 1 on rows 16..31, 0 on rows 0..15
    if constexpr (G == 4) {
      if constexpr (sizeof(T) == 8)
        bf4_src<J>(V[C], V[C + 1], V[C + 2], V[C + 3], W[0], W[1], W[2], W[3], m);
      else
        bf4f_src<J>(V[C], V[C + 1], V[C + 2], V[C + 3], W[0], W[1], W[2], W[3], m);
    } else {
      if constexpr (sizeof(T) == 8)
        bf2_src<J>(V[C], V[C + 1], W[0], W[1], m);
      else
        bf2f_src<J>(V[C], V[C + 1], W[0], W[1], m);
    }
    rhs_from_lo<C, G, J + 1, NC>(V, W, S, hm);
  }
}
__device__ __forceinline__ double lo_row_both(double v) { return from_lo_row(v); }
__device__ __forceinline__ float lo_row_both(float v) {
  const auto sw = __builtin_amdgcn_permlane16_swap(__float_as_uint(v), __float_as_uint(v), false,
                                                   false);
  return __uint_as_float(sw[0]);
}
template <int C, int NC, class T>
__device__ __forceinline__ void rhs_lo_to_hi_from(T (&V)[NC], double (&S)[32], double hm) {
  if constexpr (C < NC) {
    constexpr int G = (NC - C >= 4) ? 4 : 2;
    T W[4];
#pragma unroll
    for (int u = 0; u < G; ++u) W[u] = lo_row_both(V[C + u]);
    rhs_from_lo<C, G, 0, NC>(V, W, S, hm);
    __builtin_amdgcn_sched_barrier(0);
    rhs_lo_to_hi_from<C + G, NC>(V, S, hm);
  }
}
template <int NC, class T>
__device__ __forceinline__ void rhs_lo_to_hi(T (&V)[NC], double (&S)[32], double hm) {
  rhs_lo_to_hi_from<0, NC>(V, S, hm);
}

// Step k of the unit LDL^T factorisation of the rows kernel, with the rows of
// L'^-1 [y Z] (bv) carried along.  Lane r holds row r; rows 0..15 of a block
// sit in the first 16-lane row of its half, rows 16..31 in the second.
// Column k (lane c's S[k] = S_ck) reaches the other lanes by DPP row
// broadcasts: for c < 16 from the first row's values copied to both rows
// (one v_permlane16_swap pair per step, k < 16 only), for c >= 16 from the
// second row directly.  Right-hand sides: rows 0..15 update at steps 0..14;
// rows 16..31 take steps 0..15 together after step 15 (rhs_lo_to_hi), then
// steps 16..30 -- the same operations in the same order per row as a plain
// right-looking sweep.
template <int k, int NC>
__device__ __forceinline__ void fac_step(double (&S)[32], double (&bv)[NC], double& pr,
                                         bool& fail, int r, int K, double lo, double hm) {
  double src;
  if constexpr (k < 16)
    src = from_lo_row(S[k]);
  else
    src = S[k];
  const double pk = row_bc<k & 15>(src);   // S_kk (exact in the second row for every k)
  fail |= (k < K) && !(pk > 0.0);
  const double ik = rcp_nr(pk);
  if (r == k) pr = pk;
  // multipliers of rows below k only: rows r <= k get 0, so their strict
  // upper part stays zero and the diagonal entry is cleared (the pivot is
  // kept in pr); later masks reduce to products with S
  const double lrk = S[k] * ((r > k) ? ik : 0.0);
  if constexpr (k < 15) fac_trail<k + 1, 16, 0>(S, src, lrk);
  fac_trail<(k + 1 > 16 ? k + 1 : 16), 32, 16>(S, S[k], lrk);
  S[k] = lrk;
  asm volatile("" : "+v"(S[k]));
  if constexpr (k < 16) {
    rhs_self<k>(bv, lrk * lo);   // rows 16..31 take steps 0..15 after step 15
    if constexpr (k == 15) rhs_lo_to_hi(bv, S, hm);
  } else {
    rhs_self<k - 16>(bv, lrk);
  }
  __builtin_amdgcn_sched_barrier(0);
  if constexpr (k + 1 < 32) fac_step<k + 1>(S, bv, pr, fail, r, K, lo, hm);
}

// W = L'^-1 R for the rows kernel's NC columns (unit lower L' in S), in
// float64 or float32 (mixed mode: the factor's entries rounded to float):
// rows 0..15 with steps 0..14, rows 16..31 with steps 0..15 from the finished
// first rows, then steps 16..30.  The multipliers are formed at their step
// (S[K] passed through an empty asm there): computed ahead, all 31 would be
// live at once.
template <int K, int NC, class T>
__device__ __forceinline__ void wsw_lo(T (&R)[NC], double (&S)[32], double lo) {
  if constexpr (K < 15) {
    asm volatile("" : "+v"(S[K]));
    rhs_self<K>(R, (T)(S[K] * lo));   // S[K] is zero on rows r <= K
    __builtin_amdgcn_sched_barrier(0);
    wsw_lo<K + 1>(R, S, lo);
  }
}
template <int K, int NC, class T>
__device__ __forceinline__ void wsw_hi(T (&R)[NC], double (&S)[32]) {
  if constexpr (K < 31) {
    asm volatile("" : "+v"(S[K]));
    rhs_self<K - 16>(R, (T)S[K]);   // zero on rows r <= K
    __builtin_amdgcn_sched_barrier(0);
    wsw_hi<K + 1>(R, S);
  }
}
template <int NC, class T>
__device__ __forceinline__ void w_sweep_dpp(T (&R)[NC], double (&S)[32], double lo, double hm) {
  wsw_lo<0>(R, S, lo);
  rhs_lo_to_hi(R, S, hm);
  wsw_hi<16>(R, S);
}

// Inputs of one block pair for lane (half h, row r): the block's size, chunk
// and this row's point (PairIdx), then the chunk's jitter level and the row's
// raw coordinates and response (PairX).  The rows kernel loads them ahead --
// the indices two pairs, the values one pair before their pair -- so the
// dependent global loads (kcnt -> nbr -> X, chunk -> level) overlap the work
// of the pairs in between.
struct PairIdx {
  int K, ch, pt;
};
template <int DM>
struct PairX {
  double x[DM];
  double y;
  int lev;
};

__device__ __forceinline__ PairIdx pair_idx(const LLArgs& a, int64_t b0, int lane) {
  const int h = lane >> 5, r = lane & 31;
  const int64_t b = b0 + h;
  PairIdx o{1, 0, 0};
  if (b < a.nblk) {
    const int kc = a.kcnt[b];
    const int rw = a.rows[b];
    const int nb = a.mmax > 0 ? a.nbr[b * a.mmax + (r < a.mmax ? r : a.mmax - 1)] : rw;
    o.K = kc + 1;
    o.ch = a.chunk[b];
    o.pt = r < kc ? nb : rw;
  }
  return o;
}

template <int DM>
__device__ __forceinline__ void pair_x(const LLArgs& a, int64_t b0, const PairIdx& pi, int lane,
                                       PairX<DM>& o) {
  const int h = lane >> 5, r = lane & 31;
  const bool bvalid = b0 + h < a.nblk;
  const bool act = bvalid && r < pi.K;
  const int d = a.d;
#pragma unroll
  for (int l = 0; l < DM; ++l) o.x[l] = (act && l < d) ? a.X[(int64_t)pi.pt * d + l] : 0.0;
  o.y = act ? a.y[pi.pt] : 0.0;
  o.lev = bvalid ? a.level[pi.ch] : 0;
}

template <int NU, bool MIX, int DM, bool PROF>
__device__ __forceinline__ void rows_pair(const LLArgs& a, int64_t b0, double* lds,
                                          double (&accF)[4], double (&accX)[3],
                                          unsigned long long (&ph)[8], const PairIdx& pi,
                                          const PairX<DM>& px) {
  // right-hand-side columns: [y Z] with q <= DM + 1 trend columns, and the
  // P = d + 2 <= DM + 2 derivative columns [sigma^2, ranges, nugget]
  constexpr int NC = DM + 2;
  unsigned long long ph_t = 0;
  if constexpr (PROF) ph_t = __builtin_amdgcn_s_memtime();
  int lane_o = threadIdx.x;
  asm volatile("" : "+v"(lane_o));   // per pair: lane-derived offsets are not kept live across pairs
  const int lane = lane_o, h = lane >> 5, r = lane & 31;
  const int d = a.d, q = a.q, P = a.P;
  const bool grad = a.need_grad != 0;
  double* hb = lds + h * rows_half_doubles(DM);       // this half's region
  double* ab = hb;                                    // a
  double* xs = hb + 32;                               // coordinates of this half's block
  double* gt = lds + 2 * rows_half_doubles(DM);       // packed lower g(r), both halves
  double* gh = gt + h * kGTri;                        // this half's g
  double* misc = gt + 2 * kGTri;                      // per half: [8h] fail, [8h+1..2] pivot
  const int64_t b = b0 + h;   // first block of the pair: b0
  const bool bvalid = b < a.nblk;
  const int K = pi.K;
  const int lev = px.lev;
  const bool act = bvalid && r < K;
  const int pt = act ? pi.pt : 0;
  // Padding rows (r >= K) sit far apart from everything: first coordinate
  // kPadFar * (r + 1), the others 0.  Every pair with a padding point is then
  // at least kPadFar apart, so its covariance and g(r) underflow to exactly 0
  // (e^-c with c clamped, times a finite polynomial; squares stay finite in
  // float32 too) and need no mask.
#pragma unroll
  for (int l = 0; l < DM; ++l) {
    double v = (l < d) ? px.x[l] * a.inv_lam[l] : 0.0;
    if (l == 0) v = act ? v : kPadFar * (double)(r + 1);
    xs[xs_at(l, r)] = v;
  }
  const double yv = px.y;
  const double var = a.var, nug = a.nug;
  double jit = 0.0;
  if (lev > 0) {
    double tr = 0.0;
    for (int t = 0; t < K; ++t) tr += var * (1.0 + nug);
    jit = kJitDev[lev] * (tr / K);
  }
  wave_sync();
  SVH_PHASE(0)
  // ---- Sigma and g(r), each unordered pair once.  The 496 lower pairs of
  // block hh are spread over the whole wave: lane (r, e), e = lane >> 5,
  // evaluates (r, (r + t) & 31) for t = 8e + 1 .. 8e + 8 (t = 16 for r < 16
  // only).  Sigma entries go to the packed triangle of block hh and are read
  // back row by row; g(r) waits in registers and then takes their place
  // (zero for pairs with a padding point).
  double gk[16];
  {
    const int rr = opaque(lane_o & 31), e = lane_o >> 5;
#pragma unroll
    for (int hh = 0; hh < 2; ++hh) {
      const double* xh = lds + hh * rows_half_doubles(DM) + 32;   // block hh's coordinates
      double xi[DM];
#pragma unroll
      for (int l = 0; l < DM; ++l) xi[l] = xh[xs_at(l, rr)];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int t = 8 * e + 1 + u;
        const int c = (rr + t) & 31;
        const int i = rr > c ? rr : c, j = rr > c ? c : rr;
        double v, gs;
        {
          if (MIX && a.sig32) {
            // float32 assembly of the off-diagonal entries (coordinate
            // differences taken in float64, then rounded)
            float s2 = 0.0f;
#pragma unroll
            for (int l = 0; l < DM; ++l) {
              const float dd = (float)(xi[l] - xh[xs_at(l, c)]);
              s2 = fmaf(dd, dd, s2);
            }
            float M, g;
            matern_f<NU>(__fsqrt_rn(s2), M, g, a.nu_code);
            v = var * (double)M;
            gs = (double)g;
          } else {
            double s2 = 0.0;
#pragma unroll
            for (int l = 0; l < DM; ++l) {
              const double dd = xi[l] - xh[xs_at(l, c)];
              s2 = fma(dd, dd, s2);
            }
            double M, g;
            matern_s2<NU>(s2, M, g, a.nu_code);
            v = var * M;
            gs = g;
          }
        }
        // t = 16 (u = 7, upper half of the wave) is stored by rows 0..15
        // only; u < 7 gives t <= 15 and needs no test (no branch)
        if (u < 7 || t < 16 || rr < 16) gt[hh * kGTri + g_base_rt(j) + i] = v;
        gk[8 * hh + u] = gs;
        __builtin_amdgcn_sched_barrier(0);   // keep the iterations apart (register pressure)
      }
    }
  }
  // right-hand sides [y Z] of row r (formed here, not held through the pair
  // phase): trend columns gathered, or formed from the inputs (Z = [1, x] or
  // Z = 1), the inputs read again (an opaque index: not merged with the
  // gather's loads, whose values would then stay live)
  double bv[NC];
  {
    bv[0] = yv;
    const int zmode = a.zmode;
    const int p2 = opaque(pt);
#pragma unroll
    for (int c = 0; c < NC - 1; ++c) {
      double z = 0.0;
      if (act && c < q) {
        if (zmode == 0)
          z = a.Z[(int64_t)p2 * q + c];
        else if (c == 0)
          z = 1.0;
        else if (c - 1 < DM)
          z = a.X[(int64_t)p2 * d + (c - 1)];
      }
      bv[1 + c] = z;
    }
  }
  // the diagonal of this half's block (every slot of the triangle is now
  // written: the 496 pairs and the 32 diagonal entries)
  {
    const int rr = opaque(r);
    gh[g_base_rt(rr) + rr] = act ? var * (1.0 + nug) + jit : 1.0;
  }
  wave_sync();
  // ---- Sigma row r of this half's block: the lower part and the diagonal
  // from the triangle, the upper part other (finite) entries of it -- the
  // factorisation clears the upper part at each step without reading it
  // (its multipliers are 0 there)
  double S[32];
  {
    const int rr = opaque(r);
#pragma unroll
    for (int c = 0; c < 32; ++c) S[c] = gh[g_base(c) + rr];
  }
  wave_sync();   // every row is read before g takes the triangle
  {
    const int rr = opaque(lane_o & 31), e = lane_o >> 5;
#pragma unroll
    for (int hh = 0; hh < 2; ++hh)
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int t = 8 * e + 1 + u;
        const int c = (rr + t) & 31;
        const int i = rr > c ? rr : c, j = rr > c ? c : rr;
        if (u < 7 || t < 16 || rr < 16) gt[hh * kGTri + g_base_rt(j) + i] = gk[8 * hh + u];
      }
    gh[g_base_rt(rr) + rr] = 0.0;   // g(r, r)
  }
  SVH_PHASE(1)
  // ---- factorisation (unit lower L', pivots) with L'^-1 [y Z] on the fly,
  // cross-lane values by DPP row broadcasts (fac_step)
  double pr = 1.0;
  bool fail = false;
  // row masks: 1 on rows 0..15 (lo) or 16..31 (hm) of the half, else 0
  const double lo = (opaque(r) < 16) ? 1.0 : 0.0, hm = 1.0 - lo;
  asm volatile("s_nop 4");   // EXEC restored by the branches above: DPP needs 5 wait states
  fac_step<0>(S, bv, pr, fail, opaque(r), K, lo, hm);
  const double dinv = rsqrt_nr(pr);   // 1 / L_rr
#pragma unroll
  for (int c = 0; c < NC; ++c) bv[c] *= dinv;   // rows of L^-1 [y Z]
  // L_KK of this half (from lane K-1)
  if (r == K - 1) {
    misc[8 * h + 1] = pr;
    misc[8 * h + 2] = dinv;
    misc[8 * h + 3] = pr;   // L_KK^2 of the block record
  }
  if (r == 31) misc[8 * h] = fail ? 1.0 : 0.0;   // the second 16-lane row saw every pivot
  wave_sync();
  const double pK = misc[8 * h + 1], dK = misc[8 * h + 2];
  SVH_PHASE(2)
  double R[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) R[c] = 0.0;
  if (grad) {
    // ---- t = L'^-T e_{K-1}, a = t / L_KK
    const int ra = opaque(r);
    double t = (ra == K - 1) ? 1.0 : 0.0;
    // for c >= 16 only rows 16..31 (the second 16-lane row) contribute
#pragma unroll
    for (int c = 30; c >= 16; --c) {
      const double s = row_sum(S[c] * t);   // S[c] is zero on rows ra <= c
      if (ra == c) t -= s;
      __builtin_amdgcn_sched_barrier(0);
    }
#pragma unroll
    for (int c = 15; c >= 0; --c) {
      const double s = half_sum(S[c] * t);
      if (ra == c) t -= s;
      __builtin_amdgcn_sched_barrier(0);
    }
    const double av = t * dK;
    ab[r] = av;
    wave_sync();
    SVH_PHASE(3)
    // ---- dS_p a, row r, with the g(r) stored by the Sigma phase: entry
    // (max(r, c), min(r, c)) of the packed triangle (c = r reads the finite
    // diagonal slot and multiplies a zero difference)
    const int rd = opaque(r);
    const int gb_r = g_base(rd);
    double x[DM];
#pragma unroll
    for (int l = 0; l < DM; ++l) x[l] = xs[xs_at(l, rd)];
    // every lane runs the sums (padding rows have zero g and a), results kept
    // for the block's rows only
    if constexpr (MIX) {
      // float32 terms and accumulation
      float Rf[DM];
#pragma unroll
      for (int l = 0; l < DM; ++l) Rf[l] = 0.0f;
#pragma unroll 2
      for (int c = 0; c < 32; ++c) {
        const float ac = (float)ab[c];
        const int gi = (c < rd) ? g_base(c) + rd : gb_r + c;
        const float g = (float)gh[gi];
        float dl[DM];
#pragma unroll
        for (int l = 0; l < DM; ++l) dl[l] = (float)(x[l] - xs[xs_at(l, c)]);
        const float w = g * ac;
#pragma unroll
        for (int l = 0; l < DM; ++l) Rf[l] = fmaf(w, dl[l] * dl[l], Rf[l]);
        __builtin_amdgcn_sched_barrier(0);
      }
#pragma unroll
      for (int l = 0; l < DM; ++l) R[1 + l] = act ? -var * (double)Rf[l] : 0.0;
    } else {
      // sum_c w_c (x_rl - x_cl)^2, w_c = g(r, c) a_c, expanded about the
      // block's own point (x~ = x - x_{K-1}, small within a block):
      // x~_rl^2 W0 - 2 x~_rl W1_l + W2_l, W0 = sum w, W1 = sum w x~_c,
      // W2 = sum w x~_c^2, the centred coordinates and their squares in LDS
      // (every lane reads the same c: broadcasts).  g(r, r) is 0.
      double* xs2 = xs + 32 * DM;
      {
        double xt[DM];
#pragma unroll
        for (int l = 0; l < DM; ++l) xt[l] = x[l] - xs[xs_at(l, K - 1)];
        wave_sync();   // the block's own point is read before the slots change
#pragma unroll
        for (int l = 0; l < DM; ++l) {
          xs[xs_at(l, rd)] = xt[l];
          xs2[xs_at(l, rd)] = xt[l] * xt[l];
        }
      }
      wave_sync();
      double W0 = 0.0, W1[DM], W2[DM];
#pragma unroll
      for (int l = 0; l < DM; ++l) W1[l] = W2[l] = 0.0;
      // coordinates l, l + 1 of point c are adjacent (16-byte aligned pairs):
      // one 128-bit read each, constant offsets from one base
      typedef double dpair __attribute__((ext_vector_type(2)));
      const dpair* xp = reinterpret_cast<const dpair*>(xs);
      const dpair* xp2 = reinterpret_cast<const dpair*>(xs2);
#pragma unroll
      for (int c = 0; c < 32; ++c) {
        const double ac = ab[c];
        const int gi = (c < rd) ? g_base(c) + rd : gb_r + c;
        const double w = gh[gi] * ac;
        W0 += w;
#pragma unroll
        for (int l = 0; l < DM; l += 2) {
          const dpair v = xp[xs_at(l, c) / 2], v2 = xp2[xs_at(l, c) / 2];
          W1[l] = fma(w, v.x, W1[l]);
          W1[l + 1] = fma(w, v.y, W1[l + 1]);
          W2[l] = fma(w, v2.x, W2[l]);
          W2[l + 1] = fma(w, v2.y, W2[l + 1]);
        }
        __builtin_amdgcn_sched_barrier(0);
      }
#pragma unroll
      for (int l = 0; l < DM; ++l) {
        const double xr = xs[xs_at(l, rd)];
        const double v = fma(xr, fma(xr, W0, -2.0 * W1[l]), W2[l]);
        R[1 + l] = act ? v * -var : 0.0;
      }
    }
    R[0] = act ? ((r == K - 1) ? pK * dK : 0.0) - jit * av : 0.0;
    R[NC - 1] = act ? var * nug * av : 0.0;
    SVH_PHASE(4)
    // ---- W = L^-1 (dS a): unit solve with L', then scale by 1/L_rr
    if constexpr (MIX) {
      float Rf[NC];
#pragma unroll
      for (int c = 0; c < NC; ++c) Rf[c] = (float)R[c];
      asm volatile("s_nop 4");   // EXEC restored by the branches above: DPP needs 5 wait states
      w_sweep_dpp(Rf, S, lo, hm);
      const float df = (float)dinv;
#pragma unroll
      for (int c = 0; c < NC; ++c) R[c] = (double)(Rf[c] * df);
    } else {
      asm volatile("s_nop 4");   // EXEC restored by the branches above: DPP needs 5 wait states
      w_sweep_dpp(R, S, lo, hm);
#pragma unroll
      for (int c = 0; c < NC; ++c) R[c] *= dinv;
    }
    SVH_PHASE(5)
  }
  // ---- products and records, one block at a time on the whole wave
  const int cyl = grad ? P : 0;
  const int nX = q * q + q;
#pragma unroll 1
  for (int hh = 0; hh < 2; ++hh) {
    const int64_t bh = b0 + hh;
    if (bh >= a.nblk) break;
    const int Kh = a.kcnt[bh] + 1, km = Kh - 1;
    const int chh = a.chunk[bh];
    double* Mt = gt;   // 24 x 33 product operands (g is consumed)
    wave_sync();
    if (h == hh) {
      if (grad) {
        Mt[0 * 33 + r] = R[0];
#pragma unroll
        for (int l = 0; l < DM; ++l)
          if (l < d) Mt[(1 + l) * 33 + r] = R[1 + l];
        Mt[(d + 1) * 33 + r] = R[NC - 1];
      }
      Mt[cyl * 33 + r] = bv[0];
#pragma unroll
      for (int c = 0; c < NC - 1; ++c)
        if (c < q) Mt[(cyl + 1 + c) * 33 + r] = bv[1 + c];
    }
    wave_sync();
    const bool okh = misc[8 * hh] == 0.0;
    if (!okh) {
      if (lane == 0) atomicOr(a.fail + chh, 1);
      continue;
    }
    double* rec = a.cache + bh * a.rec_stride;
    for (int e = lane; e < 2 + q + (grad ? P : 0); e += kWave) {
      double v;
      if (e == 0)
        v = Mt[cyl * 33 + km];
      else if (e <= q)
        v = Mt[(cyl + e) * 33 + km];
      else if (e == q + 1)
        v = misc[8 * hh + 3];
      else
        v = Mt[(e - q - 2) * 33 + km];
      rec[e] = v;
    }
    if (grad) {
      typedef Mfma<MIX> MF;
      const int li = lane & 15, lk = lane >> 4;
      typename MF::acc_t D0 = MF::zero(), D1 = MF::zero();
#pragma unroll 2
      for (int st = 0; st < 8; ++st) {
        const int i = 4 * st + lk;
        const double w0 = (li < cyl + 1 + q) ? Mt[li * 33 + i] : 0.0;
        const double w1 = (16 + li < cyl + 1 + q) ? Mt[(16 + li) * 33 + i] : 0.0;
        D0 = MF::mma(w0, w0, D0);
        D1 = MF::mma(w0, w1, D1);
      }
      const int N = cyl + 1 + q;
      const double wKc0 = (li < N) ? Mt[li * 33 + km] : 0.0;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int p = MF::row(lk, v);
        if (p < P) {
          const double d0 = (double)D0[v], d1 = (double)D1[v];
          if (li < P) {
            accF[v] += d0 - 0.5 * Mt[p * 33 + km] * wKc0;
          } else if (li == cyl) {
            rec[2 + q + P + p] = d0;
          } else if (li < N) {
            rec[2 + q + 2 * P + p * q + (li - cyl - 1)] = d0;
          }
          if (16 + li < N) rec[2 + q + 2 * P + p * q + (16 + li - cyl - 1)] = d1;
        }
      }
    }
    for (int e = lane; e < nX; e += kWave) {
      double v;
      if (e < q * q) {
        const int c1 = e / q, c2 = e - c1 * q;
        v = Mt[(cyl + 1 + c1) * 33 + km] * Mt[(cyl + 1 + c2) * 33 + km];
      } else {
        const int c = e - q * q;
        v = Mt[(cyl + 1 + c) * 33 + km] * Mt[cyl * 33 + km];
      }
      if (e < kWave)
        accX[0] += v;
      else if (e < 2 * kWave)
        accX[1] += v;
      else
        accX[2] += v;
    }
  }
  SVH_PHASE(6)
}

template <int NU, bool MIX, int DM, int OCC, bool PROF = false>
__global__ __launch_bounds__(kWave, OCC) void ll_pass1_rows(LLArgs a) {
  extern __shared__ double lds[];
  const int lane = threadIdx.x;
  const int P = a.P, q = a.q;
  const bool grad = a.need_grad != 0;
  const int nX = q * q + q;
  double accF[4] = {0.0, 0.0, 0.0, 0.0};
  double accX[3] = {0.0, 0.0, 0.0};
  const int64_t npair = (a.nblk + 1) / 2;
  unsigned long long ph[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  const int64_t gs = gridDim.x;
  PairIdx i0 = pair_idx(kernarg_fresh<LLArgs>(), 2 * (int64_t)blockIdx.x, lane);
  PairIdx i1 = pair_idx(kernarg_fresh<LLArgs>(), 2 * ((int64_t)blockIdx.x + gs), lane);
  PairX<DM> x0;
  pair_x<DM>(kernarg_fresh<LLArgs>(), 2 * (int64_t)blockIdx.x, i0, lane, x0);
  for (int64_t pp = blockIdx.x; pp < npair; pp += gs) {
    const PairIdx ic = i0;
    const PairX<DM> xc = x0;
    i0 = i1;
    i1 = pair_idx(kernarg_fresh<LLArgs>(), 2 * (pp + 2 * gs), lane);
    pair_x<DM>(kernarg_fresh<LLArgs>(), 2 * (pp + gs), i0, lane, x0);
    rows_pair<NU, MIX, DM, PROF>(kernarg_fresh<LLArgs>(), 2 * pp, lds, accF, accX, ph, ic, xc);
  }
  if (PROF && lane == 0)
    for (int k = 0; k < 7; ++k) atomicAdd(a.prof + k, ph[k]);
  const int NACC = grad ? P * P + q * q + q : q * q + q;
  const int X0 = grad ? P * P : 0;
  double* part = a.partial + (int64_t)blockIdx.x * NACC;
  if (grad) write_fisher<MIX>(part, accF, P, lane);
  for (int e = lane; e < nX; e += kWave)
    part[X0 + e] = (e < kWave) ? accX[0] : (e < 2 * kWave ? accX[1] : accX[2]);
}

// two waves per SIMD (at most 256 registers; the kernel needs about 219)
template <int NU, int DM>
void launch_rows_nu_dm(hipStream_t st, int G, size_t lds, const LLArgs& a, bool mixed) {
  if constexpr (NU == 2 && DM == 8) {   // phase clocks: FP64, nu = 5/2, d in 5..8
    if (a.prof) {
      hipLaunchKernelGGL((ll_pass1_rows<NU, false, DM, 2, true>), dim3(G), dim3(kWave), lds, st,
                         a);
      return;
    }
  }
  if (mixed)
    hipLaunchKernelGGL((ll_pass1_rows<NU, true, DM, 2>), dim3(G), dim3(kWave), lds, st, a);
  else
    hipLaunchKernelGGL((ll_pass1_rows<NU, false, DM, 2>), dim3(G), dim3(kWave), lds, st, a);
}

template <int NU>
void launch_rows_nu(hipStream_t st, int G, int dm, const LLArgs& a, bool mixed) {
  const size_t lds = (size_t)rows_lds_doubles(dm) * sizeof(double);
  if (dm == 4)
    launch_rows_nu_dm<NU, 4>(st, G, lds, a, mixed);
  else if (dm == 8)
    launch_rows_nu_dm<NU, 8>(st, G, lds, a, mixed);
  else
    launch_rows_nu_dm<NU, 10>(st, G, lds, a, mixed);
}

// ---------------------------------------------------------------- reduce 1
// (a) one workgroup per accumulator entry: fixed-order strided sums of the
//     pass-1 partials, then an LDS tree;
// (b) one workgroup: Fisher to the output, beta = solve(XtX + 1e-12 I, Zty) by
//     Gaussian elimination with partial pivoting, parallel over the trailing
//     submatrix in LDS.
// out layout: [0] ll, [1..P] grad, [1+P .. 1+P+P*P) Fisher, then beta[q], nfail.
__global__ __launch_bounds__(256) void ll_sum_partials(const double* __restrict__ partial, int G,
                                                       int NACC, double* __restrict__ tot) {
  __shared__ double red[256];
  const int e = blockIdx.x;
  double s = 0.0;
  for (int g = threadIdx.x; g < G; g += 256) s += partial[(int64_t)g * NACC + e];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) tot[e] = red[0];
}

// Two-level sum of the pass-1 partials (G rows of NACC values): workgroup b
// adds rows [b*kSumRows, (b+1)*kSumRows) in order, reading whole rows
// (consecutive threads, consecutive entries); ll_sum_solve adds the level
// sums in order.  Reading a column per workgroup instead touched every
// 128-byte line once per entry it holds.
constexpr int kSumRows = 64;
__global__ __launch_bounds__(256) void ll_sum_rows(const double* __restrict__ partial, int G,
                                                   int NACC, double* __restrict__ lvl) {
  const int g0 = blockIdx.x * kSumRows;
  const int g1 = g0 + kSumRows < G ? g0 + kSumRows : G;
  for (int e = threadIdx.x; e < NACC; e += 256) {
    double s = 0.0;
#pragma unroll 16
    for (int g = g0; g < g1; ++g) s += partial[(int64_t)g * NACC + e];
    lvl[(int64_t)blockIdx.x * NACC + e] = s;
  }
}

// Fisher to `out`, beta = solve(XtX + 1e-12 I, Zty) by Gaussian elimination
// with partial pivoting, parallel over the trailing submatrix in LDS.
// Fisher entries are stored with row stride `fst` (P for one structure).
__device__ void ll_solve_beta_body(const double* __restrict__ tot, int P, int q, int need_grad,
                                   double* __restrict__ beta, double* __restrict__ out, int Pout,
                                   int fst) {
  __shared__ double A[32][34];
  __shared__ int piv_s;
  const int X0 = need_grad ? P * P : 0, Y0 = X0 + q * q;
  const int t = threadIdx.x;
  if (need_grad)
    for (int e = t; e < P * P; e += blockDim.x) out[1 + Pout + e] = tot[e];
  if (q == 0) return;
  for (int e = t; e < q * (q + 1); e += blockDim.x) {
    int i = e / (q + 1), j = e - i * (q + 1);
    A[i][j] = (j < q) ? tot[X0 + i * q + j] + (i == j ? 1e-12 : 0.0) : tot[Y0 + i];
  }
  __syncthreads();
  for (int k = 0; k < q; ++k) {
    if (t == 0) {
      int p = k;
      double best = fabs(A[k][k]);
      for (int i = k + 1; i < q; ++i)
        if (fabs(A[i][k]) > best) {
          best = fabs(A[i][k]);
          p = i;
        }
      piv_s = p;
    }
    __syncthreads();
    const int p = piv_s;
    if (p != k)
      for (int j = t; j <= q; j += blockDim.x) {
        double tmp = A[k][j];
        A[k][j] = A[p][j];
        A[p][j] = tmp;
      }
    __syncthreads();
    const int rem = q - 1 - k, w = q + 1 - k;
    for (int e = t; e < rem * w; e += blockDim.x) {
      int i = k + 1 + e / w, j = k + (e % w);
      if (j == k) continue;
      A[i][j] -= (A[i][k] / A[k][k]) * A[k][j];
    }
    __syncthreads();
    for (int i = k + 1 + t; i < q; i += blockDim.x) A[i][k] = 0.0;
    __syncthreads();
  }
  if (t == 0) {
    for (int i = q - 1; i >= 0; --i) {
      double s = A[i][q];
      for (int j = i + 1; j < q; ++j) s -= A[i][j] * A[j][33];
      A[i][33] = s / A[i][i];
    }
    for (int i = 0; i < q; ++i) {
      beta[i] = A[i][33];
      out[1 + Pout + Pout * Pout + i] = A[i][33];
    }
  }
  (void)fst;
}

__global__ __launch_bounds__(256) void ll_solve_beta(const double* __restrict__ tot, int P, int q,
                                                     int need_grad, double* __restrict__ beta,
                                                     double* __restrict__ out) {
  ll_solve_beta_body(tot, P, q, need_grad, beta, out, P, P);
}

// beta = solve(XtX + 1e-12 I, Zty), q <= 32, by one wave (lane i holds row i
// of the augmented system in LDS): the pivot (first largest |A[i][k]|,
// i >= k) found by every lane from broadcast reads, rows exchanged
// column-parallel, each row eliminated by its lane with
// A[i][j] -= (A[i][k] / A[k][k]) * A[k][j], back substitution by columns
// (x_c formed, then every row above subtracts A[r][c] x_c).  The caller has
// written the Fisher block of `out`.
__device__ void ll_solve_beta_wave(const double* __restrict__ tot, int P, int q, int need_grad,
                                   double* __restrict__ beta, double* __restrict__ out) {
  __shared__ double A[32][34];
  const int X0 = need_grad ? P * P : 0, Y0 = X0 + q * q;
  const int lane = threadIdx.x;   // one wave
  for (int e = lane; e < q * (q + 1); e += kWave) {
    const int i = e / (q + 1), j = e - i * (q + 1);
    A[i][j] = (j < q) ? tot[X0 + i * q + j] + (i == j ? 1e-12 : 0.0) : tot[Y0 + i];
  }
  wave_sync();
  for (int k = 0; k < q; ++k) {
    // pivot: every lane scans column k (the same LDS addresses: broadcast
    // reads) for the first largest |A[i][k]|, i >= k
    int p = k;
    double best = fabs(A[k][k]);
    for (int i = k + 1; i < q; ++i) {
      const double v = fabs(A[i][k]);
      if (v > best) {
        best = v;
        p = i;
      }
    }
    if (p != k && lane <= q) {
      const double t = A[k][lane];
      A[k][lane] = A[p][lane];
      A[p][lane] = t;
    }
    wave_sync();
    if (lane > k && lane < q) {
      const double f = A[lane][k] / A[k][k];
#pragma unroll 4
      for (int j = k + 1; j <= q; ++j) A[lane][j] -= f * A[k][j];
    }
    wave_sync();
  }
  double sr = lane < q ? A[lane][q] : 0.0, x = 0.0;
  for (int c = q - 1; c >= 0; --c) {
    const double xc = __shfl(sr, c, kWave) / A[c][c];
    if (lane == c) x = xc;
    if (lane < c) sr -= A[lane][c] * xc;
  }
  if (lane < q) {
    beta[lane] = x;
    out[1 + P + P * P + lane] = x;
  }
}

// Second level of the partial sums, then the solve of ll_solve_beta.
__global__ __launch_bounds__(1024) void ll_sum_solve(const double* __restrict__ lvl, int GS,
                                                     int NACC, int P, int q, int need_grad,
                                                     double* __restrict__ beta,
                                                     double* __restrict__ out) {
  __shared__ double stot[40 * 40 + 32 * 32 + 32];   // NACC at P <= 40, q <= 32 (set-up limits)
  __shared__ double part4[256][4];
  // four threads per entry (every fourth level sum each, loads in flight),
  // combined in a fixed order
  const int sub = threadIdx.x & 3;
  for (int e0 = 0; e0 < NACC; e0 += 256) {
    const int e = e0 + (threadIdx.x >> 2);
    double s = 0.0;
    if (e < NACC) {
#pragma unroll 8
      for (int b = sub; b < GS; b += 4) s += lvl[(int64_t)b * NACC + e];
    }
    part4[threadIdx.x >> 2][sub] = s;
    __syncthreads();
    if (sub == 0 && e < NACC) {
      const double t = (part4[threadIdx.x >> 2][0] + part4[threadIdx.x >> 2][1]) +
                       (part4[threadIdx.x >> 2][2] + part4[threadIdx.x >> 2][3]);
      stot[e] = t;
      if (need_grad && e < P * P) out[1 + P + e] = t;   // Fisher
    }
    __syncthreads();
  }
  if (threadIdx.x < kWave && q > 0) ll_solve_beta_wave(stot, P, q, need_grad, beta, out);
}

// ---------------------------------------------------------------- pass 2
// Pass 2: G consecutive lanes per record (G = 1 without the gradient, else
// the power of two >= 1 + P): lane t of a record's group owns the value term
// (t = 0) or derivative p = t - 1; every lane of the group forms zK (its loads
// of the record head hit the same addresses).  Records are read by G lanes in
// parallel instead of one thread walking 2 + q + 2P + Pq doubles.
__global__ __launch_bounds__(256) void ll_pass2(const double* __restrict__ cache, int64_t nblk,
                                                int P, int q, int need_grad, int G, int RPI,
                                                const double* __restrict__ beta,
                                                double* __restrict__ partial2) {
  // RPI consecutive records per iteration are one contiguous span of the
  // cache: copied to LDS by consecutive threads (coalesced), then lane group
  // g (G lanes) evaluates record g from LDS.  beta sits in LDS.
  extern __shared__ double srec[];   // RPI * NC records, then beta
  __shared__ double red[256];
  const int NC = need_grad ? 2 + q + 2 * P + P * q : 2 + q;
  const int NV = need_grad ? 1 + P : 1;
  double* sb = srec + RPI * NC;
  for (int c = threadIdx.x; c < q; c += 256) sb[c] = beta[c];
  const int g = threadIdx.x / G, t = threadIdx.x & (G - 1);
  double acc = 0.0;
  for (int64_t b0 = (int64_t)blockIdx.x * RPI; b0 < nblk; b0 += (int64_t)gridDim.x * RPI) {
    const int nr = (int)(nblk - b0 < RPI ? nblk - b0 : RPI);
    const double* src = cache + b0 * NC;
    __syncthreads();   // the previous records are consumed (and beta is in place)
    for (int i = threadIdx.x; i < nr * NC; i += 256) srec[i] = src[i];
    __syncthreads();
    // log L_KK^2 of the staged records, one record per lane of the first wave
    if (threadIdx.x < nr) srec[threadIdx.x * NC + 1 + q] = log(srec[threadIdx.x * NC + 1 + q]);
    __syncthreads();
    if (g < nr) {
      const double* rec = srec + g * NC;
      double zK = rec[0];
      for (int c = 0; c < q; ++c) zK -= rec[1 + c] * sb[c];
      if (t == 0) {
        acc += -0.5 * (rec[1 + q] + zK * zK);
      } else if (t < NV) {
        const int p = t - 1;
        const double* wK = rec + 2 + q;
        const double* sv = wK + P;
        const double* Tm = sv + P + p * q;
        double wz = sv[p];
        for (int c = 0; c < q; ++c) wz -= Tm[c] * sb[c];
        acc += -0.5 * (wK[p] * (1.0 + zK * zK) - 2.0 * zK * wz);
      }
    }
  }
  // sum over the workgroup's records, per lane-in-group t
  __syncthreads();
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s >= G; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x < NV) partial2[(int64_t)blockIdx.x * NV + threadIdx.x] = red[threadIdx.x];
}

__global__ __launch_bounds__(256) void ll_reduce2(const double* __restrict__ partial2, int G2,
                                                  int NV, const int* __restrict__ fail,
                                                  int nchunk, int P, int q,
                                                  double* __restrict__ out) {
  // one workgroup per output entry (blockIdx.x < NV) plus one for the failed-chunk count
  __shared__ double red[256];
  const int e = blockIdx.x;
  double s = 0.0;
  if (e < NV) {
    for (int g = threadIdx.x; g < G2; g += 256) s += partial2[(int64_t)g * NV + e];
  } else {
    for (int c = threadIdx.x; c < nchunk; c += 256) s += (fail[c] != 0) ? 1.0 : 0.0;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[e < NV ? e : 1 + P + P * P + q] = red[0];
}

__global__ void ll_escalate(int* __restrict__ level, int* __restrict__ fail, int nchunk) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < nchunk) {
    if (fail[c]) level[c] += 1;
    fail[c] = 0;
  }
}

__global__ void fill_int(int* p, int v, int64_t n) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// ======================================================================
// Batched fits: many independent datasets, every conditioning block of every
// active dataset in one launch per kernel, and a per-dataset Fisher-scoring
// state machine on the device (restates optimize.py:11-73 per dataset).
// ======================================================================
constexpr int kBP = 12;     // max parameters per dataset (d <= 10)
constexpr int kBQ = 11;     // max trend columns
constexpr int kBNACC = kBP * kBP + kBQ * kBQ + kBQ;
constexpr int kBNC = 2 + kBQ + 2 * kBP + kBP * kBQ;   // record stride
constexpr int kBNV = 1 + kBP;
constexpr int kBOUT = 1 + kBP + kBP * kBP + kBQ + 1;  // ll, grad, F, beta, nfail

struct DSDesc {
  int d, q, P, nu_code, mmax, nchunk, fixed_nug, pad;
  int64_t xoff, yoff, zoff, blk0, nblk, chunk0;
  double lvt;        // log variance target of the sigma^2 penalty
  double log_nug;    // fixed nugget (log) when fixed_nug
  double lam_hi;     // log(lambda_max) (or +inf)
};

enum FSStatus : int {
  FS_EVAL_INIT = 0,   // evaluate at theta (first objective call)
  FS_SEARCH = 1,      // evaluate the trial point of the line search
  FS_REORDER = 2,     // host: reorder hook due
  FS_DONE = 3,
  FS_FAILED = 4,      // Cholesky failed at the start or after a reorder
  FS_REEVAL = 5,      // evaluate at theta after a rebuild
  FS_STEP = 6,        // no evaluation: compute the next step (reorder skipped)
};

struct FSState {
  double theta[kBP], trial[kBP], step[kBP], g[kBP], M[kBP * kBP], beta[kBQ];
  double ll, alpha;
  int status, k, next_reorder, halv, retry, nevals, max_iter, pad;
  double tol;
};

__device__ __forceinline__ bool fs_evaluates(int st) {
  return st == FS_EVAL_INIT || st == FS_SEARCH || st == FS_REEVAL;
}

// theta -> clipped evaluation point -> (var, nug, 1/lambda) per dataset
__global__ void fs_params(const DSDesc* __restrict__ ds, const FSState* __restrict__ fs,
                          double* __restrict__ par, double* __restrict__ th_eval, int ndata) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ndata) return;
  const DSDesc D = ds[i];
  const FSState& F = fs[i];
  const double* src = (F.status == FS_SEARCH) ? F.trial : F.theta;
  double th[kBP];
  for (int p = 0; p < D.P; ++p) th[p] = src[p];
  for (int l = 0; l < D.d; ++l) th[1 + l] = fmin(fmax(th[1 + l], -18.420680743952367), D.lam_hi);
  th[D.P - 1] = fmin(fmax(th[D.P - 1], -27.631021115928547), 6.907755278982137);
  if (D.fixed_nug) th[D.P - 1] = D.log_nug;
  double* pr = par + i * 16;
  pr[0] = exp(th[0]);
  pr[1] = exp(th[D.P - 1]);
  for (int l = 0; l < D.d; ++l) pr[2 + l] = 1.0 / exp(th[1 + l]);
  for (int p = 0; p < D.P; ++p) th_eval[i * kBP + p] = th[p];
}

// chunk levels back to 0 for datasets that are not retrying a jittered point
__global__ void fs_reset_chunks(const int* __restrict__ chunk_ds, const FSState* __restrict__ fs,
                                int* __restrict__ level, int* __restrict__ fail, int nchunk) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nchunk) return;
  if (fs[chunk_ds[c]].retry == 0) level[c] = 0;
  fail[c] = 0;
}

struct BatchArgs {
  const DSDesc* ds;
  const FSState* fs;
  const double* par;
  const double *X, *y, *Z;
  const int *rows, *nbr, *kcnt, *chunk;
  const int* level;
  int* fail;
  double* cache;
  double* partial;   // (tiles, kBNACC)
  int mmax_stride;
};

// Work list of the batched pass 1: tile = (dataset, first pair, end pair).
// Tiles hold a fixed number of block pairs, so a dataset gets workgroups in
// proportion to its size; tiles are grouped by the coordinate registers
// their dataset needs (DM = 4, 8, 10), one launch per group.
struct BTile {
  int ds, p0, p1, pad;
};

// The rows-in-lanes pass 1 over one tile: pairs p0..p1-1 of dataset ds (the
// pairs start at the dataset's first block), nu at run time.
template <int DM>
__global__ __launch_bounds__(kWave, 2) void ll_pass1_batch_tiles(BatchArgs B,
                                                                const BTile* __restrict__ tiles,
                                                                int tile0) {
  extern __shared__ double lds[];
  const int lane = threadIdx.x;
  const int tix = tile0 + blockIdx.x;
  const BTile t = tiles[tix];
  const int i = t.ds;
  if (!fs_evaluates(B.fs[i].status)) return;
  const DSDesc D = B.ds[i];
  LLArgs a;
  a.X = B.X + D.xoff;
  a.y = B.y + D.yoff;
  a.Z = B.Z + D.zoff;
  a.rows = B.rows;
  a.nbr = B.nbr;
  a.kcnt = B.kcnt;
  a.chunk = B.chunk;
  a.level = B.level;
  a.fail = B.fail;
#pragma unroll
  for (int l = 0; l < 14; ++l) a.inv_lam[l] = B.par[i * 16 + 2 + l];
  a.cache = B.cache;
  a.partial = nullptr;
  a.nblk = D.blk0 + D.nblk;
  a.d = D.d;
  a.q = D.q;
  a.mmax = B.mmax_stride;
  a.P = D.P;
  a.need_grad = 1;
  a.nu_code = D.nu_code;
  a.var = B.par[i * 16 + 0];
  a.nug = B.par[i * 16 + 1];
  a.rec_stride = kBNC;
  a.prof = nullptr;
  a.zmode = 0;
  a.sig32 = 0;
  a.est_nu = 0;
  double accF[4] = {0.0, 0.0, 0.0, 0.0};
  double accX[3] = {0.0, 0.0, 0.0};
  unsigned long long ph[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int pp = t.p0; pp < t.p1; ++pp) {
    const int64_t b0 = D.blk0 + 2 * (int64_t)pp;
    const PairIdx pi = pair_idx(a, b0, lane);
    PairX<DM> px;
    pair_x<DM>(a, b0, pi, lane, px);
    rows_pair<-1, false, DM, false>(a, b0, lds, accF, accX, ph, pi, px);
  }
  const int P = D.P, q = D.q;
  const int nX = q * q + q;
  double* part = B.partial + (int64_t)tix * kBNACC;
  write_fisher<false>(part, accF, P, lane);
  for (int e = lane; e < nX; e += kWave)
    part[P * P + e] = (e < kWave) ? accX[0] : (e < 2 * kWave ? accX[1] : accX[2]);
}

// (entry, dataset) workgroups: fixed-order sums of the pass-1 partials
__global__ __launch_bounds__(256) void ll_sum_batch(const DSDesc* __restrict__ ds,
                                                    const FSState* __restrict__ fs,
                                                    const double* __restrict__ partial,
                                                    const int* __restrict__ dtile,
                                                    double* __restrict__ tot) {
  __shared__ double red[256];
  const int e = blockIdx.x, i = blockIdx.y;
  if (!fs_evaluates(fs[i].status)) return;
  const DSDesc D = ds[i];
  if (e >= D.P * D.P + D.q * D.q + D.q) return;
  const int t0 = dtile[2 * i], nt = dtile[2 * i + 1];   // the dataset's tiles
  double s = 0.0;
  for (int x = threadIdx.x; x < nt; x += 256) s += partial[(int64_t)(t0 + x) * kBNACC + e];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) tot[(int64_t)i * kBNACC + e] = red[0];
}

__global__ __launch_bounds__(256) void ll_beta_batch(const DSDesc* __restrict__ ds,
                                                     const FSState* __restrict__ fs,
                                                     const double* __restrict__ tot,
                                                     double* __restrict__ beta,
                                                     double* __restrict__ out) {
  const int i = blockIdx.x;
  if (!fs_evaluates(fs[i].status)) return;
  const DSDesc D = ds[i];
  ll_solve_beta_body(tot + (int64_t)i * kBNACC, D.P, D.q, 1, beta + i * kBQ, out + i * kBOUT,
                     D.P, kBP);
}

// Pass 2 of the batch: kB2L lanes per record as in ll_pass2 (lane t: value
// term for t = 0, derivative t - 1 below 1 + P).
constexpr int kB2L = 16;
static_assert(kB2L >= 1 + kBP, "pass-2 lanes per record");

__global__ __launch_bounds__(256) void ll_pass2_batch(const DSDesc* __restrict__ ds,
                                                      const FSState* __restrict__ fs,
                                                      const double* __restrict__ cache,
                                                      const double* __restrict__ beta, int G2X,
                                                      double* __restrict__ partial2) {
  __shared__ double red[256];
  const int i = blockIdx.y;
  if (!fs_evaluates(fs[i].status)) return;
  const DSDesc D = ds[i];
  const int P = D.P, q = D.q;
  const double* bet = beta + i * kBQ;
  const int t = threadIdx.x & (kB2L - 1);
  const int64_t rstride = (int64_t)G2X * (256 / kB2L);
  double acc = 0.0;
  for (int64_t b = D.blk0 + ((int64_t)blockIdx.x * 256 + threadIdx.x) / kB2L;
       b < D.blk0 + D.nblk; b += rstride) {
    const double* rec = cache + b * kBNC;
    double zK = rec[0];
    for (int c = 0; c < q; ++c) zK -= rec[1 + c] * bet[c];
    if (t == 0) {
      acc += -0.5 * (log(rec[1 + q]) + zK * zK);
    } else if (t <= P) {
      const int p = t - 1;
      const double* wK = rec + 2 + q;
      const double* sv = wK + P;
      const double* Tm = sv + P + p * q;
      double wz = sv[p];
      for (int c = 0; c < q; ++c) wz -= Tm[c] * bet[c];
      acc += -0.5 * (wK[p] * (1.0 + zK * zK) - 2.0 * zK * wz);
    }
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s >= kB2L; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x < kBNV)
    partial2[((int64_t)i * G2X + blockIdx.x) * kBNV + threadIdx.x] =
        threadIdx.x <= P ? red[threadIdx.x] : 0.0;
}

__global__ __launch_bounds__(256) void ll_reduce2_batch(const DSDesc* __restrict__ ds,
                                                        const FSState* __restrict__ fs,
                                                        const double* __restrict__ partial2,
                                                        int G2X, const int* __restrict__ fail,
                                                        double* __restrict__ out) {
  __shared__ double red[256];
  const int e = blockIdx.x, i = blockIdx.y;
  if (!fs_evaluates(fs[i].status)) return;
  const DSDesc D = ds[i];
  if (e > 1 + D.P) return;
  double s = 0.0;
  if (e <= D.P) {
    for (int g = threadIdx.x; g < G2X; g += 256) s += partial2[((int64_t)i * G2X + g) * kBNV + e];
  } else {
    for (int c = threadIdx.x; c < D.nchunk; c += 256) s += fail[D.chunk0 + c] ? 1.0 : 0.0;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[(int64_t)i * kBOUT + (e <= D.P ? e : kBOUT - 1)] = red[0];
}

// ---- small dense solves for the Fisher step (P <= 12), one thread
__device__ bool lu_solve(double* A, double* b, int n) {
  // Gaussian elimination with partial pivoting; false on an exactly zero pivot
  for (int k = 0; k < n; ++k) {
    int p = k;
    double best = fabs(A[k * n + k]);
    for (int i = k + 1; i < n; ++i)
      if (fabs(A[i * n + k]) > best) {
        best = fabs(A[i * n + k]);
        p = i;
      }
    if (best == 0.0) return false;
    if (p != k) {
      for (int j = 0; j < n; ++j) {
        double t = A[k * n + j];
        A[k * n + j] = A[p * n + j];
        A[p * n + j] = t;
      }
      double t = b[k];
      b[k] = b[p];
      b[p] = t;
    }
    for (int i = k + 1; i < n; ++i) {
      const double f = A[i * n + k] / A[k * n + k];
      for (int j = k; j < n; ++j) A[i * n + j] -= f * A[k * n + j];
      b[i] -= f * b[k];
    }
  }
  for (int i = n - 1; i >= 0; --i) {
    double s = b[i];
    for (int j = i + 1; j < n; ++j) s -= A[i * n + j] * b[j];
    b[i] = s / A[i * n + i];
  }
  return true;
}

// Fisher-scoring direction, ridged until it is an ascent direction
// (optimize.py:28-45), capped at norm 3 (optimize.py:47-49); sets the trial.
__device__ void fs_begin_search(FSState& F, int P) {
  // one thread runs this: its work arrays live in LDS
  __shared__ double A[kBP * kBP], x[kBP], dg[kBP];
  for (int p = 0; p < P; ++p) dg[p] = fmax(F.M[p * P + p], 1e-8);
  double ridge = 0.0;
  bool have = false;
  for (int t = 0; t < 12 && !have; ++t) {
    for (int e = 0; e < P * P; ++e) A[e] = F.M[e];
    for (int p = 0; p < P; ++p) {
      A[p * P + p] += ridge * dg[p];
      x[p] = F.g[p];
    }
    if (lu_solve(A, x, P)) {
      double dot = 0.0;
      bool fin = true;
      for (int p = 0; p < P; ++p) {
        dot += x[p] * F.g[p];
        fin = fin && isfinite(x[p]);
      }
      if (dot > 0.0 && fin) {
        have = true;
        break;
      }
    }
    ridge = fmax(1e-6, ridge * 10.0);
  }
  if (!have) {
    double nrm = 0.0;
    for (int p = 0; p < P; ++p) nrm += F.g[p] * F.g[p];
    nrm = fmax(sqrt(nrm), 1e-12);
    for (int p = 0; p < P; ++p) x[p] = F.g[p] / nrm;
  }
  double nrm = 0.0;
  for (int p = 0; p < P; ++p) nrm += x[p] * x[p];
  nrm = sqrt(nrm);
  const double sc = nrm > 3.0 ? 3.0 / nrm : 1.0;
  for (int p = 0; p < P; ++p) {
    F.step[p] = x[p] * sc;
    F.trial[p] = F.theta[p] + F.step[p];
  }
  F.alpha = 1.0;
  F.halv = 0;
  F.status = FS_SEARCH;
}

__device__ void fs_begin_iteration(FSState& F, int P) {
  if (F.k > F.max_iter) {
    F.status = FS_DONE;
    return;
  }
  if (F.k == F.next_reorder) {
    F.next_reorder *= 2;
    F.status = FS_REORDER;
    return;
  }
  fs_begin_search(F, P);
}

// One workgroup per dataset: escalate failed chunks, then advance the state.
__device__ __forceinline__ void fs_update_body(FSState& F, const DSDesc* __restrict__ ds, int i,
                                               const double* __restrict__ out,
                                               const double* __restrict__ beta,
                                               int* __restrict__ level, int* __restrict__ fail,
                                               double var_penalty, int* __restrict__ status_out) {
  const DSDesc D = ds[i];
  const int P = D.P;
  __shared__ int st0;
  if (threadIdx.x == 0) st0 = F.status;
  __syncthreads();
  const int st = st0;
  if (st == FS_STEP) {
    if (threadIdx.x == 0) {
      fs_begin_search(F, P);
      status_out[i] = F.status;
    }
    return;
  }
  if (!fs_evaluates(st)) {
    if (threadIdx.x == 0) status_out[i] = st;
    return;
  }
  const double* o = out + (int64_t)i * kBOUT;
  const int nfail = (int)o[kBOUT - 1];
  if (nfail > 0) {
    // jitter ladder: failed chunks go one power up and the point is re-evaluated
    __shared__ int exhausted;
    if (threadIdx.x == 0) exhausted = 0;
    __syncthreads();
    for (int c = threadIdx.x; c < D.nchunk; c += 64) {
      const int gc = D.chunk0 + c;
      if (fail[gc]) {
        if (level[gc] >= 9) atomicOr(&exhausted, 1);
        level[gc] += 1;
      }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      if (!exhausted) {
        F.retry = 1;
      } else {
        // LinAlgError: fatal for the first evaluation / a re-evaluation,
        // halves the step inside the line search (optimize.py:54-58)
        F.retry = 0;
        if (st == FS_SEARCH) {
          F.alpha *= 0.5;
          F.halv += 1;
          if (F.halv >= 20) {
            F.status = FS_DONE;
          } else {
            for (int p = 0; p < P; ++p) F.trial[p] = F.theta[p] + F.alpha * F.step[p];
          }
        } else {
          F.status = FS_FAILED;
        }
      }
      status_out[i] = F.status;
    }
    return;
  }
  if (threadIdx.x != 0) return;
  F.retry = 0;
  F.nevals += 1;
  // objective value at the evaluated point: penalty and fixed-nugget freeze
  const double* th = (st == FS_SEARCH) ? F.trial : F.theta;
  double thc0 = th[0];
  const double LOG2PI = 1.8378770664093453;
  double ll = -0.5 * (double)D.nblk * LOG2PI + o[0];
  __shared__ double g[kBP], M[kBP * kBP];
  for (int p = 0; p < P; ++p) g[p] = o[1 + p];
  for (int e = 0; e < P * P; ++e) M[e] = o[1 + P + e];
  if (var_penalty > 0.0) {
    const double ex = fmax(0.0, thc0 - D.lvt);
    ll -= 0.5 * var_penalty * ex * ex;
    g[0] -= var_penalty * ex;
    if (ex > 0.0) M[0] += var_penalty;
  }
  if (D.fixed_nug) {
    g[P - 1] = 0.0;
    for (int p = 0; p < P; ++p) {
      M[(P - 1) * P + p] = 0.0;
      M[p * P + (P - 1)] = 0.0;
    }
    M[P * P - 1] = 1.0;
  }
  if (st == FS_EVAL_INIT || st == FS_REEVAL) {
    F.ll = ll;
    for (int p = 0; p < P; ++p) F.g[p] = g[p];
    for (int e = 0; e < P * P; ++e) F.M[e] = M[e];
    for (int c = 0; c < D.q; ++c) F.beta[c] = beta[i * kBQ + c];
    if (st == FS_EVAL_INIT) {
      F.k = 1;
      F.next_reorder = 2;
      fs_begin_iteration(F, P);
    } else {
      fs_begin_search(F, P);
    }
  } else {  // line-search trial
    if (isfinite(ll) && ll > F.ll) {
      for (int p = 0; p < P; ++p) F.theta[p] = F.trial[p];
      F.ll = ll;
      for (int p = 0; p < P; ++p) F.g[p] = g[p];
      for (int e = 0; e < P * P; ++e) F.M[e] = M[e];
      for (int c = 0; c < D.q; ++c) F.beta[c] = beta[i * kBQ + c];
      double sg = 0.0;
      for (int p = 0; p < P; ++p) sg += F.alpha * F.step[p] * g[p];
      if (fabs(sg) < F.tol) {
        F.status = FS_DONE;
      } else {
        F.k += 1;
        fs_begin_iteration(F, P);
      }
    } else {
      F.alpha *= 0.5;
      F.halv += 1;
      if (F.halv >= 20) {
        F.status = FS_DONE;
      } else {
        for (int p = 0; p < P; ++p) F.trial[p] = F.theta[p] + F.alpha * F.step[p];
      }
    }
  }
  status_out[i] = F.status;
}

// One workgroup per dataset: the dataset's state is copied into LDS by the
// whole workgroup, advanced there (thread 0 for the serial parts), and copied
// back.
__global__ __launch_bounds__(64) void fs_update(const DSDesc* __restrict__ ds,
                                                FSState* __restrict__ fs,
                                                const double* __restrict__ out,
                                                const double* __restrict__ beta,
                                                int* __restrict__ level, int* __restrict__ fail,
                                                double var_penalty, int* __restrict__ status_out) {
  static_assert(sizeof(FSState) % sizeof(unsigned long long) == 0, "state copy in words");
  constexpr int kW = sizeof(FSState) / sizeof(unsigned long long);
  __shared__ FSState Fs;
  const int i = blockIdx.x;
  unsigned long long* gw = reinterpret_cast<unsigned long long*>(fs + i);
  unsigned long long* lw = reinterpret_cast<unsigned long long*>(&Fs);
  for (int w = threadIdx.x; w < kW; w += 64) lw[w] = gw[w];
  __syncthreads();
  fs_update_body(Fs, ds, i, out, beta, level, fail, var_penalty, status_out);
  __syncthreads();
  for (int w = threadIdx.x; w < kW; w += 64) gw[w] = lw[w];
}

template <int KP>
void launch_pass1(hipStream_t st, int G, size_t lds, const LLArgs& a) {
  hipLaunchKernelGGL((ll_pass1<KP>), dim3(G), dim3(kWave), lds, st, a);
}

}  // namespace

// ---------------------------------------------------------------- host side

static void int_fill(svh_ctx* ctx, int* p, int v, int64_t n) {
  if (n <= 0) return;
  hipLaunchKernelGGL(fill_int, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, p,
                     v, n);
  SVH_HIP(hipGetLastError());
}

// Trend columns the likelihood kernels can form from the inputs instead of
// gathering them: 1 for Z = [1, X] exactly (linear trend), 2 for Z = 1
// (constant trend), 0 otherwise.
static int trend_mode(const double* X, const double* Z, int64_t n, int d, int q) {
  if (q == 1) {
    for (int64_t i = 0; i < n; ++i)
      if (!(Z[i] == 1.0)) return 0;
    return 2;
  }
  if (q == d + 1) {
    for (int64_t i = 0; i < n; ++i) {
      if (!(Z[i * q] == 1.0)) return 0;
      for (int l = 0; l < d; ++l)
        if (!(Z[i * q + 1 + l] == X[i * d + l])) return 0;
    }
    return 1;
  }
  return 0;
}

void vecchia_setup_device(svh_ctx* ctx, const double* Xo, const double* yo, const double* Zo,
                          int64_t n, int d, int q, int64_t nblk, const int64_t* rows,
                          const int64_t* nbr, const int32_t* kcnt, int mmax,
                          const int32_t* chunk_grad, const int32_t* chunk_val, int nchunk_grad,
                          int nchunk_val) {
  VecchiaState& vs = ctx->vs;
  vs.ready = false;
  SVH_REQUIRE(n > 0 && d > 0 && d <= 32 && q >= 0 && q <= 32, "setup: bad n/d/q");
  SVH_REQUIRE(d + 2 + 1 <= 40, "setup: d too large");
  SVH_REQUIRE(mmax >= 0 && mmax + 1 <= kMaxKP, "setup: conditioning sets above 127 points");
  SVH_REQUIRE(n < (1ll << 31), "setup: n too large");
  // host-side conversion to int32
  std::vector<int> hrows(nblk), hnbr((size_t)nblk * (mmax > 0 ? mmax : 1), -1);
  for (int64_t b = 0; b < nblk; ++b) {
    SVH_REQUIRE(rows[b] >= 0 && rows[b] < n, "setup: row index out of range");
    SVH_REQUIRE(kcnt[b] >= 0 && kcnt[b] <= mmax, "setup: neighbour count out of range");
    hrows[b] = (int)rows[b];
    for (int t = 0; t < kcnt[b]; ++t) {
      int64_t v = nbr[b * mmax + t];
      SVH_REQUIRE(v >= 0 && v < n, "setup: neighbour index out of range");
      hnbr[(size_t)b * mmax + t] = (int)v;
    }
    SVH_REQUIRE(chunk_grad[b] >= 0 && chunk_grad[b] < nchunk_grad, "setup: bad chunk id");
    SVH_REQUIRE(chunk_val[b] >= 0 && chunk_val[b] < nchunk_val, "setup: bad chunk id");
  }
  hipStream_t s = ctx->stream;
  auto up = [&](const char* name, const void* src, size_t bytes) {
    void* dst = ctx->arena.get(name, bytes);
    if (bytes) SVH_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
  };
  up("v.X", Xo, (size_t)n * d * sizeof(double));
  up("v.y", yo, (size_t)n * sizeof(double));
  if (q > 0) up("v.Z", Zo, (size_t)n * q * sizeof(double));
  else ctx->arena.get("v.Z", 8);
  up("v.rows", hrows.data(), (size_t)nblk * sizeof(int));
  up("v.nbr", hnbr.data(), hnbr.size() * sizeof(int));
  up("v.kcnt", kcnt, (size_t)nblk * sizeof(int));
  up("v.cg", chunk_grad, (size_t)nblk * sizeof(int));
  up("v.cgn", chunk_grad, (size_t)nblk * sizeof(int));
  up("v.cv", chunk_val, (size_t)nblk * sizeof(int));
  SVH_HIP(hipStreamSynchronize(s));
  vs.n = n;
  vs.d = d;
  vs.q = q;
  vs.nblk = nblk;
  vs.mmax = mmax;
  vs.kp = (mmax + 1 <= 32) ? 32 : (mmax + 1 <= 64 ? 64 : 128);
  vs.nchunk_grad = nchunk_grad;
  vs.nchunk_grad_nu = nchunk_grad;   // the caller's gradient chunks serve both layouts
  vs.nchunk_val = nchunk_val;
  vs.zmode = q > 0 ? trend_mode(Xo, Zo, n, d, q) : 0;
  vs.ready = true;
}

// ---------------------------------------------------------------- device rebuild
// Row i of the rebuilt structure: inputs, response and design of point
// order[i]; block i conditions row i on nn[i, 1..min(i, m)] (the groups of
// k = min(i, m) are consecutive rows, so blocks run in row order); the jitter
// chunk of block i is i below m (one-row groups) and m + (i - m) / step in the
// k = m group (steps for the gradient and value passes).
__global__ void rebuild_rows(const double* __restrict__ Xe, const double* __restrict__ ye,
                             const double* __restrict__ Ze, const int64_t* __restrict__ order,
                             const int64_t* __restrict__ nn, int64_t n, int d, int q, int m,
                             int mcol, int64_t stepg, int64_t stepv, int64_t stepn,
                             double* __restrict__ X, double* __restrict__ y,
                             double* __restrict__ Z, int* __restrict__ rows,
                             int* __restrict__ nbr, int* __restrict__ kcnt, int* __restrict__ cg,
                             int* __restrict__ cv, int* __restrict__ cn) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t src = order[i];
  for (int l = 0; l < d; ++l) X[i * d + l] = Xe[src * d + l];
  y[i] = ye[src];
  for (int c = 0; c < q; ++c) Z[i * q + c] = Ze[src * q + c];
  const int k = (int)(i < m ? i : m);
  rows[i] = (int)i;
  kcnt[i] = k;
  for (int t = 0; t < mcol; ++t) nbr[i * mcol + t] = t < k ? (int)nn[i * (m + 1) + 1 + t] : -1;
  cg[i] = (int)(i < m ? i : m + (i - m) / stepg);
  cv[i] = (int)(i < m ? i : m + (i - m) / stepv);
  cn[i] = (int)(i < m ? i : m + (i - m) / stepn);
}

void vecchia_rebuild_device(svh_ctx* ctx, const double* Xe, const double* ye, const double* Ze,
                            int64_t n, int d, int q, const double* lam, int64_t first, int m,
                            int upload_data, int64_t* order) {
  VecchiaState& vs = ctx->vs;
  vs.ready = false;
  SVH_REQUIRE(n > 0 && d > 0 && d <= 32 && q >= 0 && q <= 32, "rebuild: bad n/d/q");
  SVH_REQUIRE(n < (1ll << 29), "rebuild: n too large");
  SVH_REQUIRE(m >= 0 && m < n && m + 1 <= kMaxKP, "rebuild: need 0 <= m < n, m <= 126");
  SVH_REQUIRE(first >= -1 && first < n, "rebuild: first out of range");   // -1: on the device
  hipStream_t s = ctx->stream;
  // unordered estimation data stay resident between rebuilds of one fit
  double* dXe = ctx->arena.get_t<double>("rb.X", (size_t)n * d);
  double* dye = ctx->arena.get_t<double>("rb.y", n);
  double* dZe = ctx->arena.get_t<double>("rb.Z", (size_t)n * (q > 0 ? q : 1));
  if (upload_data) {
    SVH_REQUIRE(Xe && ye && (q == 0 || Ze), "rebuild: null data pointer");
    SVH_HIP(hipMemcpyAsync(dXe, Xe, (size_t)n * d * sizeof(double), hipMemcpyHostToDevice, s));
    SVH_HIP(hipMemcpyAsync(dye, ye, (size_t)n * sizeof(double), hipMemcpyHostToDevice, s));
    if (q > 0)
      SVH_HIP(hipMemcpyAsync(dZe, Ze, (size_t)n * q * sizeof(double), hipMemcpyHostToDevice, s));
    ctx->rb_n = n;
    ctx->rb_d = d;
    ctx->rb_q = q;
    ctx->rb_zmode = q > 0 ? trend_mode(Xe, Ze, n, d, q) : 0;
  }
  SVH_REQUIRE(ctx->rb_n == n && ctx->rb_d == d && ctx->rb_q == q,
              "rebuild: data not uploaded for these sizes");
  double* dlam = ctx->arena.get_t<double>("rb.lam", 32);
  double* dXT = ctx->arena.get_t<double>("rb.XT", (size_t)n * d);
  double* dXTo = ctx->arena.get_t<double>("rb.XTo", (size_t)n * d);
  int64_t* dord = ctx->arena.get_t<int64_t>("rb.order", n);
  int64_t* dnn = ctx->arena.get_t<int64_t>("rb.nn", (size_t)n * (m + 1));
  SVH_HIP(hipMemcpyAsync(dlam, lam, d * sizeof(double), hipMemcpyHostToDevice, s));
  // X / lambda in SoA (the same correctly rounded quotients as the host's)
  launch_to_soa(ctx, dXe, n, d, dlam, dXT);
  launch_maximin(ctx, dXT, n, d, first, dord);
  launch_gather_soa(ctx, dXT, dord, n, d, dXTo);
  launch_ordered_knn(ctx, dXTo, n, d, m, 0, dnn);
  const int P = d + 2, K = m + 1, mcol = m > 0 ? m : 1;
  const int64_t stepg = std::max<int64_t>(1, 4000000 / std::max<int64_t>(1, (int64_t)K * K * P));
  const int64_t stepv = std::max<int64_t>(1, 4000000 / std::max<int64_t>(1, (int64_t)K * K));
  // gradient chunks with the smoothness estimated (P = d + 3)
  const int64_t stepn =
      std::max<int64_t>(1, 4000000 / std::max<int64_t>(1, (int64_t)K * K * (P + 1)));
  double* vX = ctx->arena.get_t<double>("v.X", (size_t)n * d);
  double* vy = ctx->arena.get_t<double>("v.y", n);
  double* vZ = ctx->arena.get_t<double>("v.Z", (size_t)n * (q > 0 ? q : 1));
  int* vrows = ctx->arena.get_t<int>("v.rows", n);
  int* vnbr = ctx->arena.get_t<int>("v.nbr", (size_t)n * mcol);
  int* vk = ctx->arena.get_t<int>("v.kcnt", n);
  int* vcg = ctx->arena.get_t<int>("v.cg", n);
  int* vcv = ctx->arena.get_t<int>("v.cv", n);
  int* vcn = ctx->arena.get_t<int>("v.cgn", n);
  hipLaunchKernelGGL(rebuild_rows, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, dXe, dye,
                     dZe, dord, dnn, n, d, q, m, mcol, stepg, stepv, stepn, vX, vy, vZ, vrows,
                     vnbr, vk, vcg, vcv, vcn);
  SVH_HIP(hipGetLastError());
  SVH_HIP(hipMemcpyAsync(order, dord, (size_t)n * sizeof(int64_t), hipMemcpyDeviceToHost, s));
  SVH_HIP(hipStreamSynchronize(s));
  ctx->timers.collect();
  const int64_t rest = n - m;   // rows of the k = m group (n > m)
  vs.n = n;
  vs.d = d;
  vs.q = q;
  vs.nblk = n;
  vs.mmax = m;
  vs.kp = (m + 1 <= 32) ? 32 : (m + 1 <= 64 ? 64 : 128);
  vs.nchunk_grad = (int)(m + (rest + stepg - 1) / stepg);
  vs.nchunk_val = (int)(m + (rest + stepv - 1) / stepv);
  vs.nchunk_grad_nu = (int)(m + (rest + stepn - 1) / stepn);
  vs.zmode = ctx->rb_zmode;
  vs.ready = true;
}

void vecchia_eval_device(svh_ctx* ctx, const double* theta, int P, double nu, int precision,
                         int need_grad, double var_penalty, double log_var_target, double* ll,
                         double* grad, double* fisher, double* beta, int* jitter_power_max) {
  VecchiaState& vs = ctx->vs;
  SVH_REQUIRE(vs.ready, "eval: call svh_vecchia_setup first");
  const bool est_nu = nu == 0.0;   // smoothness estimated: theta[1 + d] = log nu
  SVH_REQUIRE(P == vs.d + (est_nu ? 3 : 2),
              est_nu ? "eval: theta must have d+3 entries when nu is estimated"
                     : "eval: theta must have d+2 entries");
  SVH_REQUIRE(precision == SVH_FP64 || precision == SVH_MIXED || precision == SVH_MIXED_F32COV,
              "eval: bad precision mode");
  int nu_code;
  const double nuv = est_nu ? std::exp(theta[1 + vs.d]) : nu;
  if (est_nu)
    nu_code = 3;
  else if (nu == 0.5)
    nu_code = 0;
  else if (nu == 1.5)
    nu_code = 1;
  else if (nu == 2.5)
    nu_code = 2;
  else if (nu > 0.0 && std::isfinite(nu))
    nu_code = 3;
  else
    throw SvhError{SVH_ERR_ARG, "eval: nu must be positive (or 0: estimated)"};
  SVH_REQUIRE(std::isfinite(nuv) && nuv > kNuStep, "eval: smoothness out of range");
  const int d = vs.d, q = vs.q;
  hipStream_t s = ctx->stream;
  // parameters

  const int nchunk = need_grad ? (est_nu ? vs.nchunk_grad_nu : vs.nchunk_grad) : vs.nchunk_val;
  // jitter levels and failure flags of the chunks, cleared by one launch
  // unless the previous evaluation left this buffer zero (no escalation)
  int* level = ctx->arena.get_t<int>("v.levfail", 2 * (int64_t)nchunk);
  int* fail = level + nchunk;
  if (ctx->arena.fix_base || ctx->levfail_zero != (const void*)level ||
      ctx->levfail_zero_n < 2 * (int64_t)nchunk)
    int_fill(ctx, level, 0, 2 * (int64_t)nchunk);
  ctx->levfail_zero = nullptr;

  const int KP = vs.kp;
  const int NR = need_grad ? P + 1 + q : 1 + q;
  const int NACC = need_grad ? P * P + q * q + q : q * q + q;
  const int NC = need_grad ? 2 + q + 2 * P + P * q : 2 + q;
  const int NV = need_grad ? 1 + P : 1;
  // split-row kernel (one block per wave, rows in registers) when the block fits
  // SVH_LL_KERNEL = rows | split | lds forces a kernel variant (A/B timing)
  const char* kenv = std::getenv("SVH_LL_KERNEL");
  const std::string kv = kenv ? kenv : "";
  // the general-smoothness path (Bessel K_nu) runs in the LDS kernel only
  const bool fits32 = vs.mmax + 1 <= kSK && d <= kSD && q <= kSQ && nu_code < 3;
  const bool use_split = fits32 && kv != "lds";
  const bool use_rows = fits32 && kv != "lds" && kv != "split";
  const size_t lds_split = (size_t)split_lds_doubles(d) * sizeof(double);
  // launch-schedule candidates for the autotuner: the rows-in-lanes kernel at
  // two occupancies and several grid sizes, the split-row kernel, and the
  // general LDS kernel for blocks beyond 32 rows; the default comes first in
  // the order rows > split > LDS
  const int ncu = ctx->num_cu > 0 ? ctx->num_cu : 256;
  // kind 0: general LDS kernel, 1: split-row kernel, 2: rows-in-lanes kernel
  struct Cand {
    int kind;
    int G;
  };
  std::vector<Cand> cands;
  int dflt = 0;
  const int64_t npair = (vs.nblk + 1) / 2;
  if (use_rows) {
    // kind 2: rows kernel at 2 waves per SIMD (219 registers, no spills; the
    // 3-wave build spills about 60 registers, and its scratch traffic, up to
    // 0.7 GB per launch at C2, buys no measurable speed: within 2 %),
    // default 16 workgroups per CU
    for (int f : {4, 8}) cands.push_back({2, (int)std::min<int64_t>(npair, (int64_t)f * ncu)});
    dflt = (int)cands.size();
    for (int f : {16, 24}) cands.push_back({2, (int)std::min<int64_t>(npair, (int64_t)f * ncu)});
  }
  if (use_split) {
    for (int f : {8, 32}) cands.push_back({1, (int)std::min<int64_t>(vs.nblk, (int64_t)f * ncu)});
    if (!use_rows) dflt = (int)cands.size() - 1;
  }
  if (kv == "lds" || !fits32) {
    for (int f : {2, 4, 16}) cands.push_back({0, (int)std::min<int64_t>(vs.nblk, (int64_t)f * ncu)});
    dflt = (int)cands.size();
    cands.push_back({0, (int)std::min<int64_t>(vs.nblk, 2048)});
  }
  int Gmax = 1;
  for (const Cand& c : cands) Gmax = std::max(Gmax, c.G);
  int G = cands[dflt].G;
  // pass 2: lanes per record and workgroups
  int G2L = 1;
  if (need_grad)
    while (G2L < 1 + P) G2L <<= 1;
  SVH_REQUIRE(G2L <= 256, "pass 2: too many parameters");
  // pass 2: RPI records per workgroup iteration (256 / G2L lane groups, at
  // most 4096 doubles of records in LDS)
  const int NCr = need_grad ? 2 + q + 2 * P + P * q : 2 + q;
  const int RPI = std::max(1, std::min(256 / G2L, 4096 / NCr));
  const int G2 = (int)std::min<int64_t>((vs.nblk + RPI - 1) / RPI, 2048);
  const size_t p2lds = ((size_t)RPI * NCr + 33) * sizeof(double);
  const size_t lds = (size_t)(KP * d + KP * (KP + 1) + NR * KP + KP + NACC +
                             (est_nu ? KP * (KP + 1) : 0)) * sizeof(double) +
                     (size_t)KP * sizeof(int);
  SVH_REQUIRE(lds <= 160 * 1024, "eval: block too large for LDS");

  LLArgs a;
  a.X = (const double*)ctx->arena.get("v.X", 0);
  a.y = (const double*)ctx->arena.get("v.y", 0);
  a.Z = (const double*)ctx->arena.get("v.Z", 0);
  a.rows = (const int*)ctx->arena.get("v.rows", 0);
  a.nbr = (const int*)ctx->arena.get("v.nbr", 0);
  a.kcnt = (const int*)ctx->arena.get("v.kcnt", 0);
  a.chunk = (const int*)ctx->arena.get(need_grad ? (est_nu ? "v.cgn" : "v.cg") : "v.cv", 0);
  a.level = level;
  a.fail = fail;
  for (int l = 0; l < 16; ++l) a.inv_lam[l] = l < d ? 1.0 / std::exp(theta[1 + l]) : 0.0;
  a.cache = ctx->arena.get_t<double>("v.cache", (size_t)vs.nblk * NC);
  a.partial = ctx->arena.get_t<double>("v.partial", (size_t)Gmax * NACC + 1);
  a.nblk = vs.nblk;
  a.d = d;
  a.q = q;
  a.mmax = vs.mmax > 0 ? vs.mmax : 1;
  a.P = P;
  a.need_grad = need_grad;
  a.nu_code = nu_code;
  a.var = std::exp(theta[0]);
  a.nug = std::exp(theta[P - 1]);
  a.rec_stride = NC;
  a.gnu = svh_bessel::matern_nu(nu_code == 3 ? nuv : 2.5);
  a.gnu_p = svh_bessel::matern_nu(nu_code == 3 ? nuv + kNuStep : 2.5);
  a.gnu_m = svh_bessel::matern_nu(nu_code == 3 ? nuv - kNuStep : 2.5);
  a.est_nu = est_nu ? 1 : 0;
  a.prof = nullptr;
  a.zmode = vs.zmode;
  a.sig32 = precision == SVH_MIXED_F32COV ? 1 : 0;
  const char* penv = std::getenv("SVH_PHASE_PROFILE");
  if (penv && (use_split || use_rows)) {
    a.prof = ctx->arena.get_t<unsigned long long>("v.prof", 8);
    hipLaunchKernelGGL(fill_int, dim3(1), dim3(64), 0, s, (int*)a.prof, 0, (int64_t)16);
  }
  double* tot = ctx->arena.get_t<double>("v.tot", NACC + 1);
  double* lvl = ctx->arena.get_t<double>("v.lvl", (size_t)((Gmax + kSumRows - 1) / kSumRows) * NACC + 1);
  double* dbeta = ctx->arena.get_t<double>("v.beta", 33);
  double* partial2 = ctx->arena.get_t<double>("v.partial2", (size_t)G2 * NV);
  const int NOUT = 1 + P + P * P + q + 1;
  double* dout = ctx->arena.get_t<double>("v.out", NOUT);
  double* hout = static_cast<double*>(ctx->stage.get(NOUT * sizeof(double)));

  struct Launch {
    hipStream_t s;
    const std::vector<Cand>* cands;
    int nu_code, KP;
    size_t lds_split, lds;
    const LLArgs* a;
    bool mixed;
    static void run(void* u, int i) {
      const Launch& L = *static_cast<Launch*>(u);
      const Cand& c = (*L.cands)[i];
      if (c.kind >= 2) {
        const int dm = rows_dm(L.a->d, L.a->q);
        switch (L.nu_code) {
          case 0: launch_rows_nu<0>(L.s, c.G, dm, *L.a, L.mixed); break;
          case 1: launch_rows_nu<1>(L.s, c.G, dm, *L.a, L.mixed); break;
          default: launch_rows_nu<2>(L.s, c.G, dm, *L.a, L.mixed); break;
        }
      } else if (c.kind == 1) {
        switch (L.nu_code) {
          case 0: launch_split_nu<0>(L.s, c.G, L.lds_split, *L.a, L.mixed); break;
          case 1: launch_split_nu<1>(L.s, c.G, L.lds_split, *L.a, L.mixed); break;
          default: launch_split_nu<2>(L.s, c.G, L.lds_split, *L.a, L.mixed); break;
        }
      } else {
        switch (L.KP) {
          case 32: launch_pass1<32>(L.s, c.G, L.lds, *L.a); break;
          case 64: launch_pass1<64>(L.s, c.G, L.lds, *L.a); break;
          default: launch_pass1<128>(L.s, c.G, L.lds, *L.a); break;
        }
      }
    }
  };
  Launch lau{s, &cands, nu_code, KP, lds_split, lds, &a, precision != SVH_FP64};
  // per-phase clocks exist in the split-row kernel and in the FP64 rows kernel
  // for nu = 5/2 with 5..8 dimensions
  const bool rows_prof = use_rows && nu_code == 2 && d > 4 && d <= 8 && precision == SVH_FP64;
  if (a.prof && !rows_prof) {
    for (int i = 0; i < (int)cands.size(); ++i)
      if (cands[i].kind == 1) dflt = i;
  }
  if (ctx->tune.mode != 0 && !a.prof) {
    char key[160];
    int lg = 0;
    while ((int64_t(1) << (lg + 1)) <= vs.nblk) ++lg;
    std::snprintf(key, sizeof key, "ll_pass1:K%d:d%d:q%d:g%d:nu%d:p%d:nb2^%d", vs.mmax + 1, d, q,
                  need_grad, nu_code, precision, lg);
    const int pick = tune_choose(ctx, key, (int)cands.size(), dflt, &Launch::run, &lau);
    G = cands[pick].G;
    std::vector<Cand> one{cands[pick]};
    cands.swap(one);
    int_fill(ctx, fail, 0, nchunk);   // tuning launches may have flagged chunks
  } else {
    std::vector<Cand> one{cands[dflt]};
    cands.swap(one);
  }

  static const char* const kKindName[4] = {"kernel:ll_pass1", "kernel:ll_pass1_split",
                                           "kernel:ll_pass1_rows", "kernel:ll_pass1_rows"};
  ctx->timers.count(kKindName[cands[0].kind]);
  int rounds = 0;
  for (;;) {
    ctx->timers.begin(s, "vecchia_pass1");
    Launch::run(&lau, 0);
    SVH_HIP(hipGetLastError());
    ctx->timers.end(s);
    if (NACC > 0) {
      const int GS = (G + kSumRows - 1) / kSumRows;
      hipLaunchKernelGGL(ll_sum_rows, dim3(GS), dim3(256), 0, s, a.partial, G, NACC, lvl);
      SVH_REQUIRE(NACC <= 40 * 40 + 32 * 32 + 32, "evaluation: too many accumulator entries");
      hipLaunchKernelGGL(ll_sum_solve, dim3(1), dim3(1024), 0, s, lvl, GS, NACC, P, q, need_grad,
                         dbeta, dout);
    } else {
      hipLaunchKernelGGL(ll_solve_beta, dim3(1), dim3(256), 0, s, tot, P, q, need_grad, dbeta,
                         dout);
    }
    hipLaunchKernelGGL(ll_pass2, dim3(G2), dim3(256), p2lds, s, a.cache, vs.nblk, P, q, need_grad, G2L, RPI,
                       dbeta, partial2);
    hipLaunchKernelGGL(ll_reduce2, dim3(NV + 1), dim3(256), 0, s, partial2, G2, NV, fail, nchunk,
                       P, q, dout);
    SVH_HIP(hipGetLastError());
    SVH_HIP(hipMemcpyAsync(hout, dout, NOUT * sizeof(double), hipMemcpyDeviceToHost, s));
    SVH_HIP(hipStreamSynchronize(s));
    ctx->timers.collect();
    int nfail = (int)hout[NOUT - 1];
    if (nfail == 0) {
      if (rounds == 0 && !ctx->arena.fix_base) {
        ctx->levfail_zero = level;
        ctx->levfail_zero_n = 2 * (int64_t)nchunk;
      }
      break;
    }
    if (rounds == 9) throw SvhError{SVH_ERR_CHOLESKY, "block Cholesky failed even with jitter"};
    hipLaunchKernelGGL(ll_escalate, dim3((nchunk + 255) / 256), dim3(256), 0, s, level, fail,
                       nchunk);
    SVH_HIP(hipGetLastError());
    ++rounds;
  }
  const double LOG2PI = 1.8378770664093453;
  double llv = -0.5 * (double)vs.n * LOG2PI + hout[0];
  if (need_grad)
    for (int p = 0; p < P; ++p) grad[p] = hout[1 + p];
  if (need_grad)
    for (int e = 0; e < P * P; ++e) fisher[e] = hout[1 + P + e];
  for (int c = 0; c < q; ++c) beta[c] = hout[1 + P + P * P + c];
  if (var_penalty > 0.0) {
    double excess = std::max(0.0, theta[0] - log_var_target);
    llv -= 0.5 * var_penalty * excess * excess;
    if (need_grad) {
      grad[0] -= var_penalty * excess;
      if (excess > 0.0) fisher[0] += var_penalty;
    }
  }
  *ll = llv;
  if (jitter_power_max) *jitter_power_max = rounds > 0 ? -13 + rounds : 0;
  if (a.prof) {
    unsigned long long hp[8];
    SVH_HIP(hipMemcpy(hp, a.prof, 7 * sizeof(unsigned long long), hipMemcpyDeviceToHost));
    if (cands[0].kind >= 2)
      std::fprintf(stderr, "[svh phase clocks] rows: gather %llu sigma %llu chol %llu a %llu "
                   "dsa %llu wsolve %llu products %llu\n", hp[0], hp[1], hp[2], hp[3], hp[4],
                   hp[5], hp[6]);
    else
      std::fprintf(stderr, "[svh phase clocks] gather %llu pairs %llu chol %llu back %llu rhs %llu "
                   "solve %llu dots %llu\n", hp[0], hp[1], hp[2], hp[3], hp[4], hp[5], hp[6]);
  }
}

// ---------------------------------------------------------------- batched host side

// Pairs per tile of the batched pass 1: the autotuner's candidates.
constexpr int kTileCand[6] = {2, 4, 8, 16, 32, 64};
constexpr int kNTileCand = 6;

struct BatchState {
  bool ready = false;
  int ndata = 0, dmax = 0, G2X = 0, mmax = 0, nchunk = 0;
  // tile layouts, one per candidate: tiles of the DM = 4 / 8 / 10 groups at
  // toff[c] + goff[c][g] (goff[c][3] = total), per-dataset (first, count) at
  // dtile + c * 2 * ndata
  int64_t toff[kNTileCand] = {};
  int goff[kNTileCand][4] = {};
  int cand = 0;
  bool tuned = false;
  int64_t nblk = 0, nx = 0, ny = 0, nz = 0;
  std::vector<DSDesc> desc;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  double var_penalty = 1.0;
  double graph_var_penalty = 0.0;   // the value captured into `exec`
};

static BatchState& batch_of(svh_ctx* ctx) {
  static std::unordered_map<svh_ctx*, BatchState> states;
  return states[ctx];
}

void batch_release(svh_ctx* ctx) {
  BatchState& B = batch_of(ctx);
  if (B.exec) (void)hipGraphExecDestroy(B.exec);
  if (B.graph) (void)hipGraphDestroy(B.graph);
  B = BatchState();
}

void batch_setup_device(svh_ctx* ctx, int ndata, const int32_t* d, const int32_t* q,
                        const int32_t* nu_code, const int64_t* n, const double* X, const double* y,
                        const double* Z, const int64_t* nbr, const int32_t* kcnt,
                        const int32_t* chunk, const int32_t* nchunk, int mmax, const double* lvt,
                        const int32_t* fixed_nug, const double* log_nug, double lambda_max) {
  batch_release(ctx);
  ctx->batch_data = false;
  BatchState& B = batch_of(ctx);
  SVH_REQUIRE(ndata >= 1, "batch: no datasets");
  SVH_REQUIRE(mmax >= 1 && mmax + 1 <= 32, "batch: conditioning sets must have at most 31 points");
  B.desc.resize(ndata);
  int64_t xo = 0, yo = 0, zo = 0, bo = 0, co = 0;
  int dmax = 1;
  for (int i = 0; i < ndata; ++i) {
    SVH_REQUIRE(d[i] >= 1 && d[i] <= 10, "batch: d must be in 1..10");
    SVH_REQUIRE(q[i] >= 0 && q[i] <= 11, "batch: q must be in 0..11");
    SVH_REQUIRE(nu_code[i] >= 0 && nu_code[i] <= 2, "batch: nu code must be 0, 1 or 2");
    SVH_REQUIRE(n[i] >= 1 && n[i] < (1 << 30), "batch: bad n");
    DSDesc& D = B.desc[i];
    D.d = d[i];
    D.q = q[i];
    D.P = d[i] + 2;
    D.nu_code = nu_code[i];
    D.mmax = mmax;
    D.nchunk = nchunk[i];
    D.fixed_nug = fixed_nug[i];
    D.pad = 0;
    D.xoff = xo;
    D.yoff = yo;
    D.zoff = zo;
    D.blk0 = bo;
    D.nblk = n[i];
    D.chunk0 = co;
    D.lvt = lvt[i];
    D.log_nug = log_nug[i];
    D.lam_hi = std::isfinite(lambda_max) ? std::log(lambda_max)
                                          : std::numeric_limits<double>::infinity();
    xo += n[i] * d[i];
    yo += n[i];
    zo += n[i] * q[i];
    bo += n[i];
    co += nchunk[i];
    dmax = std::max(dmax, d[i]);
  }
  B.ndata = ndata;
  B.dmax = dmax;
  B.nblk = bo;
  B.nx = xo;
  B.ny = yo;
  B.nz = zo;
  B.mmax = mmax;
  B.nchunk = (int)co;
  int64_t maxn = 0;
  for (auto& D : B.desc) maxn = std::max(maxn, D.nblk);
  // workgroups per dataset: at least 1024 (pass 1) / 64 (pass 2), and with
  // few datasets enough for two waves per SIMD of the chip (8 one-wave
  // workgroups per CU in pass 1, as the single-dataset launch uses)
  const int ncu = ctx->num_cu > 0 ? ctx->num_cu : 256;
  B.G2X = (int)std::min<int64_t>((maxn * kB2L + 255) / 256,
                                  std::max(64, (8 * ncu + ndata - 1) / ndata));
  // tile layouts for every candidate tile size; the default is the candidate
  // whose tile count is nearest 8 workgroups per CU (two waves per SIMD)
  std::vector<BTile> htiles;
  std::vector<int> hdtile((size_t)kNTileCand * 2 * ndata);
  double best_gap = 1e300;
  for (int c = 0; c < kNTileCand; ++c) {
    const int T = kTileCand[c];
    B.toff[c] = (int64_t)htiles.size();
    for (int g = 0; g < 3; ++g) {
      B.goff[c][g] = (int)((int64_t)htiles.size() - B.toff[c]);
      for (int i = 0; i < ndata; ++i) {
        const DSDesc& D = B.desc[i];
        const int gdm = rows_dm(D.d, D.q);
        const int gi = gdm == 4 ? 0 : (gdm == 8 ? 1 : 2);
        if (gi != g) continue;
        const int np = (int)((D.nblk + 1) / 2);
        hdtile[((size_t)c * ndata + i) * 2] = (int)((int64_t)htiles.size() - B.toff[c]);
        int cnt = 0;
        for (int p0 = 0; p0 < np; p0 += T, ++cnt) htiles.push_back({i, p0, std::min(np, p0 + T), 0});
        hdtile[((size_t)c * ndata + i) * 2 + 1] = cnt;
      }
    }
    B.goff[c][3] = (int)((int64_t)htiles.size() - B.toff[c]);
    const double gap = std::fabs(std::log((double)B.goff[c][3] / (8.0 * ncu)));
    if (gap < best_gap) {
      best_gap = gap;
      B.cand = c;
    }
  }
  B.tuned = false;
  int max_tiles = 0;
  for (int c = 0; c < kNTileCand; ++c) max_tiles = std::max(max_tiles, B.goff[c][3]);
  // block tables (rows are 0..n-1 within each dataset, in order)
  // (nbr / kcnt NULL: empty lists, filled later by device rebuilds)
  const bool lists = nbr != nullptr;
  std::vector<int> hrows(bo), hnbr(lists ? (size_t)bo * mmax : 0, -1), hk(bo, 0), hc(bo),
      chunk_ds(co);
  for (int i = 0; i < ndata; ++i) {
    const DSDesc& D = B.desc[i];
    for (int64_t r = 0; r < D.nblk; ++r) {
      const int64_t b = D.blk0 + r;
      hrows[b] = (int)r;
      SVH_REQUIRE(chunk[b] >= 0 && chunk[b] < D.nchunk, "batch: bad chunk id");
      hc[b] = (int)(D.chunk0 + chunk[b]);
      if (!lists) continue;
      hk[b] = kcnt[b];
      SVH_REQUIRE(kcnt[b] >= 0 && kcnt[b] <= mmax, "batch: bad neighbour count");
      for (int t = 0; t < kcnt[b]; ++t) {
        const int64_t v = nbr[b * mmax + t];
        SVH_REQUIRE(v >= 0 && v < D.nblk, "batch: neighbour index out of range");
        hnbr[(size_t)b * mmax + t] = (int)v;
      }
    }
    for (int c = 0; c < D.nchunk; ++c) chunk_ds[D.chunk0 + c] = i;
  }
  hipStream_t s = ctx->stream;
  auto up = [&](const char* name, const void* src, size_t bytes) {
    void* dst = ctx->arena.get(name, bytes);
    if (bytes) SVH_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
  };
  if (X) {
    up("b.X", X, (size_t)xo * sizeof(double));
    up("b.y", y, (size_t)yo * sizeof(double));
    if (zo > 0) up("b.Z", Z, (size_t)zo * sizeof(double));
  } else {
    ctx->arena.get("b.X", (size_t)xo * sizeof(double));
    ctx->arena.get("b.y", (size_t)yo * sizeof(double));
    if (zo > 0) ctx->arena.get("b.Z", (size_t)zo * sizeof(double));
  }
  if (zo == 0) ctx->arena.get("b.Z", 8);
  up("b.rows", hrows.data(), hrows.size() * sizeof(int));
  if (lists) {
    up("b.nbr", hnbr.data(), hnbr.size() * sizeof(int));
  } else {
    int* dnbr = (int*)ctx->arena.get("b.nbr", (size_t)bo * mmax * sizeof(int));
    int_fill(ctx, dnbr, -1, bo * mmax);
  }
  up("b.kcnt", hk.data(), hk.size() * sizeof(int));
  up("b.chunk", hc.data(), hc.size() * sizeof(int));
  up("b.chunk_ds", chunk_ds.data(), chunk_ds.size() * sizeof(int));
  up("b.desc", B.desc.data(), B.desc.size() * sizeof(DSDesc));
  ctx->arena.get("b.level", (size_t)co * sizeof(int));
  ctx->arena.get("b.fail", (size_t)co * sizeof(int));
  ctx->arena.get("b.cache", (size_t)bo * kBNC * sizeof(double));
  up("b.tiles", htiles.data(), htiles.size() * sizeof(BTile));
  up("b.dtile", hdtile.data(), hdtile.size() * sizeof(int));
  ctx->arena.get("b.partial", (size_t)std::max(max_tiles, 1) * kBNACC * sizeof(double));
  ctx->arena.get("b.tot", (size_t)ndata * kBNACC * sizeof(double));
  ctx->arena.get("b.beta", (size_t)ndata * kBQ * sizeof(double));
  ctx->arena.get("b.partial2", (size_t)ndata * B.G2X * kBNV * sizeof(double));
  ctx->arena.get("b.out", (size_t)ndata * kBOUT * sizeof(double));
  ctx->arena.get("b.par", (size_t)ndata * 16 * sizeof(double));
  ctx->arena.get("b.theval", (size_t)ndata * kBP * sizeof(double));
  ctx->arena.get("b.fs", (size_t)ndata * sizeof(FSState));
  ctx->arena.get("b.status", (size_t)ndata * sizeof(int));
  ctx->stage.get(std::max<size_t>((size_t)ndata * sizeof(int), 64));
  SVH_HIP(hipStreamSynchronize(s));
  B.ready = true;
}

// Replace one dataset's ordered inputs and neighbour lists after a rebuild
// (same n, so every device buffer keeps its address and the graph stays valid).
void batch_update_device(svh_ctx* ctx, int i, const double* X, const double* y, const double* Z,
                         const int64_t* nbr, const int32_t* kcnt) {
  BatchState& B = batch_of(ctx);
  SVH_REQUIRE(B.ready && i >= 0 && i < B.ndata, "batch_update: bad dataset");
  const DSDesc& D = B.desc[i];
  hipStream_t s = ctx->stream;
  double* dX = (double*)ctx->arena.get("b.X", 0);
  double* dy = (double*)ctx->arena.get("b.y", 0);
  double* dZ = (double*)ctx->arena.get("b.Z", 0);
  int* dnbr = (int*)ctx->arena.get("b.nbr", 0);
  int* dk = (int*)ctx->arena.get("b.kcnt", 0);
  std::vector<int> hnbr((size_t)D.nblk * B.mmax, -1), hk(D.nblk);
  for (int64_t r = 0; r < D.nblk; ++r) {
    SVH_REQUIRE(kcnt[r] >= 0 && kcnt[r] <= B.mmax, "batch_update: bad neighbour count");
    hk[r] = kcnt[r];
    for (int t = 0; t < kcnt[r]; ++t) {
      const int64_t v = nbr[r * B.mmax + t];
      SVH_REQUIRE(v >= 0 && v < D.nblk, "batch_update: neighbour index out of range");
      hnbr[(size_t)r * B.mmax + t] = (int)v;
    }
  }
  SVH_HIP(hipMemcpyAsync(dX + D.xoff, X, (size_t)D.nblk * D.d * sizeof(double),
                         hipMemcpyHostToDevice, s));
  SVH_HIP(hipMemcpyAsync(dy + D.yoff, y, (size_t)D.nblk * sizeof(double), hipMemcpyHostToDevice, s));
  if (D.q > 0)
    SVH_HIP(hipMemcpyAsync(dZ + D.zoff, Z, (size_t)D.nblk * D.q * sizeof(double),
                           hipMemcpyHostToDevice, s));
  SVH_HIP(hipMemcpyAsync(dnbr + D.blk0 * B.mmax, hnbr.data(), hnbr.size() * sizeof(int),
                         hipMemcpyHostToDevice, s));
  SVH_HIP(hipMemcpyAsync(dk + D.blk0, hk.data(), hk.size() * sizeof(int), hipMemcpyHostToDevice, s));
  SVH_HIP(hipStreamSynchronize(s));
}

// Offsets of dataset i in the batch buffers b.* (for device-side rebuilds).
BatchLayout batch_layout(svh_ctx* ctx, int i) {
  BatchState& B = batch_of(ctx);
  SVH_REQUIRE(B.ready && i >= 0 && i < B.ndata, "batch: bad dataset");
  const DSDesc& D = B.desc[i];
  BatchLayout L;
  L.xoff = D.xoff;
  L.yoff = D.yoff;
  L.zoff = D.zoff;
  L.blk0 = D.blk0;
  L.n = D.nblk;
  L.d = D.d;
  L.q = D.q;
  L.mmax = B.mmax;
  L.nx = B.nx;
  L.ny = B.ny;
  L.nz = B.nz;
  return L;
}

// Pass 1 of every evaluating dataset with the tile layout of candidate c:
// one launch per coordinate-register group (DM = 4, 8, 10).
static void batch_pass1(svh_ctx* ctx, const BatchArgs& a, int c) {
  BatchState& B = batch_of(ctx);
  const BTile* tiles = (const BTile*)ctx->arena.get("b.tiles", 0) + B.toff[c];
  for (int g = 0; g < 3; ++g) {
    const int t0 = B.goff[c][g], nt = B.goff[c][g + 1] - t0;
    if (nt <= 0) continue;
    const int dm = g == 0 ? 4 : (g == 1 ? 8 : 10);
    const size_t lds = (size_t)rows_lds_doubles(dm) * sizeof(double);
    if (g == 0)
      hipLaunchKernelGGL(ll_pass1_batch_tiles<4>, dim3(nt), dim3(kWave), lds, ctx->stream, a,
                         tiles, t0);
    else if (g == 1)
      hipLaunchKernelGGL(ll_pass1_batch_tiles<8>, dim3(nt), dim3(kWave), lds, ctx->stream, a,
                         tiles, t0);
    else
      hipLaunchKernelGGL(ll_pass1_batch_tiles<10>, dim3(nt), dim3(kWave), lds, ctx->stream, a,
                         tiles, t0);
  }
}

// Parameters of the current evaluation point per dataset, chunk levels reset,
// and the pass-1 arguments.
static BatchArgs batch_prelude(svh_ctx* ctx) {
  BatchState& B = batch_of(ctx);
  hipStream_t s = ctx->stream;
  const DSDesc* dd = (const DSDesc*)ctx->arena.get("b.desc", 0);
  FSState* fs = (FSState*)ctx->arena.get("b.fs", 0);
  double* par = (double*)ctx->arena.get("b.par", 0);
  double* the = (double*)ctx->arena.get("b.theval", 0);
  int* level = (int*)ctx->arena.get("b.level", 0);
  int* fail = (int*)ctx->arena.get("b.fail", 0);
  const int* chunk_ds = (const int*)ctx->arena.get("b.chunk_ds", 0);
  hipLaunchKernelGGL(fs_params, dim3((B.ndata + 63) / 64), dim3(64), 0, s, dd, fs, par, the,
                     B.ndata);
  hipLaunchKernelGGL(fs_reset_chunks, dim3((B.nchunk + 255) / 256), dim3(256), 0, s, chunk_ds, fs,
                     level, fail, B.nchunk);
  BatchArgs a;
  a.ds = dd;
  a.fs = fs;
  a.par = par;
  a.X = (const double*)ctx->arena.get("b.X", 0);
  a.y = (const double*)ctx->arena.get("b.y", 0);
  a.Z = (const double*)ctx->arena.get("b.Z", 0);
  a.rows = (const int*)ctx->arena.get("b.rows", 0);
  a.nbr = (const int*)ctx->arena.get("b.nbr", 0);
  a.kcnt = (const int*)ctx->arena.get("b.kcnt", 0);
  a.chunk = (const int*)ctx->arena.get("b.chunk", 0);
  a.level = level;
  a.fail = fail;
  a.cache = (double*)ctx->arena.get("b.cache", 0);
  a.partial = (double*)ctx->arena.get("b.partial", 0);
  a.mmax_stride = B.mmax;
  return a;
}

// The autotuner over the batched pass 1: the tile size (pairs per
// workgroup) of the current batch, timed on its first evaluation point.
static void batch_tune(svh_ctx* ctx) {
  BatchState& B = batch_of(ctx);
  if (B.tuned || ctx->tune.mode == 0) return;
  B.tuned = true;
  struct TLaunch {
    svh_ctx* ctx;
    static void run(void* u, int c) {
      svh_ctx* cx = static_cast<TLaunch*>(u)->ctx;
      const BatchArgs a = batch_prelude(cx);
      batch_pass1(cx, a, c);
    }
  } lau{ctx};
  int64_t pairs = 0;
  for (auto& D : B.desc) pairs += (D.nblk + 1) / 2;
  int lg = 0;
  while ((int64_t(1) << (lg + 1)) <= pairs) ++lg;
  char key[160];
  std::snprintf(key, sizeof key, "batch_pass1:nd%d:dm%d:m%d:np2^%d", B.ndata, B.dmax, B.mmax,
                lg);
  B.cand = tune_choose(ctx, key, kNTileCand, B.cand, &TLaunch::run, &lau);
}

// Enqueue one evaluation of every evaluating dataset plus the state update.
static void batch_enqueue(svh_ctx* ctx, bool with_update) {
  BatchState& B = batch_of(ctx);
  hipStream_t s = ctx->stream;
  const DSDesc* dd = (const DSDesc*)ctx->arena.get("b.desc", 0);
  FSState* fs = (FSState*)ctx->arena.get("b.fs", 0);
  int* level = (int*)ctx->arena.get("b.level", 0);
  int* fail = (int*)ctx->arena.get("b.fail", 0);
  double* out = (double*)ctx->arena.get("b.out", 0);
  double* beta = (double*)ctx->arena.get("b.beta", 0);
  double* tot = (double*)ctx->arena.get("b.tot", 0);
  double* partial = (double*)ctx->arena.get("b.partial", 0);
  double* partial2 = (double*)ctx->arena.get("b.partial2", 0);
  double* cache = (double*)ctx->arena.get("b.cache", 0);
  int* status = (int*)ctx->arena.get("b.status", 0);
  const BatchArgs a = batch_prelude(ctx);
  batch_pass1(ctx, a, B.cand);
  const int* dtile = (const int*)ctx->arena.get("b.dtile", 0) + (size_t)B.cand * 2 * B.ndata;
  hipLaunchKernelGGL(ll_sum_batch, dim3(kBNACC, B.ndata), dim3(256), 0, s, dd, fs, partial, dtile,
                     tot);
  hipLaunchKernelGGL(ll_beta_batch, dim3(B.ndata), dim3(256), 0, s, dd, fs, tot, beta, out);
  hipLaunchKernelGGL(ll_pass2_batch, dim3(B.G2X, B.ndata), dim3(256), 0, s, dd, fs, cache, beta,
                     B.G2X, partial2);
  hipLaunchKernelGGL(ll_reduce2_batch, dim3(kBNV + 1, B.ndata), dim3(256), 0, s, dd, fs, partial2,
                     B.G2X, fail, out);
  if (with_update)
    hipLaunchKernelGGL(fs_update, dim3(B.ndata), dim3(64), 0, s, dd, fs, out, beta, level, fail,
                       B.var_penalty, status);
  SVH_HIP(hipGetLastError());
}

void batch_fit_start_device(svh_ctx* ctx, const double* theta0, int max_iter, double tol,
                            double var_penalty, int use_graph) {
  BatchState& B = batch_of(ctx);
  SVH_REQUIRE(B.ready, "batch: call svh_batch_setup first");
  std::vector<FSState> h(B.ndata);
  for (int i = 0; i < B.ndata; ++i) {
    FSState& F = h[i];
    std::memset(&F, 0, sizeof(FSState));
    for (int p = 0; p < B.desc[i].P; ++p) F.theta[p] = theta0[i * kBP + p];
    F.status = FS_EVAL_INIT;
    F.max_iter = max_iter;
    F.tol = tol;
    F.alpha = 1.0;
  }
  FSState* fs = (FSState*)ctx->arena.get("b.fs", 0);
  SVH_HIP(hipMemcpyAsync(fs, h.data(), h.size() * sizeof(FSState), hipMemcpyHostToDevice,
                         ctx->stream));
  SVH_HIP(hipStreamSynchronize(ctx->stream));
  B.var_penalty = var_penalty;
  // the state update takes var_penalty as a kernel argument: a graph captured
  // with another value is captured again
  if (B.exec && B.graph_var_penalty != var_penalty) {
    SVH_HIP(hipGraphExecDestroy(B.exec));
    SVH_HIP(hipGraphDestroy(B.graph));
    B.exec = nullptr;
    B.graph = nullptr;
  }
  batch_tune(ctx);   // the tile size is fixed before the step is captured
  if (use_graph && !B.exec) {
    B.graph_var_penalty = var_penalty;
    hipStream_t s = ctx->stream;
    SVH_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    batch_enqueue(ctx, true);
    SVH_HIP(hipStreamEndCapture(s, &B.graph));
    SVH_HIP(hipGraphInstantiate(&B.exec, B.graph, nullptr, nullptr, 0));
  }
}

// Advance every dataset by `nrep` state-machine steps (each = one objective
// evaluation per evaluating dataset); copies the statuses to `status`.
void batch_fit_step_device(svh_ctx* ctx, int nrep, int32_t* status) {
  BatchState& B = batch_of(ctx);
  SVH_REQUIRE(B.ready, "batch: call svh_batch_setup first");
  hipStream_t s = ctx->stream;
  for (int r = 0; r < nrep; ++r) {
    ctx->timers.begin(s, "batch_step");
    if (B.exec)
      SVH_HIP(hipGraphLaunch(B.exec, s));
    else
      batch_enqueue(ctx, true);
    ctx->timers.end(s);
  }
  int* dst = (int*)ctx->arena.get("b.status", 0);
  int* hst = (int*)ctx->stage.get((size_t)B.ndata * sizeof(int));
  SVH_HIP(hipMemcpyAsync(hst, dst, (size_t)B.ndata * sizeof(int), hipMemcpyDeviceToHost, s));
  SVH_HIP(hipStreamSynchronize(s));
  ctx->timers.collect();
  std::memcpy(status, hst, (size_t)B.ndata * sizeof(int));
}

void batch_get_device(svh_ctx* ctx, double* theta, double* ll, double* beta, int32_t* status,
                      int32_t* iters, int32_t* nevals) {
  BatchState& B = batch_of(ctx);
  SVH_REQUIRE(B.ready, "batch: call svh_batch_setup first");
  std::vector<FSState> h(B.ndata);
  FSState* fs = (FSState*)ctx->arena.get("b.fs", 0);
  SVH_HIP(hipMemcpyAsync(h.data(), fs, h.size() * sizeof(FSState), hipMemcpyDeviceToHost,
                         ctx->stream));
  SVH_HIP(hipStreamSynchronize(ctx->stream));
  for (int i = 0; i < B.ndata; ++i) {
    for (int p = 0; p < kBP; ++p) theta[i * kBP + p] = h[i].theta[p];
    for (int c = 0; c < kBQ; ++c) beta[i * kBQ + c] = h[i].beta[c];
    ll[i] = h[i].ll;
    status[i] = h[i].status;
    iters[i] = h[i].k;
    nevals[i] = h[i].nevals;
  }
}

// Host decision for a dataset in FS_REORDER: 1 = rebuilt (re-evaluate),
// 0 = unchanged (take the next step without evaluating).
void batch_set_reorder_device(svh_ctx* ctx, int i, int rebuilt) {
  BatchState& B = batch_of(ctx);
  SVH_REQUIRE(B.ready && i >= 0 && i < B.ndata, "batch: bad dataset");
  FSState* fs = (FSState*)ctx->arena.get("b.fs", 0);
  const int v = rebuilt ? FS_REEVAL : FS_STEP;
  SVH_HIP(hipMemcpyAsync((char*)(fs + i) + offsetof(FSState, status), &v, sizeof(int),
                         hipMemcpyHostToDevice, ctx->stream));
  SVH_HIP(hipStreamSynchronize(ctx->stream));
}

// One batched objective evaluation at the given parameters (no state
// update; the jitter ladder runs per dataset as in svh_vecchia_eval).
void batch_eval_device(svh_ctx* ctx, const double* theta, double* ll, double* grad,
                       double* fisher, double* beta, int32_t* jitter_rounds) {
  BatchState& B = batch_of(ctx);
  SVH_REQUIRE(B.ready, "batch: call svh_batch_setup first");
  hipStream_t s = ctx->stream;
  std::vector<FSState> h(B.ndata);
  for (int i = 0; i < B.ndata; ++i) {
    std::memset(&h[i], 0, sizeof(FSState));
    for (int p = 0; p < B.desc[i].P; ++p) h[i].theta[p] = theta[i * kBP + p];
    h[i].status = FS_EVAL_INIT;
  }
  FSState* fs = (FSState*)ctx->arena.get("b.fs", 0);
  const int* chunk_ds = (const int*)ctx->arena.get("b.chunk_ds", 0);
  int* level = (int*)ctx->arena.get("b.level", 0);
  int* fail = (int*)ctx->arena.get("b.fail", 0);
  double* out = (double*)ctx->arena.get("b.out", 0);
  double* dbeta = (double*)ctx->arena.get("b.beta", 0);
  std::vector<double> hout((size_t)B.ndata * kBOUT), hb((size_t)B.ndata * kBQ);
  int rounds = 0;
  for (;;) {
    SVH_HIP(hipMemcpyAsync(fs, h.data(), h.size() * sizeof(FSState), hipMemcpyHostToDevice, s));
    batch_enqueue(ctx, false);
    SVH_HIP(hipMemcpyAsync(hout.data(), out, hout.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    SVH_HIP(hipMemcpyAsync(hb.data(), dbeta, hb.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    SVH_HIP(hipStreamSynchronize(s));
    bool any = false;
    for (int i = 0; i < B.ndata; ++i) any = any || hout[(size_t)i * kBOUT + kBOUT - 1] > 0.0;
    if (!any) break;
    if (rounds == 9) throw SvhError{SVH_ERR_CHOLESKY, "block Cholesky failed even with jitter"};
    hipLaunchKernelGGL(ll_escalate, dim3((B.nchunk + 255) / 256), dim3(256), 0, s, level, fail,
                       B.nchunk);
    for (auto& F : h) F.retry = 1;   // keep the escalated levels on the next pass
    ++rounds;
    (void)chunk_ds;
  }
  const double LOG2PI = 1.8378770664093453;
  for (int i = 0; i < B.ndata; ++i) {
    const DSDesc& D = B.desc[i];
    const double* o = hout.data() + (size_t)i * kBOUT;
    ll[i] = -0.5 * (double)D.nblk * LOG2PI + o[0];
    for (int p = 0; p < D.P; ++p) grad[i * kBP + p] = o[1 + p];
    for (int e = 0; e < D.P * D.P; ++e) fisher[i * kBP * kBP + e] = o[1 + D.P + e];
    for (int c = 0; c < D.q; ++c) beta[i * kBQ + c] = hb[(size_t)i * kBQ + c];
  }
  if (jitter_rounds) *jitter_rounds = rounds;
}
