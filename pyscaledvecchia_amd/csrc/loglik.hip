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
//   ll_pass2  (grid-stride, one thread per block): z_K = u - v.beta,
//             W.z = s - T beta, loglik and gradient terms.
//   ll_reduce2 (1 workgroup): fixed-order sum, failed-chunk count.
// All reductions are deterministic.
//
// Cached record per block (need_grad):  u, v[q], log L_KK, wK[P], s[P], T[P*q]
// where u = (L^-1 y)_K, v = (L^-1 Z)_K, W_p = L^-1 dS_p a, s_p = W_p.(L^-1 y),
// T_pc = W_p.(L^-1 Z_c).  Then z = L^-1 y - (L^-1 Z) beta gives
// z_K = u - v.beta and W_p.z = s_p - T_p.beta exactly (linearity of L^-1).

#include "svh_internal.h"

#include <cmath>
#include <cstdlib>
#include <string>
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
  const double* inv_lam;
  double* cache;
  double* partial;
  int64_t nblk;
  int d, q, mmax, P, need_grad, nu_code;
  double var, nug;
};

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
  int* pts = reinterpret_cast<int*>(acc + NACC);  // KP

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
      matern(a.nu_code, sqrt(s), M, g);
      if (i == j) {
        S[i * LD + i] = a.var * (M + a.nug);
      } else {
        S[i * LD + j] = a.var * M;
        S[j * LD + i] = g;
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
        v = log(S[km * LD + km]);
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
__device__ __forceinline__ void matern_t(double r, double& M, double& g) {
  if constexpr (NU == 0) {
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

// LDS doubles per wave: Cb 32 | Bc 24 | Misc 8 | Lc 32*32 | Xs 32*d | Mt 32*33
__host__ __device__ inline int split_lds_doubles(int d) {
  return 32 + 24 + 8 + 32 * 32 + 32 * d + 32 * 33;
}

template <int NU>
__global__ __launch_bounds__(kWave) void ll_pass1_split(LLArgs a) {
  extern __shared__ double lds[];
  const int lane = threadIdx.x, h = lane >> 5, r = lane & 31;
  const int d = a.d, q = a.q, P = a.P;
  const bool grad = a.need_grad != 0;
  double* Cb = lds;          // one column of the block matrix / the vector a
  double* Bc = Cb + 32;      // one row of right-hand sides (24 slots)
  double* Misc = Bc + 24;    // scalars of row K-1
  double* Lc = Misc + 8;     // L' column-major: Lc[j*32 + i] = L'_ij (i > j)
  double* Xs = Lc + 32 * 32; // 32 x d scaled coordinates
  double* Mt = Xs + 32 * d;  // 32 x 33 scratch matrix
  const int nF = grad ? P * (P + 1) / 2 : 0;
  const int nT = grad ? P + P * q : 0;
  const int nX = q * q + q;
  const int NC = grad ? 2 + q + 2 * P + P * q : 2 + q;
  const int dlo = h ? (d + 1) / 2 : 0, dhi = h ? d : (d + 1) / 2;   // this half's dims
  double accF[2] = {0.0, 0.0};   // Fisher tasks lane, lane+64 (nF <= 78)
  double accX[3] = {0.0, 0.0, 0.0};  // GLS tasks lane, +64, +128 (nX <= 132)

  for (int64_t b = blockIdx.x; b < a.nblk; b += gridDim.x) {
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
        matern_t<NU>(sqrt(s2), Mv, gg);
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
      Misc[2] = 0.5 * log(piv);
    }
    __syncthreads();
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
    // ---- right-hand sides
    double R[kSC];
#pragma unroll
    for (int c = 0; c < kSC; ++c) R[c] = 0.0;
    if (grad) {
      // dS_{1+l} a for this half's dims, then half 1 gathers all of them
      double accl[kSD];
#pragma unroll
      for (int l = 0; l < kSD; ++l) accl[l] = 0.0;
#pragma unroll 2
      for (int j = 0; j < kSK; ++j) {
        const double gj = (j == r) ? 0.0 : ((j < r) ? Mt[j * 33 + r] : Mt[r * 33 + j]);
        const double wj = gj * Cb[j];
#pragma unroll
        for (int l = 0; l < kSD; ++l)
          if (l >= dlo && l < dhi) {
            const double dd = xr[l] - Xs[j * d + l];
            accl[l] = fma(wj * dd, dd, accl[l]);
          }
      }
#pragma unroll
      for (int l = 0; l < kSD; ++l) {
        const double other = __shfl_xor(accl[l], 32, 64);
        const double full = (l >= dlo && l < dhi) ? accl[l] : other;
        if (l < d) R[1 + l] = h ? -a.var * full : R[1 + l];
      }
      if (h == 1) {
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
    double* rec = a.cache + b * NC;
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
    for (int e = lane; e < nF + nT; e += kWave) {
      int c1, c2;
      if (e < nF) {
        pair_of(e, c1, c2);
      } else if (e < nF + P) {
        c1 = e - nF;
        c2 = cyl;
      } else {
        const int rr = e - nF - P;
        c1 = rr / q;
        c2 = cyl + 1 + (rr - c1 * q);
      }
      double sdot = 0.0;
      for (int i = 0; i < K; ++i) sdot = fma(Mt[c1 * 33 + i], Mt[c2 * 33 + i], sdot);
      if (e < nF) {
        const double f = sdot - 0.5 * Mt[c1 * 33 + km] * Mt[c2 * 33 + km];
        if (ok) {
          if (e < kWave)
            accF[0] += f;
          else
            accF[1] += f;
        }
      } else if (e < nF + P) {
        if (ok) rec[2 + q + P + (e - nF)] = sdot;
      } else {
        if (ok) rec[2 + q + 2 * P + (e - nF - P)] = sdot;
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
  }
  // this workgroup's partials (layout of ll_pass1: F (P*P), XtX (q*q), Zty (q))
  const int NACC = grad ? P * P + q * q + q : q * q + q;
  const int X0 = grad ? P * P : 0;
  double* part = a.partial + (int64_t)blockIdx.x * NACC;
  for (int e = lane; e < nF; e += kWave) {
    int c1, c2;
    pair_of(e, c1, c2);
    const double f = (e < kWave) ? accF[0] : accF[1];
    part[c1 * P + c2] = f;
    part[c2 * P + c1] = f;
  }
  for (int e = lane; e < nX; e += kWave)
    part[X0 + e] = (e < kWave) ? accX[0] : (e < 2 * kWave ? accX[1] : accX[2]);
}

template <int NU>
void launch_split_nu(hipStream_t st, int G, size_t lds, const LLArgs& a) {
  hipLaunchKernelGGL((ll_pass1_split<NU>), dim3(G), dim3(kWave), lds, st, a);
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

__global__ __launch_bounds__(256) void ll_solve_beta(const double* __restrict__ tot, int P, int q,
                                                     int need_grad, double* __restrict__ beta,
                                                     double* __restrict__ out) {
  __shared__ double A[32][34];
  __shared__ int piv_s;
  const int X0 = need_grad ? P * P : 0, Y0 = X0 + q * q;
  const int t = threadIdx.x;
  if (need_grad)
    for (int e = t; e < P * P; e += 256) out[1 + P + e] = tot[e];
  if (q == 0) return;
  for (int e = t; e < q * (q + 1); e += 256) {
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
      for (int j = t; j <= q; j += 256) {
        double tmp = A[k][j];
        A[k][j] = A[p][j];
        A[p][j] = tmp;
      }
    __syncthreads();
    const int rem = q - 1 - k, w = q + 1 - k;
    for (int e = t; e < rem * w; e += 256) {
      int i = k + 1 + e / w, j = k + (e % w);
      if (j == k) continue;
      A[i][j] -= (A[i][k] / A[k][k]) * A[k][j];
    }
    __syncthreads();
    for (int i = k + 1 + t; i < q; i += 256) A[i][k] = 0.0;
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
      out[1 + P + P * P + i] = A[i][33];
    }
  }
}

// ---------------------------------------------------------------- pass 2
__global__ __launch_bounds__(256) void ll_pass2(const double* __restrict__ cache, int64_t nblk,
                                                int P, int q, int need_grad,
                                                const double* __restrict__ beta,
                                                double* __restrict__ partial2) {
  __shared__ double red[256];
  const int NC = need_grad ? 2 + q + 2 * P + P * q : 2 + q;
  const int NV = need_grad ? 1 + P : 1;
  double bet[32];
  for (int c = 0; c < q; ++c) bet[c] = beta[c];
  double accv[40];
  for (int e = 0; e < NV; ++e) accv[e] = 0.0;
  for (int64_t b = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; b < nblk;
       b += (int64_t)gridDim.x * blockDim.x) {
    const double* rec = cache + b * NC;
    double zK = rec[0];
    for (int c = 0; c < q; ++c) zK -= rec[1 + c] * bet[c];
    accv[0] += -0.5 * (2.0 * rec[1 + q] + zK * zK);
    if (need_grad) {
      const double* wK = rec + 2 + q;
      const double* sv = wK + P;
      const double* Tm = sv + P;
      for (int p = 0; p < P; ++p) {
        double wz = sv[p];
        for (int c = 0; c < q; ++c) wz -= Tm[p * q + c] * bet[c];
        accv[1 + p] += -0.5 * (wK[p] * (1.0 + zK * zK) - 2.0 * zK * wz);
      }
    }
  }
  for (int e = 0; e < NV; ++e) {
    red[threadIdx.x] = accv[e];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
      if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
      __syncthreads();
    }
    if (threadIdx.x == 0) partial2[(int64_t)blockIdx.x * NV + e] = red[0];
    __syncthreads();
  }
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
  up("v.cv", chunk_val, (size_t)nblk * sizeof(int));
  SVH_HIP(hipStreamSynchronize(s));
  vs.n = n;
  vs.d = d;
  vs.q = q;
  vs.nblk = nblk;
  vs.mmax = mmax;
  vs.kp = (mmax + 1 <= 32) ? 32 : (mmax + 1 <= 64 ? 64 : 128);
  vs.nchunk_grad = nchunk_grad;
  vs.nchunk_val = nchunk_val;
  vs.ready = true;
}

void vecchia_eval_device(svh_ctx* ctx, const double* theta, int P, double nu, int precision,
                         int need_grad, double var_penalty, double log_var_target, double* ll,
                         double* grad, double* fisher, double* beta, int* jitter_power_max) {
  VecchiaState& vs = ctx->vs;
  SVH_REQUIRE(vs.ready, "eval: call svh_vecchia_setup first");
  SVH_REQUIRE(P == vs.d + 2, "eval: theta must have d+2 entries");
  SVH_REQUIRE(precision == SVH_FP64 || precision == SVH_MIXED, "eval: bad precision mode");
  int nu_code;
  if (nu == 0.5)
    nu_code = 0;
  else if (nu == 1.5)
    nu_code = 1;
  else if (nu == 2.5)
    nu_code = 2;
  else
    throw SvhError{SVH_ERR_ARG, "eval: nu must be 0.5, 1.5 or 2.5"};
  const int d = vs.d, q = vs.q;
  hipStream_t s = ctx->stream;
  // parameters
  double hpar[32];
  for (int l = 0; l < d; ++l) hpar[l] = 1.0 / std::exp(theta[1 + l]);
  double* inv_lam = ctx->arena.get_t<double>("v.invlam", 32);
  SVH_HIP(hipMemcpyAsync(inv_lam, hpar, d * sizeof(double), hipMemcpyHostToDevice, s));

  const int nchunk = need_grad ? vs.nchunk_grad : vs.nchunk_val;
  int* level = ctx->arena.get_t<int>("v.level", nchunk);
  int* fail = ctx->arena.get_t<int>("v.fail", nchunk);
  int_fill(ctx, level, 0, nchunk);
  int_fill(ctx, fail, 0, nchunk);

  const int KP = vs.kp;
  const int NR = need_grad ? P + 1 + q : 1 + q;
  const int NACC = need_grad ? P * P + q * q + q : q * q + q;
  const int NC = need_grad ? 2 + q + 2 * P + P * q : 2 + q;
  const int NV = need_grad ? 1 + P : 1;
  // split-row kernel (one block per wave, rows in registers) when the block fits
  const char* kenv = std::getenv("SVH_LL_KERNEL");
  const bool use_split = !(kenv && std::string(kenv) == "lds") && vs.mmax + 1 <= kSK &&
                         d <= kSD && q <= kSQ;
  const size_t lds_split = (size_t)split_lds_doubles(d) * sizeof(double);
  const int G = use_split ? (int)std::min<int64_t>(vs.nblk, 8192)
                          : (int)std::min<int64_t>(vs.nblk, 2048);
  const int G2 = (int)std::min<int64_t>((vs.nblk + 255) / 256, 512);
  const size_t lds = (size_t)(KP * d + KP * (KP + 1) + NR * KP + KP + NACC) * sizeof(double) +
                     (size_t)KP * sizeof(int);
  SVH_REQUIRE(lds <= 160 * 1024, "eval: block too large for LDS");

  LLArgs a;
  a.X = (const double*)ctx->arena.get("v.X", 0);
  a.y = (const double*)ctx->arena.get("v.y", 0);
  a.Z = (const double*)ctx->arena.get("v.Z", 0);
  a.rows = (const int*)ctx->arena.get("v.rows", 0);
  a.nbr = (const int*)ctx->arena.get("v.nbr", 0);
  a.kcnt = (const int*)ctx->arena.get("v.kcnt", 0);
  a.chunk = (const int*)ctx->arena.get(need_grad ? "v.cg" : "v.cv", 0);
  a.level = level;
  a.fail = fail;
  a.inv_lam = inv_lam;
  a.cache = ctx->arena.get_t<double>("v.cache", (size_t)vs.nblk * NC);
  a.partial = ctx->arena.get_t<double>("v.partial", (size_t)G * NACC + 1);
  a.nblk = vs.nblk;
  a.d = d;
  a.q = q;
  a.mmax = vs.mmax > 0 ? vs.mmax : 1;
  a.P = P;
  a.need_grad = need_grad;
  a.nu_code = nu_code;
  a.var = std::exp(theta[0]);
  a.nug = std::exp(theta[1 + d]);
  double* tot = ctx->arena.get_t<double>("v.tot", NACC + 1);
  double* dbeta = ctx->arena.get_t<double>("v.beta", 33);
  double* partial2 = ctx->arena.get_t<double>("v.partial2", (size_t)G2 * NV);
  const int NOUT = 1 + P + P * P + q + 1;
  double* dout = ctx->arena.get_t<double>("v.out", NOUT);
  double* hout = static_cast<double*>(ctx->stage.get(NOUT * sizeof(double)));

  int rounds = 0;
  for (;;) {
    ctx->timers.begin(s, "vecchia_pass1");
    if (use_split) {
      switch (nu_code) {
        case 0: launch_split_nu<0>(s, G, lds_split, a); break;
        case 1: launch_split_nu<1>(s, G, lds_split, a); break;
        default: launch_split_nu<2>(s, G, lds_split, a); break;
      }
    } else {
      switch (KP) {
        case 32: launch_pass1<32>(s, G, lds, a); break;
        case 64: launch_pass1<64>(s, G, lds, a); break;
        default: launch_pass1<128>(s, G, lds, a); break;
      }
    }
    SVH_HIP(hipGetLastError());
    ctx->timers.end(s);
    if (NACC > 0)
      hipLaunchKernelGGL(ll_sum_partials, dim3(NACC), dim3(256), 0, s, a.partial, G, NACC, tot);
    hipLaunchKernelGGL(ll_solve_beta, dim3(1), dim3(256), 0, s, tot, P, q, need_grad, dbeta, dout);
    hipLaunchKernelGGL(ll_pass2, dim3(G2), dim3(256), 0, s, a.cache, vs.nblk, P, q, need_grad,
                       dbeta, partial2);
    hipLaunchKernelGGL(ll_reduce2, dim3(NV + 1), dim3(256), 0, s, partial2, G2, NV, fail, nchunk,
                       P, q, dout);
    SVH_HIP(hipGetLastError());
    SVH_HIP(hipMemcpyAsync(hout, dout, NOUT * sizeof(double), hipMemcpyDeviceToHost, s));
    SVH_HIP(hipStreamSynchronize(s));
    ctx->timers.collect();
    int nfail = (int)hout[NOUT - 1];
    if (nfail == 0) break;
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
  (void)precision;
}
