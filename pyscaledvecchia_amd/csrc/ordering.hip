// (a) Neighbour structure on gfx950: exact maximin order and k-nearest
// neighbour sets.
//
// Distances are the canonical squared distance: sequential sum over the d
// coordinates of diff*diff, each product rounded before the add (explicit
// __dmul_rn / __dadd_rn, so no FMA contraction).  Sets are taken by
// (squared distance, index); the maximin arg-max takes the lowest index among
// equal sqrt-distances.  The results are therefore bit-identical to a CPU
// evaluation of the same definitions.

#include "svh_internal.h"

#include <climits>
#include <vector>

namespace {

constexpr int kWave = 64;

__device__ __forceinline__ double sq_step(double s, double a, double b) {
  double t = a - b;
  return __dadd_rn(s, __dmul_rn(t, t));
}

// (key, idx) with "larger" = larger key, then smaller idx.
__device__ __forceinline__ bool mm_greater(double ka, int ia, double kb, int ib) {
  return ka > kb || (ka == kb && ia < ib);
}

// (key, idx) with "less" = smaller key, then smaller idx.
__device__ __forceinline__ bool nn_less(double ka, int ia, double kb, int ib) {
  return ka < kb || (ka == kb && ia < ib);
}

__device__ __forceinline__ void wave_argmax(double& k, int& i) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    double k2 = __shfl_xor(k, off, kWave);
    int i2 = __shfl_xor(i, off, kWave);
    if (mm_greater(k2, i2, k, i)) {
      k = k2;
      i = i2;
    }
  }
}

__device__ __forceinline__ void wave_argmax3(double& k, int& i, int& e) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    double k2 = __shfl_xor(k, off, kWave);
    int i2 = __shfl_xor(i, off, kWave);
    int e2 = __shfl_xor(e, off, kWave);
    if (mm_greater(k2, i2, k, i)) {
      k = k2;
      i = i2;
      e = e2;
    }
  }
}

// Full-wave maximum of an unsigned value with DPP moves (result in every lane):
// quad swaps, row shifts with bank masks, row broadcasts; lane 63 ends with
// the maximum, which readlane broadcasts.  Every lane must be active.
__device__ __forceinline__ unsigned dpp_wave_max(unsigned v) {
  v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0xB1, 0xF, 0xF, false));
  v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x4E, 0xF, 0xF, false));
  v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xF, 0xE, false));
  v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xF, 0xC, false));
  v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xA, 0xF, false));
  v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xC, 0xF, false));
  return (unsigned)__builtin_amdgcn_readlane((int)v, 63);
}

// Order-preserving map of a double to an unsigned 64-bit key.
__device__ __forceinline__ unsigned long long ord_key(double x) {
  unsigned long long b = (unsigned long long)__double_as_longlong(x);
  return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}

// Exact wave arg-max of (key desc, idx asc) by three DPP reductions: high
// word of the key, low word among the tied, then the smallest index among
// exact ties.  Returns the winning idx and the winner lane's `e`.
__device__ __forceinline__ void wave_argmax_dpp(double k, int i, int e, double& bk, int& bi,
                                                int& be) {
  const unsigned long long u = ord_key(k);
  const unsigned hi = (unsigned)(u >> 32), lo = (unsigned)u;
  const unsigned m1 = dpp_wave_max(hi);
  const unsigned m2 = dpp_wave_max(hi == m1 ? lo : 0u);
  const bool tie = (hi == m1) && (lo == m2);
  const unsigned mi = dpp_wave_max(tie ? ~(unsigned)i : 0u);
  bi = (int)~mi;
  const unsigned long long win = __ballot(tie && i == bi);
  const int wl = (int)__ffsll((long long)win) - 1;
  be = __builtin_amdgcn_readlane(e, wl);
  const unsigned long long ku = ((unsigned long long)m1 << 32) | m2;
  const unsigned long long kb = (ku >> 63) ? (ku & 0x7fffffffffffffffull) : ~ku;
  bk = __longlong_as_double((long long)kb);
}

// ------------------------------------------------------------------ transpose
// XT[l*n + j] = X[j*d + l] * (scale ? 1 : 1) / lam[l] (lam may be NULL).
__global__ void to_soa_kernel(const double* __restrict__ X, int64_t n, int d,
                              const double* __restrict__ lam, double* __restrict__ XT) {
  int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  int64_t tot = n * d;
  if (t >= tot) return;
  int64_t j = t / d;
  int l = (int)(t - j * d);
  double v = X[t];
  if (lam) v = v / lam[l];
  XT[(int64_t)l * n + j] = v;
}

// ======================================================================
// Maximin
// ======================================================================
//
// Batched exact maximin.  Each round:
//   mm_apply  (grid): apply the points accepted in the previous round to every
//             unselected point (s2 = min squared distance, dist = sqrt(s2));
//             each 256-point slice publishes its top-(kSliceTop+1) points by
//             (dist desc, index asc).
//   mm_select (1 wave): V = the largest published "next" element over all
//             slices is an upper bound (in that order) on every point that is
//             not a published candidate.  Candidates above V form a pool;
//             selection repeats  "take the pool maximum, lower pool distances
//             with the new point"  while the maximum stays above V (the first
//             pick of a round is the global maximum and always valid).  Each
//             pick is exactly the serial algorithm's next step: every point
//             outside the pool is at or below V and distances only fall.
constexpr int kSlice = 256;
constexpr int kSliceTop = 4;          // published candidates per slice
constexpr int kPool = 256;            // pool entries (4 per lane)
constexpr int kAccCap = 256;          // max accepted points per round
constexpr int kSortCap = 4096;        // survivors kept for the pool sort
constexpr int kMMD = 12;              // max d of the maximin kernels

struct MMState {
  int count;  // points placed in `order`
  int nacc;   // accepted in the last round
  int rounds;
  int overflow_sorts;
};

__global__ __launch_bounds__(kSlice) void mm_apply_kernel(
    const double* __restrict__ XT, int n, int d, double* __restrict__ s2,
    double* __restrict__ dist, const int* __restrict__ acc, const MMState* __restrict__ st,
    double* __restrict__ cand_key, int* __restrict__ cand_idx) {
  extern __shared__ double smem[];
  const int A = st->nacc;
  double* accX = smem;  // A * d
  __shared__ double wk[kSlice / kWave];
  __shared__ int wi[kSlice / kWave];
  __shared__ int accI[kAccCap];
  for (int t = threadIdx.x; t < A; t += blockDim.x) accI[t] = acc[t];
  __syncthreads();
  for (int t = threadIdx.x; t < A * d; t += blockDim.x) {
    int a = t / d, l = t - a * d;
    accX[t] = XT[(int64_t)l * n + accI[a]];
  }
  __syncthreads();
  const int i = blockIdx.x * kSlice + threadIdx.x;
  double di = -1.0;
  if (i < n) {
    di = dist[i];
    if (di >= 0.0 && A > 0) {
      double s = s2[i];
      bool changed = false, picked = false;
      for (int a = 0; a < A; ++a) {
        if (accI[a] == i) {
          picked = true;
          break;
        }
        double t = 0.0;
        const double* xa = accX + a * d;
        for (int l = 0; l < d; ++l) t = sq_step(t, XT[(int64_t)l * n + i], xa[l]);
        if (t < s) {
          s = t;
          changed = true;
        }
      }
      if (picked) {
        di = -1.0;
        dist[i] = -1.0;
      } else if (changed) {
        s2[i] = s;
        di = sqrt(s);
        dist[i] = di;
      }
    }
  }
  // top-(kSliceTop+1) of this slice by (dist desc, index asc)
  double k = (i < n) ? di : -1.0;
  int ki = (i < n) ? i : INT_MAX;
  const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
  for (int r = 0; r <= kSliceTop; ++r) {
    double bk = k;
    int bi = ki;
    wave_argmax(bk, bi);
    if (lane == 0) {
      wk[wid] = bk;
      wi[wid] = bi;
    }
    __syncthreads();
    double gk = wk[0];
    int gi = wi[0];
    for (int w = 1; w < kSlice / kWave; ++w)
      if (mm_greater(wk[w], wi[w], gk, gi)) {
        gk = wk[w];
        gi = wi[w];
      }
    if (threadIdx.x == 0) {
      cand_key[blockIdx.x * (kSliceTop + 1) + r] = gk;
      cand_idx[blockIdx.x * (kSliceTop + 1) + r] = gi;
    }
    if (ki == gi) {
      k = -2.0;
      ki = INT_MAX;
    }
    __syncthreads();
  }
}

__global__ __launch_bounds__(kWave) void mm_select_kernel(
    const double* __restrict__ XT, int n, int d, const double* __restrict__ s2g,
    int nslice, const double* __restrict__ cand_key, const int* __restrict__ cand_idx,
    int* __restrict__ acc, MMState* __restrict__ st, int64_t* __restrict__ order, int sortcap) {
  // LDS: survivors (sortcap keys + sortcap indices)
  extern __shared__ double smem[];
  double* sk = smem;
  int* si = reinterpret_cast<int*>(smem + sortcap);
  const int lane = threadIdx.x;
  const int count0 = st->count;
  if (count0 >= n) {
    if (lane == 0) st->nacc = 0;
    return;
  }
  // bound V: the largest "next" element over all slices
  double Vk = -3.0;
  int Vi = INT_MAX;
  for (int s = lane; s < nslice; s += kWave) {
    double k = cand_key[s * (kSliceTop + 1) + kSliceTop];
    int i = cand_idx[s * (kSliceTop + 1) + kSliceTop];
    if (mm_greater(k, i, Vk, Vi)) {
      Vk = k;
      Vi = i;
    }
  }
  wave_argmax(Vk, Vi);
  // survivors: published candidates strictly above V, compacted into LDS.
  // Survivors that do not fit raise the bound instead (they leave the pool).
  int cnt = 0;
  double Ok = -3.0;
  int Oi = INT_MAX;
  const int ncand = nslice * kSliceTop;
  for (int base = 0; base < ncand; base += kWave) {
    int c = base + lane;
    double k = -3.0;
    int i = INT_MAX;
    bool keep = false;
    if (c < ncand) {
      int s = c / kSliceTop, r = c - s * kSliceTop;
      k = cand_key[s * (kSliceTop + 1) + r];
      i = cand_idx[s * (kSliceTop + 1) + r];
      keep = k >= 0.0 && mm_greater(k, i, Vk, Vi);
    }
    unsigned long long m = __ballot(keep);
    int pos = cnt + __popcll(m & ((1ull << lane) - 1ull));
    if (keep) {
      if (pos < sortcap) {
        sk[pos] = k;
        si[pos] = i;
      } else if (mm_greater(k, i, Ok, Oi)) {
        Ok = k;
        Oi = i;
      }
    }
    cnt += __popcll(m);
  }
  wave_argmax(Ok, Oi);
  if (mm_greater(Ok, Oi, Vk, Vi)) {
    Vk = Ok;
    Vi = Oi;
  }
  int nsurv = cnt < sortcap ? cnt : sortcap;
  __syncthreads();
  if (nsurv > kPool) {
    // keep the top kPool survivors: bitonic sort (descending) in LDS
    int L = 1;
    while (L < nsurv) L <<= 1;
    for (int t = nsurv + lane; t < L; t += kWave) {
      sk[t] = -3.0;
      si[t] = INT_MAX;
    }
    __syncthreads();
    for (int size = 2; size <= L; size <<= 1) {
      for (int stride = size >> 1; stride > 0; stride >>= 1) {
        for (int t = lane; t < L / 2; t += kWave) {
          int lo = 2 * stride * (t / stride) + (t % stride);
          int hi = lo + stride;
          bool desc = (lo & size) == 0;
          double ka = sk[lo], kb = sk[hi];
          int ia = si[lo], ib = si[hi];
          bool sw = desc ? mm_greater(kb, ib, ka, ia) : mm_greater(ka, ia, kb, ib);
          if (sw) {
            sk[lo] = kb;
            sk[hi] = ka;
            si[lo] = ib;
            si[hi] = ia;
          }
        }
        __syncthreads();
      }
    }
    // tighten the bound with the first entry left out
    double k = sk[kPool];
    int i = si[kPool];
    if (mm_greater(k, i, Vk, Vi)) {
      Vk = k;
      Vi = i;
    }
    nsurv = kPool;
    if (lane == 0) st->overflow_sorts += 1;
    __syncthreads();
  }
  // load pool: kPool / 64 slots per lane, coordinates in registers
  constexpr int kNT = kPool / kWave;
  double pv[kNT], ps[kNT], px[kNT][kMMD];
  int pi[kNT];
#pragma unroll
  for (int t = 0; t < kNT; ++t) {
    int e = lane + t * kWave;
    if (e < nsurv) {
      pv[t] = sk[e];
      pi[t] = si[e];
      ps[t] = s2g[pi[t]];
    } else {
      pv[t] = -3.0;
      pi[t] = INT_MAX;
      ps[t] = 0.0;
    }
#pragma unroll
    for (int l = 0; l < kMMD; ++l)
      px[t][l] = (l < d && e < nsurv) ? XT[(int64_t)l * n + pi[t]] : 0.0;
  }
  int nacc = 0;
  const int room = n - count0;
  while (nacc < kAccCap && nacc < room) {
    // this lane's best slot, with its coordinates
    double lk = -3.0, lx[kMMD];
    int li = INT_MAX;
#pragma unroll
    for (int l = 0; l < kMMD; ++l) lx[l] = 0.0;
#pragma unroll
    for (int t = 0; t < kNT; ++t) {
      const bool b = mm_greater(pv[t], pi[t], lk, li);
      lk = b ? pv[t] : lk;
      li = b ? pi[t] : li;
#pragma unroll
      for (int l = 0; l < kMMD; ++l) lx[l] = b ? px[t][l] : lx[l];
    }
    double bk;
    int bi, wl;
    wave_argmax_dpp(lk, li, lane, bk, bi, wl);
    if (bk < 0.0) break;
    if (nacc > 0 && !mm_greater(bk, bi, Vk, Vi)) break;
    if (lane == 0) {
      acc[nacc] = bi;
      order[count0 + nacc] = bi;
    }
    ++nacc;
    // coordinates of the new point from the winning lane (scalar registers)
    double sx[kMMD];
#pragma unroll
    for (int l = 0; l < kMMD; ++l) {
      const long long v = __double_as_longlong(lx[l]);
      const int vlo = __builtin_amdgcn_readlane((int)(v & 0xffffffffll), wl);
      const int vhi = __builtin_amdgcn_readlane((int)(v >> 32), wl);
      sx[l] = __longlong_as_double(((long long)vhi << 32) | (unsigned)vlo);
    }
#pragma unroll
    for (int t = 0; t < kNT; ++t) {
      if (pi[t] == bi) {
        pv[t] = -3.0;
        continue;
      }
      // only entries still above V can be picked later in this round
      if (pv[t] < 0.0 || !mm_greater(pv[t], pi[t], Vk, Vi)) {
        pv[t] = -3.0;
        continue;
      }
      double s = 0.0;
#pragma unroll
      for (int l = 0; l < kMMD; ++l)
        if (l < d) s = sq_step(s, px[t][l], sx[l]);
      if (s < ps[t]) {
        ps[t] = s;
        pv[t] = sqrt(s);
      }
    }
  }
  if (lane == 0) {
    st->nacc = nacc;
    st->count = count0 + nacc;
    st->rounds += 1;
  }
}

__global__ void mm_init_kernel(int n, double* s2, double* dist, int* acc, MMState* st,
                               int64_t* order, int first) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    s2[i] = __longlong_as_double(0x7ff0000000000000ll);  // +inf
    dist[i] = __longlong_as_double(0x7ff0000000000000ll);
  }
  if (i == 0) {
    acc[0] = first;
    order[0] = first;
    st->count = 1;
    st->nacc = 1;
    st->rounds = 0;
    st->overflow_sorts = 0;
  }
}

// ======================================================================
// k nearest neighbours (ordered and query forms)
// ======================================================================
//
// One wave per query.  The wave scans candidates 64 at a time (one per
// lane); a candidate strictly better than the current k-th best goes to an
// LDS buffer (ballot + prefix).  When the buffer would overflow, and at the
// end, the wave bitonic-sorts list+buffer in LDS and keeps the first k.
template <int LCAP>
__global__ __launch_bounds__(kWave) void knn_kernel(
    const double* __restrict__ XT, int64_t nref, const double* __restrict__ QT, int64_t nq,
    int d, int m, int ordered, int64_t rowstart, int64_t* __restrict__ out, int out_stride) {
  __shared__ double key[LCAP];
  __shared__ int idx[LCAP];
  const int lane = threadIdx.x;
  const int64_t q = blockIdx.x;
  int64_t hi;
  int k;
  int64_t* orow = out + q * (int64_t)out_stride;
  if (ordered) {
    hi = q;
    k = (int)((q < (int64_t)m) ? q : (int64_t)m);
    if (q < rowstart) k = 0;
    for (int t = lane; t <= m; t += kWave) orow[t] = (t == 0) ? q : -1;
  } else {
    hi = nref;
    k = m;
  }
  if (k <= 0) return;
  const double kInf = __longlong_as_double(0x7ff0000000000000ll);
  for (int t = lane; t < LCAP; t += kWave) {
    key[t] = kInf;
    idx[t] = INT_MAX;
  }
  __syncthreads();
  const double* qcol = ordered ? XT : QT;
  const int64_t qn = ordered ? nref : nq;
  double thr = kInf;
  int thr_i = INT_MAX;
  int nbuf = 0;
  const int cap = LCAP - k;

  auto merge = [&]() {
    __syncthreads();
    for (int size = 2; size <= LCAP; size <<= 1) {
      for (int stride = size >> 1; stride > 0; stride >>= 1) {
        for (int t = lane; t < LCAP / 2; t += kWave) {
          int lo = 2 * stride * (t / stride) + (t % stride);
          int hi2 = lo + stride;
          bool asc = (lo & size) == 0;
          double ka = key[lo], kb = key[hi2];
          int ia = idx[lo], ib = idx[hi2];
          bool sw = asc ? nn_less(kb, ib, ka, ia) : nn_less(ka, ia, kb, ib);
          if (sw) {
            key[lo] = kb;
            key[hi2] = ka;
            idx[lo] = ib;
            idx[hi2] = ia;
          }
        }
        __syncthreads();
      }
    }
    thr = key[k - 1];
    thr_i = idx[k - 1];
    __syncthreads();
    for (int t = k + lane; t < LCAP; t += kWave) {
      key[t] = kInf;
      idx[t] = INT_MAX;
    }
    nbuf = 0;
    __syncthreads();
  };

  for (int64_t base = 0; base < hi; base += kWave) {
    int64_t j = base + lane;
    double s = kInf;
    if (j < hi) {
      s = 0.0;
      for (int l = 0; l < d; ++l) s = sq_step(s, qcol[(int64_t)l * qn + q], XT[(int64_t)l * nref + j]);
    }
    bool pass = j < hi && nn_less(s, (int)j, thr, thr_i);
    unsigned long long msk = __ballot(pass);
    int c = __popcll(msk);
    if (c == 0) continue;
    if (nbuf + c > cap) {
      merge();
      pass = j < hi && nn_less(s, (int)j, thr, thr_i);
      msk = __ballot(pass);
      c = __popcll(msk);
    }
    if (pass) {
      int pos = k + nbuf + __popcll(msk & ((1ull << lane) - 1ull));
      key[pos] = s;
      idx[pos] = (int)j;
    }
    nbuf += c;
  }
  merge();
  if (ordered) {
    for (int t = lane; t < k; t += kWave) orow[1 + t] = idx[t];
  } else {
    for (int t = lane; t < k; t += kWave) orow[t] = idx[t];
  }
}

template <int LCAP>
void launch_knn_t(svh_ctx* ctx, const double* XT, int64_t nref, const double* QT, int64_t nq,
                  int d, int m, int ordered, int64_t rowstart, int64_t* out, int stride) {
  int64_t nblocks = ordered ? nref : nq;
  if (nblocks == 0) return;
  ctx->timers.begin(ctx->stream, "knn");
  hipLaunchKernelGGL((knn_kernel<LCAP>), dim3((unsigned)nblocks), dim3(kWave), 0, ctx->stream, XT,
                     nref, QT, nq, d, m, ordered, rowstart, out, stride);
  SVH_HIP(hipGetLastError());
  ctx->timers.end(ctx->stream);
}

void launch_knn_any(svh_ctx* ctx, const double* XT, int64_t nref, const double* QT, int64_t nq,
                    int d, int m, int ordered, int64_t rowstart, int64_t* out, int stride) {
  SVH_REQUIRE(m <= 448, "k/m above 448 not supported by the kNN kernel");
  if (m <= 64)
    launch_knn_t<128>(ctx, XT, nref, QT, nq, d, m, ordered, rowstart, out, stride);
  else if (m <= 192)
    launch_knn_t<256>(ctx, XT, nref, QT, nq, d, m, ordered, rowstart, out, stride);
  else
    launch_knn_t<512>(ctx, XT, nref, QT, nq, d, m, ordered, rowstart, out, stride);
}

}  // namespace

// ------------------------------------------------------------------ launchers

void launch_to_soa(svh_ctx* ctx, const double* dX, int64_t n, int d, const double* dlam,
                   double* dXT) {
  int64_t tot = n * d;
  if (tot == 0) return;
  hipLaunchKernelGGL(to_soa_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0,
                     ctx->stream, dX, n, d, dlam, dXT);
  SVH_HIP(hipGetLastError());
}

void launch_maximin(svh_ctx* ctx, const double* dXT, int64_t n64, int d, int64_t first,
                    int64_t* dorder) {
  SVH_REQUIRE(n64 < INT_MAX / 2, "maximin: n too large");
  const int n = (int)n64;
  if (n == 0) return;
  const int nslice = (n + kSlice - 1) / kSlice;
  double* s2 = ctx->arena.get_t<double>("mm.s2", n);
  double* dist = ctx->arena.get_t<double>("mm.dist", n);
  int* acc = ctx->arena.get_t<int>("mm.acc", kAccCap);
  MMState* st = ctx->arena.get_t<MMState>("mm.st", 1);
  double* ck = ctx->arena.get_t<double>("mm.ck", (size_t)nslice * (kSliceTop + 1));
  int* ci = ctx->arena.get_t<int>("mm.ci", (size_t)nslice * (kSliceTop + 1));
  hipLaunchKernelGGL(mm_init_kernel, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, n, s2,
                     dist, acc, st, dorder, (int)first);
  SVH_HIP(hipGetLastError());
  if (n == 1) return;
  const size_t apply_lds = (size_t)kAccCap * d * sizeof(double);
  int sortcap = 1;
  while (sortcap < nslice * kSliceTop && sortcap < kSortCap) sortcap <<= 1;
  if (sortcap < kPool) sortcap = kPool;
  const size_t select_lds =
      (size_t)sortcap * sizeof(double) + (size_t)((sortcap + 1) / 2) * sizeof(double);
  SVH_REQUIRE(select_lds <= 150 * 1024, "maximin: d too large for the selection pool");
  SVH_REQUIRE(apply_lds <= 64 * 1024 && d <= kMMD, "maximin: d above 12 not supported");
  MMState* hst = static_cast<MMState*>(ctx->stage.get(sizeof(MMState)));
  // enqueue rounds in batches; each round accepts at least one point
  int batch = 8;
  hipEvent_t mm_t0 = nullptr, mm_t1 = nullptr;
  if (ctx->timers.enabled) {
    SVH_HIP(hipEventCreate(&mm_t0));
    SVH_HIP(hipEventCreate(&mm_t1));
    SVH_HIP(hipEventRecord(mm_t0, ctx->stream));
  }
  for (;;) {
    for (int r = 0; r < batch; ++r) {
      ctx->timers.begin(ctx->stream, "maximin_apply");
      hipLaunchKernelGGL(mm_apply_kernel, dim3(nslice), dim3(kSlice), apply_lds, ctx->stream,
                         dXT, n, d, s2, dist, acc, st, ck, ci);
      ctx->timers.end(ctx->stream);
      ctx->timers.begin(ctx->stream, "maximin_select");
      hipLaunchKernelGGL(mm_select_kernel, dim3(1), dim3(kWave), select_lds, ctx->stream, dXT,
                         n, d, s2, nslice, ck, ci, acc, st, dorder, sortcap);
      ctx->timers.end(ctx->stream);
    }
    SVH_HIP(hipGetLastError());
    SVH_HIP(hipMemcpyAsync(hst, st, sizeof(MMState), hipMemcpyDeviceToHost, ctx->stream));
    SVH_HIP(hipStreamSynchronize(ctx->stream));
    if (hst->count >= n) break;
    if (batch < 64) batch *= 2;
  }
  if (mm_t0) {
    SVH_HIP(hipEventRecord(mm_t1, ctx->stream));
    SVH_HIP(hipEventSynchronize(mm_t1));
    float ms = 0.f;
    SVH_HIP(hipEventElapsedTime(&ms, mm_t0, mm_t1));
    auto& acc_t = ctx->timers.acc["maximin"];
    acc_t.ms += ms;
    acc_t.count += 1;
    (void)hipEventDestroy(mm_t0);
    (void)hipEventDestroy(mm_t1);
  }
  SVH_HIP(hipStreamSynchronize(ctx->stream));
  ctx->timers.collect();
  ctx->mm_rounds = hst->rounds;
}

void launch_ordered_knn(svh_ctx* ctx, const double* dXT, int64_t n, int d, int m, int64_t start,
                        int64_t* dnn) {
  launch_knn_any(ctx, dXT, n, nullptr, 0, d, m, 1, start < 1 ? 1 : start, dnn, m + 1);
}

void launch_knn_query(svh_ctx* ctx, const double* dXrefT, int64_t nref, const double* dXqT,
                      int64_t nq, int d, int k, int64_t* dnb, int nb_stride) {
  launch_knn_any(ctx, dXrefT, nref, dXqT, nq, d, k, 0, 0, dnb, nb_stride);
}
